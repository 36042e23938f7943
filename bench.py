"""NL4D throughput on one MI355X: 3840x2160 10-bit 4:2:0, temporal radius 3, spatial radius 12,
8x8 patches at stride 3 in groups of up to 16.

A step is one frame through the NL4D engine in steady state: ingest of Y, U and V, the motion
search of the six neighbours, grouping, transforms, threshold and aggregation of the luma and
chroma passes, and the normalise and egress of the finished frame.  Inputs are seeded synthetic
clips generated on the GPU before the timed region, so the planes are resident in HBM.

Prints one JSON line with frames/s, the roofline figure of the dominant kernel (its average
launch time measured with HIP events on the engine's stream) and a CPU baseline (the oracle's
FP32 restatement on this host's cores over one bounded frame of the same clip).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "av-denoise_amd"))

METRIC = "NL4D 4K 10-bit frames/s: MFMA mixed vs FP32 form at ΔPSNR≤0.05 dB; %roofline"
W, H, DEPTH, RADIUS, SPATIAL, REFINE = 3840, 2160, 10, 3, 12, 2
STEP, GROUP = 3, 16              # reference grid stride, members per group
SIGMA_Y, SIGMA_C = 16.0, 8.0     # noise of the clip, in 10-bit codes
LAMBDA = 2.7
CLIP_FRAMES = 16
# MI355X published peaks: the trajectory search, motion and filter kernels are vector work (v_dot2 /
# v_sad integer operations and f32 FMAs), priced against the FP32 rate, 157.3 TFLOP/s; the centre
# search runs its products on the matrix cores as f16 (dense 2.5 PFLOP/s).
VALU_PEAK_TOPS = 157.3
MFMA_F16_PEAK_TOPS = 2500.0
OP_STAGES = ["collab_search", "collab_centre", "motion", "collab_filter"]
STAGE_KERNELS = {
    "collab_search": "k_trajectory_codes",
    "collab_centre": "k_centre_prep + k_centre_mfma + k_centre_redo",
    "motion": "k_bm_coarse + k_bm_fine",
    "collab_filter": "k_collab_filter_mixed + k_tile_gather + k_tile_overflow",
}
STAGE_PEAK = {"collab_centre": MFMA_F16_PEAK_TOPS}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sample-steps", type=int, default=4, help="extra steps timed per stage with HIP events")
    ap.add_argument("--sample-all", action="store_true",
                    help="sample device time from the first step (every launch timed alone, the plane groups "
                         "one after the other; for kernel traces whose durations match the stage figures)")
    ap.add_argument("--form", choices=["both", "mixed", "fp32"], default="both",
                    help="transform form(s) to run; the headline value is the mixed form's when it runs")
    ap.add_argument("--no-quality", action="store_true", help="skip the PSNR comparison against the clean scene")
    ap.add_argument("--no-nlm", action="store_true", help="skip the NLMeans workloads")
    ap.add_argument("--nlm-only", action="store_true",
                    help="run only the NLMeans workloads and print their JSON (for kernel traces of them)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="after the run, one plain device copy of this many MiB (a known-size kernel that "
                         "hardware-counter runs calibrate FETCH_SIZE / WRITE_SIZE against)")
    return ap.parse_args()


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "nl4d_4k_stride3_k16_traffic.json")


def stage_traffic():
    """HBM bytes per stage launch (fetch + write) of this workload, from the committed counter
    summary (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, each scaled by the same counter on a
    256 MiB device copy); empty when the summary is absent."""
    try:
        with open(TRAFFIC_FILE) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return {}
    return {k: v["bytes_per_launch"] for k, v in data.get("stages", {}).items()}


FORMS = {"fp32": 0, "mixed": 1}


def make_engine(avd, form="mixed"):
    geom = avd.Geometry(W, H, 1, avd.SampleFormat.u16(DEPTH), avd.SampleFormat.u16(DEPTH), 0)
    common = dict(temporal_radius=RADIUS, spatial_radius=SPATIAL, refine=REFINE, lambda_ht=LAMBDA,
                  pooled_threshold=0, noise_map=0, motion={"estimation": 1}, field_lambda=0.0, step=STEP,
                  group_size=GROUP, transform_form=FORMS[form])
    luma = avd.nl4d_options(sigma=SIGMA_Y / 1023.0, **common)
    chroma = avd.nl4d_options(sigma=SIGMA_C / 1023.0, **common)
    return avd.build_engine("nl4d", geom, luma=luma, chroma=chroma), geom


def gpu_clip(avd, torch, sigma=True):
    sy, sc = (SIGMA_Y, SIGMA_C) if sigma else (0.0, 0.0)
    clip = avd.synth.clip(CLIP_FRAMES, W, H, DEPTH, 0xA5D00004, chroma="420", sigma_luma=sy, sigma_chroma=sc,
                          device="cuda", lo_hi=(64, 940))
    frames = [[avd.synth.pack_words(p, DEPTH) for p in f] for f in clip]
    torch.cuda.synchronize()
    return clip, frames


def plane_dims():
    return [(W, H), ((W + 1) // 2, (H + 1) // 2), ((W + 1) // 2, (H + 1) // 2)]


def quality(avd, torch, frames, clean, form, first=4, count=4):
    """PSNR (dB, per plane) against the noise-free scene of output frames first..first+count-1."""
    eng, geom = make_engine(avd, form)
    sq = [0.0, 0.0, 0.0]
    n = [0, 0, 0]
    emitted = 0
    for t in range(len(frames)):
        for _ in range(eng.push(frames[t])):
            outs = [torch.zeros(avd.plane_bytes(geom.output, w, h), dtype=torch.uint8, device="cuda")
                    for (w, h) in plane_dims()]
            eng.emit_into(outs)
            if first <= emitted < first + count:
                for k, (w, h) in enumerate(plane_dims()):
                    got = outs[k].view(torch.int16)[: w * h].to(torch.int64) & 0xFFFF
                    ref = clean[emitted][k].reshape(-1).to(torch.int64)
                    sq[k] += float(((got - ref).to(torch.float64) ** 2).sum())
                    n[k] += w * h
            emitted += 1
        if emitted >= first + count:
            break
    peak = float((1 << DEPTH) - 1)
    return [10.0 * math.log10(peak * peak / (sq[k] / n[k])) for k in range(3)]


HBM_PEAK_GBPS = 8000.0


def noise_spatial_measure(avd, torch, clip, calls=20):
    """The spatial noise measurement (both Immerkaer stages) on the clip's first 4K luma frame as
    f32 samples in HBM: average time per call from HIP events on torch's stream (the one the
    calls run on), against the HBM roofline of one read of the frame."""
    import ctypes

    lib = avd._lib.load()
    frame = (clip[0][0].to(torch.float32) / float((1 << DEPTH) - 1)).contiguous()
    partials = torch.empty(lib.avd_noise_partials_len(W, H), dtype=torch.float32, device="cuda")
    results = torch.empty(2, dtype=torch.float32, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def call():
        avd._lib.check(lib.avd_site_noise_spatial(ctypes.c_void_p(frame.data_ptr()), W, H, 1, 0,
                                                  ctypes.c_void_p(partials.data_ptr()),
                                                  ctypes.c_void_p(results.data_ptr()), 0, stream))

    for _ in range(3):
        call()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(calls):
        call()
    end.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / calls
    nbytes = W * H * 4
    gbps = nbytes / (ms * 1e-3) / 1e9
    return {"kernels": "k_noise_partial + k_noise_reduce", "workload": "one 3840x2160 luma frame, f32 samples",
            "avg_launch_ms": round(ms, 4), "bytes_per_launch": nbytes, "achieved": round(gbps, 1),
            "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(gbps / HBM_PEAK_GBPS, 4)}


def noise_temporal_measure(avd, torch, clip, calls=20):
    """The temporal block statistics of the clip's second 4K luma frame against its first (f32
    samples in a two-frame ring in HBM, quarter fields on), with the zero vector and with a motion
    field of the engine's 8-pixel grid (the gain check on): average time per call from HIP events
    on torch's stream, against the HBM roofline of one read of both frames plus the records."""
    import ctypes

    lib = avd._lib.load()
    ring = torch.stack([clip[f][0].to(torch.float32) / float((1 << DEPTH) - 1) for f in (0, 1)]).contiguous()
    n = lib.avd_noise_temporal_stats_len(W, H, 1)
    stats = torch.empty(n, dtype=torch.float32, device="cuda")
    mbx, mby = -(-W // 8), -(-H // 8)
    field = torch.randint(-2, 3, (mbx * mby * 2,), dtype=torch.int32, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    out = {}
    for name, mv in (("zero_vector", None), ("motion_field", field)):
        def call():
            avd._lib.check(lib.avd_site_noise_temporal_stats(
                ctypes.c_void_p(ring.data_ptr()), W, H, 1, 1, 0,
                None if mv is None else ctypes.c_void_p(mv.data_ptr()), 0, 8, mbx, mby, 1,
                ctypes.c_void_p(stats.data_ptr()), stream))

        for _ in range(3):
            call()
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(calls):
            call()
        end.record()
        torch.cuda.synchronize()
        ms = start.elapsed_time(end) / calls
        nbytes = 2 * W * H * 4 + n * 4
        gbps = nbytes / (ms * 1e-3) / 1e9
        out[name] = {"avg_launch_ms": round(ms, 4), "bytes_per_launch": nbytes, "achieved": round(gbps, 1),
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(gbps / HBM_PEAK_GBPS, 4)}
    return {"kernels": "k_temporal_stats", "workload": "3840x2160 luma frame against the previous one, f32 samples, "
            "quarter fields on", **out}


def grain_measure_measure(avd, torch, clip, calls=20):
    """The film-grain measurement of the clip's first 4K luma frame against its second (f32 samples
    in a two-frame ring in HBM, zero vectors at full confidence on the 8-pixel grid) with a flat
    denoised frame, so every interior cell is accepted and runs its halo and 46 lag sums (the
    kernel's heaviest path): average time per call from HIP events on torch's stream, against the
    HBM roofline of one read of both frames and the denoised frame plus the partials."""
    import ctypes

    lib = avd._lib.load()
    ring = torch.stack([clip[f][0].to(torch.float32) / float((1 << DEPTH) - 1) for f in (0, 1)]).contiguous()
    denoised = torch.full((H, W), 0.5, dtype=torch.float32, device="cuda")
    bx, by = -(-W // 8), -(-H // 8)
    mv = torch.zeros(bx * by * 2, dtype=torch.int32, device="cuda")
    conf = torch.ones(bx * by, dtype=torch.float32, device="cuda")
    lo, hi = 0.05 / 255.0, 32.0 / 255.0
    edges = torch.tensor([lo * (hi / lo) ** (i / 64) for i in range(65)], dtype=torch.float32, device="cuda")
    hist = torch.zeros(2048, dtype=torch.int32, device="cuda")
    n = lib.avd_grain_partials_len(W, H)
    partials = torch.empty(n, dtype=torch.float32, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    p = lambda t: ctypes.c_void_p(t.data_ptr())

    def call():
        avd._lib.check(lib.avd_site_grain_measure(p(ring), p(denoised), W, H, 1, 0, 0, 1, p(mv), p(conf), bx, by, 8,
                                                  0, 1, None, 0, 0, 1, p(edges), p(hist), p(partials), stream))

    for _ in range(3):
        call()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(calls):
        call()
    end.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / calls
    nbytes = 3 * W * H * 4 + n * 4
    gbps = nbytes / (ms * 1e-3) / 1e9
    chunk = torch.zeros(16 * 47, dtype=torch.float32, device="cuda")

    def reduce_call():
        avd._lib.check(lib.avd_site_grain_reduce(p(partials), n // 48, p(chunk), 0, stream))

    for _ in range(3):
        reduce_call()
    start.record()
    for _ in range(calls):
        reduce_call()
    end.record()
    torch.cuda.synchronize()
    reduce_ms = start.elapsed_time(end) / calls
    records = chunk.view(16, 47).to(torch.float64).contiguous()
    fit_out = [torch.empty(16, 47, dtype=torch.float64, device="cuda"),
               torch.empty(16, 24, dtype=torch.float64, device="cuda"),
               torch.empty(16, 24, dtype=torch.int32, device="cuda"),
               torch.empty(16, dtype=torch.int32, device="cuda"), torch.empty(16, dtype=torch.int32, device="cuda")]

    def fit_call():
        avd._lib.check(lib.avd_site_grain_fit(p(records), 16, *[p(t) for t in fit_out], stream))

    for _ in range(3):
        fit_call()
    start.record()
    for _ in range(calls):
        fit_call()
    end.record()
    torch.cuda.synchronize()
    fit_ms = start.elapsed_time(end) / calls
    # Persistence: the frame's cell projections into one of 9 history slots, then the newest slot
    # against the 8 before it.
    patterns = [0x9966669999666699, 0x9669699669969669, 0xA55AA55A5AA55AA5, 0xAA55AA55AA55AA55,
                0xA55A5AA5A55A5AA5, 0x9966996666996699, 0xAA5555AA55AAAA55, 0x9669966996699669]
    words = torch.tensor([w - (1 << 32) if w >= (1 << 31) else w for q in patterns
                          for w in (q & 0xFFFFFFFF, q >> 32)], dtype=torch.int32, device="cuda")
    cells = (W // 8) * (H // 8)
    history = torch.zeros(9 * cells * 9, dtype=torch.float32, device="cuda")
    persist_partials = torch.empty(lib.avd_grain_persist_partials_len(cells, 8), dtype=torch.float32, device="cuda")
    persist_record = torch.zeros(24, dtype=torch.float32, device="cuda")

    def project_call(slot):
        avd._lib.check(lib.avd_site_grain_project(p(ring), p(denoised), p(words), p(history), W, H, 1, 0, 0, slot, 1,
                                                  stream))

    def persist_call():
        avd._lib.check(lib.avd_site_grain_persist(p(history), p(persist_partials), p(persist_record), 0, 0, cells, 8,
                                                  stream))

    for slot in range(9):
        project_call(slot)
    timings = {}
    for name, fn in (("project", lambda: project_call(0)), ("persist", persist_call)):
        start.record()
        for _ in range(calls):
            fn()
        end.record()
        torch.cuda.synchronize()
        timings[name] = start.elapsed_time(end) / calls
    return {"kernels": "k_grain_measure", "workload": "3840x2160 luma, flat denoised frame (every interior cell "
            "accepted), zero vectors", "avg_launch_ms": round(ms, 4), "bytes_per_launch": nbytes,
            "achieved": round(gbps, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": round(gbps / HBM_PEAK_GBPS, 4),
            "reduce": {"kernels": "k_grain_reduce", "avg_launch_ms": round(reduce_ms, 4), "bytes_per_launch": n * 4,
                       "achieved": round(n * 4 / (reduce_ms * 1e-3) / 1e9, 1), "unit": "GB/s"},
            "fit": {"kernels": "k_grain_fit", "records": 16, "dtype": "f64", "avg_launch_ms": round(fit_ms, 4)},
            "project": {"kernels": "k_grain_project", "avg_launch_ms": round(timings["project"], 4),
                        "bytes_per_launch": 2 * W * H * 4 + cells * 9 * 4},
            "persist": {"kernels": "k_grain_persist + k_grain_persist_reduce", "lags": 8,
                        "avg_launch_ms": round(timings["persist"], 4), "bytes_per_launch": 9 * cells * 9 * 4}}


def run_steps(eng, frames, outs, start, count):
    done = 0
    for i in range(count):
        ready = eng.push(frames[(start + i) % len(frames)])
        for _ in range(ready):
            eng.emit_into(outs)
            done += 1
    return done


CPU_CROP = 4  # the CPU baseline's sample: the top-left 1/CPU_CROP of each axis (1/16 of the pixels)


def cpu_baseline(clip_codes, threads, runs=5):
    """One steady-state frame (luma + chroma passes) of the top-left 960x540 sixteenth of the same
    clip through the oracle's FP32 restatement, on `threads` host threads: one warm-up, then the
    median of `runs` timed runs; returns seconds per full frame (the crop's time times sixteen)."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    O.set_threads(threads)
    N = 2 * RADIUS + 1

    def one_frame():
        t_total = 0.0
        passes = []
        for plane_group in ("luma", "chroma"):
            if plane_group == "luma":
                w, h, C, sig = W // CPU_CROP, H // CPU_CROP, 1, SIGMA_Y / 1023.0
                planes = [[f[0][:h, :w]] for f in clip_codes[:N]]
            else:
                w, h, C, sig = W // (2 * CPU_CROP), H // (2 * CPU_CROP), 2, SIGMA_C / 1023.0
                planes = [[f[1][:h, :w], f[2][:h, :w]] for f in clip_codes[:N]]
            spec = O.Nl4dSpec(temporal_radius=RADIUS, refine=REFINE, spatial_radius=SPATIAL, lambda_ht=LAMBDA,
                              c_min=0.05, kaiser_beta=2.0, sigma=sig, motion=O.MotionGeom(0, 0, 16, 8, 4, 2),
                              step=STEP, group_size=GROUP)
            p = O.Nl4dPass(spec, w, h, C, DEPTH, DEPTH)
            for fr in planes:  # fill the ring without running passes
                p.ring.write([O.pack_plane(np.ascontiguousarray(x), DEPTH) for x in fr])
            passes.append(p)
        O.link_motion(passes[0], passes[1], "420")  # chroma refines the luma pass's motion
        for p in passes:
            w, h, C = p.W, p.H, p.C
            t0 = time.perf_counter()
            p._run_pass(RADIUS, "whole")
            slot = p.ring.phys(0)
            out = p._read_region(slot)
            O.egress(out, w * h, C, DEPTH)
            t_total += time.perf_counter() - t0
        return t_total

    one_frame()  # warm-up
    times = sorted(one_frame() for _ in range(runs))
    return float(CPU_CROP * CPU_CROP) * times[runs // 2]


NLM_CONFIGS = {
    # Plain NLMeans, spatial, luma only, 8-bit 640x360, patch radius 2, search radius 7.
    "spatial_640x360_y8": dict(W=640, H=360, depth=8, chroma="gray", frames=8, r=0, P=2, S=7, sigma=(10.0, 0.0),
                               steps=64),
    # Temporal NLMeans, 1920x1080 10-bit 4:2:0, temporal radius 2, patch radius 3, search radius 4.
    "temporal_1080p_420_10bit": dict(W=1920, H=1080, depth=10, chroma="420", frames=32, r=2, P=3, S=4,
                                     sigma=(12.0, 8.0), steps=32),
}


def nlm_candidate_flops(C, P):
    """Operations per candidate pixel of the NLM window (squared differences, channel scale,
    separable box sum, weight and accumulation): 5C + 2(2P + 1) + 7; the exp is counted apart."""
    return 5 * C + 2 * (2 * P + 1) + 7


def nlm_leg(avd, torch, name, cfg, sample_steps=4):
    """One NLMeans workload in steady state: frames/s, MPix/s and the window stages' achieved rate
    against the FP32 vector roofline, with the exp count reported apart."""
    W_, H_, depth, r, P, S = cfg["W"], cfg["H"], cfg["depth"], cfg["r"], cfg["P"], cfg["S"]
    chroma = cfg["chroma"]
    layout = {"gray": 0, "420": 1}[chroma]
    fmt = avd.SampleFormat.u8() if depth == 8 else avd.SampleFormat.u16(depth)
    geom = avd.Geometry(W_, H_, layout, fmt, fmt, 0)
    opts = dict(hq=0, temporal=1 if r else 0, temporal_radius=r, patch_radius=P, search_radius=S, strength=1.2,
                self_weight=1.0)
    top = (1 << depth) - 1
    clip = avd.synth.clip(cfg["frames"], W_, H_, depth, 0xA5D00000 + (1 if r == 0 else 2), chroma=chroma,
                          sigma_luma=cfg["sigma"][0], sigma_chroma=cfg["sigma"][1], device="cuda",
                          lo_hi=(0, top) if depth == 8 else (64, 940))
    frames = [[avd.synth.pack_words(p, depth) for p in f] for f in clip]
    dims = avd.plane_dims(geom)
    outs = [torch.zeros(avd.plane_bytes(fmt, w, h), dtype=torch.uint8, device="cuda") for (w, h) in dims]
    warm = max(2 * r + 1, 4)
    n_q = (2 * S + 1) ** 2
    cands = n_q * (2 * r + 1) - 1
    exps = sum(w * h * cands for (w, h) in dims[:1]) + (dims[1][0] * dims[1][1] * cands if layout else 0)
    forms = {}
    for form in ("plain", "batched"):
        eng = avd.build_engine("nlmeans", geom, luma=avd.nlmeans_options(**opts),
                               chroma=avd.nlmeans_options(**opts) if layout else None)
        eng.set_nlm_form(form)
        run_steps(eng, frames, outs, 0, warm)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        emitted = run_steps(eng, frames, outs, warm, cfg["steps"])
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        assert emitted == cfg["steps"], (emitted, cfg["steps"])
        eng.sample_device_time(1)
        run_steps(eng, frames, outs, warm + cfg["steps"], sample_steps)
        st = eng.timing()
        eng.sample_device_time(0)
        window = ("nlm_window",) if form == "batched" else ("nlm_pair", "nlm_single")
        window_ms = sum(st[k]["ms"] for k in window) / sample_steps
        flops = sum(st[k]["work"] for k in window) / sample_steps
        tflops = flops / (window_ms * 1e-3) / 1e12
        fps = cfg["steps"] / elapsed
        forms[form] = {
            "frames_per_s": round(fps, 2), "mpix_per_s": round(fps * W_ * H_ / 1e6, 1),
            "ms_per_frame": round(1000.0 / fps, 3), "window_ms_per_frame": round(window_ms, 4),
            "flops_per_frame": int(flops),
            "roofline": {"kernel": "k_nlm_window (every window and the finish)" if form == "batched"
                         else "k_nlm (pair + single windows)", "bound": "valu", "achieved": round(tflops, 3),
                         "peak": VALU_PEAK_TOPS, "unit": "TFLOP/s", "frac": round(tflops / VALU_PEAK_TOPS, 4)},
            "exp_per_s": round(exps / (window_ms * 1e-3), 1),
            "stages_ms_per_frame": {k: round(v["ms"] / sample_steps, 4) for k, v in st.items() if v["launches"]},
        }
    # The engine's default form: batched for a spatial pass, plain for a temporal one.
    b = forms["batched" if r == 0 else "plain"]
    return {
        "workload": f"NLMeans {'temporal radius %d' % r if r else 'spatial'}, {W_}x{H_} {depth}-bit "
                    f"{'luma' if layout == 0 else '4:2:0'}, patch radius {P}, search radius {S}, FP32; plain form "
                    "(one launch per pair, single window and finish, per plane group) and batched form (every "
                    "window and the finish of a plane group in one launch), bit-identical outputs; headline figures "
                    "are the engine's default form (batched for spatial, plain for temporal passes)",
        "frames_per_s": b["frames_per_s"], "mpix_per_s": b["mpix_per_s"], "ms_per_frame": b["ms_per_frame"],
        "roofline": b["roofline"], "exp_per_frame": int(exps),
        "batched_vs_plain": round(forms["batched"]["frames_per_s"] / forms["plain"]["frames_per_s"], 3),
        "forms": forms,
        "tuned": [e for e in avd.tuning_entries() if e["key"].endswith(f"{W_}x{H_}") or f"{W_}x{H_} " in e["key"]
                  or f"{W_}x{H_} r" in e["key"]],
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (one MI355X per box)", file=sys.stderr)
        sys.exit(2)
    import torch

    import avd

    torch.cuda.set_device(0)
    if args.nlm_only:
        print(json.dumps({"nlm": {name: nlm_leg(avd, torch, name, cfg) for name, cfg in NLM_CONFIGS.items()}}))
        return
    clip, frames = gpu_clip(avd, torch)
    warm = max(args.warmup, 2 * RADIUS + 1)

    def timed_run(form):
        eng, geom = make_engine(avd, form)
        outs = [torch.zeros(avd.plane_bytes(geom.output, w, h), dtype=torch.uint8, device="cuda")
                for (w, h) in plane_dims()]
        if args.sample_all:
            eng.sample_device_time(1)
        run_steps(eng, frames, outs, 0, warm)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        emitted = run_steps(eng, frames, outs, warm, args.steps)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        assert emitted == args.steps, (emitted, args.steps)
        return eng, outs, elapsed

    # The FP32 form (bit-exact with the reference's f32 arithmetic) and the mixed form (f16 matrix
    # transforms): frames/s of each, and PSNR against the noise-free scene on the same frames.
    forms = {}
    for form in (["fp32", "mixed"] if args.form == "both" else [args.form]):
        eng, outs, elapsed = timed_run(form)
        forms[form] = {"value": round(args.steps / elapsed, 3), "elapsed": elapsed, "eng": eng, "outs": outs}
    headline = "mixed" if "mixed" in forms else "fp32"
    eng, outs, elapsed = forms[headline]["eng"], forms[headline]["outs"], forms[headline]["elapsed"]
    fps = args.steps / elapsed
    psnr = {}
    if not args.no_quality:
        clean, _ = gpu_clip(avd, torch, sigma=False)
        for form in forms:
            psnr[form] = [round(x, 4) for x in quality(avd, torch, frames, clean, form)]
        del clean

    # Per-stage device time over a few more steps, with HIP events on the engine's stream; the
    # engine counts each stage's algorithmic operations (both search stages: 192 per candidate
    # patch and channel; motion: 3 per pixel of every scored block offset; filter: 16 members x
    # (C x 4928 + 64 x (C + 1)) per reference).
    if not args.sample_all:
        eng.sample_device_time(1)
    run_steps(eng, frames, outs, warm + args.steps, args.sample_steps)
    stages = eng.timing()
    eng.sample_device_time(0)
    if args.copy_mib > 0:
        src = torch.full((args.copy_mib << 20,), 7, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    rooflines = {}
    for name in OP_STAGES:
        st = stages[name]
        if not st["launches"] or st["ms"] <= 0:
            continue
        avg_ms = st["ms"] / st["launches"]
        ops = st["work"] / st["launches"]
        tops = ops / (avg_ms * 1e-3) / 1e12
        peak = STAGE_PEAK.get(name, VALU_PEAK_TOPS)
        rooflines[name] = {"kernels": STAGE_KERNELS[name], "avg_launch_ms": round(avg_ms, 4),
                           "ops_per_launch": int(ops), "achieved": round(tops, 3), "peak": peak,
                           "frac": round(tops / peak, 4)}
    dominant = max(rooflines, key=lambda k: stages[k]["ms"]) if rooflines else None
    dom = rooflines.get(dominant, {"kernels": "", "avg_launch_ms": None, "ops_per_launch": None, "achieved": None,
                                   "peak": VALU_PEAK_TOPS, "frac": None})
    traffic = stage_traffic().get(dominant) if dominant else None

    result = {
        "metric": METRIC,
        "value": round(fps, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": warm,
        "ms_per_step": round(1000.0 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("int32 search (f16 matrix-core products, exact) + " +
                  ("f16 matrix-core DCTs with f32 accumulation (mixed form)" if headline == "mixed"
                   else "f32 filter (FP32 form)")),
        "data": "synthetic (seeded moving sinusoid/gradient scene, Gaussian noise 16/8 codes, generated on GPU)",
        "config": {
            "workload": "NL4D 3840x2160 10-bit 4:2:0, 16 frames, temporal radius 3, spatial radius 12, refine 2, "
                        "8x8 patches at stride 3, groups of up to 16 (volume grids of 4 frames x 4 volumes), "
                        "hard threshold 2.7 sigma (sigma fixed at the clip's noise, 16/1023 luma and 8/1023 "
                        "chroma: noise estimation is not built), motion 16x16/8 blocks radius 4 two-level "
                        "Direct, field_lambda 0 (no truemotion stage); " + headline + " form",
            "global_batch": 1,
            "frames_in_clip": CLIP_FRAMES,
            "mpix_per_s": round(fps * W * H / 1e6, 3),
        },
        "roofline": {
            "kernel": f"{dominant} ({dom['kernels']})",
            # The centre search's products run on the matrix cores (f16); every other stage is
            # vector work (integer v_dot2 / v_sad operations, f32 FMAs) priced at the FP32 rate.
            "bound": "mfma" if dominant == "collab_centre" else "valu",
            "achieved": dom["achieved"],
            "peak": dom["peak"],
            "unit": "TFLOP/s" if dominant == "collab_centre" else "TOP/s",
            "frac": dom["frac"],
            "traffic": traffic,
            "avg_launch_ms": dom["avg_launch_ms"],
            "ops_per_launch": dom["ops_per_launch"],
        },
        "stage_rooflines": rooflines,
        "stages_ms_per_frame": {k: round(v["ms"] / max(args.sample_steps, 1), 4) for k, v in stages.items()
                                if v["launches"]},
        "dominant_stage": dominant,
        "forms": {f: {"frames_per_s": forms[f]["value"], "psnr_db_vs_clean": psnr.get(f)} for f in forms},
    }
    result["next_rows"] = {"noise_spatial": noise_spatial_measure(avd, torch, clip),
                           "noise_temporal": noise_temporal_measure(avd, torch, clip),
                           "grain_measure": grain_measure_measure(avd, torch, clip)}
    if not args.no_nlm:
        result["nlm"] = {name: nlm_leg(avd, torch, name, cfg) for name, cfg in NLM_CONFIGS.items()}
    if "mixed" in forms and "fp32" in forms:
        result["mixed_vs_fp32"] = round(forms["mixed"]["value"] / forms["fp32"]["value"], 4)
        if psnr:
            result["delta_psnr_db"] = [round(m - f, 4) for m, f in zip(psnr["mixed"], psnr["fp32"])]

    if not args.no_cpu_baseline:
        import numpy as np

        threads = min(os.cpu_count() or 1, 16)
        codes = [[p.cpu().numpy().astype(np.int32) for p in f] for f in clip[: 2 * RADIUS + 1]]
        secs = cpu_baseline(codes, threads)
        result["cpu_baseline"] = {
            "value": round(1.0 / secs, 5),
            "unit": "frames/s",
            "cores": threads,
            "kind": "port",
            "sample": "one steady-state frame (luma + chroma passes: motion of 6 neighbours, grouping at stride 3 "
                      "in groups of up to 16, filter, normalise, egress) over the top-left 960x540 sixteenth of "
                      "the same clip through the oracle's FP32 restatement: median of 5 runs after one warm-up, "
                      "time x16 for the full 3840x2160 frame",
        }
    print(json.dumps(result))


if __name__ == "__main__":
    main()
