"""Output error of the NL4D forms against the oracle's FP64 restatement, per plane.

On a crop-sized clip of the benchmark's content and parameters (10-bit 4:2:0, temporal radius 3,
spatial radius 12, stride 3, groups of up to 16, hard threshold 2.7 sigma), the FP32 form and the
mixed form (f16 matrix-core transforms) are compared with the FP64 oracle of the same method:
PSNR of each against the FP64 output, the maximum absolute difference in codes, and the PSNR loss
against the clean scene (the FP64 output's PSNR minus the form's), which must stay within
0.05 dB.  The figures are printed for the record.
"""
from __future__ import annotations

import numpy as np
import pytest

import avd
from avd import _lib as L
from gpu_helpers import run_engine, words_dev

import oracle as O


def _psnr(a, b, peak):
    mse = float(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2))
    return 99.0 if mse == 0 else 10.0 * np.log10(peak * peak / mse)


@pytest.mark.gpu
def test_forms_against_the_fp64_oracle_on_the_benchmark_method():
    W, H, depth, r, S, step, K = 192, 128, 10, 3, 12, 3, 16
    frames = 8
    sig_y, sig_c = 16.0, 8.0
    peak = float((1 << depth) - 1)
    noisy = avd.synth.clip(frames, W, H, depth, 0xA5D00004, chroma="420", sigma_luma=sig_y, sigma_chroma=sig_c,
                           lo_hi=(64, 940))
    clean = avd.synth.clip(frames, W, H, depth, 0xA5D00004, chroma="420", sigma_luma=0.0, sigma_chroma=0.0,
                           lo_hi=(64, 940))
    geom = L.Geometry(W, H, 1, L.SampleFormat.u16(depth), L.SampleFormat.u16(depth), 0)
    dims = [(W, H), (W // 2, H // 2), (W // 2, H // 2)]

    def opts(sigma, form):
        return avd.nl4d_options(temporal_radius=r, spatial_radius=S, sigma=sigma, lambda_ht=2.7, pooled_threshold=0,
                                noise_map=0, motion={"estimation": 1}, field_lambda=0.0, step=step, group_size=K,
                                transform_form=form)

    outs = {}
    for form in (0, 1):
        eng = avd.build_engine("nl4d", geom, luma=opts(sig_y / peak, form), chroma=opts(sig_c / peak, form))
        outs[form] = run_engine(eng, [[words_dev(p.numpy(), depth) for p in f] for f in noisy], depth, dims)

    def spec(sigma):
        return O.Nl4dSpec(temporal_radius=r, refine=2, spatial_radius=S, lambda_ht=2.7, c_min=0.05, kaiser_beta=2.0,
                          sigma=sigma, motion=O.MotionGeom(0, 0, 16, 8, 4, 2), step=step, group_size=K)

    pl = O.Nl4dPass(spec(sig_y / peak), W, H, 1, depth, depth, precision="f64")
    pc = O.Nl4dPass(spec(sig_c / peak), W // 2, H // 2, 2, depth, depth, precision="f64")
    O.link_motion(pl, pc, "420")
    f64_l, _ = O.run_stream(pl, [[O.pack_plane(f[0].numpy(), depth)] for f in noisy])
    f64_c, _ = O.run_stream(pc, [[O.pack_plane(f[1].numpy(), depth), O.pack_plane(f[2].numpy(), depth)] for f in noisy])
    f64 = [[f64_l[i][0], f64_c[i][0], f64_c[i][1]] for i in range(frames)]
    for k, (w, h) in enumerate(dims):
        ref = np.concatenate([clean[i][k].numpy().reshape(-1) for i in range(frames)])
        oracle64 = np.concatenate([O.unpack_plane(f64[i][k], depth, w * h) for i in range(frames)])
        p64 = _psnr(oracle64, ref, peak)
        for form, label in ((0, "FP32"), (1, "mixed")):
            got = np.concatenate([O.unpack_plane(outs[form][i][k], depth, w * h) for i in range(frames)])
            loss = p64 - _psnr(got, ref, peak)
            print(f"plane {k} {label}: PSNR vs FP64 oracle {_psnr(got, oracle64, peak):.2f} dB, max |diff| "
                  f"{int(np.max(np.abs(got - oracle64)))} codes, PSNR loss vs clean {loss:+.4f} dB "
                  f"(FP64 {p64:.3f} dB)", flush=True)
            assert loss <= 0.05, (k, label, loss)
