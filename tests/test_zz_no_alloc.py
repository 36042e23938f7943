"""The allocation-free spatial noise entry point against the allocating one, through the C ABI.

avd_site_noise_spatial_noalloc keeps its partials in a static device array; two calls in a row,
the second on a smaller frame (so the array still holds the first call's larger contents), must
each match avd_site_noise_spatial bit for bit.
"""
import ctypes

import numpy as np
import pytest

import noise_frames as F
from avd import _lib as L


def _both(ring, w, h, c, frame, slot, slots):
    import torch

    from gpu_helpers import dev, host, require_gpu

    require_gpu()
    lib = L.load()
    d_ring = dev(ring)
    d_part = torch.full((lib.avd_noise_partials_len(w, h),), -1.0, dtype=torch.float32, device="cuda")
    want = torch.full((2 * slots,), -7.0, dtype=torch.float32, device="cuda")
    got = torch.full((2 * slots,), -7.0, dtype=torch.float32, device="cuda")
    L.check(lib.avd_site_noise_spatial(ctypes.c_void_p(d_ring.data_ptr()), w, h, c, frame,
                                       ctypes.c_void_p(d_part.data_ptr()), ctypes.c_void_p(want.data_ptr()),
                                       slot, None))
    L.check(lib.avd_site_noise_spatial_noalloc(ctypes.c_void_p(d_ring.data_ptr()), w, h, c, frame,
                                               ctypes.c_void_p(got.data_ptr()), slot, None))
    return host(want), host(got)


@pytest.mark.gpu
@pytest.mark.parametrize("c", [1, 2])
def test_noalloc_matches_the_allocating_entry_on_two_calls(c):
    # A 4K frame first, then a smaller ragged one whose partials the first call's overwrite.
    for name, w, h, frames, frame, slots, slot in [("uhd", 3840, 2160, 1, 0, 1, 0),
                                                   ("ragged", 70, 20, 2, 1, 3, 2)]:
        print(f"noalloc case {name} c={c}", flush=True)
        ring = np.concatenate([F.noisy_gaussian(w, h, c, 0.3 + 0.1 * f, [5 / 255, 2 / 255])
                               for f in range(frames)]).astype(np.float32)
        want, got = _both(ring, w, h, c, frame, slot, slots)
        assert np.array_equal(got, want), (name, got, want)
        assert np.all(got[2 * slot:2 * slot + c] > 0), name
