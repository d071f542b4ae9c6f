"""The allocation-free filter entry point (ffip_to_convolve_noalloc) against
ffip_to_convolve through the C ABI: device arrays from torch, the call on
torch's current stream, two calls in a row (the second on a smaller array, so
the scratch array still holds the first call's weights beyond it), both
directions; results equal bit for bit, and sizes beyond the scratch limit are
rejected."""
import ctypes

import numpy as np
import pytest

from ffip_amd._lib import FfipError, TO_convolve, TO_convolve_transpose, lib

pytestmark = pytest.mark.gpu


def _noalloc(a, f, transpose):
    import torch
    da = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    df = torch.from_numpy(np.ascontiguousarray(f)).cuda()
    out = torch.full_like(da, np.nan)
    stream = torch.cuda.current_stream().cuda_stream
    rc = lib().ffip_to_convolve_noalloc(ctypes.c_void_p(da.data_ptr()), a.shape[0], a.shape[1],
                                       ctypes.c_void_p(df.data_ptr()), f.shape[0], f.shape[1], int(transpose),
                                       ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream))
    if rc != 0:
        raise FfipError(lib().ffip_last_error().decode())
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("transpose", [False, True])
def test_noalloc_equals_allocating_entry_point(transpose):
    rng = np.random.default_rng(11)
    ref = TO_convolve_transpose if transpose else TO_convolve
    for shape, fw in (((173, 129), 7), ((41, 37), 5)):
        a = rng.uniform(-1, 1, size=shape)
        f = rng.uniform(0.1, 1, size=(fw, fw))
        got = _noalloc(a, f, transpose)
        want = ref(a, f)
        assert np.array_equal(got, want), (shape, fw, np.abs(got - want).max())


def test_noalloc_rejects_oversized_arrays():
    f = np.ones((3, 3))
    with pytest.raises(FfipError, match="allocation-free limit"):
        _noalloc(np.zeros((2049, 1025)), f, False)
