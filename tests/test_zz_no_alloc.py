"""mind_mlp_forward_noalloc (activations in the library's static scratch array,
device input and output owned by the caller) against mind_mlp_forward, bit for
bit, on two calls in a row, the second with a smaller batch; and the refusal of
a batch beyond the scratch limit."""
import ctypes as C

import numpy as np
import pytest

import mlp_oracle as MO
from mind_amd import _abi
from mind_amd import surrogate as S

pytestmark = pytest.mark.gpu


class _Hip:
    """The HIP runtime libmind links (device buffers owned by the test, the caller)."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]

    def malloc(self, nbytes):
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def to_device(self, dst, arr):
        assert self.rt.hipMemcpy(dst, arr.ctypes.data, arr.nbytes, 1) == 0  # hipMemcpyHostToDevice

    def to_host(self, arr, src):
        assert self.rt.hipMemcpy(arr.ctypes.data, src, arr.nbytes, 2) == 0  # hipMemcpyDeviceToHost


def test_noalloc_forward_matches_forward(gpu):
    dims = (52, 256, 3, 2)
    L = S._lib()
    hip = _Hip()
    g = np.random.default_rng(11)
    p = MO.init_params(*dims, seed=3)
    p = (p + 0.02 * g.standard_normal(p.size)).astype(np.float32)
    net = S.MLP(*dims, max_batch=512)
    net.set_params(p)
    dx, dy = hip.malloc(512 * dims[0] * 4), hip.malloc(512 * dims[3] * 4)
    try:
        for batch in (300, 37):  # the second call is smaller: stale scratch rows must not leak in
            print(f"case: noalloc forward batch={batch}", flush=True)
            x = g.standard_normal((batch, dims[0])).astype(np.float32)
            want = net.forward(x)
            hip.to_device(dx, x)
            assert hip.rt.hipMemset(dy, 0xFF, 512 * dims[3] * 4) == 0
            _abi.check(L.mind_mlp_forward_noalloc(net.h, dx, batch, dy, None))
            got = np.zeros((batch, dims[3]), np.float32)
            hip.to_host(got, dy)
            assert np.array_equal(got, want), batch
            rest = np.zeros(((512 - batch) * dims[3],), np.float32)
            hip.to_host(rest, C.c_void_p(dy.value + batch * dims[3] * 4))
            assert np.all(rest.view(np.uint32) == 0xFFFFFFFF)  # nothing written past the batch
        rc = L.mind_mlp_forward_noalloc(net.h, dx, 4096, dy, None)  # beyond 2048 rows at width 256
        assert rc == 1  # MIND_EARG, refused before any launch
        assert "scratch limit" in _abi.lib().mind_last_error().decode()
    finally:
        hip.rt.hipFree(dx)
        hip.rt.hipFree(dy)
