"""mct_dot_f64_noalloc (the FP64 inner product whose partial sums live in a
static device array) against mct_dot_f64, through the C ABI: the same
kernels in the same fixed reduction tree, so the results must be equal bit
for bit, on two calls in a row (the second on a smaller input) and on a
context that never ran mct_dot_f64 (whose own scratch is never allocated).
"""
import ctypes as C

import numpy as np
import pytest
import torch

from reconstruction_program_taichi_amd import _abi

pytestmark = pytest.mark.gpu


def _dot(name, ctx_h, a, b, out):
    print(f"{name}: n={a.numel()}", flush=True)
    _abi.call(name, ctx_h, C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()), a.numel(),
              C.c_void_p(out.data_ptr()))


def test_dot_noalloc_equals_dot_on_consecutive_calls():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device visible")
    lib = _abi.load()
    fresh = C.c_void_p()
    _abi.call("mct_create", 0, C.byref(fresh))
    ctx = C.c_void_p()
    _abi.call("mct_create", 0, C.byref(ctx))
    try:
        stream = torch.cuda.current_stream(0).cuda_stream
        _abi.call("mct_set_stream", fresh, C.c_void_p(stream))
        _abi.call("mct_set_stream", ctx, C.c_void_p(stream))
        g = torch.Generator(device="cuda").manual_seed(17)
        for n in (3_000_017, 1_234):  # the second call smaller than the first
            a = torch.randn(n, generator=g, device="cuda")
            b = torch.randn(n, generator=g, device="cuda")
            got = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
            first = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
            want = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
            _dot("mct_dot_f64_noalloc", fresh, a, b, first)  # context without any scratch
            _dot("mct_dot_f64_noalloc", ctx, a, b, got)
            _dot("mct_dot_f64", ctx, a, b, want)
            torch.cuda.synchronize()
            ref = float(np.dot(a.cpu().numpy().astype(np.float64), b.cpu().numpy().astype(np.float64)))
            assert torch.equal(got, want) and torch.equal(first, want), (n, got.item(), want.item())
            assert abs(got.item() - ref) <= 1e-9 * max(1.0, abs(ref)) * np.sqrt(n)
        with pytest.raises(RuntimeError, match="NULL"):
            _abi.call("mct_dot_f64_noalloc", ctx, None, C.c_void_p(b.data_ptr()), 4,
                      C.c_void_p(got.data_ptr()))
        with pytest.raises(RuntimeError, match="negative"):
            _abi.call("mct_dot_f64_noalloc", ctx, C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()), -1,
                      C.c_void_p(got.data_ptr()))
    finally:
        torch.cuda.synchronize()
        lib.mct_destroy(fresh)
        lib.mct_destroy(ctx)
