"""GPU: the short arithmetic sequences of the collision are exact over every FP32 input.  The collision's
reciprocal (rcp estimate + one Newton step, general division outside 2^-125..2^125) must equal the
correctly rounded 1/x bit for bit, and its square root the correctly rounded sqrtf on its domain, so that
the GPU's FP32 step stays bit-identical to the oracle's FP32 mirror, which divides and takes square roots
in IEEE arithmetic."""
import ctypes

import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("what,name", [(0, "recip"), (1, "sqrt")])
def test_exhaustive_arith(gpu, what, name):
    from lbm import _lib
    lib = _lib.load()
    n = lib.lbm_selftest_threads()
    bad = torch.full((n,), 7, dtype=torch.int32, device="cuda")
    first = torch.zeros(n, dtype=torch.int32, device="cuda")
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(lib.lbm_selftest(what, ctypes.c_void_p(bad.data_ptr()), ctypes.c_void_p(first.data_ptr()), st),
               "lbm_selftest")
    torch.cuda.synchronize()
    total = int(bad.to(torch.int64).sum())
    if total:
        i = int(torch.nonzero(bad)[0])
        x = int(first[i]) & 0xFFFFFFFF
        pytest.fail(f"{name}: {total} of 2^32 inputs differ; first 0x{x:08x}")
