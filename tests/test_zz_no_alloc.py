"""dgmi_dot_noalloc (partial sums in the library's fixed device array, no allocation)
against dgmi_dot, through the C-ABI: two calls in a row, the second with a smaller input,
in FP64 and FP32; sizes beyond the array's limit are refused before any launch."""
import ctypes as C

import numpy as np
import pytest
import torch

from dg_nn_public_amd import _lib

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _dot(lib, fn, dtype, x, y, n):
    out = torch.full((1,), np.nan, dtype=torch.float64, device=DEV)
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = fn(dtype, n, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()),
            C.c_void_p(out.data_ptr()), s)
    assert rc == 0, lib.dgmi_last_error()
    return out


@pytest.mark.parametrize("tdtype,dt,tol", [(torch.float64, _lib.DGMI_F64, 1e-13),
                                           (torch.float32, _lib.DGMI_F32, 1e-12)],
                         ids=["f64", "f32"])
def test_dot_noalloc_matches_dgmi_dot_on_two_calls(tdtype, dt, tol):
    lib = _lib.load()
    gen = torch.Generator(device="cpu").manual_seed(7)
    n1, n2 = 3_000_001, 12_345  # second call smaller; neither a multiple of the chunk
    x = torch.randn(n1, generator=gen, dtype=torch.float64).to(DEV, tdtype)
    y = torch.randn(n1, generator=gen, dtype=torch.float64).to(DEV, tdtype)
    for n in (n1, n2):
        print(f"[case] dot_noalloc {tdtype} n={n}", flush=True)
        a = _dot(lib, lib.dgmi_dot_noalloc, dt, x, y, n)
        b = _dot(lib, lib.dgmi_dot, dt, x, y, n)
        ref = float(torch.dot(x[:n].double(), y[:n].double()).cpu())
        av, bv = float(a.cpu()), float(b.cpu())
        scale = float(torch.linalg.vector_norm(x[:n].double()) * torch.linalg.vector_norm(y[:n].double()))
        print(n, av, bv, ref)
        assert abs(av - bv) <= tol * scale
        assert abs(av - ref) <= tol * scale
    # n = 0 gives 0
    assert float(_dot(lib, lib.dgmi_dot_noalloc, dt, x, y, 0).cpu()) == 0.0


def test_dot_noalloc_refuses_sizes_beyond_its_array():
    lib = _lib.load()
    nmax = lib.dgmi_dot_noalloc_max_n()
    assert nmax == 32768 * 8192
    x = torch.zeros(4, dtype=torch.float64, device=DEV)
    out = torch.zeros(1, dtype=torch.float64, device=DEV)
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for bad in (nmax + 1, -1):
        rc = lib.dgmi_dot_noalloc(_lib.DGMI_F64, bad, C.c_void_p(x.data_ptr()),
                                  C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()), s)
        assert rc != 0
        assert b"dgmi_dot_noalloc" in lib.dgmi_last_error()
