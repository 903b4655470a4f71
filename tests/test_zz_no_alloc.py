"""gato_dot_batched_noalloc (reductions with a static device scratch, for
callers that cannot allocate) against gato_dot_batched (caller workspace),
called through the C ABI: identical results on consecutive calls of
decreasing size, and a clear refusal of sizes beyond its scratch."""
import ctypes as C

import pytest
import torch

pytestmark = pytest.mark.gpu


def _dots(fn_name, dtype_code, n, nk, A, lda, B, ldb, Cv, S, s_coef, scale, out, ws=None):
    from gato_amd import _lib as L
    ptr = lambda t: None if t is None else C.c_void_p(t.data_ptr())
    args = [dtype_code, n, nk, ptr(A), lda, ptr(B), ldb, ptr(Cv), ptr(S), ptr(s_coef), scale,
            ptr(out)]
    if ws is not None:
        args.append(ptr(ws))
    args.append(C.c_void_p(L.stream_ptr()))
    return getattr(L.lib(), fn_name)(*args)


def _case(dtype, N, nk, broadcast, with_c, with_s, seed):
    from gato_amd import _lib as L
    g = torch.Generator(device="cuda").manual_seed(seed)
    n = N ** 3
    A = torch.randn((nk, n), generator=g, device="cuda", dtype=dtype)
    B = torch.randn((1 if broadcast else nk, n), generator=g, device="cuda", dtype=dtype)
    Cv = torch.randn(n, generator=g, device="cuda", dtype=dtype) if with_c else None
    S = torch.randn(n, generator=g, device="cuda", dtype=dtype) if with_s else None
    s_coef = torch.tensor([0.37], device="cuda", dtype=torch.float64) if with_s else None
    ldb = 0 if broadcast else n
    code = L.dtype_code(dtype)
    ref = torch.full((nk,), float("nan"), device="cuda", dtype=torch.float64)
    got = torch.full((nk,), float("nan"), device="cuda", dtype=torch.float64)
    print(f"case dtype={dtype} N={N} nk={nk} broadcast={broadcast} C={with_c} S={with_s}",
          flush=True)
    assert _dots("gato_dot_batched", code, n, nk, A, n, B, ldb, Cv, S, s_coef, 0.125, ref,
                 ws=L.workspace(A.device)) == 0
    assert _dots("gato_dot_batched_noalloc", code, n, nk, A, n, B, ldb, Cv, S, s_coef, 0.125,
                 got) == 0, L.lib().gato_last_error()
    torch.cuda.synchronize()
    assert torch.equal(got, ref), (got - ref).abs().max()
    # and against a plain double reference of the same sums
    Bd = B.double() - (0.37 * S.double() if with_s else 0.0)
    prod = A.double() * Bd * (Cv.double() if with_c else 1.0)
    want = 0.125 * prod.sum(dim=1)
    assert torch.allclose(got, want, rtol=1e-10, atol=1e-10 * float(want.abs().max())), \
        (got - want).abs().max()


def test_noalloc_dots_match_workspace_dots(gpu):
    # consecutive calls, each smaller than the one before: the scratch keeps
    # the previous call's partials, which must not leak into the next result
    _case(torch.float32, 128, 256, True, False, True, 1)    # largest accepted float broadcast
    _case(torch.float32, 96, 40, True, True, False, 2)
    _case(torch.float64, 64, 7, False, False, False, 3)     # paired (ldb != 0) path
    _case(torch.float64, 20, 3, True, False, False, 4)
    _case(torch.float32, 17, 1, False, True, False, 5)      # odd size: scalar-load path


def test_noalloc_dots_refuse_sizes_beyond_the_scratch(gpu):
    from gato_amd import _lib as L
    n, nk = 128 ** 3, 257
    A = torch.zeros((nk, n), device="cuda", dtype=torch.float32)
    B = torch.zeros(n, device="cuda", dtype=torch.float32)
    out = torch.zeros(nk, device="cuda", dtype=torch.float64)
    st = _dots("gato_dot_batched_noalloc", L.dtype_code(torch.float32), n, nk, A, n, B, 0, None,
               None, None, 1.0, out)
    assert st == 4  # GATO_ERR_UNSUPPORTED, before any launch
    assert b"partials" in L.lib().gato_last_error()
    torch.cuda.synchronize()
    assert torch.count_nonzero(out) == 0
