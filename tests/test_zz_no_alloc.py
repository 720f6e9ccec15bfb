"""fem_norms_noalloc (static device scratch, no allocation) against fem_norms
(the context's workspace) through the C ABI: two calls in a row, the second
on a smaller input, and the size limit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call_noalloc(fa, ctx, arrs, nv, np_, out):
    return fa._lib.fem_norms_noalloc(ctx.handle, arrs[0].ptr, arrs[1].ptr, nv, arrs[2].ptr, arrs[3].ptr, np_,
                                     out.ptr)


def _call_ctx(fa, ctx, arrs, nv, np_):
    h = (C.c_double * 5)()
    ctx.check(fa._lib.fem_norms(ctx.handle, arrs[0].ptr, arrs[1].ptr, nv, arrs[2].ptr, arrs[3].ptr, np_, h))
    return np.array(h[:])


def test_noalloc_matches_fem_norms_twice_then_smaller():
    import femtools_amd as fa
    ctx = fa.default_context()
    rng = np.random.default_rng(5)
    nv, np_ = 1_574_913, 1_572_864
    host = [rng.standard_normal(nv), rng.standard_normal(nv) * 2, rng.standard_normal(np_),
            rng.uniform(0.5, 2.0, np_)]
    arrs = [fa.device_array(ctx, h) for h in host]
    out = fa.device_array(ctx, np.full(5, np.nan))
    for (n1, n2) in ((nv, np_), (nv // 3 + 7, np_ // 5 + 1)):
        ctx.check(_call_noalloc(fa, ctx, arrs, n1, n2, out))
        ctx.sync()
        got = out.numpy()
        ref = _call_ctx(fa, ctx, arrs, n1, n2)
        assert np.all(np.isfinite(got))
        assert np.allclose(got, ref, rtol=1e-14, atol=0), (got, ref)
        exact = (np.linalg.norm(host[0][:n1]), np.linalg.norm(host[1][:n1]),
                 np.linalg.norm(host[2][:n2] / host[3][:n2]))
        assert np.allclose(got[:3], exact, rtol=1e-13)


def test_noalloc_rejects_inputs_beyond_its_scratch():
    import femtools_amd as fa
    ctx = fa.default_context()
    a = fa.device_array(ctx, np.zeros(4))
    out = fa.device_array(ctx, np.zeros(5))
    rc = fa._lib.fem_norms_noalloc(ctx.handle, a.ptr, a.ptr, 2048 * 4096 + 1, a.ptr, a.ptr, 1, out.ptr)
    assert rc == fa.FEM_EARG if hasattr(fa, "FEM_EARG") else rc == -1
    assert "static scratch limit" in ctx.error()
