"""fem_norms (the residual norms of a DR check, DR.jl:556-557, 700) on the
device against numpy's two-norm of the same arrays."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu



@pytest.mark.parametrize("nv,np_", [(1, 3), (4097, 4095), (1_574_913, 1_572_864), (9_000_001, 10)])
def test_norms_match_numpy(nv, np_):
    import femtools_amd as fa
    ctx = fa.default_context()
    rng = np.random.default_rng(nv)
    hx, hy = rng.standard_normal(nv), rng.standard_normal(nv) * 3
    hP, hM = rng.standard_normal(np_), rng.uniform(0.5, 2.0, np_)
    arrs = [fa.device_array(ctx, h) for h in (hx, hy, hP, hM)]
    out = fa.norms(ctx, *arrs)
    ref = (np.linalg.norm(hx), np.linalg.norm(hy), np.linalg.norm(hP / hM))
    for a, b in zip(out[:3], ref):
        assert abs(a - b) <= 1e-13 * b, (a, b)
    assert out[3] == pytest.approx(max(ref[0], ref[1]) / (2 * np.sqrt(nv)), rel=1e-13)
    assert out[4] == pytest.approx(ref[2] / np.sqrt(np_), rel=1e-13)
    assert fa.norms(ctx, *arrs) == out      # same partition and order: bitwise reproducible
