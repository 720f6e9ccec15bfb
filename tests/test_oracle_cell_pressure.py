"""The oracle's CellPressureStokesDR restatement (viscous Hex27 with the cell
pressure modes (1, ξ, η, ζ); stokes_types.jl:545-613, DR.jl:960-1057),
pinned by the reference's own tests: the sparse saddle-point system of
test/test_stokes_3d_reference.jl:25-90 (assembled here in numpy, dense,
independently of the oracle) and the first-iteration behaviour of
test/test_stokes_3d_boundary_values.jl."""
import numpy as np
import pytest

import pyoracle as O


def dense_system(pr, eta, rho, g, ph):
    """A, rhs of test_stokes_3d_reference.jl:25-90 on the oracle's mesh."""
    coords, el2n = pr.coords, pr.el2n
    nn, ne = len(coords), len(el2n)
    Nq, _, dNdxi, _ = O.hex27_tables()
    xi, et, ze, _ = O.hex27_rule()
    invJ, dOm = O.hex27_geometry(coords, el2n)
    nv, npp = 3 * nn, 4 * ne
    A = np.zeros((nv + npp, nv + npp))
    rhs = np.zeros(nv + npp)
    I3 = np.eye(3)
    for e in range(ne):
        nodes, p = el2n[e], ph[e] - 1
        vd = (3 * nodes[:, None] + np.arange(3)[None, :]).ravel()      # (a, i) -> row
        for q in range(27):
            grad = dNdxi[q] @ invJ[e, q].reshape(3, 3)                  # [a][x]
            dO = dOm[e, q]
            modes = np.array([1.0, xi[q], et[q], ze[q]])
            gg = grad @ grad.T
            # K[(a,i),(b,j)] = η ((i==j) ∇N_a·∇N_b + ∂_j N_a ∂_i N_b − 2/3 ∂_i N_a ∂_j N_b) dΩ
            K = (np.einsum("ab,ij->aibj", gg, I3) + np.einsum("aj,bi->aibj", grad, grad)
                 - (2 / 3) * np.einsum("ai,bj->aibj", grad, grad)) * dO * eta[p]
            A[np.ix_(vd, vd)] += K.reshape(81, 81)
            rhs[vd] += (np.outer(Nq[q], np.asarray(g)) * rho[p] * dO).ravel()
            cpl = -(grad[:, :, None] * modes[None, None, :]) * dO      # [a][i][m]
            pd = nv + 4 * e + np.arange(4)
            A[np.ix_(vd, pd)] += cpl.reshape(81, 4)
            A[np.ix_(pd, vd)] += cpl.reshape(81, 4).T
    return A, rhs


def block_problem(n, seed=0, eta=(1.0, 50.0), rho=(1.0, 2.0), g=(0.0, 0.0, -1.0), bc=True):
    coords, el2n = O.hex27_mesh(*n)
    cen = coords[el2n].mean(axis=1)
    ph = np.where(np.all(np.abs(cen - 0.5) < 0.26, axis=1), 2, 1).astype(np.int32)
    if (ph == 1).all():
        ph[len(ph) // 2] = 2
    bcs = None
    if bc:   # free slip: the wall-normal velocity is fixed on each wall
        bcs = []
        for c in range(3):
            d = np.where((np.abs(coords[:, c]) < 1e-12) | (np.abs(coords[:, c] - 1) < 1e-12))[0].astype(np.int32)
            bcs.append((d, np.zeros(len(d))))
        bcs = tuple(bcs)
    pr = O.Problem3(coords, el2n, O.material(eta=eta, rho0=rho), g=g, phases=ph, bc=bcs)
    return pr, ph


def test_assembled_residual_is_the_sparse_system():
    # test_stokes_3d_reference.jl:120-130: [momentum; continuity] = A x − rhs
    pr, ph = block_problem((2, 2, 1), g=(0.3, -0.2, -1.0))
    rng = np.random.default_rng(0)
    for c in range(3):
        pr.v[c][:] = rng.normal(size=pr.nnodes)
    pr.P[:] = rng.normal(size=4 * pr.nels)
    A, rhs = dense_system(pr, (1.0, 50.0), (1.0, 2.0), (0.3, -0.2, -1.0), ph)
    x = np.concatenate([np.stack(pr.v, axis=1).ravel(), pr.P])
    ref = A @ x - rhs
    got = np.concatenate([np.stack(pr.cp3_momentum_residual(), axis=1).ravel(), pr.cp3_pressure_residual()])
    assert np.allclose(got, ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max())
    assert np.allclose(A, A.T, atol=1e-13)


def test_preconditioner_is_the_system_diagonal_bound():
    # the Jacobi diagonal η(|∇N|² + (∂_c N)²/3) dΩ is the diagonal of the
    # viscous block (i == j, a == b terms of A); the mode mass is Σ mode² dΩ
    pr, ph = block_problem((2, 1, 1))
    A, _ = dense_system(pr, (1.0, 50.0), (1.0, 2.0), (0.0, 0.0, -1.0), ph)
    dg, M = pr.cp3_preconditioner()
    nv = 3 * pr.nnodes
    assert np.allclose(np.stack(dg, axis=1).ravel(), np.diag(A)[:nv], rtol=1e-13)
    # Σ_q 1 dΩ = the cell volume, Σ_q ξ² dΩ = a third of it (∫ξ² = 2/3 on [-1, 1])
    vol = 1.0 / pr.nels
    assert np.allclose(M.reshape(-1, 4), [vol, vol / 3, vol / 3, vol / 3], rtol=1e-13)


def test_solve_satisfies_the_sparse_system():
    # test_stokes_3d_reference.jl:111-119, 145-163: the converged velocity
    # solves the constrained system (pressure gauge: one mode fixed)
    pr, ph = block_problem((3, 3, 3))
    st = pr.cp3_solve(tolerance=1e-9, max_iterations=40000)
    assert st["converged"] and st["iterations"] > 0
    A, rhs = dense_system(pr, (1.0, 50.0), (1.0, 2.0), (0.0, 0.0, -1.0), ph)
    nv = 3 * pr.nnodes
    fixed = set()
    for c in range(3):
        fixed |= {3 * int(d) + c for d in pr.bc[c][0]}
    fixed.add(nv)   # gauge
    free = np.array(sorted(set(range(len(rhs))) - fixed))
    x = np.zeros(len(rhs))
    x[free] = np.linalg.solve(A[np.ix_(free, free)], rhs[free])
    v_it = np.stack(pr.v, axis=1).ravel()
    assert np.abs(v_it - x[:nv]).max() < 1e-6 * np.abs(x[:nv]).max()
    # the iterate's own residual is small on the free DoFs
    xi = np.concatenate([v_it, pr.P])
    assert np.linalg.norm((A @ xi - rhs)[free]) < 1e-7


def _expanding(n=(1, 1, 1)):
    coords, el2n = O.hex27_mesh(*n)
    gam = O.box_boundary_nodes3(coords)
    vals = [coords[gam, 0] if c == 0 else np.zeros(len(gam)) for c in range(3)]
    pr = O.Problem3(coords, el2n, O.material(eta=(1.0,), rho0=(1.0,)), g=(0.0, 0.0, 0.0),
                    bc=tuple((gam, v) for v in vals))
    return pr, gam, vals


def test_prescribed_velocities_enter_the_first_residual():
    # test_stokes_3d_boundary_values.jl:7-51
    pr, gam, vals = _expanding()
    st = pr.cp3_solve(check_interval=1, max_iterations=1, history_cap=4)
    assert not st["converged"] and st["reached_total_iter"]
    assert st["err_P"] > 1e-5
    assert len(st["history"]) == 1 and st["history"][0][2] == st["err_P"]
    for c in range(3):
        assert np.array_equal(pr.v[c][gam], vals[c])
    pc, _, _ = _expanding()
    for c in range(3):
        pc.v[c][gam] = vals[c]
    sc = pc.cp3_solve(check_interval=1, max_iterations=1)
    assert sc["err_P"] == pytest.approx(st["err_P"])
    assert np.allclose(pc.P, pr.P)
    for c in range(3):
        assert np.allclose(pc.v[c], pr.v[c])


def test_adjoint_is_the_transposed_system_and_gradients_match_finite_differences():
    # test_stokes_3d_reference.jl:165-215: the adjoint of the relaxation is
    # A_ffᵀ λ = objective; the material gradients −λ·∂R/∂m equal centred
    # differences of J = objective·v on the (independently assembled) system
    eta, rho, g = (1.0, 50.0), (1.0, 2.0), (0.0, 0.0, -1.0)
    pr, ph = block_problem((3, 3, 2))
    st = pr.cp3_solve(tolerance=1e-11, max_iterations=80000)
    assert st["converged"]
    v_fwd = [v.copy() for v in pr.v]
    P_fwd = pr.P.copy()
    A, rhs = dense_system(pr, eta, rho, g, ph)
    nv = 3 * pr.nnodes
    fixed = {nv}   # pressure gauge
    for c in range(3):
        fixed |= {3 * int(d) + c for d in pr.bc[c][0]}
    free = np.array(sorted(set(range(len(rhs))) - fixed))
    rng = np.random.default_rng(2)
    obj = [rng.normal(size=pr.nnodes) for _ in range(3)]
    lam_v, lam_P, sa = pr.cp3_solve_adjoint(obj, tolerance=1e-11, max_iterations=80000)
    assert sa["converged"] and sa["iterations"] > 0
    # the forward state is untouched
    for c in range(3):
        assert np.array_equal(pr.v[c], v_fwd[c])
    assert np.array_equal(pr.P, P_fwd)
    load = np.concatenate([np.stack(obj, axis=1).ravel(), np.zeros(4 * pr.nels)])
    Aff = A[np.ix_(free, free)]
    exact = np.zeros(len(rhs))
    exact[free] = np.linalg.solve(Aff.T, load[free])
    lam = np.stack(lam_v, axis=1).ravel()
    assert np.abs(lam - exact[:nv]).max() < 1e-7 * np.abs(exact[:nv]).max()
    for c in range(3):
        assert not lam_v[c][pr.bc[c][0]].any()   # homogeneous conditions
    dens, visc = pr.cp3_material_gradient(v_fwd, lam_v)
    h = 1e-5
    for p in range(2):
        e_p = tuple(1.0 if q == p else 0.0 for q in range(2))
        drhs = dense_system(pr, eta, e_p, g, ph)[1][free]
        plus = np.linalg.solve(Aff, rhs[free] + h * drhs)
        minus = np.linalg.solve(Aff, rhs[free] - h * drhs)
        fd_rho = load[free] @ (plus - minus) / (2 * h)
        dA = (dense_system(pr, e_p, rho, g, ph)[0] - dense_system(pr, (0.0, 0.0), rho, g, ph)[0])[np.ix_(free, free)]
        step = h * eta[p]
        plus = np.linalg.solve(Aff + step * dA, rhs[free])
        minus = np.linalg.solve(Aff - step * dA, rhs[free])
        fd_eta = load[free] @ (plus - minus) / (2 * step)
        assert dens[p] == pytest.approx(fd_rho, rel=1e-6)
        assert visc[p] == pytest.approx(fd_eta, rel=1e-5)
        assert abs(dens[p]) > 0 and abs(visc[p]) > 0
