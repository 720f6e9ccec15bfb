"""The oracle's 3-D Stokes residuals on uploaded Hex27 meshes (Mesh(backend,
coords, el2n, Hex27), mesh.jl:147-188): graded box elements of different
sizes, nodes and elements in shuffled order. Pinned by:

* invariance under renumbering: on an ungraded mesh the shuffled upload gives
  the generated mesh's residuals, permuted, to rounding;
* the patch test: a linear, trace-free velocity has constant stress, so with
  P = 0 and no gravity every interior momentum residual and every continuity
  residual vanish on the graded mesh;
* the element volumes sum to the box's volume.

mesh_graded_shuffled is shared with the device tests of the upload.
"""
import math

import numpy as np

import pyoracle as O

INF = math.inf


def mesh_graded_shuffled(n, box=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0), grade=0.3, seed=0):
    """A Hex27 box mesh whose element corners are graded per axis by
    s -> s + grade s (1 - s) (edge, face and centre nodes at the midpoints), with
    nodes and elements renumbered at random. Returns coords, el2n, the new node
    index of every generated node and the generated element of every new one."""
    coords, el2n = O.hex27_mesh(*n, box)
    for d in range(3):
        lo, hi = box[2 * d], box[2 * d + 1]
        m = 2 * n[d]
        s = np.arange(n[d] + 1) / n[d]
        corner = lo + (hi - lo) * (s + grade * s * (1 - s))
        line = np.empty(m + 1)
        line[0::2] = corner
        line[1::2] = 0.5 * (corner[:-1] + corner[1:])
        idx = np.rint((coords[:, d] - lo) / (hi - lo) * m).astype(int)
        coords[:, d] = line[idx]
    rng = np.random.default_rng(seed)
    pn = rng.permutation(len(coords))
    pe = rng.permutation(len(el2n))
    c2 = np.empty_like(coords)
    c2[pn] = coords
    e2 = pn[el2n][pe].astype(np.int32)
    return c2, e2, pn, pe


def _problem(coords, el2n, seed, phases=None, g=(0.0, 0.0, 0.0)):
    mat = O.material(eta=(1.3, 0.4), etab=(7.0, 3.0), rho0=(1.0, 2.0))
    pr = O.Problem3(coords, el2n, mat, g=g, phases=phases)
    return pr


def test_renumbered_upload_matches_the_generated_mesh():
    n = (3, 2, 2)
    c0, e0 = O.hex27_mesh(*n)
    c1, e1, pn, pe = mesh_graded_shuffled(n, grade=0.0, seed=3)
    assert np.allclose(c1[pn], c0, rtol=0, atol=1e-15)
    ne = len(e0)
    ph0 = (np.arange(ne) % 2 + 1).astype(np.int32)
    p0 = _problem(c0, e0, 1, phases=ph0, g=(0.1, -0.3, -1.0))
    p1 = _problem(c1, e1, 1, phases=ph0[pe], g=(0.1, -0.3, -1.0))
    rng = np.random.default_rng(7)
    for c in range(3):
        p0.v[c][:] = rng.normal(size=len(c0))
        p1.v[c][pn] = p0.v[c]
    p0.P[:] = rng.normal(size=4 * ne)
    p1.P[:] = p0.P.reshape(ne, 4)[pe].ravel()
    for p in (p0, p1):
        p.pressure_scaling(20.0, 0.4, False)
        p.pressure_residual(0.4, False)
        p.momentum_residual(0.4, None)
    assert np.allclose(p1.RP.reshape(ne, 4), p0.RP.reshape(ne, 4)[pe], rtol=1e-13, atol=1e-13)
    assert np.allclose(p1.gammaP.reshape(ne, 4), p0.gammaP.reshape(ne, 4)[pe], rtol=1e-13, atol=0)
    for c in range(3):
        scale = np.abs(p0.Rv[c]).max()
        assert np.abs(p1.Rv[c][pn] - p0.Rv[c]).max() < 1e-13 * scale


def test_patch_test_on_a_graded_mesh():
    box = (0.0, 2.0, -1.0, 0.5, 0.0, 1.5)
    c, e, _, _ = mesh_graded_shuffled((3, 3, 2), box=box, grade=0.4, seed=5)
    _, dOm = O.hex27_geometry(c, e)
    assert abs(dOm.sum() - 2.0 * 1.5 * 1.5) < 1e-13
    # element sizes really differ
    ext = np.ptp(c[e][:, :, 0], axis=1)
    assert ext.max() / ext.min() > 1.3
    A = np.array([[0.3, -0.7, 0.2], [0.5, 0.1, -0.4], [0.25, 0.6, -0.4]])
    assert abs(np.trace(A)) < 1e-15
    mat = O.material(eta=(1.7,))
    pr = O.Problem3(c, e, mat)
    for k in range(3):
        pr.v[k][:] = c @ A[k]
    pr.pressure_scaling(20.0, 1.0, False)
    pr.pressure_residual(1.0, False)
    pr.momentum_residual(1.0, None)
    assert np.abs(pr.RP).max() < 1e-13
    inner = np.setdiff1d(np.arange(len(c)), O.box_boundary_nodes3(c, box))
    for k in range(3):
        assert np.abs(pr.Rv[k][inner]).max() < 1e-13
        assert np.abs(pr.Rv[k]).max() > 1e-3   # the boundary carries the traction
