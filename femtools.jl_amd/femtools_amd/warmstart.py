"""Learned warm start of the DR Stokes solve.

A small encoder-decoder CNN (the native network of femcnn.hip: 14 3x3
convolutions, 32/64/128/128 channels, BF16 weights and activations, FP32
accumulation on the matrix cores) reads per-cell features of the unit box on
a GRID x GRID image -- the scaled log viscosity, the inclusion indicator and
the cell-centre coordinates -- and predicts the velocity perturbation from
the background pure shear and the pressure, scaled by the strain rate and the
matrix viscosity. It is trained on the device by its own forward/backward
kernels, on solutions the DR solver itself produces for the viscous
inclusion generator (the unit-box pure-shear setting of the 64^2 viscous
inclusion with seeded random inclusions).

At application the features of any structured T7 box mesh are rasterised to
the GRID x GRID image (element centroids averaged per grid cell), and the
prediction is sampled bilinearly at the mesh's velocity nodes and at the
vertices of its P1-disc pressure elements. The reference takes any initial
guess and lets the DR iteration decide the answer
(stokes_2D_pure_shear_triangle.jl:159-163): a poor prediction costs
iterations, never accuracy; without weights the caller keeps its
zero/previous-step guess.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import time

import numpy as np

from . import _lib, Context, DeviceArray, device_array, default_context, VP

GRID = 128
IN_CH, OUT_CH = 16, 3
# (name, Cin, Cout, level) of the 14 convolutions in forward order (femcnn.hip)
LAYERS = [("e0a", 16, 32, 0), ("e0b", 32, 32, 0), ("e1a", 32, 64, 1), ("e1b", 64, 64, 1),
          ("e2a", 64, 128, 2), ("e2b", 128, 128, 2), ("e3a", 128, 128, 3), ("e3b", 128, 128, 3),
          ("d2a", 128, 128, 2), ("d2b", 128, 64, 2), ("d1a", 64, 64, 1), ("d1b", 64, 32, 1),
          ("d0a", 32, 32, 0), ("head", 32, 32, 0)]
LOG_ETA_SCALE = 4.0          # features carry log10(eta / eta_matrix) / 4
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "warmstart_cnn.npz")


def weight_shapes():
    return [((co, 3, 3, ci), (co,)) for _, ci, co, _ in LAYERS]


def init_params(seed=0):
    """He-normal weights (fan-in 9 Cin), zero biases; the padded input
    channels 4.. of the first layer and the padded outputs 3.. of the head are
    zero (they stay zero: their gradients vanish)."""
    rng = np.random.default_rng(seed)
    out = []
    for name, ci, co, _ in LAYERS:
        W = rng.standard_normal((co, 3, 3, ci)) * math.sqrt(2.0 / (9 * ci))
        if name == "e0a":
            W[..., 4:] = 0
        if name == "head":
            W *= 0.1
            W[3:] = 0
        out.append((W.astype(np.float32), np.zeros(co, np.float32)))
    return out


def bf16_bits(a):
    """float32 -> bfloat16 bits (round to nearest, ties to even)."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)


def bits_to_f32(h):
    return (np.asarray(h, np.uint16).astype(np.uint32) << 16).view(np.float32)


class Network:
    """The device network (fem_cnn): weights, forward, one training step."""

    def __init__(self, ctx: Context | None = None, max_batch=1, H=GRID, W=GRID):
        self.ctx = ctx or default_context()
        h = VP()
        self.ctx.check(_lib.fem_cnn_create(self.ctx.handle, int(max_batch), int(H), int(W), C.byref(h)))
        self.handle = h
        self.max_batch, self.H, self.W = int(max_batch), int(H), int(W)
        nw, nb = C.c_int64(), C.c_int64()
        _lib.fem_cnn_parameter_count(C.byref(nw), C.byref(nb))
        shapes = weight_shapes()
        assert nw.value == sum(int(np.prod(s[0])) for s in shapes) and nb.value == sum(s[1][0] for s in shapes)
        self.nw, self.nb = nw.value, nb.value
        self._x = device_array(self.ctx, shape=(self.max_batch, H, W, IN_CH), dtype=np.float32)
        self._t = device_array(self.ctx, shape=(self.max_batch, H, W, OUT_CH), dtype=np.float32)
        self._y = device_array(self.ctx, shape=(self.max_batch, H, W, OUT_CH), dtype=np.float32)

    def __del__(self):
        try:
            _lib.fem_cnn_destroy(self.handle)
        except Exception:
            pass

    def set_params(self, params):
        W = np.concatenate([np.asarray(w, np.float32).ravel() for w, _ in params])
        b = np.concatenate([np.asarray(bb, np.float32).ravel() for _, bb in params])
        assert W.size == self.nw and b.size == self.nb
        self.ctx.check(_lib.fem_cnn_set_weights(self.handle, W.ctypes.data, b.ctypes.data))

    def _unflatten(self, W, b):
        out, ow, ob = [], 0, 0
        for sw, sb in weight_shapes():
            nw, nb = int(np.prod(sw)), sb[0]
            out.append((W[ow:ow + nw].reshape(sw).copy(), b[ob:ob + nb].copy()))
            ow += nw
            ob += nb
        return out

    def get_params(self):
        W, b = np.empty(self.nw, np.float32), np.empty(self.nb, np.float32)
        self.ctx.check(_lib.fem_cnn_get_weights(self.handle, W.ctypes.data, b.ctypes.data))
        return self._unflatten(W, b)

    def get_grads(self):
        W, b = np.empty(self.nw, np.float32), np.empty(self.nb, np.float32)
        self.ctx.check(_lib.fem_cnn_get_grads(self.handle, W.ctypes.data, b.ctypes.data))
        return self._unflatten(W, b)

    def forward(self, x):
        """x [B][H][W][16] (host) -> y [B][H][W][3] (host)."""
        x = np.ascontiguousarray(x, np.float32)
        B = x.shape[0]
        self._x.copyto(np.concatenate([x, np.zeros((self.max_batch - B,) + x.shape[1:], np.float32)])
                       if B < self.max_batch else x)
        self.ctx.check(_lib.fem_cnn_forward(self.handle, self._x.ptr, B, self._y.ptr))
        return self._y.numpy()[:B]

    def forward_device(self, x: DeviceArray, B: int, y: DeviceArray):
        self.ctx.check(_lib.fem_cnn_forward(self.handle, x.ptr, int(B), y.ptr))

    def train_step(self, x: DeviceArray, t: DeviceArray, B: int, lr: float, step: int) -> float:
        loss = C.c_double()
        self.ctx.check(_lib.fem_cnn_train_step(self.handle, x.ptr, t.ptr, int(B), float(lr), int(step),
                                               C.byref(loss)))
        return loss.value


def gemm_bf16(ctx: Context, A_bits, Bt_bits):
    """C = A . Bt^T on the matrix cores for bfloat16 bit matrices (host)."""
    M, K = A_bits.shape
    N = Bt_bits.shape[0]
    a = device_array(ctx, A_bits, dtype=np.uint16)
    b = device_array(ctx, Bt_bits, dtype=np.uint16)
    c = device_array(ctx, shape=(M, N), dtype=np.float32)
    ctx.check(_lib.fem_cnn_gemm_bf16(ctx.handle, M, N, K, a.ptr, b.ptr, c.ptr))
    return c.numpy()


# ---------------------------------------------------------------------------
# features, targets and resampling on structured T7 box meshes
# ---------------------------------------------------------------------------
def _cells(mesh):
    nx = ny = int(round(math.sqrt(mesh.nels / 2)))
    if 2 * nx * ny != mesh.nels:
        raise ValueError("the warm start needs a square structured T7 box mesh")
    return nx, ny


def element_grid_index(mesh, grid=GRID):
    """Grid cell (row-major y, x) of every element's centroid on the unit box."""
    c = mesh.coords[mesh.el2n[:, :3]].mean(axis=1)
    ix = np.clip((c[:, 0] * grid).astype(np.int64), 0, grid - 1)
    iy = np.clip((c[:, 1] * grid).astype(np.int64), 0, grid - 1)
    return iy * grid + ix


def features(mesh, phases, eta_phase, grid=GRID, gidx=None):
    """[grid][grid][16] features of a mesh with 1-based element phases and a
    per-phase viscosity: mean log10(eta / eta_matrix) / 4 and the inclusion
    fraction per grid cell, and the cell-centre coordinates (x - 1/2, y - 1/2)."""
    gidx = element_grid_index(mesh, grid) if gidx is None else gidx
    le = np.log10(np.asarray(eta_phase, float)[np.asarray(phases) - 1] / eta_phase[0]) / LOG_ETA_SCALE
    cnt = np.bincount(gidx, minlength=grid * grid)
    cnt = np.maximum(cnt, 1)
    f = np.zeros((grid, grid, IN_CH), np.float32)
    f[..., 0] = (np.bincount(gidx, le, grid * grid) / cnt).reshape(grid, grid)
    f[..., 1] = (np.bincount(gidx, (np.asarray(phases) != 1).astype(float), grid * grid) / cnt).reshape(grid, grid)
    xc = (np.arange(grid) + 0.5) / grid - 0.5
    f[..., 2] = xc[None, :]
    f[..., 3] = xc[:, None]
    return f


def targets(mesh, vx, vy, P, v_bg, edot=1.0, eta_scale=1.0):
    """[n][n][3] training targets of a solution on an n x n T7 box mesh: the
    velocity perturbation from the background at the cell centres (/ edot)
    and the mean of the two triangles' pressures (/ 2 eta edot)."""
    nx, ny = _cells(mesh)
    sx = 2 * nx + 1
    iy, ix = np.meshgrid(2 * np.arange(ny) + 1, 2 * np.arange(nx) + 1, indexing="ij")
    ctr = (iy * sx + ix).ravel()
    t = np.zeros((ny, nx, 3), np.float32)
    t[..., 0] = ((vx[ctr] - v_bg[0][ctr]) / edot).reshape(ny, nx)
    t[..., 1] = ((vy[ctr] - v_bg[1][ctr]) / edot).reshape(ny, nx)
    pe = P.reshape(-1, 3).mean(axis=1)                  # element centroid values
    t[..., 2] = (0.5 * (pe[0::2] + pe[1::2]) / (2 * eta_scale * edot)).reshape(ny, nx)
    return t


def _bilinear(field, x, y):
    """Sample a cell-centred [g][g] field on the unit box at points (x, y)
    (clamped at the border)."""
    g = field.shape[0]
    fx = np.clip(x * g - 0.5, 0.0, g - 1.0)
    fy = np.clip(y * g - 0.5, 0.0, g - 1.0)
    i0 = np.minimum(fx.astype(np.int64), g - 2)
    j0 = np.minimum(fy.astype(np.int64), g - 2)
    tx, ty = fx - i0, fy - j0
    f00, f10 = field[j0, i0], field[j0, i0 + 1]
    f01, f11 = field[j0 + 1, i0], field[j0 + 1, i0 + 1]
    return (1 - ty) * ((1 - tx) * f00 + tx * f10) + ty * ((1 - tx) * f01 + tx * f11)


def guess_from_prediction(mesh, pred, v_bg, edot=1.0, eta_scale=1.0):
    """(vx, vy, P) on the mesh from a [g][g][3] prediction: bilinear samples at
    the velocity nodes (plus the background) and at the pressure elements'
    vertices (P1-disc DoF 3e + a is the value at vertex a of element e)."""
    x, y = mesh.coords[:, 0], mesh.coords[:, 1]
    vx = v_bg[0] + edot * _bilinear(pred[..., 0].astype(float), x, y)
    vy = v_bg[1] + edot * _bilinear(pred[..., 1].astype(float), x, y)
    pv = mesh.coords[mesh.el2n[:, :3].ravel()]
    P = 2 * eta_scale * edot * _bilinear(pred[..., 2].astype(float), pv[:, 0], pv[:, 1])
    return vx, vy, P


def hydrostatic_pressure(mesh, phases, rho, g):
    """The column-integrated weight p(x, y) = |g| int_y^1 rho dy' at the pressure
    vertices (per-element density rasterised to the GRID image), with zero
    mean: the pressure a gravity load adds where the flow is near rest."""
    gy = abs(g[1])
    if gy == 0:
        return np.zeros(mesh.nnodesP)
    gidx = element_grid_index(mesh)
    r = np.asarray(rho, float)[np.asarray(phases) - 1]
    cnt = np.maximum(np.bincount(gidx, minlength=GRID * GRID), 1)
    rg = (np.bincount(gidx, r, GRID * GRID) / cnt).reshape(GRID, GRID)
    # integral from the top of the box down to each cell centre (midpoint rule)
    col = np.cumsum(rg[::-1], axis=0)[::-1] - 0.5 * rg
    pc = gy * col / GRID
    pc -= pc.mean()
    pv = mesh.coords[mesh.el2n[:, :3].ravel()]
    return _bilinear(pc, pv[:, 0], pv[:, 1])


# ---------------------------------------------------------------------------
# the training generator: viscous inclusions under pure shear
# ---------------------------------------------------------------------------
def random_inclusions(seed, max_inclusions=10):
    """1..max_inclusions circles (centre U[0.1, 0.9]^2, r U[0.02, 0.1], log10
    eta U[-2, 3]) in an eta = 1 matrix; later circles win."""
    rng = np.random.default_rng(seed)
    k = int(rng.integers(1, max_inclusions + 1))
    circles = list(zip(rng.uniform(0.1, 0.9, k), rng.uniform(0.1, 0.9, k), rng.uniform(0.02, 0.1, k)))
    eta = (1.0,) + tuple(10.0 ** rng.uniform(-2, 3, k))
    return circles, eta


def _viscous_state(fa, mesh, circles, eta):
    from . import problems as PB
    nph = len(eta)
    ph = PB.inclusion_phases(mesh.coords, mesh.el2n, circles)
    inf = math.inf
    dr = fa.StokesDR(mesh, fa.StokesMaterial(eta=eta, etab=(inf,) * nph, G=(inf,) * nph, rho0=(1.0,) * nph,
                                             K=(inf,) * nph, g=(0.0, 0.0)))
    dr.set_phases("v", ph)
    dr.set_phases("P", ph)
    return dr, ph


_MESHES = {}   # n -> [(mesh, boundary conditions, background field, grid index)], one per batch slot


def generate_dataset(seeds, n=GRID, tolerance=1e-6, max_iterations=100_000, batch=64, log=None):
    """DR solutions of the viscous inclusion generator on n x n T7 meshes, one
    per seed (the FP32-bulk solver, `batch` problems per batched call):
    returns features [S][n][n][16], targets [S][n][n][3] and the per-solve
    stats."""
    import femtools_amd as fa
    from . import problems as PB
    X = np.zeros((len(seeds), n, n, IN_CH), np.float32)
    T = np.zeros((len(seeds), n, n, OUT_CH), np.float32)
    stats = []
    meshes = _MESHES.setdefault(n, [])
    for i0 in range(0, len(seeds), batch):
        probs = []
        for j, s in enumerate(seeds[i0:i0 + batch]):
            if len(meshes) <= j:
                mesh = fa.t7_box_mesh(n, n, ctx=fa.default_context())
                (bx, vx), (by, vy) = PB.pure_shear_bc(mesh.coords)
                meshes.append((mesh, (fa.DirichletBoundaryCondition(bx, vx), fa.DirichletBoundaryCondition(by, vy)),
                               PB.pure_shear_field(mesh.coords), element_grid_index(mesh, n)))
            mesh, bcs, vbg, gidx = meshes[j]
            circles, eta = random_inclusions(s)
            dr, ph = _viscous_state(fa, mesh, circles, eta)
            dr.v.x.copyto(vbg[0])
            dr.v.y.copyto(vbg[1])
            kw = dict(dt=1.0, tolerance=tolerance, check_interval=50, max_iterations=max_iterations,
                      iterMax=max_iterations, max_ph_iterations=20_000, throw_on_failure=False, precision="fp32")
            probs.append((dr, mesh, bcs, kw, ph, eta, vbg, gidx))
        res = fa.solve_batch([(dr, mesh, bcs, kw) for dr, mesh, bcs, kw, *_ in probs], raise_errors=False)
        for j, ((dr, mesh, bcs, kw, ph, eta, vbg, gidx), st) in enumerate(zip(probs, res)):
            i = i0 + j
            X[i] = features(mesh, ph, eta, n, gidx)
            T[i] = targets(mesh, dr.v.x.numpy(), dr.v.y.numpy(), dr.P.numpy(), vbg)
            stats.append({"iter": st["iter"], "converged": bool(st["converged"]), "status": st["status"],
                          "err_abs": st["err_abs"]})
        if log:
            log(f"dataset {min(i0 + batch, len(seeds))}/{len(seeds)}")
    return X, T, stats


def _vep_series_state(fa, PB, mesh, n, seed, eta_reg=None):
    """Step-0 state of the visco-elasto-plastic series generator
    (problems.vep_inclusions) on the given mesh: the state, the Drucker-Prager
    model, the workload, its phases and the boundary conditions."""
    w = PB.vep_inclusions(n, seed=seed, **({} if eta_reg is None else {"eta_reg": eta_reg}))
    ph = PB.inclusion_phases(mesh.coords, mesh.el2n, w.circles)
    dr = fa.StokesDR(mesh, fa.StokesMaterial(eta=w.eta, etab=w.etab, G=w.G, rho0=w.rho0, K=w.K, g=w.g),
                     CFL_v=w.CFL_v)
    dr.set_phases("v", ph)
    dr.set_phases("P", ph)
    dp = fa.DruckerPrager(w.plastic["phi"], w.plastic["psi"], w.plastic["C"], w.plastic["eta_reg"],
                          w.plastic["Kb"])
    return dr, dp, w, ph


def vep_guess(net, mesh, w, ph, v_bg):
    """The network's initial (vx, vy, P) for a step of the VEP series: the
    prediction for the step's visco-elastic viscosity field 1/(1/eta + 1/(G dt)),
    with the column's hydrostatic pressure."""
    eta_ve = [1.0 / (1.0 / e + 1.0 / (g * w.dt)) for e, g in zip(w.eta, w.G)]
    return predict_on_mesh(net, mesh, ph, eta_ve, v_bg, edot=1.0, eta_scale=eta_ve[0], rho=w.rho0, g=w.g)


def evaluate_gain(net, n, seeds, tolerance=1e-8, precision="fp32-tangent", max_iterations=600_000, log=None):
    """What the learned guess gains on n x n problems of the VEP series
    generator (problems.vep_inclusions with the given seeds; the network was
    trained on viscous problems only). Step 0 is solved to `tolerance` from
    (i) zero velocity (boundary values only), (ii) the background pure shear
    the reference's VEP examples start from and (iii) the network's
    prediction; step 1 (after the stress-history hand-off of step 0's
    solution) from (iv) the previous step's solution, as the reference's time
    loops do, and (v) the prediction. Inference and resampling count in the
    learned rows' time."""
    import femtools_amd as fa
    from . import problems as PB
    mesh = fa.t7_box_mesh(n, n, ctx=net.ctx)
    (bx, vx), (by, vy) = PB.pure_shear_bc(mesh.coords)
    bcs = (fa.DirichletBoundaryCondition(bx, vx), fa.DirichletBoundaryCondition(by, vy))
    vbg = PB.pure_shear_field(mesh.coords)
    zero = np.zeros(mesh.nnodes)
    out = []
    for s in seeds:
        dr, dp, w, ph = _vep_series_state(fa, PB, mesh, n, s)
        kw = dict(w.solve_kwargs(), dt=w.dt, plastic=dp, tolerance=tolerance, max_iterations=max_iterations,
                  iterMax=max_iterations, throw_on_failure=False, precision=precision)
        row = {"seed": s, "n": n}

        def run(name, guess, P_old, extra_s=0.0):
            dr.v.x.copyto(guess[0])
            dr.v.y.copyto(guess[1])
            dr.P.copyto(guess[2])
            dr.P0.copyto(P_old)
            net.ctx.sync()
            t0 = time.perf_counter()
            try:
                st = fa.solve(dr, mesh, *bcs, **kw)
                net.ctx.sync()
                row[name] = {"iterations": st["iter"], "seconds": round(time.perf_counter() - t0 + extra_s, 4),
                             "converged": bool(st["converged"])}
            except fa.FemError as e:   # a diverging solve is recorded, not raised
                row[name] = {"iterations": None, "seconds": round(time.perf_counter() - t0 + extra_s, 4),
                             "converged": False, "error": str(e)}

        zP = np.zeros(mesh.nnodesP)
        run("zero", (zero, zero, zP), zP)
        t0 = time.perf_counter()
        g0 = vep_guess(net, mesh, w, ph, vbg)
        tg = time.perf_counter() - t0
        run("learned", g0, zP, tg)
        run("pure_shear", (vbg[0], vbg[1], zP), zP)
        # step 1 from step 0's (pure-shear start) solution and stress
        s1 = (dr.v.x.numpy(), dr.v.y.numpy(), dr.P.numpy())
        fa.update_stokes_current_stress(dr, mesh, w.dt, plastic=dp)
        for c in ("xx", "yy", "xy"):
            fa.copyto(getattr(dr.tau_old, c), getattr(dr.tau, c))
        run("step1_previous_step", s1, s1[2])
        t0 = time.perf_counter()
        g1 = vep_guess(net, mesh, w, ph, vbg)
        tg = time.perf_counter() - t0
        run("step1_learned", g1, s1[2], tg)
        out.append(row)
        if log:
            log(f"gain n={n} seed {s}: " + ", ".join(f"{k} {row[k]['iterations']} it {row[k]['seconds']:.2f}s"
                                                    for k in row if isinstance(row[k], dict)))
    return out


def summarize_gain(rows):
    """Total iterations and wall time of each start over the rows, and the
    learned guess's ratios to the zero, pure-shear and previous-step starts
    over the rows where the compared starts all converged."""
    keys = ("zero", "pure_shear", "learned", "step1_previous_step", "step1_learned")
    out = {"problems": len(rows)}
    for k in keys:
        out[k] = {"iterations": sum(r[k]["iterations"] or 0 for r in rows),
                  "seconds": round(sum(r[k]["seconds"] for r in rows), 3),
                  "converged": sum(bool(r[k]["converged"]) for r in rows)}
    for a, b in (("learned", "zero"), ("learned", "pure_shear"), ("step1_learned", "step1_previous_step")):
        ok = [r for r in rows if r[a]["converged"] and r[b]["converged"]]
        if ok:
            out[f"{a}_over_{b}"] = {
                "iterations": round(sum(r[a]["iterations"] for r in ok) / sum(r[b]["iterations"] for r in ok), 3),
                "seconds": round(sum(r[a]["seconds"] for r in ok) / sum(r[b]["seconds"] for r in ok), 3)}
    return out


def train(net: Network, X, T, epochs=30, batch=16, lr=1e-3, seed=0, log=None, val=None):
    """Adam with a cosine learning-rate decay over epochs; X, T host arrays.
    Returns the per-epoch mean training loss (and validation loss if val)."""
    ctx = net.ctx
    S = len(X)
    rng = np.random.default_rng(seed)
    xd = device_array(ctx, shape=(batch,) + X.shape[1:], dtype=np.float32)
    td = device_array(ctx, shape=(batch,) + T.shape[1:], dtype=np.float32)
    hist = []
    step = 0
    nb = S // batch
    total = epochs * nb
    for ep in range(epochs):
        order = rng.permutation(S)
        tot = 0.0
        for k in range(nb):
            idx = order[k * batch:(k + 1) * batch]
            xd.copyto(X[idx])
            td.copyto(T[idx])
            step += 1
            lr_t = lr * 0.5 * (1 + math.cos(math.pi * (step - 1) / total))
            tot += net.train_step(xd, td, batch, lr_t, step)
        row = {"epoch": ep, "loss": tot / max(nb, 1)}
        if val is not None:
            Xv, Tv = val
            row["val_loss"] = evaluate(net, Xv, Tv)
        hist.append(row)
        if log:
            log(f"epoch {ep}: " + " ".join(f"{k} {v:.4e}" for k, v in row.items() if k != "epoch"))
    return hist


def evaluate(net: Network, X, T):
    """Mean-squared error of the prediction over a host set."""
    err, n = 0.0, 0
    for i0 in range(0, len(X), net.max_batch):
        y = net.forward(X[i0:i0 + net.max_batch])
        err += float(((y - T[i0:i0 + net.max_batch]) ** 2).sum())
        n += y.size
    return err / n


# ---------------------------------------------------------------------------
# persistence and application
# ---------------------------------------------------------------------------
def save(params, path=WEIGHTS, meta=None):
    """BF16 weights (raw bits) and FP32 biases."""
    os.makedirs(os.path.dirname(path), exist_ok=True)
    out = {}
    for i, (W, b) in enumerate(params):
        out[f"W{i}"] = bf16_bits(W)
        out[f"b{i}"] = np.asarray(b, np.float32)
    if meta:
        out["meta"] = np.frombuffer(repr(meta).encode(), np.uint8)
    np.savez_compressed(path, **out)


def load(path=WEIGHTS):
    z = np.load(path)
    return [(bits_to_f32(z[f"W{i}"]).reshape(s[0]), z[f"b{i}"]) for i, s in enumerate(weight_shapes())]


_default_model = None


def load_default(ctx: Context | None = None) -> Network:
    """The trained network (inference, batch 1); FileNotFoundError without weights."""
    global _default_model
    if _default_model is None:
        if not os.path.exists(WEIGHTS):
            raise FileNotFoundError(WEIGHTS)
        net = Network(ctx, max_batch=1)
        net.set_params(load(WEIGHTS))
        _default_model = net
    return _default_model


def predict_on_mesh(net: Network, mesh, phases, eta_phase, v_bg, edot=1.0, eta_scale=None, rho=None, g=None):
    """Initial (vx, vy, P) for a structured T7 box mesh with 1-based element
    phases and per-phase viscosities; with a density per phase and gravity the
    hydrostatic pressure of the column weight is added."""
    eta_scale = eta_phase[0] if eta_scale is None else eta_scale
    f = features(mesh, phases, eta_phase)
    pred = net.forward(f[None])[0]
    vx, vy, P = guess_from_prediction(mesh, pred, v_bg, edot, eta_scale)
    if rho is not None and g is not None:
        P = P + hydrostatic_pressure(mesh, phases, rho, g)
    return vx, vy, P
