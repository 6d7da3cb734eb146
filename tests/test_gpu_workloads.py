"""The five benchmark workloads at their stated sizes, against the oracle.

C1: 128^2 FP64 Zalesak disk, rigid rotation, one full revolution (628 steps)
    of SL + 9-point curvature, GPU and oracle compared every quarter turn.
C2: covered at its full size in test_gpu_sl.py (256^3 sample + 64^3 steps).
C3: one full precision-form PLS step at 384^3 with 32 particles per cell
    (7.97M particles, seed 3) in LevelSetAdvectPLS's order
    (LevelSetMethod/ParticleLS.c:299-329), FP64 against the oracle and FP32
    against the oracle on FP32-rounded inputs, on node and particle samples,
    the correction over the whole grid.
C4: the 512-cell batch (edges log-uniform in [32, 128], seed 4), autotuned,
    one graph step; 16 cells (the smallest, the largest, 14 drawn with seed 4)
    against the per-cell oracle, the border copy and finiteness on every cell.
C5: 128^3 FP32 SL adjoint over 50 steps with a checkpoint every 10, against
    the FP64 adjoint chain (one of its steps pinned to the oracle's adjoint at
    128^3) and against FP64 central differences on 64 components drawn with
    seed 7 from the gradient's support, delta = 1e-4 |u|_inf.

Tolerances: FP64 phi / kappa 1e-12 of the max (1e-11 after the 628-step
revolution, where rounding differences accumulate); FP32 against the oracle
on float32-rounded inputs 2e-6 (phi, SL) and 5e-5 (kappa); particle positions
1e-12 (FP64) / 1e-6 (FP32) of the window edge; the FP64 correction bit-exact;
C5 gradients within 1e-3 of the largest finite difference (the
differentiation-form target), the FP64 chain within 1e-4 (central
differences step over an occasional kink of the trilinear velocity
interpolation); kappa after the revolution within 1e-9 of its max (second
differences over h^2 magnify phi's accumulated rounding).
"""
import numpy as np
import pytest
import torch

import oracle as O
import dcell_amd as D
from dcell_amd import adjoint as A
from dcell_amd import configs
from dcell_amd import pls as PL
from dcell_amd.batch import Batch
from helpers import owin, rel_max, rounded, to_dev, to_np

pytestmark = pytest.mark.gpu

IM = {"linear": O.LINEAR, "cubic": O.CUBIC}
BM = {"euler": O.EULER, "rk2": O.RK2}


@pytest.fixture
def threads():
    import os
    O.set_threads(min(16, len(os.sched_getaffinity(0))))
    yield
    O.set_threads(1)


# ----------------------------------------------------------------------------- C1
@pytest.mark.parametrize("interp,bt", [("cubic", "rk2"), ("linear", "euler")])
def test_c1_full_revolution(dev, interp, bt):
    n, steps = 128, 628
    h = 1.0 / (n - 1)
    win = D.GridWindow((0, 0, 0), (n, n, 1), (h, h, h))
    ow = owin(win)
    phi = configs.zalesak_sdf(n)
    vel = configs.rotation_velocity(n)
    pd, vd = to_dev(phi), to_dev(vel)
    kd = torch.empty_like(pd)
    ref = phi
    errs = []
    for s in range(1, steps + 1):
        pd = D.advect_sl(pd, vd, win, 1.0 / steps, interp, bt)
        D.curvature(pd, win, out=kd)
        ref = O.sl_advect(ref, vel, ow, 1.0 / steps, IM[interp], BM[bt])
        if s % 157 == 0:
            got = to_np(pd)
            errs.append(rel_max(got, ref))
            kref = O.curvature(ref, ow)
            sel = np.abs(ref) < 6
            ek = rel_max(to_np(kd)[sel], kref[sel], np.max(np.abs(kref[sel])))
            print(f"C1 {interp}/{bt} step {s}: phi rel err {errs[-1]:.2e}, kappa rel err {ek:.2e}")
            assert errs[-1] <= 1e-11 and ek <= 1e-9, (s, errs[-1], ek)


# ----------------------------------------------------------------------------- C3
def _sample_nodes(n, rng, phi, count=20000):
    """count interior nodes, half of them within 6 cells of the interface."""
    near = np.argwhere((np.abs(phi) < 6))
    near = near[(near >= 2).all(1) & (near <= n - 3).all(1)]
    a = near[rng.choice(len(near), count // 2, replace=False)][:, ::-1]
    b = rng.integers(2, n - 2, size=(count - count // 2, 3))
    return np.ascontiguousarray(np.concatenate([a, b]).astype(np.int32))


@pytest.fixture(scope="module")
def c3_inputs():
    n = 384
    phi = configs.sphere_sdf(n)
    return n, phi, configs.deformation_velocity(n), configs.seed_particles(phi, 32, 3.0, 3)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_c3_pls_step_384(dev, threads, c3_inputs, dtype):
    n, phi, vel, P = c3_inputs
    win = D.GridWindow.unit_box(n)
    ow = owin(win)
    phi, vel, P = rounded(phi, dtype), rounded(vel, dtype), rounded(P, dtype)
    assert len(P) > 7.9e6
    dt = configs.deformation_dt(n, 0)
    rng = np.random.default_rng(33)
    print(f"C3 {dtype}: {len(P)} particles", flush=True)
    # one PLS step on the GPU: particles, SL, correction, curvature, radii
    Pd, vd, pd = to_dev(P, dtype), to_dev(vel, dtype), to_dev(phi, dtype)
    PL.advect_particles(Pd, vd, win, win.d, dt)
    P_adv = to_np(Pd)
    sl = D.advect_sl(pd, vd, win, dt, "cubic", "rk2")
    sl_np = to_np(sl)
    ws = PL.CorrectionWorkspace(win.nodes, 8 * len(P), dtype)
    PL.correct(Pd, sl, win, ws)
    corr = to_np(sl)
    kap = to_np(D.curvature(sl, win))
    PL.adjust_radii(Pd, sl, win)
    P_fin = to_np(Pd)
    ws.close()
    del Pd, vd, pd, sl
    torch.cuda.empty_cache()
    f64 = dtype == torch.float64
    # particles: RK2 positions on a sample
    ps = rng.choice(len(P), 20000, replace=False)
    ref = O.pls_advect(P[ps], vel, ow, win.d, dt)
    assert rel_max(P_adv[ps, :3], ref[:, :3], float(n)) <= (1e-12 if f64 else 1e-6)
    # SL on a node sample
    band = _sample_nodes(n, rng, phi)
    ref = O.sl_advect(phi, vel, ow, dt, O.CUBIC, O.RK2, band=band)
    got = sl_np[band[:, 2], band[:, 1], band[:, 0]]
    assert rel_max(got, ref, np.max(np.abs(phi))) <= (1e-12 if f64 else 2e-6)
    # correction over the whole grid, from the GPU's own SL output and particles
    ref = O.pls_correct(P_adv, sl_np, ow, None, 0, 3)
    if f64:
        assert np.array_equal(corr, ref)
    else:
        # an FP32 escape test can decide differently from the FP64 one within
        # rounding of its thresholds: such particles touch at most 8 nodes each
        d = configs.trilinear_numpy(sl_np, P_adv[:, :3])
        edge = (np.abs(d) < 1e-4) | (np.abs(np.abs(d) - np.abs(P_adv[:, 3])) < 1e-4)
        bad = np.abs(corr - ref) > 1e-6 * np.max(np.abs(sl_np))
        assert bad.sum() <= 8 * edge.sum(), (bad.sum(), edge.sum())
    # curvature of the corrected phi on the node sample
    kref = O.curvature(corr, ow, band=band)
    got = kap[band[:, 2], band[:, 1], band[:, 0]]
    sel = np.abs(corr[band[:, 2], band[:, 1], band[:, 0]]) < 6
    assert rel_max(got[sel], kref[sel], np.max(np.abs(kref[sel]))) <= (1e-12 if f64 else 5e-5)
    # radii from the corrected phi on the particle sample
    ref = O.pls_adjust_radii(P_adv[ps], corr, ow)
    assert np.array_equal(P_fin[ps, :3], P_adv[ps, :3])
    assert rel_max(P_fin[ps, 3], ref[:, 3], 0.5) <= (1e-12 if f64 else 1e-6)


# ----------------------------------------------------------------------------- C4
def test_c4_batch_512_cells(dev, threads):
    edges = configs.batch_edges(512, 4)
    assert edges.min() >= 32 and edges.max() <= 128
    g = configs.philox(4)
    wins = [D.GridWindow.unit_box(int(e)) for e in edges]
    b = Batch(wins, torch.float32, "cubic", "rk2")
    phi0, phi1, kap, vel = b.new_field(), b.new_field(), b.new_field(), b.new_field(3)
    dts = []
    for c, e in enumerate(edges):
        phi, v, dt = configs.cell_fields(int(e), g, "cuda:0", torch.float32)
        b.cell(phi0, c).copy_(phi)
        b.cell(vel, c, 3).copy_(v)
        dts.append(dt)
    b.set_dt(dts)
    b.autotune(phi0, phi1, kap, vel, reps=1)
    print(f"C4: {b.nodes} nodes, configuration {b.config()}", flush=True)
    phi1.fill_(float("nan"))
    kap.fill_(float("nan"))
    gr = b.graph(phi0, phi1, kap, vel)
    Batch.launch(gr)
    torch.cuda.synchronize()
    # every cell: border layers are phi0 bit for bit, everything finite
    for c in range(len(edges)):
        p0, p1, k = b.cell(phi0, c), b.cell(phi1, c), b.cell(kap, c)
        assert bool(torch.isfinite(p1).all()) and bool(torch.isfinite(k).all()), c
        for sl in (np.s_[:2], np.s_[-2:], np.s_[:, :2], np.s_[:, -2:], np.s_[:, :, :2], np.s_[:, :, -2:]):
            assert torch.equal(p1[sl], p0[sl]), c
    rng = np.random.default_rng(4)
    pick = [int(np.argmin(edges)), int(np.argmax(edges))]
    pick += [int(c) for c in rng.choice([c for c in range(len(edges)) if c not in pick], 14, replace=False)]
    for c in pick:
        w = owin(wins[c])
        p0 = to_np(b.cell(phi0, c))
        v = to_np(b.cell(vel, c, 3))
        ref = O.sl_advect(p0, v, w, dts[c], O.CUBIC, O.RK2)
        got = to_np(b.cell(phi1, c))
        assert rel_max(got, ref) <= 2e-6, (c, int(edges[c]))
        kr = O.curvature(rounded(got, torch.float32), w)
        sel = np.abs(ref) < 6
        assert rel_max(to_np(b.cell(kap, c))[sel], kr[sel], np.max(np.abs(kr[sel]))) <= 5e-5, c
    b.close()


# ----------------------------------------------------------------------------- C5
def test_c5_adjoint_128(dev, threads):
    n, steps, every, eps = 128, 50, 10, 1.5
    win = D.GridWindow.unit_box(n)
    h = 1.0 / (n - 1)
    phi0 = configs.sphere_sdf(n, (0.5, 0.5, 0.5), 0.2)
    vel = configs.random_fourier_velocity(n, 6, 3, 0.5, seed=5)
    vt = configs.random_fourier_velocity(n, 6, 3, 0.5, seed=6)
    dt = 0.8 * h / max(np.abs(vel).max(), np.abs(vt).max())
    p0 = to_dev(phi0)
    tgt = p0
    vtd = to_dev(vt)
    for _ in range(steps):
        tgt = D.advect_sl(tgt, vtd, win, dt, "cubic", "rk2")
    Ht = A.heaviside(tgt, eps)

    def loss64(v):
        p = p0
        for _ in range(steps):
            p = D.advect_sl(p, v, win, dt, "cubic", "rk2")
        return float(((A.heaviside(p, eps) - Ht) ** 2).sum())

    def grad(dtype):
        v = to_dev(vel, dtype).requires_grad_(True)
        phiT = A.advect_n(p0.to(dtype), v, win, dt, steps, checkpoint_every=every)
        ((A.heaviside(phiT, eps) - Ht.to(dtype)) ** 2).sum().backward()
        return to_np(v.grad)

    g32, g64 = grad(torch.float32), grad(torch.float64)
    # pin the FP64 adjoint chain: its first step against the oracle's adjoint at 128^3
    lam_next = np.random.default_rng(5).normal(size=phi0.shape)
    lam_ref, gv_ref = O.sl_adjoint(phi0, vel, lam_next, owin(win), dt, O.CUBIC, O.RK2)
    lam, gv = A.sl_adjoint(p0, to_dev(vel), to_dev(lam_next), win, dt, "cubic", "rk2")
    assert rel_max(to_np(lam), lam_ref) <= 1e-11 and rel_max(to_np(gv), gv_ref) <= 1e-11
    # central differences of the FP64 forward on 64 components of the gradient's support
    support = np.flatnonzero(np.abs(g64).ravel() > 1e-6 * np.abs(g64).max())
    pick = np.random.default_rng(7).choice(support, 64, replace=False)
    d = 1e-4 * np.abs(vel).max()
    vd = to_dev(vel)
    flat = vd.view(-1)
    fd = np.empty(64)
    for i, f in enumerate(pick):
        v0 = float(flat[f])
        flat[f] = v0 + d
        lp = loss64(vd)
        flat[f] = v0 - d
        lm = loss64(vd)
        flat[f] = v0
        fd[i] = (lp - lm) / (2 * d)
    scale = np.abs(fd).max()
    e32 = np.abs(g32.ravel()[pick] - fd).max() / scale
    e64 = np.abs(g64.ravel()[pick] - fd).max() / scale
    l2 = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    print(f"C5: support {len(support)}, max |fd| {scale:.3e}; FP32 vs FD {e32:.2e}, FP64 vs FD {e64:.2e}, "
          f"FP32 vs FP64 relative L2 {l2:.2e}")
    assert e64 <= 1e-4
    assert e32 <= 1e-3
    assert l2 <= 2e-3
