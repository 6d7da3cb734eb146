"""Parity of the gfx950 SL step (dcl_sl_advect / dcl_levelset_step) with the oracle.

FP64: max |phi_gpu - phi_oracle| <= 1e-12 max |phi| (the FP64 parity bar).
FP32: against the oracle run on the same float32-rounded inputs,
max |diff| <= 2e-6 max |phi| (weights and sums in float32).
"""
import math

import numpy as np
import pytest
import torch

import oracle as O
import dcell_amd as D
from dcell_amd import configs
from helpers import owin, random_velocity, rel_max, rounded, smooth_field, to_dev, to_np

pytestmark = pytest.mark.gpu

TOL = {torch.float64: 1e-12, torch.float32: 2e-6}
IM = {"linear": O.LINEAR, "cubic": O.CUBIC}
BM = {"euler": O.EULER, "rk2": O.RK2}


def _case(dim, n, seed, p=(3, -2, 5), amp=1.0):
    rng = np.random.default_rng(seed)
    h = 1.0 / (n - 1)
    win = D.GridWindow(p, (n, n, n) if dim == 3 else (n, n, 1), (h, 1.3 * h, 0.9 * h))
    phi0 = smooth_field(n, dim, rng)
    vel = random_velocity(n, dim, rng, amp)
    return win, phi0, vel, rng


@pytest.fixture(params=["gather", "ring", "plain"])
def sl_kernel(request, monkeypatch):
    """Every dense 3D SL kernel: per-node buffer-load gathers, phi0 staged in an
    LDS ring, and the pointer-based gather that windows of 2 GiB and more take
    (selected by DCL_SL_KERNEL at every call)."""
    monkeypatch.setenv("DCL_SL_KERNEL", request.param)
    return request.param


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("interp", ["linear", "cubic"])
@pytest.mark.parametrize("bt", ["euler", "rk2"])
@pytest.mark.parametrize("dim", [2, 3])
def test_sl_dense_parity(dev, dtype, interp, bt, dim, sl_kernel):
    n = 40 if dim == 3 else 97
    win, phi0, vel, _ = _case(dim, n, seed=dim * 10 + len(interp) + len(bt))
    phi0, vel = rounded(phi0, dtype), rounded(vel, dtype)
    dt = 1.7 / (n - 1)                                    # up to ~2.5 cells of motion
    out = D.advect_sl(to_dev(phi0, dtype), to_dev(vel, dtype), win, dt, interp, bt)
    ref = O.sl_advect(phi0, vel, owin(win), dt, IM[interp], BM[bt])
    assert rel_max(to_np(out), ref) <= TOL[dtype]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("interp", ["linear", "cubic"])
@pytest.mark.parametrize("bt", ["euler", "rk2"])
@pytest.mark.parametrize("dim", [2, 3])
def test_sl_band_parity_offset_velocity_window(dev, dtype, interp, bt, dim):
    """Band mode with a velocity window larger than and offset from phi's, band
    nodes in random order including the window's edge layers."""
    n = 24 if dim == 3 else 64
    win, phi0, _, rng = _case(dim, n, seed=100 + dim)
    m = n + 6
    vwin = D.GridWindow(tuple(c - 3 for c in win.p), (m, m, m) if dim == 3 else (m, m, 1), win.d)
    vel = random_velocity(m, dim, rng, 1.2)
    phi0, vel = rounded(phi0, dtype), rounded(vel, dtype)
    nodes = np.argwhere(np.abs(phi0) < 4.0)[:, ::-1]
    band = np.zeros((len(nodes), 3), dtype=np.int32)
    band[:, :dim] = nodes + np.array(win.p[:dim])
    band[:, 2] = band[:, 2] if dim == 3 else win.p[2]
    band = band[rng.permutation(len(band))]
    dt = 1.3 / (n - 1)
    vals = D.advect_sl(to_dev(phi0, dtype), to_dev(vel, dtype), win, dt, interp, bt,
                       band=to_dev(band, torch.int32), vel_win=vwin, band_values=True)
    ref = O.sl_advect(phi0, vel, owin(win), dt, IM[interp], BM[bt], band=band, vwin=owin(vwin))
    assert rel_max(to_np(vals), ref, np.max(np.abs(phi0))) <= TOL[dtype]
    # scatter form writes the same values into a copy of phi0, nothing else
    out = D.advect_sl(to_dev(phi0, dtype), to_dev(vel, dtype), win, dt, interp, bt,
                      band=to_dev(band, torch.int32), vel_win=vwin)
    full = phi0.copy()
    idx = tuple((band[:, a] - win.p[a]) for a in range(dim))[::-1]
    full[idx] = ref
    assert rel_max(to_np(out), full) <= TOL[dtype]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("interp", ["linear", "cubic"])
@pytest.mark.parametrize("shape,amp", [((37, 23, 50), 1.0), ((16, 16, 32), 0.3), ((70, 9, 33), 3.0),
                                       ((5, 6, 7), 1.0)])
@pytest.mark.parametrize("tile", ["0", "1", "2"])   # 16 x 16, 32 x 8, 64 x 4 columns
@pytest.mark.parametrize("vring", ["0", "1"])       # velocity from memory / staged in LDS
def test_sl_ring_equals_gather(dev, monkeypatch, dtype, interp, shape, amp, tile, vring):
    """The LDS-ring kernel gives the per-node gather kernel's values bit for bit
    (same arithmetic, stencil values read from LDS instead of memory) for
    every column shape, on windows that are not multiples of the column, several z
    chunks, a flow slow enough to stay in the ring, and one fast enough to
    leave it (those nodes gather from memory)."""
    rng = np.random.default_rng(sum(shape))
    nx, ny, nz = shape
    h = 1.0 / (max(shape) - 1)
    win = D.GridWindow((1, -4, 2), (nx, ny, nz), (h, h, h))
    phi0 = rounded(rng.normal(size=(nz, ny, nx)), dtype)
    vel = rounded(rng.uniform(-amp, amp, size=(nz, ny, nx, 3)), dtype)
    monkeypatch.setenv("DCL_RING_TILE", tile)
    monkeypatch.setenv("DCL_RING_VEL", vring)
    outs = {}
    for mode in ("gather", "ring"):
        monkeypatch.setenv("DCL_SL_KERNEL", mode)
        outs[mode] = to_np(D.advect_sl(to_dev(phi0, dtype), to_dev(vel, dtype), win, h, interp, "rk2"))
    assert np.array_equal(outs["gather"], outs["ring"])
    ref = O.sl_advect(phi0, vel, owin(win), h, IM[interp], O.RK2)
    assert rel_max(outs["ring"], ref, np.max(np.abs(phi0))) <= TOL[dtype] * 4


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("interp,bt", [("cubic", "rk2"), ("linear", "rk2"), ("cubic", "euler")])
@pytest.mark.parametrize("shape,amp", [((37, 23, 50), 1.0), ((70, 9, 33), 3.0), ((5, 6, 7), 1.0)])
@pytest.mark.parametrize("tile", ["0", "1", "2"])
def test_sl_ring_planes_equals_gather(dev, monkeypatch, dtype, interp, bt, shape, amp, tile):
    """Component planes through the LDS-ring kernel (every column shape)
    give the per-node gather kernel's values bit for bit, on windows that
    are not multiples of the column, a flow slow enough to stay in the ring
    and one fast enough to leave it."""
    rng = np.random.default_rng(sum(shape) + 1)
    nx, ny, nz = shape
    h = 1.0 / (max(shape) - 1)
    win = D.GridWindow((1, -4, 2), (nx, ny, nz), (h, h, h))
    phi0 = rounded(rng.normal(size=(nz, ny, nx)), dtype)
    vel = rounded(rng.uniform(-amp, amp, size=(nz, ny, nx, 3)), dtype)
    planes = to_dev(np.ascontiguousarray(np.moveaxis(vel, -1, 0)), dtype)
    monkeypatch.setenv("DCL_SL_KERNEL", "gather")
    ref_gpu = to_np(D.advect_sl(to_dev(phi0, dtype), to_dev(vel, dtype), win, h, interp, bt))
    monkeypatch.setenv("DCL_SL_KERNEL", "ring")
    monkeypatch.setenv("DCL_RING_TILE", tile)
    got = to_np(D.advect_sl(to_dev(phi0, dtype), planes, win, h, interp, bt, planes=True))
    assert np.array_equal(got, ref_gpu)
    ref = O.sl_advect(phi0, vel, owin(win), h, IM[interp], BM[bt])
    assert rel_max(got, ref, np.max(np.abs(phi0))) <= TOL[dtype] * 4


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("interp", ["linear", "cubic"])
@pytest.mark.parametrize("bt", ["euler", "rk2"])
@pytest.mark.parametrize("dim", [2, 3])
def test_sl_component_planes_equal_per_node(dev, dtype, interp, bt, dim, sl_kernel):
    """Velocity as component planes (dof 1, [c][z][y][x]) gives the values of
    the same velocity stored per node, bit for bit (same arithmetic, other
    loads), dense and in band mode with a velocity window offset from phi's;
    and both equal the oracle."""
    n = 24 if dim == 3 else 64
    win, phi0, _, rng = _case(dim, n, seed=300 + dim)
    m = n + 6
    vwin = D.GridWindow(tuple(c - 3 for c in win.p), (m, m, m) if dim == 3 else (m, m, 1), win.d)
    vel = rounded(random_velocity(m, dim, rng, 1.2), dtype)
    phi0 = rounded(phi0, dtype)
    planes = np.ascontiguousarray(np.moveaxis(vel, -1, 0))
    dt = 1.3 / (n - 1)
    p_dev, v_dev, pl_dev = to_dev(phi0, dtype), to_dev(vel, dtype), to_dev(planes, dtype)
    a = to_np(D.advect_sl(p_dev, v_dev, win, dt, interp, bt, vel_win=vwin))
    b = to_np(D.advect_sl(p_dev, pl_dev, win, dt, interp, bt, vel_win=vwin, planes=True))
    assert np.array_equal(a, b)
    ref = O.sl_advect(phi0, vel, owin(win), dt, IM[interp], BM[bt], vwin=owin(vwin))
    assert rel_max(b, ref, np.max(np.abs(phi0))) <= TOL[dtype]
    nodes = np.argwhere(np.abs(phi0) < 4.0)[:, ::-1]
    band = np.zeros((len(nodes), 3), dtype=np.int32)
    band[:, :dim] = nodes + np.array(win.p[:dim])
    band[:, 2] = band[:, 2] if dim == 3 else win.p[2]
    bd = to_dev(band, torch.int32)
    va = to_np(D.advect_sl(p_dev, v_dev, win, dt, interp, bt, band=bd, vel_win=vwin, band_values=True))
    vb = to_np(D.advect_sl(p_dev, pl_dev, win, dt, interp, bt, band=bd, vel_win=vwin, band_values=True,
                           planes=True))
    assert np.array_equal(va, vb)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("interp,bt", [("cubic", "rk2"), ("linear", "euler")])
def test_sl_component_planes_clamped_velocity_window(dev, dtype, interp, bt, sl_kernel):
    """A velocity window that covers only the middle planes of phi's (every
    node above or below it clamps to its end planes, so a thread walking z
    meets planes whose velocity plane does not advance) and ends inside phi's
    in x and y: planes equal the per-node layout and the oracle."""
    n = 28
    win, phi0, _, rng = _case(3, n, seed=77)
    vwin = D.GridWindow((win.p[0] + 2, win.p[1] - 1, win.p[2] + 6), (n - 4, n + 2, n - 13), win.d)
    vel = rounded(rng.uniform(-1.0, 1.0, size=(vwin.n[2], vwin.n[1], vwin.n[0], 3)), dtype)
    phi0 = rounded(phi0, dtype)
    planes = np.ascontiguousarray(np.moveaxis(vel, -1, 0))
    dt = 1.4 / (n - 1)
    a = to_np(D.advect_sl(to_dev(phi0, dtype), to_dev(vel, dtype), win, dt, interp, bt, vel_win=vwin))
    b = to_np(D.advect_sl(to_dev(phi0, dtype), to_dev(planes, dtype), win, dt, interp, bt, vel_win=vwin,
                          planes=True))
    assert np.array_equal(a, b)
    ref = O.sl_advect(phi0, vel, owin(win), dt, IM[interp], BM[bt], vwin=owin(vwin))
    assert rel_max(b, ref, np.max(np.abs(phi0))) <= TOL[dtype]


def test_sl_fluid_layout_udof(dev):
    """The fluid's [p, u, v, w] velocity (udof = 1, FluidField.c:146-151)."""
    win, phi0, vel, rng = _case(3, 20, seed=5)
    vel4 = np.concatenate([rng.normal(size=vel.shape[:-1] + (1,)), vel], axis=-1)
    out = D.advect_sl(to_dev(phi0), to_dev(vel4), win, 0.05, "cubic", "rk2", udof=1)
    ref = O.sl_advect(phi0, vel4, owin(win), 0.05, O.CUBIC, O.RK2, udof=1)
    assert rel_max(to_np(out), ref) <= 1e-12


def test_sl_edge_cases(dev):
    win, phi0, vel, rng = _case(3, 16, seed=6)
    # empty band: nothing happens, no launch
    empty = torch.zeros((0, 3), dtype=torch.int32, device="cuda:0")
    out = D.advect_sl(to_dev(phi0), to_dev(vel), win, 0.1, band=empty)
    assert np.array_equal(to_np(out), phi0)
    # departures far outside the window: clamped exactly as the oracle does
    big = vel * 50.0
    out = D.advect_sl(to_dev(phi0), to_dev(big), win, 0.2, "cubic", "rk2")
    ref = O.sl_advect(phi0, big, owin(win), 0.2, O.CUBIC, O.RK2)
    assert rel_max(to_np(out), ref) <= 1e-12
    # NaN / Inf velocity: sample points clamp into the window, so no fault and a
    # finite result everywhere.  The reference multiplies zero-weight corner faces
    # too (0 * NaN = NaN), the node path does not, so nodes whose 2x2x2 velocity
    # cube holds a bad face are excluded from the comparison.
    bad = vel.copy()
    bad_faces = [(5, 5, 5, 0), (7, 3, 9, 2)]
    bad[bad_faces[0]] = np.nan
    bad[bad_faces[1]] = np.inf
    out = to_np(D.advect_sl(to_dev(phi0), to_dev(bad), win, 0.1, "linear", "euler"))
    ref = O.sl_advect(phi0, bad, owin(win), 0.1, O.LINEAR, O.EULER)
    assert np.all(np.isfinite(out))
    keep = np.ones(out.shape, bool)
    for (z, y, x, _) in bad_faces:
        keep[z - 1:z + 1, y - 1:y + 1, x - 1:x + 1] = False
    assert rel_max(out[keep], ref[keep], np.max(np.abs(phi0))) <= 1e-12
    # argument errors come back as codes, not faults
    with pytest.raises(D._lib.DclError):
        t = to_dev(phi0)
        D.advect_sl(t, to_dev(vel), win, 0.1, out=t)


def test_levelset_step_deformation_c2_shape(dev, sl_kernel):
    """The C2 workload (sphere r 0.15 at 0.35, LeVeque deformation, dt = h,
    RK2 + tricubic, FP64) at 64^3 for 5 steps, and at its full 256^3 size on a
    random sample of 20000 nodes (the oracle's band mode, same arithmetic)."""
    n = 64
    win = D.GridWindow.unit_box(n)
    phi = configs.sphere_sdf(n)
    vel = configs.deformation_velocity(n)
    pd, vd = to_dev(phi), to_dev(vel)
    ref = phi
    for s in range(5):
        dt = configs.deformation_dt(n, s)
        pd, kd = D.levelset_step(pd, vd, win, dt, "cubic", "rk2")
        ref = O.sl_advect(ref, vel, owin(win), dt, O.CUBIC, O.RK2)
        assert rel_max(to_np(pd), ref) <= 1e-12
    kref = O.curvature(ref, owin(win))
    grad_ok = np.abs(ref) < 6
    assert rel_max(to_np(kd)[grad_ok], kref[grad_ok], np.max(np.abs(kref[grad_ok]))) <= 1e-12

    n = 256
    win = D.GridWindow.unit_box(n)
    phi = configs.sphere_sdf(n)
    vel = configs.deformation_velocity(n)
    dt = configs.deformation_dt(n, 0)
    out = to_np(D.advect_sl(to_dev(phi), to_dev(vel), win, dt, "cubic", "rk2"))
    rng = np.random.default_rng(21)
    band = rng.integers(2, n - 2, size=(20000, 3)).astype(np.int32)
    ref = O.sl_advect(phi, vel, owin(win), dt, O.CUBIC, O.RK2, band=band)
    got = out[band[:, 2], band[:, 1], band[:, 0]]
    assert rel_max(got, ref, np.max(np.abs(phi))) <= 1e-12
    # size-independent property: the border layers are phi0 bit for bit
    assert np.array_equal(out[:2], phi[:2]) and np.array_equal(out[:, -2:], phi[:, -2:])
    # the velocity as component planes: the same step bit for bit
    planes = to_dev(np.ascontiguousarray(np.moveaxis(vel, -1, 0)))
    outp = to_np(D.advect_sl(to_dev(phi), planes, win, dt, "cubic", "rk2", planes=True))
    assert np.array_equal(outp, out)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("dim", [2, 3])
def test_levelset_step_planes_equal_per_node(dev, dtype, dim):
    """dcl_levelset_step with component planes: phi and curvature equal the
    per-node velocity's bit for bit."""
    n = 48 if dim == 3 else 120
    win, phi0, vel, _ = _case(dim, n, seed=40 + dim, p=(0, 0, 0))
    phi0, vel = rounded(phi0, dtype), rounded(vel, dtype)
    pd = to_dev(phi0, dtype)
    planes = to_dev(np.ascontiguousarray(np.moveaxis(vel, -1, 0)), dtype)
    for interp, bt in (("cubic", "rk2"), ("linear", "euler"), ("linear", "rk2")):
        a, ka = D.levelset_step(pd, to_dev(vel, dtype), win, 1.5 / n, interp, bt)
        b, kb = D.levelset_step(pd, planes, win, 1.5 / n, interp, bt, planes=True)
        assert torch.equal(a, b) and torch.equal(ka, kb)


def test_zalesak_c1_revolution_fp64(dev):
    """C1 (128^2 Zalesak disk, rigid rotation, dt = 1/628) for 157 steps (a
    quarter revolution) in reference mode (Euler + Bilinear2D) and config mode
    (RK2 + bicubic); GPU and oracle agree step after step."""
    n = 128
    h = 1.0 / (n - 1)
    win = D.GridWindow((0, 0, 0), (n, n, 1), (h, h, h))
    phi = configs.zalesak_sdf(n)
    vel = configs.rotation_velocity(n)
    for interp, bt in (("linear", "euler"), ("cubic", "rk2")):
        pd, vd = to_dev(phi), to_dev(vel)
        ref = phi
        for _ in range(157):
            pd = D.advect_sl(pd, vd, win, 1.0 / 628, interp, bt)
            ref = O.sl_advect(ref, vel, owin(win), 1.0 / 628, IM[interp], BM[bt])
        assert rel_max(to_np(pd), ref) <= 1e-11


def test_c1_steps_as_graph_replay(dev):
    """The C1 step sequence (2D SL + 9-point curvature, ping-pong buffers)
    captured as one HIP graph: every replay, with eager work and
    synchronisations between the replays, equals the eager run bit for bit."""
    n = 128
    h = 1.0 / (n - 1)
    win = D.GridWindow((0, 0, 0), (n, n, 1), (h, h, h))
    start = to_dev(configs.zalesak_sdf(n))
    vel = to_dev(configs.rotation_velocity(n))
    steps = 12
    bufs = [start.clone(), torch.empty_like(start)]
    kap = torch.empty_like(start)

    def run():
        for s in range(steps):
            D.advect_sl(bufs[s % 2], vel, win, 1.0 / 628, "cubic", "rk2", out=bufs[(s + 1) % 2])
            D.curvature(bufs[(s + 1) % 2], win, out=kap)

    run()
    eager_phi, eager_kap = to_np(bufs[steps % 2]).copy(), to_np(kap).copy()
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        bufs[0].copy_(start)
        run()
    torch.cuda.current_stream().wait_stream(side)
    with torch.cuda.graph(g):
        run()
    for rep in range(3):
        other = D.advect_sl(start, vel, win, 1.0 / 628, "linear", "euler")  # eager work between replays
        torch.cuda.synchronize()
        bufs[0].copy_(start)
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(to_np(bufs[steps % 2]), eager_phi), rep
        assert np.array_equal(to_np(kap), eager_kap), rep
        del other
