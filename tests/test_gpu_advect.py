"""DeviceLevelSet.advect_and_reinit (LevelSetAdvectAndReinit,
LevelSetMethod/Advect.c:31-50) with every stage in HBM, checked stage by
stage against the oracle over eight steps of the 40^3 sphere in the
deformation field and of a 96^2 circle in a rigid rotation (dt = 1.2 h, so
the CFL count passes 3 every few steps and reinitialisations happen
mid-run), for both reinitialisations and both SL schemes:

- the SL step over the band from the step's starting phi (FP64, 1e-12 of
  max |phi|; nodes off the band untouched: GridCopy);
- maxVel at the positions of the previous step's irregular-node list (1e-12);
  the CFL and advect counts, and so the steps that reinitialise, as the
  reference's;
- the irregular-node list of the GPU's post-SL phi, bit for bit;
- at a reinitialisation, phi and the band from the GPU's post-SL phi and
  records, bit for bit (oracle fast marching / oracle sweeps)."""
import numpy as np
import pytest
import torch

import oracle as O
import dcell_amd as D
from dcell_amd import configs
from dcell_amd import metrics as M
from dcell_amd.advect import DeviceLevelSet
from helpers import owin, to_np

pytestmark = pytest.mark.gpu

SCHEMES = {("linear", "euler"): (O.LINEAR, O.EULER), ("cubic", "rk2"): (O.CUBIC, O.RK2)}


def _host_records(ls):
    return ls.records[:ls.count * M.IRREGULAR_DTYPE.itemsize].cpu().numpy().view(O.IRREG_DTYPE).copy()


@pytest.mark.parametrize("reinit", ["exact", "parallel"])
@pytest.mark.parametrize("scheme", list(SCHEMES))
@pytest.mark.parametrize("dim", [3, 2])
def test_advect_and_reinit_stages(dev, reinit, scheme, dim):
    n = 40 if dim == 3 else 96
    win = D.GridWindow.unit_box(n, dim)
    w = owin(win)
    h = win.d[0]
    if dim == 3:
        vel_np = configs.deformation_velocity(n)
        phi = torch.from_numpy(configs.sphere_sdf(n)).to("cuda:0")
    else:
        vel_np = 0.6 * configs.rotation_velocity(n)
        phi = torch.from_numpy(configs.circle_sdf(n, (0.55, 0.45), 0.2)).to("cuda:0")
    vel = torch.from_numpy(vel_np).to("cuda:0")
    ls = DeviceLevelSet(phi, win, interp=scheme[0], backtrace=scheme[1], reinit=reinit)
    ls.update_irregular_nodes()
    seen = {}
    orig = ls.reinitialize

    def spy():
        seen["phi"] = to_np(ls.phi).copy()
        seen["recs"] = _host_records(ls)
        orig()

    ls.reinitialize = spy
    cfl, adv, reinits = 0.0, 0, 0
    dt = 1.2 * h
    interp, bt = SCHEMES[scheme]
    for step in range(8):
        phi_b, band_b, recs_b = to_np(ls.phi).copy(), ls.band.cpu().numpy().copy(), _host_records(ls)
        print(f"step {step} band {len(band_b)} records {len(recs_b)}", flush=True)
        seen.clear()
        did = ls.advect_and_reinit(vel, dt)
        # SL over the band
        phi_o = phi_b.copy()
        idx = (band_b[:, 2], band_b[:, 1], band_b[:, 0]) if dim == 3 else (band_b[:, 1], band_b[:, 0])
        phi_o[idx] = O.sl_advect(phi_b, vel_np, w, dt, interp, bt, band=band_b)
        post = seen["phi"] if did else to_np(ls.phi)
        scale = np.abs(phi_o).max()
        assert np.abs(post - phi_o).max() <= 1e-12 * scale
        # CFL increment from the previous list
        mv = O.max_velocity(vel_np, w, recs_b["pos"].copy(), dim=dim) if len(recs_b) else 0.0
        assert abs(ls.max_vel - mv) <= 1e-12 * max(mv, 1.0)
        cfl += dt * mv / h
        adv += 1
        want = adv > 1000 or cfl > 3.0
        assert did == want
        if want:
            cfl, adv = 0.0, 0
            reinits += 1
        # the irregular-node list of the post-SL phi
        recs = seen["recs"] if did else _host_records(ls)
        ref_recs = O.irregular_nodes(post, w, band_b)
        assert recs.tobytes() == ref_recs.tobytes()
        if did:
            if reinit == "parallel":
                rphi, rband, _ = O.reinitialize_sweep(post, w, recs)
            else:
                rphi, rband = O.reinitialize(post, w, recs)
            assert to_np(ls.phi).tobytes() == rphi.tobytes()
            np.testing.assert_array_equal(ls.band.cpu().numpy(), rband)
    assert reinits >= 2 and ls.reinits == reinits
