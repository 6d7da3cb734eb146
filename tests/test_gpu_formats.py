"""Parity of the gfx950 format packers with the oracle's restatements:
byte for byte.  .Real64 (dcl_real64_pack, GridWrite): FP64 and FP32 grids
of 0, 1, odd and 2^20 + 3 values, special values, a 160^3 grid checked by
reading the bytes back the way Visualization/DCellVTK.py does, and
grid_write's file and .pos / .size sidecars (3D and 2D).  Node-list rows
(dcl_irregular_node_rows, LevelSetWriteIrregularNodeList): the records and
surface quantities of a sphere (3D) and a circle (2D) computed on the GPU,
with and without the IIM fields, an empty list, and the written file."""
import numpy as np
import pytest
import torch

import oracle as O
import dcell_amd as D
from dcell_amd import formats as F
from dcell_amd import metrics as M

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [0, 1, 7, (1 << 20) + 3])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_real64_pack_bytes(dev, n, dtype):
    x = torch.randn(n, dtype=torch.float64, generator=torch.Generator().manual_seed(n)).to(dtype)
    if n >= 7:
        x[:6] = torch.tensor([0.0, -0.0, float("inf"), float("-inf"), float("nan"), 1e-30], dtype=dtype)
    got = F.real64_pack(x.to("cuda:0")).cpu().numpy().tobytes()
    assert got == O.real64_bytes(x.numpy().astype(np.float64))


def test_real64_large_reads_back(dev):
    n = 160
    phi = torch.randn((n, n, n), dtype=torch.float64, device="cuda:0")
    b = F.real64_pack(phi).cpu().numpy()
    assert np.frombuffer(b[:8].tobytes(), dtype=">i4").tolist() == [1211214, n ** 3]
    back = np.frombuffer(b[8:].tobytes(), dtype=">f8")
    assert back.tobytes() == phi.cpu().numpy().astype(">f8").tobytes()


@pytest.mark.parametrize("dim", [3, 2])
def test_grid_write_files(dev, tmp_path, dim):
    if dim == 3:
        win = D.GridWindow((-3, 4, 1), (12, 10, 9), (1.0, 1.0, 1.0))
    else:
        win = D.GridWindow((5, -2, 0), (12, 10, 1), (1.0, 1.0, 1.0))
    phi = torch.randn(win.shape, dtype=torch.float64, device="cuda:0")
    path = F.grid_write(str(tmp_path), "phi", 0, phi, win)
    F.grid_write(str(tmp_path), "phi", 1, phi * 2, win)
    assert path.endswith("phi.10000.Real64")
    assert open(path, "rb").read() == O.real64_bytes(phi.cpu().numpy())
    assert open(tmp_path / "phi.10001.Real64", "rb").read() == O.real64_bytes(2 * phi.cpu().numpy())
    k = 3 if dim == 3 else 2
    pos = " ".join(map(str, win.p[:k])) + "\n"
    size = " ".join(map(str, win.n[:k])) + "\n"
    assert open(tmp_path / "phi.pos").read() == pos * 2
    assert open(tmp_path / "phi.size").read() == size * 2


def _records_and_surface(dim):
    if dim == 3:
        n = 40
        win = D.GridWindow((0, 0, 0), (n, n, n), (1.0, 1.0, 1.0))
        z, y, x = np.meshgrid(*(np.arange(n, dtype=float),) * 3, indexing="ij")
        phi = np.sqrt((x - 19.3) ** 2 + (y - 20.1) ** 2 + (z - 18.7) ** 2) - 11.2
    else:
        n = 64
        win = D.GridWindow((0, 0, 0), (n, n, 1), (1.0, 1.0, 1.0))
        y, x = np.meshgrid(np.arange(n, dtype=float), np.arange(n, dtype=float), indexing="ij")
        phi = np.hypot(x - 31.4, y - 30.2) - 17.9
    dphi = torch.from_numpy(phi).to("cuda:0")
    band, nb = M.band_compact(dphi, win, 6.0)
    recs = torch.empty(10 * nb * M.IRREGULAR_DTYPE.itemsize, dtype=torch.uint8, device="cuda:0")
    count = M.irregular_nodes_into(dphi, win, band.contiguous(), recs)
    surf = torch.empty(max(count, 1) * M.SURFACE_DTYPE.itemsize, dtype=torch.uint8, device="cuda:0")
    M.surface_quantities_into(dphi, win, recs, count, surf)
    h_recs = recs[:count * M.IRREGULAR_DTYPE.itemsize].cpu().numpy().view(O.IRREG_DTYPE)
    h_surf = surf[:count * M.SURFACE_DTYPE.itemsize].cpu().numpy().view(O.SURF_DTYPE)
    return recs, surf, count, h_recs, h_surf


@pytest.mark.parametrize("dim", [3, 2])
@pytest.mark.parametrize("with_extra", [False, True])
def test_irregular_rows_bytes(dev, tmp_path, dim, with_extra):
    recs, surf, count, h_recs, h_surf = _records_and_surface(dim)
    assert count > 100
    extra = None
    h_extra = None
    if with_extra:
        h_extra = np.random.default_rng(dim).normal(size=(count, F.EXTRA_FIELDS[dim]))
        extra = torch.from_numpy(h_extra).to("cuda:0")
    got = F.irregular_node_rows(recs, count, surf, dim, extra).cpu().numpy().tobytes()
    assert got == O.irregular_rows(h_recs, h_surf, h_extra, dim)
    path = F.write_irregular_node_list(str(tmp_path), "phi", 3, recs, count, surf, dim, extra)
    assert path.endswith("phi.irregNode.10003.array")
    assert open(path, "rb").read() == got


def test_irregular_rows_empty(dev):
    e = torch.zeros(0, dtype=torch.uint8, device="cuda:0")
    assert F.irregular_node_rows(e, 0, e, 3).numel() == 0


def test_irregular_rows_2d_many_tiles(dev):
    """2D rows of 300 000 synthetic records (about a third projections): the
    kept records' positions come from a scan over more than one scan tile."""
    n = 300_000
    rng = np.random.default_rng(11)
    recs = np.zeros(n, dtype=O.IRREG_DTYPE)
    recs["pos"] = rng.integers(-1000, 1000, (n, 3))
    recs["axis"] = rng.integers(-1, 2, n)
    recs["X"] = rng.normal(size=(n, 3))
    surf = np.zeros(n, dtype=O.SURF_DTYPE)
    surf["n"] = rng.normal(size=(n, 3))
    surf["k"] = rng.normal(size=n)
    extra = rng.normal(size=(n, 4))
    d_recs = torch.from_numpy(recs.view(np.uint8).copy()).to("cuda:0")
    d_surf = torch.from_numpy(surf.view(np.uint8).copy()).to("cuda:0")
    d_extra = torch.from_numpy(extra).to("cuda:0")
    got = F.irregular_node_rows(d_recs, n, d_surf, 2, d_extra).cpu().numpy().tobytes()
    assert len(got) == (recs["axis"] != -1).sum() * 11 * 8
    assert got == O.irregular_rows(recs, surf, extra, 2)


@pytest.mark.parametrize("dim", [3, 2])
def test_irregular_rows_capacity(dev, dim):
    """Through the C ABI with a capacity below the row count, ending inside a
    block of records: rows_out is the whole count, the first capacity rows are
    the oracle's, and the bytes past capacity are untouched."""
    import ctypes as C
    from dcell_amd import _lib as L
    from dcell_amd.levelset import _p, stream_handle
    recs, surf, count, h_recs, h_surf = _records_and_surface(dim)
    nf = F.ROW_FIELDS[dim]
    ref = np.frombuffer(O.irregular_rows(h_recs, h_surf, None, dim), dtype=np.uint8)
    total = ref.size // (nf * 8)
    cap = total // 2 + 37
    assert cap % 64 not in (0, 63) and cap < total
    out = torch.full(((cap + 5) * nf * 8,), 0xA5, dtype=torch.uint8, device="cuda:0")
    rows = C.c_int64()
    L.call("dcl_irregular_node_rows", _p(recs), _p(surf), None, int(count), dim, _p(out), cap, C.byref(rows),
           stream_handle())
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert rows.value == total
    assert got[:cap * nf * 8].tobytes() == ref[:cap * nf * 8].tobytes()
    assert np.all(got[cap * nf * 8:] == 0xA5)
