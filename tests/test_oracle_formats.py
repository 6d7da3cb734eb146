"""Pins of the oracle's output formats (oracle.real64_bytes,
oracle.irregular_rows), against known answers written out byte by byte.

- .Real64 (GridWrite, Common/Serialize.c:23-55; read by
  Visualization/DCellVTK.py:24-32): the PETSc binary vector layout, class id
  1211214 = 0x00127B4E and the count as big-endian int32, then IEEE doubles
  big-endian: 1.0 = 3FF0 0000 0000 0000, -2.5 = C004 0000 0000 0000; reading
  the bytes back the way DCellVTK.py does (skip 2 * 4 bytes, '>f8') returns
  the values, NaN and infinities included; an empty vector is the header.
- node-list rows (IrregularNodes.c:145-219): each field lands in the column
  the reference's row order gives it (a record whose fields are all
  distinct numbers, decoded column by column); 2D skips the projection
  records (axis -1), 3D keeps every record.
"""
import numpy as np

import oracle as O


def test_real64_known_bytes():
    b = O.real64_bytes([1.0, -2.5])
    assert b.hex() == "00127b4e" + "00000002" + "3ff0000000000000" + "c004000000000000"
    assert O.real64_bytes([]).hex() == "00127b4e00000000"


def test_real64_reads_back_like_the_viewer():
    v = np.array([0.0, -0.0, 1e-300, np.inf, -np.inf, np.nan, 3.141592653589793, -7.5])
    b = O.real64_bytes(v.astype(np.float32).astype(np.float64))
    back = np.frombuffer(b[8:], dtype=">f8")
    np.testing.assert_array_equal(back, v.astype(np.float32).astype(np.float64))
    assert np.frombuffer(b[:8], dtype=">i4").tolist() == [1211214, len(v)]


def _records(n, seed=1):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, dtype=O.IRREG_DTYPE)
    r["pos"] = rng.integers(-50, 50, (n, 3))
    r["axis"] = rng.integers(-1, 3, n)
    r["X"] = rng.normal(size=(n, 3))
    s = np.zeros(n, dtype=O.SURF_DTYPE)
    for f in ("n", "s", "r"):
        s[f] = rng.normal(size=(n, 3))
    s["k"] = rng.normal(size=n)
    return r, s, rng.normal(size=(n, 7))


def test_rows_3d_columns():
    r, s, ex = _records(9)
    rows = np.frombuffer(O.irregular_rows(r, s, ex, 3), dtype=">f8").reshape(-1, 23)
    assert rows.shape[0] == 9
    np.testing.assert_array_equal(rows[:, 0:3], r["X"])
    np.testing.assert_array_equal(rows[:, 3:6], r["pos"])
    np.testing.assert_array_equal(rows[:, 6:9], s["n"])
    np.testing.assert_array_equal(rows[:, 9], s["k"])
    np.testing.assert_array_equal(rows[:, 10:16], ex[:, :6])    # f1 f1_n f1_t f1_nn f1_tt f1_nt
    np.testing.assert_array_equal(rows[:, 16:19], s["s"])
    np.testing.assert_array_equal(rows[:, 19:22], s["r"])
    np.testing.assert_array_equal(rows[:, 22], ex[:, 6])       # numNei
    zero = np.frombuffer(O.irregular_rows(r, s, None, 3), dtype=">f8").reshape(-1, 23)
    assert not zero[:, 10:16].any() and not zero[:, 22].any()


def test_rows_2d_columns_skip_projections():
    r, s, ex = _records(40, seed=2)
    ex = ex[:, :4]
    rows = np.frombuffer(O.irregular_rows(r, s, ex, 2), dtype=">f8").reshape(-1, 11)
    keep = r["axis"] != -1
    assert 0 < rows.shape[0] == keep.sum() < len(r)
    np.testing.assert_array_equal(rows[:, 0:2], r["X"][keep, :2])
    np.testing.assert_array_equal(rows[:, 2:4], r["pos"][keep, :2])
    np.testing.assert_array_equal(rows[:, 4:6], s["n"][keep, :2])
    np.testing.assert_array_equal(rows[:, 6:8], ex[keep, :2])   # f1, fa1
    np.testing.assert_array_equal(rows[:, 8], s["k"][keep])
    np.testing.assert_array_equal(rows[:, 9:11], ex[keep, 2:4])  # f1_n, f1_nn
