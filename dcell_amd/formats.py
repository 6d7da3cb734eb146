"""DCell's output formats from device tensors, packed on the GPU.

- ``real64_pack`` / ``grid_write``: the .Real64 file GridWrite writes
  (Common/Serialize.c:23-55: a PETSc binary vector, big-endian int32 class
  id 1211214 and count, then big-endian doubles; read back by
  Visualization/DCellVTK.py:24-32), with the ``.pos`` / ``.size`` sidecar
  lines GridWrite appends.
- ``irregular_node_rows`` / ``write_irregular_node_list``: the rows of
  LevelSetWriteIrregularNodeList (LevelSetMethod/IrregularNodes.c:145-236).

The bytes are formed in HBM by ``dcl_real64_pack`` / ``dcl_irregular_node_rows``
and leave the GPU in one copy.
"""
from __future__ import annotations

import ctypes as C
import os

import torch

from . import _lib as L
from .levelset import GridWindow, _dev, _p, precision_of, stream_handle
from .metrics import IRREGULAR_DTYPE, SURFACE_DTYPE

FILE_COUNT_START = 10000  # Common/Common.h:59
ROW_FIELDS = {3: 23, 2: 11}
EXTRA_FIELDS = {3: 7, 2: 4}


def real64_pack(values: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
    """The .Real64 bytes of a device tensor (FP64, or FP32 widened): a uint8
    device tensor of 8 + 8 * values.numel() bytes."""
    _dev(values, "values")
    n = values.numel()
    if out is None:
        out = torch.empty(8 + 8 * n, dtype=torch.uint8, device=values.device)
    _dev(out, "out", torch.uint8)
    if out.numel() < 8 + 8 * n:
        raise ValueError(f"out holds {out.numel()} bytes, {8 + 8 * n} needed")
    L.call("dcl_real64_pack", _p(values), precision_of(values), n, _p(out), stream_handle())
    return out


def grid_write(directory: str, name: str, t: int, values: torch.Tensor, win: GridWindow) -> str:
    """GridWrite (Common/Serialize.c:23-55): ``{directory}/{name}.{t + 10000}.Real64``
    from a device grid, and one line appended to ``{name}.pos`` and
    ``{name}.size`` (the window's p and n; two numbers each for a 2D grid).
    Returns the .Real64 path."""
    if values.numel() != win.nodes:
        raise ValueError(f"values has {values.numel()} elements, window needs {win.nodes}")
    data = real64_pack(values).cpu().numpy()
    path = os.path.join(directory, f"{name}.{t + FILE_COUNT_START}.Real64")
    with open(path, "wb") as f:
        f.write(data.tobytes())
    k = 2 if win.is2d else 3
    with open(os.path.join(directory, f"{name}.pos"), "a") as f:
        f.write(" ".join(str(v) for v in win.p[:k]) + "\n")
    with open(os.path.join(directory, f"{name}.size"), "a") as f:
        f.write(" ".join(str(v) for v in win.n[:k]) + "\n")
    return path


def irregular_node_rows(records: torch.Tensor, count: int, surf: torch.Tensor, dim: int,
                        extra: torch.Tensor | None = None) -> torch.Tensor:
    """The node-list rows (big-endian doubles, 23 per record in 3D, 11 per
    crossing record in 2D) as a uint8 device tensor.  records / surf: uint8
    device tensors of 88-byte records and 104-byte surface quantities (as
    irregular_nodes_into / surface_quantities_into write them); extra: FP64
    device tensor of the IIM fields per record (7 in 3D, 4 in 2D) or None."""
    if dim not in ROW_FIELDS:
        raise ValueError("dim must be 2 or 3")
    _dev(records, "records")
    _dev(surf, "surf")
    if records.numel() < count * IRREGULAR_DTYPE.itemsize or surf.numel() < count * SURFACE_DTYPE.itemsize:
        raise ValueError("records or surf holds fewer than count entries")
    if extra is not None:
        _dev(extra, "extra", torch.float64)
        if extra.numel() < count * EXTRA_FIELDS[dim]:
            raise ValueError("extra holds fewer than count rows of fields")
    nf = ROW_FIELDS[dim]
    out = torch.empty(max(count, 1) * nf * 8, dtype=torch.uint8, device=records.device)
    rows = C.c_int64()
    L.call("dcl_irregular_node_rows", _p(records), _p(surf), _p(extra), int(count), dim, _p(out), max(count, 1),
           C.byref(rows), stream_handle())
    return out[:rows.value * nf * 8]


def write_irregular_node_list(directory: str, name: str, idx: int, records: torch.Tensor, count: int,
                              surf: torch.Tensor, dim: int, extra: torch.Tensor | None = None) -> str:
    """LevelSetWriteIrregularNodeList (LevelSetMethod/IrregularNodes.c:222-236):
    ``{directory}/{name}.irregNode.{idx + 10000}.array``.  Returns the path."""
    data = irregular_node_rows(records, count, surf, dim, extra).cpu().numpy()
    path = os.path.join(directory, f"{name}.irregNode.{idx + FILE_COUNT_START}.array")
    with open(path, "wb") as f:
        f.write(data.tobytes())
    return path
