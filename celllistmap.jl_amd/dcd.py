"""Reader for CHARMM/NAMD DCD trajectories (the fixture format of the reference's NAMD
comparison, test/applications/namd/*.dcd, which the reference reads with Chemfiles).

Format (Fortran unformatted records, each framed by its byte length as int32 before
and after; little- or big-endian):
  1. 84 bytes: b"CORD", then 20 int32 control words: [0] frames, [1] first step,
     [2] steps between frames, [3] total steps, [8] fixed atoms, [9] time step
     (float32 bits), [10] 1 if every frame carries a unit-cell record, [19] CHARMM version
  2. title block: int32 count, count * 80 characters
  3. int32 number of atoms
  then per frame: [unit cell: 6 float64 (A, gamma, B, beta, alpha, C in the CHARMM
  order)], X (natoms float32), Y, Z.
Fixed atoms (control word 8) are not supported.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

__all__ = ["DCDTrajectory", "read_dcd"]


@dataclass
class DCDTrajectory:
    natoms: int
    frames: List[np.ndarray]            # each (natoms, 3) float32
    cells: List[Optional[np.ndarray]]   # each the 6 unit-cell words, or None
    first_step: int
    step_interval: int
    titles: List[str]


def _records(buf: bytes, endian: str):
    off = 0
    n = len(buf)
    while off < n:
        if off + 4 > n:
            raise ValueError("truncated DCD record marker")
        (ln,) = struct.unpack_from(endian + "i", buf, off)
        if ln < 0 or off + 8 + ln > n:
            raise ValueError("corrupt DCD record length")
        body = buf[off + 4: off + 4 + ln]
        (tail,) = struct.unpack_from(endian + "i", buf, off + 4 + ln)
        if tail != ln:
            raise ValueError("DCD record markers disagree")
        yield body
        off += 8 + ln


def read_dcd(path: str) -> DCDTrajectory:
    with open(path, "rb") as f:
        buf = f.read()
    if len(buf) < 92:
        raise ValueError("not a DCD file")
    endian = "<" if struct.unpack_from("<i", buf, 0)[0] == 84 else ">"
    if struct.unpack_from(endian + "i", buf, 0)[0] != 84 or buf[4:8] != b"CORD":
        raise ValueError("not a DCD file (no CORD header)")
    recs = _records(buf, endian)
    head = next(recs)
    ctl = struct.unpack_from(endian + "20i", head, 4)
    nframes, first, interval = ctl[0], ctl[1], ctl[2]
    if ctl[8] != 0:
        raise ValueError("DCD files with fixed atoms are not supported")
    has_cell = ctl[10] != 0 and ctl[19] != 0
    tb = next(recs)
    (ntitle,) = struct.unpack_from(endian + "i", tb, 0)
    titles = [tb[4 + 80 * k: 4 + 80 * (k + 1)].decode("latin-1").rstrip("\x00 ") for k in range(ntitle)]
    (natoms,) = struct.unpack_from(endian + "i", next(recs), 0)
    fdt = np.dtype(endian + "f4")
    frames, cells = [], []
    for _ in range(nframes if nframes > 0 else 1 << 30):
        try:
            cell = None
            if has_cell:
                cell = np.frombuffer(next(recs), dtype=np.dtype(endian + "f8")).copy()
            xyz = [np.frombuffer(next(recs), dtype=fdt) for _ in range(3)]
        except StopIteration:
            break
        if any(c.shape[0] != natoms for c in xyz):
            raise ValueError("DCD coordinate record of the wrong length")
        frames.append(np.stack(xyz, axis=1).astype(np.float32))
        cells.append(cell)
    return DCDTrajectory(natoms, frames, cells, first, interval, titles)
