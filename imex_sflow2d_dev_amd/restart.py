"""Reading a restart state from a .q_2d solution file, as the reference's
read_solution does for RESTART = T with a .q_2d file
(src/inpout_2d.f90:5688-5721): for k = 1..ny, j = 1..nx one record
x, y, q(1..n_vars) (format (2e20.12, 100(e20.12))) and a blank line after
each row; a cell with q1 <= 0 is zeroed, and so is a negative or zero solid
mass.  The number in the file name before the extension is the output index
the run continues from."""
from __future__ import annotations

import os
import re

import numpy as np


def read_q2d(path: str, nx: int, ny: int, n_vars: int, n_solid: int) -> np.ndarray:
    """The conservative state q[k, j, var] of a .q_2d file."""
    q = np.zeros((ny, nx, n_vars))
    with open(path) as f:
        rows = [ln.split() for ln in f if ln.strip()]
    if len(rows) != nx * ny:
        raise ValueError(f"{path}: {len(rows)} records, the grid has {nx * ny} cells")
    for r, vals in enumerate(rows):
        if len(vals) < 2 + n_vars:
            raise ValueError(f"{path}: record {r + 1} has {len(vals)} values, needs {2 + n_vars}")
        k, j = divmod(r, nx)
        q[k, j, :] = [float(v) for v in vals[2:2 + n_vars]]
    dry = q[..., 0] <= 0.0
    q[dry] = 0.0
    for i in range(4, 4 + n_solid):
        s = q[..., i]
        s[s <= 0.0] = 0.0
    return q


def output_index(path: str) -> int:
    """The output index in a restart file name (the four digits before the
    extension, src/inpout_2d.f90:5716-5718)."""
    stem = os.path.splitext(os.path.basename(path))[0]
    m = re.search(r"(\d{4})$", stem)
    if not m:
        raise ValueError(f"{path}: no output index before the extension")
    return int(m.group(1))
