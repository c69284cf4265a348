"""Synthetic CT label images for the CT-geometry tests.

The reference's patient images (CT_3D_*.vtk) are not part of its tree, so the
tests write legacy ASCII VTK images of the same form: labels 1 (lumen) and 0
(wall) in a box placed inside the grid, with an inlet and an outlet disk on the
image's first plane and a branch disk on its last.
"""
from pathlib import Path

import numpy as np

# grid and placement used by the tests: 33 x 25 x 33 points, image 16 x 28 x 10
GRID = (33, 25, 33)
LENGTHS = (12.0, 9.0, 12.0)
OFFSETS = (4, 3, 11)  # x_offset (image axis 1 -> grid y), y_offset (image axis 2 -> grid x), z_offset
DIMS = (16, 28, 10)


def label_image(dims=DIMS, offsets=OFFSETS, inlet=(9, 12, 5.0), outlet=(24, 12, 4.0), branch=(17, 13, 3.0)):
    """labels[i, j, k] of the image (1-based axes as i, j, k); disks given in grid
    (x, y, radius in cells): inlet and outlet through every plane below the
    last, the branch on the last plane only."""
    d1, d2, d3 = dims
    xo, yo, _ = offsets
    lab = np.zeros((d1, d2, d3), dtype=np.int64)
    i = np.arange(1, d1 + 1)[:, None]  # grid y = i + x_offset
    j = np.arange(1, d2 + 1)[None, :]  # grid x = j + y_offset
    gy, gx = i + xo, j + yo
    def disk(cx, cy, r):
        return (gx - cx) ** 2 + (gy - cy) ** 2 <= r * r
    for k in range(d3 - 1):
        lab[:, :, k] = disk(*inlet) | disk(*outlet)
    lab[:, :, d3 - 1] = disk(*branch)
    return lab


def write_vtk(path: Path, lab: np.ndarray):
    d1, d2, d3 = lab.shape
    vals = lab.transpose(2, 1, 0).reshape(-1)  # value index (k-1) d2 d1 + (j-1) d1 + i
    with open(path, "w") as f:
        f.write("# vtk DataFile Version 3.0\nsynthetic CT labels\nASCII\nDATASET STRUCTURED_POINTS\n")
        f.write(f"DIMENSIONS {d1} {d2} {d3}\nSPACING 1 1 1\nORIGIN 0 0 0\nPOINT_DATA {vals.size}\n")
        f.write("SCALARS labels int 1\nLOOKUP_TABLE default\n")
        for s in range(0, vals.size, 16):
            f.write(" ".join(str(int(v)) for v in vals[s:s + 16]) + "\n")


def read_waveform(name):
    p = Path(__file__).resolve().parent / "golden" / name
    a = np.loadtxt(p, delimiter=",", usecols=(0, 1))
    return np.ascontiguousarray(a[:, 0]), np.ascontiguousarray(a[:, 1])
