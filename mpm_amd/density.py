"""Density raster output: the replacement of the reference's writeToOpenVDB
(include/mpm_solver3d.h:73, src/mpm_solver3d.cu:510-572).

The reference rasterises density = sum_p m_p w_ip / h^3 on the host into an
OpenVDB FloatGrid named "density" (background 0) and writes one `.vdb` file
per frame.  OpenVDB is not available here, so the device raster is read
back either dense (reference node order) or as the 4x4x4 node blocks it
touched, and written as a small block file (format below, also documented
with `mpm_write_density` in include/mpm_hip.h).

File layout (little-endian):

    0   char[8]   b"MPMDENS1"
    8   int32[3]  grid resolution
    20  float[3]  grid origin
    32  float     stride h
    36  int32     block edge (4)
    40  int64     block count B
    48  B records: int32 first node[3], float density[64] (node first + (i,j,k)
        at i + 4j + 16k)
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

MAGIC = b"MPMDENS1"
_HEADER = np.dtype([("magic", "S8"), ("res", "<i4", 3), ("origin", "<f4", 3), ("h", "<f4"), ("edge", "<i4"),
                    ("count", "<i8")])
_RECORD = np.dtype([("first", "<i4", 3), ("density", "<f4", 64)])


@dataclass
class DensityBlocks:
    """Sparse density raster: `first[b]` is the first node (x, y, z) of block b
    and `values[b, k, j, i]` the density at node first[b] + (i, j, k)."""
    resolution: tuple
    origin: tuple
    stride: float
    first: np.ndarray    # (B, 3) int32
    values: np.ndarray   # (B, 4, 4, 4) float32, indexed [k, j, i]

    @property
    def count(self) -> int:
        return int(self.first.shape[0])

    def dense(self) -> np.ndarray:
        """Dense grid in the reference node order (flat, x + rx*(y + ry*z));
        nodes outside every block hold the background 0."""
        rx, ry, rz = self.resolution
        out = np.zeros((rz, ry, rx), np.float32)
        for b in range(self.count):
            x0, y0, z0 = (int(v) for v in self.first[b])
            x1, y1, z1 = min(x0 + 4, rx), min(y0 + 4, ry), min(z0 + 4, rz)
            out[z0:z1, y0:y1, x0:x1] = self.values[b, :z1 - z0, :y1 - y0, :x1 - x0]
        return out.reshape(-1)


def read_density_file(path: str) -> DensityBlocks:
    """Read a file written by `mpm_write_density` / `Solver.write_density`."""
    with open(path, "rb") as f:
        raw = f.read()
    if len(raw) < _HEADER.itemsize:
        raise ValueError(f"{path}: truncated header")
    h = np.frombuffer(raw, _HEADER, count=1)[0]
    if bytes(h["magic"]) != MAGIC:
        raise ValueError(f"{path}: not a density block file")
    if int(h["edge"]) != 4:
        raise ValueError(f"{path}: unsupported block edge {int(h['edge'])}")
    n = int(h["count"])
    body = raw[_HEADER.itemsize:]
    if len(body) != n * _RECORD.itemsize:
        raise ValueError(f"{path}: {len(body)} bytes of records, expected {n * _RECORD.itemsize}")
    rec = np.frombuffer(body, _RECORD, count=n)
    return DensityBlocks(tuple(int(v) for v in h["res"]), tuple(float(v) for v in h["origin"]), float(h["h"]),
                         np.array(rec["first"], np.int32).reshape(n, 3),
                         np.array(rec["density"], np.float32).reshape(n, 4, 4, 4))
