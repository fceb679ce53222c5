"""Global node numbering of a structured hex box (the reference's MakeIds,
dawgsMain.cpp:49-143, at one rank).

Element (I,J,K) of an nx*ny*nz box has (p+1)^3 nodes at local index
((K*ny + J)*nx + I)*Np + (k*Nq + j)*Nq + i, with global id
(I*p+i) + (J*p+j)*(p*nx+1) + (K*p+k)*(p*nx+1)*(p*ny+1) + 1
(ids are 1-based; 0 means "not participating").  Face, edge and corner
nodes are shared by 2, 4 and 8 elements.
"""
from __future__ import annotations

import numpy as np


def make_ids(nx: int, ny: int, nz: int, p: int) -> np.ndarray:
    q = p + 1
    sx = np.int64(1)
    sy = np.int64(p * nx + 1)
    sz = sy * np.int64(p * ny + 1)
    I = np.arange(nx, dtype=np.int64) * p
    J = np.arange(ny, dtype=np.int64) * p
    K = np.arange(nz, dtype=np.int64) * p
    i = np.arange(q, dtype=np.int64)
    # axes: K, J, I, k, j, i  (local index order)
    gz = (K[:, None, None, None, None, None] + i[None, None, None, :, None, None]) * sz
    gy = (J[None, :, None, None, None, None] + i[None, None, None, None, :, None]) * sy
    gx = (I[None, None, :, None, None, None] + i[None, None, None, None, None, :]) * sx
    ids = np.empty((nz, ny, nx, q, q, q), dtype=np.int64)
    np.add(gz, gy, out=ids)
    ids += gx
    ids += 1
    return ids.reshape(-1)


def mesh_counts(nx: int, ny: int, nz: int, p: int) -> dict:
    """Closed forms: local DOFs and NgatherGlobal = (p*nx+1)(p*ny+1)(p*nz+1)."""
    return {"N": nx * ny * nz * (p + 1) ** 3,
            "NgatherGlobal": (p * nx + 1) * (p * ny + 1) * (p * nz + 1)}
