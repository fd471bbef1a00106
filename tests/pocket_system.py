"""A pressure system whose reference-form preconditioner has P == 1.0 cells.

The reference's P_ref tests the stored value against FLUID == 1
(src/solver.cpp:119-123), so a factor coupling is live only where some P is
exactly 1.0.  A FLUID cell with five WALL neighbours has diag = 6 - 5 = 1
(src/solver.cpp:103-117) and no live predecessor, so e = 1 and P = 1.0
exactly (:125-135).  Here a WALL slab (k = 2..4) under the liquid is
pierced by one-cell tunnels: the tunnel's bottom cell (k = 3) is such a
cell, and the cell above it (k = 4, four WALL neighbours, diag 2) gets
e = 2 - 1 = 1 from its live predecessor, so P == 1.0 as well.  With x0 = 0
and b = -1 at the tunnel bottoms (solve negates b), the first residual there
is exactly 1.0, so the forward sweep's value test on q (:151-156) is also
live for the cell above.
"""
import numpy as np

FLUID, WALL, AIR = 1, 2, 0
TUNNELS = ((3, 3), (3, 7), (6, 5), (8, 2), (8, 8))


def pocket_system(n=12, seed=3):
    A = np.full((n, n, n), AIR, np.int8)
    A[0, :, :] = A[-1, :, :] = WALL
    A[:, 0, :] = A[:, -1, :] = WALL
    A[:, :, 0] = A[:, :, -1] = WALL
    jj = np.arange(n)[None, :, None]
    kk = np.arange(n)[None, None, :]
    interior = A != WALL
    A[interior & np.broadcast_to(jj <= 8, A.shape)] = FLUID
    slab = interior & np.broadcast_to((kk >= 2) & (kk <= 4), A.shape)
    A[slab] = WALL
    for (i, j) in TUNNELS:
        A[i, j, 3] = FLUID
        A[i, j, 4] = FLUID
    L = np.broadcast_to(((jj - 8.5) / n).astype(np.float32), A.shape).copy()
    L[A == WALL] = 1.0
    rs = np.random.RandomState(seed)
    b = np.zeros((n, n, n), np.float32)
    fl = A == FLUID
    b[fl] = rs.normal(size=fl.sum()).astype(np.float32)
    for (i, j) in TUNNELS:
        b[i, j, 3] = -1.0
        b[i, j, 4] = 0.0
    x = np.zeros((n, n, n), np.float32)
    return A, L, x, b


def unit_cells(n=12):
    """(i, j, k) of the cells whose factor is exactly 1.0."""
    return [(i, j, k) for (i, j) in TUNNELS for k in (3, 4)]
