"""Panels of a flat square grid for press_normvel_der: the neighbour table,
gradient stencils pot_vel_stencil (edge e points to the neighbour across it,
stencil_e = d_e / (2 h^2), so that sum_e stencil_e (f_e - f) is the central
difference of a field f) and a linear surface velocity field."""
import numpy as np


def flat_grid(nx=6, ny=5, h=0.1):
    n = nx * ny
    idx = lambda i, j: i + nx * j
    cen = np.array([[(i + 0.5) * h, (j + 0.5) * h, 0.0] for j in range(ny) for i in range(nx)])
    steps = [(0, -1), (1, 0), (0, 1), (-1, 0)]  # edges 0..3: -y, +x, +y, -x
    neigh = -np.ones((n, 4), np.int32)
    st = np.zeros((n, 4, 3))
    for j in range(ny):
        for i in range(nx):
            for e, (di, dj) in enumerate(steps):
                st[idx(i, j), e, :2] = np.array([di, dj]) * h / (2 * h * h)
                ii, jj = i + di, j + dj
                if 0 <= ii < nx and 0 <= jj < ny:
                    neigh[idx(i, j), e] = idx(ii, jj)
    nor = np.tile([0.0, 0.0, 1.0], (n, 1))
    return dict(n=n, nx=nx, ny=ny, h=h, idx=idx, cen=cen, neigh=neigh, st=st, nor=nor,
                n_ver=np.full(n, 4, np.int32), R=np.tile(np.eye(3), (n, 1, 1)))


def linear_field(cen, a=0.7, b=-0.3, c=1.1, d=0.4, u0=(2.0, -1.0, 0.5)):
    x, y = cen[:, 0], cen[:, 1]
    return np.stack([a * x + u0[0], b * y + u0[1], c * x + d * y + u0[2]], axis=1)
