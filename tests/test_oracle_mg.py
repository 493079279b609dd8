"""The oracle's multigrid preconditioner (-pc_type mg) pinned by dense
linear algebra:
- the transfer operators are x-y full weighting (weights 1/4, 1/2, 1/4 per
  axis, periodic) and bilinear interpolation, built here from their stencils,
  and restriction is a quarter of the transpose of prolongation;
- the level sizes halve x and y while both stay even and at least 3;
- one application equals the V(2,2) cycle written with dense matrices: the
  level operators from orc_lhs_dense on each level's grid and restricted
  density, the smoother x += 0.8 (I - 11^T/n) Z^-1 (b - A x) with Z the
  same-column (z-line) part of A, 8 sweeps on the coarsest level."""
import numpy as np
import pytest

import orc

K = (0.0, 0.0, 88.0)


def _full_weighting(N):
    nx, ny, nz = N
    cx, cy = nx // 2, ny // 2
    R = np.zeros((cx * cy * nz, nx * ny * nz))
    w = (0.25, 0.5, 0.25)
    for k in range(nz):
        for J in range(cy):
            for I in range(cx):
                for b in range(3):
                    for a in range(3):
                        i, j = (2 * I + a - 1) % nx, (2 * J + b - 1) % ny
                        R[I + cx * (J + cy * k), i + nx * (j + ny * k)] += w[b] * w[a]
    return R


def _bilinear(N):
    nx, ny, nz = N
    cx, cy = nx // 2, ny // 2
    P = np.zeros((nx * ny * nz, cx * cy * nz))
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                xs = [(i // 2, 1.0)] if i % 2 == 0 else [(i // 2, 0.5), ((i // 2 + 1) % cx, 0.5)]
                ys = [(j // 2, 1.0)] if j % 2 == 0 else [(j // 2, 0.5), ((j // 2 + 1) % cy, 0.5)]
                for I, wx in xs:
                    for J, wy in ys:
                        P[i + nx * (j + ny * k), I + cx * (J + cy * k)] += wx * wy
    return P


@pytest.mark.parametrize("N", [(8, 6, 4), (12, 10, 3)])
def test_transfer_operators(N):
    rng = np.random.default_rng(N[0])
    R, P = _full_weighting(N), _bilinear(N)
    nf = int(np.prod(N))
    f = rng.standard_normal(nf)
    c = rng.standard_normal(R.shape[0])
    assert np.allclose(orc.mg_restrict(N, f), R @ f, rtol=0, atol=1e-14)
    assert np.allclose(orc.mg_prolong_add(N, c, f) - f, P @ c, rtol=0, atol=1e-14)
    assert np.allclose(R, 0.25 * P.T, rtol=0, atol=0)


def test_level_sizes():
    assert orc.mg_levels((48, 40, 10)) == [(48, 40, 10), (24, 20, 10), (12, 10, 10), (6, 5, 10)]
    assert orc.mg_levels((16, 8, 12)) == [(16, 8, 12), (8, 4, 12)]
    assert orc.mg_levels((9, 8, 4)) == [(9, 8, 4)]
    assert orc.mg_levels((12, 12, 5)) == [(12, 12, 5), (6, 6, 5), (3, 3, 5)]


def _vcycle(levels, b, l=0):
    A, Z, R, P = levels[l]
    n = A.shape[0]
    M = np.eye(n) - np.ones((n, n)) / n
    S = 0.8 * M @ np.linalg.inv(Z)
    x = np.zeros(n)
    coarsest = l == len(levels) - 1
    for _ in range(8 if coarsest else 2):
        x = x + S @ (b - A @ x)
    if coarsest:
        return x
    x = x + P @ _vcycle(levels, R @ (b - A @ x), l + 1)
    for _ in range(2):
        x = x + S @ (b - A @ x)
    return x


@pytest.mark.parametrize("N", [(12, 12, 5), (16, 8, 4), (9, 8, 4)])
def test_application_is_the_dense_v_cycle(N):
    rng = np.random.default_rng(sum(N))
    g = orc.Grid(N, p1=tuple(0.05 * v for v in N))
    n = 16 * (1 + 0.1 * rng.uniform(-1, 1, g.nc))
    levels = []
    dens, d = n, np.array(g.d)
    sizes = orc.mg_levels(N)
    for l, Nl in enumerate(sizes):
        gl = orc.Grid(Nl, p1=tuple(d[a] * Nl[a] for a in range(3)))
        A = orc.lhs_dense(gl, K, dens)
        col = np.arange(gl.nc) % (Nl[0] * Nl[1])
        Z = np.where(col[:, None] == col[None, :], A, 0.0)
        R = _full_weighting(Nl) if l + 1 < len(sizes) else None
        P = _bilinear(Nl) if l + 1 < len(sizes) else None
        levels.append((A, Z, R, P))
        if R is not None:
            dens = R @ dens
            d = d * np.array([2.0, 2.0, 1.0])
    r = rng.standard_normal(g.nc)
    z = orc.pc_apply(g, K, n, r, pc=orc.PC_MG)
    ref = _vcycle(levels, r)
    assert np.abs(z - ref).max() <= 1e-10 * np.abs(ref).max()
