"""Pins of the oracle forcing (reference rhs.c:8-303, calculus.c:21-518).

Single Fourier modes make every centred difference exact in closed form:
D_a cos(t.i) = -sin(t.i) sin(t_a)/d_a, D_aa = (2cos t_a - 2)/d_a^2,
D_ab = -sin t_a sin t_b /(d_a d_b).  All E0, K and Te components are non-zero
so every term (including the y-z cross derivative) is exercised.
"""
import numpy as np

import orc

K = (0.9, -1.4, 2.3)
E0 = (0.05, -0.02, 0.03)
CTH = orc.KB * 1500.0 / orc.Q
CG = orc.ME * 1e5 / orc.Q


def tensor(K):
    Kx, Ky, Kz = K
    return np.array([[1 + Kx * Kx, Ky * Kx - Kz, Kz * Kx + Ky],
                     [Kx * Ky + Kz, 1 + Ky * Ky, Kz * Ky - Kx],
                     [Kx * Kz - Ky, Ky * Kz + Kx, 1 + Kz * Kz]]), 1 + Kx * Kx + Ky * Ky + Kz * Kz


def grid_idx(g):
    k, j, i = np.meshgrid(*[np.arange(int(m)) for m in g.shape()], indexing="ij")
    return i.ravel(), j.ravel(), k.ravel()


def test_full_rhs_single_mode_closed_form():
    g = orc.Grid((8, 10, 6), p1=(0.4, 0.5, 0.3))
    i, j, k = grid_idx(g)
    R, detA = tensor(K)
    tn = [2 * np.pi * m / g.N[a] for a, m in enumerate((1, 2, 1))]
    tg = [[2 * np.pi * m / g.N[a] for a, m in enumerate(v)] for v in ((2, 1, 1), (1, 3, 2), (3, 1, 2))]
    ph_n = tn[0] * i + tn[1] * j + tn[2] * k
    n = 16.0 + 0.7 * np.cos(ph_n)
    G = [0.3 * np.cos(t[0] * i + t[1] * j + t[2] * k) for t in tg]
    # exact discrete derivatives
    Dn = [-0.7 * np.sin(ph_n) * np.sin(tn[a]) / g.d[a] for a in range(3)]
    D2n = np.zeros((3, 3, g.nc))
    for a in range(3):
        for b in range(3):
            if a == b:
                D2n[a, b] = 0.7 * np.cos(ph_n) * (2 * np.cos(tn[a]) - 2) / g.d[a] ** 2
            else:
                D2n[a, b] = -0.7 * np.cos(ph_n) * np.sin(tn[a]) * np.sin(tn[b]) / (g.d[a] * g.d[b])
    divG = sum(-0.3 * np.sin(t[0] * i + t[1] * j + t[2] * k) * np.sin(t[a]) / g.d[a]
               for a, t in enumerate(tg))
    RE = R @ np.array(E0)
    Eterm = sum(RE[a] * Dn[a] for a in range(3))
    Pterm = CTH * sum(R[a, b] * D2n[a, b] for a in range(3) for b in range(3))
    Gterm = detA * CG * divG
    expect = (Eterm + Pterm + Gterm) / detA
    expect -= expect.mean()
    b = orc.rhs(g, K, E0, CTH, CG, n, G[0], G[1], G[2])
    assert np.abs(b - expect).max() <= 1e-12 * np.abs(expect).max()
    assert abs(b.mean()) <= 1e-15 * np.abs(b).max()


def test_full_rhs_linear_and_constant_fields_give_zero():
    g = orc.Grid((5, 6, 7))
    rng = np.random.default_rng(7)
    f = [rng.standard_normal(g.nc) for _ in range(4)]
    h = [rng.standard_normal(g.nc) for _ in range(4)]
    b1 = orc.rhs(g, K, E0, CTH, CG, *f)
    b2 = orc.rhs(g, K, E0, CTH, CG, *h)
    b12 = orc.rhs(g, K, E0, CTH, CG, *[2.0 * a - 3.0 * c for a, c in zip(f, h)])
    assert np.allclose(b12, 2 * b1 - 3 * b2, rtol=0, atol=1e-12 * np.abs(b12).max())
    const = orc.rhs(g, K, E0, CTH, CG, *[np.full(g.nc, v) for v in (16.0, 1.0, 2.0, 3.0)])
    assert np.abs(const).max() <= 1e-20


def test_constant_and_sinusoidal_rhs():
    g = orc.Grid((4, 6, 8), p1=(0.2, 0.3, 0.4))
    rng = np.random.default_rng(8)
    n = 16 + rng.standard_normal(g.nc)
    _, detA = tensor(K)
    vol2 = 2.0 * float(np.prod(g.d)) / detA
    bc = orc.rhs(g, K, E0, CTH, CG, n, n, n, n, rhs_type=orc.RHS_CONSTANT)
    assert np.allclose(bc, n.mean() * vol2, rtol=1e-14)
    bs = orc.rhs(g, K, E0, CTH, CG, n, n, n, n, rhs_type=orc.RHS_SINUSOIDAL)
    i, j, k = grid_idx(g)
    v = np.cos(2 * np.pi * (i + 0.5) * g.d[0]) * np.cos(2 * np.pi * (j + 0.5) * g.d[1]) \
        * np.cos(2 * np.pi * (k + 0.5) * g.d[2])
    e = n.mean() * v * vol2
    e -= e.mean()
    assert np.allclose(bs, e, rtol=0, atol=1e-13 * np.abs(e).max())
