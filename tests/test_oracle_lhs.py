"""Pins of the oracle's 19-point operator (reference lhs.c:144-543).

Each check is independent of the row formula it tests: null spaces, symmetry,
the continuum-derived Fourier symbol, the laplacian relation and the
transpose identity against dense assembly.
"""
import numpy as np
import pytest

import orc

# B along z at the configs' field: Omega_e / nu_e = |q B / m_e| / nu_e.
KZ = orc.Q * 5e-5 / orc.ME / 1e5


def rand_n(g, rng, amp=0.01, n0=16.0):
    return n0 * (1.0 + rng.uniform(-amp, amp, g.nc))


@pytest.mark.parametrize("K", [(0.0, 0.0, KZ), (3.0, 2.0, KZ), (0.7, -1.3, 0.4)])
def test_row_and_column_sums_vanish(K):
    rng = np.random.default_rng(1)
    g = orc.Grid((5, 6, 7), p1=(0.25, 0.3, 0.35))
    n = rand_n(g, rng, amp=0.3)
    A = orc.lhs_dense(g, K, n)
    scale = np.abs(A).max()
    assert np.abs(A.sum(axis=1)).max() <= 1e-13 * scale
    assert np.abs(A.sum(axis=0)).max() <= 1e-13 * scale


def test_symmetric_for_uniform_density_asymmetric_otherwise():
    g = orc.Grid((4, 5, 6), p1=(0.2, 0.25, 0.3))
    K = (3.0, 2.0, KZ)
    A = orc.lhs_dense(g, K, np.full(g.nc, 16.0))
    assert np.abs(A - A.T).max() <= 1e-14 * np.abs(A).max()
    rng = np.random.default_rng(2)
    B = orc.lhs_dense(g, K, rand_n(g, rng, 0.1))
    assert np.abs(B - B.T).max() > 1e-8 * np.abs(B).max()


@pytest.mark.parametrize("K", [(0.0, 0.0, KZ), (3.0, -2.0, 5.0)])
def test_fourier_symbol_matches_continuum_discretisation(K):
    """A(nbar) cos(k.r) = lam(k) cos(k.r) with lam from (nbar/detA) sum_ab r_ab D_a D_b."""
    g = orc.Grid((8, 6, 10), p1=(0.4, 0.3, 0.5))
    nbar = 16.0
    Kx, Ky, Kz = K
    R = np.array([[1 + Kx * Kx, Ky * Kx - Kz, Kz * Kx + Ky],
                  [Kx * Ky + Kz, 1 + Ky * Ky, Kz * Ky - Kx],
                  [Kx * Kz - Ky, Ky * Kz + Kx, 1 + Kz * Kz]])
    detA = 1 + Kx * Kx + Ky * Ky + Kz * Kz
    k, j, i = np.meshgrid(*[np.arange(int(m)) for m in g.shape()], indexing="ij")
    idx = (i, j, k)
    for kv in [(1, 0, 0), (0, 2, 1), (3, 1, 4), (4, 3, 5)]:
        th = [2 * np.pi * kv[a] / g.N[a] for a in range(3)]
        mode = np.cos(th[0] * i + th[1] * j + th[2] * k).ravel()
        lam = 0.0
        for a in range(3):
            for b in range(3):
                if a == b:
                    lam += R[a, a] * (2 * np.cos(th[a]) - 2) / g.d[a] ** 2
                else:
                    lam -= R[a, b] * np.sin(th[a]) * np.sin(th[b]) / (g.d[a] * g.d[b])
        lam *= nbar / detA
        y = orc.lhs_apply(g, K, np.full(g.nc, nbar), mode)
        assert np.abs(y - lam * mode).max() <= 1e-11 * np.abs(lam) + 1e-12, kv
    _ = idx


def test_laplacian_is_full_at_zero_magnetisation_unit_density():
    g = orc.Grid((5, 4, 6), p1=(0.5, 0.6, 0.3))
    rng = np.random.default_rng(3)
    x = rng.standard_normal(g.nc)
    lap = orc.lhs_apply(g, (0, 0, 0), np.zeros(g.nc), x, lhs_type=orc.LHS_LAPLACIAN)
    full = orc.lhs_apply(g, (0, 0, 0), np.ones(g.nc), x, lhs_type=orc.LHS_FULL)
    vol = float(np.prod(g.d))
    assert np.allclose(lap, full * vol, rtol=1e-13, atol=1e-13 * np.abs(lap).max())


def test_identity_lhs():
    g = orc.Grid((3, 4, 5))
    x = np.arange(g.nc, dtype=float)
    assert np.array_equal(orc.lhs_apply(g, (1, 2, 3), np.ones(g.nc), x, lhs_type=orc.LHS_IDENTITY), x)


def test_apply_and_transpose_match_dense():
    g = orc.Grid((4, 5, 3), p1=(0.2, 0.25, 0.15))
    rng = np.random.default_rng(4)
    K = (1.5, -0.5, KZ)
    n = rand_n(g, rng, 0.2)
    A = orc.lhs_dense(g, K, n)
    x = rng.standard_normal(g.nc)
    y = orc.lhs_apply(g, K, n, x)
    yt = orc.lhs_apply(g, K, n, x, transpose=True)
    s = np.abs(A).max() * np.abs(x).max()
    assert np.abs(y - A @ x).max() <= 1e-13 * s
    assert np.abs(yt - A.T @ x).max() <= 1e-13 * s


def test_row_formula_column_order():
    """Stencil offsets follow lhs.c:11-104 and every row has 19 distinct columns."""
    off = orc.offsets19()
    assert off.tolist()[:6] == [[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]]
    assert off.tolist()[18] == [0, 0, 0]
    assert len({tuple(o) for o in off.tolist()}) == 19
    assert all(sum(abs(v) for v in o) <= 2 for o in off.tolist())


def _spectral_div_n_R_grad(g, n, phi, R, transpose_R):
    """(1/detA) div(n R grad phi) by exact spectral differentiation."""
    sh = g.shape()
    ks = [2j * np.pi * np.fft.fftfreq(int(g.N[a]), d=float(g.d[a])) for a in range(3)]
    kz, ky, kx = np.meshgrid(ks[2], ks[1], ks[0], indexing="ij")
    kk = (kx, ky, kz)
    P = np.fft.fftn(phi.reshape(sh))
    grad = [np.real(np.fft.ifftn(kk[a] * P)) for a in range(3)]
    Rm = R.T if transpose_R else R
    flux = [n.reshape(sh) * sum(Rm[a, b] * grad[b] for b in range(3)) for a in range(3)]
    div = sum(np.real(np.fft.ifftn(kk[a] * np.fft.fftn(flux[a]))) for a in range(3))
    return div.ravel()


def test_variable_density_operator_converges_to_div_n_R_grad():
    """Second-order consistency with (1/detA) div(n R grad phi): the Hall
    (antisymmetric r) terms only survive for non-uniform n, so a wrong sign or
    dropped term in any of the 19 coefficients breaks the O(h^2) rate."""
    K = (0.8, -0.6, 1.7)
    Kx, Ky, Kz = K
    R = np.array([[1 + Kx * Kx, Ky * Kx - Kz, Kz * Kx + Ky],
                  [Kx * Ky + Kz, 1 + Ky * Ky, Kz * Ky - Kx],
                  [Kx * Kz - Ky, Ky * Kz + Kx, 1 + Kz * Kz]])
    detA = 1 + Kx * Kx + Ky * Ky + Kz * Kz
    errs = {False: [], True: []}
    for M in (16, 32):
        g = orc.Grid((M, M, M))
        k, j, i = np.meshgrid(*[np.arange(M)] * 3, indexing="ij")
        x, y, z = (i * g.d[0], j * g.d[1], k * g.d[2])
        n = (16.0 * (1 + 0.3 * np.cos(2 * np.pi * x) * np.sin(2 * np.pi * y + 0.3)
                     + 0.2 * np.sin(2 * np.pi * z))).ravel()
        phi = (np.sin(2 * np.pi * x + 0.1) * np.cos(2 * np.pi * y)
               + 0.5 * np.cos(2 * np.pi * (y + z))).ravel()
        y_disc = orc.lhs_apply(g, K, n, phi)
        for tr in (False, True):
            ref = _spectral_div_n_R_grad(g, n, phi, R, tr) / detA
            errs[tr].append(np.abs(y_disc - ref).max() / np.abs(ref).max())
    # the operator is (1/detA) sum_a d_a(n sum_b r_ab d_b phi), second order;
    # the transposed tensor is not (this is what discriminates Hall signs)
    e16, e32 = errs[False]
    assert e32 < 0.01
    assert 3.5 < e16 / e32 < 4.5, errs
    assert min(errs[True]) > 0.1, errs
