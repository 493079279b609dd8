"""Pins of the oracle potential solve (reference potential.c:6-55 + PETSc KSP).

The FFT transform is checked against numpy, the FFT preconditioner is the
exact inverse for uniform density, and Krylov solutions match a dense
pseudo-inverse on small grids.
"""
import numpy as np
import pytest

import orc

KZ = orc.Q * 5e-5 / orc.ME / 1e5


def test_fft3d_matches_numpy():
    rng = np.random.default_rng(1)
    for N in ((8, 4, 16), (6, 5, 3)):
        g = orc.Grid(N)
        a = rng.standard_normal(g.shape()) + 1j * rng.standard_normal(g.shape())
        assert np.allclose(orc.fft3d(g, a), np.fft.fftn(a), atol=1e-11)
        assert np.allclose(orc.fft3d(g, a, inverse=True), np.fft.ifftn(a) * g.nc, atol=1e-11)


@pytest.mark.parametrize("K", [(0, 0, KZ), (2.0, -1.0, 30.0)])
def test_fft_preconditioner_is_exact_inverse_for_uniform_density(K):
    g = orc.Grid((8, 8, 16), p1=(0.4, 0.4, 0.8))
    rng = np.random.default_rng(2)
    n = np.full(g.nc, 16.0)
    r = rng.standard_normal(g.nc)
    r -= r.mean()
    z = orc.pc_apply(g, K, n, r)
    assert np.abs(orc.lhs_apply(g, K, n, z) - r).max() <= 1e-11 * np.abs(r).max()
    assert abs(z.mean()) < 1e-12 * np.abs(z).max()


@pytest.mark.parametrize("ksp,pc", [(orc.KSP_GMRES, orc.PC_FFT), (orc.KSP_GMRES, orc.PC_JACOBI),
                                    (orc.KSP_BCGS, orc.PC_FFT), (orc.KSP_GMRES, orc.PC_NONE)])
def test_solution_matches_dense_pseudo_inverse(ksp, pc):
    g = orc.Grid((4, 5, 4), p1=(0.2, 0.25, 0.2))
    K = (0.5, -0.3, 6.0)
    rng = np.random.default_rng(3)
    n = 16 * (1 + 0.05 * rng.uniform(-1, 1, g.nc))
    A = orc.lhs_dense(g, K, n)
    b = rng.standard_normal(g.nc)
    b -= b.mean()
    ref = np.linalg.pinv(A) @ b
    ref -= ref.mean()
    phi, its, conv, res = orc.solve(g, K, n, b, ksp=ksp, pc=pc, rtol=1e-13, max_it=5000)
    assert conv and res <= 1e-13
    assert np.abs(phi - ref).max() <= 1e-9 * np.abs(ref).max()


def test_fft_preconditioned_gmres_iteration_counts():
    """The FFT(nbar) preconditioner keeps iteration counts small and flat in N
    (a few iterations at 1 % density perturbation, more at 10 %)."""
    K = (0, 0, KZ)
    its = {}
    for M, amp in ((16, 0.01), (32, 0.01), (16, 0.10)):
        g = orc.Grid((M, M, M), p1=(0.05 * M,) * 3)
        rng = np.random.default_rng(M)
        n = 16 * (1 + rng.uniform(-amp, amp, g.nc))
        b = rng.standard_normal(g.nc)
        b -= b.mean()
        phi, it, conv, res = orc.solve(g, K, n, b, rtol=1e-10)
        assert conv and res <= 1e-10
        its[(M, amp)] = it
    assert its[(16, 0.01)] <= 12 and its[(32, 0.01)] <= 12
    assert its[(16, 0.10)] <= 40
