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


def _line_matrix(g, K, n, col, lhs_type=orc.LHS_FULL):
    """Dense z-line block of A(n) for column col: diagonal and the +-z entries."""
    m = int(g.N[2])
    plane = int(g.N[0] * g.N[1])
    M = np.zeros((m, m))
    for k in range(m):
        v = orc.lhs_row(g, K, n, col + plane * k) if lhs_type == orc.LHS_FULL else None
        M[k, k] += v[18]
        M[k, (k + 1) % m] += v[4]
        M[k, (k - 1) % m] += v[5]
    return M


@pytest.mark.parametrize("N,K", [((4, 3, 6), (0.5, -0.3, 6.0)), ((3, 4, 2), (0.0, 0.0, KZ)),
                                 ((5, 2, 7), (1.0, 2.0, 3.0))])
def test_zline_preconditioner_solves_each_z_line_exactly(N, K):
    """Block Jacobi over z-lines: each (i, j) column's cyclic tridiagonal
    system (a dense solve here) reproduces the oracle's Sherman-Morrison solve."""
    g = orc.Grid(N, p1=tuple(0.05 * m for m in N))
    rng = np.random.default_rng(sum(N))
    n = 16 * (1 + 0.2 * rng.uniform(-1, 1, g.nc))
    r = rng.standard_normal(g.nc)
    z = orc.pc_apply(g, K, n, r, pc=orc.PC_ZLINE)
    plane = int(g.N[0] * g.N[1])
    ref = np.zeros(g.nc)
    for col in range(plane):
        ref[col::plane] = np.linalg.solve(_line_matrix(g, K, n, col), r[col::plane])
    ref -= ref.mean()
    assert np.abs(z - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("degree", [1, 2, 4, 6])
def test_chebyshev_preconditioner_is_the_chebyshev_polynomial(degree):
    """degree steps of Chebyshev iteration from zero on D^-1 A z = D^-1 r give
    z = q(X) D^-1 r with X = D^-1 A and q(x) = (1 - T_m((th - x)/de) / T_m(th/de)) / x
    on the interval [0.2, 2.2] (th = 1.2, de = 1)."""
    from numpy.polynomial import Chebyshev, Polynomial
    g = orc.Grid((4, 5, 3), p1=(0.2, 0.25, 0.15))
    K = (0.5, -0.3, 6.0)
    rng = np.random.default_rng(degree)
    n = 16 * (1 + 0.1 * rng.uniform(-1, 1, g.nc))
    r = rng.standard_normal(g.nc)
    A = orc.lhs_dense(g, K, n)
    D = np.diag(A).copy()
    X = A / D[:, None]
    th, de = 1.2, 1.0
    Tm = Chebyshev.basis(degree).convert(kind=Polynomial)
    y = Polynomial([th / de, -1.0 / de])
    num = 1 - Tm(y) / Tm(th / de)
    q, rem = divmod(num, Polynomial([0.0, 1.0]))
    assert np.abs(rem.coef).max() < 1e-12
    v = r / D
    acc = np.zeros(g.nc)
    for c in q.coef[::-1]:      # Horner on the matrix
        acc = X @ acc + c * v
    acc -= acc.mean()
    z = orc.pc_apply(g, K, n, r, pc=orc.PC_CHEBY, degree=degree)
    assert np.abs(z - acc).max() <= 1e-11 * np.abs(acc).max()


def test_zline_and_chebyshev_preconditioned_gmres_converge():
    """Both converge on the 7735:1 anisotropic operator; the z-line blocks
    capture the stiff direction and need far fewer iterations than Jacobi."""
    g = orc.Grid((16, 16, 16), p1=(0.8, 0.8, 0.8))
    K = (0, 0, KZ)
    rng = np.random.default_rng(16)
    n = 16 * (1 + rng.uniform(-0.01, 0.01, g.nc))
    b = rng.standard_normal(g.nc)
    b -= b.mean()
    its = {}
    for pc in (orc.PC_ZLINE, orc.PC_CHEBY, orc.PC_JACOBI):
        phi, it, conv, res = orc.solve(g, K, n, b, pc=pc, rtol=1e-8, max_it=20000)
        assert conv and res <= 1e-8, pc
        its[pc] = it
    assert its[orc.PC_ZLINE] < its[orc.PC_JACOBI] / 5
    assert its[orc.PC_CHEBY] < its[orc.PC_JACOBI]


def test_fft2_equals_fft_for_uniform_density():
    """With uniform density the FFT preconditioner is exact, so the kz = 0
    correction of the two-level preconditioner vanishes."""
    g = orc.Grid((8, 6, 4), p1=(0.4, 0.3, 0.2))
    K = (0.5, -0.3, 6.0)
    rng = np.random.default_rng(31)
    n = np.full(g.nc, 16.0)
    r = rng.standard_normal(g.nc)
    z1 = orc.pc_apply(g, K, n, r, pc=orc.PC_FFT)
    z2 = orc.pc_apply(g, K, n, r, pc=orc.PC_FFT2)
    # the correction solves for the rounding residual of the exact FFT inverse
    assert np.abs(z2 - z1).max() <= 1e-10 * np.abs(z1).max()


def test_fft2_correction_is_gmres_on_the_z_averaged_operator():
    """The kz = 0 correction equals FFT2_ITS steps of flexible GMRES on the
    z-averaged system, with the 2-D operator and preconditioner built densely
    from the oracle's 3-D operator and FFT preconditioner applied to
    z-independent unit vectors (least squares by numpy)."""
    g = orc.Grid((5, 4, 6), p1=(0.25, 0.2, 0.3))
    K = (0.4, -0.7, 9.0)
    rng = np.random.default_rng(32)
    n = 16 * (1 + 0.3 * rng.uniform(-1, 1, g.nc))
    r = rng.standard_normal(g.nc)
    r -= r.mean()
    nx, ny, nz = (int(v) for v in g.N)
    nxy = nx * ny
    ext = lambda e2: np.tile(e2, nz)
    zavg = lambda f: f.reshape(nz, nxy).mean(axis=0)
    A2 = np.column_stack([zavg(orc.lhs_apply(g, K, n, ext(np.eye(nxy)[c]))) for c in range(nxy)])
    nbar = np.full(g.nc, n.mean())
    F2 = np.column_stack([zavg(orc.pc_apply(g, K, nbar, ext(np.eye(nxy)[c]))) for c in range(nxy)])
    z = orc.pc_apply(g, K, nbar, r, pc=orc.PC_FFT)
    r2 = zavg(r - orc.lhs_apply(g, K, n, z))
    r2 -= r2.mean()
    m = orc.FFT2_ITS
    beta = np.linalg.norm(r2)
    V = [r2 / beta]
    Z = []
    H = np.zeros((m + 1, m))
    for j in range(m):
        Z.append(F2 @ V[j])
        w = A2 @ Z[j]
        for i in range(j + 1):
            H[i, j] = V[i] @ w
        w = w - sum(H[i, j] * V[i] for i in range(j + 1))
        H[j + 1, j] = np.linalg.norm(w)
        V.append(w / H[j + 1, j])
    rhs = np.zeros(m + 1)
    rhs[0] = beta
    y = np.linalg.lstsq(H, rhs, rcond=None)[0]
    ref = z + ext(sum(yi * zi for yi, zi in zip(y, Z)))
    ref -= ref.mean()
    got = orc.pc_apply(g, K, n, r, pc=orc.PC_FFT2)
    assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max()


def test_fft2_preconditioner_cuts_iterations_for_structured_density():
    """Density with structure along x-y (constant along z), where the Hall term
    dominates the x-y coupling: the two-level preconditioner needs far fewer
    GMRES iterations than the mean-density FFT alone."""
    g = orc.Grid((32, 32, 16), p1=(1.6, 1.6, 0.8))
    K = (0, 0, KZ)
    rng = np.random.default_rng(33)
    k, j, i = np.meshgrid(*[np.arange(int(m)) for m in g.shape()], indexing="ij")
    n = 16 * (1 + 0.3 * np.sin(2 * np.pi * i / 32) * np.cos(4 * np.pi * j / 32)).ravel()
    n *= 1 + 0.05 * rng.uniform(-1, 1, g.nc)
    b = rng.standard_normal(g.nc)
    b -= b.mean()
    its = {}
    for pc in (orc.PC_FFT, orc.PC_FFT2):
        phi, it, conv, res = orc.solve(g, K, n, b, pc=pc, rtol=1e-10, max_it=2000)
        assert conv and res <= 1e-10
        its[pc] = it
    assert its[orc.PC_FFT2] < 0.6 * its[orc.PC_FFT], its


def test_fft2z_equals_fft_for_uniform_density():
    """Uniform density: fft is exact, so both corrections of fft2z vanish."""
    g = orc.Grid((8, 6, 4), p1=(0.4, 0.3, 0.2))
    K = (0.5, -0.3, 6.0)
    rng = np.random.default_rng(35)
    n = np.full(g.nc, 16.0)
    r = rng.standard_normal(g.nc)
    z1 = orc.pc_apply(g, K, n, r, pc=orc.PC_FFT)
    z2 = orc.pc_apply(g, K, n, r, pc=orc.PC_FFT2Z)
    assert np.abs(z2 - z1).max() <= 1e-10 * np.abs(z1).max()


def test_fft2z_adds_the_z_line_solve_of_the_fft2_residual():
    """fft2z = fft2 + the z-line block solve of fft2's residual, with the
    column systems solved densely from the operator's rows here."""
    g = orc.Grid((5, 4, 6), p1=(0.25, 0.2, 0.3))
    K = (0.4, -0.7, 9.0)
    rng = np.random.default_rng(36)
    n = 16 * (1 + 0.3 * rng.uniform(-1, 1, g.nc))
    r = rng.standard_normal(g.nc)
    z1 = orc.pc_apply(g, K, n, r, pc=orc.PC_FFT2)
    res = r - orc.lhs_apply(g, K, n, z1)
    plane = int(g.N[0] * g.N[1])
    y = np.zeros(g.nc)
    for col in range(plane):
        y[col::plane] = np.linalg.solve(_line_matrix(g, K, n, col), res[col::plane])
    ref = z1 + y
    ref -= ref.mean()
    got = orc.pc_apply(g, K, n, r, pc=orc.PC_FFT2Z)
    assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max()


def test_fft2z_cuts_iterations_below_fft2():
    """Density varying in all three directions with the workload's
    anisotropy: adding the z-line correction to the two-level preconditioner
    cuts the GMRES iterations to 1e-10."""
    g = orc.Grid((32, 32, 16), p1=(1.6, 1.6, 0.8))
    K = (17.6, 35.2, KZ)
    rng = np.random.default_rng(37)
    n = 16 * (1 + 0.25 * rng.uniform(-1, 1, g.nc))
    b = rng.standard_normal(g.nc)
    b -= b.mean()
    its = {}
    for pc in (orc.PC_FFT2, orc.PC_FFT2Z):
        phi, it, conv, res = orc.solve(g, K, n, b, pc=pc, rtol=1e-10, max_it=2000)
        assert conv and res <= 1e-10
        its[pc] = it
    assert its[orc.PC_FFT2Z] < 0.75 * its[orc.PC_FFT2], its
