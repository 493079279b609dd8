"""GPU parity of the grid kernels through the C ABI: matrix-free operator and
its transpose (reference lhs.c), forcing (rhs.c), and the potential solve
(potential.c + KSP), against the oracle on identical inputs."""
import numpy as np
import pytest

import chybrid
import orc
from gpu_cases import context, grid_of, opts, relmax, vlasov_from

pytestmark = pytest.mark.gpu


def _moments(g, rng, amp=0.05):
    n = 16 * (1 + amp * rng.uniform(-1, 1, g.nc))
    G = [rng.standard_normal(g.nc) * 200 * n for _ in range(3)]
    return [n] + G


@pytest.mark.parametrize("N,extra", [
    ((8, 8, 8), ""),
    ((6, 5, 7), "-B0x 2e-6 -B0y -3e-6"),
    ((32, 32, 32), ""),
    ((16, 8, 12), "--lhs-type laplacian"),
    ((5, 4, 3), "--lhs-type identity"),
])
def test_lhs_apply_fp64_matches_oracle(N, extra):
    rng = np.random.default_rng(sum(N))
    with context(opts(N, extra=extra)) as ctx:
        g = grid_of(ctx.prm)
        mom = _moments(g, rng)
        ctx.set_moments(vlasov_from(mom))
        x = rng.standard_normal(g.nc)
        y = ctx.lhs_apply(x)
        yt = ctx.lhs_apply(x, transpose=True)
    K = list(ctx.prm.kappa_e)
    ref = orc.lhs_apply(g, K, mom[0], x, lhs_type=ctx.prm.lhs_type)
    reft = orc.lhs_apply(g, K, mom[0], x, lhs_type=ctx.prm.lhs_type, transpose=True)
    # same operation order, no contraction: FP64 results are bitwise equal
    assert np.array_equal(y, ref), relmax(y, ref)
    assert relmax(yt, reft) <= 1e-14


def test_lhs_apply_mixed_uses_fp32_density_fp64_arithmetic():
    rng = np.random.default_rng(3)
    with context(opts((8, 6, 10), extra="-precision mixed")) as ctx:
        g = grid_of(ctx.prm)
        mom = _moments(g, rng)
        ctx.set_moments(vlasov_from(mom))
        x = rng.standard_normal(g.nc)
        y = ctx.lhs_apply(x)
    n32 = mom[0].astype(np.float32).astype(np.float64)
    ref = orc.lhs_apply(g, list(ctx.prm.kappa_e), n32, x)
    assert relmax(y, ref) <= 1e-14


@pytest.mark.parametrize("N,extra", [((8, 8, 8), ""), ((6, 10, 4), "-Te 1500 -B0x 1e-5 -E0y 0.02 -E0z 0.01"),
                                     ((8, 8, 8), "--rhs-type sinusoidal"), ((8, 8, 8), "--rhs-type constant")])
def test_rhs_matches_oracle(N, extra):
    rng = np.random.default_rng(5)
    with context(opts(N, extra=extra)) as ctx:
        g = grid_of(ctx.prm)
        mom = _moments(g, rng)
        ctx.set_moments(vlasov_from(mom))
        b = ctx.compute_rhs()
    p = ctx.prm
    cth = p.gamma_e * orc.KB * p.Te / orc.Q
    cG = p.me * p.nue / orc.Q
    ref = orc.rhs(g, list(p.kappa_e), list(p.E0), cth, cG, *mom, rhs_type=p.rhs_type)
    assert relmax(b, ref) <= 1e-13
    if p.rhs_type != chybrid.RHS_CONSTANT:
        assert abs(b.mean()) <= 1e-14 * np.abs(b).max()


@pytest.mark.parametrize("extra", ["", "-ksp_type bcgs", "-B0x 1e-5 -B0y 2e-5", "--lhs-type laplacian"])
def test_solve_matches_oracle(extra):
    rng = np.random.default_rng(7)
    with context(opts((16, 16, 16), extra="-ksp_rtol 1e-12 " + extra)) as ctx:
        g = grid_of(ctx.prm)
        mom = _moments(g, rng, amp=0.1)
        ctx.set_moments(vlasov_from(mom))
        phi, st = ctx.solve_potential()
        b = ctx.get_rhs()
    assert st.converged == 1 and st.residual <= 1e-12
    # self-certifying: the oracle's operator confirms the GPU solution's residual
    K = list(ctx.prm.kappa_e)
    r = b - orc.lhs_apply(g, K, mom[0], phi, lhs_type=ctx.prm.lhs_type)
    assert np.linalg.norm(r) <= 2e-12 * np.linalg.norm(b)
    # the FP64 attainable residual of the 19-term stencil sum is ~3e-13, so
    # both solvers stop at 1e-12; solutions then differ by at most the
    # operator's conditioning times that
    ref, its, conv, res = orc.solve(g, K, mom[0], b, lhs_type=ctx.prm.lhs_type, rtol=1e-12, max_it=20000)
    assert conv
    assert relmax(phi, ref) <= 1e-8, relmax(phi, ref)
    assert abs(phi.mean()) <= 1e-13 * np.abs(phi).max()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_jacobi_preconditioned_solve(prec):
    """Point-Jacobi (the reference-style PC): slow on the anisotropic operator
    (about 2000 GMRES(30) iterations at 16^3) but converges to the tolerance."""
    rng = np.random.default_rng(17)
    with context(opts((16, 16, 16), extra=f"-ksp_rtol 1e-9 -pc_type jacobi -ksp_max_it 40000 -precision {prec}")) as ctx:
        g = grid_of(ctx.prm)
        mom = _moments(g, rng, amp=0.01)
        ctx.set_moments(vlasov_from(mom))
        phi, st = ctx.solve_potential()
        b = ctx.get_rhs()
    assert st.converged == 1 and st.residual <= 1e-9, (st.iterations, st.outer, st.residual)
    assert st.iterations > 200
    n = mom[0].astype(np.float32).astype(np.float64) if prec == "mixed" else mom[0]
    r = b - orc.lhs_apply(g, list(ctx.prm.kappa_e), n, phi)
    assert np.linalg.norm(r) <= 2e-9 * np.linalg.norm(b)


def test_mixed_solve_reaches_fp64_accuracy():
    """FP32 inner Krylov + FP64 defect correction: phi within 1e-10 of the FP64
    solver on the same density and forcing (the mixed-precision target)."""
    rng = np.random.default_rng(8)
    g_args = opts((32, 32, 32), extra="-ksp_rtol 1e-10")
    mom = None
    out = {}
    for prec in ("fp64", "mixed"):
        with context(g_args + ["-precision", prec]) as ctx:
            g = grid_of(ctx.prm)
            if mom is None:
                mom = _moments(g, rng, amp=0.1)
                # density representable in FP32 so both solve the same system
                mom = [m.astype(np.float32).astype(np.float64) for m in mom]
            ctx.set_moments(vlasov_from(mom))
            phi, st = ctx.solve_potential()
            assert st.converged == 1 and st.residual <= 1e-10, (prec, st.iterations, st.outer, st.residual)
            out[prec] = (phi, st)
    assert out["mixed"][1].outer >= 2
    assert relmax(out["mixed"][0], out["fp64"][0]) <= 1e-9


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_gmres_restarts_reach_the_same_potential(prec):
    """GMRES(2) against GMRES(30) on the same system (one-level FFT
    preconditioner, so that each outer sweep takes several cycles): the outer
    sweep writes the inner right-hand side straight into the first basis
    slot, and a restart has to keep it before that slot is rebuilt."""
    out = {}
    for restart in (30, 2):
        extra = f"-ksp_rtol 1e-10 -pc_type fft -ksp_gmres_restart {restart} -precision {prec}"
        with context(opts((32, 32, 32), extra=extra)) as ctx:
            g = grid_of(ctx.prm)
            mom = _moments(g, np.random.default_rng(9), amp=0.1)
            mom = [m.astype(np.float32).astype(np.float64) for m in mom]
            ctx.set_moments(vlasov_from(mom))
            phi, st = ctx.solve_potential()
            assert st.converged == 1 and st.residual <= 1e-10, (restart, st.iterations, st.outer, st.residual)
            out[restart] = (phi, st)
    assert out[2][1].iterations > 2 * out[2][1].outer  # cycles restarted within the outer sweeps
    assert relmax(out[2][0], out[30][0]) <= 1e-8


def test_solve_warm_start_and_fft_iteration_counts():
    rng = np.random.default_rng(9)
    with context(opts((32, 32, 32), extra="-ksp_rtol 1e-10")) as ctx:
        g = grid_of(ctx.prm)
        mom = _moments(g, rng, amp=0.01)
        ctx.set_moments(vlasov_from(mom))
        phi1, st1 = ctx.solve_potential()
        assert st1.converged and st1.iterations <= 15
        phi2, st2 = ctx.solve_potential()   # same system, warm start: converged at once
        assert st2.iterations == 0 and relmax(phi2, phi1) <= 1e-12


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("refine", ["REFINE_NEVER", "REFINE_IFNEEDED", "REFINE_ALWAYS"])
def test_gram_schmidt_refinement_types_solve_to_the_oracle(refine, prec):
    """-ksp_gmres_cgs_refinement_type: every variant of the classical
    Gram-Schmidt step converges to the oracle's solution of a structured-density
    system, in the same number of iterations to within two."""
    its, phis = {}, {}
    for r in ("REFINE_ALWAYS", refine):
        with context(opts((32, 32, 16), extra=f"-ksp_rtol 1e-10 -precision {prec} "
                                               f"-ksp_gmres_cgs_refinement_type {r}")) as ctx:
            g = grid_of(ctx.prm)
            mom = _moments(g, np.random.default_rng(5), amp=0.2)
            ctx.set_moments(vlasov_from(mom))
            phi, st = ctx.solve_potential()
            b = ctx.get_rhs()
            K = list(ctx.prm.kappa_e)
        assert st.converged == 1, r
        its[r], phis[r] = st.iterations, phi
    n = mom[0].astype(np.float32).astype(np.float64) if prec == "mixed" else mom[0]
    ref, _, conv, _ = orc.solve(g, K, n, b, rtol=1e-12)
    assert conv
    # residuals of 1e-10 of |b|: the solutions agree to the conditioning
    assert relmax(phis[refine], phis["REFINE_ALWAYS"]) <= 1e-7
    assert relmax(phi, ref) <= 1e-6, relmax(phi, ref)
    assert abs(its[refine] - its["REFINE_ALWAYS"]) <= 2, its


@pytest.mark.parametrize("N,extra", [((8, 16, 32), ""), ((32, 8, 128), "-B0x 1e-5 -B0y -2e-5"),
                                     ((4, 4, 4), ""), ((64, 32, 16), "--lhs-type laplacian"),
                                     ((16, 32, 8), "-precision mixed")])
def test_fft_preconditioner_exact_for_uniform_density(N, extra):
    """For uniform density the FFT preconditioner is A(n)^+ itself: one GMRES
    iteration solves the system (odd and even log2 axis lengths)."""
    rng = np.random.default_rng(21)
    with context(opts(N, extra="-ksp_rtol 1e-9 " + extra)) as ctx:
        g = grid_of(ctx.prm)
        n = np.full(g.nc, 16.0)
        G = [rng.standard_normal(g.nc) * 3000 for _ in range(3)]
        ctx.set_moments(vlasov_from([n] + G))
        phi, st = ctx.solve_potential()
        b = ctx.get_rhs()
    assert st.converged == 1
    assert st.iterations <= (2 if ctx.prm.precision == chybrid.PREC_FP64 else 6), st.iterations
    ref = orc.pc_apply(g, list(ctx.prm.kappa_e), n, b, lhs_type=ctx.prm.lhs_type)
    assert relmax(phi, ref) <= (1e-9 if ctx.prm.precision == chybrid.PREC_FP64 else 1e-8)


PCS = {"jacobi": orc.PC_JACOBI, "fft": orc.PC_FFT, "zline": orc.PC_ZLINE, "cheby": orc.PC_CHEBY,
       "fft2": orc.PC_FFT2, "fft2z": orc.PC_FFT2Z, "mg": orc.PC_MG}


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("pc,N", [("jacobi", (16, 8, 12)), ("fft", (16, 8, 32)), ("fft2", (16, 8, 32)),
                                  ("fft2", (32, 16, 8)), ("fft2z", (16, 8, 32)), ("fft2z", (32, 16, 8)),
                                  # z-line correction on long columns: 8 and 16 register segments
                                  ("fft2z", (16, 8, 256)), ("fft2z", (8, 8, 512)),
                                  # 512-point x and y lines: four-step kernels in FP32
                                  ("fft", (512, 4, 4)), ("fft", (32, 512, 4)), ("fft2z", (512, 8, 4)),
                                  ("zline", (16, 8, 12)),
                                  ("zline", (8, 6, 3)), ("zline", (6, 4, 5)), ("cheby", (16, 8, 12)),
                                  # multigrid: 2 levels (16x8, 8x4), 3 levels (24x20, 12x10, 6x5)
                                  ("mg", (16, 8, 12)), ("mg", (24, 20, 9)),
                                  # a fine level of 4608 columns of 40 planes: register segments,
                                  # the top one partly filled
                                  ("mg", (72, 64, 40))])
def test_pc_apply_matches_oracle(pc, N, prec):
    """chy_pc_apply (the PCShell apply) against the oracle's preconditioners,
    modulo the constant that A(n) annihilates."""
    rng = np.random.default_rng(len(pc) + N[2])
    with context(opts(N, extra=f"-pc_type {pc} -precision {prec} -B0x 1e-5 -B0y -2e-5")) as ctx:
        g = grid_of(ctx.prm)
        mom = _moments(g, rng, amp=0.1)
        ctx.set_moments(vlasov_from(mom))
        r = rng.standard_normal(g.nc)
        z = ctx.pc_apply(r)
        K = list(ctx.prm.kappa_e)
        deg = ctx.prm.pc_cheby_degree
    n = mom[0].astype(np.float32).astype(np.float64) if prec == "mixed" else mom[0]
    ref = orc.pc_apply(g, K, n, r, pc=PCS[pc], degree=deg)
    z = z - z.mean()
    # FP64: the symbol lambda(k) = sum_q v_q cos(k . o_q) cancels near k = 0,
    # where its relative rounding is about eps sum|v_q| / |lambda(k)|, growing
    # as N^2 for the lowest mode of an N-point axis; the device and the oracle
    # sum its terms in different orders.  On 512-point axes the two
    # preconditioners differ by 1.4e-12 (z), 1.1e-11 (y) and 4.4e-11 (x).
    tol = (1e-12 if max(N) <= 256 else 1e-10) if prec == "fp64" else (1e-4 if pc == "mg" else 2e-5)
    assert relmax(z, ref) <= tol, relmax(z, ref)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("N", [(16, 8, 32), (256, 4, 16), (8, 8, 256), (32, 16, 4)])
def test_fft2z_pc_apply_b_along_z(N, prec):
    """fft2z with B along z, where the column-sum pass keeps r - A z and the
    z-line kernel subtracts A e2 and adds the 2-D correction itself, against
    the oracle's fft2z (which forms the residual of z + e2 in one pass):
    64-wide stencil tiles, whole-row FP32 tiles on a 256-wide grid, 256-plane
    columns, and the shortest columns the z-line solve takes."""
    rng = np.random.default_rng(N[0] + N[2])
    with context(opts(N, extra=f"-pc_type fft2z -precision {prec}")) as ctx:
        g = grid_of(ctx.prm)
        mom = _moments(g, rng, amp=0.1)
        ctx.set_moments(vlasov_from(mom))
        r = rng.standard_normal(g.nc)
        z = ctx.pc_apply(r)
        K = list(ctx.prm.kappa_e)
    assert K[0] == 0.0 and K[1] == 0.0
    n = mom[0].astype(np.float32).astype(np.float64) if prec == "mixed" else mom[0]
    ref = orc.pc_apply(g, K, n, r, pc=orc.PC_FFT2Z)
    z = z - z.mean()
    # FP64: the symbol's cancellation near k = 0 (see test_pc_apply_matches_oracle)
    # is largest with B along z, where the x-y coupling is weakest against the
    # terms that cancel, and grows with the x-y axes.  It comes from the FFT
    # stage, not the z-line one: on 32 x 16 x 4 the FFT preconditioner alone
    # differs from the oracle's by 3.4e-11, fft2 (no z-line stage) by 9.8e-12
    # and fft2z by 8.9e-12; on 256 x 4 x 16 by 2.2e-10, 6.0e-11 and 5.9e-11;
    # on 16 x 8 x 32 by 6.2e-13, 4.7e-13 and 4.1e-13.
    tol = (1e-12 if max(N[0], N[1]) < 32 else 1e-10) if prec == "fp64" else 2e-5
    assert relmax(z, ref) <= tol, relmax(z, ref)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_mg_pc_apply_b_along_z(prec):
    """The multigrid preconditioner with B along z (the B-along-z stencil on
    every level) on a non-power-of-two grid: 4 levels, 48x40 -> 24x20 ->
    12x10 -> 6x5."""
    rng = np.random.default_rng(31)
    with context(opts((48, 40, 10), extra=f"-pc_type mg -precision {prec}")) as ctx:
        g = grid_of(ctx.prm)
        mom = _moments(g, rng, amp=0.1)
        ctx.set_moments(vlasov_from(mom))
        r = rng.standard_normal(g.nc)
        z = ctx.pc_apply(r)
        K = list(ctx.prm.kappa_e)
    n = mom[0].astype(np.float32).astype(np.float64) if prec == "mixed" else mom[0]
    ref = orc.pc_apply(g, K, n, r, pc=orc.PC_MG)
    z = z - z.mean()
    # FP64 within 1e-10 (measured 8.5e-12): the four levels' z-line solves
    # (condition ~1e4 per line) amplify last-bit differences of the restriction
    # and residual sums, which the two-level cases above keep under 1e-12
    assert relmax(z, ref) <= (1e-10 if prec == "fp64" else 1e-4), relmax(z, ref)


def test_mg_iterations_do_not_grow_with_the_grid():
    """-pc_type mg (x-y semicoarsening multigrid, z-line smoother) at 10 %
    node-wise density variation: the iterations to 1e-10 stay the same from
    48^3 to 96^3 (the z-line preconditioner alone needs 233 at 48^3 and
    grows with the grid), as the oracle's restatement of the same cycle
    does, and the potential matches the oracle's solve at 48^3."""
    its = {}
    for N in (48, 96):
        rng = np.random.default_rng(N)
        with context(opts((N, N, N), extra="-pc_type mg -ksp_rtol 1e-10")) as ctx:
            g = grid_of(ctx.prm)
            n = 16 * (1 + 0.1 * rng.uniform(-1, 1, g.nc))
            z = np.zeros(g.nc)
            ctx.set_moments(vlasov_from([n, z, z, z]))
            b = rng.standard_normal(g.nc)
            b -= b.mean()
            phi, st = ctx.solve(b)
            K = list(ctx.prm.kappa_e)
        assert st.converged == 1 and st.residual <= 1e-10
        its[N] = st.iterations
        print(f"mg: {N}^3, {st.iterations} iterations, residual {st.residual:.2e}", flush=True)
        if N == 48:
            ref, oits, conv, _ = orc.solve(g, K, n, b, pc=orc.PC_MG, rtol=1e-10, max_it=200)
            assert conv and abs(st.iterations - oits) <= 2, (st.iterations, oits)
            assert relmax(phi, ref) <= 1e-5
    assert its[48] <= 12 and its[96] <= 12 and abs(its[96] - its[48]) <= 2, its


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("N", [(256, 256, 8), (8, 256, 256), (64, 256, 4), (32, 16, 256), (256, 8, 4)])
def test_fft_pc_with_256_point_axes(N, prec):
    """256-point axes (four-step transforms), alone and beside the radix-4
    passes, against the oracle's FFT preconditioner.  At these lengths the
    rounding of any FFT, amplified by the small eigenvalues of the anisotropic
    symbol, moves z by up to ~1e-10 of its size in FP64 (the radix-4 kernels
    at 128 points measure 1e-10 on the same kind of case), hence the
    tolerances."""
    rng = np.random.default_rng(N[0] + 3 * N[2])
    with context(opts(N, extra=f"-pc_type fft -precision {prec} -B0x 1e-5 -B0y -2e-5")) as ctx:
        g = grid_of(ctx.prm)
        mom = _moments(g, rng, amp=0.1)
        ctx.set_moments(vlasov_from(mom))
        r = rng.standard_normal(g.nc)
        z = ctx.pc_apply(r)
        K = list(ctx.prm.kappa_e)
    n = mom[0].astype(np.float32).astype(np.float64) if prec == "mixed" else mom[0]
    ref = orc.pc_apply(g, K, n, r, pc=orc.PC_FFT)
    assert relmax(z - z.mean(), ref) <= (1e-8 if prec == "fp64" else 5e-6), relmax(z - z.mean(), ref)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_fft2_pc_with_256_point_axes(prec):
    """The two-level preconditioner on a grid whose x and y passes take the
    four-step kernels, against the oracle."""
    rng = np.random.default_rng(41)
    N = (256, 256, 4)
    with context(opts(N, extra=f"-pc_type fft2 -precision {prec} -B0x 1e-5 -B0y -2e-5")) as ctx:
        g = grid_of(ctx.prm)
        mom = _moments(g, rng, amp=0.1)
        ctx.set_moments(vlasov_from(mom))
        r = rng.standard_normal(g.nc)
        z = ctx.pc_apply(r)
        K = list(ctx.prm.kappa_e)
    n = mom[0].astype(np.float32).astype(np.float64) if prec == "mixed" else mom[0]
    ref = orc.pc_apply(g, K, n, r, pc=orc.PC_FFT2)
    assert relmax(z - z.mean(), ref) <= (1e-8 if prec == "fp64" else 2e-5), relmax(z - z.mean(), ref)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("pc", ["zline", "cheby"])
def test_zline_and_chebyshev_preconditioned_solves(pc, prec):
    """GMRES with the z-line or Chebyshev-Jacobi preconditioner converges to
    the tolerance on the anisotropic operator, in about the oracle's number of
    iterations in FP64."""
    rng = np.random.default_rng(23)
    with context(opts((16, 16, 16), extra=f"-ksp_rtol 1e-9 -pc_type {pc} -ksp_max_it 20000 -precision {prec}")) as ctx:
        g = grid_of(ctx.prm)
        mom = _moments(g, rng, amp=0.01)
        ctx.set_moments(vlasov_from(mom))
        phi, st = ctx.solve_potential()
        b = ctx.get_rhs()
        K = list(ctx.prm.kappa_e)
    assert st.converged == 1 and st.residual <= 1e-9, (st.iterations, st.outer, st.residual)
    n = mom[0].astype(np.float32).astype(np.float64) if prec == "mixed" else mom[0]
    r = b - orc.lhs_apply(g, K, n, phi)
    assert np.linalg.norm(r) <= 2e-9 * np.linalg.norm(b)
    _, its, conv, _ = orc.solve(g, K, n, b, pc=PCS[pc], rtol=1e-9, max_it=20000)
    assert conv
    if prec == "fp64":
        assert st.iterations <= 1.2 * its + 5, (st.iterations, its)


def test_non_power_of_two_grid_falls_back_to_zline():
    rng = np.random.default_rng(29)
    with context(opts((12, 10, 9), extra="-ksp_rtol 1e-10")) as ctx:
        g = grid_of(ctx.prm)
        mom = _moments(g, rng, amp=0.05)
        ctx.set_moments(vlasov_from(mom))
        phi, st = ctx.solve_potential()
        b = ctx.get_rhs()
        K = list(ctx.prm.kappa_e)
    assert st.converged == 1
    r = b - orc.lhs_apply(g, K, mom[0], phi)
    assert np.linalg.norm(r) <= 2e-10 * np.linalg.norm(b)
    _, its, _, _ = orc.solve(g, K, mom[0], b, pc=orc.PC_ZLINE, rtol=1e-10, max_it=20000)
    assert st.iterations <= 1.2 * its + 5


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_two_level_fft_preconditioner_for_structured_density(prec):
    """Density with x-y structure (where the Hall term dominates the x-y
    coupling): the default two-level preconditioner converges to the
    tolerance in far fewer iterations than the mean-density FFT alone, and in
    about the oracle's count."""
    rng = np.random.default_rng(41)
    its = {}
    for pc in ("fft", "fft2"):
        with context(opts((32, 32, 16), extra=f"-ksp_rtol 1e-10 -pc_type {pc} -precision {prec}")) as ctx:
            g = grid_of(ctx.prm)
            k, j, i = np.meshgrid(*[np.arange(int(m)) for m in g.shape()], indexing="ij")
            n = 16 * (1 + 0.3 * np.sin(2 * np.pi * i / 32) * np.cos(4 * np.pi * j / 32)).ravel()
            n *= 1 + 0.05 * rng.uniform(-1, 1, g.nc)
            n = n.astype(np.float32).astype(np.float64)
            G = [rng.standard_normal(g.nc) * 200 * n for _ in range(3)]
            ctx.set_moments(vlasov_from([n] + G))
            phi, st = ctx.solve_potential()
            b = ctx.get_rhs()
            K = list(ctx.prm.kappa_e)
        assert st.converged == 1 and st.residual <= 1e-10, (pc, st.iterations, st.residual)
        r = b - orc.lhs_apply(g, K, n, phi)
        assert np.linalg.norm(r) <= 2e-10 * np.linalg.norm(b)
        its[pc] = st.iterations
        if prec == "fp64":
            _, oits, conv, _ = orc.solve(g, K, n, b, pc=PCS[pc], rtol=1e-10, max_it=2000)
            assert conv and st.iterations <= 1.2 * oits + 3, (pc, st.iterations, oits)
    assert its["fft2"] < 0.6 * its["fft"], its


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_fft2z_preconditioner_solves_in_fewer_iterations(prec):
    """Density varying in all three directions: the z-line correction of the
    two-level preconditioner converges to the tolerance in fewer iterations
    than fft2, and in about the oracle's count (FP64)."""
    rng = np.random.default_rng(43)
    its = {}
    for pc in ("fft2", "fft2z"):
        with context(opts((32, 32, 16), extra=f"-ksp_rtol 1e-10 -pc_type {pc} -precision {prec}")) as ctx:
            g = grid_of(ctx.prm)
            n = 16 * (1 + 0.25 * np.random.default_rng(44).uniform(-1, 1, g.nc))
            n = n.astype(np.float32).astype(np.float64)
            G = [np.random.default_rng(45 + a).standard_normal(g.nc) * 200 * n for a in range(3)]
            ctx.set_moments(vlasov_from([n] + G))
            phi, st = ctx.solve_potential()
            b = ctx.get_rhs()
            K = list(ctx.prm.kappa_e)
        assert st.converged == 1 and st.residual <= 1e-10, (pc, st.iterations, st.residual)
        r = b - orc.lhs_apply(g, K, n, phi)
        assert np.linalg.norm(r) <= 2e-10 * np.linalg.norm(b)
        its[pc] = st.iterations
        if prec == "fp64":
            _, oits, conv, _ = orc.solve(g, K, n, b, pc=PCS[pc], rtol=1e-10, max_it=2000)
            assert conv and st.iterations <= 1.2 * oits + 3, (pc, st.iterations, oits)
    assert its["fft2z"] < its["fft2"], its
