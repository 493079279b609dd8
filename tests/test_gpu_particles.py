"""GPU parity of the particle path through the C ABI: stable cell sort,
deterministic CIC deposit, gather + Boris, collisions, the move, device
initialisation and whole steps, against the oracle on identical inputs."""
import numpy as np
import pytest

import chybrid
import orc
from gpu_cases import (as_particles, context, grid_of, opts, oracle_sorted, push_params_of, random_particles,
                       relmax, upload, vlasov_from)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,n,weights", [((8, 8, 8), 5000, False), ((6, 5, 7), 3000, True),
                                         ((32, 32, 32), 200000, True)])
def test_sort_and_deposit_fp64_bitwise(N, n, weights):
    rng = np.random.default_rng(n)
    with context(opts(N), capacity=n, weighted=weights) as ctx:
        g = grid_of(ctx.prm)
        p = random_particles(g, n, rng, weights=weights)
        upload(ctx, p)
        got = as_particles(ctx.get_particles())
        v = ctx.collect_moments()
    ps, offs = oracle_sorted(g, p)
    # stable counting sort: identical order (pid sequence) and payload
    assert np.array_equal(got.pid, ps.pid)
    for a in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(getattr(got, a), getattr(ps, a)), a
    ref = orc.deposit_sorted(g, ps, offs)
    for a in range(4):
        assert np.array_equal(v[:, a], ref[a]), (a, relmax(v[:, a], ref[a]))


@pytest.mark.parametrize("N,weights", [((32, 8, 6), False), ((64, 4, 4), False), ((24, 8, 6), False),
                                       ((32, 8, 6), True)])
def test_mixed_deposit_matches_oracle(N, weights):
    """Mixed-precision deposit (FP32 in-cell offsets and sums) against the
    FP64 oracle on the same ions.  Rows of 32 and 64 cells take the
    x-combined cell partials (each block's x pairs added in LDS, the corners
    across block edges and the periodic wrap by the node sum); a row of 24
    takes the per-corner partials.  Repeated deposits are bitwise equal."""
    n = 40 * N[0] * N[1] * N[2]
    rng = np.random.default_rng(N[0])
    with context(opts(N, extra="-precision mixed"), capacity=n, weighted=weights) as ctx:
        g = grid_of(ctx.prm)
        upload(ctx, random_particles(g, n, rng, weights=weights))
        got = as_particles(ctx.get_particles())
        v1 = ctx.collect_moments()
        v2 = ctx.collect_moments()
    assert np.array_equal(v1, v2)
    ps, offs = oracle_sorted(g, got)
    assert np.array_equal(ps.pid, got.pid)
    ref = orc.deposit_sorted(g, ps, offs)
    for a in range(4):
        assert relmax(v1[:, a], ref[a]) <= 1e-5, (a, relmax(v1[:, a], ref[a]))
    total = got.w.sum() if weights else n
    assert v1[:, 0].sum() == pytest.approx(total, rel=1e-5)


def test_mixed_step_deposit_equals_fresh_deposit():
    """The step deposits inside the neighbour-move sort (the gather variant of
    the x-combined partials); depositing the same sorted ions again (the
    plain variant) gives bitwise the same moments."""
    extra = "-seed 6 -ppc 16 -ksp_rtol 1e-8 --perturbation uniform --perturbation-amplitude 0.01 -precision mixed"
    with context(opts((32, 16, 8), extra=extra)) as ctx:
        ctx.init_particles()
        ctx.collect_moments(download=False)
        ctx.solve_potential()
        for _ in range(2):
            ctx.step()
        a = ctx.get_moments()
        b = ctx.collect_moments()
    assert a[:, 0].sum() > 0
    assert np.array_equal(a, b)


def test_deposit_edge_cases():
    """Empty set, one crowded cell, positions on the upper bound and outside."""
    with context(opts((4, 4, 4)), capacity=4000) as ctx:
        g = grid_of(ctx.prm)
        ctx.set_particles([], [], [], [], [], [])
        v = ctx.collect_moments()
        assert not v.any()
        rng = np.random.default_rng(1)
        n = 3000
        x = np.full(n, g.p0[0] + 0.3 * g.d[0]) + rng.random(n) * 1e-3 * g.d[0]
        y = np.full(n, g.p1[1])            # == upper bound: wraps to p0
        z = g.p0[2] - 0.25 * g.d[2] + np.zeros(n)  # below p0: wraps into the last cell
        vv = rng.standard_normal((3, n))
        p = orc.Particles(x, y, z, *vv)
        upload(ctx, p)
        got = as_particles(ctx.get_particles())
        v = ctx.collect_moments()
    assert np.all(got.y == g.p0[1]) and np.all((got.z >= g.p0[2]) & (got.z < g.p1[2]))
    q = orc.Particles(got.x, got.y, got.z, got.vx, got.vy, got.vz, pid=got.pid)
    ps, offs = oracle_sorted(g, q)
    assert np.array_equal(ps.pid, got.pid)
    ref = orc.deposit_sorted(g, ps, offs)
    assert np.array_equal(v[:, 0], ref[0])
    assert v[:, 0].sum() == pytest.approx(n, rel=1e-13)


def _phi(g, rng):
    k, j, i = np.meshgrid(*[np.arange(int(m)) for m in g.shape()], indexing="ij")
    return (0.02 * np.sin(2 * np.pi * i / g.N[0]) * np.cos(2 * np.pi * j / g.N[1])
            + 0.01 * np.cos(2 * np.pi * k / g.N[2]) + 1e-3 * rng.standard_normal(g.shape())).ravel()


def test_push_and_move_fp64_match_oracle_bitwise():
    rng = np.random.default_rng(11)
    n = 20000
    with context(opts((8, 8, 8), extra="-mn 0"), capacity=n) as ctx:
        g, P = push_params_of(ctx.prm, step=0)
        p = random_particles(g, n, rng)
        upload(ctx, p)
        start = as_particles(ctx.get_particles())
        phi = _phi(g, rng)
        ctx.set_potential(phi)
        st = ctx.update_velocities()
        mid = as_particles(ctx.get_particles())
        ctx.update_positions()
        end = as_particles(ctx.get_particles())
    assert st.accepted == 0 and st.Nc == 0
    ex, ey, ez = orc.efield(g, list(ctx.prm.E0), phi)
    ref = start.copy()
    key, _, _ = orc.push(P, ex, ey, ez, ref, 1.0)
    for a in ("vx", "vy", "vz"):
        assert np.array_equal(getattr(mid, a), getattr(ref, a)), (a, relmax(getattr(mid, a), getattr(ref, a)))
    perm, offs = orc.sort_perm(g.nc, key)
    refs = ref.permute(perm)
    assert np.array_equal(end.pid, refs.pid)
    for a in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(getattr(end, a), getattr(refs, a)), a


def test_collisions_match_oracle_per_particle():
    rng = np.random.default_rng(12)
    n = 50000
    with context(opts((8, 8, 8), extra="-nui 3e3 -seed 5"), capacity=n) as ctx:
        prm = ctx.prm
        g, P = push_params_of(prm, step=0)
        Pn = orc.push_params(g, prm.qi / prm.mi, prm.dt, list(prm.B0), list(prm.E0), collide=True, mi=prm.mi,
                             mn=prm.mn, viT=prm.viT_r, vnT=prm.vnT, vn0=list(prm.vn0),
                             Nc=orc.collision_count(n, prm.nui, prm.dt, prm.viT_r, list(prm.vi0), list(prm.vn0),
                                                    prm.vnT, prm.mi, prm.mn), seed=prm.seed, step=0)
        p = random_particles(g, n, rng)
        upload(ctx, p)
        start = as_particles(ctx.get_particles())
        st = ctx.update_velocities()
        mid = as_particles(ctx.get_particles())
    S = sum(orc.mean_rel_speed((start.vx[s], start.vy[s], start.vz[s]), list(prm.vn0), prm.vnT)
            for s in range(n))
    ex, ey, ez = orc.efield(g, list(prm.E0), np.zeros(g.nc))
    ref = start.copy()
    _, _, stats = orc.push(Pn, ex, ey, ez, ref, S)
    assert st.Nc == Pn.Nc and st.accepted == stats[0] and st.refused == stats[1]
    sig = np.sqrt(Pn.Nc)
    assert abs(st.accepted + st.refused - Pn.Nc) < 4 * sig
    for a in ("vx", "vy", "vz"):
        assert relmax(getattr(mid, a), getattr(ref, a)) <= 1e-12, a


def test_device_init_matches_oracle():
    for pert, name in ((1, "sinusoidal-x"), (2, "uniform"), (3, "gaussian")):
        args = opts((8, 4, 6), extra=f"-seed 3 -ppc 5 --perturbation {name} --perturbation-amplitude 0.01")
        with context(args) as ctx:
            ctx.init_particles()
            got = as_particles(ctx.get_particles())
            g = grid_of(ctx.prm)
            ref = orc.init_particles(g, 5, 3, list(ctx.prm.vi0), list(ctx.prm.viT), pert=pert, amp=0.01)
        ps, _ = oracle_sorted(g, ref)
        assert np.array_equal(got.pid, ps.pid)
        for a in ("x", "y", "z"):
            assert np.array_equal(getattr(got, a), getattr(ps, a)), a
        for a in ("vx", "vy", "vz", "w"):
            assert relmax(getattr(got, a), getattr(ps, a)) <= 1e-14, a


@pytest.mark.parametrize("kind,name", [(orc.DENSITY_FLAT_SOBOL, "flat-sobol"), (orc.DENSITY_UNIFORM, "uniform"),
                                       (orc.DENSITY_SINUSOIDAL, "sinusoidal")])
def test_reference_density_placements_match_oracle(kind, name):
    """--density-type flat-sobol / uniform / sinusoidal (particles.c:26-45,
    distributions.c) on the device against the oracle's placement: the cell
    order and positions bitwise (FP64), velocities and weights to 1e-14; the
    mixed state holds the same cells and offsets to FP32 rounding."""
    Np = 3001
    for pert, pname in ((0, ""), (2, " --perturbation uniform --perturbation-amplitude 0.01")):
        args = opts((8, 4, 6), extra=f"-seed 3 -Np {Np} --density-type {name}" + pname)
        with context(args) as ctx:
            ctx.init_particles()
            got = as_particles(ctx.get_particles())
            g = grid_of(ctx.prm)
            vi0, viT = list(ctx.prm.vi0), list(ctx.prm.viT)
        ref = orc.init_density(g, kind, Np, 3, vi0, viT, pert=pert, amp=0.01)
        assert got.np == ref.np == (Np if kind != orc.DENSITY_UNIFORM else (Np // g.nc) * 7 * 3 * 5)
        ps, _ = oracle_sorted(g, ref)
        assert np.array_equal(got.pid, ps.pid)
        for a in ("x", "y", "z"):
            assert np.array_equal(getattr(got, a), getattr(ps, a)), a
        for a in ("vx", "vy", "vz") + (("w",) if pert else ()):
            assert relmax(getattr(got, a), getattr(ps, a)) <= 1e-14, a
        with context(args + ["-precision", "mixed"]) as ctx:
            ctx.init_particles()
            m = as_particles(ctx.get_particles())
        assert np.array_equal(m.pid, ps.pid)
        for a in ("x", "y", "z"):
            assert np.abs(getattr(m, a) - getattr(ps, a)).max() <= 1e-7 * g.d[0], a


def test_unimplemented_density_types_are_refused():
    """particles.c:28-45: flat-normal, flat-reverse and gaussian raise
    PETSC_ERR_ARG_WRONG ("Not implemented")."""
    for name in ("flat-normal", "flat-reverse", "gaussian"):
        with context(opts((4, 4, 4), extra=f"-Np 64 --density-type {name}")) as ctx:
            with pytest.raises(chybrid.ChyError) as e:
                ctx.init_particles()
            assert e.value.code == chybrid.ERR_ARG_WRONG and "Not implemented" in str(e.value), name


@pytest.mark.parametrize("collide,N,more,rtol", [
    (False, (8, 8, 8), "", 1e-12), (True, (8, 8, 8), "", 1e-12),
    # ions crossing more than one cell per step, and 300 ions per cell (more
    # than 255 going one way): the sort falls back from the neighbour-move
    # placement to the general counting sort
    (False, (8, 8, 8), "-dt 4e-4", 1e-10), (False, (4, 4, 4), "-ppc 300", 1e-12)])
def test_steps_match_oracle(collide, N, more, rtol):
    """Three whole steps (push, sort, deposit, forcing, solve) from a seeded
    state; trajectories, moments and phi against the oracle step by step (the
    particle order is the oracle's stable cell order after every step)."""
    extra = f"-seed 2 -ppc 16 -ksp_rtol {rtol} --perturbation uniform --perturbation-amplitude 0.01 " + more
    if not collide:
        extra += " -mn 0"
    with context(opts(N, extra=extra)) as ctx:
        ctx.init_particles()
        ctx.collect_moments(download=False)
        phi, _ = ctx.solve_potential()
        cur = as_particles(ctx.get_particles())
        g = grid_of(ctx.prm)
        prm = ctx.prm
        S = sum(orc.mean_rel_speed((cur.vx[s], cur.vy[s], cur.vz[s]), list(prm.vn0), prm.vnT)
                for s in range(cur.np))
        for it in range(3):
            st = ctx.step()
            got = as_particles(ctx.get_particles())
            mom_gpu = ctx.collect_moments()
            phi_gpu = ctx.get_potential()
            _, P = push_params_of(prm, step=it, collide=collide)
            nxt, keys, mom, S, stats = orc.step(g, P, cur, phi, S, list(prm.E0))
            assert np.array_equal(got.pid, nxt.pid)
            for a in ("x", "y", "z", "vx", "vy", "vz"):
                assert relmax(getattr(got, a), getattr(nxt, a)) <= 1e-12, (it, a)
            if collide:
                assert st.collisions.accepted == stats[0]
            for a in range(4):
                assert relmax(mom_gpu[:, a], mom[a]) <= 1e-12, (it, a)
            b = orc.rhs(g, list(prm.kappa_e), list(prm.E0), 0.0, prm.me * prm.nue / orc.Q, *mom)
            assert st.solve.converged == 1 and st.solve.residual <= rtol
            r = b - orc.lhs_apply(g, list(prm.kappa_e), mom[0], phi_gpu)
            assert np.linalg.norm(r) <= 2 * rtol * np.linalg.norm(b), it
            phi_ref, _, conv, _ = orc.solve(g, list(prm.kappa_e), mom[0], b, phi0=phi, rtol=rtol)
            assert conv and relmax(phi_gpu, phi_ref) <= 1e4 * rtol, (it, relmax(phi_gpu, phi_ref))
            cur, phi = nxt, phi_gpu


def test_mixed_collisions_track_fp64_statistics():
    """Collisions on, both precisions from the same seeded state: every
    step's collision count within 3 sigma of the binomial expectation Nc, and
    the mixed run's density, drift and temperature within 1e-3 of the FP64
    run's (the FP32 collision kinematics use the hardware square root,
    reciprocal, sine and cosine)."""
    extra = ("-seed 6 -ppc 32 -nui 3e3 -ksp_rtol 1e-10 --perturbation uniform --perturbation-amplitude 0.01")
    res = {}
    for prec in ("fp64", "mixed"):
        with context(opts((16, 16, 16), extra=extra + f" -precision {prec}")) as ctx:
            ctx.init_particles()
            ctx.collect_moments(download=False)
            ctx.solve_potential()
            for it in range(4):
                c = ctx.step().collisions
                n = c.accepted + c.refused
                assert c.Nc > 100 and abs(n - c.Nc) <= 3 * np.sqrt(c.Nc), (prec, it, n, c.Nc)
            res[prec] = (as_particles(ctx.get_particles()), ctx.collect_moments())
    (a, ma), (b, mb) = res["fp64"], res["mixed"]

    def drift_temp(p):
        v = np.stack([p.vx, p.vy, p.vz])
        w = p.w if p.w is not None else np.ones(p.np)
        u = (v * w).sum(axis=1) / w.sum()
        return u, ((v - u[:, None]) ** 2 * w).sum(axis=1) / w.sum()

    ua, Ta = drift_temp(a)
    ub, Tb = drift_temp(b)
    assert relmax(mb[:, 0], ma[:, 0]) <= 1e-3
    assert np.abs(ub - ua).max() <= 1e-3 * np.linalg.norm(ua), (ua, ub)
    assert relmax(Tb, Ta) <= 1e-3, (Ta, Tb)


def test_mixed_step_tracks_fp64_step():
    """Same seeded initial state in both precisions: after three steps the
    mixed (int cell + FP32 offset) run agrees with FP64 to FP32 accuracy."""
    extra = "-seed 4 -ppc 16 -ksp_rtol 1e-10 --perturbation uniform --perturbation-amplitude 0.01 -mn 0"
    res = {}
    for prec in ("fp64", "mixed"):
        with context(opts((16, 16, 16), extra=extra + f" -precision {prec}")) as ctx:
            ctx.init_particles()
            ctx.collect_moments(download=False)
            ctx.solve_potential()
            for _ in range(3):
                st = ctx.step()
                assert st.solve.converged == 1, (prec, st.solve.iterations, st.solve.outer, st.solve.residual)
            res[prec] = (as_particles(ctx.get_particles()), ctx.collect_moments(), ctx.get_potential())
    a, b = res["fp64"], res["mixed"]
    assert np.array_equal(a[0].pid, b[0].pid) or np.mean(a[0].pid == b[0].pid) > 0.999
    assert relmax(b[1][:, 0], a[1][:, 0]) <= 1e-5
    assert relmax(b[2], a[2]) <= 1e-3


def test_initial_guess_modes_reach_the_same_potential():
    """-ksp_initial_guess previous / zero (the reference's ComputeInitialPhi) /
    extrapolate: the same steps give the same phi to the solve tolerance."""
    extra = "-seed 5 -ppc 8 -ksp_rtol 1e-12 --perturbation uniform --perturbation-amplitude 0.01 -mn 0"
    phis = {}
    for mode in ("previous", "zero", "extrapolate"):
        with context(opts((8, 8, 8), extra=extra + f" -ksp_initial_guess {mode}")) as ctx:
            ctx.init_particles()
            ctx.collect_moments(download=False)
            ctx.solve_potential()
            out = []
            for _ in range(3):
                st = ctx.step()
                assert st.solve.converged == 1, mode
                out.append(ctx.get_potential())
            phis[mode] = out
    for mode in ("zero", "extrapolate"):
        for a, b in zip(phis[mode], phis["previous"]):
            assert relmax(a, b) <= 1e-8, (mode, relmax(a, b))


def test_mixed_sort_from_push_directions_matches_key_sort():
    """The moving push writes each ion's move direction and the neighbour-move
    sort ranks by it (k_trans_dir); a separate move (update_positions) leaves
    the sort to the new keys (k_trans).  From the same state the arithmetic is
    the same, so one step gives bitwise equal orders, states and moments."""
    extra = ("-seed 6 -ppc 16 -ksp_rtol 1e-10 --perturbation uniform --perturbation-amplitude 0.01 -mn 0 "
             "-precision mixed")
    out = []
    for fused in (True, False):
        with context(opts((16, 16, 16), extra=extra)) as ctx:
            ctx.init_particles()
            ctx.collect_moments(download=False)
            ctx.solve_potential()
            if fused:
                ctx.step()
            else:
                ctx.update_velocities()
                ctx.update_positions()
            out.append((as_particles(ctx.get_particles()), ctx.collect_moments()))
    a, b = out
    assert np.array_equal(a[0].pid, b[0].pid)
    for f in ("x", "y", "z", "vx", "vy", "vz"):
        assert np.array_equal(getattr(a[0], f), getattr(b[0], f)), f
    assert np.array_equal(a[1], b[1])
