"""Global FP64 diagnostics on the device (chy_diagnostics): ion weight,
momentum, kinetic energy and field energy against host sums of the
downloaded state, the gyration invariants of the push at E = 0 without
collisions (momentum along B kept bitwise, the perpendicular momentum turned
by 2 atan|t| per step at constant magnitude, kinetic energy kept), and the
per-step report of chy_step."""
import numpy as np
import pytest

from gpu_cases import context, opts

pytestmark = pytest.mark.gpu

EPS0 = 8.8541878128e-12


def _host_sums(ctx):
    d = ctx.get_particles()
    w, vx, vy, vz = d["w"], d["vx"], d["vy"], d["vz"]
    mi = ctx.prm.mi
    P = mi * np.array([np.sum(w * vx), np.sum(w * vy), np.sum(w * vz)])
    K = 0.5 * mi * np.sum(w * (vx * vx + vy * vy + vz * vz))
    phi = ctx.get_potential().reshape(ctx.prm.N[2], ctx.prm.N[1], ctx.prm.N[0])
    e2 = 0.0
    for ax, h in zip((2, 1, 0), ctx.prm.d):
        g = (np.roll(phi, -1, axis=ax) - np.roll(phi, 1, axis=ax)) * (0.5 / h)
        e2 += np.sum(g * g)
    W = 0.5 * EPS0 * e2 * ctx.prm.d[0] * ctx.prm.d[1] * ctx.prm.d[2]
    return np.sum(w), P, K, W


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_diagnostics_match_host_sums(prec):
    args = opts((16, 8, 12), extra=f"-seed 4 -ppc 6 -precision {prec} -ksp_rtol 1e-10 "
                                   "--perturbation uniform --perturbation-amplitude 0.01")
    with context(args) as ctx:
        ctx.init_particles()
        ctx.collect_moments(download=False)
        ctx.solve_potential()
        for _ in range(2):
            ctx.step()
        d = ctx.diagnostics()
        d2 = ctx.diagnostics()
        wt, P, K, W = _host_sums(ctx)
        assert d.valid == 1 and d.ions == ctx.num_particles()
    assert d.weight == pytest.approx(wt, rel=1e-13)
    assert np.abs(np.array(d.momentum) - P).max() <= 1e-12 * np.abs(P).max()
    assert d.kinetic_energy == pytest.approx(K, rel=1e-13)
    assert W > 0 and d.field_energy == pytest.approx(W, rel=1e-12)
    # fixed reduction grids: the same state gives the same bits
    assert d2.kinetic_energy == d.kinetic_energy and list(d2.momentum) == list(d.momentum)


def test_gyration_invariants_at_zero_field():
    """E0 = 0, no collisions, phi = 0: the push is a pure rotation about B
    (z).  v_z is kept exactly (s_x = s_y = 0), |P_perp| and K to rounding, and
    P_perp turns by 2 atan(|t|), t = (q dt / 2 m) B, per step."""
    args = opts((8, 8, 8), extra="-seed 9 -ppc 8 -mn 0 -E0x 0")
    with context(args) as ctx:
        ctx.init_particles()
        ctx.set_potential(np.zeros(ctx.nc))
        d0 = ctx.diagnostics()
        vz0 = ctx.get_particles()["vz"]
        nsteps = 5
        for _ in range(nsteps):
            ctx.update_velocities()
        d1 = ctx.diagnostics()
        vz1 = ctx.get_particles()["vz"]
        prm = ctx.prm
    assert np.array_equal(vz0, vz1)
    assert d1.momentum[2] == d0.momentum[2]
    assert d1.kinetic_energy == pytest.approx(d0.kinetic_energy, rel=1e-13)
    p0 = np.array(d0.momentum[:2])
    p1 = np.array(d1.momentum[:2])
    assert np.linalg.norm(p1) == pytest.approx(np.linalg.norm(p0), rel=1e-12)
    t = 0.5 * prm.qi / prm.mi * prm.dt * prm.B0[2]
    ang = np.arctan2(p0[0] * p1[1] - p0[1] * p1[0], p0 @ p1)
    assert abs(ang) == pytest.approx(nsteps * 2 * np.arctan(abs(t)), rel=1e-10)
    # a positive ion in B along +z gyrates clockwise seen from +z
    assert ang < 0


def test_step_reports_diagnostics_every_interval():
    args = opts((8, 8, 8), extra="-seed 2 -ppc 4 -ksp_rtol 1e-10 -diagnostics_interval 2")
    with context(args) as ctx:
        ctx.init_particles()
        ctx.collect_moments(download=False)
        ctx.solve_potential()
        s1 = ctx.step()
        s2 = ctx.step()
        d = ctx.diagnostics()
    assert s1.diag.valid == 0
    assert s2.diag.valid == 1 and s2.diag.step == 2
    assert s2.diag.kinetic_energy == d.kinetic_energy and s2.diag.field_energy == d.field_energy
    assert list(s2.diag.momentum) == list(d.momentum)
