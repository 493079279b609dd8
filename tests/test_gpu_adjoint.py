"""Adjoint of the step map through the C ABI (chy_adjoint_forward/backward):
gradients of the final wave energy J = sum (n - mean n)^2 with respect to E0
and the per-cell initial weight perturbation, against central finite
differences of the oracle's forward map run with the same random streams
(collision events frozen), plus a closed form and checkpoint invariance."""
import numpy as np
import pytest

import chybrid
import orc
from gpu_cases import as_particles, context, grid_of, opts, push_params_of, relmax

pytestmark = pytest.mark.gpu

BASE = "-seed 3 -ppc 8 --perturbation gaussian --perturbation-amplitude 0.01 -ksp_rtol 1e-12"


def _oracle_J(prm, parts, E0, nsteps, collide):
    """The forward map on the oracle: moments and a zero-guess solve of the
    given (cell-sorted) ions, then nsteps steps; J of the final density."""
    g = grid_of(prm)
    K = list(prm.kappa_e)
    cth = prm.gamma_e * orc.KB * prm.Te / orc.Q
    cG = prm.me * prm.nue / orc.Q
    _, offs = orc.sort_perm(g.nc, orc.cell_keys(g, parts))
    mom = orc.deposit_sorted(g, parts, offs)
    b = orc.rhs(g, K, E0, cth, cG, *mom)
    phi, _, conv, _ = orc.solve(g, K, mom[0], b, rtol=1e-12)
    assert conv
    S = sum(orc.mean_rel_speed((parts.vx[s], parts.vy[s], parts.vz[s]), list(prm.vn0), prm.vnT)
            for s in range(parts.np))
    cur = parts
    for it in range(nsteps):
        _, P = push_params_of(prm, step=it, collide=collide)
        P.E0[0], P.E0[1], P.E0[2] = E0
        cur, _, mom, S, _ = orc.step(g, P, cur, phi, S, E0)
        b = orc.rhs(g, K, E0, cth, cG, *mom)
        phi, _, conv, _ = orc.solve(g, K, mom[0], b, phi0=phi, rtol=1e-12)
        assert conv
    n = mom[0]
    return float(np.sum((n - n.mean()) ** 2))


def _with_weights(p, w):
    q = p.copy()
    q.w = w
    return q


@pytest.mark.parametrize("collide,nui", [(False, 1e3), (True, 1e3), (True, 1e4)])
def test_adjoint_matches_oracle_finite_differences(collide, nui):
    """nu_in = 1e4 s^-1: about a quarter of the ions scatter each step, so the
    scattering's transposed Jacobian carries much of the gradient."""
    nsteps, ckpt = 4, 2
    extra = BASE + f" -nui {nui:g}" + ("" if collide else " -mn 0")
    with context(opts((8, 8, 8), extra=extra)) as ctx:
        ctx.init_particles()
        p0 = as_particles(ctx.get_particles())
        prm = ctx.prm
        J = ctx.adjoint_forward(nsteps, ckpt)
        dE, dd = ctx.adjoint_backward()
    g = grid_of(prm)
    E0 = list(prm.E0)
    J_or = _oracle_J(prm, p0, E0, nsteps, collide)
    assert abs(J - J_or) <= 1e-10 * J_or, (J, J_or)
    # dJ/dE0: central differences at h = 1e-6 |E0| and h / 10; each component
    # within 1e-5 of itself, or of the difference of the two steps (the
    # finite difference's own error) when that is larger
    def fd_E(h):
        fd = np.zeros(3)
        for a in range(3):
            ep, em = list(E0), list(E0)
            ep[a] += h
            em[a] -= h
            fd[a] = (_oracle_J(prm, p0, ep, nsteps, collide) - _oracle_J(prm, p0, em, nsteps, collide)) / (2 * h)
        return fd

    h = 1e-6 * np.linalg.norm(E0)
    fd1, fd2 = fd_E(h), fd_E(0.1 * h)
    bar = np.abs(fd1 - fd2)
    for a in range(3):
        assert abs(dE[a] - fd1[a]) <= max(1e-5 * abs(fd1[a]), 2 * bar[a]), (a, dE[a], fd1[a], fd2[a])
    # dJ/d delta_c on random cells: the weights of the cell's ions at t = 0,
    # steps 1e-5 and 1e-6, each cell against itself
    key0 = orc.cell_keys(g, p0)
    rng = np.random.default_rng(5)
    cells = rng.choice(g.nc, 6, replace=False)

    def fd_w(c, hw):
        sel = key0 == c
        wp, wm = p0.w.copy(), p0.w.copy()
        wp[sel] += hw
        wm[sel] -= hw
        return (_oracle_J(prm, _with_weights(p0, wp), E0, nsteps, collide)
                - _oracle_J(prm, _with_weights(p0, wm), E0, nsteps, collide)) / (2 * hw)

    for c in cells:
        f1, f2 = fd_w(c, 1e-5), fd_w(c, 1e-6)
        assert abs(dd[c] - f1) <= max(1e-5 * abs(f1), 2 * abs(f1 - f2)), (c, dd[c], f1, f2)


def test_adjoint_zero_steps_closed_form():
    """nsteps = 0: J = sum (n - mean n)^2 of the deposit itself, so
    dJ/d delta_c = sum over the cell's ions of gather(2 (n - mean n)), and
    dJ/dE0 = 0."""
    with context(opts((6, 5, 7), extra=BASE)) as ctx:
        ctx.init_particles()
        p = as_particles(ctx.get_particles())
        J = ctx.adjoint_forward(0, 1)
        dE, dd = ctx.adjoint_backward()
        g = grid_of(ctx.prm)
    _, offs = orc.sort_perm(g.nc, orc.cell_keys(g, p))
    n = orc.deposit_sorted(g, p, offs)[0]
    assert J == pytest.approx(float(np.sum((n - n.mean()) ** 2)), rel=1e-13)
    lam = 2 * (n - n.mean())
    z = np.zeros(g.nc)
    gl = np.array([orc.gather(g, lam, z, z, p.x[s], p.y[s], p.z[s])[0] for s in range(p.np)])
    ref = np.bincount(orc.cell_keys(g, p), weights=gl, minlength=g.nc)
    assert relmax(dd, ref) <= 1e-12
    assert np.all(dE == 0.0)


def test_adjoint_checkpoint_interval_invariance_and_end_state():
    """Recomputation from checkpoints is exact: every checkpoint interval
    gives the same gradients, and the context ends at the forward end state."""
    extra = BASE
    res = []
    for ckpt in (1, 2, 5):
        with context(opts((8, 8, 8), extra=extra)) as ctx:
            ctx.init_particles()
            J = ctx.adjoint_forward(5, ckpt)
            end = as_particles(ctx.get_particles())
            phi_end = ctx.get_potential()
            dE, dd = ctx.adjoint_backward()
            after = as_particles(ctx.get_particles())
            assert np.array_equal(after.pid, end.pid) and np.array_equal(after.x, end.x)
            assert np.array_equal(ctx.get_potential(), phi_end)
            res.append((J, dE, dd))
    for J, dE, dd in res[1:]:
        assert J == res[0][0]
        assert np.array_equal(dE, res[0][1]) and np.array_equal(dd, res[0][2])


def test_adjoint_requires_fp64():
    with context(opts((8, 8, 8), extra=BASE + " -precision mixed")) as ctx:
        ctx.init_particles()
        with pytest.raises(chybrid.ChyError) as e:
            ctx.adjoint_forward(2, 1)
        assert e.value.code == chybrid.ERR_SUP


@pytest.mark.timeout(900)
def test_cfg3_adjoint_at_full_size():
    """The differentiation workload at its full size: 64^3 x 16 ions/cell
    (4.2 M), FP64, 1 % Gaussian per-cell perturbation, 50 steps, checkpoints
    every 10, collisions frozen.  dJ/dE0 against central differences at
    h = 1e-8 |E0|, each component within 1e-5 of itself, the difference's
    own bar (bench.fd_check: J is piecewise smooth, its CIC kinks about
    1e-9 |E0| apart, and its rounding limits slopes to ulp(J) / h) or its
    difference to h = 1e-7 (the averaging over kinks that the two steps
    show: E0z measures 86665.46 by the adjoint against 86663.29 at 1e-8 and
    86658.1 at 1e-7, while the rounding bar at 1e-8 is 0.93-2.8 as J(0)
    moves by an ulp), all within 1e-5 of the gradient's norm, and the
    differences at 1e-7 and 1e-6 |E0| no closer to the adjoint than 1e-8
    (they average more pieces); six cells against central differences at
    h = 1e-5, within 1e-5 of themselves, their own bar or the difference to
    h = 1e-6."""
    import os
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    if root not in sys.path:
        sys.path.insert(0, root)
    import bench

    print("cfg3: adjoint at 64^3 x 16 ions/cell, 50 steps", flush=True)
    res = bench.adjoint_check(ncoords=9)
    print({k: res[k] for k in ("dJ_dE0", "E0", "delta")}, flush=True)
    assert res["J_fd_map_rel_diff"] <= 1e-13
    dE = np.array(res["dJ_dE0"])
    c8 = res["E0"]["central_h=1e-08|E0|"]
    fd7 = np.array(res["E0"]["central_h=1e-07|E0|"]["fd"])
    fd, bar = np.array(c8["fd"]), np.array(c8["bar"])
    bar = np.maximum(bar, np.abs(fd - fd7))
    assert res["E0"]["within_1e-5_or_bar_at_1e-8"] == "3/3", res["E0"]
    for a in range(3):
        assert abs(dE[a] - fd[a]) <= max(1e-5 * abs(fd[a]), bar[a]), (a, dE[a], fd[a], bar[a])
    assert c8["max_err_rel_to_norm"] <= 1e-5
    for hr in ("1e-07", "1e-06"):
        assert res["E0"][f"central_h={hr}|E0|"]["max_err_rel_to_norm"] >= 0.5 * c8["max_err_rel_to_norm"]
    assert res["delta"]["within_1e-5_or_bar"] == "6/6", res["delta"]
    assert res["delta"]["central_h=1e-05"]["max_err_rel_to_max"] <= 1e-5, res["delta"]
