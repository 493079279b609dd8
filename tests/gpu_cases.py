"""Shared set-up for the GPU parity tests: reference-style option lists and
matching oracle objects built from the same derived parameters."""
import numpy as np

import chybrid
import orc

AMU = 1.6605e-27


def opts(N=(8, 8, 8), L=0.05, extra=""):
    """Config-0 physics on an N grid with dx = L (reference flag names)."""
    s = (f"-Nx {N[0]} -Ny {N[1]} -Nz {N[2]} -x1 {N[0] * L} -y1 {N[1] * L} -z1 {N[2] * L} "
         f"-mi 4.9816e-26 -mn 4.8156e-26 -Ti 300 -Tn 300 -B0z 5e-5 -E0x 0.05 "
         f"-nui 1e3 -nue 1e5 -dt 1e-5 " + extra)
    return s.split()


def context(args, capacity=0, weighted=False):
    prm, _ = chybrid.params_from_options(args)
    if weighted:
        prm.weighted = 1
    return chybrid.Context(prm, capacity)


def grid_of(prm):
    return orc.Grid([prm.N[a] for a in range(3)], p0=list(prm.p0), p1=list(prm.p1))


def push_params_of(prm, step=0, collide=None):
    g = grid_of(prm)
    col = (prm.mn > 0) if collide is None else collide
    Nc = 0
    if col:
        Nc = orc.collision_count(prm.Np, prm.nui, prm.dt, prm.viT_r, list(prm.vi0), list(prm.vn0),
                                 prm.vnT, prm.mi, prm.mn)
    return g, orc.push_params(g, prm.qi / prm.mi, prm.dt, list(prm.B0), list(prm.E0), collide=col and Nc > 0,
                              mi=prm.mi, mn=prm.mn, viT=prm.viT_r, vnT=prm.vnT, vn0=list(prm.vn0), Nc=Nc,
                              seed=prm.seed, step=step)


def random_particles(g, n, rng, vscale=135.93, weights=False):
    x = g.p0[0] + rng.random(n) * g.L[0]
    y = g.p0[1] + rng.random(n) * g.L[1]
    z = g.p0[2] + rng.random(n) * g.L[2]
    v = rng.standard_normal((3, n)) * vscale
    w = (1 + 0.01 * rng.uniform(-1, 1, n)) if weights else None
    return orc.Particles(x, y, z, v[0], v[1], v[2], w=w)


def upload(ctx, p):
    ctx.set_particles(p.x, p.y, p.z, p.vx, p.vy, p.vz, w=p.w, pid=p.pid)


def as_particles(d):
    return orc.Particles(d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"], pid=d["pid"], w=d["w"])


def oracle_sorted(g, p):
    perm, offs = orc.sort_perm(g.nc, orc.cell_keys(g, p))
    return p.permute(perm), offs


def vlasov_from(mom):
    return np.stack(mom, axis=1)


def relmax(a, b):
    a, b = np.asarray(a), np.asarray(b)
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / (s if s > 0 else 1.0))


def expectation_bar(Nc, Np):
    """Bound on |sum_i p_i - Nc| for one step's collision probabilities
    p_i = Nc |v_r,i| / S (particles.c:529-551).  S holds the expected relative
    speeds, from the velocities before the push, while each p_i uses one
    neutral draw and the velocity after it: the draws make sum_i |v_r,i|
    deviate from S by sd(|v_r|) / mean(|v_r|) / sqrt(Np), at most
    sqrt(3 pi / 8 - 1) = 0.422 / sqrt(Np) (the Maxwell speed distribution's;
    a drift lowers it), here allowed 5 of those, and the push moves the mean
    by under 1e-3 at the benchmark's field and time step."""
    return Nc * (1e-3 + 5 * 0.43 / np.sqrt(Np))
