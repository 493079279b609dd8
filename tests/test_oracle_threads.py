"""The oracle's threaded loops (push, deposit, forcing, stencil, FFT and
z-line preconditioners) give bitwise the one-thread results: every sum
keeps its serial order.  One full step and solve at 32^3 x 4 ions/cell,
sizes above the loops' thresholds."""
import numpy as np

import orc


def _run(threads):
    prev = orc.get_threads()
    orc.set_threads(threads)
    try:
        g = orc.Grid((32, 32, 32), p1=(1.6, 1.6, 1.6))
        K = (0.0, 0.0, orc.Q * 5e-5 / orc.ME / 1e5)
        E0 = (0.05, 0.0, 0.0)
        vT = 135.93
        p = orc.init_particles(g, 4, 9, (0, 0, 0), (vT, vT, vT), pert=2, amp=0.01)
        perm, offs = orc.sort_perm(g.nc, orc.cell_keys(g, p))
        p = p.permute(perm)
        mom = orc.deposit_sorted(g, p, offs)
        b = orc.rhs(g, K, E0, 0.0, orc.ME * 1e5 / orc.Q, *mom)
        phi, its, conv, _ = orc.solve(g, K, mom[0], b, pc=orc.PC_FFT2Z, rtol=1e-10)
        assert conv
        S = orc.normaliser(p, [0.0, 0.0, 0.0], 414.76)
        Nc = orc.collision_count(p.np, 1e3, 1e-5, np.sqrt(3) * vT, (0, 0, 0), (0, 0, 0), 414.76, 4.9816e-26,
                                 4.8156e-26)
        P = orc.push_params(g, orc.Q / 4.9816e-26, 1e-5, (0, 0, 5e-5), E0, collide=True, mi=4.9816e-26,
                            mn=4.8156e-26, viT=np.sqrt(3) * vT, vnT=414.76, Nc=Nc, seed=9, step=0)
        q, stats, mom2, S2, _ = orc.step(g, P, p, phi, S, E0)
        zl = orc.pc_apply(g, list(K), mom[0], b, pc=orc.PC_ZLINE)
        return [phi, q.x, q.vx, q.vz, *mom2, np.array([S2, its]), zl]
    finally:
        orc.set_threads(prev)


def test_threads_do_not_change_results():
    one = _run(1)
    many = _run(4)
    for a, b in zip(one, many):
        assert np.array_equal(a, b)
