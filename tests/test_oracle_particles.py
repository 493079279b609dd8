"""Pins of the oracle particle path (reference particles.c, vectors.c, random.c).

Known-answer vectors (Philox), quadrature (mean relative speed), exact
kinematics (Boris invariants, E x B drift, elastic-collision invariants),
brute force (sort, scatter deposit) and statistics (collision counts).
"""
import numpy as np
import pytest
from scipy import integrate, special

import orc

AMU = 1.6605e-27


def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32_10
    assert orc.philox([0, 0, 0, 0], [0, 0]).tolist() == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert orc.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2).tolist() == [
        0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert orc.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]).tolist() == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_u24_range_and_exactness():
    lo, hi = orc.u24(0), orc.u24(0xFFFFFFFF)
    assert lo == 0.5 / 2**24 and hi == 1 - 0.5 / 2**24
    assert np.float32(orc.u24(0x12345678)) == orc.u24(0x12345678)


@pytest.mark.parametrize("mu", [0.0, 1e-5, 3e-4, 0.2, 1.0, 3.7, 40.0])
def test_mean_relative_speed_matches_quadrature(mu):
    sigma = 414.76
    v = np.array([mu * sigma, 0.0, 0.0])
    lam = mu
    # E|X|, X ~ N(mu e_x, sigma^2 I): radial integral of the non-central chi-3 density
    def dens(r):
        if lam == 0.0:
            return np.sqrt(2 / np.pi) * r * r * np.exp(-r * r / 2)
        return r / (lam * np.sqrt(2 * np.pi)) * (np.exp(-(r - lam) ** 2 / 2) - np.exp(-(r + lam) ** 2 / 2))
    m, _ = integrate.quad(lambda r: r * dens(r), 0, lam + 40, limit=200)
    got = orc.mean_rel_speed(v, (0, 0, 0), sigma)
    assert abs(got - sigma * m) <= 1e-9 * sigma * m
    assert orc.mean_rel_speed(v + 5.0, (5.0, 5.0, 5.0), sigma) == pytest.approx(got, rel=1e-12)
    assert orc.mean_rel_speed(v, (0, 0, 0), 0.0) == pytest.approx(mu * sigma, rel=1e-15, abs=1e-300)
    _ = special


def test_collision_count_formula():
    # particles.c:510-519 with the config-0 parameters
    mi, mn = 30 * AMU, 29 * AMU
    viT = np.sqrt(3) * np.sqrt(2 * orc.KB * 300 / mi) / 3
    vnT = np.sqrt(2 * orc.KB * 300 / mn)
    Nc = orc.collision_count(262144, 1e3, 1e-5, viT, (0, 0, 0), (0, 0, 0), vnT, mi, mn)
    vrm = 4 * viT
    assert Nc == int(262144 * 1e-2 * vrm / (vnT * np.sqrt(mn / mi)))
    assert 6000 < Nc < 6100
    assert orc.collision_count(10, 1e9, 1.0, viT, (0, 0, 0), (0, 0, 0), vnT, mi, mn) == 10


def test_wrap_and_cell_coord_edges():
    assert orc.wrap(1.0, 0.0, 1.0, 1.0) == 0.0
    x = orc.wrap(-1e-18, 0.0, 1.0, 1.0)
    assert 0.0 <= x < 1.0
    assert orc.wrap(2.75, 0.0, 1.0, 1.0) == 0.75
    assert orc.wrap(-1.25, 0.0, 1.0, 1.0) == 0.75
    i, f = orc.cell_coord(np.nextafter(1.6, 0), 0.0, 0.05, 32)
    assert i == 31 and 0.0 <= f <= 1.0
    i, f = orc.cell_coord(0.075, 0.0, 0.05, 32)
    assert i == 1 and f == pytest.approx(0.5)


def _grid():
    return orc.Grid((6, 5, 4), p0=(0.1, -0.2, 0.3), p1=(0.4, 0.05, 0.5))


def test_gather_reproduces_trilinear_fields_and_partition_of_unity():
    g = _grid()
    k, j, i = np.meshgrid(*[np.arange(int(m)) for m in g.shape()], indexing="ij")
    f = (1.5 + 0.25 * i + 0.5 * j - 0.75 * k + 0.125 * i * j * k).ravel().astype(float)
    rng = np.random.default_rng(1)
    for _ in range(50):
        # interior cells only: the field is trilinear there (no periodic seam)
        ci, cj, ck = rng.integers(0, 5), rng.integers(0, 4), rng.integers(0, 3)
        u = rng.random(3)
        x = g.p0 + (np.array([ci, cj, ck]) + u) * g.d
        E = orc.gather(g, f, np.ones(g.nc), np.zeros(g.nc), *x)
        xi, yj, zk = ci + u[0], cj + u[1], ck + u[2]
        assert E[0] == pytest.approx(1.5 + 0.25 * xi + 0.5 * yj - 0.75 * zk + 0.125 * xi * yj * zk, rel=1e-14)
        assert E[1] == pytest.approx(1.0, rel=1e-15)
        assert E[2] == 0.0


def test_efield_nodes_central_difference():
    g = _grid()
    rng = np.random.default_rng(2)
    phi = rng.standard_normal(g.nc)
    E0 = (0.05, 0.01, -0.02)
    ex, ey, ez = orc.efield(g, E0, phi)
    P = phi.reshape(g.shape())
    for arr, ax, e0, d in ((ex, 2, E0[0], g.d[0]), (ey, 1, E0[1], g.d[1]), (ez, 0, E0[2], g.d[2])):
        ref = e0 - (np.roll(P, -1, axis=ax) - np.roll(P, 1, axis=ax)) / (2 * d)
        assert np.allclose(arr, ref.ravel(), rtol=1e-14, atol=1e-14)


def _random_particles(g, n, rng, weights=True):
    x = g.p0[0] + rng.random(n) * g.L[0]
    y = g.p0[1] + rng.random(n) * g.L[1]
    z = g.p0[2] + rng.random(n) * g.L[2]
    v = rng.standard_normal((3, n)) * 100
    w = 1 + 0.1 * rng.standard_normal(n) if weights else None
    return orc.Particles(x, y, z, *v, w=w)


def test_deposit_single_particle_known_answer():
    g = _grid()
    x = g.p0 + (np.array([5, 4, 3]) + [0.25, 0.5, 0.75]) * g.d  # top corner cell: wraps all axes
    p = orc.Particles([x[0]], [x[1]], [x[2]], [10.0], [-20.0], [30.0], w=[2.0])
    key = orc.cell_keys(g, p)
    perm, offs = orc.sort_perm(g.nc, key)
    n, gx, gy, gz = orc.deposit_sorted(g, p, offs)
    wx, wy, wz = [0.75, 0.25], [0.5, 0.5], [0.25, 0.75]
    expect = np.zeros(g.shape())
    for q in range(8):
        a, b, c = q & 1, (q >> 1) & 1, q >> 2
        expect[(3 + c) % 4, (4 + b) % 5, (5 + a) % 6] += 2.0 * wx[a] * wy[b] * wz[c]
    assert np.allclose(n.reshape(g.shape()), expect, rtol=1e-15, atol=0)
    assert np.allclose(gz, 30.0 * n, rtol=1e-15)
    assert n.sum() == pytest.approx(2.0, rel=1e-15)


def test_deposit_sorted_matches_scatter_conserves_and_is_gather_transpose():
    g = _grid()
    rng = np.random.default_rng(3)
    p = _random_particles(g, 5000, rng)
    perm, offs = orc.sort_perm(g.nc, orc.cell_keys(g, p))
    ps = p.permute(perm)
    a = orc.deposit_sorted(g, ps, offs)
    b = orc.deposit_scatter(g, p)
    for u, v in zip(a, b):
        assert np.abs(u - v).max() <= 1e-12 * np.abs(v).max()
    assert a[0].sum() == pytest.approx(p.w.sum(), rel=1e-13)
    for m, v in zip(a[1:], (p.vx, p.vy, p.vz)):
        assert m.sum() == pytest.approx((p.w * v).sum(), rel=1e-11)
    field = rng.standard_normal(g.nc)
    lhs = float(a[0] @ field)
    rhs = sum(ps.w[s] * orc.gather(g, field, field, field, ps.x[s], ps.y[s], ps.z[s])[0]
              for s in range(ps.np))
    assert lhs == pytest.approx(rhs, rel=1e-12)


def test_stable_sort_matches_numpy():
    rng = np.random.default_rng(4)
    key = rng.integers(0, 50, 2000).astype(np.int32)
    perm, offs = orc.sort_perm(50, key)
    assert np.array_equal(perm, np.argsort(key, kind="stable"))
    assert np.array_equal(np.diff(offs), np.bincount(key, minlength=50))


def _push(g, parts, B0, E0, qm=orc.Q / (30 * AMU), dt=1e-5, phi=None, S=1.0, **kw):
    P = orc.push_params(g, qm, dt, B0, E0, **kw)
    ex, ey, ez = orc.efield(g, E0, np.zeros(g.nc) if phi is None else phi)
    return orc.push(P, ex, ey, ez, parts, S)


def test_boris_conserves_speed_and_rotates_by_2atan_t():
    g = orc.Grid((8, 8, 8), p1=(0.4, 0.4, 0.4))
    rng = np.random.default_rng(5)
    p = _random_particles(g, 200, rng, weights=False)
    p.vz[:] = 0.0
    B0 = (0.0, 0.0, 0.5)
    qm, dt = orc.Q / (30 * AMU), 1e-5
    v0 = np.stack([p.vx, p.vy, p.vz]).copy()
    speed0 = np.linalg.norm(v0, axis=0)
    _push(g, p, B0, (0, 0, 0), qm=qm, dt=dt)
    v1 = np.stack([p.vx, p.vy, p.vz])
    assert np.allclose(np.linalg.norm(v1, axis=0), speed0, rtol=1e-14)
    ang = np.arctan2(v0[0] * v1[1] - v0[1] * v1[0], (v0[:2] * v1[:2]).sum(0))
    t = 0.5 * qm * dt * B0[2]
    assert np.allclose(ang, -2 * np.arctan(t), rtol=1e-10)


def test_boris_pure_electric_and_exb_drift():
    g = orc.Grid((8, 8, 8), p1=(0.4, 0.4, 0.4))
    qm, dt = orc.Q / (30 * AMU), 1e-5
    p = orc.Particles([0.1], [0.1], [0.1], [1.0], [2.0], [3.0])
    _push(g, p, (0, 0, 0), (0.05, -0.03, 0.02), qm=qm, dt=dt)
    assert np.allclose([p.vx[0], p.vy[0], p.vz[0]],
                       np.array([1.0, 2.0, 3.0]) + qm * dt * np.array([0.05, -0.03, 0.02]), rtol=1e-14)
    # E x B drift: the Boris map's fixed point is E x B / B^2, so the mean over
    # whole discrete gyro-periods (200 steps of angle 2 atan(t) = 2 pi / 200)
    # equals it to rounding
    B = 0.5
    dt = 2 * np.tan(np.pi / 200) / (qm * B)
    p = orc.Particles([0.2], [0.2], [0.2], [0.0], [0.0], [0.0])
    vsum = np.zeros(3)
    for _ in range(600):
        _push(g, p, (0, 0, B), (0.05, 0, 0), qm=qm, dt=dt)
        vsum += [p.vx[0], p.vy[0], p.vz[0]]
    vd = vsum / 600
    assert vd[1] == pytest.approx(-0.05 / B, rel=1e-9)
    assert abs(vd[0]) < 1e-9 * 0.05 / B and vd[2] == 0.0


def test_push_keys_and_positions():
    g = _grid()
    rng = np.random.default_rng(6)
    p = _random_particles(g, 3000, rng, weights=False)
    p.vx *= 100
    key, _, _ = _push(g, p, (0, 0, 1e-4), (0.02, 0, 0))
    for a, (arr, lo, hi) in enumerate(((p.x, g.p0[0], g.p1[0]), (p.y, g.p0[1], g.p1[1]), (p.z, g.p0[2], g.p1[2]))):
        assert (arr >= lo).all() and (arr < hi).all()
    assert np.array_equal(key, orc.cell_keys(g, p))


def _collide_setup(n, vnT, seed=11, step=3, nui=1e3, Tn_zero=False):
    g = orc.Grid((4, 4, 4), p1=(0.2, 0.2, 0.2))
    rng = np.random.default_rng(seed)
    mi, mn = 30 * AMU, 29 * AMU
    viT = np.sqrt(3) * 135.93
    p = _random_particles(g, n, rng, weights=False)
    for a in ("vx", "vy", "vz"):
        setattr(p, a, rng.standard_normal(n) * 135.93)
    Nc = orc.collision_count(n, nui, 1e-5, viT, (0, 0, 0), (0, 0, 0), vnT if vnT > 0 else 414.76, mi, mn)
    S = sum(orc.mean_rel_speed((p.vx[s], p.vy[s], p.vz[s]), (0, 0, 0), vnT) for s in range(n))
    return g, p, dict(collide=True, mi=mi, mn=mn, viT=viT, vnT=vnT, Nc=Nc, seed=seed, step=step), S


def test_elastic_collision_invariants_with_cold_neutrals():
    """vnT = 0: the neutral velocity is vn0 exactly, so each accepted event must
    keep |v - v_cm| and the CoM velocity; scattered directions are isotropic."""
    n = 20000
    g, p, kw, S = _collide_setup(n, vnT=0.0)
    v0 = np.stack([p.vx, p.vy, p.vz]).copy()
    mi, mn = kw["mi"], kw["mn"]
    key, Sn, stats = _push(g, p, (0, 0, 0), (0, 0, 0), dt=1e-12, S=S, **kw)
    v1 = np.stack([p.vx, p.vy, p.vz])
    changed = np.any(v1 != v0, axis=0)
    assert changed.sum() == stats[0] > 100
    vcm = v0 * mi / (mi + mn)
    r0 = np.linalg.norm(v0 - vcm, axis=0)[changed]
    r1 = np.linalg.norm(v1 - vcm, axis=0)[changed]
    assert np.allclose(r0, r1, rtol=1e-12)
    u = ((v1 - vcm)[:, changed]) / r1
    m = u.mean(axis=1)
    assert np.all(np.abs(m) < 4 / np.sqrt(changed.sum()))


def test_collision_count_statistics():
    """Accepted events are a Poisson-binomial with E = Nc (when uncapped)."""
    n = 40000
    vnT = 414.76
    g, p, kw, S = _collide_setup(n, vnT=vnT, nui=3e3)
    counts = []
    for step in range(6):
        kw["step"] = step
        q = p.copy()
        _, _, stats = _push(g, q, (0, 0, 0), (0, 0, 0), dt=1e-12, S=S, **kw)
        assert stats[2] == 0
        counts.append(stats[0] + stats[1])
    Nc = kw["Nc"]
    pbar = Nc / n
    sig = np.sqrt(n * pbar * (1 - pbar))
    for c in counts:
        assert abs(c - Nc) < 4 * sig, (c, Nc, sig)
    assert abs(np.mean(counts) - Nc) < 4 * sig / np.sqrt(len(counts))


def test_init_particles_layout_and_moments():
    g = orc.Grid((4, 4, 4), p1=(0.2, 0.2, 0.2))
    vT = (135.93, 135.93, 135.93)
    p = orc.init_particles(g, 16, 0, (10.0, 0.0, 0.0), vT, pert=2, amp=0.01)
    key = orc.cell_keys(g, p)
    assert np.array_equal(key, np.repeat(np.arange(g.nc), 16))
    assert np.array_equal(p.pid, np.arange(g.nc * 16))
    assert abs(p.vx.mean() - 10.0) < 4 * 135.93 / np.sqrt(p.np)
    assert np.std(p.vy) == pytest.approx(135.93, rel=0.05)
    wc = p.w.reshape(g.nc, 16)
    assert np.all(wc == wc[:, :1]) and np.all(np.abs(wc - 1) <= 0.01)
    q = orc.init_particles(g, 16, 0, (10.0, 0.0, 0.0), vT, pert=2, amp=0.01)
    assert np.array_equal(p.x, q.x) and np.array_equal(p.w, q.w)
    r = orc.init_particles(g, 16, 1, (10.0, 0.0, 0.0), vT, pert=2, amp=0.01)
    assert not np.array_equal(p.x, r.x)


# ---- the reference's density placements (particles.c:26-45, distributions.c) ----

def _textbook_sobol(npts, s, a, m_lead, bits=30):
    """Points 1..npts of a one-dimensional Sobol sequence by Bratley & Fox
    (ACM TOMS 14, 1988): primitive polynomial x^s + a_1 x^(s-1) + ... +
    a_(s-1) x + 1, odd leading integers m_1..m_s, then
    m_j = 2 a_1 m_(j-1) ^ 4 a_2 m_(j-2) ^ ... ^ 2^s m_(j-s) ^ m_(j-s), and
    point n = XOR of v_i = m_i / 2^i over the set bits i of the Gray code of n
    (Antonov & Saleev)."""
    m = list(m_lead)
    for j in range(s, bits):
        v = m[j - s] ^ (m[j - s] << s)
        for i in range(1, s):
            if a[i - 1]:
                v ^= m[j - i] << i
        m.append(v)
    V = [m[j] << (bits - 1 - j) for j in range(bits)]
    out = np.zeros(npts)
    for n in range(1, npts + 1):
        g = n ^ (n >> 1)
        x, i = 0, 0
        while g:
            if g & 1:
                x ^= V[i]
            g >>= 1
            i += 1
        out[n - 1] = x / 2.0**bits
    return out


def test_sobol_sequence_is_the_textbook_generator():
    """The reference's generator (random.c:234-290) is the Bratley-Fox Sobol
    sequence of the polynomials x + 1, x^2 + x + 1 and, because its
    recurrence XORs the previous integer for each polynomial bit
    (random.c:265), x^3 + x^2 + 1 (code 1 names x^3 + x + 1 in the table it
    comes from; that generator differs from the third point of dimension 3 on)."""
    n = 1 << 12
    r = orc.sobol_points(n)
    assert r[0].tolist() == [0.5, 0.5, 0.5]
    assert r[1].tolist() == [0.25, 0.75, 0.25]
    assert np.array_equal(r[:, 0], _textbook_sobol(n, 1, [], [1]))
    assert np.array_equal(r[:, 1], _textbook_sobol(n, 2, [1], [1, 1]))
    assert np.array_equal(r[:, 2], _textbook_sobol(n, 3, [1, 0], [1, 3, 7]))
    assert not np.array_equal(r[:, 2], _textbook_sobol(n, 3, [0, 1], [1, 3, 7]))


def test_sobol_points_stratify_every_axis():
    """With x_0 = 0 before them, the first 2^m - 1 calls fill each of the 2^m
    dyadic intervals of every axis exactly once ((0, m, 1)-nets)."""
    r = orc.sobol_points((1 << 14) - 1)
    for m in range(1, 15):
        for k in range(3):
            pts = np.concatenate([[0.0], r[:(1 << m) - 1, k]])
            assert np.array_equal(np.sort(np.floor(pts * (1 << m))), np.arange(1 << m)), (m, k)


def test_sobol_needs_at_most_30_bits():
    # call 2^30 - 1 would need the 31st direction integer (random.c:279-281)
    V = orc.sobol_dirs()
    assert V.shape == (30, 3) and np.all(V[0] == 1 << 29)
    # the leading 1 of integer j sits at bit 29 - j, so each is odd once scaled back
    for j in range(30):
        assert np.all((V[j] >> (29 - j)) & 1 == 1)


def test_flat_sobol_placement():
    g = orc.Grid((4, 5, 6), p1=(0.2, 0.25, 0.3))
    n = 3000
    p = orc.init_density(g, orc.DENSITY_FLAT_SOBOL, n, 7, (0.0, 0.0, 0.0), (100.0, 100.0, 100.0))
    r = orc.sobol_points(n)
    for a, x in enumerate((p.x, p.y, p.z)):
        v = r[:, a] * (g.L[a] + g.d[a]) - 0.5 * g.d[a]
        v = np.where(v < 0, 0.01 * g.d[a], np.where(v > g.p1[a], g.p1[a] - 0.01 * g.d[a], v))
        assert np.array_equal(x, v), a
        assert np.all((x >= g.p0[a]) & (x < g.p1[a]))
    assert np.array_equal(p.pid, np.arange(n))
    # the pulled-in points pile up 0.01 d inside both faces
    assert np.sum(p.x == 0.01 * g.d[0]) > 0 and np.sum(p.x == g.p1[0] - 0.01 * g.d[0]) > 0


def test_uniform_placement_fills_all_but_the_last_cell_of_each_axis():
    g = orc.Grid((5, 4, 6), p1=(0.25, 0.2, 0.3))
    Np = 3 * g.nc + 7
    p = orc.init_density(g, orc.DENSITY_UNIFORM, Np, 1, (0.0, 0.0, 0.0), (100.0, 100.0, 100.0))
    assert p.np == 3 * 4 * 3 * 5
    key = orc.cell_keys(g, p)
    cnt = np.bincount(key, minlength=g.nc).reshape(g.shape())  # [z][y][x]
    assert np.all(cnt[:5, :3, :4] == 3) and cnt.sum() == p.np
    # cell centres (i + 1/2) d
    for a, x in enumerate((p.x, p.y, p.z)):
        idx = np.rint(x / g.d[a] - 0.5)
        assert np.allclose(x, (idx + 0.5) * g.d[a], rtol=0, atol=1e-15)
    # k (z index) fastest, then j, then i; ion blocks repeat per ion slot
    assert p.z[1] > p.z[0] and p.y[1] == p.y[0] and p.x[1] == p.x[0]
    assert np.array_equal(p.x[:60], p.x[60:120])


def test_sinusoidal_placement_follows_its_density():
    """Rejection against 1 + 0.25 sin(2 pi x) (distributions.c:152-162) over
    x in [0, 1.6): chi-square of the x histogram against the integral of the
    density per bin, and uniform y and z."""
    from scipy import stats
    g = orc.Grid((32, 8, 8), p1=(1.6, 0.4, 0.4))
    n = 200000
    p = orc.init_density(g, orc.DENSITY_SINUSOIDAL, n, 5, (0.0, 0.0, 0.0), (100.0, 100.0, 100.0))
    edges = np.linspace(0.0, 1.6, 33)
    F = edges - 0.25 / (2 * np.pi) * np.cos(2 * np.pi * edges)
    expect = np.diff(F) / (F[-1] - F[0]) * n
    obs, _ = np.histogram(p.x, edges)
    chi2 = np.sum((obs - expect) ** 2 / expect)
    assert stats.chi2.sf(chi2, 31) > 1e-4, chi2
    for x, L in ((p.y, 0.4), (p.z, 0.4)):
        o, _ = np.histogram(x, np.linspace(0, L, 17))
        assert stats.chi2.sf(np.sum((o - n / 16) ** 2 / (n / 16)), 15) > 1e-4
    # a flat density would fail the same test
    flat = np.full(32, n / 32)
    assert stats.chi2.sf(np.sum((obs - flat) ** 2 / flat), 31) < 1e-10
