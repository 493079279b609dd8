"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: imported by tests/, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of ``bench.py``, always as the checker.  The
product library never loads this module or the shared object behind it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

D = np.float64
I64 = np.int64

LHS_IDENTITY, LHS_LAPLACIAN, LHS_FULL = 0, 1, 2
RHS_CONSTANT, RHS_SINUSOIDAL, RHS_FULL = 0, 1, 2
PC_NONE, PC_JACOBI, PC_FFT, PC_ZLINE, PC_CHEBY, PC_FFT2, PC_FFT2Z, PC_MG = 0, 1, 2, 3, 4, 5, 6, 7
FFT2_ITS = 2
CHEBY_DEGREE = 4
KSP_GMRES, KSP_BCGS = 0, 1

Q, MP, ME, KB = 1.6022e-19, 1.6726e-27, 9.1094e-31, 1.3807e-23


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _LIB = C.CDLL(path)
        _declare(_LIB)
    return _LIB


P = C.c_void_p


def _declare(L):
    L.orc_philox4x32_10.argtypes = [P, P, P]
    L.orc_u24.argtypes = [C.c_uint32]
    L.orc_u24.restype = C.c_double
    L.orc_lhs_full_row.argtypes = [P, P, P, P, C.c_int64, P]
    L.orc_stencil_offsets.argtypes = [P]
    for f in (L.orc_lhs_apply, L.orc_lhs_apply_transpose):
        f.argtypes = [C.c_int, P, P, P, P, P, P]
    L.orc_lhs_dense.argtypes = [C.c_int, P, P, P, P, P]
    L.orc_rhs.argtypes = [C.c_int, P, P, P, P, C.c_double, C.c_double, P, P, P, P, P]
    L.orc_remove_mean.argtypes = [C.c_int64, P]
    L.orc_cell_coord.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int64, P, P]
    L.orc_wrap.argtypes = [C.c_double] * 4
    L.orc_wrap.restype = C.c_double
    L.orc_efield_nodes.argtypes = [P, P, P, P, P, P, P]
    L.orc_gather.argtypes = [P, P, P, P, P, P, C.c_double, C.c_double, C.c_double, P]
    L.orc_mean_rel_speed.argtypes = [P, P, C.c_double]
    L.orc_mean_rel_speed.restype = C.c_double
    L.orc_normaliser.argtypes = [C.c_int64, P, P, P, P, C.c_double]
    L.orc_mg_levels.argtypes = [P, P]
    L.orc_mg_restrict.argtypes = [P, P, P]
    L.orc_mg_prolong_add.argtypes = [P, P, P]
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_get_threads.restype = C.c_int
    L.orc_normaliser.restype = C.c_double
    L.orc_collision_count.argtypes = [C.c_int64] + [C.c_double] * 3 + [P, P] + [C.c_double] * 3
    L.orc_collision_count.restype = C.c_int64
    L.orc_push.argtypes = [P, P, P, P, C.c_int64, P, P, P, P, P, P, P, C.c_double, P, P, P]
    L.orc_sort_perm.argtypes = [C.c_int64, C.c_int64, P, P, P]
    L.orc_permute_d.argtypes = [C.c_int64, P, P, P]
    L.orc_permute_u.argtypes = [C.c_int64, P, P, P]
    L.orc_deposit_sorted.argtypes = [P, P, P, C.c_int64] + [P] * 7 + [P] + [P] * 4
    L.orc_deposit_scatter.argtypes = [P, P, P, C.c_int64] + [P] * 7 + [P] * 4
    L.orc_init_particles.argtypes = [P, P, P, C.c_int64, C.c_uint64, P, P, C.c_int, C.c_double] + [P] * 8
    L.orc_sobol_dirs.argtypes = [P]
    L.orc_sobol_points.argtypes = [C.c_int64, P]
    L.orc_density_count.argtypes = [C.c_int, P, C.c_int64]
    L.orc_density_count.restype = C.c_int64
    L.orc_init_density.argtypes = [C.c_int, P, P, P, P, P, C.c_int64, C.c_uint64, P, P, C.c_int, C.c_double] + [P] * 8
    L.orc_init_density.restype = C.c_int64
    L.orc_pc_apply.argtypes = [C.c_int, C.c_int, P, P, P, P, P, P]
    L.orc_pc_apply_deg.argtypes = [C.c_int, C.c_int, P, P, P, P, P, P, C.c_int]
    L.orc_solve.argtypes = [C.c_int, C.c_int, C.c_int, P, P, P, P, P, P,
                            C.c_double, C.c_double, C.c_int64, C.c_int64, P, P]
    L.orc_fft3d.argtypes = [P, P, C.c_int]


def ptr(a):
    return None if a is None else a.ctypes.data


def _c(a, dt=D):
    return np.ascontiguousarray(a, dtype=dt)


def philox(ctr, key):
    c = _c(ctr, np.uint32)
    k = _c(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(ptr(c), ptr(k), ptr(out))
    return out


def u24(w):
    return lib().orc_u24(int(w))


class Grid:
    """Periodic node-centred grid: N nodes per axis on [p0, p1)."""

    def __init__(self, N, p0=(0.0, 0.0, 0.0), p1=(1.0, 1.0, 1.0)):
        self.N = _c(N, I64)
        self.p0 = _c(p0)
        self.p1 = _c(p1)
        self.L = self.p1 - self.p0
        self.d = self.L / self.N
        self.nc = int(np.prod(self.N))

    def shape(self):
        """numpy shape of a field in natural order: [Nz][Ny][Nx]."""
        return (int(self.N[2]), int(self.N[1]), int(self.N[0]))


def offsets19():
    off = np.zeros((19, 3), np.int32)
    lib().orc_stencil_offsets(ptr(off))
    return off


def lhs_row(g, K, n, c):
    K = _c(K)
    n = _c(n)
    out = np.zeros(19)
    lib().orc_lhs_full_row(ptr(g.N), ptr(g.d), ptr(K), ptr(n), int(c), ptr(out))
    return out


def lhs_apply(g, K, n, x, lhs_type=LHS_FULL, transpose=False):
    K, n, x = _c(K), _c(n), _c(x)
    y = np.zeros(g.nc)
    f = lib().orc_lhs_apply_transpose if transpose else lib().orc_lhs_apply
    f(lhs_type, ptr(g.N), ptr(g.d), ptr(K), ptr(n), ptr(x), ptr(y))
    return y


def lhs_dense(g, K, n, lhs_type=LHS_FULL):
    K, n = _c(K), _c(n)
    A = np.zeros((g.nc, g.nc))
    lib().orc_lhs_dense(lhs_type, ptr(g.N), ptr(g.d), ptr(K), ptr(n), ptr(A))
    return A


def rhs(g, K, E0, cth, cG, n, gx, gy, gz, rhs_type=RHS_FULL):
    K, E0 = _c(K), _c(E0)
    n, gx, gy, gz = _c(n), _c(gx), _c(gy), _c(gz)
    b = np.zeros(g.nc)
    lib().orc_rhs(rhs_type, ptr(g.N), ptr(g.d), ptr(K), ptr(E0), cth, cG,
                  ptr(n), ptr(gx), ptr(gy), ptr(gz), ptr(b))
    return b


def efield(g, E0, phi):
    E0, phi = _c(E0), _c(phi)
    ex, ey, ez = np.zeros(g.nc), np.zeros(g.nc), np.zeros(g.nc)
    lib().orc_efield_nodes(ptr(g.N), ptr(g.d), ptr(E0), ptr(phi), ptr(ex), ptr(ey), ptr(ez))
    return ex, ey, ez


def gather(g, ex, ey, ez, x, y, z):
    out = np.zeros(3)
    ex, ey, ez = _c(ex), _c(ey), _c(ez)
    lib().orc_gather(ptr(g.N), ptr(g.p0), ptr(g.d), ptr(ex), ptr(ey), ptr(ez),
                     float(x), float(y), float(z), ptr(out))
    return out


def cell_coord(x, p0, d, N):
    i = np.zeros(1, I64)
    f = np.zeros(1)
    lib().orc_cell_coord(float(x), float(p0), float(d), int(N), ptr(i), ptr(f))
    return int(i[0]), float(f[0])


def wrap(x, p0, p1, L):
    return lib().orc_wrap(float(x), float(p0), float(p1), float(L))


def mean_rel_speed(v, vn0, vnT):
    v, vn0 = _c(v), _c(vn0)
    return lib().orc_mean_rel_speed(ptr(v), ptr(vn0), float(vnT))


def normaliser(parts, vn0, vnT):
    """Sum over the ions of the mean relative speed (collision normaliser)."""
    vn0 = _c(vn0)
    return lib().orc_normaliser(parts.np, ptr(parts.vx), ptr(parts.vy), ptr(parts.vz), ptr(vn0), float(vnT))


def collision_count(Np, nui, dt, viT, vi0, vn0, vnT, mi, mn):
    vi0, vn0 = _c(vi0), _c(vn0)
    return lib().orc_collision_count(int(Np), nui, dt, viT, ptr(vi0), ptr(vn0), vnT, mi, mn)


class PushParams(C.Structure):
    _fields_ = [("N", C.c_int64 * 3), ("p0", C.c_double * 3), ("p1", C.c_double * 3),
                ("d", C.c_double * 3), ("L", C.c_double * 3), ("qm", C.c_double),
                ("dt", C.c_double), ("B0", C.c_double * 3), ("E0", C.c_double * 3),
                ("collide", C.c_int), ("mi", C.c_double), ("mn", C.c_double),
                ("viT", C.c_double), ("vnT", C.c_double), ("vn0", C.c_double * 3),
                ("Nc", C.c_int64), ("seed", C.c_uint64), ("step", C.c_uint32)]


def push_params(g, qm, dt, B0, E0, collide=False, mi=1.0, mn=0.0, viT=1.0, vnT=0.0,
                vn0=(0, 0, 0), Nc=0, seed=0, step=0):
    P_ = PushParams()
    for a in range(3):
        P_.N[a] = int(g.N[a]); P_.p0[a] = g.p0[a]; P_.p1[a] = g.p1[a]
        P_.d[a] = g.d[a]; P_.L[a] = g.L[a]; P_.B0[a] = B0[a]; P_.E0[a] = E0[a]
        P_.vn0[a] = vn0[a]
    P_.qm, P_.dt = qm, dt
    P_.collide, P_.mi, P_.mn, P_.viT, P_.vnT = int(collide), mi, mn, viT, vnT
    P_.Nc, P_.seed, P_.step = int(Nc), int(seed), int(step)
    return P_


class Particles:
    """FP64 SoA particle state."""

    def __init__(self, x, y, z, vx, vy, vz, pid=None, w=None):
        self.x, self.y, self.z = _c(x), _c(y), _c(z)
        self.vx, self.vy, self.vz = _c(vx), _c(vy), _c(vz)
        n = self.x.size
        self.pid = _c(np.arange(n) if pid is None else pid, np.uint32)
        self.w = None if w is None else _c(w)

    @property
    def np(self):
        return self.x.size

    def copy(self):
        return Particles(self.x.copy(), self.y.copy(), self.z.copy(), self.vx.copy(),
                         self.vy.copy(), self.vz.copy(), self.pid.copy(),
                         None if self.w is None else self.w.copy())

    def permute(self, perm):
        out = self.copy()
        for name in ("x", "y", "z", "vx", "vy", "vz"):
            setattr(out, name, getattr(self, name)[perm].copy())
        out.pid = self.pid[perm].copy()
        if self.w is not None:
            out.w = self.w[perm].copy()
        return out


def push(P_, ex, ey, ez, parts, S):
    key = np.zeros(parts.np, np.int32)
    Sn = np.zeros(1)
    stats = np.zeros(3, I64)
    ex, ey, ez = _c(ex), _c(ey), _c(ez)
    lib().orc_push(C.byref(P_), ptr(ex), ptr(ey), ptr(ez), parts.np,
                   ptr(parts.x), ptr(parts.y), ptr(parts.z),
                   ptr(parts.vx), ptr(parts.vy), ptr(parts.vz), ptr(parts.pid),
                   float(S), ptr(key), ptr(Sn), ptr(stats))
    return key, float(Sn[0]), stats


def sort_perm(ncell, key):
    key = _c(key, np.int32)
    perm = np.zeros(key.size, I64)
    offs = np.zeros(ncell + 1, I64)
    lib().orc_sort_perm(int(ncell), key.size, ptr(key), ptr(perm), ptr(offs))
    return perm, offs


def cell_keys(g, parts):
    xi = ((parts.x - g.p0[0]) / g.d[0]).astype(np.int64).clip(0, g.N[0] - 1)
    yi = ((parts.y - g.p0[1]) / g.d[1]).astype(np.int64).clip(0, g.N[1] - 1)
    zi = ((parts.z - g.p0[2]) / g.d[2]).astype(np.int64).clip(0, g.N[2] - 1)
    return (xi + g.N[0] * (yi + g.N[1] * zi)).astype(np.int32)


def deposit_sorted(g, parts, offsets):
    offs = _c(offsets, I64)
    out = [np.zeros(g.nc) for _ in range(4)]
    lib().orc_deposit_sorted(ptr(g.N), ptr(g.p0), ptr(g.d), parts.np,
                             ptr(parts.x), ptr(parts.y), ptr(parts.z),
                             ptr(parts.vx), ptr(parts.vy), ptr(parts.vz), ptr(parts.w),
                             ptr(offs), *[ptr(o) for o in out])
    return out


def deposit_scatter(g, parts):
    out = [np.zeros(g.nc) for _ in range(4)]
    lib().orc_deposit_scatter(ptr(g.N), ptr(g.p0), ptr(g.d), parts.np,
                              ptr(parts.x), ptr(parts.y), ptr(parts.z),
                              ptr(parts.vx), ptr(parts.vy), ptr(parts.vz), ptr(parts.w),
                              *[ptr(o) for o in out])
    return out


def init_particles(g, ppc, seed, v0, vT, pert=0, amp=0.0):
    n = int(g.nc * ppc)
    arrs = [np.zeros(n) for _ in range(6)]
    pid = np.zeros(n, np.uint32)
    w = np.zeros(n) if pert else None
    v0, vT = _c(v0), _c(vT)
    lib().orc_init_particles(ptr(g.N), ptr(g.p0), ptr(g.d), int(ppc), int(seed),
                             ptr(v0), ptr(vT), int(pert), float(amp),
                             *[ptr(a) for a in arrs], ptr(pid), ptr(w))
    return Particles(*arrs, pid=pid, w=w)


DENSITY_FLAT_SOBOL, DENSITY_UNIFORM, DENSITY_SINUSOIDAL = 0, 3, 4


def sobol_dirs():
    """[30][3] direction integers of the reference's Sobol generator."""
    V = np.zeros((30, 3), np.uint32)
    lib().orc_sobol_dirs(ptr(V))
    return V


def sobol_points(n):
    r = np.zeros((int(n), 3))
    if lib().orc_sobol_points(int(n), ptr(r)) != 0:
        raise ValueError("the Sobol sequence needs more than 30 bits")
    return r


def init_density(g, kind, Np, seed, v0, vT, pert=0, amp=0.0):
    """The reference's density placements (flat-sobol, uniform, sinusoidal)."""
    n = int(lib().orc_density_count(int(kind), ptr(g.N), int(Np)))
    arrs = [np.zeros(max(n, 1)) for _ in range(6)]
    pid = np.zeros(max(n, 1), np.uint32)
    w = np.zeros(max(n, 1)) if pert else None
    v0, vT = _c(v0), _c(vT)
    rc = lib().orc_init_density(int(kind), ptr(g.N), ptr(g.p0), ptr(g.p1), ptr(g.L), ptr(g.d), int(Np), int(seed),
                                ptr(v0), ptr(vT), int(pert), float(amp), *[ptr(a) for a in arrs], ptr(pid), ptr(w))
    if rc < 0:
        raise ValueError(f"orc_init_density failed ({rc})")
    return Particles(*[a[:n] for a in arrs], pid=pid[:n], w=None if w is None else w[:n])


def pc_apply(g, K, n, r, pc=PC_FFT, lhs_type=LHS_FULL, degree=CHEBY_DEGREE):
    z = np.zeros(g.nc)
    K, n, r = _c(K), _c(n), _c(r)
    lib().orc_pc_apply_deg(pc, lhs_type, ptr(g.N), ptr(g.d), ptr(K), ptr(n), ptr(r), ptr(z), int(degree))
    return z


def solve(g, K, n, b, phi0=None, ksp=KSP_GMRES, pc=PC_FFT, lhs_type=LHS_FULL,
          rtol=1e-10, atol=0.0, max_it=2000, restart=30):
    phi = np.zeros(g.nc) if phi0 is None else _c(phi0).copy()
    info = np.zeros(2, I64)
    res = np.zeros(1)
    K, n, b = _c(K), _c(n), _c(b)
    lib().orc_solve(ksp, pc, lhs_type, ptr(g.N), ptr(g.d), ptr(K), ptr(n), ptr(b),
                    ptr(phi), rtol, atol, int(max_it), int(restart), ptr(info), ptr(res))
    return phi, int(info[0]), bool(info[1]), float(res[0])


def mg_levels(N):
    """Grid sizes of the multigrid levels (x and y halved per level)."""
    out = np.zeros((8, 3), I64)
    L = lib().orc_mg_levels(ptr(_c(N, I64)), ptr(out))
    return [tuple(int(v) for v in out[l]) for l in range(L)]


def mg_restrict(N, f):
    N = _c(N, I64)
    c = np.zeros(int(N[0] // 2 * (N[1] // 2) * N[2]))
    lib().orc_mg_restrict(ptr(N), ptr(_c(f)), ptr(c))
    return c


def mg_prolong_add(N, c, f):
    N = _c(N, I64)
    out = _c(f).copy()
    lib().orc_mg_prolong_add(ptr(N), ptr(_c(c)), ptr(out))
    return out


def set_threads(n):
    """Host threads of the oracle's independent loops (no result depends on
    them: every sum keeps its serial order)."""
    lib().orc_set_threads(int(n))


def get_threads():
    return int(lib().orc_get_threads())


def fft3d(g, data, inverse=False):
    buf = np.zeros(2 * g.nc)
    buf[0::2] = np.real(data).ravel()
    buf[1::2] = np.imag(data).ravel()
    lib().orc_fft3d(ptr(g.N), ptr(buf), int(inverse))
    return (buf[0::2] + 1j * buf[1::2]).reshape(np.shape(data))


def step(g, P_, parts, phi, S, E0):
    """One oracle step for given phi: push, stable sort, deposit.
    Returns (sorted particles, keys, moments, S_next, stats)."""
    ex, ey, ez = efield(g, E0, phi)
    work = parts.copy()
    key, Sn, stats = push(P_, ex, ey, ez, work, S)
    perm, offs = sort_perm(g.nc, key)
    out = work.permute(perm)
    mom = deposit_sorted(g, out, offs)
    return out, key[perm], mom, Sn, stats
