"""ctypes binding of libchybrid.so, the C ABI declared in include/chybrid.h.

This is the host-side mirror of the reference's step interface
(UpdateVelocities, UpdatePositions, CollectVlasovQuantities, ComputePotential
and the KSP callbacks ComputeRHS / ComputeLHS, reference particles.h:7-29,
potential.h:7-13).  It carries no numerics of its own: every call goes to the
HIP library, and loading fails loudly when the library is not built.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "lib", "libchybrid.so")

# error codes (negated PETSc classes)
OK = 0
ERR_MEM, ERR_SUP, ERR_SIG, ERR_ARG_WRONG, ERR_ARG_RANGE = -55, -56, -59, -62, -63
ERR_FP, ERR_ARG_INCOMP, ERR_LIB, ERR_NOT_CONVERGED = -72, -75, -76, -91

LHS_IDENTITY, LHS_LAPLACIAN, LHS_FULL = 0, 1, 2
RHS_CONSTANT, RHS_SINUSOIDAL, RHS_FULL = 0, 1, 2
(DENSITY_FLAT_SOBOL, DENSITY_FLAT_REVERSE, DENSITY_FLAT_NORMAL, DENSITY_UNIFORM, DENSITY_SINUSOIDAL,
 DENSITY_GAUSSIAN, DENSITY_LATTICE) = range(7)
PREC_FP64, PREC_MIXED = 0, 1
KSP_GMRES, KSP_BCGS = 0, 1
PC_NONE, PC_JACOBI, PC_FFT, PC_ZLINE, PC_CHEBY, PC_FFT2, PC_FFT2Z, PC_MG = 0, 1, 2, 3, 4, 5, 6, 7
CGS_REFINE_NEVER, CGS_REFINE_IFNEEDED, CGS_REFINE_ALWAYS = 0, 1, 2
GUESS_PREVIOUS, GUESS_ZERO, GUESS_EXTRAPOLATE = 0, 1, 2
PERT_NONE, PERT_SINUSOIDAL_X, PERT_UNIFORM, PERT_GAUSSIAN = 0, 1, 2, 3

D3 = C.c_double * 3


class Params(C.Structure):
    """Mirror of chy_params (include/chybrid.h)."""
    _fields_ = [
        ("N", C.c_int64 * 3), ("p0", D3), ("p1", D3), ("L", D3), ("d", D3),
        ("bc", (C.c_int32 * 2) * 3),
        ("qe", C.c_double), ("me", C.c_double), ("nue", C.c_double), ("Te", C.c_double),
        ("gamma_e", C.c_double), ("ve0", D3), ("veT", D3), ("veT_r", C.c_double),
        ("Omega_e", D3), ("kappa_e", D3),
        ("qi", C.c_double), ("mi", C.c_double), ("nui", C.c_double), ("Ti", C.c_double),
        ("vi0", D3), ("viT", D3), ("viT_r", C.c_double), ("Omega_i", D3), ("kappa_i", D3),
        ("mn", C.c_double), ("Tn", C.c_double), ("vnT", C.c_double), ("vn0", D3), ("vn0_r", C.c_double),
        ("B0", D3), ("E0", D3), ("Np", C.c_int64),
        ("Nt", C.c_int64), ("dt", C.c_double),
        ("lhs_type", C.c_int32), ("rhs_type", C.c_int32), ("density_type", C.c_int32),
        ("view_lhs", C.c_int32),
        ("ksp_type", C.c_int32), ("pc_type", C.c_int32),
        ("ksp_rtol", C.c_double), ("ksp_atol", C.c_double),
        ("ksp_max_it", C.c_int64), ("ksp_restart", C.c_int64),
        ("ksp_monitor", C.c_int32), ("ksp_converged_reason", C.c_int32),
        ("ksp_error_if_not_converged", C.c_int32), ("pc_cheby_degree", C.c_int32),
        ("ksp_inner_rtol", C.c_double), ("ksp_gmres_cgs_refinement", C.c_int32), ("ksp_initial_guess", C.c_int32),
        ("seed", C.c_uint64), ("ppc", C.c_int64), ("perturbation", C.c_int32),
        ("perturbation_amplitude", C.c_double), ("weighted", C.c_int32),
        ("precision", C.c_int32), ("output_interval", C.c_int64), ("collision_refusal_limit", C.c_int64),
        ("flux_scale", D3), ("input", C.c_char * 4096), ("output", C.c_char * 4096),
        ("diagnostics_interval", C.c_int64),
    ]


class CollisionStats(C.Structure):
    _fields_ = [("accepted", C.c_int64), ("refused", C.c_int64), ("capped", C.c_int64), ("Nc", C.c_int64),
                ("expected", C.c_double), ("variance", C.c_double)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int64), ("outer", C.c_int64), ("residual", C.c_double),
                ("converged", C.c_int32), ("reason", C.c_int32), ("ms", C.c_double)]


class Diagnostics(C.Structure):
    _fields_ = [("ions", C.c_int64), ("step", C.c_int64), ("weight", C.c_double), ("momentum", D3),
                ("kinetic_energy", C.c_double), ("field_energy", C.c_double), ("valid", C.c_int32)]


class StepStats(C.Structure):
    _fields_ = [("collisions", CollisionStats), ("solve", SolveStats),
                ("ms_particles", C.c_double), ("ms_total", C.c_double), ("diag", Diagnostics)]


# every entry point of include/chybrid.h with its ctypes signature
P = C.c_void_p
SIGNATURES = {
    "chy_params_default": (C.c_int, [P]),
    "chy_params_from_options": (C.c_int, [C.c_int, P, P, C.c_char_p, C.c_size_t]),
    "chy_params_derive": (C.c_int, [P, C.c_char_p, C.c_size_t]),
    "chy_write_options_txt": (C.c_int, [P, C.c_int, P, C.c_char_p]),
    "chy_create": (C.c_int, [P, C.c_int64, P]),
    "chy_destroy": (C.c_int, [P]),
    "chy_last_error": (C.c_char_p, [P]),
    "chy_sync": (C.c_int, [P]),
    "chy_stream": (P, [P]),
    "chy_get_params": (C.c_int, [P, P]),
    "chy_num_particles": (C.c_int64, [P]),
    "chy_set_step": (C.c_int, [P, C.c_int64]),
    "chy_set_tolerances": (C.c_int, [P, C.c_double, C.c_double, C.c_int64]),
    "chy_set_particles": (C.c_int, [P, C.c_int64, P, P, P, P]),
    "chy_get_particles": (C.c_int, [P, P, P, P, P]),
    "chy_init_particles": (C.c_int, [P]),
    "chy_update_velocities": (C.c_int, [P, P]),
    "chy_update_positions": (C.c_int, [P]),
    "chy_collect_moments": (C.c_int, [P, P]),
    "chy_set_moments": (C.c_int, [P, P]),
    "chy_get_moments": (C.c_int, [P, P]),
    "chy_output_write": (C.c_int, [P, C.c_char_p, C.c_char_p]),
    "chy_output_write_async": (C.c_int, [P, C.c_char_p, C.c_char_p]),
    "chy_output_wait": (C.c_int, [P]),
    "chy_read_vlasov": (C.c_int, [P, C.c_char_p, P, C.c_char_p, C.c_size_t]),
    "chy_compute_rhs": (C.c_int, [P, P]),
    "chy_lhs_apply": (C.c_int, [P, P, P]),
    "chy_lhs_apply_transpose": (C.c_int, [P, P, P]),
    "chy_solve_potential": (C.c_int, [P, P, P]),
    "chy_solve": (C.c_int, [P, P, P, P]),
    "chy_get_potential": (C.c_int, [P, P]),
    "chy_set_potential": (C.c_int, [P, P]),
    "chy_get_rhs": (C.c_int, [P, P]),
    "chy_step": (C.c_int, [P, P]),
    "chy_timing_enable": (C.c_int, [P, C.c_int]),
    "chy_timing_get": (C.c_int, [P, C.c_char_p, P, P]),
    "chy_timing_reset": (C.c_int, [P]),
    "chy_copy_kernel": (C.c_int, [P, C.c_int64, C.c_int]),
    "chy_adjoint_forward": (C.c_int, [P, C.c_int64, C.c_int64, P]),
    "chy_pc_apply": (C.c_int, [P, P, P]),
    "chy_timing_select": (C.c_int, [P, C.c_char_p]),
    "chy_adjoint_backward": (C.c_int, [P, P, P]),
    "chy_diagnostics": (C.c_int, [P, P]),
    "chy_diagnostics_noalloc": (C.c_int, [P, P]),
}

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"libchybrid.so is not built ({LIBPATH}); run __graft_entry__.build()")
        L = C.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


class ChyError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"chybrid error {code}: {msg}")
        self.code = code


def _argv(args):
    enc = [a.encode() for a in args]
    arr = (C.c_char_p * len(enc))(*enc)
    return arr, enc


def params_from_options(args):
    """Parse reference-style options (a list of strings) into Params."""
    prm = Params()
    err = C.create_string_buffer(4096)
    arr, keep = _argv(args)
    rc = lib().chy_params_from_options(len(args), C.cast(arr, P), C.byref(prm), err, len(err))
    if rc:
        raise ChyError(rc, err.value.decode())
    return prm, err.value.decode()


def write_options_txt(prm, args, path):
    arr, keep = _argv(args)
    rc = lib().chy_write_options_txt(C.byref(prm), len(args), C.cast(arr, P), path.encode())
    if rc:
        raise ChyError(rc, "options.txt")


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a):
    return None if a is None else a.ctypes.data


class Context:
    """One device context (chy_ctx); closes on del/close()."""

    def __init__(self, prm, np_capacity=0):
        self._p = C.c_void_p()
        rc = lib().chy_create(C.byref(prm), int(np_capacity), C.byref(self._p))
        if rc:
            raise ChyError(rc, "chy_create failed (no GPU, bad parameters or out of memory)")
        out = Params()
        lib().chy_get_params(self._p, C.byref(out))
        self.prm = out
        self.shape = (int(out.N[2]), int(out.N[1]), int(out.N[0]))
        self.nc = int(out.N[0] * out.N[1] * out.N[2])

    def _chk(self, rc):
        if rc:
            raise ChyError(rc, lib().chy_last_error(self._p).decode())

    def close(self):
        if self._p:
            lib().chy_destroy(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # particles --------------------------------------------------------------
    def set_particles(self, x, y, z, vx, vy, vz, w=None, pid=None):
        n = len(x)
        pos = np.zeros((n, 4))
        vel = np.zeros((n, 4))
        pos[:, 0], pos[:, 1], pos[:, 2] = x, y, z
        vel[:, 0], vel[:, 1], vel[:, 2] = vx, vy, vz
        w = None if w is None else _f64(w)
        pid = None if pid is None else np.ascontiguousarray(pid, dtype=np.uint32)
        self._chk(lib().chy_set_particles(self._p, n, _ptr(pos), _ptr(vel), _ptr(w), _ptr(pid)))

    def get_particles(self):
        n = int(lib().chy_num_particles(self._p))
        pos = np.zeros((n, 4))
        vel = np.zeros((n, 4))
        w = np.zeros(n)
        pid = np.zeros(n, np.uint32)
        self._chk(lib().chy_get_particles(self._p, _ptr(pos), _ptr(vel), _ptr(w), _ptr(pid)))
        return dict(x=pos[:, 0].copy(), y=pos[:, 1].copy(), z=pos[:, 2].copy(),
                    vx=vel[:, 0].copy(), vy=vel[:, 1].copy(), vz=vel[:, 2].copy(), w=w, pid=pid)

    def num_particles(self):
        return int(lib().chy_num_particles(self._p))

    def init_particles(self):
        self._chk(lib().chy_init_particles(self._p))

    def set_step(self, it):
        self._chk(lib().chy_set_step(self._p, int(it)))

    def set_tolerances(self, rtol=-1.0, atol=-1.0, max_it=-1):
        self._chk(lib().chy_set_tolerances(self._p, float(rtol), float(atol), int(max_it)))

    # hot path ---------------------------------------------------------------
    def update_velocities(self):
        st = CollisionStats()
        self._chk(lib().chy_update_velocities(self._p, C.byref(st)))
        return st

    def update_positions(self):
        self._chk(lib().chy_update_positions(self._p))

    def collect_moments(self, download=True):
        if not download:
            self._chk(lib().chy_collect_moments(self._p, None))
            return None
        v = np.zeros((self.nc, 4))
        self._chk(lib().chy_collect_moments(self._p, _ptr(v)))
        return v

    def set_moments(self, vlasov):
        v = _f64(vlasov).reshape(self.nc, 4)
        self._chk(lib().chy_set_moments(self._p, _ptr(v)))

    def get_moments(self):
        v = np.zeros((self.nc, 4))
        self._chk(lib().chy_get_moments(self._p, _ptr(v)))
        return v

    def output_write(self, stem, insert):
        """Returns 0 when HDF5 was written, ERR_SUP when raw files were written."""
        rc = lib().chy_output_write(self._p, stem.encode(), insert.encode())
        if rc not in (0, ERR_SUP):
            self._chk(rc)
        return rc

    def output_write_async(self, stem, insert):
        """Starts the same output behind the computation; output_wait() ends it."""
        self._chk(lib().chy_output_write_async(self._p, stem.encode(), insert.encode()))

    def output_wait(self):
        """Result of the output in flight: 0 (HDF5) or ERR_SUP (raw files)."""
        rc = lib().chy_output_wait(self._p)
        if rc not in (0, ERR_SUP):
            self._chk(rc)
        return rc

    def compute_rhs(self):
        b = np.zeros(self.nc)
        self._chk(lib().chy_compute_rhs(self._p, _ptr(b)))
        return b

    def lhs_apply(self, x, transpose=False):
        x = _f64(x)
        y = np.zeros(self.nc)
        f = lib().chy_lhs_apply_transpose if transpose else lib().chy_lhs_apply
        self._chk(f(self._p, _ptr(x), _ptr(y)))
        return y

    def solve_potential(self):
        phi = np.zeros(self.nc)
        st = SolveStats()
        self._chk(lib().chy_solve_potential(self._p, _ptr(phi), C.byref(st)))
        return phi, st

    def solve(self, b, phi0=None):
        b = _f64(b)
        phi = np.zeros(self.nc) if phi0 is None else _f64(phi0).copy()
        st = SolveStats()
        self._chk(lib().chy_solve(self._p, _ptr(b), _ptr(phi), C.byref(st)))
        return phi, st

    def pc_apply(self, r):
        r = _f64(r)
        z = np.zeros(self.nc)
        self._chk(lib().chy_pc_apply(self._p, _ptr(r), _ptr(z)))
        return z

    def get_potential(self):
        phi = np.zeros(self.nc)
        self._chk(lib().chy_get_potential(self._p, _ptr(phi)))
        return phi

    def set_potential(self, phi):
        phi = _f64(phi)
        self._chk(lib().chy_set_potential(self._p, _ptr(phi)))

    def get_rhs(self):
        b = np.zeros(self.nc)
        self._chk(lib().chy_get_rhs(self._p, _ptr(b)))
        return b

    def step(self):
        st = StepStats()
        self._chk(lib().chy_step(self._p, C.byref(st)))
        return st

    def diagnostics(self):
        """FP64 device sums: ions, weight, momentum, kinetic and field energy."""
        d = Diagnostics()
        self._chk(lib().chy_diagnostics(self._p, C.byref(d)))
        return d

    def sync(self):
        self._chk(lib().chy_sync(self._p))

    # timing -----------------------------------------------------------------
    def timing(self, on=True):
        self._chk(lib().chy_timing_enable(self._p, int(on)))

    def timing_get(self, name):
        t = C.c_double()
        n = C.c_int64()
        self._chk(lib().chy_timing_get(self._p, name.encode(), C.byref(t), C.byref(n)))
        return t.value, n.value

    def timing_select(self, names=None):
        """Record only the given timers (an iterable of names; None: all)."""
        arg = None if names is None else ",".join(names).encode()
        self._chk(lib().chy_timing_select(self._p, arg))

    def timing_reset(self):
        self._chk(lib().chy_timing_reset(self._p))

    # adjoint ---------------------------------------------------------------
    def adjoint_forward(self, nsteps, ckpt):
        """Runs the step map from the current ions; returns J = sum (n - mean n)^2."""
        J = C.c_double()
        self._chk(lib().chy_adjoint_forward(self._p, int(nsteps), int(ckpt), C.byref(J)))
        return J.value

    def adjoint_backward(self, want_delta=True):
        """(dJ/dE0[3], dJ/d delta per cell or None) of the last forward run."""
        dE = np.zeros(3)
        dd = np.zeros(self.nc) if want_delta else None
        self._chk(lib().chy_adjoint_backward(self._p, _ptr(dE), _ptr(dd)))
        return dE, dd

    def copy_kernel(self, nbytes, repeat=1):
        self._chk(lib().chy_copy_kernel(self._p, int(nbytes), int(repeat)))
