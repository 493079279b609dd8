"""chy_diagnostics_noalloc, the allocation-free variant of chy_diagnostics
(its block partials in a static device array, its output in caller device
memory): two calls in a row through the C ABI, the second on fewer ions, each
bitwise equal to chy_diagnostics on the same state.  The device output array
is allocated here, by the test, through the HIP runtime."""
import ctypes as C

import numpy as np
import pytest

import chybrid
from gpu_cases import context, opts

pytestmark = pytest.mark.gpu


def _hip():
    chybrid.lib()  # the HIP runtime the library uses is then loaded
    h = None
    for name in ("libamdhip64.so", "libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"):
        try:
            h = C.CDLL(name)
            break
        except OSError:
            continue
    assert h is not None, "no HIP runtime library"
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    return h


def _noalloc(ctx, hip, dptr):
    rc = chybrid.lib().chy_diagnostics_noalloc(ctx._p, dptr)
    assert rc == 0, chybrid.lib().chy_last_error(ctx._p)
    assert hip.hipStreamSynchronize(chybrid.lib().chy_stream(ctx._p)) == 0
    out = np.zeros(6)
    assert hip.hipMemcpy(out.ctypes.data, dptr, 48, 2) == 0  # device to host
    return out


def _ref(ctx):
    d = ctx.diagnostics()
    return np.array([d.weight, d.momentum[0], d.momentum[1], d.momentum[2], d.kinetic_energy, d.field_energy])


def test_noalloc_diagnostics_equal_the_allocating_entry_point():
    print("no-alloc diagnostics: 12 x 10 x 8 grid, 6 ions/cell, FP64, then a third of the ions", flush=True)
    hip = _hip()
    dptr = C.c_void_p()
    assert hip.hipMalloc(C.byref(dptr), 48) == 0
    try:
        args = opts((12, 10, 8), extra="-seed 5 -ppc 6 -ksp_rtol 1e-10 --perturbation uniform "
                                        "--perturbation-amplitude 0.01")
        with context(args) as ctx:
            ctx.init_particles()
            ctx.collect_moments(download=False)
            ctx.solve_potential()
            ctx.step()
            a = _noalloc(ctx, hip, dptr)
            ra = _ref(ctx)
            assert np.array_equal(a, ra), (a, ra)
            assert a[4] > 0 and a[5] > 0
            # second call: fewer ions (a third of them re-uploaded)
            d = ctx.get_particles()
            n = len(d["x"]) // 3
            ctx.set_particles(d["x"][:n], d["y"][:n], d["z"][:n], d["vx"][:n], d["vy"][:n], d["vz"][:n],
                              w=d["w"][:n], pid=d["pid"][:n])
            b = _noalloc(ctx, hip, dptr)
            rb = _ref(ctx)
            assert np.array_equal(b, rb), (b, rb)
            assert b[0] < a[0] and ctx.num_particles() == n
            assert chybrid.lib().chy_diagnostics_noalloc(ctx._p, None) == chybrid.ERR_ARG_WRONG
    finally:
        hip.hipFree(dptr)
