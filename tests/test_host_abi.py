"""CPU tests of the C ABI library: it loads, exports every entry point that
include/chybrid.h declares, and its host-side options database reproduces the
reference's flag semantics (parameters.c:28-556) and derived quantities."""
import os
import re

import numpy as np
import pytest

import chybrid

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CFG0 = ("-Nx 32 -Ny 32 -Nz 32 -x1 1.6 -y1 1.6 -z1 1.6 -Np 262144 -mi 4.9816e-26 -mn 4.8156e-26 "
        "-Ti 300 -Tn 300 -B0z 5e-5 -E0x 0.05 -nui 1e3 -nue 1e5 -dt 1e-5 -Nt 100 -seed 0 -ppc 8 "
        "--perturbation sinusoidal-x --perturbation-amplitude 0.01 -precision fp64").split()


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "chybrid.h")).read()
    return sorted(set(re.findall(r"\b(chy_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = chybrid.lib()
    syms = declared_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(chybrid.SIGNATURES), set(syms) ^ set(chybrid.SIGNATURES)


def test_params_struct_layout_defaults():
    p = chybrid.Params()
    assert chybrid.lib().chy_params_default(chybrid.C.byref(p)) == 0
    # fields at both ends of the struct: a layout mismatch garbles the tail
    assert list(p.N) == [-1, -1, -1]
    assert p.output == b"results"
    assert list(p.flux_scale) == [1.0, 1.0, 1.0]
    assert p.ksp_rtol == 1e-5 and p.ksp_restart == 30 and p.output_interval == 1
    assert p.qi == 1.6022e-19 and p.mi == 1.6726e-27 and p.qe == -1.6022e-19


def test_config0_derived_quantities():
    p, msg = chybrid.params_from_options(CFG0)
    assert list(p.N) == [32, 32, 32] and p.Np == 262144 and p.ppc == 8
    assert np.allclose(list(p.d), 0.05, rtol=1e-15)
    assert p.viT[0] == pytest.approx(135.93, abs=0.01)          # sqrt(2 kB T / m) / 3
    assert p.viT_r == pytest.approx(np.sqrt(3) * 135.93, abs=0.02)
    assert p.vnT == pytest.approx(414.76, abs=0.01)
    assert p.Omega_i[2] == pytest.approx(160.8, abs=0.1)
    assert p.Omega_e[2] == pytest.approx(8.794e6, rel=1e-4)
    assert p.kappa_e[2] == pytest.approx(87.94, abs=0.01)
    assert 1 + p.kappa_e[2] ** 2 == pytest.approx(7734.8, abs=0.5)
    assert p.perturbation == chybrid.PERT_SINUSOIDAL_X and p.perturbation_amplitude == 0.01
    assert p.Tn == 300 and p.mn == 4.8156e-26 and p.dt == 1e-5 and p.Nt == 100


def test_reference_defaults_and_fallbacks():
    p, _ = chybrid.params_from_options([])
    assert list(p.N) == [7, 7, 7] and p.Np == 343          # setup.c:9-11,43-45
    assert p.dt == 1.0 and p.mn == p.mi and p.Tn == p.Ti == 0.0
    assert p.lhs_type == chybrid.LHS_FULL and p.rhs_type == chybrid.RHS_FULL
    p, _ = chybrid.params_from_options("-nui 250".split())
    assert p.dt == 1.0 / 250                                  # parameters.c:412-418


def test_vector_flags_common_then_component():
    p, _ = chybrid.params_from_options("-vi0 3 -vi0y 7 -viT 2 -viTz 5".split())
    assert list(p.vi0) == [3, 7, 3]
    assert list(p.viT) == [2, 2, 5]
    assert p.Ti == pytest.approx(0.5 * p.mi / 1.3807e-23 * (4 + 4 + 25))


def test_options_case_insensitive_last_wins_env_and_file(tmp_path, monkeypatch):
    f = tmp_path / "opts"
    f.write_text("# comment\n-Ny 12\n-E0y 0.5 # trailing\n")
    monkeypatch.setenv("PETSC_OPTIONS", "-nx 5 -Nz 9")
    p, _ = chybrid.params_from_options(["-NX", "6", "-Nx", "8", "-options_file", str(f)])
    assert list(p.N) == [8, 12, 9] and p.E0[1] == 0.5


def test_option_errors_mirror_reference_classes():
    with pytest.raises(chybrid.ChyError) as e:
        chybrid.params_from_options("--x0-bc periodic --x1-bc dirichlet".split())
    assert e.value.code == chybrid.ERR_ARG_INCOMP                # parameters.c:486
    with pytest.raises(chybrid.ChyError) as e:
        chybrid.params_from_options("--y0-bc neumann --y1-bc neumann".split())
    assert e.value.code == chybrid.ERR_ARG_WRONG                 # parameters.c:505
    with pytest.raises(chybrid.ChyError) as e:
        chybrid.params_from_options("--lhs-type bogus".split())
    assert e.value.code == chybrid.ERR_ARG_WRONG
    with pytest.raises(chybrid.ChyError):
        chybrid.params_from_options("-Nx 4 -Ny 4 -Nz 4 -Np 100 -ppc 2".split())
    with pytest.raises(chybrid.ChyError):
        chybrid.params_from_options("-Nx 2".split())


def test_density_type_options():
    """--density-type names (parameters.c:14-16, default flat-sobol); -ppc
    selects the per-cell lattice and refuses another explicit density type."""
    p, _ = chybrid.params_from_options([])
    assert p.density_type == chybrid.DENSITY_FLAT_SOBOL
    for name, v in (("flat-sobol", 0), ("flat-reverse", 1), ("flat-normal", 2), ("uniform", 3), ("sinusoidal", 4),
                    ("gaussian", 5), ("lattice", chybrid.DENSITY_LATTICE)):
        p, _ = chybrid.params_from_options(["--density-type", name])
        assert p.density_type == v, name
    p, _ = chybrid.params_from_options("-Nx 4 -Ny 4 -Nz 4 -ppc 2".split())
    assert p.density_type == chybrid.DENSITY_LATTICE and p.Np == 128
    p, _ = chybrid.params_from_options("-Nx 4 -Ny 4 -Nz 4 -ppc 2 --density-type lattice".split())
    assert p.density_type == chybrid.DENSITY_LATTICE
    with pytest.raises(chybrid.ChyError) as e:
        chybrid.params_from_options("-Nx 4 -Ny 4 -Nz 4 -ppc 2 --density-type sinusoidal".split())
    assert e.value.code == chybrid.ERR_ARG_INCOMP
    with pytest.raises(chybrid.ChyError) as e:
        chybrid.params_from_options("--density-type bogus".split())
    assert e.value.code == chybrid.ERR_ARG_WRONG
    # any -Np without -ppc
    p, _ = chybrid.params_from_options("-Nx 4 -Ny 4 -Nz 4 -Np 1001".split())
    assert p.Np == 1001 and p.density_type == chybrid.DENSITY_FLAT_SOBOL


def test_solver_option_mapping():
    p, msg = chybrid.params_from_options("-ksp_type bcgs -pc_type ilu -ksp_rtol 1e-9 -ksp_max_it 77 "
                                         "-ksp_gmres_restart 12 -ksp_monitor_short -precision mixed".split())
    assert p.ksp_type == chybrid.KSP_BCGS and p.pc_type == chybrid.PC_FFT2Z and "ilu" in msg
    assert p.ksp_rtol == 1e-9 and p.ksp_max_it == 77 and p.ksp_restart == 12 and p.ksp_monitor == 1
    assert p.precision == chybrid.PREC_MIXED
    for name, pc in (("none", chybrid.PC_NONE), ("jacobi", chybrid.PC_JACOBI), ("fft", chybrid.PC_FFT),
                     ("fft2", chybrid.PC_FFT2), ("fft2z", chybrid.PC_FFT2Z), ("zline", chybrid.PC_ZLINE),
                     ("bjacobi", chybrid.PC_ZLINE),
                     ("cheby", chybrid.PC_CHEBY), ("chebyshev", chybrid.PC_CHEBY),
                     # PETSc's multigrid preconditioners (run-case gamg-basic, basic-physics)
                     ("mg", chybrid.PC_MG), ("gamg", chybrid.PC_MG), ("hypre", chybrid.PC_MG),
                     ("ml", chybrid.PC_MG)):
        p, _ = chybrid.params_from_options(["-pc_type", name])
        assert p.pc_type == pc, name
    p, _ = chybrid.params_from_options("-pc_cheby_degree 7 -ksp_inner_rtol 1e-5".split())
    assert p.pc_cheby_degree == 7 and p.ksp_inner_rtol == 1e-5
    p, _ = chybrid.params_from_options([])
    assert p.pc_type == chybrid.PC_FFT2Z and p.pc_cheby_degree == 4 and p.ksp_inner_rtol == 2e-5
    # classical Gram-Schmidt without refinement unless asked (PETSc's default)
    assert p.ksp_gmres_cgs_refinement == chybrid.CGS_REFINE_NEVER
    for name, v in (("REFINE_NEVER", chybrid.CGS_REFINE_NEVER), ("refine_ifneeded", chybrid.CGS_REFINE_IFNEEDED),
                    ("REFINE_ALWAYS", chybrid.CGS_REFINE_ALWAYS)):
        p, _ = chybrid.params_from_options(["-ksp_gmres_cgs_refinement_type", name])
        assert p.ksp_gmres_cgs_refinement == v, name
    with pytest.raises(chybrid.ChyError):
        chybrid.params_from_options("-ksp_gmres_cgs_refinement_type sometimes".split())
    assert p.ksp_initial_guess == chybrid.GUESS_PREVIOUS
    for name, v in (("previous", chybrid.GUESS_PREVIOUS), ("zero", chybrid.GUESS_ZERO),
                    ("extrapolate", chybrid.GUESS_EXTRAPOLATE)):
        p, _ = chybrid.params_from_options(["-ksp_initial_guess", name])
        assert p.ksp_initial_guess == v, name
    # the reference's abort at the 10th refused scattering of a step, as a flag
    assert chybrid.params_from_options([])[0].collision_refusal_limit == 10
    assert chybrid.params_from_options("--collision-refusal-limit 0".split())[0].collision_refusal_limit == 0
    assert chybrid.params_from_options("--collision-refusal-limit 1000".split())[0].collision_refusal_limit == 1000
    with pytest.raises(chybrid.ChyError):
        chybrid.params_from_options("--collision-refusal-limit -1".split())
    with pytest.raises(chybrid.ChyError):
        chybrid.params_from_options("-pc_cheby_degree 0".split())


def test_options_txt_is_parseable(tmp_path):
    path = str(tmp_path / "options.txt")
    p, _ = chybrid.params_from_options(CFG0)
    chybrid.write_options_txt(p, CFG0, path)
    text = open(path).read()
    assert "Parameter Values" in text and "kappa_ez = 87.9" in text
    # scripts/tools.py reads "-key value" lines of the option table
    opts = dict(line.split(None, 1) for line in text.splitlines() if line.startswith("-") and " " in line)
    assert opts["-Nx"] == "32" and opts["-B0z"] == "5e-5"


# ---- solver input (LoadVlasovQuantities, fileio.c:6-70) ----------------------
SMALL = ("-Nx 4 -Ny 5 -Nz 6 -x1 0.2 -y1 0.25 -z1 0.3 -mi 4.9816e-26 -mn 4.8156e-26 -Ti 300 -Tn 300 "
         "-B0z 5e-5 -E0x 0.05 -nui 1e3 -nue 1e5 -dt 1e-5").split()


def _h5_lib():
    import ctypes as C
    for name in ("libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so"):
        try:
            h = C.CDLL(name)
        except OSError:
            continue
        for f in ("H5Fcreate", "H5Screate_simple", "H5Dcreate2"):
            getattr(h, f).restype = C.c_int64
        h.H5Fcreate.argtypes = [C.c_char_p, C.c_uint, C.c_int64, C.c_int64]
        h.H5Screate_simple.argtypes = [C.c_int, C.c_void_p, C.c_void_p]
        h.H5Dcreate2.argtypes = [C.c_int64, C.c_char_p, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int64]
        h.H5Dwrite.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_void_p]
        for f in ("H5Dclose", "H5Sclose", "H5Fclose"):
            getattr(h, f).argtypes = [C.c_int64]
        h.H5open()
        return h
    return None


def _h5_write(h, path, name, arr):
    """One f64 dataset `name` of arr's shape in a new HDF5 file."""
    import ctypes as C
    dbl = C.c_int64.in_dll(h, "H5T_NATIVE_DOUBLE_g").value
    f = h.H5Fcreate(str(path).encode(), 2, 0, 0)  # H5F_ACC_TRUNC
    dims = np.array(arr.shape, dtype=np.uint64)
    sp = h.H5Screate_simple(arr.ndim, dims.ctypes.data, None)
    ds = h.H5Dcreate2(f, name.encode(), dbl, sp, 0, 0, 0)
    assert f >= 0 and sp >= 0 and ds >= 0
    assert h.H5Dwrite(ds, dbl, 0, 0, 0, np.ascontiguousarray(arr, dtype=np.float64).ctypes.data) >= 0
    h.H5Dclose(ds)
    h.H5Sclose(sp)
    h.H5Fclose(f)


def _read_vlasov(prm, path):
    nc = prm.N[0] * prm.N[1] * prm.N[2]
    vl = np.full((nc, 4), np.nan)
    msg = chybrid.C.create_string_buffer(512)
    rc = chybrid.lib().chy_read_vlasov(chybrid.C.byref(prm), str(path).encode(), vl.ctypes.data, msg, len(msg))
    return rc, vl, msg.value.decode()


def test_read_vlasov_raw_density_makes_the_fluxes(tmp_path):
    """fileio.c:34-53: the density is read and each flux is its flux scale
    times the density (VecAXPY onto zero), so all four are exact."""
    prm, _ = chybrid.params_from_options(SMALL + "--x-flux-scale 3 --y-flux-scale -0.5 --z-flux-scale 7".split())
    n = np.random.default_rng(3).uniform(1, 2, 4 * 5 * 6)
    n.tofile(tmp_path / "n.f64")
    rc, vl, msg = _read_vlasov(prm, tmp_path / "n.f64")
    assert rc == 0, msg
    assert np.array_equal(vl[:, 0], n)
    for a, s in enumerate((3.0, -0.5, 7.0)):
        assert np.array_equal(vl[:, 1 + a], s * n)
    # a density of the wrong size, or no file, is refused with a message
    n[:-1].tofile(tmp_path / "short.f64")
    rc, _, msg = _read_vlasov(prm, tmp_path / "short.f64")
    assert rc == chybrid.ERR_ARG_WRONG and "size" in msg
    rc, _, msg = _read_vlasov(prm, tmp_path / "none.f64")
    assert rc == chybrid.ERR_ARG_WRONG and "none.f64" in msg


def test_read_vlasov_hdf5_density_kji(tmp_path):
    """The reference loads the Vec named "density-kji" (fileio.c:39-41), shape
    (Nz, Ny, Nx) with x fastest."""
    h = _h5_lib()
    if h is None:
        pytest.skip("no libhdf5 to write the input file")
    prm, _ = chybrid.params_from_options(SMALL)
    n = np.random.default_rng(4).uniform(1, 2, (6, 5, 4))
    _h5_write(h, tmp_path / "in.hdf", "density-kji", n)
    rc, vl, msg = _read_vlasov(prm, tmp_path / "in.hdf")
    assert rc == 0, msg
    assert np.array_equal(vl[:, 0], n.ravel())
    assert np.array_equal(vl[:, 1], n.ravel())  # flux scales default to 1
    _h5_write(h, tmp_path / "other.hdf", "density", n)
    rc, _, msg = _read_vlasov(prm, tmp_path / "other.hdf")
    assert rc == chybrid.ERR_ARG_WRONG and "density-kji" in msg


def test_entry_points_refuse_a_null_context():
    """Every context call returns PETSC_ERR_ARG_WRONG for a null context
    without touching a device (no exception or abort crosses the boundary)."""
    L = chybrid.lib()
    assert L.chy_output_write_async(None, b"x", b"") == chybrid.ERR_ARG_WRONG
    assert L.chy_output_wait(None) == chybrid.ERR_ARG_WRONG
    assert L.chy_get_rhs(None, None) == chybrid.ERR_ARG_WRONG
    assert L.chy_step(None, None) == chybrid.ERR_ARG_WRONG
    prm, _ = chybrid.params_from_options(SMALL)
    msg = chybrid.C.create_string_buffer(64)
    assert L.chy_read_vlasov(chybrid.C.byref(prm), None, None, msg, len(msg)) == chybrid.ERR_ARG_WRONG
