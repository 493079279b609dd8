"""bench.py's byte models and its reading of the committed traffic summary:
the roofline's `traffic` is the summary's scaled FETCH + WRITE column (not a
neighbouring column), and the deposit's partials are counted at the size of
the layout the library uses for the grid."""
import os
import re
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_traffic_is_the_scaled_fetch_plus_write_column():
    path = os.path.join(ROOT, "profiles", "hbm_traffic_256cubed_ppc16.md")
    rows = {}
    fs = ws = None
    for line in open(path):
        m = re.search(r"FETCH_SIZE is scaled by ([0-9.]+) and WRITE_SIZE by ([0-9]+\.[0-9]+)", line)
        if m:
            fs, ws = float(m.group(1)), float(m.group(2))
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if line.startswith("| ") and len(cells) == 7 and cells[0] not in ("timer", "---"):
            rows[cells[0]] = cells
    assert fs == 2.0 and ws == 1.0
    assert {"push", "sort_reorder", "deposit_nodes"} <= set(rows)
    for timer, c in rows.items():
        got = bench.measured_traffic(timer, 256, 16)
        # (the factors are printed rounded to 3 decimals: 1048576 / 524301 KiB)
        expect = (fs * float(c[2]) + ws * float(c[3])) * 1024
        assert abs(got - expect) <= 1e-4 * expect, (timer, got, expect)
        # the request-size columns measure the same bytes independently
        req = float(c[5]) + float(c[6])
        assert abs(got - req) <= 0.01 * got, (timer, got, req)
    assert bench.measured_traffic("no_such_timer", 256, 16) is None
    assert bench.measured_traffic("push", 64, 16) is None


def test_deposit_partials_follow_the_layout():
    np_, nc = 16 * 256 ** 3, 256 ** 3
    # mixed on a 256-wide grid: x-combined, 16 values per cell plus 16 per 32-cell block
    xc = bench.algorithmic_bytes("deposit_nodes", "mixed", np_, nc, nx=256)
    assert xc == nc * (16 * 4 + 16 * 4 / 32 + 4 * 4)
    # a width that is not a multiple of 32, and FP64: 32 values per cell
    assert bench.algorithmic_bytes("deposit_nodes", "mixed", np_, nc, nx=48) == nc * 36 * 4
    assert bench.algorithmic_bytes("deposit_nodes", "fp64", np_, nc, nx=256) == nc * 36 * 8
    gx = bench.algorithmic_bytes("sort_reorder", "mixed", np_, nc, nx=256)
    g = bench.algorithmic_bytes("sort_reorder", "mixed", np_, nc, nx=48)
    assert g - gx == nc * (32 * 4 - 16.5 * 4)


def test_preconditioner_bytes_only_for_fft2z():
    nc = 256 ** 3
    saved = list(bench.SOLVER_OPTS)
    try:
        bench.SOLVER_OPTS[:] = []
        # B along z: the residual kept by the column-sum pass, the 2-D
        # correction inside the z-line solve
        assert bench.algorithmic_bytes("pc", "mixed", 0, nc) == nc * 26 * 4
        bench.SOLVER_OPTS[:] = ["-B0x", "1e-5"]
        assert bench.algorithmic_bytes("pc", "mixed", 0, nc) == nc * 30 * 4
        bench.SOLVER_OPTS[:] = ["-pc_type", "mg"]
        assert bench.algorithmic_bytes("pc", "mixed", 0, nc) is None
    finally:
        bench.SOLVER_OPTS[:] = saved
