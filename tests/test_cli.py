"""
The lbm3d-perm / lbm2d-perm drop-ins (flow-anizotropy_amd/bin/): argument
handling, exit codes and CSV text of the reference executables
(LMB3d/lbm3d-perm.cpp:446-630, LMB2d/main.cpp:238-426).  The error paths run
on the CPU (they fail before any device work); full runs are GPU tests.
"""
import csv
import re
import subprocess
import sys

import numpy as np
import pytest

from conftest import GOLDEN, REPO, assert_printed_close, check_eigen_and_reciprocity, golden

BIN = REPO / "flow-anizotropy_amd" / "bin"


def run(prog, *args, cwd=None):
    p = subprocess.run([sys.executable, str(BIN / prog), *map(str, args)],
                       capture_output=True, text=True, cwd=cwd, timeout=600)
    return p.returncode, p.stdout, p.stderr


def test_usage_and_argument_errors(tmp_path):
    rc, _, err = run("lbm3d-perm")
    assert rc == 2 and err.startswith("Usage: lbm3d-perm")       # argc < 3 -> 2
    rc, _, err = run("lbm2d-perm")
    assert rc == 1 and err.startswith("Usage: lbm2d-perm")       # main.cpp:280
    rc, _, err = run("lbm3d-perm", "a.raw", "b.csv", "--bogus")
    assert rc == 2 and err.strip().endswith("lbm3d-perm: unknown option --bogus")
    rc, _, err = run("lbm3d-perm", "a.raw", "b.csv", "--lag")
    assert rc == 2 and "lbm3d-perm: missing value for --lag" in err
    rc, _, err = run("lbm2d-perm", "a.gif", "b.csv", "--lag", "0")
    assert rc == 2 and "lbm2d-perm: --lag and --avg-sample must be positive" in err


def test_structure_errors_are_exit_2(tmp_path):
    from flowperm import structio
    rc, _, err = run("lbm3d-perm", tmp_path / "missing.raw", tmp_path / "o.csv", "--quiet")
    assert rc == 2 and "lbm3d-perm: cannot open" in err
    structio.save_raw(tmp_path / "solid.raw", np.ones((3, 3, 3), np.uint8))
    rc, _, err = run("lbm3d-perm", tmp_path / "solid.raw", tmp_path / "o.csv", "--quiet")
    assert rc == 2 and err.strip() == "lbm3d-perm: structure has no pore space"
    raw = (tmp_path / "solid.raw").read_bytes()
    (tmp_path / "trunc.raw").write_bytes(raw[:-2])
    rc, _, err = run("lbm3d-perm", tmp_path / "trunc.raw", tmp_path / "o.csv", "--quiet")
    assert rc == 2 and "truncated payload" in err
    assert not (tmp_path / "o.csv").exists()                      # no row on a hard error


def test_numeric_flags_follow_strtol_strtod():
    from flowperm.cli import _strtod, _strtol
    assert _strtol("2000") == 2000 and _strtol("12abc") == 12 and _strtol("x") == 0
    assert _strtol("-5") == -5 and _strtol(" 7") == 7
    assert _strtod("1e-6") == 1e-6 and _strtod("40") == 40.0 and _strtod("2.5x") == 2.5
    assert _strtod("abc") == 0.0


@pytest.mark.gpu
def test_lbm3d_row_matches_the_reference_row(tmp_path):
    """the drop-in on a reference fixture: header text identical, values at
    the reference's printed precision (1e-10 x trace floor), same steps and
    exit code; a second run appends a row without a header"""
    name = "sph16_s0"
    mask, row, meta = golden(name)
    argv = meta["argv"][3:]
    out = tmp_path / "o.csv"
    rc, stdout, err = run("lbm3d-perm", GOLDEN / f"{name}.raw", out, *argv)
    print(stdout, err, flush=True)
    assert rc == meta["returncode"]
    ref_lines = (GOLDEN / f"{name}.csv").read_text().splitlines()
    lines = out.read_text().splitlines()
    assert lines[0] == ref_lines[0]
    ours = dict(zip(lines[0].split(","), next(csv.reader([lines[1]]))))
    refd = dict(zip(ref_lines[0].split(","), next(csv.reader([ref_lines[1]]))))
    trace = abs(float(refd["K_xx"]) + float(refd["K_yy"]) + float(refd["K_zz"]))
    for k, v in refd.items():
        if k == "filename":
            continue
        if k.startswith(("steps_", "conv_", "nx", "ny", "nz")):
            assert ours[k] == v, k
        elif k.startswith(("K_", "k_", "k1", "k2", "k3")):
            assert_printed_close(float(ours[k]), float(v), 1e-10 * trace, k)
    assert check_eigen_and_reciprocity(ours.__getitem__, refd, trace, name) >= 1
    rc2, _, _ = run("lbm3d-perm", GOLDEN / f"{name}.raw", out, *argv)
    assert rc2 == rc and len(out.read_text().splitlines()) == 3


def test_progress_line_text():
    """the per-check progress lines in the reference's text: 3D from a line of
    the reference binary's own stdout (fixture ell_x130y8z6_fast_s7), 2D as
    LMB2d/main.cpp:167-172 prints with ostream's default 6 digits"""
    from flowperm.cli import _line2d, _line3d
    from conftest import golden
    _, _, meta = golden("ell_x130y8z6_fast_s7")
    assert meta["stdout"][0] == _line3d(0, {"step": 66000, "q": [1.38345e-06, 2.97032e-07, 1.54424e-07],
                                            "ratio": [1.14254, 0.0], "hold": 0})
    assert _line3d(3, {"step": 400, "q": [0.0, -1e-10, 2.5], "ratio": [0.0, 0.0], "hold": 2}) == \
        "# val 400  q=(0, -1e-10, 2.5)  drift/tol=0  hold=2"
    assert _line2d(1, {"step": 15000, "q": [-2.5e-07, 0.000123456789, 0.0], "ratio": [0.0, 0.5],
                       "hold": 1}) == "# fy 15000  qx=-2.5e-07  qy=0.000123457  dx/tol=0  dy/tol=0.5  hold=1"


NUM = re.compile(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?")


def assert_stdout_matches(ours, ref, name, eps):
    """Our executable's stdout against the reference binary's, line by line:
    the same text between the numbers, integers (steps, hold, counts) equal,
    printed values (6 significant digits, ostream's default) within one unit of
    the 6th digit.  Two kinds of value are differences of nearly equal numbers,
    so the reference's reassociated (-ffast-math) sums show in them: drift/tol
    = max |q - q_prev| / (eps |q|) is off by at most 2 dq / (eps |q|), which
    for fluxes equal to 1e-10 relative (the bar the rows are held to) is
    2e-10 / eps; the reciprocity residuals |K_ij - K_ji| / trace are within
    1e-9, and the tensor test's relative error |q_val - K F| / |q_val| within
    2e-10."""
    a = ours.splitlines()
    assert len(a) == len(ref), (name, len(a), len(ref))
    for x, y in zip(a, ref):
        text = NUM.split(y)
        assert NUM.split(x) == text, (name, x, y)
        for before, p, q in zip(text, NUM.findall(x), NUM.findall(y)):
            if re.fullmatch(r"[-+]?\d+", q):
                assert p == q, (name, x, y)
            elif y.startswith("# reciprocity"):
                assert abs(float(p) - float(q)) <= 1e-9, (name, x, y)
            elif y.startswith("#   relative error"):      # |q_val - K F| / |q_val|, also a difference
                assert abs(float(p) - float(q)) <= 1.01e-5 * abs(float(q)) + 2e-10, (name, x, y)
            elif before.endswith("tol="):
                assert abs(float(p) - float(q)) <= 1.01e-5 * abs(float(q)) + 2e-10 / eps, (name, x, y)
            else:
                assert abs(float(p) - float(q)) <= 1.01e-5 * abs(float(q)) + 1e-300, (name, x, y)


PROGRESS_FIXTURES = ["sph16_s0", "sph20_fast_s2", "ell_x24y20z16_fast_s5", "grf_x36y20z24_fast_s8",
                     "ell24_validate_s3"]


@pytest.mark.gpu
@pytest.mark.parametrize("name", PROGRESS_FIXTURES)
def test_lbm3d_stdout_matches_the_reference_binary(tmp_path, name):
    """Without --quiet the executable prints what the reference prints: the
    structure line, one line per convergence check of every run
    (LMB3d/lbm3d-perm.cpp:353-357: step, q, drift/tol, hold), the tensor test
    with --validate, K, the reciprocity residuals and the eigenvalues.  The
    reference binary's stdout of the same invocation is in the fixture; every
    check's flux and drift/tol must agree to the printed digits."""
    mask, row, meta = golden(name)
    argv = meta["argv"][3:]
    assert "--quiet" not in argv
    (tmp_path / f"{name}.raw").write_bytes((GOLDEN / f"{name}.raw").read_bytes())
    rc, stdout, err = run("lbm3d-perm", f"{name}.raw", "o.csv", *argv, cwd=tmp_path)
    assert rc == meta["returncode"], err
    eps = float(argv[argv.index("--eps") + 1]) if "--eps" in argv else 1e-5
    assert_stdout_matches(stdout, meta["stdout"], name, eps)


@pytest.mark.gpu
def test_lbm3d_velocity_files(tmp_path):
    name = "sph12_velocity_s4"
    _, _, meta = golden(name)
    argv = [a for a in meta["argv"][3:]]
    i = argv.index("--velocity")
    argv[i + 1] = str(tmp_path / "v")
    rc, _, err = run("lbm3d-perm", GOLDEN / f"{name}.raw", tmp_path / "o.csv", *argv)
    assert rc == meta["returncode"], err
    for t in ("fx", "fy", "fz"):
        ours = (tmp_path / f"v.{t}.vel").read_bytes()
        ref = (GOLDEN / f"{name}.{t}.vel").read_bytes()
        assert ours[:16] == ref[:16]
        a = np.frombuffer(ours, "<f4", offset=16)
        b = np.frombuffer(ref, "<f4", offset=16)
        assert np.abs(a - b).max() <= 1e-6 * np.abs(b).max()


@pytest.mark.gpu
def test_lbm2d_row_and_velocity_export(tmp_path, oracle):
    """the 2D executable on a GIF: its CSV row equals the in-process solve's
    text, and its values those of the CPU oracle's whole 2D invocation (the
    reference's float arithmetic, L0 = 4 * 64 / 256) to the printed digits;
    the .dat velocity file holds the exported field"""
    from flowperm import masks, solver, structio
    m = masks.ellipses((64, 64), 0.6, seed=1)
    structio.save_gif2d(tmp_path / "m.gif", m)
    flags = ["--lag", "500", "--min-steps", "1000", "--max-steps", "8000", "--avg-steps", "1000",
             "--quiet", "--velocity", str(tmp_path / "v")]
    rc, _, err = run("lbm2d-perm", tmp_path / "m.gif", tmp_path / "o.csv", *flags)
    assert rc in (0, 3), err
    lines = (tmp_path / "o.csv").read_text().splitlines()
    assert lines[0] == ",".join(solver.COLUMNS2D)
    r = solver.solve_permeability2d(m, lag=500, min_steps=1000, max_steps=8000, avg_steps=1000,
                                    velocity=True, filename=str(tmp_path / "m.gif"))
    assert lines[1] + "\n" == solver.csv_row(r, solver.COLUMNS2D)
    ours = dict(zip(lines[0].split(","), next(csv.reader([lines[1]]))))
    _, T = oracle.solve_tensor2d(m, L0=1.0, lag=500, min_steps=1000, max_steps=8000, avg_steps=1000)
    for col, want in (("steps_fx", T.run[0].steps), ("steps_fy", T.run[1].steps),
                      ("conv_fx", T.run[0].converged), ("conv_fy", T.run[1].converged)):
        assert int(ours[col]) == want, col
    for col, want in (("k_xx", T.k_xx), ("k_yx", T.k_yx), ("k_xy", T.k_xy), ("k_yy", T.k_yy),
                      ("qx_fx", T.run[0].qx), ("qy_fx", T.run[0].qy), ("qx_fy", T.run[1].qx),
                      ("qy_fy", T.run[1].qy), ("fieldres_fx", T.run[0].fieldres),
                      ("fieldres_fy", T.run[1].fieldres), ("anisotropy", T.aniso),
                      ("theta_deg", T.theta)):
        assert abs(float(ours[col]) - want) <= 5e-10 * abs(want), col
    txt = (tmp_path / "v.fx.dat").read_text().split()
    vals = np.array([float(x) for x in txt]).reshape(64, 64, 2)
    assert np.allclose(vals, r["velocity"][0], rtol=1e-5, atol=1e-12)
