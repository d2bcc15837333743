"""
Command-line drop-ins for the reference solver executables.

    lbm3d-perm <structure.raw> <results.csv> [options]      (LMB3d/lbm3d-perm.cpp:446-630)
    lbm2d-perm <structure.gif|ppm|dat> <results.csv> [options]   (LMB2d/main.cpp:238-426)

Same arguments, same CSV header and row (appended; header only when the file
is empty or new), same exit codes: 0 all directions converged, 3 some
direction hit --max-steps (the row is still written), 2 a hard error with
"lbm3d-perm: <message>" on stderr (2D: missing arguments exit 1, as the
reference does).  The reference drivers (scripts/solve_permeability3d.py,
scripts/solve_permeability.py) take the solver path as --solver, so the
executables in flow-anizotropy_amd/bin/ replace the binaries without changing
the drivers.  One addition: --precision f64|f32 (3D; default f64, the
reference's arithmetic) and f32ref|f64|f32 (2D; default f32ref).

The solve runs on the GPU through libflowperm.so; there is no CPU path.
"""
from __future__ import annotations

import os
import sys

import numpy as np

USAGE3 = """Usage: lbm3d-perm <structure.raw> <results.csv> [options]

Solves a triply periodic cell three times (force along +x, +y, +z) on the GPU
and appends one row with the full 3x3 permeability tensor to <results.csv>.

  --ff F            force factor; body force = 2.5e-07 * F   (default 1)
  --eps E           drift tolerance, relative to |q|         (default 1e-5)
  --lag N           steps between convergence checks         (default 2000)
  --hold K          consecutive passing checks required      (default 2)
  --min-steps N     no convergence before this step          (default 4000)
  --max-steps N     give up after this many steps            (default 200000)
  --avg-steps N     averaging phase after convergence        (default 2000)
  --velocity PREFIX write PREFIX.fx.vel, PREFIX.fy.vel, PREFIX.fz.vel
  --validate        extra solve along (1,1,1)/sqrt(3); reports the flux
                    predicted by K against the measured one
  --quiet           suppress progress
  --precision P     f64 (default: the reference's arithmetic) or f32
                    (shifted FP32 populations, FP64 sums)
"""

USAGE2 = """Usage: lbm2d-perm <structure.gif|ppm|dat> <results.csv> [options]

Solves the doubly periodic cell twice (force along +x, then +y) on the GPU
and appends one row with the full permeability tensor to <results.csv>.

  --ff F            force factor; body force = 2.5e-07 * F   (default 1)
  --eps E           per-component relative tolerance         (default 1e-3)
  --lag N           steps between convergence checks         (default 5000)
  --hold K          consecutive passing checks required      (default 3)
  --min-steps N     no convergence before this step          (default 10000)
  --max-steps N     give up after this many steps            (default 500000)
  --avg-steps N     length of the averaging phase            (default 20000)
  --avg-sample N    flux sampling interval                   (default 50)
  --qfloor R        floor for the relative test, as a fraction of |q| (default 1e-2)
  --ksigma K        also accept a change within K sigma of the block mean's
                    own standard error                       (default 3)
  --velocity PREFIX write PREFIX.fx.dat and PREFIX.fy.dat
  --validate DEG    extra solve with the force at DEG degrees (diagnostic)
  --quiet           suppress progress on stdout
  --precision P     f32ref (default: the reference's float arithmetic),
                    f64 or f32 (shifted)
"""


class _Die(Exception):
    def __init__(self, msg, code=2):
        super().__init__(msg)
        self.code = code


def _strtol(s):
    """strtol(s, nullptr, 10): leading integer, 0 when there is none"""
    s = s.strip()
    i = 0
    if i < len(s) and s[i] in "+-":
        i += 1
    j = i
    while j < len(s) and s[j].isdigit():
        j += 1
    return int(s[:j]) if j > i else 0


def _strtod(s):
    """strtod(s, nullptr): the longest leading float, 0.0 when there is none"""
    s = s.strip()
    for n in range(len(s), 0, -1):
        try:
            return float(s[:n])
        except ValueError:
            continue
    return 0.0


def _parse(argv, spec, usage, prog):
    """flag parsing of the reference main()s: spec maps a flag to (key, kind)
    with kind in {"l", "d", "s", "b"}"""
    cfg = {}
    i = 0
    while i < len(argv):
        a = argv[i]
        if a in ("-h", "--help") and prog == "lbm2d-perm":
            sys.stderr.write(usage)
            raise _Die(None, 0)
        if a not in spec:
            sys.stderr.write(usage)
            raise _Die(f"unknown option {a}")
        key, kind = spec[a]
        if kind == "b":
            cfg[key] = True
            i += 1
            continue
        if i + 1 >= len(argv):
            sys.stderr.write(usage)
            raise _Die(f"missing value for {a}")
        v = argv[i + 1]
        cfg[key] = _strtol(v) if kind == "l" else _strtod(v) if kind == "d" else v
        i += 2
    return cfg


def _append_row(path, header, line):
    new = not os.path.exists(path) or os.path.getsize(path) == 0
    try:
        with open(path, "a") as fh:
            if new:
                fh.write(header)
            fh.write(line)
    except OSError:
        raise _Die(f"cannot open {path} for writing")


def _write_vel3d(path, vel):
    nz, ny, nx, _ = vel.shape
    try:
        with open(path, "wb") as fh:
            fh.write(np.array([nx, ny, nz, 3], "<i4").tobytes())
            fh.write(np.ascontiguousarray(vel, "<f4").tobytes())
    except OSError:
        raise _Die(f"cannot write {path}")


def _write_vel2d(path, vel):
    # exportvelocity (LMB2d/lbm.cpp:587-594): y outer, x inner, "U V" per
    # line at the stream's default precision (6 significant digits)
    with open(path, "w") as fh:
        for u, v in vel.reshape(-1, 2):
            fh.write("%g %g\n" % (u, v))


SPEC3 = {"--ff": ("ff", "d"), "--eps": ("eps", "d"), "--lag": ("lag", "l"),
         "--hold": ("hold", "l"), "--min-steps": ("min_steps", "l"),
         "--max-steps": ("max_steps", "l"), "--avg-steps": ("avg_steps", "l"),
         "--velocity": ("velocity", "s"), "--validate": ("validate", "b"),
         "--quiet": ("quiet", "b"), "--precision": ("precision", "s")}

SPEC2 = {"--ff": ("ff", "d"), "--eps": ("eps", "d"), "--lag": ("lag", "l"),
         "--hold": ("hold", "l"), "--min-steps": ("min_steps", "l"),
         "--max-steps": ("max_steps", "l"), "--avg-steps": ("avg_steps", "l"),
         "--avg-sample": ("avg_sample", "l"), "--qfloor": ("qfloor", "d"),
         "--ksigma": ("ksigma", "d"), "--velocity": ("velocity", "s"),
         "--validate": ("validate", "d"), "--quiet": ("quiet", "b"),
         "--precision": ("precision", "s")}


def _progress_lines(fmt):
    """a progress callback collecting each run's check lines, and a function
    printing them run by run (the order the reference's sequential runs print
    them in)"""
    lines = {}

    def progress(run, c):
        lines.setdefault(run, []).append(fmt(run, c))

    def flush():
        for run in sorted(lines):
            for line in lines[run]:
                print(line)
        sys.stdout.flush()
    return progress, flush


def _line3d(run, c):
    """LMB3d/lbm3d-perm.cpp:353-357"""
    q = c["q"]
    return (f"# {('fx', 'fy', 'fz', 'val')[run]} {c['step']}  q=({q[0]:g}, {q[1]:g}, {q[2]:g})  "
            f"drift/tol={c['ratio'][0]:g}  hold={c['hold']}")


def _line2d(run, c):
    """LMB2d/main.cpp:167-172"""
    q, r = c["q"], c["ratio"]
    return (f"# {('fx', 'fy', 'val')[run]} {c['step']}  qx={q[0]:g}  qy={q[1]:g}  "
            f"dx/tol={r[0]:g}  dy/tol={r[1]:g}  hold={c['hold']}")


def lbm3d_main(argv):
    from . import structio
    from ._lib import FlowpermError
    from .solver import COLUMNS3D, csv_header, csv_row, solve_permeability3d
    if len(argv) < 2:
        sys.stderr.write(USAGE3)
        return 2
    structure, results = argv[0], argv[1]
    cfg = _parse(argv[2:], SPEC3, USAGE3, "lbm3d-perm")
    quiet = cfg.pop("quiet", False)
    prefix = cfg.pop("velocity", None)
    mask = structio.load_raw(structure)          # StructureError -> exit 2 (main)
    nz, ny, nx = mask.shape
    if not quiet:
        F = 2.5e-7 * cfg.get("ff", 1.0)
        print(f"# structure {structure}  {nx}x{ny}x{nz}  porosity {1 - mask.mean():g}  "
              f"force {F:g}  mu {1.0 / 6.0:g}", flush=True)
    progress, flush = _progress_lines(_line3d)
    try:
        row = solve_permeability3d(mask, filename=structure, velocity=prefix is not None,
                                   progress=None if quiet else progress, **cfg)
    except FlowpermError as e:
        flush()
        raise _Die(str(e).rsplit(" (status", 1)[0])
    flush()
    if prefix is not None:
        for d, t in enumerate(("fx", "fy", "fz")):
            _write_vel3d(f"{prefix}.{t}.vel", row["velocity"][d])
    for t in ("fx", "fy", "fz"):
        if not row[f"conv_{t}"]:
            sys.stderr.write(f"# {t}: max-steps reached without convergence\n")
    if cfg.get("validate"):
        print(f"# tensor test along (1,1,1)  (converged {row['val_converged']}, "
              f"{row['val_steps']} steps)")
        print("#   predicted q = (%g, %g, %g)" % tuple(row["val_pred"]))
        print("#   measured  q = (%g, %g, %g)" % tuple(row["val_q"]))
        print("#   relative error %g" % row["val_err"], flush=True)
    _append_row(results, csv_header(COLUMNS3D), csv_row(row, COLUMNS3D))
    if not quiet:
        K = np.array([[row[f"K_{a}{b}"] for b in "xyz"] for a in "xyz"])
        print("# K =")
        for r in range(3):
            print("#   [ %12g %12g %12g ]" % tuple(K[r]))
        print(f"# reciprocity residuals  xy {row['recip_xy']:g}  xz {row['recip_xz']:g}  "
              f"yz {row['recip_yz']:g}")
        print(f"# k1 {row['k1']:g}  k2 {row['k2']:g}  k3 {row['k3']:g}  "
              f"k_mean {row['k_mean']:g}  FA {row['fa']:g}", flush=True)
    return row["exit_code"]


def lbm2d_main(argv):
    from . import structio
    from ._lib import FlowpermError
    from .solver import COLUMNS2D, csv_header, csv_row, solve_permeability2d
    if len(argv) < 2:
        sys.stderr.write(USAGE2)
        return 1
    structure, results = argv[0], argv[1]
    cfg = _parse(argv[2:], SPEC2, USAGE2, "lbm2d-perm")
    if cfg.get("lag", 1) <= 0 or cfg.get("avg_sample", 1) <= 0:
        raise _Die("--lag and --avg-sample must be positive")
    quiet = cfg.pop("quiet", False)
    prefix = cfg.pop("velocity", None)
    try:
        mask = structio.load_structure2d(structure)
    except (OSError, ValueError) as e:
        raise _Die(str(e))
    if not quiet:                                 # main.cpp:317-319 (porosity and mu are floats)
        por = np.float32((mask == 0).sum()) / np.float32(mask.size)
        ff = cfg.get("ff", 1.0)
        print(f"# structure {structure}  porosity {float(por):g}  force {2.5e-7 * ff:g} "
              f"(ff={ff:g})  mu {float(np.float32(0.33 * 0.5)):g}", flush=True)
    progress, flush = _progress_lines(_line2d)
    try:
        row = solve_permeability2d(mask, filename=structure, velocity=prefix is not None,
                                   progress=None if quiet else progress, **cfg)
    except FlowpermError as e:
        flush()
        raise _Die(str(e).rsplit(" (status", 1)[0])
    flush()
    if prefix is not None:
        for d, t in enumerate(("fx", "fy")):
            _write_vel2d(f"{prefix}.{t}.dat", row["velocity"][d])
    for t in ("fx", "fy"):
        if not row[f"conv_{t}"]:
            sys.stderr.write(f"# {t}: max-steps reached without convergence\n")
    if "val_err" in row:
        print(f"# tensor test at {cfg['validate']:g} deg  (converged {row['val_converged']}, "
              f"{row['val_steps']} steps)")
        print("#   predicted q = (%g, %g)" % tuple(row["val_pred"]))
        print("#   measured  q = (%g, %g)" % tuple(row["val_q"]))
        print("#   relative error %g" % row["val_err"], flush=True)
    _append_row(results, csv_header(COLUMNS2D), csv_row(row, COLUMNS2D))
    if not quiet:
        print(f"# K = [ {row['k_xx']:g} {row['k_xy']:g} ; {row['k_yx']:g} {row['k_yy']:g} ]")
        print(f"# reciprocity residual {row['recip_resid']:g}")
        print(f"# k1 {row['k1']:g}  k2 {row['k2']:g}  anisotropy {row['anisotropy']:g}  "
              f"theta {row['theta_deg']:g} deg", flush=True)
    return row["exit_code"]


def run(prog, argv):
    """Run one of the two programs; returns the process exit code."""
    from .structio import StructureError
    main = lbm3d_main if prog == "lbm3d-perm" else lbm2d_main
    try:
        return main(argv)
    except _Die as e:
        if e.args[0] is not None:
            sys.stderr.write(f"{prog}: {e.args[0]}\n")
        return e.code
    except StructureError as e:
        sys.stderr.write(f"{prog}: {e}\n")
        return 2


def main(argv=None):
    argv = list(sys.argv[1:] if argv is None else argv)
    if not argv or argv[0] not in ("lbm3d", "lbm2d", "lbm3d-perm", "lbm2d-perm"):
        sys.stderr.write("usage: python -m flowperm.cli {lbm3d|lbm2d} <structure> <results.csv> "
                         "[options]\n")
        return 2
    prog = "lbm3d-perm" if argv[0].startswith("lbm3d") else "lbm2d-perm"
    return run(prog, argv[1:])


if __name__ == "__main__":
    sys.exit(main())
