"""
Make the golden fixtures for the D3Q19 solver from the reference binary.

Each case writes, under tests/golden/:
  <name>.raw    the voxel mask (own seeded generators, flowperm.masks)
  <name>.csv    the row the reference solver appended (its own output format)
  <name>.json   the flags, exit code, stderr tail and progress lines it printed
plus <name>.f{x,y,z}.vel for the velocity-export case.

The binary is oracle/_ref/lbm3d-perm, compiled from LMB3d/lbm3d-perm.cpp with
the reference's own flags by oracle/Makefile; it must exist (run `make -C
oracle` where the reference tree is mounted).  Nothing else of the reference
is read.  Fixtures are committed; this script only regenerates them.

    python tests/golden/make_golden.py --jobs 8            # every case
    python tests/golden/make_golden.py --only c1_64_s0     # one case
"""
from __future__ import annotations

import argparse
import json
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
sys.path.insert(0, str(REPO / "flow-anizotropy_amd"))
from flowperm import masks, structio          # noqa: E402

REF = REPO / "oracle" / "_ref" / "lbm3d-perm"

C1 = ["--ff", "40", "--lag", "1000", "--eps", "1e-6"]
FAST = ["--ff", "40", "--lag", "200", "--eps", "1e-6", "--min-steps", "400", "--avg-steps", "200"]


def percolating_spheres(shape, seed, porosity=0.6, rmin=3.0, rmax=6.0):
    """Spheres mask from the first seed >= `seed` that wraps along x, y and z."""
    s = seed
    while True:
        m = masks.spheres(shape, porosity, rmin, rmax, seed=s)
        if all(masks.percolates(m)):
            return m, s
        s += 1000


def percolating_ellipsoids(shape, seed, porosity):
    s = seed
    while True:
        m = masks.ellipsoids(shape, porosity, rmin=2.0, rmax=3.5, seed=s)
        if all(masks.percolates(m)):
            return m, s
        s += 1000


def percolating_grf(shape, seed, porosity):
    """A Gaussian random field reduced to a pore cluster that wraps along x, y and z."""
    s = seed
    while True:
        m = masks.grf(shape, porosity, corr=(3.0, 3.0, 6.0), seed=s, require_all=True)
        if all(masks.percolates(m)):
            return m, s
        s += 1000


def case_list():
    """name -> (mask builder, flags).  Shapes are (nz, ny, nx)."""
    cases = {
        # analytic slit, scripts/validate_solver3d.py:87-96 (32^3, --max-steps 40000)
        "slit32_h10": (lambda: (masks.slit(32, 10), None), ["--max-steps", "40000"]),
        "slit32_h20": (lambda: (masks.slit(32, 20), None), ["--max-steps", "40000"]),
        # reference defaults on small sphere packs, cubic and non-cubic
        "sph16_s0": (lambda: percolating_spheres((16, 16, 16), 0), []),
        "sph_x28y24z20_s1": (lambda: percolating_spheres((20, 24, 28), 1), []),
        # the 64^3 parity settings, shortened so the CPU suite can rerun them
        "sph20_fast_s2": (lambda: percolating_spheres((20, 20, 20), 2), FAST),
        "ell_x24y20z16_fast_s5": (lambda: percolating_ellipsoids((16, 20, 24), 5, 0.55), FAST),
        # anisotropic, with off-diagonals, and the (1,1,1) tensor test
        "ell24_validate_s3": (lambda: percolating_ellipsoids((24, 24, 24), 3, 0.55), ["--validate"]),
        # velocity export (export_velocity, lbm3d-perm.cpp:305-317)
        "sph12_velocity_s4": (lambda: percolating_spheres((12, 12, 12), 4, rmin=2.0, rmax=3.0),
                              ["--velocity", "{prefix}"]),
        # rows longer than one 64-cell segment (x-runs that cross segment
        # boundaries and wrap), and a low-porosity field
        "ell_x70y12z10_fast_s6": (lambda: percolating_ellipsoids((10, 12, 70), 6, 0.55), FAST),
        "ell_x130y8z6_fast_s7": (lambda: percolating_ellipsoids((6, 8, 130), 7, 0.6), FAST),
        "grf_x36y20z24_fast_s8": (lambda: percolating_grf((24, 20, 36), 8, 0.3), FAST),
        # the 64^3 parity workload: spheres r~U[3,6], porosity 0.6, F = 1e-5
        "c1_64_s0": (lambda: percolating_spheres((64, 64, 64), 0), C1),
        "c1_64_s1": (lambda: percolating_spheres((64, 64, 64), 1), C1),
    }
    return cases


def run_case(name, builder, flags, quiet_progress=False):
    mask, seed = builder()
    raw = HERE / f"{name}.raw"
    csv = HERE / f"{name}.csv"
    structio.save_raw(raw, mask)
    if csv.exists():
        csv.unlink()
    argv = [str(REF), raw.name, csv.name]
    for f in flags:
        argv.append(f.replace("{prefix}", name))
    t0 = time.time()
    p = subprocess.run(argv, cwd=HERE, capture_output=True, text=True)
    dt = time.time() - t0
    meta = {
        "name": name,
        "argv": [Path(argv[0]).name] + argv[1:],
        "shape_zyx": list(mask.shape),
        "seed": seed,
        "porosity": float(1.0 - mask.mean()),
        "percolates_xyz": list(masks.percolates(mask)),
        "returncode": p.returncode,
        "stdout": p.stdout.splitlines()[-200:],
        "stderr": p.stderr.splitlines()[-20:],
        "seconds": round(dt, 1),
    }
    (HERE / f"{name}.json").write_text(json.dumps(meta, indent=1) + "\n")
    return name, p.returncode, dt


def error_cases():
    """Inputs the reference rejects (read_structure, :174-193): record rc + message."""
    out = {}
    good = np.zeros((4, 4, 4), np.uint8)
    good[0] = 1
    hdr = np.array([4, 4, 4], "<i4").tobytes()
    variants = {
        "truncated": hdr + good.tobytes()[:-3],
        "trailing": hdr + good.tobytes() + b"\x00",
        "nonbinary": hdr + (good + 2 * (good == 0)).astype(np.uint8).tobytes(),
        "nopore": hdr + np.ones((4, 4, 4), np.uint8).tobytes(),
        "baddims": np.array([4, 0, 4], "<i4").tobytes(),
        "noheader": b"\x01\x02",
    }
    tmp = HERE / "_err.raw"
    for k, data in variants.items():
        tmp.write_bytes(data)
        p = subprocess.run([str(REF), tmp.name, "_err.csv", "--quiet"], cwd=HERE,
                           capture_output=True, text=True)
        out[k] = {"returncode": p.returncode, "stderr": p.stderr.strip().replace(str(HERE) + "/", "")}
        if (HERE / "_err.csv").exists():
            (HERE / "_err.csv").unlink()
    tmp.unlink()
    # a flag without its value and an unknown flag (:467-505)
    for k, extra in {"missing_value": ["--lag"], "unknown_flag": ["--bogus"]}.items():
        structio.save_raw(tmp, good)
        p = subprocess.run([str(REF), tmp.name, "_err.csv"] + extra, cwd=HERE,
                           capture_output=True, text=True)
        out[k] = {"returncode": p.returncode, "stderr_last": p.stderr.strip().splitlines()[-1]}
        tmp.unlink()
    # a diverging run: huge force, checks every 10 steps
    m = masks.slit(8, 4)
    structio.save_raw(tmp, m)
    p = subprocess.run([str(REF), tmp.name, "_err.csv", "--quiet", "--ff", "1e9", "--lag", "10",
                        "--max-steps", "200", "--avg-steps", "0"], cwd=HERE,
                       capture_output=True, text=True)
    out["diverging_ff1e9"] = {"returncode": p.returncode, "stderr": p.stderr.strip()[-300:],
                              "csv": (HERE / "_err.csv").read_text() if (HERE / "_err.csv").exists() else ""}
    for f in (tmp, HERE / "_err.csv"):
        if f.exists():
            f.unlink()
    (HERE / "errors.json").write_text(json.dumps(out, indent=1) + "\n")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", nargs="*")
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("--errors", action="store_true", help="also record the error cases")
    a = ap.parse_args()
    if not REF.exists():
        sys.exit(f"{REF} missing: build it with `make -C oracle ref`")
    cases = case_list()
    names = a.only if a.only else list(cases)
    if a.errors:
        print(json.dumps(error_cases(), indent=1))
    with ThreadPoolExecutor(a.jobs) as ex:
        futs = [ex.submit(run_case, n, *cases[n]) for n in names]
        for f in futs:
            n, rc, dt = f.result()
            print(f"{n}: rc={rc} {dt:.1f}s", flush=True)


if __name__ == "__main__":
    main()
