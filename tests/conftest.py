import json
import math
import sys
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parent.parent
GOLDEN = REPO / "tests" / "golden"
sys.path.insert(0, str(REPO / "flow-anizotropy_amd"))
sys.path.insert(0, str(REPO))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


def golden(name):
    """(mask [z,y,x] uint8, reference CSV row as dict of strings/floats, meta dict)."""
    from flowperm import structio
    import pandas as pd
    mask = structio.load_raw(GOLDEN / f"{name}.raw", check=False)
    row = pd.read_csv(GOLDEN / f"{name}.csv").iloc[0].to_dict()
    meta = json.loads((GOLDEN / f"{name}.json").read_text())
    return np.array(mask), row, meta


def golden_flags(meta):
    """Reference argv flags -> keyword config."""
    argv = meta["argv"][3:]
    cfg = {}
    names = {"--ff": ("ff", float), "--eps": ("eps", float), "--lag": ("lag", int),
             "--hold": ("hold", int), "--min-steps": ("min_steps", int),
             "--max-steps": ("max_steps", int), "--avg-steps": ("avg_steps", int)}
    i = 0
    while i < len(argv):
        a = argv[i]
        if a in names:
            k, t = names[a]
            cfg[k] = t(argv[i + 1])
            i += 2
        else:
            i += 1
    return cfg


def printed_tol(ref):
    """Half a unit in the 10th significant digit: the precision of a value the
    reference printed with setprecision(10)."""
    if ref == 0 or not math.isfinite(ref):
        return 0.0
    e = math.floor(math.log10(abs(ref)))
    return 0.5 * 10.0 ** (e - 9) * 1.000001


def assert_printed_close(ours, ref, floor=0.0, what=""):
    """|ours - true reference value| <= floor, where the reference value is
    known only to half a unit of its 10th printed digit: the two bounds add."""
    tol = printed_tol(ref) + floor
    assert abs(ours - ref) <= tol, f"{what}: ours {ours!r} ref {ref!r} |d|={abs(ours-ref):.3e} tol {tol:.3e}"


def check_eigen_and_reciprocity(get, row, trace, what=""):
    """The CSV's principal vectors v1..v3, fa and recip_* (LMB3d/lbm3d-perm.cpp:
    531-553, 605-613) against a reference row.  get(key) gives our value.

    A principal vector is defined (with the sign jacobi3 gives it) only when
    its value is separated from the other two: K known to ~1e-10 trace per
    component moves it by about |dK| / gap, so it is compared with that bound
    (plus its printed precision) and skipped when the gap is below 1e-6 trace
    (a degenerate pair, e.g. K_xx = K_yy of a slit, has no preferred basis).
    fa is a smooth function of the principal values (1e-9); recip_* are
    differences of off-diagonals over the trace (3e-10 absolute)."""
    k = [float(row["k1"]), float(row["k2"]), float(row["k3"])]
    checked = 0
    for j in range(3):
        gap = min(abs(k[j] - k[i]) for i in range(3) if i != j)
        if gap < 1e-6 * trace:
            continue
        for a in "xyz":
            key = f"v{j + 1}_{a}"
            assert_printed_close(float(get(key)), float(row[key]), 1e-9 * trace / gap, f"{what} {key}")
        checked += 1
    assert_printed_close(float(get("fa")), float(row["fa"]), 1e-9, f"{what} fa")
    for key in ("recip_xy", "recip_xz", "recip_yz", "recip_resid"):
        assert_printed_close(float(get(key)), float(row[key]), 3e-10, f"{what} {key}")
    return checked


# GOLDEN_FAST: also re-solved by the CPU oracle (seconds each); rows of 70 cells
# (a segment boundary) and a porosity-0.3 field.  GOLDEN_ALL adds 130-cell rows
# and the longer cases, compared with the reference rows only.
GOLDEN_FAST = ["sph16_s0", "sph20_fast_s2", "ell_x24y20z16_fast_s5",
               "ell_x70y12z10_fast_s6", "grf_x36y20z24_fast_s8"]
GOLDEN_ALL = ["sph16_s0", "sph20_fast_s2", "ell_x24y20z16_fast_s5", "ell_x70y12z10_fast_s6",
              "ell_x130y8z6_fast_s7", "grf_x36y20z24_fast_s8", "sph_x28y24z20_s1",
              "slit32_h10", "slit32_h20", "c1_64_s0", "c1_64_s1", "ell24_validate_s3"]


@pytest.fixture(scope="session")
def oracle():
    from oracle import binding
    binding.lib()
    return binding


@pytest.fixture(scope="session")
def gpu_ctx():
    from flowperm import solver
    ctx = solver.Context()
    yield ctx
    ctx.close()
