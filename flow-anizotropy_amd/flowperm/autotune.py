"""
Launch-shape autotuning of the step kernels, cached per device.

``fp_autotune`` (libflowperm.so) times every launch shape of the local and the
neighbour kernel on a lattice and keeps the fastest of each.  This module runs
it once per (device name, lattice, precision, lattice-size class) and keeps
the answer in a JSON cache, so later contexts start with the tuned shapes.

Cache: $FLOWPERM_TUNING_CACHE, else ~/.cache/flowperm/tuning.json.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from pathlib import Path

import numpy as np

from . import _lib


FIELDS = ("local_cpt", "nbr_waves", "chunk_steps", "nontemporal", "local_bpc", "nbr_bpc",
          "nbr_mode", "nbr_segs", "local_wgs", "nbr_wgs")
N_LOCAL, N_NBR = 24, 16              # FP_TUNE_NLOCAL, FP_TUNE_NNBR


def shapes(lib):
    """The shapes fp_autotune tries: (local list of (cpt, bpc, nt, wgs),
    neighbour list of (waves, bpc, mode, segs, wgs))."""
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    loc, nbr = [], []
    for i in range(N_LOCAL):
        lib.fp_tune_shape(0, i, C.byref(a), C.byref(b), C.byref(c))
        cpt, bb, nt = a.value, b.value, c.value
        lib.fp_tune_shape(3, i, C.byref(a), C.byref(b), C.byref(c))
        loc.append((cpt, bb, nt, a.value))
    for i in range(N_NBR):
        lib.fp_tune_shape(1, i, C.byref(a), C.byref(b), C.byref(c))
        w, bb, m = a.value, b.value, c.value
        lib.fp_tune_shape(2, i, C.byref(a), C.byref(b), C.byref(c))
        sg = a.value
        lib.fp_tune_shape(4, i, C.byref(a), C.byref(b), C.byref(c))
        nbr.append((w, bb, m, sg, a.value))
    return loc, nbr


def nbr_name(w, b, m, s=1, g=0):
    """waves4_segs4_bpc0, compact_bpc0; _wgs5 when held to 5 workgroups per CU"""
    cap = f"_wgs{g}" if g else ""
    return (f"compact_bpc{b}" if m == 1 else f"waves{w}_segs{s}_bpc{b}") + cap


def local_name(c, b, n, g=0):
    """cpt2_bpc0_nt1; _wgs3 when held to 3 workgroups per CU"""
    return f"cpt{c}_bpc{b}_nt{n}" + (f"_wgs{g}" if g else "")


def cache_path() -> Path:
    p = os.environ.get("FLOWPERM_TUNING_CACHE")
    return Path(p) if p else Path.home() / ".cache" / "flowperm" / "tuning.json"


def device_name() -> str:
    try:
        out = Path("/sys/class/kfd/kfd/topology/nodes")
        names = []
        for n in sorted(out.iterdir()):
            prop = (n / "properties").read_text()
            for line in prop.splitlines():
                if line.startswith("gfx_target_version") and not line.endswith(" 0"):
                    names.append(line.split()[-1])
        return "gfx" + names[0] if names else "unknown"
    except Exception:
        return "unknown"


def _key(lattice, precision, n_cells, porosity):
    # the best neighbour-step form depends on how sparse the lattice is
    size_class = "small" if n_cells < 2_000_000 else "large"
    # (the shapes gained nbr_segs, then the workgroup caps: entries of older
    # libraries lack them)
    return f"{device_name()}/{lattice}/{precision}/{size_class}/phi{round(porosity, 1):.1f}/w"


def tune(ctx, mask, lattice=3, precision=1, steps=40, use_cache=True):
    """Return the best fp_tuning for this lattice class and apply it to ctx."""
    mask = np.ascontiguousarray(mask, np.uint8)
    key = _key(lattice, precision, mask.size, float(1.0 - mask.mean()))
    cp = cache_path()
    cache = {}
    if use_cache and cp.exists():
        try:
            cache = json.loads(cp.read_text())
        except (OSError, ValueError):
            cache = {}
    if use_cache and key in cache:
        t = _lib.Tuning(**{k: cache[key][k] for k in FIELDS})
        ctx.lib.fp_set_tuning(ctx.handle, C.byref(t))
        return cache[key]
    if lattice == 3:
        nz, ny, nx = mask.shape
    else:
        (ny, nx), nz = mask.shape, 1
    best = _lib.Tuning()
    lms = (C.c_double * N_LOCAL)()
    nms = (C.c_double * N_NBR)()
    st = ctx.lib.fp_autotune(ctx.handle, _lib.u8ptr(mask), nx, ny, nz, lattice, precision, steps,
                             C.byref(best), lms, nms)
    if st != 0:
        raise _lib.FlowpermError(st, ctx.last_error())
    rec = {k: getattr(best, k) for k in FIELDS}
    loc, nbr = shapes(ctx.lib)
    rec["local_ms"] = {local_name(*sh): round(v, 5) for sh, v in zip(loc, lms)}
    rec["nbr_ms"] = {nbr_name(*sh): round(v, 5) for sh, v in zip(nbr, nms) if v < 1e29}
    if use_cache:
        cache[key] = rec
        try:
            cp.parent.mkdir(parents=True, exist_ok=True)
            cp.write_text(json.dumps(cache, indent=1))
        except OSError:
            pass
    return rec
