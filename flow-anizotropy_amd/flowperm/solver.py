"""
Single-structure permeability solves: the call surface of the reference
solvers, without the process boundary.

``solve_permeability3d(mask_or_path, **cfg)`` does what one ``lbm3d-perm``
invocation does (LMB3d/lbm3d-perm.cpp:478-630): three forced D3Q19 solves, K,
reciprocity, principal values and FA, and returns the CSV row as a dict keyed
by the reference's column names (:582-597).  ``solve_permeability2d`` is the
same for ``lbm2d-perm`` (LMB2d/main.cpp:278-426, header :393-399).

Failures raise ``FlowpermError`` carrying the status (its ``exit_code`` is the
reference's 2); a non-converged direction is not a failure, exactly as in the
reference: the row is returned with ``conv_* = 0`` and ``exit_code = 3``.
"""
from __future__ import annotations

import ctypes as C
import threading
from pathlib import Path

import numpy as np

from . import _lib
from ._lib import FlowpermError
from .structio import StructureError, load_raw, load_structure2d

__all__ = ["Context", "solve_permeability3d", "solve_permeability2d", "solve_direction3d",
           "COLUMNS3D", "COLUMNS2D", "csv_header", "csv_row", "format_value", "jacobi3",
           "FlowpermError"]

_TAGS = ("fx", "fy", "fz")
_AX = ("x", "y", "z")

# LMB3d/lbm3d-perm.cpp:584-597
COLUMNS3D = (["filename", "ff", "force", "mu", "nx", "ny", "nz", "porosity",
              "steps_fx", "conv_fx", "steps_fy", "conv_fy", "steps_fz", "conv_fz"]
             + [f"q{c}_{t}" for t in _TAGS for c in _AX]
             + [f"se_q{c}_{t}" for t in _TAGS for c in _AX]
             + [f"K_{r}{c}" for r in _AX for c in _AX]
             + ["k_xy", "k_xz", "k_yz", "recip_xy", "recip_xz", "recip_yz", "recip_resid",
                "k1", "k2", "k3", "k_mean", "fa"]
             + [f"v{j + 1}_{c}" for j in range(3) for c in _AX])

# LMB2d/main.cpp:393-399
COLUMNS2D = ["filename", "ff", "force", "mu", "porosity",
             "steps_fx", "conv_fx", "steps_fy", "conv_fy",
             "qx_fx", "qy_fx", "qx_fy", "qy_fy",
             "se_qx_fx", "se_qy_fx", "se_qx_fy", "se_qy_fy",
             "k_xx", "k_yx", "k_xy", "k_yy", "k_off", "recip_resid",
             "k1", "k2", "k_mean", "anisotropy", "theta_deg", "alpha_x_deg", "alpha_y_deg",
             "fieldres_fx", "fieldres_fy"]

_INT_COLS = {"nx", "ny", "nz", "steps_fx", "steps_fy", "steps_fz", "conv_fx", "conv_fy", "conv_fz"}


def format_value(v) -> str:
    """How an ostream with setprecision(10) prints the value (:600)."""
    if isinstance(v, (int, np.integer)) and not isinstance(v, bool):
        return str(int(v))
    if isinstance(v, str):
        return v
    return "%.10g" % float(v)


def csv_header(columns) -> str:
    return ",".join(columns) + "\n"


def csv_row(row: dict, columns) -> str:
    return ",".join(format_value(row[c]) for c in columns) + "\n"


class Context:
    """A device context (fp_create).  Use from one thread at a time."""

    def __init__(self, device: int | None = None):
        self.lib = _lib.load()
        st = C.c_int(0)
        dev = _lib.device_index() if device is None else device
        self.handle = self.lib.fp_create(dev, C.byref(st))
        if not self.handle:
            raise FlowpermError(st.value, "fp_create failed: no usable HIP device")
        self.lock = threading.Lock()

    def last_error(self) -> str:
        return (self.lib.fp_last_error(self.handle) or b"").decode()

    def close(self):
        if self.handle:
            self.lib.fp_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None
_default_lock = threading.Lock()


def default_context() -> Context:
    global _default_ctx
    with _default_lock:
        if _default_ctx is None:
            _default_ctx = Context()
        return _default_ctx


def _mask3d(mask_or_path):
    if isinstance(mask_or_path, (str, Path)):
        return np.ascontiguousarray(load_raw(mask_or_path, check=False)), str(mask_or_path)
    arr = np.asarray(mask_or_path)
    if arr.ndim != 3:
        raise StructureError(f"expected a (nz, ny, nx) mask, got shape {arr.shape}")
    return np.ascontiguousarray(arr.astype(np.uint8, copy=False)), "<array>"


def _precision(p):
    if isinstance(p, str):
        try:
            return _lib.PRECISIONS[p.lower()]
        except KeyError:
            raise ValueError(f"unknown precision {p!r}") from None
    return int(p)


def config3d(ff=1.0, eps=1e-5, lag=2000, hold=2, min_steps=4000, max_steps=200000,
             avg_steps=2000, validate=False, precision="f64", concurrent=True):
    c = _lib.Config3D()
    c.ff, c.eps, c.lag, c.hold = float(ff), float(eps), int(lag), int(hold)
    c.min_steps, c.max_steps, c.avg_steps = int(min_steps), int(max_steps), int(avg_steps)
    c.validate = 1 if validate else 0
    c.precision = _precision(precision)
    c.concurrent = 1 if concurrent else 0
    return c


def _raise(ctx, st):
    raise FlowpermError(st, ctx.last_error())


def _check_of(o):
    """the last first-phase convergence check of a run (fp_solve_out check_*)"""
    return {"step": int(o.check_step), "checks": int(o.checks), "q": list(o.check_q),
            "ratio": list(o.check_ratio), "hold": int(o.check_hold)}


def _call(ctx, progress, fn, *args):
    """fn(*args) under the context's lock; with progress (a callable taking
    the run index and a check dict, plus 'steps' and 'converged') called after
    every first-phase convergence check (fp_set_progress)"""
    with ctx.lock:
        if progress is None:
            return fn(*args)

        def cb(user, run, out):
            o = out.contents
            d = _check_of(o)
            d["steps"], d["converged"] = int(o.steps), int(o.converged)
            progress(int(run), d)
        hook = _lib.PROGRESS_FN(cb)
        ctx.lib.fp_set_progress(ctx.handle, hook, None)
        try:
            return fn(*args)
        finally:
            ctx.lib.fp_set_progress(ctx.handle, _lib.PROGRESS_FN(), None)


def jacobi3(A):
    """The reference's cyclic Jacobi (LMB3d/lbm3d-perm.cpp:401-443) via the library."""
    lib = _lib.load()
    a = np.ascontiguousarray(np.asarray(A, float).reshape(9))
    val = np.zeros(3)
    vec = np.zeros(9)
    lib.fp_jacobi3(_lib.f64ptr(a), _lib.f64ptr(val), _lib.f64ptr(vec))
    return val, vec.reshape(3, 3)


def solve_direction3d(mask_or_path, F, *, ctx=None, velocity=False, progress=None, **cfg):
    """One forced solve (solve(), LMB3d/lbm3d-perm.cpp:320-398).  progress:
    see solve_permeability3d."""
    ctx = ctx or default_context()
    mask, _ = _mask3d(mask_or_path)
    nz, ny, nx = mask.shape
    c = config3d(**cfg)
    out = _lib.SolveOut()
    Fa = np.ascontiguousarray(np.asarray(F, dtype=np.float64).reshape(3))
    vel = np.zeros((nz, ny, nx, 3), np.float32) if velocity else None
    st = _call(ctx, progress, ctx.lib.fp_solve_dir3d, ctx.handle, _lib.u8ptr(mask), nx, ny, nz,
               C.byref(c), _lib.f64ptr(Fa), C.byref(out), _lib.f32ptr(vel))
    if st not in (_lib.FP_OK, _lib.FP_NOT_CONVERGED):
        _raise(ctx, st)
    res = {"q": np.array(out.q[:]), "se": np.array(out.se[:]), "steps": int(out.steps),
           "converged": int(out.converged), "status": st, "check": _check_of(out)}
    if velocity:
        res["velocity"] = vel
    return res


def solve_permeability3d(mask_or_path, *, ctx=None, filename=None, velocity=False, progress=None,
                         **cfg):
    """One lbm3d-perm invocation on a mask (array [z,y,x] or .raw path).

    Keyword arguments: ff, eps, lag, hold, min_steps, max_steps, avg_steps
    (the reference flags), validate, precision ("f64" | "f32"), concurrent (default
    True: the three forcing runs step together; the results are bitwise those of
    running them one after another, as the reference does).
    Returns the CSV row as a dict plus 'exit_code', 'status', the last
    first-phase check of every run ('check_fx' ...: step, q, ratio =
    [drift/tol, 0], hold) and, with validate, 'val_*' entries; with
    velocity=True also 'velocity' (3, nz, ny, nx, 3) float32.  progress, when
    given, is called as progress(run, check) after every first-phase check of
    a run (run 0..2 = fx, fy, fz, 3 = val; check as above plus 'steps' and
    'converged'): the values of the reference's per-check progress line.
    """
    ctx = ctx or default_context()
    mask, name = _mask3d(mask_or_path)
    nz, ny, nx = mask.shape
    c = config3d(**cfg)
    r = _lib.Result3D()
    vel = np.zeros((3, nz, ny, nx, 3), np.float32) if velocity else None
    st = _call(ctx, progress, ctx.lib.fp_solve3d, ctx.handle, _lib.u8ptr(mask), nx, ny, nz,
               C.byref(c), C.byref(r), _lib.f32ptr(vel))
    if st not in (_lib.FP_OK, _lib.FP_NOT_CONVERGED):
        _raise(ctx, st)
    row = row3d(r, st, c.ff, filename if filename is not None else name)
    if velocity:
        row["velocity"] = vel
    return row


def row3d(r, st, ff, filename):
    """The lbm3d-perm CSV row (LMB3d/lbm3d-perm.cpp:574-615) of a Result3D, as
    a dict with the reference's column names, plus exit_code, status, timing
    and the --validate values."""
    row = {"filename": filename,
           "ff": ff, "force": r.force, "mu": r.mu, "nx": r.nx, "ny": r.ny, "nz": r.nz,
           "porosity": r.porosity}
    for d, t in enumerate(_TAGS):
        row[f"steps_{t}"] = int(r.run[d].steps)
        row[f"conv_{t}"] = int(r.run[d].converged)
    for d, t in enumerate(_TAGS):
        row[f"check_{t}"] = _check_of(r.run[d])
    for d, t in enumerate(_TAGS):
        for k, a in enumerate(_AX):
            row[f"q{a}_{t}"] = r.run[d].q[k]
    for d, t in enumerate(_TAGS):
        for k, a in enumerate(_AX):
            row[f"se_q{a}_{t}"] = r.run[d].se[k]
    for i, a in enumerate(_AX):
        for j, b in enumerate(_AX):
            row[f"K_{a}{b}"] = r.K[3 * i + j]
    for key in ("k_xy", "k_xz", "k_yz", "recip_xy", "recip_xz", "recip_yz", "recip_resid"):
        row[key] = getattr(r, key)
    row["k1"], row["k2"], row["k3"] = r.k[0], r.k[1], r.k[2]
    row["k_mean"], row["fa"] = r.k_mean, r.fa
    for j in range(3):                       # :613-614, v<j>_<i> = vec[i][j]
        for i, a in enumerate(_AX):
            row[f"v{j + 1}_{a}"] = r.v[3 * i + j]
    row["n_pore"] = int(r.n_pore)
    row["exit_code"] = _lib.exit_code(st)
    row["status"] = st
    row["seconds"] = r.seconds
    row["fluid_updates"] = int(r.fluid_updates)
    if r.validated:
        row["val_q"] = list(r.val.q)
        row["val_pred"] = list(r.val_pred)
        row["val_err"] = r.val_err
        row["val_steps"] = int(r.val.steps)
        row["val_converged"] = int(r.val.converged)
    return row


# ---------------------------------------------------------------------------
# 2D
# ---------------------------------------------------------------------------

def config2d(ff=1.0, eps=1e-3, lag=5000, hold=3, min_steps=10000, max_steps=500000,
             avg_steps=20000, avg_sample=50, qfloor=1e-2, ksigma=3.0, L0=None,
             validate=None, precision="f32ref", concurrent=True, shape=None):
    c = _lib.Config2D()
    c.ff, c.eps, c.qfloor, c.ksigma = float(ff), float(eps), float(qfloor), float(ksigma)
    if L0 is None:                     # lbm.h:24-27: L0 = 4 at 256, scaled with the lattice
        L0 = 4.0 * (shape[1] / 256.0) if shape is not None else 4.0
    c.L0 = float(L0)
    c.validate_deg = -1.0 if validate is None else float(validate)
    c.lag, c.min_steps, c.max_steps = int(lag), int(min_steps), int(max_steps)
    c.avg_steps, c.avg_sample, c.hold = int(avg_steps), int(avg_sample), int(hold)
    c.precision = _precision(precision)
    c.concurrent = 1 if concurrent else 0
    return c


def _mask2d(mask_or_path):
    if isinstance(mask_or_path, (str, Path)):
        return np.ascontiguousarray(load_structure2d(mask_or_path)), str(mask_or_path)
    arr = np.asarray(mask_or_path)
    if arr.ndim != 2:
        raise StructureError(f"expected a (ly, lx) mask, got shape {arr.shape}")
    return np.ascontiguousarray(arr.astype(np.uint8, copy=False)), "<array>"


def solve_permeability2d(mask_or_path, *, ctx=None, filename=None, velocity=False, progress=None,
                         **cfg):
    """One lbm2d-perm invocation on a mask ([y, x] array or GIF/PPM/DAT path).

    Keyword arguments: ff, eps, lag, hold, min_steps, max_steps, avg_steps,
    avg_sample, qfloor, ksigma, validate (degrees), L0, precision ("f32ref" =
    the reference's float arithmetic, "f64", "f32" = shifted), concurrent.
    The row carries the last first-phase check of each run ('check_fx',
    'check_fy': step, q = block means, ratio = [dx/tol, dy/tol], hold);
    progress(run, check) is called after every check (run 0 fx, 1 fy, 2 val).
    """
    ctx = ctx or default_context()
    mask, name = _mask2d(mask_or_path)
    ly, lx = mask.shape
    c = config2d(shape=mask.shape, **cfg)
    r = _lib.Result2D()
    vel = np.zeros((2, ly, lx, 2), np.float32) if velocity else None
    st = _call(ctx, progress, ctx.lib.fp_solve2d, ctx.handle, _lib.u8ptr(mask), lx, ly, C.byref(c),
               C.byref(r), _lib.f32ptr(vel))
    if st not in (_lib.FP_OK, _lib.FP_NOT_CONVERGED):
        _raise(ctx, st)
    row = row2d(r, st, c.ff, filename if filename is not None else name)
    if velocity:
        row["velocity"] = vel
    return row


def row2d(r, st, ff, filename):
    """The lbm2d-perm CSV row (LMB2d/main.cpp:382-411) of a Result2D as a dict,
    plus exit_code, status, timing and the --validate values."""
    rx, ry = r.run[0], r.run[1]
    row = {"filename": filename,
           "ff": ff, "force": r.force, "mu": r.mu, "porosity": r.porosity,
           "steps_fx": int(rx.steps), "conv_fx": int(rx.converged),
           "steps_fy": int(ry.steps), "conv_fy": int(ry.converged),
           "qx_fx": rx.q[0], "qy_fx": rx.q[1], "qx_fy": ry.q[0], "qy_fy": ry.q[1],
           "se_qx_fx": rx.se[0], "se_qy_fx": rx.se[1], "se_qx_fy": ry.se[0], "se_qy_fy": ry.se[1],
           "k_xx": r.k_xx, "k_yx": r.k_yx, "k_xy": r.k_xy, "k_yy": r.k_yy,
           "k_off": r.k_off, "recip_resid": r.recip_resid,
           "k1": r.k1, "k2": r.k2, "k_mean": r.k_mean, "anisotropy": r.anisotropy,
           "theta_deg": r.theta_deg, "alpha_x_deg": r.alpha_x_deg, "alpha_y_deg": r.alpha_y_deg,
           "fieldres_fx": rx.fieldres, "fieldres_fy": ry.fieldres,
           "check_fx": _check_of(rx), "check_fy": _check_of(ry),
           "exit_code": _lib.exit_code(st), "status": st, "seconds": r.seconds,
           "fluid_updates": int(r.fluid_updates)}
    if r.validated:
        row["val_q"] = list(r.val.q[:2])
        row["val_pred"] = list(r.val_pred)
        row["val_err"] = r.val_err
        row["val_steps"] = int(r.val.steps)
        row["val_converged"] = int(r.val.converged)
    return row
