"""
Request server over the C-ABI (fp_server_*): whole lbm3d-perm / lbm2d-perm
invocations admitted by device-memory footprint into one continuously batched
stepper, lattice state resident on the device until each request finishes.

The reference has no such layer: its drivers start one solver process per
structure from a process pool (scripts/solve_permeability3d.py:250) and
collect the CSV rows; here each request's row comes back from ``poll`` with
the same column names (flowperm.solver.row3d / row2d).

    srv = Server(lattice=3, precision="f32")
    rid = srv.submit(mask, ff=40, lag=1000, eps=1e-6)
    while srv.step():          # admit, one chunk of steps, retire
        ...
    row, info = srv.poll(rid)
"""
from __future__ import annotations

import ctypes as C
import threading

import numpy as np

from . import _lib
from ._lib import FlowpermError
from .solver import Context, _mask2d, _mask3d, _precision, config2d, config3d, row2d, row3d


class Server:
    def __init__(self, lattice: int = 3, precision="f32", budget_bytes: int = 0,
                 chunk_steps: int = 0, ctx: Context | None = None, max_bypass: int = 0):
        self.ctx = ctx or Context()
        self._own_ctx = ctx is None
        self.lib = self.ctx.lib
        self.lattice = int(lattice)
        self.precision = precision
        cfg = _lib.ServerConfig(self.lattice, _precision(precision), int(budget_bytes),
                                int(chunk_steps), int(max_bypass))
        st = C.c_int(0)
        self.handle = self.lib.fp_server_create(self.ctx.handle, C.byref(cfg), C.byref(st))
        if not self.handle:
            raise FlowpermError(st.value, "fp_server_create failed")
        self._meta = {}                  # id -> (ff, filename)
        self._lock = threading.Lock()

    def now(self) -> float:
        """The library clock (request times are on it)."""
        return self.lib.fp_now()

    def submit(self, mask_or_path, *, filename=None, **cfg) -> int:
        """Queue one solve; the keyword arguments are those of
        solve_permeability3d / solve_permeability2d (precision is the server's)."""
        cfg = dict(cfg)
        cfg["precision"] = self.precision
        if self.lattice == 3:
            mask, name = _mask3d(mask_or_path)
            nz, ny, nx = mask.shape
            c = config3d(**cfg)
            rid = self.lib.fp_server_submit3d(self.handle, _lib.u8ptr(mask), nx, ny, nz, C.byref(c))
        else:
            mask, name = _mask2d(mask_or_path)
            ly, lx = mask.shape
            c = config2d(shape=mask.shape, **cfg)
            rid = self.lib.fp_server_submit2d(self.handle, _lib.u8ptr(mask), lx, ly, C.byref(c))
        if rid < 0:
            raise FlowpermError(-rid, f"request rejected: {self.ctx.last_error()}")
        with self._lock:
            self._meta[rid] = (c.ff, filename if filename is not None else name)
        return int(rid)

    def step(self) -> int:
        """Admit what fits, run one chunk, retire finished requests; returns the
        number of requests still queued or resident (0 = idle)."""
        pending = C.c_int64(0)
        st = self.lib.fp_server_step(self.handle, C.byref(pending))
        if st != _lib.FP_OK:
            raise FlowpermError(st, f"server step failed: {self.ctx.last_error()}")
        return int(pending.value)

    def poll(self, rid: int):
        """None while the request is pending; else (row, info) where row is the
        CSV row dict (raises FlowpermError for a failed request) and info holds
        t_submit / t_admit / t_done (library clock), device_bytes,
        fluid_updates."""
        info = _lib.RequestInfo()
        if self.lattice == 3:
            r = _lib.Result3D()
            st = self.lib.fp_server_poll3d(self.handle, rid, C.byref(r), C.byref(info))
        else:
            r = _lib.Result2D()
            st = self.lib.fp_server_poll2d(self.handle, rid, C.byref(r), C.byref(info))
        if st == _lib.FP_PENDING:
            return None
        if st == _lib.FP_ERR_NOT_FOUND:
            raise FlowpermError(st, f"unknown request {rid}")
        with self._lock:
            ff, name = self._meta.pop(rid)
        d = {k: getattr(info, k) for k, _ in info._fields_}
        if st not in (_lib.FP_OK, _lib.FP_NOT_CONVERGED):
            raise FlowpermError(st, f"request {rid} failed")
        row = row3d(r, st, ff, name) if self.lattice == 3 else row2d(r, st, ff, name)
        return row, d

    def run(self, ids=None):
        """Step until idle; returns {id: (row, info)} for ids (default: all
        submitted and not yet polled)."""
        while self.step():
            pass
        with self._lock:
            want = list(self._meta) if ids is None else list(ids)
        return {i: self.poll(i) for i in want}

    def stats(self) -> dict:
        s = _lib.ServerStats()
        self.lib.fp_server_get_stats(self.handle, C.byref(s))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def close(self):
        if getattr(self, "handle", None):
            self.lib.fp_server_destroy(self.handle)
            self.handle = None
        if getattr(self, "_own_ctx", False) and self.ctx is not None:
            self.ctx.close()
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poisson_arrivals(rate_per_s: float, n: int, seed: int = 0) -> np.ndarray:
    """Arrival times (seconds from 0) of n requests of a Poisson process."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_per_s, size=n))


def serve_schedule(server: Server, masks, arrivals, cfg: dict, sleep=None):
    """Drive the server with requests arriving at the given times (seconds
    from the start): a request is submitted once its time has passed, the
    server steps whenever it has work and waits for the next arrival
    otherwise.  Returns a list of dicts per request: arrival, t_submit,
    t_admit, t_done (all relative to the start), latency = t_done - arrival,
    and the row."""
    import time
    sleep = sleep or time.sleep
    t0 = server.now()
    n = len(masks)
    ids = [None] * n
    out = [None] * n
    nxt = 0
    while True:
        now = server.now() - t0
        while nxt < n and arrivals[nxt] <= now:
            ids[nxt] = server.submit(masks[nxt], filename=f"req{nxt}", **cfg)
            nxt += 1
        busy = server.step()                 # admit, one chunk (if anything is resident), retire
        for i in range(nxt):
            if out[i] is None:
                res = server.poll(ids[i])
                if res is not None:
                    row, info = res
                    out[i] = {"arrival": float(arrivals[i]),
                              "t_submit": info["t_submit"] - t0, "t_admit": info["t_admit"] - t0,
                              "t_done": info["t_done"] - t0,
                              "latency": info["t_done"] - t0 - float(arrivals[i]),
                              "device_bytes": info["device_bytes"],
                              "fluid_updates": info["fluid_updates"], "row": row}
        if busy == 0:
            if nxt >= n:
                break
            wait = arrivals[nxt] - (server.now() - t0)
            if wait > 0:
                sleep(wait)
    return out
