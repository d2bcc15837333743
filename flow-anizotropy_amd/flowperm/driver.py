"""
Dataset driver: solve a generated structure set and collect the tensors, in
one process, through the batched request server.

This is the workflow of scripts/solve_permeability3d.py (and, with --dim 2,
scripts/solve_permeability.py) with the process pool replaced by continuous
batching on the GPU.  The on-disk contract is the reference's:

* <dataset>/structures.csv lists sample_id and filename (+ generator columns);
* every solve writes <dataset>/perm/<sample_id:06d>.csv, the solver's header
  and row, through <out>.tmp.<pid> and os.replace, so a kill at any instant
  leaves either a complete file or nothing
  (scripts/solve_permeability3d.py:113-146);
* a file counts as done only if valid_result() parses it
  (scripts/solve_permeability3d.py:87-110);
* --verify / --repair audit and clean perm/, --shard I/N and --interleave COL
  split and order the work (:161-211);
* --merge-only (and the end of every unsharded run) joins the rows onto
  structures.csv with the derived targets, writes permeability.csv and its
  manifest, and prints the label-noise ceilings (:269-368;
  scripts/metrics.py:85-108; 2D: scripts/solve_permeability.py:100-141).

    python -m flowperm.driver --dataset DATA/aniso3d --verify
    python -m flowperm.driver --dataset DATA/aniso3d --precision f32
    python -m flowperm.driver --dataset DATA/aniso3d --merge-only
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time
from datetime import datetime

import numpy as np

REQUIRED_COLS3 = ("K_xx", "K_yy", "K_zz", "k_xy", "k_xz", "k_yz", "conv_fx", "conv_fy", "conv_fz")
REQUIRED_COLS2 = ("k_xx", "k_yy", "k_xy", "k_yx", "conv_fx", "conv_fy")
PAIRS = (("x", "y"), ("x", "z"), ("y", "z"))


def valid_result(path, required=REQUIRED_COLS3) -> bool:
    """True when `path` holds a complete solve: one data row whose required
    columns are all present and finite (scripts/solve_permeability3d.py:87-110)."""
    import pandas as pd
    try:
        if os.path.getsize(path) <= 0:
            return False
        one = pd.read_csv(path)
    except Exception:
        return False
    if len(one) < 1 or not set(required) <= set(one.columns):
        return False
    return bool(np.isfinite(one[list(required)].to_numpy(float)).all())


def koff_ceilings3d(df):
    """Upper bound on the R^2 achievable for each 3D off-diagonal target, from
    K = K^T: |K_ij - K_ji| / 2 estimates the error of the symmetrised k_ij,
    so R^2 <= 1 - mean(err^2) / var(k_ij) (scripts/metrics.py:85-108).
    None when the columns are not 3D ones."""
    out = {}
    for a, b in PAIRS:
        need = {f"K_{a}{b}", f"K_{b}{a}", f"k_{a}{b}"}
        if not need <= set(df.columns):
            return None
        err = (df[f"K_{a}{b}"] - df[f"K_{b}{a}"]).abs() / 2.0
        var_sig = float(df[f"k_{a}{b}"].var())
        out[f"k_{a}{b}"] = (float(1.0 - float((err ** 2).mean()) / var_sig)
                            if var_sig > 0 else float("nan"))
    return out


def derive3d(out):
    """Derived targets of the 3D merge (scripts/solve_permeability3d.py:294-322)."""
    trace = out["K_xx"] + out["K_yy"] + out["K_zz"]
    out["trace"] = trace
    out["log_k_mean"] = np.log(out["k_mean"].where(out["k_mean"] > 0))
    out["k_ratio"] = out["k1"] / out["k3"].where(out["k3"] > 0)
    out["log_k_ratio"] = np.log(out["k_ratio"].where(out["k_ratio"] > 0))
    for a, b in PAIRS:
        sym = out[f"k_{a}{b}"]
        asym = (out[f"K_{a}{b}"] - out[f"K_{b}{a}"]).abs()
        out[f"koff_rel_err_{a}{b}"] = asym / (2 * sym).abs()
        out[f"koff_over_trace_{a}{b}"] = sym.abs() / trace.abs()
    km = out["k_mean"].where(out["k_mean"] > 0)
    out["dev_xx"] = out["K_xx"] / km - 1.0
    out["dev_yy"] = out["K_yy"] / km - 1.0
    for a, b in PAIRS:
        out[f"dev_{a}{b}"] = out[f"k_{a}{b}"] / km
    out = out.sort_values("sample_id").reset_index(drop=True)
    out["conv_all"] = ((out.conv_fx == 1) & (out.conv_fy == 1) & (out.conv_fz == 1)).astype(int)
    return out


def derive2d(out):
    """Derived targets of the 2D merge (scripts/solve_permeability.py:119-131)."""
    t = np.radians(out["theta_deg"])
    out["cos2theta"] = np.cos(2 * t)
    out["sin2theta"] = np.sin(2 * t)
    out["log_k_mean"] = np.log(out["k_mean"].where(out["k_mean"] > 0))
    out["k_ratio"] = out["k1"] / out["k2"].where(out["k2"] > 0)
    out["log_k_ratio"] = np.log(out["k_ratio"].where(out["k_ratio"] > 0))
    out["koff_rel_err"] = (out["k_xy"] - out["k_yx"]).abs() / (2 * out["k_off"]).abs()
    out["koff_over_trace"] = out["k_off"].abs() / (out["k_xx"] + out["k_yy"]).abs()
    return out


def manifest(dataset, dest, df, out):
    """What a merge snapshot holds (scripts/solve_permeability3d.py:324-346)."""
    ids = out.sample_id.to_numpy()
    meta = {
        "dataset": dataset,
        "labels": dest,
        "merged": datetime.now().isoformat(timespec="seconds"),
        "n_structures": int(len(df)),
        "n_rows": int(len(out)),
        "n_converged": int(out.conv_all.sum()),
        "sample_id_min": int(ids.min()),
        "sample_id_max": int(ids.max()),
        "contiguous": bool(len(ids) == ids.max() - ids.min() + 1),
        "sha256_ids": hashlib.sha256(",".join(str(int(i)) for i in ids).encode()).hexdigest(),
    }
    if "stratum" in out.columns:
        meta["strata"] = {str(k): int(v) for k, v in out.stratum.value_counts().items()}
    return meta


def per_sample_path(perm_dir, sid):
    return os.path.join(perm_dir, f"{int(sid):06d}.csv")


def interleaved(df, column):
    """Round-robin order over `column`: the i-th sample of every group before
    the (i+1)-th of any, stable within a group (:204-207)."""
    rank = df.groupby(column, sort=False).cumcount()
    return df.assign(_r=rank).sort_values(["_r", column], kind="stable").drop(columns="_r")


def verify(perm_dir, expected, required, repair=False, shard_ids=None, log=print):
    good = [sid for sid, p in expected if os.path.exists(p) and valid_result(p, required)]
    bad = [sid for sid, p in expected if os.path.exists(p) and not valid_result(p, required)]
    missing = [sid for sid, p in expected if not os.path.exists(p)]
    strays = [f for f in os.listdir(perm_dir) if ".tmp." in f]
    if shard_ids is not None:
        strays = [f for f in strays if int(f.split(".")[0]) in shard_ids]
    log(f"  complete   {len(good)}")
    log(f"  incomplete {len(bad)}" + (f"  e.g. {bad[:5]}" if bad else ""))
    log(f"  missing    {len(missing)}"
        + (f"  first {missing[:3]} ... last {missing[-3:]}" if missing else ""))
    log(f"  stray .tmp {len(strays)}")
    if repair:
        for sid in bad:
            os.remove(per_sample_path(perm_dir, sid))
        for f in strays:
            os.remove(os.path.join(perm_dir, f))
        log(f"  repaired: removed {len(bad)} incomplete + {len(strays)} stray files")
    elif bad or strays:
        log("  rerun with --repair to clear them, then resume normally")
    return {"complete": good, "incomplete": bad, "missing": missing, "strays": strays}


def write_row(path, header, line):
    """header + row into <path>.tmp.<pid>, then os.replace (atomic)."""
    tmp = f"{path}.tmp.{os.getpid()}"
    with open(tmp, "w") as fh:
        fh.write(header)
        fh.write(line)
    os.replace(tmp, path)


def solve_set(tasks, dim, cfg, precision="f64", budget_bytes=0, window=512, ctx=None,
              progress_every=30.0, log=print):
    """Solve (sample_id, structure_path, out_csv) tasks through one request
    server; each finished row is written with write_row as it completes.
    Returns (n_ok, n_unconverged, failed list of (sample_id, status, message))."""
    from . import solver, structio
    from ._lib import FlowpermError
    from .server import Server
    cols = solver.COLUMNS3D if dim == 3 else solver.COLUMNS2D
    header = solver.csv_header(cols)
    n_ok = n_unconv = 0
    failed = []
    inflight = {}                        # request id -> (sample_id, out path)
    it = iter(tasks)
    exhausted = False
    t_last = time.time()
    done = 0
    with Server(dim, precision, budget_bytes=budget_bytes, ctx=ctx) as srv:
        while True:
            while not exhausted and len(inflight) < window:
                try:
                    sid, structure, out = next(it)
                except StopIteration:
                    exhausted = True
                    break
                try:
                    mask = (structio.load_raw(structure) if dim == 3
                            else structio.load_structure2d(structure))
                    rid = srv.submit(mask, filename=structure, **cfg)
                except (FlowpermError, structio.StructureError, OSError, ValueError) as e:
                    failed.append((sid, 2, str(e)))
                    continue
                inflight[rid] = (sid, out)
            busy = srv.step()
            for rid in list(inflight):
                try:
                    res = srv.poll(rid)
                except FlowpermError as e:
                    sid, _ = inflight.pop(rid)
                    failed.append((sid, 2, str(e)))
                    continue
                if res is None:
                    continue
                sid, out = inflight.pop(rid)
                row, _ = res
                write_row(out, header, solver.csv_row(row, cols))
                done += 1
                if row["exit_code"] == 3:
                    n_unconv += 1
                else:
                    n_ok += 1
            if time.time() - t_last > progress_every:
                t_last = time.time()
                st = srv.stats()
                log(f"  ... {done} solved, {len(inflight)} in flight "
                    f"({st['resident']} resident, {st['queued']} queued), "
                    f"{st['fluid_updates'] / max(st['busy_seconds'], 1e-9) / 1e6:.0f} MLUPS")
            if exhausted and not inflight and busy == 0:
                break
    return n_ok, n_unconv, failed


def merge(dataset, df, perm_dir, dim, log=print):
    import pandas as pd
    required = REQUIRED_COLS3 if dim == 3 else REQUIRED_COLS2
    parts, unreadable = [], []
    for r in df.itertuples():
        pth = per_sample_path(perm_dir, r.sample_id)
        if not os.path.exists(pth):
            continue
        if not valid_result(pth, required):
            unreadable.append(r.sample_id)
            continue
        one = pd.read_csv(pth)
        one["sample_id"] = r.sample_id
        parts.append(one)
    if unreadable:
        log(f"WARNING: {len(unreadable)} per-sample CSVs are incomplete and were skipped "
            f"(e.g. {unreadable[:5]}).")
        log("         Run with --verify --repair, then resume, to redo them.")
    if not parts:
        raise SystemExit("nothing to merge")
    perm = pd.concat(parts, ignore_index=True).drop(columns=["filename"])
    out = df.merge(perm, on="sample_id", how="inner", suffixes=("", "_solver"))
    dest = os.path.join(dataset, "permeability.csv")
    if dim == 3:
        out = derive3d(out)
        out.to_csv(dest, index=False)
        meta = manifest(dataset, dest, df, out)
        man = os.path.splitext(dest)[0] + "_manifest.json"
        with open(man, "w") as fh:
            json.dump(meta, fh, indent=2)
        log(f"merged {len(out)}/{len(df)} rows -> {dest}")
        log(f"  manifest -> {man}")
        conv = out.conv_all == 1
        log(f"all three directions converged: {int(conv.sum())}/{len(out)}")
        ceil = koff_ceilings3d(out.loc[conv])
        if ceil and conv.sum() > 1:
            log("  label-noise ceiling on R^2 (from K = K^T)")
            for t, c in ceil.items():
                log(f"    {t}: R^2 <= {c:.6f}")
    else:
        out = derive2d(out)
        out.to_csv(dest, index=False)
        log(f"merged {len(out)} rows -> {dest}")
        conv = (out.conv_fx == 1) & (out.conv_fy == 1)
        log(f"both directions converged: {int(conv.sum())}/{len(out)}")
    return out


def parse_extra(extra, dim):
    """The reference drivers' --extra solver flags as keyword arguments."""
    from .cli import SPEC2, SPEC3, _parse
    spec = dict(SPEC3 if dim == 3 else SPEC2)
    for k in ("--velocity", "--quiet"):
        spec.pop(k, None)
    return _parse(extra.split(), spec, "", "driver")


def main(argv=None):
    import pandas as pd
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--dataset", required=True)
    ap.add_argument("--dim", type=int, default=3, choices=[2, 3])
    ap.add_argument("--ff", type=float, default=1.0)
    ap.add_argument("--limit", type=int, default=0, help="solve only the first N")
    ap.add_argument("--merge-only", action="store_true")
    ap.add_argument("--shard", default="", help="I/N: only samples with sample_id %% N == I; no merge")
    ap.add_argument("--interleave", default="", help="solve order round-robin over this column")
    ap.add_argument("--verify", action="store_true")
    ap.add_argument("--repair", action="store_true")
    ap.add_argument("--extra", default="", help='solver flags, e.g. "--eps 5e-4"')
    ap.add_argument("--precision", default=None,
                    help="f64 (3D default), f32 (shifted), f32ref (2D default)")
    ap.add_argument("--budget-gb", type=float, default=0.0,
                    help="device memory for resident lattices (default 90%% of free)")
    ap.add_argument("--window", type=int, default=512, help="structures read ahead")
    a = ap.parse_args(argv)

    struct_csv = os.path.join(a.dataset, "structures.csv")
    if not os.path.exists(struct_csv):
        raise SystemExit(f"no structures.csv in {a.dataset}")
    df = pd.read_csv(struct_csv)
    if a.limit:
        df = df.head(a.limit)
    shard_ids = None
    if a.shard:
        if a.merge_only:
            raise SystemExit("--shard and --merge-only exclude each other: merge runs over all shards")
        si, sn = (int(v) for v in a.shard.split("/"))
        if not 0 <= si < sn:
            raise SystemExit(f"bad --shard {a.shard}")
        df = df[df.sample_id % sn == si].reset_index(drop=True)
        shard_ids = set(int(v) for v in df.sample_id)
        print(f"shard {si}/{sn}: {len(df)} samples")
    perm_dir = os.path.join(a.dataset, "perm")
    os.makedirs(perm_dir, exist_ok=True)
    structures_dir = os.path.join(a.dataset, "structures")
    required = REQUIRED_COLS3 if a.dim == 3 else REQUIRED_COLS2
    expected = [(int(r.sample_id), per_sample_path(perm_dir, r.sample_id)) for r in df.itertuples()]

    if a.verify:
        print(f"{a.dataset}: {len(expected)} structures")
        verify(perm_dir, expected, required, a.repair, shard_ids)
        return 0

    if not a.merge_only:
        order = interleaved(df, a.interleave) if a.interleave else df
        cfg = parse_extra(a.extra, a.dim)
        cfg["ff"] = a.ff
        tasks = [(int(r.sample_id), os.path.join(structures_dir, r.filename),
                  per_sample_path(perm_dir, r.sample_id))
                 for r in order.itertuples()
                 if not (os.path.exists(per_sample_path(perm_dir, r.sample_id))
                         and valid_result(per_sample_path(perm_dir, r.sample_id), required))]
        print(f"{len(expected)} structures: {len(expected) - len(tasks)} already solved, "
              f"{len(tasks)} to do")
        precision = a.precision or ("f64" if a.dim == 3 else "f32ref")
        t0 = time.time()
        n_ok, n_unconv, failed = solve_set(tasks, a.dim, cfg, precision,
                                           int(a.budget_gb * 2 ** 30), a.window)
        print(f"solved {n_ok + n_unconv} in {time.time() - t0:.1f} s   non-converged (kept, "
              f"flagged) {n_unconv}   failed {len(failed)}")
        for sid, st, msg in failed[:10]:
            print(f"  {sid:06d}: status {st} {msg}")
        if shard_ids is not None:
            print("shard finished -- not merging; run --merge-only once every shard is done")
            return 0
    merge(a.dataset, df, perm_dir, a.dim)
    return 0


if __name__ == "__main__":
    sys.exit(main())
