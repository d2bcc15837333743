"""
Axis-permutation covariance of the 3D solve, every orientation in one batch.

This is the check of scripts/rotation_covariance3d.py. The reference runs one solver
process per orientation (`:237-251`). Here the identity and the chosen permutations of
one structure are submitted together to a request server, so all 3 x (1 + n_perms)
forcing runs step side by side on the device.

Relabelling the axes of a periodic grid changes no voxel. The tensor of the permuted
structure must therefore be the permuted tensor, K' = P K P^T, i.e.
K'_ab = K_{sigma[a] sigma[b]} (`:103-105`).

* **Residual.** max |K' - P K P^T| / trace K must stay below `tol` (default 1e-6, as
  `:194-195`). In FP64 it is round-off: the permuted run visits the directions and the
  velocity components in another order.
* **Discriminating scale.** max |K - P K P^T| / trace K says how far the permutation
  moves K at all. A run whose scale is below `min_scale` has tested nothing and fails
  (`:196-199`, `:274-280`).
* **Steps.** The permuted run's step counts are the reference's, permuted. A mismatch is
  reported as a convergence difference (`:291-301`).

    python -m flowperm.covariance                       # 40^3 strong structure, all 5 perms
    python -m flowperm.covariance --size 32 --perms yxz
    python -m flowperm.covariance --structure x.raw --precision f32
    python -m flowperm.covariance --self-test           # no GPU needed
"""
from __future__ import annotations

import argparse
import itertools
import sys
import tempfile
from pathlib import Path

import numpy as np

AXES = "xyz"
# sigma[a] is the old solver axis that becomes new axis a; the label spells the old
# axes in new-axis order ("yzx": new x is old y), as rotation_covariance3d.py:66-71
PERMS = {"".join(AXES[s] for s in sigma): tuple(sigma)
         for sigma in itertools.permutations(range(3))}
IDENTITY = (0, 1, 2)


def permute_structure(arr, sigma):
    """The structure with its solver axes relabelled.

    `arr` is indexed [z, y, x] (the .raw order). New voxel (x, y, z) = i holds old voxel
    j with j[sigma[a]] = i[a]. Solver axis a sits at array position 2 - a, so new
    array axis 2 - a is old array axis 2 - sigma[a].
    """
    arr = np.asarray(arr)
    order = [2 - sigma[2 - d] for d in range(3)]
    return np.ascontiguousarray(arr.transpose(order))


def inverse(sigma):
    inv = [0, 0, 0]
    for a, s in enumerate(sigma):
        inv[s] = a
    return tuple(inv)


def expected(K, sigma):
    """P K P^T: K'_ab = K_{sigma[a] sigma[b]}."""
    s = list(sigma)
    return np.asarray(K)[np.ix_(s, s)]


def k_matrix(row):
    """The full (unsymmetrised) K of a CSV row: K[r][c] = flux r under force c."""
    return np.array([[row[f"K_{r}{c}"] for c in AXES] for r in AXES], float)


def steps_of(row):
    return tuple(int(row[f"steps_f{a}"]) for a in AXES)


def compare(results, labels, tol=1e-6, min_scale=1e-3):
    """Per-permutation verdicts against the identity run.

    results: label -> CSV row dict (must hold "xyz"). Returns (ok, lines) where
    each line is a dict: label, resid, scale, steps, steps_expected, good."""
    K0 = k_matrix(results["xyz"])
    trace = float(np.trace(K0))
    s0 = np.array(steps_of(results["xyz"]))
    ok = True
    lines = []
    for label in labels:
        if label == "xyz":
            continue
        sigma = PERMS[label]
        K = k_matrix(results[label])
        want = expected(K0, sigma)
        resid = float(np.abs(K - want).max()) / trace
        scale = float(np.abs(K0 - want).max()) / trace
        good = bool(resid < tol and scale >= min_scale)
        ok &= good
        lines.append({"label": label, "resid": resid, "scale": scale,
                      "steps": steps_of(results[label]),
                      "steps_expected": tuple(int(v) for v in s0[list(sigma)]),
                      "good": good})
    return ok, lines


def self_test(log=print) -> int:
    """Solver-free checks of the permutation bookkeeping (the checks of
    rotation_covariance3d.py:112-152), on a non-cubic grid so that swapping two equal
    axes cannot hide a mistake."""
    ok = True

    def check(label, cond):
        nonlocal ok
        ok = ok and bool(cond)
        log(f"[{'ok' if cond else 'FAIL'}] {label}")

    rng = np.random.default_rng(0)
    arr = (rng.random((5, 7, 11)) > 0.5).astype(np.uint8)       # [z, y, x]
    n = np.array([11, 7, 5])                                    # (nx, ny, nz)
    src = arr.transpose(2, 1, 0)                                # [x, y, z]
    for label, sigma in PERMS.items():
        p = permute_structure(arr, sigma)
        want_shape = tuple(int(v) for v in n[list(sigma)][::-1])
        check(f"{label}: shape {p.shape} == {want_shape}", p.shape == want_shape)
        check(f"{label}: porosity preserved", int(p.sum()) == int(arr.sum()))
        check(f"{label}: inverse round-trips",
              np.array_equal(permute_structure(p, inverse(sigma)), arr))
        got = p.transpose(2, 1, 0)
        idx = np.indices(got.shape)
        j = [None] * 3
        for a, s in enumerate(sigma):
            j[s] = idx[a]
        check(f"{label}: every voxel maps to old coords", np.array_equal(got, src[tuple(j)]))
    K = rng.random((3, 3))
    for label, sigma in PERMS.items():
        check(f"{label}: expected() inverts",
              np.array_equal(expected(expected(K, sigma), inverse(sigma)), K))
        P = np.eye(3)[list(sigma)]                              # row a = e_{sigma[a]}
        check(f"{label}: expected() == P K P^T", np.allclose(expected(K, sigma), P @ K @ P.T))
    log("\nself-test " + ("PASSED" if ok else "FAILED"))
    return 0 if ok else 1


def make_structure(size=40, seed=7, porosity=0.75):
    """A strongly anisotropic periodic structure that percolates along x, y and z.

    The long axes of 3:1 ellipsoids lie along an oblique direction with three unequal
    components, so that every permutation moves K (rotation_covariance3d.py:159-182
    picks its strongest-stratum draw for the same reason). The first seed from `seed`
    on whose mask percolates is used."""
    from . import masks
    axis = np.array([0.82, 0.49, 0.30])
    for s in range(seed, seed + 50):
        m = masks.ellipsoids((size,) * 3, porosity, rmin=2.0, rmax=4.0, aspect=3.0,
                             jitter_deg=10.0, seed=s, axis=axis)
        if all(masks.percolates(m)):
            return m
    raise RuntimeError("no percolating structure in 50 seeds")


def solve_orientations(mask, labels, *, precision="f64", ctx=None, workdir=None, **cfg):
    """Solve the identity and every permutation in `labels` in one request server.

    Each variant is written to a .raw file and read back by the solver, so that the
    identity goes through the same path as the others (rotation_covariance3d.py:227-234).
    Returns label -> CSV row dict."""
    from . import structio
    from .server import Server
    labels = ["xyz"] + [l for l in labels if l != "xyz"]
    own = workdir is None
    tmp = Path(tempfile.mkdtemp(prefix="flowperm_cov_")) if own else Path(workdir)
    try:
        paths = {}
        for label in labels:
            paths[label] = structio.save_raw(tmp / f"perm_{label}.raw",
                                             permute_structure(mask, PERMS[label]))
        with Server(3, precision, ctx=ctx) as srv:
            ids = {label: srv.submit(str(paths[label]), **cfg) for label in labels}
            while srv.step():
                pass
            return {label: srv.poll(rid)[0] for label, rid in ids.items()}
    finally:
        if own:
            for f in tmp.iterdir():
                f.unlink()
            tmp.rmdir()


def main(argv=None) -> int:
    p = argparse.ArgumentParser(description=__doc__,
                                formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--structure", help="use this .raw instead of generating one")
    p.add_argument("--size", type=int, default=40, help="cubic edge when generating")
    p.add_argument("--seed", type=int, default=7)
    p.add_argument("--perms", nargs="+", default=None,
                   help="permutations to run (default: all five non-identity ones)")
    p.add_argument("--tol", type=float, default=1e-6,
                   help="largest |K' - P K P^T| / trace(K) accepted")
    p.add_argument("--min-scale", type=float, default=1e-3,
                   help="a permutation that moves K by less than this proves nothing "
                        "and fails")
    p.add_argument("--max-steps", type=int, default=200000)
    p.add_argument("--eps", type=float, default=1e-5)
    p.add_argument("--ff", type=float, default=1.0)
    p.add_argument("--precision", default="f64", choices=["f64", "f32"])
    p.add_argument("--keep", help="write the permuted structures into this directory")
    p.add_argument("--self-test", action="store_true", dest="selftest")
    a = p.parse_args(argv)
    if a.selftest:
        return self_test()

    labels = a.perms or [k for k, s in PERMS.items() if s != IDENTITY]
    bad = [l for l in labels if l not in PERMS]
    if bad:
        print(f"not a permutation of xyz: {bad}", file=sys.stderr)
        return 2
    if a.structure:
        from . import structio
        mask = structio.load_raw(a.structure)
    else:
        mask = make_structure(a.size, a.seed)
    nz, ny, nx = mask.shape
    print(f"domain     {nx}x{ny}x{nz} (x,y,z)  porosity {1.0 - mask.mean():.6f}")
    if a.keep:
        Path(a.keep).mkdir(parents=True, exist_ok=True)
    print(f"solving {1 + len([l for l in labels if l != 'xyz'])} orientations in one batch "
          f"({a.precision})", flush=True)
    results = solve_orientations(mask, labels, precision=a.precision, workdir=a.keep,
                                 ff=a.ff, eps=a.eps, max_steps=a.max_steps)
    K0 = k_matrix(results["xyz"])
    r0 = results["xyz"]
    print(f"\nreference K (trace {np.trace(K0):.6f}, recip_resid {r0['recip_resid']:.1e}, "
          f"steps {'/'.join(str(s) for s in steps_of(r0))})")
    for r in range(3):
        print("   [ " + "  ".join(f"{K0[r][c]:12.8f}" for c in range(3)) + " ]")
    ok, lines = compare(results, labels, a.tol, a.min_scale)
    print(f"\n{'perm':>6} {'steps x/y/z':>20} {'resid':>10} {'scale':>10} {'resid/scale':>12}")
    print("-" * 64)
    for ln in lines:
        steps = "/".join(str(s) for s in ln["steps"])
        rs = ln["resid"] / ln["scale"] if ln["scale"] else float("nan")
        note = "" if ln["scale"] >= a.min_scale else "  <- vacuous, K too isotropic"
        print(f"{ln['label']:>6} {steps:>20} {ln['resid']:10.2e} {ln['scale']:10.2e} "
              f"{rs:12.2e}  [{'ok' if ln['good'] else 'FAIL'}]{note}")
    for ln in lines:
        if ln["steps"] != ln["steps_expected"]:
            print(f"\nnote: {ln['label']} stopped at steps {ln['steps']}, expected "
                  f"{ln['steps_expected']}: the residual is then partly a convergence "
                  f"difference")
    print("\nrotation covariance", "PASSED" if ok else "FAILED")
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
