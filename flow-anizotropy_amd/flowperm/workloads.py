"""
Seeded structure sets of the solver's measured workloads.

* ``batched_structure(i, seed)``: the batched-solve set: a cube with edge drawn
  uniform from {48, 64, 96, 128, 160}, overlapping 3:1 ellipsoids along a
  seeded random axis with 15-degree jitter, porosity ~ U(0.3, 0.8).
* ``serving_structure(i, seed)``: the serving set: cube edge log-uniform in
  [32, 256], spheres (radius U[3, 6]) or a Gaussian random field (correlation
  lengths 8, 8, 24 voxels, its percolating cluster at the target porosity) with
  equal odds, porosity ~ U(0.2, 0.8).
* ``large_grf(seed)``: the 512^3 low-porosity Gaussian-random-field lattice:
  its largest cluster that percolates, with porosity ~ U(0.15, 0.25) (the
  threshold is corrected for the pockets the cluster reduction removes).

Every mask is kept only if its pore space wraps the torus along x, y and z
(the test of scripts/gen_aniso_structures3d.py:184-201): a direction without a
through-path has no steady flux to converge to.  A structure that fails is
redrawn from the next sub-seed.
"""
from __future__ import annotations

import numpy as np

from . import masks

BATCHED_EDGES = (48, 64, 96, 128, 160)


def _percolating(m):
    return all(masks.percolates(m))


def batched_structure(i: int, seed: int = 0) -> np.ndarray:
    for attempt in range(64):
        rng = np.random.default_rng([seed, i, attempt])
        edge = int(rng.choice(BATCHED_EDGES))
        por = float(rng.uniform(0.3, 0.8))
        m = masks.ellipsoids((edge,) * 3, por, rmin=2.0, rmax=4.0, aspect=3.0, jitter_deg=15.0,
                             seed=int(rng.integers(1 << 31)))
        if _percolating(m):
            return m
    raise RuntimeError(f"no percolating batched structure for index {i}")


def serving_structure(i: int, seed: int = 0) -> np.ndarray:
    for attempt in range(64):
        rng = np.random.default_rng([seed, 1000003 + i, attempt])
        edge = int(round(np.exp(rng.uniform(np.log(32.0), np.log(256.0)))))
        por = float(rng.uniform(0.2, 0.8))
        sub = int(rng.integers(1 << 31))
        if rng.random() < 0.5:
            m = masks.spheres((edge,) * 3, por, 3.0, 6.0, seed=sub)
            cl = masks.percolating_cluster(m, require_all=True)
            m = cl if cl is not None else m
        else:
            m = masks.grf((edge,) * 3, porosity=por, corr=(8.0, 8.0, 24.0), seed=sub,
                          require_all=True, cluster_porosity=True)
        if _percolating(m) and (m == 0).any():
            return m
    raise RuntimeError(f"no percolating serving structure for index {i}")


def large_grf(seed: int = 0, n: int = 512) -> np.ndarray:
    for attempt in range(16):
        rng = np.random.default_rng([seed, 7, attempt])
        por = float(rng.uniform(0.15, 0.25))
        m = masks.grf((n, n, n), porosity=por, corr=(8.0, 8.0, 24.0),
                      seed=int(rng.integers(1 << 31)), require_all=True, cluster_porosity=True)
        if _percolating(m) and 0.15 <= 1.0 - float(m.mean()) <= 0.25:
            return m
    raise RuntimeError("no percolating large GRF lattice")


def _gen(args):
    kind, i, seed = args
    return batched_structure(i, seed) if kind == "batched" else serving_structure(i, seed)


def generate(kind: str, n: int, seed: int = 0, workers: int = 1):
    """The first n structures of a set, generated by `workers` processes."""
    jobs = [(kind, i, seed) for i in range(n)]
    if workers <= 1 or n <= 1:
        return [_gen(j) for j in jobs]
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(_gen, jobs))
