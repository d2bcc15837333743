"""
Seeded voxel-mask generators for the permeability workloads, and the torus
percolation filter.

All masks are uint8, 1 = solid, indexed [z, y, x] (3D) or [y, x] (2D), and
periodic: every object wraps across the cell faces, matching the solver's
periodic boundaries.

* ``spheres``      overlapping spheres, radius ~ U[rmin, rmax], centres uniform,
                   added until the porosity reaches the target (the 64^3 parity
                   workload uses r in [3, 6], porosity 0.6).
* ``ellipsoids``   overlapping prolate ellipsoids with aspect ratio 3:1 whose
                   long axes follow one seeded random unit axis with a per-object
                   angular jitter (default 15 degrees); ``ellipses`` is the 2D
                   counterpart.
* ``grf``          Gaussian random field (FFT-filtered white noise with
                   anisotropic Gaussian correlation lengths), thresholded to a
                   porosity and reduced to its largest percolating pore cluster
                   (optionally with the threshold corrected so that the kept
                   cluster has the target porosity).
* ``percolates``   for each axis, whether some pore cluster wraps around the
                   torus along it (6-connectivity in 3D, 4 in 2D).  This is what
                   scripts/gen_aniso_structures3d.py:184-201 tests by labelling
                   a 3x3x3 tiling; here the wrap is detected directly by a
                   union-find over the periodic faces that tracks winding, so
                   memory stays one label array even at 512^3.
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage

__all__ = ["spheres", "ellipsoids", "ellipses", "grf", "grf_field", "percolates",
           "percolating_cluster", "slit", "random_unit_vector"]


def random_unit_vector(rng, dim=3):
    v = rng.normal(size=dim)
    return v / np.linalg.norm(v)


def _stamp(solid, centre, half, inside):
    """OR into `solid` the voxels of the periodic box around `centre` where
    inside(d) holds, d = (dz, dy, dx) minimum-image offsets (array per axis)."""
    shape = solid.shape
    idx, offs = [], []
    for ax, n in enumerate(shape):
        c = centre[ax]
        lo = int(np.floor(c - half[ax]))
        hi = int(np.ceil(c + half[ax]))
        r = np.arange(lo, hi + 1)
        if r.size > n:                  # object larger than the cell: clip to one period
            r = np.arange(int(np.floor(c - n / 2)), int(np.floor(c - n / 2)) + n)
        idx.append(np.mod(r, n))
        offs.append(r - c)
    grids = np.meshgrid(*offs, indexing="ij")
    m = inside(grids)
    sub = solid[np.ix_(*idx)]
    solid[np.ix_(*idx)] = sub | m


def spheres(shape=(64, 64, 64), porosity=0.6, rmin=3.0, rmax=6.0, seed=0):
    """Overlapping periodic spheres added until pore fraction <= porosity."""
    rng = np.random.default_rng(seed)
    shape = tuple(int(s) for s in shape)
    solid = np.zeros(shape, dtype=bool)
    total = solid.size
    target_solid = (1.0 - porosity) * total
    n_solid = 0
    while n_solid < target_solid:
        c = rng.uniform(0, shape)
        r = rng.uniform(rmin, rmax)
        _stamp(solid, c, (r,) * len(shape),
               lambda g: sum(x * x for x in g) <= r * r)
        n_solid = int(np.count_nonzero(solid))
    return solid.astype(np.uint8)


def _jittered_axis(rng, axis, jitter_deg):
    """`axis` rotated by an angle ~ U[0, jitter] about a random perpendicular."""
    dim = axis.size
    ang = np.deg2rad(rng.uniform(0.0, jitter_deg))
    p = rng.normal(size=dim)
    p -= p.dot(axis) * axis
    p /= np.linalg.norm(p)
    a = np.cos(ang) * axis + np.sin(ang) * p
    return a / np.linalg.norm(a)


def ellipsoids(shape=(64, 64, 64), porosity=0.5, rmin=2.0, rmax=4.0, aspect=3.0,
               jitter_deg=15.0, seed=0, axis=None):
    """Overlapping prolate ellipsoids, semi-axes (aspect*r, r, r), long axes
    along one seeded unit axis (given as (x, y, z)) with per-object jitter."""
    rng = np.random.default_rng(seed)
    shape = tuple(int(s) for s in shape)
    dim = len(shape)
    if axis is None:
        axis = random_unit_vector(rng, dim)
    axis = np.asarray(axis, float)
    axis = axis / np.linalg.norm(axis)
    # axis is given in (x, y[, z]) order; grids come in array order (z, y, x)
    axis_arr = axis[::-1]
    solid = np.zeros(shape, dtype=bool)
    target_solid = (1.0 - porosity) * solid.size
    n_solid = 0
    batch = 0
    while n_solid < target_solid:
        c = rng.uniform(0, shape)
        r = rng.uniform(rmin, rmax)
        a = _jittered_axis(rng, axis_arr, jitter_deg)
        L = aspect * r

        def inside(g, a=a, r=r, L=L):
            t = sum(a[i] * g[i] for i in range(dim))
            d2 = sum(x * x for x in g)
            return (t * t) / (L * L) + (d2 - t * t) / (r * r) <= 1.0

        _stamp(solid, c, (L,) * dim, inside)
        batch += 1
        if batch % 16 == 0 or n_solid + 4 * L * r * r >= target_solid:
            n_solid = int(np.count_nonzero(solid))
    return solid.astype(np.uint8)


def ellipses(shape=(256, 256), porosity=0.5, rmin=2.0, rmax=4.0, aspect=3.0,
             jitter_deg=15.0, seed=0, axis=None):
    """2D counterpart of ``ellipsoids``: [y, x] mask."""
    return ellipsoids(shape, porosity, rmin, rmax, aspect, jitter_deg, seed, axis)


def slit(n, h, axis=2, dims=3):
    """Poiseuille slit: h fluid layers normal to `axis` (0=x,1=y,2=z), the rest
    solid (scripts/validate_solver3d.py:87-90 puts the slab at the low end)."""
    shape = (n,) * dims
    solid = np.zeros(shape, dtype=np.uint8)
    arr_axis = dims - 1 - axis
    sl = [slice(None)] * dims
    sl[arr_axis] = slice(0, n - h)
    solid[tuple(sl)] = 1
    return solid


# ---------------------------------------------------------------------------
# percolation on the torus
# ---------------------------------------------------------------------------

def _wrap_analysis(pore):
    """Label pore clusters (face connectivity) and find, per cluster root, the
    set of axes along which it wraps.  Returns (labels, root_of_label, wraps)
    with wraps[root] a bool vector over array axes."""
    dim = pore.ndim
    structure = ndimage.generate_binary_structure(dim, 1)
    lab, n = ndimage.label(pore, structure=structure)
    parent = np.arange(n + 1)
    offset = np.zeros((n + 1, dim), dtype=np.int64)
    wraps = np.zeros((n + 1, dim), dtype=bool)

    # offset[a] = lift position of label a relative to its root, in periods;
    # a root always has offset 0.
    def find(a):
        path = []
        while parent[a] != a:
            path.append(a)
            a = parent[a]
        root = a
        for node in reversed(path):          # nearest the root first
            p = parent[node]
            if p != root:
                offset[node] += offset[p]
            parent[node] = root
        return root

    for ax in range(dim):
        lo = np.take(lab, 0, axis=ax).ravel()
        hi = np.take(lab, -1, axis=ax).ravel()
        both = (lo > 0) & (hi > 0)
        if not both.any():
            continue
        pairs = np.unique(np.stack([hi[both], lo[both]], axis=1), axis=0)
        d = np.zeros(dim, dtype=np.int64)
        d[ax] = 1              # the lo neighbour of a hi cell lies one period on
        for a, b in pairs:
            ra, rb = find(a), find(b)
            oa, ob = offset[a].copy(), offset[b].copy()
            if ra == rb:
                wraps[ra] |= (oa + d - ob) != 0
            else:
                parent[rb] = ra
                offset[rb] = oa + d - ob
                wraps[ra] |= wraps[rb]
    roots = np.array([find(i) for i in range(n + 1)])
    return lab, roots, wraps


def percolates(solid):
    """Per solver axis (x, y[, z]) whether a pore cluster wraps the torus."""
    pore = np.asarray(solid) == 0
    if not pore.any():
        return (False,) * pore.ndim
    lab, roots, wraps = _wrap_analysis(pore)
    any_wrap = wraps[np.unique(roots[1:])].any(axis=0)
    return tuple(bool(v) for v in any_wrap[::-1])     # array axes -> (x, y, z)


def percolating_cluster(solid, require_all=False):
    """Keep only the largest pore cluster that wraps (along every axis if
    require_all, else along at least one); everything else becomes solid.
    Returns None when no cluster qualifies."""
    pore = np.asarray(solid) == 0
    lab, roots, wraps = _wrap_analysis(pore)
    root_lab = roots[lab]
    sizes = np.bincount(root_lab.ravel(), minlength=roots.size)
    sizes[0] = 0
    ok = wraps.all(axis=1) if require_all else wraps.any(axis=1)
    cand = np.where(ok & (sizes > 0))[0]
    if cand.size == 0:
        return None
    best = cand[np.argmax(sizes[cand])]
    return (root_lab != best).astype(np.uint8)


def grf_field(shape=(128, 128, 128), corr=(8.0, 8.0, 24.0), seed=0):
    """FFT-filtered white noise with anisotropic Gaussian correlation lengths
    corr, given per solver axis (x, y, z); float32, array order."""
    rng = np.random.default_rng(seed)
    shape = tuple(int(s) for s in shape)
    dim = len(shape)
    noise = rng.standard_normal(shape, dtype=np.float32)
    spec = np.fft.rfftn(noise)
    del noise
    corr_arr = tuple(corr[:dim])[::-1]                # (x, y, z) -> array order
    filt = None
    for ax, (n, ell) in enumerate(zip(shape, corr_arr)):
        k = np.fft.rfftfreq(n) if ax == dim - 1 else np.fft.fftfreq(n)
        g = np.exp(-0.5 * (2 * np.pi * k * ell) ** 2).astype(np.float32)
        sh = [1] * dim
        sh[ax] = g.size
        g = g.reshape(sh)
        filt = g if filt is None else filt * g
    spec *= filt
    field = np.fft.irfftn(spec, s=shape, axes=list(range(dim))).astype(np.float32)
    return field


def grf(shape=(128, 128, 128), porosity=0.2, corr=(8.0, 8.0, 24.0), seed=0,
        keep_percolating=True, require_all=False, cluster_porosity=False, tol=0.002):
    """Gaussian random field mask: pore where the field is below the quantile
    `porosity`, then (keep_percolating) only the largest percolating pore
    cluster.  The isolated pockets that reduction removes lower the porosity;
    with cluster_porosity the threshold is raised until the kept cluster itself
    has porosity within tol of the target (a few passes of
    threshold += target - achieved)."""
    field = grf_field(shape, corr, seed)
    flat = field.ravel()

    def threshold(p):
        return (field > np.quantile(flat, min(max(p, 0.0), 1.0))).astype(np.uint8)

    solid = threshold(porosity)
    if not keep_percolating:
        return solid
    kept = percolating_cluster(solid, require_all=require_all)
    if not cluster_porosity:
        return kept if kept is not None else solid
    # the kept cluster's porosity grows with the threshold level p: secant
    # steps p += target - achieved, bisecting once the target is bracketed
    best, p, lo, hi = None, porosity, None, None
    for _ in range(16):
        got = 1.0 - float(kept.mean()) if kept is not None else 0.0
        if kept is not None and (best is None or abs(got - porosity) < best[0]):
            best = (abs(got - porosity), kept)
            if best[0] <= tol:
                break
        if got < porosity:
            lo = p if lo is None else max(lo, p)
        else:
            hi = p if hi is None else min(hi, p)
        nxt = p + (porosity - got) if kept is not None else p + 0.02
        if lo is not None and hi is not None and not lo < nxt < hi:
            nxt = 0.5 * (lo + hi)
        p = nxt
        kept = percolating_cluster(threshold(p), require_all=require_all)
    return best[1] if best is not None else solid
