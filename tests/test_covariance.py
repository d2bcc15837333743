"""
Axis-permutation covariance (flowperm.covariance, the check of
scripts/rotation_covariance3d.py): the permutation bookkeeping on the CPU, the
verdict logic against planted mistakes, and on the GPU the five permuted
orientations of one anisotropic structure solved in one batch.
"""
import numpy as np
import pytest

from flowperm import covariance as cv


def test_self_test_passes():
    lines = []
    assert cv.self_test(log=lines.append) == 0
    assert sum(l.startswith("[ok]") for l in lines) == 6 * 4 + 6 * 2


def test_permuted_structure_keeps_every_voxel_under_index_map():
    """brute force over every voxel of a non-cubic grid: new (x,y,z) = i holds
    old voxel j with j[sigma[a]] = i[a]"""
    rng = np.random.default_rng(3)
    arr = rng.integers(0, 2, size=(3, 4, 6), dtype=np.uint8)    # [z, y, x]
    for label, sigma in cv.PERMS.items():
        p = cv.permute_structure(arr, sigma)
        nz, ny, nx = p.shape
        for z in range(nz):
            for y in range(ny):
                for x in range(nx):
                    i = (x, y, z)
                    j = [0, 0, 0]
                    for a in range(3):
                        j[sigma[a]] = i[a]
                    assert p[z, y, x] == arr[j[2], j[1], j[0]], (label, i)


def rows_from(K, steps=(8000, 9000, 7000)):
    row = {f"K_{r}{c}": K[i][j] for i, r in enumerate("xyz") for j, c in enumerate("xyz")}
    row.update({f"steps_f{a}": s for a, s in zip("xyz", steps)})
    return row


def test_compare_passes_covariant_results_and_flags_planted_mistakes():
    K0 = np.array([[3.0, 0.4, 0.2], [0.41, 2.0, 0.1], [0.19, 0.11, 1.0]])
    labels = [l for l in cv.PERMS if l != "xyz"]
    good = {"xyz": rows_from(K0)}
    for l in labels:
        s = cv.PERMS[l]
        good[l] = rows_from(cv.expected(K0, s), tuple(np.array([8000, 9000, 7000])[list(s)]))
    ok, lines = cv.compare(good, labels)
    assert ok and all(ln["steps"] == ln["steps_expected"] for ln in lines)
    # planted: the permuted runs report K unpermuted (an axis-labelling bug)
    bad = {"xyz": rows_from(K0), **{l: rows_from(K0) for l in labels}}
    ok, lines = cv.compare(bad, labels)
    assert not ok and all(not ln["good"] for ln in lines)
    # planted: transposed instead of permuted (P^T K P for the 3-cycles)
    tr = {"xyz": rows_from(K0), **{l: rows_from(cv.expected(K0, cv.inverse(cv.PERMS[l])))
                                   for l in labels}}
    ok, lines = cv.compare(tr, labels)
    assert not ok and {ln["label"] for ln in lines if not ln["good"]} == {"yzx", "zxy"}
    # an isotropic K passes for free: reported as vacuous, i.e. failed
    iso = {l: rows_from(np.eye(3)) for l in ["xyz"] + labels}
    ok, lines = cv.compare(iso, labels)
    assert not ok and all(ln["scale"] == 0 for ln in lines)


def test_structure_is_anisotropic_and_percolates():
    from flowperm import masks
    m = cv.make_structure(24, 1)
    assert all(masks.percolates(m))
    assert m.shape == (24, 24, 24) and 0.5 < (m == 0).mean() < 0.85


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("f64", 1e-9), ("f32", 1e-4)])
def test_permuted_orientations_solve_covariantly_in_one_batch(precision, tol):
    mask = cv.make_structure(24, 1)
    labels = [l for l in cv.PERMS if l != "xyz"]
    res = cv.solve_orientations(mask, labels, precision=precision, ff=40, eps=1e-6,
                                lag=500, min_steps=2000, avg_steps=500, max_steps=40000)
    ok, lines = cv.compare(res, labels, tol=tol, min_scale=1e-3)
    for ln in lines:
        print(ln, flush=True)
    assert ok, lines
    if precision == "f64":
        assert all(ln["steps"] == ln["steps_expected"] for ln in lines)
        assert max(ln["resid"] for ln in lines) < 1e-11
