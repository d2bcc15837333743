"""
Pins of the CPU oracle (oracle/), before anything is compared against it.

3D: against rows the reference binary itself produced (tests/golden/, made by
tests/golden/make_golden.py from LMB3d/lbm3d-perm.cpp built with its own
flags), at the CSV's printed precision, with identical step counts; against
the slit closed form; against axis-permutation covariance and the tensor test.
2D (the reference cannot be built here: it needs Magick++): against closed
forms of this forcing convention, rot90 covariance, and a loop-by-loop NumPy
restatement of LMB2d/lbm.cpp on tiny lattices.
The exact flux sum is pinned against Python's Fraction arithmetic.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

from conftest import GOLDEN_FAST, assert_printed_close, check_eigen_and_reciprocity, golden, golden_flags

pytestmark = pytest.mark.filterwarnings("ignore")


# ---------------------------------------------------------------------------
# exact flux sum
# ---------------------------------------------------------------------------

def _exact(values):
    s = 0
    for v in values:
        f = Fraction(float(v)) * (1 << 96)
        s += math.trunc(f)                        # toward zero
    return s


def test_fix96_sum_equals_fraction_arithmetic(oracle):
    rng = np.random.default_rng(0)
    parts = [rng.normal(0, 1e-4, 400), rng.normal(0, 0.3, 100).clip(-0.999, 0.999),
             rng.normal(0, 1e-20, 50), np.array([0.0, -0.0, 5e-324, -5e-324, 0.999999999999,
                                                 -0.999999999999, 2.0 ** -96, -(2.0 ** -96),
                                                 3 * 2.0 ** -97])]
    v = np.concatenate(parts)
    rng.shuffle(v)
    total, as_double, bad = oracle.fix96_sum(v)
    assert bad == 0
    want = _exact(v)
    assert total == want
    assert as_double == float(Fraction(want, 1 << 96))     # one correctly rounded conversion
    # order independence
    total2, d2, _ = oracle.fix96_sum(v[::-1].copy())
    assert (total2, d2) == (total, as_double)


@pytest.mark.parametrize("bad", [1.0, -1.0, 3.0, float("nan"), float("inf")])
def test_fix96_sum_flags_out_of_range(oracle, bad):
    _, _, flag = oracle.fix96_sum(np.array([0.1, bad, 0.2]))
    assert flag == 1


# ---------------------------------------------------------------------------
# 3D against the reference binary's own output
# ---------------------------------------------------------------------------

def _check_tensor_against_row(T, row, what):
    K = np.array(T.K[:]).reshape(3, 3)
    trace = abs(np.trace(K))
    tags = ("fx", "fy", "fz")
    for d, t in enumerate(tags):
        assert T.run[d].steps == row[f"steps_{t}"], f"{what} steps_{t}"
        assert T.run[d].converged == row[f"conv_{t}"], f"{what} conv_{t}"
        qn = float(np.linalg.norm(T.run[d].q[:]))
        for k, a in enumerate("xyz"):
            assert_printed_close(T.run[d].q[k], row[f"q{a}_{t}"], 1e-10 * qn, f"{what} q{a}_{t}")
            # se = sqrt((sumsq - n q^2)/(n-1)/n) cancels catastrophically: with a
            # (near-)constant q, sequential rounding of sum and sumsq over n = 2000
            # samples leaves up to ~1e-8 |q| of noise (measured 3.5e-9 |q| on the
            # slits), so only differences above that are meaningful
            ref_se = row[f"se_q{a}_{t}"]
            assert abs(T.run[d].se[k] - ref_se) <= max(3e-8 * qn, 1e-3 * abs(ref_se)), \
                f"{what} se_q{a}_{t}"
    for i, a in enumerate("xyz"):
        for j, b in enumerate("xyz"):
            assert_printed_close(K[i, j], row[f"K_{a}{b}"], 1e-10 * trace, f"{what} K_{a}{b}")
    for key in ("k_xy", "k_xz", "k_yz"):
        assert_printed_close(getattr(T, key), row[key], 1e-10 * trace, f"{what} {key}")
    assert_printed_close(T.k_mean, row["k_mean"], 1e-10 * trace, what)
    for j in range(3):
        assert_printed_close(T.kval[j], row[f"k{j + 1}"], 1e-10 * trace, f"{what} k{j + 1}")
    assert_printed_close(T.porosity, row["porosity"], 0, f"{what} porosity")
    assert row["nx"] == T.nx and row["ny"] == T.ny and row["nz"] == T.nz
    vec = np.array(T.vec[:]).reshape(3, 3)             # vec[component][eigen index]
    ours = {f"v{j + 1}_{a}": vec[i, j] for j in range(3) for i, a in enumerate("xyz")}
    ours.update(fa=T.fa, recip_xy=T.recip_xy, recip_xz=T.recip_xz, recip_yz=T.recip_yz,
                recip_resid=T.recip_resid)
    return check_eigen_and_reciprocity(ours.__getitem__, row, trace, what)


@pytest.mark.parametrize("name", GOLDEN_FAST + ["sph_x28y24z20_s1"])
def test_oracle_matches_reference_rows(oracle, name):
    mask, row, meta = golden(name)
    rc, T = oracle.solve_tensor3d(mask, **golden_flags(meta))
    assert rc == meta["returncode"]
    n_vectors = _check_tensor_against_row(T, row, name)
    assert n_vectors >= 1, "no principal vector separated enough to compare"


def test_oracle_error_codes_match_reference(oracle):
    """read_structure's rejections (:182-192): the binary's recorded exit codes."""
    import json
    from conftest import GOLDEN
    errs = json.loads((GOLDEN / "errors.json").read_text())
    assert errs["nonbinary"]["returncode"] == 2 and errs["nopore"]["returncode"] == 2
    m = np.zeros((4, 4, 4), np.uint8)
    m[0] = 1
    assert oracle.check_mask3d(m) == (0, 48)
    m2 = m.copy()
    m2[1, 1, 1] = 2
    assert oracle.check_mask3d(m2)[0] == 2
    assert oracle.check_mask3d(np.ones((4, 4, 4), np.uint8))[0] == 3


# ---------------------------------------------------------------------------
# 3D against mathematics
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("h", [6, 10])
def test_oracle_slit_closed_form(oracle, h):
    """Guo-forced BGK at tau = 1 with halfway bounce-back reproduces the slit
    exactly: K = h (h^2 + 1) / (12 nz) (the parabola plus a uniform F/4 slip),
    the value the reference recorded for h = 10, 20, 28 at nz = 32."""
    from flowperm import masks
    n = 16
    mask = masks.slit(n, h)
    rc, T = oracle.solve_tensor3d(mask, eps=1e-10, lag=500, min_steps=2000, max_steps=20000,
                                  avg_steps=200)
    K = np.array(T.K[:]).reshape(3, 3)
    exact = h * (h * h + 1) / (12.0 * n)
    assert abs(K[0, 0] - exact) / exact < 1e-9
    assert abs(K[1, 1] - exact) / exact < 1e-9
    assert abs(K[2, 2]) < 1e-9 * exact                  # blocked
    off = np.abs(K - np.diag(np.diag(K))).max()
    assert off < 1e-9 * exact


def _permute(arr, sigma):
    """relabel solver axes: new axis a <- old axis sigma[a] (arr is [z,y,x])."""
    a = arr.transpose(2, 1, 0).transpose(sigma)
    return np.ascontiguousarray(a.transpose(2, 1, 0))


def test_oracle_axis_permutation_covariance(oracle):
    """K' = P K P^T for every relabelling of the axes (scripts/rotation_covariance3d.py)."""
    from flowperm import masks
    mask = masks.ellipsoids((10, 12, 14), 0.55, rmin=1.5, rmax=2.5, seed=11, axis=(1, 2, 3))
    cfg = dict(ff=40, eps=1e-9, lag=200, min_steps=400, avg_steps=10, max_steps=20000)
    _, T0 = oracle.solve_tensor3d(mask, **cfg)
    K0 = np.array(T0.K[:]).reshape(3, 3)
    trace = np.trace(K0)
    assert np.abs(K0 - np.diag(np.diag(K0))).max() > 1e-3 * trace   # not vacuous
    for sigma in [(1, 0, 2), (2, 1, 0), (1, 2, 0)]:
        _, T = oracle.solve_tensor3d(_permute(mask, sigma), **cfg)
        K = np.array(T.K[:]).reshape(3, 3)
        want = K0[np.ix_(sigma, sigma)]
        assert np.abs(K - want).max() / trace < 1e-11, sigma


def test_oracle_tensor_test_and_reciprocity(oracle):
    """The (1,1,1) solve matches K F (Koza09 test, lbm3d-perm.cpp:555-572).
    The residual, like the asymmetry of K, is the O(Re) non-linearity of the
    lattice: it scales with the force (measured 1.2e-5 at ff=40, 3.1e-7 at
    ff=1), so it is checked at ff=1 and for its scaling between ff=1 and 4."""
    mask, _, _ = golden("ell_x24y20z16_fast_s5")
    cfg = dict(validate=True, eps=1e-10, lag=200, min_steps=400, avg_steps=10, max_steps=20000)
    res = {}
    for ff in (1.0, 4.0):
        rc, T = oracle.solve_tensor3d(mask, ff=ff, **cfg)
        assert T.validated == 1
        K = np.array(T.K[:]).reshape(3, 3)
        c = 1.0 / np.sqrt(3.0)
        pred = K @ (T.force * c * np.ones(3)) / T.mu       # full K: superposition
        q = np.array(T.val.q[:])
        res[ff] = (np.linalg.norm(q - pred) / np.linalg.norm(q), T.val_err, T.recip_resid)
    lin1, err1, rec1 = res[1.0]
    assert lin1 < 1e-6 and err1 < 1e-6 and rec1 < 1e-6
    assert 3.6 < res[4.0][0] / lin1 < 4.4                    # linear in F


def test_oracle_jacobi_matches_eigh(oracle):
    rng = np.random.default_rng(3)
    for _ in range(50):
        B = rng.normal(size=(3, 3))
        A = B @ B.T + 0.1 * np.eye(3)
        val, vec = oracle.jacobi3(A)
        w, V = np.linalg.eigh(A)
        assert np.allclose(val, w[::-1], rtol=1e-12, atol=1e-14)
        for j in range(3):                         # columns are eigenvectors (sign free)
            assert np.allclose(A @ vec[:, j], val[j] * vec[:, j], atol=1e-12)


# ---------------------------------------------------------------------------
# 2D
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("h", [6, 10])
def test_oracle2d_slit_closed_forms(oracle, h):
    """LMB2d convention (equilibrium-velocity shift, raw moment reported) at
    tau = 1 on a slit of h fluid rows: the raw velocity sums to exactly F h^3/2
    per column, so k_xx = mu h^3 / (2 L0^2); across the slab the fluid is at
    rest and the raw moment reads -F/(2 rho), so k_yy = -mu h LX / (2 LY L0^2).
    (A Guo-style velocity would give k_yy = 0 and k_xx larger by mu h/(2 L0^2).)"""
    L = 32
    mask = np.zeros((L, L), np.uint8)
    mask[:L - h, :] = 1
    rc, T = oracle.solve_tensor2d(mask, double=True, eps=1e-9, lag=2000, min_steps=4000,
                                  max_steps=60000, avg_steps=2000, avg_sample=50)
    mu = float(np.float32(0.165))
    assert abs(T.k_xx - mu * h ** 3 / 32.0) / (mu * h ** 3 / 32.0) < 1e-8
    blocked = -mu * h / 32.0
    assert abs(T.k_yy - blocked) / abs(blocked) < 1e-8
    assert abs(T.k_xy) < 1e-9 * T.k_xx


def test_oracle2d_float_slit_within_float_noise(oracle):
    """The float (reference) variant reaches the same closed form to the
    resolution of float populations at a force large enough to be resolved."""
    L, h = 32, 10
    mask = np.zeros((L, L), np.uint8)
    mask[:L - h, :] = 1
    _, T = oracle.solve_tensor2d(mask, ff=200, eps=1e-6, lag=2000, min_steps=4000,
                                 max_steps=40000, avg_steps=2000, avg_sample=50)
    mu = float(np.float32(0.165))
    assert abs(T.k_xx - mu * h ** 3 / 32.0) / (mu * h ** 3 / 32.0) < 2e-4


def test_oracle2d_axis_swap_and_rot90_covariance(oracle):
    """Swapping x and y maps K to P K P^T exactly (forces stay positive).
    A 90-degree rotation also sends one force to its negative, and the lattice
    response is odd in F only to first order (the O(Re) term measured in 3D),
    so it is checked at a small force."""
    from flowperm import masks
    m = masks.ellipses((24, 24), 0.7, rmin=1.5, rmax=2.5, seed=4, axis=(1, 2))
    assert all(masks.percolates(m))
    # ksigma = 0: the block-SE tolerance would stop a deterministic double run
    # while its transient is still moving
    base = dict(double=True, eps=1e-10, ksigma=0.0, lag=500, min_steps=1000,
                max_steps=60000, avg_steps=50, avg_sample=1)
    _, T0 = oracle.solve_tensor2d(m, ff=40, **base)
    tr = abs(T0.k_xx + T0.k_yy)
    assert abs(T0.k_xy) > 1e-3 * tr                  # not vacuous
    _, T1 = oracle.solve_tensor2d(np.ascontiguousarray(m.T), ff=40, **base)
    assert abs(T1.k_xx - T0.k_yy) < 1e-11 * tr
    assert abs(T1.k_yy - T0.k_xx) < 1e-11 * tr
    assert abs(T1.k_xy - T0.k_yx) < 1e-11 * tr
    assert abs(T1.k_yx - T0.k_xy) < 1e-11 * tr
    # rot90 of [y, x]: (x, y) -> (y, -x), K' = R K R^T with R = [[0, 1], [-1, 0]]
    _, S0 = oracle.solve_tensor2d(m, ff=0.05, **base)
    _, S1 = oracle.solve_tensor2d(np.ascontiguousarray(np.rot90(m)), ff=0.05, **base)
    tr = abs(S0.k_xx + S0.k_yy)
    assert abs(S1.k_xx - S0.k_yy) < 1e-7 * tr
    assert abs(S1.k_yy - S0.k_xx) < 1e-7 * tr
    assert abs(S1.k_xy + S0.k_yx) < 1e-7 * tr
    assert abs(S1.k_yx + S0.k_xy) < 1e-7 * tr


def _numpy_lbm2d(solid, fx, fy, steps, real=np.float32):
    """Loop-by-loop restatement of LMB2d/lbm.cpp:516-585 + 665-679 (tiny lattices)."""
    ly, lx = solid.shape
    F = solid.T.astype(int)                          # F[x][y]
    ex = [0, 1, 0, -1, 0, 1, -1, -1, 1]
    ey = [0, 0, 1, 0, -1, 1, 1, -1, -1]
    inv = [0, 3, 4, 1, 2, 7, 8, 5, 6]
    w = [real(4.0 / 9.0)] + [real(1.0 / 9.0)] * 4 + [real(1.0 / 36.0)] * 4
    df = np.zeros((2, lx, ly, 9), real)
    for k in range(9):
        df[:, :, :, k] = w[k]
    U = np.zeros((lx, ly), real); V = np.zeros((lx, ly), real); R = np.ones((lx, ly), real)
    fx, fy, tau = real(fx), real(fy), real(1)
    c = 1
    for _ in range(steps):
        c = 1 - c
        for i in range(lx):
            for j in range(ly):
                if F[i, j] == 0:
                    rho = real(0); ux = real(0); uy = real(0)
                    for k in range(9):
                        rho = real(rho + df[c, i, j, k])
                        ux = real(ux + real(df[c, i, j, k] * real(ex[k])))
                        uy = real(uy + real(df[c, i, j, k] * real(ey[k])))
                    U[i, j] = real(ux / rho); V[i, j] = real(uy / rho); R[i, j] = rho
        for i in range(lx):
            for j in range(ly):
                if F[i, j] == 0:
                    rho = R[i, j]
                    ux = real(U[i, j] + real(real(fx * tau) / rho))
                    uy = real(V[i, j] + real(real(fy * tau) / rho))
                    for k in range(9):
                        eu = real(real(real(ex[k]) * ux) + real(real(ey[k]) * uy))
                        t = real(real(1) - real(real(1.5) * real(real(ux * ux) + real(uy * uy))))
                        t = real(t + real(real(3) * eu))
                        t = real(t + real(real(real(4.5) * eu) * eu))
                        feq = real(real(w[k] * rho) * t)
                        d = df[c, i, j, k]
                        df[c, i, j, k] = real(d - real(real(1) * real(d - feq)))
        for i in range(lx):
            for j in range(ly):
                if F[i, j] == 0:
                    for k in range(9):
                        ip = (i + ex[k] + lx) % lx; jp = (j + ey[k] + ly) % ly
                        if F[ip, jp] == 1:
                            df[1 - c, i, j, inv[k]] = df[c, i, j, k]
                        else:
                            df[1 - c, ip, jp, k] = df[c, i, j, k]
    return df[1 - c], U, V


def test_oracle2d_matches_numpy_loop_restatement(oracle):
    rng = np.random.default_rng(5)
    solid = (rng.random((6, 7)) < 0.3).astype(np.uint8)
    solid[0, :] = 0
    for steps in (1, 2, 5):
        df_np, U, V = _numpy_lbm2d(solid, 1e-3, -4e-4, steps)
        df_or, uv = oracle.run_steps2d(solid, (1e-3, -4e-4), steps)
        fluid = solid.T == 0                                  # [x][y]
        assert np.array_equal(df_or[fluid], df_np[fluid]), steps
        assert np.array_equal(uv[:, :, 0], U.T) and np.array_equal(uv[:, :, 1], V.T)
