"""
Host-side logic that needs no GPU: the CSV text the reference prints, the
structure file formats and their checks, the mask generators' percolation
test, and the serving arrival process.
"""
import csv
from collections import deque

import numpy as np
import pytest

from conftest import GOLDEN, GOLDEN_ALL


@pytest.mark.parametrize("name", GOLDEN_ALL[:5])
def test_csv_text_reproduces_reference_rows(name):
    """header and row text exactly as lbm3d-perm wrote them
    (LMB3d/lbm3d-perm.cpp:575-615, setprecision(10))"""
    from flowperm.solver import COLUMNS3D, _INT_COLS, csv_header, csv_row
    lines = (GOLDEN / f"{name}.csv").read_text().splitlines(keepends=True)
    assert lines[0] == csv_header(COLUMNS3D)
    fields = next(csv.reader([lines[1]]))
    row = {}
    for c, v in zip(COLUMNS3D, fields):
        row[c] = v if c == "filename" else (int(v) if c in _INT_COLS else float(v))
    assert csv_row(row, COLUMNS3D) == lines[1]


def test_raw_round_trip_and_reader_checks(tmp_path):
    """the reader's rejections (LMB3d/lbm3d-perm.cpp:174-193; the
    scripts/structure_io3d.py:75-127 self-test)"""
    from flowperm import structio
    rng = np.random.default_rng(0)
    m = (rng.random((5, 6, 7)) > 0.5).astype(np.uint8)
    p = structio.save_raw(tmp_path / "a.raw", m)
    raw = p.read_bytes()
    assert np.frombuffer(raw[:12], "<i4").tolist() == [7, 6, 5]
    assert np.array_equal(structio.load_raw(p), m)
    bad = {"trunc.raw": raw[:-1], "trail.raw": raw + b"\0", "hdr.raw": raw[:8],
           "dims.raw": np.array([0, 6, 5], "<i4").tobytes() + raw[12:]}
    for fn, data in bad.items():
        (tmp_path / fn).write_bytes(data)
        with pytest.raises(structio.StructureError):
            structio.load_raw(tmp_path / fn)
    nb = bytearray(raw)
    nb[20] = 2
    (tmp_path / "nb.raw").write_bytes(bytes(nb))
    with pytest.raises(structio.StructureError, match="non-binary"):
        structio.load_raw(tmp_path / "nb.raw")
    structio.save_raw(tmp_path / "solid.raw", np.ones((2, 2, 2), np.uint8))
    with pytest.raises(structio.StructureError, match="no pore"):
        structio.load_raw(tmp_path / "solid.raw")


def test_gif_round_trip(tmp_path):
    """pore = red > 0 (LMB2d/lbm.cpp:412-426); array [y, x]"""
    pytest.importorskip("PIL")
    from flowperm import masks, structio
    m = masks.ellipses((40, 56), 0.6, seed=3)
    structio.save_gif2d(tmp_path / "m.gif", m)
    assert np.array_equal(structio.load_structure2d(tmp_path / "m.gif"), m)


def test_gif_red_channel_decides(tmp_path):
    """read_from_gif (LMB2d/lbm.cpp:412-426) tests only the decoded red channel:
    a palette colour with red 0 is solid whatever its green and blue (a
    luminance or any-channel reading would call it pore), any red > 0 is pore,
    and a transparent palette index still decodes to its palette colour"""
    pytest.importorskip("PIL")
    from PIL import Image
    from flowperm import structio
    palette = [(0, 0, 0), (1, 0, 0), (0, 255, 0), (0, 0, 255), (0, 255, 255), (255, 255, 255),
               (200, 0, 0), (0, 1, 1)]
    pore = np.array([r > 0 for r, g, b in palette])
    rng = np.random.default_rng(11)
    idx = rng.integers(0, len(palette), size=(9, 13)).astype(np.uint8)
    idx[0, :len(palette)] = np.arange(len(palette))               # every colour appears
    im = Image.fromarray(idx, mode="P")
    flat = [c for rgb in palette for c in rgb]
    im.putpalette(flat + [0] * (768 - len(flat)))
    im.save(tmp_path / "p.gif", transparency=2)                    # (0, 255, 0) transparent
    m = structio.load_structure2d(tmp_path / "p.gif")
    assert m.shape == (9, 13)
    assert np.array_equal(m, (~pore[idx]).astype(np.uint8))


def test_atoi_semantics():
    """the C library's atoi on one extracted token (LMB2d/lbm.cpp:443, :464)"""
    from flowperm.structio import atoi
    for tok, want in [("0", 0), ("255", 255), ("12abc", 12), ("abc", 0), ("-3", -3),
                      ("+7", 7), ("007", 7), ("0.5", 0), ("3.9", 3), ("-0", 0), ("+-1", 0),
                      ("0x1F", 0), ("", 0), ("1e3", 1)]:
        assert atoi(tok) == want, tok


# a 5 x 3 PPM in the reference's layout: four header lines (read_from_ppm
# skips them without parsing), then rows y = 0..LY-1 of x = 0..LX-1 (R, G, B)
# tokens; pore where atoi(R) != 0 -- "12abc", "-3" and "+1" are pores, "abc",
# "0" and "0.9" are solid
PPM_R = [["255", "0", "12abc", "abc", "-3"],
         ["0", "0", "0.9", "+1", "255"],
         ["7", "0", "0", "0", "1"]]


def _ppm_text(rows, comment_first=False):
    head = ["P3", "# Created by IrfanView", "5 3", "255"] if comment_first else \
           ["P3", "5 3", "# Created by IrfanView", "255"]
    body = []
    for row in rows:
        body.append(" ".join(f"{r} 17 0" for r in row))
    return "\n".join(head + body) + "\n"


@pytest.mark.parametrize("comment_first", [False, True])
def test_ppm_reader_follows_read_from_ppm(tmp_path, comment_first):
    from flowperm import structio
    p = tmp_path / "m.ppm"
    p.write_text(_ppm_text(PPM_R, comment_first))
    m = structio.load_structure2d(p)
    want = np.array([[0 if structio.atoi(r) != 0 else 1 for r in row] for row in PPM_R], np.uint8)
    assert m.shape == (3, 5) and m.dtype == np.uint8
    assert np.array_equal(m, want)
    assert want.tolist() == [[0, 1, 0, 1, 0], [1, 1, 1, 0, 0], [0, 1, 1, 1, 0]]
    # values wrapped over lines other than one row per line read the same
    p.write_text(_ppm_text(PPM_R, comment_first).replace(" 17 0 ", "\n17\n0\n"))
    assert np.array_equal(structio.load_structure2d(p), want)
    # fewer triples than the header's lattice
    p.write_text(_ppm_text(PPM_R[:2], comment_first))
    with pytest.raises(structio.StructureError, match="truncated"):
        structio.load_structure2d(p)


def test_dat_reader_follows_read_from_dat(tmp_path):
    """pore where atoi(token) == 0 (LMB2d/lbm.cpp:456-470); any other file
    extension than .ppm / .gif is read as DAT (:472-483)"""
    from flowperm import structio
    toks = [["0", "1", "2", "0.5", "-1"],
            ["abc", "0", "00", "3x", "+0"],
            ["1", "1", "0", "9", "0"]]
    want = np.array([[1 if structio.atoi(t) != 0 else 0 for t in row] for row in toks], np.uint8)
    assert want.tolist() == [[0, 1, 1, 0, 1], [0, 0, 0, 1, 0], [1, 1, 0, 1, 0]]
    p = tmp_path / "m.txt"
    p.write_text("\n".join(" ".join(r) for r in toks) + "\n")
    assert np.array_equal(structio.load_structure2d(p, shape=(3, 5)), want)
    with pytest.raises(structio.StructureError, match="square"):
        structio.load_structure2d(p)
    q = tmp_path / "sq.dat"
    q.write_text(" ".join(["0", "1", "0", "1"] * 4))
    assert structio.load_structure2d(q).tolist() == [[0, 1, 0, 1]] * 4
    with pytest.raises(structio.StructureError, match="truncated"):
        structio.load_structure2d(p, shape=(4, 5))


def test_grf_cluster_porosity_hits_its_target():
    """the Gaussian random field's kept percolating cluster, not the threshold
    level, has the requested porosity when cluster_porosity is asked for.  Below
    the level at which a cluster first wraps along every axis no threshold can
    give the target: then the percolating cluster closest to it comes back."""
    from flowperm import masks
    for por, seed in ((0.45, 1), (0.55, 2), (0.25, 5)):
        m = masks.grf((48, 48, 48), porosity=por, corr=(4.0, 4.0, 8.0), seed=seed,
                      require_all=True, cluster_porosity=True)
        assert abs((1 - m.mean()) - por) <= 0.002, (por, seed, 1 - m.mean())
        assert all(masks.percolates(m))
    m = masks.grf((48, 48, 48), porosity=0.2, corr=(4.0, 4.0, 8.0), seed=1, require_all=True,
                  cluster_porosity=True)
    assert all(masks.percolates(m)) and 0.2 < 1 - m.mean() < 0.3


def _brute_wraps(sol):
    """BFS over the lifted torus: an axis wraps when a cluster reaches a copy
    of one of its own cells displaced along it"""
    pore = sol == 0
    shp, dim = sol.shape, sol.ndim
    lift, wr = {}, np.zeros(dim, bool)
    for start in zip(*np.nonzero(pore)):
        if start in lift:
            continue
        lift[start] = np.zeros(dim, int)
        dq = deque([start])
        while dq:
            c = dq.popleft()
            for ax in range(dim):
                for s in (-1, 1):
                    nc, L = list(c), lift[c].copy()
                    nc[ax] += s
                    if nc[ax] < 0:
                        nc[ax] += shp[ax]; L[ax] -= 1
                    elif nc[ax] >= shp[ax]:
                        nc[ax] -= shp[ax]; L[ax] += 1
                    nc = tuple(nc)
                    if not pore[nc]:
                        continue
                    if nc in lift:
                        wr |= lift[nc] != L
                    else:
                        lift[nc] = L
                        dq.append(nc)
    return tuple(bool(v) for v in wr[::-1])


def test_percolation_matches_brute_force_on_the_torus():
    from flowperm import masks
    rng = np.random.default_rng(1)
    for _ in range(120):
        shp = tuple(int(v) for v in rng.integers(3, 8, size=3))
        sol = (rng.random(shp) > rng.uniform(0.3, 0.7)).astype(np.uint8)
        assert masks.percolates(sol) == _brute_wraps(sol)
    # a straight channel along x percolates along x only
    s = np.ones((6, 6, 6), np.uint8)
    s[2, 3, :] = 0
    assert masks.percolates(s) == (True, False, False)


def test_generators_hit_their_porosity_and_percolate():
    from flowperm import masks
    m = masks.spheres((32, 32, 32), 0.6, 3, 6, seed=0)
    assert abs((1 - m.mean()) - 0.6) < 0.03
    e = masks.ellipsoids((32, 32, 32), 0.5, seed=1)
    assert abs((1 - e.mean()) - 0.5) < 0.05
    g = masks.grf((48, 48, 48), porosity=0.3, corr=(4, 4, 8), seed=2)
    assert 0 < 1 - g.mean() <= 0.3 + 1e-3
    assert any(masks.percolates(g))
    # percolating_cluster keeps a single wrapping cluster
    from scipy import ndimage
    lab, n = ndimage.label(g == 0)
    assert n >= 1


def test_slit_mask_layout():
    from flowperm import masks
    s = masks.slit(8, 3)
    assert s.shape == (8, 8, 8)
    assert int((s == 0).all(axis=(1, 2)).sum()) == 3      # h fluid z-planes


def test_poisson_arrivals_rate():
    from flowperm.server import poisson_arrivals
    a = poisson_arrivals(20.0, 4000, seed=0)
    assert np.all(np.diff(a) > 0)
    assert abs(len(a) / a[-1] - 20.0) < 1.0
