"""
The flux-partial layout of a batch, checked on the host (fp_verify_batch_tables,
no GPU): the tables Batch::rebuild uploads -- every kernel's work items, the
partial slot each item writes, each instance's slot range and the reduction
items over it -- for batches that mix both neighbour-step forms, many
instances, instances with more than REDUCE_SPAN (1024) items and lattices
large enough for the y-band item order.  Every slot a step kernel writes must
lie in its array and in its own instance's range, the instances' ranges must
partition each array and each instance's reduction items must tile its range.

This is the arithmetic a mixed-form batch depends on: row-segment items' slots
first, compact items' slots after them (the neighbour kernels write
part[it] and part[n_rowseg + it]), one allocation of n_rowseg + n_compact
slots, reduction ranges in slot units.
"""
import ctypes as C

import numpy as np
import pytest


def verify(masks, lattice=3, precision=1, **tune):
    from flowperm import _lib
    lib = _lib.load()
    arrs = [np.ascontiguousarray(m, np.uint8) for m in masks]
    ptrs = (_lib._u8p * len(arrs))(*[_lib.u8ptr(a) for a in arrs])
    dims = []
    for a in arrs:
        dims += [a.shape[-1], a.shape[-2], a.shape[0] if a.ndim == 3 else 1]
    d = (C.c_int32 * len(dims))(*dims)
    t = None
    if tune:
        t = _lib.Tuning(tune.get("local_cpt", 2), tune.get("nbr_waves", 4), 200,
                        tune.get("nontemporal", 1), tune.get("local_bpc", 0),
                        tune.get("nbr_bpc", 0), tune.get("nbr_mode", 2), tune.get("nbr_segs", 0))
    bad = C.c_int64(-1)
    st = lib.fp_verify_batch_tables(ptrs, d, len(arrs), lattice, precision,
                                    C.byref(t) if t is not None else None, C.byref(bad))
    assert st == 0, st
    return bad.value


def mixed_pair():
    """the dense and sparse structures of the mixed-form batch GPU test"""
    from flowperm import masks
    dense = masks.spheres((20, 20, 24), 0.8, 3, 5, seed=5)
    sparse = masks.grf((24, 28, 70), porosity=0.3, corr=(3, 3, 6), seed=6)
    assert 1 - sparse.mean() < 0.68 < 1 - dense.mean()
    return [dense, sparse]


@pytest.mark.parametrize("tune", [{}, {"nbr_waves": 1}, {"nbr_waves": 2}, {"nbr_waves": 8},
                                  {"local_cpt": 1}, {"local_cpt": 4}, {"nbr_mode": 0},
                                  {"nbr_mode": 1}, {"nbr_bpc": 8, "local_bpc": 16},
                                  {"nbr_segs": 1}, {"nbr_segs": 2}, {"nbr_waves": 8, "nbr_segs": 8},
                                  {"nbr_waves": 6, "nbr_segs": 4}])
def test_mixed_form_batch_layout(tune):
    assert verify(mixed_pair(), **tune) == 0
    assert verify(mixed_pair()[::-1], **tune) == 0


def test_many_instances_with_long_reduction_ranges():
    """twelve structures of both forms, two of them with more than 1024 items
    per kernel (several reduction items each)"""
    from flowperm import masks
    ms = [masks.spheres((n, n, n), por, 2, 4, seed=i)
          for i, (n, por) in enumerate([(12, 0.9), (40, 0.5), (16, 0.7), (64, 0.6), (20, 0.3),
                                        (24, 0.8), (8, 0.5), (48, 0.75), (30, 0.4), (10, 0.95),
                                        (36, 0.65), (14, 0.55)])]
    for tune in ({}, {"local_cpt": 1, "nbr_waves": 1}):
        assert verify(ms, **tune) == 0
        assert verify(ms, precision=0, **tune) == 0


def test_banded_lattice_layout():
    """a lattice whose populations reach 400 MB deals its neighbour items in y
    bands (both forms), beside a small instance"""
    from flowperm import masks
    rng = np.random.default_rng(3)
    big = (rng.random((200, 224, 256)) < 0.35).astype(np.uint8)
    small = masks.spheres((16, 16, 16), 0.6, seed=2)
    for mode in (0, 1, 2):
        assert verify([small, big], nbr_mode=mode) == 0


def test_2d_layout():
    from flowperm import masks
    ms = [masks.ellipses((64, 80), 0.6, seed=1), masks.ellipses((300, 257), 0.4, seed=2)]
    for mode in (0, 1, 2):
        assert verify(ms, lattice=2, precision=2, nbr_mode=mode) == 0


@pytest.mark.parametrize("lattice", [3, 2])
@pytest.mark.parametrize("tune", [{"nbr_segs": 1}, {"nbr_segs": 2}, {}, {"nbr_waves": 8, "nbr_segs": 8},
                                  {"nbr_waves": 8, "nbr_segs": 4}, {"nbr_waves": 2, "nbr_segs": 4},
                                  {"nbr_waves": 3, "nbr_segs": 2}, {"nbr_waves": 6, "nbr_segs": 4}])
def test_row_segments_side_by_side_cover_every_unit_once(lattice, tune):
    """items of S x-adjacent segments: rows of 1 to 9 segments (x 40 ... 530),
    so that a row's last item has segments past its end (idle wavefronts),
    and slabs whose last one is short; every (row, segment) unit is stepped by
    exactly one wavefront of a workgroup that has it"""
    from flowperm import masks
    rng = np.random.default_rng(4)
    if lattice == 3:
        ms = [(rng.random((nz, ny, nx)) < 0.35).astype(np.uint8)
              for nz, ny, nx in ((7, 5, 40), (9, 6, 130), (5, 3, 200), (6, 4, 530), (3, 3, 300))]
    else:
        ms = [(rng.random((ny, nx)) < 0.35).astype(np.uint8)
              for ny, nx in ((13, 40), (11, 190), (7, 530), (5, 260))]
    assert verify(ms, lattice=lattice, nbr_mode=0, **tune) == 0
