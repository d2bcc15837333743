"""
The C-ABI boundary without a GPU: libflowperm.so loads, exports every
function include/flowperm.h declares, the ctypes mirrors have the C layout
(checked against a probe compiled from the header with gcc), and the entry
points that do no device work behave as the reference does.
"""
import ctypes as C
import re
import subprocess

import numpy as np
import pytest

from conftest import REPO

HEADER = REPO / "include" / "flowperm.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", " ", HEADER.read_text(), flags=re.S)
    names = re.findall(r"\b(fp_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


@pytest.fixture(scope="module")
def lib():
    from flowperm import _lib
    return _lib.load()


def test_library_exports_every_declared_function(lib):
    from flowperm import _lib
    names = declared_functions()
    assert len(names) >= 36
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (fp_[a-z0-9_]+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    # and every one of them has a ctypes signature with the header's arity
    text = re.sub(r"/\*.*?\*/", " ", HEADER.read_text(), flags=re.S)
    for n in names:
        assert n in _lib.SIGNATURES, n
        m = re.search(r"\b" + n + r"\s*\(([^)]*)\)", text)
        args = [a for a in m.group(1).split(",") if a.strip() and a.strip() != "void"]
        assert len(args) == len(_lib.SIGNATURES[n][1]), n


STRUCTS = ["fp_config3d", "fp_solve_out", "fp_result3d", "fp_config2d", "fp_result2d",
           "fp_bench_out", "fp_tuning", "fp_server_config", "fp_request_info", "fp_server_stats"]


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    """sizeof and offsetof of every field, as gcc lays the header out"""
    from flowperm import _lib
    mirrors = {"fp_config3d": _lib.Config3D, "fp_solve_out": _lib.SolveOut,
               "fp_result3d": _lib.Result3D, "fp_config2d": _lib.Config2D,
               "fp_result2d": _lib.Result2D, "fp_bench_out": _lib.BenchOut,
               "fp_tuning": _lib.Tuning, "fp_server_config": _lib.ServerConfig,
               "fp_request_info": _lib.RequestInfo, "fp_server_stats": _lib.ServerStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for s, M in mirrors.items():
        lines.append(f'printf("{s} size %zu\\n", sizeof({s}));')
        for f, _ in M._fields_:
            lines.append(f'printf("{s} {f} %zu\\n", offsetof({s}, {f}));')
    lines.append("return 0;}")
    d = tmp_path_factory.mktemp("abi")
    src = d / "probe.c"
    src.write_text("\n".join(lines))
    exe = d / "probe"
    subprocess.run(["gcc", "-std=c99", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    table = {}
    for line in out.splitlines():
        s, f, v = line.split()
        table[(s, f)] = int(v)
    return mirrors, table


@pytest.mark.parametrize("name", STRUCTS)
def test_ctypes_mirror_matches_c_layout(c_layout, name):
    mirrors, table = c_layout
    M = mirrors[name]
    assert C.sizeof(M) == table[(name, "size")]
    for f, _ in M._fields_:
        assert getattr(M, f).offset == table[(name, f)], f


def test_abi_version_and_exit_codes(lib):
    text = HEADER.read_text()
    assert lib.fp_abi_version() == int(re.search(r"FP_ABI_VERSION (\d+)", text).group(1))
    # LMB3d/lbm3d-perm.cpp:140-144, :630 -- 0 converged, 3 not converged, 2 error
    assert lib.fp_exit_code(0) == 0
    assert lib.fp_exit_code(3) == 3
    for st in range(10, 18):
        assert lib.fp_exit_code(st) == 2


def test_default_configs_are_the_reference_defaults(lib):
    from flowperm import _lib
    c = _lib.Config3D()
    lib.fp_config3d_default(C.byref(c))
    # LMB3d/lbm3d-perm.cpp:118-130
    assert (c.ff, c.eps, c.lag, c.hold, c.min_steps, c.max_steps, c.avg_steps) == \
        (1.0, 1e-5, 2000, 2, 4000, 200000, 2000)
    assert c.validate == 0 and c.precision == _lib.FP_F64
    c2 = _lib.Config2D()
    lib.fp_config2d_default(C.byref(c2))
    # LMB2d/main.cpp:64-79
    assert (c2.ff, c2.eps, c2.qfloor, c2.ksigma, c2.lag, c2.hold, c2.min_steps, c2.max_steps,
            c2.avg_steps, c2.avg_sample) == (1.0, 1e-3, 1e-2, 3.0, 5000, 3, 10000, 500000,
                                             20000, 50)
    assert c2.L0 == 4.0 and c2.validate_deg < 0 and c2.precision == _lib.FP_F32_REF2D


def test_check_mask_codes(lib):
    """read_structure's checks (LMB3d/lbm3d-perm.cpp:182-192) as statuses"""
    from flowperm import _lib
    n = C.c_int64(0)
    m = np.zeros((3, 4, 5), np.uint8)
    m[0] = 1
    assert lib.fp_check_mask3d(_lib.u8ptr(m), 5, 4, 3, C.byref(n)) == 0 and n.value == 40
    m[1, 1, 1] = 2
    assert lib.fp_check_mask3d(_lib.u8ptr(m), 5, 4, 3, None) == _lib.FP_ERR_NON_BINARY
    one = np.ones((3, 4, 5), np.uint8)
    assert lib.fp_check_mask3d(_lib.u8ptr(one), 5, 4, 3, None) == _lib.FP_ERR_NO_PORE
    assert lib.fp_check_mask3d(_lib.u8ptr(one), 0, 4, 3, None) == _lib.FP_ERR_BAD_DIMS


def test_jacobi3_matches_eigh(lib):
    rng = np.random.default_rng(5)
    for _ in range(20):
        A = rng.normal(size=(3, 3))
        A = A + A.T
        val = (C.c_double * 3)()
        vec = (C.c_double * 9)()
        lib.fp_jacobi3(np.ascontiguousarray(A).ctypes.data_as(C.POINTER(C.c_double)), val, vec)
        w = np.linalg.eigvalsh(A)[::-1]
        assert np.allclose(list(val), w, atol=1e-12)
        V = np.array(list(vec)).reshape(3, 3)
        assert np.allclose(A @ V, V * np.array(list(val)), atol=1e-10)


def test_tune_shapes_are_valid(lib):
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    for i in range(24):
        lib.fp_tune_shape(0, i, C.byref(a), C.byref(b), C.byref(c))
        assert a.value in (1, 2, 4) and 0 <= b.value <= 32 and c.value in (0, 1)
    modes, segs = set(), set()
    for i in range(16):
        lib.fp_tune_shape(1, i, C.byref(a), C.byref(b), C.byref(c))
        assert 1 <= a.value <= 8 and 0 <= b.value <= 32 and c.value in (0, 1)
        modes.add(c.value)
        waves = a.value
        lib.fp_tune_shape(2, i, C.byref(a), C.byref(b), C.byref(c))
        assert a.value in (1, 2, 4, 8) and a.value <= waves
        segs.add(a.value)
    assert modes == {0, 1}                      # both neighbour-step forms are tried
    assert {1, 2, 4} <= segs                    # and rows 1, 2 and 4 segments wide


def test_now_is_monotonic(lib):
    t = [lib.fp_now() for _ in range(100)]
    assert all(b >= a for a, b in zip(t, t[1:]))


def test_neighbour_link_tables_match_a_dense_lookup(lib):
    """the per-segment link records of the row-segment neighbour step and the
    per-cell records of the compact one (decoded by the kernel's own
    compact_offsets) both give, for every fluid cell and direction, the slot of
    the periodic neighbour at cell - e_k (or the cell's own slot when that
    neighbour is solid): random lattices of every width class (x-runs of 1, 2,
    < 64, 64, > 64 cells, where x - 1 and x + 1 wrap) in 3D and 2D"""
    from flowperm import _lib
    rng = np.random.default_rng(0)
    shapes = [(1, 1, 1), (2, 3, 1), (3, 3, 2), (3, 2, 63), (2, 2, 64), (3, 4, 65), (2, 3, 128),
              (5, 4, 130), (9, 7, 11), (16, 16, 16)]
    for shp in shapes:
        for dim in (3, 2):
            s = shp if dim == 3 else shp[1:]
            for por in (0.3, 0.7, 1.0):
                m = (rng.random(s) > por).astype(np.uint8)
                m.flat[0] = 0
                bad = C.c_int64(-1)
                if dim == 3:
                    nz, ny, nx = s
                else:
                    (ny, nx), nz = s, 1
                st = lib.fp_verify_links(_lib.u8ptr(np.ascontiguousarray(m)), dim, nx, ny, nz,
                                         C.byref(bad))
                assert st == 0 and bad.value == 0, (s, dim, por, bad.value)


C_CALLER = r"""
#include <stdio.h>
#include <stdlib.h>
#include "flowperm.h"
int main(void)
{
    printf("abi %d\n", fp_abi_version());
    fp_config3d c;
    fp_config3d_default(&c);
    printf("defaults %g %g %lld %d %lld %lld %lld\n", c.ff, c.eps, (long long)c.lag, c.hold,
           (long long)c.min_steps, (long long)c.max_steps, (long long)c.avg_steps);
    printf("exit %d %d %d %d\n", fp_exit_code(FP_OK), fp_exit_code(FP_NOT_CONVERGED),
           fp_exit_code(FP_ERR_NON_BINARY), fp_exit_code(FP_ERR_NO_PORE));
    uint8_t m[8] = {0, 1, 1, 0, 1, 1, 1, 0};
    int64_t np = -1;
    int ms = fp_check_mask3d(m, 2, 2, 2, &np);
    printf("mask %d %lld\n", ms, (long long)np);
    m[3] = 7;
    printf("mask %d\n", fp_check_mask3d(m, 2, 2, 2, &np));
    /* a constant flux under lag 100, hold 2, min 200, avg 100: checks at 100
     * (recorded), 200 and 300 (hold 1, 2), then 100 steps of averaging */
    c.lag = 100; c.min_steps = 200; c.max_steps = 2000; c.avg_steps = 100; c.eps = 1e-6;
    enum { N = 1000 };
    double *q = malloc(sizeof(double) * 3 * N);
    for(int s = 0; s < N; s++) { q[3*s] = 1.0 / 1024; q[3*s+1] = -1.0 / 2048; q[3*s+2] = 0.0; }
    fp_solve_out o;
    int st = fp_replay_convergence3d(&c, q, N, &o);
    printf("replay %d %lld %d %.17g %lld %lld %d\n", st, (long long)o.steps, o.converged, o.q[0],
           (long long)o.check_step, (long long)o.checks, o.check_hold);
    free(q);
    return 0;
}
"""


def test_a_c_caller_links_and_calls_the_host_entry_points(tmp_path):
    """the boundary as a C program sees it: the header compiles as C99, the
    program links against libflowperm.so and the entry points that need no
    GPU (ABI version, reference defaults, exit codes, the mask check, the
    convergence rules on a planted flux) return what the reference defines"""
    from flowperm import _lib
    src = tmp_path / "caller.c"
    src.write_text(C_CALLER)
    exe = tmp_path / "caller"
    libdir = _lib.LIB_PATH.parent
    inc = REPO / "include"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", str(inc), "-o", str(exe), str(src),
                    "-L", str(libdir), "-lflowperm", f"-Wl,-rpath,{libdir}"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True, timeout=120).stdout
    lines = out.splitlines()
    version = re.search(r"FP_ABI_VERSION (\d+)", (inc / "flowperm.h").read_text()).group(1)
    assert lines[0] == f"abi {version}"
    assert lines[1] == "defaults 1 1e-05 2000 2 4000 200000 2000"     # LMB3d/lbm3d-perm.cpp:118-130
    assert lines[2] == "exit 0 3 2 2"
    assert lines[3] == "mask 0 3"
    assert lines[4].startswith("mask ") and lines[4] != "mask 0"
    assert lines[5] == f"replay 0 400 1 {1.0 / 1024!r} 300 3 2"
