"""CPU checks of the native library: it loads, exports every function that
include/cellgpu_amd.h declares, and its host-callable helpers (counter-based
normal generator, periodic wrap) give the published / reference answers.
No device compute is launched here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "cellgpu_amd.h")
SO = os.path.join(ROOT, "cellgpu_amd", "libcellgpu_amd.so")


def _declared():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b(cgpu_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


@pytest.fixture(scope="module")
def so():
    if not os.path.exists(SO):
        pytest.skip("libcellgpu_amd.so not built")
    return C.CDLL(SO)


def test_every_declared_symbol_is_exported(so):
    names = _declared()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_python_binding_table_resolves(so):
    """Every signature the Python front end declares names an exported symbol
    (loading binds them all; nothing is called)."""
    from cellgpu_amd import _lib
    L = _lib.lib()
    assert L is not None


def test_philox_known_answers(so):
    # Philox4x32-10 known-answer vectors of the Random123 distribution (Salmon et al. 2011)
    kat = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    f = so.cgpu_philox4x32_10
    f.argtypes = [C.c_void_p] * 3
    for ctr, key, want in kat:
        c = np.array(ctr, np.uint32)
        k = np.array(key, np.uint32)
        out = np.zeros(4, np.uint32)
        f(c.ctypes.data, k.ctypes.data, out.ctypes.data)
        assert tuple(int(v) for v in out) == want


def test_normal_pairs_are_standard_normal(so):
    f = so.cgpu_normal_pair
    f.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]
    out = np.zeros(2)
    z = []
    for i in range(20000):
        f(1234, i, 7, 0, out.ctypes.data)
        z.extend(out)
    z = np.array(z)
    assert abs(z.mean()) < 0.02 and abs(z.var() - 1.0) < 0.03
    assert abs(np.corrcoef(z[0::2], z[1::2])[0, 1]) < 0.03
    # a different step or stream gives a different deviate
    a, b = np.zeros(2), np.zeros(2)
    f(1234, 5, 7, 0, a.ctypes.data)
    f(1234, 5, 8, 0, b.ctypes.data)
    assert a[0] != b[0]


def test_wrap_matches_reference_loops(so, oracle):
    f = so.cgpu_put_in_box
    f.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    rng = np.random.default_rng(0)
    vals = list(rng.uniform(-60, 60, 3000)) + [-1e-300, -1e-17, -0.0, 0.0, 1e-20, -2e-16, 1e5 + 0.3,
                                               -1e5 - 0.7, 31.622776601683793, -31.622776601683793]
    for L in (31.622776601683793, 8.0, 512.0, 181.01933598375618):
        for v in vals:
            x, y = C.c_double(v * L / 10), C.c_double(-0.37 * v * L / 10)
            f(L, C.byref(x), C.byref(y))
            ox, oy = oracle.put_in_box(L, v * L / 10, -0.37 * v * L / 10)
            assert x.value == ox and y.value == oy


def test_serving_workload_is_deterministic():
    from cellgpu_amd.serving import poisson_requests
    a = poisson_requests(20, 5.0, seed=3)
    b = poisson_requests(20, 5.0, seed=3)
    assert [(r.n, r.steps, r.arrival) for r in a] == [(r.n, r.steps, r.arrival) for r in b]
    assert all(100 <= r.n <= 65536 and 1000 <= r.steps <= 100000 for r in a)
    assert all(b2.arrival > b1.arrival for b1, b2 in zip(a, a[1:]))
