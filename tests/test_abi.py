"""The C ABI library loads and exports every entry point include/gqf.h
declares (no compute: runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    txt = open(os.path.join(ROOT, "include", "gqf.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b(\w+)\s*\([^;{]*\)\s*;", txt, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_reference_entry_points():
    names = _declared_functions()
    for n in ("qf_malloc_device", "qf_destroy_device", "bulk_insert", "bulk_insert_reduce",
              "bulk_delete", "bulk_get", "bulk_get_misses_wrapper", "point_get_wrapper"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from cqf_amd import gqf
    lib = ctypes.CDLL(gqf.LIB_PATH)
    missing = [n for n in _declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_declared_functions()) <= set(gqf.EXPORTS) | set(_declared_functions())


def test_python_mirror_lists_all_exports():
    from cqf_amd import gqf
    assert set(_declared_functions()) == set(gqf.EXPORTS)


def test_batch_pieces_arithmetic():
    """How a bulk mutation longer than the pipeline's 32-bit indices allow is
    cut: consecutive pieces of at most max_piece keys that cover [0, n)
    exactly (host arithmetic only)."""
    from cqf_amd import gqf
    P = 1 << 30
    assert gqf.batch_pieces(0, P) == [0]
    assert gqf.batch_pieces(1, P) == [0, 1]
    assert gqf.batch_pieces(P, P) == [0, P]
    assert gqf.batch_pieces(P + 1, P) == [0, P, P + 1]
    n = (1 << 32) + 12345  # 2^32 keys or more would wrap a 32-bit region start
    st = gqf.batch_pieces(n, P)
    assert st[0] == 0 and st[-1] == n and len(st) == 6
    assert all(0 < b - a <= P for a, b in zip(st, st[1:]))
    assert all(b - a < (1 << 31) for a, b in zip(st, st[1:]))
    for n, m in ((10, 3), (9, 3), (7, 1), (5, 100)):
        st = gqf.batch_pieces(n, m)
        assert st == list(range(0, n, m)) + [n]


def test_drop_in_caller_library_loads():
    """The reference-style caller (tests/drop_in) built by build() loads
    with libgqf.so and exports its entry points (no compute)."""
    from cqf_amd import gqf
    from tests.drop_in import build as B
    gqf.load_library()
    if not os.path.exists(B.LIB):
        B.build()
    lib = ctypes.CDLL(B.LIB)
    for n in ("dropin_gqf_init", "dropin_point_bulk_insert", "dropin_count_kmers", "dropin_gqf_bulk_get"):
        assert hasattr(lib, n)
