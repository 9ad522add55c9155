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


def test_autotune_set_validates_and_table_is_well_formed():
    """gqf_autotune_set takes the choices a measurement could make and
    rejects the rest (host cache only, no GPU call); the committed table of
    choices, when present, holds only such entries."""
    from cqf_amd import gqf, tuning
    L = gqf.load_library()
    assert L.gqf_autotune_set(20, 0.9, 48, 0) == 0
    assert L.gqf_autotune_set(20, 0.9, 40, 8) == 0
    assert L.gqf_autotune_set(7, 0.9, 48, 0) == -2     # nbits below 8
    assert L.gqf_autotune_set(20, 0.0, 48, 0) == -2    # load not in (0, 1]
    assert L.gqf_autotune_set(20, 1.5, 48, 0) == -2
    assert L.gqf_autotune_set(20, 0.9, 12, 0) == -2    # region size outside 16..56
    assert L.gqf_autotune_set(20, 0.9, 64, 0) == -2
    assert L.gqf_autotune_set(20, 0.9, 48, 5) == -2    # digit width other than 0 or 6..10
    if os.path.exists(tuning.TABLE):
        choices = tuning.read()
        assert choices
        for (q, ld), (r, d) in choices.items():
            assert 8 <= q <= 32 and 0 < ld <= 1.0
            assert 16 <= r <= 56 and (d == 0 or 6 <= d <= 10)

def test_tuning_table_round_trip(tmp_path):
    """A table written by cqf_amd.tuning reads back to the same choices."""
    from cqf_amd import tuning
    choices = {(16, 0.5): (48, 0), (20, 0.95): (40, 8), (28, 0.9): (48, 0)}
    path = str(tmp_path / "t.json")
    tuning.write(choices, path, "test")
    assert tuning.read(path) == choices
