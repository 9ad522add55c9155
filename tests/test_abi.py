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
