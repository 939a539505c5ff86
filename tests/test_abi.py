"""The C-ABI library loads, exports every symbol include/scaletrack.h declares, and its
argument checks return clean status codes (no GPU work is issued by these tests)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from scaletrack import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "scaletrack.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(st_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported():
    names = declared()
    assert len(names) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (st_[A-Za-z0-9_]+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, f"declared but not exported: {missing}"
    assert set(names) == set(_lib.SIGNATURES), "ctypes signatures out of sync with the header"


def test_version_and_defaults():
    assert _lib.lib.st_version() == 2
    p = _lib.default_params()
    assert (p.boundary, p.method, p.flush, p.substeps_per_pass) == (0, 1, 0, 0)
    assert p.cd == 1.0 and list(p.brick) == [32, 8, 2] and p.block == 512
    assert 0 < p.resort_fraction < 1
    assert p.passes_per_launch == 1


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors of st_mesh / st_params have the C compiler's layout."""
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "scaletrack.h"\n'
        "int main(void) {\n"
        '  printf("%zu %zu %zu\\n", sizeof(st_mesh), offsetof(st_mesh, origin), offsetof(st_mesh, ending));\n'
        '  printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(st_params), offsetof(st_params, cd),\n'
        "         offsetof(st_params, brick), offsetof(st_params, block), offsetof(st_params, resort_fraction),\n"
        "         offsetof(st_params, passes_per_launch));\n"
        "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    l1, l2 = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")[:2]
    M, P = _lib.st_mesh, _lib.st_params
    assert [int(v) for v in l1.split()] == [C.sizeof(M), M.origin.offset, M.ending.offset]
    assert [int(v) for v in l2.split()] == [C.sizeof(P), P.cd.offset, P.brick.offset, P.block.offset,
                                            P.resort_fraction.offset, P.passes_per_launch.offset]


def test_null_arguments_are_rejected():
    lib = _lib.lib
    assert lib.st_ctx_create(0, None) == _lib.ST_E_INVALID
    assert b"null" in lib.st_last_error()
    assert lib.st_ctx_synchronize(None) == _lib.ST_E_INVALID
    assert lib.st_evolve(None, None, None, 1.0, 10, None, None) == _lib.ST_E_INVALID
    assert lib.st_cloud_upload(None, *([None] * 8)) == _lib.ST_E_INVALID
    assert lib.st_field_upload_U(None, None, 8) == _lib.ST_E_INVALID
    assert lib.st_cloud_count(None, None) == _lib.ST_E_INVALID
    # destroying null handles is a no-op
    for fn in ("st_ctx_destroy", "st_cloud_destroy", "st_field_destroy", "st_graph_destroy"):
        assert getattr(lib, fn)(None) == _lib.ST_OK
    # page-locked host memory: sizes below one byte and a null output are refused
    p = C.c_void_p(1234)
    assert lib.st_host_alloc(0, C.byref(p)) == _lib.ST_E_INVALID and not p.value
    assert lib.st_host_alloc(-8, C.byref(p)) == _lib.ST_E_INVALID
    assert lib.st_host_alloc(64, None) == _lib.ST_E_INVALID
    assert lib.st_host_free(None) == _lib.ST_OK


def test_no_device_is_a_clean_error():
    n = C.c_int32(-1)
    r = _lib.lib.st_device_count(C.byref(n))
    if r == _lib.ST_OK and n.value > 0:
        pytest.skip("a GPU is visible")
    h = C.c_void_p()
    r = _lib.lib.st_ctx_create(0, C.byref(h))
    assert r == _lib.ST_E_DEVICE and not h.value
    assert _lib.lib.st_last_error()
    # page-locked host memory needs the device runtime as well
    p = C.c_void_p()
    assert _lib.lib.st_host_alloc(64, C.byref(p)) == _lib.ST_E_DEVICE and not p.value
    import scaletrack as stk
    with pytest.raises(_lib.ScaleTrackError) as ei:
        stk.pinned_empty(8)
    assert ei.value.status == _lib.ST_E_DEVICE


def test_error_raises_with_message():
    with pytest.raises(_lib.ScaleTrackError) as ei:
        _lib.check(_lib.lib.st_ctx_synchronize(None))
    assert "ST_E_INVALID" in str(ei.value) and "null" in str(ei.value)
