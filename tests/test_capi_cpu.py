"""CPU checks of the C-ABI library: it loads and exports every entry point that
include/kylinv.h declares; no compute call is made (no GPU here)."""
import os
import re
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import __graft_entry__ as GE  # noqa: E402


def declared_entry_points():
    src = open(os.path.join(ROOT, "include", "kylinv.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(kv_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_entry_point():
    path = os.path.join(ROOT, "kylin-v_amd", "libkylinv.so")
    if not os.path.exists(path):
        pytest.skip("libkylinv.so not built (run __graft_entry__.build())")
    kv = GE.load_package()
    lib = kv.library()
    names = declared_entry_points()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding declares a signature for every entry point
    assert set(names) == set(kv.SIGNATURES)


def test_parse_stage_matches_reference_format():
    kv = GE.load_package()
    assert kv.parse_stage("[0,40,0,0]") == [0, 40, 0, 0]
    assert kv.parse_stage("[3,50,2447,0]") == [3, 50, 2447, 0]


def test_header_is_plain_c(tmp_path):
    """include/kylinv.h compiles as strict C99 and as C++ (the boundary carries no C++ or
    HIP types)."""
    import shutil
    import subprocess
    src = tmp_path / "use.c"
    src.write_text('#include "kylinv.h"\nint main(void) { kv_ctx* c = 0; (void)c; return 0; }\n')
    inc = os.path.join(ROOT, "include")
    if shutil.which("gcc"):
        subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-fsyntax-only",
                        "-I", inc, str(src)], check=True)
    if shutil.which("g++"):
        subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-x", "c++", "-I", inc,
                        str(src)], check=True)
