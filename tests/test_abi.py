"""The C-ABI library on the CPU: it loads, exports every entry point include/obvhs_hip.h declares, the header is
plain C (a C11 program compiles against it with every warning an error, links the library and calls its host-only
entry points), and its host-side CWBVH exponent table equals what the host libm's log2f gives
(bvh2_to_cwbvh.rs:85-88).  No GPU calls."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle_lib as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "obvhs_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(obvhs_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "obvhs_ploc_build_aabbs" in names and "obvhs_cwbvh_ray_traverse" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    from obvhs_amd import native
    lib = native.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # every symbol the Python binding types is declared in the header
    assert set(native.SIGNATURES) <= set(declared_functions())


def test_status_strings():
    from obvhs_amd import native
    lib = native.lib()
    for code in native.STATUS:
        assert lib.obvhs_status_string(code)


C_PROGRAM = r"""
#include <stdio.h>
#include <string.h>
#include "obvhs_hip.h"
int main(void) {
    uint32_t table[256];
    obvhs_status s = obvhs_log2_threshold_table(table);
    if (s != OBVHS_OK) return 1;
    if (strlen(obvhs_status_string(OBVHS_ERR_INVALID_ARG)) == 0) return 2;
    printf("%u %u\n", table[127], table[200]);
    return 0;
}
"""


def test_header_is_c_and_links(tmp_path):
    import shutil
    import subprocess
    from obvhs_amd import native
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "abi.c"
    src.write_text(C_PROGRAM)
    exe = tmp_path / "abi"
    libdir = os.path.dirname(native.LIB_PATH)
    r = subprocess.run([gcc, "-std=c11", "-Wall", "-Wextra", "-Werror", "-pedantic",
                        "-I", os.path.join(REPO, "include"), str(src), "-L", libdir, "-lobvhs_hip",
                        "-Wl,-rpath," + libdir, "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, (r.returncode, r.stderr)
    t127, t200 = map(int, r.stdout.split())
    assert t127 == 0 and t200 > 0


def f32_from_bits(b):
    return np.array([b], dtype=np.uint32).view(np.float32)[0]


def test_log2_threshold_table_matches_libm():
    from obvhs_amd import log2_threshold_table
    table = log2_threshold_table()
    log2f = O.lib().orc_log2f
    for E in range(1, 255):
        k = float(E - 127)
        t = int(table[E])
        assert log2f(f32_from_bits((E << 23) | t)) == k
        assert log2f(f32_from_bits((E << 23) | (t + 1))) != k
    # large exponents round up to exactly k for a few ulps above 2^k; k = 0 only at 1.0 exactly
    assert table[127] == 0 and table[200] > 0


def test_context_without_device_reports_no_device():
    import obvhs_amd as ob
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(ob.ObvhsError) as e:
        ob.Context(0)
    assert e.value.name in ("OBVHS_ERR_NO_DEVICE", "OBVHS_ERR_DEVICE")
