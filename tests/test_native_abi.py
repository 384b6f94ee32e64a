"""The C ABI library: it loads without a GPU, exports every function that
include/imex_hip.h declares, and its host build of the path's x**y agrees
with a correctly rounded quad-precision pow (no GPU call is made here)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from oracle import oraclepy
from imex_sflow2d_dev_amd import native


def declared():
    text = open(os.path.join(ROOT, "include", "imex_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(imex_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    L = ctypes.CDLL(native.LIB_PATH)
    names = declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) <= set(native.EXPORTS)


def test_binding_structs_match_header_sizes():
    # imex_config: 10 int32 then doubles; the ctypes mirror must not pad differently
    assert ctypes.sizeof(native.ImexConfig) % 8 == 0
    assert native.ImexConfig.dx.offset == 40
    assert ctypes.sizeof(native.ImexStatic) == 7 * ctypes.sizeof(ctypes.c_void_p)


def test_config_layout_matches_compiled_header(tmp_path):
    """sizeof and field offsets of imex_config as gcc lays out the header,
    against the ctypes mirror."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    fields = [f[0] for f in native.ImexConfig._fields_]
    src = tmp_path / "layout.c"
    body = "".join(f'  printf("%zu\\n", offsetof(imex_config, {f}));\n' for f in fields)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "imex_hip.h"\n'
                   'int main(void) {\n  printf("%zu\\n", sizeof(imex_config));\n'
                   + body + '  return 0;\n}\n')
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", inc, "-o", str(exe), str(src)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True,
                                          text=True).stdout.split()]
    want = [ctypes.sizeof(native.ImexConfig)] + [getattr(native.ImexConfig, f).offset
                                                 for f in fields]
    assert got == want


@pytest.mark.parametrize("y, lo, hi", [(0.687, -8, 3), (4.65, -6, 0), (2.4, -12, 15),
                                       (None, -300, 300)])
def test_host_pow_is_correctly_rounded(y, lo, hi):
    L = native.lib()
    rng = np.random.default_rng(17)
    n = 20000
    x = 10.0 ** rng.uniform(lo, hi, n)
    ys = np.full(n, y) if y is not None else rng.uniform(-2.0, 2.0, n)
    out = np.zeros(n)
    assert L.imex_host_pow(ctypes.c_int(n), native._dp(x), native._dp(ys), native._dp(out)) == 0
    want = np.array([oraclepy.pow_cr(a, b) for a, b in zip(x, ys)])
    np.testing.assert_array_equal(out, want)


def test_host_pow_special_cases():
    L = native.lib()
    x = np.array([0.0, 1.0, 2.0, 0.0, 5.0])
    y = np.array([4.65, 2.4, 0.0, 0.0, 1.0])
    out = np.zeros(5)
    L.imex_host_pow(ctypes.c_int(5), native._dp(x), native._dp(y), native._dp(out))
    np.testing.assert_array_equal(out, [0.0, 1.0, 1.0, 1.0, 5.0])
