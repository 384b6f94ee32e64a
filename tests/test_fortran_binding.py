"""The Fortran side of the boundary, on the CPU: the BIND(C) types of
imex_hip_binding (imex_sflow2d_dev_amd/fortran/imex_hip_binding.f90) have
the sizes and member offsets of the C structs of include/imex_hip.h (and
of the ctypes mirrors in native.py); the replacement MODULE solver_2d
(fortran/solver_2d.f90) compiles against the reference's own modules, and the
reference's init_2d.f90 and stochastic_module.f90, which take q, qp, Z,
conv_kernel, solve_cells, j_cent and k_cent from solver_2d, compile
unchanged against it; the Fortran host program imex_fdriver is linked
against libimexhip.so."""
import ctypes
import os
import shutil
import subprocess

import pytest

from conftest import ROOT
from imex_sflow2d_dev_amd import build, native

FDIR = os.path.join(ROOT, "tests", "fortran")
REFSRC = "/root/reference/src"
FC = build.FC

needs_fc = pytest.mark.skipif(not os.path.exists(FC) and shutil.which("amdflang") is None,
                              reason="amdflang not installed")


def _layout(lines):
    out = {}
    for ln in lines.split("\n"):
        if ln.strip():
            k, v = ln.split()
            out[k] = int(v)
    return out


@needs_fc
def test_bind_c_types_match_the_c_structs(tmp_path):
    mod = os.path.join(build.LIBDIR, "imex_hip_binding.mod")
    if not os.path.exists(mod):
        build.build_fortran()
    subprocess.run([FC, os.path.join(FDIR, "layout_check.f90"), "-I", build.LIBDIR, "-o",
                    str(tmp_path / "lf")], check=True, capture_output=True)
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"),
                    os.path.join(FDIR, "layout_check.c"), "-o", str(tmp_path / "lc")],
                   check=True, capture_output=True)
    f = _layout(subprocess.run([str(tmp_path / "lf")], check=True, capture_output=True,
                               text=True).stdout)
    c = _layout(subprocess.run([str(tmp_path / "lc")], check=True, capture_output=True,
                               text=True).stdout)
    assert len(c) == 4 + 45 + 9 + 3 + 7 + 9 + 17 + 1
    assert f == c
    # the ctypes mirrors of the Python binding have the same layout
    for name, cls in (("imex_config", native.ImexConfig), ("imex_static", native.ImexStatic),
                      ("imex_step_stats", native.ImexStepStats),
                      ("imex_runout_red", native.ImexRunoutRed)):
        assert ctypes.sizeof(cls) == c[f"{name}.size"], name
        for fname, _ in cls._fields_:
            assert getattr(cls, fname).offset == c[f"{name}.{fname}"], (name, fname)


@needs_fc
@pytest.mark.skipif(not os.path.isdir(REFSRC), reason="reference sources not present")
def test_replacement_solver_module_compiles_with_the_reference_modules(tmp_path):
    if not os.path.exists(os.path.join(build.LIBDIR, "imex_hip_binding.mod")):
        build.build_fortran()
    inc = ["-I", build.LIBDIR, "-J", str(tmp_path)]

    def fc(src):
        r = subprocess.run([FC, "-O1", "-c", src, "-o",
                            str(tmp_path / (os.path.basename(src)[:-4] + ".o"))] + inc,
                           cwd=tmp_path, capture_output=True, text=True)
        assert r.returncode == 0, src + "\n" + r.stdout + r.stderr

    for m in ("parameters_2d", "complexify", "geometry_2d", "constitutive_2d"):
        fc(os.path.join(REFSRC, m + ".f90"))
    fc(os.path.join(build.FDIR, "solver_2d.f90"))
    # the reference's modules that USE solver_2d, unchanged
    fc(os.path.join(REFSRC, "init_2d.f90"))
    fc(os.path.join(REFSRC, "stochastic_module.f90"))


def test_fortran_driver_is_linked_against_the_library():
    if not os.path.exists(build.FDRIVER):
        pytest.skip("imex_fdriver not built")
    r = subprocess.run(["readelf", "-d", build.FDRIVER], capture_output=True, text=True)
    assert "libimexhip.so" in r.stdout
    assert "$ORIGIN" in r.stdout
