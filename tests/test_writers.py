"""The host writers of the output files (imex_write_records,
imex_write_restart, imex_format_value; no GPU involved) against the Fortran
runtime: byte-for-byte the files oracle/_ref/records_writer writes with the
reference's format statements from the same records, and the per-value
fixture tests/golden/fortran_formats.json."""
import json
import os
import struct
import subprocess

import numpy as np
import pytest

from conftest import ROOT, golden_path
from imex_sflow2d_dev_amd import native

WRITER = os.path.join(ROOT, "oracle", "_ref", "records_writer")


def _need_writer():
    if not os.path.exists(WRITER):
        pytest.skip("oracle/_ref/records_writer not built (make -C oracle oracle)")


def _values(rng, n):
    """Magnitudes across the formats' exponent range, signed zeros, values
    the reference zeroes (|x| < 1e-20) and decimal ties."""
    v = rng.standard_normal(n) * 10.0 ** rng.integers(-25, 25, n)
    special = [0.0, -0.0, 1e-21, -3e-22, 0.125, 12345678.5, 9.99999995, 1e99, 1e100, 1e-100,
               -1e-99, 300.0, 1.5, 2.5e-5]
    v[:len(special)] = special
    return v


@pytest.mark.parametrize("fmt, key", [(native.FMT_E20_12, "E20.12"), (native.FMT_ES15_7, "ES15.7E2"),
                                      (native.FMT_ES14_7, "ES14.7E2")])
def test_format_value_matches_fortran_runtime(fmt, key):
    fix = json.load(open(golden_path("fortran_formats.json")))
    bad = []
    for case in fix["values"]:
        x = struct.unpack("<d", struct.pack("<q", case["bits"]))[0]
        got = native.format_value(fmt, x)
        if got != case["out"][key]:
            bad.append((x, got, case["out"][key]))
    assert not bad, bad[:5]


@pytest.mark.parametrize("mode, fmt, ncol", [("q2d", native.FMT_E20_12, 8),
                                             ("p2d", native.FMT_ES15_7, 27)])
@pytest.mark.parametrize("threads", [1, 5])
def test_records_file_matches_fortran_runtime(tmp_path, mode, fmt, ncol, threads):
    _need_writer()
    nx, ny = 13, 11
    rng = np.random.default_rng(3)
    rec = _values(rng, nx * ny * ncol).reshape(nx * ny, ncol)
    raw = tmp_path / "in.raw"
    rec.tofile(raw)
    want = tmp_path / "want"
    subprocess.run([WRITER, mode, str(raw), str(want), str(nx), str(ny), str(ncol)], check=True)
    got = tmp_path / "got"
    native.write_records(str(got), rec, nx, fmt, threads)
    assert got.read_bytes() == want.read_bytes()


def test_restart_file_matches_fortran_runtime(tmp_path):
    _need_writer()
    nx, ny, nv, ns = 9, 7, 6, 2
    N = nx * ny
    rng = np.random.default_rng(5)
    t, dt = 123.456, 0.0123
    q = rng.standard_normal(N * nv)
    B, er, dep, ero = (rng.standard_normal(N), np.zeros(N * ns), rng.random(N * ns), np.zeros(N * ns))
    hm, pd, mv = rng.random(N), rng.random(N), rng.random(N)
    raw = tmp_path / "in.raw"
    np.concatenate([[t, dt], q, B, er, dep, ero, hm, pd, mv]).tofile(raw)
    want = tmp_path / "want.bin"
    subprocess.run([WRITER, "restart", str(raw), str(want), str(nx), str(ny), str(nv), str(ns)],
                   check=True)
    got = tmp_path / "restart.bin"
    native.write_restart(str(got), nx, ny, nv, ns, t, dt, q, B, er, dep, ero, hm, pd, mv)
    assert got.read_bytes() == want.read_bytes()
    # a stochastic_flag run appends the Z record
    Z = rng.standard_normal(N)
    np.concatenate([[t, dt], q, B, er, dep, ero, hm, pd, mv, Z]).tofile(raw)
    subprocess.run([WRITER, "restart_z", str(raw), str(want), str(nx), str(ny), str(nv), str(ns)],
                   check=True)
    native.write_restart(str(got), nx, ny, nv, ns, t, dt, q, B, er, dep, ero, hm, pd, mv, Z)
    assert got.read_bytes() == want.read_bytes()
