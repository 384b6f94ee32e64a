"""Writes tests/golden/fortran_formats.json: the output of the Fortran
runtime (amdflang) for the edit descriptors of the reference's output files
(F12.3, F14.3, F17.3, F10.2, ES14.7E2, ES15.7E2, E20.12, A13/A15/A17/A18/A4)
on a fixed set of values -- random magnitudes, decimal ties, values that
round up to the next power of ten, signed zeros, tiny and huge exponents.

Run from the repository root: python3 tests/golden/make_fortran_formats.py
(needs amdflang; the fixture is committed)."""
import json
import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FC = os.environ.get("FC", "amdflang")
FMTS = ["F12.3", "F14.3", "F17.3", "F10.2", "ES14.7E2", "ES15.7E2", "E20.12"]


def values():
    rng = np.random.default_rng(20261203)
    v = [0.0, -0.0, 1.0, -1.0, 0.5, 1e-20, -1e-20, 1.5e-21, 123.4565, 12345678.5, 0.12345675,
         9.99999995, 9.9999999949, 99999999.95, 0.0005, -0.0005, 0.0015, -0.0004, 2.5, 0.125,
         1e99, 9.9999999e99, 1e100, 1e-99, 1e-100, 1.2345678901234e-5, 5e-324, 1.7976931348623157e308,
         300.0, 273.15, 6.02214076e23, 1234567890.123, 999999.9995, -999999.9995, 12345.6785,
         0.1, 0.2, 0.3, 1.0 / 3.0, 2.0 / 3.0, 1e12, 123456789012.5, 99999999999.9995]
    for e in range(-30, 31, 3):
        v += list(rng.uniform(1.0, 10.0, 3) * 10.0 ** e)
    v += list(-rng.uniform(0, 5000, 10))
    v += list(rng.uniform(0, 1e4, 10).round(3) + 0.0005)
    return [float(x) for x in v]


def main():
    vals = values()
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "ff")
        subprocess.run([FC, "-O0", "-o", exe, os.path.join(HERE, "fortran_formats.f90")], check=True)
        with open(os.path.join(d, "fortran_formats_in.txt"), "w") as f:
            for x in vals:
                f.write(f"{struct.unpack('<q', struct.pack('<d', x))[0]}\n")
        out = subprocess.run([exe], cwd=d, check=True, capture_output=True, text=True).stdout
    lines = [ln[ln.index("[") + 1:ln.rindex("]")] for ln in out.splitlines()]
    n = len(FMTS)
    cases = []
    for i, x in enumerate(vals):
        cases.append({"bits": struct.unpack('<q', struct.pack('<d', x))[0],
                      "out": dict(zip(FMTS, lines[i * n:(i + 1) * n]))})
    strings = lines[len(vals) * n:]
    json.dump({"compiler": FC, "values": cases, "strings": strings},
              open(os.path.join(HERE, "fortran_formats.json"), "w"), indent=0)
    print(f"{len(vals)} values, {len(strings)} strings")


if __name__ == "__main__":
    main()
