"""The host writers' Fortran edit descriptors against the Fortran runtime's
own output of the same descriptors on the same values
(tests/golden/fortran_formats.json, written by make_fortran_formats.py)."""
import json
import struct

import pytest

from conftest import golden_path
from imex_sflow2d_dev_amd import fortran_format as ff

FIX = json.load(open(golden_path("fortran_formats.json")))


def _fmt(spec, x):
    if spec.startswith("F"):
        w, d = spec[1:].split(".")
        return ff.f_fmt(x, int(w), int(d))
    if spec.startswith("ES"):
        wd, e = spec[2:].split("E")
        w, d = wd.split(".")
        return ff.es_fmt(x, int(w), int(d), int(e))
    w, d = spec[1:].split(".")
    return ff.e_fmt(x, int(w), int(d))


@pytest.mark.parametrize("spec", ["F12.3", "F14.3", "F17.3", "F10.2", "ES14.7E2", "ES15.7E2",
                                  "E20.12"])
def test_numeric_descriptors_match_fortran_runtime(spec):
    bad = []
    for case in FIX["values"]:
        x = struct.unpack("<d", struct.pack("<q", case["bits"]))[0]
        got, want = _fmt(spec, x), case["out"][spec]
        if got != want:
            bad.append((x, got, want))
    assert not bad, bad[:10]


def test_character_descriptor_matches_fortran_runtime():
    want = FIX["strings"]
    got = [ff.a_fmt("Time (s),", 13), ff.a_fmt("X_runout (m),", 17),
           ff.a_fmt("maximum runout =", 18), ff.a_fmt("alfa_s(01),", 15), ff.a_fmt("abcdefgh", 4)]
    assert got == want
