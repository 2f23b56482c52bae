"""Checkpoint files in the reference's format (checkpoint.f90:38-75, :146-423), read the way
CSS-Python/read_checkpoint.py reads them.

Pins: the Fortran Ew.d and Iw edit descriptors by the standard's rules (mantissa 0.ddddddd,
two-digit exponent with E, three-digit exponent without it, sign and rounding), the i7.7
directory name, the raw REAL*8 files in Fortran (nth, nphi, nr) order, and bit-exact round trips.
"""
import os

import numpy as np
import pytest

from cssx import checkpoint as ck


@pytest.mark.parametrize("x,s", [
    (1.5, " 0.1500000E+01"),
    (-1.5, "-0.1500000E+01"),
    (0.0, " 0.0000000E+00"),
    (1.234567e-5, " 0.1234567E-04"),
    (9.99999999, " 0.1000000E+02"),      # rounds up into the next decade
    (0.72, " 0.7200000E+00"),
    (np.pi / 4, " 0.7853982E+00"),
    (1e100, " 0.1000000+101"),           # three-digit exponent: no E
    (-2.5e-120, "-0.2500000-119"),
])
def test_fortran_e14_7(x, s):
    assert ck.fortran_e(x) == s
    assert len(s) == 14


def test_fortran_i9_and_directory():
    assert ck.fortran_i(128) == "      128"
    assert ck.checkpoint_dir("run/Checkpoints/", 1234) == "run/Checkpoints/0001234/"


def test_header_roundtrip_and_touching_negatives(tmp_path):
    h = ck.Header(128, 256, 512, 8, 77, 1.0414630036549161e-4, -0.5, 0.72, 0.97, np.pi / 4, 3 * np.pi / 4,
                  -1.5, np.pi / 2)
    text = h.lines()
    line2 = text.splitlines()[1]
    assert len(line2) == 8 * 14 and "-0.5000000E+00-0.7200000" not in line2
    assert "-0.1500000E+01" in line2
    p = tmp_path / "header"
    p.write_text(text)
    g = ck.read_header(str(p))
    assert (g.nr, g.nth, g.nphi, g.nv, g.istep, g.extra) == (128, 256, 512, 8, 77, [])
    want = [h.deltat, h.sim_time, h.r1, h.r2, h.th1, h.th2, h.phi1, h.phi2]
    got = [g.deltat, g.sim_time, g.r1, g.r2, g.th1, g.th2, g.phi1, g.phi2]
    assert np.allclose(got, want, rtol=5e-7, atol=0)   # 7 significant digits
    # two negative fields touch when written with (8e14.7); the reader still splits them
    p.write_text("       16       12        8        5        3\n-0.1000000E+01-0.2000000E+01 0.3000000E+01"
                 " 0.4000000E+01-0.5000000E+01 0.6000000E+01 0.7000000E+01 0.8000000E+01\n")
    g = ck.read_header(str(p))
    assert [g.deltat, g.sim_time, g.r1, g.r2, g.th1] == [-1.0, -2.0, 3.0, 4.0, -5.0]


def test_state_roundtrip_and_fortran_layout(tmp_path):
    rng = np.random.default_rng(1)
    nr, nphi, nth = 6, 5, 7
    st = rng.standard_normal((8, nr, nphi, nth))
    d = ck.write(str(tmp_path) + "/Checkpoints/", 42, st, deltat=1e-4, sim_time=0.25, r1=0.72, r2=0.97,
                 th1=0.7, th2=2.4, phi1=0.0, phi2=1.5)
    assert d.endswith("0000042/")
    assert sorted(os.listdir(d)) == sorted(["header"] + ["checkpoint." + n for n in ck.NAMES])
    for iv, name in enumerate(ck.NAMES):
        raw = np.fromfile(os.path.join(d, "checkpoint." + name), dtype="float64")
        assert raw.size == nth * nphi * nr
        # CSS-Python/read_checkpoint.py: reshape((nth, nph, nr), order='F') is f(it, ip, ir)
        f = raw.reshape((nth, nphi, nr), order="F")
        assert np.array_equal(f, st[iv].transpose(2, 1, 0))
    hdr, back = ck.read(d)
    assert (hdr.nr, hdr.nth, hdr.nphi, hdr.nv, hdr.istep) == (nr, nth, nphi, 8, 42)
    assert np.array_equal(back, st)
    _, five = ck.read(d, nv=5)
    assert five.shape[0] == 5 and np.array_equal(five, st[:5])


def test_rejects_wrong_sizes(tmp_path):
    st = np.zeros((8, 3, 4, 5))
    d = ck.write(str(tmp_path) + "/c", 1, st, deltat=1.0, sim_time=0.0, r1=0, r2=1, th1=0, th2=1, phi1=0, phi2=1)
    np.zeros(7).tofile(os.path.join(d, "checkpoint.u"))
    with pytest.raises(ValueError):
        ck.read(d)
    with pytest.raises(ValueError):
        ck.write(str(tmp_path) + "/c", 2, np.zeros((3, 2, 2, 2)), deltat=1.0, sim_time=0.0, r1=0, r2=1, th1=0,
                 th2=1, phi1=0, phi2=1)
