"""scalar.dat formatting of the reference's Scalar_Output (report_mpi.f90:253-257): the Fortran
Dw.d descriptor by the standard's rules, and the line layout (i10 then 9 or 7 d15.7 fields)."""
import pytest

from cssx import diagnostics as dg


@pytest.mark.parametrize("x,s", [
    (1.5, "  0.1500000D+01"),
    (-2.25e-3, " -0.2250000D-02"),
    (0.0, "  0.0000000D+00"),
    (9.99999999, "  0.1000000D+02"),
    (1e100, "  0.1000000+101"),
])
def test_fortran_d15_7(x, s):
    assert dg.fortran_d(x) == s and len(s) == 15


def test_scalar_line_layout():
    q = dict(kinener=1.0, cke=0.25, mcke=0.5, drke=0.125, umach=0.01, magener=2e-3, bmax=3.0)
    line = dg.scalar_line(42, 0.5, 1e-4, q)
    assert len(line) == 10 + 9 * 15 and line[:10] == "        42"
    assert line[10:25] == "  0.5000000D+00" and line[25:40] == "  0.1000000D-03"
    assert len(dg.scalar_line(42, 0.5, 1e-4, q, magnetic=False)) == 10 + 7 * 15
