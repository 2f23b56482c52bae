"""Record files of the reference's report layer (report_mpi.f90:441-1291) as CSS-Python reads them.

The CSS-Python readers (read_shell_avg.py, read_az_avg.py, read_shell_slice.py) are Python 2;
their decoding is restated here (each step cites the reader's lines): the file is read whole as
float64, reshaped in Fortran order to (nr, nq+2, nrecs), (nth, nr+2, nq, nrecs) or
(nth, nphi+1, nradii, nq, nrecs), and the header values and data are sliced out of it.  The CPU
tests check cssx.records.write_record and the header layout against that decoding; the GPU
tests check the quantities the records carry (torch float64 on the same state).
"""
from __future__ import annotations

import math
import os

import numpy as np
import pytest


def read_shell_avg(path):
    """read_shell_avg.py:16-76 (not oldfile)."""
    data = np.fromfile(path, dtype="<f8")
    nr, nq = int(data[0]), int(data[1])
    nrecs = max(1, len(data) // (nr * (nq + 2)))
    data = data.reshape((nr, nq + 2, nrecs), order="F")
    return {"nr": nr, "nq": nq, "r1": data[2, 0, 0], "r2": data[3, 0, 0], "values": data[:, 2:nq + 2, :],
            "radius": data[:, 1, 0], "quantities": data[4:nq + 4, 0, 0].astype(int), "nrecs": nrecs}


def read_az_avg(path):
    """read_az_avg.py:15-60."""
    data = np.fromfile(path, dtype="<f8")
    nth, nr, nq = int(data[0]), int(data[1]), int(data[2])
    nrecs = max(1, len(data) // (nth * (nr + 2) * nq))
    data = data.reshape((nth, nr + 2, nq, nrecs), order="F")
    return {"nth": nth, "nr": nr, "nq": nq, "r1": data[3, 0, 0, 0], "r2": data[4, 0, 0, 0],
            "th1": data[5, 0, 0, 0], "th2": data[6, 0, 0, 0], "values": data[:, 2:, :, :],
            "quantities": data[7:nq + 7, 0, 0, 0].astype(int), "radius": data[0:nr, 1, 0, 0], "nrecs": nrecs}


def read_shell_slice(path):
    """read_shell_slice.py:16-92."""
    data = np.fromfile(path, dtype="<f8")
    nradii, nr, nth, nphi, nq = (int(x) for x in data[:5])
    nrecs = max(1, len(data) // (nth * (nphi + 1) * nradii * nq))
    data = data.reshape((nth, nphi + 1, nradii, nq, nrecs), order="F")
    return {"nradii": nradii, "nr": nr, "nth": nth, "nphi": nphi, "nq": nq,
            "bounds": data[5:11, 0, 0, 0, 0], "quantities": data[11:nq + 11, 0, 0, 0, 0].astype(int),
            "r_ind": data[nq + 11:nradii + nq + 11, 0, 0, 0, 0].astype(int), "values": data[:, 1:nphi + 1, :, :, :],
            "nrecs": nrecs}


def test_direct_access_records(tmp_path):
    from cssx.records import write_record
    rng = np.random.default_rng(3)
    a, b = rng.standard_normal((6, 5)), rng.standard_normal((6, 5))
    p = os.path.join(tmp_path, "f")
    write_record(p, b, 2)          # records may be written out of order, as with rec=
    write_record(p, a, 1)
    raw = np.fromfile(p, dtype="<f8")
    assert raw.size == 60
    assert np.array_equal(raw[:30].reshape((6, 5), order="F"), a)
    assert np.array_equal(raw[30:].reshape((6, 5), order="F"), b)
    write_record(p, -a, 2)         # rewriting a record leaves the others
    raw = np.fromfile(p, dtype="<f8")
    assert np.array_equal(raw[:30].reshape((6, 5), order="F"), a) and raw.size == 60
    with pytest.raises(ValueError):
        write_record(p, a, 0)


def test_shell_avg_layout_decodes(tmp_path):
    """A record built like Shell_Average_Output's header (report_mpi.f90:1005-1011) decodes."""
    from cssx.records import write_record
    nr, q = 12, [1, 5, 9]
    rec = np.zeros((nr, len(q) + 2))
    rec[:, 2:] = np.arange(nr * len(q)).reshape(nr, len(q), order="F")
    rec[0:4, 0] = [nr, len(q), 0.72, 0.97]
    rec[4:7, 0] = q
    rec[:, 1] = np.linspace(0.72, 0.97, nr)
    p = os.path.join(tmp_path, "Shell_Avgs")
    write_record(p, rec, 1)
    write_record(p, 2 * rec, 2)
    d = read_shell_avg(p)
    assert (d["nr"], d["nq"], d["nrecs"]) == (nr, 3, 2)
    assert list(d["quantities"]) == q and d["r1"] == 0.72 and d["r2"] == 0.97
    assert np.array_equal(d["values"][:, :, 0], rec[:, 2:])
    assert np.array_equal(d["radius"], rec[:, 1])


@pytest.fixture(scope="module")
def run():
    import torch
    from cssx.mhd import Solver
    from cssx.model import Config, Model
    if not torch.cuda.is_available():
        pytest.skip("needs the GPU")
    c = Config(nr=24, nth=40, nphi=32, omega0=0.2, bamp=0.5)   # nr >= 18 quantities + 4 header rows
    m = Model(c)
    st = m.initial_state()
    rng = np.random.default_rng(11)
    st[1:4] = 0.05 * rng.standard_normal(st[1:4].shape)
    sol = Solver(m)
    y = torch.as_tensor(st, device="cuda").contiguous()
    _, B = sol.rhs(y, want_b=True)
    yield m, st, sol, y, B.cpu().numpy()
    sol.close()


def _point_fields(m, st, B):
    rho, w, v, u, s = st[:5]
    T = m.invtconst * rho ** m.gamm1 * np.exp(s / m.Cv)
    P = m.gamm1 * m.Cv * rho * T
    bt, bp, br = B
    return rho, w, v, u, s, T, P, bt, bp, br


@pytest.mark.gpu
def test_shell_average_record(run, tmp_path):
    from cssx.records import SHELL_QUANTITIES, shell_average_record, write_record
    m, st, sol, y, B = run
    q = [k for k in SHELL_QUANTITIES if k != 9] + [9]
    rec = shell_average_record(sol, y, q)
    p = os.path.join(tmp_path, "Shell_Avgs")
    write_record(p, rec, 1)
    d = read_shell_avg(p)
    assert (d["nr"], d["nq"]) == (m.c.nr, len(q)) and list(d["quantities"]) == q
    assert np.allclose(d["radius"], m.r, rtol=0, atol=0)
    rho, w, v, u, s, T, P, bt, bp, br = _point_fields(m, st, B)
    mean = lambda x: x.mean(axis=(1, 2))              # noqa: E731  (Sum / (nth nphi))
    usq = 0.5 * rho * (u * u + v * v + w * w)
    ref = {1: mean(rho), 2: mean(T), 3: mean(s), 4: mean(P), 5: mean(u), 6: mean(w), 7: mean(v), 8: mean(usq),
           10: mean(br), 11: mean(bt), 12: mean(bp), 13: mean((br ** 2 + bt ** 2 + bp ** 2) * 0.125 / math.pi),
           14: mean(np.sqrt(br ** 2 + bt ** 2)), 15: np.zeros(m.c.nr),
           16: m.c.Cp * mean((u - mean(u)[:, None, None]) * rho * (T - mean(T)[:, None, None])),
           17: mean(usq * (u - mean(u)[:, None, None])),
           18: mean((P - mean(P)[:, None, None]) * (u - mean(u)[:, None, None]))}
    for j, k in enumerate(q):
        if k == 9:
            continue
        got = d["values"][:, j, 0]
        assert np.abs(got - ref[k]).max() <= 1e-12 * max(np.abs(ref[k]).max(), 1e-300), k
    # ds/dr: the compact derivative of the oracle restatement with the flux-wall value
    import oracle
    lo = np.zeros((4, m.c.nphi, m.c.nth))
    lo[0] = (m.dsdr1 / m.dxdr[0])[None, :]
    ds = oracle.deriv_field(0, 1, s, bdry=m.ibcr[5], ghost_lo=lo, scale=m.dri) * m.dxdr[:, None, None]
    got = d["values"][:, q.index(9), 0]
    # rounding bound of the compact derivative: 128 eps max|s| dri max(dx/dr); the entropy (the
    # absolute state) is far larger than its radial derivative, and the wall rows combine it
    # with coefficient sums near 46
    lim = 128 * np.finfo(np.float64).eps * np.abs(s).max() * m.dri * np.abs(m.dxdr).max()
    assert np.abs(got - mean(ds)).max() <= lim


@pytest.mark.gpu
def test_azimuthal_average_record(run, tmp_path):
    from cssx.records import AZ_QUANTITIES, azimuthal_average_record, write_record
    m, st, sol, y, B = run
    q = list(AZ_QUANTITIES)
    rec = azimuthal_average_record(sol, y, q)
    p = os.path.join(tmp_path, "AZ_Avgs")
    write_record(p, rec, 1)
    write_record(p, rec, 2)
    d = read_az_avg(p)
    c = m.c
    assert (d["nth"], d["nr"], d["nq"], d["nrecs"]) == (c.nth, c.nr, len(q), 2)
    assert list(d["quantities"]) == q and (d["th1"], d["th2"]) == (c.th1, c.th2)
    assert np.array_equal(d["radius"], m.r)
    rho, w, v, u, s, T, P, bt, bp, br = _point_fields(m, st, B)
    am = lambda x: x.mean(axis=1).T                    # noqa: E731  (nth, nr)
    sm = lambda x: x.mean(axis=(1, 2))[:, None, None]  # noqa: E731
    rs = (m.r[:, None] * np.sin(m.theta)[None, :]).T   # (nth, nr)
    vb = am(v)
    usq = 0.5 * rho * (u * u + v * v + w * w)
    ref = {1: am(rho - sm(rho)), 2: am(T - sm(T)), 3: am(s - sm(s)), 4: am(P - sm(P)), 5: am(u), 6: am(w),
           7: vb, 8: am(usq), 9: am(br), 10: am(bt), 11: am(bp),
           12: am((br ** 2 + bt ** 2 + bp ** 2) / (8 * math.pi)), 13: am(np.sqrt(br ** 2 + bt ** 2)),
           14: rs * am(rho * u) * (vb + rs * c.omega0), 15: rs * am(rho * w) * (vb + rs * c.omega0),
           16: rs * am(rho) * am((u - am(u).T[:, None, :]) * (v - vb.T[:, None, :])),
           17: rs * am(rho) * am((w - am(w).T[:, None, :]) * (v - vb.T[:, None, :])),
           20: rs / (4 * math.pi) * am(br) * am(bp), 21: rs / (4 * math.pi) * am(bt) * am(bp),
           22: rs / (4 * math.pi) * c.nphi * am((br - am(br).T[:, None, :]) * (bp - am(bp).T[:, None, :])),
           23: rs / (4 * math.pi) * c.nphi * am((bt - am(bt).T[:, None, :]) * (bp - am(bp).T[:, None, :]))}
    # fluctuations about the shell mean (1-4) can vanish (rho here is the background alone):
    # their scale includes the field's own magnitude, whose rounding they carry
    field_scale = {1: np.abs(rho).max(), 2: np.abs(T).max(), 3: np.abs(s).max(), 4: np.abs(P).max()}
    for j, k in enumerate(q):
        got = d["values"][:, :, j, 0]
        sc = max(np.abs(ref[k]).max(), 1e-3 * field_scale.get(k, 0.0), 1e-300)
        assert np.abs(got - ref[k]).max() <= 1e-11 * sc, k


@pytest.mark.gpu
def test_shell_slice_record(run, tmp_path):
    from cssx.records import SLICE_QUANTITIES, shell_slice_record, write_record
    m, st, sol, y, B = run
    q, radii = list(SLICE_QUANTITIES), [1, 7, 20]
    rec = shell_slice_record(sol, y, q, radii)
    p = os.path.join(tmp_path, "Shell_Slices")
    write_record(p, rec, 1)
    d = read_shell_slice(p)
    c = m.c
    assert (d["nradii"], d["nr"], d["nth"], d["nphi"], d["nq"]) == (3, c.nr, c.nth, c.nphi, len(q))
    assert list(d["quantities"]) == q and list(d["r_ind"]) == radii
    assert np.array_equal(d["bounds"], [c.r1, c.r2, c.th1, c.th2, c.phi1, c.phi2])
    rho, w, v, u, s, T, P, bt, bp, br = _point_fields(m, st, B)
    fields = {1: rho, 2: T, 3: s, 4: P, 5: u, 6: w, 7: v, 8: 0.5 * rho * (u * u + v * v + w * w), 10: br, 11: bt,
              12: bp, 13: (br ** 2 + bt ** 2 + bp ** 2) / (8 * math.pi), 14: np.sqrt(br ** 2 + bt ** 2)}
    for j, k in enumerate(q):
        for kk, ir in enumerate(radii):
            want = fields[k][ir - 1].T                 # (nth, nphi)
            got = d["values"][:, :, kk, j, 0]
            assert np.abs(got - want).max() <= 1e-13 * max(np.abs(want).max(), 1e-300), (k, ir)
