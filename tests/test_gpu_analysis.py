"""GPU-backed CSS-Python analysis routines (cssx.analysis) against the oracle's restatement of
the numpy compact_fd6 applied pencil by pencil, as CSS-Python/vectorcurl.py does it, and
shell_avg's Simpson rule against scipy and closed forms."""
import numpy as np
import pytest
import scipy.integrate

import oracle

pytestmark = pytest.mark.gpu


def _curl_loops(data, rad, theta, phi):
    """vectorcurl.py:55-153 with the oracle's compact_fd6 restatement in the same loops."""
    iphi, ith, irad = 0, 1, 2
    f6 = lambda *a, **k: oracle.compact_fd6_py(*a, **k)[0]
    dri = 1. / (rad[-1] - rad[0])
    dti = 1. / (theta[-1] - theta[0])
    dpi = 1. / (phi[-1] - phi[0])
    sines, cosec, rinv = np.sin(theta), 1. / np.sin(theta), 1. / rad
    nth, nphi, nr = len(theta), len(phi), len(rad)
    curl = np.empty((nth, nphi, nr, 3))
    for ir in range(nr):
        for ip in range(nphi):
            curl[:, ip, ir, irad] = rinv[ir] * cosec * f6(dti, sines * data[:, ip, ir, iphi], 0., 0., 4)
            curl[:, ip, ir, iphi] = -rinv[ir] * f6(dti, data[:, ip, ir, irad], 0., 0., 0)
    for ir in range(nr):
        for it in range(nth):
            b1, b2 = data[it, nphi - 4:, ir, irad], data[it, :4, ir, irad]
            curl[it, :, ir, ith] = rinv[ir] * cosec[it] * f6(dpi, data[it, :, ir, irad], b1, b2, 4, dtype=3)
            b1, b2 = data[it, nphi - 4:, ir, ith], data[it, :4, ir, ith]
            curl[it, :, ir, irad] += -rinv[ir] * cosec[it] * f6(dpi, data[it, :, ir, ith], b1, b2, 4, dtype=3)
    for ip in range(nphi):
        for it in range(nth):
            curl[it, ip, :, ith] += -rinv * f6(dri, rad * data[it, ip, :, iphi], 0., 0., 4)
            curl[it, ip, :, iphi] += rinv * f6(dri, rad * data[it, ip, :, ith], 0., 0., 0)
    return curl


def test_vectorcurl_matches_pencil_loops():
    from cssx import analysis
    print("vectorcurl", flush=True)
    rng = np.random.default_rng(2)
    nth, nphi, nr = 12, 16, 10
    rad = np.linspace(0.72, 0.97, nr)
    theta = np.linspace(np.pi / 4, 3 * np.pi / 4, nth)
    phi = np.linspace(0, np.pi / 2, nphi)
    data = rng.standard_normal((nth, nphi, nr, 3))
    got, *_ = analysis.vectorcurl(0, "", test=[data, rad, theta, phi])
    ref = _curl_loops(data, rad, theta, phi)
    assert np.abs(got - ref).max() < 1e-11 * np.abs(ref).max()


def test_simpson_weights_and_shell_avg():
    from cssx import analysis
    x = np.linspace(0.3, 1.7, 11)            # even slab count: the textbook composite rule
    f = np.exp(x)
    w = analysis.simpson_weights(11, x[1] - x[0])
    assert abs(w @ f - scipy.integrate.simpson(f, x=x)) < 1e-14
    x = np.linspace(0.3, 1.7, 12)            # odd slab count: the end slab rule is exact on quadratics
    q = 2 * x ** 2 - x + 0.5
    w = analysis.simpson_weights(12, x[1] - x[0])
    exact = (2 / 3) * (1.7 ** 3 - 0.3 ** 3) - 0.5 * (1.7 ** 2 - 0.3 ** 2) + 0.5 * 1.4
    assert abs(w @ q - exact) < 1e-13
    theta = np.linspace(np.pi / 4, 3 * np.pi / 4, 33)
    phi = np.linspace(0, np.pi / 2, 17)
    nr = 5
    c = np.arange(1, nr + 1, dtype=float)
    data = np.broadcast_to(c[None, None, :, None], (33, 17, nr, 2)).copy()
    avg = analysis.shell_avg(data, theta, phi)
    # a field constant on each shell averages to itself up to the Simpson error of the area
    assert np.allclose(avg[:, 0], c, rtol=1e-7)
    one = analysis.shell_avg(data[..., 0], theta, phi, numq=1)
    assert one.shape == (nr, 1) and np.allclose(one[:, 0], avg[:, 0])


def test_read_checkpoint_layout(tmp_path):
    from cssx import analysis
    from cssx import checkpoint as ck
    rng = np.random.default_rng(3)
    st = rng.standard_normal((8, 9, 10, 11))   # (nv, nr, nphi, nth)
    ck.write(str(tmp_path) + "/run1/Checkpoints/", 5, st, deltat=1e-3, sim_time=2.0, r1=0.7, r2=0.9, th1=0.5,
             th2=2.5, phi1=0.0, phi2=1.0)
    data, rad, theta, phi, tm, fname, ierr = analysis.read_checkpoint(5, "run1", vel_curl=True,
                                                                      base_dir=str(tmp_path) + "/")
    assert not ierr and data.shape == (11, 10, 9, 3)
    for k, iv in enumerate((1, 2, 3)):
        assert np.array_equal(data[..., k], st[iv].transpose(2, 1, 0))
    assert np.isclose(rad[0], 0.7) and np.isclose(rad[-1], 0.9) and np.allclose(tm, [1e-3, 2.0], rtol=1e-6)
    *_, ierr = analysis.read_checkpoint(6, "run1", base_dir=str(tmp_path) + "/")
    assert ierr


def test_slice_vorticity_matches_script_loops():
    """plot_shell_slice.py:484-510 with the oracle's compact_fd6 restatement in the script's
    loops (theta ibc 1 dtype 0; phi ibc 3 dtype 0 with the unused periodic ghosts)."""
    from cssx import analysis
    print("slice vorticity", flush=True)
    rng = np.random.default_rng(5)
    nth, nphi, nrad, nq = 14, 20, 3, 4
    theta = np.linspace(np.pi / 4, 3 * np.pi / 4, nth)
    phi = np.linspace(0, np.pi / 2, nphi)
    r = np.array([0.75, 0.8, 0.9])
    data = rng.standard_normal((nth, nphi, nrad, nq, 2))
    iut, iup, irec = 1, 3, 1
    got = analysis.slice_vorticity(data, theta, phi, r, iut, iup, irec)
    f6 = lambda *a, **k: oracle.compact_fd6_py(*a, **k)[0]
    wr = np.empty((nth, nphi, nrad))
    sin_t = np.sin(theta)
    up, ut = data[:, :, :, iup, irec], data[:, :, :, iut, irec]
    dpi, dti = 1. / (phi.max() - phi.min()), 1. / (theta.max() - theta.min())
    for ir in range(nrad):
        for ip in range(nphi):
            wr[:, ip, ir] = f6(dti, sin_t * up[:, ip, ir], 0., 0., 1, 0) / r[ir] / sin_t
        for it in range(nth):
            b1, b2 = ut[it, nphi - 4:, ir], ut[it, 0:4, ir]
            wr[it, :, ir] = wr[it, :, ir] - f6(dpi, ut[it, :, ir], b1, b2, 3, 0) / (r[ir] * sin_t[it])
    assert np.abs(got - wr).max() < 1e-11 * np.abs(wr).max()


@pytest.mark.parametrize("method", ["adv", "scale", "turn", "ens"])
def test_rossby_from_checkpoint(tmp_path, method):
    """rossby_from_checkpoint.py:95-190 on a written checkpoint, against the script's arithmetic
    on the same data with the oracle's compact_fd6 restatement and the Simpson shell average."""
    from cssx import analysis
    from cssx import checkpoint as ck
    print(f"rossby {method}", flush=True)
    rng = np.random.default_rng(7)
    nr, nphi, nth = 9, 16, 13
    st = rng.standard_normal((8, nr, nphi, nth)) * 0.1
    st[0] = 1.0 + 0.2 * np.linspace(1, 0, nr)[:, None, None] + 0.01 * rng.standard_normal((nr, nphi, nth))
    st[4] = 0.3 + 0.01 * rng.standard_normal((nr, nphi, nth))
    ck.write(str(tmp_path) + "/run/Checkpoints/", 3, st, deltat=1e-3, sim_time=1.0, r1=0.72, r2=0.97,
             th1=np.pi / 4, th2=3 * np.pi / 4, phi1=0.0, phi2=np.pi / 2)
    base = str(tmp_path) + "/"
    om = 2.6e-6
    got = analysis.rossby_from_chk(3, "run", om, method=method, base_dir=base)
    f6 = lambda *a, **k: oracle.compact_fd6_py(*a, **k)[0]
    if method == "ens":
        curl, rad, theta, phi = analysis.vectorcurl(3, "run", base_dir=base)
        ref_curl = _curl_loops(analysis.read_checkpoint(3, "run", vel_curl=True, base_dir=base)[0], rad, theta, phi)
        assert np.abs(curl - ref_curl).max() < 1e-11 * np.abs(ref_curl).max()
        ens = np.sqrt((ref_curl ** 2).sum(axis=3))
        ref = analysis.shell_avg(ens, theta, phi, numq=1).reshape(nr) / (2 * om)
    else:
        data, rad, theta, phi, *_ = analysis.read_checkpoint(3, "run", vel_curl=(method == "adv"), base_dir=base)
        avg = analysis.shell_avg(data if method == "adv" else data[..., :5], theta, phi)
        if method == "adv":
            ref = np.sqrt(avg[:, 2] ** 2) / (2 * (rad[-1] - rad[0]) * om)
        elif method == "scale":
            H = -avg[:, 0] / f6(1. / (rad[-1] - rad[0]), avg[:, 0], 0., 0., 1, 0)
            ref = np.sqrt(avg[:, 3] ** 2) / (2 * H * om)
        else:
            gamma, Cp = 1.525, 3.5e8
            logp = gamma * np.log(avg[:, 0]) + avg[:, 4] / (Cp / gamma)
            H = -1. / f6(1. / (rad[1] - rad[0]), logp, 0., 0., 1, 0)
            ref = 2 * np.pi / (1.4 * H / np.sqrt(avg[:, 3] ** 2) * om)
    assert got.shape == (nr,)
    assert np.abs(got - ref).max() < 1e-10 * np.abs(ref).max()
