"""CPU-side checks of the C ABI: the library loads, exports every function include/cssx.h
declares with the signature the ctypes layer uses, and its host-only plan builder produces a
segmented solve that equals a dense solve of the documented tridiagonal system.  No device
work happens here."""
import ctypes
import os
import re

import numpy as np
import pytest

import cssx
from cssx import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "cssx.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cssx_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = cssx.lib()
    names = _header_functions()
    assert names, "no declarations parsed"
    for name in names:
        assert hasattr(L, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"


def test_version_and_errors():
    L = cssx.lib()
    assert L.cssx_version() >= 100
    assert L.cssx_strerror(_lib.CSSX_EINVAL).decode().startswith("invalid")


def test_invalid_arguments_rejected_without_device():
    L = cssx.lib()
    buf = (ctypes.c_double * 100)()
    assert L.cssx_plan_tables(4, 1, 0, 0, 1, 0, buf, None) == _lib.CSSX_EINVAL      # too short
    assert L.cssx_plan_tables(40, 3, 0, 0, 1, 0, buf, None) == _lib.CSSX_EINVAL     # bad order
    assert L.cssx_plan_tables(40, 1, 0, 0, 7, 0, buf, None) == _lib.CSSX_EINVAL     # bad code
    assert L.cssx_plan_tables(40, 1, 0, 0, 1, 2, buf, None) == _lib.CSSX_EINVAL     # bad variant
    assert L.cssx_ctx_create(None, None) == _lib.CSSX_EINVAL


def _plan(n, order, lo, hi, bdry, variant=0):
    L = cssx.lib()
    buf = (ctypes.c_double * (5 * n + 960))()
    S = ctypes.c_int()
    assert L.cssx_plan_tables(n, order, lo, hi, bdry, variant, buf, ctypes.byref(S)) == 0
    return np.frombuffer(buf, dtype=np.float64).copy(), S.value


def _seg_start(n, S, k):
    base, rem = divmod(n, S)
    return k * base + min(k, rem)


def _segmented_solve(plan, S, n, d):
    """Numpy restatement of steps 3-5 of k_deriv (cssx_deriv.hip) for one pencil."""
    cf, cl, cu, vL, vR = (plan[i * n:(i + 1) * n] for i in range(5))
    K = 16
    rl, ru, rf, rcl, rcu = (plan[5 * n + i * K:5 * n + (i + 1) * K] for i in range(5))
    xh = np.zeros(n)
    first, last, dI = np.zeros(S), np.zeros(S), np.zeros(S)
    for k in range(S):
        st, en = _seg_start(n, S, k), _seg_start(n, S, k + 1)
        mi = (en - st - 1) if k < S - 1 else (en - st)
        y = 0.0
        for j in range(st, st + mi):
            y = cf[j] * d[j] - cl[j] * y
            xh[j] = y
        x = 0.0
        for j in range(st + mi - 1, st - 1, -1):
            x = xh[j] - cu[j] * x
            xh[j] = x
        first[k], last[k], dI[k] = xh[st], xh[st + mi - 1], d[en - 1]
    z = np.zeros(max(S - 1, 1))
    zp = 0.0
    for q in range(S - 1):
        R = dI[q] - rl[q] * last[q] - ru[q] * first[q + 1]
        zp = rf[q] * R - rcl[q] * zp
        z[q] = zp
    xp = 0.0
    for q in range(S - 2, -1, -1):
        xp = z[q] - rcu[q] * xp
        z[q] = xp
    x = np.zeros(n)
    for k in range(S):
        st, en = _seg_start(n, S, k), _seg_start(n, S, k + 1)
        mi = (en - st - 1) if k < S - 1 else (en - st)
        xl = z[k - 1] if k > 0 else 0.0
        xr = z[k] if k < S - 1 else 0.0
        for j in range(st, st + mi):
            x[j] = xh[j] + vL[j] * xl + vR[j] * xr
        if k < S - 1:
            x[en - 1] = xr
    return x


def _dense_lhs(order, n, lo, hi, bdry):
    a = 1 / 3 if order == 1 else 2 / 11
    A = np.eye(n) + a * (np.eye(n, k=1) + np.eye(n, k=-1))
    if order == 1:
        up1 = 4.0 if bdry in (1, 3) else 0.0
        up2 = 4.0 if bdry in (1, 2) else 0.0
        r = (1 / 8, 3 / 4)
    else:
        up1 = 137 / 13 if bdry in (1, 3) else 707 / 193
        up2 = 137 / 13 if bdry in (1, 2) else 707 / 193
        r = (0.1, -0.35)
    if lo == 0:
        A[0, 1] = up1
        A[1, 0], A[1, 2] = r
    if hi == 0:
        A[-1, -2] = up2
        A[-2, -1], A[-2, -3] = r
    return A


@pytest.mark.parametrize("n", [8, 9, 17, 40, 61, 64, 65, 128, 257, 300, 512, 777, 1024])
@pytest.mark.parametrize("order,lo,hi,bdry", [(1, 0, 0, 1), (1, 0, 0, 4), (1, 1, 1, 0), (1, 0, 1, 2),
                                              (2, 0, 0, 1), (2, 0, 0, 0), (2, 1, 1, 0), (2, 1, 0, 3)])
def test_segmented_solve_equals_dense(n, order, lo, hi, bdry):
    plan, S = _plan(n, order, lo, hi, bdry)
    assert 1 <= S <= 16
    rng = np.random.default_rng(n * 31 + order)
    d = rng.standard_normal(n)
    x = _segmented_solve(plan, S, n, d)
    ref = np.linalg.solve(_dense_lhs(order, n, lo, hi, bdry), d)
    assert np.abs(x - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


def _contig_solve(plan, nl, n, MR, d):
    """Numpy restatement of k_deriv_contig's solve: per-lane segments of MR rows, PCR on the
    interface rows with the precomputed multipliers (cssx_deriv_contig.hip)."""
    K, P = 64, 6
    cf, cl, cu, vL, vR = (plan[i * n:(i + 1) * n] for i in range(5))
    base = 5 * n
    rl, ru = plan[base:base + K], plan[base + K:base + 2 * K]
    k1 = plan[base + 2 * K:base + 2 * K + P * K].reshape(P, K)
    k2 = plan[base + 2 * K + P * K:base + 2 * K + 2 * P * K].reshape(P, K)
    invb = plan[base + 2 * K + 2 * P * K:base + 3 * K + 2 * P * K]
    xh = np.zeros(n)
    first, last, dI = np.zeros(K), np.zeros(K), np.zeros(K)
    for j in range(nl):
        st, en = j * MR, min((j + 1) * MR, n)
        mi = (en - st - 1) if j < nl - 1 else (en - st)
        y = 0.0
        for r in range(st, st + mi):
            y = cf[r] * d[r] - cl[r] * y
            xh[r] = y
        x = 0.0
        for r in range(st + mi - 1, st - 1, -1):
            x = xh[r] - cu[r] * x
            xh[r] = x
        first[j], last[j], dI[j] = xh[st], xh[st + mi - 1], d[en - 1]
    nI = nl - 1
    R = np.zeros(K)
    for j in range(nI):
        R[j] = dI[j] - rl[j] * last[j] - ru[j] * first[j + 1]
    steps = 0
    while (1 << steps) < nI:
        steps += 1
    for s in range(steps):
        dd = 1 << s
        Rm = np.array([R[j - dd] if j - dd >= 0 else 0.0 for j in range(K)])
        Rp = np.array([R[j + dd] if j + dd < K else 0.0 for j in range(K)])
        R = np.where(np.arange(K) < nI, R - k1[s] * Rm - k2[s] * Rp, 0.0)
    xI = R * invb
    x = np.zeros(n)
    for j in range(nl):
        st, en = j * MR, min((j + 1) * MR, n)
        mi = (en - st - 1) if j < nl - 1 else (en - st)
        xl = xI[j - 1] if j > 0 else 0.0
        xr = xI[j] if j < nI else 0.0
        for r in range(st, st + mi):
            x[r] = xh[r] + vL[r] * xl + vR[r] * xr
        if j < nI:
            x[en - 1] = xr
    return x


@pytest.mark.parametrize("n", [8, 9, 13, 40, 64, 65, 128, 256, 257, 300, 512, 777, 1024])
@pytest.mark.parametrize("order,lo,hi,bdry", [(1, 0, 0, 1), (1, 0, 0, 4), (1, 1, 1, 0), (1, 0, 1, 2),
                                              (2, 0, 0, 1), (2, 0, 0, 0), (2, 1, 1, 0), (2, 1, 0, 3)])
def test_contig_pcr_solve_equals_dense(n, order, lo, hi, bdry):
    MR = 4 if n <= 256 else (8 if n <= 512 else 16)
    plan, nl = _plan(n, order, lo, hi, bdry, variant=1)
    assert nl == (n + MR - 1) // MR <= 64
    rng = np.random.default_rng(n * 17 + order)
    d = rng.standard_normal(n)
    x = _contig_solve(plan, nl, n, MR, d)
    ref = np.linalg.solve(_dense_lhs(order, n, lo, hi, bdry), d)
    assert np.abs(x - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


def _c_layout(struct, fields, tmp_path):
    """sizeof and offsetof of a header struct, compiled with the host C compiler."""
    src = tmp_path / "layout.c"
    body = "\n".join(f'    printf("%zu\\n", offsetof({struct}, {f}));' for f in fields)
    src.write_text(f'#include <stddef.h>\n#include <stdio.h>\n#include "{ROOT}/include/cssx.h"\n'
                   f'int main(void) {{\n    printf("%zu\\n", sizeof({struct}));\n{body}\n    return 0;\n}}\n')
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-std=c99", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    return int(out[0]), [int(x) for x in out[1:]]


@pytest.mark.parametrize("struct,cls", [("cssx_grid", "Grid"), ("cssx_patch", "Patch"),
                                        ("cssx_mhd_params", "MhdParams"), ("cssx_mhd_profiles", "MhdProfiles")])
def test_struct_layouts_match_header(struct, cls, tmp_path):
    """The ctypes mirrors have the header's size and field offsets (a field added to one side
    only would shift every later field the library reads)."""
    from cssx import model
    C = getattr(_lib, cls, None) or getattr(model, cls)
    names = [f[0] for f in C._fields_]
    cnames = ["in" if n == "in_" else n for n in names]
    size, offs = _c_layout(struct, cnames, tmp_path)
    assert ctypes.sizeof(C) == size
    assert [getattr(C, n).offset for n in names] == offs
