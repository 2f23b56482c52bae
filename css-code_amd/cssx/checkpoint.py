"""Checkpoint reader/writer in the reference's on-disk format.

Checkpoint_Output / Checkpoint_Input (checkpoint.f90:146-423) and Make_Checkpoint_Filename
(checkpoint.f90:38-75) store one step as a directory `<prefix><istep:07d>/` holding

  header             two formatted lines: (5i9) nr, nth, nphi, nv, istep  [+ nsqualls and/or
                     iashtime, (6i9)/(7i9), when entropy rain / ASH inflow are on]
                     and (8e14.7) deltat, sim_time, r1, r2, th1, th2, phi1, phi2
  checkpoint.<var>   one raw native REAL*8 array per variable, the global (nth, nphi, nr) field
                     in Fortran order (MPI_FILE_WRITE_AT_ALL with a subarray view, :336-363),
                     var = rho, w, v, u, s, bt, bp, br

The magnetic files are named bt/bp/br but hold the vector potential A (the state's slots 6-8).
CSS-Python/read_checkpoint.py reads the same files with numpy.fromfile + reshape(order='F').
A field f(nth, nphi, nr) in Fortran order is, byte for byte, the C-order (nr, nphi, nth) array
this package uses, so the data files are written and read without any transpose.

Host-side I/O; `to_device` / `from_device` move a state to and from a CUDA tensor.
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass, field

import numpy as np

NAMES = ("rho", "w", "v", "u", "s", "bt", "bp", "br")


def fortran_e(x: float, w: int = 14, d: int = 7) -> str:
    """Fortran Ew.d edit descriptor (mantissa 0.ddddddd, exponent E+xx; three-digit exponents
    drop the E as the standard prescribes), right-justified in w characters."""
    x = float(x)
    if not np.isfinite(x):
        s = "NaN" if np.isnan(x) else ("-Infinity" if x < 0 else "Infinity")
        return s.rjust(w)
    sign = "-" if (x < 0 or (x == 0 and np.signbit(x))) else ""
    a = abs(x)
    if a == 0.0:
        mant, e = "0" * d, 0
    else:
        s = f"{a:.{d - 1}e}"              # d significant digits: D.DDDDDDe+XX
        digits, ex = s.split("e")
        mant = digits.replace(".", "")
        e = int(ex) + 1                    # 0.DDDDDDD x 10^(e)
    if abs(e) <= 99:
        es = f"E{'+' if e >= 0 else '-'}{abs(e):02d}"
    else:
        es = f"{'+' if e >= 0 else '-'}{abs(e):03d}"
    body = f"{sign}0.{mant}{es}"
    if len(body) > w:                      # no room for the leading zero
        body = f"{sign}.{mant}{es}"
    if len(body) > w:
        return "*" * w
    return body.rjust(w)


def fortran_i(v: int, w: int = 9) -> str:
    s = str(int(v))
    return s.rjust(w) if len(s) <= w else "*" * w


_NUM = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[EeDd][-+]?\d+|[-+]\d{3})?")


def _parse_reals(line: str) -> list[float]:
    """Reals of one formatted line; fields may touch (a negative value fills its 14 columns)."""
    out = []
    for tok in _NUM.findall(line):
        t = tok.replace("D", "E").replace("d", "e")
        if re.search(r"\d[-+]\d{3}$", t) and "E" not in t.upper():
            t = t[:-4] + "E" + t[-4:]      # 0.1234567+100
        out.append(float(t))
    return out


@dataclass
class Header:
    nr: int
    nth: int
    nphi: int
    nv: int
    istep: int
    deltat: float
    sim_time: float
    r1: float
    r2: float
    th1: float
    th2: float
    phi1: float
    phi2: float
    extra: list = field(default_factory=list)   # nsqualls and/or iashtime when present

    def lines(self) -> str:
        ints = [self.nr, self.nth, self.nphi, self.nv, self.istep] + list(self.extra)
        reals = [self.deltat, self.sim_time, self.r1, self.r2, self.th1, self.th2, self.phi1, self.phi2]
        return "".join(fortran_i(v) for v in ints) + "\n" + "".join(fortran_e(v) for v in reals) + "\n"


def checkpoint_dir(prefix: str, istep: int) -> str:
    """Make_Checkpoint_Filename: <prefix><istep as i7.7>/"""
    return f"{prefix}{int(istep):07d}/"


def write(prefix: str, istep: int, state, *, deltat: float, sim_time: float, r1: float, r2: float, th1: float,
          th2: float, phi1: float, phi2: float, extra=()) -> str:
    """Checkpoint_Output: writes state (nv, nr, nphi, nth) FP64 to <prefix><istep:07d>/ and returns
    that directory."""
    st = np.asarray(_host(state), dtype=np.float64)
    if st.ndim != 4 or st.shape[0] not in (5, 8):
        raise ValueError("state must be (nv, nr, nphi, nth) with nv 5 or 8")
    nv, nr, nphi, nth = st.shape
    d = checkpoint_dir(prefix, istep)
    os.makedirs(d, exist_ok=True)
    hdr = Header(nr, nth, nphi, nv, int(istep), deltat, sim_time, r1, r2, th1, th2, phi1, phi2, list(extra))
    with open(os.path.join(d, "header"), "w") as f:
        f.write(hdr.lines())
    for iv in range(nv):
        np.ascontiguousarray(st[iv]).astype("=f8", copy=False).tofile(os.path.join(d, "checkpoint." + NAMES[iv]))
    return d


def read_header(path: str) -> Header:
    lines = [ln for ln in open(path).read().splitlines() if ln.strip() and not ln.lstrip().startswith("#")]
    if len(lines) < 2:
        raise ValueError(f"{path}: header needs two lines")
    ints = [int(t) for t in lines[0].split()]
    reals = _parse_reals(lines[1])
    if len(ints) < 5 or len(ints) > 7 or len(reals) != 8:
        raise ValueError(f"{path}: expected 5-7 integers and 8 reals, got {len(ints)} and {len(reals)}")
    return Header(*ints[:5], *reals, extra=ints[5:])


def read(directory: str, nv: int | None = None):
    """Checkpoint_Input: returns (Header, state (nv, nr, nphi, nth) FP64).  nv limits the fields
    read to min(nv, the header's nv), as the reference does (checkpoint.f90:219)."""
    hdr = read_header(os.path.join(directory, "header"))
    nvr = hdr.nv if nv is None else min(nv, hdr.nv)
    shape = (hdr.nr, hdr.nphi, hdr.nth)
    st = np.empty((nvr,) + shape)
    for iv in range(nvr):
        fn = os.path.join(directory, "checkpoint." + NAMES[iv])
        a = np.fromfile(fn, dtype="=f8")
        if a.size != np.prod(shape):
            raise ValueError(f"{fn}: {a.size} values, header says {np.prod(shape)}")
        st[iv] = a.reshape(shape)
    return hdr, st


def _host(x):
    try:
        import torch
        if isinstance(x, torch.Tensor):
            return x.detach().to("cpu", torch.float64).numpy()
    except ImportError:
        pass
    return x


def to_device(state, device="cuda"):
    """FP64 state -> contiguous CUDA tensor (the layout cssx.mhd.Solver takes)."""
    import torch
    return torch.as_tensor(np.ascontiguousarray(state), dtype=torch.float64, device=device).contiguous()


def from_device(t):
    return _host(t)
