"""cssx: MI355X-native compact-FD6 derivatives, MHD right-hand side and RK4 stages of the
CSS code (orvedahl/CSS-Code), computed by hand-written gfx950 HIP kernels in libcssx.so.

Modules:
  derivatives  drop-in for CSS-Python's derivatives.compact_fd6
  operators    Dbydr/Dbydt/Dbydp and second derivatives on a subdomain (derivatives.f90)
"""
from ._lib import CssxError, LIB_PATH, lib  # noqa: F401

__version__ = "0.1.0"
