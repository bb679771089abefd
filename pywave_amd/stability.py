"""Time-step stability (CFL) of the engine's two discretisations, checked on the host.

The reference checks nothing: a model and time step beyond the limit run to Inf/NaN.  Here a
CFLWarning is issued before such a run, and every plan's result watchdog (PW_ERR_NONFINITE in the
C ABI) raises NonFiniteError when a result holds Inf or NaN.

Leapfrog in time is stable while v^2 dt^2 times the largest eigenvalue magnitude of the discrete
Laplacian stays below 4.  Both Laplacians are sums of symmetric 1D stencils whose symbol
c0 + 2 sum_k ck cos(k theta) peaks in magnitude at the Nyquist wavenumber theta = pi:
  afd/aps engine, 4th order (pywave/src/wave_fdm.c:15): |(-30 - 32 - 2)/12| = 16/3 per 1/h^2,
      so v dt sqrt(sum 1/h^2) <= 2/sqrt(16/3) = 0.8660 (0.612 v dt/h in 2D, 0.5 in 3D);
  FWI engine, 8th order (pywave/src/wave_fwiutil.c:17): |K0 - 2K1 + 2K2 - 2K3 + 2K4| = 6.5016,
      so v dt sqrt(sum 1/h^2) <= 2/sqrt(6.5016) = 0.7844.
The absorbing boundaries multiply by factors <= 1 and do not move the limit.  (The pseudo-spectral
Laplacian of aps2d/aps3d reaches pi^2/h^2 per axis, so its limit is 2/pi = 0.637; it is checked
with that value.)
"""
import math
import warnings

import numpy as np

__all__ = ["CFLWarning", "courant", "cfl_limit", "check_cfl"]

_K = (-205.0 / 72, 8.0 / 5, -1.0 / 5, 8.0 / 315, -1.0 / 560)  # pywave/src/wave_fwiutil.c:17
_C = (-30.0 / 12, 16.0 / 12, -1.0 / 12)  # pywave/src/wave_fdm.c:15


class CFLWarning(RuntimeWarning):
    """The time step exceeds the scheme's stability limit for the model's largest velocity."""


def _nyquist(coefs):
    return abs(coefs[0] + 2 * sum(c * (-1) ** k for k, c in enumerate(coefs[1:], start=1)))


def cfl_limit(scheme):
    """Largest stable v dt sqrt(sum 1/h^2) of 'afd' (4th order), 'fwi' (8th order) or 'spectral'."""
    if scheme == "afd":
        return 2.0 / math.sqrt(_nyquist(_C))
    if scheme == "fwi":
        return 2.0 / math.sqrt(_nyquist(_K))
    if scheme == "spectral":
        return 2.0 / math.pi
    raise ValueError("unknown scheme %r" % scheme)


def courant(vmax, dt, spacings):
    """v dt sqrt(sum_axes 1/h^2)."""
    return float(vmax) * float(dt) * math.sqrt(sum(1.0 / (float(h) * float(h)) for h in spacings))


def check_cfl(vel, dt, spacings, scheme, what):
    """Warn (CFLWarning) when the model's largest velocity and dt exceed the scheme's limit, or when
    the model holds a non-finite value.  Returns the Courant number."""
    v = np.asarray(vel)
    if v.size == 0:
        return 0.0
    vmax = float(np.max(np.abs(v)))
    if not math.isfinite(vmax):
        warnings.warn("%s: the velocity model holds Inf or NaN" % what, CFLWarning, stacklevel=3)
        return vmax
    c = courant(vmax, dt, spacings)
    lim = cfl_limit(scheme)
    if c > lim:
        warnings.warn("%s: v_max dt sqrt(sum 1/h^2) = %.4g exceeds the stability limit %.4g of the %s scheme "
                      "(v_max = %.6g, dt = %.6g, spacings %s): the run will blow up"
                      % (what, c, lim, scheme, vmax, dt, list(spacings)), CFLWarning, stacklevel=3)
    return c
