"""aps2d / aps3d: drop-ins for pywave.aps2d / pywave.aps3d (pywave/aps2d.py:4-90,
pywave/aps3d.py:4-94) on the MI355X engine.

The reference's pseudo-spectral entry points (apscfun, pywave/src/aps.c) take the afd arguments
and set-up unchanged and run psm2d/psm (pywave/src/wave_psp.c), whose time loop differs from the
finite-difference one only in its Laplacian: IFFT(-(2 pi |k|)^2 FFT(u)) over the padded grid.
Signatures, argument meaning and returns are afd2d's / afd3d's (pywave_amd/afd.py).  The
transforms run as HIP kernels that reproduce the reference's KISS FFT arithmetic
(pywave_amd/csrc/pw_psp.hip).
"""
from .afd import _run2d, _run3d

__all__ = ["aps2d", "aps3d"]


def aps2d(vel, nt=1501, dt=0.001, ax=[0, 20, 81], az=[0, 20, 81], ns=2, sx=[30, 40], sz=[30, 40],
          f=[10, 10], t=[0.2, 0.35], A=[1, 2], nbt=30, ct=0.01, jsnap=4, abc=True, ifsnaps=False,
          ps=True, tri=False, dat=None, verb=True):
    """2D acoustic pseudo-spectral modelling / time-reversal imaging."""
    return _run2d("pw_aps2d", "aps2d", vel, nt, dt, ax, az, ns, sx, sz, f, t, A, nbt, ct, jsnap, abc, ifsnaps,
                  tri, dat, verb)


def aps3d(vel, nt=1501, dt=0.001, ax=[0, 20, 81], ay=[0, 20, 81], az=[0, 20, 81], ns=3,
          sx=[30, 40, 50], sy=[30, 40, 50], sz=[30, 40, 50], f=[10, 10, 10], t=[0.2, 0.35, 0.5],
          A=[1, 2, 2], nbt=30, ct=0.01, jsnap=4, abc=True, ifsnaps=False, ps=True, tri=False, dat=None,
          verb=True):
    """3D acoustic pseudo-spectral modelling / time-reversal imaging."""
    return _run3d("pw_aps3d", "aps3d", vel, nt, dt, ax, ay, az, ns, sx, sy, sz, f, t, A, nbt, ct, jsnap, abc,
                  ifsnaps, tri, dat, verb)
