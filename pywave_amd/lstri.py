"""Passive modelling and least-squares time-reversal imaging (pfwi mode 4's lstric) on the
MI355X engine (pywave_amd/csrc/pw_lstri.hip).

LstriProblem wraps one pw_lstri plan:
  model(src)          timerev_lop modelling (pywave/src/wave_triutil.c:146-176): source field
                      (nz, nx, nt, ns) -> data (nt, nx, ns) recorded at every column one cell
                      below the top of the model; all shots in one launch per step
  image(data)         its adjoint (wave_triutil.c:118-145): data -> image field (nz, nx, nt, ns)
  invert(data, niter) lstri with inversion (pywave/src/wave_fwi.c:121-210, wave_fwigradient.c:454-525):
                      the reference's conjugate-gradient solver (np_solver + np_cgstep,
                      pywave/src/wave_bigsolver.c:1170-1450, wave_cgstep.c:17-67) per shot from a
                      zero field, its step vectors carried from one shot to the next
Geometry comes from the 20-float pfwi pararray (or a par dict): nz, nx, dz, dx, nt, dt, ns, nb,
coef (niter for invert's default).
"""
import ctypes

import numpy as np

from . import _native as N
from .stability import check_cfl
from .fwi import make_pararray

__all__ = ["LstriProblem"]


class LstriProblem:
    def __init__(self, par, vel, device=0):
        self.par = make_pararray(par)
        self.nz, self.nx = int(self.par[0]), int(self.par[1])
        self.nt, self.ns = int(self.par[6]), int(self.par[10])
        self.niter = int(self.par[19])
        self._L = N.lib()
        if self._L.pw_device_count() <= device:
            raise N.PywaveError("pywave_amd: no HIP device %d visible (the engine has no CPU path)" % device)
        self._ctx = self._L.pw_create(device)
        if not self._ctx:
            raise N.PywaveError("pw_create failed")
        self._p = self._L.pw_lstri_create(self._ctx, N.ptr(self.par))
        if not self._p:
            msg = self._L.pw_last_error(self._ctx).decode(errors="replace")
            self._L.pw_destroy(self._ctx)
            self._ctx = None
            raise ValueError("LstriProblem: %s" % msg)
        self.set_model(vel)

    def set_model(self, vel):
        v = np.asarray(vel)
        if v.shape != (self.nz, self.nx):
            raise ValueError("velocity must be (nz, nx) = (%d, %d)" % (self.nz, self.nx))
        check_cfl(v, float(self.par[7]), (float(self.par[2]), float(self.par[3])), "fwi", "LstriProblem")
        v = np.ascontiguousarray(v.astype(np.float32).ravel(order="F"))
        N.check(self._L.pw_lstri_set_model(self._p, N.ptr(v)), self._ctx, "set_model")

    def _field(self, src):
        f = np.ascontiguousarray(np.asarray(src, np.float32).ravel(order="F"))
        if f.size != self.nz * self.nx * self.nt * self.ns:
            raise ValueError("source field must hold nz*nx*nt*ns samples")
        return f

    def _data(self, data):
        d = np.ascontiguousarray(np.asarray(data, np.float32).ravel(order="F"))
        if d.size != self.nt * self.nx * self.ns:
            raise ValueError("data must hold nt*nx*ns samples")
        return d

    def model(self, src):
        """Data (nt, nx, ns) of the source field (nz, nx, nt, ns)."""
        f = self._field(src)
        out = np.empty(self.nt * self.nx * self.ns, np.float32)
        N.check(self._L.pw_lstri_model(self._p, N.ptr(f), N.ptr(out)), self._ctx, "model")
        return out.reshape(self.nt, self.nx, self.ns, order="F")

    def image(self, data):
        """Adjoint of model(): image field (nz, nx, nt, ns) of data (nt, nx, ns)."""
        d = self._data(data)
        out = np.empty(self.nz * self.nx * self.nt * self.ns, np.float32)
        N.check(self._L.pw_lstri_image(self._p, N.ptr(d), N.ptr(out)), self._ctx, "image")
        return out.reshape(self.nz, self.nx, self.nt, self.ns, order="F")

    def invert(self, data, niter=None, return_iters=False):
        """Least-squares source field (nz, nx, nt, ns) after niter CG iterations per shot."""
        niter = self.niter if niter is None else int(niter)
        d = self._data(data)
        out = np.empty(self.nz * self.nx * self.nt * self.ns, np.float32)
        its = np.zeros(self.ns, np.int32)
        N.check(self._L.pw_lstri_invert(self._p, N.ptr(d), niter, N.ptr(out), N.ptr(its)), self._ctx, "invert")
        src = out.reshape(self.nz, self.nx, self.nt, self.ns, order="F")
        return (src, its) if return_iters else src

    def use_graph(self, on=True):
        """Replay the inversion's sweeps from captured hipGraphs (default) or launch step by step."""
        N.check(self._L.pw_lstri_use_graph(self._p, int(bool(on))), self._ctx, "use_graph")

    def steps_per_launch(self, n=0):
        """Time steps per sweep launch: 1, 2, 4, or 0 for the plan's choice (timed on the plan's grid
        before its first sweep).  The results do not depend on it."""
        N.check(self._L.pw_lstri_steps_per_launch(self._p, int(n)), self._ctx, "steps_per_launch")

    def set_tile_width(self, w=0):
        """Single-step tile width: 1, 2, 4, 8 or 16 columns, or 0 for the plan's timed choice."""
        N.check(self._L.pw_lstri_set_tile_width(self._p, int(w)), self._ctx, "set_tile_width")

    def set_four_step_columns(self, c=0):
        """Output columns of the four-step launches' tiles: 4, 8, or 0 for the timed choice."""
        N.check(self._L.pw_lstri_set_four_step_columns(self._p, int(c)), self._ctx, "set_four_step_columns")

    def launch_choice(self):
        """(single-step tile width, steps per launch, four-step tile columns) of the last sweep, and
        the timing sweep's [(kernel, width, ms per launch), ...]."""
        names = ["model_step", "adjoint_step", "two_step_model", "two_step_adjoint", "four_step_model",
                 "four_step_adjoint"]
        w = np.zeros(3, np.int32)
        n = self._L.pw_lstri_launch_choice(self._p, N.ptr(w), None, 0)
        N.check(min(n, 0), self._ctx, "launch_choice")
        buf = np.zeros(3 * max(n, 1), np.float32)
        self._L.pw_lstri_launch_choice(self._p, N.ptr(w), N.ptr(buf), n)
        log = [(names[int(buf[3 * i])], int(buf[3 * i + 1]), round(float(buf[3 * i + 2]), 4)) for i in range(n)]
        return (int(w[0]), int(w[1]), int(w[2])), log

    def timing(self):
        """Device milliseconds of the last model() / invert() call."""
        ms = ctypes.c_double()
        self._L.pw_lstri_timing(self._p, ctypes.byref(ms))
        return ms.value

    def close(self):
        if getattr(self, "_p", None):
            self._L.pw_lstri_destroy(self._p)
            self._p = None
        if getattr(self, "_ctx", None):
            self._L.pw_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
