"""afd2d / afd3d: drop-ins for pywave.afd2d / pywave.afd3d on the MI355X engine.

Same signatures, argument meaning and return conventions as the reference wrappers
(pywave/afd2d.py:4-90, pywave/afd3d.py:4-94):
  * vel is flattened in Fortran order to float32, the source pack is
    [sx.., sz.., f.., t.., A..] (2D) / [sx.., sy.., sz.., f.., t.., A..] (3D), or, with tri=True,
    dat flattened in Fortran order;
  * with jsnap > 0 the result is (dout (nt,nx[,ny]), wfd (nz,nx[,ny],ntsnap)), both Fortran-order
    reshapes of the engine's flat return; with jsnap == 0 the flat vector itself is returned;
  * abc, ifsnaps, ps and verb are accepted and have no effect (the reference forces abc on and
    ignores the rest);
  * a velocity length that does not match nz*nx[*ny] raises SystemError, as the reference's
    NULL return without an exception does (pywave/src/afd.c:89-93,381-385).
The engine itself runs as one fused HIP kernel per time step (pywave_amd/csrc/pw_afd.hip).
"""
import ctypes

import numpy as np

from . import _native as N
from .stability import check_cfl

__all__ = ["afd2d", "afd3d", "AfdPlan", "count_snaps"]


def count_snaps(nt, jsnap):
    """ntsnap of pywave/afd2d.py:77-81: number of it in [0, nt) with it % jsnap == 0."""
    if not jsnap:
        return 0
    return (int(nt) - 1) // int(jsnap) + 1


def _check_len(vel, n, what):
    if vel.size != n:
        raise SystemError("%s: Dimension mismatch, N_input = %d, N_data = %d" % (what, vel.size, n))


def afd2d(vel, nt=1501, dt=0.001, ax=[0, 20, 81], az=[0, 20, 81], ns=2, sx=[30, 40], sz=[30, 40],
          f=[10, 10], t=[0.2, 0.35], A=[1, 2], nbt=30, ct=0.01, jsnap=4, abc=True, ifsnaps=False,
          ps=True, tri=False, dat=None, verb=True):
    """2D acoustic finite-difference modelling / time-reversal imaging (see module docstring)."""
    return _run2d("pw_afd2d", "afd2d", vel, nt, dt, ax, az, ns, sx, sz, f, t, A, nbt, ct, jsnap, abc, ifsnaps,
                  tri, dat, verb)


def _run2d(cfun, what, vel, nt, dt, ax, az, ns, sx, sz, f, t, A, nbt, ct, jsnap, abc, ifsnaps, tri, dat, verb):
    vel = np.asarray(vel).flatten(order="F").astype(np.float32)
    ox, dx, nx = ax[0], ax[1], int(ax[2])
    oz, dz, nz = az[0], az[1], int(az[2])
    nt, jsnap, nbt = int(nt), int(jsnap), int(nbt)
    _check_len(vel, nx * nz, what)
    check_cfl(vel, dt, (dz, dx), "spectral" if cfun.startswith("pw_aps") else "afd", what)
    if tri:
        source = np.asarray(dat).flatten(order="F").astype(np.float32)
        if source.size != nt * nx:
            raise ValueError("%s: dat must hold nt*nx = %d samples, got %d" % (what, nt * nx, source.size))
        tri = 1
    else:
        source = np.concatenate([np.array(sx), np.array(sz), np.array(f), np.array(t), np.array(A)],
                                axis=0, dtype="float32")
        if source.size < 5 * int(ns):
            raise ValueError("%s: the source pack holds fewer than 5*ns values" % what)
        tri = 0
    n_out = nt * nx + (nz * nx * count_snaps(nt, jsnap) if jsnap > 0 else 0)
    dout = np.empty(n_out, np.float32)
    L = N.lib()
    ctx = N.context()
    rc = getattr(L, cfun)(ctx, N.ptr(vel), N.ptr(source), tri, nt, nx, nz, int(ns), int(bool(verb)), jsnap,
                          int(bool(ifsnaps)), int(bool(abc)), nbt, float(ct), float(dt), float(ox), float(dx),
                          float(oz), float(dz), N.ptr(dout), n_out)
    N.check(rc, ctx, what)
    if jsnap > 0:
        ntsnap = count_snaps(nt, jsnap)
        wfd = dout[nt * nx:].reshape(nz, nx, ntsnap, order="F")
        dout = dout[0:nt * nx].reshape(nt, nx, order="F")
        return dout, wfd
    return dout


def afd3d(vel, nt=1501, dt=0.001, ax=[0, 20, 81], ay=[0, 20, 81], az=[0, 20, 81], ns=3,
          sx=[30, 40, 50], sy=[30, 40, 50], sz=[30, 40, 50], f=[10, 10, 10], t=[0.2, 0.35, 0.5],
          A=[1, 2, 2], nbt=30, ct=0.01, jsnap=4, abc=True, ifsnaps=False, ps=True, tri=False, dat=None,
          verb=True):
    """3D acoustic finite-difference modelling / time-reversal imaging (see module docstring)."""
    return _run3d("pw_afd3d", "afd3d", vel, nt, dt, ax, ay, az, ns, sx, sy, sz, f, t, A, nbt, ct, jsnap, abc,
                  ifsnaps, tri, dat, verb)


def _run3d(cfun, what, vel, nt, dt, ax, ay, az, ns, sx, sy, sz, f, t, A, nbt, ct, jsnap, abc, ifsnaps, tri, dat,
           verb):
    vel = np.asarray(vel).flatten(order="F").astype(np.float32)
    ox, dx, nx = ax[0], ax[1], int(ax[2])
    oy, dy, ny = ay[0], ay[1], int(ay[2])
    oz, dz, nz = az[0], az[1], int(az[2])
    nt, jsnap, nbt = int(nt), int(jsnap), int(nbt)
    _check_len(vel, nx * ny * nz, what)
    check_cfl(vel, dt, (dz, dx, dy), "spectral" if cfun.startswith("pw_aps") else "afd", what)
    if tri:
        source = np.asarray(dat).flatten(order="F").astype(np.float32)
        if source.size != nt * nx * ny:
            raise ValueError("%s: dat must hold nt*nx*ny samples" % what)
        tri = 1
    else:
        source = np.concatenate([np.array(sx), np.array(sy), np.array(sz), np.array(f), np.array(t),
                                 np.array(A)], axis=0, dtype="float32")
        if source.size < 6 * int(ns):
            raise ValueError("%s: the source pack holds fewer than 6*ns values" % what)
        tri = 0
    n_out = nt * nx * ny + (nz * nx * ny * count_snaps(nt, jsnap) if jsnap > 0 else 0)
    dout = np.empty(n_out, np.float32)
    L = N.lib()
    ctx = N.context()
    rc = getattr(L, cfun)(ctx, N.ptr(vel), N.ptr(source), tri, nt, nx, ny, nz, int(ns), int(bool(verb)), jsnap,
                          int(bool(ifsnaps)), int(bool(abc)), nbt, float(ct), float(dt), float(ox), float(dx),
                          float(oy), float(dy), float(oz), float(dz), N.ptr(dout), n_out)
    N.check(rc, ctx, what)
    if jsnap > 0:
        ntsnap = count_snaps(nt, jsnap)
        wfd = dout[nt * nx * ny:].reshape(nz, nx, ny, ntsnap, order="F")
        dout = dout[0:nt * nx * ny].reshape(nt, nx, ny, order="F")
        return dout, wfd
    return dout


class AfdPlan:
    """A modelling / time-reversal problem kept resident in HBM (pw_afd_* entry points).

    vel: (nz, nx) or (nz, nx, ny) velocity; d = (dz, dx[, dy]).  The step loop runs with
    run(); output() returns the reference's flat return vector.  spectral=True uses the
    pseudo-spectral Laplacian of aps2d/aps3d (pywave/src/wave_psp.c) instead of the stencil.
    """

    def __init__(self, vel, d, nt, dt, nbt=30, ct=0.01, jsnap=0, device=0, spectral=False):
        vel = np.asarray(vel, np.float32)
        self.ndim = vel.ndim
        if self.ndim not in (2, 3):
            raise ValueError("AfdPlan: vel must be 2D or 3D")
        self.shape = vel.shape
        nz, nx = vel.shape[:2]
        ny = vel.shape[2] if self.ndim == 3 else 1
        dz, dx = float(d[0]), float(d[1])
        dy = float(d[2]) if self.ndim == 3 else 1.0
        self.nt, self.jsnap = int(nt), int(jsnap)
        self.nz, self.nx, self.ny = nz, nx, ny
        check_cfl(vel, dt, (dz, dx, dy)[:self.ndim], "spectral" if spectral else "afd", "AfdPlan")
        self._L = N.lib()
        self._ctx = N.context(device)
        self._p = self._L.pw_afd_create(self._ctx, self.ndim, nz, nx, ny, dz, dx, dy, self.nt, float(dt),
                                        int(nbt), float(ct), self.jsnap)
        if not self._p:
            N.check(-1, self._ctx, "AfdPlan")
        if spectral:  # the pseudo-spectral Laplacian of aps2d/aps3d
            N.check(self._L.pw_afd_set_spectral(self._p, 1), self._ctx, "set_spectral")
        v = np.ascontiguousarray(vel.ravel(order="F"))
        N.check(self._L.pw_afd_set_velocity(self._p, N.ptr(v)), self._ctx, "set_velocity")

    def set_sources(self, src_pack, ns, wavelet=None):
        sp = N.f32c(src_pack)
        wv = N.f32c(wavelet) if wavelet is not None else None
        N.check(self._L.pw_afd_set_sources(self._p, int(ns), N.ptr(sp), N.ptr(wv)), self._ctx, "set_sources")

    def set_tri_data(self, dat):
        """dat: (nt, nx[, ny]) receiver data."""
        dd = np.ascontiguousarray(np.asarray(dat, np.float32).ravel(order="F"))
        N.check(self._L.pw_afd_set_tri_data(self._p, N.ptr(dd)), self._ctx, "set_tri_data")

    def set_variant(self, variant, ny_chunk=0):
        N.check(self._L.pw_afd_set_variant(self._p, int(variant), int(ny_chunk)), self._ctx, "set_variant")

    def autotune(self, steps=20):
        """Pick the fastest 3D step kernel and y-chunk for this grid (cached per shape)."""
        v, c = ctypes.c_int(0), ctypes.c_int(0)
        N.check(self._L.pw_afd_autotune(self._p, int(steps), ctypes.byref(v), ctypes.byref(c)), self._ctx,
                "autotune")
        return v.value, c.value

    def run(self, tri=False):
        N.check(self._L.pw_afd_run(self._p, int(bool(tri))), self._ctx, "run")

    def run_timed(self, tri=False, per_step=False):
        tot, step = ctypes.c_double(0), ctypes.c_double(0)
        N.check(self._L.pw_afd_run_timed(self._p, int(bool(tri)), int(bool(per_step)), ctypes.byref(tot),
                                         ctypes.byref(step)), self._ctx, "run_timed")
        return tot.value, step.value

    def output_len(self):
        return self._L.pw_afd_output_len(self.nt, self.nz, self.nx, self.ny, self.jsnap)

    def output(self):
        out = np.empty(self.output_len(), np.float32)
        N.check(self._L.pw_afd_output(self._p, N.ptr(out), out.size), self._ctx, "output")
        return out

    def final_field(self):
        out = np.empty(self.nz * self.nx * self.ny, np.float32)
        N.check(self._L.pw_afd_final_field(self._p, N.ptr(out), out.size), self._ctx, "final_field")
        shape = (self.nz, self.nx) if self.ndim == 2 else (self.nz, self.nx, self.ny)
        return out.reshape(shape, order="F")

    def close(self):
        if getattr(self, "_p", None):
            self._L.pw_afd_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
