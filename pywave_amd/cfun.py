"""The extension-module functions that pywave's package namespace re-exports, on the MI355X engine.

pywave's ``__init__`` star-imports its wrappers, which star-import the CPython modules ``afdcfun``
(pywave/src/afd.c:585-607), ``apscfun`` (pywave/src/aps.c:585-586) and ``pfwicfun``
(pywave/src/pfwi.c:2041-2047), so ``pywave.afd2dc`` and its siblings are part of its public surface.
Each function here takes the same positional arguments and returns the same flat float32 array:

  afd2dc(vel, src, tri, nt, nx, nz, ns, verb, jsnap, ifsnaps, abc, nbt, ct, dt, ox, dx, oz, dz)
  afd3dc(vel, src, tri, nt, nx, ny, nz, ns, verb, jsnap, ifsnaps, abc, nbt, ct, dt, ox, dx, oy, dy, oz, dz)
  aps2dc, aps3dc                                    as afd2dc, afd3dc with the spectral Laplacian
  forward_modeling_ac(vel, q, wav, pararray)        -> nt*nx*ns
  forward_modelingc(vel, q, wav, pararray)          -> nt*nx*ns, visco-acoustic (the name
                                                       pywave/pfwi.py:72 calls; the reference's
                                                       extension does not define it)
  fwic(vel, q, wav, data, pararray)                 -> nz*nx*niter
  lstric(vel, q, wav, datasrc, pararray)            -> data or source field (pararray[9])
  pfwic(vel, q, wav, datasrc, pararray)             -> nz*nx*niter

A velocity whose length is not nx*nz[*ny] raises SystemError, as the reference's NULL return does
(pywave/src/afd.c:89-93, 381-385); other length mismatches raise ValueError before the engine
reads the arrays (the reference reads them unchecked).
"""
import numpy as np

from . import _native as N
from .stability import check_cfl
from .afd import count_snaps

__all__ = ["afd2dc", "afd3dc", "aps2dc", "aps3dc", "forward_modeling_ac", "forward_modelingc", "fwic", "lstric",
           "pfwic"]


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).ravel())


def _fd2(cfun, name, vel, src, tri, nt, nx, nz, ns, verb, jsnap, ifsnaps, abc, nbt, ct, dt, ox, dx, oz, dz):
    vel, src = _f32(vel), _f32(src)
    nt, nx, nz, ns, jsnap = int(nt), int(nx), int(nz), int(ns), int(jsnap)
    if vel.size != nx * nz:
        raise SystemError("%s: Dimension mismatch, N_input = %d, N_data = %d" % (name, vel.size, nx * nz))
    check_cfl(vel, dt, (dz, dx), "spectral" if cfun.startswith("pw_aps") else "afd", name)
    need = nt * nx if tri else 5 * ns
    if src.size < need:
        raise ValueError("%s: the second argument holds %d values, %d expected" % (name, src.size, need))
    n_out = nt * nx + (nz * nx * count_snaps(nt, jsnap) if jsnap > 0 else 0)
    out = np.empty(n_out, np.float32)
    L, ctx = N.lib(), N.context()
    N.check(getattr(L, cfun)(ctx, N.ptr(vel), N.ptr(src), int(tri), nt, nx, nz, ns, int(verb), jsnap, int(ifsnaps),
                             int(abc), int(nbt), float(ct), float(dt), float(ox), float(dx), float(oz), float(dz),
                             N.ptr(out), n_out), ctx, name)
    return out


def _fd3(cfun, name, vel, src, tri, nt, nx, ny, nz, ns, verb, jsnap, ifsnaps, abc, nbt, ct, dt, ox, dx, oy, dy,
         oz, dz):
    vel, src = _f32(vel), _f32(src)
    nt, nx, ny, nz, ns, jsnap = int(nt), int(nx), int(ny), int(nz), int(ns), int(jsnap)
    if vel.size != nx * ny * nz:
        raise SystemError("%s: Dimension mismatch, N_input = %d, N_data = %d" % (name, vel.size, nx * ny * nz))
    check_cfl(vel, dt, (dz, dx, dy), "spectral" if cfun.startswith("pw_aps") else "afd", name)
    need = nt * nx * ny if tri else 6 * ns
    if src.size < need:
        raise ValueError("%s: the second argument holds %d values, %d expected" % (name, src.size, need))
    n_out = nt * nx * ny + (nz * nx * ny * count_snaps(nt, jsnap) if jsnap > 0 else 0)
    out = np.empty(n_out, np.float32)
    L, ctx = N.lib(), N.context()
    N.check(getattr(L, cfun)(ctx, N.ptr(vel), N.ptr(src), int(tri), nt, nx, ny, nz, ns, int(verb), jsnap,
                             int(ifsnaps), int(abc), int(nbt), float(ct), float(dt), float(ox), float(dx), float(oy),
                             float(dy), float(oz), float(dz), N.ptr(out), n_out), ctx, name)
    return out


def afd2dc(*args):
    """afdcfun.afd2dc (pywave/src/afd.c:316-578)."""
    return _fd2("pw_afd2d", "afd2dc", *args)


def afd3dc(*args):
    """afdcfun.afd3dc (pywave/src/afd.c:23-314)."""
    return _fd3("pw_afd3d", "afd3dc", *args)


def aps2dc(*args):
    """apscfun.aps2dc (pywave/src/aps.c:315)."""
    return _fd2("pw_aps2d", "aps2dc", *args)


def aps3dc(*args):
    """apscfun.aps3dc (pywave/src/aps.c:22)."""
    return _fd3("pw_aps3d", "aps3dc", *args)


def _par(pararray):
    p = _f32(pararray)
    if p.size < 20:
        raise ValueError("pararray holds %d values, 20 expected" % p.size)
    return p, int(p[0]), int(p[1]), int(p[6]), int(p[10]), int(p[19])


def _exact(a, n, what, at_least=False):
    v = _f32(a)
    if (v.size < n) if at_least else (v.size != n):
        raise ValueError("%s holds %d samples, %s%d expected" % (what, v.size, "at least " if at_least else "", n))
    return v


def _modelling(cfun, name, vel, q, wav, pararray, need_q):
    p, nz, nx, nt, ns, _ = _par(pararray)
    vel = _exact(vel, nz * nx, name + ": vel")
    q = _exact(q, nz * nx, name + ": q") if need_q else _f32(q)
    wav = _exact(wav, nt, name + ": wav", at_least=True)
    out = np.empty(nt * nx * ns, np.float32)
    L, ctx = N.lib(), N.context()
    N.check(getattr(L, cfun)(ctx, N.ptr(vel), N.ptr(q), N.ptr(wav), N.ptr(p), N.ptr(out), out.size), ctx, name)
    return out


def forward_modeling_ac(vel, q, wav, pararray):
    """pfwicfun.forward_modeling_ac (pywave/src/pfwi.c:27-495)."""
    return _modelling("pw_forward_modeling_ac", "forward_modeling_ac", vel, q, wav, pararray, False)


def forward_modelingc(vel, q, wav, pararray):
    """The visco-acoustic forward_modeling (pywave/src/wave_fwimodeling.c:116-220) under the name
    pywave/pfwi.py:72 calls for media != 1; q is the quality factor, f0 = pararray[16]."""
    return _modelling("pw_forward_modeling_c", "forward_modelingc", vel, q, wav, pararray, True)


def fwic(vel, q, wav, data, pararray):
    """pfwicfun.fwic (pywave/src/pfwi.c:497-1028): vinv flattened (nz, nx, niter)."""
    p, nz, nx, nt, ns, niter = _par(pararray)
    vel = _exact(vel, nz * nx, "fwic: vel")
    wav = _exact(wav, nt, "fwic: wav", at_least=True)
    data = _exact(data, nt * nx * ns, "fwic: data")
    q = _f32(q)
    out = np.empty(nz * nx * niter, np.float32)
    L, ctx = N.lib(), N.context()
    N.check(L.pw_fwic(ctx, N.ptr(vel), N.ptr(q), N.ptr(wav), N.ptr(data), N.ptr(p), N.ptr(out), out.size), ctx,
            "fwic")
    return out


def lstric(vel, q, wav, datasrc, pararray):
    """pfwicfun.lstric (pywave/src/pfwi.c:1031-1505): pararray[9] (inv) != 0 takes data and returns
    the source field, else takes the source field and returns data."""
    p, nz, nx, nt, ns, _ = _par(pararray)
    vel = _exact(vel, nz * nx, "lstric: vel")
    ds = _f32(datasrc)
    out = np.empty(nz * nx * nt * ns if p[9] != 0 else nt * nx * ns, np.float32)
    q, wav = _f32(q), _f32(wav)
    L, ctx = N.lib(), N.context()
    N.check(L.pw_lstric(ctx, N.ptr(vel), N.ptr(q), N.ptr(wav), N.ptr(ds), N.ptr(p), N.ptr(out), out.size, ds.size),
            ctx, "lstric")
    return out


def pfwic(vel, q, wav, datasrc, pararray):
    """pfwicfun.pfwic (pywave/src/pfwi.c:1507-2034): passive FWI for velocity with a known source
    field; datasrc = [data | source field] (or the source field alone when pararray[9] is 0)."""
    p, nz, nx, nt, ns, niter = _par(pararray)
    vel = _exact(vel, nz * nx, "pfwic: vel")
    wav = _exact(wav, nt, "pfwic: wav", at_least=True)
    ds = _f32(datasrc)
    q = _f32(q)
    out = np.empty(nz * nx * niter, np.float32)
    L, ctx = N.lib(), N.context()
    N.check(L.pw_pfwic(ctx, N.ptr(vel), N.ptr(q), N.ptr(wav), N.ptr(ds), N.ptr(p), N.ptr(out), out.size, ds.size),
            ctx, "pfwic")
    return out
