"""pfwi / fillpar: drop-ins for pywave.pfwi (pywave/pfwi.py:5-301) on the MI355X engine.

mode 1  forward modelling (pfwicfun.forward_modeling_ac, pywave/src/pfwi.c:27-495):
        returns (data (nt, nx, ns), [], [], [], [])
mode 2  active-source FWI (pfwicfun.fwic, pywave/src/pfwi.c:497-1028): L-BFGS over gradient_av;
        returns (data flattened in Fortran order, vinv (nz, nx, niter), None, None, None)
mode 3  the reference's RTM branch is empty and its return then fails (UnboundLocalError)
mode 4  inv and not onlysrc: passive FWI for velocity with the known source field src
        (pfwicfun.pfwic, pywave/src/pfwi.c:1507-2034, through pywave/pfwi.py:168-204): with
        onlyvel returns (data flattened in Fortran order, vinv (nz, nx, niter), None, None, None);
        without onlyvel the reference's return reaches unassigned grad/mwt (pywave/pfwi.py:205-209)
        and raises UnboundLocalError, raised here before the inversion runs
        inv and onlysrc: least-squares time-reversal source inversion (pfwicfun.lstric,
        pywave/src/pfwi.c:1031-1505, through pywave/pfwi.py:122-165): returns
        ([], None, None, src (nz, nx, nt, ns), [])
        not inv: passive modelling of the source field src (lstric, pywave/pfwi.py:212-250):
        returns (data (nt, nx, ns), [], [], [], [])
        The branch tests compare with True as the reference does (par['inv'] == True).
other   prints "Wrong mode" and returns 0
media != 1 in mode 1 calls forward_modelingc (pywave/pfwi.py:72), which the reference's extension
does not define (NameError there); here it is the visco-acoustic forward_modeling
(pywave/src/wave_fwimodeling.c:116-220) with q the quality factor and f0 the reference
frequency, returning data (nt, nx, ns) as media=1 does.  Units follow the reference demos: km and km/s (the optimiser clips to [0, 10]).
Slots of vinv after an early stop (line-search failure or convergence) are zero; the reference
leaves them uninitialised.
"""
import numpy as np

from . import _native as N
from .fwi import make_pararray

__all__ = ["pfwi", "fillpar"]


def _col(a, n, name, at_least=False):
    """a flattened in Fortran order to f32, checked to hold n samples (at least n when at_least):
    the engine entry points read exactly that many (the reference reads them unchecked)."""
    v = np.asarray(a).flatten(order="F").astype(np.float32)
    if (v.size < n) if at_least else (v.size != n):
        raise ValueError("pfwi: %s holds %d samples, %s%d expected" % (name, v.size, "at least " if at_least else "", n))
    return v


def pfwi(vel, q, wav, src, data=None, mode=1, media=1, inv=0, verb=1, nb=60, coef=0.005, acq=1, ss=[0, 0.2, 1],
         f0=10, waterz=-30, niter=30, conv_e=0.01, par=None):
    par = fillpar(par)
    if mode == 1:
        nz, nt, nx, ns = int(par["nz"]), int(par["nt"]), int(par["nx"]), int(par["ns"])
        vel = _col(vel, nz * nx, "vel")
        q = _col(q, nz * nx, "q") if media != 1 else np.asarray(q).flatten(order="F").astype(np.float32)
        wav = _col(wav, nt, "wav", at_least=True)
        pararray = make_pararray(par)
        out = np.empty(nt * nx * ns, np.float32)
        L, ctx = N.lib(), N.context()
        if media == 1:
            N.check(L.pw_forward_modeling_ac(ctx, N.ptr(vel), N.ptr(q), N.ptr(wav), N.ptr(pararray), N.ptr(out),
                                             out.size), ctx, "forward_modeling_ac")
        else:
            N.check(L.pw_forward_modeling_c(ctx, N.ptr(vel), N.ptr(q), N.ptr(wav), N.ptr(pararray), N.ptr(out),
                                            out.size), ctx, "forward_modelingc")
        data = out.reshape(par["nt"], par["nx"], par["ns"], order="F")
        vinv = []; grad = []; mwt = []; src = []
    elif mode == 2:
        nz, nx, niter = int(par["nz"]), int(par["nx"]), int(par["niter"])
        nt, ns = int(par["nt"]), int(par["ns"])
        vel = _col(vel, nz * nx, "vel")
        q = np.asarray(q).flatten(order="F").astype(np.float32)
        wav = _col(wav, nt, "wav", at_least=True)
        data = _col(data, nt * nx * ns, "data")
        pararray = make_pararray(par)
        out = np.empty(nz * nx * niter, np.float32)
        L, ctx = N.lib(), N.context()
        N.check(L.pw_fwic(ctx, N.ptr(vel), N.ptr(q), N.ptr(wav), N.ptr(data), N.ptr(pararray), N.ptr(out),
                          out.size), ctx, "fwic")
        vinv = out.reshape(par["nz"], par["nx"], par["niter"], order="F")
        src = None; mwt = None; grad = None
    elif mode == 3:
        raise UnboundLocalError("local variable 'vinv' referenced before assignment")
    elif mode == 4:
        if par["inv"] != True:  # noqa: E712 (the reference's test)
            return _lstric_modelling(vel, q, wav, src, par)
        if par["onlysrc"] == True:  # noqa: E712
            return _lstric_inversion(vel, q, wav, data, par)
        if par["onlyvel"] != True:  # noqa: E712
            raise UnboundLocalError("local variable 'grad' referenced before assignment")
        nz, nx, niter = int(par["nz"]), int(par["nx"]), int(par["niter"])
        nt, ns = int(par["nt"]), int(par["ns"])
        vel = _col(vel, nz * nx, "vel")
        q = np.asarray(q).flatten(order="F").astype(np.float32)
        wav = _col(wav, nt, "wav", at_least=True)
        par["inv"] = 1
        pararray = make_pararray(par)
        data = _col(data, nt * nx * ns, "data")
        src = _col(src, nz * nx * nt * ns, "src")
        datasrc = np.concatenate([data, src], axis=0)
        out = np.empty(nz * nx * niter, np.float32)
        L, ctx = N.lib(), N.context()
        N.check(L.pw_pfwic(ctx, N.ptr(vel), N.ptr(q), N.ptr(wav), N.ptr(datasrc), N.ptr(pararray), N.ptr(out),
                           out.size, datasrc.size), ctx, "pfwic")
        vinv = out.reshape(par["nz"], par["nx"], par["niter"], order="F")
        src = None; mwt = None; grad = None
    else:
        print("Wrong mode")
        return 0
    return data, vinv, grad, src, mwt


def _lstric(vel, q, wav, datasrc, par, out_len):
    vel = _col(vel, int(par["nz"]) * int(par["nx"]), "vel")
    q = q.flatten(order="F").astype(np.float32)
    wav = wav.flatten(order="F").astype(np.float32)
    datasrc = np.ascontiguousarray(datasrc.flatten(order="F").astype(np.float32))
    pararray = make_pararray(par)
    out = np.empty(out_len, np.float32)
    L, ctx = N.lib(), N.context()
    N.check(L.pw_lstric(ctx, N.ptr(vel), N.ptr(q), N.ptr(wav), N.ptr(datasrc), N.ptr(pararray), N.ptr(out),
                        out.size, datasrc.size), ctx, "lstric")
    return out


def _lstric_modelling(vel, q, wav, src, par):
    nt, nx, ns = int(par["nt"]), int(par["nx"]), int(par["ns"])
    out = _lstric(vel, q, wav, src, par, nt * nx * ns)
    return out.reshape(par["nt"], par["nx"], par["ns"], order="F"), [], [], [], []


def _lstric_inversion(vel, q, wav, data, par):
    par["inv"] = 1
    nz, nx, nt, ns = int(par["nz"]), int(par["nx"]), int(par["nt"]), int(par["ns"])
    out = _lstric(vel, q, wav, data, par, nz * nx * nt * ns)
    src = out.reshape(par["nz"], par["nx"], par["nt"], par["ns"], order="F")
    return [], None, None, src, []


def fillpar(par):
    """Fill missing keys with the reference's defaults (pywave/pfwi.py:262-301)."""
    kw = {
        "f0": 10,
        "nb": 100,
        "coef": 0.003,
        "acqui_type": 1,
        "interval": 1,
        "ns": 1,
        "ds": 0.2,
        "os": 0.0,
        "sz": 5,
        "nx": 100,
        "nz": 100,
        "nr": par["nx"],
        "dr": par["dx"],
        "r0": par["ox"],
        "fhi": 0.5 / par["dt"],
        "flo": 0,
        "onlysrc": True,
        "onlyvel": False,
        "inv": False,
        "niter": 10,
    }
    kw.update(par)
    return kw
