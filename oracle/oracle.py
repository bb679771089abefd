"""ctypes bindings to the oracle (oracle/liborc.so) plus its numpy pieces.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

What pins each function (tests/test_oracle.py):
  next_fast_size, padded_sizes   brute-force enumeration of 2^a 3^b 5^c; golden afd outputs
  abc_weights                    closed form + w[nb-1] == 1; golden afd outputs (bit-exact)
  ricker                         golden wavelets from the compiled reference (rel <= 1e-6),
                                 zero-phase symmetry, peak at the onset index
  afd2d, afd3d                   golden traces/snapshots from the compiled reference, bit-exact
                                 when fed the reference's wavelet; linearity in amplitude
  afd_f64, lap_f64               the FP64 twin: golden traces of the compiled reference within f32
                                 rounding (tests/test_oracle_f64.py); its Laplacian exact on
                                 quintics and equal to the 4th-order symbol on plane waves; linear
                                 in amplitude to 1e-13 where the f32 scheme is not
  fwi_prepare geometry           closed-form integer formulas (pywave/src/wave_fwiutil.c:45-60)
  fwi_forward, gradient_av, fwi  golden outputs from the compiled reference, bit-exact
  the same with src= (passive)   compiled reference's gradient_pas_av / pfwi, bit-exact; a field
                                 equal to the point-source term reproduces the active golden;
                                 zero misfit of the reference at the true model; linearity
  source_map, gradient_smooth2   numpy convolution with the reflected triangle (independent form)
  timerev, lstri                 golden outputs of the compiled reference's lstric (modelling
                                 bit-exact; CG inversion to 1e-6); the adjoint identity
                                 <A m, d> = <m, A^T d>; linearity; a point source against afd2d-free
                                 closed checks (zero field -> zero data, causality)
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class OrcAcq(ctypes.Structure):
    """Mirror of orc_acq in oracle/orc.h."""
    _fields_ = [
        ("nz", ctypes.c_int), ("nx", ctypes.c_int), ("nt", ctypes.c_int), ("ns", ctypes.c_int),
        ("nb", ctypes.c_int),
        ("dz", ctypes.c_float), ("dx", ctypes.c_float), ("dt", ctypes.c_float),
        ("z0", ctypes.c_float), ("x0", ctypes.c_float), ("t0", ctypes.c_float),
        ("coef", ctypes.c_float), ("acqui_type", ctypes.c_int),
        ("ds", ctypes.c_float), ("s0", ctypes.c_float), ("r0", ctypes.c_float),
        ("dr", ctypes.c_float), ("nr", ctypes.c_int), ("interval", ctypes.c_int),
        ("padnz", ctypes.c_int), ("padnx", ctypes.c_int), ("sz", ctypes.c_int),
        ("rz", ctypes.c_int), ("ds_v", ctypes.c_int), ("s0_v", ctypes.c_int),
        ("dr_v", ctypes.c_int),
        ("r0_v", ctypes.POINTER(ctypes.c_int)), ("r02", ctypes.POINTER(ctypes.c_int)),
        ("nr2", ctypes.POINTER(ctypes.c_int)), ("bc", ctypes.POINTER(ctypes.c_float)),
        ("srcf", ctypes.c_void_p), ("q", ctypes.c_void_p), ("f0", ctypes.c_float),
    ]


def lib():
    """Load oracle/liborc.so (built by ``make -C oracle``)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liborc.so missing: run `make -C oracle`")
        L = ctypes.CDLL(path)
        L.orc_next_fast_size.argtypes = [ctypes.c_int]
        L.orc_next_fast_size.restype = ctypes.c_int
        L.orc_abc_weights.argtypes = [ctypes.c_int, ctypes.c_float, _f32p]
        L.orc_afd2d.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_float,
                                ctypes.c_int, _i32p, _i32p, _f32p, ctypes.c_int, _f32p,
                                ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_afd3d.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                ctypes.c_int, ctypes.c_float, ctypes.c_int, _i32p, _i32p, _i32p,
                                _f32p, ctypes.c_int, _f32p, ctypes.c_int, ctypes.c_void_p,
                                ctypes.c_void_p]
        _f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        L.orc_afd_f64.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int,
                                  ctypes.c_float, ctypes.c_int, _i32p, _i32p, _i32p, _f32p, _f64p, ctypes.c_int,
                                  ctypes.c_void_p]
        L.orc_lap_f64.argtypes = [_f64p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_double, _f64p]
        L.orc_aps2d.argtypes = L.orc_afd2d.argtypes
        L.orc_aps3d.argtypes = L.orc_afd3d.argtypes
        L.orc_cfft.argtypes = [ctypes.c_int, ctypes.c_int, _f32p, _f32p]
        L.orc_rfft.argtypes = [ctypes.c_int, ctypes.c_int, _f32p, _f32p]
        L.orc_fwi_prepare.argtypes = [ctypes.POINTER(OrcAcq)]
        L.orc_fwi_release.argtypes = [ctypes.POINTER(OrcAcq)]
        L.orc_fwi_forward.argtypes = [ctypes.POINTER(OrcAcq), _f32p, _f32p, _f32p]
        L.orc_gradient_av.argtypes = [ctypes.POINTER(OrcAcq), _f32p, _f32p, _f32p, _f32p,
                                      ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
        L.orc_fwi.argtypes = [ctypes.POINTER(OrcAcq), _f32p, _f32p, _f32p, ctypes.c_int, _f32p]
        L.orc_fwi_forward_visco.argtypes = [ctypes.POINTER(OrcAcq), _f32p, _f32p, ctypes.c_float, _f32p, _f32p]
        L.orc_fwi.restype = ctypes.c_int
        L.orc_fcost_f32.argtypes = [_f32p, _f32p, ctypes.c_longlong]
        L.orc_fcost_f32.restype = ctypes.c_float
        L.orc_source_map.argtypes = [ctypes.c_int] * 6 + [_f32p]
        L.orc_gradient_smooth2.argtypes = [ctypes.c_int] * 5 + [ctypes.c_float, _f32p]
        L.orc_timerev.argtypes = [ctypes.c_int] * 4 + [ctypes.c_float] * 4 + [_f32p, ctypes.c_int, ctypes.c_int,
                                                                               _f32p, _f32p]
        L.orc_lstri.argtypes = [ctypes.c_int] * 5 + [ctypes.c_float] * 4 + [_f32p, ctypes.c_int, ctypes.c_int,
                                                                              _f32p, _f32p]
        _LIB = L
    return _LIB


# ------------------------------------------------------------------ geometry

def next_fast_size(n):
    return lib().orc_next_fast_size(int(n))


def padded_sizes(nz, nx, ny=1):
    """(nz2, nx2, ny2) of pywave/src/wave_fft3.c:45-60 (first axis real-to-complex)."""
    return 2 * next_fast_size((nz + 1) // 2), next_fast_size(nx), next_fast_size(ny)


def abc_weights(nb, c):
    w = np.zeros(max(nb, 1), np.float32)
    lib().orc_abc_weights(int(nb), np.float32(c), w)
    return w[:nb]


def count_snaps(nt, jsnap):
    return sum(1 for it in range(nt) if jsnap and it % jsnap == 0)


# ------------------------------------------------------------------ wavelet

def ricker(nt, dt, f0, t0, A):
    """Zero-phase Ricker built like pywave/src/wave_ricker.c:22-59 + wave_freqfilt.c:141-164:
    a spike of height A at int(f32(t0)/f32(dt)), circularly filtered over nfft samples by the
    real spectrum w^2 e^(1-w^2)/nfft (w = iw/(nfft*f0*dt), formed in float32).  The transform
    itself is numpy's double-precision FFT (the reference uses KISS FFT in float32)."""
    nt = int(nt)
    dt32, f32_, t32 = np.float32(dt), np.float32(f0), np.float32(t0)
    k0 = int(t32 / dt32)
    spike = np.zeros(nt, np.float64)
    spike[k0] = np.float32(A)
    freq = np.float32(f32_ * dt32)
    nfft = 2 * next_fast_size((nt + 1) // 2)
    nw = nfft // 2 + 1
    dw = np.float32(1.0 / (nfft * float(freq)))
    w = np.arange(nw, dtype=np.float32) * dw
    w = w * w
    shape = (w * np.exp(np.float32(1) - w) / np.float32(nfft)).astype(np.float64)
    buf = np.zeros(nfft, np.float64)
    buf[:nt] = spike
    spec = np.fft.rfft(buf) * shape
    out = np.fft.irfft(spec, nfft) * nfft
    return out[:nt].astype(np.float32)


def ricker_bank(nt, dt, f, t, A):
    return np.stack([ricker(nt, dt, f[i], t[i], A[i]) for i in range(len(f))]).astype(np.float32)


# ------------------------------------------------------------------ afd engine

def afd2d(vel, nt, dt, dx, dz, nbt, ct, sx, sz, rick, tri=False, dat=None, jsnap=0, ps=False):
    """Oracle afd2d (ps: aps2d).  vel (nz,nx); rick (ns,nt).  Returns (traces[nx,nt], snaps, img) in
    the reference's internal layouts: traces [ix][it], snaps [isnap][ix][iz], img [ix][iz]."""
    vel = np.asfortranarray(vel, dtype=np.float32)
    nz1, nx1 = vel.shape
    velf = np.ascontiguousarray(vel.ravel(order="F"))
    ns = len(sx)
    spx = np.ascontiguousarray(np.asarray(sx, np.float32).astype(np.int32))
    spz = np.ascontiguousarray(np.asarray(sz, np.float32).astype(np.int32))
    rick = np.ascontiguousarray(rick if rick is not None else np.zeros((1, nt)), np.float32)
    d = np.zeros((nx1, nt), np.float32) if not tri else np.ascontiguousarray(dat, np.float32).copy()
    nsnap = count_snaps(nt, jsnap)
    snaps = np.zeros((nsnap, nx1, nz1), np.float32) if jsnap > 0 else None
    img = np.zeros((nx1, nz1), np.float32) if tri else None
    fn = lib().orc_aps2d if ps else lib().orc_afd2d
    fn(velf, nz1, nx1, nt, dt, dx, dz, nbt, ct, ns if not tri else 0,
                    spx if ns else np.zeros(1, np.int32), spz if ns else np.zeros(1, np.int32),
                    rick, int(bool(tri)), d, int(jsnap),
                    snaps.ctypes.data if snaps is not None else None,
                    img.ctypes.data if img is not None else None)
    return d, snaps, img


def afd3d(vel, nt, dt, dx, dy, dz, nbt, ct, sx, sy, sz, rick, tri=False, dat=None, jsnap=0, ps=False):
    """Oracle afd3d (ps: aps3d).  vel (nz,nx,ny).  traces [iy][ix][it], snaps [isnap][iy][ix][iz]."""
    vel = np.asarray(vel, np.float32)
    nz1, nx1, ny1 = vel.shape
    velf = np.ascontiguousarray(vel.ravel(order="F"))
    ns = len(sx)
    spx = np.ascontiguousarray(np.asarray(sx, np.float32).astype(np.int32))
    spy = np.ascontiguousarray(np.asarray(sy, np.float32).astype(np.int32))
    spz = np.ascontiguousarray(np.asarray(sz, np.float32).astype(np.int32))
    rick = np.ascontiguousarray(rick if rick is not None else np.zeros((1, nt)), np.float32)
    d = np.zeros((ny1, nx1, nt), np.float32) if not tri else np.ascontiguousarray(dat, np.float32).copy()
    nsnap = count_snaps(nt, jsnap)
    snaps = np.zeros((nsnap, ny1, nx1, nz1), np.float32) if jsnap > 0 else None
    img = np.zeros((ny1, nx1, nz1), np.float32) if tri else None
    z1 = np.zeros(1, np.int32)
    fn = lib().orc_aps3d if ps else lib().orc_afd3d
    fn(velf, nz1, nx1, ny1, nt, dt, dx, dy, dz, nbt, ct, ns if not tri else 0,
                    spx if ns else z1, spy if ns else z1, spz if ns else z1, rick,
                    int(bool(tri)), d, int(jsnap),
                    snaps.ctypes.data if snaps is not None else None,
                    img.ctypes.data if img is not None else None)
    return d, snaps, img


def afd_f64(vel, nt, dt, d, nbt, ct, src, rick, jsnap=0):
    """FP64 twin of afd2d/afd3d modelling (oracle/orc_afd64.c).  vel (nz,nx) or (nz,nx,ny) f32;
    d = (dz, dx) or (dz, dx, dy); src: the reference's source pack ([sx..,sz..,...] in 2D,
    [sx..,sy..,sz..,...] in 3D, truncated to int as afd2dc/afd3dc do); rick (ns, nt) f32.
    Returns (traces (nt, nx[, ny]) float64 in the reference's F order, snaps (nz, nx[, ny],
    ntsnap) float64 or None)."""
    vel = np.asarray(vel, np.float32)
    is3d = vel.ndim == 3
    nz1, nx1 = vel.shape[:2]
    ny1 = vel.shape[2] if is3d else 1
    velf = np.ascontiguousarray(vel.ravel(order="F"))
    rick = np.ascontiguousarray(np.atleast_2d(rick), np.float32)
    ns = rick.shape[0]
    src = np.asarray(src, np.float32)
    spx = np.ascontiguousarray(src[0:ns].astype(np.int32))
    if is3d:
        spy = np.ascontiguousarray(src[ns:2 * ns].astype(np.int32))
        spz = np.ascontiguousarray(src[2 * ns:3 * ns].astype(np.int32))
        dz, dx, dy = d
    else:
        spy = np.zeros(max(ns, 1), np.int32)
        spz = np.ascontiguousarray(src[ns:2 * ns].astype(np.int32))
        (dz, dx), dy = d, 1.0
    dat = np.zeros(ny1 * nx1 * nt, np.float64)
    nsnap = count_snaps(nt, jsnap)
    snaps = np.zeros(nsnap * nz1 * nx1 * ny1, np.float64) if nsnap else None
    rc = lib().orc_afd_f64(velf, nz1, nx1, ny1, int(is3d), int(nt), dt, dx, dy, dz, int(nbt), ct, ns,
                           spx if ns else np.zeros(1, np.int32), spy, spz if ns else np.zeros(1, np.int32), rick,
                           dat, int(jsnap), snaps.ctypes.data if snaps is not None else None)
    if rc:
        raise MemoryError("orc_afd_f64: allocation failed")
    shp = (nt, nx1, ny1) if is3d else (nt, nx1)
    tr = dat.reshape(shp, order="F")
    if snaps is not None:
        snaps = snaps.reshape(((nz1, nx1, ny1) if is3d else (nz1, nx1)) + (nsnap,), order="F")
    return tr, snaps


def lap_f64(u, d):
    """The FP64 twin's 4th-order Laplacian of a padded field u (nz2, nx2[, ny2]) with spacings
    d = (dz, dx[, dy]); zero on the two-cell rim."""
    u = np.asarray(u, np.float64)
    nz2, nx2 = u.shape[:2]
    ny2 = u.shape[2] if u.ndim == 3 else 1
    dz, dx = d[0], d[1]
    dy = d[2] if len(d) > 2 else 1.0
    out = np.zeros(u.size, np.float64)
    lib().orc_lap_f64(np.ascontiguousarray(u.ravel(order="F")), nz2, nx2, ny2, dz, dx, dy, out)
    return out.reshape(u.shape, order="F")


# ------------------------------------------------------------------ FWI

def make_acq(pararray, src=None, q=None):
    """orc_acq from the 20-float pararray of pywave/pfwi.py:40-61 (fields as pfwi.c:697-716,
    with the hard-coded sz=5, rz=1, nr=nx, dr=dx, r0=ox of pfwi.c:781-806).  src: the passive
    source field (nz, nx, nt, ns) of pfwi mode 4, or None; q: the quality factor (nz, nx) of the
    visco-acoustic gradient (gradient_v, f0 = pararray[16]), or None.  The returned object keeps
    its float32 copies (a._src, a._q) alive while it is used."""
    p = np.asarray(pararray, np.float32)
    a = OrcAcq()
    a._src = None
    a._q = None
    if src is not None:
        a._src = np.ascontiguousarray(np.asarray(src, np.float32).ravel(order="F"))
        a.srcf = a._src.ctypes.data
    if q is not None:
        a._q = np.ascontiguousarray(np.asarray(q, np.float32).ravel(order="F"))
        a.q = a._q.ctypes.data
        a.f0 = p[16]
    a.nz, a.nx = int(p[0]), int(p[1])
    a.dz, a.dx = p[2], p[3]
    a.z0, a.x0 = p[4], p[5]
    a.nt, a.dt, a.t0 = int(p[6]), p[7], p[8]
    a.ns, a.ds, a.s0 = int(p[10]), p[11], p[12]
    a.nb, a.coef = int(p[14]), p[15]
    a.acqui_type, a.interval = int(p[17]), int(p[18])
    a.nr, a.dr, a.r0 = a.nx, a.dx, a.x0
    lib().orc_fwi_prepare(ctypes.byref(a))
    return a


def acq_geometry(a):
    ns = a.ns
    return dict(padnz=a.padnz, padnx=a.padnx, sz=a.sz, rz=a.rz, ds_v=a.ds_v, s0_v=a.s0_v,
                dr_v=a.dr_v, r0_v=[a.r0_v[i] for i in range(ns)], r02=[a.r02[i] for i in range(ns)],
                nr2=[a.nr2[i] for i in range(ns)], bc=np.array([a.bc[i] for i in range(a.nb)], np.float32))


def fwi_forward(pararray, vel, wav, src=None):
    """forward_modeling_a: returns data (nt, nx, ns) like pywave/pfwi.py:70-74.  With src (nz, nx,
    nt, ns) the passive source field replaces the point sources."""
    a = make_acq(pararray, src)
    velf = np.ascontiguousarray(np.asarray(vel, np.float32).ravel(order="F"))
    out = np.zeros(a.ns * a.nr * a.nt, np.float32)
    lib().orc_fwi_forward(ctypes.byref(a), velf, np.ascontiguousarray(wav, np.float32), out)
    res = out.reshape(a.nt, a.nr, a.ns, order="F")
    lib().orc_fwi_release(ctypes.byref(a))
    return res


def fcost_f32(data, pred):
    """gradient_av's misfit as the reference accumulates it (f32 accumulator, f64 terms) over data
    (nt, nr, ns) in its loop order."""
    d = np.ascontiguousarray(np.asarray(data, np.float32).ravel(order="F"))
    p = np.ascontiguousarray(np.asarray(pred, np.float32).ravel(order="F"))
    return float(lib().orc_fcost_f32(d, p, d.size))


def gradient_av(pararray, vel, wav, data, scaling=0.0, src=None, q=None):
    """One gradient_av call (gradient_pas_av with src, gradient_v with the quality factor q).
    data (nt, nx, ns).  Returns (grad (nz,nx), fcost, scaling)."""
    a = make_acq(pararray, src, q)
    velf = np.ascontiguousarray(np.asarray(vel, np.float32).ravel(order="F"))
    dflat = np.ascontiguousarray(np.asarray(data, np.float32).ravel(order="F"))
    g = np.zeros(a.nz * a.nx, np.float32)
    fc = ctypes.c_float(0)
    sc = ctypes.c_float(scaling)
    lib().orc_gradient_av(ctypes.byref(a), velf, np.ascontiguousarray(wav, np.float32), dflat, g,
                          ctypes.byref(fc), ctypes.byref(sc))
    nz, nx = a.nz, a.nx
    lib().orc_fwi_release(ctypes.byref(a))
    return g.reshape(nz, nx, order="F"), fc.value, sc.value


def fwi(pararray, vel, wav, data, src=None, q=None):
    """fwi() (pfwi() with src; media = 2, i.e. gradient_v, with the quality factor q): returns vinv
    (nz, nx, niter) and the number of iterates stored."""
    a = make_acq(pararray, src, q)
    niter = int(np.float32(pararray[19]))
    velf = np.ascontiguousarray(np.asarray(vel, np.float32).ravel(order="F"))
    dflat = np.ascontiguousarray(np.asarray(data, np.float32).ravel(order="F"))
    vinv = np.zeros(a.nz * a.nx * niter, np.float32)
    k = lib().orc_fwi(ctypes.byref(a), velf, np.ascontiguousarray(wav, np.float32), dflat, niter, vinv)
    nz, nx = a.nz, a.nx
    lib().orc_fwi_release(ctypes.byref(a))
    return vinv.reshape(nz, nx, niter, order="F"), k


def source_map(sx, sz, rectx, rectz, padnx, padnz):
    rr = np.zeros(padnx * padnz, np.float32)
    lib().orc_source_map(sx, sz, rectx, rectz, padnx, padnz, rr)
    return rr.reshape(padnz, padnx, order="F")


def gradient_smooth2(grad, rectx, rectz, waterz, scaling):
    nz, nx = grad.shape
    g = np.ascontiguousarray(np.asarray(grad, np.float32).ravel(order="F")).copy()
    lib().orc_gradient_smooth2(rectx, rectz, nx, nz, waterz, np.float32(scaling), g)
    return g.reshape(nz, nx, order="F")


def _tri_args(pararray):
    p = np.asarray(pararray, np.float32)
    return dict(nz=int(p[0]), nx=int(p[1]), nt=int(p[6]), ns=int(p[10]), nb=int(p[14]), dt=float(p[7]),
                dx=float(p[3]), dz=float(p[2]), coef=float(p[15]), niter=int(p[19]), inv=int(p[9] != 0))


def timerev(pararray, vel, arr, adj):
    """timerev_lop for one shot: adj=False maps a source field (nz, nx, nt) to data (nt, nx);
    adj=True maps data (nt, nx) to an image field (nz, nx, nt)."""
    t = _tri_args(pararray)
    v = np.ascontiguousarray(np.asarray(vel, np.float32).ravel(order="F"))
    mod = np.zeros(t["nz"] * t["nx"] * t["nt"], np.float32)
    dat = np.zeros(t["nx"] * t["nt"], np.float32)
    if adj:
        dat[:] = np.asarray(arr, np.float32).ravel(order="F")
    else:
        mod[:] = np.asarray(arr, np.float32).ravel(order="F")
    lib().orc_timerev(t["nz"], t["nx"], t["nt"], t["nb"], t["dt"], t["dx"], t["dz"], t["coef"], v, int(adj), 0,
                       mod, dat)
    if adj:
        return mod.reshape(t["nz"], t["nx"], t["nt"], order="F")
    return dat.reshape(t["nt"], t["nx"], order="F")


def lstri(pararray, vel, arr):
    """lstric: with pararray's inv flag, arr = data (nt, nx, ns) -> source field (nz, nx, nt, ns);
    else arr = source field -> data (nt, nx, ns)."""
    t = _tri_args(pararray)
    v = np.ascontiguousarray(np.asarray(vel, np.float32).ravel(order="F"))
    nm, nd = t["nz"] * t["nx"] * t["nt"] * t["ns"], t["nx"] * t["nt"] * t["ns"]
    src = np.zeros(nm, np.float32)
    dat = np.zeros(nd, np.float32)
    if t["inv"]:
        dat[:] = np.asarray(arr, np.float32).ravel(order="F")
    else:
        src[:] = np.asarray(arr, np.float32).ravel(order="F")
    lib().orc_lstri(t["nz"], t["nx"], t["nt"], t["ns"], t["nb"], t["dt"], t["dx"], t["dz"], t["coef"], v, t["inv"],
                     t["niter"], dat, src)
    if t["inv"]:
        return src.reshape(t["nz"], t["nx"], t["nt"], t["ns"], order="F")
    return dat.reshape(t["nt"], t["nx"], t["ns"], order="F")


def cfft(x, inverse=False):
    """Complex transform (unscaled) of a complex64 vector, the KISS algorithm in f32."""
    x = np.ascontiguousarray(np.asarray(x, np.complex64))
    out = np.zeros_like(x)
    lib().orc_cfft(x.size, int(bool(inverse)), x.view(np.float32), out.view(np.float32))
    return out


def rfft(x):
    """Real transform of an even-length float32 vector: n/2+1 bins (complex64)."""
    x = np.ascontiguousarray(np.asarray(x, np.float32))
    out = np.zeros(x.size // 2 + 1, np.complex64)
    lib().orc_rfft(x.size, 0, x, out.view(np.float32))
    return out


def irfft(X, n):
    """Inverse real transform (unscaled) of n/2+1 bins to n reals."""
    X = np.ascontiguousarray(np.asarray(X, np.complex64))
    out = np.zeros(n, np.float32)
    lib().orc_rfft(n, 1, X.view(np.float32), out)
    return out


def fwi_forward_visco(pararray, vel, q, wav):
    """visco-acoustic forward_modeling: data (nt, nx, ns); q (nz, nx) quality factor, f0 = pararray[16]."""
    a = make_acq(pararray)
    p = np.asarray(pararray, np.float32)
    velf = np.ascontiguousarray(np.asarray(vel, np.float32).ravel(order="F"))
    qf = np.ascontiguousarray(np.asarray(q, np.float32).ravel(order="F"))
    out = np.zeros(a.ns * a.nr * a.nt, np.float32)
    lib().orc_fwi_forward_visco(ctypes.byref(a), velf, qf, float(p[16]), np.ascontiguousarray(wav, np.float32), out)
    res = out.reshape(a.nt, a.nr, a.ns, order="F")
    lib().orc_fwi_release(ctypes.byref(a))
    return res
