"""Runner for oracle/_ref/pwref: the reference's own engine C compiled from its sources by
``make -C oracle ref`` (TEST INFRASTRUCTURE ONLY, see oracle/__init__.py).

Each call writes float32 inputs to a temporary directory, runs the harness as a child
process and reads the outputs back.  ``elapsed`` is the engine time the harness measured
around the engine call (file I/O excluded).
"""
import os
import re
import subprocess
import tempfile

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PWREF = os.path.join(_HERE, "_ref", "pwref")


def available():
    return os.path.exists(PWREF)


def _run(args, tmp):
    r = subprocess.run([PWREF] + [str(a) for a in args], capture_output=True, text=True, cwd=tmp)
    if r.returncode != 0:
        raise RuntimeError("pwref %s failed (%d): %s" % (args[0], r.returncode, r.stderr[-2000:]))
    m = re.search(r"PWREF_ELAPSED_S=([0-9.eE+-]+)", r.stderr)
    return float(m.group(1)) if m else float("nan")


def _put(tmp, name, arr):
    p = os.path.join(tmp, name)
    np.ascontiguousarray(arr, np.float32).tofile(p)
    return p


def afd2d(vel, nt, dt, dx, dz, nbt, ct, src_pack=None, ns=1, tri=False, dat=None, jsnap=0, ps=False):
    """Returns (flat output exactly as afd2dc returns it, elapsed seconds).
    vel (nz,nx); src_pack = [sx..,sz..,f..,t..,A..]; dat (nt,nx) for tri.  ps: aps2dc (psm2d)."""
    vel = np.asarray(vel, np.float32)
    nz, nx = vel.shape
    nsnap = sum(1 for it in range(nt) if jsnap and it % jsnap == 0)
    with tempfile.TemporaryDirectory() as tmp:
        fv = _put(tmp, "vel.f32", vel.ravel(order="F"))
        if tri:
            fs = _put(tmp, "src.f32", np.asarray(dat, np.float32).ravel(order="F"))
        else:
            fs = _put(tmp, "src.f32", np.asarray(src_pack, np.float32))
        fo = os.path.join(tmp, "out.f32")
        el = _run(["aps2d" if ps else "afd2d", nz, nx, nt, ns, nbt, repr(float(ct)), repr(float(dt)), repr(float(dx)),
                   repr(float(dz)), jsnap, int(bool(tri)), fv, fs, fo], tmp)
        out = np.fromfile(fo, np.float32)
    assert out.size == nt * nx + (nz * nx * nsnap if jsnap > 0 else 0)
    return out, el


def afd3d(vel, nt, dt, dx, dy, dz, nbt, ct, src_pack=None, ns=1, tri=False, dat=None, jsnap=0, ps=False):
    """Returns (flat output exactly as afd3dc returns it, elapsed seconds).
    vel (nz,nx,ny); src_pack = [sx..,sy..,sz..,f..,t..,A..]; dat (nt,nx,ny) for tri.  ps: aps3dc (psm)."""
    vel = np.asarray(vel, np.float32)
    nz, nx, ny = vel.shape
    with tempfile.TemporaryDirectory() as tmp:
        fv = _put(tmp, "vel.f32", vel.ravel(order="F"))
        if tri:
            fs = _put(tmp, "src.f32", np.asarray(dat, np.float32).ravel(order="F"))
        else:
            fs = _put(tmp, "src.f32", np.asarray(src_pack, np.float32))
        fo = os.path.join(tmp, "out.f32")
        el = _run(["aps3d" if ps else "afd3d", nz, nx, ny, nt, ns, nbt, repr(float(ct)), repr(float(dt)), repr(float(dx)),
                   repr(float(dy)), repr(float(dz)), jsnap, int(bool(tri)), fv, fs, fo], tmp)
        out = np.fromfile(fo, np.float32)
    return out, el


def ricker(nt, dt, f0, t0, A):
    with tempfile.TemporaryDirectory() as tmp:
        fo = os.path.join(tmp, "rick.f32")
        _run(["ricker", nt, repr(float(dt)), repr(float(f0)), repr(float(t0)), repr(float(A)), fo], tmp)
        return np.fromfile(fo, np.float32)


def fwimod(pararray, vel, q, wav, visco=False):
    """forward_modeling_ac (visco: the visco-acoustic forward_modeling): returns
    (data (nt,nx,ns), elapsed)."""
    p = np.asarray(pararray, np.float32)
    nz, nx, nt, ns = int(p[0]), int(p[1]), int(p[6]), int(p[10])
    with tempfile.TemporaryDirectory() as tmp:
        args = ["fwimodv" if visco else "fwimod", _put(tmp, "par.f32", p), _put(tmp, "vel.f32", np.asarray(vel).ravel(order="F")),
                _put(tmp, "q.f32", np.asarray(q).ravel(order="F")), _put(tmp, "wav.f32", wav),
                os.path.join(tmp, "out.f32")]
        el = _run(args, tmp)
        out = np.fromfile(args[-1], np.float32)
    return out.reshape(nt, nx, ns, order="F"), el


def fwigrad(pararray, vel, q, wav, data, vel_second=None, visco=False):
    """gradient_init + gradient_av (visco: gradient_v) (x1 or x2).  Returns (grads [(nz,nx)...],
    fcosts, elapsed)."""
    p = np.asarray(pararray, np.float32)
    nz, nx = int(p[0]), int(p[1])
    with tempfile.TemporaryDirectory() as tmp:
        args = ["fwigradv" if visco else "fwigrad", _put(tmp, "par.f32", p), _put(tmp, "vel.f32", np.asarray(vel).ravel(order="F")),
                _put(tmp, "q.f32", np.asarray(q).ravel(order="F")), _put(tmp, "wav.f32", wav),
                _put(tmp, "data.f32", np.asarray(data).ravel(order="F")),
                os.path.join(tmp, "grad.f32"), os.path.join(tmp, "fc.f32")]
        if vel_second is not None:
            args.append(_put(tmp, "vel2.f32", np.asarray(vel_second).ravel(order="F")))
        el = _run(args, tmp)
        g = np.fromfile(args[6], np.float32)
        fc = np.fromfile(args[7], np.float32)
    k = g.size // (nz * nx)
    return [g[i * nz * nx:(i + 1) * nz * nx].reshape(nz, nx, order="F") for i in range(k)], fc, el


def fwi(pararray, vel, q, wav, data, visco=False):
    """fwic (visco: fwi() with media = 2, i.e. gradient_v): returns (vinv (nz,nx,niter), elapsed)."""
    p = np.asarray(pararray, np.float32)
    nz, nx, niter = int(p[0]), int(p[1]), int(p[19])
    with tempfile.TemporaryDirectory() as tmp:
        args = ["fwiv" if visco else "fwi", _put(tmp, "par.f32", p), _put(tmp, "vel.f32", np.asarray(vel).ravel(order="F")),
                _put(tmp, "q.f32", np.asarray(q).ravel(order="F")), _put(tmp, "wav.f32", wav),
                _put(tmp, "data.f32", np.asarray(data).ravel(order="F")), os.path.join(tmp, "vinv.f32")]
        el = _run(args, tmp)
        v = np.fromfile(args[-1], np.float32)
    return v.reshape(nz, nx, niter, order="F"), el


def pfwigrad(pararray, vel, q, wav, data, src, vel_second=None):
    """Passive path (known source field, velocity only): gradient_pas_init + gradient_pas_av
    (x1 or x2).  src is (nz, nx, nt, ns).  Returns (grads [(nz,nx)...], fcosts, elapsed)."""
    p = np.asarray(pararray, np.float32)
    nz, nx = int(p[0]), int(p[1])
    with tempfile.TemporaryDirectory() as tmp:
        args = ["pfwigrad", _put(tmp, "par.f32", p), _put(tmp, "vel.f32", np.asarray(vel).ravel(order="F")),
                _put(tmp, "q.f32", np.asarray(q).ravel(order="F")), _put(tmp, "wav.f32", wav),
                _put(tmp, "data.f32", np.asarray(data).ravel(order="F")),
                _put(tmp, "src.f32", np.asarray(src).ravel(order="F")),
                os.path.join(tmp, "grad.f32"), os.path.join(tmp, "fc.f32")]
        if vel_second is not None:
            args.append(_put(tmp, "vel2.f32", np.asarray(vel_second).ravel(order="F")))
        el = _run(args, tmp)
        g = np.fromfile(args[7], np.float32)
        fc = np.fromfile(args[8], np.float32)
    k = g.size // (nz * nx)
    return [g[i * nz * nx:(i + 1) * nz * nx].reshape(nz, nx, order="F") for i in range(k)], fc, el


def pfwi(pararray, vel, q, wav, data, src):
    """pfwic (passive FWI, velocity only): returns (vinv (nz,nx,niter), elapsed)."""
    p = np.asarray(pararray, np.float32)
    nz, nx, niter = int(p[0]), int(p[1]), int(p[19])
    with tempfile.TemporaryDirectory() as tmp:
        args = ["pfwi", _put(tmp, "par.f32", p), _put(tmp, "vel.f32", np.asarray(vel).ravel(order="F")),
                _put(tmp, "q.f32", np.asarray(q).ravel(order="F")), _put(tmp, "wav.f32", wav),
                _put(tmp, "data.f32", np.asarray(data).ravel(order="F")),
                _put(tmp, "src.f32", np.asarray(src).ravel(order="F")), os.path.join(tmp, "vinv.f32")]
        el = _run(args, tmp)
        v = np.fromfile(args[-1], np.float32)
    return v.reshape(nz, nx, niter, order="F"), el


def lstri(pararray, vel, q, wav, arr):
    """lstric: pararray[9] (inv) != 0 -> arr = data (nt, nx, ns), returns the inverted source
    field (nz, nx, nt, ns); else arr = source field, returns data (nt, nx, ns).  (out, elapsed)"""
    p = np.asarray(pararray, np.float32)
    nz, nx, nt, ns = int(p[0]), int(p[1]), int(p[6]), int(p[10])
    with tempfile.TemporaryDirectory() as tmp:
        args = ["lstri", _put(tmp, "par.f32", p), _put(tmp, "vel.f32", np.asarray(vel).ravel(order="F")),
                _put(tmp, "q.f32", np.asarray(q).ravel(order="F")), _put(tmp, "wav.f32", wav),
                _put(tmp, "in.f32", np.asarray(arr).ravel(order="F")), os.path.join(tmp, "out.f32")]
        el = _run(args, tmp)
        o = np.fromfile(args[-1], np.float32)
    if p[9]:
        return o.reshape(nz, nx, nt, ns, order="F"), el
    return o.reshape(nt, nx, ns, order="F"), el
