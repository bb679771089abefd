"""Write full-size reference fixtures under tests/golden/ from the reference's own C.

The benchmark workloads are too large to commit whole, so each fixture holds the SHA-256 of the
reference's complete f32 output plus samples and reductions small enough to commit:

  full_plain3d_201     afd3d of the 201^3 model (+30-cell ABC = 261^3), nt = 1500, 15 Hz source
                       at the centre, surface receivers 201 x 201 (pywave/src/wave_fdm.c:28-381)
  full_plain3d_201_f64 the same run through the oracle's FP64 twin of the engine (samples only)
  full_batch2d_96      afd2d of four shots of the 96-shot ragged batch at their full sizes
                       (pywave/src/wave_fdm.c:397-686)
  full_fwi2d_grad      forward_modeling_ac of the 400x800 true model, 32 shots, nt = 3000
                       (pywave/src/wave_fwimodeling.c:21-114), then gradient_av of the smoothed
                       start model at interval 1 (pywave/src/wave_fwigradient.c:93-262)
  full_fwi2d_lbfgs     fwi() with 5 L-BFGS iterations from the start model on the same data
                       (pywave/src/wave_fwi.c:24-119, wave_fwilbfgs.c:108-165)

Inputs come from pywave_amd.workloads (seeded), so a test regenerates them bit for bit.

Run:  make -C oracle ref && python tests/golden/make_fullsize.py [plain3d] [plain3d_f64] [batch2d] [fwi]
(all four by default; the fwi case takes about an hour of one core, most of it the L-BFGS run).
"""
import hashlib
import os
import sys
import threading
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import refrun  # noqa: E402
from pywave_amd import workloads as W  # noqa: E402

BATCH_SHOTS = (0, 31, 64, 95)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, np.float32).tobytes()).hexdigest()


def save(name, **arrays):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrays)
    print("wrote", path, os.path.getsize(path), "bytes", flush=True)


def sumsq(a, axis):
    return np.sum(np.asarray(a, np.float64) ** 2, axis=axis)


def plain3d():
    w = W.plain3d()
    n, nt = w["vel"].shape[0], w["nt"]
    t0 = time.time()
    out, el = refrun.afd3d(w["vel"], nt, w["dt"], *w["d"], w["nbt"], w["ct"], w["src"], w["ns"], jsnap=0)
    print("plain3d: reference afd3d %.1f s (engine %.1f s)" % (time.time() - t0, el), flush=True)
    tr = out.reshape(nt, n, n, order="F")  # [it, ix, iy]
    save("full_plain3d_201", sha256=np.array(sha(out)), size=np.array(out.size),
         rx=np.arange(0, n, 20), traces=tr[:, ::20, ::20].copy(),  # every 20th receiver on both axes
         sumsq_t=sumsq(tr, (1, 2)), sumsq_r=sumsq(tr, 0).astype(np.float32), elapsed=np.array(el))


def plain3d_f64():
    """The same run through the oracle's FP64 twin of the engine (oracle/orc_afd64.c), fed the
    reference's f32 wavelet: the error of f32 traces against exact arithmetic of the scheme."""
    from oracle import oracle
    w = W.plain3d()
    n, nt = w["vel"].shape[0], w["nt"]
    s = w["src"]
    rick = refrun.ricker(nt, w["dt"], s[3], s[4], s[5])[None, :]
    t0 = time.time()
    tr, _ = oracle.afd_f64(w["vel"], nt, w["dt"], w["d"], w["nbt"], w["ct"], s, rick)
    print("plain3d f64 twin: %.1f s" % (time.time() - t0), flush=True)
    save("full_plain3d_201_f64", rx=np.arange(0, n, 20), traces=tr[:, ::20, ::20].copy(), sumsq_t=sumsq(tr, (1, 2)))


def batch2d():
    shots = W.batch2d()
    res = {}
    for i in BATCH_SHOTS:
        s = shots[i]
        nz, nx = s["vel"].shape
        out, el = refrun.afd2d(s["vel"], s["nt"], s["dt"], s["d"][1], s["d"][0], s["nbt"], s["ct"], s["src"],
                               s["ns"], jsnap=0)
        tr = out.reshape(s["nt"], nx, order="F")
        cols = np.linspace(0, nx - 1, 16).astype(int)
        print("batch2d shot %d (%dx%d, nt=%d): %.1f s" % (i, nz, nx, s["nt"], el), flush=True)
        res.update({"sha256_%d" % i: np.array(sha(out)), "cols_%d" % i: cols, "traces_%d" % i: tr[:, cols].copy(),
                    "sumsq_t_%d" % i: sumsq(tr, 1), "sumsq_r_%d" % i: sumsq(tr, 0), "elapsed_%d" % i: np.array(el)})
    save("full_batch2d_96", shots=np.array(BATCH_SHOTS), **res)


def fwi_wavelet(w, nt, dt):
    return (w["amp"] * refrun.ricker(nt, dt, w["f0"], w["t0"], 1.0)).astype(np.float32)


def fwi():
    from pywave_amd.fwi import make_pararray
    w = W.fwi2d()
    par = make_pararray(w["par"])
    nz, nx, nt, ns = int(par[0]), int(par[1]), int(par[6]), int(par[10])
    wav = fwi_wavelet(w, nt, float(par[7]))
    q = np.full((nz, nx), 1e4, np.float32)
    t0 = time.time()
    data, el_mod = refrun.fwimod(par, w["vtrue"], q, wav)
    print("fwi2d: reference forward_modeling_ac %.1f s" % (time.time() - t0), flush=True)
    out = {}

    def grad():
        g, fc, el = refrun.fwigrad(par, w["v0"], q, wav, data)
        out["grad"] = (g[0], fc, el)
        print("fwi2d: reference gradient_av %.1f s" % el, flush=True)

    def lbfgs():
        v, el = refrun.fwi(par, w["v0"], q, wav, data)
        out["lbfgs"] = (v, el)
        print("fwi2d: reference fwi (5 iterations) %.1f s" % el, flush=True)

    th = [threading.Thread(target=grad), threading.Thread(target=lbfgs)]
    for t in th:
        t.start()
    th[0].join()
    g, fc, el_g = out["grad"]
    dsel = data[:, ::100, ::8]  # every 100th receiver of every 8th shot
    save("full_fwi2d_grad", par=par, wav=wav, sha256_data=np.array(sha(data)), data_sel=dsel.copy(),
         data_sumsq_t=sumsq(data, 1), grad=g, sha256_grad=np.array(sha(g)), fcost=fc,
         elapsed_mod=np.array(el_mod), elapsed_grad=np.array(el_g))
    th[1].join()
    v, el_l = out["lbfgs"]
    save("full_fwi2d_lbfgs", par=par, niter=np.array(v.shape[2]),
         sha256_iter=np.array([sha(v[:, :, k]) for k in range(v.shape[2])]),
         vinv_sel=v[::4, ::4, :].copy(), vinv_last=v[:, :, -1].copy(), elapsed=np.array(el_l))


if __name__ == "__main__":
    if not refrun.available():
        sys.exit("oracle/_ref/pwref missing: run `make -C oracle ref` where the reference tree exists")
    which = sys.argv[1:] or ["plain3d", "plain3d_f64", "batch2d", "fwi"]
    for name in which:
        {"plain3d": plain3d, "plain3d_f64": plain3d_f64, "batch2d": batch2d, "fwi": fwi}[name]()
