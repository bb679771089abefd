"""Many afd2d modelling shots of uneven grid size and length advanced together on the GPU.

Each shot is the reference's afd2d modelling problem (pywave/afd2d.py:4-90); Afd2dBatch keeps all
shots' fields resident, advances every running shot in one launch per time step (shots sorted by
length so the running ones form a prefix), and can capture the whole step loop in a hipGraph.
Per-shot results are bit-identical to afd2d.
"""
import ctypes

import numpy as np

from . import _native as N
from .stability import check_cfl

__all__ = ["Afd2dBatch", "afd2d_batch"]


class Afd2dBatch:
    """shots: list of dicts with vel (nz, nx), d=(dz, dx), nt, dt, src (5*ns pack), ns and optional
    wavelet (ns, nt).  nbt/ct shared by all shots."""

    def __init__(self, shots, nbt=30, ct=0.01, device=0):
        self.n = len(shots)
        dims = np.array([[s["vel"].shape[0], s["vel"].shape[1], s["nt"]] for s in shots], np.int32)
        d = np.array([[s["d"][0], s["d"][1], s["dt"]] for s in shots], np.float32)
        self._dims = dims
        self._L = N.lib()
        self._ctx = N.context(device)
        self._p = self._L.pw_batch2d_create(self._ctx, self.n, N.ptr(np.ascontiguousarray(dims)),
                                            N.ptr(np.ascontiguousarray(d)), int(nbt), float(ct))
        if not self._p:
            N.check(-1, self._ctx, "Afd2dBatch")
        for i, s in enumerate(shots):
            check_cfl(s["vel"], s["dt"], s["d"][:2], "afd", "Afd2dBatch shot %d" % i)
            v = np.ascontiguousarray(np.asarray(s["vel"], np.float32).ravel(order="F"))
            sp = N.f32c(s["src"])
            wv = N.f32c(s["wavelet"]) if s.get("wavelet") is not None else None
            N.check(self._L.pw_batch2d_set_shot(self._p, i, N.ptr(v), int(s["ns"]), N.ptr(sp), N.ptr(wv)),
                    self._ctx, "set_shot %d" % i)

    def run(self, use_graph=True):
        ms = ctypes.c_double(0)
        N.check(self._L.pw_batch2d_run(self._p, int(bool(use_graph)), ctypes.byref(ms)), self._ctx, "batch run")
        return ms.value

    def traces(self, i):
        nt, nx = int(self._dims[i, 2]), int(self._dims[i, 1])
        out = np.empty(nt * nx, np.float32)
        N.check(self._L.pw_batch2d_traces(self._p, int(i), N.ptr(out), out.size), self._ctx, "traces")
        return out.reshape(nt, nx, order="F")

    def cell_updates(self):
        return int(self._L.pw_batch2d_cell_updates(self._p))

    def set_width(self, xc=0):
        """Tile width: 0 times widths 2, 4, 8, 16 before the next run and keeps the fastest; else
        that width.  Bit-identical results at every width."""
        N.check(self._L.pw_batch2d_set_width(self._p, int(xc)), self._ctx, "set_width")

    def width(self):
        """(width in use, [(width, ms per launch), ...] of the timing sweep that chose it)."""
        n = ctypes.c_int(0)
        self._L.pw_batch2d_width(self._p, None, 0, ctypes.byref(n))
        buf = np.zeros(2 * max(n.value, 1), np.float32)
        w = self._L.pw_batch2d_width(self._p, N.ptr(buf), n.value, ctypes.byref(n))
        N.check(min(w, 0), self._ctx, "width")
        return int(w), [(int(buf[2 * i]), round(float(buf[2 * i + 1]), 4)) for i in range(n.value)]

    def close(self):
        if getattr(self, "_p", None):
            self._L.pw_batch2d_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def afd2d_batch(shots, nbt=30, ct=0.01, use_graph=True):
    """Model all shots; returns the list of (nt, nx) trace panels (afd2d's dout, jsnap=0)."""
    b = Afd2dBatch(shots, nbt, ct)
    b.run(use_graph)
    out = [b.traces(i) for i in range(b.n)]
    b.close()
    return out
