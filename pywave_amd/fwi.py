"""FWI on the MI355X engine: gradients, L-BFGS and a torch.autograd.Function.

FwiProblem wraps one pw_fwi plan (pywave_amd/csrc/pw_fwi.hip):
  forward(vel)                    forward_modeling_a (pywave/src/wave_fwimodeling.c:21-114)
  gradient(vel, mode="reference") gradient_av (pywave/src/wave_fwigradient.c:93-262): imaging
                                  condition, windowed residual, first-call 0.1/max scaling,
                                  triangle smoothing -- the reference's gradient
  gradient(vel, mode="exact")     exact discrete adjoint of 0.5*||F(v) - d||^2
  lbfgs(vel0, niter)              fwi() (pywave/src/wave_fwi.c:24-119) with f64 dot products
  set_source_field(src)           passive shots (pfwi mode 4): gradient_pas_av
                                  (pywave/src/wave_fwigradient.c:672-870) and pfwi()
                                  (pywave/src/wave_fwi.c:212-322) for a known source field
Shots advance `batch` at a time in one launch per time step; forward wavefields are checkpointed
every `checkpoint` steps (0 = automatic: no recomputation when everything fits in half the free
HBM, else 50) and recomputed segment by segment during the backward sweep.

Models are (nz, nx) arrays in km/s (the reference clips the L-BFGS model to [0, 10]); data are
(nt, nr, ns) as pfwi returns them.
"""
import ctypes

import numpy as np

from . import _native as N
from .stability import check_cfl

__all__ = ["FwiProblem", "FwiForward", "fwi_traces", "make_pararray"]

PAR_KEYS = ["nz", "nx", "dz", "dx", "oz", "ox", "nt", "dt", "ot", "inv", "ns", "ds", "os", "sz", "nb", "coef",
            "f0", "acqui_type", "interval", "niter"]


def make_pararray(par):
    """The 20-float parameter vector of pywave/pfwi.py:86-107."""
    if isinstance(par, dict):
        return np.array([par[k] for k in PAR_KEYS], dtype=np.float32)
    p = np.asarray(par, np.float32).ravel()
    if p.size < 20:
        raise ValueError("pararray needs 20 values")
    return np.ascontiguousarray(p[:20])


class FwiProblem:
    def __init__(self, par, wav, data=None, batch=0, precision="f32", checkpoint=0, device=0, src=None):
        self.par = make_pararray(par)
        self.nz, self.nx = int(self.par[0]), int(self.par[1])
        self.nt, self.ns = int(self.par[6]), int(self.par[10])
        self.nr = self.nx
        self.niter = int(self.par[19])
        self.prec = {"f32": 4, "f64": 8}[precision]
        self._L = N.lib()
        # a context of its own, so its stream can follow a torch stream without affecting others
        if self._L.pw_device_count() <= device:
            raise N.PywaveError("pywave_amd: no HIP device %d visible (the engine has no CPU path)" % device)
        self._ctx = self._L.pw_create(device)
        if not self._ctx:
            raise N.PywaveError("pw_create failed")
        self._p = self._L.pw_fwi_create(self._ctx, N.ptr(self.par), int(batch), self.prec)
        if not self._p:
            msg = self._L.pw_last_error(self._ctx).decode(errors="replace")
            self._L.pw_destroy(self._ctx)
            self._ctx = None
            raise ValueError("FwiProblem: %s" % msg)
        if checkpoint:
            self.set_checkpointing(checkpoint)
        w = N.f32c(np.asarray(wav).ravel(order="F"))
        if w.size < self.nt:
            raise ValueError("wavelet shorter than nt")
        N.check(self._L.pw_fwi_set_wavelet(self._p, N.ptr(w)), self._ctx, "set_wavelet")
        if data is not None:
            self.set_observed(data)
        if src is not None:
            self.set_source_field(src)

    # ------------------------------------------------------------------ set-up
    def set_observed(self, data):
        d = np.ascontiguousarray(np.asarray(data, np.float32).ravel(order="F"))
        if d.size != self.nt * self.nr * self.ns:
            raise ValueError("observed data must hold nt*nx*ns samples")
        N.check(self._L.pw_fwi_set_observed(self._p, N.ptr(d)), self._ctx, "set_observed")

    def set_source_field(self, src):
        """Passive source field (nz, nx, nt, ns) added on the model interior at every step in place
        of the point sources (gradient_pas_init's ww3); receivers of every shot at r0_v[0] + ir*dr_v.
        A torch device tensor is taken as [ns][nt][nx][nz] of the plan's precision, copied on the
        current torch stream.  None returns to the point sources."""
        if src is None:
            N.check(self._L.pw_fwi_set_source_field(self._p, None), self._ctx, "set_source_field")
            return
        n = self.nz * self.nx * self.nt * self.ns
        if hasattr(src, "data_ptr") and getattr(src, "is_cuda", False):
            if src.numel() != n or not src.is_contiguous() or src.element_size() != self.prec:
                raise ValueError("device source field must be contiguous [ns][nt][nx][nz] of the plan precision")
            torch = _torch()
            self._bind_stream(torch.cuda.current_stream(src.device).cuda_stream)
            N.check(self._L.pw_fwi_set_source_field_dev(self._p, ctypes.c_void_p(src.data_ptr())), self._ctx,
                    "set_source_field")
            return
        f = np.ascontiguousarray(np.asarray(src, np.float32).ravel(order="F"))
        if f.size != n:
            raise ValueError("source field must hold nz*nx*nt*ns samples")
        N.check(self._L.pw_fwi_set_source_field(self._p, N.ptr(f)), self._ctx, "set_source_field")

    def set_visco(self, q):
        """Visco-acoustic plan with the quality factor q (nz, nx) and the pararray's f0: forward()
        then follows forward_modeling (pywave/src/wave_fwimodeling.c:116-220), gradient(mode=
        "reference"), misfit() and lbfgs() follow gradient_v (pywave/src/wave_fwigradient.c:264-452),
        i.e. fwi() with media != 1 (pywave/src/wave_fwi.c:40-41).  f32 plans with point sources;
        the exact gradient and the torch path refuse a visco plan.  None returns to acoustic."""
        if q is None:
            N.check(self._L.pw_fwi_set_visco(self._p, None), self._ctx, "set_visco")
            return
        qq = np.asarray(q)
        if qq.shape != (self.nz, self.nx):
            raise ValueError("q must be (nz, nx) = (%d, %d)" % (self.nz, self.nx))
        qq = np.ascontiguousarray(qq.astype(np.float32).ravel(order="F"))
        N.check(self._L.pw_fwi_set_visco(self._p, N.ptr(qq)), self._ctx, "set_visco")

    def set_checkpointing(self, steps):
        N.check(self._L.pw_fwi_set_checkpointing(self._p, int(steps)), self._ctx, "set_checkpointing")

    def info(self):
        v = np.zeros(8, np.int32)
        N.check(self._L.pw_fwi_info(self._p, N.ptr(v)), self._ctx, "info")
        keys = ["padnz", "padnx", "batch", "checkpoint_steps", "segments", "ns", "nt", "nr"]
        return dict(zip(keys, (int(x) for x in v)))

    def widths(self):
        """Tile widths of the step kernels picked by the plan's timing sweep (its first sweep runs
        it): forward, visco forward, exact adjoint, reference adjoint, visco reference adjoint, and
        the forward steps that store their term for the exact adjoint."""
        w = np.zeros(9, np.int32)
        N.check(self._L.pw_fwi_widths(self._p, N.ptr(w)), self._ctx, "widths")
        keys = ["forward", "visco_forward", "exact_adjoint", "reference_adjoint", "visco_reference_adjoint"]
        return dict(zip(keys, (int(x) for x in w[:5])), tuned=bool(w[5]), two_step_forward=int(w[6]),
                    two_step_exact_adjoint=int(w[7]), forward_storing_terms=int(w[8]))

    def tuning_log(self):
        """The width-timing sweep's measurements: [(kernel, width, ms per launch), ...]."""
        names = ["forward", "visco_forward", "exact_adjoint", "reference_adjoint", "visco_reference_adjoint",
                 "two_step_forward", "two_step_exact_adjoint", "forward_storing_terms"]
        n = self._L.pw_fwi_tuning_log(self._p, None, 0)
        buf = np.zeros(3 * max(n, 1), np.float32)
        self._L.pw_fwi_tuning_log(self._p, N.ptr(buf), n)
        return [(names[int(buf[3 * i])], int(buf[3 * i + 1]), round(float(buf[3 * i + 2]), 4)) for i in range(n)]

    def set_widths(self, forward=8, exact_adjoint=8, reference_adjoint=8, visco_forward=8,
                   visco_reference_adjoint=8, timed=False):
        """Force the step kernels' tile widths (4, 8 or 16 columns, one step per launch) instead of
        the plan's timing sweep; timed=True returns to the sweep."""
        if timed:
            N.check(self._L.pw_fwi_set_widths(self._p, None), self._ctx, "set_widths")
            return
        w = np.array([forward, visco_forward, exact_adjoint, reference_adjoint, visco_reference_adjoint], np.int32)
        N.check(self._L.pw_fwi_set_widths(self._p, N.ptr(w)), self._ctx, "set_widths")

    def set_steps_per_launch(self, n):
        """Forward time steps one launch may advance: 2 (default: two-step launches where the plan's
        timing sweep finds them faster), 1 (one kernel per time step, the plain method) or -2 (two
        steps per launch whatever the timing, on acoustic point-source plans)."""
        N.check(self._L.pw_fwi_set_steps_per_launch(self._p, int(n)), self._ctx, "set_steps_per_launch")

    def timing(self):
        f, b = ctypes.c_double(), ctypes.c_double()
        self._L.pw_fwi_timing(self._p, ctypes.byref(f), ctypes.byref(b))
        return f.value, b.value

    def _vel(self, vel):
        v = np.asarray(vel)
        if v.shape != (self.nz, self.nx):
            raise ValueError("velocity must be (nz, nx) = (%d, %d)" % (self.nz, self.nx))
        check_cfl(v, float(self.par[7]), (float(self.par[2]), float(self.par[3])), "fwi", "FwiProblem")
        return np.ascontiguousarray(v.astype(np.float32).ravel(order="F"))

    # ------------------------------------------------------------------ work
    def forward(self, vel):
        """Modelled data (nt, nr, ns)."""
        out = np.empty(self.nt * self.nr * self.ns, np.float32)
        v = self._vel(vel)  # keep the converted array alive across the call
        N.check(self._L.pw_fwi_forward(self._p, N.ptr(v), N.ptr(out)), self._ctx, "forward")
        return out.reshape(self.nt, self.nr, self.ns, order="F")

    def gradient(self, vel, mode="reference", scaling=0.0):
        """(grad (nz, nx), fcost, scaling).  mode 'reference' follows gradient_av; pass the
        returned scaling back in to reproduce the reference's later calls."""
        m = {"reference": 0, "exact": 1}[mode]
        g = np.empty(self.nz * self.nx, np.float32)
        fc = ctypes.c_double(0)
        sc = ctypes.c_float(scaling if scaling else 0.0)
        if self.prec == 8 and m == 1:
            gd = np.empty(self.nz * self.nx, np.float64)
            v = np.ascontiguousarray(np.asarray(vel, np.float64).ravel(order="F"))
            N.check(self._L.pw_fwi_gradient_f64(self._p, N.ptr(v), 1, N.ptr(gd), ctypes.byref(fc)), self._ctx,
                    "gradient")
            return gd.reshape(self.nz, self.nx, order="F"), fc.value, None
        v = self._vel(vel)
        N.check(self._L.pw_fwi_gradient(self._p, N.ptr(v), m, N.ptr(g), ctypes.byref(fc), ctypes.byref(sc)),
                self._ctx, "gradient")
        return g.reshape(self.nz, self.nx, order="F"), fc.value, (sc.value if m == 0 else None)

    def misfit(self, vel):
        """0.5*||d - F(v)||^2 reduced in f64 (f64 model for f64 plans)."""
        v = np.ascontiguousarray(np.asarray(vel, np.float64).ravel(order="F"))
        fc = ctypes.c_double(0)
        N.check(self._L.pw_fwi_misfit_f64(self._p, N.ptr(v), ctypes.byref(fc)), self._ctx, "misfit")
        return fc.value

    def lbfgs(self, vel0, niter=None):
        """fwi(): returns (vinv (nz, nx, niter), number of iterates stored)."""
        niter = self.niter if niter is None else int(niter)
        out = np.zeros(self.nz * self.nx * max(niter, 1), np.float32)
        k = ctypes.c_int(0)
        v = self._vel(vel0)
        N.check(self._L.pw_fwi_lbfgs(self._p, N.ptr(v), niter, N.ptr(out), ctypes.byref(k)), self._ctx, "lbfgs")
        return out[:self.nz * self.nx * niter].reshape(self.nz, self.nx, niter, order="F"), k.value

    # device-pointer entry points (torch)
    def _bind_stream(self, stream_ptr):
        N.check(self._L.pw_set_stream(self._ctx, ctypes.c_void_p(stream_ptr)), self._ctx, "set_stream")

    def forward_dev(self, vel_ptr, traces_ptr, stream_ptr):
        self._bind_stream(stream_ptr)
        N.check(self._L.pw_fwi_forward_dev(self._p, ctypes.c_void_p(vel_ptr), ctypes.c_void_p(traces_ptr)),
                self._ctx, "forward_dev")

    def adjoint_dev(self, vel_ptr, gtr_ptr, gvel_ptr, stream_ptr):
        self._bind_stream(stream_ptr)
        N.check(self._L.pw_fwi_adjoint_dev(self._p, ctypes.c_void_p(vel_ptr), ctypes.c_void_p(gtr_ptr),
                                           ctypes.c_void_p(gvel_ptr)), self._ctx, "adjoint_dev")

    def close(self):
        if getattr(self, "_p", None):
            self._L.pw_fwi_destroy(self._p)
            self._p = None
        if getattr(self, "_ctx", None):
            self._L.pw_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _torch():
    import torch  # noqa: WPS433 (torch is only needed for the autograd wrapper)
    return torch


def _make_function():
    torch = _torch()

    class FwiForward(torch.autograd.Function):
        """traces = F(vel): vel (nz, nx) on the GPU -> traces (ns, nt, nr); the backward pass is the
        hand-written exact adjoint (checkpointed recomputation) of the same discretisation."""

        @staticmethod
        def forward(ctx, vel, problem):
            dtype = torch.float32 if problem.prec == 4 else torch.float64
            if vel.dtype != dtype or not vel.is_cuda:
                raise TypeError("vel must be a %s CUDA tensor for this problem" % dtype)
            vz = vel.detach().t().contiguous()  # z fastest
            tr = torch.empty((problem.ns, problem.nt, problem.nr), dtype=dtype, device=vel.device)
            stream = torch.cuda.current_stream(vel.device).cuda_stream
            problem.forward_dev(vz.data_ptr(), tr.data_ptr(), stream)
            ctx.problem = problem
            ctx.save_for_backward(vz)
            return tr

        @staticmethod
        def backward(ctx, gout):
            (vz,) = ctx.saved_tensors
            problem = ctx.problem
            g = gout.contiguous()
            gv = torch.empty_like(vz)
            stream = torch.cuda.current_stream(vz.device).cuda_stream
            problem.adjoint_dev(vz.data_ptr(), g.data_ptr(), gv.data_ptr(), stream)
            return gv.t(), None

    return FwiForward


_FN = None


def FwiForward():
    """The torch.autograd.Function class (created on first use so torch loads lazily)."""
    global _FN
    if _FN is None:
        _FN = _make_function()
    return _FN


def fwi_traces(vel, problem):
    """Differentiable traces (ns, nt, nr) of a (nz, nx) CUDA velocity tensor."""
    return FwiForward().apply(vel, problem)
