"""The FWI gradient check of the benchmark configuration at its full size: the 400x800 model
(+60-cell sponge), 32 shots in one batch, nt = 3000, checkpoints every 50 steps.  The exact-adjoint
gradient (f64 plan) is compared with central finite differences of the f64 misfit at 20 interior
cells drawn with seed 4 at least 10 cells from the edges, h = 1e-3 v; bar: 1e-3 relative.  The
f32 gradient of the same batched, checkpointed path is then compared with the f64 one (rel-L2 1e-3)."""
import numpy as np
import pytest

from pywave_amd import FwiProblem
from pywave_amd import _native as N
from pywave_amd.workloads import fwi2d

pytestmark = pytest.mark.gpu


def _wavelet(nt, dt, f0, t0, amp):
    w = np.zeros(nt, np.float32)
    N.check(N.lib().pw_ricker(nt, dt, f0, t0, 1.0, N.ptr(w)), None, "ricker")
    return (amp * w).astype(np.float32)


@pytest.mark.timeout(600)
def test_full_size_gradient_vs_finite_differences():
    w = fwi2d()
    par = w["par"]
    nt, ns, nz, nx = par["nt"], par["ns"], par["nz"], par["nx"]
    wav = _wavelet(nt, par["dt"], w["f0"], w["t0"], w["amp"])
    gen = FwiProblem(par, wav, batch=ns)
    obs = gen.forward(w["vtrue"])
    gen.close()
    p = FwiProblem(par, wav, data=obs, batch=ns, checkpoint=50, precision="f64")
    v0 = w["v0"].astype(np.float64)
    g, f, _ = p.gradient(v0, "exact")
    assert np.isfinite(g).all() and f > 0
    rng = np.random.default_rng(4)
    cells = [(int(rng.integers(10, nz - 10)), int(rng.integers(10, nx - 10))) for _ in range(20)]
    gmax = np.abs(g).max()
    worst = 0.0
    for iz, ix in cells:
        h = 1e-3 * v0[iz, ix]
        vp = v0.copy(); vp[iz, ix] += h
        vm = v0.copy(); vm[iz, ix] -= h
        fd = (p.misfit(vp) - p.misfit(vm)) / (2 * h)
        scale = max(abs(fd), 1e-3 * gmax)
        assert abs(fd - g[iz, ix]) <= 1e-3 * scale, (iz, ix, fd, g[iz, ix])
        worst = max(worst, abs(fd - g[iz, ix]) / scale)
    print("full-size FD check: 20 cells, worst relative difference %.2e" % worst)
    p.close()
    # the f32 gradient of the benchmark's path (same batching and checkpoints) against the f64 one
    q = FwiProblem(par, wav, data=obs, batch=ns, checkpoint=50)
    g32, _, _ = q.gradient(w["v0"], "exact")
    q.close()
    rel = float(np.linalg.norm(g32.astype(np.float64) - g) / np.linalg.norm(g))
    print("full-size f32 gradient vs f64: rel-L2 %.2e" % rel)
    assert rel <= 1e-3


@pytest.mark.timeout(600)
def test_full_size_lbfgs_five_iterations_lower_the_misfit():
    # fwi() (pywave/src/wave_fwi.c:24-119) on the benchmark configuration: 5 L-BFGS iterations of
    # the reference gradient (gradient_av) with Wolfe line searches; every stored model lowers the
    # misfit and stays inside the reference's clip [0, 10] km/s
    w = fwi2d()
    par = dict(w["par"])
    nt, ns = par["nt"], par["ns"]
    wav = _wavelet(nt, par["dt"], w["f0"], w["t0"], w["amp"])
    gen = FwiProblem(par, wav, batch=ns)
    obs = gen.forward(w["vtrue"])
    gen.close()
    p = FwiProblem(par, wav, data=obs, batch=ns, checkpoint=50)
    vinv, k = p.lbfgs(w["v0"], 5)
    assert k == 5 and np.isfinite(vinv).all()
    assert vinv.min() >= 0.0 and vinv.max() <= 10.0
    f = [p.misfit(w["v0"])] + [p.misfit(vinv[:, :, i]) for i in range(5)]
    p.close()
    print("full-size L-BFGS misfits:", ["%.6g" % x for x in f])
    assert all(f[i + 1] < f[i] for i in range(5))
