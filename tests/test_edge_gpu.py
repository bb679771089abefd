"""Edge cases of the drop-ins on the HIP engine: no sources, a single time step, a snapshot
interval beyond nt, models of a few cells, single-shot FWI and LS-TRI plans.  Bars: bit-exact
against the oracle where it computes the case, exact zeros where the reference produces zeros."""
import numpy as np
import pytest

import pywave_amd
from oracle import oracle as O
from pywave_amd import FwiProblem, LstriProblem

pytestmark = pytest.mark.gpu


def _ricker(nt, dt, f0, t0, A):
    # the engine's own wavelet (pinned bit for bit against the reference's in the afd tests), so
    # that the comparison below is of the propagation
    from pywave_amd import _native as N
    w = np.zeros(nt, np.float32)
    N.check(N.lib().pw_ricker(nt, dt, f0, t0, A, N.ptr(w)), None, "ricker")
    return w


def test_afd2d_without_sources_is_zero():
    vel = np.full((20, 30), 2000.0, np.float32)
    d, w = pywave_amd.afd2d(vel, nt=50, dt=0.001, ax=[0, 10, 30], az=[0, 10, 20], ns=0, sx=[], sz=[], f=[],
                            t=[], A=[], nbt=5, ct=0.01, jsnap=10, verb=False)
    assert d.shape == (50, 30) and w.shape == (20, 30, 5)
    assert not d.any() and not w.any()


def test_afd3d_without_sources_is_zero():
    vel = np.full((12, 14, 10), 2000.0, np.float32)
    d = pywave_amd.afd3d(vel, nt=20, dt=0.001, ax=[0, 10, 14], ay=[0, 10, 10], az=[0, 10, 12], ns=0, sx=[],
                         sy=[], sz=[], f=[], t=[], A=[], nbt=4, ct=0.01, jsnap=0, verb=False)
    assert d.shape == (20 * 14 * 10,) and not d.any()


@pytest.mark.parametrize("nt,jsnap", [(1, 1), (2, 5), (40, 41), (40, 39)])
def test_afd2d_short_runs_and_sparse_snapshots_vs_oracle(nt, jsnap):
    # nt = 1 (record then one step), jsnap beyond nt (only it = 0 is stored), jsnap just below nt
    rng = np.random.default_rng(nt + jsnap)
    vel = (1800 + 600 * rng.random((16, 18))).astype(np.float32)
    dt, dx, dz, nbt, ct = 0.001, 10.0, 10.0, 4, 0.02
    sx, sz, f0, t0, amp = [9], [8], [25.0], [0.0], [1.0]
    rick = _ricker(nt, dt, f0[0], t0[0], amp[0])[None, :]
    d_ref, snaps, _ = O.afd2d(vel, nt, dt, dx, dz, nbt, ct, sx, sz, rick, jsnap=jsnap)
    d, w = pywave_amd.afd2d(vel, nt=nt, dt=dt, ax=[0, dx, 18], az=[0, dz, 16], ns=1, sx=sx, sz=sz, f=f0, t=t0,
                            A=amp, nbt=nbt, ct=ct, jsnap=jsnap, verb=False)
    assert np.array_equal(d, d_ref.T)
    assert w.shape[2] == len(snaps) == pywave_amd.count_snaps(nt, jsnap)
    assert all(np.array_equal(w[..., k], snaps[k].T) for k in range(w.shape[2]))


def test_afd2d_model_of_a_few_cells_vs_oracle():
    # a 3x3 model: almost the whole padded grid is absorbing boundary and FFT slack
    vel = np.array([[1500, 1600, 1700], [1800, 1900, 2000], [2100, 2200, 2300]], np.float32)
    nt, dt, nbt = 30, 0.001, 3
    rick = _ricker(nt, dt, 30.0, 0.01, 1.0)[None, :]
    d_ref, snaps, _ = O.afd2d(vel, nt, dt, 10.0, 10.0, nbt, 0.05, [1], [1], rick, jsnap=7)
    d, w = pywave_amd.afd2d(vel, nt=nt, dt=dt, ax=[0, 10.0, 3], az=[0, 10.0, 3], ns=1, sx=[1], sz=[1], f=[30.0],
                            t=[0.01], A=[1.0], nbt=nbt, ct=0.05, jsnap=7, verb=False)
    assert np.array_equal(d, d_ref.T)
    assert all(np.array_equal(w[..., k], snaps[k].T) for k in range(w.shape[2]))


def test_single_shot_fwi_plan(golden):
    # one shot of the golden acquisition: forward bit-exact, gradient of the one-shot data
    g = golden("fwi_small")
    par = g["par"].copy()
    par[10] = 1
    d = FwiProblem(par, g["wav"]).forward(g["vtrue"])
    assert np.array_equal(d[:, :, 0], g["data"][:, :, 0])
    ref = O.gradient_av(par, g["v0"], g["wav"], d)[0]
    gr, _, _ = FwiProblem(par, g["wav"], data=d, batch=1).gradient(g["v0"], "reference")
    assert np.array_equal(gr, ref)


def test_single_shot_lstri_plan(golden):
    g = golden("lstri_small")
    par = g["par"].copy()
    par[10] = 1
    p = LstriProblem(par, g["vel"])
    d = p.model(g["src"][..., :1])
    assert np.array_equal(d[:, :, 0], O.timerev(par, g["vel"], g["src"][..., 0], adj=False))


def test_cfl_violation_warns_and_the_watchdog_raises():
    # v dt sqrt(2)/h = 1.2 x the 4th-order scheme's limit: a CFLWarning before the run, and the
    # result watchdog reports the blown-up traces and wavefield instead of returning Inf/NaN
    from pywave_amd._native import NonFiniteError
    from pywave_amd.stability import CFLWarning, cfl_limit
    h, v = 10.0, 3000.0
    dt = 1.2 * cfl_limit("afd") * h / (v * np.sqrt(2))
    vel = np.full((40, 50), v, np.float32)
    with pytest.warns(CFLWarning):
        with pytest.raises(NonFiniteError, match="Inf or NaN"):
            pywave_amd.afd2d(vel, nt=600, dt=dt, ax=[0, h, 50], az=[0, h, 40], ns=1, sx=[25], sz=[20], f=[20],
                             t=[0.01], A=[1], nbt=10, ct=0.01, jsnap=0, verb=False)
    # the same model within the limit runs
    dt_ok = 0.9 * cfl_limit("afd") * h / (v * np.sqrt(2))
    out = pywave_amd.afd2d(vel, nt=600, dt=dt_ok, ax=[0, h, 50], az=[0, h, 40], ns=1, sx=[25], sz=[20], f=[20],
                           t=[0.01], A=[1], nbt=10, ct=0.01, jsnap=0, verb=False)
    assert np.isfinite(out).all()


def test_fwi_watchdog_raises_on_a_blown_up_forward():
    from pywave_amd._native import NonFiniteError
    from pywave_amd.stability import CFLWarning, cfl_limit
    nz, nx, nt, h, v = 30, 40, 400, 0.01, 3.0
    dt = 1.2 * cfl_limit("fwi") * h / (v * np.sqrt(2))
    par = np.array([nz, nx, h, h, 0, 0, nt, dt, 0, 0, 2, 0.1, 0.1, 5, 8, 0.01, 10, 1, 1, 1], np.float32)
    wav = (1e3 * _ricker(nt, dt, 10.0, 0.02, 1.0)).astype(np.float32)
    p = FwiProblem(par, wav, batch=2)
    with pytest.warns(CFLWarning):
        with pytest.raises(NonFiniteError):
            p.forward(np.full((nz, nx), v, np.float32))
    p.close()


def test_stream_bandwidth_probe():
    # the streaming probe that bench.py reads the step kernels' traffic against: a plausible rate
    # for each mix of streams (64 MB each), and argument checks before any device work
    import ctypes
    from pywave_amd import _native as N
    L, ctx = N.lib(), N.context()
    for reads in (1, 2, 3):
        for cached in (0, 1):
            g = ctypes.c_double()
            N.check(L.pw_stream_bandwidth(ctx, 1 << 24, reads, 3, cached, ctypes.byref(g)), ctx, "stream_bandwidth")
            assert 100.0 < g.value < 40000.0, (reads, cached, g.value)
    g = ctypes.c_double()
    for n, reads, reps, cached in ((1 << 24, 4, 3, 0), ((1 << 24) + 2, 1, 3, 0), (1 << 24, 1, 0, 0), (0, 1, 3, 0),
                                   (1 << 24, 1, 3, 2)):
        with pytest.raises(ValueError):
            N.check(L.pw_stream_bandwidth(ctx, n, reads, reps, cached, ctypes.byref(g)), ctx, "stream_bandwidth")
