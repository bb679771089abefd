"""The CFL limits of the two discretisations (pywave_amd/stability.py), checked on the CPU by
running the oracle just below and just above them, and the host-side warning."""
import warnings

import numpy as np
import pytest

from oracle import oracle as O
from pywave_amd.stability import CFLWarning, cfl_limit, check_cfl, courant


def test_limits_are_the_stencil_symbols_maximum():
    # the largest |symbol| over all wavenumbers, sampled, against the closed forms
    th = np.linspace(0, np.pi, 20001)
    c = (-30 / 12, 16 / 12, -1 / 12)
    k = (-205 / 72, 8 / 5, -1 / 5, 8 / 315, -1 / 560)
    sym4 = np.abs(c[0] + 2 * sum(c[j] * np.cos(j * th) for j in (1, 2))).max()
    sym8 = np.abs(k[0] + 2 * sum(k[j] * np.cos(j * th) for j in range(1, 5))).max()
    assert abs(cfl_limit("afd") - 2 / np.sqrt(sym4)) < 1e-9
    assert abs(cfl_limit("fwi") - 2 / np.sqrt(sym8)) < 1e-9
    assert abs(cfl_limit("afd") / np.sqrt(3) - 0.5) < 1e-12  # the textbook 3D value at equal spacings


@pytest.mark.parametrize("factor, stable", [(0.97, True), (1.03, False)])
def test_afd_scheme_blows_up_just_beyond_the_limit(factor, stable):
    # homogeneous 2D model, no absorbing boundary; dt set to factor x the limit
    n, h, v, nt = 48, 10.0, 2000.0, 400
    dt = factor * cfl_limit("afd") / (v * np.sqrt(2) / h)
    vel = np.full((n, n), v, np.float32)
    rick = np.zeros((1, nt), np.float32)
    rick[0, 5] = 1.0  # an impulse excites every wavenumber
    _, snaps, _ = O.afd2d(vel, nt, dt, h, h, 0, 0.0, [24], [24], rick, jsnap=nt // 4)
    last = snaps[-1].astype(np.float64)
    finite = np.isfinite(last).all()
    if stable:  # the impulse response stays bounded by its early amplitude
        assert finite and 0 < np.abs(last).max() < np.abs(snaps[1]).max() * 10
    else:  # the Nyquist mode grows by about 1.5x per step
        assert not finite or np.abs(last).max() > 1e6


@pytest.mark.parametrize("factor, stable", [(0.97, True), (1.03, False)])
def test_fwi_scheme_blows_up_just_beyond_the_limit(factor, stable):
    nz, nx, nt, h, v = 40, 48, 500, 0.01, 2.0
    dt = factor * cfl_limit("fwi") / (v * np.sqrt(2) / h)
    par = np.array([nz, nx, h, h, 0, 0, nt, dt, 0, 0, 1, 0.1, 0.2, 5, 8, 0.0, 10, 1, 1, 1], np.float32)
    wav = np.zeros(nt, np.float32)
    wav[3] = 1e3
    data = O.fwi_forward(par, np.full((nz, nx), v, np.float32), wav)
    finite = np.isfinite(data).all()
    if stable:
        assert finite and np.abs(data).max() < 1e3
    else:
        assert not finite or np.abs(data).max() > 1e6


def test_check_cfl_warns_beyond_the_limit_only():
    vel = np.full((30, 40), 3000.0, np.float32)
    dt_ok = 0.99 * cfl_limit("afd") / courant(3000.0, 1.0, (10.0, 10.0))
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        check_cfl(vel, dt_ok, (10.0, 10.0), "afd", "test")  # no warning
    with pytest.warns(CFLWarning, match="exceeds the stability limit"):
        check_cfl(vel, 1.02 * dt_ok / 0.99, (10.0, 10.0), "afd", "test")
    vel[3, 4] = np.nan
    with pytest.warns(CFLWarning, match="Inf or NaN"):
        check_cfl(vel, dt_ok, (10.0, 10.0), "afd", "test")


def test_workloads_are_within_their_limits():
    from pywave_amd import workloads as W
    w = W.plain3d(n=41, nt=10)
    assert courant(np.abs(W.plain3d_model(201)).max(), w["dt"], w["d"]) < cfl_limit("afd")
    f = W.fwi2d(nz=40, nx=80, ns=1, nt=10)
    v = W.fwi2d()["vtrue"]
    assert courant(v.max(), 0.001, (0.01, 0.01)) < cfl_limit("fwi")
    assert courant(4500.0, 0.001, (10.0, 10.0)) < cfl_limit("afd")  # batch2d's fastest layer
    assert f["par"]["dt"] == 0.001
