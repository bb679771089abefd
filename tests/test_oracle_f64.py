"""Pins the oracle's FP64 twin of the afd engine (oracle/orc_afd64.c) on the CPU: against the
compiled reference's golden outputs (which differ from it only by the reference's own f32
rounding), against what the mathematics fixes for its Laplacian (exactness on polynomials of
degree <= 5, the 4th-order symbol on plane waves), and by linearity to f64 rounding."""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle as O


def _golden_run(g):
    vel = g["vel"]
    d = (float(g["dz"]), float(g["dx"]), float(g["dy"])) if vel.ndim == 3 else (float(g["dz"]), float(g["dx"]))
    return O.afd_f64(vel, int(g["nt"]), float(g["dt"]), d, int(g["nbt"]), float(g["ct"]), g["src"], g["rick"],
                     jsnap=int(g["jsnap"]))


@pytest.mark.parametrize("name", ["afd2d_cfg0", "afd2d_two_sources", "afd2d_noslack", "afd3d_two_sources"])
def test_f64_twin_vs_reference_golden(golden, name):
    g = golden(name)
    tr, sn = _golden_run(g)
    out = g["out"]
    ref_tr = out[:tr.size].reshape(tr.shape, order="F")
    e = rel_l2(ref_tr, tr)
    # the reference's f32 traces sit a few 1e-6 from exact arithmetic of the same scheme
    # (observed 2.4e-6 .. 5.9e-6); a dropped term or a wrong sign or index gives O(1)
    assert 1e-8 < e < 2e-5, e
    if sn is not None:
        es = rel_l2(out[tr.size:].reshape(sn.shape, order="F"), sn)
        # the 101x101 case's last snapshot is the weak field left after 1 s of absorption: the
        # same absolute rounding is 2e-4 of its norm
        assert es < (1e-3 if name == "afd2d_cfg0" else 2e-5), es


def test_f64_twin_early_steps_agree_with_f32_oracle(golden):
    g = golden("afd2d_cfg0")
    vel, rick = g["vel"], g["rick"]
    nt = 200
    tr, sn = O.afd_f64(vel, nt, 0.001, (10.0, 10.0), 30, 0.01, g["src"], rick[:, :nt], jsnap=100)
    _, sn32, _ = O.afd2d(vel, nt, 0.001, 10.0, 10.0, 30, 0.01, [50], [5], rick[:, :nt], jsnap=100)
    # one hundred steps after the source: f32 rounding only (observed 3.5e-7)
    assert rel_l2(sn32[1].T, sn[:, :, 1]) < 1e-6


def _poly_field(shape, d, three_d):
    z = np.arange(shape[0])[:, None] * d[0]
    x = np.arange(shape[1])[None, :] * d[1]
    if not three_d:
        u = z ** 5 - 2 * z ** 3 + 3 * x ** 4 + z * z * x - 4 * x ** 2
        lap = 20 * z ** 3 - 12 * z + 36 * x ** 2 + 2 * x - 8
        return u, lap
    z, x = z[:, :, None], x[:, :, None]
    y = np.arange(shape[2])[None, None, :] * d[2]
    u = z ** 5 - 2 * z ** 3 + 3 * x ** 4 + z * z * x + y ** 5 - 4 * x * y * y
    lap = 20 * z ** 3 - 12 * z + 36 * x ** 2 + 2 * x + 20 * y ** 3 - 8 * x
    return u, lap


@pytest.mark.parametrize("three_d", [False, True])
def test_lap_f64_exact_on_quintics(three_d):
    # the 4th-order second difference (-1, 16, -30, 16, -1)/12h^2 is exact on polynomials of
    # degree <= 5 (truncation error -h^4 u''''''/90), so the Laplacian of a quintic is exact
    shape, d = ((20, 17, 15), (0.7, 1.3, 0.9)) if three_d else ((23, 19), (0.7, 1.3))
    u, lap = _poly_field(shape, d, three_d)
    out = O.lap_f64(u, d)
    inner = (slice(2, -2), slice(2, -2)) + ((slice(2, -2),) if three_d else ())
    np.testing.assert_allclose(out[inner], lap[inner], rtol=1e-9, atol=1e-9 * np.abs(lap).max())
    # the two-cell rim is left zero (the reference never writes it)
    assert not out[:2].any() and not out[-2:].any() and not out[:, :2].any() and not out[:, -2:].any()


def test_lap_f64_plane_wave_symbol():
    # cos(k.x) is an eigenfunction; its eigenvalue is the 4th-order symbol per axis,
    # -(30 - 32 cos(kh) + 2 cos(2kh)) / (12 h^2)
    n, d = (26, 31, 22), (10.0, 12.0, 9.0)
    k = (0.021, -0.013, 0.034)
    idx = np.meshgrid(*[np.arange(m) * h for m, h in zip(n, d)], indexing="ij")
    u = np.cos(sum(ki * xi for ki, xi in zip(k, idx)))
    sym = sum((30 - 32 * np.cos(ki * h) + 2 * np.cos(2 * ki * h)) / (12 * h * h) for ki, h in zip(k, d))
    out = O.lap_f64(u, d)
    np.testing.assert_allclose(out[2:-2, 2:-2, 2:-2], -sym * u[2:-2, 2:-2, 2:-2], rtol=1e-8, atol=1e-12)
    # the symbol tends to |k|^2 for k h -> 0 (consistency)
    assert abs(sym / sum(ki * ki for ki in k) - 1) < 1e-3


def test_f64_twin_linear_and_f32_is_not(golden):
    g = golden("afd2d_two_sources")
    vel = g["vel"]
    # wavelet samples with their last four mantissa bits cleared, so that 3 * rick is exact in f32
    rick = (g["rick"].view(np.int32) & ~np.int32(15)).view(np.float32)
    assert np.array_equal((3 * rick).astype(np.float64), 3 * rick.astype(np.float64))
    nt = int(g["nt"])
    args = (vel, nt, float(g["dt"]), (float(g["dz"]), float(g["dx"])), int(g["nbt"]), float(g["ct"]), g["src"])
    t1, _ = O.afd_f64(*args, rick)
    t3, _ = O.afd_f64(*args, 3 * rick)
    e64 = np.abs(t3 - 3 * t1).max() / np.abs(t1).max()
    assert e64 < 1e-13, e64
    sx, sz = g["src"][0:2].astype(int), g["src"][2:4].astype(int)
    d1, _, _ = O.afd2d(vel, nt, float(g["dt"]), float(g["dx"]), float(g["dz"]), int(g["nbt"]), float(g["ct"]), sx, sz,
                       rick)
    d3, _, _ = O.afd2d(vel, nt, float(g["dt"]), float(g["dx"]), float(g["dz"]), int(g["nbt"]), float(g["ct"]), sx, sz,
                       3 * rick)
    e32 = np.abs(d3.astype(np.float64) - 3 * d1.astype(np.float64)).max() / np.abs(d1).max()
    assert e32 > 100 * e64  # the f32 scheme is linear only to its rounding
