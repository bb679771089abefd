"""Pins the oracle (CPU restatement) against the compiled reference's golden vectors and
against what the mathematics fixes.  CPU only."""
import itertools

import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle as O


# ------------------------------------------------------------------ geometry

def test_next_fast_size_brute_force():
    smooth = sorted({2 ** a * 3 ** b * 5 ** c for a, b, c in itertools.product(range(14), range(9), range(7))})
    for n in range(1, 3000):
        expect = next(m for m in smooth if m >= n)
        assert O.next_fast_size(n) == expect


def test_padded_sizes_table():
    # pywave/src/wave_fft3.c:45-60 applied to the configurations' updated domains
    assert O.padded_sizes(161, 161)[:2] == (162, 162)
    assert O.padded_sizes(261, 261, 261) == (270, 270, 270)
    assert O.padded_sizes(188, 188, 188) == (192, 192, 192)
    assert O.padded_sizes(141, 141)[:2] == (144, 144)
    assert O.padded_sizes(64, 60)[:2] == (64, 60)  # no slack


def test_abc_weights_closed_form():
    w = O.abc_weights(30, 0.01)
    assert w[-1] == 1.0  # innermost cell undamped
    k = 29 - np.arange(30)
    expect = np.exp(-(0.01 ** 2) * k.astype(np.float64) ** 2)
    np.testing.assert_allclose(w, expect, rtol=1e-6)
    assert abs(float(w[0]) - 0.9193) < 1e-4
    assert np.all(np.diff(w) > 0)


# ------------------------------------------------------------------ wavelet

def test_ricker_vs_reference(golden):
    g = golden("ricker")
    for (nt, dt, f0, t0, A), ref in zip(g["params"], g["waves"]):
        nt = int(nt)
        w = O.ricker(nt, dt, f0, t0, A)
        ref = ref[:nt]
        # the reference filters in f32 KISS FFT, the oracle in f64: agree to f32 noise
        assert np.abs(w - ref).max() <= 1e-6 * np.abs(ref).max()
        k0 = int(np.float32(t0) / np.float32(dt))
        assert int(np.argmax(np.abs(w))) == k0
        # zero phase: symmetric about the onset where both sides fit
        m = min(k0, nt - 1 - k0, 60)
        np.testing.assert_allclose(w[k0 - m:k0], w[k0 + m:k0:-1], atol=1e-6 * abs(A))


def test_ricker_matches_analytic_shape():
    # the filtered spike approaches the continuous Ricker A(1-2x)e^{-x}, x=(pi f (t-t0))^2
    # times the filter's peak value; check the shape (normalised) to 1e-3
    nt, dt, f0, t0 = 2000, 0.001, 10.0, 0.5
    w = O.ricker(nt, dt, f0, t0, 1.0).astype(np.float64)
    k0 = int(np.float32(t0) / np.float32(dt))  # 499, not 500: f32 0.5/0.001 = 499.99997
    assert k0 == 499
    t = (np.arange(nt) - k0) * dt
    x = (np.pi * f0 * t) ** 2
    r = (1 - 2 * x) * np.exp(-x)
    np.testing.assert_allclose(w / w.max(), r / r.max(), atol=2e-3)


# ------------------------------------------------------------------ afd engine

def _afd2d_oracle_from_golden(g):
    src = g["src"]
    ns = len(src) // 5
    return O.afd2d(g["vel"], int(g["nt"]), float(g["dt"]), float(g["dx"]), float(g["dz"]), int(g["nbt"]),
                   float(g["ct"]), src[:ns], src[ns:2 * ns], g["rick"], jsnap=int(g["jsnap"]))


@pytest.mark.parametrize("name", ["afd2d_cfg0", "afd2d_two_sources", "afd2d_noslack"])
def test_afd2d_bit_exact(golden, name):
    g = golden(name)
    nt, jsnap = int(g["nt"]), int(g["jsnap"])
    nz, nx = g["vel"].shape
    d, snaps, _ = _afd2d_oracle_from_golden(g)
    out = g["out"]
    tr = out[:nt * nx].reshape(nt, nx, order="F")
    assert np.array_equal(tr, d.T)
    if jsnap > 0:
        w = out[nt * nx:].reshape(nz, nx, -1, order="F")
        for k in range(w.shape[-1]):
            assert np.array_equal(w[..., k], snaps[k].T)


def test_afd2d_tri_bit_exact(golden):
    g = golden("afd2d_tri")
    nt = int(g["nt"])
    nz, nx = g["vel"].shape
    d, snaps, img = O.afd2d(g["vel"], nt, float(g["dt"]), float(g["dx"]), float(g["dz"]), int(g["nbt"]),
                            float(g["ct"]), [], [], None, tri=True, dat=np.ascontiguousarray(g["dat"].T),
                            jsnap=int(g["jsnap"]))
    out = g["out"]
    assert np.array_equal(out[:nt * nx].reshape(nt, nx, order="F"), g["dat"])  # data echoed
    w = out[nt * nx:].reshape(nz, nx, -1, order="F")
    assert all(np.array_equal(w[..., k], snaps[k].T) for k in range(w.shape[-1]))
    assert np.array_equal(img.T, w[..., 0])  # final backward step is it = 0 = snapshot 0


def test_afd3d_bit_exact(golden):
    g = golden("afd3d_two_sources")
    src, nt = g["src"], int(g["nt"])
    ns = len(src) // 6
    nz, nx, ny = g["vel"].shape
    d, snaps, _ = O.afd3d(g["vel"], nt, float(g["dt"]), float(g["dx"]), float(g["dy"]), float(g["dz"]),
                          int(g["nbt"]), float(g["ct"]), src[:ns], src[ns:2 * ns], src[2 * ns:3 * ns],
                          g["rick"], jsnap=int(g["jsnap"]))
    out = g["out"]
    assert np.array_equal(out[:nt * nx * ny].reshape(nt, nx, ny, order="F"), d.transpose(2, 1, 0))
    w = out[nt * nx * ny:].reshape(nz, nx, ny, -1, order="F")
    assert all(np.array_equal(w[..., k], snaps[k].transpose(2, 1, 0)) for k in range(w.shape[-1]))


def test_afd3d_tri_bit_exact(golden):
    g = golden("afd3d_tri")
    nt = int(g["nt"])
    nz, nx, ny = g["vel"].shape
    _, snaps, _ = O.afd3d(g["vel"], nt, float(g["dt"]), float(g["dx"]), float(g["dy"]), float(g["dz"]),
                          int(g["nbt"]), float(g["ct"]), [], [], [], None, tri=True,
                          dat=np.ascontiguousarray(g["dat"].transpose(2, 1, 0)), jsnap=int(g["jsnap"]))
    w = g["out"][nt * nx * ny:].reshape(nz, nx, ny, -1, order="F")
    assert all(np.array_equal(w[..., k], snaps[k].transpose(2, 1, 0)) for k in range(w.shape[-1]))


def test_afd2d_own_wavelet_within_tolerance(golden):
    g = golden("afd2d_cfg0")
    nt = int(g["nt"])
    rick = O.ricker(nt, float(g["dt"]), 20, 0.1, 1)[None, :]
    d, _, _ = O.afd2d(g["vel"], nt, float(g["dt"]), 10, 10, 30, 0.01, [50], [5], rick, jsnap=0)
    tr = g["out"][:nt * 101].reshape(nt, 101, order="F")
    assert rel_l2(d.T, tr) <= 1e-5


def test_afd_linear_in_amplitude():
    rng = np.random.default_rng(3)
    vel = (1500 + 1000 * rng.random((30, 34))).astype(np.float32)
    r1 = O.ricker(200, 0.001, 25, 0.03, 1.0)[None, :]
    d1, _, _ = O.afd2d(vel, 200, 0.001, 10, 10, 8, 0.03, [15], [12], r1)
    # a power-of-two scale commutes with every f32 operation (up to denormal rounding)
    d2, _, _ = O.afd2d(vel, 200, 0.001, 10, 10, 8, 0.03, [15], [12], (2 * r1).astype(np.float32))
    assert rel_l2(d2, 2 * d1) < 1e-12
    d3, _, _ = O.afd2d(vel, 200, 0.001, 10, 10, 8, 0.03, [15], [12], (3 * r1).astype(np.float32))
    assert rel_l2(d3, 3 * d1) < 1e-5
    d0, _, _ = O.afd2d(vel, 200, 0.001, 10, 10, 8, 0.03, [15], [12], 0 * r1)
    assert not d0.any()


# ------------------------------------------------------------------ FWI

def test_fwi_forward_bit_exact(golden):
    g = golden("fwi_small")
    d = O.fwi_forward(g["par"], g["vtrue"], g["wav"])
    assert np.array_equal(d, g["data"])
    m = golden("fwi_marine")
    assert np.array_equal(O.fwi_forward(m["par"], m["vel"], m["wav"]), m["data"])


def test_fwi_geometry_closed_form():
    # preparation (pywave/src/wave_fwiutil.c:45-60) for the demo-like geometry
    par = np.array([400, 800, 0.01, 0.01, 0, 0, 3000, 0.001, 0, 0, 32, 0.24, 0.2, 5, 60, 0.005, 8, 1, 1, 5],
                   np.float32)
    a = O.make_acq(par)
    geo = O.acq_geometry(a)
    assert (geo["padnz"], geo["padnx"]) == (520, 920)
    assert geo["sz"] == 65 and geo["rz"] == 61
    assert geo["ds_v"] == 24 and geo["s0_v"] == 80 and geo["dr_v"] == 1
    assert geo["r0_v"] == [60] * 32 and geo["nr2"] == [800] * 32 and geo["r02"] == [0] * 32
    np.testing.assert_allclose(geo["bc"], np.exp(-(0.005 * (60 - np.arange(60))) ** 2), rtol=1e-6)


def test_gradient_av_bit_exact_and_scaling(golden):
    g = golden("fwi_small")
    g1, f1, sc = O.gradient_av(g["par"], g["v0"], g["wav"], g["data"])
    assert np.array_equal(g1, g["grad1"])
    assert np.float32(f1) == g["fcost"][0]
    assert abs(np.abs(g1).max() - 0.1) < 0.1  # scaled to 0.1/max before smoothing
    # a second call reuses the first call's scaling (static 'first' in wave_fwigradient.c:245)
    g2, f2, _ = O.gradient_av(g["par"], (g["v0"] * 1.01).astype(np.float32), g["wav"], g["data"], scaling=sc)
    assert np.array_equal(g2, g["grad2"]) and np.float32(f2) == g["fcost"][1]
    g4, f4, _ = O.gradient_av(g["par4"], g["v0"], g["wav"], g["data"])
    assert np.array_equal(g4, g["grad_int4"])


def test_fwi_lbfgs_bit_exact(golden):
    g = golden("fwi_small")
    vinv, k = O.fwi(g["par"], g["v0"], g["wav"], g["data"])
    assert k == 3
    assert np.array_equal(vinv, g["vinv"])


def _point_source_field(par, wav):
    """The active point-source term written as a passive source field (nz, nx, nt, ns):
    source_map's 3x3 stencil at each shot position times the wavelet."""
    a = O.make_acq(par)
    nz, nx, nt, ns, nb = int(par[0]), int(par[1]), int(par[6]), int(par[10]), int(par[14])
    src = np.zeros((nz, nx, nt, ns), np.float32)
    for s in range(ns):
        rr = O.source_map(a.s0_v + s * a.ds_v, a.sz, 2, 2, a.padnx, a.padnz)
        inner = rr[nb:nb + nz, nb:nb + nx]
        assert abs(float(np.abs(rr).sum()) - float(np.abs(inner).sum())) == 0.0
        src[:, :, :, s] = inner[:, :, None] * np.asarray(wav, np.float32)[None, None, :]
    return src


def test_passive_field_equal_to_point_sources_reproduces_active(golden):
    # gradient_pas_av adds the field on the interior where gradient_av adds the 3x3 point source;
    # with the same term both must give the active golden bit for bit (indexing of iz/ix/it/is)
    g = golden("fwi_small")
    src = _point_source_field(g["par"], g["wav"])
    assert np.array_equal(O.fwi_forward(g["par"], g["vtrue"], g["wav"], src=src), g["data"])
    g1, f1, _ = O.gradient_av(g["par"], g["v0"], g["wav"], g["data"], src=src)
    assert np.array_equal(g1, g["grad1"]) and np.float32(f1) == g["fcost"][0]
    # a field shifted by one sample in z is not the same term
    sh = np.roll(src, 1, axis=0)
    assert not np.array_equal(O.fwi_forward(g["par"], g["vtrue"], g["wav"], src=sh), g["data"])


def test_passive_forward_linear_in_field(golden):
    g = golden("pfwi_small")
    d1 = O.fwi_forward(g["par"], g["vtrue"], g["wav"], src=g["src"])
    assert np.array_equal(d1, g["data"])
    d2 = O.fwi_forward(g["par"], g["vtrue"], g["wav"], src=2 * g["src"])
    assert np.array_equal(d2, 2 * d1)  # power-of-two scaling is exact in binary floating point
    assert not O.fwi_forward(g["par"], g["vtrue"], g["wav"], src=0 * g["src"]).any()


def test_passive_gradient_bit_exact(golden):
    g = golden("pfwi_small")
    g1, f1, sc = O.gradient_av(g["par"], g["v0"], g["wav"], g["data"], src=g["src"])
    assert np.array_equal(g1, g["grad1"]) and np.float32(f1) == g["fcost"][0]
    g2, f2, _ = O.gradient_av(g["par"], (g["v0"] * 1.01).astype(np.float32), g["wav"], g["data"],
                              scaling=sc, src=g["src"])
    assert np.array_equal(g2, g["grad2"]) and np.float32(f2) == g["fcost"][1]
    # at the true model the residual vanishes exactly
    _, f0, _ = O.gradient_av(g["par"], g["vtrue"], g["wav"], g["data"], src=g["src"])
    assert f0 == 0.0


def test_passive_fwi_lbfgs_bit_exact(golden):
    # pfwi resets the trial step to 1 every 5 iterations (pywave/src/wave_fwi.c:282, repeat=5 from
    # pywave/src/pfwi.c:1846); 7 iterations run through one reset
    g = golden("pfwi_small")
    vinv, k = O.fwi(g["par"], g["v0"], g["wav"], g["data"], src=g["src"])
    assert k == 7
    assert np.array_equal(vinv, g["vinv"])


def _tri_adj_1d(x, r):
    """Adjoint triangle smoothing, independent form: full convolution with the triangle
    (r-|k|)/r^2 then half-sample reflection of the overhang back into the line."""
    nx = len(x)
    T = (r - np.abs(np.arange(-r + 1, r))) / float(r * r)
    full = np.convolve(np.asarray(x, np.float64), T)
    y = full[r - 1:r - 1 + nx].copy()
    for m in range(r - 1):
        y[m] += full[r - 2 - m]
        y[nx - 1 - m] += full[r - 1 + nx + m]
    return y


def test_source_map_is_separable_121():
    rr = O.source_map(10, 7, 2, 2, 20, 16)
    expect = np.zeros((16, 20))
    k = np.array([1, 2, 1]) / 4.0
    expect[6:9, 9:12] = np.outer(k, k)
    np.testing.assert_allclose(rr, expect, atol=1e-7)
    # at an edge the overhang folds back (reflection)
    rr = O.source_map(0, 0, 2, 2, 8, 8)
    assert abs(float(rr.sum()) - 1.0) < 1e-6


def test_gradient_smooth2_matches_convolution():
    rng = np.random.default_rng(5)
    g = rng.standard_normal((13, 17)).astype(np.float32)
    out = O.gradient_smooth2(g, 3, 3, 0, 2.5)
    ref = np.apply_along_axis(_tri_adj_1d, 0, g.astype(np.float64), 3)
    ref = np.apply_along_axis(_tri_adj_1d, 1, ref, 3) * 2.5
    np.testing.assert_allclose(out, ref, rtol=2e-5, atol=2e-6)
    # the smoother preserves the sum (adjoint of a normalised reflecting smoother)
    s = O.gradient_smooth2(g, 3, 3, 0, 1.0)
    assert abs(float(s.sum()) - float(g.sum())) < 1e-4


# ------------------------------------------------------------------ least-squares TRI (lstric)

def test_lstri_modelling_bit_exact(golden):
    g = golden("lstri_small")
    assert np.array_equal(O.lstri(g["par"], g["vel"], g["src"]), g["data"])


def test_lstri_inversion_bit_exact(golden):
    # six CG iterations per shot; the second shot starts from the first shot's last step
    # (np_cgstep keeps its step vectors between calls, pywave/src/wave_cgstep.c:26-31)
    g = golden("lstri_small")
    assert np.array_equal(O.lstri(g["par_inv"], g["vel"], g["data"]), g["src_inv"])


def test_timerev_adjoint_identity(golden):
    # <A m, d> == <m, A^T d>: the migration branch is the transpose of the modelling branch
    g = golden("lstri_small")
    rng = np.random.default_rng(2)
    par = g["par"]
    m = rng.standard_normal(g["src"].shape[:3]).astype(np.float32)
    d = rng.standard_normal(g["data"].shape[:2]).astype(np.float32)
    am = O.timerev(par, g["vel"], m, adj=False).astype(np.float64)
    atd = O.timerev(par, g["vel"], d, adj=True).astype(np.float64)
    lhs, rhs = float((am * d).sum()), float((m * atd).sum())
    assert abs(lhs - rhs) <= 1e-5 * max(abs(lhs), abs(rhs))
    # a planted slip (the data injected one sample late) breaks the identity
    d_late = np.roll(d, 1, axis=0)
    bad = float((m * O.timerev(par, g["vel"], d_late, adj=True).astype(np.float64)).sum())
    assert abs(lhs - bad) > 1e-2 * abs(lhs)


def test_lstri_modelling_linear_and_causal(golden):
    g = golden("lstri_small")
    d1 = O.lstri(g["par"], g["vel"], g["src"])
    assert np.array_equal(O.lstri(g["par"], g["vel"], 2 * g["src"]), 2 * d1)
    assert not O.lstri(g["par"], g["vel"], 0 * g["src"]).any()
    # a field that starts at sample k gives no data before k + 1 (record, then step, then inject)
    late = g["src"].copy()
    late[:, :, :60, :] = 0
    d = O.lstri(g["par"], g["vel"], late)
    assert not d[:61].any() and d[61:].any()


def test_lstri_inversion_reduces_residual(golden):
    g = golden("lstri_small")
    d_fit = O.lstri(g["par"], g["vel"], g["src_inv"])
    r = np.linalg.norm(d_fit - g["data"]) / np.linalg.norm(g["data"])
    assert r < 0.5


# ------------------------------------------------------------------ pseudo-spectral engine

@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 8, 9, 12, 15, 16, 20, 25, 36, 45, 60, 72, 90, 100, 7, 14])
def test_cfft_against_numpy(n):
    # the KISS-algorithm transform agrees with numpy's to f32 round-off; radix 2, 3, 4, 5 and a
    # generic prime (7) all occur
    rng = np.random.default_rng(n)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    ref = np.fft.fft(x.astype(np.complex128))
    got = O.cfft(x)
    assert np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max() * max(1, np.log2(n))
    back = O.cfft(got, inverse=True)  # unscaled: n x
    assert np.abs(back / n - x).max() <= 1e-5 * np.abs(x).max()
    # a planted transposition of one twiddle sign would show here
    assert not np.allclose(got, np.fft.ifft(x) * n, atol=1e-3) or n <= 2


@pytest.mark.parametrize("n", [4, 10, 18, 36, 80, 90, 144])
def test_rfft_against_numpy(n):
    rng = np.random.default_rng(n + 1)
    x = rng.standard_normal(n).astype(np.float32)
    X = O.rfft(x)
    ref = np.fft.rfft(x.astype(np.float64))
    assert np.abs(X - ref).max() <= 2e-6 * np.abs(ref).max() * np.log2(n)
    assert np.abs(O.irfft(X, n) / n - x).max() <= 1e-5 * np.abs(x).max()


def test_spectral_laplacian_of_a_plane_wave():
    # through the 2D engine's operator: the Laplacian of cos(2 pi (kx x + kz z)) on the periodic
    # padded grid is -(2 pi)^2 |k|^2 times it (kx, kz on the grid's wavenumbers), which pins the
    # wavenumber table, the centering and the scaling
    import ctypes
    n1, n2, dz, dx = 36, 40, 10.0, 12.0
    L = O.lib()
    L.orc_psl_new.restype = ctypes.c_void_p
    L.orc_psl_new.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                              ctypes.c_float]
    L.orc_psl_apply.argtypes = [ctypes.c_void_p, np.ctypeslib.ndpointer(np.float32),
                                np.ctypeslib.ndpointer(np.float32)]
    L.orc_psl_free.argtypes = [ctypes.c_void_p]
    P = L.orc_psl_new(n1, n2, 1, dz, dx, 1.0)
    z = np.arange(n1) * dz
    x = np.arange(n2) * dx
    kz, kx = 3 / (n1 * dz), 5 / (n2 * dx)
    f = np.cos(2 * np.pi * (kx * x[:, None] + kz * z[None, :])).astype(np.float32)  # [x][z]
    out = np.zeros_like(f)
    L.orc_psl_apply(P, np.ascontiguousarray(f), out)
    L.orc_psl_free(P)
    expect = -(2 * np.pi) ** 2 * (kx * kx + kz * kz) * f
    assert np.abs(out - expect).max() <= 1e-4 * np.abs(expect).max()


def _aps_src(g, dims):
    src = g["src"]
    ns = len(src) // (dims + 3)
    return [src[k * ns:(k + 1) * ns] for k in range(dims)]


def test_aps2d_bit_exact(golden):
    g = golden("aps2d_two_sources")
    nt = int(g["nt"])
    nz, nx = g["vel"].shape
    sx, sz = _aps_src(g, 2)
    d, snaps, _ = O.afd2d(g["vel"], nt, float(g["dt"]), float(g["dx"]), float(g["dz"]), int(g["nbt"]),
                          float(g["ct"]), sx, sz, g["rick"], jsnap=int(g["jsnap"]), ps=True)
    out = g["out"]
    assert np.array_equal(out[:nt * nx].reshape(nt, nx, order="F"), d.T)
    w = out[nt * nx:].reshape(nz, nx, -1, order="F")
    assert all(np.array_equal(w[..., k], snaps[k].T) for k in range(w.shape[-1]))


def test_aps2d_tri_bit_exact(golden):
    g = golden("aps2d_tri")
    nt = int(g["nt"])
    nz, nx = g["vel"].shape
    _, snaps, img = O.afd2d(g["vel"], nt, float(g["dt"]), float(g["dx"]), float(g["dz"]), int(g["nbt"]),
                            float(g["ct"]), [], [], None, tri=True, dat=np.ascontiguousarray(g["dat"].T),
                            jsnap=int(g["jsnap"]), ps=True)
    w = g["out"][nt * nx:].reshape(nz, nx, -1, order="F")
    assert all(np.array_equal(w[..., k], snaps[k].T) for k in range(w.shape[-1]))
    assert np.array_equal(img.T, w[..., 0])


def test_aps3d_bit_exact(golden):
    g = golden("aps3d_small")
    nt = int(g["nt"])
    nz, nx, ny = g["vel"].shape
    sx, sy, sz = _aps_src(g, 3)
    d, snaps, _ = O.afd3d(g["vel"], nt, float(g["dt"]), float(g["dx"]), float(g["dy"]), float(g["dz"]),
                          int(g["nbt"]), float(g["ct"]), sx, sy, sz, g["rick"], jsnap=int(g["jsnap"]), ps=True)
    out = g["out"]
    assert np.array_equal(out[:nt * nx * ny].reshape(nt, nx, ny, order="F"), d.transpose(2, 1, 0))
    w = out[nt * nx * ny:].reshape(nz, nx, ny, -1, order="F")
    assert all(np.array_equal(w[..., k], snaps[k].transpose(2, 1, 0)) for k in range(w.shape[-1]))


# ------------------------------------------------------------------ visco-acoustic modelling

def test_visco_forward_bit_exact(golden):
    g = golden("fwi_visco")
    assert np.array_equal(O.fwi_forward_visco(g["par"], g["vtrue"], g["q"], g["wav"]), g["data"])


def test_visco_forward_tends_to_acoustic_at_high_q(golden):
    # tau = taue/taus - 1 = 2/(sqrt(Q^2+1) - 1) -> 0 as Q grows: the memory variable vanishes and
    # the visco-acoustic data approach the acoustic data; at Q = 20-60 they differ by ~10%
    g = golden("fwi_small")
    q_hi = np.full(g["vtrue"].shape, 1e5, np.float32)
    d_hi = O.fwi_forward_visco(g["par"], g["vtrue"], q_hi, g["wav"])
    assert rel_l2(d_hi, g["data"]) < 1e-3
    v = golden("fwi_visco")
    assert rel_l2(v["data"], g["data"]) > 0.05
    # attenuation: the low-Q data carry less energy than the acoustic data
    assert np.linalg.norm(v["data"]) < np.linalg.norm(g["data"])


def test_visco_forward_marine_bit_exact(golden):
    # moving-spread acquisition (acqui_type 2: per-shot receiver windows) with Q 15-45
    g = golden("fwi_marine")
    d = O.fwi_forward_visco(g["par"], g["vel"], g["q_visco"], g["wav"])
    assert np.array_equal(d, g["data_visco"])
    assert rel_l2(g["data_visco"], g["data"]) > 0.05


def test_visco_gradient_bit_exact(golden):
    # gradient_v (the gradient fwi() takes for media != 1): two calls sharing the first call's
    # scaling, and interval 4, against the compiled reference
    g = golden("fwi_visco")
    g1, f1, sc = O.gradient_av(g["par"], g["v0"], g["wav"], g["data"], q=g["q"])
    assert np.array_equal(g1, g["grad1"]) and np.float32(f1) == g["fcost"][0]
    g2, f2, _ = O.gradient_av(g["par"], (g["v0"] * 1.01).astype(np.float32), g["wav"], g["data"], scaling=sc,
                              q=g["q"])
    assert np.array_equal(g2, g["grad2"]) and np.float32(f2) == g["fcost"][1]
    g4, f4, _ = O.gradient_av(g["par4"], g["v0"], g["wav"], g["data"], q=g["q"])
    assert np.array_equal(g4, g["grad_int4"]) and np.float32(f4) == g["fcost_int4"][0]


def test_visco_fwi_trajectory_bit_exact(golden):
    g = golden("fwi_visco")
    vinv, k = O.fwi(g["par"], g["v0"], g["wav"], g["data"], q=g["q"])
    assert k == int(g["par"][19])
    assert np.array_equal(vinv, g["vinv"])


def test_visco_gradient_tends_to_acoustic_at_high_q(golden):
    # with tau -> 0 the backward memory term vanishes and tmp -> p1: gradient_v approaches
    # gradient_av on the same data; at Q 20-60 the two differ
    g = golden("fwi_small")
    q_hi = np.full(g["v0"].shape, 1e5, np.float32)
    gv, fv, _ = O.gradient_av(g["par"], g["v0"], g["wav"], g["data"], q=q_hi)
    ga, fa, _ = O.gradient_av(g["par"], g["v0"], g["wav"], g["data"])
    assert rel_l2(gv, ga) < 2e-3 and abs(fv - fa) < 2e-3 * fa
    v = golden("fwi_visco")
    gl, _, _ = O.gradient_av(g["par"], g["v0"], g["wav"], g["data"], q=v["q"])
    assert rel_l2(gl, ga) > 0.05


def test_fcost_f32_reproduces_the_reference_misfit(golden):
    # gradient_av accumulates 0.5 r^2 into an f32 scalar (pywave/src/wave_fwigradient.c:179-184);
    # the oracle's restatement of that accumulation gives the golden misfit bit for bit, and the
    # exact (f64) sum differs from it by the accumulator's rounding
    g = golden("fwi_small")
    pred = O.fwi_forward(g["par"], g["v0"], g["wav"])
    f32 = O.fcost_f32(g["data"], pred)
    assert np.float32(f32) == g["fcost"][0]
    r = g["data"].astype(np.float64) - pred.astype(np.float32).astype(np.float64)
    exact = 0.5 * float(np.sum(r * r))
    assert abs(exact - f32) <= 1e-4 * exact
