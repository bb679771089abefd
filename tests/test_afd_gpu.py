"""HIP afd2d/afd3d against the compiled reference's golden vectors and the oracle.

Bar: bit-exact (np.array_equal, +0 == -0) against the reference's golden outputs, both through
the resident-plan API fed the reference's wavelet and through the drop-in API, whose f32 wavelet
reproduces the reference's; the stated floating-point tolerance (rel-L2 <= 1e-5) is asserted too."""
import numpy as np
import pytest

import pywave_amd
from conftest import rel_l2
from oracle import oracle as O
from pywave_amd import AfdPlan
from pywave_amd.workloads import plain3d_model

pytestmark = pytest.mark.gpu

TOL = 1e-5  # rel-L2 for floating-point results through the drop-in


def _plan2d(g, jsnap=None):
    return AfdPlan(g["vel"], (float(g["dz"]), float(g["dx"])), int(g["nt"]), float(g["dt"]), int(g["nbt"]),
                   float(g["ct"]), int(g["jsnap"]) if jsnap is None else jsnap)


def _plan3d(g):
    return AfdPlan(g["vel"], (float(g["dz"]), float(g["dx"]), float(g["dy"])), int(g["nt"]), float(g["dt"]),
                   int(g["nbt"]), float(g["ct"]), int(g["jsnap"]))


@pytest.mark.parametrize("name", ["afd2d_cfg0", "afd2d_two_sources", "afd2d_noslack"])
def test_afd2d_golden_bit_exact(golden, name):
    g = golden(name)
    p = _plan2d(g)
    p.set_sources(g["src"], len(g["src"]) // 5, wavelet=g["rick"])
    p.run()
    out = p.output()
    assert out.shape == g["out"].shape
    assert np.array_equal(out, g["out"])


def test_afd2d_tri_golden_bit_exact(golden):
    g = golden("afd2d_tri")
    p = _plan2d(g)
    p.set_tri_data(g["dat"])
    p.run(tri=True)
    assert np.array_equal(p.output(), g["out"])
    # the reference's (unreturned) img is the final field; snapshot 0 is taken at it = 0
    nz, nx = g["vel"].shape
    nt = int(g["nt"])
    w0 = g["out"][nt * nx:nt * nx + nz * nx].reshape(nz, nx, order="F")
    assert np.array_equal(p.final_field(), w0)


NVAR = 28  # pw_afd_variant_count(): 6 general, 6 lean, 12 pipelined and 4 two-step 3D step kernels
TWO_STEP = (24, 25, 26, 27)  # variants that advance two time steps per launch


@pytest.mark.parametrize("variant", range(NVAR))
def test_afd3d_golden_bit_exact_all_variants(golden, variant):
    g = golden("afd3d_two_sources")
    p = _plan3d(g)
    p.set_sources(g["src"], len(g["src"]) // 6, wavelet=g["rick"])
    for ny_chunk in (0, 5, 17):
        p.set_variant(variant, ny_chunk)
        p.run()
        assert np.array_equal(p.output(), g["out"]), (variant, ny_chunk)


@pytest.mark.parametrize("variant", [-1, 2, 6, 12, 13, 20] + list(TWO_STEP))
def test_afd3d_tri_golden_bit_exact(golden, variant):
    g = golden("afd3d_tri")
    p = _plan3d(g)
    p.set_tri_data(g["dat"])
    if variant >= 0:
        p.set_variant(variant, 7)
    p.run(tri=True)
    assert np.array_equal(p.output(), g["out"])


@pytest.mark.parametrize("variant", TWO_STEP)
@pytest.mark.parametrize("tri", [False, True])
def test_two_step_odd_nt_vs_oracle(variant, tri):
    # odd nt: the last step runs on the single-step kernel; snapshots fall on both steps of a launch
    rng = np.random.default_rng(300 + variant + 7 * tri)
    vel, nbt, ns, (pz, px, py) = _random_case(rng, 3)
    nz, nx, ny = vel.shape
    nt, dt, ct = 71, 0.0008, float(rng.uniform(0.005, 0.06))
    jsnap = 5
    p = AfdPlan(vel, (12.0, 10.0, 8.0), nt, dt, nbt, ct, jsnap)
    if tri:
        dat = rng.standard_normal((nt, nx, ny)).astype(np.float32)
        d, snaps, _ = O.afd3d(vel, nt, dt, 10.0, 8.0, 12.0, nbt, ct, [], [], [], None, tri=True,
                              dat=np.ascontiguousarray(dat.transpose(2, 1, 0)), jsnap=jsnap)
        p.set_tri_data(dat)
    else:
        f = rng.uniform(10, 40, ns).astype(np.float32)
        t = rng.uniform(0.0, 0.05, ns).astype(np.float32)
        A = rng.uniform(-2, 2, ns).astype(np.float32)
        rick = O.ricker_bank(nt, dt, f, t, A)
        d, snaps, _ = O.afd3d(vel, nt, dt, 10.0, 8.0, 12.0, nbt, ct, px, py, pz, rick, jsnap=jsnap)
        p.set_sources(np.concatenate([px, py, pz, f, t, A]), ns, wavelet=rick)
    p.set_variant(variant, int(rng.integers(0, 12)))
    p.run(tri=tri)
    out = p.output()
    if not tri:
        assert np.array_equal(out[:nt * nx * ny], d.ravel())
    assert np.array_equal(out[nt * nx * ny:], snaps.ravel())


def test_dropin_afd2d_cfg0(golden):
    g = golden("afd2d_cfg0")
    nt = 1000
    dout, wfd = pywave_amd.afd2d(g["vel"], nt=nt, dt=0.001, ax=[0, 10, 101], az=[0, 10, 101], ns=1, sx=[50],
                                 sz=[5], f=[20], t=[0.1], A=[1], nbt=30, ct=0.01, jsnap=999, verb=False)
    assert dout.shape == (1000, 101) and wfd.shape == (101, 101, 2)
    ref = g["out"]
    assert rel_l2(dout, ref[:nt * 101].reshape(nt, 101, order="F")) <= TOL
    wref = ref[nt * 101:].reshape(101, 101, 2, order="F")
    assert rel_l2(wfd[..., 1], wref[..., 1]) <= TOL
    assert np.array_equal(dout, ref[:nt * 101].reshape(nt, 101, order="F"))
    assert np.array_equal(wfd, wref)
    # jsnap == 0 returns the flat vector, like the reference
    flat = pywave_amd.afd2d(g["vel"], nt=nt, dt=0.001, ax=[0, 10, 101], az=[0, 10, 101], ns=1, sx=[50],
                            sz=[5], f=[20], t=[0.1], A=[1], nbt=30, ct=0.01, jsnap=0)
    assert flat.ndim == 1 and flat.size == nt * 101
    assert np.array_equal(flat, dout.ravel(order="F"))


def test_dropin_afd3d_matches_reference(golden):
    g = golden("afd3d_two_sources")
    src = g["src"]
    nz, nx, ny = g["vel"].shape
    nt = int(g["nt"])
    dout, wfd = pywave_amd.afd3d(g["vel"], nt=nt, dt=float(g["dt"]), ax=[0, float(g["dx"]), nx],
                                 ay=[0, float(g["dy"]), ny], az=[0, float(g["dz"]), nz], ns=2, sx=src[0:2],
                                 sy=src[2:4], sz=src[4:6], f=src[6:8], t=src[8:10], A=src[10:12],
                                 nbt=int(g["nbt"]), ct=float(g["ct"]), jsnap=int(g["jsnap"]))
    ref = g["out"]
    assert rel_l2(dout, ref[:nt * nx * ny].reshape(nt, nx, ny, order="F")) <= TOL
    wref = ref[nt * nx * ny:].reshape(nz, nx, ny, -1, order="F")
    assert rel_l2(wfd, wref) <= TOL
    assert np.array_equal(dout, ref[:nt * nx * ny].reshape(nt, nx, ny, order="F"))
    assert np.array_equal(wfd, wref)


def test_dropin_tri_echo_and_snapshots(golden):
    g = golden("afd2d_tri")
    nz, nx = g["vel"].shape
    dout, wfd = pywave_amd.afd2d(g["vel"], nt=300, dt=0.001, ax=[0, 10, nx], az=[0, 10, nz], nbt=20, ct=0.02,
                                 jsnap=7, tri=True, dat=g["dat"])
    assert np.array_equal(dout, g["dat"])
    ref = g["out"][300 * nx:].reshape(nz, nx, -1, order="F")
    assert np.array_equal(wfd, ref)  # no wavelet involved: bit-exact through the drop-in


def _random_case(rng, ndim):
    if ndim == 2:
        nz, nx = rng.integers(20, 90, 2)
        shape = (nz, nx)
    else:
        nz, nx, ny = rng.integers(12, 40, 3)
        shape = (nz, nx, ny)
    vel = (1500 + 2000 * rng.random(shape)).astype(np.float32)
    nbt = int(rng.integers(2, 15))
    ns = int(rng.integers(1, 4))
    pos = [rng.integers(0, n, ns).astype(np.float32) for n in shape]
    return vel, nbt, ns, pos


@pytest.mark.parametrize("seed", range(6))
def test_random_2d_vs_oracle(seed):
    rng = np.random.default_rng(100 + seed)
    vel, nbt, ns, (pz, px) = _random_case(rng, 2)
    nt, dt, ct = 180, 0.0009, float(rng.uniform(0.005, 0.06))
    f = rng.uniform(10, 40, ns).astype(np.float32)
    t = rng.uniform(0.0, 0.1, ns).astype(np.float32)
    A = rng.uniform(-2, 2, ns).astype(np.float32)
    rick = O.ricker_bank(nt, dt, f, t, A)
    jsnap = int(rng.integers(0, 40))
    d, snaps, _ = O.afd2d(vel, nt, dt, 9.0, 11.0, nbt, ct, px, pz, rick, jsnap=jsnap)
    p = AfdPlan(vel, (11.0, 9.0), nt, dt, nbt, ct, jsnap)
    p.set_sources(np.concatenate([px, pz, f, t, A]), ns, wavelet=rick)
    p.run()
    out = p.output()
    nz, nx = vel.shape
    assert np.array_equal(out[:nt * nx], d.ravel())
    if jsnap > 0:
        assert np.array_equal(out[nt * nx:], snaps.ravel())


@pytest.mark.parametrize("seed", range(6))
def test_random_3d_vs_oracle(seed):
    rng = np.random.default_rng(200 + seed)
    vel, nbt, ns, (pz, px, py) = _random_case(rng, 3)
    nt, dt, ct = 90, 0.0008, float(rng.uniform(0.005, 0.06))
    f = rng.uniform(10, 40, ns).astype(np.float32)
    t = rng.uniform(0.0, 0.05, ns).astype(np.float32)
    A = rng.uniform(-2, 2, ns).astype(np.float32)
    rick = O.ricker_bank(nt, dt, f, t, A)
    jsnap = int(rng.integers(0, 30))
    d, snaps, _ = O.afd3d(vel, nt, dt, 10.0, 8.0, 12.0, nbt, ct, px, py, pz, rick, jsnap=jsnap)
    p = AfdPlan(vel, (12.0, 10.0, 8.0), nt, dt, nbt, ct, jsnap)
    p.set_sources(np.concatenate([px, py, pz, f, t, A]), ns, wavelet=rick)
    p.set_variant((seed * 5) % NVAR, int(rng.integers(0, 12)))
    p.run()
    out = p.output()
    nz, nx, ny = vel.shape
    assert np.array_equal(out[:nt * nx * ny], d.ravel())
    if jsnap > 0:
        assert np.array_equal(out[nt * nx * ny:], snaps.ravel())


def _cfg1_model():
    return plain3d_model(201)


@pytest.fixture(scope="module")
def cfg1_early():
    # the benchmark grid, 13 steps: traces and the last wavefield (a smooth pulse of thirteen
    # samples as the wavelet: a 15 Hz Ricker that short is all but zero)
    vel = _cfg1_model()
    nt = 13
    rick = (np.sin(np.arange(nt) * np.pi / nt) ** 2)[None, :].astype(np.float32)
    d, snaps, _ = O.afd3d(vel, nt, 0.001, 10, 10, 10, 30, 0.01, [100], [100], [100], rick, jsnap=nt - 1)
    return vel, nt, rick, d, snaps


@pytest.mark.parametrize("variant", [-1, 12, 13, 15, 24, 25])
def test_cfg1_size_early_steps_vs_oracle(cfg1_early, variant):
    vel, nt, rick, d, snaps = cfg1_early
    assert np.abs(snaps).max() > 1e-3
    p = AfdPlan(vel, (10, 10, 10), nt, 0.001, 30, 0.01, nt - 1)
    p.set_sources(np.array([100, 100, 100, 15, 0.0, 1.0], np.float32), 1, wavelet=rick)
    if variant >= 0:
        p.set_variant(variant, 0)
    p.run()
    out = p.output()
    p.close()
    n = vel.shape[0]
    assert np.array_equal(out[:nt * n * n], d.ravel())
    assert np.array_equal(out[nt * n * n:], snaps.ravel())


@pytest.fixture(scope="module")
def large_grid():
    # 241^3 + 30-cell ABC = 301^3 updated (304^3 padded): three fields of more than 240 MB, which
    # the plan lays out with rows on 128 B boundaries (the 261^3 benchmark grid, whose fields fit
    # the Infinity Cache, gets 64 B ones); the wavefield after a few steps around a central source
    vel = plain3d_model(241)
    nt = 12
    # (a Ricker of twelve samples is all but zero: the wavelet is a smooth pulse of its own, which
    # both the plan and the oracle take as given)
    rick = O.ricker_bank(nt, 0.001, [15], [0.006], [1.0])
    rick = (np.sin(np.arange(nt) * np.pi / nt) ** 2)[None, :].astype(np.float32).reshape(rick.shape)
    d, snaps, _ = O.afd3d(vel, nt, 0.001, 10, 10, 10, 30, 0.01, [120], [120], [120], rick, jsnap=nt - 1)
    return vel, nt, rick, d, snaps


@pytest.mark.parametrize("variant", [-1, 15, 17, 24])
def test_large_grid_layout_vs_oracle(large_grid, variant):
    vel, nt, rick, d, snaps = large_grid
    assert np.abs(snaps).max() > 1e-3
    p = AfdPlan(vel, (10, 10, 10), nt, 0.001, 30, 0.01, nt - 1)
    p.set_sources(np.array([120, 120, 120, 15, 0.006, 1.0], np.float32), 1, wavelet=rick)
    if variant >= 0:
        p.set_variant(variant, 0)
    p.run()
    out = p.output()
    p.close()
    n = vel.shape[0]
    assert np.array_equal(out[:nt * n * n], d.ravel())
    assert np.array_equal(out[nt * n * n:], snaps.ravel())


def test_fields_beyond_2gib_pipelined_equals_lean():
    # 781^3 + 30-cell ABC = 841^3 updated: each field holds more than 2 GiB, so the pipelined
    # single-step kernels address one y chunk per buffer window (the two-step kernels, whose
    # windows span the whole field, do not apply); their traces equal bit for bit those of the lean
    # kernel (32-bit offsets over the whole field, pinned against the oracle on the small grids),
    # with sources just below the surface at both y ends and in the middle
    vel = plain3d_model(781)
    nt = 8
    pulse = (np.sin(np.arange(nt) * np.pi / nt) ** 2).astype(np.float32)
    src = np.array([[3, 390, 1, 15, 0.0, 1.0], [3, 100, 779, 15, 0.0, 1.0], [4, 700, 390, 15, 0.0, 1.0]],
                   np.float32)
    p = AfdPlan(vel, (10, 10, 10), nt, 0.001, 30, 0.01, 0)
    p.set_sources(src.ravel(), 3, wavelet=np.repeat(pulse[None, :], 3, axis=0))
    outs = {}
    for variant in (8, 15, 17):
        p.set_variant(variant, 0)
        p.run()
        outs[variant] = p.output().copy()
    with pytest.raises(ValueError):
        p.set_variant(24, 0)
        p.run()
    p.close()
    assert np.abs(outs[8]).max() > 1e-3
    assert np.array_equal(outs[15], outs[8])
    assert np.array_equal(outs[17], outs[8])


def test_cfg1_full_run_linear_in_amplitude():
    # Linear in the source amplitude up to gradual underflow: ahead of the wavefront the field
    # passes through the f32 denormal range, where doubling is not exact (the CPU oracle shows
    # the same 6e-7 deviation over 400 steps at 161^3), so the bar is the float tolerance.
    vel = _cfg1_model()
    nt = 1500
    p = AfdPlan(vel, (10, 10, 10), nt, 0.001, 30, 0.01, 0)
    p.set_sources(np.array([100, 100, 100, 15, 0.1, 1.0], np.float32), 1)
    p.run()
    a = p.output()
    p.set_sources(np.array([100, 100, 100, 15, 0.1, 2.0], np.float32), 1)
    p.run()
    b = p.output()
    assert np.isfinite(a).all() and np.abs(a).max() > 0
    assert rel_l2(b, 2 * a) < TOL


def test_deep_underflow_2d_vs_oracle():
    # a large grid and long run: far from the source the field grows out of the denormal range,
    # so bit-exactness here shows denormals are kept (not flushed) exactly as on the CPU
    rng = np.random.default_rng(9)
    vel = (1600 + 1800 * rng.random((300, 420))).astype(np.float32)
    nt = 700
    rick = O.ricker_bank(nt, 0.001, [18.0], [0.05], [1.0])
    d, _, _ = O.afd2d(vel, nt, 0.001, 10.0, 10.0, 30, 0.01, [30], [20], rick)
    p = AfdPlan(vel, (10.0, 10.0), nt, 0.001, 30, 0.01, 0)
    p.set_sources(np.array([30, 20, 18, 0.05, 1.0], np.float32), 1, wavelet=rick)
    p.run()
    assert np.array_equal(p.output(), d.ravel())


def test_deep_underflow_3d_vs_oracle():
    vel = plain3d_model(64, seed=5)
    nt = 220
    rick = O.ricker_bank(nt, 0.001, [15.0], [0.03], [1.0])
    d, _, _ = O.afd3d(vel, nt, 0.001, 10, 10, 10, 30, 0.01, [8], [8], [8], rick)
    p = AfdPlan(vel, (10, 10, 10), nt, 0.001, 30, 0.01, 0)
    p.set_sources(np.array([8, 8, 8, 15, 0.03, 1.0], np.float32), 1, wavelet=rick)
    p.run()
    assert np.array_equal(p.output(), d.ravel())


def test_variant_count_matches():
    from pywave_amd import _native as N
    assert N.lib().pw_afd_variant_count() == NVAR


def test_errors_follow_reference():
    vel = np.full((20, 30), 2000.0, np.float32)
    with pytest.raises(SystemError):
        pywave_amd.afd2d(vel, nt=10, ax=[0, 10, 31], az=[0, 10, 20], ns=1, sx=[5], sz=[5], f=[10], t=[0.001],
                         A=[1], nbt=5, jsnap=0)
    with pytest.raises(ValueError):  # stencil would leave the padded domain
        pywave_amd.afd2d(vel, nt=10, ax=[0, 10, 30], az=[0, 10, 20], ns=1, sx=[5], sz=[-1], f=[10],
                         t=[0.001], A=[1], nbt=0, jsnap=0)
