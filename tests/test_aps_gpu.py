"""HIP pseudo-spectral engine (aps2d/aps3d) against the compiled reference's golden vectors and the
oracle.

Bar: bit-exact (np.array_equal, +0 == -0): the device transforms reproduce the reference's KISS FFT
arithmetic, so traces, snapshots and time-reversal fields equal the reference's, through the
resident plan fed the reference's wavelet and through the drop-in aps2d/aps3d."""
import numpy as np
import pytest

import pywave_amd
from oracle import oracle as O
from pywave_amd import AfdPlan

pytestmark = pytest.mark.gpu


def _plan(g, ndim):
    d = (float(g["dz"]), float(g["dx"])) + ((float(g["dy"]),) if ndim == 3 else ())
    return AfdPlan(g["vel"], d, int(g["nt"]), float(g["dt"]), int(g["nbt"]), float(g["ct"]), int(g["jsnap"]),
                   spectral=True)


def test_aps2d_golden_bit_exact(golden):
    g = golden("aps2d_two_sources")
    p = _plan(g, 2)
    p.set_sources(g["src"], len(g["src"]) // 5, wavelet=g["rick"])
    p.run()
    assert np.array_equal(p.output(), g["out"])
    # the drop-in, with its own wavelet
    src = g["src"]
    ns = len(src) // 5
    nz, nx = g["vel"].shape
    dout, wfd = pywave_amd.aps2d(g["vel"], nt=int(g["nt"]), dt=float(g["dt"]), ax=[0, float(g["dx"]), nx],
                                 az=[0, float(g["dz"]), nz], ns=ns, sx=src[:ns], sz=src[ns:2 * ns],
                                 f=src[2 * ns:3 * ns], t=src[3 * ns:4 * ns], A=src[4 * ns:], nbt=int(g["nbt"]),
                                 ct=float(g["ct"]), jsnap=int(g["jsnap"]), verb=False)
    nt = int(g["nt"])
    assert np.array_equal(dout, g["out"][:nt * nx].reshape(nt, nx, order="F"))
    assert np.array_equal(wfd, g["out"][nt * nx:].reshape(nz, nx, -1, order="F"))


def test_aps2d_tri_golden_bit_exact(golden):
    g = golden("aps2d_tri")
    p = _plan(g, 2)
    p.set_tri_data(g["dat"])
    p.run(tri=True)
    assert np.array_equal(p.output(), g["out"])
    nz, nx = g["vel"].shape
    nt = int(g["nt"])
    w0 = g["out"][nt * nx:nt * nx + nz * nx].reshape(nz, nx, order="F")
    assert np.array_equal(p.final_field(), w0)


def test_aps3d_golden_bit_exact(golden):
    g = golden("aps3d_small")
    p = _plan(g, 3)
    p.set_sources(g["src"], len(g["src"]) // 6, wavelet=g["rick"])
    p.run()
    assert np.array_equal(p.output(), g["out"])
    src = g["src"]
    nz, nx, ny = g["vel"].shape
    nt = int(g["nt"])
    dout, wfd = pywave_amd.aps3d(g["vel"], nt=nt, dt=float(g["dt"]), ax=[0, float(g["dx"]), nx],
                                 ay=[0, float(g["dy"]), ny], az=[0, float(g["dz"]), nz], ns=1, sx=src[:1],
                                 sy=src[1:2], sz=src[2:3], f=src[3:4], t=src[4:5], A=src[5:6], nbt=int(g["nbt"]),
                                 ct=float(g["ct"]), jsnap=int(g["jsnap"]), verb=False)
    assert np.array_equal(dout, g["out"][:nt * nx * ny].reshape(nt, nx, ny, order="F"))


@pytest.mark.parametrize("shape,nbt", [((61, 83), 20), ((30, 44), 7), ((50, 31), 12)])
def test_aps2d_other_factorisations_vs_oracle(shape, nbt):
    # padded sizes 108x125 (2^2 3^3 x 5^3), 44x60, 80x60 ... exercise every butterfly radix on both
    # axes, against the oracle's transforms
    rng = np.random.default_rng(sum(shape))
    nz, nx = shape
    vel = (1500 + 1200 * rng.random(shape)).astype(np.float32)
    nt, dt, dx, dz = 180, 0.001, 10.0, 9.0
    sx, sz = [nx // 3, 2 * nx // 3], [nz // 2, nz // 4]
    rick = np.stack([O.ricker(nt, dt, 20, 0.03, 1.0), O.ricker(nt, dt, 15, 0.05, -0.7)]).astype(np.float32)
    d, snaps, _ = O.afd2d(vel, nt, dt, dx, dz, nbt, 0.02, sx, sz, rick, jsnap=45, ps=True)
    p = AfdPlan(vel, (dz, dx), nt, dt, nbt, 0.02, 45, spectral=True)
    pack = np.array(sx + sz + [20, 15] + [0.03, 0.05] + [1.0, -0.7], np.float32)
    p.set_sources(pack, 2, wavelet=rick)
    p.run()
    out = p.output()
    assert np.array_equal(out[:nt * nx].reshape(nt, nx, order="F"), d.T)
    w = out[nt * nx:].reshape(nz, nx, -1, order="F")
    assert all(np.array_equal(w[..., k], snaps[k].T) for k in range(w.shape[-1]))


def test_aps3d_other_factorisation_vs_oracle():
    rng = np.random.default_rng(21)
    nz, nx, ny = 17, 21, 25
    vel = (1700 + 900 * rng.random((nz, nx, ny))).astype(np.float32)
    nt, dt = 60, 0.001
    rick = O.ricker(nt, dt, 25, 0.02, 1.0)[None, :].astype(np.float32)
    d, _, _ = O.afd3d(vel, nt, dt, 10.0, 11.0, 12.0, 6, 0.03, [10], [12], [8], rick, ps=True)
    p = AfdPlan(vel, (12.0, 10.0, 11.0), nt, dt, 6, 0.03, 0, spectral=True)
    p.set_sources(np.array([10, 12, 8, 25, 0.02, 1.0], np.float32), 1, wavelet=rick)
    p.run()
    out = p.output()
    assert np.array_equal(out.reshape(nt, nx, ny, order="F"), d.transpose(2, 1, 0))
