"""Write the golden vectors under tests/golden/ from the reference's own C.

The reference ships no tests, fixtures or known-answer vectors for this path, so the golden
vectors are produced here from the reference's engine sources compiled unchanged by
``make -C oracle ref`` (oracle/_ref/pwref; it exists only where the reference tree is
mounted).  Inputs are seeded synthetic models.  Each fixture stores its inputs, the
reference outputs and the reference's own wavelet(s).

Run:  make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle, refrun  # noqa: E402


def save(name, **arrays):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrays)
    print("wrote", path, os.path.getsize(path), "bytes")


def layered2d(nz, nx, vals):
    v = np.zeros((nz, nx), np.float32)
    edges = np.linspace(0, nz, len(vals) + 1).astype(int)
    for k, val in enumerate(vals):
        v[edges[k]:edges[k + 1]] = val
    return v


def afd2d_cases():
    # cfg0: three flat layers 101x101, nt=1000, source (z=5,x=50) 20 Hz at 0.1 s, jsnap=999
    vel = np.zeros((101, 101), np.float32)
    vel[:34] = 1500; vel[34:67] = 2000; vel[67:] = 2500
    src = np.array([50, 5, 20, 0.1, 1], np.float32)
    out, _ = refrun.afd2d(vel, 1000, 0.001, 10, 10, 30, 0.01, src, 1, jsnap=999)
    rick = refrun.ricker(1000, 0.001, 20, 0.1, 1)[None, :]
    save("afd2d_cfg0", vel=vel, src=src, nt=1000, dt=0.001, dx=10., dz=10., nbt=30, ct=0.01,
         jsnap=999, out=out, rick=rick)

    # two overlapping sources, uneven dims with FFT slack on both axes, fractional positions
    rng = np.random.default_rng(11)
    vel = (1500 + 1500 * rng.random((57, 83))).astype(np.float32)
    src = np.array([40.7, 41.2, 20.0, 21.9, 25, 18, 0.03, 0.05, 1.0, -2.5], np.float32)
    nt = 300
    out, _ = refrun.afd2d(vel, nt, 0.0008, 10, 9, 12, 0.03, src, 2, jsnap=23)
    rick = np.stack([refrun.ricker(nt, 0.0008, 25, 0.03, 1.0), refrun.ricker(nt, 0.0008, 18, 0.05, -2.5)])
    save("afd2d_two_sources", vel=vel, src=src, nt=nt, dt=0.0008, dx=10., dz=9., nbt=12, ct=0.03,
         jsnap=23, out=out, rick=rick)

    # padded extents equal the updated domain (no slack): the bottom/right rim has a zero Laplacian
    vel = layered2d(44, 40, [1800, 2600]).astype(np.float32)
    src = np.array([20, 0, 30, 0.02, 1], np.float32)  # source on the first model row
    nt = 250
    out, _ = refrun.afd2d(vel, nt, 0.001, 8, 8, 10, 0.02, src, 1, jsnap=0)
    rick = refrun.ricker(nt, 0.001, 30, 0.02, 1)[None, :]
    save("afd2d_noslack", vel=vel, src=src, nt=nt, dt=0.001, dx=8., dz=8., nbt=10, ct=0.02,
         jsnap=0, out=out, rick=rick)

    # time-reversal: back-propagate seeded receiver data
    vel = (1600 + 900 * rng.random((40, 50))).astype(np.float32)
    dat = rng.standard_normal((300, 50)).astype(np.float32)
    out, _ = refrun.afd2d(vel, 300, 0.001, 10, 10, 20, 0.02, tri=True, dat=dat, jsnap=7)
    save("afd2d_tri", vel=vel, dat=dat, nt=300, dt=0.001, dx=10., dz=10., nbt=20, ct=0.02, jsnap=7,
         out=out)


def afd3d_cases():
    rng = np.random.default_rng(7)
    vel = (1500 + 1000 * rng.random((30, 26, 22))).astype(np.float32)
    src = np.array([10, 15, 12, 9, 8, 7, 30, 25, 0.02, 0.05, 1, 2], np.float32)
    nt = 150
    out, _ = refrun.afd3d(vel, nt, 0.001, 10, 12, 9, 8, 0.05, src, 2, jsnap=40)
    rick = np.stack([refrun.ricker(nt, 0.001, 30, 0.02, 1), refrun.ricker(nt, 0.001, 25, 0.05, 2)])
    save("afd3d_two_sources", vel=vel, src=src, nt=nt, dt=0.001, dx=10., dy=12., dz=9., nbt=8,
         ct=0.05, jsnap=40, out=out, rick=rick)

    vel = (1700 + 800 * rng.random((24, 20, 18))).astype(np.float32)
    dat = (0.1 * rng.standard_normal((120, 20, 18))).astype(np.float32)
    out, _ = refrun.afd3d(vel, 120, 0.001, 10, 10, 10, 6, 0.05, tri=True, dat=dat, jsnap=30)
    save("afd3d_tri", vel=vel, dat=dat, nt=120, dt=0.001, dx=10., dy=10., dz=10., nbt=6, ct=0.05,
         jsnap=30, out=out)


def aps_cases():
    """Pseudo-spectral engine (apscfun: pywave/src/aps.c -> psm2d/psm, pywave/src/wave_psp.c).
    Padded sizes with radix-2, 3, 4 and 5 factors on every axis."""
    rng = np.random.default_rng(13)
    # 46x57 + 20: 86x97 -> padded 90x100 (2*3*3*5 x 2*2*5*5); two sources, snapshots
    vel = layered2d(46, 57, [1600, 2100, 2700]).astype(np.float32)
    vel = (vel * (1 + 0.05 * rng.standard_normal(vel.shape))).astype(np.float32)
    src = np.array([20.4, 40.0, 15.0, 30.6, 22, 16, 0.03, 0.06, 1.0, -1.5], np.float32)
    nt = 400
    out, _ = refrun.afd2d(vel, nt, 0.001, 10, 10, 20, 0.02, src, 2, jsnap=50, ps=True)
    rick = np.stack([refrun.ricker(nt, 0.001, 22, 0.03, 1.0), refrun.ricker(nt, 0.001, 16, 0.06, -1.5)])
    save("aps2d_two_sources", vel=vel, src=src, nt=nt, dt=0.001, dx=10., dz=10., nbt=20, ct=0.02,
         jsnap=50, out=out, rick=rick)

    # time reversal through the spectral engine
    vel = (1700 + 700 * rng.random((40, 50))).astype(np.float32)
    dat = rng.standard_normal((250, 50)).astype(np.float32)
    out, _ = refrun.afd2d(vel, 250, 0.001, 10, 10, 20, 0.02, tri=True, dat=dat, jsnap=60, ps=True)
    save("aps2d_tri", vel=vel, dat=dat, nt=250, dt=0.001, dx=10., dz=10., nbt=20, ct=0.02, jsnap=60,
         out=out)

    # 3D: 20x22x18 + 8 -> 36x38x34 -> padded 36x40x36
    vel = (1800 + 600 * rng.random((20, 22, 18))).astype(np.float32)
    src = np.array([10, 9, 8, 30, 0.03, 1.0], np.float32)
    nt = 100
    out, _ = refrun.afd3d(vel, nt, 0.001, 10, 12, 9, 8, 0.03, src, 1, jsnap=40, ps=True)
    rick = refrun.ricker(nt, 0.001, 30, 0.03, 1.0)[None, :]
    save("aps3d_small", vel=vel, src=src, nt=nt, dt=0.001, dx=10., dy=12., dz=9., nbt=8, ct=0.03,
         jsnap=40, out=out, rick=rick)


def ricker_cases():
    cases = [(1000, 0.001, 20, 0.1, 1), (1500, 0.001, 15, 0.1, 1), (2000, 0.001, 10, 0.25, 1),
             (3000, 0.001, 8, 0.15, 1e6), (333, 0.002, 12, 0.35, -3), (1501, 0.001, 10, 0.2, 2)]
    waves = [refrun.ricker(*c) for c in cases]
    n = max(len(w) for w in waves)
    arr = np.zeros((len(cases), n), np.float32)
    for i, w in enumerate(waves):
        arr[i, :len(w)] = w
    save("ricker", params=np.array(cases, np.float64), waves=arr)


def fwi_cases():
    rng = np.random.default_rng(4)
    nz, nx, nt = 40, 60, 500
    vtrue = np.full((nz, nx), 2.0, np.float32)
    vtrue[15:] = 2.5
    vtrue[28:] = 3.0
    vtrue += (0.1 * rng.standard_normal((nz, nx))).astype(np.float32)
    vtrue = vtrue.astype(np.float32)
    # smooth start model: separable running mean (radius 3), computed in float64
    k = np.ones(7) / 7.0
    vs = np.apply_along_axis(lambda c: np.convolve(np.pad(c, 3, mode="edge"), k, "valid"), 0, vtrue)
    vs = np.apply_along_axis(lambda r: np.convolve(np.pad(r, 3, mode="edge"), k, "valid"), 1, vs)
    v0 = vs.astype(np.float32)
    q = np.full((nz, nx), 1e4, np.float32)
    wav = (1e6 * refrun.ricker(nt, 0.001, 10, 0.1, 1)).astype(np.float32)
    par = np.array([nz, nx, 0.01, 0.01, 0, 0, nt, 0.001, 0, 0, 3, 0.2, 0.05, 5, 20, 0.01, 10, 1, 1, 3],
                   np.float32)
    data, _ = refrun.fwimod(par, vtrue, q, wav)
    grads, fc, _ = refrun.fwigrad(par, v0, q, wav, data, vel_second=(v0 * 1.01).astype(np.float32))
    par4 = par.copy(); par4[18] = 4
    grads4, fc4, _ = refrun.fwigrad(par4, v0, q, wav, data)
    vinv, _ = refrun.fwi(par, v0, q, wav, data)
    save("fwi_small", par=par, par4=par4, vtrue=vtrue, v0=v0, q=q, wav=wav, data=data,
         grad1=grads[0], grad2=grads[1], fcost=fc, grad_int4=grads4[0], fcost_int4=fc4, vinv=vinv)

    # visco-acoustic modelling (forward_modeling, the media=2 path) on the same acquisition
    qv = (20 + 40 * np.random.default_rng(17).random((nz, nx))).astype(np.float32)
    data_visco, _ = refrun.fwimod(par, vtrue, qv, wav, visco=True)
    # its gradient (gradient_v: the gradient fwi() takes for media != 1), two calls, interval 1 and
    # 4, and the fwi() trajectory with it
    gv, fcv, _ = refrun.fwigrad(par, v0, qv, wav, data_visco, vel_second=(v0 * 1.01).astype(np.float32),
                                visco=True)
    gv4, fcv4, _ = refrun.fwigrad(par4, v0, qv, wav, data_visco, visco=True)
    vinv_v, _ = refrun.fwi(par, v0, qv, wav, data_visco, visco=True)
    save("fwi_visco", par=par, par4=par4, vtrue=vtrue, v0=v0, q=qv, wav=wav, data=data_visco, grad1=gv[0],
         grad2=gv[1], fcost=fcv, grad_int4=gv4[0], fcost_int4=fcv4, vinv=vinv_v)

    # marine-style acquisition (acqui_type 2), sources marching across the model
    par2 = np.array([30, 48, 0.01, 0.01, 0, 0, 300, 0.001, 0, 0, 2, 0.15, 0.0, 5, 16, 0.01, 10, 2, 1, 1],
                    np.float32)
    v2 = (2.2 + 0.4 * rng.random((30, 48))).astype(np.float32)
    q2 = np.full((30, 48), 1e4, np.float32)
    wav2 = (1e6 * refrun.ricker(300, 0.001, 12, 0.06, 1)).astype(np.float32)
    data2, _ = refrun.fwimod(par2, v2, q2, wav2)
    # and its visco-acoustic modelling (forward_modeling) with a Q of 15-45
    q2v = (15 + 30 * np.random.default_rng(23).random((30, 48))).astype(np.float32)
    data2v, _ = refrun.fwimod(par2, v2, q2v, wav2, visco=True)
    save("fwi_marine", par=par2, vel=v2, wav=wav2, data=data2, q_visco=q2v, data_visco=data2v)


def passive_field(nz, nx, nt, events, dt=0.001):
    """Space-time source field (nz, nx, nt, ns): each event is a list of buried point sources
    (iz, ix, f0, t0, amp), each radiating a Ricker wavelet."""
    src = np.zeros((nz, nx, nt, len(events)), np.float32)
    for e, pts in enumerate(events):
        for iz, ix, f0, t0, amp in pts:
            src[iz, ix, :, e] += (1e6 * refrun.ricker(nt, dt, f0, t0, amp)).astype(np.float32)
    return src


def passive_cases():
    """pfwi mode 4 (passive, velocity only, known source field): pywave/src/pfwi.c:1507-2034."""
    rng = np.random.default_rng(11)
    nz, nx, nt = 30, 40, 400
    vtrue = np.full((nz, nx), 2.0, np.float32)
    vtrue[12:] = 2.4
    vtrue[22:] = 2.8
    vtrue = (vtrue + 0.05 * rng.standard_normal((nz, nx))).astype(np.float32)
    v0 = np.repeat(np.linspace(2.0, 2.8, nz, dtype=np.float32)[:, None], nx, 1).astype(np.float32)
    q = np.full((nz, nx), 1e4, np.float32)
    wav = (1e6 * refrun.ricker(nt, 0.001, 10, 0.1, 1)).astype(np.float32)
    events = [[(24, 9, 15, 0.05, 1.0), (20, 27, 12, 0.08, -0.7)],
              [(26, 33, 14, 0.04, 0.8), (17, 15, 11, 0.1, 1.2), (25, 21, 16, 0.06, -0.5)]]
    src = passive_field(nz, nx, nt, events)
    par = np.array([nz, nx, 0.01, 0.01, 0, 0, nt, 0.001, 0, 1, 2, 0.2, 0.05, 5, 16, 0.01, 10, 1, 1, 7],
                   np.float32)
    # observed data: the oracle's passive forward at the true model (pinned below by a zero misfit
    # of the reference's own gradient_pas_av at that model)
    data = oracle.fwi_forward(par, vtrue, wav, src=src)
    _, fc_true, _ = refrun.pfwigrad(par, vtrue, q, wav, data, src)
    assert fc_true[0] == 0.0, fc_true
    grads, fc, _ = refrun.pfwigrad(par, v0, q, wav, data, src, vel_second=(v0 * 1.01).astype(np.float32))
    vinv, _ = refrun.pfwi(par, v0, q, wav, data, src)
    save("pfwi_small", par=par, vtrue=vtrue, v0=v0, q=q, wav=wav, src=src, data=data,
         grad1=grads[0], grad2=grads[1], fcost=fc, vinv=vinv)


def lstri_cases():
    """pfwi mode 4 lstric: passive modelling (inv=0) and least-squares time-reversal source
    inversion (inv=1, onlysrc) through the tri2d propagator (pywave/src/wave_triutil.c:14-180,
    wave_fwi.c:121-210, wave_fwigradient.c:454-525)."""
    nz, nx, nt = 16, 24, 120
    vel = np.full((nz, nx), 2.0, np.float32)
    vel[6:] = 2.3
    vel[11:] = 2.7
    q = np.full((nz, nx), 1e4, np.float32)
    wav = np.zeros(nt, np.float32)  # unused by lstric
    events = [[(12, 7, 25, 0.03, 1.0), (9, 16, 20, 0.05, -0.6)],
              [(13, 18, 22, 0.04, 0.9)]]
    src = passive_field(nz, nx, nt, events)
    par = np.array([nz, nx, 0.01, 0.01, 0, 0, nt, 0.001, 0, 0, 2, 0.2, 0.05, 5, 8, 0.02, 10, 1, 1, 6],
                   np.float32)
    data, _ = refrun.lstri(par, vel, q, wav, src)
    par_inv = par.copy()
    par_inv[9] = 1
    src_inv, _ = refrun.lstri(par_inv, vel, q, wav, data)
    save("lstri_small", par=par, par_inv=par_inv, vel=vel, src=src, data=data, src_inv=src_inv)


if __name__ == "__main__":
    if not refrun.available():
        sys.exit("oracle/_ref/pwref missing: run `make -C oracle ref` where the reference tree exists")
    ricker_cases()
    afd2d_cases()
    afd3d_cases()
    fwi_cases()
    passive_cases()
    lstri_cases()
    aps_cases()
