"""Seeded synthetic models for the benchmark workloads (no external data is needed).

  plain2d    2D 101x101, three flat layers 1500/2000/2500 m/s, nt=1000, 20 Hz source at (5,50)
  plain3d    3D 201^3, 10 m, v = 1500 m/s at the top rising linearly to 3500 m/s at the bottom,
             times 8 seeded Gaussian perturbations (sigma 100 m, +-10%); nt=1500, 15 Hz source
             at the centre, surface receivers 201x201, 30-cell absorbing boundary
  tri3d      128^3 version of plain3d, three seeded sources (10 Hz, onsets in 0-0.3 s), nt=2000,
             4 interleaved receiver groups
  batch2d    96 shots of uneven size nz in [200,600], nx in [400,1600], nt in [2000,4000], dipping
             layered velocities 1500-4500 m/s, 12 Hz surface sources
  fwi2d      400x800 at 10 m (km units), 4 layers 1.5-3.5 km/s plus 6 Gaussian anomalies; start
             model smoothed with a 200 m Gaussian; 32 shots, nt=3000, 8 Hz Ricker at 0.15 s
  lstri2d    the geometry of the reference's passive demos (demos/test_pfwi_mod2d.py): 130x334 at
             8 m x 12 m (km units), 60-cell decay boundary, nt=3001, 3 events of 6 buried point
             sources each (10 Hz Ricker, amplitude 1e6, onsets 0.1-1.1 s at the demo's positions),
             30 CG iterations; 5 layers 1.5-4.0 km/s plus 5 Gaussian anomalies in place of the
             demo's Marmousi cut (the model file is external data)
"""
import numpy as np


def _gauss1d(n, c, sigma):
    i = np.arange(n, dtype=np.float64)
    return np.exp(-((i - c) ** 2) / (2.0 * sigma ** 2))


def plain3d_model(n=201, seed=1, nblob=8, sigma_cells=10.0):
    """Velocity (nz, nx, ny) float32 of the 3D workload."""
    rng = np.random.default_rng(seed)
    z = np.arange(n, dtype=np.float64)
    base = 1500.0 + 2000.0 * z / (n - 1)
    pert = np.zeros((n, n, n), np.float64)
    for _ in range(nblob):
        c = rng.uniform(0, n - 1, 3)
        a = rng.uniform(-0.1, 0.1)
        pert += a * (_gauss1d(n, c[0], sigma_cells)[:, None, None] * _gauss1d(n, c[1], sigma_cells)[None, :, None]
                     * _gauss1d(n, c[2], sigma_cells)[None, None, :])
    np.clip(pert, -0.1, 0.1, out=pert)
    return (base[:, None, None] * (1.0 + pert)).astype(np.float32)


def plain3d(n=201, nt=1500):
    vel = plain3d_model(n)
    c = n // 2
    return dict(vel=vel, d=(10.0, 10.0, 10.0), nt=nt, dt=0.001, nbt=30, ct=0.01,
                src=np.array([c, c, c, 15.0, 0.1, 1.0], np.float32), ns=1)


def plain2d():
    vel = np.zeros((101, 101), np.float32)
    vel[:34] = 1500
    vel[34:67] = 2000
    vel[67:] = 2500
    return dict(vel=vel, d=(10.0, 10.0), nt=1000, dt=0.001, nbt=30, ct=0.01,
                src=np.array([50, 5, 20.0, 0.1, 1.0], np.float32), ns=1, jsnap=999)


def tri3d(n=128, nt=2000, seed=2, ngroups=4):
    rng = np.random.default_rng(seed)
    vel = plain3d_model(n, seed=seed)
    pos = rng.integers(20, 108, (3, 3)).astype(np.float32)  # rows: sx, sy, sz
    t0 = rng.uniform(0.0, 0.3, 3).astype(np.float32)
    src = np.concatenate([pos[0], pos[1], pos[2], np.full(3, 10.0, np.float32), t0,
                          np.ones(3, np.float32)]).astype(np.float32)
    groups = [np.arange(g, n, ngroups) for g in range(ngroups)]
    return dict(vel=vel, d=(10.0, 10.0, 10.0), nt=nt, dt=0.001, nbt=30, ct=0.01, src=src, ns=3,
                groups=groups, jsnap=4)


def dipping_layers(nz, nx, rng, vmin=1500.0, vmax=4500.0):
    nint = int(rng.integers(3, 7))
    vals = np.sort(rng.uniform(vmin, vmax, nint + 1))
    depth0 = np.sort(rng.uniform(0.1, 0.95, nint)) * nz
    dips = np.tan(np.deg2rad(rng.uniform(-15, 15, nint)))
    x = np.arange(nx, dtype=np.float64) - nx / 2
    z = np.arange(nz, dtype=np.float64)[:, None]
    v = np.full((nz, nx), vals[0])
    for k in range(nint):
        v = np.where(z >= depth0[k] + dips[k] * x[None, :], vals[k + 1], v)
    return v.astype(np.float32)


def batch2d(nshot=96, seed=3):
    rng = np.random.default_rng(seed)
    shots = []
    for _ in range(nshot):
        nz = int(rng.integers(200, 601))
        nx = int(rng.integers(400, 1601))
        nt = int(rng.integers(2000, 4001))
        vel = dipping_layers(nz, nx, rng)
        shots.append(dict(vel=vel, d=(10.0, 10.0), nt=nt, dt=0.001, nbt=30, ct=0.01,
                          src=np.array([nx // 2, 2, 12.0, 0.1, 1.0], np.float32), ns=1))
    return shots


def smooth_gaussian(v, sigma):
    """Separable Gaussian smoothing with edge replication (float64)."""
    r = int(np.ceil(3 * sigma))
    k = np.exp(-0.5 * (np.arange(-r, r + 1) / sigma) ** 2)
    k /= k.sum()
    out = np.asarray(v, np.float64)
    for ax in range(out.ndim):
        out = np.apply_along_axis(lambda s: np.convolve(np.pad(s, r, mode="edge"), k, "valid"), ax, out)
    return out


def fwi2d(nz=400, nx=800, ns=32, nt=3000, seed=4):
    """True/start models (km/s) and the pfwi parameter set of the FWI workload."""
    rng = np.random.default_rng(seed)
    v = np.zeros((nz, nx), np.float64)
    edges = np.linspace(0, nz, 5).astype(int)
    for k, val in enumerate([1.5, 2.2, 2.8, 3.5]):
        v[edges[k]:edges[k + 1]] = val
    zz, xx = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
    for _ in range(6):
        cz, cx = rng.uniform(0.15 * nz, 0.9 * nz), rng.uniform(0.1 * nx, 0.9 * nx)
        s = rng.uniform(8, 15)  # 80-150 m
        a = rng.uniform(-0.08, 0.08)
        v *= 1 + a * np.exp(-((zz - cz) ** 2 + (xx - cx) ** 2) / (2 * s * s))
    vtrue = v.astype(np.float32)
    v0 = smooth_gaussian(vtrue, 20.0).astype(np.float32)
    par = dict(nz=nz, nx=nx, dz=0.01, dx=0.01, oz=0, ox=0, ns=ns, ds=0.24, os=0.2, sz=5, nt=nt, dt=0.001, ot=0,
               nb=60, coef=0.005, acqui_type=1, interval=1, f0=8, inv=0, niter=5)
    return dict(vtrue=vtrue, v0=v0, par=par, f0=8.0, t0=0.15, amp=1e6)


def ricker_np(nt, dt, f0, t0, amp=1.0):
    """Analytic Ricker wavelet (1 - 2 a) exp(-a), a = (pi f0 (t - t0))^2."""
    t = np.arange(nt, dtype=np.float64) * dt - t0
    a = (np.pi * f0 * t) ** 2
    return amp * (1.0 - 2.0 * a) * np.exp(-a)


def lstri2d(nz=130, nx=334, nt=3001, seed=5):
    """Velocity (km/s), source field (nz, nx, nt, 3) and the pfwi parameter set of the passive
    workload."""
    rng = np.random.default_rng(seed)
    v = np.zeros((nz, nx), np.float64)
    edges = np.linspace(0, nz, 6).astype(int)
    for k, val in enumerate([1.5, 2.0, 2.6, 3.2, 4.0]):
        v[edges[k]:edges[k + 1]] = val
    zz, xx = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
    for _ in range(5):
        cz, cx = rng.uniform(0.2 * nz, 0.9 * nz), rng.uniform(0.1 * nx, 0.9 * nx)
        sg = rng.uniform(6, 12)
        a = rng.uniform(-0.06, 0.06)
        v *= 1 + a * np.exp(-((zz - cz) ** 2 + (xx - cx) ** 2) / (2 * sg * sg))
    vel = np.clip(v, 1.5, 4.0).astype(np.float32)
    # the demo's event positions (z, x, onset sample) of demos/test_pfwi_mod2d.py
    ev = [[(110, 100, 100), (120, 110, 300), (117, 165, 500), (120, 205, 700), (112, 220, 900), (115, 260, 1100)],
          [(115, 95, 100), (111, 120, 300), (120, 170, 500), (122, 200, 700), (118, 230, 900), (115, 255, 1100)],
          [(112, 90, 100), (118, 110, 300), (111, 145, 500), (120, 180, 700), (116, 200, 900), (110, 240, 1100)]]
    dt = 0.001
    src = np.zeros((nz, nx, nt, len(ev)), np.float32)
    for e, pts in enumerate(ev):
        for iz, ix, it in pts:
            src[iz, ix, :, e] += ricker_np(nt, dt, 10.0, it * dt, 1e6).astype(np.float32)
    par = dict(nz=nz, nx=nx, dz=0.008, dx=0.012, oz=0, ox=0, ns=len(ev), ds=0.2, os=0.0, sz=5, nt=nt, dt=dt, ot=0,
               nb=60, coef=0.005, acqui_type=1, interval=1, f0=10, inv=0, niter=30)
    return dict(vel=vel, src=src, par=par)
