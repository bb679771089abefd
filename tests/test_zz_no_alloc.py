"""The allocation-free 2D modelling entry point (pw_afd2d_dev_noalloc) against the ordinary one
(pw_afd2d, the afdcfun.afd2dc drop-in): bit for bit, on two calls in a row through the C ABI, the
second one smaller than the first (the scratch array keeps the first call's contents)."""
import ctypes

import numpy as np
import pytest

from pywave_amd import _native as N

pytestmark = pytest.mark.gpu


def _ricker(nt, dt, f0, t0, A):
    w = np.zeros(nt, np.float32)
    N.check(N.lib().pw_ricker(nt, dt, f0, t0, A, N.ptr(w)), None, "ricker")
    return w


def _case(rng, nz, nx, nt, ns, nbt):
    vel = (1500 + 1500 * rng.random((nz, nx))).astype(np.float32)
    sx = rng.integers(1, nx - 1, ns)
    sz = rng.integers(1, nz - 1, ns)
    f = rng.uniform(10, 30, ns)
    t = rng.uniform(0.01, 0.05, ns)
    A = rng.uniform(-2, 2, ns)
    pack = np.concatenate([sx, sz, f, t, A]).astype(np.float32)
    return vel, pack, nbt


def _ordinary(vel, pack, ns, nt, dt, d, nbt, ct):
    nz, nx = vel.shape
    L, ctx = N.lib(), N.context()
    out = np.empty(nt * nx, np.float32)
    v = np.ascontiguousarray(vel.ravel(order="F"))
    N.check(L.pw_afd2d(ctx, N.ptr(v), N.ptr(pack), 0, nt, nx, nz, ns, 0, 0, 0, 1, nbt, ct, dt, 0.0, d, 0.0, d,
                       N.ptr(out), out.size), ctx, "pw_afd2d")
    return out


def _noalloc(vel, pack, ns, nt, dt, d, nbt, ct):
    import torch
    nz, nx = vel.shape
    L, ctx = N.lib(), N.context()
    dev = torch.device("cuda:0")
    d_vel = torch.tensor(np.ascontiguousarray(vel.ravel(order="F")), device=dev)
    wav = np.stack([_ricker(nt, dt, pack[2 * ns + i], pack[3 * ns + i], pack[4 * ns + i]) for i in range(ns)])
    d_wav = torch.tensor(np.ascontiguousarray(wav), device=dev)
    d_tr = torch.full((nt * nx,), float("nan"), dtype=torch.float32, device=dev)
    src_xz = np.ascontiguousarray(np.concatenate([pack[0:ns], pack[ns:2 * ns]]).astype(np.int32))
    torch.cuda.synchronize()
    N.check(L.pw_use_own_stream(ctx), ctx, "own stream")
    N.check(L.pw_afd2d_dev_noalloc(ctx, ctypes.c_void_p(d_vel.data_ptr()), nz, nx, d, d, nt, dt, nbt, ct, ns,
                                   src_xz.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(d_wav.data_ptr()),
                                   ctypes.c_void_p(d_tr.data_ptr())), ctx, "pw_afd2d_dev_noalloc")
    N.check(L.pw_synchronize(ctx), ctx, "sync")
    return d_tr.cpu().numpy()


def test_noalloc_equals_pw_afd2d_on_two_calls():
    pytest.importorskip("torch")
    rng = np.random.default_rng(21)
    cap = N.lib().pw_afd2d_noalloc_capacity()
    assert cap == 64 << 20
    for (nz, nx, nt, ns, nbt) in [(150, 210, 700, 3, 25), (61, 77, 300, 1, 12)]:  # the second call smaller
        vel, pack, nbt = _case(rng, nz, nx, nt, ns, nbt)
        print("noalloc case nz=%d nx=%d nt=%d ns=%d nbt=%d" % (nz, nx, nt, ns, nbt), flush=True)
        want = _ordinary(vel, pack, ns, nt, 0.001, 10.0, nbt, 0.02)
        got = _noalloc(vel, pack, ns, nt, 0.001, 10.0, nbt, 0.02)
        assert np.isfinite(want).all() and np.abs(want).max() > 0
        assert np.array_equal(got, want)


def test_noalloc_rejects_a_run_beyond_its_scratch():
    import torch
    L, ctx = N.lib(), N.context()
    dev = torch.device("cuda:0")
    nz, nx, nt = 4000, 4000, 10  # three 16M-cell fields: more than 64 MiB
    d_vel = torch.full((nz * nx,), 2000.0, device=dev)
    d_tr = torch.empty((nt * nx,), device=dev)
    rc = L.pw_afd2d_dev_noalloc(ctx, ctypes.c_void_p(d_vel.data_ptr()), nz, nx, 10.0, 10.0, nt, 0.001, 10, 0.02, 0,
                                None, None, ctypes.c_void_p(d_tr.data_ptr()))
    assert rc == -1
    assert b"scratch" in L.pw_last_error(ctx)
