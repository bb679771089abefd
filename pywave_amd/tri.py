"""Time-reversal source-location imaging of receiver groups (3D), fused on the GPU.

The reference forms the image in its demo (demos/test_third_tri.py:57-77): each receiver group's
data is back-propagated with afd3d(tri=True, jsnap=4), the returned snapshots are kept, and the
image is  sum_snap prod_g u_g^2,  normalised by its maximum.  tri3d_image runs all groups in one
back-propagation (one launch per step reads the model once for all groups) and accumulates the
image on the fly in f64, so no snapshot ever leaves HBM.
"""
import ctypes

import numpy as np

from . import _native as N
from .stability import check_cfl

__all__ = ["tri3d_image", "group_masks"]


def group_masks(nx, ny, groups):
    """Masks [g][iy][ix] from a list of x-index arrays (all iy) or from boolean (nx, ny) masks."""
    m = np.zeros((len(groups), ny, nx), np.uint8)
    for k, gsel in enumerate(groups):
        gsel = np.asarray(gsel)
        if gsel.dtype == bool and gsel.shape == (nx, ny):
            m[k] = gsel.T
        else:
            m[k][:, gsel.astype(int)] = 1
    return m


def tri3d_image(vel, dat, groups, nt=None, dt=0.001, ax=[0, 20, 81], ay=[0, 20, 81], az=[0, 20, 81], nbt=30,
                ct=0.01, jsnap=4, normalize=True, return_plan_time=False, launch_info=None, launch=None):
    """vel (nz, nx, ny); dat (nt, nx, ny) surface data; groups: list of x-index arrays or (nx, ny)
    boolean masks (up to 64; more than 4 run in passes of 4, with each snapshot's running product
    kept in HBM).  Returns the image (nz, nx, ny) (divided by its maximum when normalize).  A dict
    passed as launch_info receives the back-propagation kernel's launch shape (rows per block, y
    chunk, non-temporal cache policy) and the timing sweep that chose it; launch=(rows, y_chunk[,
    non_temporal]) forces that shape (rows 4, 8 or 16; the image is the same bit for bit)."""
    vel = np.asarray(vel, np.float32)
    nz, nx, ny = int(az[2]), int(ax[2]), int(ay[2])
    if vel.shape != (nz, nx, ny):
        raise ValueError("vel must be (nz, nx, ny)")
    dat = np.asarray(dat, np.float32)
    nt = dat.shape[0] if nt is None else int(nt)
    if dat.shape != (nt, nx, ny):
        raise ValueError("dat must be (nt, nx, ny)")
    masks = group_masks(nx, ny, groups)
    check_cfl(vel, dt, (float(az[1]), float(ax[1]), float(ay[1])), "afd", "tri3d_image")
    L = N.lib()
    ctx = N.context()
    p = L.pw_tri_create(ctx, nz, nx, ny, float(az[1]), float(ax[1]), float(ay[1]), nt, float(dt), int(nbt),
                        float(ct), len(groups), int(jsnap))
    if not p:
        N.check(-1, ctx, "tri3d_image")
    try:
        v = np.ascontiguousarray(vel.ravel(order="F"))
        N.check(L.pw_tri_set_velocity(p, N.ptr(v)), ctx, "tri set_velocity")
        d = np.ascontiguousarray(dat.ravel(order="F"))
        mk = np.ascontiguousarray(masks)
        N.check(L.pw_tri_set_data(p, N.ptr(d), N.ptr(mk)), ctx, "tri set_data")
        if launch is not None:
            nt_ = int(launch[2]) if len(launch) > 2 else 0
            N.check(L.pw_tri_set_launch(p, int(launch[0]), int(launch[1]), nt_), ctx, "tri set_launch")
        ms = ctypes.c_double(0)
        N.check(L.pw_tri_run_timed(p, ctypes.byref(ms)), ctx, "tri run")
        img = np.empty(nz * nx * ny, np.float64)
        N.check(L.pw_tri_image(p, N.ptr(img), img.size), ctx, "tri image")
        if launch_info is not None:
            w = np.zeros(3, np.int32)
            n = L.pw_tri_launch_choice(p, N.ptr(w), None, 0)
            buf = np.zeros(4 * max(n, 1), np.float32)
            L.pw_tri_launch_choice(p, N.ptr(w), N.ptr(buf), n)
            launch_info.update(rows_per_block=int(w[0]), y_chunk=int(w[1]), non_temporal=bool(w[2]),
                               timing_ms_per_launch=[(int(buf[4 * i]), int(buf[4 * i + 1]), bool(buf[4 * i + 2]),
                                                      round(float(buf[4 * i + 3]), 4)) for i in range(n)])
    finally:
        L.pw_tri_destroy(p)
    img = img.reshape(nz, nx, ny, order="F")
    if normalize:
        img = img / img.max()
    img = img.astype(np.float32)
    return (img, ms.value) if return_plan_time else img
