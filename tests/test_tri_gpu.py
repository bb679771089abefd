"""Fused multi-group time-reversal imaging against the reference's imaging recipe.

The reference demo (demos/test_third_tri.py:57-77) back-propagates each receiver group with
afd3d(tri=True, jsnap) and forms  sum_snap prod_g u_g^2 / max.  Here the same recipe is computed
from the drop-in afd3d (bit-exact against the reference C, see test_afd_gpu.py) and from the
oracle; the fused kernel's f32 products are identical, and its f64 sum over snapshots differs from
an f64 numpy sum only by summation order (rel 1e-12)."""
import numpy as np
import pytest

import pywave_amd
from conftest import rel_l2
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _case(seed=3, n=(22, 20, 18), nt=140, ng=3):
    rng = np.random.default_rng(seed)
    nz, nx, ny = n
    vel = (1700 + 900 * rng.random(n)).astype(np.float32)
    src = np.array([6, 13, 7, 11, 10, 12, 30, 25, 0.01, 0.03, 1.0, 1.0], np.float32)  # two sources
    dout = pywave_amd.afd3d(vel, nt=nt, dt=0.001, ax=[0, 10, nx], ay=[0, 10, ny], az=[0, 10, nz], ns=2,
                            sx=src[0:2], sy=src[2:4], sz=src[4:6], f=src[6:8], t=src[8:10], A=src[10:12],
                            nbt=8, ct=0.04, jsnap=0)
    dat = dout.reshape(nt, nx, ny, order="F")
    groups = [np.arange(g, nx, ng) for g in range(ng)]
    return vel, dat, groups


def _recipe_image(vel, dat, groups, jsnap, use_oracle=False):
    nt, nx, ny = dat.shape
    nz = vel.shape[0]
    img = None
    for gsel in groups:
        dg = np.zeros_like(dat)
        dg[:, gsel, :] = dat[:, gsel, :]
        if use_oracle:
            _, snaps, _ = O.afd3d(vel, nt, 0.001, 10, 10, 10, 8, 0.04, [], [], [], None, tri=True,
                                  dat=np.ascontiguousarray(dg.transpose(2, 1, 0)), jsnap=jsnap)
            wfd = snaps.transpose(3, 2, 1, 0)  # (nz, nx, ny, nsnap)
        else:
            _, wfd = pywave_amd.afd3d(vel, nt=nt, dt=0.001, ax=[0, 10, nx], ay=[0, 10, ny], az=[0, 10, nz],
                                      nbt=8, ct=0.04, jsnap=jsnap, tri=True, dat=dg)
        img = wfd * wfd if img is None else img * wfd * wfd  # f32 products, as the demo
    return img.astype(np.float64).sum(axis=3)


@pytest.mark.parametrize("jsnap", [1, 4])
def test_tri_image_matches_recipe(jsnap):
    vel, dat, groups = _case()
    ref = _recipe_image(vel, dat, groups, jsnap)
    nz, nx, ny = vel.shape
    img = pywave_amd.tri3d_image(vel, dat, groups, dt=0.001, ax=[0, 10, nx], ay=[0, 10, ny], az=[0, 10, nz], nbt=8,
                                 ct=0.04, jsnap=jsnap, normalize=False)
    assert rel_l2(img, ref) <= 1e-6
    imn = pywave_amd.tri3d_image(vel, dat, groups, dt=0.001, ax=[0, 10, nx], ay=[0, 10, ny], az=[0, 10, nz],
                                 nbt=8, ct=0.04, jsnap=jsnap)
    assert abs(float(imn.max()) - 1.0) < 1e-6
    assert rel_l2(imn, ref / ref.max()) <= 1e-6


def test_tri_image_matches_oracle_recipe():
    vel, dat, groups = _case(seed=5, nt=100, ng=2)
    ref = _recipe_image(vel, dat, groups, 5, use_oracle=True)
    nz, nx, ny = vel.shape
    img = pywave_amd.tri3d_image(vel, dat, groups, dt=0.001, ax=[0, 10, nx], ay=[0, 10, ny], az=[0, 10, nz], nbt=8,
                                 ct=0.04, jsnap=5, normalize=False)
    assert rel_l2(img, ref) <= 1e-6


def test_tri_image_focuses_on_sources():
    # three subsurface sources: the normalised image peaks within 2 cells of one of them laterally;
    # a surface array resolves depth only to about a wavelength (15 Hz at 2000 m/s: 133 m = 13 cells)
    # and the back-propagated energy leans toward the receivers, so depth is allowed 7 cells
    nz = nx = ny = 40
    vel = np.full((nz, nx, ny), 2000.0, np.float32)
    pos = np.array([[12, 20, 24], [28, 14, 18], [22, 30, 10]])  # (z, x, y)
    nt = 320
    src = np.concatenate([pos[:, 1], pos[:, 2], pos[:, 0], [15.0] * 3, [0.02, 0.06, 0.04], [1.0] * 3]).astype(
        np.float32)
    dout = pywave_amd.afd3d(vel, nt=nt, dt=0.001, ax=[0, 10, nx], ay=[0, 10, ny], az=[0, 10, nz], ns=3,
                            sx=src[0:3], sy=src[3:6], sz=src[6:9], f=src[9:12], t=src[12:15], A=src[15:18],
                            nbt=20, ct=0.01, jsnap=0)
    dat = dout.reshape(nt, nx, ny, order="F")
    groups = [np.arange(g, nx, 4) for g in range(4)]
    img = pywave_amd.tri3d_image(vel, dat, groups, dt=0.001, ax=[0, 10, nx], ay=[0, 10, ny], az=[0, 10, nz],
                                 nbt=20, ct=0.01, jsnap=2)
    peak = np.array(np.unravel_index(np.argmax(img), img.shape))
    d = np.abs(pos - peak)
    assert np.any((d[:, 1] <= 2) & (d[:, 2] <= 2) & (d[:, 0] <= 7)), peak


def _recipe_image_f64(vel, dat, groups, jsnap):
    """The demo's recipe from the oracle's per-group back-propagations, products in f64."""
    nt, nx, ny = dat.shape
    img = None
    for gsel in groups:
        dg = np.zeros_like(dat)
        dg[:, gsel, :] = dat[:, gsel, :]
        _, snaps, _ = O.afd3d(vel, nt, 0.001, 10, 10, 10, 8, 0.04, [], [], [], None, tri=True,
                              dat=np.ascontiguousarray(dg.transpose(2, 1, 0)), jsnap=jsnap)
        u2 = snaps.transpose(3, 2, 1, 0).astype(np.float64) ** 2  # (nz, nx, ny, nsnap)
        img = u2 if img is None else img * u2
    return img.sum(axis=3)


@pytest.mark.parametrize("ng", [5, 6, 9])
def test_tri_image_more_than_four_groups(ng):
    # groups beyond four run in passes of four with each snapshot's running product kept in HBM;
    # the data are scaled so that the 2*ng-th powers stay far from f32 underflow
    vel, dat, groups = _case(seed=7, nt=120, ng=ng)
    dat = (dat / np.abs(dat).max()).astype(np.float32)
    ref = _recipe_image_f64(vel, dat, groups, 4)
    nz, nx, ny = vel.shape
    img = pywave_amd.tri3d_image(vel, dat, groups, dt=0.001, ax=[0, 10, nx], ay=[0, 10, ny], az=[0, 10, nz], nbt=8,
                                 ct=0.04, jsnap=4, normalize=False)
    print("tri %d groups: rel-L2 against the f64 recipe %.2e" % (ng, rel_l2(img, ref)))
    assert np.isfinite(img).all() and img.max() > 0
    assert rel_l2(img, ref) <= 1e-5


def test_every_launch_shape_gives_the_same_image():
    # rows per block 4, 8, 16 and several y chunks: the same per-cell arithmetic in the same order,
    # so the images are equal bit for bit; the timed choice is one of the swept shapes
    vel, dat, groups = _case(seed=11, n=(40, 36, 30), nt=60, ng=4)
    nz, nx, ny = vel.shape
    kw = dict(dt=0.001, ax=[0, 10, nx], ay=[0, 10, ny], az=[0, 10, nz], nbt=8, ct=0.04, jsnap=3, normalize=False)
    info = {}
    ref = pywave_amd.tri3d_image(vel, dat, groups, launch_info=info, **kw)
    assert info["rows_per_block"] in (4, 8, 16) and 1 <= info["y_chunk"] <= 24
    assert len(info["timing_ms_per_launch"]) == 36
    assert ((info["rows_per_block"], info["y_chunk"], info["non_temporal"]) in
            [(r, c, n) for r, c, n, _ in info["timing_ms_per_launch"]])
    for rows, chunk, nt_ in [(4, 3, 0), (8, 7, 1), (16, 5, 0), (16, 46, 1), (4, 6, 1)]:
        img = pywave_amd.tri3d_image(vel, dat, groups, launch=(rows, chunk, nt_), **kw)
        assert np.array_equal(img, ref), (rows, chunk, nt_)
    with pytest.raises(ValueError):
        pywave_amd.tri3d_image(vel, dat, groups, launch=(5, 4), **kw)
