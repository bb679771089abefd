"""Benchmark configurations at their full size, through properties that do not depend on an oracle
run of that size: the 96-shot ragged batch equals single-shot runs of the same shots bit for bit
(the single-shot path is pinned against the reference's golden vectors), and the 128^3
time-reversal image of the three-source workload focuses on a true source."""
import numpy as np
import pytest

import pywave_amd
from pywave_amd import AfdPlan, Afd2dBatch
from pywave_amd.workloads import batch2d, tri3d

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(600)
def test_batch_96_shots_equal_single_shot_runs():
    shots = batch2d()
    b = Afd2dBatch(shots)
    b.run(use_graph=True)
    for i in (0, 31, 64, 95):  # shots of different sizes and lengths
        s = shots[i]
        p = AfdPlan(s["vel"], s["d"], s["nt"], s["dt"], s["nbt"], s["ct"], 0)
        p.set_sources(s["src"], 1)
        p.run()
        nz, nx = s["vel"].shape
        single = p.output()[:s["nt"] * nx].reshape(s["nt"], nx, order="F")
        p.close()
        assert np.array_equal(b.traces(i), single), i
    b.close()


@pytest.mark.timeout(600)
def test_tri_128_image_focuses_on_a_source():
    w = tri3d()
    vel, n = w["vel"], w["vel"].shape[0]
    plan = AfdPlan(vel, w["d"], w["nt"], w["dt"], w["nbt"], w["ct"], 0)
    plan.set_sources(w["src"], w["ns"])
    plan.run()
    dat = plan.output().reshape(w["nt"], n, n, order="F")
    plan.close()
    ax = [0, w["d"][1], n]
    img = pywave_amd.tri3d_image(vel, dat, w["groups"], dt=w["dt"], ax=ax, ay=ax, az=ax, nbt=w["nbt"],
                                 ct=w["ct"], jsnap=w["jsnap"])
    assert img.shape == (n, n, n) and np.isfinite(img).all()
    src = w["src"]
    pos = np.stack([src[6:9], src[0:3], src[3:6]], axis=1)  # (z, x, y) of the three sources
    peak = np.array(np.unravel_index(np.argmax(img), img.shape))
    d = np.abs(pos - peak)
    print("tri 128^3: image peak (z, x, y) =", peak.tolist(), "sources:", pos.astype(int).tolist())
    # the product of the four groups' back-propagated fields peaks where they all focus (observed
    # one cell above a source); allowed: 2 cells laterally, 4 in depth
    assert np.any((d[:, 1] <= 2) & (d[:, 2] <= 2) & (d[:, 0] <= 4)), peak


@pytest.mark.timeout(600)
def test_plain3d_201_is_linear_in_the_source_amplitude():
    # the headline configuration (201^3 + 30-cell ABC, nt = 1500): doubling the amplitude doubles
    # every wavelet sample exactly; the traces double to within the f32 scheme's own departure
    # from exact linearity, which is the reference's too: its operation sequence (the oracle's,
    # bit-identical to the compiled reference) gives max |t(2A) - 2 t(A)| / max |t(A)| = 2e-6 on a
    # 41^3 model, from 1-ulp differences that first appear next to the absorbing corners
    from pywave_amd.workloads import plain3d
    w = plain3d()
    runs = []
    for amp in (1.0, 2.0):
        src = w["src"].copy()
        src[5] = amp
        p = AfdPlan(w["vel"], w["d"], w["nt"], w["dt"], w["nbt"], w["ct"], 0)
        p.set_sources(src, w["ns"])
        p.run()
        runs.append(p.output().astype(np.float64))
        p.close()
    t1, t2 = runs
    assert np.isfinite(t1).all() and np.abs(t1).max() > 0
    err = np.abs(t2 - 2.0 * t1).max() / np.abs(t1).max()
    print("plain3d 201^3: max |t(2A) - 2 t(A)| / max |t(A)| = %.2e" % err)
    assert err <= 1e-5  # observed 4.4e-6
