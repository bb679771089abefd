"""The benchmark workloads at their full sizes against the reference's own C run at the same
sizes (fixtures written by tests/golden/make_fullsize.py from the compiled reference): the
SHA-256 of each complete f32 output, samples of it and per-step reductions.

  201^3 3D modelling (nt = 1500)          bit-exact (one shot, the reference's operation order)
  four shots of the 96-shot ragged batch  bit-exact
  400x800 FWI, 32 shots, nt = 3000        forward data bit-exact; gradient_av one shot per launch
                                          bit-exact, 32 shots per launch rel-L2 <= 2e-4 (the
                                          per-shot gradients are summed in another order); the
                                          reference's f32 misfit accumulation reproduced exactly
  its 5-iteration L-BFGS (fwi())          each stored model within 1e-3 km/s of the reference's
                                          (f64 dot products here, f32 there), misfit decreasing
"""
import hashlib
import os

import numpy as np
import pytest

from conftest import GOLDEN, rel_l2
from oracle import oracle as O
from conftest import load_golden as _load
from pywave_amd import AfdPlan, Afd2dBatch, FwiProblem
from pywave_amd import _native as N
from pywave_amd import workloads as W

pytestmark = pytest.mark.gpu


def load_golden(name):
    # the fixtures come from hour-long runs of the reference C (tests/golden/make_fullsize.py)
    if not os.path.exists(os.path.join(GOLDEN, name + ".npz")):
        pytest.skip("fixture %s not generated (python tests/golden/make_fullsize.py)" % name)
    return _load(name)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, np.float32).tobytes()).hexdigest()


@pytest.mark.timeout(300)
def test_plain3d_201_bit_exact_against_reference():
    ref = load_golden("full_plain3d_201")
    w = W.plain3d()
    n, nt = w["vel"].shape[0], w["nt"]
    p = AfdPlan(w["vel"], w["d"], w["nt"], w["dt"], w["nbt"], w["ct"], 0)
    p.set_sources(w["src"], w["ns"])
    p.run()
    out = p.output()
    p.close()
    tr = out.reshape(nt, n, n, order="F")
    e = rel_l2(tr[:, ::20, ::20], ref["traces"])
    es = float(np.max(np.abs(np.sum(tr.astype(np.float64) ** 2, axis=(1, 2)) - ref["sumsq_t"])) /
               np.max(ref["sumsq_t"]))
    print("plain3d 201^3 nt=1500: sampled traces rel-L2 %.2e, per-step energy %.2e, sha256 %s" %
          (e, es, "equal" if sha(out) == str(ref["sha256"]) else "differs"))
    assert e <= 1e-5 and es <= 1e-5
    assert sha(out) == str(ref["sha256"])


@pytest.mark.timeout(300)
def test_plain3d_201_error_against_fp64_twin():
    """The f32 traces' error against the FP64 twin of the scheme (oracle/orc_afd64.c) run on the
    same model and wavelet; the GPU's error equals the reference's own (bit-exact traces)."""
    ref = load_golden("full_plain3d_201")
    f64 = load_golden("full_plain3d_201_f64")
    w = W.plain3d()
    n, nt = w["vel"].shape[0], w["nt"]
    p = AfdPlan(w["vel"], w["d"], w["nt"], w["dt"], w["nbt"], w["ct"], 0)
    p.set_sources(w["src"], w["ns"])
    p.run()
    tr = p.output().reshape(nt, n, n, order="F")
    p.close()
    e_gpu = rel_l2(tr[:, ::20, ::20], f64["traces"])
    e_ref = rel_l2(ref["traces"], f64["traces"])
    print("plain3d 201^3: f32 traces vs FP64 twin rel-L2: GPU %.3e, reference C %.3e" % (e_gpu, e_ref))
    assert e_gpu <= 1e-4
    assert abs(e_gpu - e_ref) <= 1e-3 * e_ref


@pytest.mark.timeout(600)
def test_batch2d_shots_bit_exact_against_reference():
    ref = load_golden("full_batch2d_96")
    shots = W.batch2d()
    b = Afd2dBatch(shots)
    b.run(use_graph=True)
    for i in ref["shots"]:
        i = int(i)
        tr = b.traces(i)
        e = rel_l2(tr[:, ref["cols_%d" % i]], ref["traces_%d" % i])
        print("batch2d shot %d: sampled traces rel-L2 %.2e, sha256 %s" %
              (i, e, "equal" if sha(tr.ravel(order="F")) == str(ref["sha256_%d" % i]) else "differs"))
        assert e <= 1e-5
        assert sha(tr.ravel(order="F")) == str(ref["sha256_%d" % i]), i
    b.close()


def _fwi_setup():
    ref = load_golden("full_fwi2d_grad")
    w = W.fwi2d()
    par = ref["par"]
    nt = int(par[6])
    wav = np.zeros(nt, np.float32)
    N.check(N.lib().pw_ricker(nt, float(par[7]), w["f0"], w["t0"], 1.0, N.ptr(wav)), None, "ricker")
    wav = (w["amp"] * wav).astype(np.float32)
    assert np.array_equal(wav, ref["wav"])  # the reference's wavelet, bit for bit
    return ref, w, par, wav


@pytest.mark.timeout(600)
def test_fwi2d_forward_and_gradient_against_reference():
    ref, w, par, wav = _fwi_setup()
    ns = int(par[10])
    gen = FwiProblem(par, wav, batch=ns)
    data = gen.forward(w["vtrue"])
    gen.close()
    e = rel_l2(data[:, ::100, ::8], ref["data_sel"])
    same = sha(data) == str(ref["sha256_data"])  # (nt, nr, ns) as the fixture hashed it
    print("fwi2d forward, 32 shots: sampled data rel-L2 %.2e, sha256 %s" % (e, "equal" if same else "differs"))
    assert e <= 1e-5 and same
    # gradient_av one shot per launch: the reference's summation order
    p = FwiProblem(par, wav, data=data, batch=1)
    g1, f1, _ = p.gradient(w["v0"], "reference")
    p.close()
    e1 = rel_l2(g1, ref["grad"])
    # the reference's misfit is an f32 accumulation over 77 M samples: reproduced bit for bit by
    # the oracle's restatement of that sum from the GPU's predicted data, while the engine's f64
    # sum is the accurate value (about 2 % above it here)
    gen = FwiProblem(par, wav, batch=ns)
    pred = gen.forward(w["v0"])
    gen.close()
    f32 = O.fcost_f32(data, pred)
    print("fwi2d gradient_av, one shot per launch: rel-L2 %.2e, sha256 %s; misfit f64 %.6f, reference %.6f (its f32 "
          "accumulation restated: %.6f)" % (e1, "equal" if sha(g1) == str(ref["sha256_grad"]) else "differs", f1,
                                            float(ref["fcost"][0]), f32))
    assert sha(g1) == str(ref["sha256_grad"])
    assert np.float32(f32) == ref["fcost"][0]
    r = (data - pred).astype(np.float64)  # the residual in f32, as the engine forms it
    assert abs(f1 - 0.5 * float(np.sum(r * r))) <= 1e-7 * f1
    # all 32 shots in one launch, checkpoints every 50 steps: each shot's imaging terms are summed in
    # its own f32 accumulator and the 32 sums added in shot order, where the reference adds every
    # term of every shot to one f32 accumulator; the f32 reassociation of about 96,000 terms per
    # cell (which cancel heavily) is the whole difference: 16 shots per launch, another
    # association, lands as far from the reference
    q = FwiProblem(par, wav, data=data, batch=ns, checkpoint=50)
    g32, _, _ = q.gradient(w["v0"], "reference")
    q.close()
    q = FwiProblem(par, wav, data=data, batch=ns // 2, checkpoint=50)
    g16, _, _ = q.gradient(w["v0"], "reference")
    q.close()
    e32, e16 = rel_l2(g32, ref["grad"]), rel_l2(g16, ref["grad"])
    print("fwi2d gradient_av, 32 shots per launch: rel-L2 %.2e; 16 shots per launch: %.2e; between the two: %.2e"
          % (e32, e16, rel_l2(g32, g16)))
    assert e32 <= 2e-4 and e16 <= 2e-4


@pytest.mark.timeout(900)
def test_fwi2d_lbfgs_five_iterations_against_reference():
    ref, w, par, wav = _fwi_setup()
    lb = load_golden("full_fwi2d_lbfgs")
    ns = int(par[10])
    gen = FwiProblem(par, wav, batch=ns)
    data = gen.forward(w["vtrue"])
    gen.close()
    p = FwiProblem(par, wav, data=data, batch=ns, checkpoint=50)
    niter = int(lb["niter"])
    vinv, k = p.lbfgs(w["v0"], niter)
    assert k == niter
    misfits = [p.misfit(w["v0"])] + [p.misfit(vinv[:, :, i]) for i in range(niter)]
    p.close()
    diffs = [float(np.abs(vinv[::4, ::4, i] - lb["vinv_sel"][:, :, i]).max()) for i in range(niter)]
    print("fwi2d L-BFGS, 5 iterations: misfit %s; max |v - v_ref| per iterate (km/s) %s" %
          (["%.6g" % m for m in misfits], ["%.1e" % d for d in diffs]))
    assert all(misfits[i + 1] < misfits[i] for i in range(niter))
    assert max(diffs) <= 1e-3
    assert float(np.abs(vinv[:, :, -1] - lb["vinv_last"]).max()) <= 1e-3
