"""HIP FWI engine against the compiled reference's golden vectors and the mathematics.

Bars: forward modelling and the reference imaging-condition gradient (one shot per launch) are
bit-exact against the reference C; batched launches change only the order per-shot gradients
are summed (rel-L2 <= 1e-5); the exact adjoint matches central finite differences in f64 to
1e-3 relative (observed ~1e-7), and the autograd path matches the direct gradient."""
import numpy as np
import pytest

import pywave_amd
from conftest import rel_l2
from pywave_amd import FwiProblem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small(request):
    from conftest import load_golden
    return load_golden("fwi_small")


@pytest.mark.parametrize("batch", [1, 2, 3])
def test_forward_bit_exact(small, batch):
    p = FwiProblem(small["par"], small["wav"], batch=batch)
    d = p.forward(small["vtrue"])
    assert d.shape == small["data"].shape
    assert np.array_equal(d, small["data"])


def test_forward_marine_geometry(golden):
    m = golden("fwi_marine")
    p = FwiProblem(m["par"], m["wav"], batch=2)
    assert np.array_equal(p.forward(m["vel"]), m["data"])
    # visco-acoustic modelling on the same moving-spread acquisition
    p.set_visco(m["q_visco"])
    assert np.array_equal(p.forward(m["vel"]), m["data_visco"])


@pytest.mark.parametrize("ckpt", [0, 7, 50, 130])
def test_reference_gradient_bit_exact(small, ckpt):
    p = FwiProblem(small["par"], small["wav"], data=small["data"], batch=1, checkpoint=ckpt)
    g1, f1, sc = p.gradient(small["v0"], "reference")
    assert np.array_equal(g1, small["grad1"])
    assert abs(f1 - float(small["fcost"][0])) <= 1e-3 * abs(f1)  # reference accumulates in f32
    # later calls reuse the first call's scaling (wave_fwigradient.c:245-252)
    g2, f2, _ = p.gradient((small["v0"] * 1.01).astype(np.float32), "reference", scaling=sc)
    assert np.array_equal(g2, small["grad2"])


def test_reference_gradient_interval4(small):
    p = FwiProblem(small["par4"], small["wav"], data=small["data"], batch=1, checkpoint=12)
    g4, _, _ = p.gradient(small["v0"], "reference")
    assert np.array_equal(g4, small["grad_int4"])


def test_reference_gradient_batched(small):
    p = FwiProblem(small["par"], small["wav"], data=small["data"], batch=3, checkpoint=40)
    g, f, _ = p.gradient(small["v0"], "reference")
    assert rel_l2(g, small["grad1"]) <= 1e-5


def test_pfwi_dropin_modes(small):
    par = dict(zip(["nz", "nx", "dz", "dx", "oz", "ox", "nt", "dt", "ot", "inv", "ns", "ds", "os", "sz", "nb",
                    "coef", "f0", "acqui_type", "interval", "niter"], [float(x) for x in small["par"]]))
    for k in ["nz", "nx", "nt", "ns", "nb", "acqui_type", "interval", "niter"]:
        par[k] = int(par[k])
    q = small["q"]
    data, vinv, grad, src, mwt = pywave_amd.pfwi(small["vtrue"], q, small["wav"], None, mode=1, par=par)
    assert np.array_equal(data, small["data"]) and vinv == [] and grad == []
    d2, vinv, grad, src, mwt = pywave_amd.pfwi(small["v0"], q, small["wav"], None, data=small["data"], mode=2,
                                               par=par)
    assert grad is None and src is None and vinv.shape == (40, 60, 3)
    ref = small["vinv"]
    # f64 dot products in the optimiser: same trajectory to f32 round-off of the models
    assert np.abs(vinv - ref).max() < 1e-4


def test_visco_forward_bit_exact(golden):
    # forward_modeling (wave_fwimodeling.c:116-220): memory variable r, tau/taus from Q and f0
    g = golden("fwi_visco")
    for batch in (1, 0):
        p = FwiProblem(g["par"], g["wav"], data=g["data"], batch=batch)
        p.set_visco(g["q"])
        assert np.array_equal(p.forward(g["vtrue"]), g["data"]), batch
        # back to acoustic on the same plan, and visco again
        p.set_visco(None)
        ac = p.forward(g["vtrue"])
        assert rel_l2(ac, g["data"]) > 0.05
        p.set_visco(g["q"])
        assert np.array_equal(p.forward(g["vtrue"]), g["data"]), batch
    with pytest.raises(ValueError, match="visco"):
        p.gradient(g["vtrue"], "exact")


@pytest.mark.parametrize("ckpt", [0, 7, 130])
def test_visco_gradient_bit_exact(golden, ckpt):
    # gradient_v (the gradient fwi() takes for media != 1), one shot per launch: bit-exact, also
    # when the forward memory variable is restored from checkpoints for recomputed segments
    g = golden("fwi_visco")
    p = FwiProblem(g["par"], g["wav"], data=g["data"], batch=1, checkpoint=ckpt)
    p.set_visco(g["q"])
    g1, f1, sc = p.gradient(g["v0"], "reference")
    assert np.array_equal(g1, g["grad1"])
    assert abs(f1 - float(g["fcost"][0])) <= 1e-3 * abs(f1)
    g2, _, _ = p.gradient((g["v0"] * 1.01).astype(np.float32), "reference", scaling=sc)
    assert np.array_equal(g2, g["grad2"])


def test_visco_gradient_interval4_batched_and_misfit(golden):
    g = golden("fwi_visco")
    p = FwiProblem(g["par4"], g["wav"], data=g["data"], batch=1, checkpoint=12)
    p.set_visco(g["q"])
    g4, _, _ = p.gradient(g["v0"], "reference")
    assert np.array_equal(g4, g["grad_int4"])
    b = FwiProblem(g["par"], g["wav"], data=g["data"], batch=3, checkpoint=40)
    b.set_visco(g["q"])
    gb, fb, _ = b.gradient(g["v0"], "reference")
    assert rel_l2(gb, g["grad1"]) <= 1e-5
    assert abs(b.misfit(g["v0"]) - fb) <= 1e-6 * fb
    # switching the plan back to acoustic gives the acoustic gradient of the same data
    b.set_visco(None)
    ga, _, _ = b.gradient(g["v0"], "reference")
    assert rel_l2(ga, gb) > 0.05


def test_visco_lbfgs_matches_reference(golden):
    # fwi() with media = 2: the reference's trajectory to f32 round-off of the models (the
    # optimiser's dot products are f64 here)
    g = golden("fwi_visco")
    p = FwiProblem(g["par"], g["wav"], data=g["data"], batch=3)
    p.set_visco(g["q"])
    vinv, k = p.lbfgs(g["v0"])
    assert k == g["vinv"].shape[2]
    assert np.abs(vinv - g["vinv"]).max() < 1e-4


def test_pfwi_media2_forward_modelingc(golden):
    # the media=2 branch of mode 1 (pywave/pfwi.py:72) runs the visco-acoustic modelling
    g = golden("fwi_visco")
    par = dict(zip(["nz", "nx", "dz", "dx", "oz", "ox", "nt", "dt", "ot", "inv", "ns", "ds", "os", "sz", "nb",
                    "coef", "f0", "acqui_type", "interval", "niter"], [float(x) for x in g["par"]]))
    for k in ["nz", "nx", "nt", "ns", "nb", "acqui_type", "interval", "niter"]:
        par[k] = int(par[k])
    data, vinv, grad, src, mwt = pywave_amd.pfwi(g["vtrue"], g["q"], g["wav"], None, mode=1, media=2, par=par)
    assert np.array_equal(data, g["data"]) and vinv == [] and grad == []


def test_lbfgs_decreases_misfit(small):
    p = FwiProblem(small["par"], small["wav"], data=small["data"], batch=3)
    vinv, k = p.lbfgs(small["v0"], 3)
    assert k == 3
    f0 = p.misfit(small["v0"])
    f3 = p.misfit(vinv[:, :, -1])
    assert f3 < f0


def _fd_case():
    rng = np.random.default_rng(8)
    nz, nx, nt = 24, 36, 260
    vt = np.full((nz, nx), 2.0)
    vt[10:] = 2.6
    vt += 0.08 * rng.standard_normal((nz, nx))
    v0 = vt.copy()
    v0[8:16, 12:24] += 0.15
    par = np.array([nz, nx, 0.01, 0.01, 0, 0, nt, 0.001, 0, 0, 2, 0.12, 0.06, 5, 12, 0.01, 12, 1, 1, 1],
                   np.float32)
    from oracle import oracle as O
    wav = (1e6 * O.ricker(nt, 0.001, 14, 0.05, 1.0)).astype(np.float32)
    return par, wav, vt.astype(np.float32), v0


def test_exact_adjoint_matches_finite_differences_f64():
    par, wav, vt, v0 = _fd_case()
    obs = FwiProblem(par, wav).forward(vt)
    p = FwiProblem(par, wav, data=obs, precision="f64", checkpoint=40)
    g, f, _ = p.gradient(v0, "exact")
    rng = np.random.default_rng(4)
    cells = [(int(rng.integers(2, 22)), int(rng.integers(2, 34))) for _ in range(8)]
    for iz, ix in cells:
        h = 1e-4 * v0[iz, ix]
        vp = v0.copy(); vp[iz, ix] += h
        vm = v0.copy(); vm[iz, ix] -= h
        fd = (p.misfit(vp) - p.misfit(vm)) / (2 * h)
        assert abs(fd - g[iz, ix]) <= 1e-3 * max(abs(fd), 1e-3 * np.abs(g).max()), (iz, ix, fd, g[iz, ix])


def test_exact_gradient_checkpointing_and_batching_consistent():
    par, wav, vt, v0 = _fd_case()
    obs = FwiProblem(par, wav).forward(vt)
    v0 = v0.astype(np.float32)
    ga, fa, _ = FwiProblem(par, wav, data=obs, batch=1, checkpoint=0).gradient(v0, "exact")
    gb, fb, _ = FwiProblem(par, wav, data=obs, batch=1, checkpoint=17).gradient(v0, "exact")
    assert np.array_equal(ga, gb) and fa == fb
    gc, fc, _ = FwiProblem(par, wav, data=obs, batch=2, checkpoint=50).gradient(v0, "exact")
    assert rel_l2(gc, ga) <= 1e-6 and abs(fc - fa) <= 1e-9 * abs(fa)
    # the f32 gradient agrees with the f64 one within the gradient tolerance (1e-3): the residual
    # d - F(v0) is a few percent of the data, so the ~1e-6 relative rounding of f32 traces
    # reaches the gradient amplified (measured 1.9e-4)
    gd, _, _ = FwiProblem(par, wav, data=obs, precision="f64").gradient(v0.astype(np.float64), "exact")
    assert rel_l2(ga, gd) <= 1e-3


def test_autograd_matches_direct_gradient():
    torch = pytest.importorskip("torch")
    par, wav, vt, v0 = _fd_case()
    p = FwiProblem(par, wav, batch=2, checkpoint=30)
    obs = p.forward(vt)
    p.set_observed(obs)
    g_direct, f_direct, _ = p.gradient(v0.astype(np.float32), "exact")
    dev = torch.device("cuda:0")
    v = torch.tensor(v0.astype(np.float32), device=dev, requires_grad=True)
    d = torch.tensor(np.ascontiguousarray(obs.transpose(2, 0, 1)), device=dev)  # (ns, nt, nr)
    tr = pywave_amd.fwi_traces(v, p)
    loss = 0.5 * ((tr.double() - d.double()) ** 2).sum()
    loss.backward()
    torch.cuda.synchronize()
    assert abs(loss.item() - f_direct) <= 1e-6 * f_direct
    assert rel_l2(v.grad.cpu().numpy(), g_direct) <= 1e-5


def test_autograd_on_side_stream_matches_host_forward():
    # the engine follows torch's current stream, also a non-default one
    torch = pytest.importorskip("torch")
    par, wav, vt, v0 = _fd_case()
    p = FwiProblem(par, wav, batch=2)
    ref = p.forward(v0)  # (nt, nr, ns)
    dev = torch.device("cuda:0")
    side = torch.cuda.Stream(dev)
    with torch.cuda.stream(side):
        v = torch.tensor(v0.astype(np.float32), device=dev)
        tr = pywave_amd.fwi_traces(v, p)
        host = tr.cpu().numpy()
    torch.cuda.synchronize()
    assert np.array_equal(host, np.ascontiguousarray(ref.transpose(2, 0, 1)))


def test_autograd_checkpoint_reuse_and_repeat():
    # one batch holds both shots: the backward reuses the forward's checkpoints; a second backward
    # through the same graph recomputes them and must give the same gradient; with one shot per
    # batch nothing is reused
    torch = pytest.importorskip("torch")
    par, wav, vt, v0 = _fd_case()
    obs = FwiProblem(par, wav).forward(vt)
    g_direct, _, _ = FwiProblem(par, wav, data=obs, batch=2, checkpoint=30).gradient(v0.astype(np.float32), "exact")
    dev = torch.device("cuda:0")
    d = torch.tensor(np.ascontiguousarray(obs.transpose(2, 0, 1)), device=dev)
    for batch in (2, 1):
        p = FwiProblem(par, wav, batch=batch, checkpoint=30)
        v = torch.tensor(v0.astype(np.float32), device=dev, requires_grad=True)
        loss = 0.5 * ((pywave_amd.fwi_traces(v, p).double() - d.double()) ** 2).sum()
        loss.backward(retain_graph=True)
        g1 = v.grad.clone()
        v.grad = None
        loss.backward()
        torch.cuda.synchronize()
        g2 = v.grad
        assert torch.equal(g1, g2), batch
        assert rel_l2(g1.cpu().numpy(), g_direct) <= 1e-5, batch


def test_adjoint_dev_recomputes_after_in_place_model_change():
    # pw_fwi_forward_dev keeps its checkpoints for the next pw_fwi_adjoint_dev on the same model
    # pointer; a model changed in place between the two calls must not reuse them
    torch = pytest.importorskip("torch")
    par, wav, vt, v0 = _fd_case()
    obs = FwiProblem(par, wav).forward(vt)
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev).cuda_stream
    g = torch.tensor(np.ascontiguousarray((0.01 * obs).transpose(2, 0, 1)), device=dev)  # cotangent
    nz, nx = v0.shape
    v1 = (1.02 * v0).astype(np.float32)

    def adjoint(p, vz):
        gv = torch.empty_like(vz)
        p.adjoint_dev(vz.data_ptr(), g.data_ptr(), gv.data_ptr(), stream)
        torch.cuda.synchronize()
        return gv.cpu().numpy()

    # the gradient at v1 from a plan that never saw another model
    fresh = FwiProblem(par, wav, batch=2, checkpoint=30)
    vz1 = torch.tensor(np.ascontiguousarray(v1.T), device=dev)  # z fastest
    g_ref = adjoint(fresh, vz1)
    fresh.close()
    p = FwiProblem(par, wav, batch=2, checkpoint=30)
    vz = torch.tensor(np.ascontiguousarray(v0.astype(np.float32).T), device=dev)
    tr = torch.empty((p.ns, p.nt, p.nr), dtype=torch.float32, device=dev)
    p.forward_dev(vz.data_ptr(), tr.data_ptr(), stream)
    vz.copy_(vz1)  # same pointer, new values
    g_got = adjoint(p, vz)
    p.close()
    assert np.array_equal(g_got, g_ref)


@pytest.mark.parametrize("ckpt", [0, 17, 50])
def test_two_step_forward_equals_single_steps(small, ckpt):
    # the temporally blocked forward (two steps per launch, forced) against one kernel per time
    # step: modelled data, the reference gradient (its recomputed snapshots) and the exact gradient
    # (its recomputed terms) bit for bit, also with an odd checkpoint interval
    out = {}
    for n in (1, -2):
        p = FwiProblem(small["par"], small["wav"], data=small["data"], batch=2, checkpoint=ckpt)
        p.set_steps_per_launch(n)
        d = p.forward(small["vtrue"])
        w = p.widths()
        g0, f0, _ = p.gradient(small["v0"], "reference")
        g1, f1, _ = p.gradient(small["v0"], "exact")
        p.close()
        out[n] = (d, g0, f0, g1, f1, w)
    assert out[-2][5]["two_step_forward"] > 0 and out[1][5]["two_step_forward"] == 0
    assert out[-2][5]["two_step_exact_adjoint"] > 0 and out[1][5]["two_step_exact_adjoint"] == 0
    for k in range(3):
        assert np.array_equal(out[1][k], out[-2][k]), k
    # the exact gradient: two-step forward terms and two-step adjoint launches; the adjoint
    # contracts to FMAs, so its rounding may differ from single launches in the last bits
    assert out[1][4] == out[-2][4]
    e = rel_l2(out[-2][3], out[1][3])
    print("two-step exact gradient vs single steps: rel-L2 %.2e" % e)
    assert e <= 1e-6
    # and against the compiled reference's forward (golden)
    assert np.array_equal(out[-2][0], small["data"])


def test_two_step_forward_odd_nt_and_marine_geometry(golden):
    g = golden("fwi_marine")
    par = g["par"].copy()
    par[6] = 299  # odd nt: the last step runs alone
    res = []
    for n in (1, -2):
        p = FwiProblem(par, g["wav"], batch=2)
        p.set_steps_per_launch(n)
        res.append(p.forward(g["vel"]))
        p.close()
    assert np.array_equal(res[0], res[1])
    assert np.array_equal(res[1], g["data"][:299])


def test_forced_widths_give_the_reference_results(small):
    # every forced tile width: forward data and the reference gradient bit-exact against the
    # reference C, the exact gradient within f32 rounding of the timed widths' one
    par, wav, vt, v0 = _fd_case()
    obs = FwiProblem(par, wav).forward(vt)
    v0 = v0.astype(np.float32)
    ge, _, _ = FwiProblem(par, wav, data=obs, batch=2).gradient(v0, "exact")
    for wd in (4, 8, 16):
        p = FwiProblem(small["par"], small["wav"], data=small["data"], batch=1)
        p.set_widths(forward=wd, exact_adjoint=wd, reference_adjoint=wd)
        assert np.array_equal(p.forward(small["vtrue"]), small["data"]), wd
        g1, _, _ = p.gradient(small["v0"], "reference")
        assert np.array_equal(g1, small["grad1"]), wd
        w = p.widths()
        assert (w["forward"], w["exact_adjoint"], w["reference_adjoint"]) == (wd, wd, wd) and w["tuned"]
        q = FwiProblem(par, wav, data=obs, batch=2)
        q.set_widths(forward=wd, exact_adjoint=wd, reference_adjoint=wd)
        gq, _, _ = q.gradient(v0, "exact")
        assert rel_l2(gq, ge) <= 1e-6, wd
        q.close()
        p.close()
    with pytest.raises(ValueError):
        FwiProblem(small["par"], small["wav"], batch=1).set_widths(forward=5)
