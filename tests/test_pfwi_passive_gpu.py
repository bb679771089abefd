"""Passive FWI (pfwi mode 4, known source field, velocity only) on the HIP engine.

Bars: forward data and the reference imaging-condition gradient (one shot per launch) are
bit-exact against the compiled reference's gradient_pas_av (pywave/src/wave_fwigradient.c:672-870);
batched launches only reorder the per-shot sums (rel-L2 <= 1e-5); the L-BFGS iterates of pfwi
(pywave/src/wave_fwi.c:212-322) follow the reference's to f32 round-off (f64 dot products here,
max |dv| < 1e-4 km/s); the exact adjoint with a source field matches central finite differences
in f64 to 1e-3 relative."""
import numpy as np
import pytest

import pywave_amd
from conftest import rel_l2
from pywave_amd import FwiProblem

pytestmark = pytest.mark.gpu

PAR_KEYS = ["nz", "nx", "dz", "dx", "oz", "ox", "nt", "dt", "ot", "inv", "ns", "ds", "os", "sz", "nb", "coef",
            "f0", "acqui_type", "interval", "niter"]


@pytest.fixture(scope="module")
def pas():
    from conftest import load_golden
    return load_golden("pfwi_small")


def _par_dict(par):
    d = dict(zip(PAR_KEYS, [float(x) for x in par]))
    for k in ["nz", "nx", "nt", "ns", "nb", "acqui_type", "interval", "niter"]:
        d[k] = int(d[k])
    return d


@pytest.mark.parametrize("batch", [1, 2])
def test_passive_forward_bit_exact(pas, batch):
    p = FwiProblem(pas["par"], pas["wav"], src=pas["src"], batch=batch)
    assert np.array_equal(p.forward(pas["vtrue"]), pas["data"])


@pytest.mark.parametrize("ckpt", [0, 37])
def test_passive_reference_gradient_bit_exact(pas, ckpt):
    p = FwiProblem(pas["par"], pas["wav"], data=pas["data"], src=pas["src"], batch=1, checkpoint=ckpt)
    g1, f1, sc = p.gradient(pas["v0"], "reference")
    assert np.array_equal(g1, pas["grad1"])
    assert abs(f1 - float(pas["fcost"][0])) <= 1e-3 * abs(f1)  # reference accumulates in f32
    g2, _, _ = p.gradient((pas["v0"] * 1.01).astype(np.float32), "reference", scaling=sc)
    assert np.array_equal(g2, pas["grad2"])
    # at the true model the residual is exactly zero
    _, f0, _ = p.gradient(pas["vtrue"], "reference", scaling=sc)
    assert f0 == 0.0


def test_passive_reference_gradient_batched(pas):
    p = FwiProblem(pas["par"], pas["wav"], data=pas["data"], src=pas["src"], batch=2, checkpoint=50)
    g, _, _ = p.gradient(pas["v0"], "reference")
    assert rel_l2(g, pas["grad1"]) <= 1e-5


def test_point_source_field_reproduces_active_gradient(golden):
    # the active point-source term written as a source field: same data and gradient, bit for bit
    from oracle import oracle as O
    g = golden("fwi_small")
    par = g["par"]
    a = O.make_acq(par)
    nz, nx, nt, ns, nb = int(par[0]), int(par[1]), int(par[6]), int(par[10]), int(par[14])
    src = np.zeros((nz, nx, nt, ns), np.float32)
    for s in range(ns):
        rr = O.source_map(a.s0_v + s * a.ds_v, a.sz, 2, 2, a.padnx, a.padnz)
        src[:, :, :, s] = rr[nb:nb + nz, nb:nb + nx][:, :, None] * g["wav"][None, None, :]
    p = FwiProblem(par, g["wav"], data=g["data"], src=src, batch=1)
    assert np.array_equal(p.forward(g["vtrue"]), g["data"])
    g1, _, _ = p.gradient(g["v0"], "reference")
    assert np.array_equal(g1, g["grad1"])
    # back to the point sources
    p.set_source_field(None)
    assert np.array_equal(p.forward(g["vtrue"]), g["data"])


def test_pfwi_mode4_dropin(pas):
    par = _par_dict(pas["par"])
    par.update(inv=True, onlysrc=False, onlyvel=True)
    data, vinv, grad, src, mwt = pywave_amd.pfwi(pas["v0"], pas["q"], pas["wav"], pas["src"], data=pas["data"],
                                                 mode=4, par=par)
    assert grad is None and src is None and mwt is None
    assert vinv.shape == pas["vinv"].shape == (30, 40, 7)
    # seven iterations run through the step reset at iteration 5; f64 dot products in the
    # optimiser keep the trajectory to f32 round-off of the models
    assert np.abs(vinv - pas["vinv"]).max() < 1e-4
    assert np.array_equal(data, pas["data"].flatten(order="F"))
    par2 = dict(par, onlyvel=False)
    with pytest.raises(UnboundLocalError):
        pywave_amd.pfwi(pas["v0"], pas["q"], pas["wav"], pas["src"], data=pas["data"], mode=4, par=par2)


def test_passive_lbfgs_matches_pfwi_entry(pas):
    p = FwiProblem(pas["par"], pas["wav"], data=pas["data"], src=pas["src"], batch=2)
    vinv, k = p.lbfgs(pas["v0"])
    assert k == 7
    assert np.abs(vinv - pas["vinv"]).max() < 1e-4
    assert p.misfit(vinv[:, :, -1]) < p.misfit(pas["v0"])


def test_passive_exact_adjoint_matches_finite_differences_f64(pas):
    par = pas["par"].copy()
    v0 = pas["v0"].astype(np.float64)
    obs = FwiProblem(par, pas["wav"], src=pas["src"]).forward(pas["vtrue"])
    p = FwiProblem(par, pas["wav"], data=obs, src=pas["src"], precision="f64", checkpoint=60)
    g, _, _ = p.gradient(v0, "exact")
    rng = np.random.default_rng(6)
    cells = [(int(rng.integers(2, 28)), int(rng.integers(2, 38))) for _ in range(8)]
    for iz, ix in cells:
        h = 1e-4 * v0[iz, ix]
        vp = v0.copy(); vp[iz, ix] += h
        vm = v0.copy(); vm[iz, ix] -= h
        fd = (p.misfit(vp) - p.misfit(vm)) / (2 * h)
        assert abs(fd - g[iz, ix]) <= 1e-3 * max(abs(fd), 1e-3 * np.abs(g).max()), (iz, ix, fd, g[iz, ix])


def test_passive_autograd_matches_direct_gradient(pas):
    torch = pytest.importorskip("torch")
    p = FwiProblem(pas["par"], pas["wav"], data=pas["data"], batch=2, checkpoint=40)
    dev = torch.device("cuda:0")
    # device source field [ns][nt][nx][nz]
    fs = torch.tensor(np.ascontiguousarray(pas["src"].transpose(3, 2, 1, 0)), device=dev)
    p.set_source_field(fs)
    g_direct, f_direct, _ = p.gradient(pas["v0"], "exact")
    v = torch.tensor(pas["v0"], device=dev, requires_grad=True)
    d = torch.tensor(np.ascontiguousarray(pas["data"].transpose(2, 0, 1)), device=dev)
    loss = 0.5 * ((pywave_amd.fwi_traces(v, p).double() - d.double()) ** 2).sum()
    loss.backward()
    torch.cuda.synchronize()
    assert abs(loss.item() - f_direct) <= 1e-6 * f_direct
    assert rel_l2(v.grad.cpu().numpy(), g_direct) <= 1e-5
