"""Passive modelling and least-squares time-reversal imaging (pfwi mode 4 lstric) on the HIP engine.

Bars: modelled data and the adjoint image are bit-exact against the compiled reference's
timerev_lop (golden vectors) and the oracle; the CG source inversion matches the reference's
iterates to rel-L2 <= 1e-5 (its dot products are f64 tree reductions here, sequential f64 sums in
the reference)."""
import numpy as np
import pytest

import pywave_amd
from conftest import rel_l2
from pywave_amd import LstriProblem

pytestmark = pytest.mark.gpu

PAR_KEYS = ["nz", "nx", "dz", "dx", "oz", "ox", "nt", "dt", "ot", "inv", "ns", "ds", "os", "sz", "nb", "coef",
            "f0", "acqui_type", "interval", "niter"]


@pytest.fixture(scope="module")
def ls():
    from conftest import load_golden
    return load_golden("lstri_small")


def _par_dict(par):
    d = dict(zip(PAR_KEYS, [float(x) for x in par]))
    for k in ["nz", "nx", "nt", "ns", "nb", "acqui_type", "interval", "niter"]:
        d[k] = int(d[k])
    return d


def test_lstri_modelling_bit_exact(ls):
    p = LstriProblem(ls["par"], ls["vel"])
    d = p.model(ls["src"])
    assert d.shape == ls["data"].shape
    assert np.array_equal(d, ls["data"])


def test_lstri_image_bit_exact_against_oracle(ls):
    from oracle import oracle as O
    rng = np.random.default_rng(3)
    data = rng.standard_normal(ls["data"].shape).astype(np.float32)
    p = LstriProblem(ls["par"], ls["vel"])
    img = p.image(data)
    for s in range(data.shape[2]):
        ref = O.timerev(ls["par"], ls["vel"], data[:, :, s], adj=True)
        assert np.array_equal(img[:, :, :, s], ref), s


def test_lstri_adjoint_identity_on_device(ls):
    rng = np.random.default_rng(9)
    p = LstriProblem(ls["par"], ls["vel"])
    m = rng.standard_normal(ls["src"].shape).astype(np.float32)
    d = rng.standard_normal(ls["data"].shape).astype(np.float32)
    lhs = float((p.model(m).astype(np.float64) * d).sum())
    rhs = float((m.astype(np.float64) * p.image(d)).sum())
    assert abs(lhs - rhs) <= 1e-5 * max(abs(lhs), abs(rhs))


def test_lstri_inversion_matches_reference(ls):
    p = LstriProblem(ls["par_inv"], ls["vel"])
    src, its = p.invert(ls["data"], return_iters=True)
    assert list(its) == [6, 6]
    assert rel_l2(src, ls["src_inv"]) <= 1e-5
    # per shot too (the second shot starts from the first shot's last CG step)
    for s in range(2):
        assert rel_l2(src[..., s], ls["src_inv"][..., s]) <= 1e-5, s
    fit = p.model(src)
    assert rel_l2(fit, ls["data"]) < 0.5


def test_pfwi_mode4_lstric_dropins(ls):
    par = _par_dict(ls["par"])
    z = np.zeros((16, 24), np.float32)
    wav = np.zeros(120, np.float32)
    data, vinv, grad, src, mwt = pywave_amd.pfwi(ls["vel"], z, wav, ls["src"], mode=4, par=dict(par, inv=0))
    assert np.array_equal(data, ls["data"]) and vinv == [] and grad == [] and src == [] and mwt == []
    par_i = dict(par, inv=True, onlysrc=True)
    d2, vinv, grad, src, mwt = pywave_amd.pfwi(ls["vel"], z, wav, None, data=ls["data"], mode=4, par=par_i)
    assert d2 == [] and vinv is None and grad is None and mwt == []
    assert src.shape == (16, 24, 120, 2)
    assert rel_l2(src, ls["src_inv"]) <= 1e-5


def test_lstri_graph_replay_equals_step_launches(ls):
    # the inversion's sweeps replayed from captured hipGraphs (with the CG's eager work and host
    # synchronisations between replays) give the step-by-step launches' result bit for bit, also
    # on a second call and after a modelling call has re-sized the field buffers
    p = LstriProblem(ls["par_inv"], ls["vel"])
    p.use_graph(False)
    eager = p.invert(ls["data"])
    q = LstriProblem(ls["par_inv"], ls["vel"])
    g1 = q.invert(ls["data"])
    g2 = q.invert(ls["data"])
    assert np.array_equal(g1, eager) and np.array_equal(g2, eager)
    q.model(ls["src"])  # two shots in one launch: larger field buffers, captured sweeps dropped
    assert np.array_equal(q.invert(ls["data"]), eager)


@pytest.mark.parametrize("nt", [120, 121])
def test_two_step_launches_equal_single_steps(ls, nt):
    # the sweeps advance two steps per launch on small grids (odd nt ends with one single step):
    # modelled data, images and the CG inversion equal the one-step launches bit for bit, and the
    # modelling and adjoint equal the oracle's timerev_lop
    from oracle import oracle as O
    par = ls["par"].copy()
    par[6] = nt
    rng = np.random.default_rng(nt)
    src = rng.standard_normal((16, 24, nt, 2)).astype(np.float32)
    dat = rng.standard_normal((nt, 24, 2)).astype(np.float32)
    one = LstriProblem(par, ls["vel"])
    one.steps_per_launch(1)
    two = LstriProblem(par, ls["vel"])
    two.steps_per_launch(2)
    d2 = two.model(src)
    assert np.array_equal(one.model(src), d2)
    i2 = two.image(dat)
    assert np.array_equal(one.image(dat), i2)
    for s in range(2):
        assert np.array_equal(d2[:, :, s], O.timerev(par, ls["vel"], src[..., s], adj=False)), s
        assert np.array_equal(i2[..., s], O.timerev(par, ls["vel"], dat[:, :, s], adj=True)), s
    pi = par.copy()
    pi[9], pi[19] = 1, 3
    a = LstriProblem(pi, ls["vel"])
    a.steps_per_launch(1)
    b = LstriProblem(pi, ls["vel"])
    b.steps_per_launch(2)
    assert np.array_equal(a.invert(d2), b.invert(d2))
    with pytest.raises(ValueError):
        b.steps_per_launch(3)


def test_timed_launch_choice_and_every_tile_width_bit_exact(ls):
    # the plan times single-step widths 1-16 and two-step launches before its first sweep; every
    # width gives the same modelled data, image and inversion bit for bit
    par = ls["par"].copy()
    rng = np.random.default_rng(9)
    nt = int(par[6])
    src = rng.standard_normal((16, 24, nt, 2)).astype(np.float32)
    dat = rng.standard_normal((nt, 24, 2)).astype(np.float32)
    p = LstriProblem(par, ls["vel"])
    d0, i0 = p.model(src), p.image(dat)
    (w, spl, cols), log = p.launch_choice()
    assert w in (1, 2, 4, 8, 16) and spl in (1, 2, 4) and cols in (4, 8)
    kinds = [k for k, _, _ in log]
    assert kinds.count("model_step") == 5 and kinds.count("adjoint_step") == 5 and all(ms > 0 for _, _, ms in log)
    assert kinds.count("two_step_model") == 1 and kinds.count("two_step_adjoint") == 1
    assert kinds.count("four_step_model") == 2 and kinds.count("four_step_adjoint") == 2
    pi = par.copy()
    pi[9], pi[19] = 1, 3
    q = LstriProblem(pi, ls["vel"])
    x0 = q.invert(d0)
    for width in (1, 2, 4, 8, 16):
        p.set_tile_width(width)
        p.steps_per_launch(1)
        assert np.array_equal(p.model(src), d0), width
        assert np.array_equal(p.image(dat), i0), width
        assert p.launch_choice()[0][:2] == (width, 1)
        q.set_tile_width(width)
        q.steps_per_launch(1)
        assert np.array_equal(q.invert(d0), x0), width
    with pytest.raises(ValueError):
        p.set_tile_width(3)


@pytest.mark.parametrize("nt", [120, 121, 122, 123])
@pytest.mark.parametrize("cols", [4, 8])
def test_four_step_launches_equal_single_steps(ls, nt, cols):
    # four steps per launch (remainders of 1-3 steps in two-step and single-step launches): modelled
    # data, images and the CG inversion equal the one-step launches bit for bit, and the modelling
    # and adjoint equal the oracle's timerev_lop
    from oracle import oracle as O
    par = ls["par"].copy()
    par[6] = nt
    rng = np.random.default_rng(1000 + nt + cols)
    src = rng.standard_normal((16, 24, nt, 2)).astype(np.float32)
    dat = rng.standard_normal((nt, 24, 2)).astype(np.float32)
    one = LstriProblem(par, ls["vel"])
    one.steps_per_launch(1)
    four = LstriProblem(par, ls["vel"])
    four.steps_per_launch(4)
    four.set_four_step_columns(cols)
    d4 = four.model(src)
    assert four.launch_choice()[0][1:] == (4, cols)
    assert np.array_equal(one.model(src), d4)
    i4 = four.image(dat)
    assert np.array_equal(one.image(dat), i4)
    for s in range(2):
        assert np.array_equal(d4[:, :, s], O.timerev(par, ls["vel"], src[..., s], adj=False)), s
        assert np.array_equal(i4[..., s], O.timerev(par, ls["vel"], dat[:, :, s], adj=True)), s
    pi = par.copy()
    pi[9], pi[19] = 1, 3
    a = LstriProblem(pi, ls["vel"])
    a.steps_per_launch(1)
    b = LstriProblem(pi, ls["vel"])
    b.steps_per_launch(4)
    b.set_four_step_columns(cols)
    assert np.array_equal(a.invert(d4), b.invert(d4))
    with pytest.raises(ValueError):
        b.set_four_step_columns(6)
