"""The extension-module functions of pywave's namespace (afdcfun, apscfun, pfwicfun) on the HIP
engine, called as the reference's wrappers call them.  The golden vectors hold the compiled
reference's own flat returns, so the bar is bit-exact equality with them (the optimisers' iterates:
1e-4 km/s and rel-L2 1e-5, as for the high-level drop-ins)."""
import numpy as np
import pytest

import pywave_amd
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _f(a):
    return np.asarray(a, np.float32).flatten(order="F")


def test_afd2dc_modelling_and_tri(golden):
    g = golden("afd2d_cfg0")
    nz, nx = g["vel"].shape
    out = pywave_amd.afd2dc(_f(g["vel"]), g["src"], 0, int(g["nt"]), nx, nz, 1, 0, int(g["jsnap"]), 0, 1,
                            int(g["nbt"]), float(g["ct"]), float(g["dt"]), 0.0, float(g["dx"]), 0.0, float(g["dz"]))
    assert np.array_equal(out, g["out"])
    t = golden("afd2d_tri")
    nz, nx = t["vel"].shape
    out = pywave_amd.afd2dc(_f(t["vel"]), _f(t["dat"]), 1, int(t["nt"]), nx, nz, 0, 0, int(t["jsnap"]), 0, 1,
                            int(t["nbt"]), float(t["ct"]), float(t["dt"]), 0.0, float(t["dx"]), 0.0, float(t["dz"]))
    assert np.array_equal(out, t["out"])


@pytest.mark.parametrize("name,fn", [("afd3d_two_sources", "afd3dc"), ("aps3d_small", "aps3dc")])
def test_3d_cfuns(golden, name, fn):
    g = golden(name)
    nz, nx, ny = g["vel"].shape
    ns = len(g["src"]) // 6
    out = getattr(pywave_amd, fn)(_f(g["vel"]), g["src"], 0, int(g["nt"]), nx, ny, nz, ns, 0, int(g["jsnap"]), 0, 1,
                                  int(g["nbt"]), float(g["ct"]), float(g["dt"]), 0.0, float(g["dx"]), 0.0,
                                  float(g["dy"]), 0.0, float(g["dz"]))
    assert np.array_equal(out, g["out"])


def test_aps2dc(golden):
    g = golden("aps2d_two_sources")
    nz, nx = g["vel"].shape
    ns = len(g["src"]) // 5
    out = pywave_amd.aps2dc(_f(g["vel"]), g["src"], 0, int(g["nt"]), nx, nz, ns, 0, int(g["jsnap"]), 0, 1,
                            int(g["nbt"]), float(g["ct"]), float(g["dt"]), 0.0, float(g["dx"]), 0.0, float(g["dz"]))
    assert np.array_equal(out, g["out"])


def test_pfwicfun_modelling_and_fwi(golden):
    s = golden("fwi_small")
    d = pywave_amd.forward_modeling_ac(_f(s["vtrue"]), _f(s["q"]), s["wav"], s["par"])
    assert np.array_equal(d, _f(s["data"]))
    v = pywave_amd.fwic(_f(s["v0"]), _f(s["q"]), s["wav"], _f(s["data"]), s["par"])
    assert np.abs(v.reshape(s["vinv"].shape, order="F") - s["vinv"]).max() < 1e-4
    vi = golden("fwi_visco")
    dv = pywave_amd.forward_modelingc(_f(vi["vtrue"]), _f(vi["q"]), vi["wav"], vi["par"])
    assert np.array_equal(dv, _f(vi["data"]))


def test_pfwicfun_passive(golden):
    ls = golden("lstri_small")
    z = np.zeros(16 * 24, np.float32)
    w = np.zeros(120, np.float32)
    d = pywave_amd.lstric(_f(ls["vel"]), z, w, _f(ls["src"]), ls["par"])
    assert np.array_equal(d, _f(ls["data"]))
    src = pywave_amd.lstric(_f(ls["vel"]), z, w, _f(ls["data"]), ls["par_inv"])
    assert rel_l2(src, _f(ls["src_inv"])) <= 1e-5
    pas = golden("pfwi_small")
    par = pas["par"].copy()
    par[9] = 1
    ds = np.concatenate([_f(pas["data"]), _f(pas["src"])])
    v = pywave_amd.pfwic(_f(pas["v0"]), _f(pas["q"]), pas["wav"], ds, par)
    assert np.abs(v.reshape(pas["vinv"].shape, order="F") - pas["vinv"]).max() < 1e-4
