"""The C-ABI library loads, exports every symbol include/pywave_amd.h declares, and its
host-side set-up helpers agree with the oracle.  CPU only (no device work)."""
import ctypes
import os
import re

import numpy as np
import pytest

import pywave_amd
from conftest import ROOT
from oracle import oracle as O
from pywave_amd import _native as N

HEADER = os.path.join(ROOT, "include", "pywave_amd.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pw_[a-z0-9_]+)\s*\(", txt)))


def test_header_symbols_exported_and_bound():
    L = N.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), "missing export %s" % s
    assert set(syms) == set(N.SIGNATURES), "ctypes table and header disagree"


def test_next_fast_size_and_padding_match_oracle():
    L = N.lib()
    for n in list(range(1, 400)) + [1501, 2999, 3001]:
        assert L.pw_next_fast_size(n) == O.next_fast_size(n)
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    for dims in [(161, 161, 1), (261, 261, 261), (188, 188, 188), (64, 60, 1), (77, 91, 43)]:
        L.pw_padded_sizes(*dims, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        expect = O.padded_sizes(*dims)
        assert (a.value, b.value) == expect[:2]
        if dims[2] > 1:
            assert c.value == expect[2]


def test_abc_weights_bit_exact():
    L = N.lib()
    for nb, c in [(30, 0.01), (12, 0.03), (60, 0.005), (1, 0.5)]:
        w = np.zeros(nb, np.float32)
        L.pw_abc_weights(nb, c, N.ptr(w))
        assert np.array_equal(w, O.abc_weights(nb, c))


def test_ricker_host_vs_reference(golden):
    L = N.lib()
    g = golden("ricker")
    for (nt, dt, f0, t0, A), ref in zip(g["params"], g["waves"]):
        nt = int(nt)
        w = np.zeros(nt, np.float32)
        assert L.pw_ricker(nt, dt, f0, t0, A, N.ptr(w)) == 0
        ref = ref[:nt]
        # the f32 mixed-radix transform reproduces the reference's f32 FFT rounding
        assert np.array_equal(w, ref)
        # independent f64 FFT evaluation of the same filter
        assert np.abs(w - O.ricker(nt, dt, f0, t0, A)).max() <= 1e-6 * np.abs(ref).max()


def test_ricker_rejects_onset_outside_trace():
    L = N.lib()
    w = np.zeros(100, np.float32)
    assert L.pw_ricker(100, 0.001, 10, 0.5, 1, N.ptr(w)) == -1


def test_output_length_formula():
    L = N.lib()
    assert L.pw_afd_output_len(1000, 101, 101, 1, 999) == 1000 * 101 + 101 * 101 * 2
    assert L.pw_afd_output_len(1500, 201, 201, 201, 0) == 1500 * 201 * 201
    assert L.pw_afd_output_len(10, 5, 6, 7, 3) == 10 * 42 + 5 * 42 * 4


def test_binread_binwrite_round_trip(tmp_path):
    # first axis fastest, as the demos' binread(n1=..., n2=...) reads model files
    rng = np.random.default_rng(0)
    a = rng.standard_normal((7, 5)).astype(np.float32)
    f = tmp_path / "m.bin"
    pywave_amd.binwrite(str(f), a)
    raw = np.fromfile(str(f), dtype="<f4")
    assert np.array_equal(raw[:7], a[:, 0])  # n1 contiguous
    assert np.array_equal(pywave_amd.binread(str(f), n1=7, n2=5), a)
    b = rng.standard_normal((3, 4, 2)).astype(np.float32)
    pywave_amd.binwrite(str(f), b)
    assert np.array_equal(pywave_amd.binread(str(f), 3, 4, 2), b)
    with pytest.raises(ValueError):
        pywave_amd.binread(str(f), 5, 5)


def test_pfwi_mode4_branches_without_device():
    # branches decided on the host before any device work (pywave/pfwi.py:120-256)
    par = {"nz": 8, "nx": 8, "dz": 0.01, "dx": 0.01, "oz": 0, "ox": 0, "nt": 10, "dt": 0.001, "ot": 0,
           "ns": 1, "nb": 8, "niter": 1}
    v = np.full((8, 8), 2.0, np.float32)
    w = np.zeros(10, np.float32)
    src = np.zeros((8, 8, 10, 1), np.float32)
    with pytest.raises(UnboundLocalError):  # joint inversion: the reference's return fails
        pywave_amd.pfwi(v, v, w, src, data=np.zeros((10, 8, 1)), mode=4,
                        par=dict(par, inv=True, onlysrc=False, onlyvel=False))


def test_pfwi_checks_input_lengths_before_device_work():
    # the engine entry points read nz*nx model samples, nt wavelet samples and nt*nx*ns data
    # samples unchecked (as the reference does); the wrapper rejects other lengths first
    par = {"nz": 8, "nx": 8, "dz": 0.01, "dx": 0.01, "oz": 0, "ox": 0, "nt": 10, "dt": 0.001, "ot": 0,
           "ns": 1, "nb": 8, "niter": 1}
    v = np.full((8, 8), 2.0, np.float32)
    w = np.zeros(10, np.float32)
    with pytest.raises(ValueError, match="vel"):
        pywave_amd.pfwi(v[:, :7], v, w, None, mode=1, par=dict(par))
    with pytest.raises(ValueError, match="q"):
        pywave_amd.pfwi(v, v[:4], w, None, mode=1, media=2, par=dict(par))
    with pytest.raises(ValueError, match="wav"):
        pywave_amd.pfwi(v, v, w[:9], None, mode=1, par=dict(par))
    with pytest.raises(ValueError, match="data"):
        pywave_amd.pfwi(v, v, w, None, data=np.zeros((10, 8, 2)), mode=2, par=dict(par))
    with pytest.raises(ValueError, match="src"):
        pywave_amd.pfwi(v, v, w, np.zeros((8, 8, 10, 2)), data=np.zeros((10, 8, 1)), mode=4,
                        par=dict(par, inv=True, onlysrc=False, onlyvel=True))


def test_cfun_length_checks_before_device_work():
    # afdcfun's NULL-without-exception on a velocity length mismatch surfaces as SystemError
    with pytest.raises(SystemError):
        pywave_amd.afd2dc(np.zeros(10, np.float32), np.zeros(5, np.float32), 0, 10, 4, 3, 1, 0, 0, 0, 1, 2, 0.01,
                          0.001, 0, 10, 0, 10)
    with pytest.raises(SystemError):
        pywave_amd.aps3dc(np.zeros(10, np.float32), np.zeros(6, np.float32), 0, 10, 2, 2, 2, 1, 0, 0, 0, 1, 2, 0.01,
                          0.001, 0, 10, 0, 10, 0, 10)
    with pytest.raises(ValueError):  # a source pack shorter than 5*ns
        pywave_amd.afd2dc(np.zeros(12, np.float32), np.zeros(4, np.float32), 0, 10, 4, 3, 1, 0, 0, 0, 1, 2, 0.01,
                          0.001, 0, 10, 0, 10)
    par = np.array([8, 8, 0.01, 0.01, 0, 0, 10, 0.001, 0, 0, 1, 0.2, 0, 5, 8, 0.01, 10, 1, 1, 1], np.float32)
    with pytest.raises(ValueError):
        pywave_amd.forward_modeling_ac(np.zeros(63, np.float32), np.zeros(64, np.float32), np.zeros(10), par)
    with pytest.raises(ValueError):
        pywave_amd.fwic(np.zeros(64, np.float32), np.zeros(64), np.zeros(10), np.zeros(79), par)
