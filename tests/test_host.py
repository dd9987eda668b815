"""Host-side logic that needs no GPU: set_data validation, parameter normalisation
and the descriptors handed to the C-ABI (through a recording stand-in for the
library), and that the built library exports every symbol of include/rr_hip.h."""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

import pytransit_amd
from pytransit_amd import RoadRunnerModel, _lib
from pytransit_amd import rrmodel as rrmod

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, 'include', 'rr_hip.h')).read()
    src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
    return sorted(set(re.findall(r'^\s*(?:const\s+)?[a-z_0-9]+\**\s+\**([a-z_0-9]+)\s*\(', src, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = header_functions()
    assert 'rr_evaluate' in names and 'rr_last_error' in names and len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_loader_binds_every_signature():
    lib = _lib.load(require_gpu=False)
    assert lib.rr_abi_version() >= 2
    for name in _lib._SIGNATURES:
        assert hasattr(lib, name)


def test_z_grid_host_call():
    ze, zm, mu = _lib.z_grid(0.7, 20, 20)
    np.testing.assert_allclose(ze[:3], [0.035, 0.07, 0.105], rtol=0, atol=1e-15)
    assert ze[19] == pytest.approx(0.7, abs=1e-15) and ze[-1] == 1.0 and zm[0] == 0.0 and mu[0] == 1.0


def test_c_abi_rejects_bad_arguments():
    lib = _lib.load(require_gpu=False)
    ze = np.zeros(4)
    dp = ctypes.POINTER(ctypes.c_double)
    rc = lib.rr_z_grid(1.5, 2, 2, ze.ctypes.data_as(dp), ze.ctypes.data_as(dp), None)
    assert rc == -1 and b'zcut' in lib.rr_last_error()
    rc = lib.rr_ld_profiles(42, None, 1, 1, 2, None, 4, None, None, None)
    assert rc == -1 and b'unknown law' in lib.rr_last_error()
    rc = lib.rr_ld_profiles(2, None, 1, 1, 1, None, 4, None, None, None)
    assert rc == -1 and b'coefficients' in lib.rr_last_error()
    assert lib.rr_workspace_bytes(4, 1000, 1, 100, 1) >= 4 * 128 + 4 * 100 * 12
    assert lib.ts_workspace_bytes(4, 8, 100, 40, 1) >= 4 * 8 * 100 * 4 + 4 * 40 * 112 * 4
    # batches: size limit and an empty batch are checked before any device work
    assert lib.rr_max_batch() == 16
    rc = lib.rr_evaluate_batch(1, None, 17, None, None)
    assert rc == -1 and b'at most 16' in lib.rr_last_error()
    assert lib.rr_evaluate_batch(1, None, 0, None, None) == 0
    reqs = (_lib.RRRequest * 2)()
    reqs[1].pv.npv = -1
    rc = lib.rr_evaluate_batch(1, None, 2, reqs, None)
    assert rc == -1 and b'rr_evaluate_batch[0]' in lib.rr_last_error()
    # profile-grid interpolation: axes of one node, a bad stride, a zero step and
    # null pointers are rejected before any device work; an empty population is a no-op
    rc = lib.rr_ldtk_interpolate(None, 1, 2, 2, 1, 4, 0.0, 1.0, 0.0, 1.0, 0.0, 1.0, None, 3, 1, None, None)
    assert rc == -1 and b'>= 2 nodes' in lib.rr_last_error()
    rc = lib.rr_ldtk_interpolate(None, 2, 2, 2, 1, 4, 0.0, 1.0, 0.0, 1.0, 0.0, 1.0, None, 2, 1, None, None)
    assert rc == -1 and b'par_stride' in lib.rr_last_error()
    rc = lib.rr_ldtk_interpolate(None, 2, 2, 2, 1, 4, 0.0, 0.0, 0.0, 1.0, 0.0, 1.0, None, 3, 1, None, None)
    assert rc == -1 and b'steps positive' in lib.rr_last_error()
    rc = lib.rr_ldtk_interpolate(None, 2, 2, 2, 1, 4, 0.0, 1.0, 0.0, 1.0, float('nan'), 1.0, None, 3, 1, None, None)
    assert rc == -1 and b'finite' in lib.rr_last_error()
    rc = lib.rr_ldtk_interpolate(None, 2, 2, 2, 1, 4, 0.0, 1.0, 0.0, 1.0, 0.0, 1.0, None, 3, 1, None, None)
    assert rc == -1 and b'null pointer' in lib.rr_last_error()
    assert lib.rr_ldtk_interpolate(None, 2, 2, 2, 1, 4, 0.0, 1.0, 0.0, 1.0, 0.0, 1.0, None, 3, 0, None, None) == 0


# ---- set_data (transitmodel.py:56-125) ----------------------------------------------

def test_set_data_defaults_and_validation():
    tm = RoadRunnerModel('quadratic')
    tm.set_data(np.linspace(0, 1, 10))
    assert (tm.nlc, tm.npb, tm.nep) == (1, 1, 1)
    assert tm.nsamples.tolist() == [1] and tm.exptimes.tolist() == [0]
    with pytest.raises(ValueError, match='integers'):
        tm.set_data(np.linspace(0, 1, 10), lcids=np.zeros(10))
    with pytest.raises(ValueError, match='should equal'):
        tm.set_data(np.linspace(0, 1, 10), lcids=np.zeros(9, int))
    with pytest.raises(ValueError, match='ligt curves'):
        tm.set_data(np.linspace(0, 1, 10), lcids=np.array([0, 1] * 5), pbids=[0])
    with pytest.raises(ValueError, match='between 0 and'):
        tm.set_data(np.linspace(0, 1, 10), lcids=np.array([0, 1] * 5), pbids=[0, 2])


def test_unknown_ld_model_raises_keyerror(capsys):
    with pytest.raises(KeyError):
        RoadRunnerModel('no-such-law')
    assert 'Unknown limb darkening model' in capsys.readouterr().out
    with pytest.raises(NotImplementedError):
        RoadRunnerModel(42)


# ---- the evaluate path against a recording stand-in for the library -----------------

class FakeLib:
    """Records the descriptors rr_setup / rr_flux receive; host-only entry points
    (rr_z_grid) go to the real library."""

    def __init__(self, real):
        self.calls = []
        self._real = real

    def __getattr__(self, name):
        if name == 'rr_z_grid':
            return getattr(self._real, name)
        raise AttributeError(name)

    def rr_workspace_bytes(self, npv, npt, npb, ng, nlc):
        return 4096 + npv * (128 + npb * 8 + npb * ng * 12)

    def rr_weights_3d(self, nk, kmin, kmax, ze, nz, ng, W, W32, dk, dg, stream):
        dk._obj.value = (kmax - kmin) / nk
        dg._obj.value = (1 - 1e-7) / (ng - 1)
        return 0

    def rr_setup(self, prec, tab, lc, prm, ld, ws, nbytes, stream):
        t, l, p, d = tab._obj, lc._obj, prm._obj, ld._obj
        self.calls.append(('setup', prec, dict(npv=p.npv, kcols=p.kcols, t0cols=p.t0cols, npt=l.npt, nlc=l.nlc,
                                               npb=l.npb, nep=l.nep, sorted=l.sorted, uns=l.uniform_nsamples,
                                               law=d.law, nldc=d.nldc,
                                               nk=t.nk, ng=t.ng, nz=t.nz)))
        return 0

    def rr_flux(self, prec, tab, lc, prm, ws, nbytes, flux, stream):
        self.calls.append(('flux', prec, None))
        return 0

    def rr_last_error(self):
        return b''


@pytest.fixture
def fake(monkeypatch):
    fl = FakeLib(_lib.load(require_gpu=False))
    monkeypatch.setattr(_lib, 'load', lambda require_gpu=True: fl)
    monkeypatch.setattr(rrmod, 'stream_handle', lambda device: 0)
    return fl


def test_evaluate_descriptors_population(fake):
    tm = RoadRunnerModel('power-2', precision='mixed', device='cpu')
    tm.set_data(np.linspace(-0.1, 0.1, 50), nsamples=10, exptimes=0.02)
    npv = 6
    out = tm.evaluate(np.full(npv, 0.1), np.tile([[[0.5, 0.6]]], (npv, 1, 1)), np.zeros(npv), np.full(npv, 3.0),
                      np.full(npv, 8.0), np.full(npv, 1.5), copy=False)
    assert out.shape == (npv, 50) and out.dtype == torch.float32
    (kind, prec, d), (kind2, _, _) = fake.calls
    assert kind == 'setup' and kind2 == 'flux' and prec == _lib.RR_MIXED
    assert d == dict(npv=6, kcols=1, t0cols=1, npt=50, nlc=1, npb=1, nep=1, sorted=7, uns=10, law=9, nldc=2,
                     nk=256, ng=100, nz=40)


def test_evaluate_scalars_and_shapes(fake):
    tm = RoadRunnerModel('quadratic', device='cpu')
    tm.set_data(np.linspace(-0.1, 0.1, 20)[::-1].copy())
    f = tm.evaluate(0.1, [0.3, 0.1], 0.0, 3.0, 8.0, 1.5)
    assert isinstance(f, np.ndarray) and f.shape == (20,) and f.dtype == np.float64
    d = fake.calls[0][2]
    assert d['npv'] == 1 and d['sorted'] == 0 and d['law'] == 2


def test_evaluate_multi_passband_and_epochs(fake):
    tm = RoadRunnerModel('quadratic', device='cpu')
    lcids = np.array([0] * 10 + [1] * 10 + [2] * 10)
    tm.set_data(np.linspace(0, 1, 30), lcids=lcids, pbids=[0, 1, 1], epids=[0, 1, 0], nsamples=[1, 2, 3],
                exptimes=[0, 0.01, 0.02])
    npv = 3
    tm.evaluate(np.full((npv, 2), 0.1), np.full((npv, 2, 2), 0.2), np.zeros((npv, 2)), np.full(npv, 3.0),
                np.full(npv, 8.0), np.full(npv, 1.5))
    d = fake.calls[0][2]
    assert (d['kcols'], d['t0cols'], d['nlc'], d['npb'], d['nep'], d['uns']) == (2, 2, 3, 2, 2, 0)
    with pytest.raises(ValueError, match='Radius ratios'):
        tm.evaluate(np.full((npv, 3), 0.1), np.full((npv, 2, 2), 0.2), np.zeros((npv, 2)), np.full(npv, 3.0),
                    np.full(npv, 8.0), np.full(npv, 1.5))
    with pytest.raises(ValueError, match='t0 has 1 columns'):
        tm.evaluate(np.full((npv, 2), 0.1), np.full((npv, 2, 2), 0.2), np.zeros((npv, 1)), np.full(npv, 3.0),
                    np.full(npv, 8.0), np.full(npv, 1.5))


def test_evaluate_host_callable_ld(fake):
    tm = RoadRunnerModel(lambda mu, c: 1 - c[0] * (1 - mu), device='cpu')
    tm.set_data(np.linspace(-0.1, 0.1, 20))
    tm.evaluate(0.1, [0.3], 0.0, 3.0, 8.0, 1.5)
    assert fake.calls[0][2]['law'] == _lib.RR_LD_TABULATED


def test_product_path_fails_loudly_without_gpu():
    if torch.cuda.is_available():
        pytest.skip('a GPU is present')
    tm = RoadRunnerModel('quadratic')
    tm.set_data(np.linspace(-0.1, 0.1, 20))
    with pytest.raises(RuntimeError, match='GPU'):
        tm.evaluate(0.1, [0.3, 0.1], 0.0, 3.0, 8.0, 1.5)


def test_layout_flags():
    from pytransit_amd.transitmodel import layout_flags
    t = np.linspace(0.0, 1.0, 12)
    one = np.zeros(12, int)
    two = np.repeat([0, 1], 6)
    e1, x1 = np.zeros(1, int), np.zeros(1)
    e2, x2 = np.zeros(2, int), np.zeros(2)
    # one sorted light curve: all three bits
    assert layout_flags(t, one, e1, x1) == 7
    # unsorted times: nothing else counts
    assert layout_flags(t[::-1].copy(), one, e1, x1) == 0
    # two light curves in time order, one window
    assert layout_flags(t, two, e2, x2) == 7
    # the same times for each passband: runs sorted, times not
    assert layout_flags(np.concatenate([t[:6], t[:6]]), two, e2, x2) == 1 | 4
    # different exposure or epoch: no common window
    assert layout_flags(t, two, e2, np.array([0.0, 0.01])) == 1 | 2
    assert layout_flags(t, two, np.array([0, 1]), x2) == 1 | 2
    # interleaved light curves: 0
    assert layout_flags(t, np.tile([0, 1], 6), e2, x2) == 0
