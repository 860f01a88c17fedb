"""GPU parity tests: the HIP path (through the C ABI, via the package) against the float64
oracle and the committed golden fixtures.

Tolerances (stated per test): FFTs and meshes are FP32, so they are compared in norm relative
to the largest value; spectra follow |dP_l| <= 1e-5 max(|P_l|, |P_0|) per bin; integer mode
counts are compared exactly.
"""
import os
import sys

import numpy as np
import pytest
import torch

from oracle import power_oracle as po

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DEV = 'cuda'


def say(*args):
    print('[case]', *args, flush=True)
    sys.stdout.flush()


def to_dev(x, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(x)).to(DEV)
    return t.to(dtype) if dtype is not None else t


def host(x):
    return x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)


def assert_poles_close(got, want, rtol=1e-5, label=''):
    """|dP_l| <= rtol * max(|P_l|, |P_0|) per bin; NaN (empty bins) must match."""
    got, want = np.asarray(got), np.asarray(want)
    p0 = np.abs(want[0])
    errs = []
    for g, w in zip(got, want):
        ok = ~np.isnan(w)
        errs.append(float(np.max(np.abs(g - w)[ok] / np.maximum(np.maximum(np.abs(w), p0)[ok], 1e-300), initial=0.)))
    print(f'[err] {label} max |dP_l|/max(|P_l|,|P_0|) per multipole: ' + ' '.join(f'{e:.2e}' for e in errs), flush=True)
    for ill, (g, w) in enumerate(zip(got, want)):
        nan = np.isnan(w)
        assert np.array_equal(nan, np.isnan(g)), f'{label} ell index {ill}: NaN pattern differs'
        scale = np.maximum(np.abs(w), p0)[~nan]
        err = np.abs(g - w)[~nan]
        bad = err > rtol * scale + 1e-30
        assert not np.any(bad), f'{label} ell index {ill}: max rel err {np.max(err / np.maximum(scale, 1e-300)):.3e}'


# ----------------------------------------------------------------------------- FFT

@pytest.mark.parametrize('shape', [(8, 8, 8), (4, 16, 32), (32, 32, 32), (64, 64, 64), (128, 32, 64), (16, 16, 256), (2, 2, 2)])
def test_r2c_matches_numpy(shape):
    from jaxpower_amd.mesh import _r2c
    say('r2c', shape)
    rng = np.random.default_rng(sum(shape))
    x = rng.normal(size=shape).astype('f4')
    got = host(_r2c(to_dev(x), shape))
    want = np.fft.rfftn(x.astype('f8'))
    err = np.abs(got - want).max() / np.abs(want).max()
    assert err < 2e-6, err
    # rocFFT (through torch) as a cross-check of the same transform
    roc = host(torch.fft.rfftn(to_dev(x)))
    assert np.abs(got - roc).max() / np.abs(roc).max() < 2e-6


@pytest.mark.parametrize('shape', [(8, 8, 8), (32, 16, 64), (64, 64, 64), (16, 128, 32)])
def test_c2r_matches_numpy(shape):
    from jaxpower_amd.mesh import _c2r
    say('c2r', shape)
    rng = np.random.default_rng(7)
    c = (rng.normal(size=(shape[0], shape[1], shape[2] // 2 + 1)) + 1j * rng.normal(size=(shape[0], shape[1], shape[2] // 2 + 1)))
    got = host(_c2r(to_dev(c.astype('c8')), shape))
    want = np.fft.irfftn(c, s=shape, axes=(0, 1, 2))
    assert np.abs(got - want).max() / np.abs(want).max() < 2e-6


def test_fft_roundtrip():
    from jaxpower_amd import RealMeshField
    say('roundtrip 64^3')
    rng = np.random.default_rng(3)
    x = rng.normal(size=(64, 64, 64)).astype('f4')
    mesh = RealMeshField(to_dev(x), boxsize=1000.)
    back = host(mesh.r2c().c2r().value)
    assert np.abs(back - x).max() < 1e-5


# ----------------------------------------------------------------------------- paint

def _grid(attrs, pos):
    return po.to_grid(po.Attrs(attrs.meshsize, attrs.boxsize, attrs.boxcenter), pos)


@pytest.mark.parametrize('shape', [(8, 8, 8), (32, 32, 32), (48, 16, 64), (20, 12, 10)])
@pytest.mark.parametrize('order', [1, 2, 3, 4])
@pytest.mark.parametrize('ni', [1, 2, 3])
def test_paint_meshes_match_oracle(shape, order, ni):
    from jaxpower_amd import _lib, MeshAttrs
    say('paint', shape, order, ni)
    rng = np.random.default_rng(order * 10 + ni)
    attrs = MeshAttrs(meshsize=shape, boxsize=np.array(shape) * 3.5, boxcenter=(1., -2., 3.))
    npart = 3000
    lo = attrs.boxcenter - attrs.boxsize / 2.
    pos = rng.uniform(-0.1, 1.1, (npart, 3)) * attrs.boxsize + lo  # some outside the box: periodic wrap
    pos[:10] = lo  # exactly on the low corner
    w = rng.uniform(0.5, 1.5, npart).astype('f4')
    out = torch.empty((ni,) + shape, dtype=torch.float32, device=DEV)
    n = _lib.i32x3(shape)
    ws = _lib.workspace(_lib.load().mp_paint_workspace_bytes(n, npart, order, ni), DEV)
    dpos, dw = to_dev(pos), to_dev(w)
    _lib.call('mp_paint', _lib.stream_handle(), _lib.ptr(dpos), 1, _lib.ptr(dw), npart, n, _lib.f64x3(attrs.boxsize),
              _lib.f64x3(attrs.boxcenter), order, ni, _lib.ptr(out), _lib.ptr(ws), ws.numel())
    atomic = torch.empty_like(out)
    _lib.call('mp_paint_atomic', _lib.stream_handle(), _lib.ptr(dpos), 1, _lib.ptr(dw), npart, n, _lib.f64x3(attrs.boxsize),
              _lib.f64x3(attrs.boxcenter), order, ni, _lib.ptr(atomic))
    got, got_atomic = host(out), host(atomic)
    xg = _grid(attrs, pos)
    for j in range(ni):
        want = po.paint_grid(shape, xg + j / ni, w.astype('f8'), order)
        scale = np.abs(want).max()
        assert np.abs(got[j] - want).max() <= 2e-6 * scale, (j, np.abs(got[j] - want).max() / scale)
        assert np.abs(got_atomic[j] - want).max() <= 2e-6 * scale
        assert np.isclose(got[j].sum(dtype='f8'), w.sum(dtype='f8'), rtol=1e-5)


def test_paint_float32_positions_and_unit_weights():
    from jaxpower_amd import MeshAttrs, ParticleField
    say('paint f32 tsc 64^3')
    rng = np.random.default_rng(11)
    attrs = MeshAttrs(meshsize=64, boxsize=2000., boxcenter=1000.)
    pos = rng.uniform(0., 2000., (20000, 3)).astype('f4')
    mesh = ParticleField(to_dev(pos), attrs=attrs).paint(resampler='tsc', interlacing=0)
    want = po.paint_grid(attrs.meshsize, _grid(attrs, pos.astype('f8')), 1., 3)
    # float32 grid coordinates near x~ = 64 carry ~4e-6 cells of rounding: 1e-4 of the peak
    assert np.abs(host(mesh.value) - want).max() <= 1e-4 * np.abs(want).max()
    assert np.isclose(float(mesh.sum()), 20000., rtol=1e-5)


@pytest.mark.parametrize('ni', [1, 2, 3])
def test_ngp_exact_ties_round_half_even(ni):
    """Particles exactly on cell centres and cell boundaries: NGP (and the interlaced shift of
    1/2) lands on exact halves, which jnp.round sends to the even cell.  Checked for the
    real-space row paint and the fused paint + z-FFT path."""
    from jaxpower_amd import _lib, MeshAttrs, ParticleField
    say('ngp ties', ni)
    shape = (16, 16, 32)
    attrs = MeshAttrs(meshsize=shape, boxsize=(160., 160., 320.), boxcenter=(80., 80., 160.))
    axes = [np.arange(n) * 10. for n in shape]
    centres = np.stack(np.meshgrid(*axes, indexing='ij'), -1).reshape(-1, 3)
    pos = np.concatenate([centres, centres + 5., centres + np.array([5., 0., 5.])])
    rng = np.random.default_rng(3)
    w = rng.uniform(0.5, 1.5, len(pos)).astype('f4')
    xg = _grid(attrs, pos)
    assert np.all(xg * 2. == np.round(xg * 2.))  # exact halves
    out = torch.empty((ni,) + shape, dtype=torch.float32, device=DEV)
    n = _lib.i32x3(shape)
    ws = _lib.workspace(_lib.load().mp_paint_workspace_bytes(n, len(pos), 1, ni), DEV)
    dpos, dw = to_dev(pos), to_dev(w)  # held until the call has run
    _lib.call('mp_paint', _lib.stream_handle(), _lib.ptr(dpos), 1, _lib.ptr(dw), len(pos), n,
              _lib.f64x3(attrs.boxsize), _lib.f64x3(attrs.boxcenter), 1, ni, _lib.ptr(out), _lib.ptr(ws), ws.numel())
    got = host(out)
    for j in range(ni):
        want = po.paint_grid(shape, xg + j / ni, w.astype('f8'), 1)
        assert np.abs(got[j] - want).max() <= 2e-6 * np.abs(want).max(), j
    p = ParticleField(to_dev(pos), to_dev(w), attrs=attrs)
    got = host(p.paint(resampler='ngp', interlacing=ni if ni > 1 else 0, compensate=True, out='complex').value)
    want = po.paint(po.Attrs(shape, attrs.boxsize, attrs.boxcenter), pos, w.astype('f8'), resampler='ngp',
                    interlacing=ni if ni > 1 else 0, compensate=True, out='complex')
    assert np.abs(got - want).max() <= 3e-6 * np.abs(want).max()


@pytest.mark.parametrize('resampler', ['cic', 'tsc', 'pcs'])
@pytest.mark.parametrize('interlacing', [0, 2, 3])
@pytest.mark.parametrize('compensate', [False, True])
def test_paint_spectrum_matches_oracle(resampler, interlacing, compensate):
    from jaxpower_amd import MeshAttrs, ParticleField
    say('paint spectrum', resampler, interlacing, compensate)
    rng = np.random.default_rng(5)
    attrs = MeshAttrs(meshsize=(32, 16, 64), boxsize=(320., 160., 640.), boxcenter=(0., 50., 320.))
    lo = attrs.boxcenter - attrs.boxsize / 2.
    pos = rng.uniform(0., 1., (5000, 3)) * attrs.boxsize + lo
    w = rng.uniform(0.5, 1.5, 5000).astype('f4')
    p = ParticleField(to_dev(pos), to_dev(w), attrs=attrs)
    oattrs = po.Attrs(attrs.meshsize, attrs.boxsize, attrs.boxcenter)
    for out in ('complex', 'real'):
        got = host(p.paint(resampler=resampler, interlacing=interlacing, compensate=compensate, out=out).value)
        want = po.paint(oattrs, pos, w.astype('f8'), resampler=resampler, interlacing=interlacing, compensate=compensate, out=out)
        assert got.shape == want.shape
        err = np.abs(got - want).max() / np.abs(want).max()
        assert err < 3e-6, (out, err)


def test_read_matches_oracle():
    from jaxpower_amd import MeshAttrs, RealMeshField
    say('read')
    rng = np.random.default_rng(9)
    attrs = MeshAttrs(meshsize=(16, 32, 8), boxsize=(100., 200., 50.), boxcenter=0.)
    mesh = rng.normal(size=(16, 32, 8)).astype('f4')
    pos = rng.uniform(-60., 60., (1000, 3))
    oattrs = po.Attrs(attrs.meshsize, attrs.boxsize, attrs.boxcenter)
    for resampler in ['ngp', 'cic', 'tsc', 'pcs']:
        for compensate in (False, True):
            got = host(RealMeshField(to_dev(mesh), attrs=attrs).read(to_dev(pos), resampler=resampler, compensate=compensate))
            want = po.read(oattrs, mesh.astype('f8'), pos, resampler=resampler, compensate=compensate)
            assert np.abs(got - want).max() < 1e-5 * np.abs(want).max()


# ----------------------------------------------------------------------------- binning

@pytest.mark.parametrize('meshsize,boxsize,edges', [
    (64, 1000., {'step': 0.01}), (32, 2000., {'step': 0.005}), ((16, 32, 8), (100., 200., 50.), {'step': 0.02}),
    (16, 1000., {}), (32, 500., np.array([0., 0.01, 0.05, 0.1, 0.2, 0.3])),
    (64, 64., {'step': 0.0005})])  # 6283 bins: more than the LDS histogram holds (global path)
def test_bin_modes_exact(meshsize, boxsize, edges):
    from jaxpower_amd import MeshAttrs, BinMesh2SpectrumPoles
    say('bin modes', meshsize, boxsize)
    bin = BinMesh2SpectrumPoles(MeshAttrs(meshsize=meshsize, boxsize=boxsize), edges=edges, ells=(0, 2, 4))
    ref = po.BinPoles(po.Attrs(meshsize, boxsize), edges=edges, ells=(0, 2, 4))
    assert np.array_equal(bin.edges, ref.edges)
    assert bin.nmodes.dtype == np.int64 and np.array_equal(bin.nmodes, ref.nmodes)
    assert np.allclose(bin.xavg, ref.xavg, rtol=1e-12, equal_nan=True)


@pytest.mark.parametrize('los', ['x', 'y', 'z', (0.3, -0.5, np.sqrt(1 - 0.34))])
@pytest.mark.parametrize('cross', [False, True])
def test_spectrum_global_los(los, cross):
    from jaxpower_amd import MeshAttrs, RealMeshField, BinMesh2SpectrumPoles, compute_mesh2_spectrum
    say('spectrum global', los, cross)
    rng = np.random.default_rng(21)
    shape = (32, 64, 16)
    attrs = MeshAttrs(meshsize=shape, boxsize=(500., 1000., 250.))
    m1 = rng.normal(size=shape).astype('f4')
    m2 = (m1 + 0.5 * rng.normal(size=shape)).astype('f4')
    ells = (0, 1, 2, 3, 4)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.02}, ells=ells)
    meshes = [RealMeshField(to_dev(m1), attrs=attrs)] + ([RealMeshField(to_dev(m2), attrs=attrs)] if cross else [])
    got = compute_mesh2_spectrum(*meshes, bin=bin, los=los)
    oattrs = po.Attrs(shape, attrs.boxsize)
    obin = po.BinPoles(oattrs, edges={'step': 0.02}, ells=ells)
    want = po.spectrum_poles(oattrs, obin, m1.astype('f8'), m2.astype('f8') if cross else None, los=los)
    assert got.ells == ells
    assert_poles_close([got.get(ell).values('num_raw') for ell in ells], want['num_raw'], label=f'{los} cross={cross}')
    assert np.array_equal(got.get(0).values('nmodes'), want['nmodes'])
    assert np.allclose(got.get(0).values('norm'), want['norm'])


@pytest.mark.parametrize('los', ['firstpoint', 'endpoint'])
@pytest.mark.parametrize('cross', [False, True])
@pytest.mark.parametrize('boxcenter', [(300., -100., 900.), 0.])
def test_spectrum_local_los(los, cross, boxcenter):
    """Local line of sight against the oracle; with boxcenter 0 the origin is a node, where the
    reference's Y_40 is its constant term (not 0)."""
    from jaxpower_amd import MeshAttrs, RealMeshField, BinMesh2SpectrumPoles, compute_mesh2_spectrum
    say('spectrum local', los, cross, boxcenter)
    rng = np.random.default_rng(22)
    shape = (32, 16, 32)
    attrs = MeshAttrs(meshsize=shape, boxsize=(800., 400., 800.), boxcenter=boxcenter)
    m1 = rng.normal(size=shape).astype('f4')
    m2 = (0.7 * m1 + 0.5 * rng.normal(size=shape)).astype('f4')
    ells = (0, 2, 4)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.01}, ells=ells)
    meshes = [RealMeshField(to_dev(m1), attrs=attrs)] + ([RealMeshField(to_dev(m2), attrs=attrs)] if cross else [])
    got = compute_mesh2_spectrum(*meshes, bin=bin, los=los)
    oattrs = po.Attrs(shape, attrs.boxsize, attrs.boxcenter)
    obin = po.BinPoles(oattrs, edges={'step': 0.01}, ells=ells)
    want = po.spectrum_poles(oattrs, obin, m1.astype('f8'), m2.astype('f8') if cross else None, los=los)
    assert_poles_close([got.get(ell).values('num_raw') for ell in ells], want['num_raw'], rtol=1e-5, label=f'{los} cross={cross}')


def test_bin_call_plain_values():
    from jaxpower_amd import MeshAttrs, BinMesh2SpectrumPoles
    say('bin call')
    attrs = MeshAttrs(meshsize=16, boxsize=100.)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.1})
    obin = po.BinPoles(po.Attrs(16, 100.), edges={'step': 0.1})
    rng = np.random.default_rng(4)
    v = rng.normal(size=(16, 16, 9)) + 1j * rng.normal(size=(16, 16, 9))
    got = bin(to_dev(v.astype('c8')), remove_zero=True)
    want = obin(v, remove_zero=True)
    assert np.allclose(got, want, rtol=1e-5, atol=1e-6, equal_nan=True)


# ----------------------------------------------------------------------------- golden fixtures and end to end

def _load(name):
    return dict(np.load(os.path.join(HERE, 'golden', name), allow_pickle=False))


@pytest.mark.parametrize('name', ['box_cic_n32.npz', 'box_tsc3_n32.npz', 'box_pcs2_n16x32x16_losx.npz'])
def test_golden_box(name):
    from jaxpower_amd import (MeshAttrs, ParticleField, BinMesh2SpectrumPoles, compute_mesh2_spectrum,
                              compute_box2_normalization, compute_fkp2_shotnoise)
    say('golden', name)
    g = _load(name)
    attrs = MeshAttrs(meshsize=g['meshsize'], boxsize=g['boxsize'], boxcenter=g['boxcenter'])
    p = ParticleField(to_dev(g['positions']), to_dev(g['weights']), attrs=attrs)
    mesh = p.paint(resampler=str(g['resampler']), interlacing=int(g['interlacing']), compensate=True, out='complex')
    mesh = mesh.at[0, 0, 0].set(0.)
    ells = tuple(int(e) for e in g['ells'])
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': float(g['step'])}, ells=ells)
    spec = compute_mesh2_spectrum(mesh, bin=bin, los=str(g['los']))
    spec = spec.clone(norm=compute_box2_normalization(p, bin=bin), num_shotnoise=compute_fkp2_shotnoise(p, bin=bin))
    assert np.array_equal(spec.get(0).values('nmodes'), g['nmodes'])
    assert np.allclose(spec.get(0).coords('k'), g['k'], rtol=1e-12, equal_nan=True)
    assert_poles_close([spec.get(ell).values('num_raw') for ell in ells], g['num_raw'], label=name)
    assert np.isclose(spec.get(0).values('norm')[0], float(g['norm']), rtol=1e-6)
    assert np.isclose(spec.get(0).values('num_shotnoise')[0], float(g['shotnoise']), rtol=1e-6)


def test_golden_fkp_firstpoint():
    from jaxpower_amd import (MeshAttrs, ParticleField, FKPField, BinMesh2SpectrumPoles, compute_mesh2_spectrum,
                              compute_fkp2_normalization, compute_fkp2_shotnoise)
    say('golden fkp firstpoint')
    g = _load('fkp_tsc3_n32_firstpoint.npz')
    attrs = MeshAttrs(meshsize=g['meshsize'], boxsize=g['boxsize'], boxcenter=g['boxcenter'])
    data = ParticleField(to_dev(g['data']), to_dev(g['data_weights']), attrs=attrs)
    randoms = ParticleField(to_dev(g['randoms']), to_dev(g['randoms_weights']), attrs=attrs)
    fkp = FKPField(data, randoms)
    ells = tuple(int(e) for e in g['ells'])
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': float(g['step'])}, ells=ells)
    mesh = fkp.paint(resampler='tsc', interlacing=3, compensate=True, out='real')
    spec = compute_mesh2_spectrum(mesh, bin=bin, los='firstpoint')
    norm = compute_fkp2_normalization(fkp, bin=bin)
    sn = compute_fkp2_shotnoise(fkp, bin=bin)
    assert np.array_equal(spec.get(0).values('nmodes'), g['nmodes'])
    assert_poles_close([spec.get(ell).values('num_raw') for ell in ells], g['num_raw'], rtol=1e-5, label='fkp')
    assert np.isclose(norm[0], float(g['norm']), rtol=1e-5)
    assert np.isclose(sn[0][0], float(g['shotnoise']), rtol=1e-6)


def test_box_64_cic_end_to_end():
    """The CPU-runnable box: 64^3, L=1000, 1e5 float64 uniform particles, CIC, no interlacing,
    compensated, los z, ells (0,2,4), step 0.01."""
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh2SpectrumPoles, compute_mesh2_spectrum
    say('box 64 cic e2e')
    rng = np.random.default_rng(42)
    pos = rng.uniform(0., 1000., (100000, 3))
    attrs = MeshAttrs(meshsize=64, boxsize=1000., boxcenter=500.)
    p = ParticleField(to_dev(pos), attrs=attrs)
    mesh = p.paint(resampler='cic', interlacing=0, compensate=True, out='real')
    mesh = mesh - mesh.mean()
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.01}, ells=(0, 2, 4))
    spec = compute_mesh2_spectrum(mesh, bin=bin, los='z')
    oattrs = po.Attrs(64, 1000., 500.)
    omesh = po.paint(oattrs, pos, None, resampler='cic', interlacing=0, compensate=True)
    omesh = omesh - omesh.mean()
    want = po.spectrum_poles(oattrs, po.BinPoles(oattrs, edges={'step': 0.01}, ells=(0, 2, 4)), omesh, los='z')
    assert np.array_equal(spec.get(0).values('nmodes'), want['nmodes'])
    assert_poles_close([spec.get(ell).values('num_raw') for ell in (0, 2, 4)], want['num_raw'], label='cfg1')


def test_fkp_normalization_matches_oracle():
    from jaxpower_amd import MeshAttrs, ParticleField, FKPField, compute_fkp2_normalization
    say('fkp normalization')
    rng = np.random.default_rng(12)
    attrs = MeshAttrs(meshsize=64, boxsize=1005., boxcenter=(10., 20., 30.))
    d = rng.uniform(-400., 400., (20000, 3)) + attrs.boxcenter
    r = rng.uniform(-400., 400., (60000, 3)) + attrs.boxcenter
    wd, wr = rng.uniform(0.5, 1., 20000).astype('f4'), rng.uniform(0.5, 1., 60000).astype('f4')
    fkp = FKPField(ParticleField(to_dev(d), to_dev(wd), attrs=attrs), ParticleField(to_dev(r), to_dev(wr), attrs=attrs))
    for cellsize in (10., None):
        got = compute_fkp2_normalization(fkp, cellsize=cellsize)
        want = po.fkp_normalization(po.Attrs(64, 1005., (10., 20., 30.)), d, wd.astype('f8'), r, wr.astype('f8'), cellsize=cellsize)
        assert np.isclose(got, want, rtol=1e-5), (cellsize, got, want)


def test_fkp3_normalization_matches_oracle():
    from jaxpower_amd import MeshAttrs, ParticleField, FKPField, compute_fkp3_normalization, BinMesh3SpectrumPoles
    say('fkp3 normalization')
    rng = np.random.default_rng(13)
    attrs = MeshAttrs(meshsize=64, boxsize=1005., boxcenter=(10., 20., 30.))
    d = rng.uniform(-400., 400., (20000, 3)) + attrs.boxcenter
    r = rng.uniform(-400., 400., (60000, 3)) + attrs.boxcenter
    wd, wr = rng.uniform(0.5, 1., 20000).astype('f4'), rng.uniform(0.5, 1., 60000).astype('f4')
    fkp = FKPField(ParticleField(to_dev(d), to_dev(wd), attrs=attrs), ParticleField(to_dev(r), to_dev(wr), attrs=attrs))
    for cellsize in (10., None):
        got = float(compute_fkp3_normalization(fkp, cellsize=cellsize))
        want = po.fkp3_normalization(po.Attrs(64, 1005., (10., 20., 30.)), d, wd.astype('f8'), r, wr.astype('f8'), cellsize=cellsize)
        assert np.isclose(got, want, rtol=1e-5), (cellsize, got, want)
    bin = BinMesh3SpectrumPoles(attrs, edges={'step': 0.02}, ells=(0, 2), basis='scoccimarro')
    norms = compute_fkp3_normalization(fkp, fkp, fkp, bin=bin)
    assert len(norms) == 2 and np.isclose(float(norms[1]), float(compute_fkp3_normalization(fkp)), rtol=1e-12)


def test_lognormal_reduced_tsc3():
    """The headline recipe at reduced size (128^3, L=500, ~4e5 particles), TSC x3, los z."""
    from jaxpower_amd import MeshAttrs, BinMesh2SpectrumPoles, compute_mesh2_spectrum, generate_lognormal_particles
    say('lognormal 128 tsc3')
    attrs = MeshAttrs(meshsize=128, boxsize=500., boxcenter=250.)
    p = generate_lognormal_particles(attrs, size=400000, seed=3)
    mesh = p.paint(resampler='tsc', interlacing=3, compensate=True, out='complex').at[0, 0, 0].set(0.)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.01}, ells=(0, 2, 4))
    spec = compute_mesh2_spectrum(mesh, bin=bin, los='z')
    pos = host(p.positions).astype('f8')
    oattrs = po.Attrs(128, 500., 250.)
    omesh = po.paint(oattrs, pos, None, resampler='tsc', interlacing=3, compensate=True, out='complex')
    omesh[0, 0, 0] = 0.
    want = po.spectrum_poles(oattrs, po.BinPoles(oattrs, edges={'step': 0.01}, ells=(0, 2, 4)), omesh, los='z')
    assert_poles_close([spec.get(ell).values('num_raw') for ell in (0, 2, 4)], want['num_raw'], label='lognormal')
