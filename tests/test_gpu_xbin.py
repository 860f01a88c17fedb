"""GPU parity of the lazily combined interlaced mesh and of the x pass fused with the binning
(mp_paint_partial -> mp_combine_partial / mp_bin_poles_partial), against the float64 oracle
and against the materialized path (combine written out, then mp_bin_poles).

Tolerances: spectra |dP_l| <= 1e-5 max(|P_l|, |P_0|) per bin; meshes 1e-5 relative to the
largest value; mode counts exact.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import power_oracle as po

from test_gpu_parity import DEV, assert_poles_close, host, say, to_dev

pytestmark = pytest.mark.gpu


def _particles(attrs, size, seed, clumps=True):
    rng = np.random.default_rng(seed)
    lo = attrs.boxcenter - attrs.boxsize / 2.
    pos = rng.uniform(0., 1., (size, 3)) * attrs.boxsize + lo
    if clumps:  # some structure so that the ell > 0 multipoles are not pure noise
        n = size // 3
        centers = rng.uniform(0., 1., (6, 3)) * attrs.boxsize + lo
        pos[:n] = centers[rng.integers(0, 6, n)] + rng.normal(0., 0.04, (n, 3)) * attrs.boxsize
        pos[:n] = lo + np.mod(pos[:n] - lo, attrs.boxsize)
    w = rng.uniform(0.5, 1.5, size).astype('f4')
    return pos, w


CASES = [
    # meshsize, boxsize, resampler, interlacing, compensate, zero_dc, los
    ((64, 64, 64), 1000., 'tsc', 3, True, True, 'z'),
    ((64, 64, 64), 1000., 'cic', 2, True, False, 'x'),
    ((128, 32, 64), (1280., 320., 640.), 'tsc', 3, False, True, (0.3, -0.5, np.sqrt(1 - 0.34))),
    ((32, 64, 16), (320., 640., 160.), 'pcs', 2, True, True, 'y'),
    ((16, 8, 32), (160., 80., 320.), 'cic', 1, True, True, 'z'),
    ((256, 16, 16), (2560., 160., 160.), 'tsc', 3, True, True, 'z'),
    ((4, 8, 8), (40., 80., 80.), 'tsc', 2, True, False, 'z'),
]


@pytest.mark.parametrize('meshsize,boxsize,resampler,interlacing,compensate,zero_dc,los', CASES)
def test_fused_bin_matches_oracle_and_materialized(meshsize, boxsize, resampler, interlacing, compensate, zero_dc, los):
    from jaxpower_amd import MeshAttrs, ParticleField, ComplexMeshField, BinMesh2SpectrumPoles, compute_mesh2_spectrum
    say('fused bin', meshsize, resampler, interlacing, compensate, zero_dc, los)
    attrs = MeshAttrs(meshsize=meshsize, boxsize=boxsize, boxcenter=0.)
    pos, w = _particles(attrs, 20000, sum(meshsize))
    p = ParticleField(to_dev(pos), to_dev(w), attrs=attrs)
    ells = (0, 2, 4)
    kmax = np.min(attrs.kfun) * np.min(attrs.meshsize) / 2.  # default last edge (mesh2.py:39-44)
    step = kmax / 12.
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': step}, ells=ells)
    lazy = p.paint(resampler=resampler, interlacing=interlacing, compensate=compensate, out='complex')
    assert lazy._pending_combine() is not None
    if zero_dc:
        lazy = lazy.at[0, 0, 0].set(0.)
        assert lazy._pending_combine() is not None and lazy._pending_combine().zero_dc
    fused = compute_mesh2_spectrum(lazy, bin=bin, los=los)
    assert lazy._pending_combine() is not None  # the fused path did not write the mesh
    # the same mesh materialized (combine written out), binned by mp_bin_poles
    value = lazy.value
    assert lazy._pending_combine() is None
    plain = compute_mesh2_spectrum(ComplexMeshField(value, attrs=attrs), bin=bin, los=los)
    got = np.array([fused.get(ell).values('num_raw') for ell in ells])
    ref = np.array([plain.get(ell).values('num_raw') for ell in ells])
    assert_poles_close(got, ref, label='fused vs materialized')
    oattrs = po.Attrs(meshsize, attrs.boxsize, attrs.boxcenter)
    omesh = po.paint(oattrs, pos, w.astype('f8'), resampler=resampler, interlacing=interlacing, compensate=compensate,
                     out='complex')
    if zero_dc:
        omesh[0, 0, 0] = 0.
    # the fused paint accumulates in per-block int32 fixed point: with strong clumps and the PCS
    # compensation (up to (pi/2)^12 at the corner mode) the mesh agrees to 1e-5 of its maximum
    err = np.abs(host(value) - omesh).max() / np.abs(omesh).max()
    assert err < 1e-5, err
    want = po.spectrum_poles(oattrs, po.BinPoles(oattrs, edges={'step': step}, ells=ells), omesh, los=los)
    assert np.array_equal(fused.get(0).values('nmodes'), want['nmodes'])
    assert_poles_close(got, want['num_raw'], label='fused vs oracle')


def test_lazy_mesh_semantics():
    from jaxpower_amd import MeshAttrs, ParticleField
    say('lazy mesh semantics')
    attrs = MeshAttrs(meshsize=32, boxsize=320., boxcenter=0.)
    pos, w = _particles(attrs, 5000, 1, clumps=False)
    p = ParticleField(to_dev(pos), to_dev(w), attrs=attrs)
    mesh = p.paint(resampler='tsc', interlacing=2, compensate=True, out='complex')
    zeroed = mesh.at[0, 0, 0].set(0.)
    assert zeroed._pending_combine() is not None and mesh._pending_combine() is not None
    full = host(mesh.value)
    z = host(zeroed.value)
    assert z[0, 0, 0] == 0. and np.array_equal(z[1:], full[1:]) and np.array_equal(z[0, 0, 1:], full[0, 0, 1:])
    assert np.isclose(full[0, 0, 0].real, w.astype('f8').sum(), rtol=1e-6)
    # any other update materializes and applies it
    other = p.paint(resampler='tsc', interlacing=2, compensate=True, out='complex').at[1, 0, 0].set(0.)
    assert other._pending_combine() is None
    o = host(other.value)
    assert o[1, 0, 0] == 0. and np.array_equal(o[2:], full[2:])
    # arithmetic and clone act on the combined values
    assert np.allclose(host((mesh * 2.).value), 2. * full)
    lazy = p.paint(resampler='tsc', interlacing=2, compensate=True, out='complex')
    assert lazy.clone()._pending_combine() is not None
    assert lazy.shape == (32, 32, 17)


def _bin_abi(lib, stream, partial, pitch, n, order, ni, comp, zero_dc, t, nedges, los, ells, rz, out, ws):
    i32 = lambda v: (ctypes.c_int32 * len(v))(*v)
    return lib.mp_bin_poles_partial(stream, partial.data_ptr(), pitch, i32(n), order, ni, comp, zero_dc, t['kx'].data_ptr(),
                                    t['ky'].data_ptr(), t['kz'].data_ptr(), t['thr'].data_ptr(), nedges,
                                    (ctypes.c_double * 3)(*los), len(ells), i32(ells), i32(rz), out.data_ptr(), ws.data_ptr(),
                                    ws.numel())


def test_bin_poles_partial_abi():
    """The C entry point directly: six products (ells 0..4 and 6), output layout [p][b][re, im]
    with im = 0, against mp_combine_partial + mp_bin_poles; argument errors."""
    from jaxpower_amd import MeshAttrs, BinMesh2SpectrumPoles, _lib
    say('bin_poles_partial abi')
    lib = _lib.load()
    n = (64, 32, 32)
    attrs = MeshAttrs(meshsize=n, boxsize=(640., 320., 320.))
    pos, w = _particles(attrs, 8000, 9)
    pitch = lib.mp_spectrum_pitch(n[2])
    ni, order = 3, 3
    partial = torch.empty((ni, n[0], n[1], pitch), dtype=torch.complex64, device=DEV)
    pos_t, w_t = to_dev(pos), to_dev(w)
    stream = _lib.stream_handle()
    n3, box, center = _lib.i32x3(n), _lib.f64x3(attrs.boxsize), _lib.f64x3(attrs.boxcenter)
    ws = _lib.workspace(lib.mp_paint_partial_workspace_bytes(n3, len(pos), order, ni), DEV)
    _lib.check(lib.mp_paint_partial(stream, pos_t.data_ptr(), 1, w_t.data_ptr(), len(pos), n3, box, center, order, ni,
                                    partial.data_ptr(), pitch, ws.data_ptr(), ws.numel()))
    ells = [0, 1, 2, 3, 4, 6]
    rz = [0, 1, 1, 1, 1, 1]
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.02}, ells=0)
    t, nedges = bin._tables, len(bin._edges1d)
    los = (0., 0.6, 0.8)
    out = torch.full((len(ells), bin.nbins, 2), np.nan, dtype=torch.float64, device=DEV)
    ws = _lib.workspace(lib.mp_bin_partial_workspace_bytes(n3, ni, len(ells), nedges), DEV)
    _lib.check(_bin_abi(lib, stream, partial, pitch, n, order, ni, 1, 1, t, nedges, los, ells, rz, out, ws))
    got = host(out)
    assert np.all(got[..., 1] == 0.)
    combined = torch.empty((n[0], n[1], pitch), dtype=torch.complex64, device=DEV)
    cws = _lib.workspace(lib.mp_combine_partial_workspace_bytes(n3, ni), DEV)
    _lib.check(lib.mp_combine_partial(stream, partial.data_ptr(), pitch, n3, order, ni, 1, 1, combined.data_ptr(), pitch,
                                      cws.data_ptr(), cws.numel()))
    ref = bin.products([combined[..., :n[2] // 2 + 1]], [(0, 0, ell, r) for ell, r in zip(ells, rz)], los=los)
    assert_poles_close(got[..., 0], ref.real, label='abi')
    # errors: too many products, too many bins for the LDS histogram, bad order
    assert _bin_abi(lib, stream, partial, pitch, n, order, ni, 1, 1, t, nedges, los, ells + [2], rz + [1], out, ws) == -1
    big = torch.zeros(5000, dtype=torch.float64, device=DEV)
    tb = dict(t, thr=big)
    assert _bin_abi(lib, stream, partial, pitch, n, order, ni, 1, 1, tb, 5000, los, ells, rz, out, ws) == -2
    assert _bin_abi(lib, stream, partial, pitch, n, 7, ni, 1, 1, t, nedges, los, ells, rz, out, ws) == -1
    torch.cuda.synchronize()


@pytest.mark.parametrize('los,cross,boxcenter', [('firstpoint', False, (500., -200., 1400.)), ('endpoint', True, (500., -200., 1400.)),
                                                 ('firstpoint', True, (500., -200., 1400.)), ('firstpoint', False, 0.)])
def test_local_los_fused_matches_stored(los, cross, boxcenter, monkeypatch):
    """The fused local line of sight (mp_ylm_bin: C_ell never stored) against the path that
    stores C_ell and bins it with mp_bin_poles, and against the oracle (boxcenter 0: the origin
    is a node)."""
    from jaxpower_amd import MeshAttrs, RealMeshField, BinMesh2SpectrumPoles, compute_mesh2_spectrum
    from jaxpower_amd import mesh2 as m2
    say('local fused', los, cross, boxcenter)
    rng = np.random.default_rng(44)
    shape = (64, 32, 32)
    attrs = MeshAttrs(meshsize=shape, boxsize=(1280., 640., 640.), boxcenter=boxcenter)
    m1 = rng.normal(size=shape).astype('f4')
    m2_ = (0.6 * m1 + 0.5 * rng.normal(size=shape)).astype('f4')
    ells = (0, 2, 4)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.005}, ells=ells)
    meshes = [RealMeshField(to_dev(m1), attrs=attrs)] + ([RealMeshField(to_dev(m2_), attrs=attrs)] if cross else [])
    fused = compute_mesh2_spectrum(*meshes, bin=bin, los=los)
    orig = m2._ylm_bin_all
    monkeypatch.setattr(m2, '_ylm_bin_all', lambda *args, **kw: None)
    stored = compute_mesh2_spectrum(*meshes, bin=bin, los=los)
    monkeypatch.setattr(m2, '_ylm_bin_all', orig)
    got = np.array([fused.get(ell).values('num_raw') for ell in ells])
    ref = np.array([stored.get(ell).values('num_raw') for ell in ells])
    assert_poles_close(got, ref, rtol=1e-5, label='fused vs stored')
    oattrs = po.Attrs(shape, attrs.boxsize, attrs.boxcenter)
    want = po.spectrum_poles(oattrs, po.BinPoles(oattrs, edges={'step': 0.005}, ells=ells), m1.astype('f8'),
                             m2_.astype('f8') if cross else None, los=los)
    assert_poles_close(got, want['num_raw'], rtol=1e-5, label='fused vs oracle')


@pytest.mark.parametrize('factor', [0.7, 0.3, 0.12])
@pytest.mark.parametrize('ells', [(0, 2, 4), (0, 1, 2)])
def test_fused_bin_bins_narrower_than_kx_spacing(factor, ells):
    """Bins of factor * kf: the fused x pass moves the bin by two conditional thresholds per kx
    step (0.7 kf), by a fixed number of them (0.3 kf: four), or walks (0.12 kf); even multipoles
    take the lean binning, odd ones the general loop.  Exact mode counts and the oracle's
    multipoles."""
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh2SpectrumPoles, compute_mesh2_spectrum
    say('fused bin narrow bins', factor, ells)
    attrs = MeshAttrs(meshsize=64, boxsize=1000., boxcenter=0.)
    pos, w = _particles(attrs, 20000, 7)
    p = ParticleField(to_dev(pos), to_dev(w), attrs=attrs)
    step = factor * float(np.min(attrs.kfun))
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': step}, ells=ells)
    lazy = p.paint(resampler='tsc', interlacing=3, compensate=True, out='complex').at[0, 0, 0].set(0.)
    assert lazy._pending_combine() is not None
    fused = compute_mesh2_spectrum(lazy, bin=bin, los='z')
    got = np.array([fused.get(ell).values('num_raw') for ell in ells])
    oattrs = po.Attrs(attrs.meshsize, attrs.boxsize, attrs.boxcenter)
    omesh = po.paint(oattrs, pos, w.astype('f8'), resampler='tsc', interlacing=3, compensate=True, out='complex')
    omesh[0, 0, 0] = 0.
    want = po.spectrum_poles(oattrs, po.BinPoles(oattrs, edges={'step': step}, ells=ells), omesh, los='z')
    assert np.array_equal(fused.get(0).values('nmodes'), want['nmodes'])
    assert_poles_close(got, want['num_raw'], label=f'narrow bins {factor} kf')
