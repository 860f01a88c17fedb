"""GPU parity of the correlation-function multipoles (compute_mesh2_correlation,
BinMesh2CorrelationPoles; jaxpower/mesh2.py:120-184, 377-504) against the float64 oracle.

Tolerances: separation-bin cell counts bit-exact, bin-averaged |s| to 1e-12 relative.  The
multipoles come from float32 meshes and float32 FFTs, whose rounding is a fraction ~1e-7 of
the zero-lag value spread over every lag, so each bin is compared as
|dxi_l(s)| <= 1e-5 |xi_l(s)| + 2e-6 max_s |xi_0(s)| (the first bin holds the zero lag).
"""
import numpy as np
import pytest
import torch

from oracle import power_oracle as po

from test_gpu_parity import DEV, host, say, to_dev

pytestmark = pytest.mark.gpu


def assert_xi_close(got, want, rtol=1e-5, ftol=2e-6, label=''):
    got, want = np.asarray(got), np.asarray(want)
    floor = ftol * np.nanmax(np.abs(want[0]))
    for ill, (g, w) in enumerate(zip(got, want)):
        nan = np.isnan(w)
        assert np.array_equal(nan, np.isnan(g)), f'{label} ell index {ill}: NaN pattern differs'
        err = np.abs(g - w)[~nan]
        bad = err > rtol * np.abs(w)[~nan] + floor
        assert not np.any(bad), f'{label} ell index {ill}: max err {err.max():.3e}, floor {floor:.3e}'


def _mock(attrs, size, seed):
    rng = np.random.default_rng(seed)
    lo = attrs.boxcenter - attrs.boxsize / 2.
    pos = rng.uniform(0., 1., (size, 3)) * attrs.boxsize + lo
    n = size // 2
    centers = rng.uniform(0., 1., (5, 3)) * attrs.boxsize + lo
    pos[:n] = centers[rng.integers(0, 5, n)] + rng.normal(0., 0.04, (n, 3)) * attrs.boxsize
    pos[:n] = lo + np.mod(pos[:n] - lo, attrs.boxsize)
    w = rng.uniform(0.5, 1.5, size).astype('f4')
    return pos, w


@pytest.mark.parametrize('meshsize,boxsize,edges', [
    ((32, 32, 32), 320., {'step': 7.}),
    ((16, 32, 24), (160., 320., 240.), {'step': 9., 'min': 3.}),
    ((16, 16, 16), 160., None),  # unique edges
    ((32, 16, 64), (320., 240., 640.), np.linspace(0., 150., 31)),
])
def test_bin_real_modes_exact(meshsize, boxsize, edges):
    from jaxpower_amd import MeshAttrs, BinMesh2CorrelationPoles
    say('corr modes', meshsize, boxsize)
    attrs = MeshAttrs(meshsize=meshsize, boxsize=boxsize, boxcenter=0.)
    bin = BinMesh2CorrelationPoles(attrs, edges=edges, ells=(0, 2))
    obin = po.BinCorrelationPoles(po.Attrs(meshsize, boxsize, 0.), edges=edges, ells=(0, 2))
    assert np.array_equal(bin.edges, obin.edges)
    assert np.array_equal(bin.nmodes, obin.nmodes)
    ok = obin.nmodes > 0
    assert np.allclose(bin.xavg[ok], obin.xavg[ok], rtol=1e-12, atol=0)


@pytest.mark.parametrize('resampler,interlacing,los', [('cic', 0, 'z'), ('tsc', 3, 'x'), ('pcs', 2, (0.6, 0., 0.8)),
                                                       ('tsc', 2, 'y')])
def test_correlation_global_los_matches_oracle(resampler, interlacing, los):
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh2CorrelationPoles, compute_mesh2_correlation
    say('corr global', resampler, interlacing, los)
    attrs = MeshAttrs(meshsize=(32, 32, 64), boxsize=(320., 320., 640.), boxcenter=(0., 50., 300.))
    pos, w = _mock(attrs, 20000, 3)
    p = ParticleField(to_dev(pos), to_dev(w), attrs=attrs)
    mesh = p.paint(resampler=resampler, interlacing=interlacing, compensate=True, out='complex').at[0, 0, 0].set(0.)
    ells = (0, 2, 4)
    bin = BinMesh2CorrelationPoles(attrs, edges={'step': 12.}, ells=ells)
    xi = compute_mesh2_correlation(mesh, bin=bin, los=los)
    oattrs = po.Attrs(attrs.meshsize, attrs.boxsize, attrs.boxcenter)
    omesh = po.paint(oattrs, pos, w.astype('f8'), resampler=resampler, interlacing=interlacing, compensate=True, out='complex')
    omesh[0, 0, 0] = 0.
    want = po.correlation_poles(oattrs, po.BinCorrelationPoles(oattrs, edges={'step': 12.}, ells=ells), omesh, los=los)
    got = np.array([pole.values('num_raw') for pole in xi])
    assert_xi_close(got, np.array(want['num_raw']), label=f'{resampler} {interlacing} {los}')
    assert np.array_equal(xi.get(0).values('nmodes'), want['nmodes'])
    assert np.allclose(xi.get(2).values('norm'), want['norm'])
    assert np.allclose(xi.get(0).coords('s')[want['nmodes'] > 0], want['s'][want['nmodes'] > 0], rtol=1e-12)


def test_correlation_fused_combine_matches_stored_mesh():
    """|A|^2 combined straight from the interlaced partial spectra equals |A|^2 of the stored
    combined mesh (the same float32 operations, up to contraction choices)."""
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh2CorrelationPoles, compute_mesh2_correlation, ComplexMeshField
    say('corr fused vs stored')
    attrs = MeshAttrs(meshsize=64, boxsize=640., boxcenter=0.)
    pos, w = _mock(attrs, 40000, 5)
    p = ParticleField(to_dev(pos), to_dev(w), attrs=attrs)
    bin = BinMesh2CorrelationPoles(attrs, edges={'step': 15.}, ells=(0, 2, 4))
    lazy = p.paint(resampler='tsc', interlacing=3, compensate=True, out='complex')
    stored = ComplexMeshField(lazy.value.clone(), attrs=attrs)
    a = compute_mesh2_correlation(lazy, bin=bin, los='z')
    b = compute_mesh2_correlation(stored, bin=bin, los='z')
    assert_xi_close([pa.values('num_raw') for pa in a], [pb.values('num_raw') for pb in b], rtol=1e-6, ftol=1e-7)


def test_correlation_cross_and_real_input():
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh2CorrelationPoles, compute_mesh2_correlation
    say('corr cross')
    attrs = MeshAttrs(meshsize=(32, 64, 32), boxsize=(320., 640., 320.), boxcenter=(10., 0., -5.))
    pos1, w1 = _mock(attrs, 15000, 7)
    pos2, w2 = _mock(attrs, 12000, 8)
    m1 = ParticleField(to_dev(pos1), to_dev(w1), attrs=attrs).paint(resampler='cic', interlacing=0, out='real')
    m2 = ParticleField(to_dev(pos2), to_dev(w2), attrs=attrs).paint(resampler='cic', interlacing=0, out='real')
    ells = (0, 1, 2)
    bin = BinMesh2CorrelationPoles(attrs, edges={'step': 10., 'max': 120.}, ells=ells)
    xi = compute_mesh2_correlation(m1, m2, bin=bin, los='z')
    oattrs = po.Attrs(attrs.meshsize, attrs.boxsize, attrs.boxcenter)
    o1 = po.paint(oattrs, pos1, w1.astype('f8'), resampler='cic')
    o2 = po.paint(oattrs, pos2, w2.astype('f8'), resampler='cic')
    want = po.correlation_poles(oattrs, po.BinCorrelationPoles(oattrs, edges={'step': 10., 'max': 120.}, ells=ells), o1, o2, los='z')
    assert_xi_close([pole.values('num_raw') for pole in xi], want['num_raw'], label='cross')


@pytest.mark.parametrize('los,boxcenter', [('firstpoint', (900., -300., 600.)), ('endpoint', (900., -300., 600.)), ('firstpoint', 0.)])
def test_correlation_local_los_matches_oracle(los, boxcenter):
    """Observer outside the box, or on a node (boxcenter 0: the reference's Y_40 constant term)."""
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh2CorrelationPoles, compute_mesh2_correlation
    say('corr local', los, boxcenter)
    attrs = MeshAttrs(meshsize=32, boxsize=500., boxcenter=boxcenter)
    pos, w = _mock(attrs, 20000, 11)
    mesh = ParticleField(to_dev(pos), to_dev(w), attrs=attrs).paint(resampler='tsc', interlacing=2, compensate=True, out='real')
    ells = (0, 2, 4)
    bin = BinMesh2CorrelationPoles(attrs, edges={'step': 25.}, ells=ells)
    xi = compute_mesh2_correlation(mesh, bin=bin, los=los)
    oattrs = po.Attrs(attrs.meshsize, attrs.boxsize, attrs.boxcenter)
    omesh = po.paint(oattrs, pos, w.astype('f8'), resampler='tsc', interlacing=2, compensate=True, out='real')
    want = po.correlation_poles(oattrs, po.BinCorrelationPoles(oattrs, edges={'step': 25.}, ells=ells), omesh, los=los)
    # the l > 0 terms sum 2l + 1 float32 FFT pairs
    assert_xi_close([pole.values('num_raw') for pole in xi], want['num_raw'], rtol=2e-5, ftol=5e-6, label=los)


def test_correlation_bin_call_plain_mean():
    from jaxpower_amd import MeshAttrs, BinMesh2CorrelationPoles
    say('corr bin call')
    attrs = MeshAttrs(meshsize=(16, 24, 32), boxsize=(160., 240., 320.), boxcenter=0.)
    rng = np.random.default_rng(2)
    value = rng.normal(size=tuple(attrs.meshsize)).astype('f4')
    bin = BinMesh2CorrelationPoles(attrs, edges={'step': 11.}, ells=(0,))
    obin = po.BinCorrelationPoles(po.Attrs(attrs.meshsize, attrs.boxsize, 0.), edges={'step': 11.}, ells=(0,))
    for rz in (False, True):
        got = bin(to_dev(value), remove_zero=rz)
        want = obin(value.astype('f8'), remove_zero=rz)
        assert np.allclose(got, want, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize('meshsize', [(32, 32, 64), (16, 8, 2), (8, 16, 1024)])
def test_fused_c2r_binning_matches_stored_real_mesh(meshsize):
    """The z c2r pass fused with the separation binning equals c2r to a stored real mesh
    followed by the binning kernel."""
    from jaxpower_amd import MeshAttrs, BinMesh2CorrelationPoles
    from jaxpower_amd.mesh import _c2r, _r2c
    say('corr fused c2r', meshsize)
    attrs = MeshAttrs(meshsize=meshsize, boxsize=np.array(meshsize) * 10., boxcenter=0.)
    rng = np.random.default_rng(13)
    x = rng.normal(size=meshsize).astype('f4')
    P = _r2c(to_dev(x), attrs.meshsize)
    ells, rz = (0, 2, 4, 1), (False, True, True, False)
    bin = BinMesh2CorrelationPoles(attrs, edges={'step': 13., 'max': 100.}, ells=ells)
    los = (0.3, -0.5, np.sqrt(1 - 0.34))
    fused = host(bin._spectrum_sums_dev(P, ells, rz, los=los))
    stored = host(bin._sums_dev(_c2r(P, attrs.meshsize), ells, rz, los=los))
    scale = np.abs(stored).max()
    assert np.abs(fused - stored).max() <= 1e-6 * scale
