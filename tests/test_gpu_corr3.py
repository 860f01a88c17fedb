"""GPU parity of the three-point correlation function in the Sugiyama basis
(BinMesh3CorrelationPoles, compute_mesh3_correlation; jaxpower/mesh3.py:316-409, 632-727)
against the float64 oracle, itself pinned in test_oracle.py to direct sums over pairs of
full-grid modes, Nyquist planes included.

Tolerances: edges, pairs and cell counts exact, mean |s| to 1e-12.  The multipoles come from
float32 meshes, float32 inverse FFTs and float32 products per cell (FP64 across cells): each pair
sum is compared as |d| <= 1e-5 |want| + ftol max|want| per multipole.
"""
import numpy as np
import pytest

from oracle import power_oracle as po

from test_gpu_bispec import _mock, assert_close
from test_gpu_parity import host, say, to_dev

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('meshsize,boxsize,edges,basis', [
    ((24, 24, 24), 240., {'step': 15.}, 'sugiyama'),
    ((16, 24, 20), (160., 240., 200.), {'step': 12., 'min': 4.}, 'sugiyama-diagonal'),
])
def test_corr3_bin_setup(meshsize, boxsize, edges, basis):
    from jaxpower_amd import MeshAttrs, BinMesh3CorrelationPoles
    say('corr3 bin', meshsize, basis)
    attrs = MeshAttrs(meshsize=meshsize, boxsize=boxsize, boxcenter=0.)
    bin = BinMesh3CorrelationPoles(attrs, edges=edges, ells=[(0, 0, 0), (2, 0, 2)], basis=basis)
    obin = po.BinCorrelation3Sugiyama(po.Attrs(meshsize, boxsize, 0.), edges=edges, ells=[(0, 0, 0), (2, 0, 2)],
                                      diagonal='diagonal' in basis)
    for a in range(2):
        assert np.array_equal(bin.nmodes1d[a], obin.nmodes1d[a])
        assert np.allclose(bin.xavg1d[a], obin.xavg1d[a], rtol=1e-12, atol=0., equal_nan=True)
    assert np.array_equal(bin._iedges, obin.iedges)
    assert np.array_equal(bin.edges, obin.edges)
    assert all(np.array_equal(nm, obin.nmodes) for nm in bin.nmodes)


@pytest.mark.parametrize('los,klimit', [('z', None), ('x', True), ((0.6, 0., 0.8), None), ('local', None)])
def test_corr3_matches_oracle(los, klimit):
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh3CorrelationPoles, compute_mesh3_correlation, compute_mesh3
    say('corr3', los, klimit)
    attrs = MeshAttrs(meshsize=(20, 20, 16), boxsize=(200., 200., 160.), boxcenter=(30., -20., 150.))
    pos, w = _mock(attrs, 20000, 7)
    mesh = ParticleField(to_dev(pos), to_dev(w), attrs=attrs).paint(resampler='tsc', interlacing=2, compensate=True, out='real')
    ells = [(0, 0, 0), (2, 0, 2), (1, 1, 0), (1, 1, 2)]
    edges = {'step': 12.}
    bin = BinMesh3CorrelationPoles(attrs, edges=edges, ells=ells, klimit=klimit)
    got = compute_mesh3_correlation(mesh, bin=bin, los=los)
    again = compute_mesh3(mesh, bin=bin, los=los)
    oattrs = po.Attrs(attrs.meshsize, attrs.boxsize, attrs.boxcenter)
    obin = po.BinCorrelation3Sugiyama(oattrs, edges=edges, ells=ells, klimit=klimit)
    want = po.correlation3_sugiyama(oattrs, obin, host(mesh.value).astype('f8'), los=los)
    for ill, ell in enumerate(ells):
        pole = got.get(ell)
        assert pole.ell == ell and pole.basis == 'sugiyama'
        assert np.array_equal(pole.values('nmodes'), want['nmodes'])
        assert_close(pole.values('num_raw'), want['num_raw'][ill], ftol=3e-6, label=f'{los} {ell}')
        assert np.array_equal(again.get(ell).values('num_raw'), pole.values('num_raw'))
    assert np.isclose(got.get((0, 0, 0)).values('norm')[0], want['norm'])


def test_corr3_bin_call_remove_zero(tmp_path):
    from jaxpower_amd import MeshAttrs, RealMeshField, BinMesh3CorrelationPoles, compute_mesh3_correlation, read_observable
    say('corr3 bin call')
    rng = np.random.default_rng(12)
    shape = (16, 16, 16)
    attrs = MeshAttrs(meshsize=shape, boxsize=160., boxcenter=0.)
    ms = [(rng.normal(size=shape) + 0.2).astype('f4') for _ in range(3)]
    bin = BinMesh3CorrelationPoles(attrs, edges={'step': 15.}, ells=[(0, 0, 0)], mask_edges='')
    obin = po.BinCorrelation3Sugiyama(po.Attrs(shape, 160., 0.), edges={'step': 15.}, ells=[(0, 0, 0)], mask_edges='')
    A = [RealMeshField(to_dev(m), attrs=attrs).r2c() for m in ms[:2]]
    for ell in ((0, 0), (2, 1)):
        for rz in (False, True):
            got = bin(A[0], A[1], RealMeshField(to_dev(ms[2]), attrs=attrs), ell=ell + (0,), remove_zero=rz)
            F = [np.fft.fftn(m.astype('f8')) for m in ms[:2]]
            third = ms[2].astype('f8')
            if rz:
                for f in F:
                    f[0, 0, 0] = 0.
                third = third - third.mean()
            want = obin.reduce(F[0], F[1], third, ell=ell)
            err = np.abs(got - want)
            assert np.all(err <= 1e-5 * np.abs(want) + 3e-6 * np.abs(want).max()), (ell, rz, err.max() / np.abs(want).max())
    xi = compute_mesh3_correlation(*[RealMeshField(to_dev(m), attrs=attrs) for m in ms], bin=bin, los='z')
    fn = str(tmp_path / 'corr3.npz')
    xi.write(fn)
    back = read_observable(fn)
    assert type(back).__name__ == 'Mesh3CorrelationPoles' and back.ells == ((0, 0, 0),)
    assert np.array_equal(back.get((0, 0, 0)).values('num_raw'), xi.get((0, 0, 0)).values('num_raw'))


@pytest.mark.parametrize('los,resampler,interlacing,compensate,ells', [
    ('z', 'cic', 2, True, [(0, 0, 0), (2, 0, 2), (0, 2, 2), (1, 1, 0)]),
    ('local', 'tsc', 0, True, [(0, 0, 0), (2, 0, 2), (1, 1, 2)]),
    ('x', 'ngp', 0, False, [(0, 0, 0), (0, 1, 1), (2, 2, 2)]),
])
def test_corr3_shotnoise_matches_oracle(los, resampler, interlacing, compensate, ells):
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh3CorrelationPoles, compute_fkp3_shotnoise
    say('corr3 shot noise', los, resampler, interlacing, compensate)
    attrs = MeshAttrs(meshsize=(16, 16, 12), boxsize=(160., 160., 120.), boxcenter=(20., -10., 90.))
    pos, w = _mock(attrs, 6000, 23)
    p = ParticleField(to_dev(pos), to_dev(w), attrs=attrs)
    edges = {'step': 10.}
    bin = BinMesh3CorrelationPoles(attrs, edges=edges, ells=ells)
    got = compute_fkp3_shotnoise(p, bin=bin, los=los, resampler=resampler, interlacing=interlacing, compensate=compensate)
    oattrs = po.Attrs(attrs.meshsize, attrs.boxsize, attrs.boxcenter)
    obin = po.BinCorrelation3Sugiyama(oattrs, edges=edges, ells=ells)
    want = po.sugiyama_shotnoise(oattrs, obin, pos, w.astype('f8'), los=los, resampler=resampler, interlacing=interlacing,
                                 compensate=compensate)
    scale = np.abs(want[0]).max()
    for ill, ell in enumerate(ells):
        g, wv = np.asarray(got[ill]), np.asarray(want[ill])
        assert np.all(np.abs(g - wv) <= 1e-5 * np.abs(wv) + 1e-5 * scale), (ell, np.abs(g - wv).max() / scale)
