"""Mode oversampling on the GPU (BinMesh2SpectrumPoles / BinMesh2CorrelationPoles with
mode_oversampling = m; jaxpower/mesh2.py:48-61, 94, 153-156, and _bincount over the per-shift
bin indices, jaxpower/mesh.py:2874-2910) against the float64 oracle.

The boxes are small (wavenumbers of order 1 h/Mpc, separations of order 1 Mpc/h), so that the
integer shifts the reference adds to the coordinates move modes between bins.

Tolerances: nmodes equal (integer counts divided by the number of shifts, formed the same way),
xavg to 1e-12 relative; multipoles as in the other spectrum and correlation parity tests
(float32 meshes and FFTs).
"""
import numpy as np
import pytest

from oracle import power_oracle as po

from test_gpu_corr import assert_xi_close
from test_gpu_parity import assert_poles_close, host, say, to_dev

pytestmark = pytest.mark.gpu

KEDGES = np.arange(0., 3.01, 0.1)


@pytest.mark.parametrize('m', [1, 2])
@pytest.mark.parametrize('meshsize,boxsize', [((16, 16, 16), 40.), ((12, 16, 10), (30., 40., 25.))])
def test_oversampled_modes(m, meshsize, boxsize):
    from jaxpower_amd import MeshAttrs, BinMesh2SpectrumPoles
    say('oversampled modes', m, meshsize, boxsize)
    attrs = MeshAttrs(meshsize=meshsize, boxsize=boxsize, boxcenter=0.)
    bin = BinMesh2SpectrumPoles(attrs, edges=KEDGES, ells=(0, 2), mode_oversampling=m)
    obin = po.BinPoles(po.Attrs(meshsize, boxsize, 0.), edges=KEDGES, ells=(0, 2), mode_oversampling=m)
    assert bin.nmodes.dtype == np.float64
    assert np.array_equal(bin.nmodes, obin.nmodes)
    ok = obin.nmodes > 0
    assert np.allclose(bin.xavg[ok], obin.xavg[ok], rtol=1e-12, atol=0)
    assert np.array_equal(np.isnan(bin.xavg), np.isnan(obin.xavg))


def test_oversampled_bin_call():
    from jaxpower_amd import MeshAttrs, BinMesh2SpectrumPoles
    say('oversampled bin call')
    attrs = MeshAttrs(meshsize=16, boxsize=40.)
    bin = BinMesh2SpectrumPoles(attrs, edges=KEDGES, mode_oversampling=1)
    obin = po.BinPoles(po.Attrs(16, 40.), edges=KEDGES, mode_oversampling=1)
    rng = np.random.default_rng(4)
    v = rng.normal(size=(16, 16, 9)) + 1j * rng.normal(size=(16, 16, 9))
    for rz in (False, True):
        got = bin(to_dev(v.astype('c8')), remove_zero=rz)
        want = obin(v, remove_zero=rz)
        assert np.allclose(got, want, rtol=1e-5, atol=1e-6, equal_nan=True)


@pytest.mark.parametrize('los', ['z', (0.3, -0.5, np.sqrt(1 - 0.34)), 'firstpoint'])
@pytest.mark.parametrize('cross', [False, True])
def test_oversampled_spectrum(los, cross):
    """compute_mesh2_spectrum with an oversampled bin: global line of sight (mu of the unshifted
    k) and local line of sight (the stored C_ell route)."""
    from jaxpower_amd import MeshAttrs, RealMeshField, BinMesh2SpectrumPoles, compute_mesh2_spectrum
    say('oversampled spectrum', los, cross)
    rng = np.random.default_rng(31)
    shape = (16, 24, 16)
    attrs = MeshAttrs(meshsize=shape, boxsize=(40., 60., 40.), boxcenter=(5., -3., 70.))
    m1 = rng.normal(size=shape).astype('f4')
    m2 = (m1 + 0.5 * rng.normal(size=shape)).astype('f4')
    ells = (0, 2, 4)
    bin = BinMesh2SpectrumPoles(attrs, edges=KEDGES, ells=ells, mode_oversampling=1)
    meshes = [RealMeshField(to_dev(m1), attrs=attrs)] + ([RealMeshField(to_dev(m2), attrs=attrs)] if cross else [])
    got = compute_mesh2_spectrum(*meshes, bin=bin, los=los)
    oattrs = po.Attrs(shape, attrs.boxsize, attrs.boxcenter)
    obin = po.BinPoles(oattrs, edges=KEDGES, ells=ells, mode_oversampling=1)
    want = po.spectrum_poles(oattrs, obin, m1.astype('f8'), m2.astype('f8') if cross else None, los=los)
    assert_poles_close([got.get(ell).values('num_raw') for ell in ells], want['num_raw'], label=f'oversampled {los} cross={cross}')
    assert np.array_equal(got.get(0).values('nmodes'), want['nmodes'])


def test_oversampled_spectrum_of_painted_mesh():
    """An interlaced, compensated paint whose combine is otherwise fused with the binning: the
    oversampled bin takes the stored combined mesh instead."""
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh2SpectrumPoles, compute_mesh2_spectrum
    say('oversampled painted spectrum')
    rng = np.random.default_rng(8)
    attrs = MeshAttrs(meshsize=32, boxsize=80., boxcenter=40.)
    pos = rng.uniform(0., 80., (30000, 3))
    mesh = ParticleField(to_dev(pos), attrs=attrs).paint(resampler='tsc', interlacing=2, compensate=True, out='complex')
    ells = (0, 2, 4)
    bin = BinMesh2SpectrumPoles(attrs, edges=KEDGES, ells=ells, mode_oversampling=1)
    got = compute_mesh2_spectrum(mesh, bin=bin, los='x')
    oattrs = po.Attrs(32, 80., 40.)
    omesh = po.paint(oattrs, pos, None, resampler='tsc', interlacing=2, compensate=True, out='complex')
    want = po.spectrum_poles(oattrs, po.BinPoles(oattrs, edges=KEDGES, ells=ells, mode_oversampling=1), omesh, los='x')
    assert_poles_close([got.get(ell).values('num_raw') for ell in ells], want['num_raw'], label='oversampled painted')


@pytest.mark.parametrize('m', [1])
def test_oversampled_correlation_modes_and_call(m):
    from jaxpower_amd import MeshAttrs, BinMesh2CorrelationPoles
    say('oversampled corr modes', m)
    meshsize, boxsize = (16, 12, 20), (16., 12., 20.)
    sedges = np.arange(0., 12.01, 0.5)
    attrs = MeshAttrs(meshsize=meshsize, boxsize=boxsize, boxcenter=0.)
    bin = BinMesh2CorrelationPoles(attrs, edges=sedges, ells=(0,), mode_oversampling=m)
    obin = po.BinCorrelationPoles(po.Attrs(meshsize, boxsize, 0.), edges=sedges, ells=(0,), mode_oversampling=m)
    assert np.array_equal(bin.nmodes, obin.nmodes)
    ok = obin.nmodes > 0
    assert np.allclose(bin.xavg[ok], obin.xavg[ok], rtol=1e-12, atol=0)
    rng = np.random.default_rng(2)
    value = rng.normal(size=meshsize).astype('f4')
    for rz in (False, True):
        got = bin(to_dev(value), remove_zero=rz)
        want = obin(value.astype('f8'), remove_zero=rz)
        assert np.allclose(got[ok], want[ok], rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize('los', ['z', 'firstpoint'])
def test_oversampled_correlation(los):
    from jaxpower_amd import MeshAttrs, RealMeshField, BinMesh2CorrelationPoles, compute_mesh2_correlation
    say('oversampled corr', los)
    rng = np.random.default_rng(12)
    shape = (16, 16, 24)
    attrs = MeshAttrs(meshsize=shape, boxsize=(16., 16., 24.), boxcenter=(2., 1., 40.))
    m1 = rng.normal(size=shape).astype('f4')
    sedges = np.arange(0., 10.01, 0.5)
    ells = (0, 2)
    bin = BinMesh2CorrelationPoles(attrs, edges=sedges, ells=ells, mode_oversampling=1)
    xi = compute_mesh2_correlation(RealMeshField(to_dev(m1), attrs=attrs), bin=bin, los=los)
    oattrs = po.Attrs(shape, attrs.boxsize, attrs.boxcenter)
    want = po.correlation_poles(oattrs, po.BinCorrelationPoles(oattrs, edges=sedges, ells=ells, mode_oversampling=1), m1.astype('f8'),
                                los=los)
    assert_xi_close([pole.values('num_raw') for pole in xi], want['num_raw'], rtol=2e-5, ftol=5e-6, label=f'oversampled {los}')
    assert np.array_equal(xi.get(0).values('nmodes'), want['nmodes'])


def test_oversampled_bin_refused_by_fused_paths():
    """The window and mean paths bin inside their own fused kernels, one shift only."""
    from jaxpower_amd import MeshAttrs, BinMesh2SpectrumPoles, compute_mesh2_spectrum_window
    say('oversampled window refused')
    attrs = MeshAttrs(meshsize=16, boxsize=40.)
    bin = BinMesh2SpectrumPoles(attrs, edges=KEDGES, ells=(0,), mode_oversampling=1)
    with pytest.raises(NotImplementedError, match='mode_oversampling'):
        compute_mesh2_spectrum_window(attrs, edgesin=np.linspace(0., 1., 5), ellsin=(0,), bin=bin, los='z')
    assert host(bin._nmodes_dev).dtype == np.float64
