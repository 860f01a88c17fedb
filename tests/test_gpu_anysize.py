"""GPU parity for meshes whose sizes are not all powers of two (the any-length FFT passes):
sizes from next_fft_size(primes=(2, 3, 5, 7)) (jaxpower/mesh.py:2226-2262) and the
reference's own (128, 124, 122) geometry (tests/test_mesh2.py:36), whose factors include the
primes 31 and 61, plus odd lengths along the hermitian axis.

Tolerances: transforms in norm relative to the largest value (float32); spectra
|dP_l| <= 1e-5 max(|P_l|, |P_0|) per bin; mode counts exact.
"""
import numpy as np
import pytest
import torch

from oracle import power_oracle as po

from test_gpu_parity import assert_poles_close, host, say, to_dev

pytestmark = pytest.mark.gpu

SHAPES = [(128, 124, 122), (30, 42, 50), (12, 10, 18), (7, 9, 15), (16, 20, 61), (250, 20, 36), (2, 3, 5), (1000, 4, 6),
          (6, 4, 1000)]


@pytest.mark.parametrize('shape', SHAPES)
def test_r2c_c2r_any_length(shape):
    from jaxpower_amd.mesh import _r2c, _c2r
    say('r2c/c2r any', shape)
    rng = np.random.default_rng(sum(shape))
    x = rng.normal(size=shape).astype('f4')
    got = host(_r2c(to_dev(x), shape))
    want = np.fft.rfftn(x.astype('f8'))
    err = np.abs(got - want).max() / np.abs(want).max()
    print('r2c rel err', err)
    assert err < 3e-6, err
    h = (shape[0], shape[1], shape[2] // 2 + 1)
    c = rng.normal(size=h) + 1j * rng.normal(size=h)
    back = host(_c2r(to_dev(c.astype('c8')), shape))
    want = np.fft.irfftn(c, s=shape, axes=(0, 1, 2))
    err = np.abs(back - want).max() / np.abs(want).max()
    print('c2r rel err', err)
    assert err < 3e-6, err


@pytest.mark.parametrize('shape', [(128, 124, 122), (30, 42, 50), (7, 9, 15), (36, 40, 45)])
@pytest.mark.parametrize('resampler,interlacing', [('tsc', 3), ('cic', 2), ('pcs', 0)])
def test_paint_spectrum_any_length(shape, resampler, interlacing):
    """paint(out='complex') (real paints, any-length z/y passes, the combine with its any-length
    x pass) and compute_mesh2_spectrum against the oracle; the combined mesh itself too."""
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh2SpectrumPoles, compute_mesh2_spectrum
    say('paint spectrum any', shape, resampler, interlacing)
    rng = np.random.default_rng(len(resampler) + interlacing + sum(shape))
    attrs = MeshAttrs(meshsize=shape, boxsize=np.array(shape) * 8., boxcenter=(10., -20., 30.))
    lo = attrs.boxcenter - attrs.boxsize / 2.
    pos = rng.uniform(0., 1., (20000, 3)) * attrs.boxsize + lo
    pos[:5000] = lo + np.mod(pos[:5000] * 0.3, attrs.boxsize)  # some structure
    w = rng.uniform(0.5, 1.5, len(pos)).astype('f4')
    p = ParticleField(to_dev(pos), to_dev(w), attrs=attrs)
    ells = (0, 2, 4)
    step = np.min(attrs.kfun) * np.min(attrs.meshsize) / 2. / 8.
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': step}, ells=ells)
    mesh = p.paint(resampler=resampler, interlacing=interlacing, compensate=True, out='complex').at[0, 0, 0].set(0.)
    spec = compute_mesh2_spectrum(mesh, bin=bin, los='z')
    value = host(mesh.value)
    oattrs = po.Attrs(shape, attrs.boxsize, attrs.boxcenter)
    omesh = po.paint(oattrs, pos, w.astype('f8'), resampler=resampler, interlacing=interlacing, compensate=True, out='complex')
    omesh[0, 0, 0] = 0.
    err = np.abs(value - omesh).max() / np.abs(omesh).max()
    print('combined mesh rel err', err)
    # PCS compensation reaches (pi/2)^12 at the corner mode, where float32 paint rounding shows
    assert err < (2e-5 if resampler == 'pcs' else 1e-5)
    want = po.spectrum_poles(oattrs, po.BinPoles(oattrs, edges={'step': step}, ells=ells), omesh, los='z')
    assert np.array_equal(spec.get(0).values('nmodes'), want['nmodes'])
    assert_poles_close([spec.get(ell).values('num_raw') for ell in ells], want['num_raw'], label=f'any {shape}')
    # the real-space mesh (c2r of the combined spectrum)
    real = host(p.paint(resampler=resampler, interlacing=interlacing, compensate=True, out='real').value)
    oreal = po.paint(oattrs, pos, w.astype('f8'), resampler=resampler, interlacing=interlacing, compensate=True, out='real')
    assert np.abs(real - oreal).max() < (2e-5 if resampler == 'pcs' else 1e-5) * np.abs(oreal).max()


@pytest.mark.parametrize('los', ['x', 'endpoint'])
def test_reference_closed_form_128_124_122(los):
    """The reference's closed-form test (tests/test_mesh2.py:23-61) at its non-power-of-two
    geometry: a unitary-amplitude Gaussian mesh with P(k) = 1e4 (k/0.03)^3 exp(-k/0.03),
    meshsize (128, 124, 122), boxsize 1000, ells (0, 2, 4), step 0.01: the monopole equals P(k)
    at the bin centres to 1e-2 from the third bin on; and the spectrum equals the oracle's on the
    same mesh to 1e-5."""
    from jaxpower_amd import MeshAttrs, BinMesh2SpectrumPoles, compute_mesh2_spectrum
    from jaxpower_amd.mock import generate_gaussian_mesh
    say('closed form (128, 124, 122)', los)

    def pkvec(kvec):
        k = torch.sqrt(sum(kk.to(torch.float64)**2 for kk in kvec))
        return (1e4 * (k / 0.03)**3 * torch.exp(-k / 0.03)).to(torch.float32)

    attrs = MeshAttrs(meshsize=(128, 124, 122), boxsize=1000.)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.01}, ells=(0, 2, 4))
    mesh = generate_gaussian_mesh(attrs, pkvec, seed=43, unitary_amplitude=True, device='cuda')
    spec = compute_mesh2_spectrum(mesh, los=los, bin=bin)
    k = spec.get(0).coords('k')
    pk = 1e4 * (k / 0.03)**3 * np.exp(-k / 0.03)
    assert tuple(spec.ells) == (0, 2, 4)
    assert np.allclose(spec.get(0).value()[2:], pk[2:], rtol=1e-2)
    oattrs = po.Attrs((128, 124, 122), 1000.)
    want = po.spectrum_poles(oattrs, po.BinPoles(oattrs, edges={'step': 0.01}, ells=(0, 2, 4)), host(mesh.value).astype('f8'), los=los)
    assert np.array_equal(spec.get(0).values('nmodes'), want['nmodes'])
    assert_poles_close([spec.get(ell).values('num_raw') for ell in (0, 2, 4)], want['num_raw'], label=f'closed form {los}')


def test_cellsize_geometry_end_to_end():
    """A cellsize-driven geometry (get_mesh_attrs with primes (2, 3, 5, 7)) from particles,
    FKP-like weights with both signs, TSC with interlacing 2, local line of sight: against the
    oracle."""
    from jaxpower_amd import ParticleField, BinMesh2SpectrumPoles, compute_mesh2_spectrum, get_mesh_attrs
    say('cellsize geometry')
    rng = np.random.default_rng(7)
    pos = rng.uniform(-300., 400., (30000, 3)) + np.array([1000., 500., 800.])
    w = np.where(rng.uniform(size=len(pos)) < 0.7, 1., -0.4).astype('f4')
    attrs = get_mesh_attrs(pos, cellsize=13., boxpad=1.3, primes=(2, 3, 5, 7))
    assert not all(int(n) & (int(n) - 1) == 0 for n in attrs.meshsize)  # not all powers of two
    oattrs = po.get_mesh_attrs(pos, cellsize=13., boxpad=1.3, primes=(2, 3, 5, 7))
    assert np.array_equal(attrs.meshsize, oattrs.meshsize) and np.allclose(attrs.boxsize, oattrs.boxsize)
    p = ParticleField(to_dev(pos), to_dev(w), attrs=attrs)
    mesh = p.paint(resampler='tsc', interlacing=2, compensate=True, out='real')
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.01}, ells=(0, 2, 4))
    spec = compute_mesh2_spectrum(mesh, bin=bin, los='firstpoint')
    omesh = po.paint(oattrs, pos, w.astype('f8'), resampler='tsc', interlacing=2, compensate=True, out='real')
    want = po.spectrum_poles(oattrs, po.BinPoles(oattrs, edges={'step': 0.01}, ells=(0, 2, 4)), omesh, los='firstpoint')
    assert np.array_equal(spec.get(0).values('nmodes'), want['nmodes'])
    assert_poles_close([spec.get(ell).values('num_raw') for ell in (0, 2, 4)], want['num_raw'], label='cellsize firstpoint')
