"""GPU parity at the transform lengths of the full-size workloads: axes of 512 and 1024
points, which select their own kernel configurations (x-pass tiles of 4096/L columns, the
column-pass tiles, the 1024-point local line-of-sight x pass), and one case with the full
512^3 geometry of the headline box (boxsize 2000, TSC, interlacing 3, k step 0.005).

Tolerances: FFTs and meshes in norm relative to the largest value (float32 transforms);
spectra |dP_l| <= 1e-5 max(|P_l|, |P_0|) per bin; mode counts exact.  Each test prints the
measured errors per multipole.
"""
import numpy as np
import pytest
import torch

from oracle import power_oracle as po

from test_gpu_parity import DEV, assert_poles_close, host, say, to_dev

pytestmark = pytest.mark.gpu

LONG = [(512, 16, 16), (16, 512, 16), (16, 16, 512), (1024, 16, 16), (16, 1024, 16), (16, 16, 1024)]


def _max_rel(got, want):
    """max over bins of |dP_l| / max(|P_l|, |P_0|), per multipole."""
    got, want = np.asarray(got), np.asarray(want)
    p0 = np.abs(want[0])
    out = []
    for g, w in zip(got, want):
        ok = ~np.isnan(w)
        out.append(float(np.max(np.abs(g - w)[ok] / np.maximum(np.maximum(np.abs(w), p0)[ok], 1e-300))))
    return out


def _particles(attrs, size, seed):
    rng = np.random.default_rng(seed)
    lo = attrs.boxcenter - attrs.boxsize / 2.
    pos = rng.uniform(0., 1., (size, 3)) * attrs.boxsize + lo
    n = size // 3  # clumps, so that the ell > 0 multipoles carry signal
    centers = rng.uniform(0., 1., (8, 3)) * attrs.boxsize + lo
    pos[:n] = centers[rng.integers(0, 8, n)] + rng.normal(0., 0.03, (n, 3)) * attrs.boxsize
    pos[:n] = lo + np.mod(pos[:n] - lo, attrs.boxsize)
    w = rng.uniform(0.5, 1.5, size).astype('f4')
    return pos, w


@pytest.mark.parametrize('shape', LONG)
def test_r2c_c2r_long_axes(shape):
    from jaxpower_amd.mesh import _r2c, _c2r
    say('r2c/c2r long', shape)
    rng = np.random.default_rng(sum(shape))
    x = rng.normal(size=shape).astype('f4')
    got = host(_r2c(to_dev(x), shape))
    want = np.fft.rfftn(x.astype('f8'))
    err = np.abs(got - want).max() / np.abs(want).max()
    print('r2c rel err', err)
    assert err < 2e-6, err
    h = (shape[0], shape[1], shape[2] // 2 + 1)
    c = rng.normal(size=h) + 1j * rng.normal(size=h)
    back = host(_c2r(to_dev(c.astype('c8')), shape))
    want = np.fft.irfftn(c, s=shape, axes=(0, 1, 2))
    err = np.abs(back - want).max() / np.abs(want).max()
    print('c2r rel err', err)
    assert err < 2e-6, err


@pytest.mark.parametrize('shape', LONG)
@pytest.mark.parametrize('resampler,interlacing', [('tsc', 3), ('cic', 2)])
def test_fused_bin_long_axes(shape, resampler, interlacing):
    """paint(out='complex') -> the fused x pass + FP64 binning, and the same mesh combined and
    written (the combine x pass), against the oracle."""
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh2SpectrumPoles, compute_mesh2_spectrum
    say('fused bin long', shape, resampler, interlacing)
    boxsize = np.array(shape) * 4.
    attrs = MeshAttrs(meshsize=shape, boxsize=boxsize, boxcenter=0.)
    pos, w = _particles(attrs, 30000, sum(shape) + interlacing)
    p = ParticleField(to_dev(pos), to_dev(w), attrs=attrs)
    ells = (0, 2, 4)
    step = np.min(attrs.kfun) * np.min(attrs.meshsize) / 2. / 8.
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': step}, ells=ells)
    lazy = p.paint(resampler=resampler, interlacing=interlacing, compensate=True, out='complex').at[0, 0, 0].set(0.)
    fused = compute_mesh2_spectrum(lazy, bin=bin, los='z')
    value = host(lazy.value)  # the combine x pass, written
    oattrs = po.Attrs(shape, attrs.boxsize, attrs.boxcenter)
    omesh = po.paint(oattrs, pos, w.astype('f8'), resampler=resampler, interlacing=interlacing, compensate=True, out='complex')
    omesh[0, 0, 0] = 0.
    err = np.abs(value - omesh).max() / np.abs(omesh).max()
    print('combined mesh rel err', err)
    assert err < 1e-5, err
    want = po.spectrum_poles(oattrs, po.BinPoles(oattrs, edges={'step': step}, ells=ells), omesh, los='z')
    assert np.array_equal(fused.get(0).values('nmodes'), want['nmodes'])
    got = np.array([fused.get(ell).values('num_raw') for ell in ells])
    print('per-ell max rel err', _max_rel(got, want['num_raw']))
    assert_poles_close(got, want['num_raw'], label=f'fused {shape}')


@pytest.mark.parametrize('shape', [(1024, 16, 16), (16, 1024, 16), (512, 16, 32)])
def test_local_fused_long_axes(shape):
    """The fused local line-of-sight binning (mp_ylm_bin: Ylm z pass, y passes, Ylm(khat) x pass)
    at 512 and 1024 points, against the oracle."""
    from jaxpower_amd import MeshAttrs, RealMeshField, BinMesh2SpectrumPoles, compute_mesh2_spectrum
    say('local fused long', shape)
    rng = np.random.default_rng(sum(shape))
    boxsize = np.array(shape) * 5.
    attrs = MeshAttrs(meshsize=shape, boxsize=boxsize, boxcenter=(1500., -700., 2600.))
    m1 = rng.normal(size=shape).astype('f4')
    ells = (0, 2, 4)
    step = np.min(attrs.kfun) * np.min(attrs.meshsize) / 2. / 8.
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': step}, ells=ells)
    got = compute_mesh2_spectrum(RealMeshField(to_dev(m1), attrs=attrs), bin=bin, los='firstpoint')
    oattrs = po.Attrs(shape, attrs.boxsize, attrs.boxcenter)
    want = po.spectrum_poles(oattrs, po.BinPoles(oattrs, edges={'step': step}, ells=ells), m1.astype('f8'), los='firstpoint')
    got = np.array([got.get(ell).values('num_raw') for ell in ells])
    print('per-ell max rel err', _max_rel(got, want['num_raw']))
    assert_poles_close(got, want['num_raw'], rtol=1e-5, label=f'local {shape}')


def test_box_512_headline_geometry():
    """The headline box geometry at full mesh size: 512^3, boxsize 2000, boxcenter 1000, TSC,
    interlacing 3, compensated, DC removed, ells (0, 2, 4) along z, k step 0.005 (160 bins);
    5e5 particles keep the float64 oracle at ~20 s.  nmodes bit-exact (this geometry has
    shells within 1e-7 of an edge), every bin within 1e-5."""
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh2SpectrumPoles, compute_mesh2_spectrum
    say('box 512 headline geometry')
    attrs = MeshAttrs(meshsize=512, boxsize=2000., boxcenter=1000.)
    pos, w = _particles(attrs, 500000, 2024)
    pos = pos.astype('f4')
    p = ParticleField(to_dev(pos), attrs=attrs)
    ells = (0, 2, 4)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.005}, ells=ells)
    assert bin.nbins == 160
    mesh = p.paint(resampler='tsc', interlacing=3, compensate=True, out='complex').at[0, 0, 0].set(0.)
    spec = compute_mesh2_spectrum(mesh, bin=bin, los='z')
    got = np.array([spec.get(ell).values('num_raw') for ell in ells])
    nmodes = spec.get(0).values('nmodes')
    del mesh, spec
    torch.cuda.empty_cache()
    oattrs = po.Attrs(512, 2000., 1000.)
    obin = po.BinPoles(oattrs, edges={'step': 0.005}, ells=ells)
    assert np.array_equal(nmodes, obin.nmodes)
    omesh = po.paint(oattrs, pos.astype('f8'), None, resampler='tsc', interlacing=3, compensate=True, out='complex')
    omesh[0, 0, 0] = 0.
    want = po.spectrum_poles(oattrs, obin, omesh, los='z')
    print('per-ell max rel err', _max_rel(got, want['num_raw']))
    assert_poles_close(got, want['num_raw'], label='512^3 headline geometry')
