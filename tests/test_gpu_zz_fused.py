"""GPU parity of the fused paint + z-FFT path (paint_zfft.hip) against the oracle and
against the separate paint + z-pass path."""
import numpy as np
import pytest
import torch

from oracle import power_oracle as po

pytestmark = pytest.mark.gpu


def say(*args):
    print('[case]', *args, flush=True)


@pytest.fixture
def fused_path():
    from jaxpower_amd import _lib
    old = _lib.load().mp_set_paint_path(2)
    yield
    _lib.load().mp_set_paint_path(old)


@pytest.mark.parametrize('meshsize', [(32, 32, 32), (64, 32, 64), (64, 64, 16), (32, 32, 16)])
@pytest.mark.parametrize('resampler', ['ngp', 'cic', 'tsc', 'pcs'])
@pytest.mark.parametrize('interlacing', [0, 2, 3])
def test_fused_paint_spectrum_matches_oracle(fused_path, meshsize, resampler, interlacing):
    from jaxpower_amd import MeshAttrs, ParticleField
    say('fused paint spectrum', meshsize, resampler, interlacing)
    rng = np.random.default_rng(17)
    attrs = MeshAttrs(meshsize=meshsize, boxsize=np.array(meshsize) * 7., boxcenter=(5., -5., 20.))
    lo = attrs.boxcenter - attrs.boxsize / 2.
    pos = rng.uniform(-0.05, 1.05, (6000, 3)) * attrs.boxsize + lo
    w = rng.uniform(0.5, 1.5, 6000).astype('f4')
    p = ParticleField(torch.from_numpy(pos).cuda(), torch.from_numpy(w).cuda(), attrs=attrs)
    got = p.paint(resampler=resampler, interlacing=interlacing, compensate=True, out='complex').value.cpu().numpy()
    oattrs = po.Attrs(attrs.meshsize, attrs.boxsize, attrs.boxcenter)
    want = po.paint(oattrs, pos, w.astype('f8'), resampler=resampler, interlacing=interlacing, compensate=True, out='complex')
    err = np.abs(got - want).max() / np.abs(want).max()
    assert err < 3e-6, err


def test_fused_matches_separate_path_f32_positions():
    from jaxpower_amd import MeshAttrs, ParticleField, _lib, generate_lognormal_particles
    say('fused vs separate 128^3 lognormal')
    attrs = MeshAttrs(meshsize=128, boxsize=500., boxcenter=250.)
    p = generate_lognormal_particles(attrs, size=300000, seed=5)
    lib = _lib.load()
    old = lib.mp_set_paint_path(1)
    try:
        a = p.paint(resampler='tsc', interlacing=3, compensate=True, out='complex').value.cpu().numpy()
        lib.mp_set_paint_path(2)
        b = p.paint(resampler='tsc', interlacing=3, compensate=True, out='complex').value.cpu().numpy()
    finally:
        lib.mp_set_paint_path(old)
    assert np.abs(a - b).max() / np.abs(a).max() < 1e-5


def test_fused_rejects_unsupported_shape(fused_path):
    from jaxpower_amd import MeshAttrs, ParticleField
    say('fused unsupported shape')
    attrs = MeshAttrs(meshsize=(32, 4, 64), boxsize=100.)  # no row-block shape divides 4 rows
    p = ParticleField(torch.rand((100, 3), device='cuda') * 100., attrs=attrs)
    with pytest.raises(NotImplementedError):
        p.paint(resampler='tsc', interlacing=3, compensate=True, out='complex')


@pytest.mark.parametrize('resampler', ['cic', 'tsc', 'pcs'])
def test_fused_dense_clumps(fused_path, resampler):
    """Many particles per cell (clumps a few cells wide, as in a strongly clustered mock):
    exercises the per-cell accumulation and the z-sorted rows."""
    from jaxpower_amd import MeshAttrs, ParticleField
    say('fused dense clumps', resampler)
    rng = np.random.default_rng(23)
    attrs = MeshAttrs(meshsize=64, boxsize=640., boxcenter=320.)
    centers = rng.uniform(0., 640., (6, 3))
    pos = centers[rng.integers(0, 6, 40000)] + rng.normal(0., 8., (40000, 3))
    pos = np.concatenate([pos, rng.uniform(0., 640., (5000, 3))])
    w = rng.uniform(0.5, 1.5, len(pos)).astype('f4')
    p = ParticleField(torch.from_numpy(pos).cuda(), torch.from_numpy(w).cuda(), attrs=attrs)
    got = p.paint(resampler=resampler, interlacing=3, compensate=True, out='complex').value.cpu().numpy()
    oattrs = po.Attrs(attrs.meshsize, attrs.boxsize, attrs.boxcenter)
    want = po.paint(oattrs, pos, w.astype('f8'), resampler=resampler, interlacing=3, compensate=True, out='complex')
    err = np.abs(got - want).max() / np.abs(want).max()
    assert err < 1e-5, err  # many float32 weights summed per cell


@pytest.mark.parametrize('meshsize', [(64, 64, 512), (32, 32, 256), (64, 32, 128), (32, 32, 64), (32, 16, 32)])
@pytest.mark.parametrize('resampler', ['cic', 'tsc', 'pcs'])
def test_fused_every_row_block_shape(fused_path, meshsize, resampler):
    """Every row-block shape the autotuner may pick (forced one at a time) paints the oracle's mesh."""
    from jaxpower_amd import MeshAttrs, ParticleField, _lib
    say('fused shapes', meshsize, resampler)
    rng = np.random.default_rng(31)
    attrs = MeshAttrs(meshsize=meshsize, boxsize=np.array(meshsize) * 5., boxcenter=0.)
    lo = attrs.boxcenter - attrs.boxsize / 2.
    pos = rng.uniform(0., 1., (4000, 3)) * attrs.boxsize + lo
    w = rng.uniform(0.5, 1.5, 4000).astype('f4')
    p = ParticleField(torch.from_numpy(pos).cuda(), torch.from_numpy(w).cuda(), attrs=attrs)
    want = po.paint(po.Attrs(attrs.meshsize, attrs.boxsize, attrs.boxcenter), pos, w.astype('f8'), resampler=resampler,
                    interlacing=3, compensate=True, out='complex')
    lib = _lib.load()
    old = lib.mp_set_paint_tuning(-1)
    try:
        for q in range(3):
            lib.mp_set_paint_tuning(100 + q)
            got = p.paint(resampler=resampler, interlacing=3, compensate=True, out='complex').value.cpu().numpy()
            err = np.abs(got - want).max() / np.abs(want).max()
            assert err < 1e-5, (q, err)
    finally:
        lib.mp_set_paint_tuning(old)
