"""GPU tests of Mesh2SpectrumBatch: the batch replayed as one captured HIP graph gives the same
spectra, bit for bit, as the per-mock eager calls (the kernels are the same and the paint is
order independent), replay after replay and after the inputs are refilled in place; the
eager batch agrees with the float64 oracle to |dP| <= 1e-5 max(|P_l|, |P_0|)."""
import numpy as np
import pytest
import torch

from oracle import power_oracle as po

from test_gpu_parity import assert_poles_close, host, say, to_dev

pytestmark = pytest.mark.gpu


def _fields(attrs, counts, seed):
    from jaxpower_amd import ParticleField
    rng = np.random.default_rng(seed)
    out = []
    for c in counts:
        pos = rng.uniform(0., 1., (c, 3)) * attrs.boxsize + attrs.boxcenter - attrs.boxsize / 2.
        out.append(ParticleField(to_dev(pos.astype('f4')), to_dev(rng.uniform(0.5, 1.5, c).astype('f4')), attrs=attrs))
    return out


def _values(spec, ells):
    return np.array([spec.get(ell).value() for ell in ells])


@pytest.mark.parametrize('resampler,interlacing', [('cic', 2), ('tsc', 3)])
def test_batch_graph_replays_match_eager(resampler, interlacing):
    from jaxpower_amd import MeshAttrs, BinMesh2SpectrumPoles, Mesh2SpectrumBatch
    say('batch graph', resampler, interlacing)
    attrs = MeshAttrs(meshsize=(64, 32, 32), boxsize=(640., 320., 320.), boxcenter=0.)
    ells = (0, 2, 4)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.02}, ells=ells)
    fields = _fields(attrs, [3000, 1200, 5000], 3)
    eager = Mesh2SpectrumBatch(fields, bin, resampler=resampler, interlacing=interlacing, compensate=True, graph=False)()
    batch = Mesh2SpectrumBatch(fields, bin, resampler=resampler, interlacing=interlacing, compensate=True)
    for _ in range(3):
        got = batch()
        torch.cuda.synchronize()
        for g, e in zip(got, eager):
            assert np.array_equal(_values(g, ells), _values(e, ells), equal_nan=True)
    # refill field 1 with new particles (same count): the replay measures the new mock
    rng = np.random.default_rng(9)
    newpos = (rng.uniform(0., 1., (1200, 3)) * attrs.boxsize - attrs.boxsize / 2.).astype('f4')
    batch.load(1, to_dev(newpos))
    got = batch()
    ref = Mesh2SpectrumBatch([fields[1]], bin, resampler=resampler, interlacing=interlacing, compensate=True, graph=False)()[0]
    assert np.array_equal(_values(got[1], ells), _values(ref, ells), equal_nan=True)
    assert np.array_equal(_values(got[0], ells), _values(eager[0], ells), equal_nan=True)
    # and the eager result is the oracle's
    w = host(fields[1].weights).astype('f8')
    oattrs = po.Attrs(attrs.meshsize, attrs.boxsize, attrs.boxcenter)
    omesh = po.paint(oattrs, newpos.astype('f8'), w, resampler=resampler, interlacing=interlacing, compensate=True, out='complex')
    omesh[0, 0, 0] = 0.
    want = po.spectrum_poles(oattrs, po.BinPoles(oattrs, edges={'step': 0.02}, ells=ells), omesh, los='z')
    assert_poles_close([ref.get(ell).values('num_raw') for ell in ells], want['num_raw'], label='batch vs oracle')


def test_batch_streams_graph_matches_eager():
    """Mocks issued on several forked streams (parallel branches of the captured graph) give
    the per-mock spectra; FP64 bin sums may differ in their last bits where blocks of other
    mocks change the order of LDS additions."""
    from jaxpower_amd import MeshAttrs, BinMesh2SpectrumPoles, Mesh2SpectrumBatch
    say('batch streams')
    attrs = MeshAttrs(meshsize=(32, 64, 32), boxsize=(320., 640., 320.), boxcenter=0.)
    ells = (0, 2, 4)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.02}, ells=ells)
    fields = _fields(attrs, [3000, 1200, 5000, 800, 2600], 7)
    eager = Mesh2SpectrumBatch(fields, bin, resampler='tsc', interlacing=2, compensate=True, graph=False)()
    for graph in (False, True):
        batch = Mesh2SpectrumBatch(fields, bin, resampler='tsc', interlacing=2, compensate=True, graph=graph, streams=3)
        for _ in range(2):
            got = batch()
            torch.cuda.synchronize()
            for g, e in zip(got, eager):
                a, b = _values(g, ells), _values(e, ells)
                assert np.array_equal(np.isnan(a), np.isnan(b))
                ok = ~np.isnan(b)
                assert np.allclose(a[ok], b[ok], rtol=1e-12, atol=1e-12 * np.nanmax(np.abs(b)))
