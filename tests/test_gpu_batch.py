"""GPU tests of Mesh2SpectrumBatch: the batch replayed as one captured HIP graph gives the
spectra of the same batch run eagerly, replay after replay and after the inputs are refilled in
place: the painted meshes are the same bit for bit (integer atomics), and the FP64 bin sums
agree to 1e-12 relative (the waves of an x-pass block add into one LDS histogram in an order that
the scheduling decides); the eager batch agrees with the float64 oracle to
|dP| <= 1e-5 max(|P_l|, |P_0|).  The merged form
(one launch per phase for every mock) gives the per-mock spectra: its painted meshes are the
per-mock ones bit for bit, and its FP64 bin sums are added in another order (fewer x-pass blocks
per mock), so they agree to 1e-12 relative."""
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


@pytest.mark.parametrize('resampler,interlacing,merged', [('cic', 2, False), ('tsc', 3, False), ('cic', 2, True),
                                                          ('tsc', 3, True)])
def test_batch_graph_replays_match_eager(resampler, interlacing, merged):
    from jaxpower_amd import MeshAttrs, BinMesh2SpectrumPoles, Mesh2SpectrumBatch
    say('batch graph', resampler, interlacing, 'merged' if merged else 'per mock')
    attrs = MeshAttrs(meshsize=(64, 32, 32), boxsize=(640., 320., 320.), boxcenter=0.)
    ells = (0, 2, 4)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.02}, ells=ells)
    fields = _fields(attrs, [3000, 1200, 5000], 3)
    kw = dict(resampler=resampler, interlacing=interlacing, compensate=True, merged=merged)
    eager = Mesh2SpectrumBatch(fields, bin, graph=False, **kw)()
    batch = Mesh2SpectrumBatch(fields, bin, **kw)
    assert batch.merged == merged
    for _ in range(3):
        got = batch()
        torch.cuda.synchronize()
        for g, e in zip(got, eager):
            _close(_values(g, ells), _values(e, ells))
    # refill field 1 with new particles (same count): the replay measures the new mock
    rng = np.random.default_rng(9)
    newpos = (rng.uniform(0., 1., (1200, 3)) * attrs.boxsize - attrs.boxsize / 2.).astype('f4')
    batch.load(1, to_dev(newpos))
    got = batch()
    refs = Mesh2SpectrumBatch(fields, bin, graph=False, **kw)()  # the same batch, eager, after the refill
    ref = refs[1]
    _close(_values(got[1], ells), _values(ref, ells))
    _close(_values(got[0], ells), _values(eager[0], ells))
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
    eager = Mesh2SpectrumBatch(fields, bin, resampler='tsc', interlacing=2, compensate=True, graph=False, merged=False)()
    for graph in (False, True):
        batch = Mesh2SpectrumBatch(fields, bin, resampler='tsc', interlacing=2, compensate=True, graph=graph, streams=3,
                                   merged=False)
        for _ in range(2):
            got = batch()
            torch.cuda.synchronize()
            for g, e in zip(got, eager):
                a, b = _values(g, ells), _values(e, ells)
                assert np.array_equal(np.isnan(a), np.isnan(b))
                ok = ~np.isnan(b)
                assert np.allclose(a[ok], b[ok], rtol=1e-12, atol=1e-12 * np.nanmax(np.abs(b)))


def _close(a, b, rtol=1e-12):
    assert np.array_equal(np.isnan(a), np.isnan(b))
    ok = ~np.isnan(b)
    assert np.allclose(a[ok], b[ok], rtol=rtol, atol=rtol * np.nanmax(np.abs(b))), np.nanmax(np.abs(a - b))


@pytest.mark.parametrize('resampler,interlacing,compensate,los,pos_f64', [
    ('cic', 2, True, 'z', False), ('tsc', 3, True, 'x', False), ('pcs', 2, False, (0.6, 0., 0.8), False),
    ('ngp', 2, True, 'y', True), ('tsc', 1, True, 'z', False)])
def test_merged_batch_matches_per_mock(resampler, interlacing, compensate, los, pos_f64):
    """One launch per phase for every mock = the per-mock sequence: the multipoles to 1e-12
    relative, the box normalization and the shot noise to 1e-13."""
    from jaxpower_amd import MeshAttrs, BinMesh2SpectrumPoles, Mesh2SpectrumBatch, ParticleField
    say('merged batch', resampler, interlacing, compensate, los, pos_f64)
    attrs = MeshAttrs(meshsize=(32, 64, 32), boxsize=(320., 640., 320.), boxcenter=(10., -20., 30.))
    ells = (0, 1, 2, 4)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.02}, ells=ells)
    fields = _fields(attrs, [3000, 1, 5000, 800, 26000, 12], 11)
    if pos_f64:
        fields = [ParticleField(f.positions.double(), f.weights, attrs=attrs) for f in fields]
    kw = dict(resampler=resampler, interlacing=interlacing, compensate=compensate, los=los)
    per = Mesh2SpectrumBatch(fields, bin, graph=False, merged=False, **kw)()
    for graph in (False, True):
        batch = Mesh2SpectrumBatch(fields, bin, graph=graph, merged=True, **kw)
        got = batch()
        for g, e in zip(got, per):
            for ell in ells:
                _close(g.get(ell).values('num_raw'), e.get(ell).values('num_raw'))
                _close(g.get(ell).values('norm'), e.get(ell).values('norm'), rtol=1e-13)
                _close(g.get(ell).values('num_shotnoise'), e.get(ell).values('num_shotnoise'), rtol=1e-13)
            assert np.array_equal(g.get(0).values('nmodes'), e.get(0).values('nmodes'))
            assert g.attrs == e.attrs


def test_merged_batch_uneven_and_empty_mocks():
    """Mocks of very different sizes, an empty one among them, at a mesh whose mocks need
    several row-sort chunks: each mock's spectrum is the float64 oracle's, the empty mock's
    sums are 0."""
    from jaxpower_amd import MeshAttrs, BinMesh2SpectrumPoles, Mesh2SpectrumBatch
    say('merged batch uneven')
    attrs = MeshAttrs(meshsize=64, boxsize=640., boxcenter=0.)
    ells = (0, 2, 4)
    edges = {'step': 0.02}
    bin = BinMesh2SpectrumPoles(attrs, edges=edges, ells=ells)
    counts = [70000, 0, 17, 40000, 0, 123457]
    fields = _fields(attrs, counts, 5)
    got = Mesh2SpectrumBatch(fields, bin, resampler='tsc', interlacing=2, compensate=True, graph=False, merged=True)()
    oattrs = po.Attrs(attrs.meshsize, attrs.boxsize, attrs.boxcenter)
    obin = po.BinPoles(oattrs, edges=edges, ells=ells)
    for m, (f, c) in enumerate(zip(fields, counts)):
        num = np.array([got[m].get(ell).values('num_raw') for ell in ells])
        if c == 0:
            assert np.all(num == 0.) and float(got[m].get(0).values('norm')[0]) == 0.
            continue
        omesh = po.paint(oattrs, host(f.positions).astype('f8'), host(f.weights).astype('f8'), resampler='tsc', interlacing=2,
                         compensate=True, out='complex')
        omesh[0, 0, 0] = 0.
        want = po.spectrum_poles(oattrs, obin, omesh, los='z')
        assert_poles_close(num, want['num_raw'], label=f'merged mock {m} ({c} particles)')
