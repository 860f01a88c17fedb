"""The autotuner: every candidate of the x-pass and y-pass tile widths and of the row-sort
scatter chunk, forced one at a time, gives the float64 oracle's multipoles; measured choices
are recorded with their timings, and the table survives a JSON round trip."""
import json

import numpy as np
import pytest
import torch

from oracle import power_oracle as po

pytestmark = pytest.mark.gpu


def say(*args):
    print('[case]', *args, flush=True)


def _spectrum(p, attrs, bin):
    from jaxpower_amd import compute_mesh2_spectrum
    mesh = p.paint(resampler='tsc', interlacing=3, compensate=True, out='complex')
    return compute_mesh2_spectrum(mesh, bin=bin, los='z')


@pytest.fixture(scope='module')
def case():
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh2SpectrumPoles
    rng = np.random.default_rng(41)
    attrs = MeshAttrs(meshsize=(128, 64, 64), boxsize=(640., 320., 320.), boxcenter=0.)
    pos = rng.uniform(-0.5, 0.5, (300000, 3)) * attrs.boxsize  # >= 2^18 particles: the sort chunk is tunable
    p = ParticleField(torch.from_numpy(pos).cuda(), attrs=attrs)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': 0.02}, ells=(0, 2, 4))
    oattrs = po.Attrs(attrs.meshsize, attrs.boxsize, attrs.boxcenter)
    omesh = po.paint(oattrs, pos, None, resampler='tsc', interlacing=3, compensate=True, out='complex')
    want = po.spectrum_poles(oattrs, po.BinPoles(oattrs, edges={'step': 0.02}, ells=(0, 2, 4)), omesh, los='z')
    return p, attrs, bin, want


def _check(spec, want, label):
    assert np.array_equal(spec.get(0).values('nmodes'), want['nmodes']), label
    scale = np.abs(np.asarray(want['num_raw'][0]))
    for ell, w in zip((0, 2, 4), want['num_raw']):
        g = spec.get(ell).values('num_raw')
        err = np.max(np.abs(g - w) / np.maximum(np.abs(w), scale))
        assert err <= 1e-5, (label, ell, err)


@pytest.mark.parametrize('knob', ['xpass_cols', 'ypass_cols', 'sort_chunk'])
def test_every_candidate_matches_oracle(case, knob):
    from jaxpower_amd import tuning
    p, attrs, bin, want = case
    for q in range(3):
        say('tuning candidate', knob, q)
        with tuning.force(knob, q):
            _check(_spectrum(p, attrs, bin), want, (knob, q))


def test_measured_choices_and_json_round_trip(case, tmp_path):
    from jaxpower_amd import tuning
    p, attrs, bin, want = case
    say('tuning measurement and JSON round trip')
    tuning.clear()
    _check(_spectrum(p, attrs, bin), want, 'measured')
    entries = tuning.table()
    knobs = {e['knob'] for e in entries.values()}
    assert {'xpass_cols', 'ypass_cols', 'sort_chunk'} <= knobs, entries
    for key, e in entries.items():
        assert e['measured'] and 0 <= e['choice'] < len(e['ms']), (key, e)
        assert e['ms'][e['choice']] == min(t for t in e['ms'] if t >= 0.), (key, e)
    path = str(tmp_path / 'tuning.json')
    assert tuning.save(path) == len(entries)
    doc = json.load(open(path))
    assert doc['entries'] == entries
    tuning.clear()
    assert tuning.table() == {}
    assert tuning.load(path) == len(entries)
    again = tuning.table()
    assert {k: e['choice'] for k, e in again.items()} == {k: e['choice'] for k, e in entries.items()}
    assert not any(e['measured'] for e in again.values())
    _check(_spectrum(p, attrs, bin), want, 'imported')
