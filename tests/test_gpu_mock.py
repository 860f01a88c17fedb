"""The seeded input generators against the distributions they state (jaxpower_amd/mock.py).

* Gaussian field with unitary amplitude: its monopole is the bin average of P(k) exactly
  (the closed form of the reference's tests/test_mesh2.py:23-61), to float32 precision.
* Gaussian field with random amplitude: per-bin monopoles within 5 sigma of the bin average of
  P(k) (sigma = P sqrt(2 / nmodes), Gaussian variance of a mode average).
* Lognormal Poisson mock: with the Gaussian field regenerated from the same seed, the per-cell
  counts against lambda = nbar Vcell exp(G - var/2): mean count, and the Pearson statistic
  sum (N - lambda)^2 / lambda within 5 sigma of the number of cells (Poisson sampling);
  positions uniform within their cell.
"""
import numpy as np
import pytest
import torch

from oracle import power_oracle as po

from test_gpu_parity import host, say

pytestmark = pytest.mark.gpu


def _bin_average(attrs, edges, pk):
    """sum over the modes of each bin of w P(|k|) / nmodes with the hermitian weights."""
    oattrs = po.Attrs(attrs.meshsize, attrs.boxsize, attrs.boxcenter)
    obin = po.BinPoles(oattrs, edges=edges, ells=(0,))
    return obin(pk(po.knorm(oattrs))).real, obin.nmodes


def _pk_numpy(k, amplitude=2e4, k0=0.1, index=-1.5, kcut=0.5):
    with np.errstate(divide='ignore'):
        p = amplitude * np.where(k > 0, k / k0, 1.)**index * np.exp(-(k / kcut)**2)
    return np.where(k > 0, p, 0.)


@pytest.mark.parametrize('unitary', [True, False])
def test_gaussian_mesh_power(unitary):
    from jaxpower_amd import MeshAttrs, BinMesh2SpectrumPoles, compute_mesh2_spectrum
    from jaxpower_amd.mock import generate_gaussian_mesh, power_law_cutoff
    say('gaussian mesh', unitary)
    attrs = MeshAttrs(meshsize=128, boxsize=1000., boxcenter=500.)
    mesh = generate_gaussian_mesh(attrs, power=power_law_cutoff(), seed=7, unitary_amplitude=unitary, device='cuda')
    edges = {'step': 0.02}
    bin = BinMesh2SpectrumPoles(attrs, edges=edges, ells=(0,))
    p0 = compute_mesh2_spectrum(mesh, bin=bin, los='z').get(0).value()
    want, nmodes = _bin_average(attrs, edges, _pk_numpy)
    ok = nmodes > 0
    if unitary:
        rel = np.abs(p0 - want)[ok] / want[ok]
        print('unitary max rel err', rel.max())
        assert rel.max() < 2e-5
    else:
        pull = (p0 - want)[ok] / (want[ok] * np.sqrt(2. / nmodes[ok]))
        print('max |pull|', np.abs(pull).max())
        assert np.abs(pull).max() < 5.
        assert abs(pull.mean()) < 5. / np.sqrt(pull.size)


def test_lognormal_mock_counts():
    from jaxpower_amd import MeshAttrs
    from jaxpower_amd.mock import generate_gaussian_mesh, generate_lognormal_particles, power_law_cutoff
    say('lognormal counts')
    attrs = MeshAttrs(meshsize=64, boxsize=500., boxcenter=250.)
    size, seed = 400_000, 11
    particles = generate_lognormal_particles(attrs, size=size, seed=seed)
    pos = host(particles.positions).astype('f8')
    ncell = 64**3
    xg = pos / attrs.cellsize  # boxcenter - boxsize / 2 = 0
    assert np.all((xg >= 0.) & (xg < 64.))
    idx = np.floor(xg).astype('i8')
    counts = np.bincount((idx[:, 0] * 64 + idx[:, 1]) * 64 + idx[:, 2], minlength=ncell)
    G = host(generate_gaussian_mesh(attrs, power=power_law_cutoff(), seed=seed, device='cuda').value).astype('f8').ravel()
    lam = size / ncell * np.exp(G - 0.5 * G.var())
    # the expected total is sum(lambda): `size` is its mean over realizations of G
    assert abs(counts.sum() - lam.sum()) < 5. * np.sqrt(lam.sum())
    assert abs(lam.sum() / size - 1.) < 0.2
    pearson = np.sum((counts - lam)**2 / lam)
    print('pearson / ncell', pearson / ncell)
    assert abs(pearson - ncell) < 5. * np.sqrt(2. * ncell)
    # a lognormal field, not a uniform one: the counts correlate with lambda
    assert np.corrcoef(counts, lam)[0, 1] > 0.3
    frac = xg - idx  # uniform within the cell
    assert np.all(np.abs(frac.mean(axis=0) - 0.5) < 5. * np.sqrt(1. / 12. / len(frac)))
    assert np.all(np.abs(frac.var(axis=0) - 1. / 12.) < 5. * np.sqrt(1. / 180. / len(frac)))
