"""Pins the float64 oracle against the reference's own known-answer tests, closed forms,
invariants and brute force.  Each check is chosen so that a plausible slip in the oracle
(dropped term, wrong sign or index, transposed operand) fails at least one of them.

Reference tests mirrored: tests/test_mesh.py:24-48 (get_mesh_attrs KATs), 110-121 and
216-232 (mass conservation), 89-107 (round trip); tests/test_mesh2.py:23-61 (closed-form P0
of a unitary-amplitude Gaussian field), 75-91 (shot noise), 459-467 (Ylm);
tests/test_particle2.py:58-63 (Legendre).
"""
import itertools
import math

import numpy as np
import pytest
from scipy import special

from oracle import power_oracle as po


# ---------------------------------------------------------------- resamplers

@pytest.mark.parametrize('order,values', [
    (1, {0.: 1., 0.3: 1.}),
    (2, {0.: 1., 0.25: 0.75, 1.: 0.}),
    (3, {0.: 0.75, 0.5: 0.5, 1.: 0.125, 1.5: 0.}),
    (4, {0.: 2. / 3., 1.: 1. / 6., 2.: 0.}),
])
def test_kernel_closed_forms(order, values):
    for s, expected in values.items():
        assert np.isclose(po.resampler_kernel(order, s), expected, rtol=0, atol=1e-15)


@pytest.mark.parametrize('order', [1, 2, 3, 4])
def test_partition_of_unity(order):
    rng = np.random.default_rng(order)
    x = rng.uniform(-3, 3, 1000)
    id0 = (np.round if order % 2 else np.floor)(x)
    offs = np.arange(order) - (order - 1) // 2
    tot = sum(po.resampler_kernel(order, np.abs(id0 + o - x)) for o in offs)
    assert np.allclose(tot, 1., atol=1e-14)


@pytest.mark.parametrize('order', [1, 2, 3, 4])
def test_paint_vs_bruteforce_and_invariants(order):
    rng = np.random.default_rng(10 + order)
    shape = (4, 3, 5)  # tests/test_mesh.py:112
    xg = rng.uniform(-1, 6, (50, 3))
    w = 1. + rng.uniform(size=50)
    mesh = po.paint_grid(shape, xg, w, order)
    assert np.allclose(mesh, po.paint_grid_bruteforce(shape, xg, w, order), atol=1e-13)
    assert np.isclose(mesh.sum(), w.sum())  # tests/test_mesh.py:119
    # shifting every particle by one cell along y rolls the mesh by one along y
    shifted = po.paint_grid(shape, xg + np.array([0., 1., 0.]), w, order)
    assert np.allclose(shifted, np.roll(mesh, 1, axis=1), atol=1e-13)


def test_paint_integer_position_kats():
    # a CIC particle exactly on node (1,2,3) lands only there; TSC spreads 3/4 and 1/8 per axis
    mesh = po.paint_grid((6, 6, 6), np.array([[1., 2., 3.]]), np.array([2.]), 2)
    assert mesh[1, 2, 3] == 2. and np.count_nonzero(mesh) == 1
    mesh = po.paint_grid((6, 6, 6), np.array([[1., 2., 3.]]), np.array([1.]), 3)
    assert np.isclose(mesh[1, 2, 3], 0.75**3) and np.isclose(mesh[0, 2, 3], 0.125 * 0.75**2)
    # periodic wrap: x~ = -0.25 with CIC paints 0.25 into cell N-1 and 0.75 into cell 0
    mesh = po.paint_grid((4, 4, 4), np.array([[-0.25, 0., 0.]]), np.array([1.]), 2)
    assert np.isclose(mesh[3, 0, 0], 0.25) and np.isclose(mesh[0, 0, 0], 0.75)


@pytest.mark.parametrize('resampler', ['ngp', 'cic', 'tsc', 'pcs'])
@pytest.mark.parametrize('interlacing', [0, 2, 3])
@pytest.mark.parametrize('compensate', [False, True])
def test_paint_conserves_sum(resampler, interlacing, compensate):
    # tests/test_mesh.py:216-232: paint().sum() == sum(weights) for every option
    rng = np.random.default_rng(3)
    attrs = po.Attrs(meshsize=(8, 6, 10), boxsize=(80., 60., 100.), boxcenter=(1., 2., 3.))
    pos = rng.uniform(-40, 40, (200, 3)) + attrs.boxcenter
    w = 1. + rng.uniform(size=200)
    mesh = po.paint(attrs, pos, w, resampler=resampler, interlacing=interlacing, compensate=compensate)
    assert mesh.shape == (8, 6, 10)
    assert np.isclose(mesh.sum(), w.sum(), rtol=1e-12)


def _direct_interlaced_ngp(meshsize, xg, w, ni):
    """Interlaced NGP spectrum written from its meaning: the particle painted at x~ + s lands on
    node n = round(x~ + s), which stands for the physical grid point n - s of the unshifted
    mesh; the n_i meshes are averaged as Fourier sums over those points,
    A(k) = 1/n_i sum_j sum_p w_p exp(-i k.(n_pj - s_j)), with k_a = 2 pi m_a / N_a and the
    reference's index -> frequency map (numpy fftfreq on the first axes, rfftfreq on the last).
    A direct sum over particles and half-spectrum modes; no FFT, no phase factor table."""
    n = np.array(meshsize)
    freqs = [np.fft.fftfreq(n[0]) * 2. * np.pi, np.fft.fftfreq(n[1]) * 2. * np.pi, np.fft.rfftfreq(n[2]) * 2. * np.pi]
    K = np.stack(np.meshgrid(*freqs, indexing='ij'), axis=-1)  # (nx, ny, nz//2+1, 3)
    out = np.zeros(K.shape[:3], dtype='c16')
    for j in range(ni):
        s = j / ni
        nodes = np.round(xg + s) - s  # physical grid points, in grid units (half-even ties excluded below)
        for x, wp in zip(nodes, w):
            out += wp * np.exp(-1j * (K @ x)) / ni
    return out


@pytest.mark.parametrize('ni', [2, 3])
@pytest.mark.parametrize('meshsize', [(8, 8, 8), (6, 4, 10)])
def test_interlacing_phase_against_direct_sum(ni, meshsize):
    """The interlacing combine (mesh.py:2738-2747: exp(+i s sum k_circ) / n_i, summed over the
    shifts) against the direct Fourier sum over the physical node positions of every shifted
    NGP paint, on all modes including the Nyquist planes.  A wrong sign of the phase, a shift of
    -s instead of +s, a missing 1/n_i or a wrong Nyquist frequency fails it."""
    rng = np.random.default_rng(ni + sum(meshsize))
    n = np.array(meshsize)
    attrs = po.Attrs(meshsize=n, boxsize=n * 2.5, boxcenter=(0.3, -0.2, 0.1))
    pos = (rng.uniform(0., 1., (40, 3)) - 0.5) * attrs.boxsize + attrs.boxcenter
    xg = po.to_grid(attrs, pos)
    for j in range(ni):  # keep clear of exact halves, where NGP rounding would be a tie
        assert np.all(np.abs(np.mod(xg + j / ni, 1.) - 0.5) > 1e-6)
    w = rng.uniform(0.5, 1.5, len(pos))
    got = po.paint(attrs, pos, w, resampler='ngp', interlacing=ni, out='complex')
    want = _direct_interlaced_ngp(n, xg, w, ni)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()
    # a plausible mistake, planted in the oracle's phase: the opposite sign fails
    kx, ky, kz = attrs.kcoords('circular')
    ks = kx[:, None, None] + ky[None, :, None] + kz[None, None, :]
    wrong = sum(po.r2c(po.paint_grid(n, xg + j / ni, w, 1)) * np.exp(-1j * (j / ni) * ks) / ni for j in range(ni))
    assert np.abs(wrong - want).max() > 1e-3 * np.abs(want).max()


def test_circular_nyquist_convention():
    """Known answer of the circular frequencies (mesh.py:377-381: jnp.fft.fftfreq on the first
    axes, rfftfreq on the hermitian last axis, times 2 pi): the Nyquist index N/2 is -pi along
    x and y and +pi along z."""
    attrs = po.Attrs(meshsize=(8, 6, 10), boxsize=100.)
    kx, ky, kz = attrs.kcoords('circular')
    assert kx[4] == -np.pi and ky[3] == -np.pi and kz[5] == np.pi
    assert np.allclose(kx[:4], 2. * np.pi * np.arange(4) / 8.) and np.allclose(kx[5:], 2. * np.pi * np.arange(-3, 0) / 8.)
    assert len(kz) == 6 and np.all(np.diff(kz) > 0.)


def test_pole_value_single_particle_is_zero():
    """value() = (num_raw - num_shotnoise) / norm (lsstypes Mesh2SpectrumPole): a single particle
    of weight w painted with NGP on a node has |A(k)|^2 = w^2 in every mode and shot noise
    sum w^2 = w^2 (mesh2.py:603), so the monopole value is exactly 0 in every bin; with the
    shot noise added instead of subtracted it would be 2 w^2 / norm."""
    attrs = po.Attrs(meshsize=16, boxsize=160., boxcenter=0.)
    w = np.array([1.7])
    pos = np.array([[20., -30., 50.]])  # a node: grid units (10, 5, 13)
    mesh = po.paint(attrs, pos, w, resampler='ngp', interlacing=0, out='complex')
    assert np.allclose(np.abs(mesh)**2, w[0]**2)
    bin = po.BinPoles(attrs, edges={'step': 0.02}, ells=(0, 2, 4))
    res = po.spectrum_poles(attrs, bin, mesh, los='z')
    sn = po.fkp_shotnoise(w)
    assert sn == w[0]**2
    norm = po.box_normalization(attrs, w)
    value = po.pole_value(res['num_raw'][0], norm, sn)
    assert np.all(np.abs(value) < 1e-12 * w[0]**2 / norm)


def test_pole_value_two_particles_closed_form():
    """Two particles on nodes x1, x2: |A(k)|^2 = w1^2 + w2^2 + 2 w1 w2 cos(k.(x1 - x2)), so
    value() in each bin is 2 w1 w2 <cos(k.d)>_bin / norm with the box normalization
    (sum w)^2 / (N^3 V_cell) (mesh.py:3121-3129), written here as an explicit sum over the
    bin's modes (hermitian weights 1 + (kz > 0))."""
    attrs = po.Attrs(meshsize=16, boxsize=160., boxcenter=0.)
    w = np.array([1.3, 0.6])
    pos = np.array([[20., -30., 50.], [-40., 0., 10.]])
    mesh = po.paint(attrs, pos, w, resampler='ngp', interlacing=0, out='complex')
    bin = po.BinPoles(attrs, edges={'step': 0.03}, ells=(0,))
    res = po.spectrum_poles(attrs, bin, mesh, los='z')
    norm = po.box_normalization(attrs, w)
    assert np.isclose(norm, w.sum()**2 / (16**3 * 10.**3))
    value = po.pole_value(res['num_raw'][0], norm, po.fkp_shotnoise(w))
    kx, ky, kz = attrs.kcoords('wavenumber')
    d = pos[0] - pos[1]
    want = np.zeros(bin.edges.shape[0])
    count = np.zeros(bin.edges.shape[0])
    for a in kx:
        for b in ky:
            for ic, c in enumerate(kz):
                k = np.sqrt(a * a + b * b + c * c)
                ib = np.searchsorted(bin.edges[:, 0], k, side='right') - 1
                if ib < 0 or k >= bin.edges[ib, 1]:
                    continue
                wm = 1 + (c > 0)
                want[ib] += wm * 2. * w[0] * w[1] * np.cos(a * d[0] + b * d[1] + c * d[2])
                count[ib] += wm
    with np.errstate(invalid='ignore'):
        want = want / count / norm
    ok = count > 0
    assert np.array_equal(count[ok], bin.nmodes[ok])
    assert np.allclose(value[ok], want[ok], rtol=1e-10, atol=1e-12 * np.abs(want).max())


# ---------------------------------------------------------------- FFT

@pytest.mark.parametrize('shape', [(2, 2, 2), (4, 2, 6), (3, 5, 4), (4, 4, 4)])
def test_dft_matches_definition(shape):
    rng = np.random.default_rng(7)
    x = rng.normal(size=shape)
    assert np.allclose(po.r2c(x), po.dft3_half(x), atol=1e-10)
    assert np.allclose(po.c2r(po.r2c(x), shape), x, atol=1e-12)  # tests/test_mesh.py:94


def test_compensation_closed_forms():
    attrs = po.Attrs(meshsize=8, boxsize=100.)
    for order in (1, 2, 3, 4):
        comp = po.compensation(attrs, order)
        assert comp[0, 0, 0] == 1.
        # Nyquist along z: k = pi -> sinc(1/2) = 2/pi -> (pi/2)^order
        assert np.isclose(comp[0, 0, 4], (np.pi / 2)**order)
        # k = 2pi/8 along x: sin(pi/8)/(pi/8)
        assert np.isclose(comp[1, 0, 0], (np.sin(np.pi / 8) / (np.pi / 8))**(-order))


# ---------------------------------------------------------------- binning

def _lattice_counts(attrs, edges):
    """Full-grid (not half) integer enumeration of modes per bin."""
    n = attrs.meshsize
    m = [np.where(np.arange(k) < (k + 1) // 2, np.arange(k), np.arange(k) - k) for k in n]
    kf = attrs.kfun
    k = np.sqrt((m[0][:, None, None] * kf[0])**2 + (m[1][None, :, None] * kf[1])**2 + (m[2][None, None, :] * kf[2])**2)
    return np.histogram(k.ravel(), bins=edges)[0]


@pytest.mark.parametrize('meshsize,boxsize,step', [(16, 1000., 0.01), (32, 2000., 0.005), ((16, 8, 12), (100., 80., 90.), 0.02)])
def test_nmodes_full_grid_enumeration(meshsize, boxsize, step):
    attrs = po.Attrs(meshsize=meshsize, boxsize=boxsize)
    bin = po.BinPoles(attrs, edges={'step': step}, ells=(0, 2, 4))
    edges = np.append(bin.edges[:, 0], bin.edges[-1, 1])
    assert edges[-1] < attrs.knyq.min()  # edges stop below Nyquist (mesh2.py:39-44)
    full = _lattice_counts(attrs, edges)
    # hermitian weights 1 + (kz > 0) reproduce the full-grid count below Nyquist
    assert np.array_equal(bin.nmodes, full)
    assert bin.nmodes.dtype.kind == 'i'
    # first bin holds the DC mode (count 1 + its neighbours)
    assert bin.ibin[0, 0, 0] == 1


def test_nmodes_integer_thresholds_cubic():
    # cubic box: bin membership is a function of the integer n^2 = i^2 + j^2 + l^2
    attrs = po.Attrs(meshsize=64, boxsize=1000.)
    bin = po.BinPoles(attrs, edges={'step': 0.01})
    edges = np.append(bin.edges[:, 0], bin.edges[-1, 1])
    kf = attrs.kfun[0]
    nmax = 32
    counts = np.zeros(len(edges) - 1, dtype='i8')
    r = np.arange(-nmax, nmax)
    n2 = (r[:, None, None]**2 + r[None, :, None]**2 + r[None, None, :]**2).ravel()
    thr = (edges / kf)**2
    ib = np.searchsorted(thr, n2, side='right')
    valid = (ib >= 1) & (ib < len(edges))
    np.add.at(counts, ib[valid] - 1, 1)
    assert np.array_equal(counts, bin.nmodes)


def test_xavg_and_bincount_kats():
    attrs = po.Attrs(meshsize=8, boxsize=2 * np.pi * 8)  # kfun = 1/8
    bin = po.BinPoles(attrs, edges=np.array([0., 0.1, 0.13, 0.2]))
    # bin 0 [0,0.1): DC only; bin 1 [0.1,0.13): |k|=1/8 (6 modes); bin 2 [0.13, 0.2): sqrt2/8 (12 modes)
    assert list(bin.nmodes) == [1, 6, 12]
    assert np.allclose(bin.xavg, [0., 0.125, np.sqrt(2) / 8])
    ones = np.ones((8, 8, 5))
    assert np.allclose(bin(ones), 1.)
    assert np.allclose(bin(ones, remove_zero=True), [0., 1., 1.])  # nmodes keeps the DC mode


def _bruteforce_oversampled(vec, edges, wz, m, value):
    """Mode oversampling by explicit loops (mesh2.py:48-61, mesh.py:2874-2910): every integer
    shift s in [-m, m]^3, every mode, bin = bisect of |v + s| in the edges, weight wz[iz];
    returns nmodes (count / number of shifts when there are several), xavg = sum w |v + s| /
    nmodes, and the binned value = sum w value / number of shifts / nmodes."""
    import bisect
    shifts = list(itertools.product(range(-m, m + 1), repeat=3))
    nb = len(edges) - 1
    count, xs, sums = [0] * nb, [0.] * nb, [0j] * nb
    for sh in shifts:
        for i, x in enumerate(vec[0]):
            for j, y in enumerate(vec[1]):
                for l, z in enumerate(vec[2]):
                    a, b, c = x + sh[0], y + sh[1], z + sh[2]
                    r = math.sqrt((a * a + b * b) + c * c)
                    k = bisect.bisect_right(list(edges), r) - 1
                    if 0 <= k < nb:
                        w = int(wz[l])
                        count[k] += w
                        xs[k] += w * r
                        sums[k] += w * complex(value[i, j, l])
    S = len(shifts)
    nmodes = np.array(count, dtype='i8') / (S if S > 1 else 1)
    with np.errstate(invalid='ignore', divide='ignore'):
        return nmodes, np.array(xs) / nmodes, np.array(sums) / S / nmodes


@pytest.mark.parametrize('m', [1, 2])
def test_mode_oversampling_bruteforce(m):
    """BinPoles(mode_oversampling=m) against explicit loops on a small box whose wavenumbers are
    of order 1 h/Mpc, so that the integer shifts move modes between bins and out of range."""
    attrs = po.Attrs((4, 6, 6), (12., 18., 15.), 0.)
    edges = np.arange(0., 3.01, 0.2)
    rng = np.random.default_rng(m)
    shape = (4, 6, 4)
    value = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    bin = po.BinPoles(attrs, edges=edges, mode_oversampling=m)
    wz = po.hermitian_weights(attrs)[0, 0]
    nmodes, xavg, binned = _bruteforce_oversampled(attrs.kcoords('wavenumber'), edges, wz, m, value)
    assert len(bin.ibin) == (2 * m + 1)**3 and bin.nmodes.dtype == np.float64
    assert np.array_equal(bin.nmodes, nmodes)
    ok = nmodes > 0
    assert ok.sum() >= 8 and np.allclose(bin.xavg[ok], xavg[ok], rtol=1e-13)
    assert np.allclose(bin(value)[ok], binned[ok], rtol=1e-12)
    # the reference divides the summed counts by the number of shifts before xavg = xsum / nmodes
    # (mesh2.py:58-61), so xavg is that many times the mean shifted norm of the bin
    S = (2 * m + 1)**3
    assert np.all((bin.xavg[ok] / S >= bin.edges[ok, 0]) & (bin.xavg[ok] / S < bin.edges[ok, 1]))
    # remove_zero zeroes the (0,0,0) mode in every shifted binning
    v0 = value.copy()
    v0[0, 0, 0] = 0.
    assert np.allclose(bin(value, remove_zero=True)[ok], _bruteforce_oversampled(attrs.kcoords('wavenumber'), edges, wz, m, v0)[2][ok],
                       rtol=1e-12)


def test_mode_oversampling_wide_bin_and_no_shift():
    """One bin holding every shifted mode: nmodes is the weighted mode count (a float), and the
    binned value is the w-weighted mean whatever the shifts.  m = 0 keeps the integer counts and a
    single index array."""
    attrs = po.Attrs((4, 6, 6), (12., 18., 15.), 0.)
    rng = np.random.default_rng(5)
    value = rng.normal(size=(4, 6, 4))
    w = po.hermitian_weights(attrs)
    bin = po.BinPoles(attrs, edges=np.array([0., 1e3]), mode_oversampling=1)
    assert bin.nmodes.dtype == np.float64 and bin.nmodes[0] == w.sum()
    assert np.isclose(bin(value)[0], (w * value).sum() / w.sum(), rtol=1e-14)
    plain = po.BinPoles(attrs, edges=np.arange(0., 3.01, 0.2), mode_oversampling=0)
    assert plain.nmodes.dtype == np.int64 and isinstance(plain.ibin, np.ndarray)


@pytest.mark.parametrize('m', [1])
def test_correlation_mode_oversampling_bruteforce(m):
    """BinCorrelationPoles(mode_oversampling=m): the same loops over the full real mesh, every
    cell weighted 1, shifts in Mpc/h."""
    attrs = po.Attrs((6, 4, 5), (9., 6., 7.5), 0.)
    edges = np.arange(0., 6.01, 0.5)
    rng = np.random.default_rng(7)
    value = rng.normal(size=(6, 4, 5))
    bin = po.BinCorrelationPoles(attrs, edges=edges, mode_oversampling=m)
    nmodes, xavg, binned = _bruteforce_oversampled(po.separation_coords(attrs), edges, np.ones(5), m, value)
    assert np.array_equal(bin.nmodes, nmodes)
    ok = nmodes > 0
    assert ok.sum() >= 8 and np.allclose(bin.xavg[ok], xavg[ok], rtol=1e-13)
    assert np.allclose(bin(value)[ok], binned[ok].real, rtol=1e-12)


def test_find_unique_edges():
    attrs = po.Attrs(meshsize=8, boxsize=2 * np.pi * 8)
    edges = po.find_unique_edges(attrs, xmax=0.3)
    x = np.sqrt(np.array([0, 1, 2, 3, 4, 5]) / 64.)
    mid = (x[:-1] + x[1:]) / 2
    assert np.allclose(edges[1:-1], mid)
    assert np.isclose(edges[0], mid[0] - (x[1] - x[0]))


# ---------------------------------------------------------------- special functions

def test_legendre_vs_scipy():
    x = np.linspace(-1, 1, 101)
    for ell in range(7):
        assert np.allclose(po.legendre(ell, x), special.eval_legendre(ell, x), rtol=1e-9, atol=1e-12)


def test_ylm_normalisation_and_addition_theorem():
    rng = np.random.default_rng(42)
    a = rng.normal(size=(200, 3))
    b = rng.normal(size=(200, 3))
    ah = a / np.linalg.norm(a, axis=1)[:, None]
    bh = b / np.linalg.norm(b, axis=1)[:, None]
    theta = np.arccos(ah[:, 2])
    phi = np.arctan2(ah[:, 1], ah[:, 0])
    for ell in (0, 1, 2, 3, 4):
        s = sum(po.real_ylm(ell, m, *a.T) * po.real_ylm(ell, m, *(3. * b).T) for m in range(-ell, ell + 1))
        assert np.allclose(s, (2 * ell + 1) / (4 * np.pi) * special.eval_legendre(ell, np.sum(ah * bh, axis=1)), atol=1e-12)
        for m in range(-ell, ell + 1):
            # |real Y| matches the complex spherical harmonic of scipy (tests/test_mesh2.py:459-467)
            Y = special.sph_harm_y(ell, abs(m), theta, phi)
            ref = Y.real if m >= 0 else Y.imag
            if m != 0:
                ref = np.sqrt(2.) * ref
            assert np.allclose(np.abs(po.real_ylm(ell, m, *a.T)), np.abs(ref), atol=1e-12)
            assert po.real_ylm(ell, m, 0., 0., 0.) == po.ylm_at_zero(ell, m)


def _sympy_real_ylm(ell, m):
    """The real Y_lm polynomial in (xhat, yhat, zhat) built the way utils.py:299-327 builds it
    (associated Legendre of cos(theta), cos/sin(m phi) expanded, phi and theta replaced by
    Cartesian ratios, x^2 + y^2 + z^2 -> r^2, then x/r -> xhat ...), with the normalisation of
    utils.py:227-238.  Written here from those steps; the reference itself is not imported."""
    sp = pytest.importorskip('sympy')
    amp = np.sqrt((2 * ell + 1) / (4 * np.pi))
    if m != 0:
        amp *= np.sqrt(math.factorial(ell - abs(m)) / math.factorial(ell + abs(m))) * np.sqrt(2.) * (-1)**m
    x, y, z, r = sp.symbols('x y z r', real=True, positive=True)
    xh, yh, zh = sp.symbols('xhat yhat zhat', real=True, positive=True)
    phi, theta = sp.symbols('phi theta', real=True)
    defs = [(sp.sin(phi), y / sp.sqrt(x**2 + y**2)), (sp.cos(phi), x / sp.sqrt(x**2 + y**2)),
            (sp.cos(theta), z / sp.sqrt(x**2 + y**2 + z**2))]
    expr = sp.assoc_legendre(ell, abs(m), sp.cos(theta))
    expr *= sp.expand_trig(sp.sin(abs(m) * phi) if m < 0 else (sp.cos(m * phi) if m > 0 else 1))
    expr = sp.together(expr.subs(defs)).subs(x**2 + y**2 + z**2, r**2)
    expr = amp * expr.expand().subs([(x / r, xh), (y / r, yh), (z / r, zh)])
    return sp.lambdify((xh, yh, zh), expr, modules=['numpy']), (xh, yh, zh), expr


def test_ylm_zero_vector_matches_sympy_construction():
    """utils.py:252-254 sends the zero vector to xhat = 0 and evaluates the sympy polynomial
    there: its constant term.  The oracle's ylm_at_zero must equal that for every l <= 4 (nonzero
    only for Y_00 and Y_40), and on unit vectors the polynomial must be the oracle's Y_lm up to
    the per-m sign."""
    rng = np.random.default_rng(5)
    u = rng.normal(size=(50, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    for ell in range(5):
        for m in range(-ell, ell + 1):
            f, syms, expr = _sympy_real_ylm(ell, m)
            at0 = float(expr.subs({s_: 0 for s_ in syms}))
            assert np.isclose(po.ylm_at_zero(ell, m), at0, rtol=1e-12, atol=1e-15), (ell, m, at0)
            vals = np.broadcast_to(np.asarray(f(*u.T), dtype='f8'), (len(u),))
            assert np.allclose(np.abs(vals), np.abs(po.real_ylm(ell, m, *u.T)), atol=1e-12)
    assert po.ylm_at_zero(4, 0) > 0.3 and po.ylm_at_zero(2, 0) == 0.


def test_addition_at_zero_vectors():
    """addition_at(l, u, v) = 4 pi/(2l+1) sum_m Y_lm(u) Y_lm(v) with the oracle's Y_lm, zero
    vectors included."""
    rng = np.random.default_rng(6)
    u = rng.normal(size=(30, 3))
    v = rng.normal(size=(30, 3))
    u[:5] = 0.
    v[3:8] = 0.
    for ell in range(5):
        s = sum(po.real_ylm(ell, m, *u.T) * po.real_ylm(ell, m, *v.T) for m in range(-ell, ell + 1)) * 4. * np.pi / (2 * ell + 1)
        assert np.allclose(po.addition_at(ell, u, v), s, atol=1e-12), ell


# ---------------------------------------------------------------- estimator

def _unitary_gaussian(attrs, pk, seed=0):
    """Oracle version of generate_gaussian_mesh(unitary_amplitude=True) (mock.py:15-47)."""
    rng = np.random.default_rng(seed)
    c = po.r2c(rng.normal(size=tuple(attrs.meshsize)))
    k = po.knorm(attrs)
    ker = np.sqrt(pk(k) / np.prod(attrs.cellsize)) * np.sqrt(float(np.prod(attrs.meshsize))) / np.abs(c)
    return po.c2r(c * ker, attrs.meshsize)


def test_p0_closed_form_gaussian_field():
    # tests/test_mesh2.py:23-61: a unitary-amplitude field gives P0(k) = mode-average of P(k)
    attrs = po.Attrs(meshsize=32, boxsize=1000.)
    pk = lambda k: np.where(k > 0, 1e4 * (k / 0.03)**3 * np.exp(-k / 0.03), 0.)
    mesh = _unitary_gaussian(attrs, pk)
    bin = po.BinPoles(attrs, edges={'step': 0.01}, ells=(0, 2, 4))
    res = po.spectrum_poles(attrs, bin, mesh, los='x')
    p0 = po.pole_value(res['num_raw'][0], res['norm'])
    # independent lattice sum over the full grid
    n = 32
    m = np.where(np.arange(n) < n // 2, np.arange(n), np.arange(n) - n)
    kf = attrs.kfun[0]
    kk = kf * np.sqrt(m[:, None, None]**2 + m[None, :, None]**2 + m[None, None, :]**2)
    edges = np.append(bin.edges[:, 0], bin.edges[-1, 1])
    num = np.histogram(kk.ravel(), bins=edges, weights=pk(kk).ravel())[0]
    cnt = np.histogram(kk.ravel(), bins=edges)[0]
    assert np.allclose(p0, num / cnt, rtol=1e-9)
    # isotropic power: higher multipoles vanish only on average; P2 is the lattice average of 5 P L2
    mu = np.where(kk > 0, m[:, None, None] * kf / np.where(kk > 0, kk, 1.), 0.)
    num2 = np.histogram(kk.ravel(), bins=edges, weights=(5 * pk(kk) * special.eval_legendre(2, mu)).ravel())[0]
    p2 = po.pole_value(res['num_raw'][1], res['norm'])
    assert np.allclose(p2, num2 / cnt, rtol=1e-8, atol=1e-8 * np.abs(num / cnt).max())


def test_los_axis_permutation():
    # the multipoles along x of a field equal those along z of the axis-swapped field
    rng = np.random.default_rng(5)
    attrs = po.Attrs(meshsize=16, boxsize=500.)
    mesh = rng.normal(size=(16, 16, 16))
    bin = po.BinPoles(attrs, edges={'step': 0.02}, ells=(0, 2, 4))
    rx = po.spectrum_poles(attrs, bin, mesh, los='x')
    rz = po.spectrum_poles(attrs, bin, np.ascontiguousarray(np.swapaxes(mesh, 0, 2)), los='z')
    for a, b in zip(rx['num_raw'], rz['num_raw']):
        assert np.allclose(a, b, rtol=1e-10, atol=1e-12 * np.abs(rx['num_raw'][0]).max())


def test_cross_equals_auto_for_same_mesh():
    rng = np.random.default_rng(8)
    attrs = po.Attrs(meshsize=8, boxsize=100.)
    mesh = rng.normal(size=(8, 8, 8))
    bin = po.BinPoles(attrs, edges={'step': 0.1}, ells=(0, 1, 2))
    auto = po.spectrum_poles(attrs, bin, mesh, los='z')
    cross = po.spectrum_poles(attrs, bin, mesh, mesh.copy(), los='z')
    for a, b in zip(auto['num_raw'], cross['num_raw']):
        assert np.allclose(a, b, equal_nan=True, atol=1e-9)


@pytest.mark.parametrize('ell', [2, 4])
@pytest.mark.parametrize('boxcenter', [(100., -20., 45.), 0.])
def test_local_los_vs_bruteforce(ell, boxcenter):
    # firstpoint estimator (Ylm sums + FFTs) equals the direct double sum with P_l(khat.xhat);
    # with boxcenter 0 the origin is a node, where the reference's Y_40 is not 0
    rng = np.random.default_rng(ell)
    attrs = po.Attrs(meshsize=6, boxsize=60., boxcenter=boxcenter)
    mesh = rng.normal(size=(6, 6, 6))
    bin = po.BinPoles(attrs, edges={'step': 0.05}, ells=(0, ell))
    res = po.spectrum_poles(attrs, bin, mesh, los='firstpoint')
    direct = bin(po.local_los_bruteforce(attrs, mesh, ell), remove_zero=True)
    assert np.allclose(res['num_raw'][1], direct.real, rtol=1e-9, atol=1e-9 * np.nanmax(np.abs(direct)), equal_nan=True)
    # endpoint swaps the meshes and conjugates (mesh2.py:296, 337): for an auto spectrum
    # of even ell it equals firstpoint
    res_end = po.spectrum_poles(attrs, bin, mesh, los='endpoint')
    assert np.allclose(res_end['num_raw'][1], res['num_raw'][1], equal_nan=True)


# ---------------------------------------------------------------- mesh attrs, norm, shot noise

def test_get_mesh_attrs_kats():
    # tests/test_mesh.py:24-48
    attrs = po.get_mesh_attrs(boxsize=1000., boxcenter=0., cellsize=10.)
    assert np.all(attrs.meshsize == 100) and np.allclose(attrs.cellsize, 10.)
    attrs = po.get_mesh_attrs(boxsize=1000., boxcenter=0., meshsize=128)
    assert np.all(attrs.meshsize == 128) and np.allclose(attrs.cellsize, 1000. / 128)
    rng = np.random.default_rng(84)
    pos = rng.uniform(0., 100., (100, 3))
    attrs = po.get_mesh_attrs(pos, meshsize=128)
    assert np.all(attrs.meshsize == 128)
    attrs = po.get_mesh_attrs(pos, cellsize=12.)
    assert np.allclose(attrs.cellsize, 12.)
    attrs = po.get_mesh_attrs(pos, cellsize=12., primes=(2, 3, 5), divisors=(2, 3))
    assert np.all(attrs.meshsize == 18) and np.allclose(attrs.cellsize, 12.)
    assert po.next_fft_size(17, primes=(2, 3, 5, 7)) == 18
    assert po.next_fft_size(97) == 98 and po.next_fft_size(121, primes=(2,)) == 128


def test_regrid_parity_rule():
    # mesh.py:3024-3033: new meshsize ceil(L/10); if any axis changes parity add 1 to every axis
    attrs = po.Attrs(meshsize=64, boxsize=1000., boxcenter=500.)
    new = po.regrid_attrs(attrs, 10.)
    assert np.all(new.meshsize == 100) and np.allclose(new.boxsize, 1000.)
    attrs = po.Attrs(meshsize=64, boxsize=995., boxcenter=500.)  # ceil(99.5) = 100: parity kept
    new = po.regrid_attrs(attrs, 10.)
    assert np.all(new.meshsize == 100) and np.allclose(new.boxsize, 1000.)
    attrs = po.Attrs(meshsize=64, boxsize=1005., boxcenter=500.)  # ceil(100.5) = 101: odd change
    new = po.regrid_attrs(attrs, 10.)
    assert np.all(new.meshsize == 102) and np.allclose(new.boxsize, 1020.) and np.allclose(new.cellsize, 10.)


def test_normalization_closed_forms():
    rng = np.random.default_rng(1)
    attrs = po.Attrs(meshsize=10, boxsize=100., boxcenter=0.)
    # particles on nodes with CIC: each lands in exactly one cell, so sum(D R) is a lattice sum
    nodes = rng.integers(0, 10, (300, 3))
    pos_d = nodes[:150] * 10. - 50.
    pos_r = nodes[150:] * 10. - 50.
    wd, wr = rng.uniform(1, 2, 150), rng.uniform(1, 2, 150)
    D = np.zeros((10, 10, 10)); np.add.at(D, tuple(nodes[:150].T), wd)
    R = np.zeros((10, 10, 10)); np.add.at(R, tuple(nodes[150:].T), wr)
    norm = po.fkp_normalization(attrs, pos_d, wd, pos_r, wr, cellsize=None)
    assert np.isclose(norm, wd.sum() / wr.sum() * np.sum(D * R) / 1000.)
    # bispectrum: alpha^3 sum(R^3) / Vcell^2 (mesh3.py:738-783)
    norm3 = po.fkp3_normalization(attrs, pos_d, wd, pos_r, wr, cellsize=None)
    assert np.isclose(norm3, (wd.sum() / wr.sum())**3 * np.sum(R**3) / 1000.**2)
    # box normalisation (sum w)^2 / (N^3 Vcell)
    assert np.isclose(po.box_normalization(attrs, wd), wd.sum()**2 / (1000 * 1000.))
    # shot noise: sum of squared FKP weights [w_d, -alpha w_r] (tests/test_mesh2.py:75-91)
    alpha = wd.sum() / wr.sum()
    assert np.isclose(po.fkp_shotnoise(wd, wr), np.sum(np.concatenate([wd, -alpha * wr])**2))
    assert np.isclose(po.fkp_shotnoise(wd), np.sum(wd**2))


@pytest.mark.parametrize('order', [1, 2, 3, 4])
def test_read_is_adjoint_of_paint(order):
    # <paint(x, w), m> = sum_i w_i read(m, x_i): pins the gather against the brute-force-pinned
    # scatter (resamplers.py:106-131 is the transpose of 73-103)
    rng = np.random.default_rng(40 + order)
    shape = (5, 4, 6)
    xg = rng.uniform(-2, 7, (40, 3))
    w = rng.normal(size=40)
    m = rng.normal(size=shape)
    lhs = np.sum(po.paint_grid_bruteforce(shape, xg, w, order) * m)
    rhs = np.sum(w * po.read_grid(m, xg, order))
    assert np.isclose(lhs, rhs, rtol=1e-12, atol=1e-12)
    # constant mesh reads back the constant (partition of unity)
    assert np.allclose(po.read_grid(np.full(shape, 2.5), xg, order), 2.5, atol=1e-13)


def test_read_compensated_plane_wave():
    # the compensation divides each Fourier mode by prod sinc^p: a plane wave read back with
    # compensate=True at the grid nodes equals the wave times prod sinc^-p at its wavenumber
    attrs = po.Attrs(8, 8., 4.)
    x = np.arange(8.)
    wave = np.cos(2. * np.pi * 1. * x / 8.)[:, None, None] * np.ones((8, 8, 8))
    nodes = np.array([[i, 3., 5.] for i in range(8)]) + attrs.boxcenter - attrs.boxsize / 2.
    for order, name in [(2, 'cic'), (3, 'tsc')]:
        got = po.read(attrs, wave, nodes, resampler=name, compensate=True)
        k = 2. * np.pi / 8.
        expected = np.cos(k * np.arange(8.)) * np.sinc(k / (2. * np.pi))**(-order)
        # at nodes the CIC kernel is exact; TSC spreads 3/4, 1/8, 1/8 -> multiply by its transfer
        transfer = 1. if order == 2 else 0.75 + 0.25 * np.cos(k)
        assert np.allclose(got, expected * transfer, atol=1e-12)


# ---------------------------------------------------------------- correlation function

def test_correlation_c2r_is_circular_correlation():
    """c2r(conj(A1) A2) is sum_x m1(x) m2(x + s), the definition of the estimator's xi
    (mesh2.py:472-476); pins the c2r normalisation and the conjugation side."""
    rng = np.random.default_rng(3)
    shape = (6, 4, 8)
    m1, m2 = rng.normal(size=shape), rng.normal(size=shape)
    got = po.c2r(np.conj(po.r2c(m1)) * po.r2c(m2), shape)
    assert np.allclose(got, po.circular_correlation_bruteforce(m1, m2), rtol=0, atol=1e-12)
    auto = po.c2r(np.abs(po.r2c(m1))**2, shape)
    assert np.allclose(auto, po.circular_correlation_bruteforce(m1), rtol=0, atol=1e-12)
    assert np.isclose(auto[0, 0, 0], np.sum(m1**2))  # zero lag


def _lags(attrs):
    """Lag vectors (i, j, k) -> separation, from the index definition of fftfreq."""
    out = {}
    for idx in np.ndindex(*attrs.meshsize):
        out[idx] = np.array([(i if i < n / 2 else i - n) * c for i, n, c in zip(idx, attrs.meshsize, attrs.cellsize)])
    return out


def test_correlation_nmodes_and_edges_bruteforce():
    """Separation bins: counts and mean |s| against an explicit loop over lags."""
    attrs = po.Attrs((6, 8, 4), (60., 80., 40.), 0.)
    bin = po.BinCorrelationPoles(attrs, edges={'step': 7.}, ells=(0,))
    edges = np.append(bin.edges[:, 0], bin.edges[-1, 1])
    assert np.isclose(edges[-1] - edges[-2], 7.) and edges[0] == 0. and edges[-1] < 10. * 4 / 2 + 7.
    count, ssum = np.zeros(len(bin.nmodes), 'i8'), np.zeros(len(bin.nmodes))
    for idx, s in _lags(attrs).items():
        sn = np.linalg.norm(s)
        b = np.searchsorted(edges, sn, side='right') - 1
        if 0 <= b < len(count):
            count[b] += 1
            ssum[b] += sn
    assert np.array_equal(count, bin.nmodes)
    ok = count > 0
    assert np.allclose(ssum[ok] / count[ok], bin.xavg[ok], rtol=1e-14)


@pytest.mark.parametrize('los', ['x', 'z', (0.6, 0., 0.8)])
def test_correlation_global_los_bruteforce(los):
    """Global line of sight: (2l+1) <xi(s) L_l(mu_s)> over separation shells, xi from explicit
    shifts and L_l from scipy, against the oracle (mesh2.py:465-504)."""
    rng = np.random.default_rng(5)
    attrs = po.Attrs((6, 6, 6), (60., 60., 60.), (5., -3., 2.))
    mesh = rng.normal(size=attrs.meshsize)
    ells = (0, 2, 4)
    bin = po.BinCorrelationPoles(attrs, edges={'step': 9.}, ells=ells)
    got = po.correlation_poles(attrs, bin, mesh, los=los)
    xi = po.circular_correlation_bruteforce(mesh)
    vlos = po.los_vector(los)
    vlos = vlos / np.linalg.norm(vlos)
    edges = np.append(bin.edges[:, 0], bin.edges[-1, 1])
    vcell = np.prod(attrs.cellsize)
    for ill, ell in enumerate(ells):
        acc, cnt = np.zeros(len(bin.nmodes)), np.zeros(len(bin.nmodes))
        for idx, s in _lags(attrs).items():
            sn = np.linalg.norm(s)
            b = np.searchsorted(edges, sn, side='right') - 1
            if not 0 <= b < len(acc):
                continue
            cnt[b] += 1
            if ell and sn == 0.:
                continue  # remove_zero for l > 0
            mu = s @ vlos / sn if sn > 0 else 0.
            acc[b] += xi[idx] * special.eval_legendre(ell, mu)
        want = (2 * ell + 1) * acc / cnt / vcell
        assert np.allclose(got['num_raw'][ill], want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    assert np.allclose(got['norm'], np.prod(attrs.meshsize) / vcell)


@pytest.mark.parametrize('boxcenter', [(70., -20., 40.), 0.])
def test_correlation_local_los_bruteforce(boxcenter):
    """Local (firstpoint) line of sight: the Y_lm route of mesh2.py:445-460 against the
    explicit double sum over cells and lags with P_l(xhat . shat) (origin a node or not)."""
    rng = np.random.default_rng(9)
    attrs = po.Attrs((6, 6, 6), (60., 60., 60.), boxcenter)
    mesh = rng.normal(size=attrs.meshsize)
    ells = (0, 2, 4)
    bin = po.BinCorrelationPoles(attrs, edges={'step': 9.}, ells=ells)
    got = po.correlation_poles(attrs, bin, mesh, los='firstpoint')
    vcell = np.prod(attrs.cellsize)
    assert np.allclose(got['num_raw'][0], bin(po.circular_correlation_bruteforce(mesh)) / vcell, rtol=1e-12)
    for ill, ell in enumerate(ells[1:], start=1):
        want = 4. * np.pi * bin(po.local_correlation_bruteforce(attrs, mesh, ell), remove_zero=True) / vcell
        assert np.allclose(got['num_raw'][ill], want, rtol=1e-10, atol=1e-10 * np.abs(want).max())


# ---------------------------------------------------------------- bispectrum (Scoccimarro basis)

@pytest.mark.parametrize('meshsize,boxsize,step', [((8, 8, 8), 80., 0.06), ((8, 6, 8), (80., 60., 80.), 0.05)])
def test_bispectrum_counts_bruteforce(meshsize, boxsize, step):
    """nmodes = N^6 sum_x prod c2r(shell indicators) equals the count of closed full-grid
    triangles k1 + k2 + k3 = 0 (mod N) with k_i in shell b_i (mesh3.py:229-236)."""
    attrs = po.Attrs(meshsize, boxsize, 0.)
    bin = po.BinBispectrumPoles(attrs, edges={'step': step}, ells=0)
    counts = po.bispectrum_bruteforce(attrs, bin, [None, None, None])
    assert len(bin.iedges) > 3
    assert np.allclose(bin.nmodes, counts.real, rtol=1e-12, atol=1e-9)
    assert np.sum(bin.nmodes > 0.5) >= 3


def test_bispectrum_mask_and_triangle_order():
    """Default mask (mesh3.py:31-35): sorted sides, triangle inequality on the bin centres and
    the sum of upper edges within 2 k_nyq; triangles in lexicographic order; mask_edges=''
    keeps every combination; the diagonal basis keeps (b, b, b)."""
    attrs = po.Attrs(16, 160., 0.)
    bin = po.BinBispectrumPoles(attrs, edges={'step': 0.05})
    mid = bin.edges.mean(axis=-1)
    nyq = attrs.kfun.min() * 8
    assert np.all(mid[:, 0] <= mid[:, 1]) and np.all(mid[:, 1] <= mid[:, 2])
    assert np.all(mid[:, 2] <= mid[:, 0] + mid[:, 1]) and np.all(mid[:, 2] >= np.abs(mid[:, 0] - mid[:, 1]))
    assert np.all(bin.edges[:, :, 1].sum(axis=1) <= 2 * nyq)
    assert [tuple(t) for t in bin.iedges] == sorted(tuple(t) for t in bin.iedges)
    nb = len(bin.xavg1d[0])
    full = po.BinBispectrumPoles(attrs, edges={'step': 0.05}, mask_edges='')
    assert len(full.iedges) == nb**3 and len(bin.iedges) < nb**3
    diag = po.BinBispectrumPoles(attrs, edges={'step': 0.05}, mask_edges='', diagonal=True)
    assert np.array_equal(diag.iedges, np.repeat(np.arange(nb)[:, None], 3, axis=1))


@pytest.mark.parametrize('los,cross', [('z', False), ((0.6, 0., 0.8), False), ('x', True)])
def test_bispectrum_poles_bruteforce(los, cross):
    """compute_mesh3_spectrum's FFT route equals the explicit sum over closed triangles of
    F1 F2 F3 P_ell(mu3), divided by the triangle counts, for ell = 0, 2, 4 (odd ell: 0)."""
    attrs = po.Attrs((8, 8, 8), 80., 0.)
    rng = np.random.default_rng(4)
    meshes = [rng.normal(size=(8, 8, 8)) + 0.3 * rng.normal(size=(8, 8, 8))**2 for _ in range(3)]
    if not cross:
        meshes = [meshes[0]] * 3
    bin = po.BinBispectrumPoles(attrs, edges={'step': 0.07}, ells=(0, 1, 2, 4))
    res = po.bispectrum_poles(attrs, bin, *meshes, los=los)
    counts = po.bispectrum_bruteforce(attrs, bin, [None, None, None]).real
    for i, ell in enumerate(bin.ells):
        if ell % 2:
            assert np.all(res['num_raw'][i] == 0.)
            continue
        b = po.bispectrum_bruteforce(attrs, bin, meshes, ell=ell, los=los)
        want = (2 * ell + 1) * b.real / counts
        scale = np.abs(want).max()
        assert np.allclose(res['num_raw'][i], want, rtol=1e-9, atol=1e-10 * scale), ell
    assert np.isclose(res['norm'], 512 / 1000.**2)


# ---------------------------------------------------------------- window matrix (global line of sight)

def test_window_periodic_matches_common_edges_route():
    """mesh2.py:1091-1095: with input edges equal to the output edges, the reference bins
    P_ellin(mu) N^3 over the whole mesh once and puts the result on the diagonal; the per-shell
    route of the oracle must give that matrix (zeros off the diagonal, the DC mode dropped)."""
    a = po.Attrs(16, 400., 0.)
    bin = po.BinPoles(a, edges={'step': 2. * np.pi / 400.}, ells=(0, 2, 4))
    ellsin = (0, 2, 4)
    W = po.spectrum_window(a, bin, bin.edges, ellsin=ellsin)
    nb = len(bin.nmodes)
    kx, ky, kz = a.kcoords('wavenumber')
    k2 = kx[:, None, None]**2 + ky[None, :, None]**2 + kz[None, None, :]**2
    kn = np.sqrt(k2)
    mu = np.broadcast_to(kz[None, None, :], kn.shape) / np.where(kn == 0, 1., kn)
    n3 = np.prod(a.meshsize)
    for iin, ellin in enumerate(ellsin):
        for il, ell in enumerate(bin.ells):
            block = W[il * nb:(il + 1) * nb, iin * nb:(iin + 1) * nb]
            diag = (2 * ell + 1) * bin(po.legendre(ellin, mu) * po.legendre(ell, mu) * n3, remove_zero=True).real / n3
            np.testing.assert_allclose(np.diag(block), diag, rtol=1e-12, atol=1e-14)
            np.testing.assert_array_equal(block - np.diag(np.diag(block)), 0.)
    # monopole to monopole: 1 in every bin, minus the DC mode in the first
    block = W[:nb, :nb]
    np.testing.assert_allclose(np.diag(block), 1. - (np.arange(nb) == 0) / bin.nmodes[0], rtol=1e-13)


def test_window_selection_vs_bruteforce_convolution():
    """r2c(Q c2r(A)) with Q = c2r(A1 conj A2)/N^3 against the convolution written out:
    Q(s) from explicit circular shifts, A on the full grid from fftfreq wavenumbers (even ellin:
    A is real and even, hence hermitian), Qhat by direct DFT sums."""
    rng = np.random.default_rng(3)
    n = 6
    a = po.Attrs(n, 60., 0.)
    bin = po.BinPoles(a, edges={'step': 2. * np.pi / 60.}, ells=(0, 2, 4))
    m1, m2 = rng.normal(size=(2, n, n, n))
    edgesin = np.array([[0., 0.15], [0.15, 0.25], [0.25, 0.4]])
    ellsin = (0, 2)
    W = po.spectrum_window(a, bin, edgesin, ellsin=ellsin, meshes=[m1, m2], los='y')
    n3 = n**3
    Q = po.circular_correlation_bruteforce(m2, m1) / n3
    kk = [np.fft.fftfreq(n) * a.kfun[i] * n for i in range(3)]
    kn = np.sqrt(kk[0][:, None, None]**2 + kk[1][None, :, None]**2 + kk[2][None, None, :]**2)
    mu = np.broadcast_to(kk[1][None, :, None], kn.shape) / np.where(kn == 0, 1., kn)
    nb = len(bin.nmodes)
    kxh, kyh, kzh = a.kcoords('wavenumber')
    knh = np.sqrt(kxh[:, None, None]**2 + kyh[None, :, None]**2 + kzh[None, None, :]**2)
    muh = np.broadcast_to(kyh[None, :, None], knh.shape) / np.where(knh == 0, 1., knh)
    for iin, ellin in enumerate(ellsin):
        for ie, (e0, e1) in enumerate(edgesin):
            A_full = ((kn >= e0) & (kn < e1)) * po.legendre(ellin, mu) * n3
            B = po.convolve_bruteforce(Q, A_full)[:, :, :n // 2 + 1]
            for il, ell in enumerate(bin.ells):
                want = (2 * ell + 1) * bin(B * po.legendre(ell, muh), remove_zero=True).real / n3
                got = W[il * nb:(il + 1) * nb, iin * len(edgesin) + ie]
                np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-12 * np.max(np.abs(W)))


def test_window_constant_selection_is_periodic():
    """A constant selection mesh gives Q = 1 (mesh2.py:1022-1029), so the window must be the
    periodic one; with norm given, the matrix scales by N^3/V_cell / norm (mesh2.py:973-979)."""
    a = po.Attrs(12, 300., 0.)
    bin = po.BinPoles(a, edges={'step': 0.03}, ells=(0, 2))
    edgesin = np.arange(0., 0.2, 0.025)
    W0 = po.spectrum_window(a, bin, edgesin, ellsin=(0, 2, 4), los=[0.6, 0., 0.8])
    W1 = po.spectrum_window(a, bin, edgesin, ellsin=(0, 2, 4), meshes=[np.ones(tuple(a.meshsize))], los=[0.6, 0., 0.8])
    np.testing.assert_allclose(W1, W0, rtol=1e-9, atol=1e-12)
    norm = 7.5
    W2 = po.spectrum_window(a, bin, edgesin, ellsin=(0, 2, 4), los=[0.6, 0., 0.8], norm=norm)
    _norm = np.prod(a.meshsize) / np.prod(a.cellsize)
    np.testing.assert_allclose(W2, W0 * _norm / norm, rtol=1e-13)
    assert np.abs(np.diag(W0[:len(bin.nmodes), :])).max() > 0.5


@pytest.mark.parametrize('ell', [0, 1, 2, 3, 4, 5])
def test_tophat_bessel_vs_quadrature(ell):
    """The closed forms and low-x series of utils.py:403-415 are x^-3 int_0^x t^2 j_ell(t) dt
    (numerical quadrature of scipy's spherical Bessel functions), on both sides of the x = 0.1
    switch.  Just above the switch the reference's closed forms cancel catastrophically for
    ell >= 3 (terms ~x^-(ell+1) summing to ~x^ell), which the oracle reproduces: below x = 3
    the check is loose (1e-6) for ell >= 3, and at x = 0.1001 it is skipped for ell >= 4."""
    from scipy import integrate
    cases = [(x, 1e-9 if (x < 0.1 or x >= 3. or ell <= 2) else 1e-6) for x in (0.03, 0.0999, 0.7, 3., 11.3, 40., 95.)]
    if ell <= 3:
        cases.append((0.1001, 1e-9 if ell <= 2 else 1e-6))
    for x, rtol in cases:
        want = integrate.quad(lambda t: t**2 * special.spherical_jn(ell, t), 0., x, limit=400, epsabs=1e-14, epsrel=1e-12)[0] / x**3
        got = po.tophat_bessel(ell, x)
        assert abs(got - want) <= rtol * max(abs(want), 1e-6 / x**2), (ell, x, got, want)


def test_window_infinite_close_to_mesh_window():
    """The configuration-space (infinite-resolution) theory of a well-resolved shell, sampled on
    the mesh and binned, must nearly reproduce the mesh-defined window on the shell's own bins:
    1 for ell = ellin = 0 and ell = ellin = 2 (a sign error in (-1)^(ell/2), a missing 4 pi or
    (2 pi)^3 fails this), and little leakage from ellin = 0 into ell = 2."""
    a = po.Attrs(48, 1000., 0.)
    kf = 2. * np.pi / 1000.
    bin = po.BinPoles(a, edges=np.arange(0., 0.13, 4 * kf), ells=(0, 2))
    edgesin = np.array([[12 * kf, 16 * kf]])
    nb = len(bin.nmodes)
    Wi = po.spectrum_window_infinite(a, bin, edgesin, ellsin=(0, 2), los='z')
    Wm = po.spectrum_window(a, bin, edgesin, ellsin=(0, 2), los='z')
    b = 3  # the output bin [12 kf, 16 kf)
    assert abs(Wm[b, 0] - 1.) < 1e-12 and abs(Wm[nb + b, 1] - 1.) < 1e-2  # lattice average of 5 P_2^2
    assert abs(Wi[b, 0] - 1.) < 0.05, Wi[b, 0]
    assert abs(Wi[nb + b, 1] - 1.) < 0.08, Wi[nb + b, 1]
    assert abs(Wi[nb + b, 0]) < 0.05


@pytest.mark.parametrize('ell', [0, 1, 2, 3, 4, 5])
def test_spherical_jn_vs_scipy(ell):
    """utils.py:418-450 closed forms and series against scipy.special.spherical_jn.  Just above
    the x = 0.1 switch the reference's closed forms cancel for ell >= 3 (terms ~x^-(ell+1)
    summing to ~x^ell); there only the well-conditioned points are held to 1e-9."""
    x = np.concatenate([np.linspace(0., 0.0999, 7), [0.5, 2., 9.7, 33., 150.]])
    want = special.spherical_jn(ell, x)
    got = po.spherical_jn(ell, x)
    rtol = np.where((x < 0.1) | (x >= 2.) | (ell <= 3), 1e-9, 1e-5)
    assert np.all(np.abs(got - want) <= rtol * np.abs(want) + 1e-15), (ell, np.max(np.abs(got - want)))


def test_window_smooth_close_to_infinite():
    """The smooth window (Hankel transform of the binned selection multipoles,
    mesh2.py:1045-1076) approximates the infinite-resolution one: on a periodic box the
    monopole-to-monopole block must agree within a few per cent on the shell's own bins, and
    for a constant selection mesh Q = 1 the smooth window equals the periodic one exactly."""
    a = po.Attrs(32, 800., 0.)
    kf = 2. * np.pi / 800.
    bin = po.BinPoles(a, edges=np.arange(0., 0.2, 3 * kf), ells=(0, 2))
    edgesin = np.array([[6 * kf, 9 * kf], [9 * kf, 12 * kf]])
    Ws = po.spectrum_window_smooth(a, bin, edgesin, ellsin=(0, 2), los='z')
    Wi = po.spectrum_window_infinite(a, bin, edgesin, ellsin=(0, 2), los='z')
    nb = len(bin.nmodes)
    for col, b in ((0, 2), (1, 3)):
        assert abs(Ws[b, col] - Wi[b, col]) < 0.08 * abs(Wi[b, col]), (col, Ws[b, col], Wi[b, col])
    assert abs(Ws[nb + 2, 2] - Wi[nb + 2, 2]) < 0.15 * abs(Wi[nb + 2, 2])
    W1 = po.spectrum_window_smooth(a, bin, edgesin, ellsin=(0, 2), meshes=[np.ones((32,) * 3)], los='z')
    np.testing.assert_allclose(W1, Ws, rtol=1e-9, atol=1e-12 * np.abs(Ws).max())


@pytest.mark.parametrize('ell,m,ellin', [(0, 0, 0), (2, 1, 0), (0, 0, 2), (2, -2, 2), (4, 3, 2)])
def test_window_local_selection_vs_bruteforce(ell, m, ellin):
    """Q_lm(s) of the local-line-of-sight window (sum over m1 of Y_lin,m1(shat) times the
    r2c/c2r cross-correlation of rho Y_lm Y_lin,m1 with rho) against the explicit double sum
    over cells and lags with P_lin(xhat . shat): pins the Y normalisation, the conjugation and
    the direction of the lag, independently of the Y sign conventions."""
    rng = np.random.default_rng(5)
    a = po.Attrs(6, 60., 20.)
    rho = rng.uniform(0.5, 1.5, (6, 6, 6))
    got = po.window_local_selection(a, rho, ell, m, ellin)
    want = po.window_local_selection_bruteforce(a, rho, ell, m, ellin)
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-11 * np.abs(want).max())


@pytest.mark.parametrize('ell,m,ellin', [(0, 0, 0), (2, -1, 0), (0, 0, 2), (2, 2, 2)])
def test_window_local_terms_vs_bruteforce(ell, m, ellin):
    """A_lm(s) = sum_m1 c2r(shell Y_lin,m1(khat)) Q_lm,m1(s) of the local window with the theory
    on the mesh, against the addition-theorem double sum over cells and modes with
    P_lin(xhat . khat): pins the pairing of the two Y_lin,m1, the conjugation, the lag
    direction and the closed right edge of the shell."""
    rng = np.random.default_rng(8)
    a = po.Attrs(6, 60., 20.)
    rho = rng.uniform(0.5, 1.5, (6, 6, 6))
    kf = 2. * np.pi / 60.
    e0, e1 = kf, 2. * kf  # closed: |k| = 2 kf is in, as in mesh2.py:1276
    got = po.window_local_terms(a, rho, ell, m, ellin, e0, e1)
    want = po.window_local_terms_bruteforce(a, rho, ell, m, ellin, e0, e1)
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-11 * np.abs(want).max())


@pytest.mark.parametrize('ell,ellin', [(0, 0), (2, 0), (0, 2), (2, 2), (4, 2)])
def test_window_local_smooth_terms_vs_bruteforce(ell, ellin):
    """sum_{m, m1} Q_lm,m1(s) Y_lm(shat) Y_lin,m1(shat) of the local smooth window against the
    addition-theorem sum over cells with P_l(xhat.shat) P_lin(xhat.shat)."""
    rng = np.random.default_rng(9)
    a = po.Attrs(6, 60., 20.)
    rho = rng.uniform(0.5, 1.5, (6, 6, 6))
    got = po.window_local_smooth_terms(a, rho, ell, ellin)
    want = po.window_local_smooth_terms_bruteforce(a, rho, ell, ellin)
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-11 * np.abs(want).max())


@pytest.mark.parametrize('los', ['x', 'firstpoint'])
def test_spectrum_mean_equals_window_times_theory(los):
    """For a theory tabulated on the input shells, the mean spectrum (mesh2.py:1485-1680) is the
    window (mesh2.py:921-1305, pinned above) applied to it, in every output bin that does not hold
    the k = 0 mode (the window drops it, the mean keeps it): two restatements of different code
    paths of the reference checked against each other."""
    rng = np.random.default_rng(12)
    a = po.Attrs(8, 160., (300., 100., 200.))
    kf = 2. * np.pi / 160.
    bin = po.BinPoles(a, edges=np.arange(0., 4.5 * kf, kf), ells=(0, 2))
    edgesin = np.column_stack([np.arange(0., 4.) * kf, np.arange(1., 5.) * kf])
    edges1d = np.append(edgesin[:, 0], edgesin[-1, 1])
    poles = {0: rng.uniform(1., 2., 4), 2: rng.uniform(-0.5, 0.5, 4)}
    rho = rng.uniform(0.5, 1.5, (8, 8, 8))
    num = po.spectrum_mean(a, bin, edges1d, poles, edges=True, mesh=rho, los=los)
    norm = np.prod(a.meshsize) / np.prod(a.cellsize)
    if los == 'firstpoint':
        # closed right edges in the window: move them down by one ulp to match digitize's half-open shells
        wedges = np.column_stack([edgesin[:, 0], np.nextafter(edgesin[:, 1], 0.)])
        W = po.spectrum_window_local(a, bin, wedges, ellsin=(0, 2), mesh=rho)
    else:
        W = po.spectrum_window(a, bin, edgesin, ellsin=(0, 2), meshes=[rho], los=los)
    theta = np.concatenate([poles[0], poles[2]])
    got = W @ theta  # num_raw / norm with the window's default norm
    nb = len(bin.nmodes)
    for il in range(2):
        want = num[il] / norm
        np.testing.assert_allclose(got[il * nb + 1:(il + 1) * nb], want[1:], rtol=1e-9, atol=1e-12 * np.abs(want).max())


def test_real_gaunt_vs_reference_table():
    """The oracle's real Gaunt coefficients against entries of the reference's table
    (jaxpower/utils.py:600-), whose Y_lm differ from the oracle's by (-1)^|m| per factor, and the
    convention-free invariant sum_{m1 m2 m3} G^2 = (2l1+1)(2l2+1)(2l3+1)/(4 pi) (l1 l2 l3; 0 0 0)^2."""
    from scipy.special import factorial
    table = {((0, 0), (2, -2), (2, -2)): 0.28209479177387814, ((2, -2), (2, -1), (2, 1)): 0.15607834722743974,
             ((2, -1), (2, -1), (2, 0)): 0.0901118757864343, ((2, 0), (2, 2), (2, 2)): -0.18022375157286857,
             ((2, 1), (2, 1), (2, 2)): 0.15607834722743988, ((4, -4), (2, -2), (2, 2)): 0.23841361350444812,
             ((4, -3), (2, -2), (2, 1)): 0.16858388283618375}
    for key, value in table.items():
        sign = (-1)**sum(abs(m) for _, m in key)
        got = po.real_gaunt(*[v for lm in key for v in lm])
        assert abs(sign * got - value) < 1e-12, (key, got, value)

    def three_j_000(a, b, c):
        g = (a + b + c) // 2
        if (a + b + c) % 2 or c > a + b or c < abs(a - b):
            return 0.
        return (-1)**g * np.sqrt(factorial(2 * g - 2 * a) * factorial(2 * g - 2 * b) * factorial(2 * g - 2 * c) / factorial(2 * g + 1)) \
            * factorial(g) / (factorial(g - a) * factorial(g - b) * factorial(g - c))

    for l1, l2, l3 in [(0, 2, 2), (2, 2, 2), (4, 2, 2), (2, 2, 0), (4, 4, 2)]:
        tot = sum(po.real_gaunt(l1, a, l2, b, l3, c)**2 for a in range(-l1, l1 + 1) for b in range(-l2, l2 + 1) for c in range(-l3, l3 + 1))
        want = (2 * l1 + 1) * (2 * l2 + 1) * (2 * l3 + 1) / (4. * np.pi) * three_j_000(l1, l2, l3)**2
        assert abs(tot - want) < 1e-12, (l1, l2, l3, tot, want)


@pytest.mark.parametrize('ell,m,ellin,wa', [(2, 1, 0, 1), (0, 0, 2, 1), (2, -2, 2, 2)])
def test_window_local_selection_wide_angle_vs_bruteforce(ell, m, ellin, wa):
    """The wide-angle factors |x|^-wa and |s|^wa of the local window (mesh2.py:1244) against the
    brute-force double sum (positions kept off the origin, as a survey's are)."""
    rng = np.random.default_rng(6)
    a = po.Attrs(6, 60., 70.)
    rho = rng.uniform(0.5, 1.5, (6, 6, 6))
    got = po.window_local_selection(a, rho, ell, m, ellin, wa=wa)
    want = po.window_local_selection_bruteforce(a, rho, ell, m, ellin, wa=wa)
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-11 * np.abs(want).max())


def test_bispectrum_shotnoise_monopole_is_power_spectrum_sum():
    """With unit weights paint(w^2) = paint(w), so the l = 0 bispectrum shot noise
    (mesh3.py:830-832) is the sum over the three sides of the (k = 0 removed) power-spectrum
    monopole at the side's shell, minus 2 N: checked against the pinned spectrum oracle."""
    a = po.Attrs(16, 400., 0.)
    bin = po.BinBispectrumPoles(a, edges={'step': 0.04}, ells=(0, 2))
    pos = np.random.default_rng(1).uniform(-200., 200., (3000, 3))
    sn = po.bispectrum_shotnoise(a, bin, pos, resampler='tsc', interlacing=2, compensate=True, los='x')
    mesh = po.paint(a, pos, None, resampler='tsc', interlacing=2, compensate=True, out='complex')
    mesh[0, 0, 0] = 0.
    pk = po.spectrum_poles(a, po.BinPoles(a, edges={'step': 0.04}, ells=(0,)), mesh, los='x')['num_raw'][0]
    want = sum(pk[bin.iedges[:, i]] for i in range(3)) - 2. * len(pos)
    np.testing.assert_allclose(sn[0], want, rtol=1e-12)


@pytest.mark.parametrize('ell', [0, 1, 2, 3, 4])
def test_bispectrum_local_mesh_vs_direct_dft(ell):
    """The third mesh of the local-line-of-sight bispectrum, sum_m r2c(rho Y_lm(xhat)) Y_lm(khat)
    (mesh3.py:563-572), against a direct DFT weighted by (2l+1)/(4 pi) P_l(xhat.khat), with the
    reference's Y values at zero vectors (the origin is a node here, and k = 0 is included)."""
    rng = np.random.default_rng(4)
    a = po.Attrs(6, 60., 20.)
    rho = rng.uniform(0.5, 1.5, (6, 6, 6))
    got = po.bispectrum_local_mesh(a, rho, ell)
    want = po.bispectrum_local_mesh_bruteforce(a, rho, ell)
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-10 * np.abs(want).max())
    if ell == 4:
        # k = 0: sum_m Y_4m(0) sum_x rho Y_4m(xhat) = Y_40(0) sum_x rho Y_40(xhat), nonzero
        xs = a.rcoords()
        X = np.stack(np.meshgrid(*xs, indexing='ij'), axis=-1)
        c0 = po.ylm_at_zero(4, 0) * np.sum(rho * po.real_ylm(4, 0, X[..., 0], X[..., 1], X[..., 2]))
        assert abs(c0) > 1e-3 and np.isclose(got[0, 0, 0].real, c0, rtol=1e-12)


def test_bispectrum_shotnoise_local_monopole_and_ell2():
    """Local line of sight of the bispectrum shot noise (mesh3.py:823-834): l = 0 does not
    depend on the line of sight; for l = 2 the first line of eq. 58 with H_2 built by
    bispectrum_local_mesh equals the same line with P_2(xhat.khat) applied by a direct DFT."""
    rng = np.random.default_rng(8)
    a = po.Attrs(8, 80., 40.)
    pos = rng.uniform(0., 80., (300, 3))
    w = rng.uniform(0.5, 1.5, 300)
    b = po.BinBispectrumPoles(a, edges={'step': 0.2}, ells=(0, 2))
    loc = po.bispectrum_shotnoise(a, b, pos, w, los='local')
    glo = po.bispectrum_shotnoise(a, b, pos, w, los='z')
    np.testing.assert_allclose(loc[0], glo[0], rtol=1e-12)
    cw = po.paint(a, pos, w, resampler='cic', compensate=True, out='complex')
    cw[0, 0, 0] = 0.
    rho = po.c2r(cw, a.meshsize)
    h = 4. * np.pi / 5. * po.bispectrum_local_mesh(a, rho, 2)
    want = 4. * np.pi / 5. * po.bispectrum_local_mesh_bruteforce(a, rho, 2)
    np.testing.assert_allclose(h, want, rtol=1e-9, atol=1e-10 * np.abs(want).max())
    assert np.all(np.isfinite(loc[1])) and np.abs(loc[1] - glo[1]).max() > 0.


# ----------------------------------------------------------------------------- Sugiyama basis

def _sympy_ylm(ell, m, real, reduced, conj):
    """get_Ylm's polynomial (utils.py:223-334) redone here step by step: normalisation of
    utils.py:227-238, associated Legendre of cos(theta), cos/sin(m phi) (real) or
    cos(m phi) + i sin(m phi), conjugated if asked (complex), expanded, phi and theta replaced by
    Cartesian ratios, x^2 + y^2 + z^2 -> r^2, x/r -> xhat.  The reference is not imported."""
    sp = pytest.importorskip('sympy')
    amp = 1. if reduced else np.sqrt((2 * ell + 1) / (4 * np.pi))
    if m != 0:
        amp *= np.sqrt(math.factorial(ell - abs(m)) / math.factorial(ell + abs(m)))
    if real and m != 0:
        amp *= np.sqrt(2.) * (-1)**m
    if not real and m < 0:
        amp *= (-1)**m
    x, y, z, r = sp.symbols('x y z r', real=True, positive=True)
    xh, yh, zh = sp.symbols('xhat yhat zhat', real=True, positive=True)
    phi, theta = sp.symbols('phi theta', real=True)
    defs = [(sp.sin(phi), y / sp.sqrt(x**2 + y**2)), (sp.cos(phi), x / sp.sqrt(x**2 + y**2)),
            (sp.cos(theta), z / sp.sqrt(x**2 + y**2 + z**2))]
    expr = sp.assoc_legendre(ell, abs(m), sp.cos(theta))
    if real:
        eimphi = sp.sin(abs(m) * phi) if m < 0 else (sp.cos(m * phi) if m > 0 else 1)
    else:
        eimphi = sp.cos(m * phi) + sp.sin(m * phi) * sp.I
        if conj:
            eimphi = sp.conjugate(eimphi)
    expr *= sp.expand_trig(eimphi)
    expr = sp.together(expr.subs(defs)).subs(x**2 + y**2 + z**2, r**2)
    expr = amp * expr.expand().subs([(x / r, xh), (y / r, yh), (z / r, zh)])
    return sp.lambdify((xh, yh, zh), expr, modules=['numpy']), (xh, yh, zh), expr


def test_reference_real_and_complex_ylm_match_sympy_construction():
    """ref_real_ylm (sign included) and complex_ylm_conj (get_Ylm(reduced=True, real=False,
    conj=True)) against the sympy construction, on unit vectors and at the zero vector."""
    rng = np.random.default_rng(11)
    u = rng.normal(size=(40, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    for ell in range(5):
        for m in range(-ell, ell + 1):
            f, syms, expr = _sympy_ylm(ell, m, real=True, reduced=False, conj=False)
            want = np.broadcast_to(np.asarray(f(*u.T), dtype=complex), (len(u),)).real
            assert np.allclose(po.ref_real_ylm(ell, m, *u.T), want, atol=1e-12), (ell, m)
            assert np.isclose(po.ref_real_ylm(ell, m, 0., 0., 0.), float(expr.subs({s_: 0 for s_ in syms})), atol=1e-15)
            f, syms, expr = _sympy_ylm(ell, m, real=False, reduced=True, conj=True)
            want = np.broadcast_to(np.asarray(f(*u.T), dtype=complex), (len(u),))
            assert np.allclose(po.complex_ylm_conj(ell, m, *u.T), want, atol=1e-12), (ell, m)
            at0 = complex(expr.subs({s_: 0 for s_ in syms}))
            assert np.isclose(po.complex_ylm_conj(ell, m, 0., 0., 0.), at0, atol=1e-15), (ell, m, at0)


def test_wigner_3j_and_triposh_rows_against_sympy():
    sp_wigner = pytest.importorskip('sympy.physics.wigner')
    for j1, j2, j3 in itertools.product(range(4), repeat=3):
        for m1 in range(-j1, j1 + 1):
            for m2 in range(-j2, j2 + 1):
                m3 = -m1 - m2
                want = float(sp_wigner.wigner_3j(j1, j2, j3, m1, m2, m3)) if abs(m3) <= j3 else 0.
                assert np.isclose(po.wigner_3j(j1, j2, j3, m1, m2, m3), want, atol=1e-14), (j1, j2, j3, m1, m2)
    # rows: every (m1, m2, M) with a nonzero coefficient; M = 0 only for los 'z'
    rows = po.triposh_rows(2, 2, 2, los_z=False)
    assert all(m1 + m2 + m3 == 0 for m1, m2, m3, _ in rows) and len(rows) == 19
    assert [r[:3] for r in po.triposh_rows(2, 0, 2, los_z=True)] == [(0, 0, 0)]
    assert np.isclose(po.triposh_rows(0, 0, 0)[0][3], 1.)
    # sum over all m of W(l1 l2 L; m1 m2 M)^2 = 1, so sum of g^2 over the rows = ((2l1+1)(2l2+1)(2L+1) W(000))^2
    g = np.array([r[3] for r in rows])
    assert np.isclose(np.sum(g**2), (125. * float(sp_wigner.wigner_3j(2, 2, 2, 0, 0, 0)))**2)


@pytest.mark.parametrize('ell,los', [((0, 0, 0), 'z'), ((2, 0, 2), 'z'), ((2, 0, 2), 'x'), ((1, 1, 0), 'x'), ((1, 1, 2), 'local'),
                                     ((2, 0, 2), 'local')])
def test_sugiyama_bispectrum_vs_mode_pair_sums(ell, los):
    """The oracle's FFT reduction against direct sums over pairs of full-grid modes."""
    attrs = po.Attrs((6, 8, 6), (60., 80., 60.), (10., -20., 35.))
    rng = np.random.default_rng(17)
    meshes = [rng.normal(size=(6, 8, 6)) for _ in range(3)]
    bin = po.BinBispectrumSugiyama(attrs, edges=np.array([0., 0.06, 0.12, 0.2]), ells=[ell], mask_edges='')
    got = po.bispectrum_sugiyama(attrs, bin, *meshes, los=los)['num_raw'][0]
    pair = po.bispectrum_sugiyama_bruteforce(attrs, bin, meshes, ell=ell, los=los) / bin.nmodes
    want = pair.real if (ell[0] + ell[1]) % 2 == 0 else pair.imag
    assert np.allclose(got, want, rtol=1e-10, atol=1e-12 * np.abs(want).max())
    assert np.abs(want).max() > 0


def test_sugiyama_counts_and_mask():
    """Full-grid shell counts equal the hermitian-weighted counts below the Nyquist planes, and
    the default mask keeps mid1 <= mid2 with the sum of upper edges within the Nyquist."""
    attrs = po.Attrs(16, 160., 0.)
    edges = np.arange(0.01, 0.3, 0.02)
    bin = po.BinBispectrumSugiyama(attrs, edges=edges, ells=[(0, 0, 0)])
    herm = po.BinPoles(attrs, edges=edges)
    assert np.array_equal(bin.nmodes1d[0], herm.nmodes.astype('f8'))
    assert np.allclose(bin.xavg1d[0], herm.xavg, rtol=1e-14, equal_nan=True)
    nyq = attrs.kfun.min() * 16 / 2.
    mid = bin.edges.mean(axis=-1)
    assert np.all(mid[:, 0] <= mid[:, 1]) and np.all(bin.edges[:, :, 1].sum(axis=1) <= nyq)
    assert np.array_equal(bin.nmodes, bin.nmodes1d[0][bin.iedges[:, 0]] * bin.nmodes1d[1][bin.iedges[:, 1]])
    diag = po.BinBispectrumSugiyama(attrs, edges=edges, ells=[(0, 0, 0)], diagonal=True)
    assert np.all(diag.iedges[:, 0] == diag.iedges[:, 1])


def test_sugiyama_shotnoise_s111_known_answer():
    """S111 with NGP, no interlacing, no compensation: the painted w^3 mesh is exactly the per-cell
    sum, so s111_lm = sum over particles of w^3 Y*_lm(xhat of the particle's node)."""
    attrs = po.Attrs(8, 80., (15., -10., 60.))
    rng = np.random.default_rng(4)
    nodes = rng.integers(0, 8, (500, 3))
    pos = attrs.boxcenter - attrs.boxsize / 2. + nodes * attrs.cellsize + 0.3 * attrs.cellsize
    w = rng.uniform(0.5, 1.5, 500)
    bin = po.BinBispectrumSugiyama(attrs, edges={'step': 0.08}, ells=[(0, 0, 0), (2, 0, 2)])
    parts = {}
    po.sugiyama_shotnoise(attrs, bin, pos, w, los='z', resampler='ngp', interlacing=0, compensate=False, parts=parts)
    xn = attrs.boxcenter - attrs.boxsize / 2. + nodes * attrs.cellsize
    for (l, m), v in parts['s111'].items():
        assert np.isclose(v, np.sum(w**3 * po.complex_ylm_conj(l, m, *xn.T)), rtol=1e-12, atol=1e-9), (l, m)
    assert np.isclose(parts['s111'][0, 0], np.sum(w**3))


@pytest.mark.parametrize('ell', [0, 1, 2, 3])
def test_sugiyama_s122_kernel_addition_theorem(ell):
    """sum_m Y_lm(khat) conj(FFT(rho Y_lm(xhat)))(k) = sum_x rho(x) exp(+i k.x) P_l(khat.xhat), the
    addition theorem of the reduced harmonics that the S122 term rests on, by direct sums on a
    grid without the origin node (k = 0 excluded)."""
    attrs = po.Attrs((6, 4, 5), (60., 40., 50.), (13., -7., 40.))
    rng = np.random.default_rng(ell)
    rho = rng.normal(size=(6, 4, 5))
    X = np.meshgrid(*attrs.rcoords(), indexing='ij')
    kk = po.full_kvec(attrs)
    K = np.meshgrid(*kk, indexing='ij')
    got = 0.
    for m in range(-ell, ell + 1):
        got = got + np.conj(po.complex_ylm_conj(ell, m, *K)) * np.conj(np.fft.fftn(rho * np.conj(po.complex_ylm_conj(ell, m, *X))))
    xv = np.stack([x.ravel() for x in X], axis=1)
    xh = xv / np.linalg.norm(xv, axis=1)[:, None]
    kv = np.stack([k.ravel() for k in K], axis=1)
    idx = np.stack(np.meshgrid(*[np.arange(n) for n in (6, 4, 5)], indexing='ij'), axis=-1).reshape(-1, 3)
    for i in range(1, len(kv), 7):
        kh = kv[i] / np.linalg.norm(kv[i])
        phase = np.exp(2j * np.pi * (idx @ (idx[i] / np.array([6, 4, 5]))))
        want = np.sum(rho.ravel() * phase * special.eval_legendre(ell, xh @ kh))
        assert np.isclose(got.ravel()[i], want, rtol=1e-10, atol=1e-10 * np.abs(rho).sum())


@pytest.mark.parametrize('ell,los,klimit', [((0, 0, 0), 'z', None), ((2, 0, 2), 'x', None), ((1, 1, 0), 'local', None),
                                            ((2, 2, 2), 'local', True)])
def test_correlation3_sugiyama_vs_mode_pair_sums(ell, los, klimit):
    """The three-point correlation function's FFT reduction against direct sums over pairs of
    full-grid modes (Nyquist planes included)."""
    attrs = po.Attrs((6, 4, 6), (60., 40., 60.), (10., -20., 35.))
    rng = np.random.default_rng(31)
    meshes = [rng.normal(size=(6, 4, 6)) for _ in range(3)]
    bin = po.BinCorrelation3Sugiyama(attrs, edges=np.array([0., 12., 25., 40.]), ells=[ell], mask_edges='', klimit=klimit)
    got = po.correlation3_sugiyama(attrs, bin, *meshes, los=los)['num_raw'][0]
    want = po.correlation3_sugiyama_bruteforce(attrs, bin, meshes, ell=ell, los=los).real
    assert np.allclose(got, want, rtol=1e-10, atol=1e-12 * np.abs(want).max())
    assert np.abs(want).max() > 0


def test_sugiyama_correlation_shotnoise_full_grid_sums():
    """Three-point correlation shot noise with one separation bin holding every cell: summing
    ifftn(P) over all cells gives P(k = 0), so for (0, 0, 0) the S122 sum is
    C(0) conj(rho_c~(0)) - s111 W(0) = -s111 (k = 0 of the paint removed, W(0) = 1) and the S113
    sum is -s111 likewise; the S111 term is s111 / Vcell^2."""
    attrs = po.Attrs((8, 6, 8), (80., 60., 80.), (5., 3., 40.))
    rng = np.random.default_rng(8)
    pos = rng.uniform(-1., 1., (800, 3)) * attrs.boxsize / 2. + attrs.boxcenter
    w = rng.uniform(0.5, 1.5, 800)
    bin = po.BinCorrelation3Sugiyama(attrs, edges=np.array([0., 1e3]), ells=[(0, 0, 0)], mask_edges='')
    parts = {}
    total = po.sugiyama_shotnoise(attrs, bin, pos, w, los='z', resampler='tsc', interlacing=2, compensate=True, parts=parts)[0]
    s111 = parts['s111'][0, 0]
    ncell = np.prod(attrs.meshsize)
    vc2 = np.prod(attrs.cellsize)**2
    assert bin.nmodes1d[0][0] == ncell
    assert np.isclose(parts['s122'][0, 0][0] * ncell, -s111, rtol=1e-9)
    assert np.isclose(parts['s122'][1, 0][0] * ncell, -s111, rtol=1e-9)
    assert np.isclose(parts['s113'][0, 0, 0][0], -s111 / ncell**2 / vc2, rtol=1e-9)
    want = s111 / vc2 + 2 * (-s111 / ncell) / vc2 + (-s111 / ncell**2 / vc2)
    assert np.isclose(total[0], want, rtol=1e-9)
