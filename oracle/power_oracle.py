"""
Float64 NumPy restatement of jax-power's mesh power-spectrum path (the oracle).

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped package imports this module:
it is used by ``tests/``, by ``__graft_entry__.smoke()`` as the checker, and by
``bench.py``'s ``cpu_baseline`` leg as the timed CPU port.  Every function cites
the reference file:line it restates (paths relative to the jax-power repo,
``jaxpower/`` omitted for package files).

Conventions followed (all from the reference):
  * grid units   x~ = (x + L/2 - c) / cellsize                     mesh.py:2684
  * resamplers   id0 = round (odd order) / floor (even order),
                 stencil offsets arange(p) - (p-1)//2, periodic wrap  resamplers.py:73-103
  * interlacing  shifts j/n_i on positions, phase exp(+i s sum k_circ)/n_i,
                 compensation after the interlacing sum             mesh.py:2676-2752
  * spectra      hermitian mode weights 1 + (kz > 0), digitize(right=False),
                 DC mode kept in the first l=0 bin, removed for l>0  mesh.py:2812-2910,
                                                                     mesh2.py:25-119, 253-374
Pinning of every function is in tests/test_oracle.py; a function listed as
"parity unpinned" there has no independent check.
"""
from functools import lru_cache
from itertools import product
import math

import os

import numpy as np
from scipy import fft as scipy_fft
from scipy import special

# threads of the oracle's FFTs (the host cores this process may use)
FFT_WORKERS = len(os.sched_getaffinity(0)) if hasattr(os, 'sched_getaffinity') else (os.cpu_count() or 1)


# ----------------------------------------------------------------------------
# Geometry: MeshAttrs / get_mesh_attrs / next_fft_size
# ----------------------------------------------------------------------------

def fill3(value, dtype=float):
    """Broadcast a scalar or 3-sequence to a length-3 array (mesh.py _np_array_fill)."""
    value = np.asarray(value, dtype=dtype)
    return np.broadcast_to(value, (3,)).copy()


class Attrs:
    """Host geometry of a mesh: MeshAttrs (mesh.py:1453-1571), float64 only."""

    def __init__(self, meshsize, boxsize=None, boxcenter=0.):
        self.meshsize = fill3(meshsize, dtype='i8')
        if boxsize is None:
            boxsize = 1. * self.meshsize
        self.boxsize = fill3(boxsize)
        self.boxcenter = fill3(boxcenter)

    @property
    def cellsize(self):  # mesh.py:1558-1561
        return self.boxsize / self.meshsize

    @property
    def kfun(self):  # mesh.py:1563-1566
        return 2. * np.pi / self.boxsize

    @property
    def knyq(self):  # mesh.py:1568-1571
        return np.pi * self.meshsize / self.boxsize

    def rcoords(self):
        """Node positions, kind='position' (mesh.py:1573-1603, 360-403)."""
        offset = self.boxcenter - self.boxsize / 2.
        return tuple(np.arange(n) * d + o for n, d, o in zip(self.meshsize, self.cellsize, offset))

    def kcoords(self, kind='wavenumber'):
        """Half-spectrum Fourier coordinates (mesh.py:1605-1636, fftfreq 360-403).
        kind='wavenumber': fftfreq(n) * kfun * n ; kind='circular': fftfreq(n) * 2pi/n * n.
        Last axis uses rfftfreq (hermitian)."""
        toret = []
        for axis, n in enumerate(self.meshsize):
            spacing = self.kfun[axis] if kind == 'wavenumber' else 2. * np.pi / n
            f = np.fft.rfftfreq(n) if axis == 2 else np.fft.fftfreq(n)
            toret.append(f * spacing * n)
        return toret


def next_fft_size(n, primes=(2, 3, 5, 7), divisors=()):
    """mesh.py:2226-2262: smallest m >= n with prime factors in primes and divisible by divisors."""
    n = int(n)
    primes = tuple(p for p in primes if p > 1)
    divisors = tuple(int(d) for d in divisors if d > 1)
    D = 1
    for d in divisors:
        D = D * d // math.gcd(D, d)
    m = ((n + D - 1) // D) * D
    if not primes:
        return m
    while True:
        x = m
        for p in primes:
            while x % p == 0:
                x //= p
        if x == 1:
            return m
        m += D


def get_mesh_attrs(*positions, meshsize=None, boxsize=None, boxcenter=None, cellsize=None,
                   boxpad=2., approximate=False, primes=None, divisors=None):
    """mesh.py:2270-2381 (single process, 3 dimensions)."""
    if cellsize is not None and meshsize is not None:
        if boxsize is not None:
            raise ValueError('cannot specify boxsize, cellsize and meshsize simultaneously')
        meshsize = fill3(meshsize, 'i8')
        boxsize = meshsize * fill3(cellsize)
    if boxsize is None or boxcenter is None:
        if not positions:
            raise ValueError('positions must be provided if boxsize and boxcenter are not specified')
        pos_min = np.min([np.min(p, axis=0) for p in positions if p.size], axis=0)
        pos_max = np.max([np.max(p, axis=0) for p in positions if p.size], axis=0)
        delta = np.abs(pos_max - pos_min)
        if boxcenter is None:
            boxcenter = 0.5 * (pos_min + pos_max)
        if boxsize is None:
            boxsize = delta.max() * boxpad
    boxsize = fill3(boxsize)
    if meshsize is None:
        if cellsize is None:
            raise ValueError('meshsize (or cellsize) must be specified')
        cellsize = fill3(cellsize)
        meshsize = np.ceil(boxsize / cellsize).astype('i8')
        primes = tuple(primes or ())
        divisors = tuple(divisors or ())
        meshsize = np.array([next_fft_size(m, primes=primes, divisors=divisors) for m in meshsize])
        newboxsize = meshsize * cellsize
        if not positions and not approximate and not np.allclose(newboxsize, boxsize):
            raise ValueError('cannot enforce cellsize')
        boxsize = newboxsize
    return Attrs(meshsize, boxsize, fill3(boxcenter))


# ----------------------------------------------------------------------------
# Resamplers (resamplers.py)
# ----------------------------------------------------------------------------

def resampler_kernel(order, s):
    """1-D assignment kernels, resamplers.py:64-70."""
    s = np.asarray(s, dtype='f8')
    if order == 1:
        return np.ones_like(s)
    if order == 2:
        return 1. - s
    if order == 3:
        return (s <= 0.5) * (0.75 - s**2) + (0.5 < s) / 2. * (1.5 - s)**2
    if order == 4:
        return (s <= 1.) / 6. * (4. - 6. * s**2 + 3. * s**3) + (1. < s) / 6. * (2. - s)**3
    raise ValueError(order)


ORDERS = {'ngp': 1, 'cic': 2, 'tsc': 3, 'pcs': 4}


def _stencil(order):
    ishifts = np.arange(order) - (order - 1) // 2
    return np.array(list(product(ishifts, repeat=3)), dtype='i8')


def paint_grid(meshsize, xg, weights, order, threads=1):
    """Scatter-add particles given in grid units, resamplers.py:73-103: per axis the p stencil
    indices (periodic) and 1-D kernel values, then one unbuffered np.add.at per 3-D stencil
    offset of the product weight (the same sum as the reference's scatter, in another order).
    threads > 1 splits the offsets over that many threads, each with its own mesh."""
    shape = tuple(int(n) for n in meshsize)
    xg = np.asarray(xg, dtype='f8')
    w = np.broadcast_to(np.asarray(weights, dtype='f8'), xg.shape[:1])
    size = int(np.prod(shape))
    if not len(xg):
        return np.zeros(shape)
    id0 = (np.round if order % 2 else np.floor)(xg).astype('i8')
    ishifts = np.arange(order) - (order - 1) // 2
    idx = [[(id0[:, a] + s) % shape[a] for s in ishifts] for a in range(3)]
    ker = [[resampler_kernel(order, np.abs(id0[:, a] + s - xg[:, a])) for s in ishifts] for a in range(3)]
    offsets = list(product(range(order), repeat=3))

    def scatter(part):
        mesh = np.zeros(size, dtype='f8')
        for ox, oy, oz in part:
            flat = (idx[0][ox] * shape[1] + idx[1][oy]) * shape[2] + idx[2][oz]
            np.add.at(mesh, flat, w * ker[0][ox] * ker[1][oy] * ker[2][oz])
        return mesh

    threads = max(1, min(int(threads), len(offsets)))
    if threads == 1:
        return scatter(offsets).reshape(shape)
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(threads) as ex:
        meshes = list(ex.map(scatter, [offsets[i::threads] for i in range(threads)]))
    return sum(meshes[1:], meshes[0]).reshape(shape)


def paint_grid_bruteforce(meshsize, xg, weights, order):
    """Pure-Python loop version of paint_grid (tiny inputs only), resamplers.py:73-103."""
    shape = tuple(int(n) for n in meshsize)
    mesh = np.zeros(shape)
    for x, w in zip(np.asarray(xg, dtype='f8'), np.broadcast_to(weights, (len(xg),))):
        i0 = [int(np.round(v)) if order % 2 else int(np.floor(v)) for v in x]
        for off in _stencil(order):
            idx = [i0[a] + off[a] for a in range(3)]
            k = 1.
            for a in range(3):
                k *= float(resampler_kernel(order, abs(idx[a] - x[a])))
            mesh[idx[0] % shape[0], idx[1] % shape[1], idx[2] % shape[2]] += w * k
    return mesh


def read_grid(mesh, xg, order):
    """Gather, resamplers.py:106-131."""
    shape = np.array(mesh.shape)
    xg = np.asarray(xg, dtype='f8')
    out = np.zeros(len(xg))
    id0 = (np.round if order % 2 else np.floor)(xg).astype('i8')
    for ishift in _stencil(order):
        idx = id0 + ishift
        ker = resampler_kernel(order, np.abs(idx - xg)).prod(axis=-1)
        idx = idx % shape
        out += mesh[tuple(idx.T)] * ker
    return out


def compensation(attrs, order):
    """prod_i sinc(k_i/2pi)^(-order) on the half spectrum, k circular: resamplers.py:16-26."""
    kx, ky, kz = attrs.kcoords('circular')
    cx, cy, cz = (np.sinc(k / (2. * np.pi)) for k in (kx, ky, kz))
    return (cx[:, None, None] * cy[None, :, None] * cz[None, None, :])**(-order)


def to_grid(attrs, positions):
    """mesh.py:2684."""
    return (np.asarray(positions, dtype='f8') + attrs.boxsize / 2. - attrs.boxcenter) / attrs.cellsize


def r2c(value):
    """MeshAttrs.r2c with fft_backend='jax': rfftn (mesh.py:1679-1697).  scipy's pocketfft
    (numpy's own FFT algorithm) on FFT_WORKERS threads."""
    return scipy_fft.rfftn(value, workers=FFT_WORKERS)


def c2r(value, meshsize):
    """MeshAttrs.c2r: irfftn(s=meshsize) (mesh.py:1699-1719)."""
    return scipy_fft.irfftn(value, s=tuple(int(n) for n in meshsize), axes=(0, 1, 2), workers=FFT_WORKERS)


def paint(attrs, positions, weights=None, resampler='cic', interlacing=0, compensate=False, out='real', threads=1):
    """_paint, mesh.py:2673-2754 (single device).  threads > 1 paints the interlaced meshes
    concurrently and splits each scatter over threads / n_i threads (same sums)."""
    order = ORDERS[resampler]
    if weights is None:
        weights = np.ones(len(positions))
    xg = to_grid(attrs, positions)
    ni = max(interlacing, 1)
    if ni <= 1:
        mesh = paint_grid(attrs.meshsize, xg, weights, order, threads=threads)
        if compensate:
            cmesh = r2c(mesh) * compensation(attrs, order)
            return c2r(cmesh, attrs.meshsize) if out == 'real' else cmesh
        return mesh if out == 'real' else r2c(mesh)
    kx, ky, kz = attrs.kcoords('circular')
    ksum = kx[:, None, None] + ky[None, :, None] + kz[None, None, :]

    def shifted(shift):
        return r2c(paint_grid(attrs.meshsize, xg + shift, weights, order, threads=max(1, threads // ni))) * np.exp(shift * 1j * ksum) / ni

    shifts = np.arange(ni) * 1. / ni
    if threads > 1:
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(ni) as ex:
            terms = list(ex.map(shifted, shifts))
    else:
        terms = [shifted(shift) for shift in shifts]
    acc = 0.
    for term in terms:
        acc = acc + term
    if compensate:
        acc = acc * compensation(attrs, order)
    return c2r(acc, attrs.meshsize) if out == 'real' else acc


def read(attrs, mesh, positions, resampler='cic', compensate=False):
    """RealMeshField.read / _read, mesh.py:2030-2115 (single device)."""
    order = ORDERS[resampler]
    if compensate:
        mesh = c2r(r2c(mesh) * compensation(attrs, order), attrs.meshsize)
    return read_grid(mesh, to_grid(attrs, positions), order)


# ----------------------------------------------------------------------------
# Binning (mesh.py:2788-2910, mesh2.py:25-119)
# ----------------------------------------------------------------------------

def knorm(attrs):
    """|k| on the half spectrum, summed x, y, z in that order as in mesh2.py:54."""
    kx, ky, kz = attrs.kcoords('wavenumber')
    k2 = kx[:, None, None]**2 + ky[None, :, None]**2
    k2 = k2 + kz[None, None, :]**2
    return np.sqrt(k2)


def hermitian_weights(attrs):
    """1 + (kz > 0), mesh.py:2812-2828 (the Nyquist plane kz=+knyq also gets 2)."""
    kz = attrs.kcoords('wavenumber')[2]
    return np.broadcast_to((1 + (kz > 0)).astype('i8')[None, None, :],
                           (attrs.meshsize[0], attrs.meshsize[1], len(kz)))


def make_edges(attrs, edges=None):
    """Edge array from dict {'min','max','step'} (mesh2.py:39-47); arrays pass through."""
    if edges is None:
        edges = {}
    if isinstance(edges, dict):
        step = edges.get('step', None)
        vecmax = attrs.kfun.min() * np.min(attrs.meshsize) / 2.
        if step is None:
            return find_unique_edges(attrs, edges.get('min', 0.), edges.get('max', np.inf))
        return np.arange(edges.get('min', 0.), edges.get('max', vecmax), step)
    edges = np.asarray(edges, dtype='f8')
    if edges.ndim == 2:
        edges = np.append(edges[:, 0], edges[-1, 1])
    return edges


def find_unique_edges(attrs, xmin=0., xmax=np.inf):
    """mesh.py:2788-2809 (single device), on the wavenumbers."""
    return unique_edges(attrs.kcoords('wavenumber'), attrs.kfun.min(), xmin, xmax)


def unique_edges(vec, x0, xmin=0., xmax=np.inf):
    """_find_unique_edges, mesh.py:2788-2809: bin edges midway between the distinct norms of
    the lattice vectors vec (three 1-D coordinate arrays), x0 the smallest spacing."""
    kx, ky, kz = vec
    x2 = (kx[:, None, None]**2 + ky[None, :, None]**2 + kz[None, None, :]**2).ravel()
    x2 = x2[(x2 >= xmin**2) & (x2 <= xmax**2)]
    _, index = np.unique(np.int64(x2 / (0.5 * x0)**2 + 0.5), return_index=True)
    x = np.sqrt(x2[index])
    tmp = (x[:-1] + x[1:]) / 2.
    return np.concatenate([[tmp[0] - (x[1] - x[0])], tmp, [tmp[-1] + (x[-1] - x[-2])]])


def oversampling_shifts(mode_oversampling):
    """The shifts of _make_edges2 (mesh2.py:48-49): every integer triple in [-m, m]^3, last
    axis fastest (itertools.product order)."""
    r = range(-int(mode_oversampling), int(mode_oversampling) + 1)
    return [(a, b, c) for a in r for b in r for c in r]


def shifted_norm(vec, shift):
    """sqrt(sum((x + s)^2)) over the three axes in order (mesh2.py:51): ((x^2 + y^2) + z^2)."""
    x, y, z = (v + s for v, s in zip(vec, shift))
    k2 = x[:, None, None]**2 + y[None, :, None]**2
    k2 = k2 + z[None, None, :]**2
    return np.sqrt(k2)


def oversampled_bins(vec, edges, weights, mode_oversampling):
    """_make_edges2's loop (mesh2.py:48-61): one digitize per shift, nmodes and xsum summed over
    the shifts; nmodes divided by the number of shifts when there are several (a float then)
    and xavg = xsum / that nmodes.  With weights None every mode counts 1."""
    shifts = oversampling_shifts(mode_oversampling)
    length = len(edges) + 1
    ibin, nmodes, xsum = [], 0, 0.
    for shift in shifts:
        coords = shifted_norm(vec, shift)
        ib = np.digitize(coords, edges, right=False)  # mesh.py:2840
        w = None if weights is None else np.broadcast_to(weights, coords.shape).ravel()
        ibin.append(ib)
        counts = np.bincount(ib.ravel(), weights=w, minlength=length)[1:-1]
        nmodes = nmodes + counts.round().astype('i8')
        xsum = xsum + np.bincount(ib.ravel(), weights=coords.ravel() * (1 if w is None else w), minlength=length)[1:-1]
    if len(shifts) > 1:
        nmodes = nmodes / len(shifts)
    with np.errstate(invalid='ignore', divide='ignore'):
        xavg = xsum / nmodes  # mesh2.py:62 (NaN for empty bins)
    return ibin, nmodes, xavg


def bincount_shifts(ibin, value, nbins):
    """_bincount over the list of per-shift bin indices (mesh.py:2874-2910): the per-shift
    bincounts summed, then divided by their number."""
    n = nbins + 2
    s = 0.
    for ib in ibin:
        ib = ib.ravel()
        if np.iscomplexobj(value):
            s = s + np.bincount(ib, weights=value.real.ravel(), minlength=n) + 1j * np.bincount(ib, weights=value.imag.ravel(), minlength=n)
        else:
            s = s + np.bincount(ib, weights=value.ravel(), minlength=n)
    return s[1:-1] / len(ibin)


class BinPoles:
    """BinMesh2SpectrumPoles (mesh2.py:65-119) with mode_oversampling (mesh2.py:48-61)."""

    def __init__(self, attrs, edges=None, ells=0, mode_oversampling=0):
        self.attrs = attrs
        edges = make_edges(attrs, edges)
        self.wmodes = hermitian_weights(attrs)
        self.ibin, self.nmodes, self.xavg = oversampled_bins(attrs.kcoords('wavenumber'), edges, self.wmodes, mode_oversampling)
        if mode_oversampling == 0:
            self.ibin = self.ibin[0]
        self.edges = np.column_stack([edges[:-1], edges[1:]])
        self.ells = tuple(sorted(ells if np.ndim(ells) else (ells,)))

    def __call__(self, value, remove_zero=False):
        """mesh2.py:100-119 and mesh.py:2874-2910."""
        value = np.array(value)
        if remove_zero:
            value[0, 0, 0] = 0.
        value = value * self.wmodes
        ibin = self.ibin if isinstance(self.ibin, list) else [self.ibin]
        with np.errstate(invalid='ignore', divide='ignore'):
            return bincount_shifts(ibin, value, len(self.xavg)) / self.nmodes


# ----------------------------------------------------------------------------
# Legendre and real spherical harmonics (utils.py:190-360)
# ----------------------------------------------------------------------------

_LEGENDRE = {
    0: lambda x: np.ones_like(x),
    1: lambda x: x,
    2: lambda x: 3 * x**2 / 2 - 1 / 2,
    3: lambda x: 5 * x**3 / 2 - 3 * x / 2,
    4: lambda x: 35 * x**4 / 8 - 15 * x**2 / 4 + 3 / 8,
    5: lambda x: 63 * x**5 / 8 - 35 * x**3 / 4 + 15 * x / 8,
    6: lambda x: 231 * x**6 / 16 - 315 * x**4 / 16 + 105 * x**2 / 16 - 5 / 16,
}


def legendre(ell, x):
    """Closed-form Legendre polynomials, utils.py:341-360."""
    return _LEGENDRE[ell](np.asarray(x, dtype='f8'))


def ylm_at_zero(ell, m):
    """Value of the reference's real Y_lm at the zero vector.  utils.py:252-254 maps the zero
    vector to xhat = (0, 0, 0) and evaluates the sympy polynomial built at utils.py:299-327, in
    which the r^2 substitution leaves a constant term for (l, m) = (0, 0) (Y_00 itself) and
    (4, 0) (3/8 sqrt(9/4pi), since 35 z^4 - 30 z^2 r^2 + 3 r^4 stays 35 zhat^4 - 30 zhat^2 + 3);
    every other l <= 4 polynomial is homogeneous and gives 0.  Pinned in tests/test_oracle.py by
    redoing the sympy construction."""
    if ell == 0:
        return 1. / np.sqrt(4. * np.pi)
    if ell == 4 and m == 0:
        return 3. / 8. * np.sqrt(9. / (4. * np.pi))
    return 0.


def addition_at(ell, u, v):
    """4 pi / (2l+1) sum_m Y_lm(u) Y_lm(v) for arrays of vectors u (..., 3) and v (..., 3), with
    the reference's values at zero vectors (ylm_at_zero): P_l(uhat . vhat) when both are nonzero
    (addition theorem); with u = 0, 4 pi/(2l+1) Y_l0(0) Y_l0(vhat), i.e. 1 for l = 0,
    P_4(0) P_4(vhat_z) for l = 4 and 0 otherwise; both zero: 1, P_4(0)^2 or 0.  Independent of
    the per-m sign convention of the Y_lm."""
    u, v = np.broadcast_arrays(np.asarray(u, dtype='f8'), np.asarray(v, dtype='f8'))
    un, vn = np.linalg.norm(u, axis=-1), np.linalg.norm(v, axis=-1)
    uh = u / np.where(un == 0, 1., un)[..., None]
    vh = v / np.where(vn == 0, 1., vn)[..., None]
    both = legendre(ell, np.sum(uh * vh, axis=-1))
    if ell == 0:
        return np.ones_like(both)
    if ell != 4:
        return np.where((un == 0) | (vn == 0), 0., both)
    p40 = legendre(4, 0.)
    out = np.where(un == 0, p40 * legendre(4, vh[..., 2]), both)
    out = np.where(vn == 0, p40 * legendre(4, uh[..., 2]), out)
    return np.where((un == 0) & (vn == 0), p40 * p40, out)


def real_ylm(ell, m, x, y, z):
    """Real spherical harmonic of (x, y, z)/|.| with the reference normalisation
    sqrt((2l+1)/4pi) sqrt((l-|m|)!/(l+|m|)!) sqrt(2) (utils.py:227-238), and the reference's
    value at the zero vector (ylm_at_zero; utils.py:252-254).  Built from scipy's associated
    Legendre function; the per-m sign convention may differ from the reference's sympy
    expression, which cancels in sum_m Y_lm(a) Y_lm(b) (the only combination the estimator
    uses, mesh2.py:319-329)."""
    x, y, z = np.broadcast_arrays(*(np.asarray(v, dtype='f8') for v in (x, y, z)))
    r = np.sqrt(x**2 + y**2 + z**2)
    with np.errstate(invalid='ignore', divide='ignore'):
        zh = np.where(r == 0, 0., z / np.where(r == 0, 1., r))
    phi = np.arctan2(y, x)
    amp = np.sqrt((2 * ell + 1) / (4 * np.pi))
    if m != 0:
        amp *= np.sqrt(math.factorial(ell - abs(m)) / math.factorial(ell + abs(m))) * np.sqrt(2.)
    toret = amp * special.lpmv(abs(m), ell, zh)
    if m < 0:
        toret = toret * np.sin(abs(m) * phi)
    elif m > 0:
        toret = toret * np.cos(m * phi)
    return np.where(r == 0, ylm_at_zero(ell, m), toret)


# ----------------------------------------------------------------------------
# Estimator (mesh2.py:253-374) and normalisation / shot noise (mesh2.py:505-621)
# ----------------------------------------------------------------------------

def los_vector(los):
    """mesh2.py:195-203."""
    if isinstance(los, str):
        v = [0., 0., 0.]
        v['xyz'.index(los)] = 1.
        return np.array(v)
    return np.asarray(los, dtype='f8')


def spectrum_poles(attrs, bin, mesh1, mesh2=None, los='z'):
    """compute_mesh2_spectrum, mesh2.py:253-374.  mesh1/mesh2 are real (N,N,N) or complex
    half spectra (N,N,N//2+1).  Returns dict with k, nmodes, num_raw (per ell), norm."""
    ells = bin.ells
    same = mesh2 is None
    norm = np.prod(attrs.meshsize) / np.prod(attrs.cellsize) * np.ones_like(bin.xavg)  # mesh2.py:277

    def c(m):
        return m if np.iscomplexobj(m) else r2c(m)

    def r(m):
        return c2r(m, attrs.meshsize) if np.iscomplexobj(m) else m

    num = []
    if isinstance(los, str) and los in ('firstpoint', 'local', 'endpoint'):  # mesh2.py:293-344
        meshes = [mesh1, mesh1 if same else mesh2]
        swap = los == 'endpoint'
        if swap:
            meshes = meshes[::-1]
        rmesh1 = meshes[0]
        A0 = c(rmesh1 if same else meshes[1])
        if 0 in ells:
            Aell = (A0.real**2 + A0.imag**2) if same else np.conj(c(rmesh1)) * A0
            num.append(bin(Aell))
        rmesh1 = r(rmesh1)
        xs = attrs.rcoords()
        xx, xy, xz = xs[0][:, None, None], xs[1][None, :, None], xs[2][None, None, :]
        kx, ky, kz = attrs.kcoords('wavenumber')
        kx, ky, kz = kx[:, None, None], ky[None, :, None], kz[None, None, :]
        for ell in ells:
            if ell == 0:
                continue
            C = 0.
            for m in range(-ell, ell + 1):
                C = C + r2c(rmesh1 * real_ylm(ell, m, xx, xy, xz)) * real_ylm(ell, m, kx, ky, kz)
            num.append(4. * np.pi * bin(np.conj(C) * A0, remove_zero=True))
        if swap:
            num = [np.conj(n) for n in num]
        los_attr = 'firstpoint' if los == 'local' else los
    else:  # mesh2.py:346-374
        vlos = los_vector(los)
        A1 = c(mesh1)
        Aell = (A1.real**2 + A1.imag**2) if same else np.conj(A1) * c(mesh2)
        if 0 in ells:
            num.append(bin(Aell))
        kx, ky, kz = attrs.kcoords('wavenumber')
        kx, ky, kz = kx[:, None, None], ky[None, :, None], kz[None, None, :]
        kn = np.sqrt(kx**2 + ky**2 + kz**2)
        kn[0, 0, 0] = 1.
        mu = (kx * vlos[0] + ky * vlos[1] + kz * vlos[2]) / kn
        for ell in ells:
            if ell == 0:
                continue
            num.append((2 * ell + 1) * bin(Aell * legendre(ell, mu), remove_zero=True))
        los_attr = vlos
    num = [np.real(n) if ell % 2 == 0 else np.imag(n) for ell, n in zip(ells, num)]
    return dict(k=bin.xavg, k_edges=bin.edges, nmodes=bin.nmodes, num_raw=num, norm=norm,
                volume=np.prod(attrs.kfun) * bin.nmodes, ells=ells, los=los_attr)


def box_normalization(attrs, weights1, weights2=None):
    """compute_box2_normalization -> compute_box_normalization, mesh.py:3121-3129 / mesh2.py:505-513."""
    if weights2 is None:
        weights2 = weights1
    size = float(np.prod(attrs.meshsize))
    return (np.sum(weights1) / size) * (np.sum(weights2) / size) * size * np.prod(attrs.cellsize)**(-1)


def regrid_attrs(attrs, cellsize=10.):
    """get_halo_for_new_mattrs with cellsize (mesh.py:3002-3039), single device."""
    new = get_mesh_attrs(boxsize=attrs.boxsize, boxcenter=attrs.boxcenter, cellsize=cellsize, approximate=True)
    if np.any((new.meshsize - attrs.meshsize) % 2):
        meshsize = new.meshsize + 1
        new = Attrs(meshsize, new.cellsize * meshsize, new.boxcenter)
    return new


def fkp_normalization(attrs, data_positions, data_weights, randoms_positions, randoms_weights, cellsize=10.):
    """compute_fkp2_normalization (mesh2.py:516-574), auto spectrum, split=None:
    alpha * sum(D R) / cellvolume with D, R CIC-painted on a regridded mesh (mesh.py:3092-3118)."""
    alpha = np.sum(data_weights) / np.sum(randoms_weights)
    new = attrs if cellsize is None else regrid_attrs(attrs, cellsize)
    d = paint(new, data_positions, data_weights, resampler='cic', interlacing=0, compensate=False)
    r = paint(new, randoms_positions, randoms_weights, resampler='cic', interlacing=0, compensate=False)
    return alpha * np.sum(d * r) * np.prod(new.cellsize)**(-1)


def fkp3_normalization(attrs, data_positions, data_weights, randoms_positions, randoms_weights, cellsize=10.):
    """compute_fkp3_normalization (mesh3.py:738-783), one field, split=None: alpha^3 *
    sum(R^3) / cellvolume^2 with R CIC-painted on a regridded mesh (mesh.py:3092-3118), alpha =
    sum(w_data) / sum(w_randoms) once per each of the three (identical) inputs."""
    alpha = np.sum(data_weights) / np.sum(randoms_weights)
    new = attrs if cellsize is None else regrid_attrs(attrs, cellsize)
    r = paint(new, randoms_positions, randoms_weights, resampler='cic', interlacing=0, compensate=False)
    return alpha**3 * np.sum(r * r * r) * np.prod(new.cellsize)**(-2)


def fkp_shotnoise(data_weights, randoms_weights=None):
    """compute_fkp2_shotnoise (mesh2.py:577-621): sum of squared FKP weights
    [w_d, -alpha w_r] (mesh.py:2965-2971)."""
    sn = np.sum(np.asarray(data_weights, dtype='f8')**2)
    if randoms_weights is not None:
        alpha = np.sum(data_weights) / np.sum(randoms_weights)
        sn += alpha**2 * np.sum(np.asarray(randoms_weights, dtype='f8')**2)
    return sn


def pole_value(num_raw, norm, num_shotnoise=0.):
    """Mesh2SpectrumPole.value(): (num_raw - num_shotnoise) / norm (lsstypes; semantics inferred
    from tests/test_mesh2.py:60, 954-958)."""
    return (np.asarray(num_raw) - num_shotnoise) / norm


# ----------------------------------------------------------------------------
# Brute-force checkers (tiny N): direct DFT and direct local-LOS estimator
# ----------------------------------------------------------------------------

def dft3_half(value):
    """Direct O(N^6) r2c DFT, unnormalised e^{-ikx}: rfftn by definition."""
    nx, ny, nz = value.shape
    out = np.zeros((nx, ny, nz // 2 + 1), dtype='c16')
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing='ij')
    for a in range(nx):
        for b in range(ny):
            for c in range(nz // 2 + 1):
                ph = np.exp(-2j * np.pi * (a * ix / nx + b * iy / ny + c * iz / nz))
                out[a, b, c] = np.sum(value * ph)
    return out


def local_los_bruteforce(attrs, rmesh, ell):
    """Direct form of the firstpoint estimator before binning: for every half-spectrum mode,
    conj(sum_x rho(x) e^{-ikx} (2l+1) P_l(khat.xhat)) * A0(k) with A0 = rfftn(rho)
    (addition theorem applied to mesh2.py:319-333, 4pi * sum_m Y_lm Y_lm = (2l+1) P_l; at zero
    vectors the reference's Y values, addition_at)."""
    nx, ny, nz = rmesh.shape
    xs = attrs.rcoords()
    X = np.stack(np.meshgrid(*xs, indexing='ij'), axis=-1).reshape(-1, 3)
    ix = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing='ij'), axis=-1).reshape(-1, 3)
    kx, ky, kz = attrs.kcoords('wavenumber')
    A0 = r2c(rmesh)
    rho = rmesh.ravel()
    out = np.zeros(A0.shape, dtype='c16')
    for a in range(nx):
        for b in range(ny):
            for c in range(nz // 2 + 1):
                kv = np.array([kx[a], ky[b], kz[c]])
                pl = addition_at(ell, X, kv)  # P_l(xhat.khat), reference values at zero vectors
                ph = np.exp(-2j * np.pi * (a * ix[:, 0] / nx + b * ix[:, 1] / ny + c * ix[:, 2] / nz))
                C = (2 * ell + 1) * np.sum(rho * pl * ph)
                out[a, b, c] = np.conj(C) * A0[a, b, c]
    return out


# ----------------------------------------------------------------------------
# Correlation function multipoles (mesh2.py:120-184, 377-504)
# ----------------------------------------------------------------------------

def separation_coords(attrs):
    """xcoords(kind='separation'): fftfreq(N) * cellsize * N per axis (mesh.py:1573-1603,
    361-403)."""
    return [np.fft.fftfreq(int(n)) * c * n for n, c in zip(attrs.meshsize, attrs.cellsize)]


def snorm(attrs):
    """|s| on the full real mesh, summed x, y, z in that order (mesh2.py:54)."""
    sx, sy, sz = separation_coords(attrs)
    s2 = sx[:, None, None]**2 + sy[None, :, None]**2
    s2 = s2 + sz[None, None, :]**2
    return np.sqrt(s2)


def make_sedges(attrs, edges=None):
    """_make_edges2(kind='real'), mesh2.py:25-47: vec0 = smallest cellsize, default max
    vec0 * min(N) / 2; arrays pass through."""
    if edges is None:
        edges = {}
    if isinstance(edges, dict):
        step = edges.get('step', None)
        vec0 = attrs.cellsize.min()
        vecmax = vec0 * np.min(attrs.meshsize) / 2.
        if step is None:
            return unique_edges(separation_coords(attrs), vec0, edges.get('min', 0.), edges.get('max', np.inf))
        return np.arange(edges.get('min', 0.), edges.get('max', vecmax), step)
    edges = np.asarray(edges, dtype='f8')
    if edges.ndim == 2:
        edges = np.append(edges[:, 0], edges[-1, 1])
    return edges


class BinCorrelationPoles:
    """BinMesh2CorrelationPoles (mesh2.py:120-184), basis=None, with mode_oversampling
    (mesh2.py:48-61): every cell of the real mesh weighted 1, bins by digitize(|s + shift|,
    edges, right=False)."""

    def __init__(self, attrs, edges=None, ells=0, mode_oversampling=0):
        self.attrs = attrs
        edges = make_sedges(attrs, edges)
        self.ibin, self.nmodes, self.xavg = oversampled_bins(separation_coords(attrs), edges, None, mode_oversampling)
        if mode_oversampling == 0:
            self.ibin = self.ibin[0]
        self.edges = np.column_stack([edges[:-1], edges[1:]])
        self.ells = tuple(sorted(ells if np.ndim(ells) else (ells,)))

    def __call__(self, value, remove_zero=False):
        """mesh2.py:175-181 with _bincount, mesh.py:2874-2910."""
        value = np.array(value)
        if remove_zero:
            value[0, 0, 0] = 0.
        ibin = self.ibin if isinstance(self.ibin, list) else [self.ibin]
        with np.errstate(invalid='ignore', divide='ignore'):
            return bincount_shifts(ibin, value, len(self.xavg)) / self.nmodes


def correlation_poles(attrs, bin, mesh1, mesh2=None, los='z'):
    """compute_mesh2_correlation, mesh2.py:377-504 (basis=None).  mesh1/mesh2 are real
    (N,N,N) or complex half spectra.  Returns dict with s, nmodes, num_raw (per ell), norm."""
    ells = bin.ells
    same = mesh2 is None
    vcell = np.prod(attrs.cellsize)
    norm = np.prod(attrs.meshsize) / vcell * np.ones_like(bin.xavg)  # mesh2.py:400

    def c(m):
        return m if np.iscomplexobj(m) else r2c(m)

    def r(m):
        return c2r(m, attrs.meshsize) if np.iscomplexobj(m) else m

    num = []
    if isinstance(los, str) and los in ('firstpoint', 'local', 'endpoint'):  # mesh2.py:416-469
        meshes = [mesh1, mesh1 if same else mesh2]
        if los == 'endpoint':
            meshes = meshes[::-1]
        rmesh1 = meshes[0]
        A0 = c(rmesh1 if same else meshes[1])
        if 0 in ells:
            Aell = (A0.real**2 + A0.imag**2) if same else np.conj(c(rmesh1)) * A0
            num.append(bin(c2r(Aell, attrs.meshsize)))
        rmesh1 = r(rmesh1)
        xs = attrs.rcoords()
        xx, xy, xz = xs[0][:, None, None], xs[1][None, :, None], xs[2][None, None, :]
        ss = separation_coords(attrs)
        sx, sy, sz = ss[0][:, None, None], ss[1][None, :, None], ss[2][None, None, :]
        for ell in ells:
            if ell == 0:
                continue
            C = 0.
            for m in range(-ell, ell + 1):
                B = r2c(rmesh1 * real_ylm(ell, m, xx, xy, xz))
                C = C + c2r(np.conj(B) * A0, attrs.meshsize) * real_ylm(ell, m, sx, sy, sz)
            num.append(4. * np.pi * bin(C, remove_zero=True))
        los_attr = 'firstpoint' if los == 'local' else los
    else:  # mesh2.py:471-504
        vlos = los_vector(los)
        A1 = c(mesh1)
        Aell = c2r((A1.real**2 + A1.imag**2) if same else np.conj(A1) * c(mesh2), attrs.meshsize)
        if 0 in ells:
            num.append(bin(Aell))
        ss = separation_coords(attrs)
        sx, sy, sz = ss[0][:, None, None], ss[1][None, :, None], ss[2][None, None, :]
        sn = np.sqrt(sx**2 + sy**2 + sz**2)
        sn[0, 0, 0] = 1.
        mu = (sx * vlos[0] + sy * vlos[1] + sz * vlos[2]) / sn
        for ell in ells:
            if ell == 0:
                continue
            num.append((2 * ell + 1) * bin(Aell * legendre(ell, mu), remove_zero=True))
        los_attr = vlos
    num = [np.real(n) for n in num]
    return dict(s=bin.xavg, s_edges=bin.edges, nmodes=bin.nmodes, num_raw=[n / vcell for n in num], norm=norm,
                volume=vcell * bin.nmodes, ells=ells, los=los_attr)


def circular_correlation_bruteforce(mesh1, mesh2=None):
    """xi(s) = sum_x mesh1(x) mesh2(x + s), periodic, by explicit shifts (the definition the
    c2r(conj(A1) A2) route must reproduce); for small meshes only."""
    mesh2 = mesh1 if mesh2 is None else mesh2
    out = np.empty(mesh1.shape)
    for idx in np.ndindex(*mesh1.shape):
        out[idx] = np.sum(mesh1 * np.roll(mesh2, [-i for i in idx], axis=(0, 1, 2)))
    return out


def local_correlation_bruteforce(attrs, rmesh, ell):
    """sum over cells x and lags s of rho(x) rho(x + s) (2l+1)/(4pi) P_l(xhat . shat), per
    lag s, with the reference's Y values at zero vectors (addition_at; utils.py:252-254): the
    local line-of-sight term of mesh2.py:445-460 written without Y_lm (addition theorem);
    small meshes only."""
    xs = attrs.rcoords()
    X = np.stack(np.meshgrid(*xs, indexing='ij'), axis=-1)
    ss = separation_coords(attrs)
    out = np.zeros(rmesh.shape)
    for idx in np.ndindex(*rmesh.shape):
        svec = np.array([ss[a][idx[a]] for a in range(3)])
        w = (2 * ell + 1) / (4 * np.pi) * addition_at(ell, X, svec)
        out[idx] = np.sum(rmesh * w * np.roll(rmesh, [-i for i in idx], axis=(0, 1, 2)))
    return out


# ----------------------------------------------------------------------------
# Bispectrum, Scoccimarro basis, global line of sight (mesh3.py:23-177, 180-313, 509-629)
# ----------------------------------------------------------------------------

class BinBispectrumPoles:
    """BinMesh3SpectrumPoles with basis 'scoccimarro' or 'scoccimarro-diagonal'
    (mesh3.py:23-177, 180-313).  Side i bins its modes as BinPoles does; triangles are the
    lexicographic (b1, b2, b3) combinations (mesh3.py:107-124) that pass the mask, by default
    mid1 <= mid2 <= mid3, |mid1 - mid2| <= mid3 <= mid1 + mid2 and the sum of the upper edges
    <= 2 k_nyq (mesh3.py:31-35); mask_edges='' keeps every combination.  nmodes are the
    closed-triangle counts N^6 sum_x prod_i c2r([bin_i == b_i]) (mesh3.py:229-236)."""

    def __init__(self, attrs, edges=None, ells=0, mask_edges=None, diagonal=False):
        ndim = 3
        if not isinstance(edges, (list, tuple)):
            edges = [edges] * ndim
        self.attrs = attrs
        self.axes = [BinPoles(attrs, edges=e) for e in edges[:ndim]]
        self.axes += [self.axes[-1]] * (ndim - len(self.axes))
        self.ibin1d = [ax.ibin - 1 for ax in self.axes]  # mesh3.py:92-93
        self.edges1d = [ax.edges for ax in self.axes]
        self.xavg1d = [ax.xavg for ax in self.axes]
        self.nmodes1d = [ax.nmodes for ax in self.axes]
        self.ells = [int(ell) for ell in np.atleast_1d(ells)]

        def prod(arrays):
            if diagonal:
                return np.column_stack([np.asarray(arrays[0])] * ndim)
            grid = np.meshgrid(*arrays, indexing='ij')
            return np.column_stack([g.ravel() for g in grid])

        lo = prod([e[:, 0] for e in self.edges1d])
        hi = prod([e[:, 1] for e in self.edges1d])
        edges3 = np.stack([lo, hi], axis=-1)
        mid = edges3.mean(axis=-1)
        nyq = attrs.kfun.min() * np.min(attrs.meshsize) / 2.
        if mask_edges is None:
            m1, m2, m3 = mid.T
            mask = (m1 <= m2) & (m2 <= m3) & (m3 >= np.abs(m1 - m2)) & (m3 <= np.abs(m1 + m2))
            mask &= edges3[:, :, 1].sum(axis=1) <= 2 * nyq
        else:
            assert mask_edges == '', 'the oracle takes the default mask or none'
            mask = np.ones(len(edges3), dtype=bool)
        self.edges = edges3[mask]
        self.xavg = prod(self.xavg1d)[mask]
        self.iedges = prod([np.arange(len(x)) for x in self.xavg1d])[mask]
        ones = np.ones(self.ibin1d[0].shape, dtype='f8')
        self.nmodes = self.reduce(ones, ones, ones) * float(np.prod(attrs.meshsize))**2

    def shells(self, axis, value):
        """c2r(value [bin == b]) for every shell b of side axis (mesh3.py:300-305)."""
        return [c2r(value * (self.ibin1d[axis] == b), self.attrs.meshsize) for b in range(len(self.xavg1d[axis]))]

    def reduce(self, A1, A2, A3):
        """sum_x c2r(A1 [b1]) c2r(A2 [b2]) c2r(A3 [b3]) per triangle (mesh3.py:238-275, 307-311)."""
        s = [self.shells(i, A) for i, A in enumerate((A1, A2, A3))]
        return np.array([np.sum(s[0][b1] * s[1][b2] * s[2][b3]) for b1, b2, b3 in self.iedges])


def bispectrum_poles(attrs, bin, mesh1, mesh2=None, mesh3=None, los='x'):
    """compute_mesh3_spectrum, Scoccimarro basis, global line of sight (mesh3.py:509-629):
    num_raw = (2 ell + 1) N^6 reduce(A1, A2, A3 P_ell(mu)) / nmodes (real part; odd ell keep the
    imaginary part of a real sum, i.e. 0); norm = N^3 / Vcell^2.  Missing meshes repeat the last
    one given (mesh.py:3138-3156)."""
    meshes = [mesh1, mesh2, mesh3]
    for i in (1, 2):
        if meshes[i] is None:
            meshes[i] = meshes[i - 1]
    A = [m if np.iscomplexobj(m) else r2c(m) for m in meshes]
    n6 = float(np.prod(attrs.meshsize))**2
    norm = np.prod(attrs.meshsize) / np.prod(attrs.cellsize)**2
    if isinstance(los, str) and los in ('local', 'firstpoint'):  # mesh3.py:563-576
        rho3 = c2r(A[2], attrs.meshsize) if np.iscomplexobj(meshes[2]) else meshes[2]
        num = [np.real(4. * np.pi * bin.reduce(A[0], A[1], bispectrum_local_mesh(attrs, rho3, ell)) * n6 / bin.nmodes) for ell in bin.ells]
        return dict(k=bin.xavg, k_edges=bin.edges, nmodes=bin.nmodes, num_raw=num, norm=norm, ells=bin.ells, los='local')
    vlos = los_vector(los)
    kx, ky, kz = attrs.kcoords('wavenumber')
    kx, ky, kz = kx[:, None, None], ky[None, :, None], kz[None, None, :]
    kn = np.sqrt(kx**2 + ky**2 + kz**2)
    kn[0, 0, 0] = 1.
    mu = (kx * vlos[0] + ky * vlos[1] + kz * vlos[2]) / kn
    num = []
    for ell in bin.ells:
        if ell % 2:
            num.append(np.zeros(len(bin.iedges)))
            continue
        num.append((2 * ell + 1) * bin.reduce(A[0], A[1], A[2] * legendre(ell, mu)) * n6 / bin.nmodes)
    return dict(k=bin.xavg, k_edges=bin.edges, nmodes=bin.nmodes, num_raw=num, norm=norm, ells=bin.ells, los=vlos)


def _full_grid_bins(attrs, edges):
    """Shell index (digitize - 1) of every mode of the full FFT grid, and the integer mode
    indices (fftfreq order) per axis."""
    n = [int(x) for x in attrs.meshsize]
    kk = [np.fft.fftfreq(n[a]) * attrs.kfun[a] * n[a] for a in range(3)]
    k2 = kk[0][:, None, None]**2 + kk[1][None, :, None]**2
    k2 = k2 + kk[2][None, None, :]**2
    return np.digitize(np.sqrt(k2), edges, right=False) - 1, kk


def bispectrum_bruteforce(attrs, bin, rmeshes, ell=0, los='z'):
    """Closed-triangle sums over pairs of full-grid modes: for each triangle (b1, b2, b3),
    sum of F1(k1) F2(k2) F3(k3) P_ell(mu3) over k1 in shell b1, k2 in shell b2 and
    k3 = -(k1 + k2) (mod N) in shell b3, with F = fftn of the real meshes (counts with F = 1
    and ell = 0).  This is what N^6 sum_x prod c2r(...) must equal; small meshes only."""
    n = np.array([int(x) for x in attrs.meshsize])
    edges = np.append(bin.edges1d[0][:, 0], bin.edges1d[0][-1, 1])
    assert all(np.array_equal(e, bin.edges1d[0]) for e in bin.edges1d), 'one set of edges for all sides'
    ib, kk = _full_grid_bins(attrs, edges)
    F = [np.fft.fftn(m) if m is not None else np.ones(tuple(n)) for m in rmeshes]
    vlos = los_vector(los)
    idx = np.stack(np.meshgrid(*[np.arange(x) for x in n], indexing='ij'), axis=-1).reshape(-1, 3)
    ibf = ib.ravel()
    valid = ibf >= 0
    out = np.zeros(len(bin.iedges), dtype=complex)
    key = {tuple(t): i for i, t in enumerate(bin.iedges)}
    i1 = np.flatnonzero(valid)
    for a in i1:
        k1 = idx[a]
        k3 = (-(k1[None, :] + idx[i1])) % n
        b3 = ib[k3[:, 0], k3[:, 1], k3[:, 2]]
        kv = np.stack([kk[0][k3[:, 0]], kk[1][k3[:, 1]], kk[2][k3[:, 2]]], axis=-1)
        kn = np.linalg.norm(kv, axis=-1)
        mu = np.where(kn > 0, kv @ vlos / np.where(kn > 0, kn, 1.), 0.)
        f = F[0].ravel()[a] * F[1].ravel()[i1] * F[2][k3[:, 0], k3[:, 1], k3[:, 2]] * legendre(ell, mu)
        for j, (bb2, bb3) in enumerate(zip(ibf[i1], b3)):
            t = key.get((ibf[a], bb2, bb3))
            if t is not None:
                out[t] += f[j]
    return out


# ----------------------------------------------------------------------------
# Window matrix, global line of sight (mesh2.py:921-1118, 1453-1482)
# ----------------------------------------------------------------------------

def spectrum_window(attrs, bin, edgesin, ellsin=0, meshes=None, los='z', norm=None):
    """compute_mesh2_spectrum_window with a global line of sight and no flags
    (mesh2.py:921-1118, 1453-1482): for each input multipole ellin and input shell
    [e0, e1), the mesh-defined theory A = [e0 <= |k| < e1] P_ellin(mu) N^3 is (for a
    selection) convolved as r2c(Q c2r(A)) with Q = c2r(A1 conj(A2)) / N^3 (mesh2.py:1022-1029,
    1106-1107), then binned as (2l+1)/(1+odd) bin(A (P_l(mu) + odd P_l(-mu)), remove_zero=True)
    rnorm (mesh2.py:1037-1045), rnorm = N^3/V_cell / norm / N^3 (mesh2.py:973-979).
    meshes: None (periodic box, mesh2.py:1030-1034) or a list of one or two real meshes /
    half spectra.  Returns the (n_ells n_bins, n_ellsin n_in) real matrix: the reference
    keeps the real part of every row (mesh2.py:1466-1472)."""
    n3 = float(np.prod(attrs.meshsize))
    _norm = n3 / np.prod(attrs.cellsize)
    ells = bin.ells
    norm = np.broadcast_to(_norm if norm is None else np.asarray(norm, dtype='f8'), (len(ells),))
    rnorm = _norm / norm / n3
    ellsin = (ellsin,) if np.ndim(ellsin) == 0 else tuple(ellsin)
    edgesin = np.asarray(edgesin, dtype='f8')
    if edgesin.ndim != 2:
        edgesin = np.column_stack([edgesin[:-1], edgesin[1:]])
    vlos = los_vector(los)
    Q = None
    if meshes is not None:
        cm = [m if np.iscomplexobj(m) else r2c(m) for m in meshes]
        Q = c2r(cm[0] * np.conj(cm[-1]), attrs.meshsize) / n3
    kx, ky, kz = attrs.kcoords('wavenumber')
    kx, ky, kz = kx[:, None, None], ky[None, :, None], kz[None, None, :]
    kn = knorm(attrs)
    mu = (kx * vlos[0] + ky * vlos[1] + kz * vlos[2]) / np.where(kn == 0., 1., kn)
    cols = []
    for ellin in ellsin:
        legin = legendre(ellin, mu)
        for e0, e1 in edgesin:
            A = ((kn >= e0) & (kn < e1)) * legin * n3
            if Q is not None:
                A = r2c(Q * c2r(A, attrs.meshsize))
            col = []
            for ill, ell in enumerate(ells):
                odd = ell % 2
                leg = legendre(ell, mu) + (legendre(ell, -mu) if odd else 0.)
                col.append((2 * ell + 1) / (1 + odd) * bin(A * leg, remove_zero=True) * rnorm[ill])
            cols.append(np.concatenate(col))
    return np.real(np.array(cols).T)


def convolve_bruteforce(Q, A_full):
    """fftn(Q ifftn(A)) by the convolution theorem written out: (1/N^3) sum_q Qhat(k - q) A(q),
    with Qhat the direct DFT of Q (explicit exponential sums, no FFT).  Full grids; tiny
    meshes only."""
    n = Q.shape
    W = [np.exp(-2j * np.pi * np.outer(np.arange(m), np.arange(m)) / m) for m in n]
    Qhat = np.einsum('ia,jb,kc,abc->ijk', W[0], W[1], W[2], Q)
    out = np.zeros(n, dtype=complex)
    idx = [np.arange(m) for m in n]
    for q in np.ndindex(*n):
        if A_full[q] == 0:
            continue
        shifted = Qhat[np.ix_((idx[0] - q[0]) % n[0], (idx[1] - q[1]) % n[1], (idx[2] - q[2]) % n[2])]
        out += shifted * A_full[q]
    return out / np.prod(n)


def tophat_bessel(ell, x):
    """F_ell(x) = x^-3 int_0^x t^2 j_ell(t) dt by the reference's closed forms for x > 0.1 and
    its low-x series otherwise (utils.py:403-415, 453-462); Si from scipy.special.sici
    (utils.py:370-372)."""
    x = np.asarray(x, dtype='f8')
    si = special.sici(np.where(x > 0.1, x, 1.))[0]
    c, s = np.cos(x), np.sin(x)
    with np.errstate(divide='ignore', invalid='ignore'):
        big = {0: lambda: (-c / x + s / x**2) / x,
               1: lambda: (-s - 2 * c / x) / x**2 + 2 / x**3,
               2: lambda: (x * c - 4 * s + 3 * si) / x**3,
               3: lambda: 8 / x**3 + (x**2 * s + 7 * x * c - 15 * s) / x**4,
               4: lambda: (-x**3 * c + 11 * x**2 * s + 15 * x**2 * si / 2 + 105 * x * c / 2 - 105 * s / 2) / x**5,
               5: lambda: 16 / x**3 + (-x**4 * s - 16 * x**3 * c + 105 * x**2 * s + 315 * x * c - 315 * s) / x**6}[ell]()
    small = {0: -x**10 / 518918400 + x**8 / 3991680 - x**6 / 45360 + x**4 / 840 - x**2 / 30 + 1 / 3,
             1: x**9 / 47900160 - x**7 / 453600 + x**5 / 6720 - x**3 / 180 + x / 12,
             2: x**10 / 674593920 - x**8 / 5488560 + x**6 / 68040 - x**4 / 1470 + x**2 / 75,
             3: -x**9 / 77837760 + x**7 / 831600 - x**5 / 15120 + x**3 / 630,
             4: -x**10 / 1264863600 + x**8 / 11891880 - x**6 / 187110 + x**4 / 6615,
             5: x**9 / 194594400 - x**7 / 2702700 + x**5 / 83160}[ell]
    return np.where(x > 0.1, big, small)


def tophat_shell_xi(ell, k0, k1, s):
    """BesselIntegral(edges=(k0, k1), s, ell, method='exact', mode='backward').w
    (utils.py:470-501): (-1)^(ell//2) / (2 pi)^3 * 4 pi [F(k1 s) k1^3 - F(k0 s) k0^3]."""
    return (-1)**(ell // 2) / (2. * np.pi)**3 * 4. * np.pi * (tophat_bessel(ell, k1 * s) * k1**3 - tophat_bessel(ell, k0 * s) * k0**3)


def spectrum_window_infinite(attrs, bin, edgesin, ellsin=0, meshes=None, los='z', norm=None):
    """compute_mesh2_spectrum_window(flags=('infinite',)) with a global line of sight
    (mesh2.py:1078-1092): for each input shell the theory is taken to configuration space
    analytically, A = BesselIntegral(shell, |s|, ellin).w P_ellin(mu_s) V (s the separation
    vectors, mu_s = 0 at s = 0), multiplied by Q for a selection, then binned as in
    spectrum_window after r2c."""
    n3 = float(np.prod(attrs.meshsize))
    _norm = n3 / np.prod(attrs.cellsize)
    ells = bin.ells
    norm = np.broadcast_to(_norm if norm is None else np.asarray(norm, dtype='f8'), (len(ells),))
    rnorm = _norm / norm / n3
    ellsin = (ellsin,) if np.ndim(ellsin) == 0 else tuple(ellsin)
    edgesin = np.asarray(edgesin, dtype='f8')
    if edgesin.ndim != 2:
        edgesin = np.column_stack([edgesin[:-1], edgesin[1:]])
    vlos = los_vector(los)
    Q = None
    if meshes is not None:
        cm = [m if np.iscomplexobj(m) else r2c(m) for m in meshes]
        Q = c2r(cm[0] * np.conj(cm[-1]), attrs.meshsize) / n3
    sx, sy, sz = separation_coords(attrs)
    sx, sy, sz = sx[:, None, None], sy[None, :, None], sz[None, None, :]
    sn = np.sqrt(sx**2 + sy**2 + sz**2)
    smu = (sx * vlos[0] + sy * vlos[1] + sz * vlos[2]) / np.where(sn == 0., 1., sn)
    kx, ky, kz = attrs.kcoords('wavenumber')
    kx, ky, kz = kx[:, None, None], ky[None, :, None], kz[None, None, :]
    kn = knorm(attrs)
    mu = (kx * vlos[0] + ky * vlos[1] + kz * vlos[2]) / np.where(kn == 0., 1., kn)
    volume = np.prod(attrs.boxsize)
    cols = []
    for ellin in ellsin:
        legin = legendre(ellin, smu)
        for e0, e1 in edgesin:
            A = tophat_shell_xi(ellin, e0, e1, sn) * legin * volume
            if Q is not None:
                A = A * Q
            A = r2c(A)
            col = []
            for ill, ell in enumerate(ells):
                odd = ell % 2
                leg = legendre(ell, mu) + (legendre(ell, -mu) if odd else 0.)
                col.append((2 * ell + 1) / (1 + odd) * bin(A * leg, remove_zero=True) * rnorm[ill])
            cols.append(np.concatenate(col))
    return np.real(np.array(cols).T)


def spherical_jn(ell, x):
    """get_spherical_jn (utils.py:418-450): closed forms for x > 0.1, Taylor series below."""
    x = np.asarray(x, dtype='f8')
    s, c = np.sin(x), np.cos(x)
    with np.errstate(divide='ignore', invalid='ignore'):
        big = {0: lambda: s / x,
               1: lambda: -c / x + s / x**2,
               2: lambda: (-1 / x + 3 / x**3) * s - 3 * c / x**2,
               3: lambda: (-6 / x**2 + 15 / x**4) * s + (1 / x - 15 / x**3) * c,
               4: lambda: (10 / x**2 - 105 / x**4) * c + (1 / x - 45 / x**3 + 105 / x**5) * s,
               5: lambda: (15 / x**2 - 420 / x**4 + 945 / x**6) * s + (-1 / x + 105 / x**3 - 945 / x**5) * c}[ell]()
    small = {0: -x**10 / 39916800 + x**8 / 362880 - x**6 / 5040 + x**4 / 120 - x**2 / 6 + 1,
             1: x**9 / 3991680 - x**7 / 45360 + x**5 / 840 - x**3 / 30 + x / 3,
             2: x**10 / 51891840 - x**8 / 498960 + x**6 / 7560 - x**4 / 210 + x**2 / 15,
             3: -x**9 / 6486480 + x**7 / 83160 - x**5 / 1890 + x**3 / 105,
             4: -x**10 / 97297200 + x**8 / 1081080 - x**6 / 20790 + x**4 / 945,
             5: x**9 / 16216200 - x**7 / 270270 + x**5 / 10395}[ell]
    return np.where(x > 0.1, big, small)


def spectrum_window_smooth(attrs, bin, edgesin, ellsin=0, meshes=None, los='z', norm=None):
    """compute_mesh2_spectrum_window(flags=('smooth',)) with a global line of sight
    (mesh2.py:1045-1076): the selection (Q, or 1 for a periodic box) times P_ell(mu_s)
    P_ellin(mu_s) is binned in separation shells with the default (unique) edges, the first
    shell zeroed for ell > 0 and empty shells dropped; each input shell then gives
    (2l+1) rnorm (-1)^(l//2) sum_s nmodes(s) j_l(k savg(s)) xi_ellin(savg(s)) V Qs(s) at the
    output k = xavg (0 for empty output bins)."""
    n3 = float(np.prod(attrs.meshsize))
    _norm = n3 / np.prod(attrs.cellsize)
    ells = bin.ells
    norm = np.broadcast_to(_norm if norm is None else np.asarray(norm, dtype='f8'), (len(ells),))
    rnorm = _norm / norm / n3
    ellsin = (ellsin,) if np.ndim(ellsin) == 0 else tuple(ellsin)
    edgesin = np.asarray(edgesin, dtype='f8')
    if edgesin.ndim != 2:
        edgesin = np.column_stack([edgesin[:-1], edgesin[1:]])
    vlos = los_vector(los)
    Q = 1.
    if meshes is not None:
        cm = [m if np.iscomplexobj(m) else r2c(m) for m in meshes]
        Q = c2r(cm[0] * np.conj(cm[-1]), attrs.meshsize) / n3
    sbin = BinCorrelationPoles(attrs, edges=None)
    with np.errstate(invalid='ignore'):
        kout = np.where(bin.nmodes == 0, 0., bin.xavg)
    sx, sy, sz = separation_coords(attrs)
    sx, sy, sz = sx[:, None, None], sy[None, :, None], sz[None, None, :]
    sn = np.sqrt(sx**2 + sy**2 + sz**2)
    smu = (sx * vlos[0] + sy * vlos[1] + sz * vlos[2]) / np.where(sn == 0., 1., sn)
    volume = np.prod(attrs.boxsize)
    blocks = []
    for ellin in ellsin:
        tmp = []
        for ill, ell in enumerate(ells):
            Qs = sbin(Q * legendre(ell, smu) * legendre(ellin, smu))
            if ell != 0:
                Qs[0] = 0.
            empty = sbin.nmodes == 0
            Qs = np.where(empty, 0., Qs)
            savg = np.where(empty, 0., sbin.xavg)
            rows = []
            for e0, e1 in edgesin:
                tophat_Qs = tophat_shell_xi(ellin, e0, e1, savg) * volume * Qs
                spec = (-1)**(ell // 2) * np.sum(sbin.nmodes[None, :] * spherical_jn(ell, kout[:, None] * savg[None, :]) * tophat_Qs[None, :], axis=-1)
                rows.append((2 * ell + 1) * spec * rnorm[ill])
            tmp.append(np.array(rows))
        blocks.append(np.concatenate(tmp, axis=-1))
    return np.real(np.concatenate(blocks, axis=0).T)


def radial_power(coords, power):
    """|c|^power on the mesh of the three 1-D coordinate arrays (0^p as numpy gives it)."""
    cx, cy, cz = coords
    r = np.sqrt(cx[:, None, None]**2 + cy[None, :, None]**2 + cz[None, None, :]**2)
    with np.errstate(divide='ignore'):
        return r**power


def window_local_selection(attrs, rmesh, ell, m, ellin, wa=0):
    """Q_lm(s) of the local (firstpoint) window with a configuration-space theory
    (mesh2.py:1241-1247): V_cell sum_m1 Y_lin,m1(shat) 4 pi / (2 lin + 1)
    c2r(conj(r2c(rho |x|^-wa Y_lm(xhat) Y_lin,m1(xhat))) A0) |s|^wa, A0 = r2c(rho); positions
    from the origin, Y at the zero vector as the reference (ylm_at_zero), wa the wide-angle
    order."""
    xs = attrs.rcoords()
    xx, xy, xz = xs[0][:, None, None], xs[1][None, :, None], xs[2][None, None, :]
    sx, sy, sz = separation_coords(attrs)
    sx, sy, sz = sx[:, None, None], sy[None, :, None], sz[None, None, :]
    A0 = r2c(rmesh)
    base = rmesh * real_ylm(ell, m, xx, xy, xz) * (radial_power(xs, -wa) if wa else 1.)
    spow = radial_power(separation_coords(attrs), wa) if wa else 1.
    Qs = 0.
    for m1 in range(-ellin, ellin + 1):
        Q = 4. * np.pi / (2 * ellin + 1) * c2r(np.conj(r2c(base * real_ylm(ellin, m1, xx, xy, xz))) * A0, attrs.meshsize) * spow
        Qs = Qs + real_ylm(ellin, m1, sx, sy, sz) * Q * np.prod(attrs.cellsize)
    return Qs


def window_local_selection_bruteforce(attrs, rmesh, ell, m, ellin, wa=0):
    """The same Q_lm(s) as an explicit double sum over cells x and lags s,
    V_cell sum_x rho(x) rho(x + s) Y_lm(xhat) P_lin(xhat . shat) (addition theorem; the
    reference's Y values at zero vectors, addition_at), independent of the Y_lin,m1
    conventions; tiny meshes."""
    n = [int(v) for v in attrs.meshsize]
    xs = attrs.rcoords()
    ss = separation_coords(attrs)
    X = np.stack(np.meshgrid(*xs, indexing='ij'), axis=-1)
    xn = np.linalg.norm(X, axis=-1)
    ylm = real_ylm(ell, m, X[..., 0], X[..., 1], X[..., 2])
    if wa:
        with np.errstate(divide='ignore'):
            ylm = ylm * xn**(-wa)
    out = np.zeros(n)
    for idx in np.ndindex(*n):
        svec = np.array([ss[a][idx[a]] for a in range(3)])
        sn = np.linalg.norm(svec)
        leg = addition_at(ellin, X, svec)
        shifted = np.roll(rmesh, [-i for i in idx], axis=(0, 1, 2))  # rho(x + s)
        out[idx] = np.sum(rmesh * shifted * ylm * leg) * sn**wa
    return out * np.prod(attrs.cellsize)


def spectrum_window_local_infinite(attrs, bin, edgesin, ellsin=0, mesh=None, norm=None):
    """compute_mesh2_spectrum_window(flags=('infinite',)) with the local (firstpoint) line of
    sight and theory line of sight 'firstpoint', wide-angle order 0 (mesh2.py:1235-1260): for each
    input multipole lin, output (l, m) and input shell, 4 pi bin(Y_lm(khat) r2c(xi_lin(|s|) Q_lm(s)),
    remove_zero = (l == 0)) rnorm, summed over m; mesh None is the periodic box (rho = 1).
    Returns the (n_ells n_bins, n_ellsin n_in) real matrix."""
    n3 = float(np.prod(attrs.meshsize))
    _norm = n3 / np.prod(attrs.cellsize)
    ells = bin.ells
    norm = np.broadcast_to(_norm if norm is None else np.asarray(norm, dtype='f8'), (len(ells),))
    rnorm = _norm / norm / n3
    ellsin = (ellsin,) if np.ndim(ellsin) == 0 else tuple(ellsin)
    edgesin = np.asarray(edgesin, dtype='f8')
    if edgesin.ndim != 2:
        edgesin = np.column_stack([edgesin[:-1], edgesin[1:]])
    rmesh = np.ones(tuple(int(v) for v in attrs.meshsize)) if mesh is None else (c2r(mesh, attrs.meshsize) if np.iscomplexobj(mesh) else mesh)
    sx, sy, sz = separation_coords(attrs)
    sn = np.sqrt(sx[:, None, None]**2 + sy[None, :, None]**2 + sz[None, None, :]**2)
    kx, ky, kz = attrs.kcoords('wavenumber')
    kx, ky, kz = kx[:, None, None], ky[None, :, None], kz[None, None, :]
    nb = len(bin.xavg)
    blocks = []
    for ellin in ellsin:
        ellin, wa = ellin if isinstance(ellin, tuple) else (ellin, 0)
        W = np.zeros((len(edgesin), len(ells) * nb))
        tophats = [tophat_shell_xi(ellin, e0, e1, sn) for e0, e1 in edgesin]
        for ill, ell in enumerate(ells):
            for m in range(-ell, ell + 1):
                Qs = window_local_selection(attrs, rmesh, ell, m, ellin, wa=wa)
                ykm = real_ylm(ell, m, kx, ky, kz)
                for ie, tophat in enumerate(tophats):
                    spec = 4. * np.pi * bin(ykm * r2c(tophat * Qs), remove_zero=ell == 0) * rnorm[ill]
                    W[ie, ill * nb:(ill + 1) * nb] += np.real(spec)
        blocks.append(W)
    return np.concatenate(blocks, axis=0).T


def window_local_terms(attrs, rmesh, ell, m, ellin, e0, e1, wa=0):
    """A_lm(s) = sum_m1 xi_m1(s) Q_lm,m1(s) of the local (firstpoint) window with the theory on the
    mesh (mesh2.py:1267-1281, wide-angle order 0): xi_m1 = c2r([e0 <= |k| <= e1] Y_lin,m1(khat))
    and Q_lm,m1 = 4 pi / (2 lin + 1) c2r(conj(r2c(rho Y_lm(xhat) Y_lin,m1(xhat))) A0)."""
    xs = attrs.rcoords()
    xx, xy, xz = xs[0][:, None, None], xs[1][None, :, None], xs[2][None, None, :]
    kx, ky, kz = attrs.kcoords('wavenumber')
    kx, ky, kz = kx[:, None, None], ky[None, :, None], kz[None, None, :]
    kn = knorm(attrs)
    shell = (kn >= e0) & (kn <= e1)
    A0 = r2c(rmesh)
    base = rmesh * real_ylm(ell, m, xx, xy, xz) * (radial_power(xs, -wa) if wa else 1.)
    spow = radial_power(separation_coords(attrs), wa) if wa else 1.
    out = 0.
    for m1 in range(-ellin, ellin + 1):
        xi = c2r(shell * real_ylm(ellin, m1, kx, ky, kz), attrs.meshsize)
        Q = 4. * np.pi / (2 * ellin + 1) * c2r(np.conj(r2c(base * real_ylm(ellin, m1, xx, xy, xz))) * A0, attrs.meshsize) * spow
        out = out + xi * Q
    return out


def window_local_terms_bruteforce(attrs, rmesh, ell, m, ellin, e0, e1):
    """The same A_lm(s) from the addition theorem, without any Y_lin,m1:
    N^-3 sum_x sum_k rho(x) rho(x + s) Y_lm(xhat) [e0 <= |k| <= e1] P_lin(xhat . khat) cos(k . s)
    over the full k grid (the shell and P_lin are even in k), the reference's Y values at zero
    vectors (addition_at); tiny meshes only."""
    n = [int(v) for v in attrs.meshsize]
    xs = attrs.rcoords()
    X = np.stack(np.meshgrid(*xs, indexing='ij'), axis=-1).reshape(-1, 3)
    ylm = real_ylm(ell, m, X[:, 0], X[:, 1], X[:, 2])
    kk = [np.fft.fftfreq(n[a]) * attrs.kfun[a] * n[a] for a in range(3)]
    K = np.stack(np.meshgrid(*kk, indexing='ij'), axis=-1).reshape(-1, 3)
    kn = np.linalg.norm(K, axis=-1)
    # the half-spectrum shell uses |k| summed x, y, z as knorm does; on the full grid that is the same
    sel = (kn >= e0) & (kn <= e1)
    K, kn = K[sel], kn[sel]
    leg = addition_at(ellin, X[:, None, :], K[None, :, :])  # (x, k)
    idx = [np.arange(v) for v in n]
    grid = np.stack(np.meshgrid(*idx, indexing='ij'), axis=-1).reshape(-1, 3)
    ss = [np.fft.fftfreq(n[a]) * attrs.cellsize[a] * n[a] for a in range(3)]
    flat = rmesh.ravel()
    out = np.zeros(n)
    for sidx in np.ndindex(*n):
        svec = np.array([ss[a][sidx[a]] for a in range(3)])
        shifted = np.roll(rmesh, [-i for i in sidx], axis=(0, 1, 2)).ravel()  # rho(x + s)
        w = flat * shifted * ylm  # (x,)
        out[sidx] = np.sum((w @ leg) * np.cos(K @ svec)) / np.prod(n)
    return out


def spectrum_window_local(attrs, bin, edgesin, ellsin=0, mesh=None, norm=None):
    """compute_mesh2_spectrum_window without flags, local (firstpoint) line of sight, theory line
    of sight 'firstpoint', wide-angle order 0 (mesh2.py:1262-1305): for each input multipole lin
    and closed input shell [e0, e1], sum over (l, m) of 4 pi bin(r2c(A_lm) Y_lm(khat),
    remove_zero = (l == 0)) rnorm.  mesh None is the periodic box (rho = 1)."""
    n3 = float(np.prod(attrs.meshsize))
    _norm = n3 / np.prod(attrs.cellsize)
    ells = bin.ells
    norm = np.broadcast_to(_norm if norm is None else np.asarray(norm, dtype='f8'), (len(ells),))
    rnorm = _norm / norm / n3
    ellsin = (ellsin,) if np.ndim(ellsin) == 0 else tuple(ellsin)
    edgesin = np.asarray(edgesin, dtype='f8')
    if edgesin.ndim != 2:
        edgesin = np.column_stack([edgesin[:-1], edgesin[1:]])
    rmesh = np.ones(tuple(int(v) for v in attrs.meshsize)) if mesh is None else (c2r(mesh, attrs.meshsize) if np.iscomplexobj(mesh) else mesh)
    kx, ky, kz = attrs.kcoords('wavenumber')
    kx, ky, kz = kx[:, None, None], ky[None, :, None], kz[None, None, :]
    nb = len(bin.xavg)
    blocks = []
    for ellin in ellsin:
        ellin, wa = ellin if isinstance(ellin, tuple) else (ellin, 0)
        W = np.zeros((len(edgesin), len(ells) * nb))
        for ie, (e0, e1) in enumerate(edgesin):
            for ill, ell in enumerate(ells):
                acc = 0.
                for m in range(-ell, ell + 1):
                    acc = acc + r2c(window_local_terms(attrs, rmesh, ell, m, ellin, e0, e1, wa=wa)) * real_ylm(ell, m, kx, ky, kz)
                W[ie, ill * nb:(ill + 1) * nb] = np.real(4. * np.pi * bin(acc, remove_zero=ell == 0) * rnorm[ill])
        blocks.append(W)
    return np.concatenate(blocks, axis=0).T


def window_local_smooth_terms(attrs, rmesh, ell, ellin, wa=0):
    """sum_{m, m1} 4 pi / (2 lin + 1) c2r(conj(r2c(rho Y_lm Y_lin,m1)) A0)(s) Y_lm(shat) Y_lin,m1(shat)
    of the local smooth window (mesh2.py:1205-1210, wide-angle order 0), before the separation
    binning, the 4 pi and V_cell."""
    xs = attrs.rcoords()
    xx, xy, xz = xs[0][:, None, None], xs[1][None, :, None], xs[2][None, None, :]
    sx, sy, sz = separation_coords(attrs)
    sx, sy, sz = sx[:, None, None], sy[None, :, None], sz[None, None, :]
    A0 = r2c(rmesh)
    xpow = radial_power(xs, -wa) if wa else 1.
    spow = radial_power(separation_coords(attrs), wa) if wa else 1.
    out = 0.
    for m1 in range(-ellin, ellin + 1):
        for m in range(-ell, ell + 1):
            Q = 4. * np.pi / (2 * ellin + 1) * c2r(np.conj(r2c(rmesh * xpow * real_ylm(ell, m, xx, xy, xz) * real_ylm(ellin, m1, xx, xy, xz))) * A0,
                                                  attrs.meshsize) * spow
            out = out + Q * real_ylm(ell, m, sx, sy, sz) * real_ylm(ellin, m1, sx, sy, sz)
    return out


def window_local_smooth_terms_bruteforce(attrs, rmesh, ell, ellin):
    """The same sum from the addition theorem: (2l+1)/(4 pi) sum_x rho(x) rho(x+s) P_l(xhat.shat)
    P_lin(xhat.shat), with the reference's Y values at zero vectors (addition_at); tiny
    meshes."""
    n = [int(v) for v in attrs.meshsize]
    xs = attrs.rcoords()
    ss = separation_coords(attrs)
    X = np.stack(np.meshgrid(*xs, indexing='ij'), axis=-1)
    out = np.zeros(n)
    for idx in np.ndindex(*n):
        svec = np.array([ss[a][idx[a]] for a in range(3)])
        f = addition_at(ell, X, svec) * addition_at(ellin, X, svec)
        shifted = np.roll(rmesh, [-i for i in idx], axis=(0, 1, 2))
        out[idx] = (2 * ell + 1) / (4. * np.pi) * np.sum(rmesh * shifted * f)
    return out


def spectrum_window_local_smooth(attrs, bin, edgesin, ellsin=0, mesh=None, norm=None):
    """compute_mesh2_spectrum_window(flags=('smooth',)) with a local (firstpoint) line of sight,
    theory line of sight 'firstpoint', wide-angle order 0 (mesh2.py:1195-1233): Qs(s) =
    4 pi V_cell sbin(sum_{m, m1} Q_lm,m1 Y_lm(shat) Y_lin,m1(shat)), first shell zeroed for l > 0,
    then (-1)^(l//2) sum_s nmodes(s) j_l(k savg) xi_lin(savg) Qs(s) rnorm per input shell."""
    n3 = float(np.prod(attrs.meshsize))
    _norm = n3 / np.prod(attrs.cellsize)
    ells = bin.ells
    norm = np.broadcast_to(_norm if norm is None else np.asarray(norm, dtype='f8'), (len(ells),))
    rnorm = _norm / norm / n3
    ellsin = (ellsin,) if np.ndim(ellsin) == 0 else tuple(ellsin)
    edgesin = np.asarray(edgesin, dtype='f8')
    if edgesin.ndim != 2:
        edgesin = np.column_stack([edgesin[:-1], edgesin[1:]])
    rmesh = np.ones(tuple(int(v) for v in attrs.meshsize)) if mesh is None else (c2r(mesh, attrs.meshsize) if np.iscomplexobj(mesh) else mesh)
    sbin = BinCorrelationPoles(attrs, edges=None)
    with np.errstate(invalid='ignore'):
        kout = np.where(bin.nmodes == 0, 0., bin.xavg)
    empty = sbin.nmodes == 0
    savg = np.where(empty, 0., sbin.xavg)
    nb = len(bin.xavg)
    blocks = []
    for ellin in ellsin:
        ellin, wa = ellin if isinstance(ellin, tuple) else (ellin, 0)
        W = np.zeros((len(edgesin), len(ells) * nb))
        for ill, ell in enumerate(ells):
            Qs = 4. * np.pi * sbin(window_local_smooth_terms(attrs, rmesh, ell, ellin, wa=wa)) * np.prod(attrs.cellsize)
            if ell != 0:
                Qs[0] = 0.
            Qs = np.where(empty, 0., Qs)
            for ie, (e0, e1) in enumerate(edgesin):
                tq = tophat_shell_xi(ellin, e0, e1, savg) * Qs
                spec = (-1)**(ell // 2) * np.sum(sbin.nmodes[None, :] * spherical_jn(ell, kout[:, None] * savg[None, :]) * tq[None, :], axis=-1)
                W[ie, ill * nb:(ill + 1) * nb] = spec * rnorm[ill]
        blocks.append(W)
    return np.concatenate(blocks, axis=0).T


def theory_on_mesh(attrs, kin, values, edges):
    """Interpolator1D(kin, |k|, edges=edges) of order 0 applied to values (utils.py:507-537):
    piecewise constant by digitize(right=False) on the edges (points: mid-point edges with the
    ends extrapolated), 0 outside."""
    kin = np.asarray(kin, dtype='f8')
    if not edges:
        tmp = (kin[:-1] + kin[1:]) / 2.
        kin = np.concatenate([[tmp[0] - (kin[1] - kin[0])], tmp, [tmp[-1] + (kin[-1] - kin[-2])]])
    idx = np.digitize(knorm(attrs), kin, right=False) - 1
    mask = (idx >= 0) & (idx <= len(kin) - 2)
    return np.where(mask, np.asarray(values, dtype='f8')[np.where(mask, idx, 0)], 0.)


def spectrum_mean(attrs, bin, kin, poles, edges=True, mesh=None, los='z', theory_los='firstpoint'):
    """compute_mesh2_spectrum_mean (mesh2.py:1485-1680) for a theory tabulated per multipole
    (poles: dict ell -> values on kin), wide-angle order 0: global line of sight
    num_l = (2l+1) bin(r2c(Q c2r(P rnorm)) P_l(mu)) with P = sum_lin P_lin(|k|) L_lin(mu),
    Q = c2r(A conj A) (times N^3 for the periodic box, mesh None); local (firstpoint) line of
    sight num_l = 4 pi bin(sum_m r2c(Q_m) Y_lm(khat)) with
    Q_m = sum_{lin, m1} 4 pi/(2 lin + 1) c2r(P_lin rnorm Y_lin,m1(khat)) c2r(conj(r2c(rho Y_lm Y_lin,m1)) A0).
    Returns num_raw per output ell (norm = N^3 / V_cell)."""
    n3 = float(np.prod(attrs.meshsize))
    rnorm = 1. / np.prod(attrs.cellsize)
    kx, ky, kz = attrs.kcoords('wavenumber')
    kx, ky, kz = kx[:, None, None], ky[None, :, None], kz[None, None, :]
    kn = knorm(attrs)
    rmesh = np.ones(tuple(int(v) for v in attrs.meshsize)) if mesh is None else (c2r(mesh, attrs.meshsize) if np.iscomplexobj(mesh) else mesh)
    tables = {(key if isinstance(key, tuple) else (key, 0)): theory_on_mesh(attrs, kin, values, edges) for key, values in poles.items()}
    num = []
    if not (isinstance(los, str) and los in ('firstpoint', 'local', 'endpoint')):
        vlos = los_vector(los)
        mu = (kx * vlos[0] + ky * vlos[1] + kz * vlos[2]) / np.where(kn == 0., 1., kn)
        A = sum(tables[key] * legendre(key[0], mu) for key in tables) * rnorm
        if mesh is None:
            A = A * n3
        else:
            A1 = r2c(rmesh)
            A = r2c(c2r(A1 * np.conj(A1), attrs.meshsize) * c2r(A, attrs.meshsize))
        for ell in bin.ells:
            odd = ell % 2
            leg = legendre(ell, mu) + (legendre(ell, -mu) if odd else 0.)
            num.append(np.real((2 * ell + 1) / (1 + odd) * bin(A * leg)))
        return num
    xs = attrs.rcoords()
    xx, xy, xz = xs[0][:, None, None], xs[1][None, :, None], xs[2][None, None, :]
    A0 = r2c(rmesh)
    if theory_los == 'local':  # mesh2.py:1647-1662 (one selection mesh: rho2 = rho1)
        for ell in bin.ells:
            acc = 0.
            for m in range(-ell, ell + 1):
                Q = 0.
                for ell1, ell2 in product((0, 2), (0, 2)):
                    kernel = sum(c * tables[l, 0] for l, c in LOCAL_THEORY_COEFF2[ell1, ell2]) * rnorm
                    for m1 in range(-ell1, ell1 + 1):
                        F = r2c(rmesh * real_ylm(ell, m, xx, xy, xz) * real_ylm(ell1, m1, xx, xy, xz))
                        for m2 in range(-ell2, ell2 + 1):
                            xi = c2r(kernel * real_ylm(ell1, m1, kx, ky, kz) * real_ylm(ell2, m2, -kx, -ky, -kz), attrs.meshsize)
                            Q = Q + (4. * np.pi)**2 / ((2 * ell1 + 1) * (2 * ell2 + 1)) * xi * c2r(np.conj(F) * r2c(rmesh * real_ylm(ell2, m2, xx, xy, xz)),
                                                                                                    attrs.meshsize)
                acc = acc + r2c(Q) * real_ylm(ell, m, kx, ky, kz)
            num.append(np.real(4. * np.pi * bin(acc)))
        return num
    for ell in bin.ells:
        acc = 0.
        for m in range(-ell, ell + 1):
            Q = 0.
            for (ell1, wa), table in tables.items():
                xpow = radial_power(xs, -wa) if wa else 1.
                spow = radial_power(separation_coords(attrs), wa) if wa else 1.
                for m1 in range(-ell1, ell1 + 1):
                    xi = c2r(table * rnorm * real_ylm(ell1, m1, kx, ky, kz), attrs.meshsize) * spow
                    Q = Q + 4. * np.pi / (2 * ell1 + 1) * xi * c2r(np.conj(r2c(rmesh * xpow * real_ylm(ell, m, xx, xy, xz) * real_ylm(ell1, m1, xx, xy, xz))) * A0,
                                                                 attrs.meshsize)
            acc = acc + r2c(Q) * real_ylm(ell, m, kx, ky, kz)
        num.append(np.real(4. * np.pi * bin(acc)))
    return num


@lru_cache(maxsize=None)
def real_gaunt(l1, m1, l2, m2, l3, m3):
    """int Y_l1,m1 Y_l2,m2 Y_l3,m3 dOmega with the oracle's real_ylm (utils.py:540-600 computes it
    with sympy in the reference's convention, which differs per m by (-1)^|m|); Gauss-Legendre in
    cos(theta) times a uniform rule in phi, exact for l1 + l2 + l3 <= 12."""
    x, w = np.polynomial.legendre.leggauss(16)
    phi = 2. * np.pi * np.arange(32) / 32
    st = np.sqrt(1. - x**2)
    ux, uy, uz = st[:, None] * np.cos(phi)[None, :], st[:, None] * np.sin(phi)[None, :], np.broadcast_to(x[:, None], (16, 32))
    f = real_ylm(l1, m1, ux, uy, uz) * real_ylm(l2, m2, ux, uy, uz) * real_ylm(l3, m3, ux, uy, uz)
    return float(np.sum(w[:, None] * f) * 2. * np.pi / 32)


LOCAL_THEORY_COEFF2 = {(0, 0): [(0, 1.), (4, -7. / 18.)], (0, 2): [(2, 1. / 2.), (4, -5. / 18.)], (2, 2): [(4, 35. / 18.)]}
LOCAL_THEORY_COEFF2[2, 0] = LOCAL_THEORY_COEFF2[0, 2]


def spectrum_window_local_theory_infinite(attrs, bin, edgesin, ellsin=(0, 2, 4), mesh=None, mesh2=None, norm=None):
    """compute_mesh2_spectrum_window(flags=('infinite',)) with a local line of sight and theory line
    of sight 'local' (mesh2.py:1382-1413, 1446-1457): for (l1, l2) in {2, 0}^2 and output (l, m),
    Qs_p = V_cell sum_{m1, m2, mp} G(p mp; l1 m1; l2 m2) Y_p,mp(shat)
    (4 pi)^2/((2 l1 + 1)(2 l2 + 1)) c2r(conj(r2c(rho1 Y_lm Y_l1,m1)) r2c(rho2 Y_l2,m2)), then per
    input shell 4 pi bin(Y_lm(khat) r2c(sum_p xi_p(|s|) Qs_p), remove_zero = (l == 0)) rnorm, and
    the input multipoles are the combinations of mesh2.py:1449-1457."""
    n3 = float(np.prod(attrs.meshsize))
    _norm = n3 / np.prod(attrs.cellsize)
    ells = bin.ells
    norm = np.broadcast_to(_norm if norm is None else np.asarray(norm, dtype='f8'), (len(ells),))
    rnorm = _norm / norm / n3
    edgesin = np.asarray(edgesin, dtype='f8')
    if edgesin.ndim != 2:
        edgesin = np.column_stack([edgesin[:-1], edgesin[1:]])
    ones = np.ones(tuple(int(v) for v in attrs.meshsize))
    rho1 = ones if mesh is None else mesh
    rho2 = rho1 if mesh2 is None else mesh2
    xs = attrs.rcoords()
    xx, xy, xz = xs[0][:, None, None], xs[1][None, :, None], xs[2][None, None, :]
    sx, sy, sz = separation_coords(attrs)
    sx, sy, sz = sx[:, None, None], sy[None, :, None], sz[None, None, :]
    sn = np.sqrt(sx**2 + sy**2 + sz**2)
    kx, ky, kz = attrs.kcoords('wavenumber')
    kx, ky, kz = kx[:, None, None], ky[None, :, None], kz[None, None, :]
    nb = len(bin.xavg)
    tmp = {}
    for ell1, ell2 in product((2, 0), (2, 0)):
        W = np.zeros((len(edgesin), len(ells) * nb))
        G2 = {m2: r2c(rho2 * real_ylm(ell2, m2, xx, xy, xz)) for m2 in range(-ell2, ell2 + 1)}
        for ill, ell in enumerate(ells):
            ps = [p for p in (0, 2, 4) if abs(real_gaunt(p, 0, ell1, 0, ell2, 0)) > 1e-10]
            for m in range(-ell, ell + 1):
                Qs = {p: 0. for p in ps}
                for m1 in range(-ell1, ell1 + 1):
                    F = r2c(rho1 * real_ylm(ell, m, xx, xy, xz) * real_ylm(ell1, m1, xx, xy, xz))
                    for m2 in range(-ell2, ell2 + 1):
                        Q = (4. * np.pi)**2 / ((2 * ell1 + 1) * (2 * ell2 + 1)) * c2r(np.conj(F) * G2[m2], attrs.meshsize)
                        for p in ps:
                            for mp in range(-p, p + 1):
                                rg = real_gaunt(p, mp, ell1, m1, ell2, m2)
                                if abs(rg) > 1e-10:
                                    Qs[p] = Qs[p] + rg * real_ylm(p, mp, sx, sy, sz) * Q
                ykm = real_ylm(ell, m, kx, ky, kz)
                for ie, (e0, e1) in enumerate(edgesin):
                    xi = sum(tophat_shell_xi(p, e0, e1, sn) * Qs[p] * np.prod(attrs.cellsize) for p in ps)
                    spec = 4. * np.pi * bin(ykm * r2c(xi), remove_zero=ell == 0) * rnorm[ill]
                    W[ie, ill * nb:(ill + 1) * nb] += np.real(spec)
        tmp[ell1, ell2] = W
    out = np.zeros((len(ellsin), len(edgesin), len(ells) * nb))
    for illin, ellin in enumerate(ellsin):
        for key, lst in LOCAL_THEORY_COEFF2.items():
            coeff = sum(c * (ell == ellin) for ell, c in lst)
            out[illin] += coeff * tmp[key]
    return out.reshape(len(ellsin) * len(edgesin), -1).T


def spectrum_window_local_theory_smooth(attrs, bin, edgesin, ellsin=(0, 2, 4), mesh=None, mesh2=None, norm=None):
    """compute_mesh2_spectrum_window(flags=('smooth',)) with a local line of sight and theory line of
    sight 'local' (mesh2.py:1346-1380, 1446-1457): per (l1, l2) and output l,
    Q_p = 4 pi V_cell sbin(sum_{m, m1, m2, mp} G(p mp; l1 m1; l2 m2) Y_lm(shat) Y_p,mp(shat)
    (4 pi)^2/((2 l1 + 1)(2 l2 + 1)) c2r(conj(r2c(rho1 Y_lm Y_l1,m1)) r2c(rho2 Y_l2,m2))), first shell
    zeroed for l > 0, Hankel-summed with xi_p of each input shell, then combined into the input
    multipoles."""
    n3 = float(np.prod(attrs.meshsize))
    _norm = n3 / np.prod(attrs.cellsize)
    ells = bin.ells
    norm = np.broadcast_to(_norm if norm is None else np.asarray(norm, dtype='f8'), (len(ells),))
    rnorm = _norm / norm / n3
    edgesin = np.asarray(edgesin, dtype='f8')
    if edgesin.ndim != 2:
        edgesin = np.column_stack([edgesin[:-1], edgesin[1:]])
    ones = np.ones(tuple(int(v) for v in attrs.meshsize))
    rho1 = ones if mesh is None else mesh
    rho2 = rho1 if mesh2 is None else mesh2
    xs = attrs.rcoords()
    xx, xy, xz = xs[0][:, None, None], xs[1][None, :, None], xs[2][None, None, :]
    sx, sy, sz = separation_coords(attrs)
    sx, sy, sz = sx[:, None, None], sy[None, :, None], sz[None, None, :]
    sbin = BinCorrelationPoles(attrs, edges=None)
    with np.errstate(invalid='ignore'):
        kout = np.where(bin.nmodes == 0, 0., bin.xavg)
    empty = sbin.nmodes == 0
    savg = np.where(empty, 0., sbin.xavg)
    nb = len(bin.xavg)
    tmp = {}
    for ell1, ell2 in product((2, 0), (2, 0)):
        W = np.zeros((len(edgesin), len(ells) * nb))
        G2 = {m2: r2c(rho2 * real_ylm(ell2, m2, xx, xy, xz)) for m2 in range(-ell2, ell2 + 1)}
        for ill, ell in enumerate(ells):
            ps = [p for p in (0, 2, 4) if abs(real_gaunt(p, 0, ell1, 0, ell2, 0)) > 1e-10]
            acc = {p: 0. for p in ps}
            for m in range(-ell, ell + 1):
                for m1 in range(-ell1, ell1 + 1):
                    F = r2c(rho1 * real_ylm(ell, m, xx, xy, xz) * real_ylm(ell1, m1, xx, xy, xz))
                    for m2 in range(-ell2, ell2 + 1):
                        Q = (4. * np.pi)**2 / ((2 * ell1 + 1) * (2 * ell2 + 1)) * c2r(np.conj(F) * G2[m2], attrs.meshsize)
                        for p in ps:
                            for mp in range(-p, p + 1):
                                rg = real_gaunt(p, mp, ell1, m1, ell2, m2)
                                if abs(rg) > 1e-10:
                                    acc[p] = acc[p] + rg * real_ylm(ell, m, sx, sy, sz) * real_ylm(p, mp, sx, sy, sz) * Q
            for p in ps:
                Qp = 4. * np.pi * sbin(acc[p]) * np.prod(attrs.cellsize)
                if ell != 0:
                    Qp[0] = 0.
                Qp = np.where(empty, 0., Qp)
                for ie, (e0, e1) in enumerate(edgesin):
                    tq = tophat_shell_xi(p, e0, e1, savg) * Qp
                    spec = (-1)**(ell // 2) * np.sum(sbin.nmodes[None, :] * spherical_jn(ell, kout[:, None] * savg[None, :]) * tq[None, :], axis=-1)
                    W[ie, ill * nb:(ill + 1) * nb] += spec * rnorm[ill]
        tmp[ell1, ell2] = W
    out = np.zeros((len(ellsin), len(edgesin), len(ells) * nb))
    for illin, ellin in enumerate(ellsin):
        for key, lst in LOCAL_THEORY_COEFF2.items():
            out[illin] += sum(c * (ell == ellin) for ell, c in lst) * tmp[key]
    return out.reshape(len(ellsin) * len(edgesin), -1).T


def spectrum_window_local_theory(attrs, bin, edgesin, ellsin=(0, 2, 4), mesh=None, mesh2=None, norm=None):
    """compute_mesh2_spectrum_window without flags, local line of sight, theory line of sight
    'local' (mesh2.py:1415-1457): per (l1, l2) and closed input shell [e0, e1],
    A_lm = sum_{m1, m2} c2r(shell Y_l1,m1(khat) Y_l2,m2(-khat)) (4 pi)^2/((2 l1 + 1)(2 l2 + 1))
    c2r(conj(r2c(rho1 Y_lm Y_l1,m1)) r2c(rho2 Y_l2,m2)), then 4 pi bin(sum_m r2c(A_lm) Y_lm(khat),
    remove_zero = (l == 0)) rnorm, combined into the input multipoles."""
    n3 = float(np.prod(attrs.meshsize))
    _norm = n3 / np.prod(attrs.cellsize)
    ells = bin.ells
    norm = np.broadcast_to(_norm if norm is None else np.asarray(norm, dtype='f8'), (len(ells),))
    rnorm = _norm / norm / n3
    edgesin = np.asarray(edgesin, dtype='f8')
    if edgesin.ndim != 2:
        edgesin = np.column_stack([edgesin[:-1], edgesin[1:]])
    ones = np.ones(tuple(int(v) for v in attrs.meshsize))
    rho1 = ones if mesh is None else mesh
    rho2 = rho1 if mesh2 is None else mesh2
    xs = attrs.rcoords()
    xx, xy, xz = xs[0][:, None, None], xs[1][None, :, None], xs[2][None, None, :]
    kx, ky, kz = attrs.kcoords('wavenumber')
    kx, ky, kz = kx[:, None, None], ky[None, :, None], kz[None, None, :]
    kn = knorm(attrs)
    nb = len(bin.xavg)
    tmp = {}
    for ell1, ell2 in product((2, 0), (2, 0)):
        W = np.zeros((len(edgesin), len(ells) * nb))
        G2 = {m2: r2c(rho2 * real_ylm(ell2, m2, xx, xy, xz)) for m2 in range(-ell2, ell2 + 1)}
        Q = {}
        for ell in ells:
            for m in range(-ell, ell + 1):
                for m1 in range(-ell1, ell1 + 1):
                    F = r2c(rho1 * real_ylm(ell, m, xx, xy, xz) * real_ylm(ell1, m1, xx, xy, xz))
                    for m2 in range(-ell2, ell2 + 1):
                        Q[ell, m, m1, m2] = (4. * np.pi)**2 / ((2 * ell1 + 1) * (2 * ell2 + 1)) * c2r(np.conj(F) * G2[m2], attrs.meshsize)
        for ie, (e0, e1) in enumerate(edgesin):
            shell = (kn >= e0) & (kn <= e1)
            xi = {(m1, m2): c2r(shell * real_ylm(ell1, m1, kx, ky, kz) * real_ylm(ell2, m2, -kx, -ky, -kz), attrs.meshsize)
                  for m1 in range(-ell1, ell1 + 1) for m2 in range(-ell2, ell2 + 1)}
            for ill, ell in enumerate(ells):
                acc = 0.
                for m in range(-ell, ell + 1):
                    A = sum(xi[m1, m2] * Q[ell, m, m1, m2] for (m1, m2) in xi)
                    acc = acc + r2c(A) * real_ylm(ell, m, kx, ky, kz)
                W[ie, ill * nb:(ill + 1) * nb] = np.real(4. * np.pi * bin(acc, remove_zero=ell == 0) * rnorm[ill])
        tmp[ell1, ell2] = W
    out = np.zeros((len(ellsin), len(edgesin), len(ells) * nb))
    for illin, ellin in enumerate(ellsin):
        for key, lst in LOCAL_THEORY_COEFF2.items():
            out[illin] += sum(c * (ell == ellin) for ell, c in lst) * tmp[key]
    return out.reshape(len(ellsin) * len(edgesin), -1).T


def bispectrum_shotnoise(attrs, bin, positions, weights=None, resampler='cic', interlacing=0, compensate=True, los='x'):
    """compute_fkp3_shotnoise, Scoccimarro basis, one field, global line of sight
    (mesh3.py:999-1062, _compute_scoccimarro_S 786-852, eq. 58 of arXiv:1506.02729):
    A = paint(w) conj(paint(w^2)) with both k = 0 modes removed; l = 0:
    sum_axis bin_axis(A)[b_axis] - 2 sum w^3; l > 0: (2l+1) [bin_3(A L_l)[b3] +
    4 pi sum_m sum_{axis < 3} bin_axis(A L_l Y_lm)[b_axis] bin_3(Y_lm)[b3]], where bin_axis is the
    hermitian-weighted shell mean of side axis.  Returns one array per ell (real for even ell).
    los='local' replaces P_l(mu) c by H_l(c) (mesh3.py:823-834), in the first line of eq. 58 on
    paint(w) and in the second on paint(w^2); H_l is built from bispectrum_local_mesh, pinned by
    a direct DFT, and the rest is the global-line-of-sight code."""
    weights = np.ones(len(positions)) if weights is None else np.asarray(weights, dtype='f8')
    cw = paint(attrs, positions, weights, resampler=resampler, interlacing=interlacing, compensate=compensate, out='complex')
    cw2 = paint(attrs, positions, weights**2, resampler=resampler, interlacing=interlacing, compensate=compensate, out='complex')
    cw[0, 0, 0] = 0.
    cw2[0, 0, 0] = 0.
    sumw3 = np.sum(weights**3)
    A = cw * np.conj(cw2)
    local = isinstance(los, str) and los == 'local'
    vlos = los_vector('z' if local else los)
    kx, ky, kz = attrs.kcoords('wavenumber')
    kx, ky, kz = kx[:, None, None], ky[None, :, None], kz[None, None, :]
    kn = np.sqrt(kx**2 + ky**2 + kz**2)
    mu = (kx * vlos[0] + ky * vlos[1] + kz * vlos[2]) / np.where(kn == 0., 1., kn)
    out = []
    for ell in bin.ells:
        if ell == 0:
            sn = sum(bin.axes[a](A)[bin.iedges[:, a]] for a in range(3)) - 2. * sumw3
        else:
            if local:  # mesh3.py:823-834: P_l(mu) c -> H_l(c) = 4 pi / (2l+1) sum_m r2c(c2r(c) Y_lm(xhat)) Y_lm(khat)
                h = 4. * np.pi / (2 * ell + 1)
                sn = bin.axes[2](h * bispectrum_local_mesh(attrs, c2r(cw, attrs.meshsize), ell) * np.conj(cw2))[bin.iedges[:, 2]]
                Al = cw * np.conj(h * bispectrum_local_mesh(attrs, c2r(cw2, attrs.meshsize), ell))
            else:
                Al = A * legendre(ell, mu)
                sn = bin.axes[2](Al)[bin.iedges[:, 2]]
            for m in range(-ell, ell + 1):
                y = real_ylm(ell, m, kx, ky, kz)
                b3 = bin.axes[2](y * np.ones_like(Al))[bin.iedges[:, 2]]
                sn = sn + 4. * np.pi * sum(bin.axes[a](Al * y)[bin.iedges[:, a]] for a in range(2)) * b3
            sn = (2 * ell + 1) * sn
        out.append(np.real(sn) if ell % 2 == 0 else np.imag(sn))
    return out


def bispectrum_local_mesh(attrs, rho, ell):
    """Third mesh of the local-line-of-sight Scoccimarro bispectrum (mesh3.py:563-572):
    sum_m r2c(rho Y_lm(xhat)) Y_lm(khat), positions from the origin, Y at zero vectors as the
    reference (ylm_at_zero): Y_00 is the constant 1/sqrt(4 pi) everywhere, so the l = 0 mesh is
    r2c(rho) / (4 pi), k = 0 included."""
    if ell == 0:
        return r2c(rho) / (4. * np.pi)
    xs = attrs.rcoords()
    xx, xy, xz = xs[0][:, None, None], xs[1][None, :, None], xs[2][None, None, :]
    kx, ky, kz = attrs.kcoords('wavenumber')
    kx, ky, kz = kx[:, None, None], ky[None, :, None], kz[None, None, :]
    return sum(r2c(rho * real_ylm(ell, m, xx, xy, xz)) * real_ylm(ell, m, kx, ky, kz) for m in range(-ell, ell + 1))


def bispectrum_local_mesh_bruteforce(attrs, rho, ell):
    """The same mesh as a direct DFT with the addition theorem:
    sum_n rho(n) exp(-2 pi i k.n / N) (2l+1)/(4 pi) P_l(xhat_n . khat), with the reference's Y
    values at zero vectors (addition_at); tiny meshes only."""
    n = [int(v) for v in attrs.meshsize]
    xs = attrs.rcoords()
    X = np.stack(np.meshgrid(*xs, indexing='ij'), axis=-1).reshape(-1, 3)
    idx = np.stack(np.meshgrid(*[np.arange(v) for v in n], indexing='ij'), axis=-1).reshape(-1, 3)
    kk = attrs.kcoords('wavenumber')
    freq = [np.fft.fftfreq(n[0]) * n[0], np.fft.fftfreq(n[1]) * n[1], np.fft.rfftfreq(n[2]) * n[2]]
    out = np.zeros((n[0], n[1], n[2] // 2 + 1), dtype=complex)
    flat = rho.ravel()
    for i in range(n[0]):
        for j in range(n[1]):
            for l in range(n[2] // 2 + 1):
                kv = np.array([kk[0][i], kk[1][j], kk[2][l]])
                leg = addition_at(ell, X, kv)
                phase = np.exp(-2j * np.pi * (idx @ np.array([freq[0][i] / n[0], freq[1][j] / n[1], freq[2][l] / n[2]])))
                out[i, j, l] = (2 * ell + 1) / (4. * np.pi) * np.sum(flat * phase * leg)
    return out


# ----------------------------------------------------------------------------
# Bispectrum, Sugiyama basis (mesh3.py:23-177, 180-313, 443-471, 509-629)
# ----------------------------------------------------------------------------

def ref_real_ylm(ell, m, x, y, z):
    """The reference's real Y_lm (get_Ylm(ell, m, real=True), utils.py:191-334) including its
    per-m sign: the normalisation of utils.py:227-238 carries (-1)^m and sympy's associated
    Legendre function the Condon-Shortley phase, so it is (-1)^m real_ylm for m != 0.  Pinned
    against the sympy construction in tests/test_oracle.py."""
    out = real_ylm(ell, m, x, y, z)
    return out if m == 0 else (-1.)**m * out


def complex_ylm_conj(ell, m, x, y, z):
    """get_Ylm(ell, m, reduced=True, real=False, conj=True) (utils.py:191-334): the conjugate of
    sqrt((l-|m|)!/(l+|m|)!) [(-1)^m for m < 0] P_l^|m|(cos theta) exp(i m phi), written through
    the real harmonics: with a = sqrt((2l+1)/4pi), m = 0: Y_l0 / a; m = p > 0:
    (-1)^p (Y_lp - i Y_l,-p) / (a sqrt 2); m = -p: (Y_lp + i Y_l,-p) / (a sqrt 2).  At the zero
    vector this is the constant term of the sympy polynomial, as for the real harmonics.  Pinned
    against the sympy construction in tests/test_oracle.py."""
    a = np.sqrt((2 * ell + 1) / (4. * np.pi))
    if m == 0:
        return ref_real_ylm(ell, 0, x, y, z) / a + 0j
    p = abs(m)
    re, im = ref_real_ylm(ell, p, x, y, z), ref_real_ylm(ell, -p, x, y, z)
    if m > 0:
        return (-1.)**p * (re - 1j * im) / (a * np.sqrt(2.))
    return (re + 1j * im) / (a * np.sqrt(2.))


def wigner_3j(j1, j2, j3, m1, m2, m3):
    """Wigner 3j symbol of integers by Racah's formula (the reference calls sympy's,
    utils.py:634-637).  Pinned against sympy in tests/test_oracle.py."""
    j1, j2, j3, m1, m2, m3 = (int(v) for v in (j1, j2, j3, m1, m2, m3))
    if m1 + m2 + m3 != 0 or not (abs(j1 - j2) <= j3 <= j1 + j2) or abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.
    f = math.factorial
    delta = f(j1 + j2 - j3) * f(j1 - j2 + j3) * f(-j1 + j2 + j3) / f(j1 + j2 + j3 + 1)
    pre = (-1.)**(j1 - j2 - m3) * math.sqrt(delta * f(j1 + m1) * f(j1 - m1) * f(j2 + m2) * f(j2 - m2) * f(j3 + m3) * f(j3 - m3))
    kmin = max(0, j2 - j3 - m1, j1 - j3 + m2)
    kmax = min(j1 + j2 - j3, j1 - m1, j2 + m2)
    s = 0.
    for k in range(kmin, kmax + 1):
        s += (-1.)**k / (f(k) * f(j3 - j2 + k + m1) * f(j3 - j1 + k - m2) * f(j1 + j2 - j3 - k) * f(j1 - k - m1) * f(j2 - k + m2))
    return pre * s


def triposh_rows(ell1, ell2, ell3, los_z=False):
    """_iter_triposh (mesh3.py:443-471): every (m1, m2, M) with coefficient
    (2l1+1)(2l2+1)(2L+1) W(l1 l2 L; 0 0 0) W(l1 l2 L; m1 m2 M) of modulus >= 1e-7, M = 0 only for
    the line of sight 'z'.  The reference's pairing of (m1, m2, M) with (-m1, -m2, -M) compares a
    tuple with lists (mesh3.py:462-468) and never matches, so every row is kept with sym = 0."""
    rows = []
    ms3 = [0] if los_z else range(-ell3, ell3 + 1)
    for m1 in range(-ell1, ell1 + 1):
        for m2 in range(-ell2, ell2 + 1):
            for m3 in ms3:
                g = (2 * ell1 + 1) * (2 * ell2 + 1) * (2 * ell3 + 1) * wigner_3j(ell1, ell2, ell3, 0, 0, 0) * \
                    wigner_3j(ell1, ell2, ell3, m1, m2, m3)
                if abs(g) < 1e-7:
                    continue
                rows.append((m1, m2, m3, g))
    return rows


def full_kvec(attrs):
    """Wavenumbers of the full (non-hermitian) FFT grid, fftfreq on every axis (mesh.py:361-403
    with hermitian=False, as the complex-dtype attributes of the Sugiyama basis give)."""
    n = [int(x) for x in attrs.meshsize]
    return [np.fft.fftfreq(n[a]) * attrs.kfun[a] * n[a] for a in range(3)]


class BinBispectrumSugiyama:
    """BinMesh3SpectrumPoles with basis 'sugiyama' or 'sugiyama-diagonal' (mesh3.py:23-177,
    180-313): two sides, each binning the modes of the full complex grid (the Sugiyama basis
    clones the attributes to the complex dtype, mesh3.py:29-30, so no hermitian weights) by
    digitize(|k|, edges) - 1; pairs (b1, b2) in lexicographic order (or the diagonal) that pass
    the mask, by default mid1 <= mid2 and upper1 + upper2 <= nyq (mesh3.py:36-38); nmodes is the
    product of the two one-dimensional counts and xavg the pair of their mean |k|."""

    def __init__(self, attrs, edges=None, ells=(0, 0, 0), mask_edges=None, diagonal=False):
        ndim = 2
        if not isinstance(edges, (list, tuple)):
            edges = [edges] * ndim
        self.attrs = attrs
        kk = full_kvec(attrs)
        k2 = kk[0][:, None, None]**2 + kk[1][None, :, None]**2
        knorm = np.sqrt(k2 + kk[2][None, None, :]**2)
        self.ibin1d, self.nmodes1d, self.xavg1d, self.edges1d = [], [], [], []
        for e in edges[:ndim]:
            e = make_edges(attrs, e)
            ib = np.digitize(knorm, e, right=False) - 1
            nb = len(e) - 1
            ok = (ib >= 0) & (ib < nb)
            cnt = np.bincount(ib[ok], minlength=nb).astype('f8')
            xs = np.bincount(ib[ok], weights=knorm[ok], minlength=nb)
            with np.errstate(invalid='ignore', divide='ignore'):
                self.xavg1d.append(xs / cnt)
            self.ibin1d.append(ib)
            self.nmodes1d.append(cnt)
            self.edges1d.append(np.column_stack([e[:-1], e[1:]]))
        self.ells = [tuple(int(v) for v in ell) for ell in (ells if np.ndim(ells[0]) else [ells])]

        def prod(arrays):
            if diagonal:
                return np.column_stack([np.asarray(arrays[0])] * ndim)
            grid = np.meshgrid(*arrays, indexing='ij')
            return np.column_stack([g.ravel() for g in grid])

        lo = prod([e[:, 0] for e in self.edges1d])
        hi = prod([e[:, 1] for e in self.edges1d])
        edges2 = np.stack([lo, hi], axis=-1)
        nyq = attrs.kfun.min() * np.min(attrs.meshsize) / 2.
        if mask_edges is None:
            mid = edges2.mean(axis=-1)
            mask = (mid[:, 0] <= mid[:, 1]) & (edges2[:, :, 1].sum(axis=1) <= nyq)
        else:
            assert mask_edges == '', 'the oracle takes the default mask or none'
            mask = np.ones(len(edges2), dtype=bool)
        self.edges = edges2[mask]
        self.xavg = prod(self.xavg1d)[mask]
        self.iedges = prod([np.arange(len(x)) for x in self.xavg1d])[mask]
        self.nmodes = np.prod(prod(self.nmodes1d)[mask], axis=-1)

    def shells(self, axis, value):
        """ifftn(value [bin == b]) for every shell b of side axis, value on the full grid
        (mesh3.py:300-305 with the complex-dtype c2r)."""
        return [np.fft.ifftn(value * (self.ibin1d[axis] == b)) for b in range(len(self.xavg1d[axis]))]

    def reduce(self, F0, F1, third):
        """N^6 sum_x ifftn(F0 [b1]) ifftn(F1 [b2]) third per pair (mesh3.py:277-313)."""
        s0, s1 = self.shells(0, F0), self.shells(1, F1)
        n6 = float(np.prod(self.attrs.meshsize))**2
        return np.array([np.sum(s0[b1] * s1[b2] * third) for b1, b2 in self.iedges]) * n6


def bispectrum_sugiyama(attrs, bin, mesh1, mesh2=None, mesh3=None, los='z'):
    """compute_mesh3_spectrum in the Sugiyama basis (mesh3.py:595-621): for each (l1, l2, L)
    and each row (m1, m2, M, g) of triposh_rows, g bin(A1 Y*_l1m1(khat), A2 Y*_l2m2(khat),
    Y*_LM(los) rho3), with A the full FFTs of the first two (real) meshes, rho3 the third mesh in
    configuration space and los the line-of-sight vector (global) or the node positions from the
    origin ('local'); summed over the rows, divided by nmodes, real part for even l1 + l2 and
    imaginary part otherwise; norm = N^3 / Vcell^2."""
    meshes = [mesh1, mesh2, mesh3]
    for i in (1, 2):
        if meshes[i] is None:
            meshes[i] = meshes[i - 1]
    real = [c2r(m, attrs.meshsize) if np.iscomplexobj(m) else np.asarray(m, dtype='f8') for m in meshes]
    A = [np.fft.fftn(r) for r in real[:2]]
    kk = full_kvec(attrs)
    kx, ky, kz = kk[0][:, None, None], kk[1][None, :, None], kk[2][None, None, :]
    local = isinstance(los, str) and los == 'local'
    if local:
        xs = attrs.rcoords()
        lx, ly, lz = xs[0][:, None, None], xs[1][None, :, None], xs[2][None, None, :]
    else:
        lx, ly, lz = los_vector(los)
    norm = np.prod(attrs.meshsize) / np.prod(attrs.cellsize)**2
    num = []
    for ell1, ell2, ell3 in bin.ells:
        total = np.zeros(len(bin.iedges), dtype=complex)
        for m1, m2, m3, g in triposh_rows(ell1, ell2, ell3, los_z=isinstance(los, str) and los == 'z'):
            third = complex_ylm_conj(ell3, m3, lx, ly, lz) * real[2]
            total += g * bin.reduce(A[0] * complex_ylm_conj(ell1, m1, kx, ky, kz), A[1] * complex_ylm_conj(ell2, m2, kx, ky, kz), third)
        with np.errstate(invalid='ignore', divide='ignore'):
            total = total / bin.nmodes
        num.append(total.real if (ell1 + ell2) % 2 == 0 else total.imag)
    return dict(k=bin.xavg, k_edges=bin.edges, nmodes=bin.nmodes, num_raw=num, norm=norm, ells=bin.ells,
                los='local' if local else los_vector(los))


def bispectrum_sugiyama_bruteforce(attrs, bin, rmeshes, ell=(0, 0, 0), los='z'):
    """The pair sums of bispectrum_sugiyama (before the division by nmodes) as direct sums over
    pairs of full-grid modes: sum over rows of g sum_{k1 in b1, k2 in b2} A1(k1) Y*(k1) A2(k2)
    Y*(k2) H(-(k1 + k2)), with H = fftn(Y*_LM(los) rho3), from N^6 sum_x ifftn(.) ifftn(.) h =
    sum_{k1, k2} a(k1) b(k2) fftn(h)(-(k1 + k2)).  Tiny meshes only."""
    n = np.array([int(x) for x in attrs.meshsize])
    A = [np.fft.fftn(m) for m in rmeshes[:2]]
    kk = full_kvec(attrs)
    KX, KY, KZ = np.meshgrid(*kk, indexing='ij')
    local = isinstance(los, str) and los == 'local'
    if local:
        xs = attrs.rcoords()
        LX, LY, LZ = np.meshgrid(*xs, indexing='ij')
    else:
        LX, LY, LZ = los_vector(los)
    idx = np.stack(np.meshgrid(*[np.arange(x) for x in n], indexing='ij'), axis=-1).reshape(-1, 3)
    ell1, ell2, ell3 = ell
    out = np.zeros(len(bin.iedges), dtype=complex)
    key = {tuple(t): i for i, t in enumerate(bin.iedges)}
    ib0, ib1 = bin.ibin1d[0].ravel(), bin.ibin1d[1].ravel()
    for m1, m2, m3, g in triposh_rows(ell1, ell2, ell3, los_z=isinstance(los, str) and los == 'z'):
        a = (A[0] * complex_ylm_conj(ell1, m1, KX, KY, KZ)).ravel()
        b = (A[1] * complex_ylm_conj(ell2, m2, KX, KY, KZ)).ravel()
        H = np.fft.fftn(complex_ylm_conj(ell3, m3, LX, LY, LZ) * rmeshes[2])
        for i in np.flatnonzero(ib0 >= 0):
            k3 = (-(idx[i][None, :] + idx)) % n
            h = H[k3[:, 0], k3[:, 1], k3[:, 2]]
            for j in np.flatnonzero(ib1 >= 0):
                t = key.get((ib0[i], ib1[j]))
                if t is not None:
                    out[t] += g * a[i] * b[j] * h[j]
    return out


def aliasing_shotnoise_window(order, kcirc):
    """Aliased shot-noise factor prod_i K_order(k_i) on circular wavenumbers
    (_aliasing_shotnoise_kernel, resamplers.py:29-55)."""
    def one(a):
        s2 = np.sin(a / 2.)**2
        return {1: np.ones_like(a), 2: 1. - 2. / 3. * s2, 3: 1. - s2 + 2. / 15. * s2**2,
                4: 1. - 4. / 3. * s2 + 2. / 5. * s2 * s2 - 4. / 315. * s2**3}[order]
    out = 1.
    for a in kcirc:
        out = out * one(a)
    return out


def sugiyama_shotnoise(attrs, bin, positions, weights=None, los='z', resampler='cic', interlacing=0, compensate=True, parts=None):
    """compute_fkp3_shotnoise in the Sugiyama basis for one field (mesh3.py:1091-1180 with
    _compute_sugiyama_spectrum_S122 855-883 and _S113 886-915, eqs. 45-46 of arXiv:1803.02132),
    on the full complex grid as the reference's complex-dtype attributes give:
      S111: s111_lm = sum_x paint(w^3) Y*_lm(xhat) (conjugated reduced harmonics, xhat from the
        origin), added to the (0, 0, 0) multipole;
      S122 (and S121 on the second side): for (l, 0, l) [(0, l, l)], (2l + 1) times the shell
        mean over the full grid of sum_m Y_lm(khat) [C(k) conj(FFT(rho_c Y_lm(los))(k)) -
        s111_lm W(k)], with C = FFT(paint(w)) (k = 0 removed), rho_c = paint(w^2) - mean;
      S113: i^(l1 + l2) sum over the rows (m1, m2, M, g) of g sum_s
        ifftn(FFT(rho_c Y*_LM(los)) conj(C_2) - s111_LM W)(s) Y*_l1m1(shat) Y*_l2m2(shat)
        j_l1(|s| k1) j_l2(|s| k2), rho_c = paint(w) - mean, C_2 = FFT(paint(w^2)) (k = 0
        removed), (k1, k2) the pair's mean |k|;
    W = 1 (compensated) or the squared window (not compensated) with interlacing, else the
    aliased shot-noise factor times the squared compensation when compensated (mesh3.py:1081-1089).
    Real part for even l1 + l2, imaginary part otherwise.  With a BinCorrelation3Sugiyama (the
    three-point correlation function, mesh3.py:918-976 and the is_correlation branches of
    1064-1180): S111 counts only pairs whose two bins contain s = 0, S122 bins
    sum_m Y_lm(shat) ifftn(.) in separation and counts where the other bin contains s = 0, S113
    sums U(s) over the intersection of the two bins divided by both cell counts with sign
    (-1)^(l1 + l2), all divided by Vcell^2; real part.  A dict given as parts receives the
    pieces: 's111' {(l, m): value}, 's122' {(axis, l): per-shell array}, 's113' {ell: per-pair
    array}."""
    order = ORDERS[resampler]
    w = np.ones(len(positions)) if weights is None else np.asarray(weights, dtype='f8')
    n = [int(x) for x in attrs.meshsize]

    def painted(ww):
        return paint(attrs, positions, ww, resampler=resampler, interlacing=interlacing, compensate=compensate, out='real')

    kk = full_kvec(attrs)
    KX, KY, KZ = np.meshgrid(*kk, indexing='ij')
    kcirc = np.meshgrid(*[np.fft.fftfreq(x) * 2. * np.pi for x in n], indexing='ij')
    comp = 1.
    for a in kcirc:
        comp = comp * np.sinc(a / (2. * np.pi))**(-order)
    if max(interlacing, 1) >= 2:
        W = np.ones(n) if compensate else 1. / comp**2
    else:
        W = aliasing_shotnoise_window(order, kcirc) * (comp**2 if compensate else 1.)
    xs = attrs.rcoords()
    X = np.meshgrid(*xs, indexing='ij')
    ssv = np.meshgrid(*separation_coords(attrs), indexing='ij')
    snorm_ = np.sqrt((ssv[0]**2 + ssv[1]**2) + ssv[2]**2)
    local = isinstance(los, str) and los == 'local'
    L = X if local else list(los_vector(los))
    ells = [tuple(e) for e in bin.ells]
    out = [np.zeros(len(bin.iedges), dtype=complex) for _ in ells]
    uells = sorted({l for e in ells for l in e})
    correlation = isinstance(bin, BinCorrelation3Sugiyama)  # mesh3.py:1064-1180 with is_correlation
    vc2 = np.prod(attrs.cellsize)**2
    rho3 = painted(w**3)
    s111 = {(l, m): np.sum(rho3 * complex_ylm_conj(l, m, *X)) for l in uells for m in range(-l, l + 1)}
    if (0, 0, 0) in ells:
        mask0 = np.all((bin.edges[..., 0] <= 0.) & (bin.edges[..., 1] >= 0.), axis=1) / vc2 if correlation else 1.
        out[ells.index((0, 0, 0))] += s111[0, 0] * mask0

    def shell_mean(value, axis):
        ib = bin.ibin1d[axis]
        nb = len(bin.nmodes1d[axis])
        ok = (ib >= 0) & (ib < nb)
        s = np.bincount(ib[ok], weights=value.real[ok], minlength=nb) + 1j * np.bincount(ib[ok], weights=value.imag[ok], minlength=nb)
        with np.errstate(invalid='ignore', divide='ignore'):
            return s / bin.nmodes1d[axis]

    C = np.fft.fftn(painted(w))
    C[0, 0, 0] = 0.
    rc = painted(w * w)
    rc = rc - rc.mean()
    for axis in (0, 1):
        sel = [e for e in ells if (e[2] == e[0] and e[1] == 0) if axis == 0] or [] if axis == 0 else \
              [e for e in ells if e[2] == e[1] and e[0] == 0]
        for e in sel:
            l = e[0] if axis == 0 else e[1]
            carry = 0.
            for m in range(-l, l + 1):
                ylm_l = np.conj(complex_ylm_conj(l, m, *L))
                prod_ = C * np.conj(np.fft.fftn(rc * ylm_l)) - s111[l, m] * W
                if correlation:  # Y_lm(shat) c2r(.), binned in separation (mesh3.py:918-944)
                    carry = carry + np.conj(complex_ylm_conj(l, m, *ssv)) * np.fft.ifftn(prod_)
                else:
                    carry = carry + np.conj(complex_ylm_conj(l, m, KX, KY, KZ)) * prod_
            piece = (2 * l + 1) * shell_mean(carry, axis)
            if parts is not None:
                parts.setdefault('s122', {})[axis, l] = piece
            mask = ((bin.edges[:, 1 - axis, 0] <= 0.) & (bin.edges[:, 1 - axis, 1] >= 0.)) / vc2 if correlation else 1.
            out[ells.index(e)] += piece[bin.iedges[:, axis]] * mask
    rc = painted(w)
    rc = rc - rc.mean()
    C2 = np.fft.fftn(painted(w * w))
    C2[0, 0, 0] = 0.
    k1 = bin.xavg1d[0][bin.iedges[:, 0]]
    k2 = bin.xavg1d[1][bin.iedges[:, 1]]
    for ill, (l1, l2, l3) in enumerate(ells):
        U = 0.
        for m1, m2, m3, g in triposh_rows(l1, l2, l3, los_z=isinstance(los, str) and los == 'z'):
            t = np.fft.ifftn(np.fft.fftn(rc * complex_ylm_conj(l3, m3, *L)) * np.conj(C2) - s111[l3, m3] * W)
            U = U + g * t * complex_ylm_conj(l1, m1, *ssv) * complex_ylm_conj(l2, m2, *ssv)
        if np.ndim(U) == 0:
            continue
        if correlation:  # sums over the intersection of the pair's two separation bins (mesh3.py:947-976, 1118-1128)
            lo = np.max(bin.edges[..., 0], axis=1)
            hi = np.min(bin.edges[..., 1], axis=1)
            total = np.array([np.sum(U[(snorm_ >= a) & (snorm_ < b)]) if b > a else 0. for a, b in zip(lo, hi)])
            total = (-1.)**(l1 + l2) * total / (bin.nmodes1d[0][bin.iedges[:, 0]] * bin.nmodes1d[1][bin.iedges[:, 1]]) / vc2
        else:
            total = (1j)**(l1 + l2) * np.array([np.sum(U * special.spherical_jn(l1, snorm_ * a) * special.spherical_jn(l2, snorm_ * b))
                                                for a, b in zip(k1, k2)])
        out[ill] += total
        if parts is not None:
            parts.setdefault('s113', {})[l1, l2, l3] = total
    if parts is not None:
        parts['s111'] = s111
    if correlation:
        return [o.real for o in out]
    return [o.real if (e[0] + e[1]) % 2 == 0 else o.imag for e, o in zip(ells, out)]


# ----------------------------------------------------------------------------
# Three-point correlation function, Sugiyama basis (mesh3.py:316-409, 632-727)
# ----------------------------------------------------------------------------

class BinCorrelation3Sugiyama:
    """BinMesh3CorrelationPoles, basis 'sugiyama' or 'sugiyama-diagonal' (mesh3.py:316-409 with
    _make_edges3(kind='real'), mesh3.py:23-177): per side, separation bins of the full real grid
    (every cell weighted 1, as BinCorrelationPoles); pairs in lexicographic order (or the
    diagonal) that pass the mask (mid1 <= mid2 and upper1 + upper2 <= cellsize_min N_min / 2 by
    default); nmodes the product of the cell counts; xavg the mean |s| of each side's bin.  The
    call's filter of bin b is j_l(|k| xavg_b) (times [klimit[0] <= |k| < klimit[1]])."""

    def __init__(self, attrs, edges=None, ells=(0, 0, 0), mask_edges=None, diagonal=False, klimit=None):
        ndim = 2
        if not isinstance(edges, (list, tuple)):
            edges = [edges] * ndim
        self.attrs = attrs
        axes = [BinCorrelationPoles(attrs, edges=e) for e in edges[:ndim]]
        self.ibin1d = [ax.ibin - 1 for ax in axes]  # separation bin of every cell (digitize - 1)
        self.edges1d = [ax.edges for ax in axes]
        self.nmodes1d = [ax.nmodes for ax in axes]
        self.xavg1d = [ax.xavg for ax in axes]
        self.ells = [tuple(int(v) for v in ell) for ell in (ells if np.ndim(ells[0]) else [ells])]

        def prod(arrays):
            if diagonal:
                return np.column_stack([np.asarray(arrays[0])] * ndim)
            grid = np.meshgrid(*arrays, indexing='ij')
            return np.column_stack([g.ravel() for g in grid])

        lo = prod([e[:, 0] for e in self.edges1d])
        hi = prod([e[:, 1] for e in self.edges1d])
        edges2 = np.stack([lo, hi], axis=-1)
        nyq = attrs.cellsize.min() * np.min(attrs.meshsize) / 2.
        if mask_edges is None:
            mid = edges2.mean(axis=-1)
            mask = (mid[:, 0] <= mid[:, 1]) & (edges2[:, :, 1].sum(axis=1) <= nyq)
        else:
            assert mask_edges == '', 'the oracle takes the default mask or none'
            mask = np.ones(len(edges2), dtype=bool)
        self.edges = edges2[mask]
        self.xavg = prod(self.xavg1d)[mask]
        self.iedges = prod([np.arange(len(x)) for x in self.xavg1d])[mask]
        self.nmodes = np.prod(prod(self.nmodes1d)[mask], axis=-1)
        self.klimit = (0., float(np.min(attrs.knyq))) if klimit is True else klimit

    def filters(self, axis, value, ell):
        """ifftn(value j_ell(|k| xavg_b) [klimit]) for every bin b of side axis, value on the full
        complex grid (mesh3.py:383-388 with the complex-dtype c2r)."""
        kk = full_kvec(self.attrs)
        kn = np.sqrt((kk[0][:, None, None]**2 + kk[1][None, :, None]**2) + kk[2][None, None, :]**2)
        lim = 1. if self.klimit is None else ((kn >= self.klimit[0]) & (kn < self.klimit[1]))
        return [np.fft.ifftn(value * special.spherical_jn(ell, kn * s) * lim) for s in self.xavg1d[axis]]

    def reduce(self, F0, F1, third, ell=(0, 0)):
        """i^(l1 + l2) / Vcell^2 sum_x filters(0, F0, l1)[b1] filters(1, F1, l2)[b2] third per pair
        (mesh3.py:362-395)."""
        f0, f1 = self.filters(0, F0, ell[0]), self.filters(1, F1, ell[1])
        norm = (1j)**(ell[0] + ell[1]) / np.prod(self.attrs.cellsize)**2
        return np.array([np.sum(f0[b1] * f1[b2] * third) for b1, b2 in self.iedges]) * norm


def correlation3_sugiyama(attrs, bin, mesh1, mesh2=None, mesh3=None, los='z'):
    """compute_mesh3_correlation, Sugiyama basis (mesh3.py:632-727): for each (l1, l2, L) and each
    row (m1, m2, M, g) of triposh_rows, g bin.reduce(A1 Y*_l1m1(khat), A2 Y*_l2m2(khat),
    Y*_LM(los) rho3, ell=(l1, l2)), summed over the rows, real part; A the full FFTs of the first
    two (real) meshes, rho3 the third in configuration space; norm = N^3 / Vcell^2."""
    meshes = [mesh1, mesh2, mesh3]
    for i in (1, 2):
        if meshes[i] is None:
            meshes[i] = meshes[i - 1]
    real = [c2r(m, attrs.meshsize) if np.iscomplexobj(m) else np.asarray(m, dtype='f8') for m in meshes]
    A = [np.fft.fftn(r) for r in real[:2]]
    K = np.meshgrid(*full_kvec(attrs), indexing='ij')
    local = isinstance(los, str) and los == 'local'
    L = np.meshgrid(*attrs.rcoords(), indexing='ij') if local else list(los_vector(los))
    num = []
    for ell1, ell2, ell3 in bin.ells:
        total = np.zeros(len(bin.iedges), dtype=complex)
        for m1, m2, m3, g in triposh_rows(ell1, ell2, ell3, los_z=isinstance(los, str) and los == 'z'):
            total += g * bin.reduce(A[0] * complex_ylm_conj(ell1, m1, *K), A[1] * complex_ylm_conj(ell2, m2, *K),
                                    complex_ylm_conj(ell3, m3, *L) * real[2], ell=(ell1, ell2))
        num.append(total.real)
    return dict(s=bin.xavg, s_edges=bin.edges, nmodes=bin.nmodes, num_raw=num,
                norm=np.prod(attrs.meshsize) / np.prod(attrs.cellsize)**2, ells=bin.ells)


def correlation3_sugiyama_bruteforce(attrs, bin, rmeshes, ell=(0, 0, 0), los='z'):
    """The row sums of correlation3_sugiyama before the real part, as direct sums over pairs of
    full-grid modes: sum over rows of g i^(l1 + l2) / (Vcell^2 N^6) sum_{k1, k2} a(k1) b(k2)
    H(-(k1 + k2)), with a = A1 Y*(k1) j_l1(|k1| s1), b = A2 Y*(k2) j_l2(|k2| s2) and
    H = fftn(Y*_LM(los) rho3), from sum_x ifftn(a) ifftn(b) h = N^-6 sum a b fftn(h)(-(k1 + k2))
    (N^3 cells).  Tiny meshes only."""
    n = np.array([int(x) for x in attrs.meshsize])
    ncell = float(np.prod(n))
    A = [np.fft.fftn(m) for m in rmeshes[:2]]
    kk = full_kvec(attrs)
    KX, KY, KZ = np.meshgrid(*kk, indexing='ij')
    kn = np.sqrt((KX**2 + KY**2) + KZ**2).ravel()
    local = isinstance(los, str) and los == 'local'
    L = np.meshgrid(*attrs.rcoords(), indexing='ij') if local else list(los_vector(los))
    idx = np.stack(np.meshgrid(*[np.arange(x) for x in n], indexing='ij'), axis=-1).reshape(-1, 3)
    ell1, ell2, ell3 = ell
    out = np.zeros(len(bin.iedges), dtype=complex)
    lim = np.ones_like(kn) if bin.klimit is None else ((kn >= bin.klimit[0]) & (kn < bin.klimit[1])).astype(float)
    for m1, m2, m3, g in triposh_rows(ell1, ell2, ell3, los_z=isinstance(los, str) and los == 'z'):
        a0 = (A[0] * complex_ylm_conj(ell1, m1, KX, KY, KZ)).ravel() * lim
        b0 = (A[1] * complex_ylm_conj(ell2, m2, KX, KY, KZ)).ravel() * lim
        H = np.fft.fftn(complex_ylm_conj(ell3, m3, *L) * rmeshes[2])
        k3 = (-(idx[:, None, :] + idx[None, :, :])) % n
        Hk = H[k3[..., 0], k3[..., 1], k3[..., 2]]
        for t, (b1, b2) in enumerate(bin.iedges):
            a = a0 * special.spherical_jn(ell1, kn * bin.xavg1d[0][b1])
            b = b0 * special.spherical_jn(ell2, kn * bin.xavg1d[1][b2])
            out[t] += g * np.sum(a[:, None] * b[None, :] * Hk)
    return out * (1j)**(ell1 + ell2) / np.prod(attrs.cellsize)**2 / ncell**2
