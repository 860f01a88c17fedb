"""Bispectrum multipoles in the Scoccimarro basis (global or local line of sight): the binning
operator BinMesh3SpectrumPoles (jaxpower/mesh3.py:23-177, 180-313), compute_mesh3_spectrum
(mesh3.py:509-629), compute_mesh3 (mesh3.py:475-506) and compute_box3_normalization
(mesh3.py:730-735).

Each |k| shell b of a half spectrum A becomes the real mesh I_b = irfftn(A [bin == b]) on
the GPU (``mp_bispec_shells``), and the closed-triangle sums sum_x I_b1 I_b2 I_b3 over the
allowed (b1, b2, b3) are taken in one pass over the cells (``mp_bispec_triangles``), FP32 per
cell and FP64 across cells.  The mode counts are the same triangle sums of the shell
indicators (mesh3.py:229-236).  With the local line of sight the third side is
sum_m r2c(rho_3 Y_lm(xhat)) Y_lm(khat) (mesh3.py:563-576).

The Sugiyama basis (the reference's default) bins pairs of shells (k1, k2): for each
(l1, l2, L) and each (m1, m2, M) of nonzero Gaunt coefficient g (mesh3.py:443-471),
g N^6 sum_x ifftn(A1 [b1] Y*_l1m1) ifftn(A2 [b2] Y*_l2m2) Y*_LM(los) rho_3 (mesh3.py:595-621).
The complex harmonics are combinations of two real ones, and ifftn(A [b] Y_lm) of a real
harmonic is a real mesh (even l) or i times one (odd l), so every term is a sum of real
triple products taken by the same triangle pass.  The three-point correlation function
(BinMesh3CorrelationPoles, compute_mesh3_correlation, mesh3.py:316-409, 632-727) replaces the
shell indicators by spherical Bessel filters j_l(|k| s_b) on the full complex grid.
"""
import math

import numpy as np
import torch

from . import _lib
from .mesh import (MeshAttrs, RealMeshField, ComplexMeshField, ParticleField, FKPField, format_meshes, compute_box_normalization,
                   compute_normalization,
                   _as_spectrum, _default_device, _empty_spectrum, _spectrum_pitch, _sum_f64)
from .mesh2 import BinMesh2SpectrumPoles, _deferred, _get_los_vector, _ylm_accumulate
from .types import Deferred, Mesh3SpectrumPole, Mesh3SpectrumPoles

_BASES = ('sugiyama', 'sugiyama-diagonal', 'scoccimarro', 'scoccimarro-diagonal')
_MAX_SHELL_ROWS = 96  # distinct shell meshes one triangle pass stages (three stacks share this)


def _format_ells(ells, basis='scoccimarro'):
    """mesh3.py:411-430: a list of ints (Scoccimarro) or of (ell1, ell2, L) (Sugiyama)."""
    if 'scoccimarro' in basis:
        return [int(ell) for ell in (np.atleast_1d(ells) if np.ndim(ells) else [ells])]
    if np.ndim(ells) == 0:
        raise ValueError('ells must be (a list of) (ell1, ell2, L)')
    if np.ndim(ells[0]) == 0:
        return [tuple(ells)]
    return [tuple(ell) for ell in ells]


def _default_mask(basis):
    """Default triangle selection, mesh3.py:31-39 (after eq. 53 of arXiv:1512.07295)."""
    if 'scoccimarro' in basis:
        return ['mid1 <= mid2', 'mid2 <= mid3', '(mid3 >= jnp.abs(mid1 - mid2)) & (mid3 <= jnp.abs(mid1 + mid2))',
                '(edge1 + edge2 + edge3)[:, 1] <= 2 * nyq']
    return ['mid1 <= mid2', '(edge1 + edge2)[:, 1] <= nyq']


def _mask_function(mask_edges, nyq):
    """The triangle mask of mesh3.py:40-57: conditions on mid{i} (bin centres) and edge{i}
    ((n, 2) bin edges) of each side, as ';'-separated expressions (evaluated with numpy
    bound to both ``np`` and ``jnp``), a list of them, or a callable."""
    if isinstance(mask_edges, str):
        mask_edges = tuple(mask_edges.strip().split(';'))
    if callable(mask_edges):
        return mask_edges
    conditions = [c.strip() for c in mask_edges if c.strip()]

    def mask(*edges):
        d = {'nyq': nyq}
        for i, edge in enumerate(edges):
            d[f'mid{i + 1:d}'] = np.mean(edge, axis=-1)
            d[f'edge{i + 1:d}'] = edge
        out = np.ones(edges[0].shape[0], dtype=bool)
        for c in conditions:
            out &= np.asarray(eval(c, {'np': np, 'jnp': np}, d), dtype=bool)  # the reference's expression interface
        return out

    return mask


def _wigner_3j(j1, j2, j3, m1, m2, m3):
    """Wigner 3j symbol of integers by Racah's formula (the reference calls sympy's,
    utils.py:634-637)."""
    if m1 + m2 + m3 != 0 or not (abs(j1 - j2) <= j3 <= j1 + j2) or abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.
    f = math.factorial
    tri = f(j1 + j2 - j3) * f(j1 - j2 + j3) * f(-j1 + j2 + j3) / f(j1 + j2 + j3 + 1)
    pre = (-1.)**(j1 - j2 - m3) * math.sqrt(tri * f(j1 + m1) * f(j1 - m1) * f(j2 + m2) * f(j2 - m2) * f(j3 + m3) * f(j3 - m3))
    total = 0.
    for k in range(max(0, j2 - j3 - m1, j1 - j3 + m2), min(j1 + j2 - j3, j1 - m1, j2 + m2) + 1):
        total += (-1.)**k / (f(k) * f(j3 - j2 + k + m1) * f(j3 - j1 + k - m2) * f(j1 + j2 - j3 - k) * f(j1 - k - m1) * f(j2 - k + m2))
    return pre * total


def _triposh_rows(ell1, ell2, ell3, los_z=False):
    """(m1, m2, M, g) with g = (2l1+1)(2l2+1)(2L+1) W(l1 l2 L; 000) W(l1 l2 L; m1 m2 M) of modulus
    >= 1e-7, M = 0 only for the line of sight 'z' (_iter_triposh, mesh3.py:443-471).  The
    reference's pairing of opposite m triples (mesh3.py:462-468) compares a tuple with lists and
    never matches, so every row enters once with its own coefficient."""
    rows = []
    for m1 in range(-ell1, ell1 + 1):
        for m2 in range(-ell2, ell2 + 1):
            for m3 in ([0] if los_z else range(-ell3, ell3 + 1)):
                g = (2 * ell1 + 1) * (2 * ell2 + 1) * (2 * ell3 + 1) * _wigner_3j(ell1, ell2, ell3, 0, 0, 0) * \
                    _wigner_3j(ell1, ell2, ell3, m1, m2, m3)
                if abs(g) >= 1e-7:
                    rows.append((m1, m2, m3, g))
    return rows


def _ylm_components(ell, m):
    """get_Ylm(ell, m, reduced=True, real=False, conj=True) (utils.py:191-334) as
    sum_mu c_mu Y_l,mu of the real harmonics (get_Ylm(real=True)): with a = sqrt((2l+1)/4pi),
    m = 0: Y_l0 / a; m = p > 0: (-1)^p (Y_lp - i Y_l,-p) / (a sqrt 2); m = -p:
    (Y_lp + i Y_l,-p) / (a sqrt 2).  The zero-vector values agree: only m = 0 harmonics are
    nonzero there."""
    a = math.sqrt((2 * ell + 1) / (4. * math.pi))
    if m == 0:
        return {0: 1. / a + 0j}
    p = abs(m)
    c = 1. / (a * math.sqrt(2.))
    if m > 0:
        return {p: (-1.)**p * c + 0j, -p: -1j * (-1.)**p * c}
    return {p: c + 0j, -p: 1j * c}


def _real_ylm_at(ell, m, v):
    """get_Ylm(ell, m, real=True) (utils.py:191-334) at a nonzero vector v: sqrt((2l+1)/4pi)
    [sqrt((l-|m|)!/(l+|m|)!) sqrt(2) (-1)^m for m != 0] P_l^|m|(cos theta) {cos m phi | sin |m| phi},
    P_l^|m| with the Condon-Shortley phase."""
    from scipy.special import lpmv
    v = np.asarray(v, dtype='f8')
    x, y, z = v / np.linalg.norm(v)
    amp = math.sqrt((2 * ell + 1) / (4. * math.pi))
    if m != 0:
        amp *= math.sqrt(math.factorial(ell - abs(m)) / math.factorial(ell + abs(m))) * math.sqrt(2.) * (-1.)**m
    phi = math.atan2(y, x)
    trig = math.sin(abs(m) * phi) if m < 0 else (math.cos(m * phi) if m > 0 else 1.)
    return float(amp * lpmv(abs(m), ell, z) * trig)


def _complex_ylm_conj_at(ell, m, v):
    """get_Ylm(ell, m, reduced=True, real=False, conj=True) at a nonzero vector v:
    sqrt((l-|m|)!/(l+|m|)!) [(-1)^m if m < 0] P_l^|m|(cos theta) exp(-i m phi), with the
    Condon-Shortley phase in P_l^|m| as sympy's assoc_legendre (utils.py:223-238, 300-325)."""
    from scipy.special import lpmv
    v = np.asarray(v, dtype='f8')
    x, y, z = v / np.linalg.norm(v)
    amp = math.sqrt(math.factorial(ell - abs(m)) / math.factorial(ell + abs(m)))
    if m < 0:
        amp *= (-1.)**m
    return complex(amp * lpmv(abs(m), ell, z) * np.exp(-1j * m * math.atan2(y, x)))


def _full_grid_modes(axis, mattrs):
    """Counts and sum |k| per shell over the full complex grid (the Sugiyama basis clones the
    attributes to the complex dtype, mesh3.py:29-30, so no hermitian weights): the hermitian
    counts of the half spectrum less one copy of the kz = +Nyquist plane, which the weights
    1 + (kz > 0) count twice and the full grid holds once (at -Nyquist, same |k|)."""
    counts = np.asarray(axis.nmodes, dtype='f8')
    with np.errstate(invalid='ignore'):
        xsum = np.where(counts > 0, np.nan_to_num(axis.xavg) * counts, 0.)
    n = [int(x) for x in mattrs.meshsize]
    if n[2] % 2 == 0:
        from .mesh2 import _kvectors, _k2_thresholds
        kx, ky, kz = _kvectors(mattrs)
        k2 = (kx[:, None]**2 + ky[None, :]**2) + kz[-1]**2
        b = np.searchsorted(_k2_thresholds(axis._edges1d), k2.ravel(), side='right') - 1
        ok = (b >= 0) & (b < axis.nbins)
        counts = counts - np.bincount(b[ok], minlength=axis.nbins)
        xsum = xsum - np.bincount(b[ok], weights=np.sqrt(k2.ravel()[ok]), minlength=axis.nbins)
    with np.errstate(invalid='ignore', divide='ignore'):
        return counts, xsum / counts


def _product(arrays, diagonal):
    """mesh3.py:107-117: all (i, j, k) combinations in lexicographic order, or the diagonal."""
    if diagonal:
        return np.column_stack([np.asarray(arrays[0])] * len(arrays))
    grid = np.meshgrid(*arrays, sparse=False, indexing='ij')
    return np.column_stack([g.ravel() for g in grid])


class BinMesh3SpectrumPoles:
    """Binning operator for the bispectrum (mesh3.py:180-313), basis 'sugiyama' (the default),
    'sugiyama-diagonal', 'scoccimarro' or 'scoccimarro-diagonal'.

    Attributes mirror the reference: ``edges`` (nbins, ndim, 2), ``xavg`` (nbins, ndim),
    ``nmodes`` (one float array per ell: the closed-triangle counts in the Scoccimarro basis,
    the products of the two shells' full-grid mode counts in the Sugiyama basis), ``edges1d``,
    ``nmodes1d``, ``xavg1d``, ``ells``, ``basis``; ndim = 3 (Scoccimarro) or 2 (Sugiyama).
    ``batch_size`` and ``buffer_size`` (the reference's memory/throughput knobs for
    jax.lax.map) are accepted and stored; the GPU path stages the shells in groups of its own.
    """

    def __init__(self, mattrs, edges=None, ells=0, basis='sugiyama', batch_size=None, buffer_size=0, mask_edges=None, device=None):
        if not isinstance(mattrs, MeshAttrs):
            mattrs = mattrs.attrs
        if basis not in _BASES:
            raise ValueError(f'basis must be one of {_BASES}, got {basis!r}')
        device = device or _default_device()
        sugiyama = 'sugiyama' in basis
        ndim = 2 if sugiyama else 3
        diagonal = 'diagonal' in basis
        self.mattrs, self.basis, self._device = mattrs, basis, device
        self.batch_size = 1 if batch_size is None else int(batch_size)
        self.buffer_size = int(buffer_size)
        self.ells = _format_ells(ells, basis=basis)
        if edges is None:
            edges = {}
        if not isinstance(edges, (tuple, list)):
            edges = [edges] * ndim
        edges = list(edges)[:ndim]
        # one-dimensional shells per side (digitize and hermitian weights of mesh.py:2831-2863)
        axes = []
        for e in edges:
            axis = BinMesh2SpectrumPoles(mattrs, edges=e, device=device)
            same = [a for a in axes if np.array_equal(a._edges1d, axis._edges1d)]
            axes.append(same[0] if same else axis)  # sides with equal edges share their shells
        axes = axes + [axes[-1]] * (ndim - len(axes))
        self._axes = axes
        self.edges1d = tuple(a.edges for a in axes)
        if sugiyama:
            full = [_full_grid_modes(a, mattrs) for a in axes]
            self.nmodes1d = tuple(f[0] for f in full)
            self.xavg1d = tuple(f[1] for f in full)
        else:
            self.nmodes1d = tuple(a.nmodes for a in axes)
            self.xavg1d = tuple(a.xavg for a in axes)
        nyq = mattrs.kfun.min() * np.min(mattrs.meshsize) / 2.
        if mask_edges is None:
            mask_edges = _default_mask(basis)
        mask = _mask_function(mask_edges, nyq)
        lo = _product([e[:, 0] for e in self.edges1d], diagonal)
        hi = _product([e[:, 1] for e in self.edges1d], diagonal)
        edges3 = np.stack([lo, hi], axis=-1)  # (nall, 3, 2)
        keep = np.asarray(mask(*[edges3[:, i, :] for i in range(ndim)]), dtype=bool)
        self.edges = edges3[keep]
        self.xavg = _product(self.xavg1d, diagonal)[keep]
        self._iedges = _product([np.arange(len(x)) for x in self.xavg1d], diagonal)[keep].astype('i4')
        if not len(self._iedges):
            raise ValueError('no triangle passes the edge mask')
        self.wmodes = None
        if sugiyama:
            # The complex-grid shells are reproduced from half spectra, which holds below the
            # Nyquist planes (the default mask keeps every used shell there, mesh3.py:36-38).
            knyq = float(np.min(np.pi / np.asarray(mattrs.cellsize, dtype='f8')))
            top = max(float(self.edges1d[a][self._iedges[:, a], 1].max()) for a in range(ndim))
            if top > knyq:
                raise NotImplementedError(f'Sugiyama shells reaching {top:.4g} > the Nyquist wavenumber {knyq:.4g} are not '
                                          'implemented on this path')
            nmodes = np.prod(np.stack([self.nmodes1d[a][self._iedges[:, a]] for a in range(ndim)]), axis=0)
            self._nmodes_dev = torch.from_numpy(nmodes).to(device)
            self.nmodes = [nmodes] * len(self.ells)
            return
        self._tri = torch.from_numpy(np.ascontiguousarray(self._iedges.ravel())).to(device)
        # closed-triangle counts: triangle sums of the shell indicators times N^6 (mesh3.py:229-236)
        stacks = self._shell_stacks([None] * ndim, [0] * ndim)
        counts = self._triangles(stacks)
        norm = float(np.prod(mattrs.meshsize, dtype='f8'))**2
        del stacks
        self._nmodes_dev = counts * norm
        nmodes = self._nmodes_dev.cpu().numpy()
        self.nmodes = [nmodes] * len(self.ells)

    @property
    def nbins(self):
        return len(self.xavg)

    def _shell_stacks(self, spectra, ells, los=None):
        """Real shell meshes (nshell, N0, N1, N2) for each triangle side: spectra[i] is a
        (value, pitch) half spectrum or None (shell indicators), times P_ells[i](mu) for
        ells[i] > 0.  Sides with the same input, shells and ell share one stack."""
        n = _lib.i32x3(self.mattrs.meshsize)
        shape = tuple(int(x) for x in self.mattrs.meshsize)
        ws = _lib.workspace(_lib.load().mp_bispec_shells_workspace_bytes(n), self._device)
        made, stacks = {}, []
        for side in range(3):
            spec, ell, axis = spectra[side], ells[side], self._axes[side]
            key = (None if spec is None else spec[0].data_ptr(), ell, id(axis))
            if key not in made:
                nshell = axis.nbins
                out = torch.empty((nshell,) + shape, dtype=torch.float32, device=self._device)
                t = axis._tables
                _lib.call('mp_bispec_shells', _lib.stream_handle(self._device), None if spec is None else _lib.ptr(spec[0]),
                          0 if spec is None else spec[1], n, _lib.ptr(t['kx']), _lib.ptr(t['ky']), _lib.ptr(t['kz']),
                          _lib.ptr(t['thr']), len(axis._edges1d), 0, nshell, _lib.f64x3(los if los is not None else (0., 0., 0.)),
                          int(ell), _lib.ptr(out), _lib.ptr(ws), ws.numel())
                made[key] = out
            stacks.append(made[key])
        return stacks

    def _triangles(self, stacks):
        """FP64 triangle sums over the cells (device tensor, one per triangle)."""
        rows = sum(s.shape[0] for s in {s.data_ptr(): s for s in stacks}.values())
        if rows > _MAX_SHELL_ROWS:
            raise NotImplementedError(f'{rows} shell meshes in one triangle pass (at most {_MAX_SHELL_ROWS})')
        ntri = len(self._iedges)
        ncells = int(np.prod(self.mattrs.meshsize))
        out = torch.empty(ntri, dtype=torch.float64, device=self._device)
        ws = _lib.workspace(_lib.load().mp_bispec_triangles_workspace_bytes(ntri), self._device)
        _lib.call('mp_bispec_triangles', _lib.stream_handle(self._device), _lib.ptr(stacks[0]), _lib.ptr(stacks[1]), _lib.ptr(stacks[2]),
                  ncells, stacks[0].shape[0], stacks[1].shape[0], stacks[2].shape[0], _lib.ptr(self._tri), ntri, _lib.ptr(out),
                  _lib.ptr(ws), ws.numel())
        return out

    def _sums_dev(self, spectra, ells=(0, 0, 0), los=None):
        return self._triangles(self._shell_stacks(spectra, ells, los=los))

    def _ylm_stack(self, side, spec, ell, mu, cache):
        """Real meshes Q_b, one per shell b of the side, with ifftn(A [b] Y_l,mu(khat)) = Q_b
        (even l) or i Q_b (odd l), for the half spectrum spec = (value, pitch) (None: the shell
        indicators); ell = -1 gives the plain shells ifftn(A [b])."""
        axis = self._axes[side]
        key = (None if spec is None else spec[0].data_ptr(), id(axis), ell, mu)
        if key in cache:
            return cache[key]
        n = _lib.i32x3(self.mattrs.meshsize)
        shape = tuple(int(x) for x in self.mattrs.meshsize)
        ws = _lib.workspace(_lib.load().mp_bispec_shells_workspace_bytes(n), self._device)
        out = torch.empty((axis.nbins,) + shape, dtype=torch.float32, device=self._device)
        t = axis._tables
        args = (_lib.stream_handle(self._device), None if spec is None else _lib.ptr(spec[0]), 0 if spec is None else spec[1], n,
                _lib.ptr(t['kx']), _lib.ptr(t['ky']), _lib.ptr(t['kz']), _lib.ptr(t['thr']), len(axis._edges1d), 0, axis.nbins)
        if ell < 0:
            _lib.call('mp_bispec_shells', *args, _lib.f64x3((0., 0., 0.)), 0, _lib.ptr(out), _lib.ptr(ws), ws.numel())
        else:
            _lib.call('mp_bispec_shells_ylm', *args, int(ell), int(mu), _lib.ptr(out), _lib.ptr(ws), ws.numel())
        cache[key] = out
        return out

    def _pair_sums(self, Q0, Q1, H):
        """sum_x Q0[b1](x) Q1[b2](x) H[j](x) for every pair (b1, b2) of the bin and every row j
        of H: a (npairs, len(H)) float64 device tensor.  The shells are staged in groups that
        fit one triangle pass."""
        n3 = H.shape[0]
        nb = (Q0.shape[0], Q1.shape[0])
        budget = _MAX_SHELL_ROWS - n3
        if budget < 2:
            raise NotImplementedError(f'{n3} third-side meshes in one triangle pass')
        if nb[0] + nb[1] <= budget:
            groups = [(0, nb[0])], [(0, nb[1])]
        else:
            g = budget // 2
            groups = [(lo, min(lo + g, nb[0])) for lo in range(0, nb[0], g)], [(lo, min(lo + g, nb[1])) for lo in range(0, nb[1], g)]
        ie = self._iedges
        out = torch.zeros((len(ie), n3), dtype=torch.float64, device=self._device)
        ncells = int(np.prod(self.mattrs.meshsize))
        for lo0, hi0 in groups[0]:
            for lo1, hi1 in groups[1]:
                sel = np.flatnonzero((ie[:, 0] >= lo0) & (ie[:, 0] < hi0) & (ie[:, 1] >= lo1) & (ie[:, 1] < hi1))
                if not len(sel):
                    continue
                tri = np.stack([np.repeat(ie[sel, 0] - lo0, n3), np.repeat(ie[sel, 1] - lo1, n3), np.tile(np.arange(n3), len(sel))],
                               axis=1).astype('i4')
                tri_dev = torch.from_numpy(np.ascontiguousarray(tri.ravel())).to(self._device)
                ntri = len(tri)
                res = torch.empty(ntri, dtype=torch.float64, device=self._device)
                ws = _lib.workspace(_lib.load().mp_bispec_triangles_workspace_bytes(ntri), self._device)
                _lib.call('mp_bispec_triangles', _lib.stream_handle(self._device), _lib.ptr(Q0[lo0]), _lib.ptr(Q1[lo1]), _lib.ptr(H), ncells,
                          hi0 - lo0, hi1 - lo1, n3, _lib.ptr(tri_dev), ntri, _lib.ptr(res), _lib.ptr(ws), ws.numel())
                out[torch.from_numpy(sel).to(self._device)] = res.view(len(sel), n3)
        return out

    def __call__(self, *meshes, remove_zero=False):
        """Triangle sums of three half spectra times N^6 (mesh3.py:277-313), as a float64
        array; remove_zero sets each mesh's k = 0 mode to 0 first.  In the Sugiyama basis the
        third mesh is a configuration-space mesh, multiplied in as is (remove_zero subtracts
        its mean) and the sums run over the pairs of shells of the first two."""
        if 'sugiyama' in self.basis:
            return self._call_sugiyama(meshes, remove_zero)
        if len(meshes) != 3:
            raise ValueError('the scoccimarro binning takes three meshes')
        spectra = []
        for mesh in meshes:
            value = mesh.value if isinstance(mesh, ComplexMeshField) else mesh
            _lib.require_device(value, 'mesh')
            if remove_zero:
                value = value.clone()
                value[0, 0, 0] = 0.
            spectra.append(_as_spectrum(value))
        norm = float(np.prod(self.mattrs.meshsize, dtype='f8'))**2
        return (self._sums_dev(spectra) * norm).cpu().numpy()


    def _call_sugiyama(self, meshes, remove_zero):
        if len(meshes) != 3:
            raise ValueError('the sugiyama binning takes three meshes')
        spectra = []
        for mesh in meshes[:2]:
            value = mesh.value if isinstance(mesh, ComplexMeshField) else mesh
            _lib.require_device(value, 'mesh')
            if remove_zero:
                value = value.clone()
                value[0, 0, 0] = 0.
            spectra.append(_as_spectrum(value))
        third = meshes[2].value if isinstance(meshes[2], RealMeshField) else meshes[2]
        third = third.to(torch.float32).contiguous()
        if remove_zero:
            third = third - third.double().mean().float()
        cache = {}
        Q0 = self._ylm_stack(0, spectra[0], -1, 0, cache)
        Q1 = self._ylm_stack(1, spectra[1], -1, 0, cache)
        norm = float(np.prod(self.mattrs.meshsize, dtype='f8'))**2
        return (self._pair_sums(Q0, Q1, third[None])[:, 0] * norm).cpu().numpy()


def _2c(mesh):
    return mesh if isinstance(mesh, ComplexMeshField) else mesh.r2c()


def _2r(mesh):
    return mesh if isinstance(mesh, RealMeshField) else mesh.c2r()


def _mesh3_spectrum_sugiyama(meshes, bin, mattrs, los):
    """compute_mesh3_spectrum in the Sugiyama basis (mesh3.py:595-621), global or local line of
    sight.  With Y*_lm = sum_mu c_mu Y_l,mu (real harmonics) and ifftn(A [b] Y_l,mu) = i^(l mod 2)
    Q_b (mp_bispec_shells_ylm), each row (m1, m2, M, g) of each (l1, l2, L) contributes
    g N^6 sum_mu1,mu2 c1 c2 i^(l1 mod 2 + l2 mod 2) sum_x Q1 Q2 H, H = Y*_LM(los) rho_3 (global)
    or sum_mu3 c3 Y_L,mu3(xhat) rho_3 (local, positions from the origin); the real triple sums
    come from the triangle pass, one per (mu1, mu2), and the complex coefficients are applied
    on the host.  num_raw = real part (even l1 + l2) or imaginary part of the row sum / nmodes;
    norm = N^3 / Vcell^2."""
    local = isinstance(los, str) and los == 'local'
    vlos = None if local else _get_los_vector(los)
    device = bin._device
    spectra = [_as_spectrum(_2c(m).value) for m in meshes[:2]]
    rho = _2r(meshes[2]).value.to(torch.float32).contiguous()
    norm = float(np.prod(mattrs.meshsize, dtype='f8')) / float(np.prod(mattrs.cellsize))**2
    n6 = float(np.prod(mattrs.meshsize, dtype='f8'))**2
    attrs = dict(los='local' if local else vlos.tolist())
    cache, thirds = {}, {}
    if local:
        from .mesh2 import _mul_ylm
        offset = np.asarray(mattrs.boxcenter) - np.asarray(mattrs.boxsize) / 2.
        xs = [torch.from_numpy(np.arange(int(n)) * float(d) + float(o)).to(device)
              for n, d, o in zip(mattrs.meshsize, mattrs.cellsize, offset)]  # node positions (mesh.py:1594-1603)
    poles = []
    for ill, (ell1, ell2, ell3) in enumerate(bin.ells):
        rows = _triposh_rows(ell1, ell2, ell3, los_z=isinstance(los, str) and los == 'z')
        coef = {}  # (mu1, mu2) -> {row of H: complex coefficient}
        phase = 1j**(ell1 % 2 + ell2 % 2)
        for m1, m2, m3, g in rows:
            h = _ylm_components(ell3, m3) if local else {0: _complex_ylm_conj_at(ell3, m3, vlos)}
            for mu1, c1 in _ylm_components(ell1, m1).items():
                for mu2, c2 in _ylm_components(ell2, m2).items():
                    d = coef.setdefault((mu1, mu2), {})
                    for mu3, c3 in h.items():
                        j = mu3 + ell3 if local else 0
                        d[j] = d.get(j, 0.) + g * c1 * c2 * c3 * phase
        if local:
            if ell3 not in thirds:
                H = torch.empty((2 * ell3 + 1,) + tuple(rho.shape), dtype=torch.float32, device=device)
                for mu3 in range(-ell3, ell3 + 1):
                    _mul_ylm(rho, H[mu3 + ell3], mattrs, xs, ell3, mu3, False)
                thirds[ell3] = H
            H = thirds[ell3]
        else:
            H = rho[None]
        total = np.zeros(bin.nbins, dtype=complex)
        for (mu1, mu2), d in coef.items():
            if all(abs(c) < 1e-14 for c in d.values()):
                continue
            Q0 = bin._ylm_stack(0, spectra[0], ell1, mu1, cache)
            Q1 = bin._ylm_stack(1, spectra[1], ell2, mu2, cache)
            T = bin._pair_sums(Q0, Q1, H).cpu().numpy()
            for j, c in d.items():
                total += c * T[:, j]
        with np.errstate(invalid='ignore', divide='ignore'):
            total = total * n6 / bin.nmodes[ill]
        num = total.real if (ell1 + ell2) % 2 == 0 else total.imag
        poles.append(Mesh3SpectrumPole(k=bin.xavg, k_edges=bin.edges, nmodes=bin.nmodes[ill], num_raw=num, norm=norm,
                                       ell=(ell1, ell2, ell3), attrs=attrs, basis=bin.basis))
    return Mesh3SpectrumPoles(poles, attrs=attrs)


def compute_mesh3_spectrum(*meshes, bin=None, los='x'):
    """Bispectrum multipoles of one to three meshes (mesh3.py:509-629).  Scoccimarro basis,
    global line of sight: num_raw_ell3 = (2 ell3 + 1) sum over closed triangles of
    A1 A2 (A3 P_ell3(mu)), divided by the triangle counts; odd ell3 give 0 (the reference keeps
    the imaginary part of a real sum); local line of sight: see _mesh3_spectrum_local.  Sugiyama
    basis: see _mesh3_spectrum_sugiyama.  norm = N^3 / Vcell^2."""
    if not isinstance(bin, BinMesh3SpectrumPoles):
        raise TypeError('bin must be a BinMesh3SpectrumPoles')
    if isinstance(los, str) and los in ('firstpoint', 'endpoint'):
        raise ValueError(f"the bispectrum line of sight is 'local' or a global one, got {los!r} (mesh3.py:433-440)")
    meshes, _ = format_meshes(*meshes, nmeshes=3)
    mattrs = meshes[0].attrs
    if 'sugiyama' in bin.basis:
        return _mesh3_spectrum_sugiyama(meshes, bin, mattrs, los)
    if isinstance(los, str) and los == 'local':
        return _mesh3_spectrum_local(meshes, bin, mattrs)
    vlos = _get_los_vector(los)
    cs = [_2c(m) for m in meshes]
    spectra = [_as_spectrum(c.value) for c in cs]
    norm = float(np.prod(mattrs.meshsize, dtype='f8')) / float(np.prod(mattrs.cellsize))**2
    n6 = float(np.prod(mattrs.meshsize, dtype='f8'))**2
    attrs = dict(los=vlos.tolist())
    poles = []
    for ill, ell in enumerate(bin.ells):
        if ell % 2:
            num = torch.zeros(bin.nbins, dtype=torch.float64, device=bin._device)
        else:
            sums = bin._sums_dev(spectra, ells=(0, 0, ell), los=vlos)
            num = (2 * ell + 1) * sums * n6 / bin._nmodes_dev
        poles.append(Mesh3SpectrumPole(k=bin.xavg, k_edges=bin.edges, nmodes=bin.nmodes[ill], num_raw=Deferred(num), norm=norm, ell=ell,
                                       attrs=attrs, basis=bin.basis))
    return Mesh3SpectrumPoles(poles, attrs=attrs)


def _mesh3_spectrum_local(meshes, bin, mattrs):
    """Local line of sight (mesh3.py:563-576): the third side becomes
    C_ell = sum_m r2c(rho_3 Y_ell,m(xhat)) Y_ell,m(khat), xhat the direction of the node position
    from the origin, and num_raw_ell = 4 pi sum over closed triangles of A1 A2 C_ell divided by the
    triangle counts.  For ell = 0, 4 pi Y_00^2 = 1 and the third side is r2c(rho_3) itself."""
    spectra = [_as_spectrum(_2c(m).value) for m in meshes[:2]]
    rho3 = meshes[2] if isinstance(meshes[2], RealMeshField) else meshes[2].c2r()
    norm = float(np.prod(mattrs.meshsize, dtype='f8')) / float(np.prod(mattrs.cellsize))**2
    n6 = float(np.prod(mattrs.meshsize, dtype='f8'))**2
    attrs = dict(los='local')
    poles = []
    for ill, ell in enumerate(bin.ells):
        if ell == 0:
            C, scale = _2c(rho3).value, 1.
        else:
            C = mattrs.create('complex', device=bin._device).value
            _ylm_accumulate(rho3, ell, C)
            scale = 4. * np.pi
        sums = bin._sums_dev(spectra + [_as_spectrum(C)], ells=(0, 0, 0))
        num = scale * sums * n6 / bin._nmodes_dev
        poles.append(Mesh3SpectrumPole(k=bin.xavg, k_edges=bin.edges, nmodes=bin.nmodes[ill], num_raw=Deferred(num), norm=norm, ell=ell,
                                       attrs=attrs, basis=bin.basis))
    return Mesh3SpectrumPoles(poles, attrs=attrs)


class BinMesh3CorrelationPoles:
    """Binning operator for the three-point correlation function (mesh3.py:316-409): separation
    bins of the full real grid per side (every cell weighted 1, as BinMesh2CorrelationPoles),
    pairs (Sugiyama) or triangles (Scoccimarro) that pass the mask, ``nmodes`` the products of
    the per-side cell counts, ``xavg`` the per-side mean |s|.  The call filters a mesh's half
    spectrum by j_l(|k| s_b) for each bin b (s_b its mean |s|, mesh3.py:383-395) instead of a
    shell indicator, optionally restricted to klimit[0] <= |k| < klimit[1] (True: up to the
    smallest Nyquist wavenumber).  Only the Sugiyama basis has an estimator, as in the reference
    (compute_mesh3_correlation raises for Scoccimarro, mesh3.py:680-682)."""

    def __init__(self, mattrs, edges=None, ells=0, basis='sugiyama', batch_size=None, buffer_size=0, mask_edges=None, klimit=None,
                 device=None):
        from .mesh2 import BinMesh2CorrelationPoles
        if not isinstance(mattrs, MeshAttrs):
            mattrs = mattrs.attrs
        if basis not in _BASES:
            raise ValueError(f'basis must be one of {_BASES}, got {basis!r}')
        device = device or _default_device()
        ndim = 2 if 'sugiyama' in basis else 3
        diagonal = 'diagonal' in basis
        self.mattrs, self.basis, self._device = mattrs, basis, device
        self.batch_size = 1 if batch_size is None else int(batch_size)
        self.buffer_size = int(buffer_size)
        self.ells = _format_ells(ells, basis=basis)
        if edges is None:
            edges = {}
        if not isinstance(edges, (tuple, list)):
            edges = [edges] * ndim
        axes = []
        for e in list(edges)[:ndim]:
            axis = BinMesh2CorrelationPoles(mattrs, edges=e, device=device)
            same = [a for a in axes if np.array_equal(a._edges1d, axis._edges1d)]
            axes.append(same[0] if same else axis)
        axes = axes + [axes[-1]] * (ndim - len(axes))
        self._axes = axes
        self.edges1d = tuple(a.edges for a in axes)
        self.nmodes1d = tuple(a.nmodes for a in axes)
        self.xavg1d = tuple(a.xavg for a in axes)
        nyq = float(np.min(mattrs.cellsize)) * np.min(mattrs.meshsize) / 2.
        mask = _mask_function(_default_mask(basis) if mask_edges is None else mask_edges, nyq)
        lo = _product([e[:, 0] for e in self.edges1d], diagonal)
        hi = _product([e[:, 1] for e in self.edges1d], diagonal)
        edges3 = np.stack([lo, hi], axis=-1)
        keep = np.asarray(mask(*[edges3[:, i, :] for i in range(ndim)]), dtype=bool)
        self.edges = edges3[keep]
        self.xavg = _product(self.xavg1d, diagonal)[keep]
        self._iedges = _product([np.arange(len(x)) for x in self.xavg1d], diagonal)[keep].astype('i4')
        if not len(self._iedges):
            raise ValueError('no pair passes the edge mask')
        nmodes = np.prod(np.stack([np.asarray(self.nmodes1d[a])[self._iedges[:, a]] for a in range(ndim)]), axis=0)
        self.nmodes = [nmodes] * len(self.ells)
        if isinstance(klimit, bool):
            klimit = (0., float(np.min(mattrs.knyq))) if klimit else None
        self.klimit = klimit
        self.wmodes = None

    @property
    def nbins(self):
        return len(self.xavg)

    _pair_sums = BinMesh3SpectrumPoles._pair_sums

    def _bessel_stack(self, side, spec, ellj, ell, mu, cache):
        """(R_h, R_a): real meshes per separation bin b of the side with
        ifftn(A j_ellj(|k| s_b) [klimit] Y_l,mu(khat)) = R_h[b] + i R_a[b] (ell = -1: no harmonic)."""
        key = (spec[0].data_ptr(), side, ellj, ell, mu)
        if key in cache:
            return cache[key]
        n = _lib.i32x3(self.mattrs.meshsize)
        shape = tuple(int(x) for x in self.mattrs.meshsize)
        s = np.ascontiguousarray(self.xavg1d[side], dtype='f8')
        ws = _lib.workspace(_lib.load().mp_bispec_shells_workspace_bytes(n), self._device)
        rh = torch.empty((len(s),) + shape, dtype=torch.float32, device=self._device)
        ra = torch.empty_like(rh)
        klo, khi = (-1., np.inf) if self.klimit is None else (float(self.klimit[0]), float(self.klimit[1]))
        _lib.call('mp_bessel_shells_ylm', _lib.stream_handle(self._device), _lib.ptr(spec[0]), spec[1], n, _lib.f64x3(self.mattrs.kfun),
                  s.ctypes.data_as(_lib.P_F64), len(s), int(ell), int(mu), int(ellj), klo, khi, _lib.ptr(rh), _lib.ptr(ra), _lib.ptr(ws),
                  ws.numel())
        cache[key] = (rh, ra)
        return cache[key]

    def _complex_pair_sums(self, F0, F1, H):
        """sum_x F0[b1] F1[b2] H[j] for complex F = (R_h, R_a): (npairs, len(H)) complex128."""
        hh = self._pair_sums(F0[0], F1[0], H)
        aa = self._pair_sums(F0[1], F1[1], H)
        ha = self._pair_sums(F0[0], F1[1], H)
        ah = self._pair_sums(F0[1], F1[0], H)
        return torch.complex(hh - aa, ha + ah)

    def __call__(self, *meshes, ell=(0, 0, 0), remove_zero=False):
        """i^(l1 + l2) / Vcell^2 sum_x c2r(A1 j_l1(|k| s_b1)) c2r(A2 j_l2(|k| s_b2)) m3 for every pair
        (mesh3.py:362-395), complex; remove_zero sets the k = 0 mode of the first two meshes to 0
        and subtracts the mean of the third."""
        if 'sugiyama' not in self.basis:
            raise NotImplementedError('the three-point correlation function is built in the Sugiyama basis only')
        if len(meshes) != 3:
            raise ValueError('the binning takes three meshes')
        spectra = []
        for mesh in meshes[:2]:
            value = mesh.value if isinstance(mesh, ComplexMeshField) else mesh
            _lib.require_device(value, 'mesh')
            if remove_zero:
                value = value.clone()
                value[0, 0, 0] = 0.
            spectra.append(_as_spectrum(value))
        third = meshes[2].value if isinstance(meshes[2], RealMeshField) else meshes[2]
        third = third.to(torch.float32).contiguous()
        if remove_zero:
            third = third - third.double().mean().float()
        cache = {}
        ell = tuple(int(v) for v in ell)[:2]
        F0 = self._bessel_stack(0, spectra[0], ell[0], -1, 0, cache)
        F1 = self._bessel_stack(1, spectra[1], ell[1], -1, 0, cache)
        norm = (1j)**sum(ell) / float(np.prod(self.mattrs.cellsize))**2
        return (self._complex_pair_sums(F0, F1, third[None])[:, 0] * norm).cpu().numpy()


def compute_mesh3_correlation(*meshes, bin=None, los='x'):
    """Three-point correlation multipoles in the Sugiyama basis (mesh3.py:632-727): for each
    (l1, l2, L) and each row (m1, m2, M, g) of nonzero Gaunt coefficient,
    g i^(l1 + l2) / Vcell^2 sum_x c2r(A1 Y*_l1m1 j_l1(|k| s1)) c2r(A2 Y*_l2m2 j_l2(|k| s2)) Y*_LM(los) rho_3,
    real part; norm = N^3 / Vcell^2.  The complex harmonics are combinations of real ones and each
    filtered mesh is R_h + i R_a on the full complex grid (mp_bessel_shells_ylm), so the sums are
    real triple sums from the triangle pass with complex coefficients applied on the host."""
    from .types import Mesh3CorrelationPole, Mesh3CorrelationPoles
    if not isinstance(bin, BinMesh3CorrelationPoles):
        raise TypeError('bin must be a BinMesh3CorrelationPoles')
    if 'scoccimarro' in bin.basis:
        raise NotImplementedError('the three-point correlation function in the Scoccimarro basis (mesh3.py:680-682)')
    if isinstance(los, str) and los in ('firstpoint', 'endpoint'):
        raise ValueError(f"the line of sight is 'local' or a global one, got {los!r} (mesh3.py:433-440)")
    meshes, _ = format_meshes(*meshes, nmeshes=3)
    mattrs = meshes[0].attrs
    local = isinstance(los, str) and los == 'local'
    vlos = None if local else _get_los_vector(los)
    device = bin._device
    spectra = [_as_spectrum(_2c(m).value) for m in meshes[:2]]
    rho = _2r(meshes[2]).value.to(torch.float32).contiguous()
    norm = float(np.prod(mattrs.meshsize, dtype='f8')) / float(np.prod(mattrs.cellsize))**2
    vcell2 = float(np.prod(mattrs.cellsize))**2
    attrs = dict(los='local' if local else vlos.tolist())
    cache, thirds = {}, {}
    if local:
        from .mesh2 import _mul_ylm
        offset = np.asarray(mattrs.boxcenter) - np.asarray(mattrs.boxsize) / 2.
        xs = [torch.from_numpy(np.arange(int(n)) * float(d) + float(o)).to(device)
              for n, d, o in zip(mattrs.meshsize, mattrs.cellsize, offset)]
    poles = []
    for ill, (ell1, ell2, ell3) in enumerate(bin.ells):
        coef = {}
        for m1, m2, m3, g in _triposh_rows(ell1, ell2, ell3, los_z=isinstance(los, str) and los == 'z'):
            h = _ylm_components(ell3, m3) if local else {0: _complex_ylm_conj_at(ell3, m3, vlos)}
            for mu1, c1 in _ylm_components(ell1, m1).items():
                for mu2, c2 in _ylm_components(ell2, m2).items():
                    d = coef.setdefault((mu1, mu2), {})
                    for mu3, c3 in h.items():
                        j = mu3 + ell3 if local else 0
                        d[j] = d.get(j, 0.) + g * c1 * c2 * c3
        if local:
            if ell3 not in thirds:
                H = torch.empty((2 * ell3 + 1,) + tuple(rho.shape), dtype=torch.float32, device=device)
                for mu3 in range(-ell3, ell3 + 1):
                    _mul_ylm(rho, H[mu3 + ell3], mattrs, xs, ell3, mu3, False)
                thirds[ell3] = H
            H = thirds[ell3]
        else:
            H = rho[None]
        total = np.zeros(bin.nbins, dtype=complex)
        for (mu1, mu2), d in coef.items():
            if all(abs(c) < 1e-14 for c in d.values()):
                continue
            F0 = bin._bessel_stack(0, spectra[0], ell1, ell1, mu1, cache)
            F1 = bin._bessel_stack(1, spectra[1], ell2, ell2, mu2, cache)
            T = bin._complex_pair_sums(F0, F1, H).cpu().numpy()
            for j, c in d.items():
                total += c * T[:, j]
        num = ((1j)**(ell1 + ell2) * total / vcell2).real
        poles.append(Mesh3CorrelationPole(s=bin.xavg, s_edges=bin.edges, nmodes=bin.nmodes[ill], num_raw=num, norm=norm,
                                          ell=(ell1, ell2, ell3), attrs=attrs, basis=bin.basis))
    return Mesh3CorrelationPoles(poles, attrs=attrs)


def compute_mesh3(*meshes, bin=None, los='z'):
    """Dispatch on the binning operator (mesh3.py:475-506)."""
    if isinstance(bin, BinMesh3SpectrumPoles):
        return compute_mesh3_spectrum(*meshes, bin=bin, los=los)
    if isinstance(bin, BinMesh3CorrelationPoles):
        return compute_mesh3_correlation(*meshes, bin=bin, los=los)
    raise TypeError(f'bin must be a BinMesh3SpectrumPoles or BinMesh3CorrelationPoles, not {type(bin)}')


def compute_box3_normalization(*inputs, bin=None):
    """Constant-density normalization of the bispectrum (mesh3.py:730-735): the box
    normalization of the three (padded) inputs, once per ell when bin is given."""
    inputs, _ = format_meshes(*inputs, nmeshes=3)
    norm = compute_box_normalization(*inputs)
    if bin is not None:
        return [norm] * len(bin.ells)
    return norm


def compute_fkp3_normalization(*fkps, bin=None, cellsize=10., split=None, fields=None, **kwargs):
    """FKP normalization of the bispectrum (mesh3.py:738-783), split=None: alpha, the product
    over the three (padded) inputs of sum(data w) / sum(randoms w) (1 for a ParticleField),
    times compute_normalization of the three randoms, CIC-painted on a mesh regridded to
    ``cellsize`` (sum R1 R2 R3 / cellvolume^2).  Once per ell when bin is given."""
    if split is not None:
        raise NotImplementedError('split normalisation needs the reference random streams; use split=None')
    kwargs.update(cellsize=cellsize)
    fkps, fields = format_meshes(*fkps, fields=fields, nmeshes=3)
    alpha = 1.
    for fkp in fkps:
        if isinstance(fkp, FKPField):
            alpha = alpha * (fkp.data.sum() / fkp.randoms.sum())
    randoms = [fkp.randoms if isinstance(fkp, FKPField) else fkp for fkp in fkps]
    norm = _deferred(alpha * compute_normalization(*randoms, **kwargs))
    if bin is not None:
        return [norm] * len(bin.ells)
    return norm


def compute_fkp3_shotnoise(*fkps, bin=None, los='z', resampler='cic', interlacing=False, compensate=True, fields=None, **kwargs):
    """Bispectrum shot noise, Scoccimarro basis, one field, global or local line of sight
    (mesh3.py:999-1062, _compute_scoccimarro_S 786-852; eq. 58 of arXiv:1506.02729): with
    A = paint(w) conj(paint(w^2)) (both k = 0 modes removed) and bin_a the shell means of side a,
    ell = 0: sum_a bin_a(A)[b_a] - 2 sum w^3; ell > 0: (2 ell + 1) [bin_3(A P_ell)[b3]
    + 4 pi sum_m sum_{a < 3} bin_a(A P_ell Y_ell,m)[b_a] bin_3(Y_ell,m)[b3]].  Returns one host
    array per ell (real part for even ell, imaginary part for odd ell).  los='local' replaces A P_ell by
    H_ell(paint(w)) conj(paint(w^2)) in the first term and paint(w) conj(H_ell(paint(w^2))) in the
    second, H_ell(c) = 4 pi / (2 ell + 1) sum_m r2c(c2r(c) Y_lm(xhat)) Y_lm(khat) (mesh3.py:823-834)."""
    if not isinstance(bin, (BinMesh3SpectrumPoles, BinMesh3CorrelationPoles)):
        raise TypeError('bin must be a BinMesh3SpectrumPoles or BinMesh3CorrelationPoles')
    if isinstance(los, str) and los in ('firstpoint', 'endpoint'):
        raise ValueError(f"the bispectrum line of sight is 'local' or a global one, got {los!r} (mesh3.py:433-440)")
    if isinstance(bin, BinMesh3CorrelationPoles) and 'scoccimarro' in bin.basis:
        raise NotImplementedError('the Scoccimarro shot noise of the three-point correlation function (mesh3.py:1064-1066)')
    local = isinstance(los, str) and los == 'local'
    fkps, fields = format_meshes(*fkps, fields=fields, nmeshes=3)
    if fields[2] == fields[1] + 1 == fields[0] + 2:
        return tuple(np.zeros(bin.nbins) for _ in bin.ells)
    if 'sugiyama' in bin.basis:
        particles = []
        for fkp, f in zip(fkps, fields):
            particles.append(particles[f] if f < len(particles) else (fkp.particles if isinstance(fkp, FKPField) else fkp))
        return _sugiyama_shotnoise(particles, fields, bin, los, dict(kwargs, resampler=resampler, interlacing=interlacing,
                                                                         compensate=compensate))
    if not (fields[0] == fields[1] == fields[2]):
        raise NotImplementedError('the Scoccimarro shot noise takes one field (mesh3.py:1050-1051)')
    particles = fkps[0].particles if isinstance(fkps[0], FKPField) else fkps[0]
    vlos = _get_los_vector('z' if local else los)
    mattrs = particles.attrs
    kw = dict(kwargs, resampler=resampler, interlacing=interlacing, compensate=compensate, out='complex')
    fw = particles.paint(**kw).at[0, 0, 0].set(0.)
    w = particles.weights
    fw2 = particles.clone(weights=w * w).paint(**kw).at[0, 0, 0].set(0.)
    cw, cw2 = fw.value, fw2.value
    sumw3 = float(_sum_f64(w * w, w).cpu().numpy()) if w.dtype == torch.float32 else float(torch.sum(w.double()**3).cpu())
    a, pa = _as_spectrum(cw)
    b, pb = _as_spectrum(cw2)
    device = a.device
    n = _lib.i32x3(mattrs.meshsize)
    A = _empty_spectrum(mattrs.meshsize, device)
    _lib.call('mp_spectrum_product', _lib.stream_handle(device), _lib.ptr(b), pb, _lib.ptr(a), pa, n, _lib.ptr(A), _spectrum_pitch(A))
    axes = bin._axes
    ie = bin._iedges
    with np.errstate(invalid='ignore', divide='ignore'):
        def shells(axis, meshes, prods, los_=(0., 0., 0.)):
            return axes[axis]._products_dev(meshes, prods, los=los_).cpu().numpy() / axes[axis].nmodes
        out = []
        for ell in bin.ells:
            if ell == 0:
                sn = sum(shells(ax, [A], [(-1, 0, 0, False)])[0][ie[:, ax]] for ax in range(3)) - 2. * sumw3
            else:
                Al = A
                if local:
                    # mesh3.py:823-834: P_ell(mu) c -> (4 pi / (2 ell + 1)) sum_m r2c(c2r(c) Y_lm(xhat)) Y_lm(khat);
                    # the factor is applied to sn at the end, both terms being linear in it
                    H = mattrs.create('complex', device=device).value
                    _ylm_accumulate(fw.c2r(), ell, H)
                    h, ph = _as_spectrum(H)
                    P = _empty_spectrum(mattrs.meshsize, device)
                    _lib.call('mp_spectrum_product', _lib.stream_handle(device), _lib.ptr(b), pb, _lib.ptr(h), ph, n, _lib.ptr(P),
                              _spectrum_pitch(P))
                    sn = shells(2, [P], [(-1, 0, 0, False)])[0][ie[:, 2]]
                    _ylm_accumulate(fw2.c2r(), ell, H)  # overwrites H
                    Al = _empty_spectrum(mattrs.meshsize, device)
                    _lib.call('mp_spectrum_product', _lib.stream_handle(device), _lib.ptr(h), ph, _lib.ptr(a), pa, n, _lib.ptr(Al),
                              _spectrum_pitch(Al))
                    del H, P
                else:
                    sn = shells(2, [A], [(-1, 0, ell, False)], vlos)[0][ie[:, 2]]
                W = _empty_spectrum(mattrs.meshsize, device)
                V = _empty_spectrum(mattrs.meshsize, device)
                t = axes[0]._tables
                for m in range(-ell, ell + 1):
                    _lib.call('mp_ylm_legendre_spectrum', _lib.stream_handle(device), n, _lib.ptr(t['kx']), _lib.ptr(t['ky']),
                              _lib.ptr(t['kz']), int(ell), int(m), 0 if local else int(ell), _lib.f64x3(vlos), _lib.ptr(W),
                              _spectrum_pitch(W))
                    _lib.call('mp_ylm_legendre_spectrum', _lib.stream_handle(device), n, _lib.ptr(t['kx']), _lib.ptr(t['ky']),
                              _lib.ptr(t['kz']), int(ell), int(m), 0, _lib.f64x3(vlos), _lib.ptr(V), _spectrum_pitch(V))
                    b3 = shells(2, [V], [(-1, 0, 0, False)])[0][ie[:, 2]]
                    # conj(W) A with W real: A P_ell Y_lm binned on the first two sides
                    sn = sn + 4. * np.pi * sum(shells(ax, [W, Al], [(0, 1, 0, False)])[0][ie[:, ax]] for ax in range(2)) * b3
                sn = (4. * np.pi if local else 2 * ell + 1) * sn
            out.append(np.real(sn) if ell % 2 == 0 else np.imag(sn))
    return tuple(out)


def _shotnoise_window(mattrs, resampler, interlacing, compensate, device):
    """compensate_shotnoise of mesh3.py:1081-1089 on the half spectrum: with interlacing (n_i >= 2)
    1 if compensate else 1 / compensate(1, kcirc)^2, without it aliasing_shotnoise(1, kcirc)
    times compensate(1, kcirc)^2 if compensate; a complex64 half spectrum (imaginary part 0)."""
    from . import resamplers
    r = resamplers.get_resampler(resampler)
    kcirc = mattrs.kcoords(kind='circular', sparse=True, device=device)
    one = torch.ones((), dtype=torch.float32, device=device)
    if max(int(interlacing), 1) >= 2:
        w = one if compensate else 1. / r.compensate(one, kcirc)**2
    else:
        w = r.aliasing_shotnoise(one, kcirc) * (r.compensate(one, kcirc)**2 if compensate else 1.)
    out = _empty_spectrum(mattrs.meshsize, device)
    out.copy_(torch.broadcast_to(w, out.shape).to(torch.complex64))
    return out


def _sugiyama_shotnoise(particles, fields, bin, los, paint_kw):
    """compute_fkp3_shotnoise in the Sugiyama basis (mesh3.py:1091-1180, eqs. 45-46 of
    arXiv:1803.02132), spectrum, global or local line of sight.  With rho_c the mean-subtracted
    configuration-space paint, C the complex paint with its k = 0 mode removed, W the shot-noise
    window of _shotnoise_window, Y the reduced complex harmonics and Y* their conjugates:
      S111 (all three fields equal): s111_lm = sum_x paint(w^3)(x) Y*_lm(xhat); the (0, 0, 0)
        multipole gets s111_00;
      S122 (fields 2 and 3 equal; multipoles (l, 0, l)): (2l + 1) times the full-grid shell mean of
        sum_m Y_lm(khat) [C_w1 conj(FFT(rho_c[w2 w3] Y_lm(los))) - s111_lm W], taken on side 1 (and
        S121 likewise on side 2 for (0, l, l) when fields 1 and 3 are equal, mesh3.py:855-883);
      S113 (fields 1 and 2 equal): i^(l1 + l2) sum over the rows (m1, m2, M, g) of
        g sum_s ifftn(FFT(rho_c[w3] Y*_LM(los)) conj(C_w1w2) - s111_LM W)(s) Y*_l1m1(shat)
        Y*_l2m2(shat) j_l1(|s| k1) j_l2(|s| k2) at the pairs' mean |k| (mesh3.py:886-915).
    With real harmonics Y_l,mu: sum_m Y_lm(khat) Y*_lm(u) = 4 pi/(2l+1) sum_mu Y_l,mu(khat) Y_l,mu(u)
    (= P_l(khat.u) for nonzero vectors), so the S122 sums are real-harmonic spectra binned with the
    hermitian weights (real part for even l, i times the imaginary part for odd l, the full-grid
    sum of an (anti)hermitian product); in S113 every ifftn is of a hermitian product and the
    angular factor is a sum of products of two real harmonics of shat, all on the device.  Returns
    one host array per multipole (real part for even l1 + l2, imaginary part otherwise)."""
    from .mesh2 import _mul_ylm, _svectors, BinMesh2CorrelationPoles
    mattrs = particles[0].attrs
    device = bin._device
    ells = bin.ells
    correlation = isinstance(bin, BinMesh3CorrelationPoles)
    vc2 = float(np.prod(mattrs.cellsize))**2
    local = isinstance(los, str) and los == 'local'
    vlos = None if local else _get_los_vector(los)
    n = _lib.i32x3(mattrs.meshsize)
    shape = tuple(int(x) for x in mattrs.meshsize)
    nbins = bin.nbins
    out = [np.zeros(nbins, dtype=complex) for _ in ells]
    offset = np.asarray(mattrs.boxcenter) - np.asarray(mattrs.boxsize) / 2.
    xs = [torch.from_numpy(np.arange(int(nn)) * float(d) + float(o)).to(device) for nn, d, o in zip(mattrs.meshsize, mattrs.cellsize, offset)]
    ss = [torch.from_numpy(v).to(device) for v in _svectors(mattrs)]
    W = _shotnoise_window(mattrs, paint_kw['resampler'], paint_kw['interlacing'], paint_kw['compensate'], device)
    stream = _lib.stream_handle(device)

    def paint(p, weights=None, out_kind='real'):
        q = p if weights is None else p.clone(weights=weights)
        return q.paint(**paint_kw, out=out_kind)

    def centered(mesh):
        v = mesh.value.to(torch.float32)
        return (v - v.double().mean().float()).contiguous()

    def complex_zero_dc(mesh):
        v = mesh.value.clone()
        v[0, 0, 0] = 0.
        return v

    def prod_spectrum(a, b):
        """conj(a) b on half spectra."""
        pa, pb = _as_spectrum(a), _as_spectrum(b)
        res = _empty_spectrum(mattrs.meshsize, device)
        _lib.call('mp_spectrum_product', stream, _lib.ptr(pa[0]), pa[1], _lib.ptr(pb[0]), pb[1], n, _lib.ptr(res), _spectrum_pitch(res))
        return res

    def c2r(a):
        from .mesh import _c2r
        return _c2r(a, mattrs.meshsize)

    # S111: s111_lm = sum_x paint(w^3) Y*_lm(xhat), through the real harmonics
    uells = sorted({l for ell in ells for l in ell})
    S = {l: np.zeros(2 * l + 1) for l in uells}  # S[l][mu + l] = sum_x rho3 Y_l,mu(xhat)
    all_equal = fields[0] == fields[1] == fields[2]
    if all_equal:
        rho3 = paint(particles[0], particles[0].weights**3).value.to(torch.float32).contiguous()
        tmp = torch.empty_like(rho3)
        for l in uells:
            for mu in range(-l, l + 1):
                _mul_ylm(rho3, tmp, mattrs, xs, l, mu, False)
                S[l][mu + l] = float(tmp.double().sum())
        del rho3, tmp

    def s111(l, m):
        return sum(c * S[l][mu + l] for mu, c in _ylm_components(l, m).items())

    if all_equal and (0, 0, 0) in ells:
        if correlation:  # both separation bins must contain s = 0 (mesh3.py:1075-1077)
            mask = np.all((bin.edges[..., 0] <= 0.) & (bin.edges[..., 1] >= 0.), axis=1)
            out[ells.index((0, 0, 0))] += s111(0, 0) * mask / vc2
        else:
            out[ells.index((0, 0, 0))] += s111(0, 0)

    def s122_correlation(ps, ells1, axis):
        """(2l+1) separation-bin means on side axis of sum_m Y_lm(shat) c2r(C conj(FFT(rho_c
        Y_lm(los))) - s111_lm W) (mesh3.py:918-944), per l in ells1: by the addition theorem
        4 pi/(2l+1) sum_mu Y_l,mu(shat) [c2r(C conj(FFT(rho_c Y_l,mu(xhat))))  (local) or
        Y_l,mu(los) c2r(C conj(FFT(rho_c))) (global)] - 4 pi/(2l+1) sum_mu S_l,mu Y_l,mu(shat) c2r(W)."""
        side = bin._axes[axis]
        rc = centered(paint(ps[1], ps[1].weights * ps[2].weights))
        C = complex_zero_dc(paint(ps[0], out_kind='complex'))
        R = _as_spectrum(_2c(RealMeshField(rc, attrs=mattrs)).value)[0]
        omega = c2r(W) if all_equal else None
        res = {}
        for l in ells1:
            carry = torch.zeros(shape, dtype=torch.float32, device=device)
            tmp = torch.empty(shape, dtype=torch.float32, device=device)
            if local and l > 0:
                for mu in range(-l, l + 1):
                    _mul_ylm(rc, tmp, mattrs, xs, l, mu, False)
                    t = c2r(prod_spectrum(_2c(RealMeshField(tmp, attrs=mattrs)).value, C))
                    _mul_ylm(t, tmp, mattrs, ss, l, mu, False)
                    carry.add_(tmp)
            else:
                t = c2r(prod_spectrum(R, C))
                for mu in range(-l, l + 1):
                    y = _real_ylm_at(l, mu, vlos) if l else 1. / math.sqrt(4. * math.pi)
                    if y == 0.:
                        continue
                    _mul_ylm(t, tmp, mattrs, ss, l, mu, False)
                    carry.add_(tmp, alpha=y)
            if all_equal:
                for mu in range(-l, l + 1):
                    if S[l][mu + l] == 0.:
                        continue
                    _mul_ylm(omega, tmp, mattrs, ss, l, mu, False)
                    carry.add_(tmp, alpha=-float(S[l][mu + l]))
            carry.mul_(4. * math.pi / (2 * l + 1))
            sums = side._sums_dev(carry, [0], [False])[0].cpu().numpy()
            with np.errstate(invalid='ignore', divide='ignore'):
                res[l] = (2 * l + 1) * sums / np.asarray(bin.nmodes1d[axis], dtype='f8')
        return res

    def s122(ps, ells1, axis):
        """(2l+1) full-grid shell means on side axis, per l in ells1 (mesh3.py:855-883)."""
        side = bin._axes[axis]
        rc = centered(paint(ps[1], ps[1].weights * ps[2].weights))
        C = complex_zero_dc(paint(ps[0], out_kind='complex'))
        res = {}
        R = _as_spectrum(_2c(RealMeshField(rc, attrs=mattrs)).value)[0]
        for l in ells1:
            if local and l > 0:
                G = mattrs.create('complex', device=device).value
                _ylm_accumulate(RealMeshField(rc, attrs=mattrs), l, G)
                t1 = side._products_dev([G, C], [(0, 1, 0, False)])[0] * (4. * np.pi / (2 * l + 1))
                del G
            else:  # global line of sight, or l = 0 where 4 pi Y_00(khat) Y_00(xhat) = 1
                t1 = side._products_dev([R, C], [(0, 1, l, False)], los=vlos if l else (0., 0., 0.))[0]
            t1 = t1.cpu().numpy()
            full = t1.real + 0j if l % 2 == 0 else 1j * t1.imag
            if all_equal and l % 2 == 0:
                V = torch.zeros_like(W)
                Y = _empty_spectrum(mattrs.meshsize, device)
                t = side._tables
                for mu in range(-l, l + 1):
                    if S[l][mu + l] == 0.:
                        continue
                    _lib.call('mp_ylm_legendre_spectrum', stream, n, _lib.ptr(t['kx']), _lib.ptr(t['ky']), _lib.ptr(t['kz']), int(l), int(mu),
                              0, _lib.f64x3((0., 0., 1.)), _lib.ptr(Y), _spectrum_pitch(Y))
                    V += Y * float(S[l][mu + l])
                V = V * W * (4. * np.pi / (2 * l + 1))
                full = full - side._products_dev([V], [(-1, 0, 0, False)])[0].cpu().numpy().real
            with np.errstate(invalid='ignore', divide='ignore'):
                res[l] = (2 * l + 1) * full / bin.nmodes1d[axis]
        return res

    def s122_mask(axis):  # the other side's bin must contain s = 0 (mesh3.py:1098-1099)
        if not correlation:
            return 1.
        return ((bin.edges[:, 1 - axis, 0] <= 0.) & (bin.edges[:, 1 - axis, 1] >= 0.)) / vc2

    if fields[1] == fields[2]:
        ells1 = [ell[0] for ell in ells if ell[2] == ell[0] and ell[1] == 0]
        if ells1:
            res = (s122_correlation if correlation else s122)(particles, ells1, 0)
            for ill, ell in enumerate(ells):
                if ell[2] == ell[0] and ell[1] == 0:
                    out[ill] += res[ell[0]][bin._iedges[:, 0]] * s122_mask(0)
    if fields[0] == fields[2]:
        ells2 = [ell[1] for ell in ells if ell[2] == ell[1] and ell[0] == 0]
        if ells2:
            res = (s122_correlation if correlation else s122)([particles[1], particles[0], particles[2]], ells2, 1)
            for ill, ell in enumerate(ells):
                if ell[2] == ell[1] and ell[0] == 0:
                    out[ill] += res[ell[1]][bin._iedges[:, 1]] * s122_mask(1)
    if fields[0] == fields[1]:
        rc = centered(paint(particles[2]))
        Cw2 = complex_zero_dc(paint(particles[0], particles[0].weights * particles[1].weights, out_kind='complex'))
        omega = c2r(W)
        k1 = torch.from_numpy(np.ascontiguousarray(bin.xavg1d[0], dtype='f8')).to(device)
        k2 = torch.from_numpy(np.ascontiguousarray(bin.xavg1d[1], dtype='f8')).to(device)
        taus = {}

        def tau(mu):  # c2r(conj(C_w1w2) FFT(rho_c Y_L,mu(xhat))) (local) or c2r(conj(C_w1w2) FFT(rho_c)) (global)
            if mu not in taus:
                if mu is None:
                    taus[mu] = c2r(prod_spectrum(Cw2, _2c(RealMeshField(rc, attrs=mattrs)).value))
                else:
                    L, m = mu
                    tmp = torch.empty_like(rc)
                    _mul_ylm(rc, tmp, mattrs, xs, L, m, False)
                    taus[mu] = c2r(prod_spectrum(Cw2, _2c(RealMeshField(tmp, attrs=mattrs)).value))
            return taus[mu]

        ws = _lib.workspace(_lib.load().mp_hankel_pairs_workspace_bytes(min(bin.nbins, 1024)), device)
        if correlation:
            # sums over the intersection of the two separation bins of each pair, by fine bins on
            # the unique edges and a rebinning matrix, divided by both cell counts (mesh3.py:1118-1128, 955-957)
            lo = np.max(bin.edges[..., 0], axis=1)
            hi = np.min(bin.edges[..., 1], axis=1)
            empty = hi <= lo
            lo, hi = np.where(empty, 0., lo), np.where(empty, 0., hi)
            fine = np.unique(np.concatenate([lo, hi]))
            rebin = ((fine[:-1][None, :] >= lo[:, None]) & (fine[1:][None, :] <= hi[:, None])).astype('f8')
            inter_bin = BinMesh2CorrelationPoles(mattrs, edges=fine, device=device) if len(fine) >= 2 else None
            counts = (np.asarray(bin.nmodes1d[0], dtype='f8')[bin._iedges[:, 0]] * np.asarray(bin.nmodes1d[1], dtype='f8')[bin._iedges[:, 1]])

        def inter_sums(F):
            if inter_bin is None:
                return np.zeros(bin.nbins)
            sums = inter_bin._sums_dev(F, [0], [False])[0].cpu().numpy()
            with np.errstate(invalid='ignore', divide='ignore'):
                return (rebin @ sums) / counts

        def hankel(F, l1, l2):
            res = np.empty(bin.nbins)
            for lo in range(0, bin.nbins, 1024):
                pr = np.ascontiguousarray(bin._iedges[lo:lo + 1024], dtype='i4')
                pd = torch.from_numpy(pr.ravel()).to(device)
                o = torch.empty(len(pr), dtype=torch.float64, device=device)
                _lib.call('mp_hankel_pairs', stream, _lib.ptr(F), n, _lib.ptr(ss[0]), _lib.ptr(ss[1]), _lib.ptr(ss[2]), int(l1), _lib.ptr(k1),
                          len(bin.xavg1d[0]), int(l2), _lib.ptr(k2), len(bin.xavg1d[1]), _lib.ptr(pd), len(pr), _lib.ptr(o), _lib.ptr(ws),
                          ws.numel())
                res[lo:lo + len(pr)] = o.cpu().numpy()
            return res

        for ill, (l1, l2, L) in enumerate(ells):
            coef = {}  # (field key, mu1, mu2) -> complex coefficient of field(s) Y_l1,mu1(shat) Y_l2,mu2(shat)
            for m1, m2, M, g in _triposh_rows(l1, l2, L, los_z=isinstance(los, str) and los == 'z'):
                base = {('tau', (L, mu)): c for mu, c in _ylm_components(L, M).items()} if local else \
                    {('tau', None): _complex_ylm_conj_at(L, M, vlos)}
                if all_equal:
                    base[('omega', None)] = base.get(('omega', None), 0.) - s111(L, M)
                for mu1, c1 in _ylm_components(l1, m1).items():
                    for mu2, c2 in _ylm_components(l2, m2).items():
                        for key, cb in base.items():
                            k = (key, mu1, mu2)
                            coef[k] = coef.get(k, 0.) + g * cb * c1 * c2
            Ure = torch.zeros(shape, dtype=torch.float32, device=device)
            Uim = torch.zeros(shape, dtype=torch.float32, device=device)
            tmp = torch.empty(shape, dtype=torch.float32, device=device)
            for ((kind, arg), mu1, mu2), c in coef.items():
                if abs(c) < 1e-14:
                    continue
                field = omega if kind == 'omega' else tau(arg)
                _mul_ylm(field, tmp, mattrs, ss, l1, mu1, False)
                _mul_ylm(tmp, tmp, mattrs, ss, l2, mu2, False)
                if c.real != 0.:
                    Ure.add_(tmp, alpha=float(c.real))
                if c.imag != 0.:
                    Uim.add_(tmp, alpha=float(c.imag))
            if correlation:
                total = inter_sums(Ure) + 1j * inter_sums(Uim)
                out[ill] += (-1.)**(l1 + l2) * total / vc2
            else:
                total = hankel(Ure, l1, l2) + 1j * hankel(Uim, l1, l2)
                out[ill] += (1j)**(l1 + l2) * total
    if correlation:
        return tuple(o.real for o in out)
    return tuple(o.real if (ell[0] + ell[1]) % 2 == 0 else o.imag for ell, o in zip(ells, out))
