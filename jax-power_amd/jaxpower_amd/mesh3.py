"""Bispectrum multipoles in the Scoccimarro basis (global or local line of sight): the binning
operator BinMesh3SpectrumPoles (jaxpower/mesh3.py:23-177, 180-313), compute_mesh3_spectrum
(mesh3.py:509-629), compute_mesh3 (mesh3.py:475-506) and compute_box3_normalization
(mesh3.py:730-735).

Each |k| shell b of a half spectrum A becomes the real mesh I_b = irfftn(A [bin == b]) on
the GPU (``mp_bispec_shells``), and the closed-triangle sums sum_x I_b1 I_b2 I_b3 over the
allowed (b1, b2, b3) are taken in one pass over the cells (``mp_bispec_triangles``), FP32 per
cell and FP64 across cells.  The mode counts are the same triangle sums of the shell
indicators (mesh3.py:229-236).  With the local line of sight the third side is
sum_m r2c(rho_3 Y_lm(xhat)) Y_lm(khat) (mesh3.py:563-576).  The Sugiyama basis and the
three-point correlation function are not implemented on this path and raise
NotImplementedError.
"""
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


def _product(arrays, diagonal):
    """mesh3.py:107-117: all (i, j, k) combinations in lexicographic order, or the diagonal."""
    if diagonal:
        return np.column_stack([np.asarray(arrays[0])] * len(arrays))
    grid = np.meshgrid(*arrays, sparse=False, indexing='ij')
    return np.column_stack([g.ravel() for g in grid])


class BinMesh3SpectrumPoles:
    """Binning operator for the bispectrum (mesh3.py:180-313), basis 'scoccimarro' or
    'scoccimarro-diagonal'.

    Attributes mirror the reference: ``edges`` (ntri, 3, 2), ``xavg`` (ntri, 3), ``nmodes``
    (one float array per ell: the closed-triangle counts), ``edges1d``, ``nmodes1d``,
    ``xavg1d``, ``ells``, ``basis``.  ``batch_size`` and ``buffer_size`` (the reference's
    memory/throughput knobs for jax.lax.map) are accepted and stored; the GPU path stages all
    shells at once.
    """

    def __init__(self, mattrs, edges=None, ells=0, basis='sugiyama', batch_size=None, buffer_size=0, mask_edges=None, device=None):
        if not isinstance(mattrs, MeshAttrs):
            mattrs = mattrs.attrs
        if basis not in _BASES:
            raise ValueError(f'basis must be one of {_BASES}, got {basis!r}')
        if 'scoccimarro' not in basis:
            raise NotImplementedError(f"basis={basis!r} (the reference's default is 'sugiyama') is not implemented on this "
                                      "path: pass basis='scoccimarro' or 'scoccimarro-diagonal'")
        device = device or _default_device()
        ndim = 3
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
        self._tri = torch.from_numpy(np.ascontiguousarray(self._iedges.ravel())).to(device)
        # closed-triangle counts: triangle sums of the shell indicators times N^6 (mesh3.py:229-236)
        stacks = self._shell_stacks([None] * ndim, [0] * ndim)
        counts = self._triangles(stacks)
        norm = float(np.prod(mattrs.meshsize, dtype='f8'))**2
        del stacks
        self._nmodes_dev = counts * norm
        nmodes = self._nmodes_dev.cpu().numpy()
        self.nmodes = [nmodes] * len(self.ells)
        self.wmodes = None

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

    def __call__(self, *meshes, remove_zero=False):
        """Triangle sums of three half spectra times N^6 (mesh3.py:277-313), as a float64
        array; remove_zero sets each mesh's k = 0 mode to 0 first."""
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


def _2c(mesh):
    return mesh if isinstance(mesh, ComplexMeshField) else mesh.r2c()


def compute_mesh3_spectrum(*meshes, bin=None, los='x'):
    """Bispectrum multipoles (Scoccimarro basis, global line of sight) of one to three meshes
    (mesh3.py:509-629): num_raw_ell3 = (2 ell3 + 1) sum over closed triangles of
    A1 A2 (A3 P_ell3(mu)), divided by the triangle counts; odd ell3 give 0 (the reference keeps
    the imaginary part of a real sum); norm = N^3 / Vcell^2."""
    if not isinstance(bin, BinMesh3SpectrumPoles):
        raise TypeError('bin must be a BinMesh3SpectrumPoles')
    if isinstance(los, str) and los in ('firstpoint', 'endpoint'):
        raise ValueError(f"the bispectrum line of sight is 'local' or a global one, got {los!r} (mesh3.py:433-440)")
    meshes, _ = format_meshes(*meshes, nmeshes=3)
    mattrs = meshes[0].attrs
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


def compute_mesh3(*meshes, bin=None, los='z'):
    """Dispatch on the binning operator (mesh3.py:475-506); only the bispectrum is built."""
    if isinstance(bin, BinMesh3SpectrumPoles):
        return compute_mesh3_spectrum(*meshes, bin=bin, los=los)
    raise NotImplementedError('the three-point correlation function is not implemented on this path')


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
    if not isinstance(bin, BinMesh3SpectrumPoles):
        raise TypeError('bin must be a BinMesh3SpectrumPoles')
    if isinstance(los, str) and los in ('firstpoint', 'endpoint'):
        raise ValueError(f"the bispectrum line of sight is 'local' or a global one, got {los!r} (mesh3.py:433-440)")
    local = isinstance(los, str) and los == 'local'
    fkps, fields = format_meshes(*fkps, fields=fields, nmeshes=3)
    if fields[2] == fields[1] + 1 == fields[0] + 2:
        return tuple(np.zeros(bin.nbins) for _ in bin.ells)
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
