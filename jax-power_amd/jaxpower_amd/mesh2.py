"""Power-spectrum multipoles from meshes (jaxpower/mesh2.py:25-119, 195-374, 505-621).

``BinMesh2SpectrumPoles`` keeps the bin geometry: edges, exact integer mode counts, mode-averaged
|k|, and the device tables the binning kernel reads (per-axis wavenumbers and per-edge k^2
thresholds).  It never stores the per-mode bin index: the kernel recomputes it from k^2.
"""
import numpy as np
import torch

from . import _lib
from .mesh import (MeshAttrs, BaseMeshField, RealMeshField, ComplexMeshField, ParticleField, FKPField, format_meshes,
                   compute_normalization, compute_box_normalization, _as_spectrum, _sum_f64, _as_device_tensor,
                   _default_device, _empty_spectrum, _spectrum_pitch, _c2r, _r2c)
from .types import (Deferred, Mesh2SpectrumPole, Mesh2SpectrumPoles, Mesh2CorrelationPole, Mesh2CorrelationPoles,
                    WindowMatrix, _host)


def _format_ells(ells):
    if np.ndim(ells) == 0:
        ells = (ells,)
    return tuple(sorted(int(ell) for ell in ells))


def _k2_thresholds(edges):
    """t_e = min{t : fl(sqrt(t)) >= e}: digitize(|k|, edges) == number of t_e <= k^2."""
    thr = np.empty(len(edges), dtype='f8')
    for i, e in enumerate(np.asarray(edges, dtype='f8')):
        if not e > 0.:
            thr[i] = -1. if e <= 0. else np.inf
            continue
        t = e * e
        while np.sqrt(t) >= e:
            t = np.nextafter(t, 0.)
        while np.sqrt(t) < e:
            t = np.nextafter(t, np.inf)
        thr[i] = t
    return thr


def _kvectors(mattrs):
    """Per-axis wavenumbers fftfreq(n) * kfun * n (rfftfreq on the last axis), float64
    (mesh.py:360-403, 1605-1636)."""
    toret = []
    for axis, n in enumerate(mattrs.meshsize):
        n = int(n)
        f = np.fft.rfftfreq(n) if axis == 2 else np.fft.fftfreq(n)
        toret.append(f * mattrs.kfun[axis] * n)
    return toret


def _svectors(mattrs):
    """Per-axis separations fftfreq(n) * cellsize * n, float64 (xcoords(kind='separation'),
    mesh.py:361-403, 1573-1603)."""
    return [np.fft.fftfreq(int(n)) * mattrs.cellsize[axis] * int(n) for axis, n in enumerate(mattrs.meshsize)]


def _find_unique_edges(mattrs, xmin=0., xmax=np.inf, device=None, vectors=None, x0=None):
    """mesh.py:2788-2809: edges at mid-points between the distinct norms of the lattice
    vectors (wavenumbers by default, x0 the fundamental mode)."""
    vectors = _kvectors(mattrs) if vectors is None else vectors
    kx, ky, kz = (torch.from_numpy(k).to(device) for k in vectors)
    x2 = (kx[:, None, None]**2 + ky[None, :, None]**2 + kz[None, None, :]**2).reshape(-1)
    x2 = x2[(x2 >= xmin**2) & (x2 <= xmax**2)]
    x0 = float(mattrs.kfun.min() if x0 is None else x0)
    keys = (x2 / (0.5 * x0)**2 + 0.5).to(torch.int64)
    keys, inverse = torch.unique(keys, return_inverse=True)
    first = torch.full((len(keys),), x2.numel(), dtype=torch.int64, device=x2.device)
    first = first.scatter_reduce(0, inverse, torch.arange(x2.numel(), device=x2.device), reduce='amin')
    x = torch.sqrt(x2[first]).cpu().numpy()
    tmp = (x[:-1] + x[1:]) / 2.
    return np.concatenate([[tmp[0] - (x[1] - x[0])], tmp, [tmp[-1] + (x[-1] - x[-2])]])


def _oversampling_shifts(mode_oversampling):
    """Integer shifts of _make_edges2 (mesh2.py:48-49): every triple in [-m, m]^3, last axis
    fastest.  They are added to the coordinates themselves (h/Mpc for wavenumbers, Mpc/h for
    separations), not to mode indices."""
    m = int(mode_oversampling)
    if m < 0:
        raise ValueError(f'mode_oversampling must be >= 0, got {mode_oversampling}')
    r = range(-m, m + 1)
    return [(float(a), float(b), float(c)) for a in r for b in r for c in r]


def _shifted_modes(call, shifts, nbins, device):
    """nmodes, xavg of _make_edges2 (mesh2.py:48-61) from call(shift, nmodes, xsum), which
    fills the counts and the sum of shifted norms of one shift: counts summed over the shifts,
    divided by their number when there are several (float then), xavg = xsum / nmodes."""
    total = torch.zeros(nbins, dtype=torch.int64, device=device)
    xsum = torch.zeros(nbins, dtype=torch.float64, device=device)
    nmodes = torch.empty(nbins, dtype=torch.int64, device=device)
    part = torch.empty(nbins, dtype=torch.float64, device=device)
    for shift in shifts:
        call(shift, nmodes, part)
        total += nmodes
        xsum += part
    if len(shifts) > 1:
        host = total.cpu().numpy() / len(shifts)
    else:
        host = total.cpu().numpy()
    with np.errstate(invalid='ignore', divide='ignore'):
        xavg = xsum.cpu().numpy() / host
    return host, torch.from_numpy(np.asarray(host, dtype='f8')).to(device), xavg


def _single_shift(bin, what):
    """Paths that bin with their own fused kernels take bins without mode oversampling."""
    if len(getattr(bin, '_shifts', ((0., 0., 0.),))) > 1:
        raise NotImplementedError(f'{what}: bins with mode_oversampling > 0 are taken by compute_mesh2_spectrum, '
                                  'compute_mesh2_correlation and the bin call, not here')


class BinMesh2SpectrumPoles:
    """Binning operator for mesh to power spectrum (mesh2.py:65-119).

    mode_oversampling = m bins every mode once per integer shift s in [-m, m]^3 by |k + s|
    (mesh2.py:48-61): nmodes is the mean over the shifts of the counts (a float when m > 0),
    xavg = (sum over the shifts of sum |k + s|) / nmodes, and a binned sum is the mean over
    the shifts of the per-shift sums (mesh.py:2874-2910); mu and the hermitian weights stay
    those of k."""

    def __init__(self, mattrs, edges=None, ells=0, mode_oversampling=0, device=None):
        if not isinstance(mattrs, MeshAttrs):
            mattrs = mattrs.attrs
        self.mode_oversampling = int(mode_oversampling)
        self._shifts = _oversampling_shifts(mode_oversampling)
        device = device or _default_device()
        if edges is None:
            edges = {}
        if isinstance(edges, dict):
            step = edges.get('step', None)
            vecmax = mattrs.kfun.min() * np.min(mattrs.meshsize) / 2.
            if step is None:
                edges = _find_unique_edges(mattrs, edges.get('min', 0.), edges.get('max', np.inf), device=device)
            else:
                edges = np.arange(edges.get('min', 0.), edges.get('max', vecmax), step)
        edges = np.asarray(edges, dtype='f8')
        if edges.ndim == 2:
            assert np.allclose(edges[1:, 0], edges[:-1, 1])
            edges = np.append(edges[:, 0], edges[-1, 1])
        if len(edges) < 2:
            raise ValueError('need at least two edges')
        self.mattrs = mattrs
        self.ells = _format_ells(ells)
        self._edges1d = edges
        self.edges = np.column_stack([edges[:-1], edges[1:]])
        self._device = device
        kx, ky, kz = _kvectors(mattrs)
        self._tables = dict(kx=torch.from_numpy(kx).to(device), ky=torch.from_numpy(ky).to(device),
                            kz=torch.from_numpy(kz).to(device),
                            thr=torch.from_numpy(_k2_thresholds(edges)).to(device))
        t = self._tables

        def count(shift, nmodes, ksum):
            _lib.call('mp_bin_modes', _lib.stream_handle(device), _lib.i32x3(mattrs.meshsize), _lib.ptr(t['kx']),
                      _lib.ptr(t['ky']), _lib.ptr(t['kz']), _lib.ptr(t['thr']), len(edges), _lib.f64x3(shift), _lib.ptr(nmodes),
                      _lib.ptr(ksum))

        self.nmodes, self._nmodes_dev, self.xavg = _shifted_modes(count, self._shifts, len(edges) - 1, device)
        self.wmodes = None  # hermitian weights 1 + (kz > 0) are applied inside the kernel

    @property
    def nbins(self):
        return len(self.xavg)

    def products(self, meshes, products, los=(0., 0., 0.)):
        """Raw FP64 sums for products (i, j, ell, remove_zero) of conj(mesh_i) * mesh_j
        (i = -1: the constant 1) weighted by P_ell(k.los/|k|); returns complex (nprod, nbins)."""
        return self._products_dev(meshes, products, los=los).cpu().numpy()

    def _products_dev(self, meshes, products, los=(0., 0., 0.)):
        """products() left on the device (complex128 tensor)."""
        device = meshes[0].device
        arrays, pitches = [], []
        for m in meshes:
            m, pitch = _as_spectrum(m)
            arrays.append(m)
            pitches.append(pitch)
        nprod = len(products)
        out = torch.empty((nprod, self.nbins, 2), dtype=torch.float64, device=device)
        ws = _lib.workspace(_lib.load().mp_bin_workspace_bytes(nprod, len(self._edges1d)), device)
        ptrs = (_lib.c_void_p * len(arrays))(*[a.data_ptr() for a in arrays])
        pitch_arr = (_lib.ctypes.c_int64 * len(arrays))(*pitches)
        col = lambda k: (_lib.ctypes.c_int32 * nprod)(*[int(p[k]) for p in products])
        t = self._tables
        total = None
        for shift in self._shifts:  # _bincount over the ibin list: per-shift sums, then their mean
            _lib.call('mp_bin_poles', _lib.stream_handle(device), len(arrays), ptrs, pitch_arr, _lib.i32x3(self.mattrs.meshsize),
                      _lib.ptr(t['kx']), _lib.ptr(t['ky']), _lib.ptr(t['kz']), _lib.ptr(t['thr']), len(self._edges1d),
                      _lib.f64x3(shift), _lib.f64x3(los), nprod, col(0), col(1), col(2), col(3), _lib.ptr(out), _lib.ptr(ws),
                      ws.numel())
            if len(self._shifts) == 1:
                return torch.view_as_complex(out)
            total = out.clone() if total is None else total + out
        return torch.view_as_complex((total / len(self._shifts)).contiguous())

    def _pending_products(self, pending, ells, remove_zero, los):
        """_products_dev() of conj(A) A L_ell for the combined mesh A of a pending interlacing
        combine, binned without writing A; None when the fused kernel does not cover the case."""
        nprod = len(ells)
        device = pending.partial.device
        n = _lib.i32x3(self.mattrs.meshsize)
        nedges = len(self._edges1d)
        if nprod > 6 or nprod * (nedges - 1) > 4096 or len(self._shifts) > 1:
            return None
        out = torch.empty((nprod, self.nbins, 2), dtype=torch.float64, device=device)
        ws = _lib.workspace(_lib.load().mp_bin_partial_workspace_bytes(n, pending.interlacing, nprod, nedges), device)
        col = lambda values: (_lib.ctypes.c_int32 * nprod)(*[int(v) for v in values])
        t = self._tables
        try:
            _lib.call('mp_bin_poles_partial', _lib.stream_handle(device), _lib.ptr(pending.partial), pending.pitch, n,
                      pending.order, pending.interlacing, int(pending.compensate), int(pending.zero_dc), _lib.ptr(t['kx']),
                      _lib.ptr(t['ky']), _lib.ptr(t['kz']), _lib.ptr(t['thr']), nedges, _lib.f64x3(los), nprod, col(ells),
                      col(remove_zero), _lib.ptr(out), _lib.ptr(ws), ws.numel())
        except NotImplementedError:
            return None
        return torch.view_as_complex(out)

    def __call__(self, mesh, remove_zero=False):
        """Sum_{modes in bin} w * value / nmodes (mesh2.py:100-119)."""
        value = mesh.value if isinstance(mesh, BaseMeshField) else _as_device_tensor(mesh, device=self._device)
        if not value.is_complex():
            value = value.to(torch.complex64)
        s = self.products([value], [(-1, 0, 0, bool(remove_zero))])[0]
        with np.errstate(invalid='ignore', divide='ignore'):
            s = s / self.nmodes
        return s if np.iscomplexobj(s) and np.any(s.imag) else s.real


def _get_los_vector(los):
    if isinstance(los, str):
        vlos = [0., 0., 0.]
        vlos['xyz'.index(los)] = 1.
        return np.array(vlos)
    return np.asarray(los, dtype='f8')


def _format_los(los):
    """mesh2.py:215-222."""
    vlos, swap = None, False
    if isinstance(los, str) and los in ['firstpoint', 'local', 'endpoint']:
        if los == 'local':
            los = 'firstpoint'
        swap = los == 'endpoint'
    else:
        vlos = _get_los_vector(los)
    return los, vlos, swap


def _2c(mesh):
    return mesh if isinstance(mesh, ComplexMeshField) else mesh.r2c()


def _2r(mesh):
    return mesh if isinstance(mesh, RealMeshField) else mesh.c2r()


def compute_mesh2_spectrum(*meshes, bin=None, los='z'):
    """Power-spectrum multipoles from one (auto) or two (cross) meshes (mesh2.py:253-374)."""
    meshes, fields = format_meshes(*meshes, nmeshes=2)
    mattrs = meshes[0].attrs
    norm = np.prod(mattrs.meshsize, dtype='f8') / np.prod(mattrs.cellsize) * np.ones_like(bin.xavg)
    los, vlos, swap = _format_los(los)
    attrs = dict(los=vlos.tolist() if vlos is not None else los)
    ells = bin.ells
    same = fields[1] == fields[0]
    nonzeroells = [ell for ell in ells if ell != 0]

    if vlos is None:  # local (varying) line of sight, mesh2.py:293-344
        if swap:
            meshes = meshes[::-1]
        rmesh1 = meshes[0]
        A0 = _2c(rmesh1 if same else meshes[1])
        rreal = _2r(rmesh1) if nonzeroells else None
        sums = _ylm_bin_all(bin, rreal, A0, nonzeroells, with0=(0 in ells and same)) if nonzeroells else None
        if sums is not None:  # fused: C_ell never stored
            if 0 in ells and 0 not in sums:
                sums[0] = bin._products_dev([A0.value, _2c(rmesh1).value], [(1, 0, 0, False)])[0]
            num = [sums[ell] * ((4. * np.pi if ell else 1.) / bin._nmodes_dev) for ell in ells]
        else:
            bmeshes, products, scales = [A0.value], [], []
            if 0 in ells:
                if same:
                    products.append((0, 0, 0, False))
                else:
                    bmeshes.append(_2c(rmesh1).value)
                    products.append((len(bmeshes) - 1, 0, 0, False))
                scales.append(1.)
            for ell in nonzeroells:
                C = mattrs.create('complex', device=A0.value.device).value
                _ylm_accumulate(rreal, ell, C)
                bmeshes.append(C)
                products.append((len(bmeshes) - 1, 0, 0, True))
                scales.append(4. * np.pi)
            sums = bin._products_dev(bmeshes, products)
            num = [sums[i] * (scale / bin._nmodes_dev) for i, scale in enumerate(scales)]
        if swap:
            num = [torch.conj(n).resolve_conj() for n in num]
    else:  # fixed line of sight, mesh2.py:346-374
        A1 = _2c(meshes[0])
        sums = None
        pending = A1._pending_combine() if same else None
        if pending is not None:  # interlacing combine fused with the binning
            sums = bin._pending_products(pending, ells, [ell != 0 for ell in ells], vlos)
        if sums is None:
            bmeshes = [A1.value]
            j = 0
            if not same:
                bmeshes.append(_2c(meshes[1]).value)
                j = 1
            products = [(0, j, ell, ell != 0) for ell in ells]
            sums = bin._products_dev(bmeshes, products, los=vlos)
        num = [sums[i] * ((2 * ell + 1 if ell else 1) / bin._nmodes_dev) for i, ell in enumerate(ells)]
    # left on the device: the spectrum reads them back on first access
    # an auto spectrum along a fixed line of sight bins |A|^2, a real array, whose .imag the
    # reference takes for odd ell: 0 in every bin, empty ones included (mesh2.py:350-351, 368)
    auto_real = vlos is not None and same
    num = [(n.real if ell % 2 == 0 else (torch.zeros_like(n.real) if auto_real else n.imag)).contiguous()
           for ell, n in zip(ells, num)]
    poles = [Mesh2SpectrumPole(k=bin.xavg, k_edges=bin.edges, nmodes=bin.nmodes, num_raw=num[i], num_shotnoise=0.,
                               norm=norm, volume=np.prod(mattrs.kfun) * bin.nmodes, ell=ell, attrs=attrs)
             for i, ell in enumerate(ells)]
    return Mesh2SpectrumPoles(poles, attrs=attrs)


def _ylm_bin_all(bin, rreal, A0, ells, with0=False, memory_fraction=0.5):
    """{ell: complex sums (nbins,)} of conj(C_ell) A0 over every ell > 0 with the fused kernel
    (mp_ylm_bin), plus |A0|^2 under key 0 when with0; None when the fused kernel does not apply
    (shape, bins, or its 2 ell + 1 partial spectra would take more than memory_fraction of the
    free device memory), in which case the caller stores C_ell."""
    if len(bin._shifts) > 1:
        return None
    a = rreal.attrs
    device = rreal.value.device
    n = _lib.i32x3(a.meshsize)
    nedges = len(bin._edges1d)
    lib = _lib.load()
    need = max(int(lib.mp_ylm_bin_workspace_bytes(n, ell, nedges)) for ell in ells)
    if need > memory_fraction * torch.cuda.mem_get_info(device)[0]:
        return None
    value = rreal.value.contiguous()
    a0, pitch = _as_spectrum(A0.value)
    t = bin._tables
    out = {}
    for i, ell in enumerate(ells):
        w0 = bool(with0 and i == 0)
        res = torch.empty((2 if w0 else 1, bin.nbins, 2), dtype=torch.float64, device=device)
        ws = _lib.workspace(lib.mp_ylm_bin_workspace_bytes(n, ell, nedges), device)
        try:
            _lib.call('mp_ylm_bin', _lib.stream_handle(device), _lib.ptr(value), _lib.ptr(a0), pitch, n, _lib.f64x3(a.boxsize),
                      _lib.f64x3(a.boxcenter), ell, int(w0), _lib.ptr(t['kx']), _lib.ptr(t['ky']), _lib.ptr(t['kz']),
                      _lib.ptr(t['thr']), nedges, _lib.ptr(res), _lib.ptr(ws), ws.numel())
        except NotImplementedError:
            return None
        res = torch.view_as_complex(res)
        out[ell] = res[0]
        if w0:
            out[0] = res[1]
    return out


def _ylm_accumulate(rmesh, ell, out):
    a = rmesh.attrs
    value = rmesh.value.contiguous()
    n = _lib.i32x3(a.meshsize)
    ws = _lib.workspace(_lib.load().mp_ylm_workspace_bytes(n), value.device)
    out_, pitch = _as_spectrum(out)
    assert out_.data_ptr() == out.data_ptr()
    _lib.call('mp_ylm_accumulate', _lib.stream_handle(value.device), _lib.ptr(value), n, _lib.f64x3(a.boxsize),
              _lib.f64x3(a.boxcenter), ell, _lib.ptr(out), pitch, _lib.ptr(ws), ws.numel())


class BinMesh2CorrelationPoles:
    """Binning operator for mesh to correlation function (mesh2.py:120-184): separation shells
    over the full real mesh, every cell weighted 1.  Edges as for the spectrum with the
    smallest cellsize as the unit (mesh2.py:25-47, kind='real').  Only the default
    configuration-space basis is on this path."""

    def __init__(self, mattrs, edges=None, ells=0, mode_oversampling=0, basis=None, klimit=None, batch_size=None, device=None):
        if not isinstance(mattrs, MeshAttrs):
            mattrs = mattrs.attrs
        self.mode_oversampling = int(mode_oversampling)
        self._shifts = _oversampling_shifts(mode_oversampling)
        if basis is not None:
            raise NotImplementedError(f"basis={basis!r} is not implemented on this path (only the default, basis=None)")
        device = device or _default_device()
        if edges is None:
            edges = {}
        if isinstance(edges, dict):
            step = edges.get('step', None)
            vec0 = float(np.min(mattrs.cellsize))
            vecmax = vec0 * np.min(mattrs.meshsize) / 2.
            if step is None:
                edges = _find_unique_edges(mattrs, edges.get('min', 0.), edges.get('max', np.inf), device=device,
                                           vectors=_svectors(mattrs), x0=vec0)
            else:
                edges = np.arange(edges.get('min', 0.), edges.get('max', vecmax), step)
        edges = np.asarray(edges, dtype='f8')
        if edges.ndim == 2:
            assert np.allclose(edges[1:, 0], edges[:-1, 1])
            edges = np.append(edges[:, 0], edges[-1, 1])
        if len(edges) < 2:
            raise ValueError('need at least two edges')
        self.mattrs = mattrs
        self.ells = _format_ells(ells)
        self.basis, self.klimit, self.batch_size = basis, klimit, batch_size
        self._edges1d = edges
        self.edges = np.column_stack([edges[:-1], edges[1:]])
        self._device = device
        sx, sy, sz = _svectors(mattrs)
        self._tables = dict(sx=torch.from_numpy(sx).to(device), sy=torch.from_numpy(sy).to(device),
                            sz=torch.from_numpy(sz).to(device), thr=torch.from_numpy(_k2_thresholds(edges)).to(device))
        t = self._tables

        def count(shift, nmodes, ssum):
            _lib.call('mp_bin_real_modes', _lib.stream_handle(device), _lib.i32x3(mattrs.meshsize), _lib.ptr(t['sx']),
                      _lib.ptr(t['sy']), _lib.ptr(t['sz']), _lib.ptr(t['thr']), len(edges), _lib.f64x3(shift), _lib.ptr(nmodes),
                      _lib.ptr(ssum))

        self.nmodes, self._nmodes_dev, self.xavg = _shifted_modes(count, self._shifts, len(edges) - 1, device)

    @property
    def nbins(self):
        return len(self.xavg)

    def _sums_dev(self, value, ells, remove_zero, los=(0., 0., 0.)):
        """FP64 sums over each separation bin of value * P_ell(mu_s) for each ell, (nell, nbins),
        left on the device; value is a real (N0, N1, N2) float32 tensor."""
        value = value.contiguous()
        if value.dtype != torch.float32:
            value = value.to(torch.float32)
        device = value.device
        nell = len(ells)
        n = _lib.i32x3(self.mattrs.meshsize)
        nedges = len(self._edges1d)
        out = torch.empty((nell, self.nbins), dtype=torch.float64, device=device)
        ws = _lib.workspace(_lib.load().mp_bin_real_workspace_bytes(n, nell, nedges), device)
        col = lambda values: (_lib.ctypes.c_int32 * nell)(*[int(v) for v in values])
        t = self._tables
        total = None
        for shift in self._shifts:  # per-shift sums, then their mean (mesh.py:2874-2910)
            _lib.call('mp_bin_real_poles', _lib.stream_handle(device), _lib.ptr(value), n, _lib.ptr(t['sx']), _lib.ptr(t['sy']),
                      _lib.ptr(t['sz']), _lib.ptr(t['thr']), nedges, _lib.f64x3(shift), _lib.f64x3(los), nell, col(ells),
                      col(remove_zero), _lib.ptr(out), _lib.ptr(ws), ws.numel())
            if len(self._shifts) == 1:
                return out
            total = out.clone() if total is None else total + out
        return total / len(self._shifts)

    def _spectrum_sums_dev(self, spectrum, ells, remove_zero, los=(0., 0., 0.), x_done=False):
        """_sums_dev() of irfftn(spectrum) for a half spectrum, with the last pass of the
        inverse FFT fused with the binning (the real mesh is never stored); x_done: the
        spectrum already carries its inverse FFT along x."""
        if len(self._shifts) > 1:  # the fused pass bins one shift: store the real mesh instead
            if x_done:
                raise ValueError('an oversampled bin takes the product spectrum without its inverse x pass')
            return self._sums_dev(_c2r(spectrum, self.mattrs.meshsize), ells, remove_zero, los=los)
        value, pitch = _as_spectrum(spectrum)
        device = value.device
        nell = len(ells)
        n = _lib.i32x3(self.mattrs.meshsize)
        nedges = len(self._edges1d)
        out = torch.empty((nell, self.nbins), dtype=torch.float64, device=device)
        ws = _lib.workspace(_lib.load().mp_irfft3d_bin_real_workspace_bytes(n, nell, nedges), device)
        col = lambda values: (_lib.ctypes.c_int32 * nell)(*[int(v) for v in values])
        t = self._tables
        _lib.call('mp_irfft3d_bin_real', _lib.stream_handle(device), _lib.ptr(value), pitch, int(x_done), n, _lib.ptr(t['sx']),
                  _lib.ptr(t['sy']),
                  _lib.ptr(t['sz']), _lib.ptr(t['thr']), nedges, _lib.f64x3(los), nell, col(ells), col(remove_zero),
                  _lib.ptr(out), _lib.ptr(ws), ws.numel())
        return out

    def __call__(self, mesh, ell=0, remove_zero=False):
        """Sum_{cells in bin} value / nmodes (mesh2.py:175-184); mesh is real."""
        value = mesh.value if isinstance(mesh, BaseMeshField) else _as_device_tensor(mesh, device=self._device)
        if value.is_complex():
            raise ValueError('BinMesh2CorrelationPoles bins a real (configuration-space) mesh')
        s = self._sums_dev(value, [0], [bool(remove_zero)])[0].cpu().numpy()
        with np.errstate(invalid='ignore', divide='ignore'):
            return s / self.nmodes


def _product_spectrum(A1, A2=None, inverse_x=False):
    """(P, x_done): the half spectrum P = |A1|^2 (A2 None) or conj(A1) A2 on the device.  A
    pending interlacing combine is combined straight into |A|^2 without storing A, and with
    inverse_x P then also carries its inverse FFT along x (x_done True)."""
    mattrs = A1.attrs
    n = _lib.i32x3(mattrs.meshsize)
    if A2 is None:
        pending = A1._pending_combine()
        if pending is not None:
            device = pending.partial.device
            out = _empty_spectrum(mattrs.meshsize, device)
            ws = _lib.workspace(_lib.load().mp_combine_partial_workspace_bytes(n, pending.interlacing), device)
            _lib.call('mp_combine_partial_power', _lib.stream_handle(device), _lib.ptr(pending.partial), pending.pitch, n,
                      pending.order, pending.interlacing, int(pending.compensate), int(pending.zero_dc), int(inverse_x),
                      _lib.ptr(out), _spectrum_pitch(out), _lib.ptr(ws), ws.numel())
            return out, bool(inverse_x)
    a, pa = _as_spectrum(A1.value)
    out = _empty_spectrum(mattrs.meshsize, a.device)
    if A2 is None:
        b, pb = None, 0
    else:
        b, pb = _as_spectrum(A2.value)
    _lib.call('mp_spectrum_product', _lib.stream_handle(a.device), _lib.ptr(a), pa, _lib.ptr(b) if b is not None else None, pb, n,
              _lib.ptr(out), _spectrum_pitch(out))
    return out, False


def _mul_radial_power(value, out, mattrs, coords, power):
    """out = value |c|^power on real meshes; coords are three device float64 tables."""
    _lib.call('mp_mul_radial_power', _lib.stream_handle(value.device), _lib.ptr(value), _lib.ptr(out), _lib.i32x3(mattrs.meshsize),
              _lib.ptr(coords[0]), _lib.ptr(coords[1]), _lib.ptr(coords[2]), float(power))


def _mul_ylm(value, out, mattrs, coords, ell, m, accumulate):
    """out (+)= value * Y_lm(coords) on real meshes; coords are three device float64 tables."""
    _lib.call('mp_mul_ylm', _lib.stream_handle(value.device), _lib.ptr(value), _lib.ptr(out), _lib.i32x3(mattrs.meshsize),
              _lib.ptr(coords[0]), _lib.ptr(coords[1]), _lib.ptr(coords[2]), int(ell), int(m), int(accumulate))


def compute_mesh2_correlation(*meshes, bin=None, los='z'):
    """Correlation-function multipoles from one (auto) or two (cross) meshes (mesh2.py:377-504):
    xi = c2r(|A|^2) (or c2r(conj(A1) A2)) binned in separation shells with L_ell(mu_s) for a
    fixed line of sight; for a local (firstpoint/endpoint) line of sight the ell > 0 terms are
    4 pi sum_m c2r(conj(r2c(rho Y_lm(xhat))) A0) Y_lm(shat)."""
    if not isinstance(bin, BinMesh2CorrelationPoles):
        raise ValueError(f'bin must be BinMesh2CorrelationPoles, not {type(bin)}')
    meshes, fields = format_meshes(*meshes, nmeshes=2)
    mattrs = meshes[0].attrs
    vcell = float(np.prod(mattrs.cellsize))
    norm = np.prod(mattrs.meshsize, dtype='f8') / vcell * np.ones_like(bin.xavg)
    los, vlos, swap = _format_los(los)
    attrs = dict(los=vlos.tolist() if vlos is not None else los)
    ells = bin.ells
    same = fields[1] == fields[0]
    nonzeroells = [ell for ell in ells if ell != 0]
    fused = len(bin._shifts) == 1  # inverse FFT fused with the binning of one shift

    if vlos is None:  # local (varying) line of sight, mesh2.py:416-469
        if swap:
            meshes = meshes[::-1]
        rmesh1 = meshes[0]
        A0 = _2c(rmesh1 if same else meshes[1])
        num = {}
        if 0 in ells:
            P, x_done = _product_spectrum(A0, inverse_x=fused) if same else _product_spectrum(_2c(rmesh1), A0)
            num[0] = bin._spectrum_sums_dev(P, [0], [False], x_done=x_done)[0]
            del P
        if nonzeroells:
            rreal = _2r(rmesh1).value.contiguous()
            if rreal.dtype != torch.float32:
                rreal = rreal.to(torch.float32)
            device = rreal.device
            offset = mattrs.boxcenter - mattrs.boxsize / 2.
            xs = [torch.from_numpy(np.arange(int(n)) * float(d) + float(o)).to(device)
                  for n, d, o in zip(mattrs.meshsize, mattrs.cellsize, offset)]  # positions (mesh.py:1594-1603)
            ss = [bin._tables['sx'], bin._tables['sy'], bin._tables['sz']]
            a0 = A0.value
            weighted = torch.empty_like(rreal)
            for ell in nonzeroells:
                carry = torch.zeros_like(rreal)
                for m in range(-ell, ell + 1):
                    _mul_ylm(rreal, weighted, mattrs, xs, ell, m, False)
                    B = _r2c(weighted, mattrs.meshsize)
                    _lib.call('mp_spectrum_product', _lib.stream_handle(device), _lib.ptr(B), _spectrum_pitch(B),
                              _lib.ptr(_as_spectrum(a0)[0]), _as_spectrum(a0)[1], _lib.i32x3(mattrs.meshsize), _lib.ptr(B),
                              _spectrum_pitch(B))
                    _mul_ylm(_c2r(B, mattrs.meshsize), carry, mattrs, ss, ell, m, True)
                    del B
                num[ell] = bin._sums_dev(carry, [0], [True])[0] * (4. * np.pi)
                del carry
        num = [num[ell] / (bin._nmodes_dev * vcell) for ell in ells]
    else:  # fixed line of sight, mesh2.py:471-504
        A1 = _2c(meshes[0])
        P, x_done = _product_spectrum(A1, inverse_x=fused) if same else _product_spectrum(A1, _2c(meshes[1]))
        sums = bin._spectrum_sums_dev(P, ells, [ell != 0 for ell in ells], los=vlos, x_done=x_done)
        del P
        num = [sums[i] * ((2 * ell + 1) / (bin._nmodes_dev * vcell)) for i, ell in enumerate(ells)]
    num = [n.contiguous() for n in num]
    poles = [Mesh2CorrelationPole(s=bin.xavg, s_edges=bin.edges, nmodes=bin.nmodes, num_raw=num[i], num_shotnoise=0., norm=norm,
                                  volume=vcell * bin.nmodes, ell=ell, attrs=attrs)
             for i, ell in enumerate(ells)]
    return Mesh2CorrelationPoles(poles, attrs=attrs)


def compute_mesh2_spectrum_window(*meshes, edgesin, ellsin=None, bin=None, los='z', buffer=None, batch_size=None, pbar=False,
                                  norm=None, flags=tuple()):
    """Window matrix of the power-spectrum multipoles from a selection mesh, global line of
    sight, theory defined on the mesh (mesh2.py:921-1118 without flags, 1453-1482).

    For each input multipole ellin and input shell [e0, e1): A = [e0 <= |k| < e1]
    P_ellin(mu) N^3, convolved by the selection as r2c(Q c2r(A)) with Q = c2r(A1 conj(A2)) / N^3
    (nothing for a periodic box, given as a MeshAttrs), then binned as
    (2l+1)/(1+odd) bin(A (P_l(mu) + odd P_l(-mu)), remove_zero=True) N^3/V_cell / norm / N^3.
    The real part of every row is kept (mesh2.py:1466-1472), and odd output multipoles vanish.
    With flags=('infinite',) each input shell is taken to configuration space analytically,
    A = BesselIntegral(shell, |s|, ellin).w P_ellin(mu_s) V, times Q for a selection, then r2c
    and binned the same way (mesh2.py:1078-1092).
    ``buffer``, ``batch_size`` and ``pbar`` only steer the reference's memory use and progress
    display and are accepted for compatibility.  With flags=('smooth',) the selection
    multipoles binned in separation shells are Hankel-transformed against each input shell
    (mesh2.py:1045-1076).  The local line of sight is not implemented on this path."""
    if not isinstance(bin, BinMesh2SpectrumPoles):
        raise ValueError(f'bin must be BinMesh2SpectrumPoles, not {type(bin)}')
    _single_shift(bin, 'compute_mesh2_spectrum_window')
    flags = tuple(flags or ())
    infinite, smooth = 'infinite' in flags, 'smooth' in flags
    if set(flags) - {'infinite', 'smooth'}:
        raise NotImplementedError(f'window flags {flags} are not implemented on this path')
    periodic = isinstance(meshes[0], MeshAttrs)
    if periodic:
        mattrs = meshes[0]
        assert all(m is None or m is mattrs for m in meshes[1:]), 'a periodic window takes one MeshAttrs'
        fields = (0, 0)
    else:
        meshes, fields = format_meshes(*meshes, nmeshes=2)
        mattrs = meshes[0].attrs
    los, vlos, swap = _format_los(los)
    if vlos is None:
        return _window_local(meshes, fields, periodic, mattrs, edgesin, ellsin, bin, swap, norm, infinite, smooth)
    ells = bin.ells
    n3 = float(np.prod(mattrs.meshsize, dtype='f8'))
    _norm = n3 / float(np.prod(mattrs.cellsize))
    if norm is None:
        norm = _norm
    norm = np.broadcast_to(np.asarray(_host(norm), dtype='f8').ravel() if np.ndim(_host(norm)) else float(_host(norm)),
                           (len(ells),)).copy()
    rnorm = _norm / norm / n3
    if hasattr(edgesin, 'ells') and hasattr(edgesin, 'edges'):  # a theory container (mesh2.py:994-996)
        ellsin = edgesin.ells
        edgesin = edgesin.edges('k')
    edgesin = np.asarray(edgesin, dtype='f8')
    if edgesin.ndim != 2:
        edgesin = np.column_stack([edgesin[:-1], edgesin[1:]])
    ellsin = [int(ellsin)] if np.ndim(ellsin) == 0 else [int(ell) for ell in ellsin]
    device = bin._device
    Q = None
    if not periodic:  # Q N^3 = c2r(A1 conj(A2)), mesh2.py:1022-1029
        A1 = _2c(meshes[0])
        P, _ = _product_spectrum(A1) if fields[1] == fields[0] else _product_spectrum(_2c(meshes[1]), A1)
        Q = _c2r(P, mattrs.meshsize)
        del P
        device = Q.device
    if smooth:  # Hankel transforms of the binned selection multipoles, mesh2.py:1045-1076
        wmat = _window_smooth(mattrs, bin, edgesin, ellsin, vlos, Q, rnorm, device)
        return _window_matrix(wmat, bin, edgesin, ellsin, vlos, norm)
    if infinite:  # the theory shells taken to configuration space analytically, mesh2.py:1078-1092
        wmat = _window_infinite(mattrs, bin, edgesin, ellsin, vlos, Q, rnorm, device)
        return _window_matrix(wmat, bin, edgesin, ellsin, vlos, norm)
    # shell membership e0 <= |k| < e1 as k^2 >= t(e0) and k^2 < t(e1), t as for the bin edges
    thr = _k2_thresholds(edgesin.ravel()).reshape(edgesin.shape)
    scale = n3 if Q is None else 1.  # the N^3 of A cancels the 1/N^3 of Q
    even = [ell for ell in ells if ell % 2 == 0]
    products = [(-1, 0, ell, True) for ell in even]
    t = bin._tables
    stream = _lib.stream_handle(device)
    n = _lib.i32x3(mattrs.meshsize)
    cols = []
    nedges = len(bin._edges1d)
    fused = bool(Q is not None and products and len(even) <= 6 and len(even) * (nedges - 1) <= 4096
                 and all(int(m) >= 2 and int(m) & (int(m) - 1) == 0 for m in mattrs.meshsize))
    B = None if fused else _empty_spectrum(mattrs.meshsize, device)
    pitch = _spectrum_pitch(B) if B is not None else 0
    if fused:  # every column in one call, four launches each, neither the shell nor its convolution written
        ncol = len(ellsin) * len(thr)
        ws = _lib.workspace(_lib.load().mp_window_bin_workspace_bytes(n, len(even), nedges), device)
        out = torch.empty((ncol, len(even), bin.nbins, 2), dtype=torch.float64, device=device)
        dbl = lambda v: (_lib.ctypes.c_double * ncol)(*[float(x) for x in v])
        _lib.call('mp_window_bin', stream, _lib.ptr(Q), n, _lib.ptr(t['kx']), _lib.ptr(t['ky']), _lib.ptr(t['kz']), _lib.ptr(t['thr']),
                  nedges, ncol, dbl(list(thr[:, 0]) * len(ellsin)), dbl(list(thr[:, 1]) * len(ellsin)),
                  (_lib.ctypes.c_int32 * ncol)(*[ellin for ellin in ellsin for _ in thr]), _lib.f64x3(vlos), scale, len(even),
                  (_lib.ctypes.c_int32 * len(even))(*even), _lib.ptr(out), _lib.ptr(ws), ws.numel())
        cols = list(torch.view_as_complex(out))
    else:
        for ellin in ellsin:
            for t_lo, t_hi in thr:
                _lib.call('mp_window_shell', stream, _lib.ptr(Q) if Q is not None else None, n, _lib.ptr(t['kx']), _lib.ptr(t['ky']),
                          _lib.ptr(t['kz']), float(t_lo), float(t_hi), int(ellin), _lib.f64x3(vlos), scale, _lib.ptr(B), pitch)
                cols.append(bin._products_dev([B], products, los=vlos) if products else None)
    nbins = bin.nbins
    ncol = len(cols)
    sums = torch.stack(cols).cpu().numpy() if products else np.zeros((ncol, 0, nbins), dtype='c16')
    wmat = np.zeros((ncol, len(ells), nbins), dtype='f8')
    with np.errstate(invalid='ignore', divide='ignore'):
        for ie, ell in enumerate(even):
            ill = ells.index(ell)
            wmat[:, ill, :] = (2 * ell + 1) * rnorm[ill] * sums[:, ie, :].real / bin.nmodes
    return _window_matrix(wmat.reshape(ncol, -1).T, bin, edgesin, ellsin, vlos, norm)


def _window_infinite(mattrs, bin, edgesin, ellsin, vlos, Q, rnorm, device):
    """(n_ells n_bins, n_ellsin n_in) matrix of the infinite-resolution theory
    (mesh2.py:1078-1092) through mp_window_bin_tophat."""
    cell = np.asarray(mattrs.cellsize, dtype='f8')
    if not np.all(cell == cell[0]):
        raise NotImplementedError("window flags=('infinite',) needs equal cellsizes on this path")
    if not all(int(m) >= 2 and int(m) & (int(m) - 1) == 0 for m in mattrs.meshsize):
        raise NotImplementedError("window flags=('infinite',) needs power-of-two meshsizes on this path")
    ells = bin.ells
    even = [ell for ell in ells if ell % 2 == 0]
    nbins, nedges = bin.nbins, len(bin._edges1d)
    ncol = len(ellsin) * len(edgesin)
    wmat = np.zeros((ncol, len(ells), nbins), dtype='f8')
    if even:
        if len(even) > 6 or len(even) * (nedges - 1) > 4096:
            raise NotImplementedError('too many multipoles or bins for the window binning pass')
        n = _lib.i32x3(mattrs.meshsize)
        volume = float(np.prod(mattrs.boxsize, dtype='f8'))
        scale = volume if Q is None else volume / float(np.prod(mattrs.meshsize, dtype='f8'))  # Q = c2r(A1 conj(A2)) / N^3
        ws = _lib.workspace(_lib.load().mp_window_tophat_workspace_bytes(n, len(even), nedges), device)
        out = torch.empty((ncol, len(even), nbins, 2), dtype=torch.float64, device=device)
        dbl = lambda v: (_lib.ctypes.c_double * ncol)(*[float(x) for x in v])
        t = bin._tables
        _lib.call('mp_window_bin_tophat', _lib.stream_handle(device), _lib.ptr(Q) if Q is not None else None, n, float(cell[0]),
                  _lib.ptr(t['kx']), _lib.ptr(t['ky']), _lib.ptr(t['kz']), _lib.ptr(t['thr']), nedges, ncol,
                  dbl(list(edgesin[:, 0]) * len(ellsin)), dbl(list(edgesin[:, 1]) * len(ellsin)),
                  (_lib.ctypes.c_int32 * ncol)(*[ellin for ellin in ellsin for _ in edgesin]), _lib.f64x3(vlos), scale, len(even),
                  (_lib.ctypes.c_int32 * len(even))(*even), _lib.ptr(out), _lib.ptr(ws), ws.numel())
        sums = out[..., 0].cpu().numpy()
        with np.errstate(invalid='ignore', divide='ignore'):
            for ie, ell in enumerate(even):
                ill = ells.index(ell)
                wmat[:, ill, :] = (2 * ell + 1) * rnorm[ill] * sums[:, ie, :] / bin.nmodes
    return wmat.reshape(ncol, -1).T


def _window_local(meshes, fields, periodic, mattrs, edgesin, ellsin, bin, swap, norm, infinite, smooth=False):
    """compute_mesh2_spectrum_window with a local (firstpoint/endpoint) line of sight, theory line
    of sight 'firstpoint', wide-angle order 0 (mesh2.py:1133-1156, 1235-1305, 1463-1482).

    flags=('infinite',): for each input multipole lin and output (l, m) the selection term
    Qs = V_cell sum_m1 Y_lin,m1(shat) 4 pi / (2 lin + 1) c2r(conj(r2c(rho Y_lm Y_lin,m1)) A0) is
    formed once on the device; every input shell then takes one z pass (the shell's xi_lin(|s|)
    generated from a table and multiplied into Qs as it is loaded), one y pass and one binning x
    pass with Y_lm(khat) weights.

    No flags (theory on the mesh): the terms Q_lm,m1 = c2r(conj(r2c(rho Y_lm Y_lin,m1)) A0) are
    kept for every (l, m, m1); each closed input shell gives xi_m1 = c2r(shell Y_lin,m1(khat))
    (generated in the inverse x pass), then A_lm = sum_m1 xi_m1 Q_lm,m1 and one binning pass of
    r2c(A_lm) Y_lm(khat) per (l, m).

    flags=('smooth',): per (lin, l) the selection terms times Y_lm(shat) Y_lin,m1(shat), summed
    over (m, m1) on the device and binned once in separation shells, then Hankel-summed against
    each input shell on the host (mesh2.py:1195-1233)."""
    theory_los = 'firstpoint'
    if isinstance(ellsin, tuple) and len(ellsin) == 2 and isinstance(ellsin[1], str):
        ellsin, theory_los = ellsin
    if theory_los not in ('firstpoint', 'local'):
        raise NotImplementedError(f'window theory line of sight {theory_los!r} is not implemented on this path')
    ellsin = [ellsin] if np.ndim(ellsin) == 0 and not isinstance(ellsin, tuple) else list(ellsin)
    pairs = []
    for ellin in ellsin:
        ellin, wa = ellin if isinstance(ellin, tuple) else (ellin, 0)
        if wa != 0 and theory_los != 'firstpoint':
            raise NotImplementedError("wide-angle orders > 0 need the 'firstpoint' theory line of sight")
        pairs.append((int(ellin), int(wa)))
    was = [wa for _, wa in pairs]
    ellsin = [ellin for ellin, _ in pairs]
    ells = bin.ells
    if max(ells) > 4 or max(ellsin) > 4:
        raise NotImplementedError('local line-of-sight window multipoles up to 4 on this path')
    cell = np.asarray(mattrs.cellsize, dtype='f8')
    if infinite and not np.all(cell == cell[0]):
        raise NotImplementedError("window flags=('infinite',) needs equal cellsizes on this path")
    if not smooth and not all(int(m) >= 2 and int(m) & (int(m) - 1) == 0 for m in mattrs.meshsize):
        raise NotImplementedError('a local line-of-sight window needs power-of-two meshsizes on this path')
    n3 = float(np.prod(mattrs.meshsize, dtype='f8'))
    _norm = n3 / float(np.prod(mattrs.cellsize))
    if norm is None:
        norm = _norm
    norm = np.broadcast_to(np.asarray(_host(norm), dtype='f8').ravel() if np.ndim(_host(norm)) else float(_host(norm)),
                           (len(ells),)).copy()
    rnorm = _norm / norm / n3
    if hasattr(edgesin, 'ells') and hasattr(edgesin, 'edges'):
        edgesin = edgesin.edges('k')
    edgesin = np.asarray(edgesin, dtype='f8')
    if edgesin.ndim != 2:
        edgesin = np.column_stack([edgesin[:-1], edgesin[1:]])
    device = bin._device
    shape = tuple(int(n) for n in mattrs.meshsize)
    if periodic:
        rmesh = torch.ones(shape, dtype=torch.float32, device=device)
        A0 = _r2c(rmesh, mattrs.meshsize)
    else:
        if swap:
            meshes = meshes[::-1]
        rmesh = _2r(meshes[0]).value.contiguous().to(torch.float32)
        A0 = _2c(meshes[0] if fields[1] == fields[0] else meshes[1]).value
        device = rmesh.device
    a0, a0_pitch = _as_spectrum(A0)
    offset = mattrs.boxcenter - mattrs.boxsize / 2.
    xs = [torch.from_numpy(np.arange(int(n)) * float(d) + float(o)).to(device)
          for n, d, o in zip(mattrs.meshsize, mattrs.cellsize, offset)]  # positions (mesh.py:1594-1603)
    ss = [torch.from_numpy(v).to(device) for v in _svectors(mattrs)]
    if theory_los == 'local':
        rmesh2 = rmesh if (periodic or fields[1] == fields[0]) else _2r(meshes[1]).value.contiguous().to(torch.float32)
        if not (infinite or smooth):
            wmat = _window_local_theory_mesh(rmesh, rmesh2, mattrs, xs, edgesin, ellsin, bin, rnorm)
        else:
            wmat = _window_local_theory(rmesh, rmesh2, mattrs, xs, ss, edgesin, ellsin, bin, rnorm, float(cell[0]), smooth)
        return _window_matrix(wmat, bin, edgesin, ellsin, np.zeros(3), norm, los='firstpoint')
    if smooth:
        wmat = _window_local_smooth(rmesh, a0, a0_pitch, mattrs, xs, ss, edgesin, ellsin, bin, rnorm, was)
        return _window_matrix(wmat, bin, edgesin, ellsin, np.zeros(3), norm, los='firstpoint', was=was)
    n = _lib.i32x3(mattrs.meshsize)
    stream = _lib.stream_handle(device)
    nbins, nedges, nin = bin.nbins, len(bin._edges1d), len(edgesin)
    ws = _lib.workspace(_lib.load().mp_window_tophat_workspace_bytes(n, 1, nedges), device)
    t = bin._tables
    k0 = (_lib.ctypes.c_double * nin)(*[float(e) for e in edgesin[:, 0]])
    k1 = (_lib.ctypes.c_double * nin)(*[float(e) for e in edgesin[:, 1]])
    base = torch.empty_like(rmesh)
    tmp = torch.empty_like(rmesh)
    Qs = torch.empty_like(rmesh)
    # per (ellin, l, m): the binned sums of every input shell, summed over m on the host
    out = torch.empty((len(ellsin), len(ells), 2 * max(ells) + 1, nin, nbins, 2), dtype=torch.float64, device=device)
    if not infinite:
        _window_local_mesh_theory(rmesh, a0, a0_pitch, mattrs, xs, ss, edgesin, ellsin, bin, out, was)
    for iin, ellin in enumerate(ellsin if infinite else ()):
        wa = was[iin]
        for ill, ell in enumerate(ells):
            for m in range(-ell, ell + 1):
                _mul_ylm(rmesh, base, mattrs, xs, ell, m, False)
                if wa:
                    _mul_radial_power(base, base, mattrs, xs, -wa)
                for m1 in range(-ellin, ellin + 1):
                    _mul_ylm(base, tmp, mattrs, xs, ellin, m1, False)
                    B = _r2c(tmp, mattrs.meshsize)
                    _lib.call('mp_spectrum_product', stream, _lib.ptr(B), _spectrum_pitch(B), _lib.ptr(a0), a0_pitch, n, _lib.ptr(B),
                              _spectrum_pitch(B))
                    q = _c2r(B, mattrs.meshsize)
                    if wa:
                        _mul_radial_power(q, q, mattrs, ss, wa)
                    _mul_ylm(q, Qs, mattrs, ss, ellin, m1, m1 != -ellin)
                    del B, q
                scale = 4. * np.pi / (2 * ellin + 1) * float(np.prod(mattrs.cellsize))
                _lib.call('mp_window_bin_tophat_local', stream, _lib.ptr(Qs), n, float(cell[0]), _lib.ptr(t['kx']), _lib.ptr(t['ky']),
                          _lib.ptr(t['kz']), _lib.ptr(t['thr']), nedges, nin, k0, k1, int(ellin), int(ell), int(m), int(ell == 0),
                          scale, _lib.ptr(out[iin, ill, m + ell]), _lib.ptr(ws), ws.numel())
    sums = out[..., 0].cpu().numpy()
    wmat = np.zeros((len(ellsin), nin, len(ells), nbins))
    for ill, ell in enumerate(ells):
        wmat[:, :, ill, :] = sums[:, ill, :2 * ell + 1].sum(axis=1)
    if not infinite:  # the 4 pi / (2 lin + 1) of Q_lm,m1 (mesh2.py:1269)
        wmat *= np.array([4. * np.pi / (2 * ellin + 1) for ellin in ellsin])[:, None, None, None]
    with np.errstate(invalid='ignore', divide='ignore'):
        wmat = wmat * (4. * np.pi * rnorm)[None, None, :, None] / bin.nmodes
    wmat = wmat.reshape(len(ellsin) * nin, -1).T
    attrs = dict(los='firstpoint')
    zeros = np.zeros(nbins)
    observable = Mesh2SpectrumPoles([Mesh2SpectrumPole(k=bin.xavg, k_edges=bin.edges, nmodes=bin.nmodes, num_raw=zeros,
                                                       num_shotnoise=zeros, norm=np.full(nbins, norm[ill]), ell=ell, attrs=attrs)
                                     for ill, ell in enumerate(ells)], attrs=attrs)
    kin = edgesin.mean(axis=-1)
    theory = Mesh2SpectrumPoles([Mesh2SpectrumPole(k=kin, k_edges=edgesin, nmodes=np.zeros(len(kin), dtype='i8'),
                                                   num_raw=np.zeros(len(kin)), ell=ell, attrs=dict(wa_order=wa))
                                 for ell, wa in zip(ellsin, was)])
    return WindowMatrix(observable=observable, theory=theory, value=wmat)


def _window_local_mesh_theory(rmesh, a0, a0_pitch, mattrs, xs, ss, edgesin, ellsin, bin, out, was=None):
    """Binned sums of the local window with the theory on the mesh (mesh2.py:1262-1305) into
    out[iin, ill, m + l, ie] (without the 4 pi / (2 lin + 1) of Q and the 4 pi rnorm / nmodes)."""
    device = rmesh.device
    n = _lib.i32x3(mattrs.meshsize)
    stream = _lib.stream_handle(device)
    ells = bin.ells
    t = bin._tables
    nedges = len(bin._edges1d)
    ncell = rmesh.numel()
    ws_shell = _lib.workspace(_lib.load().mp_window_shell_ylm_workspace_bytes(n), device)
    ws_bin = _lib.workspace(_lib.load().mp_bin_ylm_real_workspace_bytes(n, nedges), device)
    # shell [e0, e1] closed on the right (mesh2.py:1276): k >= e0 and not k >= nextafter(e1)
    lo = _k2_thresholds(edgesin[:, 0])
    hi = _k2_thresholds(np.nextafter(edgesin[:, 1], np.inf))
    base = torch.empty_like(rmesh)
    tmp = torch.empty_like(rmesh)
    acc = torch.empty_like(rmesh)
    for iin, ellin in enumerate(ellsin):
        wa = was[iin] if was else 0
        nm1 = 2 * ellin + 1
        Q = {}
        for ell in ells:
            for m in range(-ell, ell + 1):
                _mul_ylm(rmesh, base, mattrs, xs, ell, m, False)
                if wa:
                    _mul_radial_power(base, base, mattrs, xs, -wa)
                for m1 in range(-ellin, ellin + 1):
                    _mul_ylm(base, tmp, mattrs, xs, ellin, m1, False)
                    B = _r2c(tmp, mattrs.meshsize)
                    _lib.call('mp_spectrum_product', stream, _lib.ptr(B), _spectrum_pitch(B), _lib.ptr(a0), a0_pitch, n, _lib.ptr(B),
                              _spectrum_pitch(B))
                    Q[ell, m, m1] = _c2r(B, mattrs.meshsize)
                    if wa:
                        _mul_radial_power(Q[ell, m, m1], Q[ell, m, m1], mattrs, ss, wa)
                    del B
        xi = torch.empty((nm1,) + tuple(rmesh.shape), dtype=torch.float32, device=device)
        for ie in range(len(edgesin)):
            for m1 in range(-ellin, ellin + 1):
                _lib.call('mp_window_shell_ylm', stream, n, _lib.ptr(t['kx']), _lib.ptr(t['ky']), _lib.ptr(t['kz']), float(lo[ie]),
                          float(hi[ie]), int(ellin), int(m1), 1., _lib.ptr(xi[m1 + ellin]), _lib.ptr(ws_shell), ws_shell.numel())
            for ill, ell in enumerate(ells):
                for m in range(-ell, ell + 1):
                    for m1 in range(-ellin, ellin + 1):
                        _lib.call('mp_mul_acc', stream, _lib.ptr(xi[m1 + ellin]), _lib.ptr(Q[ell, m, m1]), _lib.ptr(acc), ncell,
                                  int(m1 != -ellin))
                    _lib.call('mp_bin_ylm_real', stream, _lib.ptr(acc), n, _lib.ptr(t['kx']), _lib.ptr(t['ky']), _lib.ptr(t['kz']),
                              _lib.ptr(t['thr']), nedges, int(ell), int(m), int(ell == 0), _lib.ptr(out[iin, ill, m + ell, ie]),
                              _lib.ptr(ws_bin), ws_bin.numel())
        del Q, xi


_LOCAL_THEORY_COEFF2 = {(0, 0): [(0, 1.), (4, -7. / 18.)], (0, 2): [(2, 1. / 2.), (4, -5. / 18.)], (2, 2): [(4, 35. / 18.)]}
_LOCAL_THEORY_COEFF2[2, 0] = _LOCAL_THEORY_COEFF2[0, 2]  # the local-theory multipoles, mesh2.py:1449-1452


def _window_local_theory_mesh(rmesh1, rmesh2, mattrs, xs, edgesin, ellsin, bin, rnorm):
    """(n_ells n_bins, n_ellsin n_in) matrix of the window with the 'local' theory line of sight
    and the theory on the mesh (mesh2.py:1415-1457): per (l1, l2) the terms
    Q_lm,m1,m2 = (4 pi)^2/((2l1+1)(2l2+1)) c2r(conj(r2c(rho1 Y_lm Y_l1,m1)) r2c(rho2 Y_l2,m2)) are
    kept on the device; each closed input shell gives xi_m1,m2 = c2r(shell Y_l1,m1(khat)
    Y_l2,m2(-khat)), then A_lm = sum xi Q and one Y_lm(khat) binning pass per (l, m)."""
    lib = _lib.load()
    device = rmesh1.device
    ells = bin.ells
    n = _lib.i32x3(mattrs.meshsize)
    stream = _lib.stream_handle(device)
    ncell = rmesh1.numel()
    nbins, nedges, nin = bin.nbins, len(bin._edges1d), len(edgesin)
    t = bin._tables
    ws_shell = _lib.workspace(lib.mp_window_shell_ylm_workspace_bytes(n), device)
    ws_bin = _lib.workspace(lib.mp_bin_ylm_real_workspace_bytes(n, nedges), device)
    lo = _k2_thresholds(edgesin[:, 0])
    hi = _k2_thresholds(np.nextafter(edgesin[:, 1], np.inf))  # closed right edge (mesh2.py:1436)
    tmp, acc = torch.empty_like(rmesh1), torch.empty_like(rmesh1)
    out = torch.empty((nin, len(ells), 2 * max(ells) + 1, nbins, 2), dtype=torch.float64, device=device)
    blocks = {}
    for ell1, ell2 in ((2, 2), (2, 0), (0, 2), (0, 0)):
        c12 = (4. * np.pi)**2 / ((2 * ell1 + 1) * (2 * ell2 + 1))
        G2 = {}
        for m2 in range(-ell2, ell2 + 1):
            _mul_ylm(rmesh2, tmp, mattrs, xs, ell2, m2, False)
            G2[m2] = _r2c(tmp, mattrs.meshsize)
        base = torch.empty_like(rmesh1)
        Q = {}
        for ell in ells:
            for m in range(-ell, ell + 1):
                _mul_ylm(rmesh1, base, mattrs, xs, ell, m, False)
                for m1 in range(-ell1, ell1 + 1):
                    _mul_ylm(base, tmp, mattrs, xs, ell1, m1, False)
                    F = _r2c(tmp, mattrs.meshsize)
                    for m2 in range(-ell2, ell2 + 1):
                        B = _empty_spectrum(mattrs.meshsize, device)
                        _lib.call('mp_spectrum_product', stream, _lib.ptr(F), _spectrum_pitch(F), _lib.ptr(G2[m2]), _spectrum_pitch(G2[m2]),
                                  n, _lib.ptr(B), _spectrum_pitch(B))
                        Q[ell, m, m1, m2] = _c2r(B, mattrs.meshsize)
                        del B
                    del F
        pairs = [(m1, m2) for m1 in range(-ell1, ell1 + 1) for m2 in range(-ell2, ell2 + 1)]
        xi = torch.empty((len(pairs),) + tuple(rmesh1.shape), dtype=torch.float32, device=device)
        for ie in range(nin):
            for i, (m1, m2) in enumerate(pairs):
                _lib.call('mp_window_shell_ylm2', stream, n, _lib.ptr(t['kx']), _lib.ptr(t['ky']), _lib.ptr(t['kz']), float(lo[ie]),
                          float(hi[ie]), int(ell1), int(m1), int(ell2), int(m2), c12, _lib.ptr(xi[i]), _lib.ptr(ws_shell),
                          ws_shell.numel())
            for ill, ell in enumerate(ells):
                for m in range(-ell, ell + 1):
                    for i, (m1, m2) in enumerate(pairs):
                        _lib.call('mp_mul_acc', stream, _lib.ptr(xi[i]), _lib.ptr(Q[ell, m, m1, m2]), _lib.ptr(acc), ncell, int(i > 0))
                    _lib.call('mp_bin_ylm_real', stream, _lib.ptr(acc), n, _lib.ptr(t['kx']), _lib.ptr(t['ky']), _lib.ptr(t['kz']),
                              _lib.ptr(t['thr']), nedges, int(ell), int(m), int(ell == 0), _lib.ptr(out[ie, ill, m + ell]),
                              _lib.ptr(ws_bin), ws_bin.numel())
        sums = out[..., 0].cpu().numpy()
        W = np.zeros((nin, len(ells), nbins))
        for ill, ell in enumerate(ells):
            W[:, ill] = sums[:, ill, :2 * ell + 1].sum(axis=1)
        with np.errstate(invalid='ignore', divide='ignore'):
            blocks[ell1, ell2] = W * (4. * np.pi * rnorm)[None, :, None] / bin.nmodes
        del Q, xi
    wmat = np.zeros((len(ellsin), nin, len(ells), nbins))
    for illin, ellin in enumerate(ellsin):
        for key, lst in _LOCAL_THEORY_COEFF2.items():
            wmat[illin] += sum(c * (ell == ellin) for ell, c in lst) * blocks[key]
    return wmat.reshape(len(ellsin) * nin, -1).T


def _window_local_theory(rmesh1, rmesh2, mattrs, xs, ss, edgesin, ellsin, bin, rnorm, cell, smooth=False):
    """(n_ells n_bins, n_ellsin n_in) matrix of the window with the 'local' theory line of sight and
    an infinite-resolution theory (mesh2.py:1382-1413, 1446-1457).  For (l1, l2) in {2, 0}^2 and
    output (l, m) the Gaunt-weighted selection terms
    Qs_p = V_cell sum G(p mp; l1 m1; l2 m2) Y_p,mp(shat) (4 pi)^2/((2l1+1)(2l2+1))
    c2r(conj(r2c(rho1 Y_lm Y_l1,m1)) r2c(rho2 Y_l2,m2)) are formed on the device (Gaunt
    coefficients in the library's own Y convention), then each p takes the shells through
    mp_window_bin_tophat_local with Bessel order p; the input multipoles combine the (l1, l2)
    blocks as mesh2.py:1454-1457 does.  With smooth (mesh2.py:1346-1380) the terms are multiplied
    by Y_lm(shat) instead, summed over m, binned once per p in separation shells and
    Hankel-summed on the host."""
    lib = _lib.load()
    device = rmesh1.device
    ells = bin.ells
    n = _lib.i32x3(mattrs.meshsize)
    stream = _lib.stream_handle(device)
    ncell = rmesh1.numel()
    nbins, nedges, nin = bin.nbins, len(bin._edges1d), len(edgesin)
    t = bin._tables
    ws = _lib.workspace(lib.mp_window_tophat_workspace_bytes(n, 1, nedges), device)
    k0 = (_lib.ctypes.c_double * nin)(*[float(e) for e in edgesin[:, 0]])
    k1 = (_lib.ctypes.c_double * nin)(*[float(e) for e in edgesin[:, 1]])
    vcell = float(np.prod(mattrs.cellsize, dtype='f8'))
    base, tmp, tmp2 = torch.empty_like(rmesh1), torch.empty_like(rmesh1), torch.empty_like(rmesh1)
    tmp3 = torch.empty_like(rmesh1)
    out = torch.empty((nin, nbins, 2), dtype=torch.float64, device=device)
    if smooth:
        sbin = BinMesh2CorrelationPoles(mattrs, edges=None, device=device)
        empty = sbin.nmodes == 0
        with np.errstate(invalid='ignore', divide='ignore'):
            savg = np.where(empty, 0., sbin.xavg)
            kout = np.where(bin.nmodes == 0, 0., bin.xavg)
        snm = sbin.nmodes.astype('f8')
        jn = {ell: _host_array_fn('mp_spherical_jn_array', ell, np.outer(kout, savg)) for ell in ells}
    blocks = {}
    for ell1, ell2 in ((2, 2), (2, 0), (0, 2), (0, 0)):
        G2 = {}
        for m2 in range(-ell2, ell2 + 1):
            _mul_ylm(rmesh2, tmp, mattrs, xs, ell2, m2, False)
            G2[m2] = _r2c(tmp, mattrs.meshsize)
        ps = [p for p in (0, 2, 4) if abs(lib.mp_real_gaunt(p, 0, ell1, 0, ell2, 0)) > 1e-6]
        W = np.zeros((nin, len(ells), nbins))
        for ill, ell in enumerate(ells):
            if smooth:
                Qs = {p: torch.zeros_like(rmesh1) for p in ps}
            for m in range(-ell, ell + 1):
                if not smooth:
                    Qs = {p: torch.zeros_like(rmesh1) for p in ps}
                _mul_ylm(rmesh1, base, mattrs, xs, ell, m, False)
                for m1 in range(-ell1, ell1 + 1):
                    _mul_ylm(base, tmp, mattrs, xs, ell1, m1, False)
                    F = _r2c(tmp, mattrs.meshsize)
                    for m2 in range(-ell2, ell2 + 1):
                        B = _empty_spectrum(mattrs.meshsize, device)
                        _lib.call('mp_spectrum_product', stream, _lib.ptr(F), _spectrum_pitch(F), _lib.ptr(G2[m2]), _spectrum_pitch(G2[m2]),
                                  n, _lib.ptr(B), _spectrum_pitch(B))
                        Q = _c2r(B, mattrs.meshsize)
                        del B
                        for p in ps:
                            for mp in range(-p, p + 1):
                                rg = lib.mp_real_gaunt(p, mp, ell1, m1, ell2, m2)
                                if abs(rg) > 1e-6:
                                    _mul_ylm(Q, tmp2, mattrs, ss, p, mp, False)
                                    if smooth:
                                        _mul_ylm(tmp2, tmp3, mattrs, ss, ell, m, False)
                                    _lib.call('mp_axpy', stream, float(rg * (4. * np.pi)**2 / ((2 * ell1 + 1) * (2 * ell2 + 1))),
                                              _lib.ptr(tmp3 if smooth else tmp2), _lib.ptr(Qs[p]), ncell)
                    del F
                if smooth:
                    continue
                for p in ps:
                    _lib.call('mp_window_bin_tophat_local', stream, _lib.ptr(Qs[p]), n, cell, _lib.ptr(t['kx']), _lib.ptr(t['ky']),
                              _lib.ptr(t['kz']), _lib.ptr(t['thr']), nedges, nin, k0, k1, int(p), int(ell), int(m), int(ell == 0),
                              vcell, _lib.ptr(out), _lib.ptr(ws), ws.numel())
                    W[:, ill, :] += out[..., 0].cpu().numpy()
            if smooth:  # Hankel sums of the binned terms (mesh2.py:1366-1380)
                for p in ps:
                    with np.errstate(invalid='ignore', divide='ignore'):
                        Qp = sbin._sums_dev(Qs[p], [0], [False])[0].cpu().numpy() / snm * (4. * np.pi * vcell)
                    if ell != 0:
                        Qp[0] = 0.
                    Qp = np.where(empty, 0., Qp)
                    for ie, (e0, e1) in enumerate(edgesin):
                        xi = _host_array_fn('mp_tophat_shell_xi_array', p, savg, float(e0), float(e1))
                        W[ie, ill] += (-1)**(ell // 2) * (jn[ell] @ (snm * xi * Qp)) * rnorm[ill]
        if smooth:
            blocks[ell1, ell2] = W
            continue
        with np.errstate(invalid='ignore', divide='ignore'):
            blocks[ell1, ell2] = W * (4. * np.pi * rnorm)[None, :, None] / bin.nmodes
    wmat = np.zeros((len(ellsin), nin, len(ells), nbins))
    for illin, ellin in enumerate(ellsin):
        for key, lst in _LOCAL_THEORY_COEFF2.items():
            wmat[illin] += sum(c * (ell == ellin) for ell, c in lst) * blocks[key]
    return wmat.reshape(len(ellsin) * nin, -1).T


def _window_local_smooth(rmesh, a0, a0_pitch, mattrs, xs, ss, edgesin, ellsin, bin, rnorm, was=None):
    """(n_ells n_bins, n_ellsin n_in) matrix of the local smooth window (mesh2.py:1195-1233)."""
    device = rmesh.device
    n = _lib.i32x3(mattrs.meshsize)
    stream = _lib.stream_handle(device)
    ells = bin.ells
    sbin = BinMesh2CorrelationPoles(mattrs, edges=None, device=device)
    empty = sbin.nmodes == 0
    with np.errstate(invalid='ignore', divide='ignore'):
        savg = np.where(empty, 0., sbin.xavg)
        kout = np.where(bin.nmodes == 0, 0., bin.xavg)
    snm = sbin.nmodes.astype('f8')
    vcell = float(np.prod(mattrs.cellsize, dtype='f8'))
    jn = {ell: _host_array_fn('mp_spherical_jn_array', ell, np.outer(kout, savg)) for ell in ells}
    base = torch.empty_like(rmesh)
    tmp = torch.empty_like(rmesh)
    acc = torch.empty_like(rmesh)
    blocks = []
    for iin, ellin in enumerate(ellsin):
        wa = was[iin] if was else 0
        rows = np.zeros((len(edgesin), len(ells), bin.nbins))
        xi = [_host_array_fn('mp_tophat_shell_xi_array', ellin, savg, float(e0), float(e1)) for e0, e1 in edgesin]
        for ill, ell in enumerate(ells):
            first = True
            for m1 in range(-ellin, ellin + 1):
                for m in range(-ell, ell + 1):
                    _mul_ylm(rmesh, base, mattrs, xs, ell, m, False)
                    if wa:
                        _mul_radial_power(base, base, mattrs, xs, -wa)
                    _mul_ylm(base, tmp, mattrs, xs, ellin, m1, False)
                    B = _r2c(tmp, mattrs.meshsize)
                    _lib.call('mp_spectrum_product', stream, _lib.ptr(B), _spectrum_pitch(B), _lib.ptr(a0), a0_pitch, n, _lib.ptr(B),
                              _spectrum_pitch(B))
                    Q = _c2r(B, mattrs.meshsize)
                    if wa:
                        _mul_radial_power(Q, Q, mattrs, ss, wa)
                    del B
                    _mul_ylm(Q, tmp, mattrs, ss, ell, m, False)
                    _mul_ylm(tmp, acc, mattrs, ss, ellin, m1, not first)
                    first = False
            with np.errstate(invalid='ignore', divide='ignore'):
                Qs = sbin._sums_dev(acc, [0], [False])[0].cpu().numpy() / snm
            Qs *= 4. * np.pi * 4. * np.pi / (2 * ellin + 1) * vcell
            if ell != 0:
                Qs[0] = 0.
            Qs = np.where(empty, 0., Qs)
            for ie in range(len(edgesin)):
                rows[ie, ill] = (-1)**(ell // 2) * (jn[ell] @ (snm * xi[ie] * Qs)) * rnorm[ill]
        blocks.append(rows.reshape(len(edgesin), -1))
    return np.concatenate(blocks, axis=0).T


def _theory_tables(mattrs, theory, device):
    """Parse the theory of compute_mesh2_spectrum_mean (mesh2.py:1549-1567) into
    (theory_los, {(ell, wa): values}, keys): piecewise-constant values on the squared-edge
    thresholds keys (Interpolator1D of order 0, utils.py:507-537), or, for callables of |k|,
    values at the sorted distinct k^2 of the mesh."""
    poles, kin, theory_los = theory, None, 'firstpoint'
    if isinstance(poles, tuple) and isinstance(poles[-1], str):
        theory_los = poles[-1]
        poles = poles[0] if len(poles) == 2 else poles[:-1]
    if isinstance(poles, tuple):
        kin, poles = poles
    if hasattr(poles, 'ells') and hasattr(poles, 'get'):  # a poles container with k_edges
        e = poles.edges('k')
        kin = np.append(e[:, 0], e[-1, 1])
        poles = {ell: poles.get(ell).value() for ell in poles.ells}
    if isinstance(poles, list):
        poles = {ell: pole for ell, pole in zip((0, 2, 4), poles)}
    if not isinstance(poles, dict):
        raise NotImplementedError('theory must be given per multipole (dict, list or poles container) on this path')
    poles = {ell if isinstance(ell, tuple) else (ell, 0): pole for ell, pole in poles.items()}
    if all(callable(p) for p in poles.values()):
        kx, ky, kz = (torch.from_numpy(k).to(device) for k in _kvectors(mattrs))
        k2 = (kx[:, None, None]**2 + ky[None, :, None]**2) + kz[None, None, :]**2
        keys = torch.unique(k2.reshape(-1)).cpu().numpy()  # sorted distinct k^2, summed as the kernels do
        del k2
        knorm = np.sqrt(keys)
        values = {key: np.asarray(p(knorm), dtype='f8') * np.ones_like(knorm) for key, p in poles.items()}
        return theory_los, values, keys, len(keys)
    if kin is None:
        raise ValueError('tabulated poles need their k (or k edges)')
    kin = np.asarray(kin, dtype='f8')
    values = {key: np.asarray(p, dtype='f8') for key, p in poles.items()}
    n = len(next(iter(values.values())))
    if len(kin) != n + 1:  # points: mid-point edges, ends extrapolated (utils.py:515-517)
        tmp = (kin[:-1] + kin[1:]) / 2.
        kin = np.concatenate([[tmp[0] - (kin[1] - kin[0])], tmp, [tmp[-1] + (kin[-1] - kin[-2])]])
    return theory_los, values, _k2_thresholds(kin), n


def _mean_local_theory(rmesh1, rmesh2, mattrs, xs, bin, values, dkeys, ntab, rnorm, device):
    """num_raw of compute_mesh2_spectrum_mean with the 'local' theory line of sight
    (mesh2.py:1647-1662): kernels K_l1l2 = sum c P_l (the coefficients of mesh2.py:1648-1651),
    xi = c2r(K Y_l1,m1(khat) Y_l2,m2(-khat)) (generated in the inverse x pass),
    Q_lm = sum (4 pi)^2/((2l1+1)(2l2+1)) xi c2r(conj(r2c(rho1 Y_lm Y_l1,m1)) r2c(rho2 Y_l2,m2)),
    num_l = 4 pi bin(sum_m r2c(Q_lm) Y_lm(khat))."""
    lib = _lib.load()
    ells = bin.ells
    if any(wa for _, wa in values):
        raise NotImplementedError("the 'local' theory line of sight takes wide-angle order 0")
    n = _lib.i32x3(mattrs.meshsize)
    stream = _lib.stream_handle(device)
    t = bin._tables
    nedges = len(bin._edges1d)
    ncell = rmesh1.numel()
    ws_xi = _lib.workspace(lib.mp_window_shell_ylm_workspace_bytes(n), device)
    ws_bin = _lib.workspace(lib.mp_bin_ylm_real_workspace_bytes(n, nedges), device)
    kern = {}
    for key, lst in _LOCAL_THEORY_COEFF2.items():
        v = sum(c * values[l, 0] for l, c in lst if (l, 0) in values)
        kern[key] = torch.from_numpy(np.ascontiguousarray(np.broadcast_to(v, (ntab,)), dtype='f4')).to(device)
    pairs = [(l1, m1, l2, m2) for l1, l2 in ((0, 0), (0, 2), (2, 0), (2, 2))
             for m1 in range(-l1, l1 + 1) for m2 in range(-l2, l2 + 1)]
    xi = {}
    for l1, m1, l2, m2 in pairs:
        x = torch.empty_like(rmesh1)
        scale = rnorm * (4. * np.pi)**2 / ((2 * l1 + 1) * (2 * l2 + 1))
        lib_args = (stream, n, _lib.ptr(t['kx']), _lib.ptr(t['ky']), _lib.ptr(t['kz']), len(dkeys), _lib.ptr(dkeys), ntab,
                    _lib.ptr(kern[l1, l2]), l1, m1, l2, m2, scale, _lib.ptr(x), _lib.ptr(ws_xi), ws_xi.numel())
        _lib.call('mp_theory_ylm2_real', *lib_args)
        xi[l1, m1, l2, m2] = x
    G2 = {}
    tmp, base, acc = torch.empty_like(rmesh1), torch.empty_like(rmesh1), torch.empty_like(rmesh1)
    for l2 in (0, 2):
        for m2 in range(-l2, l2 + 1):
            _mul_ylm(rmesh2, tmp, mattrs, xs, l2, m2, False)
            G2[l2, m2] = _r2c(tmp, mattrs.meshsize)
    out = torch.empty((len(ells), 2 * max(ells) + 1, bin.nbins, 2), dtype=torch.float64, device=device)
    for ill, ell in enumerate(ells):
        for m in range(-ell, ell + 1):
            _mul_ylm(rmesh1, base, mattrs, xs, ell, m, False)
            first = True
            F = {}
            for l1, m1, l2, m2 in pairs:
                if (l1, m1) not in F:
                    _mul_ylm(base, tmp, mattrs, xs, l1, m1, False)
                    F[l1, m1] = _r2c(tmp, mattrs.meshsize)
                B = _empty_spectrum(mattrs.meshsize, device)
                _lib.call('mp_spectrum_product', stream, _lib.ptr(F[l1, m1]), _spectrum_pitch(F[l1, m1]), _lib.ptr(G2[l2, m2]),
                          _spectrum_pitch(G2[l2, m2]), n, _lib.ptr(B), _spectrum_pitch(B))
                Q = _c2r(B, mattrs.meshsize)
                del B
                _lib.call('mp_mul_acc', stream, _lib.ptr(xi[l1, m1, l2, m2]), _lib.ptr(Q), _lib.ptr(acc), ncell, int(not first))
                first = False
            del F
            _lib.call('mp_bin_ylm_real', stream, _lib.ptr(acc), n, _lib.ptr(t['kx']), _lib.ptr(t['ky']), _lib.ptr(t['kz']), _lib.ptr(t['thr']),
                      nedges, int(ell), int(m), 0, _lib.ptr(out[ill, m + ell]), _lib.ptr(ws_bin), ws_bin.numel())
    host_out = out[..., 0].cpu().numpy()
    sums = np.zeros((len(ells), bin.nbins))
    with np.errstate(invalid='ignore', divide='ignore'):
        for ill, ell in enumerate(ells):
            sums[ill] = 4. * np.pi * host_out[ill, :2 * ell + 1].sum(axis=0) / bin.nmodes
    return sums


def compute_mesh2_spectrum_mean(*meshes, theory, bin=None, los='z'):
    """Mean power-spectrum multipoles of a theory seen through a selection (mesh2.py:1485-1680),
    for a theory given per multipole (dict/list of arrays on k or k edges, a poles container, or
    callables of |k|; wide-angle order 0).  Global line of sight: the theory P(k) L_lin(mu) rnorm
    on the mesh is convolved as r2c(Q c2r(.)), Q = c2r(A1 conj(A2)) (N^3 for a periodic box given
    as a MeshAttrs), and binned with (2l+1) P_l(mu) in one fused pass per input multipole.  Local
    (firstpoint/endpoint) line of sight with theory line of sight 'firstpoint':
    4 pi bin(sum_m r2c(Q_m) Y_lm(khat)) with Q_m = sum_{lin, m1} 4 pi/(2 lin + 1)
    c2r(P_lin rnorm Y_lin,m1(khat)) c2r(conj(r2c(rho Y_lm Y_lin,m1)) A0)."""
    if not isinstance(bin, BinMesh2SpectrumPoles):
        raise ValueError(f'bin must be BinMesh2SpectrumPoles, not {type(bin)}')
    _single_shift(bin, 'compute_mesh2_spectrum_mean')
    periodic = isinstance(meshes[0], MeshAttrs)
    if periodic:
        mattrs = meshes[0]
        fields = (0, 0)
    else:
        meshes, fields = format_meshes(*meshes, nmeshes=2)
        mattrs = meshes[0].attrs
    if not all(int(m) >= 2 and int(m) & (int(m) - 1) == 0 for m in mattrs.meshsize):
        raise NotImplementedError('compute_mesh2_spectrum_mean needs power-of-two meshsizes on this path')
    device = bin._device
    n3 = float(np.prod(mattrs.meshsize, dtype='f8'))
    vcell = float(np.prod(mattrs.cellsize, dtype='f8'))
    norm = n3 / vcell
    rnorm = 1. / vcell
    los, vlos, swap = _format_los(los)
    attrs = dict(los=vlos.tolist() if vlos is not None else los)
    ells = bin.ells
    theory_los, values, keys, ntab = _theory_tables(mattrs, theory, device)
    shape = tuple(int(v) for v in mattrs.meshsize)
    n = _lib.i32x3(mattrs.meshsize)
    t = bin._tables
    nedges = len(bin._edges1d)
    dkeys = torch.from_numpy(np.ascontiguousarray(keys, dtype='f8')).to(device)
    dvals = {key: torch.from_numpy(np.ascontiguousarray(v, dtype='f4')).to(device) for key, v in values.items()}
    stream = _lib.stream_handle(device)
    sums = np.zeros((len(ells), bin.nbins))
    if vlos is not None and any(wa for _, wa in values):
        raise NotImplementedError('wide-angle orders > 0 need a local line of sight')
    if vlos is not None:
        if periodic:
            Q = torch.full(shape, n3, dtype=torch.float32, device=device)
        else:
            A1 = _2c(meshes[0])
            P, _ = _product_spectrum(A1) if fields[1] == fields[0] else _product_spectrum(_2c(meshes[1]), A1)
            Q = _c2r(P, mattrs.meshsize)
            del P
        even = [ell for ell in ells if ell % 2 == 0]
        if even:
            ws = _lib.workspace(_lib.load().mp_window_bin_workspace_bytes(n, len(even), nedges), device)
            out = torch.empty((len(values), len(even), bin.nbins, 2), dtype=torch.float64, device=Q.device)
            for i, ((ellin, _), v) in enumerate(dvals.items()):
                _lib.call('mp_spectrum_mean_bin', stream, _lib.ptr(Q), n, _lib.ptr(t['kx']), _lib.ptr(t['ky']), _lib.ptr(t['kz']),
                          _lib.ptr(t['thr']), nedges, len(keys), _lib.ptr(dkeys), ntab, _lib.ptr(v), int(ellin), _lib.f64x3(vlos),
                          rnorm, len(even), (_lib.ctypes.c_int32 * len(even))(*even), _lib.ptr(out[i]), _lib.ptr(ws), ws.numel())
            tot = out[..., 0].sum(dim=0).cpu().numpy() if len(values) > 1 else out[0, ..., 0].cpu().numpy()
            with np.errstate(invalid='ignore', divide='ignore'):
                for ie, ell in enumerate(even):  # odd ells: P_l(mu) + P_l(-mu) = 0 (mesh2.py:1599-1602)
                    sums[ells.index(ell)] = (2 * ell + 1) * tot[ie] / bin.nmodes
    else:
        if theory_los not in ('firstpoint', 'local'):
            raise NotImplementedError(f'theory line of sight {theory_los!r} is not implemented on this path')
        if max(ells) > 4 or max(ell for ell, _ in values) > 4:
            raise NotImplementedError('local line-of-sight multipoles up to 4 on this path')
        if periodic:
            rmesh = torch.ones(shape, dtype=torch.float32, device=device)
            A0 = _r2c(rmesh, mattrs.meshsize)
        else:
            if swap:
                meshes = meshes[::-1]
            rmesh = _2r(meshes[0]).value.contiguous().to(torch.float32)
            A0 = _2c(meshes[0] if fields[1] == fields[0] else meshes[1]).value
            device = rmesh.device
        a0, a0_pitch = _as_spectrum(A0)
        offset = mattrs.boxcenter - mattrs.boxsize / 2.
        xs = [torch.from_numpy(np.arange(int(m)) * float(d) + float(o)).to(device)
              for m, d, o in zip(mattrs.meshsize, mattrs.cellsize, offset)]
        if theory_los == 'local':
            rmesh2 = rmesh if (periodic or fields[1] == fields[0]) else _2r(meshes[1]).value.contiguous().to(torch.float32)
            sums = _mean_local_theory(rmesh, rmesh2, mattrs, xs, bin, values, dkeys, ntab, rnorm, device)
            poles = [Mesh2SpectrumPole(k=bin.xavg, k_edges=bin.edges, nmodes=bin.nmodes, num_raw=sums[ill], num_shotnoise=np.zeros(bin.nbins),
                                       norm=np.full(bin.nbins, norm), volume=np.prod(mattrs.kfun) * bin.nmodes, ell=ell, attrs=attrs)
                     for ill, ell in enumerate(ells)]
            return Mesh2SpectrumPoles(poles, attrs=attrs)
        ws_xi = _lib.workspace(_lib.load().mp_window_shell_ylm_workspace_bytes(n), device)
        ws_bin = _lib.workspace(_lib.load().mp_bin_ylm_real_workspace_bytes(n, nedges), device)
        ss = [torch.from_numpy(v).to(device) for v in _svectors(mattrs)]
        xi = {}
        for (ell1, wa), v in dvals.items():
            for m1 in range(-ell1, ell1 + 1):
                x = torch.empty(shape, dtype=torch.float32, device=device)
                _lib.call('mp_theory_ylm_real', stream, n, _lib.ptr(t['kx']), _lib.ptr(t['ky']), _lib.ptr(t['kz']), len(keys),
                          _lib.ptr(dkeys), ntab, _lib.ptr(v), int(ell1), int(m1), rnorm * 4. * np.pi / (2 * ell1 + 1),
                          _lib.ptr(x), _lib.ptr(ws_xi), ws_xi.numel())
                if wa:  # |s|^wa (mesh2.py:1643)
                    _mul_radial_power(x, x, mattrs, ss, wa)
                xi[ell1, wa, m1] = x
        base, tmp, acc = torch.empty_like(rmesh), torch.empty_like(rmesh), torch.empty_like(rmesh)
        out = torch.empty((len(ells), 2 * max(ells) + 1, bin.nbins, 2), dtype=torch.float64, device=device)
        for ill, ell in enumerate(ells):
            for m in range(-ell, ell + 1):
                _mul_ylm(rmesh, base, mattrs, xs, ell, m, False)
                first = True
                for (ell1, wa, m1), x in xi.items():
                    _mul_ylm(base, tmp, mattrs, xs, ell1, m1, False)
                    if wa:  # |x|^-wa (mesh2.py:1644)
                        _mul_radial_power(tmp, tmp, mattrs, xs, -wa)
                    B = _r2c(tmp, mattrs.meshsize)
                    _lib.call('mp_spectrum_product', stream, _lib.ptr(B), _spectrum_pitch(B), _lib.ptr(a0), a0_pitch, n, _lib.ptr(B),
                              _spectrum_pitch(B))
                    Qt = _c2r(B, mattrs.meshsize)
                    del B
                    _lib.call('mp_mul_acc', stream, _lib.ptr(x), _lib.ptr(Qt), _lib.ptr(acc), rmesh.numel(), int(not first))
                    first = False
                _lib.call('mp_bin_ylm_real', stream, _lib.ptr(acc), n, _lib.ptr(t['kx']), _lib.ptr(t['ky']), _lib.ptr(t['kz']),
                          _lib.ptr(t['thr']), nedges, int(ell), int(m), 0, _lib.ptr(out[ill, m + ell]), _lib.ptr(ws_bin), ws_bin.numel())
        host_out = out[..., 0].cpu().numpy()
        with np.errstate(invalid='ignore', divide='ignore'):
            for ill, ell in enumerate(ells):
                sums[ill] = 4. * np.pi * host_out[ill, :2 * ell + 1].sum(axis=0) / bin.nmodes
    poles = [Mesh2SpectrumPole(k=bin.xavg, k_edges=bin.edges, nmodes=bin.nmodes, num_raw=sums[ill], num_shotnoise=np.zeros(bin.nbins),
                               norm=np.full(bin.nbins, norm), volume=np.prod(mattrs.kfun) * bin.nmodes, ell=ell, attrs=attrs)
             for ill, ell in enumerate(ells)]
    return Mesh2SpectrumPoles(poles, attrs=attrs)


def _host_array_fn(name, ell, x, *args):
    x = np.ascontiguousarray(x, dtype='f8')
    out = np.empty_like(x)
    _lib.call(name, int(ell), *args, x.ctypes.data, x.size, out.ctypes.data)
    return out


def _window_smooth(mattrs, bin, edgesin, ellsin, vlos, Q, rnorm, device):
    """(n_ells n_bins, n_ellsin n_in) matrix of the smooth window (mesh2.py:1045-1076).  The
    selection (Q, or 1) times P_l(mu_s) P_lin(mu_s) binned in separation shells for every pair
    (l, lin) is a combination of the shell sums of Q P_L(mu_s), L = |l - lin| .. l + lin (the
    Legendre series of the product), which one binning pass of the selection provides; the
    Hankel sums over the few hundred shells are then formed on the host."""
    from numpy.polynomial import legendre as npleg
    ells = bin.ells
    if max(list(ells) + list(ellsin)) > 5:
        raise NotImplementedError("window flags=('smooth',) supports multipoles up to 5 on this path")
    sbin = BinMesh2CorrelationPoles(mattrs, edges=None, device=device)
    value = Q if Q is not None else torch.ones(tuple(int(n) for n in mattrs.meshsize), dtype=torch.float32, device=device)
    qscale = 1. / float(np.prod(mattrs.meshsize, dtype='f8')) if Q is not None else 1.  # Q = c2r(A1 conj(A2)) / N^3
    unit = lambda ell: np.eye(ell + 1)[ell]
    series = {(ell, ellin): npleg.legmul(unit(ell), unit(ellin)) for ellin in ellsin for ell in ells}
    Ls = sorted({L for c in series.values() for L, v in enumerate(c) if v != 0.})
    sums = {}
    for i in range(0, len(Ls), 6):
        chunk = Ls[i:i + 6]
        out = sbin._sums_dev(value, chunk, [False] * len(chunk), los=vlos).cpu().numpy() * qscale
        sums.update(zip(chunk, out))
    empty = sbin.nmodes == 0
    with np.errstate(invalid='ignore', divide='ignore'):
        savg = np.where(empty, 0., sbin.xavg)
        kout = np.where(bin.nmodes == 0, 0., bin.xavg)
    snm = sbin.nmodes.astype('f8')
    volume = float(np.prod(mattrs.boxsize, dtype='f8'))
    jn = {ell: _host_array_fn('mp_spherical_jn_array', ell, np.outer(kout, savg)) for ell in ells}
    blocks = []
    for ellin in ellsin:
        tmp = []
        for ill, ell in enumerate(ells):
            with np.errstate(invalid='ignore', divide='ignore'):
                Qs = sum(v * sums[L] for L, v in enumerate(series[ell, ellin]) if v != 0.) / snm
            if ell != 0:
                Qs[0] = 0.
            Qs = np.where(empty, 0., Qs)
            rows = []
            for e0, e1 in edgesin:
                xi = _host_array_fn('mp_tophat_shell_xi_array', ellin, savg, float(e0), float(e1))
                spec = (-1)**(ell // 2) * (jn[ell] @ (snm * xi * volume * Qs))
                rows.append((2 * ell + 1) * spec * rnorm[ill])
            tmp.append(np.array(rows))
        blocks.append(np.concatenate(tmp, axis=-1))
    return np.real(np.concatenate(blocks, axis=0).T)


def _window_matrix(wmat, bin, edgesin, ellsin, vlos, norm, los=None, was=None):
    """WindowMatrix container of mesh2.py:1463-1482."""
    ells = bin.ells
    nbins = bin.nbins
    attrs = dict(los=vlos.tolist() if los is None else los)
    zeros = np.zeros(nbins)
    observable = Mesh2SpectrumPoles([Mesh2SpectrumPole(k=bin.xavg, k_edges=bin.edges, nmodes=bin.nmodes, num_raw=zeros,
                                                       num_shotnoise=zeros, norm=np.full(nbins, norm[ill]), ell=ell, attrs=attrs)
                                     for ill, ell in enumerate(ells)], attrs=attrs)
    kin = edgesin.mean(axis=-1)
    was = was or [0] * len(ellsin)
    theory = Mesh2SpectrumPoles([Mesh2SpectrumPole(k=kin, k_edges=edgesin, nmodes=np.zeros(len(kin), dtype='i8'),
                                                   num_raw=np.zeros(len(kin)), ell=ell, attrs=dict(wa_order=wa))
                                 for ell, wa in zip(ellsin, was)])
    return WindowMatrix(observable=observable, theory=theory, value=wmat)


def compute_mesh2(*meshes, bin=None, los='z'):
    """Dispatch on the binning type (mesh2.py:225-250)."""
    if isinstance(bin, BinMesh2SpectrumPoles):
        return compute_mesh2_spectrum(*meshes, bin=bin, los=los)
    if isinstance(bin, BinMesh2CorrelationPoles):
        return compute_mesh2_correlation(*meshes, bin=bin, los=los)
    raise ValueError(f'bin must be BinMesh2SpectrumPoles or BinMesh2CorrelationPoles, not {type(bin)}')


def compute_box2_normalization(*inputs, bin=None):
    """(sum w)^2 / (N^3 Vcell) for a constant density (mesh2.py:505-513)."""
    norm = _deferred(compute_box_normalization(*format_meshes(*inputs, nmeshes=2)[0]))
    if bin is not None:
        return [norm] * len(bin.ells)
    return norm


def compute_fkp2_normalization(*fkps, bin=None, cellsize=10., split=None, fields=None, **kwargs):
    """FKP normalisation alpha * sum(D R) / cellvolume on a regridded CIC mesh (mesh2.py:516-574)."""
    if split is not None:
        raise NotImplementedError('split normalisation needs the reference random streams; use split=None')
    kwargs.update(cellsize=cellsize)
    fkps, fields = format_meshes(*fkps, nmeshes=2)
    if fields[1] == fields[0]:
        fkp = fkps[0]
        alpha = fkp.data.sum() / fkp.randoms.sum()
        norm = alpha * compute_normalization(fkp.data, fkp.randoms, **kwargs)
    else:
        alpha2 = fkps[1].data.sum() / fkps[1].randoms.sum()
        norm = alpha2 * compute_normalization(fkps[0].data, fkps[1].randoms, **kwargs)
        alpha2 = fkps[0].data.sum() / fkps[0].randoms.sum()
        norm = norm + alpha2 * compute_normalization(fkps[1].data, fkps[0].randoms, **kwargs)
        norm = norm / 2
    norm = _deferred(norm)
    if bin is not None:
        return [norm] * len(bin.ells)
    return norm


def compute_fkp2_shotnoise(*fkps, bin=None, fields=None):
    """Sum of products of FKP weights (mesh2.py:577-621); non-zero for ell = 0 only."""
    fkps, fields = format_meshes(*fkps, fields=fields, nmeshes=2)

    def get_particles(fkp):
        return fkp.particles if isinstance(fkp, FKPField) else fkp

    if fields[1] == fields[0]:
        p = [get_particles(fkp) for fkp in fkps]
        shotnoise = _sum_f64(p[0].weights, p[1].weights)
    else:
        shotnoise = torch.zeros((), dtype=torch.float64, device=_default_device())
    if bin is not None:
        ones = torch.ones(bin.nbins, dtype=torch.float64, device=shotnoise.device)
        return [Deferred(shotnoise * ones) if ell == 0 else Deferred(0. * ones) for ell in bin.ells]
    return Deferred(shotnoise)


def _deferred(value):
    """A float64 result left on the device until it is read (see types.Deferred)."""
    if isinstance(value, torch.Tensor):
        return Deferred(value.to(torch.float64))
    return Deferred(torch.tensor(float(value), dtype=torch.float64))
