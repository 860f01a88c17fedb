"""Power-spectrum multipoles of many particle fields that share one mesh geometry.

jax-power measures many mocks by calling the same jitted function per mock (one XLA
executable, launched per call; scripts/abacus_hf.py:357-373).  Here the batch runs in one of
two forms:

- merged (the default where it applies): the catalogues are stacked in one particle array and
  every phase of the spectrum runs once for the whole batch -- one row sort, one fused paint +
  z pass and one y pass over all the interlaced meshes, one x pass + FP64 binning with a grid
  row per mock, one launch for the weight sums of every mock -- so each launch fills the GPU
  whatever the size of one mock;
- per mock: the per-mock sequence (interlaced paint, z/y FFT passes, fused x pass + binning,
  normalization and shot-noise sums) for every field, optionally on several forked streams.

Either is captured once into a HIP graph through torch's graph API and replayed.  The fields'
position and weight tensors are the inputs (in the merged form, views of the stacked arrays):
refill them in place (``load``) to measure new mocks with the same particle counts.
"""
import numpy as np
import torch

from . import _lib
from . import resamplers
from .mesh import ParticleField
from .mesh2 import (BinMesh2SpectrumPoles, compute_mesh2_spectrum, compute_box2_normalization, compute_fkp2_shotnoise,
                    _format_los)
from .types import Deferred, Mesh2SpectrumPole, Mesh2SpectrumPoles

__all__ = ['Mesh2SpectrumBatch']


def _detach_copy(value):
    if isinstance(value, torch.Tensor):
        return value.clone()
    if isinstance(value, Deferred):
        return Deferred(value._tensor.clone()) if value._tensor is not None else value
    if isinstance(value, (list, tuple)):
        return type(value)(_detach_copy(v) for v in value)
    return value


class _HostOnce:
    """A device tensor read back whole on the first access to any part of it."""

    def __init__(self, tensor):
        self._tensor, self._host = tensor, None

    def host(self):
        if self._host is None:
            self._host = self._tensor.detach().cpu().numpy()
            self._tensor = None
        return self._host


class _Part(Deferred):
    """Deferred value holder[index] of a batch-wide result: one transfer serves every mock."""

    def __init__(self, holder, index):
        super().__init__(None)
        self._holder, self._index = holder, index

    def numpy(self):
        if self._host is None:
            self._host = self._holder.host()[self._index]
        return self._host


def _copy_poles(spec):
    poles = []
    for p in spec:
        state = {name: _detach_copy(v) for name, v in p._raw.items()}
        poles.append(type(p)(**state, ell=p.ell, attrs=p.attrs))
    return Mesh2SpectrumPoles(poles, attrs=spec.attrs)


class Mesh2SpectrumBatch:
    """Spectra of ``fields`` (ParticleFields on the same MeshAttrs): for each, paint
    (resampler, interlacing, compensate, out='complex'), remove the k = 0 mode (remove_dc),
    ``compute_mesh2_spectrum(mesh, bin=bin, los=los)``, then box normalization and shot noise.

    ``graph=True`` captures the whole batch into one HIP graph on the first call and replays it
    afterwards; ``graph=False`` runs the same sequence eagerly (the per-mock reference).
    ``streams > 1`` issues the mocks round-robin on that many streams forked from the caller's
    stream (and joined back), so the kernels of independent mocks overlap on the GPU; in the
    captured graph they become parallel branches."""

    def __init__(self, fields, bin, resampler='cic', interlacing=0, compensate=False, los='z', remove_dc=True, graph=True,
                 streams=1, merged=None):
        self.fields = list(fields)
        if not self.fields or not all(isinstance(f, ParticleField) for f in self.fields):
            raise ValueError('fields must be a non-empty list of ParticleField')
        attrs = self.fields[0].attrs
        if any(f.attrs != attrs for f in self.fields):
            raise ValueError('all fields of a batch must share one MeshAttrs')
        self.bin, self.resampler, self.interlacing, self.compensate = bin, resampler, interlacing, compensate
        self.los, self.remove_dc, self.graph = los, remove_dc, graph
        self.streams = max(1, int(streams))
        self._graph, self._outputs, self._side = None, None, None
        reason = self._merge_blocker()
        if merged and reason:
            raise NotImplementedError(f'merged batch: {reason}')
        self.merged = merged is not False and reason is None
        self._auto = merged is None  # automatic: fall back to the per-mock form where the kernels refuse the case
        if self.merged:
            self._stack()

    # ------------------------------------------------------------------ merged form

    def _merge_blocker(self):
        """Why the merged form does not apply (None when it does)."""
        if not isinstance(self.bin, BinMesh2SpectrumPoles):
            return 'bin must be BinMesh2SpectrumPoles'
        if len(getattr(self.bin, '_shifts', [None])) > 1:
            return 'mode oversampling'
        if not resamplers.is_builtin(resamplers.get_resampler(self.resampler)) or not isinstance(self.compensate, bool):
            return 'custom resampler or compensation'
        if _format_los(self.los)[1] is None:
            return 'local line of sight'
        nprod, nedges = len(self.bin.ells), len(self.bin._edges1d)
        if nprod > 6 or nprod * (nedges - 1) > 4096:
            return 'too many multipoles or bins for the fused binning'
        f0 = self.fields[0]
        if any(f.positions.dtype != f0.positions.dtype or f.positions.device != f0.positions.device for f in self.fields):
            return 'positions of different dtypes or devices'
        if f0.positions.device.type != 'cuda':
            return 'positions not on a GPU'
        if sum(f.size for f in self.fields) >= 2**31 - 1:
            return 'more than 2^31 particles'
        n = [int(v) for v in f0.attrs.meshsize]
        if any(v & (v - 1) or v < 16 or v > 2048 for v in n):
            return f'meshsize {tuple(n)}: the fused paint and x pass take powers of two in 16..2048'
        return None

    def _stack(self):
        """Stack positions and weights into one array each; the fields keep views of them."""
        fields = self.fields
        counts = [int(f.size) for f in fields]
        offsets = np.concatenate([[0], np.cumsum(counts)]).astype('i8')
        pos = torch.cat([f.positions for f in fields]).contiguous()
        w = torch.cat([f.weights.to(torch.float32) for f in fields]).contiguous()
        for f, a, b in zip(fields, offsets[:-1], offsets[1:]):
            f.__dict__.update(positions=pos[a:b], weights=w[a:b])
        self._pos, self._w, self._offsets = pos, w, offsets
        self._moff = torch.from_numpy(offsets).to(pos.device)
        attrs, bin = fields[0].attrs, self.bin
        nmock, ni = len(fields), max(int(self.interlacing), 1)
        n = _lib.i32x3(attrs.meshsize)
        self._pitch = int(_lib.load().mp_spectrum_pitch(int(attrs.meshsize[2])))
        self._partial = torch.empty((nmock, ni, int(attrs.meshsize[0]), int(attrs.meshsize[1]), self._pitch), dtype=torch.complex64,
                                    device=pos.device)
        lib = _lib.load()
        order = int(resamplers.get_resampler(self.resampler).order)
        nprod, nedges = len(bin.ells), len(bin._edges1d)
        self._ws_paint = _lib.workspace(lib.mp_paint_partial_batch_workspace_bytes(n, len(pos), nmock, order, ni), pos.device)
        self._ws_bin = _lib.workspace(lib.mp_bin_partial_batch_workspace_bytes(n, ni, nprod, nedges, nmock), pos.device)
        self._ws_sums = _lib.workspace(lib.mp_catalogue_sums_workspace_bytes(nmock), pos.device)
        self._sums = torch.empty((nmock, nprod, bin.nbins, 2), dtype=torch.float64, device=pos.device)
        self._wsums = torch.empty((nmock, 2), dtype=torch.float64, device=pos.device)
        self._order, self._ni = order, ni
        # compute_mesh2_spectrum's scaling (2 ell + 1) / nmodes, and the part kept: real for even
        # ell, imaginary for odd (made here: no host-to-device copy may run under graph capture)
        ells = list(bin.ells)
        scale = torch.tensor([2. * ell + 1. if ell else 1. for ell in ells], dtype=torch.float64, device=pos.device)
        scale = scale[:, None] / bin._nmodes_dev[None, :]
        # odd ell of an auto spectrum: the imaginary part of a real sum, 0 in every bin (empty
        # ones included, as the reference's .imag of |A|^2 binned, mesh2.py:350-351, 368)
        scale[[i for i, ell in enumerate(ells) if ell % 2]] = 0.
        self._scale = scale[None, :, :, None]
        odd = torch.tensor([ell % 2 for ell in ells], dtype=torch.long, device=pos.device)
        self._part = odd[None, :, None, None].expand(nmock, nprod, bin.nbins, 1).contiguous()

    def _run_merged(self):
        """One launch per phase for the whole batch; returns the batch-wide device results."""
        fields, bin = self.fields, self.bin
        attrs = fields[0].attrs
        device = self._pos.device
        stream = _lib.stream_handle(device)
        n = _lib.i32x3(attrs.meshsize)
        nmock = len(fields)
        _lib.call('mp_paint_partial_batch', stream, _lib.ptr(self._pos), int(self._pos.dtype == torch.float64), _lib.ptr(self._w),
                  len(self._pos), _lib.ptr(self._moff), nmock, n, _lib.f64x3(attrs.boxsize), _lib.f64x3(attrs.boxcenter),
                  self._order, self._ni, _lib.ptr(self._partial), self._pitch, _lib.ptr(self._ws_paint), self._ws_paint.numel())
        ells = list(bin.ells)
        nprod = len(ells)
        col = lambda values: (_lib.ctypes.c_int32 * nprod)(*[int(v) for v in values])
        t = bin._tables
        vlos = _format_los(self.los)[1]
        _lib.call('mp_bin_poles_partial_batch', stream, _lib.ptr(self._partial), self._pitch, nmock, n, self._order, self._ni,
                  int(self.compensate), int(self.remove_dc), _lib.ptr(t['kx']), _lib.ptr(t['ky']), _lib.ptr(t['kz']),
                  _lib.ptr(t['thr']), len(bin._edges1d), _lib.f64x3(vlos), nprod, col(ells), col([ell != 0 for ell in ells]),
                  _lib.ptr(self._sums), _lib.ptr(self._ws_bin), self._ws_bin.numel())
        _lib.call('mp_catalogue_sums', stream, _lib.ptr(self._w), _lib.ptr(self._moff), nmock, _lib.ptr(self._wsums),
                  _lib.ptr(self._ws_sums), self._ws_sums.numel())
        # the spectrum's scaling and part; box normalization (sum w)^2 / (N^3 Vcell); shot noise
        # sum w^2 at ell = 0
        num = torch.gather(self._sums * self._scale, 3, self._part)[..., 0]
        size = float(np.prod(attrs.meshsize))
        norm = self._wsums[:, 0]**2 / (size * float(np.prod(attrs.cellsize)))
        return num, norm, self._wsums[:, 1].clone()  # results outlive the next call's buffers

    def _assemble(self, num, norm, shotnoise):
        """Per-mock Mesh2SpectrumPoles over the batch-wide results (one host transfer each)."""
        bin, ells = self.bin, list(self.bin.ells)
        attrs = self.fields[0].attrs
        vlos = _format_los(self.los)[1]
        pattrs = dict(los=vlos.tolist())
        hnum, hnorm, hsn = _HostOnce(num), _HostOnce(norm), _HostOnce(shotnoise)
        volume = np.prod(attrs.kfun) * bin.nmodes
        zero = np.zeros(bin.nbins)
        out = []
        for m in range(len(self.fields)):
            poles = [Mesh2SpectrumPole(k=bin.xavg, k_edges=bin.edges, nmodes=bin.nmodes, num_raw=_Part(hnum, (m, i)),
                                       num_shotnoise=_Part(hsn, m) if ell == 0 else zero, norm=_Part(hnorm, m), volume=volume,
                                       ell=ell, attrs=pattrs)
                     for i, ell in enumerate(ells)]
            out.append(Mesh2SpectrumPoles(poles, attrs=pattrs))
        return out

    # ------------------------------------------------------------------ per-mock form

    def _one(self, p):
        mesh = p.paint(resampler=self.resampler, interlacing=self.interlacing, compensate=self.compensate, out='complex')
        if self.remove_dc:
            mesh = mesh.at[0, 0, 0].set(0.)
        spec = compute_mesh2_spectrum(mesh, bin=self.bin, los=self.los)
        return spec.clone(norm=compute_box2_normalization(p, bin=self.bin), num_shotnoise=compute_fkp2_shotnoise(p, bin=self.bin))

    def _run(self):
        if self.merged:
            try:
                return self._run_merged()
            except NotImplementedError:
                if not self._auto:
                    raise
                self.merged = False  # e.g. no fused paint for this geometry and resampler
        if self.streams == 1:
            return [self._one(p) for p in self.fields]
        main = torch.cuda.current_stream()
        if self._side is None:
            self._side = [torch.cuda.Stream() for _ in range(self.streams)]  # default priority
        out = [None] * len(self.fields)
        for side in self._side:
            side.wait_stream(main)
        for k, side in enumerate(self._side):
            with torch.cuda.stream(side):
                for i in range(k, len(self.fields), self.streams):
                    out[i] = self._one(self.fields[i])
        for side in self._side:
            main.wait_stream(side)
        return out

    def capture(self):
        """Warm up (tuning, twiddles, allocator) on a side stream, then capture the batch."""
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            self._run()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self._graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self._graph):
            self._outputs = self._run()
        torch.cuda.synchronize()

    def __call__(self):
        """Spectra of the current contents of the fields (one list entry per field)."""
        if not self.graph:
            out = self._run()
            return self._assemble(*out) if self.merged else out
        if self._graph is None:
            self.capture()
        self._graph.replay()
        if self.merged:
            return self._assemble(*[t.clone() for t in self._outputs])
        return [_copy_poles(spec) for spec in self._outputs]

    def load(self, index, positions, weights=None):
        """Copy new positions (and weights) of field ``index`` into the graph's input buffers;
        the particle count must not change."""
        f = self.fields[index]
        positions = torch.as_tensor(positions, device=f.positions.device)
        if positions.shape != f.positions.shape:
            raise ValueError(f'field {index}: positions of shape {tuple(positions.shape)} != {tuple(f.positions.shape)}')
        f.positions.copy_(positions)
        if weights is not None:
            f.weights.copy_(torch.broadcast_to(torch.as_tensor(weights, device=f.weights.device), f.weights.shape))
