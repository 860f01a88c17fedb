"""Power-spectrum multipoles of many particle fields that share one mesh geometry, replayed as
one captured HIP graph.

jax-power measures many mocks by calling the same jitted function per mock (one XLA
executable, launched per call).  Here the whole per-mock sequence for every field of the
batch (interlaced paint, z/y FFT passes, fused x pass + binning, normalization and shot-noise
sums) is captured once into a HIP graph through torch's graph API and replayed: one launch
from the host per batch, no per-kernel launch overhead, and no host work between mocks.  The
fields' position and weight tensors are the graph's inputs: refill them in place
(``load``) to measure new mocks with the same particle counts.
"""
import torch

from .mesh import ParticleField
from .mesh2 import compute_mesh2_spectrum, compute_box2_normalization, compute_fkp2_shotnoise
from .types import Deferred, Mesh2SpectrumPoles

__all__ = ['Mesh2SpectrumBatch']


def _detach_copy(value):
    if isinstance(value, torch.Tensor):
        return value.clone()
    if isinstance(value, Deferred):
        return Deferred(value._tensor.clone()) if value._tensor is not None else value
    if isinstance(value, (list, tuple)):
        return type(value)(_detach_copy(v) for v in value)
    return value


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
                 streams=1):
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

    def _one(self, p):
        mesh = p.paint(resampler=self.resampler, interlacing=self.interlacing, compensate=self.compensate, out='complex')
        if self.remove_dc:
            mesh = mesh.at[0, 0, 0].set(0.)
        spec = compute_mesh2_spectrum(mesh, bin=self.bin, los=self.los)
        return spec.clone(norm=compute_box2_normalization(p, bin=self.bin), num_shotnoise=compute_fkp2_shotnoise(p, bin=self.bin))

    def _run(self):
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
            return self._run()
        if self._graph is None:
            self.capture()
        self._graph.replay()
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
