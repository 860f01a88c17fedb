"""Differentiable power-spectrum multipoles of a particle field, with a hand-written backward.

The reference differentiates ``compute_mesh2_spectrum(ParticleField(x, w).paint(...))`` with
``jax.grad`` through the scatter paint, the interlacing scan and the binning
(jaxpower/resamplers.py:73-103, jaxpower/mesh.py:2673-2754, jaxpower/mesh2.py:346-374).  Here
the same map, restricted to a global line of sight and an auto spectrum, is a
``torch.autograd.Function``: forward = interlaced paint -> fused combine + binning; backward =
binning transposed (mp_bin_adjoint) -> inverse FFT per shift -> adjoint readout with the
derivatives of the mass-assignment kernel (mp_read_grad).  The backward recomputes the painted
spectrum instead of keeping it, so a step holds one complex mesh at a time.

Derivatives of round/floor are zero, as under jax.grad.  Normalization and shot noise are
functions of the weights alone; ``box_spectrum_values`` composes them with torch operations on
the weights, so autograd handles them.
"""
import numpy as np
import torch

from . import _lib, resamplers
from .mesh import MeshAttrs, _paint_partial, _spectrum_pitch
from .mesh2 import _format_los, _single_shift

__all__ = ['mesh2_spectrum_poles', 'box_spectrum_values']


def _ell_scales(bin, device):
    nm = bin._nmodes_dev
    return [torch.where(nm > 0, (2 * ell + 1) / nm, torch.zeros_like(nm)) for ell in bin.ells]


class _SpectrumPoles(torch.autograd.Function):

    @staticmethod
    def forward(ctx, positions, weights, attrs, bin, order, ni, compensate, vlos, zero_dc):
        w = weights.detach().contiguous().to(torch.float32)
        pos = positions.detach().contiguous()
        pending = _paint_partial(attrs, pos, w, order, ni, compensate)
        if zero_dc:
            pending = pending.with_zero_dc()
        ells = bin.ells
        rz = [ell != 0 for ell in ells]
        sums = bin._pending_products(pending, ells, rz, vlos)
        if sums is None:
            sums = bin._products_dev([pending.materialize(attrs)], [(0, 0, ell, r) for ell, r in zip(ells, rz)], los=vlos)
        del pending
        # odd ell of this auto spectrum: the imaginary part of a real sum, 0 in every bin
        # (empty ones included), as the reference's .imag of |A|^2 binned (mesh2.py:350-351, 368)
        num = torch.stack([(sums[i] * s).real if ell % 2 == 0 else torch.zeros_like((sums[i] * s).real)
                           for i, (ell, s) in enumerate(zip(ells, _ell_scales(bin, pos.device)))])
        ctx.save_for_backward(pos, w)
        ctx.params = (attrs, bin, order, ni, compensate, vlos, zero_dc, positions.dtype)
        return num

    @staticmethod
    def backward(ctx, grad_num):
        pos, w = ctx.saved_tensors
        attrs, bin, order, ni, compensate, vlos, zero_dc, pos_dtype = ctx.params
        device = pos.device
        lib = _lib.load()
        n = _lib.i32x3(attrs.meshsize)
        # recompute the combined spectrum A
        pending = _paint_partial(attrs, pos, w, order, ni, compensate)
        if zero_dc:
            pending = pending.with_zero_dc()
        A = pending.materialize(attrs)
        del pending
        ells = bin.ells
        # num[p] = (2 ell + 1)/nmodes sum_b(...): chain rule onto the raw sums; odd ells of an
        # auto spectrum are identically zero
        coef = torch.stack([grad_num[i].to(torch.float64) * s if ell % 2 == 0 else torch.zeros_like(s)
                            for i, (ell, s) in enumerate(zip(ells, _ell_scales(bin, device)))]).contiguous()
        pitch = _spectrum_pitch(A)
        Y = torch.empty((ni, int(attrs.meshsize[0]), int(attrs.meshsize[1]), pitch), dtype=torch.complex64, device=device)
        ws = _lib.workspace(lib.mp_bin_adjoint_workspace_bytes(n, ni), device)
        nprod = len(ells)
        col = lambda values: (_lib.ctypes.c_int32 * nprod)(*[int(v) for v in values])
        t = bin._tables
        stream = _lib.stream_handle(device)
        _lib.call('mp_bin_adjoint', stream, _lib.ptr(A), pitch, n, order, ni, int(compensate), int(zero_dc), _lib.ptr(t['kx']),
                  _lib.ptr(t['ky']), _lib.ptr(t['kz']), _lib.ptr(t['thr']), len(bin._edges1d), _lib.f64x3(vlos), nprod,
                  col(ells), col([ell != 0 for ell in ells]), _lib.ptr(coef), _lib.ptr(Y), _lib.ptr(ws), ws.numel())
        del A
        D = torch.empty((ni,) + tuple(int(s) for s in attrs.meshsize), dtype=torch.float32, device=device)
        ws = _lib.workspace(lib.mp_irfft3d_workspace_bytes(n), device)
        for j in range(ni):
            _lib.call('mp_irfft3d', stream, _lib.ptr(Y[j]), pitch, _lib.ptr(D[j]), n, _lib.ptr(ws), ws.numel())
        del Y, ws
        np_ = pos.shape[0]
        gw = torch.empty(np_, dtype=torch.float32, device=device)
        gpos = torch.empty((np_, 3), dtype=torch.float32, device=device)
        _lib.call('mp_read_grad', stream, _lib.ptr(D), ni, n, _lib.f64x3(attrs.boxsize), _lib.f64x3(attrs.boxcenter),
                  _lib.ptr(pos), int(pos.dtype == torch.float64), _lib.ptr(w), np_, order, _lib.ptr(gw), _lib.ptr(gpos))
        return gpos.to(pos_dtype), gw, None, None, None, None, None, None, None


def mesh2_spectrum_poles(positions, weights=None, attrs=None, bin=None, resampler='tsc', interlacing=0, compensate=False,
                         los='z', remove_dc=True):
    """num_raw of compute_mesh2_spectrum(ParticleField(positions, weights).paint(resampler,
    interlacing, compensate, out='complex') with the k = 0 mode set to 0 (remove_dc), bin, los),
    as a (len(bin.ells), nbins) float64 tensor differentiable with respect to positions and
    weights.  Global line of sight only (a string axis or a vector)."""
    if not isinstance(attrs, MeshAttrs):
        raise ValueError('attrs must be a MeshAttrs')
    _single_shift(bin, 'mesh2_spectrum_poles')
    los, vlos, swap = _format_los(los)
    if vlos is None:
        raise NotImplementedError('the differentiable path takes a global line of sight')
    _lib.require_device(positions, 'positions')
    order = int(resamplers.get_resampler(resampler).order)
    ni = max(int(interlacing), 1)
    if weights is None:
        weights = torch.ones(positions.shape[0], dtype=torch.float32, device=positions.device)
    return _SpectrumPoles.apply(positions, weights, attrs, bin, order, ni, bool(compensate), tuple(float(v) for v in vlos),
                                bool(remove_dc))


def box_spectrum_values(num, weights, attrs, ells, shotnoise=True):
    """The multipoles P_ell = (num_raw - SN [ell = 0]) / norm of a periodic box, as the box
    harness of the reference forms them (scripts/abacus_hf.py:357-373): norm the box
    normalization (sum w)^2 / (N^3 V_cell) (compute_box2_normalization, mesh.py:3121-3129) and
    SN = sum w^2 (compute_fkp2_shotnoise, mesh2.py:577-621), here torch functions of the weights,
    so that gradients flow through them.  num: the (len(ells), nbins) output of
    mesh2_spectrum_poles."""
    w = weights.to(torch.float64)
    ncell = float(np.prod(attrs.meshsize))
    vcell = float(np.prod(attrs.cellsize))
    norm = w.sum()**2 / (ncell * vcell)
    sn = (w * w).sum() if shotnoise else None
    return torch.stack([((num[i] - sn) if (ell == 0 and shotnoise) else num[i]) / norm for i, ell in enumerate(ells)])
