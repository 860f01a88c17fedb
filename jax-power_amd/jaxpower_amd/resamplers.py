"""Resampler namespaces ``ngp``, ``cic``, ``tsc``, ``pcs`` with ``.order``, ``.paint``,
``.read``, ``.compensate`` and ``.aliasing_shotnoise`` (jaxpower/resamplers.py:1-144).

``paint`` and ``read`` take grid-unit positions on a 3-D mesh and run the HIP kernels;
``compensate`` and ``aliasing_shotnoise`` are Fourier kernels ``fn(value, kvec)`` tagged
``kind = 'circular'`` for ``ComplexMeshField.apply``.
"""
from functools import partial

import numpy as np
import torch

from . import _lib


def _compensate_tophat_convolution_kernel(order: int):
    """value * prod_i sinc(k_i / 2pi)^-order (resamplers.py:16-26)."""

    def fn(value, kvec):
        kernel = 1.
        for kk in kvec:
            kernel = kernel * torch.sinc(kk / (2 * np.pi))
        return value * kernel**(-order)

    fn.kind = 'circular'
    fn.order = order
    fn.is_compensation = True
    return fn


def _aliasing_shotnoise_kernel(order: int):
    """Aliased shot-noise window (resamplers.py:29-55)."""

    def k1(angle):
        return torch.ones_like(angle)

    def k2(angle):
        return 1. - 2. / 3. * torch.sin(angle / 2.)**2

    def k3(angle):
        s2 = torch.sin(angle / 2.)**2
        return 1. - s2 + 2. / 15. * s2**2

    def k4(angle):
        s2 = torch.sin(angle / 2.)**2
        return 1. - 4. / 3. * s2 + 2. / 5. * s2 * s2 - 4. / 315. * s2**3

    kernels = [None, k1, k2, k3, k4]

    def fn(value, kvec):
        kernel = 1.
        for kk in kvec:
            kernel = kernel * kernels[order](kk)
        return value * kernel

    fn.kind = 'circular'
    return fn


def paint(mesh, positions, weights=None, order: int = 2):
    """Paint grid-unit ``positions`` onto ``mesh`` (a 3-D float32 GPU tensor, added to) or onto
    a zero mesh of shape ``mesh`` if a tuple (resamplers.py:73-103)."""
    shape = tuple(mesh) if isinstance(mesh, tuple) else tuple(mesh.shape)
    if len(shape) != 3:
        raise NotImplementedError('only 3-D meshes are painted on this path')
    _lib.require_device(positions, 'positions')
    device = positions.device
    out = torch.empty(shape, dtype=torch.float32, device=device)
    pos = positions.contiguous()
    w = None
    if weights is not None:
        w = torch.as_tensor(weights, dtype=torch.float32, device=device)
        w = torch.broadcast_to(w, (pos.shape[0],)).contiguous()
    n = _lib.i32x3(shape)
    boxsize = _lib.f64x3(shape)
    boxcenter = _lib.f64x3([s / 2. for s in shape])
    nbytes = _lib.load().mp_paint_workspace_bytes(n, pos.shape[0], order, 1)
    ws = _lib.workspace(nbytes, device)
    _lib.call('mp_paint', _lib.stream_handle(device), _lib.ptr(pos), int(pos.dtype == torch.float64), _lib.ptr(w),
              pos.shape[0], n, boxsize, boxcenter, order, 1, _lib.ptr(out), _lib.ptr(ws), ws.numel())
    if not isinstance(mesh, tuple):
        out = out + mesh
    return out


def read(mesh, positions, order: int = 2, out=None):
    """Read ``mesh`` at grid-unit ``positions`` (resamplers.py:106-131)."""
    _lib.require_device(mesh, 'mesh')
    _lib.require_device(positions, 'positions')
    shape = tuple(mesh.shape)
    if len(shape) != 3:
        raise NotImplementedError('only 3-D meshes are read on this path')
    pos = positions.contiguous()
    res = torch.empty(pos.shape[0], dtype=torch.float32, device=mesh.device)
    _lib.call('mp_read', _lib.stream_handle(mesh.device), _lib.ptr(mesh.contiguous().float()), _lib.i32x3(shape),
              _lib.f64x3(shape), _lib.f64x3([s / 2. for s in shape]), _lib.ptr(pos), int(pos.dtype == torch.float64),
              pos.shape[0], order, _lib.ptr(res))
    if out is not None:
        res = res + out
    return res


for _i, _name in enumerate(['ngp', 'cic', 'tsc', 'pcs']):
    _order = _i + 1
    globals()[_name] = type(_name, (), dict(paint=staticmethod(partial(paint, order=_order)),
                                            read=staticmethod(partial(read, order=_order)),
                                            compensate=staticmethod(_compensate_tophat_convolution_kernel(_order)),
                                            aliasing_shotnoise=staticmethod(_aliasing_shotnoise_kernel(_order)),
                                            order=_order))


def is_builtin(resampler):
    """True for the four namespaces above, whose paint, read and compensation run in the fused
    HIP kernels; any other object's own ``.paint`` / ``.read`` / ``.compensate`` are called."""
    return any(resampler is globals()[name] for name in ('ngp', 'cic', 'tsc', 'pcs'))


def get_resampler(resampler):
    """Name -> namespace; objects with ``.order`` pass through (resamplers.py:134-144)."""
    if isinstance(resampler, str):
        try:
            return globals()[resampler]
        except KeyError:
            raise ValueError(f'unknown resampler {resampler}') from None
    return resampler
