"""The reference's plugin points on the GPU path: a resampler object other than the built-in
ngp/cic/tsc/pcs (resamplers.get_resampler, jaxpower/resamplers.py:134-144) has its own
.paint / .read / .compensate called by _paint and RealMeshField.read (jaxpower/mesh.py:
2673-2754, 2030-2115), never replaced by the built-in kernel of the same order.

Tolerance: meshes and read values to 1e-5 of their largest value (float32 scatter order).
"""
from itertools import product

import numpy as np
import pytest
import torch

from test_gpu_parity import host, say, to_dev

pytestmark = pytest.mark.gpu


def _torch_cic(scale=1.):
    """A user resampler written with torch scatter/gather (CIC, optionally scaled)."""
    from jaxpower_amd import resamplers

    class TorchCIC:
        order = 2
        compensate = staticmethod(resamplers.cic.compensate)
        calls = {'paint': 0, 'read': 0}

        @staticmethod
        def paint(mesh, positions, weights=None):
            TorchCIC.calls['paint'] += 1
            n = torch.tensor(mesh.shape, device=positions.device)
            i0 = torch.floor(positions)
            t = (positions - i0).to(torch.float32)
            i0 = i0.long()
            w = torch.ones(positions.shape[0], device=positions.device) if weights is None else weights.to(torch.float32)
            for off in product((0, 1), repeat=3):
                o = torch.tensor(off, device=positions.device)
                idx = (i0 + o) % n
                k = torch.prod(torch.where(o.bool(), t, 1. - t), dim=1)
                mesh.index_put_((idx[:, 0], idx[:, 1], idx[:, 2]), scale * w * k, accumulate=True)
            return mesh

        @staticmethod
        def read(mesh, positions):
            TorchCIC.calls['read'] += 1
            n = torch.tensor(mesh.shape, device=positions.device)
            i0 = torch.floor(positions)
            t = (positions - i0).to(torch.float32)
            i0 = i0.long()
            out = torch.zeros(positions.shape[0], device=positions.device)
            for off in product((0, 1), repeat=3):
                o = torch.tensor(off, device=positions.device)
                idx = (i0 + o) % n
                out += scale * mesh[idx[:, 0], idx[:, 1], idx[:, 2]] * torch.prod(torch.where(o.bool(), t, 1. - t), dim=1)
            return out

    return TorchCIC


@pytest.mark.parametrize('interlacing,compensate,out', [(0, False, 'real'), (2, True, 'complex'), (3, True, 'real'),
                                                         (0, True, 'complex')])
def test_custom_resampler_paint_is_called(interlacing, compensate, out):
    from jaxpower_amd import MeshAttrs, ParticleField
    say('custom resampler paint', interlacing, compensate, out)
    rng = np.random.default_rng(3)
    attrs = MeshAttrs(meshsize=(32, 16, 24), boxsize=(320., 160., 240.), boxcenter=(5., -5., 10.))
    pos = (rng.uniform(0., 1., (8000, 3)) - 0.5) * attrs.boxsize + attrs.boxcenter
    w = rng.uniform(0.5, 1.5, len(pos)).astype('f4')
    p = ParticleField(to_dev(pos), to_dev(w), attrs=attrs)
    want = host(p.paint(resampler='cic', interlacing=interlacing, compensate=compensate, out=out).value)
    mine = _torch_cic()
    got = host(p.paint(resampler=mine, interlacing=interlacing, compensate=compensate, out=out).value)
    assert mine.calls['paint'] == max(interlacing, 1)
    assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()
    doubled = host(p.paint(resampler=_torch_cic(2.), interlacing=interlacing, compensate=compensate, out=out).value)
    assert np.abs(doubled - 2. * want).max() <= 2e-5 * np.abs(want).max()


def test_custom_resampler_read_is_called():
    from jaxpower_amd import MeshAttrs, RealMeshField
    say('custom resampler read')
    rng = np.random.default_rng(4)
    attrs = MeshAttrs(meshsize=16, boxsize=160., boxcenter=0.)
    mesh = RealMeshField(to_dev(rng.normal(size=(16, 16, 16)).astype('f4')), attrs=attrs)
    pos = to_dev(rng.uniform(-80., 80., (2000, 3)))
    mine = _torch_cic()
    for compensate in (False, True):
        want = host(mesh.read(pos, resampler='cic', compensate=compensate))
        got = host(mesh.read(pos, resampler=mine, compensate=compensate))
        assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()
    assert mine.calls['read'] == 2
