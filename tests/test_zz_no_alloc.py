"""The entry point that allocates nothing (mp_bin_poles_partial_noalloc), through the C ABI:
its spectra must equal those of mp_bin_poles_partial (the same kernels with a caller-provided
workspace) to 1e-12 relative (the FP64 bin sums of an x-pass block are added into one LDS
histogram in an order that the wave scheduling decides, so the last bits may differ), on two
calls in a row, the second on a smaller mesh with fewer products, so that the shared scratch
array is reused with what the first call left in it.
"""
import ctypes

import numpy as np
import pytest
import torch

from test_gpu_parity import DEV, host, say, to_dev

pytestmark = pytest.mark.gpu


def _i32(v):
    return (ctypes.c_int32 * len(v))(*[int(x) for x in v])


def _case(lib, n, boxsize, ni, order, ells, rz, los, npart, seed):
    from jaxpower_amd import MeshAttrs, BinMesh2SpectrumPoles, _lib
    attrs = MeshAttrs(meshsize=n, boxsize=boxsize, boxcenter=0.)
    rng = np.random.default_rng(seed)
    pos = (rng.uniform(0., 1., (npart, 3)) - 0.5) * attrs.boxsize
    w = rng.uniform(0.5, 1.5, npart).astype('f4')
    pitch = lib.mp_spectrum_pitch(n[2])
    partial = torch.empty((ni, n[0], n[1], pitch), dtype=torch.complex64, device=DEV)
    pos_t, w_t = to_dev(pos), to_dev(w)
    stream = _lib.stream_handle()
    n3, box, center = _lib.i32x3(n), _lib.f64x3(attrs.boxsize), _lib.f64x3(attrs.boxcenter)
    ws = _lib.workspace(lib.mp_paint_partial_workspace_bytes(n3, npart, order, ni), DEV)
    _lib.check(lib.mp_paint_partial(stream, pos_t.data_ptr(), 1, w_t.data_ptr(), npart, n3, box, center, order, ni,
                                    partial.data_ptr(), pitch, ws.data_ptr(), ws.numel()))
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': float(np.min(attrs.kfun)) * 1.5}, ells=0)
    t, nedges = bin._tables, len(bin._edges1d)
    args = (stream, partial.data_ptr(), pitch, _i32(n), order, ni, 1, 1, t['kx'].data_ptr(), t['ky'].data_ptr(),
            t['kz'].data_ptr(), t['thr'].data_ptr(), nedges, (ctypes.c_double * 3)(*los), len(ells), _i32(ells), _i32(rz))
    ref = torch.full((len(ells), bin.nbins, 2), np.nan, dtype=torch.float64, device=DEV)
    wsb = _lib.workspace(lib.mp_bin_partial_workspace_bytes(n3, ni, len(ells), nedges), DEV)
    _lib.check(lib.mp_bin_poles_partial(*args, ref.data_ptr(), wsb.data_ptr(), wsb.numel()))
    got = torch.full_like(ref, np.nan)
    _lib.check(lib.mp_bin_poles_partial_noalloc(*args, got.data_ptr()))
    torch.cuda.synchronize()
    return host(got), host(ref), args


def _same(got, ref, rtol=1e-12):
    assert np.all(np.isfinite(ref)) and np.all(np.isfinite(got))
    assert np.allclose(got, ref, rtol=rtol, atol=rtol * np.max(np.abs(ref))), np.max(np.abs(got - ref))


def test_noalloc_matches_workspace_entry_point_twice():
    from jaxpower_amd import _lib
    lib = _lib.load()
    assert lib.mp_noalloc_scratch_bytes() == 16 << 20
    say('noalloc first call (128, 64, 64) tsc x3')
    got, ref, _ = _case(lib, (128, 64, 64), (1280., 640., 640.), 3, 3, [0, 2, 4], [0, 1, 1], (0., 0., 1.), 40000, 1)
    _same(got, ref)
    say('noalloc second call (32, 32, 32) cic x2')
    got, ref, args = _case(lib, (32, 32, 32), (320., 320., 320.), 2, 2, [0, 2], [0, 1], (0.6, 0., 0.8), 5000, 2)
    _same(got, ref)
    # errors are reported before any launch: too many bins for the LDS histogram, bad order
    out = torch.zeros(16, dtype=torch.float64, device=DEV)
    bad = list(args)
    bad[12] = 5000
    assert lib.mp_bin_poles_partial_noalloc(*bad, out.data_ptr()) == -2
    bad = list(args)
    bad[4] = 7
    assert lib.mp_bin_poles_partial_noalloc(*bad, out.data_ptr()) == -1
    torch.cuda.synchronize()
