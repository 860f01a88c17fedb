"""Write the small golden fixtures under tests/golden/ from the float64 oracle.

The reference (pure Python on JAX) cannot be imported in this environment and ships no
fixtures of its own, so the fixtures are produced by the oracle, whose pinning checks live in
tests/test_oracle.py.  Each fixture stores its inputs (positions, weights) and the expected
outputs, so the GPU tests compare the HIP path with fixed numbers.

Usage: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import power_oracle as po  # noqa: E402


def box_case(name, seed, meshsize, boxsize, boxcenter, size, resampler, interlacing, los, ells, step, f32):
    rng = np.random.default_rng(seed)
    attrs = po.Attrs(meshsize=meshsize, boxsize=boxsize, boxcenter=boxcenter)
    pos = rng.uniform(0., 1., (size, 3)) * attrs.boxsize + attrs.boxcenter - attrs.boxsize / 2.
    # a few clumps so the multipoles are not pure noise
    nclump = size // 4
    centers = rng.uniform(0., 1., (8, 3)) * attrs.boxsize + attrs.boxcenter - attrs.boxsize / 2.
    pos[:nclump] = centers[rng.integers(0, 8, nclump)] + rng.normal(0., 0.03 * attrs.boxsize[0], (nclump, 3))
    weights = rng.uniform(0.5, 1.5, size)
    if f32:
        pos = pos.astype('f4').astype('f8')
        weights = weights.astype('f4').astype('f8')
    mesh = po.paint(attrs, pos, weights, resampler=resampler, interlacing=interlacing, compensate=True, out='complex')
    mesh[0, 0, 0] = 0.  # the harness subtracts the mean (scripts/abacus_hf.py:357-373)
    bin = po.BinPoles(attrs, edges={'step': step}, ells=ells)
    res = po.spectrum_poles(attrs, bin, mesh, los=los)
    norm = po.box_normalization(attrs, weights)
    sn = po.fkp_shotnoise(weights)
    out = dict(positions=pos.astype('f4' if f32 else 'f8'), weights=weights.astype('f4'), meshsize=attrs.meshsize,
               boxsize=attrs.boxsize, boxcenter=attrs.boxcenter, resampler=resampler, interlacing=interlacing,
               los=los, ells=np.array(ells), step=step, k=bin.xavg, nmodes=bin.nmodes, edges=bin.edges,
               num_raw=np.array(res['num_raw']), norm=norm, shotnoise=sn)
    np.savez_compressed(os.path.join(HERE, f'{name}.npz'), **out)


def fkp_case(name, seed, meshsize, size_data, size_randoms, resampler, interlacing, los, ells, step):
    rng = np.random.default_rng(seed)

    def shell(n):
        # uniform in volume for r in [100, 200], within 60 degrees of the z axis
        u = rng.uniform(0., 1., n)
        r = (100.**3 + u * (200.**3 - 100.**3))**(1. / 3.)
        cost = rng.uniform(np.cos(np.pi / 3.), 1., n)
        phi = rng.uniform(0., 2. * np.pi, n)
        sint = np.sqrt(1. - cost**2)
        return np.column_stack([r * sint * np.cos(phi), r * sint * np.sin(phi), r * cost])

    data, randoms = shell(size_data), shell(size_randoms)
    wd = (1. / (1. + rng.uniform(0.5, 1.5, size_data))).astype('f4').astype('f8')
    wr = (1. / (1. + rng.uniform(0.5, 1.5, size_randoms))).astype('f4').astype('f8')
    attrs = po.get_mesh_attrs(data, randoms, meshsize=meshsize, boxpad=1.5)
    alpha = wd.sum() / wr.sum()
    pos = np.concatenate([data, randoms])
    w = np.concatenate([wd, -alpha * wr])
    mesh = po.paint(attrs, pos, w, resampler=resampler, interlacing=interlacing, compensate=True, out='real')
    bin = po.BinPoles(attrs, edges={'step': step}, ells=ells)
    res = po.spectrum_poles(attrs, bin, mesh, los=los)
    norm = po.fkp_normalization(attrs, data, wd, randoms, wr, cellsize=10.)
    sn = po.fkp_shotnoise(wd, wr)
    out = dict(data=data, randoms=randoms, data_weights=wd.astype('f4'), randoms_weights=wr.astype('f4'),
               meshsize=attrs.meshsize, boxsize=attrs.boxsize, boxcenter=attrs.boxcenter, resampler=resampler,
               interlacing=interlacing, los=los, ells=np.array(ells), step=step, k=bin.xavg, nmodes=bin.nmodes,
               edges=bin.edges, num_raw=np.array(res['num_raw']), norm=norm, shotnoise=sn)
    np.savez_compressed(os.path.join(HERE, f'{name}.npz'), **out)


if __name__ == '__main__':
    box_case('box_cic_n32', 1, 32, 1000., 500., 4000, 'cic', 0, 'z', (0, 2, 4), 0.01, f32=False)
    box_case('box_tsc3_n32', 2, 32, 2000., 1000., 4000, 'tsc', 3, 'z', (0, 2, 4), 0.005, f32=True)
    box_case('box_pcs2_n16x32x16_losx', 3, (16, 32, 16), (400., 800., 400.), (0., 0., 0.), 3000, 'pcs', 2, 'x', (0, 1, 2, 3, 4),
             0.02, f32=True)
    fkp_case('fkp_tsc3_n32_firstpoint', 4, 32, 2000, 8000, 'tsc', 3, 'firstpoint', (0, 2, 4), 0.02)
    print(sorted(f for f in os.listdir(HERE) if f.endswith('.npz')))
