"""Benchmark of the mesh power-spectrum path on one MI355X.

Workload (one step = one mock): a 512^3 mesh, boxsize 2000 Mpc/h, 1.6e7 particles of a seeded
lognormal Poisson mock (float32 positions, unit weights, already resident in HBM), painted with
TSC and 3-fold interlacing, compensated, the DC mode removed (mean subtraction), multipoles
ell = 0, 2, 4 along z in k-bins of 0.005 h/Mpc, with box normalisation and shot noise.

Prints one JSON line.  `value` is mocks/s on the GPU.  `roofline` describes the kernel with
the largest share of the step, from HIP events recorded around each launch on its stream
during the timed steps, against its algorithmic bytes per launch.  `cpu_baseline` times the
float64 NumPy oracle on a bounded sample (see its `sample` string).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, 'jax-power_amd'))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.
MESHSIZE, BOXSIZE, NPART, STEP, ELLS, NI = 512, 2000., 16_000_000, 0.005, (0, 2, 4), 3


def parse():
    p = argparse.ArgumentParser()
    p.add_argument('--gpus', type=int, default=1)
    p.add_argument('--steps', type=int, default=10)
    p.add_argument('--warmup', type=int, default=3)
    p.add_argument('--no-cpu-baseline', action='store_true')
    p.add_argument('--seed', type=int, default=42)
    p.add_argument('--paint-path', type=int, default=None,
                   help='0 automatic, 1 tiled paint + z pass, 2 fused paint + z-FFT, 3 atomics + z pass')
    return p.parse_args()


def algorithmic_bytes(n=MESHSIZE, np_=NPART, ni=NI, nprod=len(ELLS)):
    """Compulsory HBM bytes per launch of each kernel (read inputs once, write outputs once)."""
    ncell = n**3
    H = n * n * (n // 2 + 1)
    return {
        'paint_brick_count': np_ * 12,                     # float32 positions
        'paint_brick_scatter': np_ * 16 + np_ * 16,        # positions + weights in, sorted records out
        'paint_tile': np_ * 16 + ni * ncell * 4,           # sorted records in, n_i meshes out
        'paint_rowsort_hist': np_ * 12,                    # float32 positions
        'paint_rowsort_scatter': np_ * 16 + np_ * 18,      # positions + weights in, records + y keys out
        'paint_rowsort_fine': np_ * 18 + np_ * 16,         # records + keys in, sorted records out
        'paint_zfft': np_ * 16 + ni * H * 8,               # sorted records in, n_i z-transformed spectra out
        'fft_zpass_r2c': ncell * 4 + H * 8,
        'fft_ypass': 2 * H * 8,
        'fft_xpass_combine': ni * H * 8 + H * 8,
        'bin_poles': H * 8,
    }


def end_to_end_bytes(n=MESHSIZE, np_=NPART, ni=NI):
    """Algorithmic bytes of one mock: paint (16 B/particle + n_i 4N^3), n_i r2c (4N^3 + 8H),
    binning (n_i 8H), as counted in the project's baseline accounting."""
    H = n * n * (n // 2 + 1)
    return np_ * 16 + ni * 4 * n**3 + ni * (4 * n**3 + 8 * H) + ni * 8 * H


def cpu_baseline():
    """Float64 NumPy oracle, one core, on one 256^3 mock with 2e6 particles (1/8 of the cells
    and particles of the 512^3 workload); the 512^3 rate is taken as 1/8 of the sample rate."""
    import numpy as np
    from oracle import power_oracle as po
    n, L, npart = 256, 1000., 2_000_000
    rng = np.random.default_rng(0)
    pos = rng.uniform(0., L, (npart, 3)).astype('f4').astype('f8')
    attrs = po.Attrs(n, L, L / 2.)
    t0 = time.perf_counter()
    mesh = po.paint(attrs, pos, None, resampler='tsc', interlacing=NI, compensate=True, out='complex')
    mesh[0, 0, 0] = 0.
    bin = po.BinPoles(attrs, edges={'step': STEP}, ells=ELLS)
    po.spectrum_poles(attrs, bin, mesh, los='z')
    po.box_normalization(attrs, np.ones(npart))
    po.fkp_shotnoise(np.ones(npart))
    dt = time.perf_counter() - t0
    return {'value': 1. / (8. * dt), 'unit': 'mocks/s', 'cores': 1, 'kind': 'port',
            'sample': f'float64 NumPy oracle, 1 core: one 256^3 TSCx3 mock with 2e6 particles in {dt:.1f} s '
                      '(1/8 of the 512^3 cells and particles); value = sample rate / 8'}


def main():
    args = parse()
    if args.gpus != 1:
        print(f'bench.py runs on one GPU only (got --gpus {args.gpus}); nothing is measured for more GPUs.', file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from jaxpower_amd import (MeshAttrs, BinMesh2SpectrumPoles, compute_mesh2_spectrum, compute_box2_normalization,
                              compute_fkp2_shotnoise, generate_lognormal_particles, _lib)

    torch.cuda.set_device(0)
    if args.paint_path is not None:
        _lib.load().mp_set_paint_path(args.paint_path)
    attrs = MeshAttrs(meshsize=MESHSIZE, boxsize=BOXSIZE, boxcenter=BOXSIZE / 2.)
    particles = generate_lognormal_particles(attrs, size=NPART, seed=args.seed)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': STEP}, ells=ELLS)
    torch.cuda.synchronize()
    print(f'# setup done: {particles.size} particles, {bin.nbins} bins', file=sys.stderr, flush=True)

    def step():
        mesh = particles.paint(resampler='tsc', interlacing=NI, compensate=True, out='complex')
        mesh.value[0, 0, 0] = 0.  # mesh - mesh.mean() in Fourier space
        spec = compute_mesh2_spectrum(mesh, bin=bin, los='z')
        return spec.clone(norm=compute_box2_normalization(particles, bin=bin),
                          num_shotnoise=compute_fkp2_shotnoise(particles, bin=bin))

    for _ in range(args.warmup):
        spec = step()
    torch.cuda.synchronize()
    lib = _lib.load()
    lib.mp_timing_reset()
    lib.mp_timing_enable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        spec = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    lib.mp_timing_enable(0)
    buf = _lib.ctypes.create_string_buffer(1 << 16)
    lib.mp_timing_report(buf, len(buf))
    kernels = {}
    for line in buf.value.decode().splitlines():
        name, total, count = line.split()
        kernels[name] = {'total_ms': float(total), 'launches': int(count)}
    print(f'# {args.steps} steps in {elapsed:.3f} s', file=sys.stderr, flush=True)
    alg = algorithmic_bytes()
    table = {}
    for name, k in kernels.items():
        avg_ms = k['total_ms'] / k['launches']
        entry = {'avg_ms': round(avg_ms, 5), 'launches_per_step': k['launches'] / args.steps,
                 'share': round(k['total_ms'] / (elapsed * 1e3), 4)}
        if name in alg:
            entry['achieved_GBs'] = round(alg[name] / (avg_ms * 1e-3) / 1e9, 1)
        table[name] = entry
    dominant = max((n for n in table if n in alg), key=lambda n: kernels[n]['total_ms'])
    d = table[dominant]
    ms_per_step = elapsed * 1e3 / args.steps
    e2e_gbs = end_to_end_bytes() / (ms_per_step * 1e-3) / 1e9
    p0 = spec.get(0).value()
    result = {
        'metric': 'P_ell(k) mocks/s at 512^3 TSCx3 (batched vs plain), HBM GB/s % of peak',
        'value': round(args.steps / elapsed, 4),
        'unit': 'mocks/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(ms_per_step, 4),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32 (f64 bin sums)',
        'data': 'synthetic: seeded lognormal Poisson mock, P_G(k)=2e4 (k/0.1)^-1.5 exp(-(k/0.5)^2)',
        'config': {'workload': 'plain periodic box, one mock per step', 'meshsize': MESHSIZE, 'boxsize': BOXSIZE,
                   'particles': particles.size, 'resampler': 'tsc', 'interlacing': NI, 'compensate': True, 'los': 'z',
                   'ells': list(ELLS), 'kstep': STEP, 'nbins': bin.nbins, 'path': 'paint(out=complex) -> compute_mesh2_spectrum',
                   'paint_path': _lib.load().mp_set_paint_path(-1)},
        'roofline': {'kernel': dominant, 'bound': 'hbm', 'achieved': d['achieved_GBs'], 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': round(d['achieved_GBs'] / HBM_PEAK_GBS, 4), 'traffic': None,
                     'algorithmic_bytes_per_launch': alg[dominant], 'avg_ms': d['avg_ms']},
        'end_to_end': {'algorithmic_GB_per_mock': round(end_to_end_bytes() / 1e9, 3), 'GBs': round(e2e_gbs, 1),
                       'frac_of_peak': round(e2e_gbs / HBM_PEAK_GBS, 4)},
        'kernels': table,
        'check': {'p0_finite': bool(np.all(np.isfinite(p0))), 'p0_first_bins': [float(x) for x in p0[1:4]]},
    }
    if not args.no_cpu_baseline:
        print('# cpu baseline ...', file=sys.stderr, flush=True)
        result['cpu_baseline'] = cpu_baseline()
    print(json.dumps(result))


if __name__ == '__main__':
    main()
