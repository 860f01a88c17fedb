"""Benchmark of the mesh power-spectrum path on one MI355X.

Workload (one step = one mock): a 512^3 mesh, boxsize 2000 Mpc/h, 1.6e7 particles of a seeded
lognormal Poisson mock (float32 positions, unit weights, already resident in HBM), painted with
TSC and 3-fold interlacing, compensated, the DC mode removed (mean subtraction), multipoles
ell = 0, 2, 4 along z in k-bins of 0.005 h/Mpc, with box normalisation and shot noise.

Prints one JSON line.  `value` is mocks/s on the GPU, timed over replays of the whole step
captured as one HIP graph (`eager_ms_per_step` gives the same step issued eagerly).  `roofline`
describes the kernel with the largest share of the step, from HIP events recorded around each
launch on its stream during the eager timed steps, against its algorithmic bytes per launch.
`batch256` is the batched form of the metric (64 uneven mocks at 256^3, batched vs per-mock
calls).  `cpu_baseline` times the float64 NumPy oracle on the same particles (one whole mock),
and `parity` compares the GPU spectrum with it per multipole.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-batch]

HBM traffic: `roofline.traffic` is read from profiles/hbm_traffic_512_tsc3.txt, written by
`python bench.py --summarize-pmc FETCH_DIR WRITE_DIR --copy-bytes N` from two rocprofv3 counter runs
(`rocprofv3 --pmc FETCH_SIZE ...` and `--pmc WRITE_SIZE ...` of `bench.py --copy-bytes N`),
each scaled by what the same counter reads for a plain copy of N bytes in the same run.
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
    p.add_argument('--no-batch', action='store_true', help='skip the 64-mock 256^3 batched-vs-per-mock leg')
    p.add_argument('--seed', type=int, default=42)
    p.add_argument('--copy-bytes', type=int, default=0,
                   help='also run a plain device copy of this many bytes (counter calibration)')
    p.add_argument('--summarize-pmc', nargs=2, metavar=('FETCH_DIR', 'WRITE_DIR'),
                   help='write the HBM traffic file from two rocprofv3 --pmc output folders and exit')
    p.add_argument('--summarize-sq', metavar='SQ_DIR',
                   help='write the SQ counter summary from a rocprofv3 --pmc folder (SQ_COUNTERS) and exit')
    p.add_argument('--summarize-trace', metavar='TRACE_DIR',
                   help='write the kernel statistics summary from a rocprofv3 --kernel-trace --stats folder and exit')
    p.add_argument('--workload', default='box512', choices=['box512', 'fkp1024', 'batch256', 'corr512', 'window256', 'grad256'],
                   help='box512: the headline plain box; fkp1024: survey FKP field, 1024^3, local line of sight; '
                        'batch256: 64 lognormal mocks at 256^3 with uneven counts, batched (HIP graph) vs per-mock calls; '
                        'corr512: correlation multipoles of the 512^3 mock; window256: window matrix of a survey selection at 256^3; '
                        'grad256: forward + hand-written backward of a multipole loss at 256^3 with 2e6 particles')
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
        'fft_ypass': ni * 2 * H * 8,                      # the n_i partial spectra, in place, one launch
        'fft_xpass_combine': ni * H * 8 + H * 8,
        'fft_xpass_bin': ni * H * 8,                      # n_i partial spectra in, per-bin sums out
        'bin_poles': H * 8,
    }


# timing name -> substring of the kernel symbol in rocprofv3 traces
KERNEL_SYMBOLS = {
    'paint_brick_count': 'brick_count_kernel', 'paint_brick_scatter': 'brick_scatter_kernel',
    'paint_tile': 'paint_tile_kernel', 'paint_rowsort_hist': 'rowsort_hist_kernel',
    'paint_rowsort_scatter': 'rowsort_scatter_kernel', 'paint_rowsort_fine': 'rowsort_fine_kernel',
    'paint_zfft': 'paint_zfft_kernel', 'fft_zpass_r2c': 'zpass_r2c_kernel', 'fft_ypass': 'colpass_kernel<512, false',
    'fft_xpass_combine': 'xpass_kernel<512, false', 'fft_xpass_bin': 'xpass_kernel<512, true',
    'bin_poles': 'bin_poles_kernel',
}
COPY_SYMBOL = 'neg_kernel'  # torch.neg(src, out=dst): reads and writes copy_bytes each
TRAFFIC_FILE = os.path.join(ROOT, 'profiles', 'hbm_traffic_512_tsc3.txt')
TRACE_FILE = os.path.join(ROOT, 'profiles', 'kernels_512_tsc3.txt')
PMC_COMMAND = ('rocprofv3 --pmc {counter} -d gpurun_out/pmc_{tag} -o run --output-format csv -- '
               'python3 bench.py --steps 2 --warmup 1 --no-cpu-baseline --no-batch --copy-bytes 1073741824')
TRACE_COMMAND = ('rocprofv3 --kernel-trace --stats -d gpurun_out/kt_bench -o run --output-format csv -- '
                 'python3 bench.py --steps 10 --warmup 3 --no-cpu-baseline --no-batch')


def _pmc_rows(folder, counter):
    """(kernel name, dispatch id, value) of `counter` from the counter_collection.csv files under folder."""
    import csv
    import glob
    rows = []
    for path in glob.glob(os.path.join(folder, '**', '*counter_collection.csv'), recursive=True):
        with open(path) as f:
            for r in csv.DictReader(f):
                if r.get('Counter_Name') == counter:
                    rows.append((r['Kernel_Name'], r.get('Dispatch_Id'), float(r['Counter_Value'])))
    return rows


def summarize_pmc(fetch_dir, write_dir, copy_bytes, out=TRAFFIC_FILE):
    """Mean FETCH_SIZE and WRITE_SIZE (KiB per dispatch) per kernel, scaled by the copy kernel's
    reading for its known byte count; written as `name fetch_bytes write_bytes traffic_bytes`."""
    import collections
    table = {}
    for counter, folder in (('FETCH_SIZE', fetch_dir), ('WRITE_SIZE', write_dir)):
        per = collections.defaultdict(lambda: collections.defaultdict(float))
        for kname, disp, val in _pmc_rows(folder, counter):
            per[kname][disp] += val
        means = {k: sum(v.values()) / len(v) * 1024. for k, v in per.items()}
        copy = [v for k, v in means.items() if COPY_SYMBOL in k]
        if not copy:
            raise SystemExit(f'no {COPY_SYMBOL} dispatch in {folder}')
        # second calibrations of other access patterns: a read-only FP64 sum of the same bytes
        # (the library's dot_partial kernel) and a write-only fill (fill_f32)
        read_only = [v for k, v in means.items() if 'dot_partial_kernel' in k]
        write_only = [v for k, v in means.items() if 'fill_f32_kernel' in k]
        table[counter] = (means, max(copy), max(read_only) if read_only else None, max(write_only) if write_only else None)
    ffac = table['FETCH_SIZE'][1] / copy_bytes
    wfac = table['WRITE_SIZE'][1] / copy_bytes
    rf, ww = table['FETCH_SIZE'][2], table['WRITE_SIZE'][3]
    lines = [f'# HBM bytes per launch: rocprofv3 FETCH_SIZE and WRITE_SIZE, scaled by a {copy_bytes}-byte copy',
             '# made by: ' + PMC_COMMAND.format(counter='FETCH_SIZE', tag='fetch'),
             '#          ' + PMC_COMMAND.format(counter='WRITE_SIZE', tag='write'),
             f'#          python bench.py --summarize-pmc {fetch_dir} {write_dir} --copy-bytes {copy_bytes}',
             f'# copy read {table["FETCH_SIZE"][1]:.6g} B fetched (factor {ffac:.4f}), {table["WRITE_SIZE"][1]:.6g} B written (factor {wfac:.4f})',
             f'# check of the factors on other patterns: a read-only sum of {copy_bytes} B fetched '
             + (f'{rf:.6g} B (factor {rf / copy_bytes:.4f})' if rf else 'n/a')
             + f', a write-only fill of {copy_bytes} B wrote ' + (f'{ww:.6g} B (factor {ww / copy_bytes:.4f})' if ww else 'n/a'),
             '# name fetch_bytes write_bytes traffic_bytes']
    for name, sym in KERNEL_SYMBOLS.items():
        f = [v for k, v in table['FETCH_SIZE'][0].items() if sym in k]
        w = [v for k, v in table['WRITE_SIZE'][0].items() if sym in k]
        if f and w:
            fb, wb = sum(f) / len(f) / ffac, sum(w) / len(w) / wfac
            lines.append(f'{name} {fb:.0f} {wb:.0f} {fb + wb:.0f}')
    with open(out, 'w') as fh:
        fh.write('\n'.join(lines) + '\n')
    print('\n'.join(lines))


SQ_FILE = os.path.join(ROOT, 'profiles', 'sq_counters_512_tsc3.txt')
SQ_COUNTERS = ('SQ_WAVES SQ_INSTS_VALU SQ_INSTS_LDS SQ_LDS_BANK_CONFLICT SQ_ACTIVE_INST_LDS SQ_INSTS_SALU '
               'SQ_BUSY_CU_CYCLES SQ_WAVE_CYCLES GRBM_GUI_ACTIVE')
# MI355X: GRBM_GUI_ACTIVE is summed over the 8 XCDs; 256 CUs.  A wave64 VALU instruction
# occupies a 16-lane SIMD for 4 clocks, so a CU issues at most 1 VALU wave-instruction per clock
# (measured: 0.98 with a pure-FMA kernel).
SQ_XCDS, SQ_CUS = 8, 256
SQ_COMMAND = ('rocprofv3 --pmc ' + SQ_COUNTERS + ' -d gpurun_out/pmc_sq -o run --output-format csv -- '
              'python3 bench.py --steps 2 --warmup 1 --no-cpu-baseline --no-batch')


def summarize_sq(folder, out=SQ_FILE):
    """Per-kernel means of the SQ counters of one rocprofv3 --pmc run (summed over the chip per
    dispatch), with the derived VALU instructions per wave, VALU instructions per busy CU cycle,
    resident waves per busy CU (occupancy) and LDS bank-conflict cycles per LDS-active cycle."""
    import collections
    per = collections.defaultdict(lambda: collections.defaultdict(lambda: collections.defaultdict(float)))
    for counter in SQ_COUNTERS.split():
        for kname, disp, val in _pmc_rows(folder, counter):
            per[kname][counter][disp] += val
    lines = ['# rocprofv3 SQ counters of the 512^3 TSCx3 step kernels (means per dispatch, summed over the chip)',
             '# made by: ' + SQ_COMMAND, f'#          python bench.py --summarize-sq {folder}',
             '# VALU/clk and SALU/clk: wave-instructions per CU per clock (GRBM_GUI_ACTIVE / 8 XCDs clocks, 256 CUs; the',
             '# VALU issue limit is 1.0); LDSconf/clk: LDS bank-conflict cycles per CU per clock',
             f'# {"kernel":52s} {"waves":>8s} {"VALU/wave":>10s} {"VALU/clk":>9s} {"SALU/clk":>9s} {"LDS/wave":>9s} '
             f'{"LDSconf/clk":>11s} {"clocks":>9s}']
    for kname, cnt in sorted(per.items()):
        if 'mp::' not in kname:
            continue
        m = {c: (sum(v.values()) / len(v) if v else 0.) for c, v in cnt.items()}
        waves = m.get('SQ_WAVES', 0.)
        busy = m.get('SQ_BUSY_CU_CYCLES', 0.)
        if waves <= 0 or busy <= 0:
            continue
        name = kname.replace('void ', '').split('(')[0][:52]
        clk = m.get('GRBM_GUI_ACTIVE', 0.) / SQ_XCDS
        if clk <= 0:
            continue
        per_clk = lambda c: m.get(c, 0.) / SQ_CUS / clk  # noqa: E731
        lines.append(f'  {name:52s} {waves:8.0f} {m.get("SQ_INSTS_VALU", 0.) / waves:10.1f} {per_clk("SQ_INSTS_VALU"):9.3f} '
                     f'{per_clk("SQ_INSTS_SALU"):9.3f} {m.get("SQ_INSTS_LDS", 0.) / waves:9.1f} '
                     f'{per_clk("SQ_LDS_BANK_CONFLICT"):11.3f} {clk:9.0f}')
    with open(out, 'w') as fh:
        fh.write('\n'.join(lines) + '\n')
    print('\n'.join(lines))


WINDOW_TRACE_FILE = os.path.join(ROOT, 'profiles', 'kernels_window_256.txt')
WINDOW_TRACE_COMMAND = ('rocprofv3 --kernel-trace --stats -d gpurun_out/kt_window -o run --output-format csv -- '
                        'python3 bench.py --workload window256 --steps 5 --warmup 1 --no-cpu-baseline')
WINDOW_SYMBOLS = {'window_xpass_inv': 'xpass_inv_shell_kernel', 'window_ypass_inv': 'col_fft_kernel<256, true>',
                  'window_zpass': 'zpass_mul_real_kernel', 'window_ypass': 'col_fft_kernel<256, false>',
                  'fft_xpass_bin': 'xpass_kernel<256, true'}


def summarize_trace(folder, out=TRACE_FILE, workload='box512'):
    """Per-kernel launch statistics of one rocprofv3 --kernel-trace --stats run, with the
    algorithmic bytes and the bandwidth they imply."""
    import csv
    import glob
    import re
    path = glob.glob(os.path.join(folder, '**', '*kernel_stats.csv'), recursive=True)[0]
    if workload == 'window256':
        alg, symbols, out = window_bytes(256), WINDOW_SYMBOLS, WINDOW_TRACE_FILE
        title = ('# rocprofv3 kernel statistics of the 256^3 window matrix (57 columns per step, 6 steps: 1 warm-up + 5 timed;'
                 ' the col_fft lines include the 3 passes of each selection FFT at set-up)')
        command = WINDOW_TRACE_COMMAND
    else:
        alg, symbols = algorithmic_bytes(), KERNEL_SYMBOLS
        title = ('# rocprofv3 kernel statistics of the 512^3 TSCx3 step (13 eager steps: 3 warm-up + 10 timed, then the same step '
                 'captured as a HIP graph and replayed 13 times; the paint lines include the autotuner\'s timing launches)')
        command = TRACE_COMMAND
    lines = [title, '# made by: ' + command, f'#          python bench.py --workload {workload} --summarize-trace {folder}',
             f'# {"kernel":58s} {"calls":>6s} {"avg_us":>10s} {"total_ms":>10s} {"share":>6s} {"alg_MB":>9s} {"alg_GB/s":>9s}']
    other = [0, 0.]
    for r in csv.DictReader(open(path)):
        name = re.sub(r'\(.*', '', r['Name']).replace('void ', '')
        if not name.startswith('mp::'):
            other[0] += int(r['Calls'])
            other[1] += float(r['TotalDurationNs']) * 1e-6
            continue
        avg = float(r['AverageNs']) * 1e-3
        timing = [t for t, sym in symbols.items() if sym in name]
        b = alg.get(timing[0]) if timing and not (timing[0] == 'fft_ypass' and 'true' in name) else None
        gbs = f'{b / (avg * 1e-6) / 1e9:9.1f}' if b else f'{"":9s}'
        mb = f'{b / 1e6:9.1f}' if b else f'{"":9s}'
        lines.append(f'{name[:60]:60s} {int(r["Calls"]):6d} {avg:10.1f} {float(r["TotalDurationNs"]) * 1e-6:10.3f} '
                     f'{float(r["Percentage"]):6.2f} {mb} {gbs}')
    lines.append(f'# other kernels (torch and rocPRIM, mock generation at set-up): {other[0]} launches, {other[1]:.3f} ms')
    lines.append('# bin_modes and the plain z-pass/c2r kernels run once at set-up (bin tables, Gaussian mock)' if workload != 'window256'
                 else '# bin_modes runs once at set-up (bin tables); the plain z-pass r2c/c2r and spectrum_product are the selection\'s '
                      'Q = c2r(|A|^2) once per step')
    with open(out, 'w') as fh:
        fh.write('\n'.join(lines) + '\n')
    print('\n'.join(lines))


def read_traffic(path=TRAFFIC_FILE):
    if not os.path.exists(path):
        return {}
    out = {}
    with open(path) as f:
        for line in f:
            if line.startswith('#') or not line.strip():
                continue
            name, fb, wb, tb = line.split()
            out[name] = float(tb)
    return out


def read_issue_limit(name, path=SQ_FILE):
    """VALU wave-instructions per CU clock of kernel `name` in the committed SQ-counter summary
    (the fastest of its autotuner candidates, i.e. the one the step runs), or None.  A CU issues
    at most one VALU wave-instruction per clock, so a value near 1 marks a kernel bound by
    instruction issue rather than by HBM."""
    symbol = KERNEL_SYMBOLS.get(name)
    if symbol is None or not os.path.exists(path):
        return None
    best = None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if line.startswith('#') or len(parts) < 8 or symbol not in line:
                continue
            clocks, valu = float(parts[-1]), float(parts[-5])
            if best is None or clocks < best[0]:
                best = (clocks, valu)
    return None if best is None else {'valu_per_cu_clock': best[1], 'issue_limit': 1.0,
                                      'source': 'profiles/sq_counters_512_tsc3.txt'}


def end_to_end_bytes(n=MESHSIZE, np_=NPART, ni=NI):
    """Algorithmic bytes of one mock: paint (16 B/particle + n_i 4N^3), n_i r2c (4N^3 + 8H),
    binning (n_i 8H), as counted in the project's baseline accounting."""
    H = n * n * (n // 2 + 1)
    return np_ * 16 + ni * 4 * n**3 + ni * (4 * n**3 + 8 * H) + ni * 8 * H


def _cpu_model():
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def cpu_baseline(particles, spec, bin):
    """The float64 NumPy oracle on the identical input: the GPU mock's float32 positions copied
    to the host once, painted (TSC, interlacing 3, compensated), DC removed, binned into the
    same 160 shells (the bin set-up timed apart, as the GPU step excludes it), normalized, with
    shot noise.  scipy.fft on all the host cores this process may use, the scatter split over
    threads.  Returns the baseline record and the parity of the GPU spectrum against it."""
    import numpy as np
    from oracle import power_oracle as po
    threads = len(os.sched_getaffinity(0)) if hasattr(os, 'sched_getaffinity') else (os.cpu_count() or 1)
    po.FFT_WORKERS = threads
    pos = particles.positions.cpu().numpy().astype('f8')  # the same float32 values
    attrs = po.Attrs(MESHSIZE, BOXSIZE, BOXSIZE / 2.)
    t0 = time.perf_counter()
    obin = po.BinPoles(attrs, edges={'step': STEP}, ells=ELLS)
    tbin = time.perf_counter() - t0
    t0 = time.perf_counter()
    mesh = po.paint(attrs, pos, None, resampler='tsc', interlacing=NI, compensate=True, out='complex', threads=threads)
    mesh[0, 0, 0] = 0.
    res = po.spectrum_poles(attrs, obin, mesh, los='z')
    norm = po.box_normalization(attrs, np.ones(len(pos)))
    sn = po.fkp_shotnoise(np.ones(len(pos)))
    dt = time.perf_counter() - t0
    del mesh
    want = np.array(res['num_raw'])
    got = np.array([spec.get(ell).values('num_raw') for ell in ELLS])
    scale = np.maximum(np.abs(want), np.abs(want[0])[None, :])
    ok = ~np.isnan(want)
    rel = np.where(ok, np.abs(got - want) / np.where(ok, scale, 1.), 0.)
    vgot = np.array([spec.get(ell).value() for ell in ELLS])
    vwant = np.array([po.pole_value(want[i], norm, sn if ell == 0 else 0.) for i, ell in enumerate(ELLS)])
    vscale = np.maximum(np.abs(vwant), np.abs(vwant[0])[None, :])
    vrel = np.where(ok, np.abs(vgot - vwant) / np.where(ok, vscale, 1.), 0.)
    parity = {'against': 'float64 NumPy oracle on the same particles (the GPU float32 positions copied to the host)',
              'tolerance': '|dP_l| <= 1e-5 max(|P_l|, |P_0|) per bin',
              'max_rel_err_num_raw': {str(ell): float(rel[i].max()) for i, ell in enumerate(ELLS)},
              'max_rel_err_value': {str(ell): float(vrel[i].max()) for i, ell in enumerate(ELLS)},
              'nmodes_bitwise_equal': bool(np.array_equal(spec.get(0).values('nmodes'), obin.nmodes)),
              'within_tolerance': bool(np.all(rel <= 1e-5))}
    baseline = {'value': round(1. / dt, 5), 'unit': 'mocks/s', 'cores': threads, 'kind': 'port',
                'sample': f'one whole mock: the float64 NumPy oracle on the identical 512^3 input ({len(pos)} particles, '
                          f'TSC x3, 160 bins), paint + spectrum + normalization + shot noise in {dt:.1f} s '
                          f'(bin set-up {tbin:.1f} s, not counted); scipy.fft with {threads} workers, scatter on {threads} '
                          f'threads; JAX is absent, so this is the NumPy restatement, not XLA; host CPU: {_cpu_model()}'}
    return baseline, parity


def batch_rates(nmock=64, steps=3, warmup=1, seed=42, resamplers=('cic', 'tsc'), log=True):
    """64 lognormal mocks at 256^3 (boxsize 1000, counts uniform in [5e5, 4e6]), interlacing 2,
    compensated, ells 0, 2, 4, k step 0.005, box normalization and shot noise: mocks/s of the
    per-mock eager calls and of the batched forms, and their ratio."""
    import numpy as np
    import torch
    from jaxpower_amd import MeshAttrs, BinMesh2SpectrumPoles, Mesh2SpectrumBatch, generate_lognormal_particles
    attrs = MeshAttrs(meshsize=256, boxsize=1000., boxcenter=500.)
    counts = np.random.default_rng(seed).integers(500_000, 4_000_001, nmock)
    fields = [generate_lognormal_particles(attrs, size=int(c), seed=seed + 1 + i) for i, c in enumerate(counts)]
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': STEP}, ells=ELLS)
    torch.cuda.synchronize()
    if log:
        print(f'# batch setup done: {sum(f.size for f in fields)} particles in {nmock} mocks', file=sys.stderr, flush=True)
    from jaxpower_amd import compute_mesh2_spectrum, compute_box2_normalization, compute_fkp2_shotnoise
    out = {}
    for resampler in resamplers:
        rates = {}

        def plain(field):
            # the reference harness per mock (scripts/abacus_hf.py:357-373): paint, mean removed,
            # compute_mesh2_spectrum, normalization and shot noise, results ready before the
            # next mock (block_until_ready, then .write)
            mesh = field.paint(resampler=resampler, interlacing=2, compensate=True, out='complex').at[0, 0, 0].set(0.)
            spec = compute_mesh2_spectrum(mesh, bin=bin, los='z')
            spec = spec.clone(norm=compute_box2_normalization(field, bin=bin), num_shotnoise=compute_fkp2_shotnoise(field, bin=bin))
            return [spec.get(ell).value() for ell in ELLS]

        plain(fields[0])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            for f in fields:
                plain(f)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        rates['plain_calls'] = {'mocks_per_s': round(nmock * steps / el, 2), 'ms_per_mock': round(el * 1e3 / (nmock * steps), 4)}
        if log:
            print(f'# {resampler} plain_calls: {rates["plain_calls"]}', file=sys.stderr, flush=True)
        def consume(specs):
            # every multipole read back to the host, as the plain calls have them
            return [[spec.get(ell).value() for ell in ELLS] for spec in specs]

        modes = (('per_mock', False, 1, False), ('batched_graph', True, 1, False), ('batched_graph_4streams', True, 4, False),
                 ('batched_merged', False, 1, True), ('batched_merged_graph', True, 1, True))
        for mode, graph, streams, merged in modes:
            batch = Mesh2SpectrumBatch(fields, bin, resampler=resampler, interlacing=2, compensate=True, graph=graph, streams=streams,
                                       merged=merged)
            for _ in range(max(warmup, 1)):
                consume(batch())
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                consume(batch())
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            rates[mode] = {'mocks_per_s': round(nmock * steps / el, 2), 'ms_per_mock': round(el * 1e3 / (nmock * steps), 4)}
            if log:
                print(f'# {resampler} {mode}: {rates[mode]}', file=sys.stderr, flush=True)
            del batch
        best = max((m for m in rates if m.startswith('batched')), key=lambda m: rates[m]['mocks_per_s'])
        rates['best_batched'] = best
        # against the plain per-mock calls of the reference's harness, and against the per-mock
        # kernel sequence issued without any host synchronisation
        rates['speedup'] = round(rates[best]['mocks_per_s'] / rates['plain_calls']['mocks_per_s'], 3)
        rates['speedup_vs_async_per_mock'] = round(rates[best]['mocks_per_s'] / rates['per_mock']['mocks_per_s'], 3)
        out[resampler] = rates
    return out, fields


def run_fkp1024(args):
    """Survey FKP measurement: 5e6 data and 1e8 randoms uniform in the shell 1000 <= r < 2000
    Mpc/h within 60 degrees of the z axis (observer at the origin), FKP weights
    1/(1 + nbar 1e4), 1024^3 mesh with boxpad 2, TSC with 3-fold interlacing, compensated,
    los='firstpoint', ells 0, 2, 4 (real spherical harmonics), k step 0.005, FKP normalization
    on the cellsize-10 regridded mesh and shot noise.  One step = paint + spectrum +
    normalization + shot noise.  Prints one JSON line."""
    import numpy as np
    import torch
    from jaxpower_amd import (ParticleField, FKPField, BinMesh2SpectrumPoles, compute_mesh2_spectrum, get_mesh_attrs,
                              compute_fkp2_normalization, compute_fkp2_shotnoise, _lib)
    from jaxpower_amd.mock import generate_shell_cap_positions, shell_cap_volume
    torch.cuda.set_device(0)
    nd, nr = 5_000_000, 100_000_000
    dpos = generate_shell_cap_positions(nd, seed=args.seed, device='cuda')
    rpos = generate_shell_cap_positions(nr, seed=args.seed + 1, device='cuda')
    nbar = nd / shell_cap_volume()
    wfkp = 1. / (1. + nbar * 1e4)
    attrs = get_mesh_attrs(dpos.cpu().numpy()[::50], rpos.cpu().numpy()[::1000], boxpad=2., meshsize=1024)
    data = ParticleField(dpos, torch.full((nd,), wfkp, device='cuda'), attrs=attrs)
    randoms = ParticleField(rpos, torch.full((nr,), wfkp, device='cuda'), attrs=attrs)
    fkp = FKPField(data, randoms)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': STEP}, ells=ELLS)
    torch.cuda.synchronize()
    print(f'# fkp setup done: boxsize {attrs.boxsize[0]:.1f}, {bin.nbins} bins', file=sys.stderr, flush=True)

    def step():
        mesh = fkp.paint(resampler='tsc', interlacing=NI, compensate=True, out='complex')
        spec = compute_mesh2_spectrum(mesh, bin=bin, los='firstpoint')
        return spec.clone(norm=compute_fkp2_normalization(fkp, bin=bin), num_shotnoise=compute_fkp2_shotnoise(fkp, bin=bin))

    for _ in range(args.warmup):
        spec = step()
    torch.cuda.synchronize()
    lib = _lib.load()
    lib.mp_timing_reset()
    lib.mp_timing_enable(1)
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
        kernels[name] = {'avg_ms': round(float(total) / int(count), 4), 'launches_per_step': int(count) / args.steps,
                         'share': round(float(total) / (elapsed * 1e3), 4)}
    p0 = spec.get(0).value()
    result = {'metric': 'FKP P_ell(k) local line of sight, measurements/s', 'value': round(args.steps / elapsed, 4),
              'unit': 'measurements/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
              'ms_per_step': round(elapsed * 1e3 / args.steps, 3), 'higher_is_better': True, 'scaling': 'weak',
              'vs_baseline': None, 'dtype': 'f32 (f64 bin sums)',
              'data': 'synthetic: 5e6 data and 1e8 randoms uniform in a shell-cap, FKP weights',
              'config': {'workload': 'FKP survey field, 1024^3, local line of sight (firstpoint)', 'meshsize': 1024,
                         'boxsize': float(attrs.boxsize[0]), 'data': nd, 'randoms': nr, 'resampler': 'tsc', 'interlacing': NI,
                         'compensate': True, 'los': 'firstpoint', 'ells': list(ELLS), 'kstep': STEP, 'nbins': bin.nbins},
              'kernels': kernels,
              'check': {'p0_finite': bool(np.all(np.isfinite(p0[1:]))), 'p0_first_bins': [float(x) for x in p0[1:4]]}}
    print(json.dumps(result))


def run_corr512(args):
    """Correlation-function multipoles of the 512^3 TSCx3 lognormal mock (the headline box):
    paint(out='complex'), DC removed, compute_mesh2_correlation with los 'z', ells 0, 2, 4,
    separation bins of 4 Mpc/h up to 200 Mpc/h (the binning of the reference's
    tests/test_mesh2.py:130-132).  One step = paint + correlation.  Prints one JSON line."""
    import numpy as np
    import torch
    from jaxpower_amd import MeshAttrs, BinMesh2CorrelationPoles, compute_mesh2_correlation, generate_lognormal_particles, _lib
    torch.cuda.set_device(0)
    attrs = MeshAttrs(meshsize=MESHSIZE, boxsize=BOXSIZE, boxcenter=BOXSIZE / 2.)
    particles = generate_lognormal_particles(attrs, size=NPART, seed=args.seed)
    bin = BinMesh2CorrelationPoles(attrs, edges={'step': 4., 'max': 200.}, ells=ELLS)
    torch.cuda.synchronize()
    print(f'# corr setup done: {particles.size} particles, {bin.nbins} bins', file=sys.stderr, flush=True)

    def step():
        mesh = particles.paint(resampler='tsc', interlacing=NI, compensate=True, out='complex').at[0, 0, 0].set(0.)
        return compute_mesh2_correlation(mesh, bin=bin, los='z')

    for _ in range(args.warmup):
        xi = step()
    torch.cuda.synchronize()
    lib = _lib.load()
    lib.mp_timing_reset()
    lib.mp_timing_enable(1)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        xi = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    lib.mp_timing_enable(0)
    buf = _lib.ctypes.create_string_buffer(1 << 16)
    lib.mp_timing_report(buf, len(buf))
    kernels = {}
    for line in buf.value.decode().splitlines():
        name, total, count = line.split()
        kernels[name] = {'avg_ms': round(float(total) / int(count), 4), 'launches_per_step': int(count) / args.steps,
                         'share': round(float(total) / (elapsed * 1e3), 4)}
    x0 = xi.get(0).value()
    result = {'metric': 'xi_ell(s) mocks/s at 512^3 TSCx3', 'value': round(args.steps / elapsed, 4), 'unit': 'mocks/s',
              'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': round(elapsed * 1e3 / args.steps, 3),
              'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32 (f64 bin sums)',
              'data': 'synthetic: seeded lognormal Poisson mock, P_G(k)=2e4 (k/0.1)^-1.5 exp(-(k/0.5)^2)',
              'config': {'workload': 'plain periodic box, correlation multipoles', 'meshsize': MESHSIZE, 'boxsize': BOXSIZE,
                         'particles': particles.size, 'resampler': 'tsc', 'interlacing': NI, 'compensate': True, 'los': 'z',
                         'ells': list(ELLS), 'sstep': 4., 'smax': 200., 'nbins': bin.nbins},
              'kernels': kernels,
              'check': {'xi0_finite': bool(np.all(np.isfinite(x0[1:]))), 'xi0_bins_2_4': [float(v) for v in x0[2:5]]}}
    print(json.dumps(result))


def window_bytes(n):
    """Compulsory HBM bytes per launch of the window-matrix kernels at n^3 (one input shell)."""
    H = n * n * (n // 2 + 1)
    return {'window_xpass_inv': H * 8,                     # shell generated, spectrum written
            'window_ypass_inv': 2 * H * 8, 'window_zpass': 2 * H * 8 + n**3 * 4, 'window_ypass': 2 * H * 8,
            'fft_xpass_bin': H * 8}                          # spectrum read, per-bin sums out


def run_window256(args):
    """Window matrix of a survey selection (compute_mesh2_spectrum_window, global line of sight
    'x', theory on the mesh): 256^3 mesh, boxsize 2000, selection = CIC paint of 1e6 points of
    a Gaussian blob squashed along y and z (the selection of the reference's
    tests/test_mesh2.py:577-607), input shells linspace(0.01, 0.1, 20) for ellin = 0, 2, 4,
    output ells 0, 2, 4 in k-bins of 0.005.  One step = the whole matrix (57 columns).  Prints
    one JSON line."""
    import numpy as np
    import torch
    from jaxpower_amd import MeshAttrs, ParticleField, BinMesh2SpectrumPoles, compute_mesh2_spectrum_window, _lib
    torch.cuda.set_device(0)
    n, boxsize, size = 256, 2000., 1_000_000
    attrs = MeshAttrs(meshsize=n, boxsize=boxsize, boxcenter=0.)
    rng = np.random.default_rng(args.seed)
    pos = np.array([1., 0.2, 0.2]) * 0.2 * rng.normal(size=(size, 3))
    w = np.all((pos > -0.2) & (pos < 0.2), axis=-1).astype('f4')
    pos = (pos * boxsize).astype('f4')
    sel = ParticleField(torch.from_numpy(pos).cuda(), torch.from_numpy(w).cuda(), attrs=attrs).paint(
        resampler='cic', interlacing=1, compensate=False)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': STEP}, ells=ELLS)
    edgesin = np.linspace(0.01, 0.1, 20)
    ellsin = (0, 2, 4)
    torch.cuda.synchronize()
    print(f'# window setup done: {bin.nbins} output bins, {3 * (len(edgesin) - 1)} columns', file=sys.stderr, flush=True)

    def step():
        return compute_mesh2_spectrum_window(sel, edgesin=edgesin, ellsin=ellsin, bin=bin, los='x')

    for _ in range(args.warmup):
        wm = step()
    torch.cuda.synchronize()
    lib = _lib.load()
    lib.mp_timing_reset()
    lib.mp_timing_enable(1)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        wm = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    lib.mp_timing_enable(0)
    buf = _lib.ctypes.create_string_buffer(1 << 16)
    lib.mp_timing_report(buf, len(buf))
    alg = window_bytes(n)
    kernels = {}
    for line in buf.value.decode().splitlines():
        name, total, count = line.split()
        avg = float(total) / int(count)
        kernels[name] = {'avg_ms': round(avg, 5), 'launches_per_step': int(count) / args.steps,
                         'share': round(float(total) / (elapsed * 1e3), 4)}
        if name in alg:
            kernels[name]['achieved_GBs'] = round(alg[name] / (avg * 1e-3) / 1e9, 1)
    top = max((k for k in kernels if k in alg), key=lambda k: kernels[k]['share'])
    ncol = wm.value().shape[1]
    value = wm.value()
    cpu = None
    if not args.no_cpu_baseline:  # float64 oracle, one core: 2 input shells x 3 ellsin of the same matrix
        from oracle import power_oracle as po
        oattrs = po.Attrs(n, boxsize, 0.)
        obin = po.BinPoles(oattrs, edges={'step': STEP}, ells=ELLS)
        host_sel = sel.value.cpu().numpy().astype('f8')
        t0 = time.perf_counter()
        po.spectrum_window(oattrs, obin, edgesin[:3], ellsin=ellsin, meshes=[host_sel], los='x')
        dt = time.perf_counter() - t0
        cpu = {'value': round(6 / dt, 4), 'unit': 'columns/s', 'cores': 1, 'kind': 'port',
               'sample': f'float64 NumPy oracle, 1 core: 6 columns (2 input shells x ellin 0, 2, 4) of the same 256^3 '
                         f'matrix, including the selection product once, {dt:.1f} s'}
    result = {'metric': 'window-matrix columns/s at 256^3 (global line of sight)', 'value': round(ncol * args.steps / elapsed, 2),
              'unit': 'columns/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
              'ms_per_step': round(elapsed * 1e3 / args.steps, 3), 'higher_is_better': True, 'scaling': 'weak',
              'vs_baseline': None, 'dtype': 'f32 (f64 bin sums)',
              'data': 'synthetic: CIC paint of 1e6 Gaussian-blob points (seeded)',
              'config': {'workload': 'window matrix, survey selection, global line of sight x', 'meshsize': n, 'boxsize': boxsize,
                         'edgesin': 'linspace(0.01, 0.1, 20)', 'ellsin': list(ellsin), 'ells': list(ELLS), 'kstep': STEP,
                         'columns': ncol, 'rows': value.shape[0]},
              'roofline': {'kernel': top, 'bound': 'hbm', 'achieved': kernels[top]['achieved_GBs'], 'peak': HBM_PEAK_GBS,
                           'unit': 'GB/s', 'frac': round(kernels[top]['achieved_GBs'] / HBM_PEAK_GBS, 4), 'traffic': None,
                           'algorithmic_bytes_per_launch': alg[top], 'avg_ms': kernels[top]['avg_ms']},
              'cpu_baseline': cpu, 'kernels': kernels,
              'check': {'finite': bool(np.all(np.isfinite(value[:, :]))), 'w_max': float(np.abs(value).max())}}
    print(json.dumps(result))


def run_batch256(args):
    """64 lognormal mocks (the 512^3 recipe at 256^3, boxsize 1000) with particle counts drawn
    uniformly in [5e5, 4e6], CIC and TSC, 2-fold interlacing, compensated, DC removed, ells
    0, 2, 4, k step 0.005, box normalization and shot noise: mocks/s of the batched forms and of
    the same sequence as per-mock eager calls.  One step = the 64 mocks.  Prints one JSON line."""
    import torch
    from jaxpower_amd import _lib
    torch.cuda.set_device(0)
    nmock = 64
    out, fields = batch_rates(nmock=nmock, steps=args.steps, warmup=args.warmup, seed=args.seed)
    best = out['tsc']['best_batched']
    buf = _lib.ctypes.create_string_buffer(1 << 14)
    _lib.load().mp_paint_tuning_report(buf, len(buf))
    result = {'metric': 'mocks/s, 64 mocks at 256^3 (batched vs per-mock)', 'value': out['tsc'][best]['mocks_per_s'],
              'unit': 'mocks/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
              'ms_per_step': round(1e3 * nmock / out['tsc'][best]['mocks_per_s'], 3), 'higher_is_better': True,
              'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32 (f64 bin sums)',
              'data': 'synthetic: 64 seeded lognormal Poisson mocks, counts uniform in [5e5, 4e6]',
              'config': {'workload': '64 lognormal mocks, 256^3, boxsize 1000, interlacing 2, compensated, los z', 'meshsize': 256,
                         'mocks': nmock, 'particles': int(sum(f.size for f in fields)), 'ells': list(ELLS), 'kstep': STEP},
              'results': out, 'tuning': buf.value.decode().strip().splitlines()}
    print(json.dumps(result))


def run_grad256(args):
    """The differentiated estimator: 256^3, boxsize 1000, 2e6 lognormal particles (the 512^3
    recipe), TSC with 3-fold interlacing, compensated, ells 0, 2, k step 0.005; loss =
    sum over ell and k of (P_ell - T_ell)^2 / T_ell^2, P the normalized, shot-noise-subtracted
    multipoles and T those of a second mock (seed + 1).  One step = forward + hand-written
    backward (d loss / d positions and weights).  The gradient is checked against a central
    difference of the float64 oracle along a random displacement of every particle (and a
    random weight perturbation).  Prints one JSON line."""
    import numpy as np
    import torch
    from jaxpower_amd import (MeshAttrs, BinMesh2SpectrumPoles, generate_lognormal_particles, mesh2_spectrum_poles,
                              box_spectrum_values)
    torch.cuda.set_device(0)
    n, boxsize, size, ells = 256, 1000., 2_000_000, (0, 2)
    attrs = MeshAttrs(meshsize=n, boxsize=boxsize, boxcenter=boxsize / 2.)
    bin = BinMesh2SpectrumPoles(attrs, edges={'step': STEP}, ells=ells)
    part = generate_lognormal_particles(attrs, size=size, seed=args.seed)
    tpart = generate_lognormal_particles(attrs, size=size, seed=args.seed + 1)
    kw = dict(attrs=attrs, bin=bin, resampler='tsc', interlacing=NI, compensate=True, los='z')
    pos0 = part.positions
    w0 = torch.ones(pos0.shape[0], dtype=torch.float32, device=pos0.device)
    with torch.no_grad():
        tw = torch.ones(tpart.positions.shape[0], dtype=torch.float32, device=pos0.device)
        target = box_spectrum_values(mesh2_spectrum_poles(tpart.positions, tw, **kw), tw, attrs, ells)
    ok = target != 0.
    torch.cuda.synchronize()
    print(f'# grad setup done: {pos0.shape[0]} particles, {bin.nbins} bins', file=sys.stderr, flush=True)

    def step():
        pos = pos0.detach().clone().requires_grad_(True)
        w = w0.detach().clone().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        values = box_spectrum_values(mesh2_spectrum_poles(pos, w, **kw), w, attrs, ells)
        loss = torch.sum(((values - target)[ok] / target[ok])**2)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        return float(loss.detach()), pos.grad, w.grad, t1 - t0, t2 - t1

    for _ in range(max(args.warmup, 1)):
        step()
    torch.cuda.reset_peak_memory_stats()
    fw, bw = [], []
    for _ in range(args.steps):
        loss, gpos, gw, tf, tb = step()
        fw.append(tf)
        bw.append(tb)
    peak = torch.cuda.max_memory_allocated()
    ms_f, ms_b = 1e3 * float(np.median(fw)), 1e3 * float(np.median(bw))
    result = {'metric': 'loss-gradient evaluations/s (256^3, 2e6 particles, TSC x3)', 'value': round(1e3 / (ms_f + ms_b), 3),
              'unit': 'evaluations/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': round(ms_f + ms_b, 3),
              'ms_forward': round(ms_f, 3), 'ms_backward': round(ms_b, 3), 'higher_is_better': True, 'scaling': 'weak',
              'vs_baseline': None, 'dtype': 'f32 (f64 bin sums)',
              'data': 'synthetic: seeded lognormal Poisson mocks (target from seed + 1)',
              'config': {'workload': 'differentiated estimator', 'meshsize': n, 'boxsize': boxsize, 'particles': int(pos0.shape[0]),
                         'resampler': 'tsc', 'interlacing': NI, 'compensate': True, 'ells': list(ells), 'kstep': STEP,
                         'loss': 'sum_ell,k (P_ell - T_ell)^2 / T_ell^2'},
              'memory': {'peak_bytes_256': int(peak),
                         'plan_1024': 'the backward recomputes the painted spectrum (one interlaced paint + FFT) instead of '
                                      'storing it; peak ~ n_i half spectra (partials, then Y) + n_i real meshes (D) + one '
                                      'half spectrum: 3 x 4.3 + 3 x 4.3 + 4.3 GB ~ 30 GB at 1024^3'},
              'loss': loss}
    if not args.no_cpu_baseline:  # directional check against the float64 oracle (the same particles)
        from oracle import power_oracle as po
        po.FFT_WORKERS = len(os.sched_getaffinity(0))
        oattrs = po.Attrs(n, boxsize, boxsize / 2.)
        obin = po.BinPoles(oattrs, edges={'step': STEP}, ells=ells)
        tgt = target.cpu().numpy()
        okn = tgt != 0.
        pos_h = pos0.cpu().numpy().astype('f8')
        w_h = np.ones(len(pos_h))

        def oloss(p, ww):
            mesh = po.paint(oattrs, p, ww, resampler='tsc', interlacing=NI, compensate=True, out='complex',
                            threads=po.FFT_WORKERS)
            mesh[0, 0, 0] = 0.
            num = np.array(po.spectrum_poles(oattrs, obin, mesh, los='z')['num_raw'])
            norm, sn = po.box_normalization(oattrs, ww), po.fkp_shotnoise(ww)
            vals = np.array([po.pole_value(num[i], norm, sn if ell == 0 else 0.) for i, ell in enumerate(ells)])
            return float(np.sum(((vals - tgt)[okn] / tgt[okn])**2))

        rng = np.random.default_rng(args.seed + 7)
        v = rng.normal(size=pos_h.shape)
        u = rng.normal(size=w_h.shape)
        t0 = time.perf_counter()
        h, hw = 1e-3, 1e-5
        fdpos = (oloss(pos_h + h * v, w_h) - oloss(pos_h - h * v, w_h)) / (2 * h)
        fdw = (oloss(pos_h, w_h + hw * u) - oloss(pos_h, w_h - hw * u)) / (2 * hw)
        dt = time.perf_counter() - t0
        dpos = float(np.sum(gpos.cpu().numpy().astype('f8') * v))
        dw = float(np.sum(gw.cpu().numpy().astype('f8') * u))
        result['gradient_check'] = {'against': 'central differences of the float64 oracle loss along a random displacement v '
                                               '(h = 1e-3 Mpc/h) and a random weight perturbation u (h = 1e-5)',
                                    'grad_dot_v': dpos, 'fd_v': fdpos, 'rel_err_v': abs(dpos - fdpos) / abs(fdpos),
                                    'grad_dot_u': dw, 'fd_u': fdw, 'rel_err_u': abs(dw - fdw) / abs(fdw), 'oracle_s': round(dt, 1)}
    print(json.dumps(result))


def main():
    args = parse()
    if args.summarize_trace:
        summarize_trace(args.summarize_trace, workload=args.workload)
        return
    if args.summarize_sq:
        summarize_sq(args.summarize_sq)
        return
    if args.summarize_pmc:
        summarize_pmc(*args.summarize_pmc, copy_bytes=args.copy_bytes)
        return
    if args.workload == 'batch256':
        if args.gpus != 1:
            print('bench.py runs on one GPU only.', file=sys.stderr)
            sys.exit(2)
        run_batch256(args)
        return
    if args.workload == 'corr512':
        if args.gpus != 1:
            print('bench.py runs on one GPU only.', file=sys.stderr)
            sys.exit(2)
        run_corr512(args)
        return
    if args.workload == 'window256':
        if args.gpus != 1:
            print('bench.py runs on one GPU only.', file=sys.stderr)
            sys.exit(2)
        run_window256(args)
        return
    if args.workload == 'grad256':
        if args.gpus != 1:
            print('bench.py runs on one GPU only.', file=sys.stderr)
            sys.exit(2)
        run_grad256(args)
        return
    if args.workload == 'fkp1024':
        if args.gpus != 1:
            print('bench.py runs on one GPU only.', file=sys.stderr)
            sys.exit(2)
        run_fkp1024(args)
        return
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
    if args.copy_bytes:
        src = torch.ones(args.copy_bytes // 4, dtype=torch.float32, device='cuda')
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.neg(src, out=dst)
        lib = _lib.load()
        ws = _lib.workspace(lib.mp_reduce_workspace_bytes(), 'cuda')
        res = torch.zeros(1, dtype=torch.float64, device='cuda')
        for _ in range(3):  # read-only and write-only calibrations of the same byte count
            _lib.call('mp_sum_f64', _lib.stream_handle(), _lib.ptr(src), None, src.numel(), 0, _lib.ptr(res), _lib.ptr(ws),
                      ws.numel())
            _lib.call('mp_fill_f32', _lib.stream_handle(), _lib.ptr(dst), dst.numel(), 1.)
        torch.cuda.synchronize()
        del src, dst
    torch.cuda.synchronize()
    print(f'# setup done: {particles.size} particles, {bin.nbins} bins', file=sys.stderr, flush=True)

    def step():
        mesh = particles.paint(resampler='tsc', interlacing=NI, compensate=True, out='complex')
        mesh = mesh.at[0, 0, 0].set(0.)  # mesh - mesh.mean() in Fourier space
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
    # The same step replayed as one captured HIP graph (the whole mock: paint, FFT passes,
    # binning, normalization and shot-noise sums from the resident particles): the GPU runs the
    # step's kernels back to back whatever the host thread is doing.  `value` is taken from these
    # timed replays; the per-kernel table above comes from the eager timed steps (HIP events
    # cannot be recorded from the host inside a replayed graph).
    from jaxpower_amd import Mesh2SpectrumBatch
    graph = Mesh2SpectrumBatch([particles], bin, resampler='tsc', interlacing=NI, compensate=True, los='z', graph=True,
                               merged=False)
    for _ in range(max(args.warmup, 1)):
        gspec = graph()[0]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        gspec = graph()[0]
    torch.cuda.synchronize()
    gelapsed = time.perf_counter() - t0
    print(f'# {args.steps} graph replays in {gelapsed:.3f} s', file=sys.stderr, flush=True)
    # the FP64 bin sums of an x-pass block are added in LDS in an order the wave scheduling
    # decides, so a replay agrees with the eager step to rounding, not necessarily bit for bit
    def close(a, b):
        return np.allclose(a, b, rtol=1e-12, atol=1e-12 * np.nanmax(np.abs(b)), equal_nan=True)
    same = all(close(gspec.get(ell).values('num_raw'), spec.get(ell).values('num_raw')) for ell in ELLS)
    eager_ms = elapsed * 1e3 / args.steps
    ms_per_step = gelapsed * 1e3 / args.steps
    e2e_gbs = end_to_end_bytes() / (ms_per_step * 1e-3) / 1e9
    p0 = spec.get(0).value()
    traffic = read_traffic()
    result = {
        'metric': 'P_ell(k) mocks/s at 512^3 TSCx3 (batched vs plain), HBM GB/s % of peak',
        'value': round(args.steps / gelapsed, 4),
        'unit': 'mocks/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(ms_per_step, 4),
        'eager_ms_per_step': round(eager_ms, 4),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32 (f64 bin sums)',
        'data': 'synthetic: seeded lognormal Poisson mock, P_G(k)=2e4 (k/0.1)^-1.5 exp(-(k/0.5)^2)',
        'config': {'workload': 'plain periodic box, one mock per step', 'meshsize': MESHSIZE, 'boxsize': BOXSIZE,
                   'particles': particles.size, 'resampler': 'tsc', 'interlacing': NI, 'compensate': True, 'los': 'z',
                   'ells': list(ELLS), 'kstep': STEP, 'nbins': bin.nbins, 'path': 'paint(out=complex) -> compute_mesh2_spectrum -> clone(norm, shot noise), replayed as one HIP graph',
                   'paint_path': _lib.load().mp_set_paint_path(-1)},
        'roofline': {'kernel': dominant, 'bound': 'hbm', 'achieved': d['achieved_GBs'], 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': round(d['achieved_GBs'] / HBM_PEAK_GBS, 4), 'traffic': traffic.get(dominant),
                     'algorithmic_bytes_per_launch': alg[dominant], 'avg_ms': d['avg_ms'],
                     'instruction_issue': read_issue_limit(dominant)},
        'end_to_end': {'algorithmic_GB_per_mock': round(end_to_end_bytes() / 1e9, 3), 'GBs': round(e2e_gbs, 1),
                       'frac_of_peak': round(e2e_gbs / HBM_PEAK_GBS, 4)},
        'kernels': table,
        'check': {'p0_finite': bool(np.all(np.isfinite(p0))), 'p0_first_bins': [float(x) for x in p0[1:4]],
                  'graph_equals_eager_1e-12': bool(same)},
    }
    if not args.no_batch:
        # the batched form of the metric: 64 uneven lognormal mocks at 256^3, batched vs per-mock
        del graph
        torch.cuda.empty_cache()
        rates, bfields = batch_rates(steps=3, warmup=1, seed=args.seed)
        del bfields
        torch.cuda.empty_cache()
        result['batch256'] = {'workload': '64 lognormal mocks, 256^3, boxsize 1000, counts uniform in [5e5, 4e6], '
                                          'interlacing 2, compensated, ells 0, 2, 4, k step 0.005',
                              'baseline': 'plain_calls: the reference harness per mock (paint, spectrum, normalization, '
                                          'shot noise, results read back before the next mock); per_mock: the same kernels '
                                          'issued for all mocks without host synchronisation',
                              'results': rates, 'speedup_tsc': rates['tsc']['speedup'], 'speedup_cic': rates['cic']['speedup']}
    tbuf = _lib.ctypes.create_string_buffer(1 << 14)
    _lib.load().mp_paint_tuning_report(tbuf, len(tbuf))
    result['tuning'] = tbuf.value.decode().strip().splitlines()
    if not args.no_cpu_baseline:
        print('# cpu baseline (float64 oracle on the same particles) ...', file=sys.stderr, flush=True)
        result['cpu_baseline'], result['parity'] = cpu_baseline(particles, spec, bin)
    print(json.dumps(result))


if __name__ == '__main__':
    main()
