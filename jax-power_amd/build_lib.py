"""Build libjaxpower_amd.so (the C-ABI HIP library) in-tree for gfx950.

Usage: python jax-power_amd/build_lib.py [--jobs N]
Compiles every csrc/*.hip with hipcc and links jaxpower_amd/lib/libjaxpower_amd.so.
"""
import argparse
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OBJDIR = os.path.join(HERE, '_build')
LIBDIR = os.path.join(HERE, 'jaxpower_amd', 'lib')
LIBNAME = 'libjaxpower_amd.so'
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ARCH = os.environ.get('JAXPOWER_AMD_ARCH', 'gfx950')

CFLAGS = [
    '-O3', '-std=c++17', '-fPIC', f'--offload-arch={ARCH}', '-mcode-object-version=5',
    '-munsafe-fp-atomics', '-Wall', '-Wno-unused-function', '-Wno-unused-variable',
    # packed FP32 instructions stay off on this target
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-fvisibility=hidden',
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith('.hip'))


def compile_one(src):
    obj = os.path.join(OBJDIR, os.path.basename(src)[:-4] + '.o')
    cmd = [HIPCC, *CFLAGS, '-c', src, '-o', obj]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f'compilation of {src} failed:\n{" ".join(cmd)}\n{proc.stdout}\n{proc.stderr}')
    warn = '\n'.join(line for line in proc.stderr.splitlines() if 'packed-fp32-ops' not in line and line.strip())
    return obj, warn


def build(jobs=8, verbose=False):
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sources()
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(compile_one, srcs))
    objs = [r[0] for r in results]
    if verbose:
        for src, (_, warn) in zip(srcs, results):
            if warn:
                print(f'--- {os.path.basename(src)}\n{warn}')
    out = os.path.join(LIBDIR, LIBNAME)
    tmp = out + '.tmp'
    cmd = [HIPCC, '-shared', f'--offload-arch={ARCH}', '-fPIC', '-o', tmp, *objs]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f'link failed:\n{" ".join(cmd)}\n{proc.stdout}\n{proc.stderr}')
    os.replace(tmp, out)
    return out


if __name__ == '__main__':
    parser = argparse.ArgumentParser()
    parser.add_argument('--jobs', type=int, default=8)
    parser.add_argument('-v', '--verbose', action='store_true')
    args = parser.parse_args()
    try:
        print(build(jobs=args.jobs, verbose=args.verbose))
    except RuntimeError as exc:
        print(exc, file=sys.stderr)
        sys.exit(1)
