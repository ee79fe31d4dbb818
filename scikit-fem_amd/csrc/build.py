"""Build libfemhip.so for gfx950 with hipcc (no CMake, no JIT cache).

    python scikit-fem_amd/csrc/build.py [-j N]

Each .hip file is compiled to an object in ``_build/`` and linked into
``scikit-fem_amd/femhip/libfemhip.so`` (in-tree, so it travels with the
repository snapshot).  ``exact_*.hip`` are compiled with
``-ffp-contract=off``: its kernels must reproduce NumPy's rounding.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.normpath(os.path.join(HERE, '..', 'femhip'))
OUT = os.path.join(PKG, 'libfemhip.so')
BUILD = os.path.normpath(os.path.join(HERE, '..', '..', '_build'))
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')

COMMON = ['-O3', '-std=c++17', '-fPIC', '--offload-arch=gfx950',
          '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
          '-Wall', '-Wno-unused-function', '-Wno-unused-variable']
PER_FILE = {
    'exact_local.hip': ['-ffp-contract=off'],
    'exact_tri.hip': ['-ffp-contract=off'],
    'exact_tet.hip': ['-ffp-contract=off'],
    'exact_hex.hip': ['-ffp-contract=off'],
    'interpolate.hip': ['-ffp-contract=off'],
    'facet.hip': ['-ffp-contract=off'],
    'coords.hip': ['-ffp-contract=off'],
}


def sources():
    return sorted(f for f in os.listdir(HERE) if f.endswith('.hip') or f.endswith('.cpp'))


def compile_one(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + '.o')
    cmd = [HIPCC] + COMMON + PER_FILE.get(src, []) + ['-c', os.path.join(HERE, src), '-o', obj]
    if src.endswith('.cpp'):
        cmd = ['g++', '-O3', '-std=c++17', '-fPIC', '-Wall', '-pthread', '-c',
               os.path.join(HERE, src), '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('compile failed: %s\n%s\n%s' % (' '.join(cmd), r.stdout, r.stderr))
    return obj


def build(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = OUT + '.tmp'
    cmd = [HIPCC, '-shared', '-fPIC', '--offload-arch=gfx950', '-pthread', '-o', tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('link failed: %s\n%s' % (r.stdout, r.stderr))
    os.replace(tmp, OUT)
    return OUT


if __name__ == '__main__':
    ap = argparse.ArgumentParser()
    ap.add_argument('-j', type=int, default=min(8, os.cpu_count() or 1))
    a = ap.parse_args()
    print(build(a.j))
