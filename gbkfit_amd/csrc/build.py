"""Builds gbkfit_amd/libgbkhip.so from the HIP sources in this directory.

    python gbkfit_amd/csrc/build.py [-j N]

Every source is compiled for gfx950 only, with packed FP32 instructions
disabled, then linked into one shared library next to the package.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
BUILD = os.path.join(PKG, '_build')
OUT = os.path.join(PKG, 'libgbkhip.so')

SOURCES = ['api.hip', 'smdisk.hip', 'mcdisk.hip', 'fft.hip', 'sepconv.hip',
           'dmodel.hip',
           'objective.hip', 'batch.hip']

FLAGS = ['--offload-arch=gfx950', '-O3', '-fPIC', '-std=c++17',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-Wall', '-Wno-unused-result', '-Wno-unknown-pragmas',
         '-Werror=uninitialized', '-Werror=sometimes-uninitialized',
         '-I' + os.path.join(ROOT, 'include')]


def hipcc():
    exe = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    if not os.path.exists(exe):
        raise RuntimeError('hipcc not found')
    return exe


def compile_one(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + '.o')
    cmd = [hipcc(), *FLAGS, '-c', os.path.join(HERE, src), '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build(jobs=8, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(HERE, s))]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(compile_one, srcs))
    if verbose:
        for _, err in results:
            if err.strip():
                print(err, file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = OUT + '.tmp'
    cmd = [hipcc(), '--offload-arch=gfx950', '-shared', '-fPIC', '-o', tmp,
           *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == '__main__':
    ap = argparse.ArgumentParser()
    ap.add_argument('-j', type=int, default=8)
    ap.add_argument('-v', action='store_true')
    a = ap.parse_args()
    print(build(a.j, a.v))
