"""Build librtcd_mi355.so (HIP, gfx950) in-tree.

hipcc compiles every csrc/*.hip file for gfx950 (CDNA4, wave64), g++ the host-only csrc/*.cpp
files, and hipcc links one shared library next to this file.  Packed-FP32 instructions are disabled
(-packed-fp32-ops) and FMA contraction is left to the per-function pragmas of the FP64 decision code.
"""
import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OUT = os.path.join(HERE, 'librtcd_mi355.so')
BUILD = os.path.join(HERE, '_build')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
FLAGS = ['--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-Wall', '-Wno-unused-function',
         '-Wno-unused-variable', '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-I' + os.path.join(os.path.dirname(HERE), 'include')]
CXX = os.environ.get('CXX', 'g++')
CXX_FLAGS = ['-O2', '-std=c++17', '-fPIC', '-Wall', '-I' + os.path.join(os.path.dirname(HERE), 'include')]


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, '*.hip'))) + sorted(glob.glob(os.path.join(CSRC, '*.cpp')))

    def compile_one(src):
        obj = os.path.join(BUILD, os.path.basename(src) + '.o')
        if src.endswith('.cpp'):
            cmd = [CXX] + CXX_FLAGS + ['-c', src, '-o', obj]
        else:
            cmd = [HIPCC] + FLAGS + ['-c', src, '-o', obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f'{cmd[0]} failed for {src}:\n{r.stderr}')
        if verbose and r.stderr.strip():
            print(r.stderr)
        return obj

    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = OUT + '.tmp'
    subprocess.check_call([HIPCC, '--offload-arch=gfx950', '-shared', '-fPIC', '-o', tmp] + objs)
    os.replace(tmp, OUT)
    return OUT


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
