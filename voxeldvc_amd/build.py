"""Build libvdvc.so (the gfx950 kernels + C ABI) in-tree with hipcc.

    python -m voxeldvc_amd.build

Each source is compiled to an object file of its own (in parallel), then the
objects are linked into one shared library.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OBJ = os.path.join(HERE, '_build')
LIB = os.path.join(HERE, 'libvdvc.so')
SOURCES = ['vdvc_util.hip', 'vdvc_voxel.hip', 'vdvc_hessian.hip', 'vdvc_node.hip', 'vdvc_prealign.hip',
           'vdvc_seednet.hip', 'vdvc_post.hip', 'vdvc_affine.hip', 'vdvc_cells.hip',
           'vdvc_capi.hip']
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
# packed-FP32 instruction selection is switched off for every device pass
FLAGS = ['--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-Wno-unused-command-line-argument', '-I', CSRC]


def _jobs():
    n = os.environ.get('MAX_JOBS')
    if n and n.isdigit():
        return max(1, min(int(n), 16))
    return max(1, min(os.cpu_count() or 1, 8))


def build(verbose=True):
    """Compile every source for gfx950 and link one shared library; raises
    on failure."""
    os.makedirs(OBJ, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(OBJ, os.path.splitext(src)[0] + '.o')
        cmd = [HIPCC] + FLAGS + ['-c', os.path.join(CSRC, src), '-o', obj]
        if verbose:
            print(' '.join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError('hipcc failed on %s:\n%s%s' % (src, r.stdout, r.stderr))
        return obj

    with ThreadPoolExecutor(_jobs()) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    cmd = [HIPCC, '--offload-arch=gfx950', '-shared', '-fPIC', '-o', LIB + '.tmp'] + objs
    if verbose:
        print(' '.join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + '.tmp', LIB)
    return LIB


if __name__ == '__main__':
    build()
    sys.exit(0)
