"""Build libvdvc.so (the gfx950 kernels + C ABI) in-tree with hipcc.

    python -m voxeldvc_amd.build            # the library
    python -m voxeldvc_amd.build --probe    # also the HBM streaming probe
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libvdvc.so')
SOURCES = ['vdvc_voxel.hip', 'vdvc_hessian.hip', 'vdvc_node.hip', 'vdvc_prealign.hip',
           'vdvc_seednet.hip', 'vdvc_post.hip', 'vdvc_affine.hip', 'vdvc_cells.hip',
           'vdvc_capi.hip']
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')


def build(verbose=True):
    """Compile every source for gfx950 into one shared library; raises on
    failure.  Packed-FP32 instruction selection is switched off."""
    cmd = [HIPCC, '--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-shared',
           '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
           '-Wno-unused-command-line-argument',
           '-I', CSRC, '-o', LIB + '.tmp'] + [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(' '.join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + '.tmp', LIB)
    return LIB


def build_probe(verbose=True):
    """The stand-alone HBM streaming-bandwidth probe (probe_hbm_stream)."""
    out = os.path.join(HERE, 'probe_hbm_stream')
    cmd = [HIPCC, '--offload-arch=gfx950', '-O3', '-std=c++17',
           '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
           '-o', out, os.path.join(CSRC, 'probe_hbm_stream.hip')]
    if verbose:
        print(' '.join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return out


if __name__ == '__main__':
    build()
    if '--probe' in sys.argv[1:]:
        build_probe()
    sys.exit(0)
