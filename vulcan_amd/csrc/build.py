"""Builds libvulcan_amd.so (gfx950) in-tree with hipcc.

    python vulcan_amd/csrc/build.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
SOURCES = ['vk_chem.hip', 'vk_solve.hip', 'vk_ros2.hip', 'vk_mlp.hip']
TARGET = os.path.join(PKG, 'libvulcan_amd.so')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')

FLAGS = [
    '-O3', '--offload-arch=gfx950', '-fPIC', '-shared', '-std=c++17',
    # the kernels follow the reference's expression trees operation by
    # operation; fused multiply-adds are written explicitly where wanted
    '-ffp-contract=off',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-I' + os.path.join(ROOT, 'include'),
]


def build(verbose=False):
    srcs = [os.path.join(HERE, s) for s in SOURCES]
    tmp = TARGET + '.tmp'
    cmd = [HIPCC] + FLAGS + srcs + ['-o', tmp]
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError('hipcc failed building libvulcan_amd.so')
    os.replace(tmp, TARGET)
    return TARGET


if __name__ == '__main__':
    print(build(verbose=True))
