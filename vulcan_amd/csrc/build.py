"""Builds libvulcan_amd.so (gfx950) in-tree with hipcc, and its bounds-checked
twin libvulcan_amd_checked.so.

    python vulcan_amd/csrc/build.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
SOURCES = ['vk_solve.hip', 'vk_factor_reg.hip', 'vk_factor_blk1.hip', 'vk_factor_blk2.hip', 'vk_factor_wave.hip', 'vk_bcr.hip',
           'vk_chem.hip', 'vk_ros2.hip', 'vk_mlp.hip', 'vk_photo.hip']
TARGET = os.path.join(PKG, 'libvulcan_amd.so')
# the same library with the chemistry kernels' table indexes bounds-checked
# (VK_CHECKED: vk_chem.hip compiled a second time, the other objects shared)
CHECKED_SOURCE = 'vk_chem.hip'
TARGET_CHECKED = os.path.join(PKG, 'libvulcan_amd_checked.so')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')

FLAGS = [
    '-O3', '--offload-arch=gfx950', '-fPIC', '-std=c++17',
    # the kernels follow the reference's expression trees operation by
    # operation; fused multiply-adds are written explicitly where wanted
    '-ffp-contract=off',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-I' + os.path.join(ROOT, 'include'),
]


def build(verbose=False):
    """Compiles each source to an object in parallel (the translation units
    are independent: no relocatable device code), then links the library."""
    objdir = os.path.join(HERE, '_build')
    os.makedirs(objdir, exist_ok=True)
    procs = []
    units = [(src, os.path.join(objdir, src.replace('.hip', '.o')), []) for src in SOURCES]
    units.append((CHECKED_SOURCE, os.path.join(objdir, CHECKED_SOURCE.replace('.hip', '_checked.o')), ['-DVK_CHECKED']))
    for src, obj, extra in units:
        cmd = [HIPCC] + FLAGS + extra + ['-c', os.path.join(HERE, src), '-o', obj]
        if verbose:
            print(' '.join(cmd), flush=True)
        procs.append((src, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = []
    for src, obj, pr in procs:
        out, _ = pr.communicate()
        if pr.returncode != 0:
            sys.stderr.write(out)
            failed.append(src)
    if failed:
        raise RuntimeError('hipcc failed compiling ' + ', '.join(failed))
    normal = [o for _, o, _ in procs[:len(SOURCES)]]
    checked = [o for (src, o, _) in zip(SOURCES, normal, procs) if src != CHECKED_SOURCE] + [procs[-1][1]]
    for target, objs in ((TARGET, normal), (TARGET_CHECKED, checked)):
        tmp = target + '.tmp'
        cmd = [HIPCC, '--offload-arch=gfx950', '-fPIC', '-shared'] + objs + ['-o', tmp]
        if verbose:
            print(' '.join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError('hipcc failed linking ' + os.path.basename(target))
        os.replace(tmp, target)
    return TARGET


if __name__ == '__main__':
    print(build(verbose=True))
