"""Builds libyaff_nb.so (HIP, gfx950) in-tree with hipcc.

    python -m yaff_amd.build

-ffp-contract=off keeps the FP64 restatement of Yaff's arithmetic free of
fused multiply-adds (explicit fma() calls are still fused); packed FP32
instructions are disabled for the whole library.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
SOURCES = ['nlist.hip', 'pair.hip', 'ewald.hip', 'md.hip', 'tables.hip', 'ewald_plan.cpp']
LIB = os.path.join(HERE, 'libyaff_nb.so')
ARCH = os.environ.get('YNB_ARCH', 'gfx950')


def hipcc():
    for cand in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', 'hipcc'):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError('hipcc not found')


def flags():
    return ['--offload-arch=%s' % ARCH, '-O3', '-std=c++17', '-fPIC', '-shared', '-ffp-contract=off',
            '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops', '-Wno-unused-result',
            '-I' + os.path.join(HERE, '..', 'include')]


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + '.tmp'
    cmd = [hipcc()] + flags() + ['-o', tmp] + srcs
    if verbose:
        print(' '.join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, universal_newlines=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        raise RuntimeError('hipcc failed building %s' % LIB)
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose=True))
