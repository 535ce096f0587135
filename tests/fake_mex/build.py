"""Build the MEX gateway (voxhenry_amd/mex/vh_mex.c) against the fake MEX host
into tests/fake_mex/libvh_mex_fake.so, linked to the in-tree libvoxhenry.so.

    python tests/fake_mex/build.py
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(HERE, 'libvh_mex_fake.so')
SOURCES = [os.path.join(ROOT, 'voxhenry_amd', 'mex', 'vh_mex.c'), os.path.join(HERE, 'fake_mex.c')]


def build(verbose=True):
    vh = os.path.join(ROOT, 'voxhenry_amd')
    if not os.path.exists(os.path.join(vh, 'libvoxhenry.so')):
        raise RuntimeError('libvoxhenry.so not built: run `python -m voxhenry_amd.build` first')
    tmp = LIB + '.tmp'
    cmd = ['gcc', '-std=c11', '-O2', '-g', '-Wall', '-Wextra', '-Werror', '-Wno-unused-parameter', '-shared', '-fPIC',
           '-I', HERE, '-I', os.path.join(ROOT, 'include')] + SOURCES + \
          ['-o', tmp, '-L', vh, '-lvoxhenry', '-Wl,-rpath,$ORIGIN/../../voxhenry_amd', '-lm']
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('gcc failed:\n%s\n%s' % (' '.join(cmd), r.stderr[-4000:]))
    os.replace(tmp, LIB)
    if verbose:
        print('built', LIB)
    return LIB


if __name__ == '__main__':
    build()
