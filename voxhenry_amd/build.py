"""Build libvoxhenry.so (HIP for gfx950) in-tree.

    python -m voxhenry_amd.build          # or __graft_entry__.build()

Every source under csrc/ is compiled by hipcc into _build/ and linked into
voxhenry_amd/libvoxhenry.so.  Packed FP32 instructions are disabled
(-packed-fp32-ops) for all device code.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
BUILD = os.path.join(HERE, '..', '_build')
LIB = os.path.join(HERE, 'libvoxhenry.so')
ARCH = os.environ.get('VH_OFFLOAD_ARCH', 'gfx950')

COMMON = ['-O3', '-std=c++17', '-fPIC', '--offload-arch=' + ARCH,
          '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
          '-Wno-unused-result', '-I', os.path.join(HERE, '..', 'include')]


def hipcc():
    for c in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', 'hipcc'):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return 'hipcc'


def sources():
    # the FFT-pass instantiation units are the slowest to compile: start them first
    fs = [f for f in os.listdir(CSRC) if f.endswith(('.hip', '.cpp'))]
    return sorted(fs, key=lambda f: (not f.startswith(('vh_matvec_xy', 'vh_matvec_z')), f))


def _up_to_date(obj, src):
    """VH_INCREMENTAL=1 (development only): skip a source whose object is newer
    than it and than every header of csrc/ and include/."""
    if os.environ.get('VH_INCREMENTAL') != '1' or not os.path.exists(obj):
        return False
    inc = os.path.join(HERE, '..', 'include')
    deps = [os.path.join(CSRC, src)] + [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith('.h')] + \
        [os.path.join(inc, f) for f in os.listdir(inc)]
    return os.path.getmtime(obj) > max(os.path.getmtime(d) for d in deps)


def compile_one(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + '.o')
    if _up_to_date(obj, src):
        return obj
    cmd = [hipcc(), '-c'] + COMMON + [os.path.join(CSRC, src), '-o', obj]
    if src.endswith('.cpp'):
        cmd = [hipcc(), '-c', '-x', 'hip'] + COMMON + [os.path.join(CSRC, src), '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('hipcc failed for %s:\n%s\n%s' % (src, ' '.join(cmd), r.stderr[-6000:]))
    return obj


def build(verbose=True, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, len(srcs))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = LIB + '.tmp'
    cmd = [hipcc(), '-shared', '-fPIC', '--offload-arch=' + ARCH, '-o', tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('link failed:\n%s\n%s' % (' '.join(cmd), r.stderr[-4000:]))
    os.replace(tmp, LIB)
    if verbose:
        print('built', LIB)
    return LIB


if __name__ == '__main__':
    build()
    sys.exit(0)
