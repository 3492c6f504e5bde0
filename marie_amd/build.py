"""Builds libmarie_vie.so (all HIP sources of the package) for gfx950 in-tree.

Usage: python -m marie_amd.build  (or marie_amd.build.build()).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libmarie_vie.so')
SOURCES = ['vie_capi.hip', 'vie_mvp.hip', 'vie_mvp_reg_f64.hip', 'vie_mvp_reg_f32.hip', 'vie_krylov.hip',
           'vie_assembly.hip', 'vie_block.hip', 'vie_adjoint.hip', 'vie_excitation.hip', 'vie_bicgstab.hip', 'vie_directfn.hip',
           'vie_sstep.hip']

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
FLAGS = ['--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC',
         # packed FP32 instructions are kept out of the generated code
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-Wno-unused-result', '-Wno-unknown-warning-option']


def build(verbose=False, jobs=6):
    objs = []
    procs = []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(CSRC, s.replace('.hip', '.o'))
        cmd = [HIPCC] + FLAGS + ['-c', src, '-o', obj]
        if verbose:
            print(' '.join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    for p in procs:
        _wait(p)
    tmp = LIB + '.tmp'
    cmd = [HIPCC, '--offload-arch=gfx950', '-shared', '-fPIC', '-o', tmp] + objs
    if verbose:
        print(' '.join(cmd), flush=True)
    out = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if out.returncode:
        raise RuntimeError('link failed:\n' + out.stdout.decode())
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


def _wait(item):
    name, p = item
    out, _ = p.communicate()
    if p.returncode:
        raise RuntimeError('hipcc failed on %s:\n%s' % (name, out.decode()))


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
