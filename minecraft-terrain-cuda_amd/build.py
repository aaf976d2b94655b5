"""Build the gfx950 library (libmct.so) and its host emulation (libmct_host.so), in-tree.

  python build.py            codegen from data/*.json, then hipcc (device) and g++ (emulation,
                             plain and with operation counting: libmct_count.so)
  python build.py --no-emu   skip the host emulation libraries
  python build.py --remarks F  also write the compiler's per-kernel resource usage to F

The exactness flags are part of the semantics (Java arithmetic): no FMA contraction, no
fast-math, denormals kept, correctly rounded f32 division and square root; packed FP32
instruction selection is disabled.
"""
import argparse
import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
IR = os.path.join(HERE, 'data', 'synthetic_overworld_ir.json')
TERRAIN = os.path.join(HERE, 'data', 'synthetic_overworld_terrain.json')
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libmct.so')
EMU = os.path.join(HERE, 'libmct_host.so')
COUNT = os.path.join(HERE, 'libmct_count.so')

HIP_FLAGS = ['--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-shared',
             '-ffp-contract=off', '-fno-fast-math', '-fno-gpu-flush-denormals-to-zero',
             '-fhip-fp32-correctly-rounded-divide-sqrt',
             '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
             '-Wno-unused-command-line-argument']
HOST_FLAGS = ['-std=c++17', '-O2', '-fPIC', '-shared', '-pthread', '-ffp-contract=off', '-fno-fast-math']


def run(cmd, log=None):
    t = time.time()
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        sys.stderr.write(p.stdout[-20000:])
        raise SystemExit('build step failed: %s' % ' '.join(cmd[:3]))
    if log:
        print('%s (%.0f s)' % (log, time.time() - t), flush=True)
    return p.stdout


def codegen(ir=IR, terrain=TERRAIN):
    run([sys.executable, os.path.join(HERE, 'codegen', 'compile_ir.py'), ir, terrain, os.path.join(CSRC, 'gen')], 'codegen')


def build_device(out=LIB, extra=(), remarks=None):
    hipcc = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    tmp = out + '.tmp'
    extra = list(extra) + (['-Rpass-analysis=kernel-resource-usage'] if remarks else [])
    text = run([hipcc] + HIP_FLAGS + extra + [os.path.join(CSRC, 'mct.hip'), '-o', tmp], 'hipcc gfx950 -> %s' % os.path.basename(out))
    if remarks:
        with open(remarks, 'w') as fh:
            fh.write(text)
    os.replace(tmp, out)


def build_emu(out=EMU, extra=()):
    cxx = shutil.which('g++') or 'c++'
    tmp = out + '.tmp'
    run([cxx] + HOST_FLAGS + list(extra) + [os.path.join(CSRC, 'host_emu.cpp'), '-o', tmp], 'g++ host emulation -> %s' % os.path.basename(out))
    os.replace(tmp, out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--no-emu', action='store_true')
    ap.add_argument('--no-device', action='store_true')
    ap.add_argument('--remarks', metavar='FILE', help='write the per-kernel register / scratch / occupancy report to FILE')
    a = ap.parse_args()
    codegen()
    if not a.no_device:
        build_device(remarks=a.remarks)
    if not a.no_emu:
        build_emu()
        # the same emulation counting the operations of every evaluated scope (the per-scope
        # counts the code generator emits), for the algorithmic operations per stage
        build_emu(COUNT, ['-DMCT_COUNT_OPS'])


if __name__ == '__main__':
    main()
