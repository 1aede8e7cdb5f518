"""Builds libhipmc.so in-tree with hipcc for gfx950.

Usage: python -m probability_amd.build   (or probability_amd.build.build()).
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libhipmc.so')

SOURCES = ['hipmc.hip']
DEPS = ['hipmc.hip', 'common.h', 'targets.h', 'kernels.h']

HIPCC_FLAGS = [
    '--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-shared',
    # Packed FP32 VALU ops are kept out of every code object.
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-result',
]


def hipcc():
  for cand in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', 'hipcc'):
    if cand and (os.path.sep not in cand or os.path.exists(cand)):
      return cand
  raise RuntimeError('hipcc not found')


def build(verbose=False, extra_flags=()):
  srcs = [os.path.join(CSRC, s) for s in SOURCES]
  cmd = [hipcc()] + HIPCC_FLAGS + list(extra_flags) + [
      '-I', os.path.join(ROOT, 'include'), '-o', LIB + '.tmp'] + srcs
  if verbose:
    print(' '.join(cmd), flush=True)
  proc = subprocess.run(cmd, capture_output=True, text=True)
  if proc.returncode != 0:
    sys.stderr.write(proc.stdout + proc.stderr)
    raise RuntimeError('hipcc failed for libhipmc.so')
  os.replace(LIB + '.tmp', LIB)
  return LIB


if __name__ == '__main__':
  print(build(verbose=True))
