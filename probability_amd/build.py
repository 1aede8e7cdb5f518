"""Builds libhipmc.so in-tree with hipcc for gfx950.

The library is linked from several objects compiled in parallel: the ABI
and small kernels (hipmc.hip), the BF16 MFMA path (mfma.hip) and the
sampling kernels, one object per (kernel group, dtype, target family) from
inst.hip.

Usage: python -m probability_amd.build   (or probability_amd.build.build()).
"""

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libhipmc.so')
OBJ_DIR = os.path.join(HERE, '_objs')

HIPCC_FLAGS = [
    '--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC',
    # Packed FP32 VALU ops are kept out of every code object.
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-result',
]

_FAMILIES = {'hier': 0, 'glm': 1, 'test': 2}


def objects():
  """(object name, source, extra defines) of every object of the library."""
  out = [('hipmc', 'hipmc.hip', []), ('mfma', 'mfma.hip', [])]
  for group, nuts in (('nuts', 1), ('hmc', 0)):
    for t in ('float', 'double'):
      for fam, code in _FAMILIES.items():
        out.append((f'{group}_{t[0]}{"32" if t == "float" else "64"}_{fam}',
                    'inst.hip',
                    [f'-DHIPMC_INST_T={t}', f'-DHIPMC_INST_FAM={code}',
                     f'-DHIPMC_INST_NUTS={nuts}']))
  return out


def hipcc():
  for cand in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', 'hipcc'):
    if cand and (os.path.sep not in cand or os.path.exists(cand)):
      return cand
  raise RuntimeError('hipcc not found')


def _compile(name, src, defs, extra_flags, verbose):
  obj = os.path.join(OBJ_DIR, name + '.o')
  cmd = [hipcc()] + HIPCC_FLAGS + list(extra_flags) + defs + [
      '-I', os.path.join(ROOT, 'include'), '-c', os.path.join(CSRC, src),
      '-o', obj + '.tmp']
  if verbose:
    print(' '.join(cmd), flush=True)
  proc = subprocess.run(cmd, capture_output=True, text=True)
  if proc.returncode != 0:
    return name, proc.stdout + proc.stderr
  os.replace(obj + '.tmp', obj)
  return name, None


def build(verbose=False, extra_flags=(), jobs=None, only=None):
  """Compiles every object (or only those named in `only`, reusing the
  others' objects) and links libhipmc.so."""
  os.makedirs(OBJ_DIR, exist_ok=True)
  objs = objects()
  todo = [o for o in objs if only is None or o[0] in only]
  jobs = jobs or min(len(todo), max(1, min(16, os.cpu_count() or 1)))
  errors = []
  with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
    futs = [ex.submit(_compile, n, s, d, extra_flags, verbose)
            for n, s, d in todo]
    for f in futs:
      name, err = f.result()
      if err:
        errors.append((name, err))
  if errors:
    for name, err in errors:
      sys.stderr.write(f'--- {name}\n{err}\n')
    raise RuntimeError('hipcc failed for ' + ', '.join(n for n, _ in errors))
  cmd = [hipcc(), '--offload-arch=gfx950', '-shared', '-fPIC', '-o',
         LIB + '.tmp'] + [os.path.join(OBJ_DIR, n + '.o') for n, _, _ in objs]
  if verbose:
    print(' '.join(cmd), flush=True)
  proc = subprocess.run(cmd, capture_output=True, text=True)
  if proc.returncode != 0:
    sys.stderr.write(proc.stdout + proc.stderr)
    raise RuntimeError('link failed for libhipmc.so')
  os.replace(LIB + '.tmp', LIB)
  return LIB


if __name__ == '__main__':
  print(build(verbose=True, only=sys.argv[1:] or None))
