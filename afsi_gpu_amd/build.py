"""Build ``libafsi_hip.so`` (gfx950) in-tree with hipcc.

Every source under ``csrc/`` is compiled to an object in parallel and linked
into one shared library next to this file. Flags:

* ``-ffp-contract=off``: no multiply-add contraction, so elementwise results
  are bit-identical to the FP64 oracle's IEEE operation sequence (the
  reference disables contraction the same way, mac/cuda_ib.py:58).
* ``-munsafe-fp-atomics``: hardware FP64/FP32 atomics on coarse-grained
  device memory (the only memory the library writes).
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 math.
"""
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path
import os
import shutil
import subprocess
import sys

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
CSRC = HERE / 'csrc'
INCLUDE = ROOT / 'include'
LIBRARY = HERE / 'libafsi_hip.so'
BUILD = HERE / '_build'
ARCH = os.environ.get('AFSI_OFFLOAD_ARCH', 'gfx950')

FLAGS = ['-O3', f'--offload-arch={ARCH}', '-fPIC', '-std=c++17', '-ffp-contract=off',
         '-munsafe-fp-atomics', '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-Wno-unused-command-line-argument', f'-I{INCLUDE}', f'-I{CSRC}']


def hipcc():
    path = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    if not Path(path).exists():
        raise RuntimeError('hipcc not found; install ROCm or put hipcc on PATH')
    return path


def sources():
    return sorted(CSRC.glob('*.hip'))


def _compile(src):
    obj = BUILD / (src.stem + '.o')
    cmd = [hipcc(), *FLAGS, '-c', str(src), '-o', str(obj)]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode:
        raise RuntimeError(f'hipcc failed for {src.name}:\n{result.stderr}')
    return obj


def build(verbose=False):
    """Compile every source and link ``libafsi_hip.so``; always rebuilds."""
    BUILD.mkdir(exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(len(srcs), os.cpu_count() or 4)) as pool:
        objs = list(pool.map(_compile, srcs))
    tmp = LIBRARY.with_suffix('.so.tmp')
    cmd = [hipcc(), f'--offload-arch={ARCH}', '-shared', '-fPIC', *map(str, objs), '-o', str(tmp)]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode:
        raise RuntimeError(f'link failed:\n{result.stderr}')
    os.replace(tmp, LIBRARY)
    if verbose:
        print(f'built {LIBRARY} from {len(srcs)} sources')
    return LIBRARY


if __name__ == '__main__':
    build(verbose=True)
    sys.exit(0)
