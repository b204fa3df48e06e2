"""The C-ABI library loads and exports every entry point include/afsi_hip.h
declares (no GPU calls: the CPU suite only loads the library)."""
import ctypes
import re

import numpy as np
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / 'include' / 'afsi_hip.h'


def declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r'^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(afsi_[a-z0-9_]+)\s*\(',
                                 text, re.M)))


def test_header_declarations_are_bound():
    from afsi_gpu_amd import _abi
    names = declared()
    assert len(names) >= 40
    assert sorted(_abi.exported_names()) == names


def test_library_exports_every_declared_symbol():
    from afsi_gpu_amd import _abi
    if not _abi.LIBRARY.exists():
        pytest.fail(f'{_abi.LIBRARY} not built (run __graft_entry__.build())')
    lib = ctypes.CDLL(str(_abi.LIBRARY))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing
    lib = _abi.load()
    assert lib.afsi_abi_version() >= 1
    assert lib.afsi_reduce_partials(10**6) >= 1


def test_host_validation_without_device_work():
    # Argument validation runs on the host before any launch.
    from afsi_gpu_amd import _abi
    with pytest.raises(RuntimeError, match='invalid multigrid plan'):
        _abi.call('afsi_mg3d_cycles_f64', _abi.MGPlan3(), 1, None)
    with pytest.raises(RuntimeError, match='tile sizes must be positive'):
        g = _abi.Grid3()
        _abi.call('afsi_bin_points', None, 0, g, _abi.i32x3((0, 1, 1)), None, None, None, None, 0,
                  None, None)


def test_struct_layouts_match_the_header(tmp_path):
    """Size and field offsets of every ctypes mirror equal the C structs'
    (compiled from include/afsi_hip.h with the host compiler)."""
    import shutil
    import subprocess
    from afsi_gpu_amd import _abi
    cxx = shutil.which('g++') or shutil.which('c++')
    if cxx is None:
        pytest.skip('no host C++ compiler')
    mirrors = {'afsi_grid3': _abi.Grid3, 'afsi_grid2': _abi.Grid2, 'afsi_mg3d_plan': _abi.MGPlan3,
               'afsi_mg2d_plan': _abi.MGPlan2, 'afsi_mgb_item': _abi.MGBItem,
               'afsi_mgb_batch': _abi.MGBBatch, 'afsi_tb_item': _abi.TBItem,
               'afsi_tb_batch': _abi.TBBatch, 'afsi_fe_geometry': _abi.FEGeometry, 'afsi_csr': _abi.CSR,
               'afsi_pcg_options': _abi.PCGOptions, 'afsi_pcg_info': _abi.PCGInfo,
               'afsi_p2_group': _abi.P2Group, 'afsi_p2_tables': _abi.P2Tables, 'afsi_hash_table': _abi.HashTable}
    lines = ['#include <cstdio>', '#include <cstddef>', '#include "afsi_hip.h"', 'int main() {']
    for cname, py in mirrors.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in py._fields_:
            lines.append(f'  printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines.append('  return 0; }')
    src = tmp_path / 'layout.cpp'
    src.write_text('\n'.join(lines))
    exe = tmp_path / 'layout'
    subprocess.run([cxx, '-I', str(ROOT / 'include'), str(src), '-o', str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split('\n')
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in mirrors.items():
        assert got[cname, 'size'] == ctypes.sizeof(py), cname
        for field, _ in py._fields_:
            assert got[cname, field] == getattr(py, field).offset, (cname, field)


@pytest.mark.parametrize('shape,periodic,tile', [((4, 16, 16), True, (8, 8, 8)), ((4, 4, 4), True, (8, 8, 8)),
                                                 ((5, 7, 9), True, (4, 4, 4)), ((12, 20, 8), False, (8, 8, 8)),
                                                 ((128, 128, 128), True, (8, 8, 8)), ((6, 6, 6), True, (3, 5, 8))])
def test_tile_boxes_stay_inside_the_face_arrays(shape, periodic, tile):
    """Every face offset a tile's LDS box entry maps to (the addresses the
    tiled interpolation stages and the spread flushes) lies inside the face
    arrays, including periodic axes shorter than the tile (host-side walk of
    the device mapping, no GPU)."""
    from afsi_gpu_amd import _abi
    from afsi_gpu_amd.mac import MACGrid
    if not _abi.LIBRARY.exists():
        pytest.fail(f'{_abi.LIBRARY} not built (run __graft_entry__.build())')
    g = MACGrid(shape, (1., 1., 1.), periodic=periodic)
    lo, hi = (ctypes.c_int64 * 3)(), (ctypes.c_int64 * 3)()
    _abi.call('afsi_tile_box_bounds', g.abi(), _abi.i32x3(tile), lo, hi)
    for c in range(3):
        count = int(np.prod(g.face_shape(c)))
        assert 0 <= lo[c] and hi[c] < count, (c, lo[c], hi[c], count)
        # with the halo every face is reachable from some tile box
        assert lo[c] == 0 and hi[c] == count - 1
