"""ctypes binding of ``libafsi_hip.so`` (the C ABI declared in include/afsi_hip.h).

The library is loaded from this package directory only; there is no CPU
fallback: if it is missing every call raises. Device pointers come from
``Tensor.data_ptr()`` and kernels run on ``torch.cuda.current_stream()``.
The library binds to the HIP runtime torch already loaded (same SONAME), so
torch streams and allocations are valid handles for it.
"""
import ctypes
from ctypes import POINTER, Structure, c_char_p, c_double, c_int, c_int32, c_int64, c_size_t, c_uint32, c_void_p
from pathlib import Path
import threading

import torch

LIBRARY = Path(__file__).resolve().parent / 'libafsi_hip.so'
_lock = threading.Lock()
_lib = None

MAX_LEVELS = 16


class Grid3(Structure):
    _fields_ = [('n', c_int64 * 3), ('h', c_double * 3), ('origin', c_double * 3),
                ('center_origin', c_double * 3), ('periodic', c_int32), ('reserved', c_int32)]


class MGPlan3(Structure):
    _fields_ = [('levels', c_int32), ('periodic', c_int32), ('smoother', c_int32),
                ('nu', c_int32), ('omega', c_double),
                ('n', (c_int64 * 3) * MAX_LEVELS), ('h2', (c_double * 3) * MAX_LEVELS),
                ('inv_h2', (c_double * 3) * MAX_LEVELS),
                ('p', c_void_p * MAX_LEVELS), ('rhs', c_void_p * MAX_LEVELS),
                ('tmp', c_void_p * MAX_LEVELS), ('coarse_inverse', c_void_p),
                ('coarse_cells', c_int64), ('reduce_partials', c_void_p),
                ('reduce_result', c_void_p), ('partials_capacity', c_int64),
                ('zero_initial', c_int32), ('fuse_cells', c_int32),
                ('sweep_tile', c_int32 * MAX_LEVELS), ('sweep_blocks', c_int32 * MAX_LEVELS),
                ('relax', c_double)]


class TBItem(Structure):
    _fields_ = [('x', c_void_p), ('force', c_void_p), ('face', c_void_p * 3), ('out', c_void_p),
                ('points', c_int64), ('grid', Grid3)]


class TBBatch(Structure):
    _fields_ = [('count', c_int32), ('tile', c_int32 * 3), ('items', POINTER(TBItem)),
                ('keys_sorted', c_void_p), ('perm', c_void_p), ('cell_ptr', c_void_p),
                ('workspace', c_void_p), ('workspace_bytes', c_size_t), ('flags', c_void_p)]


class MGBItem(Structure):
    _fields_ = [('plan', MGPlan3), ('x', c_void_p * 2), ('b', c_void_p)]


class MGBBatch(Structure):
    _fields_ = [('count', c_int32), ('inner_cycles', c_int32), ('items', POINTER(MGBItem)),
                ('workspace', c_void_p), ('workspace_bytes', c_size_t), ('active', c_void_p),
                ('tol', c_void_p), ('results', c_void_p), ('partials', c_void_p)]


def cast_items(array):
    return ctypes.cast(array, POINTER(MGBItem))


class Grid2(Structure):
    _fields_ = [('n', c_int64 * 2), ('h', c_double * 2), ('lengths', c_double * 2),
                ('periodic', c_int32), ('reserved', c_int32)]


class MGPlan2(Structure):
    _fields_ = [('levels', c_int32), ('periodic', c_int32), ('smoother', c_int32),
                ('nu', c_int32), ('omega', c_double),
                ('n', (c_int64 * 2) * MAX_LEVELS), ('h2', (c_double * 2) * MAX_LEVELS),
                ('inv_h2', (c_double * 2) * MAX_LEVELS),
                ('p', c_void_p * MAX_LEVELS), ('rhs', c_void_p * MAX_LEVELS),
                ('tmp', c_void_p * MAX_LEVELS), ('coarse_inverse', c_void_p),
                ('coarse_cells', c_int64), ('reduce_partials', c_void_p),
                ('reduce_result', c_void_p), ('partials_capacity', c_int64)]




class CHBItem(Structure):
    _fields_ = [('x', c_void_p), ('points', c_int64), ('grid', Grid2), ('plan', MGPlan2)]


class CHBBatch(Structure):
    _fields_ = [('count', c_int32), ('reserved', c_int32), ('items', POINTER(CHBItem)),
                ('workspace', c_void_p), ('workspace_bytes', c_size_t)]


class IB3BItem(Structure):
    _fields_ = [('velocity', c_void_p * 3), ('force', c_void_p * 3), ('rhs', c_void_p),
                ('pressure', c_void_p * 2), ('x', c_void_p), ('point_velocity', c_void_p),
                ('points', c_int64), ('grid', Grid3)]


class IB3BBatch(Structure):
    _fields_ = [('count', c_int32), ('reserved', c_int32), ('items', POINTER(IB3BItem)),
                ('workspace', c_void_p), ('workspace_bytes', c_size_t)]

class FEGeometry(Structure):
    _fields_ = [('cells', c_void_p), ('values', c_void_p), ('weights', c_void_p), ('elements', c_int64),
                ('nodes_per_cell', c_int32), ('quad', c_int32), ('node_count', c_int64),
                ('node_ptr', c_void_p), ('node_items', c_void_p)]


class CSR(Structure):
    _fields_ = [('row_ptr', c_void_p), ('cols', c_void_p), ('vals', c_void_p), ('rows', c_int64),
                ('nnz', c_int64)]


class PCGOptions(Structure):
    _fields_ = [('rtol', c_double), ('atol', c_double), ('max_iterations', c_int32),
                ('recompute_every', c_int32), ('check_every', c_int32), ('reserved', c_int32)]


class PCGInfo(Structure):
    _fields_ = [('status', c_int32), ('iterations', c_int32), ('residual_norm', c_double),
                ('rhs_norm', c_double), ('tolerance', c_double)]


class P2Group(Structure):
    _fields_ = [('cells', c_void_p), ('values', c_void_p), ('weights', c_void_p), ('elements', c_int64),
                ('quad', c_int64), ('node_count', c_int64)]


class P2Tables(Structure):
    _fields_ = [('base', c_void_p), ('phi', c_void_p), ('total', c_int64), ('offset', c_int64)]


class HashTable(Structure):
    _fields_ = [('keys', c_void_p), ('values', c_void_p), ('capacity', c_int64), ('flag', c_void_p)]

_P = c_void_p
_I64, _F64 = POINTER(c_int64), POINTER(c_double)
_SIGNATURES = {
    'afsi_last_error': (c_char_p, []),
    'afsi_abi_version': (c_int, []),
    'afsi_prepare_compact_f64': (c_int, [_P, c_int64, POINTER(Grid3), _P, _P, _P, _P]),
    'afsi_compact_gather_f64': (c_int, [_P, _P, _P, c_int64, c_int64, c_int64, c_int64, _P, c_int64, _P, _P]),
    'afsi_compact_gather_f32': (c_int, [_P, _P, _P, c_int64, c_int64, c_int64, c_int64, _P, c_int64, _P, _P]),
    'afsi_compact_spread_workspace_bytes': (c_size_t, [c_int64, c_int64, c_int64, c_int64]),
    'afsi_compact_spread_f64': (c_int, [_P, c_int64, _P, _P, c_int64, c_int64, c_int64, c_int64, c_double, _P, _P,
                                        c_size_t, _P, _P]),
    'afsi_compact_spread_f32': (c_int, [_P, c_int64, _P, _P, c_int64, c_int64, c_int64, c_int64, c_double, _P, _P,
                                        c_size_t, _P, _P]),
    'afsi_indexed_gather_f64': (c_int, [_P, c_int64, c_int64, _P, _P, c_int64, c_int64, _P, _P, _P]),
    'afsi_indexed_gather_f32': (c_int, [_P, c_int64, c_int64, _P, _P, c_int64, c_int64, _P, _P, _P]),
    'afsi_indexed_spread_workspace_bytes': (c_size_t, [c_int64, c_int64, c_int64]),
    'afsi_indexed_spread_f64': (c_int, [_P, c_int64, c_int64, _P, _P, c_int64, c_int64, c_double, _P, _P, c_size_t,
                                        _P, _P]),
    'afsi_indexed_spread_f32': (c_int, [_P, c_int64, c_int64, _P, _P, c_int64, c_int64, c_double, _P, _P, c_size_t,
                                        _P, _P]),
    'afsi_bin_key_space': (c_int64, [POINTER(Grid3), POINTER(c_int32)]),
    'afsi_bin_workspace_bytes': (c_size_t, [c_int64]),
    'afsi_bin_points': (c_int, [_P, c_int64, POINTER(Grid3), POINTER(c_int32), _P, _P, _P, _P, c_size_t, _P, _P]),
    'afsi_spread_workspace_bytes': (c_size_t, [POINTER(Grid3), POINTER(c_int32)]),
    'afsi_spread_gather_workspace_bytes': (c_size_t, [POINTER(Grid3), POINTER(c_int32), c_int64]),
    'afsi_tile_box_bounds': (c_int, [POINTER(Grid3), POINTER(c_int32), _I64, _I64]),
    'afsi_spread_tiled_f64': (c_int, [_P, _P, c_int64, POINTER(Grid3), POINTER(c_int32), _P, _P, _P, _P, _P, _P, c_size_t, _P, _P]),
    'afsi_spread_tiled_noalloc_f64': (c_int, [_P, _P, c_int64, POINTER(Grid3), POINTER(c_int32), _P, _P, _P, _P, _P,
                                              _P, _P]),
    'afsi_spread_gather_f64': (c_int, [_P, _P, c_int64, POINTER(Grid3), POINTER(c_int32), _P, _P, _P, _P, _P, _P, c_size_t, _P, _P]),
    'afsi_interp_tiled_f64': (c_int, [_P, c_int64, POINTER(Grid3), POINTER(c_int32), _P, _P, _P, _P, _P, _P, _P, c_size_t, _P, _P]),
    'afsi_mg_residual_3d_f64': (c_int, [_P, _P, _P, POINTER(c_int64), POINTER(c_double), c_int, _P]),
    'afsi_mg_residual_3d_f32': (c_int, [_P, _P, _P, POINTER(c_int64), POINTER(c_double), c_int, _P]),
    'afsi_mg_jacobi_3d_f64': (c_int, [_P, _P, _P, POINTER(c_int64), POINTER(c_double), POINTER(c_double), c_double, c_int, _P]),
    'afsi_mg_jacobi_3d_f32': (c_int, [_P, _P, _P, POINTER(c_int64), POINTER(c_double), POINTER(c_double), c_double, c_int, _P]),
    'afsi_mg_rbgs_3d_f64': (c_int, [_P, _P, POINTER(c_int64), POINTER(c_double), POINTER(c_double), c_int, c_int, _P]),
    'afsi_mg_rbgs_3d_f32': (c_int, [_P, _P, POINTER(c_int64), POINTER(c_double), POINTER(c_double), c_int, c_int, _P]),
    'afsi_mg_restrict_3d_f64': (c_int, [_P, _P, POINTER(c_int64), _P]),
    'afsi_mg_restrict_3d_f32': (c_int, [_P, _P, POINTER(c_int64), _P]),
    'afsi_mg_residual_restrict_3d_f64': (c_int, [_P, _P, _P, POINTER(c_int64), POINTER(c_double), c_int, _P]),
    'afsi_mg_residual_restrict_3d_f32': (c_int, [_P, _P, _P, POINTER(c_int64), POINTER(c_double), c_int, _P]),
    'afsi_mg_prolong_add_3d_f64': (c_int, [_P, _P, POINTER(c_int64), c_int, _P]),
    'afsi_mg_prolong_add_3d_f32': (c_int, [_P, _P, POINTER(c_int64), c_int, _P]),
    'afsi_dense_matvec_f64': (c_int, [_P, _P, _P, c_int64, _P]),
    'afsi_dense_matvec_f32': (c_int, [_P, _P, _P, c_int64, _P]),
    'afsi_reduce_partials': (c_int64, [c_int64]),
    'afsi_reduce_f64': (c_int, [_P, c_int64, _P, _P, _P]),
    'afsi_reduce_f32': (c_int, [_P, c_int64, _P, _P, _P]),
    'afsi_sub_mean_f64': (c_int, [_P, c_int64, _P, _P]),
    'afsi_sub_mean_f32': (c_int, [_P, c_int64, _P, _P]),
    'afsi_fill_f64': (c_int, [_P, c_int64, c_double, _P]),
    'afsi_fill_f32': (c_int, [_P, c_int64, c_double, _P]),
    'afsi_divergence_3d_f64': (c_int, [_P, _P, _P, _P, POINTER(c_int64), POINTER(c_double), c_double, c_int, _P]),
    'afsi_grad_correct_3d_f64': (c_int, [_P, _P, _P, _P, POINTER(c_int64), POINTER(c_double), c_double, c_int, _P]),
    'afsi_axpy_f64': (c_int, [_P, _P, c_int64, c_double, _P]),
    'afsi_project_rhs_3d_f64': (c_int, [_P, _P, _P, _P, _P, _P, c_double, _P, POINTER(c_int64), POINTER(c_double),
                                        c_double, c_int, _P]),
    'afsi_project_correct_3d_f64': (c_int, [_P, _P, _P, _P, _P, _P, c_double, _P, POINTER(c_int64),
                                            POINTER(c_double), c_double, c_int, _P]),
    'afsi_mg3d_partials_doubles': (c_int64, [POINTER(c_int64)]),
    'afsi_mg3d_supported': (c_int, [POINTER(MGPlan3)]),
    'afsi_mg3d_sweep_f64': (c_int, [POINTER(MGPlan3), c_int32, _P]),
    'afsi_mg3d_sweep_f32': (c_int, [POINTER(MGPlan3), c_int32, _P]),
    'afsi_refine_residual_3d': (c_int, [_P, _P, _P, POINTER(c_int64), POINTER(c_double), c_int, _P, c_int64, _P, _P]),
    'afsi_refine_update': (c_int, [_P, _P, c_int64, _P]),
    'afsi_refine_step_3d': (c_int, [_P, _P, _P, _P, _P, POINTER(c_int64), POINTER(c_double), c_int, _P, c_int64, _P, _P]),
    'afsi_mg3d_cycles_f64': (c_int, [POINTER(MGPlan3), c_int32, _P]),
    'afsi_mg3d_cycles_f32': (c_int, [POINTER(MGPlan3), c_int32, _P]),
    'afsi_mg3d_residual_f64': (c_int, [POINTER(MGPlan3), _P]),
    'afsi_mg3d_residual_f32': (c_int, [POINTER(MGPlan3), _P]),
    'afsi_mg3d_cycles_unfused_f64': (c_int, [POINTER(MGPlan3), c_int32, _P]),
    'afsi_mg3d_cycles_unfused_f32': (c_int, [POINTER(MGPlan3), c_int32, _P]),
    'afsi_mg3d_residual_unfused_f64': (c_int, [POINTER(MGPlan3), _P]),
    'afsi_mg3d_residual_unfused_f32': (c_int, [POINTER(MGPlan3), _P]),
    'afsi_tb_sizes': (c_int, [POINTER(TBBatch), POINTER(c_int64), POINTER(c_int64), POINTER(c_size_t)]),
    'afsi_tb_prepare': (c_int, [POINTER(TBBatch), _P]),
    'afsi_tb_bin': (c_int, [POINTER(TBBatch), _P]),
    'afsi_tb_spread': (c_int, [POINTER(TBBatch), _P]),
    'afsi_tb_interp': (c_int, [POINTER(TBBatch), _P]),
    'afsi_mgb_workspace_bytes': (c_size_t, [POINTER(MGBBatch)]),
    'afsi_mgb_partials_doubles': (c_int64, [POINTER(MGBBatch)]),
    'afsi_mgb_prepare': (c_int, [POINTER(MGBBatch), _P]),
    'afsi_mgb_begin': (c_int, [POINTER(MGBBatch), c_double, c_double, _P]),
    'afsi_mgb_begin_at': (c_int, [POINTER(MGBBatch), c_int32, c_double, c_double, _P]),
    'afsi_ib3b_workspace_bytes': (c_size_t, [POINTER(IB3BBatch)]),
    'afsi_ib3b_prepare': (c_int, [POINTER(IB3BBatch), _P]),
    'afsi_ib3b_rhs': (c_int, [POINTER(IB3BBatch), c_double, c_double, _P]),
    'afsi_ib3b_correct': (c_int, [POINTER(IB3BBatch), c_int32, c_double, c_double, _P]),
    'afsi_ib3b_advance': (c_int, [POINTER(IB3BBatch), c_double, _P]),
    'afsi_mgb_iterate': (c_int, [POINTER(MGBBatch), c_int32, c_int32, _P]),
    'afsi_mgb_finish': (c_int, [POINTER(MGBBatch), c_int32, _P]),
    'afsi_channel_prepare_f64': (c_int, [_P, c_int64, POINTER(Grid2), _P, _P, _P, _P, _P, _P]),
    'afsi_mg_residual_2d_f64': (c_int, [_P, _P, _P, _I64, _F64, c_int, _P]),
    'afsi_mg_residual_2d_f32': (c_int, [_P, _P, _P, _I64, _F64, c_int, _P]),
    'afsi_mg_jacobi_2d_f64': (c_int, [_P, _P, _P, _I64, _F64, _F64, c_double, c_int, _P]),
    'afsi_mg_jacobi_2d_f32': (c_int, [_P, _P, _P, _I64, _F64, _F64, c_double, c_int, _P]),
    'afsi_mg_rbgs_2d_f64': (c_int, [_P, _P, _I64, _F64, _F64, c_int, _P]),
    'afsi_mg_rbgs_2d_f32': (c_int, [_P, _P, _I64, _F64, _F64, c_int, _P]),
    'afsi_mg_residual_restrict_2d_f64': (c_int, [_P, _P, _P, _I64, _F64, c_int, _P]),
    'afsi_mg_residual_restrict_2d_f32': (c_int, [_P, _P, _P, _I64, _F64, c_int, _P]),
    'afsi_mg_prolong_add_2d_f64': (c_int, [_P, _P, _I64, c_int, _P]),
    'afsi_mg_prolong_add_2d_f32': (c_int, [_P, _P, _I64, c_int, _P]),
    'afsi_divergence_2d_f64': (c_int, [_P, _P, _P, _I64, _F64, c_double, c_int, _P]),
    'afsi_grad_correct_2d_f64': (c_int, [_P, _P, _P, _I64, _F64, c_double, c_int, _P]),
    'afsi_mg2d_cycles_f64': (c_int, [POINTER(MGPlan2), c_int32, _P]),
    'afsi_mg2d_cycles_f32': (c_int, [POINTER(MGPlan2), c_int32, _P]),
    'afsi_mg2d_residual_f64': (c_int, [POINTER(MGPlan2), _P]),
    'afsi_mg2d_residual_f32': (c_int, [POINTER(MGPlan2), _P]),
    'afsi_mac_predict_f64': (c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, c_double, c_double, c_double, _P]),
    'afsi_mac_speeds_workspace_bytes': (c_size_t, []),
    'afsi_mac_speeds_f64': (c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P, c_size_t, _P]),
    'afsi_fe_evaluate_f64': (c_int, [POINTER(FEGeometry), _P, _P, _P]),
    'afsi_fe_weighted_f64': (c_int, [POINTER(FEGeometry), _P, _P, _P]),
    'afsi_fe_assemble_f64': (c_int, [POINTER(FEGeometry), _P, _P, _P]),
    'afsi_csr_spmv3_f64': (c_int, [POINTER(CSR), _P, _P, _P]),
    'afsi_mass_pcg_workspace_bytes': (c_size_t, [c_int64]),
    'afsi_mass_pcg_f64': (c_int, [POINTER(CSR), _P, _P, _P, _P, POINTER(PCGOptions), _P, c_size_t,
                                  POINTER(PCGInfo), _P]),
    **{f'afsi_p2_prepare_{t}': (c_int, [_P, POINTER(P2Group), POINTER(Grid3), POINTER(P2Tables), _P, _P])
       for t in ('f64', 'f32')},
    **{f'afsi_p2_spread_{t}': (c_int, [_P, POINTER(P2Group), POINTER(P2Tables), POINTER(Grid3), c_double, _P, _P,
                                       _P, _P, _P]) for t in ('f64', 'f32')},
    **{f'afsi_p2_gather_assemble_{t}': (c_int, [_P, _P, _P, POINTER(P2Group), POINTER(P2Tables), POINTER(Grid3),
                                                _P, _P, _P]) for t in ('f64', 'f32')},
    'afsi_hash_accumulate_f64': (c_int, [_P, _P, c_int64, POINTER(HashTable), _P]),
    'afsi_chb_sizes': (c_int, [POINTER(CHBBatch), _I64, _I64, _I64, _I64, POINTER(c_size_t)]),
    'afsi_chb_prepare': (c_int, [POINTER(CHBBatch), _P]),
    'afsi_chb_stencil_f64': (c_int, [POINTER(CHBBatch), _P, _P, _P, _P, _P, _P]),
    'afsi_chb_divergence_f64': (c_int, [POINTER(CHBBatch), _P, _P, c_double, _P]),
    'afsi_chb_correct_f64': (c_int, [POINTER(CHBBatch), _P, _P, c_double, _P]),
    'afsi_chb_cycles_f64': (c_int, [POINTER(CHBBatch), _P, c_int32, _P]),
    'afsi_chb_stats_f64': (c_int, [POINTER(CHBBatch), c_int, _P, _P]),
    'afsi_chb_spread_workspace_bytes': (c_size_t, [POINTER(CHBBatch)]),
    'afsi_chb_spread_f64': (c_int, [POINTER(CHBBatch), _P, _P, _P, _P, c_size_t, _P, _P]),
    'afsi_p1_contract_f64': (c_int, [POINTER(FEGeometry), _P, _P, c_int64, c_int64, c_int, POINTER(Grid3),
                                     POINTER(HashTable), _P]),
    'afsi_csr_spmv_f64': (c_int, [POINTER(CSR), _P, c_int64, _P, c_int64, c_double, _P]),
}


def exported_names():
    return sorted(_SIGNATURES)


def load():
    """Load the library (once) and attach signatures; raise if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIBRARY.exists():
            raise RuntimeError(f'{LIBRARY} is missing: build it with '
                               '`python -c "import __graft_entry__ as g; g.build()"`')
        lib = ctypes.CDLL(str(LIBRARY))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def call(name, *args):
    """Call an entry point; raise RuntimeError with the library's message."""
    lib = load()
    status = getattr(lib, name)(*args)
    if status:
        message = lib.afsi_last_error().decode(errors='replace')
        raise RuntimeError(f'{name}: {message}')


def stream():
    return c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    return c_void_p(t.data_ptr()) if t is not None else c_void_p(0)


def i64x3(values):
    return (c_int64 * 3)(*[int(v) for v in values])


def f64x3(values):
    return (c_double * 3)(*[float(v) for v in values])


def i32x3(values):
    return (c_int32 * 3)(*[int(v) for v in values])
