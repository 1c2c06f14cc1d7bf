"""ctypes binding of libfea_amd.so (the C ABI declared in include/fea_amd.h).

The library is loaded from fea_amd/_lib/ (built in-tree by fea_amd.build).
There is no fallback: if the library is missing or a call fails, a
FeaNativeError is raised.
"""
import ctypes
import os

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, '_lib', 'libfea_amd.so')

F32, F64 = 0, 1
F64_STAGE_F32 = 2      # element tangent: float64 arithmetic, float32 staged blocks
C_INT, C_I64, C_DBL, C_P = ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p


class FeaNativeError(RuntimeError):
    """A fea_amd C-ABI call returned an error status (or the library is absent)."""


_SIGS = {
    'fea_last_error': (ctypes.c_char_p, []),
    'fea_version': (C_INT, []),
    'fea_pattern_create': (C_INT, [C_INT, C_P, C_INT, C_P, C_P, C_P, C_P, ctypes.POINTER(C_P)]),
    'fea_pattern_sizes': (C_INT, [C_P, C_P]),
    'fea_pattern_export_csr': (C_INT, [C_P, C_P, C_P, C_P]),
    'fea_pattern_export_blocks': (C_INT, [C_P, C_P, C_P, C_P, C_P, C_P]),
    'fea_pattern_export_group': (C_INT, [C_P, C_INT, C_P, C_P]),
    'fea_pattern_destroy': (None, [C_P]),
    'fea_pattern_export_group_colpos': (C_INT, [C_P, C_INT, C_P]),
    'fea_color_elements': (C_INT, [C_INT, C_INT, C_INT, C_P, C_P, C_P]),
    'fea_elem_tangent_direct': (C_INT, [C_INT, C_INT, C_INT, C_INT, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P,
                                        C_INT, C_P, C_P, C_P]),
    'fea_elem_residual_tangent': (C_INT, [C_INT, C_INT, C_INT, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P]),
    'fea_elem_residual': (C_INT, [C_INT, C_INT, C_INT, C_P, C_P, C_P, C_P, C_P, C_P, C_P]),
    'fea_elem_energy': (C_INT, [C_INT, C_INT, C_P, C_P, C_P, C_P, C_P, C_P]),
    'fea_elem_energy_blocks': (C_INT, [C_INT]),
    'fea_elem_prepare': (C_INT, [C_INT, C_INT, C_P, C_P, C_P, C_P, C_P, C_P]),
    'fea_elem_mass': (C_INT, [C_INT, C_INT, C_P, C_P, C_DBL, C_P, C_P, C_P]),
    'fea_elem_lumped_mass': (C_INT, [C_INT, C_INT, C_P, C_P, C_P, C_DBL, C_P, C_P, C_P]),
    'fea_surface_pressure': (C_INT, [C_INT, C_INT, C_INT, C_P, C_P, C_P, C_DBL, C_P, C_P, C_P, C_P, C_P]),
    'fea_surface_volume': (C_INT, [C_INT, C_INT, C_P, C_P, C_P, C_P, C_P]),
    'fea_surface_traction': (C_INT, [C_INT, C_INT, C_P, C_P, C_P, C_P, C_DBL, C_P, C_P, C_P, C_P]),
    'fea_surface_blocks': (C_INT, [C_INT]),
    'fea_gather_csr': (C_INT, [C_INT, C_INT, C_INT, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P]),
    'fea_gather_residual': (C_INT, [C_INT, C_P, C_P, C_P, C_P, C_P]),
    'fea_reduce_sum': (C_INT, [C_INT, C_P, C_P, C_P]),
    'fea_spmv': (C_INT, [C_INT, C_INT, C_INT, C_I64, C_INT, C_P, C_P, C_P, C_P, C_P, C_P, C_P]),
    'fea_jacobi_inv': (C_INT, [C_INT, C_INT, C_INT, C_P, C_P, C_P, C_P]),
    'fea_pcg_ws_size': (C_I64, [C_INT]),
    'fea_pcg_start': (C_INT, [C_INT, C_INT, C_INT, C_I64, C_INT, C_P, C_P, C_P, C_P, C_INT, C_P, C_P, C_P,
                              C_P, C_P, C_P, C_P, C_P]),
    'fea_pcg_iterate': (C_INT, [C_INT, C_INT, C_INT, C_I64, C_INT, C_P, C_P, C_P, C_P, C_INT, C_P, C_P, C_P,
                                C_P, C_P, C_P, C_P, C_DBL, C_INT, C_INT, C_P, C_P]),
    'fea_pcg_run_noalloc': (C_INT, [C_INT, C_INT, C_INT, C_I64, C_INT, C_P, C_P, C_P, C_P, C_INT, C_P, C_P, C_P,
                                    C_P, C_P, C_P, C_P, C_DBL, C_INT, C_INT, C_P, C_P]),
    'fea_amg_build': (C_INT, [C_INT, C_P, C_P, C_P, C_INT, C_P, C_P, C_P, C_INT, C_INT,
                                     C_INT, ctypes.POINTER(C_P)]),
    'fea_amg_num_levels': (C_INT, [C_P]),
    'fea_amg_level_sizes': (C_INT, [C_P, C_INT, C_P]),
    'fea_amg_level_export': (C_INT, [C_P, C_INT] + [C_P] * 12),
    'fea_amg_destroy': (None, [C_P]),
    'fea_amg_level_size': (C_INT, []),
    'fea_amg_vcycle': (C_INT, [C_INT, C_INT, C_I64, C_INT, C_P, C_P, C_P, C_P, C_P,
                                      C_DBL, C_P, C_INT, C_INT, C_P, C_P, C_P]),
    'fea_amg_smooth0': (C_INT, [C_INT, C_INT, C_I64, C_INT, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_DBL, C_DBL,
                                C_P]),
    'fea_amg_half_values': (C_INT, [C_INT, C_I64, C_P, C_P, C_P, C_P]),
    'fea_amg_smooth0_half': (C_INT, [C_INT, C_I64, C_INT, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_DBL,
                                     C_DBL, C_P]),
    'fea_amg_power': (C_INT, [C_INT, C_INT, C_I64, C_INT, C_P, C_P, C_P, C_P, C_P, C_P,
                                     C_P, C_P, C_INT, ctypes.POINTER(C_DBL), C_P]),
    'fea_pcg_start_amg': (C_INT, [C_INT, C_INT, C_I64, C_INT, C_P, C_P, C_P, C_P, C_P,
                                         C_DBL, C_P, C_INT, C_INT] + [C_P] * 7),
    'fea_pcg_iterate_amg': (C_INT, [C_INT, C_INT, C_I64, C_INT, C_P, C_P, C_P, C_P,
                                           C_P, C_DBL, C_P, C_INT, C_INT] + [C_P] * 6 +
                            [C_DBL, C_INT, C_INT, C_P, C_P]),
    'fea_block_jacobi_inv': (C_INT, [C_INT, C_INT, C_INT, C_P, C_P, C_P, C_P, C_P, C_P, C_P]),
    'fea_pcg_spmv': (C_INT, [C_INT, C_INT, C_INT, C_I64, C_INT, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P]),
    'fea_spmv_residual_mixed': (C_INT, [C_INT, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P]),
    'fea_vec_blocks': (C_INT, [C_INT]),
    'fea_dof_axpy': (C_INT, [C_INT, C_P, C_P, C_DBL, C_P, C_P, C_P]),
    'fea_convert': (C_INT, [C_INT, C_INT, C_P, C_INT, C_P, C_P]),
    'fea_dot_partial': (C_INT, [C_INT, C_P, C_P, C_P, C_P]),
    'fea_axpy': (C_INT, [C_INT, C_DBL, C_INT, C_P, C_P, C_P]),
    'fea_vec_stats': (C_INT, [C_INT, C_P, C_P, C_P]),
    'fea_explicit_update': (C_INT, [C_INT, C_P, C_P, C_P, C_P, C_P, C_DBL, C_DBL, C_P, C_P, C_P, C_P]),
    'fea_explicit_update_tiles': (C_INT, [C_INT, C_P, C_P, C_P, C_P, C_P, C_P, C_DBL, C_DBL, C_P, C_P, C_P, C_P]),
    'fea_elem_tile_size': (C_INT, [C_INT]),
    'fea_elem_residual_tiles': (C_INT, [C_INT, C_INT, C_INT, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P,
                                        C_INT, C_P]),
    'fea_contact_set_size': (C_INT, []),
    'fea_contact_rebuild': (C_INT, [C_P, C_P, C_P, C_P]),
    'fea_contact_eval': (C_INT, [C_P, C_P, C_P, C_P, C_P, C_P]),
    'fea_contact_track': (C_INT, [C_P, C_P, C_P, C_P]),
    'fea_elem_residual_tiles_contact': (C_INT, [C_INT, C_INT, C_INT, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P,
                                                C_P]),
    'fea_explicit_update_tiles_contact': (C_INT, [C_INT, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_DBL, C_DBL,
                                                  C_P, C_P, C_P, C_P]),
    'fea_contact_grid_build': (C_INT, [C_INT, C_P, C_P, C_P, C_P, C_DBL, C_INT, C_P, C_P, C_P, C_P, C_P, C_P]),
    'fea_contact_search': (C_INT, [C_INT, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_DBL, C_INT, C_DBL, C_INT, C_P,
                                   C_P, C_P, C_P]),
    'fea_contact_blocks': (C_INT, [C_INT]),
    'fea_elem_adjoint_blocks': (C_INT, [C_INT]),
    'fea_elem_adjoint': (C_INT, [C_INT, C_INT, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P]),
    'fea_surface_pressure_hvp': (C_INT, [C_INT, C_INT, C_P, C_P, C_P, C_DBL, C_P, C_P, C_P, C_P]),
    'fea_gather_nodal': (C_INT, [C_INT, C_P, C_P, C_P, C_P]),
    'fea_pcg_system_size': (C_INT, []),
    'fea_couple_map_size': (C_INT, []),
    'fea_barrier_search': (C_INT, [C_INT, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_DBL, C_INT, C_DBL, C_INT, C_P, C_P, C_P,
                                   C_P, C_P]),
    'fea_barrier_active': (C_INT, [C_P, C_INT, C_P, C_P, C_P, C_P, C_P, C_P, C_P]),
    'fea_barrier_missing': (C_INT, [C_P, C_INT, C_P, C_P, C_P, C_P, C_P, C_P, C_INT, C_I64, C_P, C_P]),
    'fea_barrier_energy_blocks': (C_INT, [C_INT]),
    'fea_scan_i32': (C_INT, [C_INT, C_P, C_P, C_P]),
    'fea_barrier_energy': (C_INT, [C_P, C_INT, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P]),
    'fea_barrier_assemble': (C_INT, [C_P, C_INT, C_INT, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P]),
    'fea_couple_kinematics': (C_INT, [C_P, C_P, C_INT, C_P]),
    'fea_couple_condense': (C_INT, [C_P, C_INT, C_P, C_P, C_DBL, C_P, C_P, C_P]),
    'fea_pcg_batch_layout': (C_INT, [C_INT, C_P, C_P, C_P]),
    'fea_pcg_batch_start': (C_INT, [C_INT, C_INT, C_INT, C_P, C_P, C_P, C_INT, C_INT, C_P]),
    'fea_pcg_batch_spmv': (C_INT, [C_INT, C_INT, C_INT, C_P, C_P, C_INT, C_P]),
    'fea_pcg_batch_iterate': (C_INT, [C_INT, C_INT, C_INT, C_P, C_P, C_P, C_INT, C_INT, C_INT, C_P, C_P]),
    'fea_pcg_batch_iterate_dev': (C_INT, [C_INT, C_INT, C_P, C_P, C_P, C_P, C_INT, C_INT, C_INT, C_P, C_P]),
    'fea_contact_forces': (C_INT, [C_INT, C_P, C_P, C_P, C_P, C_INT, C_P, C_P, C_INT, C_DBL, C_P, C_DBL, C_INT, C_P,
                                   C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_P, C_I64, C_P, C_P, C_P]),
}

_lib = None


def lib():
    """Load (once) and return the ctypes library; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FeaNativeError('fea_amd HIP library not found at %s: run `python -m fea_amd.build`'
                             % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


TORCH_OPS_PATH = os.path.join(_HERE, '_lib', 'libfea_torch.so')
_ops = None


def ops():
    """The PyTorch-ROCm operators (torch.ops.fea, TORCH_LIBRARY in csrc/torch_ops.cpp) over
    the same C ABI, with device / dtype / contiguity / size checks; raises if not built."""
    global _ops
    if _ops is not None:
        return _ops
    lib()
    if not os.path.exists(TORCH_OPS_PATH):
        raise FeaNativeError('fea_amd torch operator library not found at %s: run `python -m fea_amd.build`'
                             % TORCH_OPS_PATH)
    torch.ops.load_library(TORCH_OPS_PATH)
    _ops = torch.ops.fea
    return _ops


def exported_symbols():
    return list(_SIGS.keys())


def check(rc, what=''):
    if rc != 0:
        msg = lib().fea_last_error()
        raise FeaNativeError('%s failed (status %d): %s' % (what, rc, msg.decode() if msg else ''))


def _arg(a):
    # pointer parameters are declared c_void_p, which takes the address as an int
    if isinstance(a, torch.Tensor):
        return a.data_ptr()
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a


_FN = {}


def call(name, *args):
    """Call a status-returning C-ABI function and raise on error.

    Tensors and numpy arrays are passed by address; ``args`` keeps every one of
    them referenced until the call returns (and the kernels are enqueued), so a
    temporary cannot be freed and its memory reused before the launch.
    """
    fn = _FN.get(name)
    if fn is None:
        if name not in _SIGS:        # an int address needs a declared pointer parameter
            raise FeaNativeError('%s has no declared signature' % name)
        fn = _FN[name] = getattr(lib(), name)
    rc = fn(*[_arg(a) for a in args])
    if rc != 0:
        check(rc, name)


def raw(name):
    """The ctypes function of a declared entry point (for pre-bound integer arguments)."""
    fn = _FN.get(name)
    if fn is None:
        if name not in _SIGS:
            raise FeaNativeError('%s has no declared signature' % name)
        fn = _FN[name] = getattr(lib(), name)
    return fn


def check_rc(rc, name):
    if rc != 0:
        check(rc, name)


def ptr(t):
    """Device/host address of a tensor (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


_raw_stream = torch._C._cuda_getCurrentRawStream
_cur_device = torch._C._cuda_getDevice


def stream_ptr(device=None):
    """The current torch stream of ``device`` (default: the current device) as a handle."""
    if device is None:
        return ctypes.c_void_p(_raw_stream(_cur_device()))
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def require_gpu():
    """The hot path runs only on a ROCm GPU; fail loudly otherwise."""
    if not torch.cuda.is_available():
        raise FeaNativeError('fea_amd requires a ROCm GPU (torch.cuda.is_available() is False)')
    lib()
