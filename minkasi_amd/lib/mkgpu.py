"""ctypes bindings for libmkgpu (declared in include/mkgpu.h).

Counterpart of the reference's argtypes tables (minkasi/lib/minkasi.py:10-225,
minkasi/lib/mkfftw.py:7-95).  Every entry point returns a status; ``check``
turns a non-zero status into ``MkError`` carrying ``mk_last_error()``.  The
reference's C functions return void and never fail; raising here is a
deliberate divergence.
"""
import ctypes

from ._mkgpu_init import libmkgpu as lib, libmkgpu_path

P = ctypes.c_void_p
I = ctypes.c_int
I32 = ctypes.c_int32
I64 = ctypes.c_int64
D = ctypes.c_double

MK_OK = 0
MK_ERR_ARG = 1
MK_ERR_HIP = 2
MK_ERR_RANGE = 3
MK_ERR_NUMERIC = 4
MK_ERR_ALLOC = 5

# device scalar block layout (MK_SC_* in mkgpu.h)
SC_ZR, SC_PAP, SC_ALPHA, SC_BETA, SC_RR, SC_FLAG, SC_DOT = 0, 1, 2, 3, 4, 5, 6
SC_COUNT = 8

# polarisation modes (MK_POL_* in mkgpu.h)
POL_QU, POL_IQU, POL_QU_PRECON, POL_IQU_PRECON = 1, 2, 3, 4


class MkError(RuntimeError):
    def __init__(self, name, status, msg):
        super().__init__("%s failed (status %d): %s" % (name, status, msg))
        self.status = status


_SIGS = {
    "mk_conv3x3_f32": ([P, I32, I32, I32, I32, P, I32, I32, P, I32, P], I),
    "mk_conv3x3_adjoint_weights": ([P, I32, I32, P, P], I),
    "mk_conv3x3_wgrad_workspace_bytes": ([I32, I32, I32, I32, I32, I32], I64),
    "mk_conv3x3_wgrad_f32": ([P, I32, P, I32, I32, I32, I32, I32, P, I32, P, I64, P], I),
    "mk_precon_finish_f32": ([P, P, P, P, P, I64, P], I),
    "mk_restrict_box_f32": ([P, I32, I32, I32, I32, P, P], I),
    "mk_prolong_box_f32": ([P, I32, I32, I32, I32, P, I32, P], I),
    "mk_restrict_tent_f32": ([P, I32, I32, I32, I32, P, P], I),
    "mk_prolong_tent_f32": ([P, I32, I32, I32, I32, P, I32, P], I),
    "mk_precon_loss_workspace_bytes": ([], I64),
    "mk_precon_aloss_f32": ([P, P, P, P, P, P, I64, P, P, P, I64, P], I),
    "mk_precon_loss_f32": ([P, P, P, P, P, P, I64, P, P, P, I64, P], I),
    "mk_adam_f32": ([P, P, P, P, I32, P, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, I32,
                     ctypes.c_float, P], I),
    "mk_last_error": ([], ctypes.c_char_p),
    "mk_version": ([], I),
    "mk_reduce_workspace_bytes": ([], I64),
    "mk_pix_from_radec_car": ([P, P, I64, P, I32, I32, P, P, P], I),
    "mk_world2pix_car": ([P, P, I64, P, P, P, P], I),
    "mk_pix_from_raster_car": ([P, P, P, P, I64, I64, P, I32, I32, P, P, P], I),
    "mk_map2tod_f32": ([P, P, I64, P, I, P], I),
    "mk_map2tod_f64": ([P, P, I64, P, I, P], I),
    "mk_tod2map_atomic_f32": ([P, P, I64, P, P], I),
    "mk_tod2map_atomic_f64": ([P, P, I64, P, P], I),
    "mk_t2m_plan_create": ([P, I64, I32, I32, I32, I32, I32, ctypes.POINTER(P), P], I),
    "mk_t2m_plan_exec_f32": ([P, P, P, I, P], I),
    "mk_t2m_plan_exec_f64": ([P, P, P, I, P], I),
    "mk_t2m_plan_exec_f32_acc64": ([P, P, I, P, I, P], I),
    "mk_t2m_plan_exec_pol_f32": ([P, P, P, P, I, I, P], I),
    "mk_t2m_plan_exec_pol_f64": ([P, P, P, P, I, I, P], I),
    "mk_t2m_plan_info": ([P, P], I),
    "mk_t2m_plan_destroy": ([P], I),
    "mk_dot_f32": ([P, P, I64, P, P, P], I),
    "mk_dot_f64": ([P, P, I64, P, P, P], I),
    "mk_axpy_f32": ([P, P, D, I64, P], I),
    "mk_mul_f32": ([P, P, P, I64, P], I),
    "mk_pcg_alpha": ([P, P, I64, P, P, P], I),
    "mk_pcg_xr": ([P, P, P, P, I64, P, P, P], I),
    "mk_pcg_xrz_jacobi": ([P, P, P, P, P, P, I64, P, P, P], I),
    "mk_pcg_beta": ([P, P, I64, P, P, P], I),
    "mk_pcg_p": ([P, P, I64, P, P], I),
    "mk_project_out_f32": ([P, P, I64, D, P, P, P], I),
    "mk_pcg_init": ([P, P, P, I64, P, P, P], I),
    "mk_pcg_xr_x64": ([P, P, P, P, P, I64, P, P, P], I),
    "mk_pcg_xrz_jacobi_x64": ([P, P, P, P, P, P, P, I64, P, P, P], I),
    "mk_pcg_set_r": ([P, P, P, I64, P, P, P], I),
    "mk_dct1_workspace_bytes": ([I64, I64, I], I64),
    "mk_dct1_f64": ([P, P, I64, I64, P, I64, P], I),
    "mk_dct1_f32": ([P, P, I64, I64, P, I64, P], I),
    "mk_dct1_f64_noalloc": ([P, P, I64, I64, P], I),
    "mk_filt_plan_create": ([I, P, I, ctypes.POINTER(P)], I),
    "mk_filt_plan_set_weights": ([P, P, P], I),
    "mk_filt_plan_set_halve": ([P, I], I),
    "mk_filt_plan_borrow_spectrum": ([P, I, P, I], I),
    "mk_filt_plan_exec_f32": ([P, P, P, P], I),
    "mk_filt_plan_exec_f64": ([P, P, P, P], I),
    "mk_filt_plan_release_work": ([P, P], I),
    "mk_filt_plan_work_bytes": ([P, P], I),
    "mk_filt_plan_set_work": ([P, P, I64, P], I),
    "mk_filt_plan_info": ([P, P], I),
    "mk_filt_plan_geometry": ([P, P], I),
    "mk_filt_plan_destroy": ([P], I),
    "mk_split_bf16_padded_dims": ([I64, I64, P, P], I64),
    "mk_split_bf16_matrix": ([P, I64, I64, I64, P, P, P, P], I),
    "mk_split_f16_matrix": ([P, I64, I64, I64, I, P, P, P], I),
    "mk_rot_plan_create": ([I, P, I, ctypes.POINTER(P)], I),
    "mk_rot_plan_exec": ([P, P, P, P], I),
    "mk_rot_plan_destroy": ([P], I),
    "mk_gemm_nn_f64": ([I64, I64, I64, P, I64, P, I, I64, P, I64, P], I),
    "mk_gram_workspace_bytes": ([I64, I64], I64),
    "mk_gram_f64": ([I64, I64, P, I, I64, P, P, I64, P], I),
    "mk_cross_workspace_bytes": ([I64, I64, I64], I64),
    "mk_cross_f64": ([I64, I64, I64, P, I64, P, I64, P, P, I64, P], I),
    "mk_smooth_rows_f64": ([P, P, P, I64, I64, P, I32, I32, I32, P], I),
    "mk_median_cols_f32": ([P, I64, I64, I64, P, P], I),
    "mk_median_cols_f64": ([P, I64, I64, I64, P, P], I),
    "mk_cm_poly_design_f32": ([P, I64, I32, I32, P, P, P], I),
    "mk_cm_poly_design_f64": ([P, I64, I32, I32, P, P, P], I),
    "mk_design_dots_workspace_bytes": ([I64, I32, I64], I64),
    "mk_design_dots_f32": ([P, I64, I32, P, I64, I64, P, P, I64, P], I),
    "mk_design_dots_f64": ([P, I64, I32, P, I64, I64, P, P, I64, P], I),
    "mk_design_apply_f32": ([P, I64, I64, I64, P, I32, I32, I32, P, P, I64, P], I),
    "mk_design_apply_f64": ([P, I64, I64, I64, P, I32, I32, I32, P, P, I64, P], I),
    "mk_axpy_f64": ([P, P, D, I64, P], I),
    "mk_mul_f64": ([P, P, P, I64, P], I),
    "mk_addmul_f32": ([P, P, P, I64, P], I),
    "mk_addmul_f64": ([P, P, P, I64, P], I),
    "mk_filt_plan_create_p1": ([I, P, I, P, P], I),
    "mk_filt_hot_p1": ([P, I], I),
    "mk_pix_count_bad": ([P, I64, I64, P, P], I),
    "mk_map_tiled_size": ([I32, I32, I32, I32], I64),
    "mk_pix_tile_remap": ([P, I64, I32, I32, I32, I32, P, P, P], I),
    "mk_map_to_tiles_f32": ([P, I32, I32, I32, I32, P, P], I),
    "mk_map_to_tiles_f64": ([P, I32, I32, I32, I32, P, P], I),
    "mk_pix_code_blocks": ([I64], I64),
    "mk_pix_code_size": ([P, I64, I32, I32, P, P, P], I),
    "mk_pix_code_write": ([P, I64, I32, P, P, P, P, P], I),
    "mk_map2tod_coded_f32": ([P, P, I64, P, P, I32, I32, I32, I, P], I),
    "mk_map2tod_coded_f64": ([P, P, I64, P, P, I32, I32, I32, I, P], I),
    "mk_defl_plan_create": ([I, P, P, P], I),
    "mk_defl_plan_exec": ([P, P, P, P], I),
    "mk_defl_plan_addback": ([P, P, P], I),
    "mk_defl_plan_destroy": ([P], I),
    "mk_tod2cuts_f32": ([P, P, I64, P, I64, I, P, P], I),
    "mk_tod2cuts_f64": ([P, P, I64, P, I64, I, P, P], I),
    "mk_cuts2tod_f32": ([P, P, I64, P, I64, I, P, P], I),
    "mk_cuts2tod_f64": ([P, P, I64, P, I64, I, P, P], I),
    "mk_polmap2tod_f32": ([P, P, P, I64, P, I64, I, I, P, P], I),
    "mk_polmap2tod_f64": ([P, P, P, I64, P, I64, I, I, P, P], I),
    "mk_tod2polmap_atomic_f32": ([P, P, P, I64, P, I64, I, P, P], I),
    "mk_tod2polmap_atomic_f64": ([P, P, P, I64, P, I64, I, P, P], I),
    "mk_polmap_invert_f32": ([P, I64, I, D, P, I64, P], I),
    "mk_polmap_invert_f64": ([P, I64, I, D, P, I64, P], I),
    "mk_polmap_precon_mul_f32": ([P, P, P, I64, I, P], I),
    "mk_polmap_precon_mul_f64": ([P, P, P, I64, I, P], I),
}

_fns = {}
for _name, (_args, _res) in _SIGS.items():
    if hasattr(lib, _name):
        _f = getattr(lib, _name)
        _f.argtypes = _args
        _f.restype = _res
        _fns[_name] = _f


def fn(name):
    try:
        return _fns[name]
    except KeyError:
        raise MkError(name, MK_ERR_ARG, "symbol not exported by %s" % libmkgpu_path)


def check(name, *args):
    """Calls ``name`` and raises MkError on a non-zero status."""
    st = fn(name)(*args)
    if st != MK_OK:
        msg = lib.mk_last_error()
        raise MkError(name, st, msg.decode() if msg else "")
    return st


def last_error():
    m = lib.mk_last_error()
    return m.decode() if m else ""
