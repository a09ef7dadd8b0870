"""pamr_amrd_amd -- MI355X (gfx950) multigrid hot path of PAMR/AMRD.

Python view of the C ABI declared in ``include/pamr_mg_gpu.h`` and
``include/pamr_amr_gpu.h`` (and the PAMR/AMRD surface in ``include/pamr.h`` /
``include/amrd_mg.h``), implemented by
``libpamr_mg.so`` in this directory.  Device memory is owned either by the
library (``Solver``) or by the caller (torch tensors passed as pointers to
the m_util_r8 / hook entry points).  There is no CPU fallback: when the
shared library is missing every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpamr_mg.so")

PMG_F_U, PMG_F_RHO, PMG_F_RHS, PMG_F_RES, PMG_F_LOP, PMG_F_A, PMG_F_C = range(7)

_lib = None
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class NativeLibraryMissing(RuntimeError):
    pass


def lib():
    """Load libpamr_mg.so (raises NativeLibraryMissing when not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryMissing(
                f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    i, d = C.c_int, C.c_double
    m_args = [_vp, _vp] + [i] * 21
    L.pmg_dmcopy3d.argtypes = m_args + [_vp]
    L.pmg_dmfwr3d.argtypes = m_args + [_vp, d, i, _vp]
    L.pmg_dmhwr3d.argtypes = m_args + [_vp, d, i, _vp]
    L.pmg_dminterp3d.argtypes = [_vp, _vp] + [i] * 21 + [i, _vp, d, i, _vp]
    L.pmg_hook_relax.argtypes = [i] + [_vp] * 7 + [d, i, i, i, d, d, d, _vp, _vp]
    L.pmg_hook_relax_noalloc.argtypes = [i] + [_vp] * 7 + [d, i, i, i, d, d, d, _vp, _vp]
    L.pmg_hook_residual.argtypes = [i] + [_vp] * 8 + [d, i, i, i, d, d, d, _vp, _vp]
    L.pmg_hook_lop.argtypes = [i] + [_vp] * 7 + [d, i, i, i, d, d, d, _vp]
    L.pmg_create.argtypes = [i, i, i, _dp, i, i, i, d]
    L.pmg_create.restype = _vp
    L.pmg_destroy.argtypes = [_vp]
    L.pmg_destroy.restype = None
    L.pmg_nlev.argtypes = [_vp]
    L.pmg_level_shape.argtypes = [_vp, i, _ip]
    L.pmg_set_params.argtypes = [_vp, i, i, i, d, d, d, i, i, i, i]
    L.pmg_field.argtypes = [_vp, i, i]
    L.pmg_field.restype = _vp
    L.pmg_upload.argtypes = [_vp, i, i, _dp]
    L.pmg_download.argtypes = [_vp, i, i, _dp]
    L.pmg_set_chr.argtypes = [_vp, i, _dp]
    L.pmg_download_cmask.argtypes = [_vp, i, i, _dp]
    L.pmg_init.argtypes = [_vp]
    L.pmg_vcycle.argtypes = [_vp, _dp, _dp, _ip]
    L.pmg_vcycle_async.argtypes = [_vp]
    L.pmg_read_norms.argtypes = [_vp, _dp, _dp, _ip]
    L.pmg_residual.argtypes = [_vp, _dp]
    L.pmg_solve.argtypes = [_vp, i, i, d, _dp, _ip]
    L.pmg_solve_rel.argtypes = [_vp, i, d, i, _dp, _ip, _dp]
    L.pmg_last_citer.argtypes = [_vp, i]
    L.pmg_zero_solution.argtypes = [_vp]
    L.pmg_time_kernel.argtypes = [_vp, i, i, i, i, _dp]
    L.pmg_set_sweep_config.argtypes = [_vp, i, i]
    L.pmg_set_level_sweep_config.argtypes = [_vp, i, i, i]
    L.pmg_set_level_sweep_config_prec.argtypes = [_vp, i, i, i, i]
    L.pmg_num_sweep_variants.argtypes = []
    L.pmg_set_graph.argtypes = [_vp, i]
    L.pmg_set_bottom_fusion.argtypes = [_vp, i]
    L.pmg_set_bottom_lds.argtypes = [_vp, i]
    L.pmg_set_fp32_exact_bottom.argtypes = [_vp, i]
    L.pmg_set_restrict_fusion.argtypes = [_vp, i]
    L.pmg_set_resid_restrict_zchunk.argtypes = [_vp, i, i]
    L.pmg_set_sweep_blocking.argtypes = [_vp, i, i]
    L.pmg_stream.argtypes = [_vp]
    L.pmg_stream.restype = _vp
    L.pmg_net_create.argtypes = [i, C.c_ulonglong]
    L.pmg_net_create.restype = _vp
    L.pmg_net_destroy.argtypes = [_vp]
    L.pmg_net_destroy.restype = None
    L.pmg_net_set_weights.argtypes = [_vp] + [_dp] * 6
    L.pmg_net_get_weights.argtypes = [_vp] + [_dp] * 6
    L.pmg_net_forward.argtypes = [_vp, _vp, _vp, i, i, i, i]
    L.pmg_net_loss_grad.argtypes = [_vp, _vp, _vp, i, i, i, i, i, _dp] + [_dp] * 6
    L.pmg_net_train_step.argtypes = [_vp, _vp, _vp, i, i, i, i, i, C.c_float, _dp]
    L.pmg_net_stream.argtypes = [_vp]
    L.pmg_net_stream.restype = _vp
    L.pmg_net_mfma_selftest.argtypes = [C.POINTER(C.c_float)] * 3
    L.pmg_net_mfma32_selftest.argtypes = [C.POINTER(C.c_float)] * 3
    L.pmg_net_correct.argtypes = [_vp, _vp, C.c_float, _vp, d, i, i, i, _vp]
    L.pmg_net_recorrect.argtypes = [_vp, _vp, d, i, i, i, _vp]
    L.pmg_net_infer.argtypes = [_vp, _vp, C.c_float, _vp, i, i, i, i]
    L.pmg_net_reserve.argtypes = [_vp, C.c_longlong]
    L.pmg_gather_crops.argtypes = [_vp, _vp, i, i, i, _ip, i, i, d, _vp, _vp]
    L.pmg_set_net.argtypes = [_vp, _vp]
    L.pmg_net_stats.argtypes = [_vp, _ip, _ip, _dp, _dp]
    L.pmg_set_solution_lincomb.argtypes = [_vp, _vp, _vp, d, d]
    L.pmg_solve_fgmres.argtypes = [_vp, i, d, i, _dp, _ip, _dp]
    L.pmg_set_fgmres_prec_cycles.argtypes = [_vp, i]
    L.pmg_history_warm_start.argtypes = [_vp, i, C.POINTER(_vp), C.POINTER(_vp), _dp, _dp]
    # learned band correction next to excised regions
    u8p = C.POINTER(C.c_ubyte)
    L.pmg_band_mask.argtypes = [_dp, i, i, i, d, i, u8p]
    L.pmg_band_mask.restype = C.c_longlong
    L.pmg_band_cores.argtypes = [u8p, i, i, i, i, i, _ip, i]
    L.pmg_band_setup.argtypes = [_vp, i, i]
    L.pmg_band_info.argtypes = [_vp, C.POINTER(C.c_longlong), _ip, _ip]
    L.pmg_band_download.argtypes = [_vp, u8p]
    L.pmg_set_band_net.argtypes = [_vp, _vp, d, i, i]
    L.pmg_band_stats.argtypes = [_vp, _ip, _ip, _dp]
    L.pmg_band_enable.argtypes = [_vp, i, i]
    L.pmg_band_set_inner.argtypes = [_vp, i]
    f = C.c_float
    L.pmg_band_gather_f32.argtypes = [_vp, i, i, i, _vp, i, i, f, _vp, _vp, _vp]
    L.pmg_band_scatter.argtypes = [_vp, _vp, _vp, _vp, _vp, i, i, i, _vp, i, i, i, d, _vp, _vp]
    L.pmg_band_resid.argtypes = [i, i, _vp, _vp, _vp, _vp, _vp, _vp, i, i, i, _vp, i, i, i, d, d, d, _vp]
    L.pmg_band_pairs.argtypes = [_vp, f, _vp, _vp, d, _vp, i, i, i, _vp, i, i, _vp, _vp, _vp, _vp]
    L.pmg_band_gather_res32.argtypes = [i, i, _vp, _vp, _vp, _vp, _vp, i, i, i, _vp, i, i, f, _vp, d, d, d, _vp,
                                        _vp]
    L.pmg_band_pairs_res32.argtypes = [i, i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, i, i, i, _vp, i, i, _vp, d, d,
                                       d, _vp, _vp, _vp, _vp]
    L.pmg_set_band_net_inner.argtypes = [_vp, _vp, d]
    L.pmg_band_sample_inner.argtypes = [_vp, _vp, i, C.c_ulonglong, _vp, _vp, _vp, _dp]
    L.pmg_band_sample.argtypes = [_vp, _vp, d, d, i, C.c_ulonglong, _vp, _vp, _vp]
    L.pmg_band_correct.argtypes = [_vp, i]
    L.pmg_download32.argtypes = [_vp, i, i, C.POINTER(C.c_float)]
    L.pmg_upload32.argtypes = [_vp, i, i, C.POINTER(C.c_float)]
    L.pmg_net_loss_grad_w.argtypes = [_vp, _vp, _vp, _vp, i, i, i, i, i, _dp] + [_dp] * 6
    L.pmg_net_train_step_w.argtypes = [_vp, _vp, _vp, _vp, i, i, i, i, i, C.c_float, _dp]
    L.pmg_net_infer_async.argtypes = [_vp, _vp, C.c_float, i, i, i, i, _vp]
    L.pmg_net_infer_async.restype = _vp
    # batched AMR hierarchy (include/pamr_amr_gpu.h)
    L.pmg_amr_create.argtypes = [_ip, _dp, i, _ip, _dp, i, i]
    L.pmg_amr_create.restype = _vp
    L.pmg_amr_destroy.argtypes = [_vp]
    L.pmg_amr_destroy.restype = None
    for f in ("pmg_amr_nlev", "pmg_amr_base_level", "pmg_amr_commit_u", "pmg_amr_init"):
        getattr(L, f).argtypes = [_vp]
    L.pmg_amr_ngrids.argtypes = [_vp, i]
    L.pmg_amr_grid_info.argtypes = [_vp, i, i, _ip, _dp]
    L.pmg_amr_field.argtypes = [_vp, i, i, i]
    L.pmg_amr_field.restype = _vp
    L.pmg_amr_upload.argtypes = [_vp, i, i, i, _dp]
    L.pmg_amr_download.argtypes = [_vp, i, i, i, _dp]
    L.pmg_amr_set_params.argtypes = [_vp, i, i, i, d, d, d, i, i, i, i]
    L.pmg_amr_vcycle.argtypes = [_vp, _dp, _dp, _ip]
    L.pmg_amr_solve.argtypes = [_vp, i, i, d, _dp, _ip]
    L.pmg_amr_solve_rel.argtypes = [_vp, i, d, i, _dp, _ip, _dp]
    L.pmg_amr_set_mixed_switch.argtypes = [_vp, d]
    L.pmg_amr_create_mgh.argtypes = [_vp, i]
    L.pmg_amr_create_mgh.restype = _vp
    L.pmg_amr_grid_coarsest.argtypes = [_vp, i, i]
    L.pmg_amr_sync_fields_coarsest.argtypes = [_vp, i, i, _ip]
    L.pmg_amr_residual.argtypes = [_vp, _dp]
    L.pmg_amr_sync.argtypes = [_vp, i, i, i]
    L.pmg_amr_inject.argtypes = [_vp, i, i, i]
    L.pmg_amr_interp.argtypes = [_vp, i, i]
    L.pmg_amr_interp_ord.argtypes = [_vp, i, i, i]
    L.pmg_amr_interp_fields_ord.argtypes = [_vp, i, i, i, _ip]
    L.pmg_amr_cmask.argtypes = [_vp, i, i, i, _dp]
    L.pmg_amr_plan_stats.argtypes = [_vp, i, _ip]
    L.pmg_amr_time_sweep.argtypes = [_vp, i, i, _dp]
    L.pmg_amr_set_graph.argtypes = [_vp, i]
    L.pmg_amr_stream.argtypes = [_vp]
    L.pmg_amr_stream.restype = _vp
    L.pmg_amr_add_field.argtypes = [_vp]
    L.pmg_amr_pool_points.argtypes = [_vp, i]
    L.pmg_amr_pool_points.restype = C.c_longlong
    L.pmg_amr_sync_fields.argtypes = [_vp, i, i, i, _ip]
    L.pmg_amr_inject_fields.argtypes = [_vp, i, i, _ip, _ip]
    L.pmg_amr_interp_fields.argtypes = [_vp, i, i, _ip]
    L.pmg_amr_fill_cmask.argtypes = [_vp, i, i, i]
    L.pmg_amr_num_sweep_variants.argtypes = []
    L.pmg_amr_set_level_sweep_config.argtypes = [_vp, i, i, i]
    L.pmg_amr_sweep_config.argtypes = [_vp, i, _ip, _ip]
    _declare_pamr(L)


def _declare_pamr(L):
    """include/pamr.h and include/amrd_mg.h (C strings are bytes)."""
    i, d, s = C.c_int, C.c_double, C.c_char_p
    L.PAMR_init_context.argtypes = [_dp, i, i, _ip, _dp, s, i]
    L.PAMR_free_context.argtypes = [i]
    L.PAMR_free_context.restype = None
    L.PAMR_compose_hierarchy.argtypes = [i, i, i, _ip, _dp, d]
    L.PAMR_build_mgh.argtypes = [i, i, i]
    L.PAMR_destroy_mgh.argtypes = []
    L.PAMR_excision_on.argtypes = [s, C.c_void_p, d, i]
    L.PAMR_excision_off.argtypes = []
    L.PAMR_excision_off.restype = None
    for f in ("PAMR_set_rho", "PAMR_get_rho"):
        getattr(L, f).argtypes = [_ip, _ip, i]
        getattr(L, f).restype = None
    L.PAMR_set_lambda.argtypes = [d]
    L.PAMR_set_lambda.restype = None
    L.PAMR_get_lambda.argtypes = [_dp]
    L.PAMR_get_lambda.restype = None
    L.PAMR_get_dxdt.argtypes = [i, _dp, _dp]
    L.PAMR_get_dxdt.restype = None
    for f in ("PAMR_set_MG_coarse_width", "PAMR_get_MG_coarse_width", "PAMR_set_ghost_width",
              "PAMR_get_ghost_width"):
        getattr(L, f).argtypes = [_ip]
        getattr(L, f).restype = None
    L.PAMR_get_global_bbox.argtypes = [_dp]
    L.PAMR_get_global_bbox.restype = None
    L.PAMR_def_var_brief.argtypes = [s]
    L.PAMR_def_var_full.argtypes = [s] + [i] * 12 + [_ip]
    L.PAMR_get_var_attribs.argtypes = [s] + [_ip] * 13
    L.PAMR_get_gfn.argtypes = [s, i, i]
    for f in ("PAMR_freeze_tf_bits", "PAMR_thaw_tf_bits", "PAMR_clear_tf_bits"):
        getattr(L, f).argtypes = []
        getattr(L, f).restype = None
    L.PAMR_set_tf_bit.argtypes = [i, i]
    L.PAMR_set_tf_bit.restype = None
    L.PAMR_set_tf_bits.argtypes = [i, i, i]
    L.PAMR_set_tf_bits.restype = None
    L.PAMR_sync.argtypes = [i, i, i, i]
    L.PAMR_inject.argtypes = [i, i, i]
    L.PAMR_interp.argtypes = [i, i, i]
    L.PAMR_set_time.argtypes = [i, d]
    L.PAMR_set_time.restype = None
    L.PAMR_get_time.argtypes = [i]
    L.PAMR_get_time.restype = d
    L.PAMR_init_s_iter.argtypes = [i, i, i]
    for f in ("PAMR_next_g", "PAMR_push_iter", "PAMR_pop_iter"):
        getattr(L, f).argtypes = []
    L.PAMR_get_g_attribs.argtypes = [_ip, _ip, _ip, _dp, _ip, _dp, _ip, C.POINTER(_dp), C.POINTER(_vp)]
    for f in ("PAMR_get_g_rank", "PAMR_get_g_dim", "PAMR_get_g_shape", "PAMR_get_g_ghost_width",
              "PAMR_get_g_ngfs", "PAMR_get_g_coarsest", "PAMR_get_g_level", "PAMR_get_g_comm"):
        getattr(L, f).argtypes = [_ip]
    L.PAMR_get_g_bbox.argtypes = [_dp]
    L.PAMR_get_g_t.argtypes = [_dp]
    L.PAMR_get_g_x.argtypes = [C.POINTER(_dp)]
    L.PAMR_get_g_gfs.argtypes = [C.POINTER(_vp)]
    L.PAMR_get_max_lev.argtypes = [i]
    L.PAMR_get_min_lev.argtypes = [i]
    L.PAMR_set_g_comm.argtypes = [i]
    L.amrd_set_app_MG_hooks.argtypes = [_vp, _vp, _vp]
    L.amrd_set_app_MG_hooks.restype = None
    L.amrd_set_app_MG_level_hooks.argtypes = [_vp, _vp, _vp]
    L.amrd_set_app_MG_level_hooks.restype = None
    L.amrd_mg_def_vars.argtypes = [s, s, s, i]
    L.amrd_mg_set_params.argtypes = [i, i, i, i, d, d, i, d, d]
    L.amrd_mg_set_params.restype = None
    L.amrd_vcycle.argtypes = [i]
    L.amrd_vcycle.restype = d
    L.amrd_solve_elliptics.argtypes = [i, i, _dp]
    L.solve_elliptics.argtypes = [i, i, i]
    L.solve_elliptics.restype = None
    L.vcycle.argtypes = [i, i]
    L.vcycle.restype = d
    L.amrd_vcycle_tre.argtypes = [i, i]
    L.amrd_vcycle_tre.restype = d
    L.amrd_solve_elliptics_tre.argtypes = [i, i, i, _dp]
    L.amrd_def_tre_vars.argtypes = [s]
    L.amrd_mg_def_extrap_vars.argtypes = [s, s]
    L.amrd_mg_def_vars_t0.argtypes = [s, s, s]
    L.amrd_evo_iterate.argtypes = [i, C.c_void_p, C.c_void_p, i, i, d, i, _dp, _dp]
    L.amrd_mg_def_cnst_data_vars.argtypes = [s]
    L.amrd_set_is_t0.argtypes = [i]
    L.amrd_set_is_t0.restype = None
    L.amrd_set_elliptic_vars_t0_init.argtypes = [C.c_void_p]
    L.amrd_set_elliptic_vars_t0_init.restype = None
    L.amrd_set_app_t0_cnst_data.argtypes = [C.c_void_p]
    L.amrd_set_app_t0_cnst_data.restype = None
    L.amrd_mg_extrap.argtypes = [i, i, d, d, d, d]
    L.amrd_set_mg_brs_vars.argtypes = [i, i]
    L.amrd_update_mg_extrap_vars.argtypes = [i, i, d]
    L.amrd_mg_last_norms.argtypes = [_dp, _ip]
    L.amrd_gpu_stream.argtypes = []
    L.amrd_gpu_stream.restype = _vp
    L.PAMR_save_gfn.argtypes = [s, i, i, i, d, s, s]
    L.PAMR_save_gfn_close.argtypes = [s, i, i, i, d, s, s]
    L.PAMR_set_gft_writer.argtypes = [_vp, _vp]
    L.PAMR_set_gft_writer.restype = None
    L.amrd_mg_set_trace.argtypes = [i, i, d, d]
    L.amrd_mg_set_trace.restype = None
    L.amrd_mg_def_dump_vars.argtypes = [s, s, s]
    L.mg_dump.argtypes = [i, s, i, i]
    L.mg_dump.restype = None


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with code {rc}")


def P(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def tptr(t):
    """device pointer of a torch tensor (None -> NULL)."""
    return None if t is None else C.c_void_p(t.data_ptr())


class Solver:
    """Owning wrapper of a ``pmg_solver`` (device-resident MG hierarchy)."""

    def __init__(self, n, bbox=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0), min_cwidth=5, op=0, do_ex=0,
                 ex=1.0, shape=None):
        L = lib()
        nx, ny, nz = shape if shape is not None else (n, n, n)
        bb = np.ascontiguousarray(bbox, dtype=np.float64)
        self._L = L
        self.h = L.pmg_create(nx, ny, nz, P(bb), min_cwidth, op, do_ex, ex)
        if not self.h:
            raise RuntimeError("pmg_create failed")
        self.nlev = L.pmg_nlev(self.h)
        self.op, self.do_ex = op, do_ex
        self.apply_tuning()

    def apply_tuning(self):
        """Per-level sweep tile/z-chunk of the FP64 and the FP32 sweeps from
        the autotuner table (tuning.json): the entry tuned for this operator
        and excision when there is one, else the plain Poisson entry."""
        from .autotune import tuned_level_config
        for lev in range(1, self.nlev + 1):
            for pi, prec in enumerate(("f64", "f32")):
                cfg = tuned_level_config(self.shape(lev)[2], prec, self.op, self.do_ex)
                if cfg is not None:
                    _check(self._L.pmg_set_level_sweep_config_prec(self.h, lev, pi, *cfg),
                           "pmg_set_level_sweep_config_prec")

    def set_level_sweep_config(self, lev, variant, zchunk):
        _check(self._L.pmg_set_level_sweep_config(self.h, lev, variant, zchunk),
               "pmg_set_level_sweep_config")

    def close(self):
        if getattr(self, "h", None):
            self._L.pmg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, lev):
        s = (C.c_int * 3)()
        _check(self._L.pmg_level_shape(self.h, lev, s), "pmg_level_shape")
        return (s[2], s[1], s[0])

    def set_params(self, pre=2, pst=2, max_citer=5000, crtol=1e-3, w0_r=1.0, w0_i=1.0, rop_u=3,
                   rop_res=3, rop_src=3, interp_ord=2):
        _check(self._L.pmg_set_params(self.h, pre, pst, max_citer, crtol, w0_r, w0_i, rop_u,
                                      rop_res, rop_src, interp_ord), "pmg_set_params")

    def upload(self, lev, which, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == self.shape(lev)
        _check(self._L.pmg_upload(self.h, lev, which, P(a)), "pmg_upload")

    def download(self, lev, which):
        a = np.zeros(self.shape(lev))
        _check(self._L.pmg_download(self.h, lev, which, P(a)), "pmg_download")
        return a

    def set_chr(self, lev, chr_):
        a = np.ascontiguousarray(chr_, dtype=np.float64)
        _check(self._L.pmg_set_chr(self.h, lev, P(a)), "pmg_set_chr")

    def cmask(self, lev, kind):
        a = np.zeros(self.shape(lev))
        _check(self._L.pmg_download_cmask(self.h, lev, kind, P(a)), "pmg_download_cmask")
        return a

    def init(self):
        _check(self._L.pmg_init(self.h), "pmg_init")

    def vcycle(self):
        r = C.c_double(0)
        norms = np.zeros(self.nlev)
        ci = C.c_int(0)
        _check(self._L.pmg_vcycle(self.h, C.byref(r), P(norms), C.byref(ci)), "pmg_vcycle")
        return r.value, norms, ci.value

    def residual(self):
        r = C.c_double(0)
        _check(self._L.pmg_residual(self.h, C.byref(r)), "pmg_residual")
        return r.value

    def solve(self, min_iter, max_iter, tol=0.0):
        out = np.zeros(max(max_iter, 1))
        it = C.c_int(0)
        _check(self._L.pmg_solve(self.h, min_iter, max_iter, tol, P(out), C.byref(it)), "pmg_solve")
        return out[: it.value]

    def solve_rel(self, mode, rtol, max_iter, r_ref=0.0):
        """(reference norm, residual history); r_ref > 0 replaces the
        residual of the initial u as the reference of the relative test."""
        hist = np.zeros(max(max_iter, 1))
        it = C.c_int(0)
        r0 = C.c_double(r_ref)
        _check(self._L.pmg_solve_rel(self.h, mode, rtol, max_iter, P(hist), C.byref(it),
                                     C.byref(r0)), "pmg_solve_rel")
        return r0.value, hist[: it.value]

    def zero_solution(self):
        _check(self._L.pmg_zero_solution(self.h), "pmg_zero_solution")

    def solve_fgmres(self, m, rtol, max_iter, r_ref=0.0):
        """(reference norm, residual history) of pmg_solve_fgmres."""
        hist = np.zeros(max(max_iter, 1))
        it = C.c_int(0)
        r0 = C.c_double(r_ref)
        _check(self._L.pmg_solve_fgmres(self.h, m, rtol, max_iter, P(hist), C.byref(it), C.byref(r0)),
               "pmg_solve_fgmres")
        return r0.value, hist[: it.value]

    def set_fgmres_prec_cycles(self, cycles):
        _check(self._L.pmg_set_fgmres_prec_cycles(self.h, cycles), "pmg_set_fgmres_prec_cycles")

    def history_warm_start(self, F, U, misfit=True):
        """u = sum c_k U[k] minimising ||rho - sum c_k F[k]|| (device
        pointers, newest first); returns (c, relative misfit or None)."""
        K = len(F)
        assert K == len(U) and 1 <= K <= 8
        fa = (C.c_void_p * K)(*F)
        ua = (C.c_void_p * K)(*U)
        c = np.zeros(K)
        rel = C.c_double(0)
        _check(self._L.pmg_history_warm_start(self.h, K, fa, ua, P(c), C.byref(rel) if misfit else None),
               "pmg_history_warm_start")
        return c, (rel.value if misfit else None)

    def time_kernel(self, which, lev, prec, nrep):
        ms = C.c_double(0)
        _check(self._L.pmg_time_kernel(self.h, which, lev, prec, nrep, C.byref(ms)), "pmg_time_kernel")
        return ms.value

    def set_sweep_config(self, variant, zchunk):
        _check(self._L.pmg_set_sweep_config(self.h, variant, zchunk), "pmg_set_sweep_config")

    def set_graph(self, on):
        _check(self._L.pmg_set_graph(self.h, 1 if on else 0), "pmg_set_graph")

    def set_bottom_fusion(self, on, max_n=None):
        """fuse the bottom levels into one workgroup (max_n: largest side fused)"""
        v = (max_n if max_n is not None else 1) if on else 0
        _check(self._L.pmg_set_bottom_fusion(self.h, v), "pmg_set_bottom_fusion")

    def set_fp32_exact_bottom(self, on):
        _check(self._L.pmg_set_fp32_exact_bottom(self.h, 1 if on else 0), "pmg_set_fp32_exact_bottom")

    def set_bottom_lds(self, on):
        """keep the fused bottom levels' arrays in LDS when they fit"""
        _check(self._L.pmg_set_bottom_lds(self.h, 1 if on else 0), "pmg_set_bottom_lds")

    def set_restrict_fusion(self, on):
        """0 off, 1 (True) the ring kernel, 2 the restriction streamed through the z march,
        3 the same with two planes in flight"""
        _check(self._L.pmg_set_restrict_fusion(self.h, int(on)), "pmg_set_restrict_fusion")

    def set_resid_restrict_zchunk(self, prec, czchunk):
        """coarse planes per workgroup of the fused finest residual + injection"""
        _check(self._L.pmg_set_resid_restrict_zchunk(self.h, prec, czchunk), "pmg_set_resid_restrict_zchunk")

    def set_sweep_blocking(self, on, zchunk=0):
        _check(self._L.pmg_set_sweep_blocking(self.h, 1 if on else 0, zchunk), "pmg_set_sweep_blocking")

    def set_net(self, net):
        """attach a trained network: solve_rel mode 2 applies its correction"""
        _check(self._L.pmg_set_net(self.h, net.h if net is not None else None), "pmg_set_net")

    def net_stats(self):
        a, r = C.c_int(0), C.c_int(0)
        tn, tv = C.c_double(0), C.c_double(0)
        _check(self._L.pmg_net_stats(self.h, C.byref(a), C.byref(r), C.byref(tn), C.byref(tv)), "pmg_net_stats")
        return {"accepted": a.value, "rejected": r.value, "t_net_ms": tn.value, "t_vcycle_ms": tv.value}

    # -- learned band correction (excision) --
    def band_setup(self, depth=3, core=26):
        nc = self._L.pmg_band_setup(self.h, depth, core)
        if nc < 0:
            raise RuntimeError("pmg_band_setup failed")
        return nc

    def band_info(self):
        np_, nc, side = C.c_longlong(0), C.c_int(0), C.c_int(0)
        _check(self._L.pmg_band_info(self.h, C.byref(np_), C.byref(nc), C.byref(side)), "pmg_band_info")
        return np_.value, nc.value, side.value

    def band_mask(self):
        out = np.zeros(self.shape(self.nlev), dtype=np.uint8)
        _check(self._L.pmg_band_download(self.h, out.ctypes.data_as(C.POINTER(C.c_ubyte))), "pmg_band_download")
        return out

    def set_band_net(self, net, ts, use_dc=True, use_fgmres=False):
        self._band_net = net
        _check(self._L.pmg_set_band_net(self.h, net.h if net is not None else None, ts, int(use_dc),
                                        int(use_fgmres)), "pmg_set_band_net")

    def band_enable(self, use_dc, use_fgmres):
        """use_fgmres: False, True (policy: kept while faster), or 2 (always)"""
        _check(self._L.pmg_band_enable(self.h, int(use_dc), int(use_fgmres)), "pmg_band_enable")

    def set_band_net_inner(self, net, ts):
        self._band_net_in = net
        _check(self._L.pmg_set_band_net_inner(self.h, net.h if net is not None else None, ts),
               "pmg_set_band_net_inner")

    def band_set_inner(self, on):
        _check(self._L.pmg_band_set_inner(self.h, int(on)), "pmg_band_set_inner")

    def band_sample_inner(self, ustar_ptr, ncrop, seed, x_ptr, t_ptr, w_ptr):
        r = C.c_double(0)
        _check(self._L.pmg_band_sample_inner(self.h, ustar_ptr, ncrop, seed, x_ptr, t_ptr, w_ptr, C.byref(r)),
               "pmg_band_sample_inner")
        return r.value

    def band_stats(self):
        a, r = C.c_int(0), C.c_int(0)
        out = np.zeros(6)
        _check(self._L.pmg_band_stats(self.h, C.byref(a), C.byref(r), P(out)), "pmg_band_stats")
        return {"accepted": a.value, "rejected": r.value, "rho_plain": out[0], "ms_plain": out[1],
                "rho_net": out[2], "ms_net": out[3], "fgmres_rate_plain": out[4], "fgmres_rate_net": out[5]}

    def band_sample(self, ustar_ptr, xs, ts, ncrop, seed, x_ptr, t_ptr, w_ptr):
        _check(self._L.pmg_band_sample(self.h, ustar_ptr, xs, ts, ncrop, seed, x_ptr, t_ptr, w_ptr),
               "pmg_band_sample")

    def band_correct(self, which=0):
        _check(self._L.pmg_band_correct(self.h, which), "pmg_band_correct")

    def download32(self, lev, which):
        out = np.zeros(self.shape(lev), dtype=np.float32)
        _check(self._L.pmg_download32(self.h, lev, which, out.ctypes.data_as(C.POINTER(C.c_float))),
               "pmg_download32")
        return out

    def upload32(self, lev, which, arr):
        a = np.ascontiguousarray(arr, dtype=np.float32)
        _check(self._L.pmg_upload32(self.h, lev, which, a.ctypes.data_as(C.POINTER(C.c_float))), "pmg_upload32")

    def field_ptr(self, lev, which):
        return self._L.pmg_field(self.h, lev, which)

    def stream(self):
        return self._L.pmg_stream(self.h)


class Amr:
    """Owning wrapper of a ``pmg_amr``: the batched multigrid solve over an
    AMR hierarchy.  ``levels`` lists, per refined AMR level, its physical
    boxes (x0, x1, y0, y1, z0, z1); the base grid covers ``bbox``."""

    def __init__(self, base_n, levels, bbox=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0), mcw=5, op=0):
        self._L = lib()
        shape = (C.c_int * 3)(base_n, base_n, base_n)
        bb = np.array(bbox, dtype=np.float64)
        nbox = (C.c_int * max(1, len(levels)))(*[len(l) for l in levels])
        boxes = np.array([b for l in levels for b in l] or [[0.0] * 6], dtype=np.float64).reshape(-1)
        self.h = self._L.pmg_amr_create(shape, P(bb), len(levels), nbox, P(boxes), mcw, op)
        if not self.h:
            raise RuntimeError("pmg_amr_create failed")
        self.nlev = self._L.pmg_amr_nlev(self.h)
        self.base = self._L.pmg_amr_base_level(self.h)

    def close(self):
        if getattr(self, "h", None):
            self._L.pmg_amr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def ngrids(self, lev):
        return self._L.pmg_amr_ngrids(self.h, lev)

    def grid(self, lev, g):
        s = (C.c_int * 3)()
        b = np.zeros(6)
        _check(self._L.pmg_amr_grid_info(self.h, lev, g, s, P(b)), "pmg_amr_grid_info")
        return (s[2], s[1], s[0]), b

    def upload(self, lev, g, which, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == self.grid(lev, g)[0]
        _check(self._L.pmg_amr_upload(self.h, lev, g, which, P(a)), "pmg_amr_upload")

    def download(self, lev, g, which):
        out = np.zeros(self.grid(lev, g)[0])
        _check(self._L.pmg_amr_download(self.h, lev, g, which, P(out)), "pmg_amr_download")
        return out

    def set_params(self, pre=2, pst=2, max_citer=5000, crtol=1e-3, w0_r=1.0, w0_i=1.0,
                   rop_u=3, rop_res=3, rop_src=3, interp_ord=2):
        _check(self._L.pmg_amr_set_params(self.h, pre, pst, max_citer, crtol, w0_r, w0_i, rop_u,
                                          rop_res, rop_src, interp_ord), "pmg_amr_set_params")

    def init(self):
        _check(self._L.pmg_amr_init(self.h), "pmg_amr_init")

    def vcycle(self):
        r = C.c_double(0)
        norms = np.zeros(self.nlev)
        cit = C.c_int(0)
        _check(self._L.pmg_amr_vcycle(self.h, C.byref(r), P(norms), C.byref(cit)), "pmg_amr_vcycle")
        return r.value, norms, cit.value

    def residual(self):
        r = C.c_double(0)
        _check(self._L.pmg_amr_residual(self.h, C.byref(r)), "pmg_amr_residual")
        return r.value

    def solve_rel(self, mode, rtol, max_iter, r_ref=0.0):
        """(reference norm, residual history) of pmg_amr_solve_rel: mode 0
        FP64 V-cycles, 1 FP32 V-cycles in FP64 defect correction"""
        hist = np.zeros(max(max_iter, 1))
        it = C.c_int(0)
        r0 = C.c_double(r_ref)
        _check(self._L.pmg_amr_solve_rel(self.h, mode, rtol, max_iter, P(hist), C.byref(it), C.byref(r0)),
               "pmg_amr_solve_rel")
        return r0.value, hist[: it.value]

    def set_mixed_switch(self, frac):
        """mode 1 runs FP64 V-cycles until ||r|| <= frac ||r0||, then FP32 inner cycles"""
        _check(self._L.pmg_amr_set_mixed_switch(self.h, frac), "pmg_amr_set_mixed_switch")

    def sync(self, lev, bdy, which):
        _check(self._L.pmg_amr_sync(self.h, lev, bdy, which), "pmg_amr_sync")

    def inject(self, lev_fine, which, op=3):
        _check(self._L.pmg_amr_inject(self.h, lev_fine, which, op), "pmg_amr_inject")

    def interp(self, lev_coarse, which, ord=2):
        _check(self._L.pmg_amr_interp_ord(self.h, lev_coarse, which, ord), "pmg_amr_interp_ord")

    def cmask(self, lev, g, which):
        out = np.zeros(self.grid(lev, g)[0])
        _check(self._L.pmg_amr_cmask(self.h, lev, g, which, P(out)), "pmg_amr_cmask")
        return out

    def plan_stats(self, lev):
        out = (C.c_int * 4)()
        _check(self._L.pmg_amr_plan_stats(self.h, lev, out), "pmg_amr_plan_stats")
        return list(out)

    def time_sweep(self, lev, nrep=10):
        ms = C.c_double(0)
        _check(self._L.pmg_amr_time_sweep(self.h, lev, nrep, C.byref(ms)), "pmg_amr_time_sweep")
        return ms.value

    def set_graph(self, on):
        _check(self._L.pmg_amr_set_graph(self.h, 1 if on else 0), "pmg_amr_set_graph")

    def set_level_sweep_config(self, lev, variant, zchunk):
        """Sweep tiling of level ``lev``; RuntimeError (code -3) when the
        level would need more tiles than the norm workspace holds."""
        _check(self._L.pmg_amr_set_level_sweep_config(self.h, lev, variant, zchunk),
               "pmg_amr_set_level_sweep_config")

    def sweep_config(self, lev):
        v, z = C.c_int(0), C.c_int(0)
        _check(self._L.pmg_amr_sweep_config(self.h, lev, C.byref(v), C.byref(z)), "pmg_amr_sweep_config")
        return v.value, z.value


class Net:
    """Owning wrapper of a ``pmg_net`` (the learned smoother network)."""

    SHAPES = {"w1": (16, 1, 27), "b1": (16,), "w2": (16, 16, 27), "b2": (16,), "w3": (1, 16, 27),
              "b3": (1,)}

    def __init__(self, seed=1):
        self._L = lib()
        self.h = self._L.pmg_net_create(16, seed)
        if not self.h:
            raise RuntimeError("pmg_net_create failed")

    def close(self):
        if getattr(self, "h", None):
            self._L.pmg_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_weights(self):
        ws = {k: np.zeros(v) for k, v in self.SHAPES.items()}
        _check(self._L.pmg_net_get_weights(self.h, *[P(ws[k]) for k in ("w1", "b1", "w2", "b2", "w3", "b3")]),
               "pmg_net_get_weights")
        return ws

    def set_weights(self, ws):
        arr = [np.ascontiguousarray(ws[k], dtype=np.float64) for k in ("w1", "b1", "w2", "b2", "w3", "b3")]
        _check(self._L.pmg_net_set_weights(self.h, *[P(a) for a in arr]), "pmg_net_set_weights")

    def forward(self, x_dev, y_dev, shape, nb=1):
        nz, ny, nx = shape
        _check(self._L.pmg_net_forward(self.h, tptr(x_dev), tptr(y_dev), nx, ny, nz, nb), "pmg_net_forward")

    def infer(self, x_dev, y_dev, shape, nb=1, scale=1.0):
        """y = N(scale x) through the inference path of pmg_net_correct."""
        nz, ny, nx = shape
        _check(self._L.pmg_net_infer(self.h, tptr(x_dev), scale, tptr(y_dev), nx, ny, nz, nb), "pmg_net_infer")

    def loss_grad(self, x_dev, t_dev, shape, nb=1, margin=0):
        nz, ny, nx = shape
        loss = C.c_double(0)
        gs = {k: np.zeros(v) for k, v in self.SHAPES.items()}
        _check(self._L.pmg_net_loss_grad(self.h, tptr(x_dev), tptr(t_dev), nx, ny, nz, nb, margin,
                                         C.byref(loss),
                                         *[P(gs[k]) for k in ("w1", "b1", "w2", "b2", "w3", "b3")]),
               "pmg_net_loss_grad")
        return loss.value, gs

    def train_step(self, x_dev, t_dev, shape, nb, lr, margin=0, w_dev=None):
        """One Adam step; w_dev (optional): 0/1 weights selecting the points
        the mean loss runs over (besides the margin rule)."""
        nz, ny, nx = shape
        loss = C.c_double(0)
        if w_dev is None:
            _check(self._L.pmg_net_train_step(self.h, tptr(x_dev), tptr(t_dev), nx, ny, nz, nb, margin,
                                              lr, C.byref(loss)), "pmg_net_train_step")
        else:
            _check(self._L.pmg_net_train_step_w(self.h, tptr(x_dev), tptr(t_dev), tptr(w_dev), nx, ny, nz, nb,
                                                margin, lr, C.byref(loss)), "pmg_net_train_step_w")
        return loss.value

    def loss_grad_w(self, x_dev, t_dev, w_dev, shape, nb=1, margin=0):
        nz, ny, nx = shape
        loss = C.c_double(0)
        gs = {k: np.zeros(v) for k, v in self.SHAPES.items()}
        _check(self._L.pmg_net_loss_grad_w(self.h, tptr(x_dev), tptr(t_dev), tptr(w_dev), nx, ny, nz, nb, margin,
                                           C.byref(loss),
                                           *[P(gs[k]) for k in ("w1", "b1", "w2", "b2", "w3", "b3")]),
               "pmg_net_loss_grad_w")
        return loss.value, gs


def mfma_selftest(A, B):
    """D = A B through one MFMA: A 16x16, B 16x16 (K = 16) or A 16x32, B 32x16 (K = 32)."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    D = np.zeros((16, 16), dtype=np.float32)
    fp = C.POINTER(C.c_float)
    f = lib().pmg_net_mfma32_selftest if A.shape[1] == 32 else lib().pmg_net_mfma_selftest
    _check(f(A.ctypes.data_as(fp), B.ctypes.data_as(fp), D.ctypes.data_as(fp)), "pmg_net_mfma_selftest")
    return D


# --------------------------------------------------------------------------
# synthetic workload generators (host side, deterministic)
# --------------------------------------------------------------------------
def splitmix64(seed):
    """Generator of uint64 values, splitmix64 (Steele, Lea, Flood 2014)."""
    state = np.uint64(seed)
    M = np.uint64(0xFFFFFFFFFFFFFFFF)
    while True:
        with np.errstate(over="ignore"):
            state = (state + np.uint64(0x9E3779B97F4A7C15)) & M
            z = state
            z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & M
            z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & M
            z = z ^ (z >> np.uint64(31))
        yield z


def uniforms(seed, n):
    g = splitmix64(seed)
    return np.array([float(int(next(g)) >> 11) * (1.0 / 9007199254740992.0) for _ in range(n)])


def gaussian_params(seed, m=8):
    """8 Gaussians: centres U[0.2,0.8]^3, widths U[0.05,0.15], amplitudes U[-1,1]."""
    u = uniforms(seed, 5 * m).reshape(m, 5)
    c = 0.2 + 0.6 * u[:, 0:3]
    w = 0.05 + 0.10 * u[:, 3]
    A = -1.0 + 2.0 * u[:, 4]
    return c, w, A


def gaussian_source(n, seed=1234, m=8, params=None, bbox=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0)):
    """f(x) = sum_m A_m exp(-(|x-c_m|/w_m)^2) on an n^3 vertex grid (numpy (z,y,x)).

    This is the reference's gauss3d form (examples/nbs/num.f:13-39) with r0=0
    and no eccentricity; boundary points are set to 0."""
    c, w, A = params if params is not None else gaussian_params(seed, m)
    x = np.linspace(bbox[0], bbox[1], n)
    X, Y, Z = x[None, None, :], x[None, :, None], x[:, None, None]
    f = np.zeros((n, n, n))
    for k in range(len(A)):
        r2 = (X - c[k, 0]) ** 2 + (Y - c[k, 1]) ** 2 + (Z - c[k, 2]) ** 2
        f += A[k] * np.exp(-r2 / (w[k] * w[k]))
    f[0, :, :] = f[-1, :, :] = 0
    f[:, 0, :] = f[:, -1, :] = 0
    f[:, :, 0] = f[:, :, -1] = 0
    return f


from .autotune import tuned_sweep_config  # noqa: E402,F401
