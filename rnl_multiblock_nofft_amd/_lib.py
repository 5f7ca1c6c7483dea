"""ctypes binding of liblhgpu.so (the C-ABI declared in include/lhgpu.h).

The HIP library is the only compute path: importing a product entry point
without a built liblhgpu.so raises, there is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblhgpu.so")

LH_U1, LH_U2, LH_U3, LH_P = 0, 1, 2, 3
LH_NU1, LH_NU2, LH_NU3 = 4, 5, 6
LH_R1, LH_R2, LH_R3 = 7, 8, 9
LH_PSI, LH_DIV = 10, 11
FIELDS = {"u1": LH_U1, "u2": LH_U2, "u3": LH_U3, "p": LH_P, "Nu1": LH_NU1,
          "Nu2": LH_NU2, "Nu3": LH_NU3, "r1": LH_R1, "r2": LH_R2, "r3": LH_R3,
          "psi": LH_PSI, "div": LH_DIV}
PHASE_NONLINEAR, PHASE_RHS, PHASE_SOLVE = 1, 2, 3
TUNE_SOLVE_COLS, TUNE_ZINV_POINTS, TUNE_XPROD_POINTS, TUNE_ZFWD_POINTS = 1, 2, 3, 4
TUNE_TEX_SEPARABLE = 5
TUNE_SOLVE_SCAN = 6
TUNE_SOLVE_WAVES = 7
TUNE_XPROD_INPLACE = 8
TUNE_ZFWD_INPLACE = 9
TUNE_XPROD_THREADS = 10
TUNE_RHS_FUSED = 11
TUNE_GHOST_IN_ZINV = 12
TUNE_ZINV_INPLACE = 13
TUNE_ZFWD_FIELDWISE = 14
TUNE_TEX_FUSED = 15
TUNE_SHAPE_GROUPS = 16
PREC_FP64, PREC_MIXED = 0, 1

# every symbol include/lhgpu.h declares
EXPORTS = ["lh_last_error", "lh_ctx_create", "lh_ctx_destroy", "lh_set_grid", "lh_get_grid",
           "lh_field_upload", "lh_field_download", "lh_get_scalars", "lh_set_scalars",
           "lh_init_flowrate", "lh_substep", "lh_step", "lh_sync", "lh_time_kernels", "lh_run_phase",
           "lh_four_to_phys_planes", "lh_phys_to_four_planes", "lh_plane_transform_noalloc",
           "lh_cft", "lh_rft",
           "lh_lusol_uw", "lh_lusol_v",
           "lh_dev_alloc", "lh_dev_free", "lh_copy_to_device", "lh_copy_to_host",
           "lh_set_tuning", "lh_get_tuning", "lh_autotune", "lh_set_texture",
           "lh_stats_accumulate", "lh_stats_reset", "lh_stats_get", "lh_spectrum_get",
           "lh_lusol_p", "lh_lusol0", "lh_four_to_phys_N_planes", "lh_phys_to_four_N_planes",
           "lh_modes_to_planes", "lh_planes_to_modes", "lh_set_triads", "lh_history", "lh_slip_stats_get",
           "lh_closure_set", "lh_closure_enable", "lh_closure_eval", "lh_closure_sample", "lh_closure_time",
           "lh_mlp_create", "lh_mlp_destroy", "lh_mlp_set_params", "lh_mlp_get_params", "lh_mlp_set_batch",
           "lh_mlp_sample", "lh_mlp_set_points", "lh_mlp_forward", "lh_mlp_grad", "lh_mlp_get_grad",
           "lh_mlp_adam", "lh_mlp_sync", "lh_band_rows", "lh_set_bands"]
CL_NIN, CL_NHID, CL_NOUT, CL_NPARAM = 135, 128, 3, 50819
ST_N = 13
SP_IDS = {"U": 0, "V": 1, "W": 2, "UV": 3, "P": 4}


class LhConfig(ctypes.Structure):
    _fields_ = [("N1", ctypes.c_int), ("N2", ctypes.c_int),
                ("Ngal1", ctypes.c_int), ("Ngal2", ctypes.c_int), ("ny", ctypes.c_int),
                ("alp", ctypes.c_double), ("bet", ctypes.c_double), ("Re", ctypes.c_double),
                ("grid_kind", ctypes.c_int), ("dyq", ctypes.c_double), ("ppp", ctypes.c_int),
                ("gamma", ctypes.c_double), ("bslpu1", ctypes.c_double),
                ("bslpu3", ctypes.c_double), ("flag_ctpress", ctypes.c_int),
                ("mpgx", ctypes.c_double), ("dt", ctypes.c_double), ("precision", ctypes.c_int),
                ("adaptive_dt", ctypes.c_int), ("CFL", ctypes.c_double)]


class LhError(RuntimeError):
    pass


_lib = None


def load():
    """Load liblhgpu.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LhError("liblhgpu.so is not built (run __graft_entry__.build() or "
                      "python rnl_multiblock_nofft_amd/build.py)")
    L = ctypes.CDLL(LIB_PATH)
    vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
    L.lh_last_error.restype = ctypes.c_char_p
    L.lh_ctx_create.argtypes = [ctypes.POINTER(LhConfig), ip, ctypes.POINTER(vp)]
    L.lh_ctx_destroy.argtypes = [vp]
    L.lh_set_grid.argtypes = [vp, ip, dp, dp, dp, dp]
    L.lh_band_rows.argtypes = [ctypes.POINTER(LhConfig), ctypes.c_double, ctypes.POINTER(ctypes.c_int)]
    L.lh_set_bands.argtypes = [vp, ip, ip, ip, ip]
    L.lh_get_grid.argtypes = [vp, ip, dp, dp, dp, dp]
    L.lh_field_upload.argtypes = [vp, ip, ip, dp]
    L.lh_field_download.argtypes = [vp, ip, ip, dp]
    L.lh_get_scalars.argtypes = [vp, ip, dp]
    L.lh_set_scalars.argtypes = [vp, ip, ctypes.c_double, ctypes.c_double, ctypes.c_double, ip]
    L.lh_init_flowrate.argtypes = [vp]
    L.lh_substep.argtypes = [vp, ip]
    L.lh_step.argtypes = [vp, ip]
    L.lh_sync.argtypes = [vp]
    L.lh_time_kernels.argtypes = [vp, ip, dp]
    L.lh_run_phase.argtypes = [vp, ip, ip]
    L.lh_four_to_phys_planes.argtypes = [ip, ip, ip, vp, ip, vp]
    L.lh_phys_to_four_planes.argtypes = [ip, ip, ip, vp, ip, vp]
    L.lh_plane_transform_noalloc.argtypes = [ip, ip, ip, ip, vp, ip, vp]
    L.lh_cft.argtypes = [vp, ip, ip, ip, ip, ip, vp]
    L.lh_rft.argtypes = [vp, ip, ip, ip, ip, vp]
    L.lh_lusol_uw.argtypes = [vp, ip, ip, vp, vp, vp, vp]
    L.lh_lusol_v.argtypes = [vp, ip, ip, vp, vp]
    L.lh_dev_alloc.argtypes = [ctypes.c_size_t, ctypes.POINTER(vp)]
    L.lh_dev_free.argtypes = [vp]
    L.lh_copy_to_device.argtypes = [vp, vp, ctypes.c_size_t]
    L.lh_copy_to_host.argtypes = [vp, vp, ctypes.c_size_t]
    L.lh_set_tuning.argtypes = [vp, ip, ip]
    L.lh_get_tuning.argtypes = [vp, ip, ctypes.POINTER(ip)]
    L.lh_autotune.argtypes = [vp, ip]
    L.lh_set_texture.argtypes = [vp, ip, ctypes.POINTER(ip), ctypes.POINTER(ip)]
    L.lh_stats_accumulate.argtypes = [vp]
    L.lh_stats_reset.argtypes = [vp]
    L.lh_stats_get.argtypes = [vp, ip, dp, ctypes.POINTER(ip)]
    L.lh_spectrum_get.argtypes = [vp, ip, ip, dp]
    L.lh_history.argtypes = [vp, ip, dp]
    L.lh_slip_stats_get.argtypes = [vp, ip, ip, dp]
    L.lh_set_triads.argtypes = [vp, ip, ip, ctypes.POINTER(ip), vp, ctypes.POINTER(ip), ctypes.POINTER(ip)]
    L.lh_lusol_p.argtypes = [vp, ip, vp]
    L.lh_lusol0.argtypes = [vp, ip, ip, vp]
    L.lh_four_to_phys_N_planes.argtypes = [ip, ip, vp, ip, vp]
    L.lh_phys_to_four_N_planes.argtypes = [ip, ip, vp, ip, vp]
    L.lh_modes_to_planes.argtypes = [ip, ip, ip, ip, vp, ip, vp, vp]
    L.lh_planes_to_modes.argtypes = [ip, ip, ip, ip, vp, ip, vp, vp]
    fp = ctypes.POINTER(ctypes.c_float)
    L.lh_closure_set.argtypes = [vp, fp, fp]
    L.lh_closure_enable.argtypes = [vp, ip, ip]
    L.lh_closure_eval.argtypes = [vp, ip, vp, vp]
    L.lh_closure_sample.argtypes = [vp, ip, vp, ip, vp, vp]
    L.lh_closure_time.argtypes = [vp, ip, dp]
    L.lh_mlp_create.argtypes = [ip, ctypes.POINTER(vp)]
    L.lh_mlp_destroy.argtypes = [vp]
    L.lh_mlp_set_params.argtypes = [vp, fp, fp]
    L.lh_mlp_get_params.argtypes = [vp, fp]
    L.lh_mlp_set_batch.argtypes = [vp, fp, fp]
    L.lh_mlp_sample.argtypes = [vp, ip, ip, ip, vp, vp, ip, ctypes.c_uint64]
    L.lh_mlp_set_points.argtypes = [vp, ip, ip, ip, vp, ctypes.POINTER(ip)]
    L.lh_mlp_forward.argtypes = [vp, fp]
    L.lh_mlp_grad.argtypes = [vp, dp]
    L.lh_mlp_get_grad.argtypes = [vp, fp]
    L.lh_mlp_adam.argtypes = [vp, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                              ctypes.c_float]
    L.lh_mlp_sync.argtypes = [vp]
    for name in EXPORTS:
        if name != "lh_last_error":
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise LhError("liblhgpu error %d: %s" % (rc, load().lh_last_error().decode()))
    return rc
