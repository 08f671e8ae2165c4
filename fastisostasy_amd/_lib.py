"""ctypes binding of libfastiso.so (the C ABI declared in include/fastiso.h).

The shared library is built in-tree by fastisostasy_amd.build.  Loading it
does not need a GPU; every compute entry point fails with FI_E_HIP when no HIP
device is present (the library has no CPU path).
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfastiso.so")
MAXL = 8

FI_OK, FI_E_ARG, FI_E_UNSUPPORTED, FI_E_HIP, FI_E_UNSTABLE, FI_E_NAN, FI_E_STATE = 0, -1, -2, -3, -4, -5, -6

FIELDS = ["z_bed", "z_ss", "dz_ss", "w", "we", "dwdt", "canom_full", "Hice",
          "maskocean", "maskgrounded", "maskcontinent",
          "canom_mantle", "Haf", "eta_eff", "D_lith", "kappa", "R", "maskactive", "He_lith",
          "ode_x", "z_ss_ref", "z_bed_ref", "w_equilibrium"]
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}
OUT_FIELDS = FIELDS[:11]
SCALARS = ["n", "t", "dt", "bsl", "deltaV", "V_af", "V_den", "attempts", "accepts", "D_lith11", "err", "tol",
           "hbm_bytes", "status", "nx_ice", "ny_ice"]
SCALAR_ID = {n: i for i, n in enumerate(SCALARS)}

# every entry point of include/fastiso.h
EXPORTS = ["fi_last_error", "fi_version", "fi_padded_size", "fi_isos_create", "fi_isos_create_gridded",
           "fi_isos_init_ref",
           "fi_isos_init_state", "fi_isos_update", "fi_isos_update_noalloc", "fi_isos_get", "fi_isos_get_mask", "fi_isos_get_full",
           "fi_isos_get_scalar", "fi_isos_restart_write", "fi_isos_restart_read", "fi_bsl_restart_write",
           "fi_bsl_restart_read", "fi_isos_destroy", "fi_bsl_create", "fi_bsl_create_series",
           "fi_bsl_read_series", "fi_bsl_update", "fi_bsl_get",
           "fi_bsl_destroy", "fi_batch_create", "fi_batch_update", "fi_batch_destroy", "fi_sync",
           "fi_prof_select", "fi_prof_collect", "fi_prof_collect_kernel", "fi_prof_fused", "fi_prof_shape", "fi_bench_copy",
           "fi_isos_write_init", "fi_isos_write_step", "fi_bsl_write_init", "fi_bsl_write_step", "fi_autotune",
           "fi_server_create", "fi_server_admit", "fi_server_release", "fi_server_submit", "fi_server_pending",
           "fi_server_run", "fi_server_poll", "fi_server_completed", "fi_server_failed", "fi_server_stats", "fi_server_destroy"]


class FiParams(ctypes.Structure):
    _fields_ = [
        ("heterogeneous_ssh", ctypes.c_int), ("interactive_sealevel", ctypes.c_int),
        ("include_elastic", ctypes.c_int), ("correct_compression", ctypes.c_int),
        ("correct_distortion", ctypes.c_int), ("method", ctypes.c_int),
        ("dt_diagnostics", ctypes.c_double), ("min_pad", ctypes.c_double),
        ("mantle", ctypes.c_int), ("lithosphere", ctypes.c_int), ("layering", ctypes.c_int),
        ("lumping", ctypes.c_int), ("viscosity_scaling_method", ctypes.c_int),
        ("viscosity_scaling", ctypes.c_double), ("nl", ctypes.c_int),
        ("zl", ctypes.c_double * MAXL), ("viscosities", ctypes.c_double * MAXL),
        ("dl", ctypes.c_double), ("eta_ref", ctypes.c_double),
        ("rho_water", ctypes.c_double), ("rho_ice", ctypes.c_double),
        ("rho_seawater", ctypes.c_double), ("rho_uppermantle", ctypes.c_double),
        ("rho_litho", ctypes.c_double), ("E", ctypes.c_double), ("nu", ctypes.c_double),
        ("g", ctypes.c_double), ("r_earth", ctypes.c_double), ("m_earth", ctypes.c_double),
        ("dt_method", ctypes.c_int), ("atol", ctypes.c_double), ("rtol", ctypes.c_double),
        ("dt_min", ctypes.c_double), ("dt_max", ctypes.c_double),
        ("dt_init", ctypes.c_double), ("q", ctypes.c_double), ("precision", ctypes.c_int),
        ("tau", ctypes.c_double),
    ]


class FiRheologyGrid(ctypes.Structure):
    """fi_rheology_grid: the arrays of the rheology_file / mask_file inputs."""
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("xc", ctypes.POINTER(ctypes.c_double)), ("yc", ctypes.POINTER(ctypes.c_double)),
        ("litho_thickness", ctypes.POINTER(ctypes.c_double)), ("litho_thickness_scaling", ctypes.c_double),
        ("nz", ctypes.c_int), ("zc", ctypes.POINTER(ctypes.c_double)),
        ("log10_eta", ctypes.POINTER(ctypes.c_double)), ("stddev_log10_eta", ctypes.POINTER(ctypes.c_double)),
        ("rheo_smoothing_radius", ctypes.c_double),
        ("mask_nx", ctypes.c_int), ("mask_ny", ctypes.c_int),
        ("mask_xc", ctypes.POINTER(ctypes.c_double)), ("mask_yc", ctypes.POINTER(ctypes.c_double)),
        ("mask", ctypes.POINTER(ctypes.c_double)),
    ]


_lib = None


def load():
    """Load libfastiso.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m fastisostasy_amd.build` "
                           "(the HIP library is required; there is no fallback)")
    L = ctypes.CDLL(LIB_PATH)
    dp = ctypes.POINTER(ctypes.c_double)
    vp = ctypes.c_void_p
    i = ctypes.c_int
    d = ctypes.c_double
    L.fi_last_error.restype = ctypes.c_char_p
    L.fi_version.restype = i
    L.fi_padded_size.argtypes = [i, i, d, d, d, ctypes.POINTER(i)]
    L.fi_isos_create.argtypes = [ctypes.POINTER(FiParams), i, i, d, d, dp, dp, dp, dp, ctypes.POINTER(vp)]
    L.fi_isos_create_gridded.argtypes = [ctypes.POINTER(FiParams), i, i, d, d, dp, dp,
                                         ctypes.POINTER(FiRheologyGrid), ctypes.POINTER(vp)]
    L.fi_isos_init_ref.argtypes = [vp, dp, dp, dp, dp]
    L.fi_isos_init_state.argtypes = [vp, dp, dp, d, vp]
    L.fi_isos_update.argtypes = [vp, dp, d, vp]
    L.fi_isos_update_noalloc.argtypes = [vp, dp, d, vp]
    L.fi_isos_get.argtypes = [vp, i, dp]
    L.fi_isos_get_mask.argtypes = [vp, i, ctypes.POINTER(ctypes.c_int8)]
    L.fi_isos_get_full.argtypes = [vp, i, dp]
    L.fi_isos_get_scalar.argtypes = [vp, i, dp]
    L.fi_isos_destroy.argtypes = [vp]
    L.fi_isos_restart_write.argtypes = [vp, ctypes.c_char_p, d]
    L.fi_isos_restart_read.argtypes = [vp, ctypes.c_char_p, d]
    L.fi_bsl_restart_write.argtypes = [vp, ctypes.c_char_p, d]
    L.fi_bsl_restart_read.argtypes = [ctypes.c_char_p, dp, dp]
    L.fi_bsl_create.argtypes = [i, d, d, i, dp, dp, d, ctypes.POINTER(vp)]
    L.fi_bsl_create_series.argtypes = [i, d, d, i, dp, dp, i, dp, dp, d, ctypes.POINTER(vp)]
    L.fi_bsl_read_series.argtypes = [ctypes.c_char_p, i, dp, dp, ctypes.POINTER(i)]
    L.fi_bsl_update.argtypes = [vp, d]
    L.fi_bsl_get.argtypes = [vp, dp, dp]
    L.fi_bsl_destroy.argtypes = [vp]
    L.fi_batch_create.argtypes = [ctypes.POINTER(vp), i, ctypes.POINTER(vp)]
    L.fi_batch_update.argtypes = [vp, ctypes.POINTER(dp), d, ctypes.POINTER(vp)]
    L.fi_batch_destroy.argtypes = [vp]
    L.fi_prof_select.argtypes = [i]
    L.fi_prof_collect.argtypes = [dp, ctypes.POINTER(ctypes.c_longlong)]
    L.fi_prof_collect_kernel.argtypes = [i, dp, ctypes.POINTER(ctypes.c_longlong)]
    L.fi_isos_write_init.argtypes = [vp, dp, dp, ctypes.c_char_p, d]
    L.fi_isos_write_step.argtypes = [vp, ctypes.c_char_p, d, dp, dp]
    L.fi_bsl_write_init.argtypes = [vp, ctypes.c_char_p, d]
    L.fi_bsl_write_step.argtypes = [vp, ctypes.c_char_p, d]
    L.fi_autotune.argtypes = [i, i, i, i, ctypes.POINTER(i)]
    L.fi_prof_fused.argtypes = []
    L.fi_prof_shape.argtypes = [ctypes.POINTER(i)]
    L.fi_bench_copy.argtypes = [d, i, dp]
    ll = ctypes.c_longlong
    L.fi_server_create.argtypes = [d, i, ctypes.POINTER(vp)]
    L.fi_server_admit.argtypes = [vp, vp, vp, ctypes.POINTER(i)]
    L.fi_server_release.argtypes = [vp, i]
    L.fi_server_submit.argtypes = [vp, i, dp, d, ll]
    L.fi_server_pending.argtypes = [vp, ctypes.POINTER(i)]
    L.fi_server_run.argtypes = [vp, ctypes.POINTER(i)]
    L.fi_server_poll.argtypes = [vp, i, ctypes.POINTER(i)]
    L.fi_server_completed.argtypes = [vp, ctypes.POINTER(ll), i, ctypes.POINTER(i)]
    L.fi_server_failed.argtypes = [vp, ctypes.POINTER(ll), i, ctypes.POINTER(i)]
    L.fi_server_stats.argtypes = [vp, ctypes.POINTER(ll), ctypes.POINTER(ll), ctypes.POINTER(ll), dp]
    L.fi_server_destroy.argtypes = [vp]
    for name in EXPORTS:
        if name not in ("fi_last_error", "fi_version"):
            getattr(L, name).restype = i
    _lib = L
    return L


class FiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def check(rc):
    if rc != 0:
        raise FiError(rc, load().fi_last_error().decode())
    return rc
