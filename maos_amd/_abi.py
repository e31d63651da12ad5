"""ctypes view of include/maos_tomo.h and the loader of libmaos_tomo.so.

The library is built in-tree by maos_amd.build; importing this module never
falls back to anything else: a missing library raises at load time.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libmaos_tomo.so")

MT_OK = 0
MT_EINVAL = -1
MT_ENOMEM = -2
MT_EHIP = -3
MT_ENOTFINITE = -4
MT_ESTATE = -5
MT_EFULL = -6
MT_NOTCONV = 1

MT_BIN_DOUBLE, MT_BIN_CELL, MT_BIN_COMPLEX, MT_BIN_FLOAT, MT_BIN_FCOMPLEX, MT_BIN_INT, MT_BIN_SPARSE, \
    MT_BIN_CSPARSE = range(8)

dp = C.POINTER(C.c_double)
lp = C.POINTER(C.c_long)
ip = C.POINTER(C.c_int)


class MtPowfs(C.Structure):
    _fields_ = [("dsa", C.c_double), ("dx", C.c_double), ("saat", C.c_double),
                ("nearecon", C.c_double), ("hs", C.c_double), ("trs", C.c_int),
                ("lo", C.c_int), ("skip", C.c_int), ("dtrat", C.c_int)]


class MtWfs(C.Structure):
    _fields_ = [("thetax", C.c_double), ("thetay", C.c_double), ("powfs", C.c_int)]


class MtParms(C.Structure):
    _fields_ = [("aper_d", C.c_double), ("aper_din", C.c_double), ("npsr", C.c_int),
                ("ht", dp), ("wt", dp), ("os", ip), ("atmr_dx", C.c_double),
                ("r0z", C.c_double), ("L0", C.c_double), ("za", C.c_double),
                ("tomo_pos", C.c_int), ("tomo_guard", C.c_int), ("tomo_cone", C.c_int),
                ("tomo_cxxscale", C.c_double), ("tomo_precond", C.c_int),
                ("recon_split", C.c_int), ("tomo_splitlrt", C.c_int),
                ("npowfs", C.c_int), ("powfs", C.POINTER(MtPowfs)),
                ("nwfs", C.c_int), ("wfs", C.POINTER(MtWfs)),
                ("ndir", C.c_int), ("dir_thetax", dp), ("dir_thetay", dp), ("dir_hs", dp),
                ("tomoscale", C.c_double)]


class MtRecon(C.Structure):
    _fields_ = [("npsr", C.c_int), ("xnx", lp), ("xny", lp), ("xox", dp), ("xoy", dp),
                ("xdx", dp), ("ht", dp), ("pnx", C.c_long), ("pny", C.c_long),
                ("pox", C.c_double), ("poy", C.c_double), ("pdx", C.c_double),
                ("nwfs", C.c_int), ("thetax", dp), ("thetay", dp), ("hs", dp), ("skip", ip),
                ("nsa", lp), ("dsa", dp), ("saloc", C.POINTER(dp)), ("saa", C.POINTER(dp)),
                ("gp_p", C.POINTER(lp)), ("gp_i", C.POINTER(lp)), ("gp_x", C.POINTER(dp)),
                ("pos", C.c_int), ("neai", C.POINTER(dp)), ("trs", ip), ("ptt", C.POINTER(dp)),
                ("ptt_lhs", C.c_int), ("l2c", dp), ("precond", C.c_int), ("fd_bs", C.c_int),
                ("fd_nb", C.c_long), ("fd_nbhalf", C.c_long), ("fd_scale", C.c_int),
                ("fd_Minv", dp), ("fd_perm", lp), ("fd_permhf", lp), ("tomoscale", C.c_double)]


class MtFdpcgIn(C.Structure):
    _fields_ = [("nps", C.c_int), ("nx", lp), ("ny", lp), ("os", lp), ("pos", C.c_int),
                ("dxp", C.c_double), ("ht", dp), ("l2cf", dp), ("nsa", C.c_long), ("dsa", C.c_double),
                ("salocx", dp), ("salocy", dp), ("saa", dp), ("nwfs", C.c_int), ("thetax", dp),
                ("thetay", dp), ("neam", dp)]


class MtFdpcgOut(C.Structure):
    _fields_ = [("bs", C.c_int), ("nb", C.c_long), ("nbhalf", C.c_long), ("scale", C.c_int),
                ("Mhat", dp), ("Minv", dp), ("perm", lp), ("permhf", lp)]


class MtDm(C.Structure):
    _fields_ = [("ht", C.c_double), ("dx", C.c_double), ("offset", C.c_double), ("guard", C.c_double),
                ("iac", C.c_double)]


class MtFitParms(C.Structure):
    _fields_ = [("ndm", C.c_int), ("dm", C.POINTER(MtDm)), ("nfit", C.c_int), ("thetax", dp), ("thetay", dp),
                ("wt", dp), ("hs", dp), ("pos", C.c_int), ("maxit", C.c_int), ("warm", C.c_int),
                ("lrt_piston", C.c_int), ("lrt_tt", C.c_int), ("actslave", C.c_int), ("actthres", C.c_double),
                ("tikcr", C.c_double), ("actinterp", C.c_int), ("annular_W", C.c_int)]


class MtFitRecon(C.Structure):
    _fields_ = [("ndm", C.c_int), ("anx", lp), ("any", lp), ("aox", dp), ("aoy", dp), ("adx", dp), ("aht", dp),
                ("fnx", C.c_long), ("fny", C.c_long), ("fox", C.c_double), ("foy", C.c_double),
                ("fdx", C.c_double), ("nfit", C.c_int), ("thetax", dp), ("thetay", dp), ("wt", dp), ("hs", dp),
                ("W0", dp), ("W1", dp), ("nnw", C.c_int), ("NW", dp), ("sl_p", lp), ("sl_i", lp),
                ("sl_x", dp), ("tikdiag", C.c_double), ("actcpl", dp), ("maxit", C.c_int), ("warm", C.c_int)]


_lib = None


def lib():
    """Load libmaos_tomo.so (built by maos_amd.build); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libmaos_tomo.so is not built (%s); run `python -m maos_amd.build`"
                           % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.mt_last_error.restype = C.c_char_p
    L.mt_setup_create.argtypes = [C.POINTER(MtParms), C.POINTER(vp)]
    L.mt_setup_update_cn2.argtypes = [vp, dp, C.c_double]
    L.mt_setup_recon.argtypes = [vp, C.POINTER(MtRecon)]
    L.mt_setup_free.argtypes = [vp]
    L.mt_setup_free.restype = None
    L.mt_setup_fdpcg_mhat.argtypes = [vp, C.POINTER(dp)]
    L.mt_setup_sizes.argtypes = [vp, lp, lp]
    L.mt_fdpcg_prepare.argtypes = [C.POINTER(MtFdpcgIn), C.POINTER(MtFdpcgOut)]
    L.mt_fdpcg_free.argtypes = [C.POINTER(MtFdpcgOut)]
    L.mt_fdpcg_free.restype = None
    L.gpu_tomo_setup.argtypes = [C.POINTER(MtRecon), C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(vp)]
    L.gpu_tomo_update_cn2.argtypes = [vp, C.POINTER(MtRecon)]
    L.gpu_tomo_reset.argtypes = [vp]
    L.gpu_tomo_free.argtypes = [vp]
    L.gpu_tomo_free.restype = None
    L.gpu_tomo_solve.argtypes = [vp, C.c_int, dp, dp, dp]
    L.gpu_tomo_solve_batch.argtypes = [vp, C.c_int, dp, dp, dp]
    L.gpu_tomo_set_grad.argtypes = [vp, C.c_int, dp]
    L.gpu_tomo_get_x.argtypes = [vp, C.c_int, dp, dp]
    L.gpu_tomo_group_create.argtypes = [C.c_int, C.POINTER(vp), ip, ip, C.POINTER(vp)]
    L.gpu_tomo_group_solve.argtypes = [vp, C.c_int, vp]
    L.gpu_tomo_group_autotune.argtypes = [vp, ip]
    L.gpu_tomo_group_free.argtypes = [vp]
    L.gpu_tomo_group_free.restype = None
    L.gpu_tomo_autotune.argtypes = [vp, C.c_int, ip]
    L.gpu_tomo_tune_set.argtypes = [vp, ip]
    L.gpu_tomo_info.argtypes = [vp, ip, ip, ip, ip, ip]
    L.gpu_tomo_server_create.argtypes = [C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.POINTER(vp)]
    L.gpu_tomo_server_add_loop.argtypes = [vp, C.POINTER(MtRecon), C.c_int, ip]
    L.gpu_tomo_server_remove_loop.argtypes = [vp, C.c_int]
    L.gpu_tomo_server_submit.argtypes = [vp, C.c_int, dp, C.c_long]
    L.gpu_tomo_server_poll.argtypes = [vp, C.c_int, ip, lp, dp, dp, ip]
    L.gpu_tomo_server_drain.argtypes = [vp]
    L.gpu_tomo_server_get_x.argtypes = [vp, C.c_int, dp, dp]
    L.gpu_tomo_server_usage.argtypes = [vp, dp, dp]
    L.gpu_tomo_server_autotune.argtypes = [vp, ip]
    L.gpu_tomo_server_update_cn2.argtypes = [vp, C.c_int, C.POINTER(MtRecon)]
    L.gpu_tomo_server_free.argtypes = [vp]
    L.gpu_tomo_server_free.restype = None
    L.gpu_tomo_solve_dev.argtypes = [vp, vp, vp, C.c_int, C.c_int, vp]
    L.gpu_tomo_solve_dev_noalloc.argtypes = [vp, vp, vp, vp, C.c_int, vp]
    L.gpu_tomo_mvm_tomo.argtypes = [vp, C.c_long, C.c_long, dp]
    L.gpu_tomo_residuals.argtypes = [vp, dp, C.c_int]
    L.gpu_tomo_L.argtypes = [vp, dp, dp, C.c_double, C.c_double]
    L.gpu_tomo_R.argtypes = [vp, dp, dp, C.c_double]
    L.gpu_tomo_P.argtypes = [vp, dp, dp]
    L.gpu_tomo_sizes.argtypes = [vp, lp, lp]
    L.gpu_tomo_precision.argtypes = [vp, C.c_int]
    L.gpu_tomo_profile.argtypes = [vp, C.c_int]
    L.gpu_tomo_profile_read.argtypes = [vp, C.POINTER(C.c_char_p), dp, dp, ip]
    L.mt_fit_create.argtypes = [vp, C.POINTER(MtFitParms), C.POINTER(vp)]
    L.mt_fit_recon.argtypes = [vp, C.POINTER(MtFitRecon)]
    L.mt_fit_sizes.argtypes = [vp, lp, lp]
    L.mt_fit_free.argtypes = [vp]
    L.mt_fit_free.restype = None
    L.gpu_fit_setup.argtypes = [C.POINTER(MtRecon), C.POINTER(MtFitRecon), C.c_int, C.c_int, C.POINTER(vp)]
    L.gpu_fit_solve.argtypes = [vp, C.c_int, dp, dp, dp]
    L.gpu_fit_solve_dev.argtypes = [vp, vp, vp, C.c_int, vp]
    L.gpu_fit_R.argtypes = [vp, dp, dp, C.c_double]
    L.gpu_fit_L.argtypes = [vp, dp, dp, C.c_double, C.c_double]
    L.gpu_fit_residuals.argtypes = [vp, dp, C.c_int]
    L.gpu_mvm_assemble.argtypes = [vp, vp, C.c_long, C.c_long, dp]
    L.gpu_fit_free.argtypes = [vp]
    L.gpu_fit_free.restype = None
    L.mt_conf_new.argtypes = [C.POINTER(vp)]
    L.mt_conf_free.argtypes = [vp]
    L.mt_conf_free.restype = None
    L.mt_conf_add_path.argtypes = [vp, C.c_char_p]
    L.mt_conf_read.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_int]
    L.mt_conf_has.argtypes = [vp, C.c_char_p]
    L.mt_conf_str.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_long, lp]
    L.mt_conf_strarr.argtypes = [vp, C.c_char_p, C.c_long, C.c_char_p, C.c_long, lp]
    L.mt_conf_dbl.argtypes = [vp, C.c_char_p, dp, C.c_long, lp, lp, lp]
    L.mt_conf_unused.argtypes = [vp, C.c_char_p, C.c_long, lp]
    L.mt_setup_load_gp.argtypes = [vp, vp]
    L.mt_bin_type.argtypes = [vp]
    L.mt_bin_idata.argtypes = [vp]
    L.mt_bin_idata.restype = lp
    L.mt_bin_sparse.argtypes = [vp, C.POINTER(lp), C.POINTER(lp), C.POINTER(dp), lp]
    L.mt_bin_new_typed.argtypes = [C.c_int, C.c_long, C.c_long, dp, C.c_char_p, C.POINTER(vp)]
    L.mt_bin_new_sparse.argtypes = [C.c_long, C.c_long, lp, lp, dp, C.c_int, C.c_char_p, C.POINTER(vp)]
    L.mt_bin_read.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.mt_bin_write.argtypes = [C.c_char_p, vp]
    L.mt_bin_free.argtypes = [vp]
    L.mt_bin_free.restype = None
    L.mt_bin_info.argtypes = [vp, ip, lp, lp]
    L.mt_bin_data.argtypes = [vp]
    L.mt_bin_data.restype = dp
    L.mt_bin_child.argtypes = [vp, C.c_long]
    L.mt_bin_child.restype = vp
    L.mt_bin_header.argtypes = [vp]
    L.mt_bin_header.restype = C.c_char_p
    L.mt_bin_new_dbl.argtypes = [C.c_long, C.c_long, dp, C.c_char_p, C.POINTER(vp)]
    L.mt_bin_new_cell.argtypes = [C.c_long, C.c_long, C.c_char_p, C.POINTER(vp)]
    L.mt_bin_set.argtypes = [vp, C.c_long, vp]
    L.mt_bin_grad.argtypes = [vp, C.POINTER(MtRecon), dp]
    L.mt_bin_opdr.argtypes = [C.POINTER(MtRecon), dp, C.POINTER(vp)]
    _lib = L
    return L


# Every symbol include/maos_tomo.h declares (checked by the CPU tests).
EXPORTS = [
    "mt_last_error", "mt_setup_create", "mt_setup_update_cn2", "mt_setup_recon",
    "mt_setup_free", "mt_setup_fdpcg_mhat", "mt_setup_sizes", "mt_fdpcg_prepare",
    "mt_fdpcg_free", "gpu_tomo_setup",
    "gpu_tomo_update_cn2", "gpu_tomo_reset", "gpu_tomo_free", "gpu_tomo_solve",
    "gpu_tomo_solve_batch", "gpu_tomo_solve_dev", "gpu_tomo_solve_dev_noalloc", "gpu_tomo_mvm_tomo", "gpu_tomo_set_grad", "gpu_tomo_get_x",
    "gpu_tomo_group_create", "gpu_tomo_group_solve", "gpu_tomo_group_autotune", "gpu_tomo_group_free",
    "gpu_tomo_server_create", "gpu_tomo_server_add_loop", "gpu_tomo_server_remove_loop",
    "gpu_tomo_server_submit", "gpu_tomo_server_poll", "gpu_tomo_server_drain", "gpu_tomo_server_get_x",
    "gpu_tomo_server_usage", "gpu_tomo_server_autotune", "gpu_tomo_server_update_cn2", "gpu_tomo_server_free", "gpu_tomo_autotune", "gpu_tomo_tune_set", "gpu_tomo_info", "gpu_tomo_residuals", "gpu_tomo_L", "gpu_tomo_R", "gpu_tomo_P",
    "gpu_tomo_sizes", "gpu_tomo_profile", "gpu_tomo_profile_read", "gpu_tomo_precision",
    "mt_fit_create", "mt_fit_recon", "mt_fit_sizes", "mt_fit_free", "gpu_fit_setup", "gpu_fit_solve",
    "gpu_fit_solve_dev", "gpu_fit_R", "gpu_fit_L", "gpu_fit_residuals", "gpu_fit_free", "gpu_mvm_assemble",
    "mt_conf_new", "mt_conf_free", "mt_conf_add_path", "mt_conf_read", "mt_conf_has", "mt_conf_str",
    "mt_conf_strarr", "mt_conf_dbl", "mt_conf_unused",
    "mt_bin_read", "mt_bin_write", "mt_bin_free", "mt_bin_info", "mt_bin_data", "mt_bin_child",
    "mt_bin_header", "mt_bin_new_dbl", "mt_bin_new_cell", "mt_bin_set", "mt_bin_grad", "mt_bin_opdr",
    "mt_setup_load_gp", "mt_bin_type", "mt_bin_idata", "mt_bin_sparse", "mt_bin_new_typed", "mt_bin_new_sparse",
]


class MtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (status %d)" % (msg, code))
        self.code = code


def check(rc, what):
    if rc < 0:
        raise MtError(rc, "%s: %s" % (what, (lib().mt_last_error() or b"").decode()))
    return rc
