"""Loader of the in-tree HIP library libpadeops_hip.so (ctypes).

There is no CPU fallback: if the library is missing the import fails loudly,
and every compute call needs a HIP device.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpadeops_hip.so")

_lib = None

# (name, argtypes) of every entry point declared in include/padeops_hip.h
_c_int, _c_double, _vp, _sz = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t
_pvp = ctypes.POINTER(ctypes.c_void_p)
_pint = ctypes.POINTER(ctypes.c_int)
_pdbl = ctypes.POINTER(ctypes.c_double)

_APPLY = [_vp, _c_int, _vp, _vp, _c_int, _c_int, _vp]
_APPLY_BC = [_vp, _c_int, _vp, _vp, _c_int, _c_int, _c_int, _c_int, _vp]
_STAGG = [_vp, _c_int, _vp, _vp, _c_int, _c_int, _c_int, _vp]

SIGNATURES = {
    "pdo_last_error": ([], ctypes.c_char_p),
    "pdo_device_count": ([_pint], _c_int),
    "pdo_sync": ([_vp], _c_int),
    "pdo_malloc": ([_pvp, _sz], _c_int),
    "pdo_free": ([_vp], _c_int),
    "pdo_memcpy_h2d": ([_vp, _vp, _sz], _c_int),
    "pdo_memcpy_d2h": ([_vp, _vp, _sz], _c_int),
    "pdo_memcpy_d2d": ([_vp, _vp, _sz, _vp], _c_int),
    "pdo_cd06_init": ([_pvp, _c_int, _c_double, _c_int, _c_int, _c_int], _c_int),
    "pdo_cd06_dd1": (_APPLY, _c_int),
    "pdo_cd06_dd2": (_APPLY, _c_int),
    "pdo_cd06_dd3": (_APPLY, _c_int),
    "pdo_cd06_d2d1": (_APPLY, _c_int),
    "pdo_cd06_d2d2": (_APPLY, _c_int),
    "pdo_cd06_d2d3": (_APPLY, _c_int),
    "pdo_cd10_init": ([_pvp, _c_int, _c_double, _c_int, _c_int, _c_int], _c_int),
    "pdo_cd10_dd1": (_APPLY_BC, _c_int),
    "pdo_cd10_dd2": (_APPLY_BC, _c_int),
    "pdo_cd10_dd3": (_APPLY_BC, _c_int),
    "pdo_cd10_d2d1": (_APPLY_BC, _c_int),
    "pdo_cd10_d2d2": (_APPLY_BC, _c_int),
    "pdo_cd10_d2d3": (_APPLY_BC, _c_int),
    "pdo_cf90_init": ([_pvp, _c_int, _c_int], _c_int),
    "pdo_cf90_filter1": (_APPLY_BC, _c_int),
    "pdo_cf90_filter2": (_APPLY_BC, _c_int),
    "pdo_cf90_filter3": (_APPLY_BC, _c_int),
    "pdo_cf90_filter3_cmplx": (_APPLY_BC, _c_int),
    "pdo_gaussian_init": ([_pvp, _c_int, _c_int], _c_int),
    "pdo_lstsq_init": ([_pvp, _c_int, _c_int], _c_int),
    "pdo_xfilter_filter1": (_APPLY_BC, _c_int),
    "pdo_xfilter_filter2": (_APPLY_BC, _c_int),
    "pdo_xfilter_filter3": (_APPLY_BC, _c_int),
    "pdo_xfilter_filter3_cmplx": (_APPLY_BC, _c_int),
    "pdo_xfilter_destroy": ([_vp], _c_int),
    "pdo_xfilter_dense": ([_vp, _c_int, _c_int, _pdbl], _c_int),
    "pdo_derivatives_init": ([_pvp, _c_int, _c_int, _c_int, _c_double, _c_double, _c_double, _pint, _pint], _c_int),
    "pdo_derivatives_method": ([_vp, _c_int, _pint], _c_int),
    "pdo_derivatives_ddx": ([_vp, _c_int, _vp, _vp, _c_int, _c_int, _vp], _c_int),
    "pdo_derivatives_ddy": ([_vp, _c_int, _vp, _vp, _c_int, _c_int, _vp], _c_int),
    "pdo_derivatives_ddz": ([_vp, _c_int, _vp, _vp, _c_int, _c_int, _vp], _c_int),
    "pdo_derivatives_d2dx2": ([_vp, _c_int, _vp, _vp, _c_int, _c_int, _vp], _c_int),
    "pdo_derivatives_d2dy2": ([_vp, _c_int, _vp, _vp, _c_int, _c_int, _vp], _c_int),
    "pdo_derivatives_d2dz2": ([_vp, _c_int, _vp, _vp, _c_int, _c_int, _vp], _c_int),
    "pdo_derivatives_destroy": ([_vp], _c_int),
    "pdo_filters_init": ([_pvp, _c_int, _c_int, _c_int, _pint, _pint], _c_int),
    "pdo_filters_method": ([_vp, _c_int, _pint], _c_int),
    "pdo_filters_filterx": ([_vp, _c_int, _vp, _vp, _c_int, _c_int, _vp], _c_int),
    "pdo_filters_filtery": ([_vp, _c_int, _vp, _vp, _c_int, _c_int, _vp], _c_int),
    "pdo_filters_filterz": ([_vp, _c_int, _vp, _vp, _c_int, _c_int, _vp], _c_int),
    "pdo_filters_destroy": ([_vp], _c_int),
    "pdo_operators_gradient": ([_vp, _c_int, _vp, _vp, _vp, _vp, _pint, _vp], _c_int),
    "pdo_operators_curl": ([_vp, _c_int, _vp, _vp, _vp, _vp, _pint, _vp], _c_int),
    "pdo_operators_divergence": ([_vp, _c_int, _vp, _vp, _vp, _vp, _pint, _vp], _c_int),
    "pdo_operators_filter3d": ([_vp, _c_int, _vp, _c_int, _pint, _vp], _c_int),
    "pdo_cd06stagg_init_periodic": ([_pvp, _c_int, _c_double], _c_int),
    "pdo_cd06stagg_ddz_E2C": (_STAGG, _c_int),
    "pdo_cd06stagg_ddz_C2E": (_STAGG, _c_int),
    "pdo_cd06stagg_d2dz2_C2C": (_STAGG, _c_int),
    "pdo_cd06stagg_d2dz2_E2E": (_STAGG, _c_int),
    "pdo_cd06stagg_interpz_E2C": (_STAGG, _c_int),
    "pdo_cd06stagg_interpz_C2E": (_STAGG, _c_int),
    "pdo_cd06stagg_init": ([_pvp, _c_int, _c_double, _c_int, _c_int, _c_int, _c_int], _c_int),
    "pdo_cd06stagg_ddz_C2C": (_STAGG, _c_int),
    "pdo_cd06stagg_ddz_E2E": (_STAGG, _c_int),
    "pdo_op_destroy": ([_vp], _c_int),
    "pdo_op_apply_noalloc": ([_vp, _c_int, _c_int, _c_int, _vp, _vp, _c_int, _c_int, _c_int, _c_int, _vp], _c_int),
    "pdo_op_dense": ([_vp, _c_int, _c_int, _c_int, _pdbl, _pdbl], _c_int),
    "pdo_op_dims": ([_vp, _c_int, _pint, _pint], _c_int),
    "pdo_op_partition": ([_vp, _c_int, _c_int, _c_int, _pint, _pint], _c_int),
    "pdo_op_tables": ([_vp, _c_int, _c_int, _c_int, _c_int, _pint, _pdbl, _pdbl, _pdbl, _pdbl], _c_int),
    "pdo_poisson_periodic_init": ([_pvp, _c_int, _c_int, _c_int, _c_double, _c_double, _c_double,
                                   _c_int, _c_int, _c_int, _c_int], _c_int),
    "pdo_poisson_periodic_init_k": ([_pvp, _c_int, _c_int, _c_int, _pdbl, _pdbl, _pdbl, _c_int], _c_int),
    "pdo_poisson_periodic_solve": ([_vp, _c_int, _vp, _vp, _vp], _c_int),
    "pdo_poisson_destroy": ([_vp], _c_int),
    "pdo_pade6stagg_init": ([_pvp, _c_int, _c_double, _c_int], _c_int),
    "pdo_pade6stagg_apply": ([_vp, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _c_int, _c_int, _c_int, _vp], _c_int),
    "pdo_pade6stagg_modk": ([_c_int, _c_int, _c_double, _pdbl, _pdbl], _c_int),
    "pdo_pade6stagg_destroy": ([_vp], _c_int),
    "pdo_duidxj_init": ([_pvp, _vp, _vp, _vp, _pint, _c_int], _c_int),
    "pdo_duidxj_compute": ([_vp, _c_int, _vp, _vp], _c_int),
    "pdo_duidxj_destroy": ([_vp], _c_int),
    "pdo_sgs_kernel": ([_c_int, _c_int, _vp, _vp, _vp, ctypes.c_long, _c_double, _c_double, _c_double, _vp],
                       _c_int),
    "pdo_sgs_scale": ([_c_int, _vp, ctypes.c_long, _c_int, _vp, _c_double, _vp], _c_int),
    "pdo_dynproc_init": ([_pvp, _vp, _c_double], _c_int),
    "pdo_dynproc_compute": ([_vp, _c_int] + [_vp] * 9 + [_vp], _c_int),
    "pdo_dynproc_destroy": ([_vp], _c_int),
    "pdo_igrid_interp_primitive": ([_vp, _c_int, _vp, _vp], _c_int),
    "pdo_sgsrhs_init": ([_pvp, _vp, _vp, _vp, _vp, _c_int], _c_int),
    "pdo_cgrid_init": ([_pvp, _vp, _c_int], _c_int),
    "pdo_cgrid_rhs": ([_vp, _c_int, _vp, _vp, _vp], _c_int),
    "pdo_cgrid_filter": ([_vp, _c_int, _vp, _c_int, _c_int, _vp], _c_int),
    "pdo_cgrid_field": ([_vp, _c_int, _pvp], _c_int),
    "pdo_cgrid_destroy": ([_vp], _c_int),
    "pdo_sgsrhs_compute_wall": ([_vp, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _c_int),
    "pdo_sgsrhs_wall_state": ([_vp, _vp], _c_int),
    "pdo_sgsrhs_compute": ([_vp, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _c_int),
    "pdo_sgsrhs_fields": ([_vp, _pvp, _pvp, _pvp, _pvp], _c_int),
    "pdo_sgsrhs_set_closure": ([_vp, _vp, _c_double, _c_double, _c_int], _c_int),
    "pdo_sgsrhs_destroy": ([_vp], _c_int),
    "pdo_igrid_init": ([_pvp, _vp, _c_int], _c_int),
    "pdo_igrid_set_velocity": ([_vp, _vp, _vp, _vp, _vp], _c_int),
    "pdo_igrid_step": ([_vp, _c_int, _vp], _c_int),
    "pdo_igrid_field": ([_vp, _c_int, _pvp], _c_int),
    "pdo_igrid_time": ([_vp, _pdbl, ctypes.POINTER(ctypes.c_long), _pdbl], _c_int),
    "pdo_igrid_set_closure": ([_vp, _c_int, _vp, _c_double, _c_double, _c_double], _c_int),
    "pdo_igrid_destroy": ([_vp], _c_int),
    "pdo_igrid_dump_restart": ([_vp, ctypes.c_char_p, _c_int, _vp], _c_int),
    "pdo_igrid_read_restart": ([_vp, ctypes.c_char_p, _c_int, ctypes.c_long, _vp], _c_int),
    "pdo_mlp_nparam": ([], _c_int),
    "pdo_mlp_create": ([_pvp, ctypes.c_long], _c_int),
    "pdo_mlp_set_params": ([_vp, _vp], _c_int),
    "pdo_mlp_get_params": ([_vp, _vp], _c_int),
    "pdo_mlp_get_grads": ([_vp, _vp], _c_int),
    "pdo_mlp_forward": ([_vp, _vp, _vp, ctypes.c_long, _vp], _c_int),
    "pdo_mlp_gradients": ([_vp, _vp, _vp, ctypes.c_long, _vp, _vp], _c_int),
    "pdo_mlp_adam_step": ([_vp, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, _vp], _c_int),
    "pdo_mlp_train_step": ([_vp, _vp, _vp, ctypes.c_long, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                            ctypes.c_float, _vp, _vp], _c_int),
    "pdo_mlp_forward_train": ([_vp, _vp, _vp, ctypes.c_long, _vp], _c_int),
    "pdo_mlp_backward": ([_vp, _vp, ctypes.c_long, _vp], _c_int),
    "pdo_mlp_destroy": ([_vp], _c_int),
    "pdo_closure_create": ([_pvp, _vp, ctypes.c_long], _c_int),
    "pdo_closure_features": ([_vp, _c_int, _vp, ctypes.c_long, _c_double, _vp, _vp, _vp], _c_int),
    "pdo_closure_stress": ([_vp, _c_int, _vp, ctypes.c_long, _c_double, _c_double, _c_int, _c_int, _c_int, _vp, _vp],
                           _c_int),
    "pdo_closure_fallback_count": ([_vp, ctypes.POINTER(ctypes.c_long), _vp], _c_int),
    "pdo_closure_scale": ([_vp, _pdbl, _vp], _c_int),
    "pdo_closure_loss": ([_vp, _vp, _vp, _vp, ctypes.c_long, _vp, _pdbl, _vp], _c_int),
    "pdo_closure_destroy": ([_vp], _c_int),
    "pdo_spectral_init": ([_pvp, _c_int, _c_int, _c_int, _c_double, _c_double, _c_double, _c_int, _c_double,
                           _c_int], _c_int),
    "pdo_spectral_fft": ([_vp, _c_int, _vp, _vp, _vp], _c_int),
    "pdo_spectral_ifft": ([_vp, _c_int, _vp, _vp, _c_int, _vp], _c_int),
    "pdo_spectral_mtimes_ik": ([_vp, _c_int, _c_int, _vp, _vp, _vp], _c_int),
    "pdo_spectral_dealias": ([_vp, _c_int, _vp, _vp], _c_int),
    "pdo_spectral_dealias_edge": ([_vp, _c_int, _vp, _vp], _c_int),
    "pdo_spectral_wavenumbers": ([_vp] + [_pdbl] * 6, _c_int),
    "pdo_spectral_set_factors": ([_vp, _pdbl, _pdbl, _pdbl], _c_int),
    "pdo_spectral_destroy": ([_vp], _c_int),
    "pdo_padepoisson_init_periodic": ([_pvp, _vp, _c_double], _c_int),
    "pdo_padepoisson_init_nonperiodic": ([_pvp, _vp, _c_double], _c_int),
    "pdo_padepoisson_projection": ([_vp, _c_int, _vp, _vp, _vp, _vp], _c_int),
    "pdo_padepoisson_divergence": ([_vp, _c_int, _vp, _vp, _vp, _vp, _vp], _c_int),
    "pdo_padepoisson_set_stokes": ([_vp, _c_double], _c_int),
    "pdo_padepoisson_pressure": ([_vp, _c_int, _vp, _vp, _vp, _vp, _c_int, _vp], _c_int),
    "pdo_padepoisson_destroy": ([_vp], _c_int),
    "pdo_rhs_init": ([_pvp, _c_int, _c_double, _c_int], _c_int),
    "pdo_rhs_eval": ([_vp] + [_vp] * 7 + [_vp], _c_int),
    "pdo_rhs_pressure": ([_vp, _pvp], _c_int),
    "pdo_rhs_destroy": ([_vp], _c_int),
    "pdo_rhs_max_divergence": ([_vp, _vp, _vp, _vp, _pdbl, _vp], _c_int),
}


class DuidxjFields(ctypes.Structure):
    """pdo_duidxj_fields of include/padeops_hip.h"""
    _fields_ = [(name, ctypes.c_void_p) for name in (
        "uhat", "vhat", "whatC", "what", "uEhat", "vEhat", "duidxjC", "duidxjE", "duidxjChat", "duidxjEhat",
        "d2wdz2hatE", "d2udz2hatC", "d2vdz2hatC")]


class InterpFields(ctypes.Structure):
    """pdo_interp_fields of include/padeops_hip.h"""
    _fields_ = [(name, ctypes.c_void_p) for name in (
        "uhat", "vhat", "what", "That", "whatC", "wC", "uEhat", "uE", "vEhat", "vE", "TEhat", "TE")] + [
        ("TBC_bottom", ctypes.c_int), ("TBC_top", ctypes.c_int), ("botBC_Temp", ctypes.c_int),
        ("Tsurf", ctypes.c_double)]


class SgsParams(ctypes.Structure):
    """pdo_sgs_params of include/padeops_hip.h"""
    _fields_ = [("model", ctypes.c_int), ("explicit_edge", ctypes.c_int), ("camd_x", ctypes.c_double),
                ("camd_y", ctypes.c_double), ("camd_z", ctypes.c_double), ("cglobal", ctypes.c_double),
                ("cmodelC", ctypes.POINTER(ctypes.c_double)), ("cmodelE", ctypes.POINTER(ctypes.c_double)),
                ("wall_model", ctypes.c_int), ("z0", ctypes.c_double)]


class CgridParams(ctypes.Structure):
    """pdo_cgrid_params of include/padeops_hip.h"""
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("dx", ctypes.c_double),
                ("dy", ctypes.c_double), ("dz", ctypes.c_double), ("periodic", ctypes.c_int * 3),
                ("x_bc", ctypes.c_int * 2), ("y_bc", ctypes.c_int * 2), ("z_bc", ctypes.c_int * 2),
                ("derivative", ctypes.c_int * 3), ("filter", ctypes.c_int * 3), ("gam", ctypes.c_double),
                ("Rgas", ctypes.c_double), ("Cmu", ctypes.c_double), ("Cbeta", ctypes.c_double),
                ("Ckap", ctypes.c_double), ("inviscid", ctypes.c_int), ("mu_ref", ctypes.c_double),
                ("T_ref", ctypes.c_double), ("visc_n", ctypes.c_double), ("Pr", ctypes.c_double)]


class IgridParams(ctypes.Structure):
    """pdo_igrid_params of include/padeops_hip.h"""
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("Lx", ctypes.c_double),
                ("Ly", ctypes.c_double), ("Lz", ctypes.c_double), ("Re", ctypes.c_double), ("scheme", ctypes.c_int),
                ("CFL", ctypes.c_double), ("dt", ctypes.c_double), ("use_sgs", ctypes.c_int), ("sgs", SgsParams),
                ("use_forcing", ctypes.c_int), ("kmin", ctypes.c_double), ("kmax", ctypes.c_double),
                ("eps_amplitude", ctypes.c_double), ("nwaves", ctypes.c_int), ("seed", ctypes.c_ulonglong),
                ("walls", ctypes.c_int), ("botWall", ctypes.c_int), ("topWall", ctypes.c_int),
                ("dpFdx", ctypes.c_double), ("stokes_pressure", ctypes.c_int)]


class LibraryMissing(ImportError):
    pass


def lib():
    """Load libpadeops_hip.so; raises LibraryMissing if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(
                "%s is not built: run `python -m padeops_amd.build` (or __graft_entry__.build())" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def last_error():
    return lib().pdo_last_error().decode(errors="replace")
