"""ctypes binding of libseismod_hip.so (declared in include/seismod_hip.h).

The library is loaded from seismod_amd/lib/.  There is no fallback: if the
library is missing or fails to load, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libseismod_hip.so"

SM_OK, SM_ERR_ARG, SM_ERR_HIP, SM_ERR_CFL, SM_ERR_NOMEM, SM_ERR_STATE, SM_ERR_NODEV = range(7)
SM_STORE_F32, SM_STORE_F64, SM_STORE_BF16, SM_STORE_F16 = range(4)
SM_VX, SM_VY, SM_VZ, SM_TXX, SM_TZZ, SM_TXZ, SM_P = 0, 1, 2, 5, 6, 7, 8
SM_QX, SM_QZ = 3, 4
SM_PVE_PML_MEDIAN = 2
SM_PSTD_HIPFFT = 1
SM_TXY, SM_TYZ = 9, 10
SM_DIV, SM_CURL, SM_VR, SM_VT, SM_TRR, SM_TRT, SM_TRZ, SM_TTT, SM_TTZ = range(11, 20)
SM_BULK_S, SM_FR, SM_FT, SM_FZ, SM_KURKJIAN = 8, 9, 10, 11, 12
SM_SX, SM_SY, SM_SZ, SM_SXY, SM_SXZ, SM_SYZ, SM_BULK, SM_SF = range(8)
SM_A3D_NAIVE = 1
SM_A3D_SROUND = 2
SM_A3D_PAD64 = 4
SM_A3D_MAT16 = 8
SM_A3D_CELL = 0
SM_A3D_CELL2 = 2
SM_A3D_TILE = 5
SM_A3D_TILE_SPLIT = 6
SM_A3D_TILE_NT = 7
SM_A3D_TILE_SPLIT_NT = 8
SM_PROBE_NT = 16
SM_E2D_ROWS = 0
SM_E2D_STAGED = 1
SM_PROBE_TILES = 32

_size_p = C.POINTER(C.c_size_t)
_dbl_p = C.POINTER(C.c_double)
_int_p = C.POINTER(C.c_int)


class A3dParams(C.Structure):
    _fields_ = [
        ("nx", C.c_size_t), ("ny", C.c_size_t), ("nz", C.c_size_t), ("np", C.c_size_t),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
        ("vmax", C.c_double), ("pml_order", C.c_double), ("Rc", C.c_double),
        ("kappa_max", C.c_double), ("alpha_max", C.c_int), ("f0", C.c_double),
        ("storage", C.c_int), ("flags", C.c_int),
    ]


class CylParams(C.Structure):
    _fields_ = [
        ("nx", C.c_size_t), ("nz", C.c_size_t), ("np", C.c_size_t),
        ("dr", C.c_double), ("dz", C.c_double), ("dt", C.c_double), ("mode", C.c_int),
        ("vmax", C.c_double), ("pml_order", C.c_double), ("Rc", C.c_double),
        ("kappa_max", C.c_double), ("alpha_max", C.c_int), ("f0", C.c_double),
        ("storage", C.c_int), ("flags", C.c_int),
    ]


class ShParams(C.Structure):
    _fields_ = [
        ("nx", C.c_size_t), ("nz", C.c_size_t), ("np", C.c_size_t),
        ("dx", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
        ("vmax", C.c_double), ("pml_order", C.c_double), ("Rc", C.c_double),
        ("kappa_max", C.c_double), ("alpha_max", C.c_int), ("f0", C.c_double),
        ("abs", C.c_int), ("storage", C.c_int), ("flags", C.c_int),
    ]


class VeParams(C.Structure):
    _fields_ = [
        ("nx", C.c_size_t), ("nz", C.c_size_t), ("np", C.c_size_t),
        ("dx", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
        ("vmax", C.c_double), ("pml_order", C.c_double), ("Rc", C.c_double),
        ("kappa_max", C.c_double), ("alpha_max", C.c_int), ("f0", C.c_double),
        ("abs", C.c_int), ("L", C.c_int), ("storage", C.c_int), ("flags", C.c_int),
    ]


# sm_pve_params has the layout of sm_sh_params
PveParams = ShParams


class E2dParams(C.Structure):
    _fields_ = [
        ("nx", C.c_size_t), ("nz", C.c_size_t), ("npml", C.c_size_t),
        ("h", C.c_double), ("dt", C.c_double), ("storage", C.c_int),
        ("q1_compat", C.c_int), ("damping_scale", C.c_double), ("flags", C.c_int),
    ]


# name -> (restype, argtypes)
_SIGS = {
    "sm_abi_version": (C.c_int, []),
    "sm_strerror": (C.c_char_p, [C.c_int]),
    "sm_last_error": (C.c_char_p, []),
    "sm_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "sm_ctx_destroy": (C.c_int, [C.c_void_p]),
    "sm_ctx_sync": (C.c_int, [C.c_void_p]),
    "sm_ctx_mem_info": (C.c_int, [C.c_void_p, _size_p, _size_p]),
    "sm_probe_stream": (C.c_int, [C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_int,
                                  _dbl_p]),
    "sm_ricker_length": (C.c_size_t, [C.c_double, C.c_double]),
    "sm_compute_src_fct": (C.c_int, [C.c_size_t, _dbl_p, _dbl_p, C.c_double, _size_p, _dbl_p]),
    "sm_nsteps": (C.c_size_t, [C.c_double, C.c_double]),
    "sm_record_due": (C.c_int, [C.c_size_t, C.c_double, C.c_double, C.c_double]),
    "sm_a3d_create": (C.c_int, [C.c_void_p, C.POINTER(A3dParams), _dbl_p, _dbl_p,
                                C.POINTER(C.c_void_p)]),
    "sm_a3d_destroy": (C.c_int, [C.c_void_p]),
    "sm_a3d_set_sources": (C.c_int, [C.c_void_p, C.c_size_t, _size_p, _size_p, _dbl_p]),
    "sm_a3d_set_receivers": (C.c_int, [C.c_void_p, C.c_size_t, _size_p, _int_p, C.c_size_t]),
    "sm_a3d_run": (C.c_int, [C.c_void_p, C.c_size_t]),
    "sm_a3d_current_step": (C.c_int, [C.c_void_p, _size_p]),
    "sm_a3d_get_traces": (C.c_int, [C.c_void_p, C.c_size_t, C.c_size_t, _dbl_p]),
    "sm_a3d_get_field": (C.c_int, [C.c_void_p, C.c_int, _dbl_p]),
    "sm_a3d_get_slice_xz": (C.c_int, [C.c_void_p, C.c_int, C.c_size_t, _dbl_p]),
    "sm_a3d_reset": (C.c_int, [C.c_void_p]),
    "sm_a3d_set_launch": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "sm_a3d_set_variant": (C.c_int, [C.c_void_p, C.c_int]),
    "sm_a3d_get_variant": (C.c_int, [C.c_void_p, _int_p]),
    "sm_a3d_get_launch": (C.c_int, [C.c_void_p, _int_p, _int_p]),
    "sm_a3d_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "sm_a3d_get_timing": (C.c_int, [C.c_void_p, _dbl_p, _size_p]),
    "sm_a3d_device_bytes": (C.c_int, [C.c_void_p, _size_p]),
    "sm_a3d_bytes_needed": (C.c_size_t, [C.POINTER(A3dParams), C.c_size_t, C.c_size_t]),
    "sm_e2d_create": (C.c_int, [C.c_void_p, C.POINTER(E2dParams), _dbl_p, _dbl_p, _dbl_p,
                                C.POINTER(C.c_void_p)]),
    "sm_e2d_destroy": (C.c_int, [C.c_void_p]),
    "sm_e2d_set_sources": (C.c_int, [C.c_void_p, C.c_size_t, _int_p, _size_p, _size_p, _dbl_p]),
    "sm_e2d_set_receivers": (C.c_int, [C.c_void_p, C.c_size_t, _size_p, _int_p, C.c_size_t]),
    "sm_e2d_run": (C.c_int, [C.c_void_p, C.c_size_t]),
    "sm_e2d_current_step": (C.c_int, [C.c_void_p, _size_p]),
    "sm_e2d_get_traces": (C.c_int, [C.c_void_p, C.c_size_t, C.c_size_t, _dbl_p]),
    "sm_e2d_get_field": (C.c_int, [C.c_void_p, C.c_int, _dbl_p]),
    "sm_e2d_get_field_padded": (C.c_int, [C.c_void_p, C.c_int, _dbl_p]),
    "sm_e2d_reset": (C.c_int, [C.c_void_p]),
    "sm_e2d_device_bytes": (C.c_int, [C.c_void_p, _size_p]),
    "sm_e2d_bytes_needed": (C.c_size_t, [C.POINTER(E2dParams), C.c_size_t, C.c_size_t]),
    "sm_e2d_set_launch": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "sm_e2d_set_variant": (C.c_int, [C.c_void_p, C.c_int]),
    "sm_e2d_get_variant": (C.c_int, [C.c_void_p, _int_p]),
    "sm_e2d_batch_set_variant": (C.c_int, [C.c_void_p, C.c_int]),
    "sm_e2d_get_launch": (C.c_int, [C.c_void_p, _int_p, _int_p]),
    "sm_e2d_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "sm_e2d_get_timing": (C.c_int, [C.c_void_p, _dbl_p, _size_p]),
    "sm_e2d_batch_create": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "sm_e2d_batch_destroy": (C.c_int, [C.c_void_p]),
    "sm_e2d_batch_set_graph_steps": (C.c_int, [C.c_void_p, C.c_int]),
    "sm_e2d_batch_set_group_bytes": (C.c_int, [C.c_void_p, C.c_size_t]),
    "sm_e2d_batch_run": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p), _size_p]),
    "sm_e2d_batch_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "sm_e2d_batch_get_timing": (C.c_int, [C.c_void_p, _dbl_p, _size_p]),
    "sm_e2d_batch_get_graphs": (C.c_int, [C.c_void_p, _size_p, _size_p]),
    "sm_e2d_batch_run_noalloc": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p), _size_p]),
    "sm_cyl_create": (C.c_int, [C.c_void_p, C.POINTER(CylParams), _dbl_p, _dbl_p, _dbl_p,
                                C.POINTER(C.c_void_p)]),
    "sm_cyl_destroy": (C.c_int, [C.c_void_p]),
    "sm_cyl_set_sources": (C.c_int, [C.c_void_p, C.c_size_t, _int_p, _size_p, _size_p, _dbl_p]),
    "sm_kurkjian_fct": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                  C.c_double, C.c_double, C.c_double, C.c_size_t, _dbl_p]),
    "sm_cyl_set_receivers": (C.c_int, [C.c_void_p, C.c_size_t, _size_p, _int_p, C.c_size_t]),
    "sm_cyl_run": (C.c_int, [C.c_void_p, C.c_size_t]),
    "sm_cyl_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "sm_cyl_get_timing": (C.c_int, [C.c_void_p, _dbl_p, _size_p]),
    "sm_cyl_current_step": (C.c_int, [C.c_void_p, _size_p]),
    "sm_cyl_get_traces": (C.c_int, [C.c_void_p, C.c_size_t, C.c_size_t, _dbl_p]),
    "sm_cyl_get_field": (C.c_int, [C.c_void_p, C.c_int, _dbl_p]),
    "sm_cyl_reset": (C.c_int, [C.c_void_p]),
    "sm_cyl_device_bytes": (C.c_int, [C.c_void_p, _size_p]),
    "sm_cyl_bytes_needed": (C.c_size_t, [C.POINTER(CylParams), C.c_size_t, C.c_size_t]),
    "sm_sh_create": (C.c_int, [C.c_void_p, C.POINTER(ShParams), _dbl_p, _dbl_p, _dbl_p, _dbl_p, _dbl_p,
                               C.POINTER(C.c_void_p)]),
    "sm_sh_destroy": (C.c_int, [C.c_void_p]),
    "sm_sh_set_sources": (C.c_int, [C.c_void_p, C.c_size_t, _int_p, _size_p, _size_p, _dbl_p]),
    "sm_sh_set_receivers": (C.c_int, [C.c_void_p, C.c_size_t, _size_p, _int_p, C.c_size_t]),
    "sm_sh_run": (C.c_int, [C.c_void_p, C.c_size_t]),
    "sm_sh_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "sm_sh_get_timing": (C.c_int, [C.c_void_p, _dbl_p, _size_p]),
    "sm_sh_current_step": (C.c_int, [C.c_void_p, _size_p]),
    "sm_sh_get_traces": (C.c_int, [C.c_void_p, C.c_size_t, C.c_size_t, _dbl_p]),
    "sm_sh_get_field": (C.c_int, [C.c_void_p, C.c_int, _dbl_p]),
    "sm_sh_reset": (C.c_int, [C.c_void_p]),
    "sm_sh_device_bytes": (C.c_int, [C.c_void_p, _size_p]),
    "sm_sh_bytes_needed": (C.c_size_t, [C.POINTER(ShParams), C.c_size_t, C.c_size_t]),
    "sm_ve_create": (C.c_int, [C.c_void_p, C.POINTER(VeParams)] + [_dbl_p] * 9 + [C.POINTER(C.c_void_p)]),
    "sm_ve_moduli": (C.c_int, [C.c_size_t, C.c_size_t, C.c_int] + [_dbl_p] * 14),
    "sm_ve_destroy": (C.c_int, [C.c_void_p]),
    "sm_ve_set_sources": (C.c_int, [C.c_void_p, C.c_size_t, _int_p, _size_p, _size_p, _dbl_p]),
    "sm_ve_set_receivers": (C.c_int, [C.c_void_p, C.c_size_t, _size_p, _int_p, C.c_size_t]),
    "sm_ve_run": (C.c_int, [C.c_void_p, C.c_size_t]),
    "sm_ve_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "sm_ve_get_timing": (C.c_int, [C.c_void_p, _dbl_p, _size_p]),
    "sm_ve_current_step": (C.c_int, [C.c_void_p, _size_p]),
    "sm_ve_get_traces": (C.c_int, [C.c_void_p, C.c_size_t, C.c_size_t, _dbl_p]),
    "sm_ve_get_field": (C.c_int, [C.c_void_p, C.c_int, _dbl_p]),
    "sm_ve_reset": (C.c_int, [C.c_void_p]),
    "sm_ve_device_bytes": (C.c_int, [C.c_void_p, _size_p]),
    "sm_ve_bytes_needed": (C.c_size_t, [C.POINTER(VeParams), C.c_size_t, C.c_size_t]),
    "sm_pve_create": (C.c_int, [C.c_void_p, C.POINTER(PveParams)] + [_dbl_p] * 12 + [C.POINTER(C.c_void_p)]),
    "sm_pve_create_vti": (C.c_int, [C.c_void_p, C.POINTER(PveParams)] + [_dbl_p] * 16 + [C.POINTER(C.c_void_p)]),
    "sm_pve_destroy": (C.c_int, [C.c_void_p]),
    "sm_pve_set_sources": (C.c_int, [C.c_void_p, C.c_size_t, _int_p, _size_p, _size_p, _dbl_p]),
    "sm_pve_set_receivers": (C.c_int, [C.c_void_p, C.c_size_t, _size_p, _int_p, C.c_size_t]),
    "sm_pve_run": (C.c_int, [C.c_void_p, C.c_size_t]),
    "sm_pve_set_energy": (C.c_int, [C.c_void_p] + [C.c_size_t] * 4 + [_dbl_p] * 3 + [C.c_size_t]),
    "sm_pve_set_energy_axes": (C.c_int, [C.c_void_p] + [C.c_size_t] * 4 + [_dbl_p] * 6 + [C.c_size_t]),
    "sm_pve_get_energy": (C.c_int, [C.c_void_p, C.c_size_t, C.c_size_t, _dbl_p]),
    "sm_pve_state_bytes": (C.c_int, [C.c_void_p, _size_p]),
    "sm_pve_get_state": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "sm_pve_set_state": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "sm_pve_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "sm_pve_get_timing": (C.c_int, [C.c_void_p, _dbl_p, _size_p]),
    "sm_pve_current_step": (C.c_int, [C.c_void_p, _size_p]),
    "sm_pve_get_traces": (C.c_int, [C.c_void_p, C.c_size_t, C.c_size_t, _dbl_p]),
    "sm_pve_get_field": (C.c_int, [C.c_void_p, C.c_int, _dbl_p]),
    "sm_pve_reset": (C.c_int, [C.c_void_p]),
    "sm_pve_device_bytes": (C.c_int, [C.c_void_p, _size_p]),
    "sm_pve_bytes_needed": (C.c_size_t, [C.POINTER(PveParams), C.c_size_t, C.c_size_t]),
}


class SeismodError(RuntimeError):
    def __init__(self, status: int, where: str, msg: str):
        super().__init__(f"{where}: status {status} ({msg})")
        self.status = status


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built; run `python -m seismod_amd.build`")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()
EXPORTED = tuple(_SIGS)


def check(status: int, where: str) -> None:
    if status != SM_OK:
        msg = lib.sm_last_error().decode(errors="replace")
        raise SeismodError(status, where, msg)
