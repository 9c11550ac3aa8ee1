"""ctypes binding of the C ABI in ``include/axs3d.h``.

Loads the in-tree ``axisem3d_amd/libaxs3d.so``.  There is no fallback: if
the library is missing or the device is absent the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libaxs3d.so")

AXS3D_FP64 = 0
AXS3D_FP32_FP64 = 1

FIELD_DISPL, FIELD_VELOC, FIELD_ACCEL, FIELD_STIFF = 0, 1, 2, 3
FIELD_FLUID_DISPL, FIELD_FLUID_VELOC, FIELD_FLUID_ACCEL = 4, 5, 6
FIELD_FLUID_STIFF = 7
KERNEL_STIFFNESS, KERNEL_POINT, KERNEL_SOURCE, KERNEL_RECORD = 0, 1, 2, 3
KERNEL_STIFFNESS_3D = 4
KERNEL_FLUID_STIFFNESS, KERNEL_FLUID_POINT = 5, 6
KERNEL_CD_GRAD, KERNEL_CD_STRESS, KERNEL_CD_QUAD = 7, 8, 9
NUM_KERNELS = 10

# every symbol declared in include/axs3d.h, with (restype, argtypes)
_p = C.c_void_p
_i = C.c_int
_d = C.c_double
SIGNATURES = {
    "axs3d_ctx_create": (_i, [_i, _i, C.POINTER(_p)]),
    "axs3d_ctx_destroy": (_i, [_p]),
    "axs3d_last_error": (C.c_char_p, [_p]),
    "axs3d_precision_of": (_i, [_p]),
    "axs3d_set_gmat": (_i, [_p, _p, _p]),
    "axs3d_set_solid_points": (_i, [_p, _i, _p, _p]),
    "axs3d_set_solid_elements": (_i, [_p, _i, _p, _p, _p]),
    "axs3d_set_solid_iso1d": (_i, [_p, _p]),
    "axs3d_set_axial_points": (_i, [_p, _i, _p]),
    "axs3d_set_solid_mass3d": (_i, [_p, _i, _p, _p]),
    "axs3d_set_ocean_load": (_i, [_p, _i, _p, _p, _p]),
    "axs3d_set_solid_prt": (_i, [_p, _i, _p, _p]),
    "axs3d_set_fluid_kinds": (_i, [_p, _p]),
    "axs3d_set_fluid_k3d": (_i, [_p, _i, _i, _p]),
    "axs3d_set_element_kinds": (_i, [_p, _p]),
    "axs3d_set_solid_iso3d": (_i, [_p, _i, _i, _p]),
    "axs3d_set_attenuation": (_i, [_p, _i, _i, _p, _p, _p]),
    "axs3d_set_attenuation3d": (_i, [_p, _i, _i, _p]),
    "axs3d_set_attenuation1d": (_i, [_p, _p]),
    "axs3d_set_material_types": (_i, [_p, _p]),
    "axs3d_set_fluid_points": (_i, [_p, _i, _p, _p]),
    "axs3d_set_fluid_elements": (_i, [_p, _i, _p, _p, _p]),
    "axs3d_set_fluid_k1d": (_i, [_p, _p]),
    "axs3d_set_fluid_axial_points": (_i, [_p, _i, _p]),
    "axs3d_set_fluid_surface_points": (_i, [_p, _i, _p]),
    "axs3d_set_solid_fluid_coupling": (_i, [_p, _i, _p, _p, _p]),
    "axs3d_num_fluid_modes": (_i, [_p, C.POINTER(C.c_int64)]),
    "axs3d_set_solid_material1d": (_i, [_p, _p]),
    "axs3d_set_solid_material3d": (_i, [_p, _i, _i, _p]),
    "axs3d_set_rtz_theta": (_i, [_p, _p]),
    "axs3d_set_moment_source": (_i, [_p, _i, _i, _p, _i, _p]),
    "axs3d_set_receivers": (_i, [_p, _i, _p, _p, _p, _i]),
    "axs3d_set_receiver_frames": (_i, [_p, _i, _p, _p, _p, _i]),
    "axs3d_add_clayton_solid": (_i, [_p, _i, _p, _p, _p, _p, _p]),
    "axs3d_add_clayton_solid3d": (_i, [_p, _i, _p, _p, _p, _p, _p]),
    "axs3d_add_clayton_fluid": (_i, [_p, _i, _p, _i, _p, _p]),
    "axs3d_add_sponge": (_i, [_p, _i, _i, _p, _i, _p]),
    "axs3d_set_dt": (_i, [_p, _d]),
    "axs3d_finalize": (_i, [_p]),
    "axs3d_apply_sources": (_i, [_p, _i]),
    "axs3d_compute_stiffness": (_i, [_p]),
    "axs3d_compute_stiffness_static": (_i, [_p]),
    "axs3d_apply_bc_before": (_i, [_p]),
    "axs3d_record": (_i, [_p, _i]),
    "axs3d_apply_bc_after": (_i, [_p]),
    "axs3d_apply_bc_after_accel": (_i, [_p]),
    "axs3d_set_time_scheme": (_i, [_p, _i]),
    "axs3d_symplectic_launch": (_i, [_p]),
    "axs3d_symplectic_update": (_i, [_p, _i]),
    "axs3d_stiff_to_accel": (_i, [_p]),
    "axs3d_newmark": (_i, [_p]),
    "axs3d_check_stability": (_i, [_p, _i]),
    "axs3d_run": (_i, [_p, _i, _i]),
    "axs3d_set_record_origin": (_i, [_p, _i]),
    "axs3d_set_graph": (_i, [_p, _i]),
    "axs3d_set_fluid_mass3d": (_i, [_p, _i, _p, _p]),
    "axs3d_set_coupling3d": (_i, [_p, _i, _p, _p, _p]),
    "axs3d_add_moment_source": (_i, [_p, _i, _i, _p, _i, _p]),
    "axs3d_set_instances": (_i, [_p, _i, _p, _p]),
    "axs3d_adjoint_run": (_i, [_p, _i, _i, _p, _p, _p]),
    "axs3d_adjoint_run_media": (_i, [_p, _i, _i, _p, _p, _p, _p]),
    "axs3d_adjoint_density": (_i, [_p, _p, _p, _p, _p]),
    "axs3d_read_seismograms": (_i, [_p, _p, _i]),
    "axs3d_num_modes": (_i, [_p, C.POINTER(C.c_int64)]),
    "axs3d_get_field": (_i, [_p, _i, _p]),
    "axs3d_set_field": (_i, [_p, _i, _p]),
    "axs3d_synchronize": (_i, [_p]),
    "axs3d_profile_enable": (_i, [_p, _i]),
    "axs3d_profile_read": (_i, [_p, _i, C.POINTER(C.c_double),
                                C.POINTER(C.c_int64)]),
    "axs3d_profile_reset": (_i, [_p]),
    "axs3d_fft_check": (_i, [_i, _i, _i, _p, C.POINTER(_i), C.POINTER(_i)]),
    "axs3d_autotune_info": (_i, [_p, C.POINTER(_i), C.POINTER(_i),
                                 C.POINTER(C.c_double)]),
    "axs3d_autotune_report": (_i, [_p, C.POINTER(C.c_double), _i,
                                   C.POINTER(_i)]),
}

_lib = None


class Axs3dError(RuntimeError):
    pass


def load():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise Axs3dError(
                "axisem3d_amd: libaxs3d.so is not built; run "
                "python -c 'import __graft_entry__ as g; g.build()'")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(ctx, rc, what=""):
    if rc != 0:
        msg = load().axs3d_last_error(ctx)
        msg = msg.decode() if msg else ""
        raise Axs3dError(f"{what}: {msg}" if what else msg)
