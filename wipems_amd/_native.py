"""ctypes bindings of the two native libraries of the package.

``libwipems_fdtd.so`` is the HIP engine (C ABI in include/wipems_fdtd.h) and
``libwipems_operator.so`` the host operator builder (include/wipems_operator.h).
Loading fails loudly when a library is missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_int32, c_size_t, c_uint32, c_uint64, c_void_p

PKG = os.path.dirname(os.path.abspath(__file__))


class FdtdError(RuntimeError):
    """Raised when a native call returns an error status."""


class GridDesc(ctypes.Structure):
    _fields_ = [("num_lines", c_uint32 * 3), ("dT", c_double), ("precision", c_int32), ("flags", c_uint32)]


class ExcitationDesc(ctypes.Structure):
    _fields_ = [
        ("length", c_uint32), ("period", c_uint32),
        ("signal_volt", POINTER(c_double)), ("signal_curr", POINTER(c_double)),
        ("volt_count", c_uint32), ("volt_pos", POINTER(c_uint32)), ("volt_dir", POINTER(c_uint32)),
        ("volt_amp", POINTER(c_double)), ("volt_delay", POINTER(c_uint32)),
        ("curr_count", c_uint32), ("curr_pos", POINTER(c_uint32)), ("curr_dir", POINTER(c_uint32)),
        ("curr_amp", POINTER(c_double)), ("curr_delay", POINTER(c_uint32)),
    ]


_ENGINE_SIGS = {
    "fdtd_create": [POINTER(GridDesc), POINTER(c_void_p)],
    "fdtd_destroy": [c_void_p],
    "fdtd_last_error": [c_char_p, c_size_t],
    "fdtd_device_name": [c_char_p, c_size_t],
    "fdtd_set_coeffs": [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p],
    "fdtd_set_coeffs_compressed": [c_void_p, POINTER(c_uint32), c_int, c_void_p, c_uint32],
    "fdtd_set_coeffs_factored": [c_void_p, POINTER(c_uint32), POINTER(c_float), c_uint32, POINTER(c_float), POINTER(c_float)],
    "fdtd_set_storage_scale": [c_void_p, POINTER(c_double), POINTER(c_double)],
    "fdtd_set_excitation": [c_void_p, POINTER(ExcitationDesc)],
    "fdtd_add_upml": [c_void_p, POINTER(c_uint32), POINTER(c_uint32), c_int] + [c_void_p] * 6,
    "fdtd_add_ade": [c_void_p, c_uint32, POINTER(c_uint32), c_int] + [c_void_p] * 6 + [c_uint32],
    "fdtd_set_tfsf": [c_void_p, c_int, c_uint32, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_uint32),
                      POINTER(c_double), POINTER(c_double), c_uint32],
    "fdtd_add_probe": [c_void_p, c_int, POINTER(c_uint32), POINTER(c_uint32), c_int, c_uint32, c_double,
                       c_uint32, c_uint32, POINTER(c_double), c_uint32, POINTER(c_int)],
    "fdtd_probe_num_samples": [c_void_p, c_int, POINTER(c_uint32)],
    "fdtd_read_probe": [c_void_p, c_int, POINTER(c_double), POINTER(c_double), c_uint32],
    "fdtd_read_probe_fd": [c_void_p, c_int, POINTER(c_double), POINTER(c_double)],
    "fdtd_set_edge_lengths": [c_void_p, POINTER(c_double), POINTER(c_double)],
    "fdtd_add_dft_box": [c_void_p, c_int, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_double), c_uint32, c_uint32,
                         POINTER(c_int)],
    "fdtd_dft_num_samples": [c_void_p, c_int, POINTER(c_uint64)],
    "fdtd_read_dft": [c_void_p, c_int, POINTER(c_double)],
    "fdtd_iterate": [c_void_p, c_uint32],
    "fdtd_synchronize": [c_void_p],
    "fdtd_get_num_timesteps": [c_void_p, POINTER(c_uint32)],
    "fdtd_set_use_graph": [c_void_p, c_int],
    "fdtd_set_fuse_hooks": [c_void_p, c_int],
    "fdtd_hooks_fused": [c_void_p, POINTER(c_int)],
    "fdtd_set_pml_compression": [c_void_p, c_int],
    "fdtd_pml_compressed": [c_void_p, POINTER(c_int)],
    "fdtd_set_pml_split": [c_void_p, c_int],
    "fdtd_set_ade_compression": [c_void_p, c_int],
    "fdtd_ade_compressed": [c_void_p, c_int, POINTER(c_int), POINTER(c_int)],
    "fdtd_add_mur": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_uint32],
    "fdtd_set_rlc": [c_void_p, c_uint32, POINTER(c_uint32), POINTER(c_uint32), c_int, c_void_p],
    "nf2ff_create": [c_double, POINTER(c_double), c_uint32, POINTER(c_double), c_uint32, POINTER(c_double), c_double,
                     c_double, c_double, POINTER(c_void_p)],
    "nf2ff_add_plane": [c_void_p, POINTER(c_double), c_uint32, POINTER(c_double), c_uint32, POINTER(c_double), c_uint32,
                        POINTER(c_double), POINTER(c_double)],
    "nf2ff_get": [c_void_p] + [POINTER(c_double)] * 7,
    "nf2ff_destroy": [c_void_p],
    "nf2ff_set_mirror": [c_void_p, c_int, c_int, c_double],
    "fdtd_add_td_box": [c_void_p, c_int, POINTER(c_uint32), POINTER(c_uint32), c_uint32, c_void_p, c_void_p, POINTER(c_int)],
    "fdtd_set_material": [c_void_p, c_int, POINTER(c_float)],
    "fdtd_add_dft_dump_box": [c_void_p, c_int, c_int, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_double), c_uint32,
                              c_uint32, POINTER(c_int)],
    "fdtd_add_td_dump_box": [c_void_p, c_int, c_int, POINTER(c_uint32), POINTER(c_uint32), c_uint32, c_void_p, c_void_p,
                             POINTER(c_int)],
    "fdtd_td_num_samples": [c_void_p, c_int, POINTER(c_uint64)],
    "fdtd_set_steady_state": [c_void_p, c_double, c_uint32, POINTER(c_uint32), POINTER(c_uint32)],
    "fdtd_get_steady_state_period": [c_void_p, POINTER(c_uint32)],
    "fdtd_steady_state_last_diff": [c_void_p, POINTER(c_double)],
    "fdtd_pml_split": [c_void_p, POINTER(c_int)],
    "fdtd_set_tile_order": [c_void_p, c_int, c_int],
    "fdtd_set_shell_align": [c_void_p, c_int],
    "fdtd_tile_order": [c_void_p, POINTER(c_int), POINTER(c_int)],
    "fdtd_set_kernel_timing": [c_void_p, c_int],
    "fdtd_get_kernel_times": [c_void_p, POINTER(c_double), POINTER(c_double), POINTER(c_uint32), POINTER(c_uint32)],
    "fdtd_calc_energy": [c_void_p, POINTER(c_double)],
    "fdtd_calc_energy_noalloc": [c_void_p, POINTER(c_double)],
    "fdtd_field_absmax": [c_void_p, c_int, POINTER(c_double)],
    "fdtd_get_field": [c_void_p, c_int, c_int, c_void_p],
    "fdtd_set_field": [c_void_p, c_int, c_int, c_void_p],
    "fdtd_get_value": [c_void_p, c_int, c_uint32, c_uint32, c_uint32, c_uint32, POINTER(c_double)],
    "fdtd_set_value": [c_void_p, c_int, c_uint32, c_uint32, c_uint32, c_uint32, c_double],
    "fdtd_device_bytes": [c_void_p, POINTER(c_uint64)],
    "fdtd_copy_bench": [c_uint64, c_int, POINTER(c_double)],
    "fdtd_batch_create": [POINTER(c_void_p), c_int, POINTER(c_void_p)],
    "fdtd_batch_destroy": [c_void_p],
    "fdtd_batch_set_block": [c_void_p, c_int],
    "fdtd_batch_set_split": [c_void_p, c_int],
    "fdtd_batch_set_use_graph": [c_void_p, c_int],
    "fdtd_batch_iterate": [c_void_p, c_uint32],
    "fdtd_batch_synchronize": [c_void_p],
    "fdtd_batch_set_kernel_timing": [c_void_p, c_int],
    "fdtd_batch_get_kernel_times": [c_void_p, POINTER(c_double), POINTER(c_double), POINTER(c_uint32)],
}

_OPERATOR_SIGS = {
    "wop_create": [POINTER(c_void_p)],
    "wop_destroy": [c_void_p],
    "wop_last_error": [c_char_p, c_size_t],
    "wop_set_mesh": [c_void_p, c_int, POINTER(c_double), c_uint32],
    "wop_set_grid_delta": [c_void_p, c_double],
    "wop_set_boundary": [c_void_p, POINTER(c_int), POINTER(c_uint32)],
    "wop_set_timestep_factor": [c_void_p, c_double],
    "wop_set_timestep": [c_void_p, c_double],
    "wop_set_threads": [c_void_p, c_int],
    "wop_set_verbose": [c_void_p, c_int],
    "wop_add_material": [c_void_p, c_double, c_double, c_double, c_double, POINTER(c_int)],
    "wop_add_metal": [c_void_p, POINTER(c_int)],
    "wop_add_lorentz_material": [c_void_p] + [c_double] * 7 + [POINTER(c_int)],
    "wop_add_lumped_element": [c_void_p, c_int, c_double, c_double, c_int, POINTER(c_int)],
    "wop_add_excitation": [c_void_p, c_int, POINTER(c_double), c_double, POINTER(c_int)],
    "wop_add_box": [c_void_p, c_int, POINTER(c_double), POINTER(c_double), c_int],
    "wop_add_sphere": [c_void_p, c_int, POINTER(c_double), c_double, c_int],
    "wop_set_plane_wave": [c_void_p, POINTER(c_double), POINTER(c_double), POINTER(c_double), POINTER(c_double), c_double],
    "wop_get_tfsf_info": [c_void_p, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_uint32)],
    "wop_get_tfsf": [c_void_p, c_int, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_uint32), POINTER(c_double),
                     POINTER(c_double)],
    "wop_set_excite_signal": [c_void_p, c_int, c_double, c_double],
    "wop_build": [c_void_p, c_int, c_uint32],
    "wop_get_num_lines": [c_void_p, POINTER(c_uint32)],
    "wop_get_timestep": [c_void_p, POINTER(c_double)],
    "wop_get_edge_lengths": [c_void_p, c_int, POINTER(c_double)],
    "wop_get_excite_info": [c_void_p, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_uint32)],
    "wop_get_signal_period": [c_void_p, POINTER(c_double), POINTER(c_uint32)],
    "wop_get_signal": [c_void_p, POINTER(c_double), POINTER(c_double)],
    "wop_get_coeffs": [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p],
    "wop_get_ec": [c_void_p, POINTER(c_double), POINTER(c_double), POINTER(c_double), POINTER(c_double)],
    "wop_set_material_store_flags": [c_void_p, c_int, c_int],
    "wop_get_material": [c_void_p, c_int, POINTER(c_float)],
    "wop_set_density": [c_void_p, c_int, c_double],
    "wop_set_row_reuse": [c_void_p, c_int],
    "wop_get_cell_materials": [c_void_p, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_float), POINTER(c_float),
                               POINTER(c_float)],
    "wop_compress": [c_void_p, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_double), c_uint32],
    "wop_factor": [c_void_p, c_int, c_uint32, POINTER(c_uint32), POINTER(c_float), POINTER(c_uint32), POINTER(c_float),
                   POINTER(c_float), POINTER(c_double)],
    "wop_build_streamed": [c_void_p, c_uint32, POINTER(c_uint32), c_uint32, POINTER(c_uint32)],
    "wop_get_streamed_table": [c_void_p, POINTER(c_double)],
    "wop_get_excitation_counts": [c_void_p, POINTER(c_uint32), POINTER(c_uint32)],
    "wop_get_excitation": [c_void_p, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_double), POINTER(c_uint32),
                           POINTER(c_uint32), POINTER(c_uint32), POINTER(c_double), POINTER(c_uint32)],
    "wop_get_num_upml": [c_void_p, POINTER(c_uint32)],
    "wop_get_num_mur": [c_void_p, POINTER(c_uint32)],
    "wop_add_lumped_rlc": [c_void_p, c_int, c_int, c_double, c_double, c_double, c_int, POINTER(c_int)],
    "wop_get_rlc_count": [c_void_p, POINTER(c_uint32)],
    "wop_get_rlc": [c_void_p, POINTER(c_uint32), POINTER(c_uint32), c_int, c_void_p],
    "wop_get_mur": [c_void_p, c_uint32, POINTER(c_int), c_int, c_void_p, c_void_p, POINTER(c_uint32)],
    "wop_set_mur_phase_velocity": [c_void_p, c_int, c_double],
    "wop_get_upml_range": [c_void_p, c_uint32, POINTER(c_uint32), POINTER(c_uint32)],
    "wop_get_upml": [c_void_p, c_uint32, c_int] + [c_void_p] * 6,
    "wop_get_num_ade": [c_void_p, POINTER(c_uint32)],
    "wop_get_ade_info": [c_void_p, c_uint32, POINTER(c_uint32), POINTER(c_uint32)],
    "wop_get_ade": [c_void_p, c_uint32, POINTER(c_uint32), c_int] + [c_void_p] * 6,
    "wop_snap_to_mesh_line": [c_void_p, c_int, c_double, c_int, POINTER(c_uint32), POINTER(c_int)],
}

_DUMP_SIGS = {
    "wdump_last_error": [c_char_p, c_size_t],
    "wdump_create": [c_char_p, POINTER(c_double), c_uint32, POINTER(c_double), c_uint32, POINTER(c_double), c_uint32,
                     c_double, POINTER(c_void_p)],
    "wdump_write_td": [c_void_p, c_uint32, c_double, c_void_p],
    "wdump_write_fd": [c_void_p, c_uint32, POINTER(c_double), POINTER(c_double)],
    "wdump_write_sar": [c_void_p, c_uint32, POINTER(c_double), POINTER(c_double), POINTER(c_float), POINTER(c_float),
                        POINTER(c_float), c_int],
    "wdump_close": [c_void_p],
}

_libs = {}


def _load(fname, sigs):
    if fname in _libs:
        return _libs[fname]
    path = os.path.join(PKG, fname)
    if not os.path.exists(path):
        raise FdtdError("native library %s is missing; build it with `python -m wipems_amd.build`" % path)
    lib = ctypes.CDLL(path)
    for name, argtypes in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = c_int
    if "fdtd_version" in dir(lib) or fname.startswith("libwipems_fdtd"):
        lib.fdtd_version.restype = c_char_p
        lib.fdtd_version.argtypes = []
    _libs[fname] = lib
    return lib


def engine_lib():
    """The HIP engine library (raises FdtdError if it is not built)."""
    return _load("libwipems_fdtd.so", _ENGINE_SIGS)


def operator_lib():
    """The host operator library (raises FdtdError if it is not built)."""
    return _load("libwipems_operator.so", _OPERATOR_SIGS)


def dump_lib():
    """The HDF5 field-dump writer (raises FdtdError if it is not built)."""
    return _load("libwipems_dump.so", _DUMP_SIGS)


def engine_path():
    return os.path.join(PKG, "libwipems_fdtd.so")


def check(lib, rc, errfn):
    if rc != 0:
        buf = ctypes.create_string_buffer(1024)
        getattr(lib, errfn)(buf, 1024)
        raise FdtdError("native call failed (%d): %s" % (rc, buf.value.decode(errors="replace")))
    return rc
