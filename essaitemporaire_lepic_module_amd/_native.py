"""ctypes binding of liblepic_hip.so (include/lepic.h).

The library is the only compute path of this package: if it is missing,
load() raises.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "liblepic_hip.so")

LEPIC_OK = 0
ERRORS = {1: "ERR_ARG", 2: "ERR_CAPACITY", 3: "ERR_DIVERGED", 4: "ERR_PMAX", 5: "ERR_HIP",
          6: "ERR_SPIN"}
MAX_COL = 128
MAX_SPECIES = 16
MCC_REPLAY_RAN2 = 0
MCC_PHILOX = 1

# Every symbol include/lepic.h declares.
EXPORTS = (
    "lepic_create", "lepic_destroy", "lepic_last_error", "lepic_set_stream", "lepic_synchronize",
    "lepic_version", "lepic_ran2", "lepic_seed", "lepic_load_particles_host",
    "lepic_load_uniform_device", "lepic_sort_particles_by_cell", "lepic_move_electrostatic",
    "lepic_move_boris", "lepic_apply_bc", "lepic_build_kq", "lepic_fill", "lepic_deposit",
    "lepic_deposit_fixed",
    "lepic_reduce_density", "lepic_build_rho", "lepic_poisson_solve", "lepic_efield",
    "lepic_efield_noalloc",
    "lepic_collisions", "lepic_collisions_deposited", "lepic_push_bc_deposit",
    "lepic_push_bc_deposit_inplace", "lepic_sort_compact", "lepic_bin_compact",
    "lepic_plane_moments", "lepic_electron_heating", "lepic_dielectric_phi",
    "lepic_pset_to32", "lepic_pset_from32", "lepic_round_field4", "lepic_push32",
    "lepic_deposit32", "lepic_fixed_density", "lepic_sort32", "lepic_collisions32",
    "lepic_poisson_solve_mixed", "lepic_poisson_solve_pcg", "lepic_adj_push",
    "lepic_adj_push_map", "lepic_adj_efield", "lepic_adj_poisson_rhs", "lepic_adj_bias",
    "lepic_adj_deposit", "lepic_plane_flux", "lepic_plane_flux_axis", "lepic_apply_deaths",
)


class LepicError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Grid(C.Structure):
    _fields_ = [("n", C.c_int32 * 3), ("h", C.c_double * 3)]


class PSet(C.Structure):
    _fields_ = [("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p),
                ("vx", C.c_void_p), ("vy", C.c_void_p), ("vz", C.c_void_p),
                ("cell", C.c_void_p), ("dead", C.c_void_p),
                ("n", C.c_int64), ("cap", C.c_int64)]


class BCParams(C.Structure):
    _fields_ = [("xmax", C.c_double), ("ymax", C.c_double), ("zmax", C.c_double),
                ("flag_pbc", C.c_int32), ("flag_nmn", C.c_int32),
                ("ptype", C.c_int32), ("tag_neg", C.c_int32),
                ("qmacro", C.c_double), ("mass", C.c_double), ("Nm", C.c_double),
                ("ngrid", C.c_int32), ("flag_die", C.c_int32), ("dtype", C.c_void_p),
                ("sum_q_xz", C.c_void_p), ("sum_q_yz", C.c_void_p)]


class Heating(C.Structure):
    _fields_ = [("nudt", C.c_double), ("Pabs", C.c_double), ("nu_h", C.c_double),
                ("xl_pow", C.c_double), ("xr_pow", C.c_double),
                ("flag_circxh", C.c_int32), ("flag_ahp", C.c_int32),
                ("R_ahp_moments", C.c_double), ("R_ahp", C.c_double),
                ("ymax", C.c_double), ("zmax", C.c_double), ("Nm", C.c_double),
                ("mass", C.c_double), ("vt0", C.c_double)]


class PSet32(C.Structure):
    _fields_ = [("cell", C.c_void_p), ("fx", C.c_void_p), ("fy", C.c_void_p),
                ("fz", C.c_void_p), ("vx", C.c_void_p), ("vy", C.c_void_p),
                ("vz", C.c_void_p), ("n", C.c_int64), ("cap", C.c_int64)]


class AdjSet(C.Structure):
    _fields_ = [("lx", C.c_void_p), ("ly", C.c_void_p), ("lz", C.c_void_p),
                ("lvx", C.c_void_p), ("lvy", C.c_void_p), ("lvz", C.c_void_p),
                ("n", C.c_int64)]


class BField32(C.Structure):
    _fields_ = [("mode", C.c_int32), ("B", C.c_double * 3), ("B4", C.c_void_p),
                ("magnetized", C.c_int32)]


class BField(C.Structure):
    _fields_ = [("mode", C.c_int32), ("B", C.c_double * 3), ("B4", C.c_void_p),
                ("magnetized", C.c_int32), ("nB", C.c_int32 * 3), ("hB", C.c_double * 3)]


class MGParams(C.Structure):
    _fields_ = [("flag_pbc", C.c_int32), ("flag_nmn", C.c_int32), ("eps0", C.c_double),
                ("ncycl", C.c_int32), ("eps", C.c_double), ("omega", C.c_double),
                ("ng", C.c_int32), ("graph", C.c_int32)]


class MGMixParams(C.Structure):
    _fields_ = [("inner_cycles", C.c_int32), ("nu", C.c_int32), ("nu_coarsest", C.c_int32),
                ("omega", C.c_double), ("precond_fp64", C.c_int32)]


class MGStats(C.Structure):
    _fields_ = [("cycles", C.c_int32), ("ktot", C.c_double), ("res", C.c_double),
                ("sweeps", C.c_int64 * 12)]


class Chem(C.Structure):
    _fields_ = [("ntype_tracked", C.c_int32), ("ntype_total", C.c_int32),
                ("mass", C.c_double * MAX_SPECIES), ("vt0", C.c_double * MAX_SPECIES),
                ("Nm", C.c_double * MAX_SPECIES), ("neutral_density", C.c_double * MAX_SPECIES),
                ("nu_uplim", C.c_double * MAX_SPECIES), ("p_ncol", C.c_int32 * MAX_SPECIES),
                ("ncol", C.c_int32), ("npt", C.c_int32),
                ("sig_list", C.c_void_p), ("sig_Er", C.c_void_p), ("sig", C.c_void_p),
                ("sig_Eth", C.c_void_p), ("col_info", C.c_void_p), ("sigv_mx", C.c_void_p)]


class MCCStats(C.Structure):
    _fields_ = [("Nc", C.c_int32 * MAX_SPECIES), ("accepted", C.c_int64 * MAX_SPECIES),
                ("err_coll", C.c_int64 * MAX_SPECIES), ("skipped_dead", C.c_int64 * MAX_SPECIES),
                ("appended", C.c_int64 * MAX_SPECIES)]


_lib = None
_lock = threading.Lock()


def load() -> C.CDLL:
    """Load liblepic_hip.so (built by essaitemporaire_lepic_module_amd.build)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with "
                              "essaitemporaire_lepic_module_amd.build.build() (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        vp, i32p, dp = C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_double)
        G, PS, P32 = C.POINTER(Grid), C.POINTER(PSet), C.POINTER(PSet32)
        AJ = C.POINTER(AdjSet)
        sig = {
            "lepic_create": (C.c_int, [C.c_int32, C.POINTER(vp)]),
            "lepic_destroy": (C.c_int, [vp]),
            "lepic_last_error": (C.c_char_p, [vp]),
            "lepic_set_stream": (C.c_int, [vp, vp]),
            "lepic_synchronize": (C.c_int, [vp]),
            "lepic_version": (C.c_char_p, []),
            "lepic_ran2": (C.c_double, [i32p]),
            "lepic_seed": (C.c_int32, [C.c_int32, C.c_int32]),
            "lepic_load_particles_host": (C.c_int, [G, vp, vp, C.c_double, C.c_int32, C.c_int32,
                                                    dp, dp, dp, i32p, PS, vp]),
            "lepic_load_uniform_device": (C.c_int, [vp, G, vp, C.c_int32, C.c_int32, dp,
                                                    C.c_uint64, PS]),
            "lepic_sort_particles_by_cell": (C.c_int, [vp, G, PS, PS, vp, vp]),
            "lepic_move_electrostatic": (C.c_int, [vp, G, PS, vp, C.c_double, C.c_double,
                                                   C.c_double]),
            "lepic_move_boris": (C.c_int, [vp, G, PS, vp, C.POINTER(BField), C.c_double,
                                           C.c_double, C.c_double]),
            "lepic_apply_bc": (C.c_int, [vp, G, PS, PS, vp, C.POINTER(BCParams), vp, vp,
                                         C.POINTER(C.c_int64)]),
            "lepic_build_kq": (C.c_int, [vp, G, vp, vp]),
            "lepic_fill": (C.c_int, [vp, vp, C.c_int64, C.c_double]),
            "lepic_deposit": (C.c_int, [vp, G, PS, vp, C.c_double, vp]),
            "lepic_deposit_fixed": (C.c_int, [vp, G, PS, vp, C.c_double, C.c_int32, vp]),
            "lepic_reduce_density": (C.c_int, [vp, G, vp, C.c_int32, vp, vp]),
            "lepic_build_rho": (C.c_int, [vp, G, C.c_int32, vp, dp, vp]),
            "lepic_poisson_solve": (C.c_int, [vp, G, vp, vp, vp, C.POINTER(MGParams),
                                              C.POINTER(MGStats)]),
            "lepic_poisson_solve_mixed": (C.c_int, [vp, G, vp, vp, vp, C.POINTER(MGParams),
                                                    C.POINTER(MGMixParams), C.POINTER(MGStats)]),
            "lepic_poisson_solve_pcg": (C.c_int, [vp, G, vp, vp, vp, C.POINTER(MGParams),
                                                    C.POINTER(MGMixParams), C.POINTER(MGStats)]),
            "lepic_efield": (C.c_int, [vp, G, vp, vp, vp]),
            "lepic_efield_noalloc": (C.c_int, [vp, G, vp, vp, vp]),
            "lepic_collisions": (C.c_int, [vp, G, PS, vp, C.POINTER(Chem), C.c_int32, C.c_double,
                                           C.c_int32, i32p, C.c_uint64, C.c_uint64, vp,
                                           C.POINTER(MCCStats)]),
            "lepic_collisions_deposited": (C.c_int, [vp, G, PS, vp, C.POINTER(Chem), C.c_int32,
                                                     C.c_double, C.c_int32, i32p, C.c_uint64,
                                                     C.c_uint64, vp, C.POINTER(MCCStats), vp,
                                                     vp]),
            "lepic_push_bc_deposit": (C.c_int, [vp, G, PS, PS, vp, C.POINTER(BField), C.c_double,
                                                C.c_double, C.c_double, vp, C.POINTER(BCParams),
                                                vp, vp, vp, vp, C.POINTER(C.c_int64)]),
            "lepic_push_bc_deposit_inplace": (C.c_int, [vp, G, PS, vp, C.POINTER(BField),
                                                        C.c_double, C.c_double, C.c_double, vp,
                                                        C.POINTER(BCParams), vp, vp, vp, vp, vp,
                                                        C.POINTER(C.c_int64)]),
            "lepic_sort_compact": (C.c_int, [vp, G, PS, PS]),
            "lepic_bin_compact": (C.c_int, [vp, G, PS, PS]),
            "lepic_plane_moments": (C.c_int, [vp, G, PS, C.c_double, C.c_double, C.c_int32,
                                              C.c_int32, C.c_int32, vp, vp, vp]),
            "lepic_electron_heating": (C.c_int, [vp, G, PS, C.POINTER(Heating), C.c_int32, i32p,
                                                 C.c_uint64, C.c_uint64, vp, dp]),
            "lepic_dielectric_phi": (C.c_int, [vp, G, vp, vp, C.c_int32, C.c_int32, vp, vp,
                                               C.c_double, vp]),
            "lepic_pset_to32": (C.c_int, [vp, G, PS, P32, C.POINTER(C.c_int64)]),
            "lepic_pset_from32": (C.c_int, [vp, G, P32, PS]),
            "lepic_round_field4": (C.c_int, [vp, vp, C.c_int64, vp]),
            "lepic_push32": (C.c_int, [vp, G, P32, vp, C.POINTER(BField32), C.c_double,
                                       C.c_double, C.c_double, vp, C.POINTER(BCParams),
                                       C.c_int32, vp, vp, vp, vp, vp]),
            "lepic_deposit32": (C.c_int, [vp, G, P32, C.c_double, C.c_int32, vp]),
            "lepic_fixed_density": (C.c_int, [vp, G, vp, vp, C.c_int32, vp]),
            "lepic_sort32": (C.c_int, [vp, G, P32, P32]),
            "lepic_collisions32": (C.c_int, [vp, G, P32, vp, C.POINTER(Chem), C.c_int32,
                                             C.c_double, C.c_uint64, C.c_uint64, vp,
                                             C.POINTER(MCCStats), vp, i32p]),
            "lepic_adj_push": (C.c_int, [vp, G, PS, AJ, vp, dp, C.c_double, C.c_double,
                                         C.c_double, vp, vp]),
            "lepic_adj_push_map": (C.c_int, [vp, G, PS, AJ, vp, vp, C.c_int32, C.c_double,
                                             C.c_double, C.c_double, vp]),
            "lepic_adj_efield": (C.c_int, [vp, G, vp, vp, vp]),
            "lepic_adj_poisson_rhs": (C.c_int, [vp, G, vp, vp, C.c_double, vp]),
            "lepic_adj_bias": (C.c_int, [vp, G, vp, C.c_int32, vp, vp, C.c_double, vp]),
            "lepic_adj_deposit": (C.c_int, [vp, G, PS, AJ, vp, vp, C.c_double, C.c_double]),
            "lepic_plane_flux": (C.c_int, [vp, G, PS, C.c_int32, C.c_double, vp, AJ,
                                           C.c_double]),
            "lepic_plane_flux_axis": (C.c_int, [vp, G, PS, C.c_int32, C.c_int32, C.c_double, vp,
                                                AJ, C.c_double]),
            "lepic_apply_deaths": (C.c_int, [vp, PS, vp, C.c_int64]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
        return L


def grid(n, h) -> Grid:
    g = Grid()
    for d in range(3):
        g.n[d] = int(n[d])
        g.h[d] = float(h[d])
    return g


class Context:
    """One lepic_ctx bound to a device; calls run on torch's current stream."""

    def __init__(self, device: int = 0):
        self.lib = load()
        self.device = int(device)
        h = C.c_void_p()
        st = self.lib.lepic_create(self.device, C.byref(h))
        if st != LEPIC_OK:
            raise LepicError(st, f"lepic_create(device={device}) failed (no usable GPU?)")
        self.h = h

    def check(self, st: int):
        if st != LEPIC_OK:
            raise LepicError(st, self.lib.lepic_last_error(self.h).decode())

    def bind_stream(self, stream_ptr):
        self.check(self.lib.lepic_set_stream(self.h, C.c_void_p(stream_ptr)))

    def synchronize(self):
        self.check(self.lib.lepic_synchronize(self.h))

    def close(self):
        if self.h:
            self.lib.lepic_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
