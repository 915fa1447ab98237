"""ctypes bindings of libblade_mi355x.so (include/blade.h, include/blade_mi355x.h).

The library is loaded from the package directory; a missing library raises
immediately - there is no fallback implementation.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libblade_mi355x.so")

NTERMS = 23
TERM_NAMES = ["bond", "angle", "urey", "dihe", "impr", "cmap", "nb14", "nbdirect", "nbrecip",
              "nbrecipself", "nbrecipexcl", "lambda", "theta", "cats", "noe", "harmonic", "mmfp",
              "resd", "mlp", "bias", "potential", "kinetic", "total"]
TERM = {n: i for i, n in enumerate(TERM_NAMES)}
CALC_ENERGY = 1

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)


class Topology(C.Structure):
    """Mirror of bm_topology (include/blade_mi355x.h)."""
    _fields_ = [
        ("n_atoms", C.c_int32), ("charge", _f64p), ("mass", _f64p), ("type_idx", _i32p),
        ("site_block", _i32p), ("n_types", C.c_int32), ("vdw_c12", _f64p), ("vdw_c6", _f64p),
        ("n_bonds", C.c_int32), ("n_bonds12", C.c_int32), ("bond_idx", _i32p), ("bond_sb", _i32p),
        ("bond_kb", _f64p), ("bond_b0", _f64p),
        ("n_angles", C.c_int32), ("angle_idx", _i32p), ("angle_sb", _i32p), ("angle_k", _f64p),
        ("angle_t0", _f64p),
        ("n_dihes", C.c_int32), ("dihe_idx", _i32p), ("dihe_sb", _i32p), ("dihe_k", _f64p),
        ("dihe_d0", _f64p), ("dihe_n", _i32p),
        ("n_imprs", C.c_int32), ("impr_idx", _i32p), ("impr_sb", _i32p), ("impr_k", _f64p),
        ("impr_d0", _f64p), ("impr_n", _i32p),
        ("n_nb14", C.c_int32), ("nb14_idx", _i32p), ("nb14_sb", _i32p), ("nb14_qq", _f64p),
        ("nb14_c12", _f64p), ("nb14_c6", _f64p), ("nb14_e14fac", _f64p),
        ("n_nbex", C.c_int32), ("nbex_idx", _i32p), ("nbex_sb", _i32p), ("nbex_qq", _f64p),
        ("n_excl", C.c_int32), ("excl_idx", _i32p),
        ("n_tri", C.c_int32), ("n_br1", C.c_int32), ("n_br2", C.c_int32), ("n_br3", C.c_int32),
        ("tri_idx", _i32p), ("tri_b0", _f64p), ("br1_idx", _i32p), ("br1_b0", _f64p),
        ("br2_idx", _i32p), ("br2_b0", _f64p), ("br3_idx", _i32p), ("br3_b0", _f64p),
        ("n_blocks", C.c_int32), ("n_sites", C.c_int32), ("block_site", _i32p),
        ("site_bound", _i32p), ("theta_mass", _f64p), ("block_fixed", _i32p),
        ("fixed_bias", _f64p), ("theta_friction", _f64p),
        ("n_varbias", C.c_int32), ("vb_i", _i32p), ("vb_j", _i32p), ("vb_type", _i32p),
        ("vb_n", _i32p), ("vb_l0", _f64p), ("vb_k", _f64p),
        ("n_cats", C.c_int32), ("cats_bound", _i32p), ("cats_idx", _i32p), ("cats_k", C.c_double),
        ("block_charge", _f64p), ("cr_k1", C.c_double), ("cr_k2", C.c_double), ("cr_k3", C.c_double),
        ("cr_q0", C.c_double), ("cr_w", C.c_double),
        ("n_cmaps", C.c_int32), ("cmap_idx", _i32p), ("cmap_sb", _i32p), ("cmap_type", _i32p),
        ("n_cmap_types", C.c_int32), ("cmap_ngrid", _i32p), ("cmap_off", _i32p), ("cmap_grid", _f64p),
        ("tcoll_k", _f64p), ("tcoll_n", _f64p), ("tinde_k", _f64p),
        ("piecewise", C.c_int32), ("pw_width", C.c_double), ("pw_k", C.c_double),
        ("edge_k", C.c_double), ("edge_n", C.c_double), ("edge_alpha", C.c_double), ("edge_phi", C.c_double),
        ("fix_lambda", C.c_int32),
        ("n_noe", C.c_int32), ("noe_idx", _i32p), ("noe_par", _f64p),
        ("n_harm", C.c_int32), ("harm_idx", _i32p), ("harm_par", _f64p),
        ("n_borest", C.c_int32), ("borest_idx", _i32p), ("borest_block", _i32p), ("borest_par", _f64p),
        ("n_anrest", C.c_int32), ("anrest_idx", _i32p), ("anrest_block", _i32p), ("anrest_par", _f64p),
        ("n_direst", C.c_int32), ("direst_idx", _i32p), ("direst_block", _i32p), ("direst_n", _i32p),
        ("direst_par", _f64p),
        ("n_resd", C.c_int32), ("resd_idx", _i32p), ("resd_par", _f64p),
        ("bond_soft", _i32p), ("angle_soft", _i32p), ("dihe_soft", _i32p), ("impr_soft", _i32p),
        ("n_virt2", C.c_int32), ("virt2_idx", _i32p), ("virt2_par", _f64p),
        ("n_virt3", C.c_int32), ("virt3_idx", _i32p), ("virt3_par", _f64p),
    ]

    def array(self, name, count, width=1):
        """numpy copy of pointer field `name` holding count*width items."""
        p = getattr(self, name)
        n = int(count) * width
        if n == 0 or not p:
            dt = np.int32 if p._type_ is C.c_int32 else np.float64
            return np.zeros((0,) if width == 1 else (0, width), dtype=dt)
        a = np.ctypeslib.as_array(p, shape=(n,)).copy()
        return a if width == 1 else a.reshape(-1, width)


class RunParams(C.Structure):
    """Mirror of bm_runparams (include/blade_mi355x.h)."""
    _fields_ = [
        ("rcut", C.c_double), ("rswitch", C.c_double), ("beta_ewald", C.c_double),
        ("vdw_method", C.c_int32), ("elec_method", C.c_int32), ("grid", C.c_int32 * 3),
        ("order", C.c_int32), ("dt", C.c_double), ("temperature", C.c_double),
        ("gamma", C.c_double), ("msld_gamma", C.c_double), ("shake_tol", C.c_double),
        ("fnex", C.c_double), ("soft_core", C.c_int32), ("soft_core14", C.c_int32),
        ("soft_bond_exp", C.c_double), ("box", C.c_double * 3), ("freq_domdec", C.c_int32),
        ("cull_pad", C.c_double), ("seed", C.c_uint64), ("fft_mode", C.c_int32),
        ("tile_chunk", C.c_int32), ("soft_bond_alpha", C.c_double), ("soft_bond_power", C.c_double),
    ]


class TuneTrial(C.Structure):
    """Mirror of bm_tune_trial."""
    _fields_ = [("fft_mode", C.c_int32), ("freq_domdec", C.c_int32), ("tile_chunk", C.c_int32),
                ("status", C.c_int32), ("ms_per_step", C.c_double)]


class TuneResult(C.Structure):
    """Mirror of bm_tune_result."""
    _fields_ = [("fft_mode", C.c_int32), ("freq_domdec", C.c_int32), ("tile_chunk", C.c_int32),
                ("ms_per_step", C.c_double), ("n_trials", C.c_int32), ("trial", TuneTrial * 32)]

    def as_dict(self):
        return {"fft_mode": self.fft_mode, "freq_domdec": self.freq_domdec, "tile_chunk": self.tile_chunk,
                "ms_per_step": round(self.ms_per_step, 5),
                "trials": [{"fft_mode": t.fft_mode, "freq_domdec": t.freq_domdec, "tile_chunk": t.tile_chunk,
                            "status": t.status, "ms_per_step": round(t.ms_per_step, 5)}
                           for t in self.trial[:self.n_trials]]}


_lib = None


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libblade_mi355x.so is not built: run `python -m blade_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, i, d, s = C.c_void_p, C.c_int, C.c_double, C.c_char_p
    sig = {
        # reference-compatible API
        "blade_init_system": (vp, [i, C.POINTER(C.c_int)]),
        "blade_dest_system": (None, [vp]),
        "blade_set_verbose": (None, [vp, i]),
        "blade_init_parameters": (None, [vp]),
        "blade_dest_parameters": (None, [vp]),
        "blade_add_parameter_atoms": (None, [vp, s, d]),
        "blade_add_parameter_bonds": (None, [vp, s, s, d, d]),
        "blade_add_parameter_angles": (None, [vp, s, s, s, d, d, d, d]),
        "blade_add_parameter_dihes": (None, [vp, s, s, s, s, d, i, d]),
        "blade_add_parameter_imprs": (None, [vp, s, s, s, s, d, i, d]),
        "blade_add_parameter_nbonds": (None, [vp, s, d, d, d, d, d, i]),
        "blade_add_parameter_nbfixs": (None, [vp, s, s, d, d, d, d]),
        "blade_init_structure": (None, [vp]),
        "blade_add_atom": (None, [vp, i, s, s, s, s, s, d, d]),
        "blade_add_bond": (None, [vp, i, i]),
        "blade_add_angle": (None, [vp, i, i, i]),
        "blade_add_dihe": (None, [vp, i, i, i, i]),
        "blade_add_impr": (None, [vp, i, i, i, i]),
        "blade_add_shake": (None, [vp, i]),
        "blade_add_virt2": (None, [vp, i, i, i, d, d]),
        "blade_add_virt3": (None, [vp, i, i, i, i, d, d, d]),
        "blade_add_noe": (None, [vp, i, i, d, d, d, d, d, d, d, d, d, d, C.c_bool]),
        "blade_add_harmonic": (None, [vp, i, d, d, d, d, d]),
        "blade_add_borest": (None, [vp, i, i, d, d, i]),
        "blade_add_anrest": (None, [vp, i, i, i, d, d, i]),
        "blade_add_direst": (None, [vp, i, i, i, i, d, i, d, d, i]),
        "blade_add_resd": (None, [vp, i, i, i, i, d, d, d, d]),
        "blade_init_coordinates": (None, [vp, i]),
        "blade_add_coordinates_position": (None, [vp, i, d, d, d]),
        "blade_add_coordinates_velocity": (None, [vp, i, d, d, d]),
        "blade_add_coordinates_box": (None, [vp, i, d, d, d, d, d, d]),
        "blade_init_msld": (None, [vp, i]),
        "blade_add_msld_atomassignment": (None, [vp, i, i]),
        "blade_add_msld_initialconditions": (None, [vp, i, i, d, d, d, d, d]),
        "blade_add_msld_termscaling": (None, [vp, i, i, i, i, i, i]),
        "blade_add_msld_flags": (None, [vp, d, d, i, i, i, d, d, d, d, i]),
        "blade_add_msld_bias": (None, [vp, i, i, i, d, d, i]),
        "blade_add_msld_block_friction": (None, [vp, i, d]),
        "blade_set_msld_block_fixed": (None, [vp, i, i]),
        "blade_add_msld_thetacollbias": (None, [vp, i, i, d, d]),
        "blade_add_msld_thetaindebias": (None, [vp, i, i, d]),
        "blade_set_msld_thetaedgebias": (None, [vp, d, d, d, d]),
        "blade_set_msld_piecewise_constraint": (None, [vp, i, d, d]),
        "blade_add_msld_softbond": (None, [vp, i, i]),
        "blade_add_msld_atomrestraint": (None, [vp]),
        "blade_add_msld_charges": (None, [vp, d, d, d, d, d]),
        "blade_add_parameter_cmaps": (None, [vp] + [s] * 8 + [i]),
        "blade_add_parameter_cmaps_fill": (None, [vp] + [s] * 8 + [i, i, d]),
        "blade_add_cmap": (None, [vp] + [i] * 8),
        "blade_add_msld_atomrestraint_element": (None, [vp, i]),
        "blade_init_run": (None, [vp]),
        "blade_add_run_flags": (None, [vp, d, d, d, d, i, i, d, i, i, i, i, d]),
        "blade_add_run_dynopts": (None, [vp, i, i, i, d, d, i, d, d]),
        "blade_run_energy": (None, [vp]),
        # extensions
        "blade_last_error": (s, []),
        "blade_system_error": (s, [vp]),
        "blade_run_dynamics": (i, [vp, i, i, i]),
        "blade_get_energy": (i, [vp, _f64p]),
        "blade_get_forces": (i, [vp, _f64p]),
        "blade_get_positions": (i, [vp, _f64p]),
        "blade_set_positions": (i, [vp, _f64p]),
        "blade_get_lambda": (i, [vp, _f64p]),
        "blade_get_dudl": (i, [vp, _f64p]),
        "blade_get_topology": (C.POINTER(Topology), [vp]),
        "blade_get_runparams": (C.POINTER(RunParams), [vp]),
        "blade_get_device_system": (vp, [vp]),
        "blade_prepare": (i, [vp]),
        "blade_set_seed": (None, [vp, C.c_uint64]),
        "blade_set_domdec_frequency": (None, [vp, i]),
        "blade_set_fft_mode": (None, [vp, i]),
        "blade_set_tile_chunk": (None, [vp, i]),
        "blade_set_output_nrg": (i, [vp, C.c_char_p, i]),
        "blade_set_output_lmd": (i, [vp, C.c_char_p, i]),
        "blade_set_output_xtc": (i, [vp, C.c_char_p, i]),
        "blade_interpretter": (None, [C.c_char_p, vp]),
        "blade_write_xtc_frame": (i, [C.c_char_p, i, i, i, C.c_float, C.POINTER(C.c_float),
                                      C.POINTER(C.c_float), C.c_float]),
        "blade_set_lmd_format": (None, [vp, i]),
        "blade_write_checkpoint": (i, [vp, C.c_char_p]),
        "blade_get_npt_stats": (i, [vp, C.POINTER(C.c_long), C.POINTER(C.c_long), _f64p]),
        "blade_read_checkpoint": (i, [vp, C.c_char_p]),
        "blade_autotune": (i, [vp, i, C.POINTER(TuneResult)]),
        "blade_batch_autotune": (i, [i, C.POINTER(vp), i, C.POINTER(TuneResult)]),
        "blade_read_parameter_file": (i, [vp, C.c_char_p]),
        "blade_read_psf": (i, [vp, C.c_char_p]),
        "blade_read_crd": (i, [vp, C.c_char_p]),
        "blade_read_pdb": (i, [vp, C.c_char_p]),
        "blade_structure_atom_count": (i, [vp]),
        "blade_structure_atom_names": (i, [vp, i, C.c_char_p, C.c_char_p, C.c_char_p, i]),
        # kernel layer
        "bm_version": (s, []),
        "bm_last_error": (s, []),
        "bm_system_create": (i, [i, C.POINTER(Topology), C.POINTER(RunParams), _f64p, _f64p, _f64p,
                                 _f64p, C.POINTER(vp)]),
        "bm_system_destroy": (i, [vp]),
        "bm_domdec_update": (i, [vp, i]),
        "bm_calc_force": (i, [vp, C.c_long, C.c_uint]),
        "bm_reset_force": (i, [vp, C.c_uint]),
        "bm_getforce_nbdirect": (i, [vp, C.c_uint]),
        "bm_getforce_restrain": (i, [vp, C.c_uint]),
        "bm_getforce_ewaldself": (i, [vp, C.c_uint]),
        "bm_getforce_nb14": (i, [vp, C.c_uint]),
        "bm_getforce_nbex": (i, [vp, C.c_uint]),
        "bm_getforce_msldbias": (i, [vp, C.c_uint]),
        "bm_holonomic_position": (i, [vp]),
        "bm_holonomic_rectify": (i, [vp]),
        "bm_holonomic_velocity": (i, [vp]),
        "bm_msld_chain_rule": (i, [vp]),
        "bm_getforce_ewald": (i, [vp, C.c_uint]),
        "bm_getforce_bonded": (i, [vp, C.c_uint]),
        "bm_update": (i, [vp, C.c_long, i]),
        "bm_dynamics": (i, [vp, C.c_long, C.c_long, i, i]),
        "bm_capture_graphs": (i, [vp]),
        "bm_set_list_capacity": (i, [vp, C.c_longlong, i]),
        "bm_get_list_capacity": (i, [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_int)]),
        "bm_profile_steps": (i, [vp, C.c_long, C.c_long, _f64p]),
        "bm_time_nbdirect": (i, [vp, C.c_int, _f64p]),
        "bm_time_pme": (i, [vp, C.c_int, _f64p]),
        "bm_autotune": (i, [i, C.POINTER(Topology), C.POINTER(RunParams), _f64p, _f64p, _f64p, _f64p, i,
                            C.POINTER(TuneResult)]),
        "bm_batch_autotune": (i, [i, i, C.POINTER(C.POINTER(Topology)), C.POINTER(C.POINTER(RunParams)),
                                  C.POINTER(_f64p), C.POINTER(_f64p), C.POINTER(_f64p), C.POINTER(_f64p), i,
                                  C.POINTER(TuneResult)]),
        "bm_calibration_copy": (i, [i, C.c_size_t, i, _f64p]),
        "bm_batch_create": (i, [i, C.POINTER(vp), C.POINTER(vp)]),
        "bm_batch_dynamics": (i, [vp, C.c_long, C.c_long, i]),
        "bm_batch_capture_graphs": (i, [vp]),
        "bm_batch_sync": (i, [vp]),
        "bm_volume_move": (i, [vp, _f64p]),
        "bm_volume_restore": (i, [vp, _f64p]),
        "bm_get_box": (i, [vp, _f64p]),
        "bm_uniform": (C.c_double, [C.c_uint64, C.c_long, C.c_uint32]),
        "bm_batch_exchange": (i, [vp, C.c_long, i, C.c_uint64, _i32p, _i32p, _f64p]),
        "bm_batch_exchange_noalloc": (i, [vp, C.c_long, i, C.c_uint64, _i32p, _i32p, _f64p]),
        "bm_batch_destroy": (i, [vp]),
        "bm_sync": (i, [vp]),
        "bm_get_positions": (i, [vp, _f64p, _f64p]),
        "bm_set_positions": (i, [vp, _f64p, _f64p]),
        "bm_get_velocities": (i, [vp, _f64p, _f64p]),
        "bm_set_velocities": (i, [vp, _f64p, _f64p]),
        "bm_get_forces": (i, [vp, _f64p, _f64p]),
        "bm_get_force_fixed": (i, [vp, C.POINTER(C.c_longlong)]),
        "bm_get_dudl": (i, [vp, _f64p]),
        "bm_get_lambda": (i, [vp, _f64p]),
        "bm_get_energy": (i, [vp, _f64p]),
        "bm_get_nan_flag": (i, [vp, C.POINTER(C.c_int)]),
        "bm_get_domdec_sizes": (i, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "bm_get_domdec": (i, [vp, _i32p, _i32p, _i32p, _i32p, C.POINTER(C.c_uint32)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def exported_symbols():
    """Names the library must export: every function declared in include/*.h."""
    import re
    names = []
    inc = os.path.join(os.path.dirname(HERE), "include")
    for fn in sorted(os.listdir(inc)):
        if fn.endswith(".h"):
            txt = open(os.path.join(inc, fn)).read()
            names += re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b((?:blade|bm)_\w+)\s*\(", txt, re.M)
    return sorted(set(names))


def ptr(a, ctype=C.c_double):
    return a.ctypes.data_as(C.POINTER(ctype))


def check(rc, what="call"):
    if rc != 0:
        msg = lib().bm_last_error().decode() or lib().blade_last_error().decode()
        raise RuntimeError("%s failed (%d): %s" % (what, rc, msg))
