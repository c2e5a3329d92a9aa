"""ctypes bindings of include/pmc_capi.h and include/pmc_solver.h.

Loading fails loudly when the built libraries are missing: there is no CPU
fallback for the step.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(PKG, "_build")
CONFIGS = os.path.join(PKG, "configs")

PMC_OK = 0
KERNEL_CLASSES = ["begin", "transport", "events", "inject", "refresh", "sort", "moments", "newton", "relax", "source",
                  "output", "step", "callaway"]
STATUS = {0: "ok", 1: "invalid argument", 2: "device error", 3: "excess transport events", 4: "boundary failure",
          5: "escaped carrier", 6: "invalid state", 7: "capacity exhausted", 8: "latched", 9: "unsupported"}


class Setup(C.Structure):
    _fields_ = [("capacity", C.c_int64), ("time_step_ps", C.c_double), ("push_distance_a", C.c_double),
                ("seed", C.c_uint64), ("instance", C.c_uint32), ("background_temperature", C.c_double),
                ("lifetime_local", C.c_int32), ("lifetime_temperature", C.c_double), ("refill_policy", C.c_int32),
                ("mode_sampling", C.c_int32), ("temperature_inversion", C.c_int32),
                ("conservative_walls", C.c_int32), ("reserved", C.c_int32 * 7)]


class ModeTable(C.Structure):
    _fields_ = [("n_active", C.c_int32), ("n_flat", C.c_int32), ("active_flat", C.POINTER(C.c_int32)),
                ("omega", C.POINTER(C.c_double)), ("velocity", C.POINTER(C.c_double)),
                ("neq_background", C.POINTER(C.c_double)), ("n_temps", C.c_int32), ("temps", C.POINTER(C.c_double)),
                ("gamma_total", C.POINTER(C.c_double)), ("energy_weight", C.c_double),
                ("background_energy_density", C.c_double), ("zero_point_density", C.c_double),
                ("lut_n", C.c_int32), ("lut_t", C.POINTER(C.c_double)), ("lut_e", C.POINTER(C.c_double)),
                ("energy_normalization", C.c_double)]


class Facet(C.Structure):
    _fields_ = [("normal", C.c_double * 3), ("shift", C.c_double * 3), ("condition", C.c_int32),
                ("rough_table", C.c_int32), ("reservoir", C.c_int32), ("periodic_pair", C.c_int32)]


class BvhNode(C.Structure):
    _fields_ = [("bmin", C.c_float * 3), ("bmax", C.c_float * 3), ("index", C.c_int32), ("count", C.c_int32)]


class Mesh(C.Structure):
    _fields_ = [("n_vertices", C.c_int32), ("vertices", C.POINTER(C.c_double)), ("n_faces", C.c_int32),
                ("faces", C.POINTER(C.c_int32)), ("face_normal", C.POINTER(C.c_double)),
                ("face_k", C.POINTER(C.c_double)), ("face_bounds", C.POINTER(C.c_double)),
                ("face_facet", C.POINTER(C.c_int32)), ("n_face_nodes", C.c_int32),
                ("face_nodes", C.POINTER(BvhNode)), ("face_order", C.POINTER(C.c_int32)),
                ("n_simplices", C.c_int32), ("simplex_vertices", C.POINTER(C.c_double)),
                ("simplex_inverse", C.POINTER(C.c_double)), ("simplex_bounds", C.POINTER(C.c_double)),
                ("simplex_volume", C.POINTER(C.c_double)), ("n_simplex_nodes", C.c_int32),
                ("simplex_nodes", C.POINTER(BvhNode)), ("simplex_order", C.POINTER(C.c_int32)),
                ("field_dims", C.c_int32 * 3), ("field_h", C.c_double), ("field_unit", C.c_double),
                ("field", C.POINTER(C.c_uint8)), ("field_inside", C.POINTER(C.c_uint8))]


class Geometry(C.Structure):
    _fields_ = [("kind", C.c_int32), ("bmin", C.c_double * 3), ("bmax", C.c_double * 3), ("n_facets", C.c_int32),
                ("facets", C.POINTER(Facet)), ("box_plane_facet", C.c_int32 * 6), ("grid", C.c_int32 * 3),
                ("inv_cell", C.c_double * 3), ("voxel_to_cell", C.POINTER(C.c_int32)), ("n_cells", C.c_int32),
                ("cell_volume", C.POINTER(C.c_double)), ("mesh", C.POINTER(Mesh))]


class RoughTable(C.Structure):
    _fields_ = [("specularity", C.POINTER(C.c_double)), ("spec_match", C.POINTER(C.c_int32)),
                ("diffuse_begin", C.POINTER(C.c_int32)), ("diffuse_end", C.POINTER(C.c_int32)),
                ("n_outgoing", C.c_int32), ("outgoing_active", C.POINTER(C.c_int32)),
                ("outgoing_sorted", C.POINTER(C.c_int32)), ("residual_prefix", C.POINTER(C.c_double)),
                ("n_roulette", C.c_int32), ("roulette_active", C.POINTER(C.c_int32)),
                ("roulette_cdf", C.POINTER(C.c_double)), ("strict_elastic", C.c_int32),
                ("mode_begin", C.c_int32), ("mode_count", C.c_int32)]


class Reservoir(C.Structure):
    _fields_ = [("facet", C.c_int32), ("temperature", C.c_double), ("normal", C.c_double * 3),
                ("expected_count", C.c_double), ("initial_count", C.c_int32), ("cdf", C.POINTER(C.c_double)),
                ("alias_prob", C.POINTER(C.c_double)), ("alias_index", C.POINTER(C.c_int32)),
                ("occupation", C.POINTER(C.c_double)), ("n_faces", C.c_int32), ("face_cdf", C.POINTER(C.c_double)),
                ("triangles", C.POINTER(C.c_double))]


class HeatSource(C.Structure):
    _fields_ = [("n_cells", C.c_int32), ("rate", C.POINTER(C.c_double)), ("mode_eligible", C.POINTER(C.c_int32)),
                ("square", C.c_int32), ("start_ps", C.c_double), ("end_ps", C.c_double), ("period_ps", C.c_double),
                ("on_duration_ps", C.c_double), ("duty_cycle", C.c_double), ("amplitude", C.c_double),
                ("spectrum", C.c_int32), ("mode_weight", C.POINTER(C.c_double)),
                ("frequency_center_thz", C.c_double), ("frequency_sigma_thz", C.c_double)]


class CallawayTable(C.Structure):
    _fields_ = [("capacity", C.POINTER(C.c_double)), ("wavevector", C.POINTER(C.c_double)),
                ("normal_rate", C.POINTER(C.c_double)), ("resistive_rate", C.POINTER(C.c_double)),
                ("heat_capacity_density", C.c_double), ("background_energy_density", C.c_double)]


class CallawayRates(C.Structure):
    _fields_ = [("n_temps", C.c_int32), ("temps", C.POINTER(C.c_double)), ("gamma_normal", C.POINTER(C.c_double)),
                ("gamma_umklapp", C.POINTER(C.c_double)), ("gamma_extra", C.POINTER(C.c_double)),
                ("wavevector", C.POINTER(C.c_double))]


class Layers(C.Structure):
    _fields_ = [("count", C.c_int32), ("axis", C.c_int32), ("planes", C.POINTER(C.c_double)),
                ("mode_begin", C.POINTER(C.c_int32)), ("cell_material", C.POINTER(C.c_int32)),
                ("background_energy_density", C.POINTER(C.c_double)), ("lut_begin", C.POINTER(C.c_int32)),
                ("spatial_volume", C.POINTER(C.c_double)), ("band_width", C.c_double),
                ("bank_begin", C.POINTER(C.c_int32)), ("bank_mode", C.POINTER(C.c_int32)),
                ("bank_omega", C.POINTER(C.c_double)), ("bank_prefix", C.POINTER(C.c_double)),
                ("amm_fraction", C.c_double), ("amm_block", C.POINTER(C.c_int32)),
                ("amm_reflected", C.POINTER(C.c_int32)), ("amm_edge_begin", C.POINTER(C.c_int32)),
                ("amm_edge_mode", C.POINTER(C.c_int32)), ("amm_edge_prob", C.POINTER(C.c_double)),
                ("facet_rough", C.POINTER(C.c_int32)), ("heat_capacity_density", C.POINTER(C.c_double))]


class LaunchConfig(C.Structure):
    _fields_ = [("event_split", C.c_int32), ("callaway_threads", C.c_int32)]


class Tuning(C.Structure):
    _fields_ = [("series_on", C.c_int32), ("series_radius", C.c_double), ("reserved", C.c_int32 * 6)]


class Ledger(C.Structure):
    _fields_ = [("step", C.c_int64), ("live_carriers", C.c_int64), ("slots", C.c_int64),
                ("step_absorbed", C.c_int64), ("step_injected", C.c_int64),
                ("step_absorbed_energy_ev", C.c_double), ("step_injected_energy_ev", C.c_double),
                ("step_lifetime_residual_ev", C.c_double), ("step_balance_residual_ev", C.c_double),
                ("total_absorbed", C.c_int64), ("total_injected", C.c_int64),
                ("total_absorbed_energy_ev", C.c_double), ("total_injected_energy_ev", C.c_double),
                ("total_lifetime_residual_ev", C.c_double), ("total_balance_residual_ev", C.c_double),
                ("thermal_energy_ev", C.c_double), ("carrier_steps", C.c_int64), ("rough_events", C.c_int64),
                ("rough_specular", C.c_int64), ("rough_diffuse", C.c_int64), ("newton_passes", C.c_int64),
                ("newton_evals", C.c_int64), ("step_source_energy_ev", C.c_double),
                ("total_source_energy_ev", C.c_double), ("source_updates", C.c_int64),
                ("source_exact_passes", C.c_int64), ("interface_events", C.c_int64),
                ("interface_transmitted", C.c_int64), ("step_source_updates", C.c_int64),
                ("source_time_ps", C.c_double), ("source_pending_ev", C.c_double)]


class Error(C.Structure):
    _fields_ = [("code", C.c_int32), ("carrier", C.c_int64), ("position", C.c_float * 3), ("facet", C.c_int32),
                ("message", C.c_char * 256)]


class Init(C.Structure):
    _fields_ = [("count", C.c_int64), ("temperature_mode", C.c_int32), ("uniform_temperature", C.c_double),
                ("cold_point", C.c_double * 3), ("hot_point", C.c_double * 3), ("cold_temperature", C.c_double),
                ("hot_temperature", C.c_double), ("cell_temperature", C.POINTER(C.c_double)),
                ("excite", C.c_int32), ("excite_fmin_thz", C.c_double), ("excite_fmax_thz", C.c_double),
                ("excite_multiplier", C.c_double)]


class MaterialView(C.Structure):
    _fields_ = [("volume_a3", C.c_double), ("reciprocal", C.c_double * 9), ("qpoints", C.c_int32),
                ("branches", C.c_int32), ("mesh", C.c_int32 * 3), ("frequency_thz", C.POINTER(C.c_double)),
                ("qpoint_fractions", C.POINTER(C.c_double)), ("group_velocity", C.POINTER(C.c_double)),
                ("ntemps", C.c_int32), ("temperatures", C.POINTER(C.c_double)), ("gamma", C.POINTER(C.c_double)),
                ("gamma_normal", C.POINTER(C.c_double)), ("gamma_umklapp", C.POINTER(C.c_double)),
                ("gamma_isotope", C.POINTER(C.c_double)), ("temperature_lookup_dt", C.c_double)]


_lib = None


def lib():
    """Loads libpmc_host.so (and through it libpmc_device.so); raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    host = os.path.join(BUILD, "libpmc_host.so")
    device = os.path.join(BUILD, "libpmc_device.so")
    for path in (device, host):
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing; build with `python -m phonomc_amd.build` first")
    C.CDLL(device, mode=C.RTLD_GLOBAL)
    L = C.CDLL(host, mode=C.RTLD_GLOBAL)
    P, V = C.c_void_p, C.c_void_p
    sig = {
        "pmcs_problem_from_toml": (C.c_int, [C.c_char_p, C.c_int32, C.POINTER(P), C.c_char_p, C.c_int32]),
        "pmcs_problem_from_text": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int32, C.POINTER(P), C.c_char_p, C.c_int32]),
        "pmcs_problem_destroy": (None, [P]),
        "pmcs_problem_views": (C.c_int, [P, C.POINTER(Setup), C.POINTER(ModeTable), C.POINTER(Geometry),
                                         C.POINTER(Init), C.POINTER(C.c_int32), C.POINTER(C.POINTER(RoughTable)),
                                         C.POINTER(C.c_int32), C.POINTER(C.POINTER(Reservoir))]),
        "pmcs_problem_material": (C.c_int, [P, C.POINTER(MaterialView)]),
        "pmcs_problem_scalars": (C.c_int, [P, C.POINTER(C.c_double)]),
        "pmcs_problem_bose": (C.c_double, [P, C.c_double, C.c_int32]),
        "pmcs_problem_relaxation_weight": (C.c_double, [P, C.c_double, C.c_int32, C.c_double]),
        "pmcs_problem_temperature_from_energy": (C.c_double, [P, C.c_double]),
        "pmcs_problem_energy_from_temperature": (C.c_double, [P, C.c_double]),
        "pmcs_problem_facets": (C.c_int, [P, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                          C.POINTER(C.c_double), C.c_char_p, C.POINTER(C.c_double),
                                          C.POINTER(C.c_double), C.POINTER(C.c_int32)]),
        "pmcs_problem_cells": (C.c_int, [P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "pmcs_problem_write_summary": (C.c_int, [P, C.c_char_p, C.c_char_p, C.c_int32]),
        "pmcs_solver_create": (C.c_int, [P, C.c_char_p, C.POINTER(P), C.c_char_p, C.c_int32]),
        "pmcs_solver_run": (C.c_int, [P, C.c_int32, C.c_char_p, C.c_int32]),
        "pmcs_solver_timestep": (C.c_int32, [P]),
        "pmcs_solver_context": (P, [P]),
        "pmcs_solver_problem": (P, [P]),
        "pmcs_solver_observables": (C.c_int, [P, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                              C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "pmcs_solver_append_summary": (C.c_int, [P, C.c_char_p]),
        "pmcs_solver_destroy": (None, [P]),
        "pmcs_synthetic_material": (C.c_int, [C.c_int32, C.c_int32, C.POINTER(MaterialView), C.POINTER(V)]),
        "pmcs_material_free": (None, [V]),
        "pmc_create": (C.c_int, [C.POINTER(Setup), C.POINTER(P)]),
        "pmc_upload_material": (C.c_int, [P, C.POINTER(ModeTable)]),
        "pmc_upload_geometry": (C.c_int, [P, C.POINTER(Geometry)]),
        "pmc_upload_boundary_tables": (C.c_int, [P, C.c_int32, C.POINTER(RoughTable), C.c_int32,
                                                 C.POINTER(Reservoir)]),
        "pmc_initialize_carriers": (C.c_int, [P, C.POINTER(Init)]),
        "pmc_seed_carriers": (C.c_int, [P, C.c_int64, C.POINTER(C.c_float), C.POINTER(C.c_int32),
                                        C.POINTER(C.c_double)]),
        "pmc_step": (C.c_int, [P, C.c_int32]),
        "pmc_read_cells": (C.c_int, [P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "pmc_read_heat_flux": (C.c_int, [P, C.POINTER(C.c_double)]),
        "pmc_read_source_pending": (C.c_int, [P, C.POINTER(C.c_double)]),
        "pmc_read_ledger": (C.c_int, [P, C.POINTER(Ledger)]),
        "pmc_read_carriers": (C.c_int, [P, C.POINTER(C.c_int64), C.POINTER(C.c_float), C.POINTER(C.c_int32),
                                        C.POINTER(C.c_double), C.POINTER(C.c_int32)]),
        "pmc_read_uids": (C.c_int, [P, C.c_int64, C.POINTER(C.c_uint32)]),
        "pmc_last_error": (C.c_int, [P, C.POINTER(Error)]),
        "pmc_kernel_times": (C.c_int, [P, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.c_int32]),
        "pmc_reset_kernel_times": (C.c_int, [P]),
        "pmc_set_timing": (C.c_int, [P, C.c_int32]),
        "pmc_set_tuning": (C.c_int, [P, C.POINTER(Tuning)]),
        "pmc_upload_heat_source": (C.c_int, [P, C.POINTER(HeatSource)]),
        "pmcs_problem_heat_source": (C.c_int, [P, C.POINTER(HeatSource)]),
        "pmcs_problem_callaway": (C.c_int, [P, C.POINTER(CallawayTable)]),
        "pmcs_write_finfet_stl": (C.c_int, [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                            C.POINTER(C.c_double), C.c_char_p, C.c_int]),
        "pmc_upload_callaway": (C.c_int, [P, C.POINTER(CallawayTable)]),
        "pmc_upload_callaway_nonlinear": (C.c_int, [P, C.POINTER(CallawayRates)]),
        "pmcs_problem_callaway_nonlinear": (C.c_int, [P, C.POINTER(CallawayRates)]),
        "pmcs_problem_amm_audit": (C.c_int, [P, C.c_int32, C.POINTER(C.c_double)]),
        "pmc_get_tuning": (C.c_int, [P, C.POINTER(Tuning)]),
        "pmc_destroy": (None, [P]),
        "pmc_status_string": (C.c_char_p, [C.c_int32]),
        "pmcs_problem_layers": (C.c_int, [P, C.POINTER(Layers)]),
        "pmcs_problem_layer_material": (C.c_int, [P, C.c_int32, C.POINTER(MaterialView)]),
        "pmc_upload_layers": (C.c_int, [P, C.POINTER(Layers)]),
        "pmc_get_launch_config": (C.c_int, [P, C.POINTER(LaunchConfig)]),
        "pmc_set_launch_config": (C.c_int, [P, C.POINTER(LaunchConfig)]),
        "pmc_autotune": (C.c_int, [P, C.c_int32, C.POINTER(LaunchConfig), C.POINTER(C.c_double), C.c_int32]),
        "pmc_batch_create": (C.c_int, [C.POINTER(P), C.c_int32, C.POINTER(P)]),
        "pmc_batch_step": (C.c_int, [P, C.c_int32]),
        "pmc_batch_failed": (C.c_int, [P]),
        "pmc_batch_set_graph": (C.c_int, [P, C.c_int32]),
        "pmc_batch_destroy": (None, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def arr(ptr, n, dtype=np.float64):
    """Copies n elements behind a ctypes pointer into a numpy array."""
    if not ptr or n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class Problem:
    """Host tables of one TOML input (pmcs_problem)."""

    def __init__(self, toml_text=None, path=None, input_directory=".", mode_sampling=0, handle=None, owner=None):
        L = lib()
        err = C.create_string_buffer(1024)
        h = C.c_void_p()
        if handle is not None:
            self.h, self._owner = handle, owner
            return
        if path is not None:
            rc = L.pmcs_problem_from_toml(path.encode(), mode_sampling, C.byref(h), err, 1024)
        else:
            rc = L.pmcs_problem_from_text(toml_text.encode(), input_directory.encode(), mode_sampling, C.byref(h),
                                          err, 1024)
        if rc != 0:
            raise RuntimeError(err.value.decode())
        self.h, self._owner = h, None

    def close(self):
        if self.h and self._owner is None:
            lib().pmcs_problem_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def views(self):
        s, m, g, i = Setup(), ModeTable(), Geometry(), Init()
        nr, ns = C.c_int32(), C.c_int32()
        r, v = C.POINTER(RoughTable)(), C.POINTER(Reservoir)()
        lib().pmcs_problem_views(self.h, C.byref(s), C.byref(m), C.byref(g), C.byref(i), C.byref(nr), C.byref(r),
                                 C.byref(ns), C.byref(v))
        return s, m, g, i, [r[k] for k in range(nr.value)], [v[k] for k in range(ns.value)]

    def heat_source(self):
        """The heat-source view, or None when the input has no heat source."""
        hs = HeatSource()
        return hs if lib().pmcs_problem_heat_source(self.h, C.byref(hs)) == 1 else None

    def callaway(self):
        """The linearized Callaway view, or None under RTA."""
        t = CallawayTable()
        return t if lib().pmcs_problem_callaway(self.h, C.byref(t)) == 1 else None

    def amm_audit(self, face=0):
        """The AMM audit of interface `face` (9 values, see pmc_solver.h), or None without AMM."""
        out = (C.c_double * 9)()
        return np.array(out[:]) if lib().pmcs_problem_amm_audit(self.h, face, out) == 1 else None

    def callaway_nonlinear(self):
        """The nonlinear Callaway rate tables, or None unless the input selects them."""
        t = CallawayRates()
        return t if lib().pmcs_problem_callaway_nonlinear(self.h, C.byref(t)) == 1 else None

    def material(self):
        mv = MaterialView()
        lib().pmcs_problem_material(self.h, C.byref(mv))
        return mv

    def layers(self):
        """The layered-materials view, or None for one material."""
        lv = Layers()
        return lv if lib().pmcs_problem_layers(self.h, C.byref(lv)) == 1 else None

    def layer_material(self, k):
        mv = MaterialView()
        if lib().pmcs_problem_layer_material(self.h, k, C.byref(mv)) != 0:
            raise IndexError(k)
        return mv

    def scalars(self):
        out = (C.c_double * 10)()
        lib().pmcs_problem_scalars(self.h, out)
        keys = ["particle_volume", "particle_density", "energy_weight", "background_energy_density",
                "zero_point_density", "energy_normalization", "volume", "cells", "carriers", "push"]
        return dict(zip(keys, list(out)))

    def facets(self):
        n = C.c_int32()
        lib().pmcs_problem_facets(self.h, C.byref(n), None, None, None, None, None, None, None)
        k = n.value
        normals, cents, shifts = np.zeros(3 * k), np.zeros(3 * k), np.zeros(3 * k)
        areas, values = np.zeros(k), np.zeros(k)
        cond = C.create_string_buffer(k + 1)
        faces = np.zeros(k, dtype=np.int32)
        lib().pmcs_problem_facets(self.h, C.byref(n), _ptr(normals, C.c_double), _ptr(cents, C.c_double),
                                  _ptr(areas, C.c_double), cond, _ptr(shifts, C.c_double), _ptr(values, C.c_double),
                                  _ptr(faces, C.c_int32))
        return dict(normals=normals.reshape(k, 3), centroids=cents.reshape(k, 3), areas=areas,
                    conditions=cond.raw[:k].decode(), shifts=shifts.reshape(k, 3), values=values, faces=faces)

    def write_summary(self, folder):
        """summary.txt and grid_centers.csv of the problem in `folder` (host only)."""
        err = C.create_string_buffer(512)
        if lib().pmcs_problem_write_summary(self.h, folder.encode(), err, 512) != 0:
            raise RuntimeError(err.value.decode())

    def cells(self):
        n = int(self.scalars()["cells"])
        centers, volumes = np.zeros(3 * n), np.zeros(n)
        lib().pmcs_problem_cells(self.h, _ptr(centers, C.c_double), _ptr(volumes, C.c_double))
        return centers.reshape(n, 3), volumes


class Context:
    """Device context (pmc_ctx) built from a Problem's views."""

    def __init__(self, problem, seed=None, mode_sampling=None, capacity=None, initialize=True, initial_field=None,
                 inversion=None):
        """initial_field: per-cell initial temperature [n_cells] (K), the warm start
        (pmc_init.temperature_mode = 2); otherwise the input's own initial temperature.
        inversion: 1 for the per-cell Newton inversion of the exact crystal energy
        instead of the lookup table (pmc_setup.temperature_inversion)."""
        L = lib()
        s, m, g, i, rough, res = problem.views()
        if inversion is not None:
            s.temperature_inversion = inversion
        if seed is not None:
            s.seed = seed
        if mode_sampling is not None:
            s.mode_sampling = mode_sampling
        if capacity is not None:
            s.capacity = capacity
        self.problem = problem
        self.init = i
        h = C.c_void_p()
        self._check(L.pmc_create(C.byref(s), C.byref(h)), None)
        self.h = h
        self._check(L.pmc_upload_material(h, C.byref(m)))
        self._check(L.pmc_upload_geometry(h, C.byref(g)))
        rough_arr = (RoughTable * max(1, len(rough)))(*rough)
        res_arr = (Reservoir * max(1, len(res)))(*res)
        self._check(L.pmc_upload_boundary_tables(h, len(rough), rough_arr, len(res), res_arr))
        cw = problem.callaway()
        if cw is not None:
            self._check(L.pmc_upload_callaway(h, C.byref(cw)))
        nl = problem.callaway_nonlinear()
        if nl is not None:
            self._check(L.pmc_upload_callaway_nonlinear(h, C.byref(nl)))
        lv = problem.layers()
        if lv is not None:
            self._check(L.pmc_upload_layers(h, C.byref(lv)))
        hs = problem.heat_source()
        if hs is not None:
            self._check(L.pmc_upload_heat_source(h, C.byref(hs)))
        self.n_cells = g.n_cells
        if initial_field is not None:
            self._field = np.ascontiguousarray(initial_field, dtype=np.float64)
            if self._field.shape != (g.n_cells,):
                raise ValueError(f"initial_field must hold {g.n_cells} cell temperatures")
            i.temperature_mode = 2
            i.cell_temperature = _ptr(self._field, C.c_double)
        if initialize:
            self._check(L.pmc_initialize_carriers(h, C.byref(i)))

    def _check(self, rc, h="self"):
        if rc != PMC_OK:
            msg = STATUS.get(rc, str(rc))
            if h is not None and getattr(self, "h", None):
                e = Error()
                lib().pmc_last_error(self.h, C.byref(e))
                msg = e.message.decode() or msg
            raise RuntimeError(f"pmc error {rc}: {msg}")

    def seed_carriers(self, positions, modes, occupations):
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1)
        md = np.ascontiguousarray(modes, dtype=np.int32)
        oc = np.ascontiguousarray(occupations, dtype=np.float64)
        self._check(lib().pmc_seed_carriers(self.h, len(md), _ptr(pos, C.c_float), _ptr(md, C.c_int32),
                                            _ptr(oc, C.c_double)))

    def step(self, n=1):
        self._check(lib().pmc_step(self.h, n))

    def launch_config(self):
        c = LaunchConfig()
        self._check(lib().pmc_get_launch_config(self.h, C.byref(c)))
        return {"event_split": c.event_split, "callaway_threads": c.callaway_threads}

    def set_launch_config(self, event_split, callaway_threads=1024):
        c = LaunchConfig(event_split, callaway_threads)
        self._check(lib().pmc_set_launch_config(self.h, C.byref(c)))

    def autotune(self, steps=4):
        """pmc_autotune: times `steps` real steps per candidate and keeps the fastest.
        Returns the chosen configuration and the mean kernel time (ms) per candidate."""
        c = LaunchConfig()
        rep = (C.c_double * 7)()
        self._check(lib().pmc_autotune(self.h, steps, C.byref(c), rep, 7))
        names = ["event_split=1", "event_split=2", "event_split=4", "event_split=8",
                 "callaway_threads=256", "callaway_threads=384", "callaway_threads=1024"]
        return ({"event_split": c.event_split, "callaway_threads": c.callaway_threads},
                {k: v for k, v in zip(names, list(rep)) if v > 0})

    def cells(self):
        t, e = np.zeros(self.n_cells), np.zeros(self.n_cells)
        self._check(lib().pmc_read_cells(self.h, _ptr(t, C.c_double), _ptr(e, C.c_double)))
        return t, e

    def heat_flux(self):
        q = np.zeros(3 * self.n_cells)
        self._check(lib().pmc_read_heat_flux(self.h, _ptr(q, C.c_double)))
        return q.reshape(-1, 3)

    def source_pending(self):
        """Heat-source energy pending per cell (eV)."""
        out = np.zeros(self.n_cells)
        self._check(lib().pmc_read_source_pending(self.h, _ptr(out, C.c_double)))
        return out

    def ledger(self):
        led = Ledger()
        self._check(lib().pmc_read_ledger(self.h, C.byref(led)))
        return {k: getattr(led, k) for k, _ in Ledger._fields_}

    def carriers(self):
        n = C.c_int64()
        self._check(lib().pmc_read_carriers(self.h, C.byref(n), None, None, None, None))
        k = n.value
        pos = np.zeros(3 * k, dtype=np.float32)
        mode = np.zeros(k, dtype=np.int32)
        occ = np.zeros(k)
        cell = np.zeros(k, dtype=np.int32)
        uid = np.zeros(k, dtype=np.uint32)
        self._check(lib().pmc_read_carriers(self.h, C.byref(n), _ptr(pos, C.c_float), _ptr(mode, C.c_int32),
                                            _ptr(occ, C.c_double), _ptr(cell, C.c_int32)))
        self._check(lib().pmc_read_uids(self.h, k, _ptr(uid, C.c_uint32)))
        return dict(position=pos.reshape(k, 3), mode=mode, occupation=occ, cell=cell, uid=uid)

    def set_timing(self, on=True):
        self._check(lib().pmc_set_timing(self.h, 1 if on else 0))

    def tuning(self):
        t = Tuning()
        self._check(lib().pmc_get_tuning(self.h, C.byref(t)))
        return {"series_on": t.series_on, "series_radius": t.series_radius}

    def set_tuning(self, series_on=None, series_radius=None):
        """Pseudo-temperature evaluation: per-cell series (default) or exact carrier sums only."""
        t = Tuning()
        self._check(lib().pmc_get_tuning(self.h, C.byref(t)))
        if series_on is not None:
            t.series_on = 1 if series_on else 0
        if series_radius is not None:
            t.series_radius = series_radius
        self._check(lib().pmc_set_tuning(self.h, C.byref(t)))

    def kernel_times(self):
        """Per-class (total ms, launches) since the last reset."""
        ms, n = (C.c_double * len(KERNEL_CLASSES))(), (C.c_int64 * len(KERNEL_CLASSES))()
        self._check(lib().pmc_kernel_times(self.h, ms, n, len(KERNEL_CLASSES)))
        return {name: (ms[k], n[k]) for k, name in enumerate(KERNEL_CLASSES)}

    def reset_kernel_times(self):
        self._check(lib().pmc_reset_kernel_times(self.h))

    def last_error(self):
        e = Error()
        lib().pmc_last_error(self.h, C.byref(e))
        return e.code, e.message.decode()

    def close(self):
        if getattr(self, "h", None):
            lib().pmc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Independent contexts stepped together (pmc_batch): one launch per kernel of
    the step over all instances, replayed as a captured hipGraph.  Each context
    stays readable (cells, ledger, carriers) and must be stepped only through the
    batch while it exists."""

    def __init__(self, contexts, graph=True):
        self.contexts = list(contexts)
        handles = (C.c_void_p * len(self.contexts))(*[c.h for c in self.contexts])
        h = C.c_void_p()
        rc = lib().pmc_batch_create(handles, len(self.contexts), C.byref(h))
        if rc != PMC_OK:
            for c in self.contexts:
                e = Error()
                lib().pmc_last_error(c.h, C.byref(e))
                if e.message:
                    raise RuntimeError(f"pmc error {rc}: {e.message.decode()}")
            raise RuntimeError(f"pmc error {rc}: {STATUS.get(rc, rc)}")
        self.h = h
        lib().pmc_batch_set_graph(h, 1 if graph else 0)

    def step(self, n=1):
        rc = lib().pmc_batch_step(self.h, n)
        if rc != PMC_OK:
            i = lib().pmc_batch_failed(self.h)
            c = self.contexts[max(i, 0)]
            e = Error()
            lib().pmc_last_error(c.h, C.byref(e))
            raise RuntimeError(f"pmc error {rc} in instance {i}: {e.message.decode() or STATUS.get(rc, rc)}")

    def close(self):
        if getattr(self, "h", None):
            lib().pmc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Solver:
    """MonteCarloSolver facade (pmcs_solver): the time-step loop with the
    reference's observables and output files.  Takes ownership of the problem."""

    def __init__(self, problem, output_folder=None):
        err = C.create_string_buffer(1024)
        h = C.c_void_p()
        ph, problem.h = problem.h, None
        rc = lib().pmcs_solver_create(ph, output_folder.encode() if output_folder else None, C.byref(h), err, 1024)
        if rc != 0:
            raise RuntimeError(err.value.decode())
        self.h = h
        self.problem = Problem(handle=lib().pmcs_solver_problem(h), owner=self)
        self.n_cells = int(self.problem.scalars()["cells"])

    def run(self, steps):
        err = C.create_string_buffer(1024)
        if lib().pmcs_solver_run(self.h, steps, err, 1024) != 0:
            raise RuntimeError(err.value.decode())

    def timestep(self):
        return lib().pmcs_solver_timestep(self.h)

    def observables(self):
        ke, ki, q = C.c_double(), C.c_double(), C.c_double()
        t = np.zeros(self.n_cells)
        lib().pmcs_solver_observables(self.h, C.byref(ke), C.byref(ki), C.byref(q), _ptr(t, C.c_double))
        return {"kappa_eff": ke.value, "kappa_int": ki.value, "mean_heat_flux": q.value, "temperature": t}

    def append_summary(self, path):
        if lib().pmcs_solver_append_summary(self.h, path.encode()) != 0:
            raise RuntimeError("summary write failed")

    def close(self):
        if getattr(self, "h", None):
            lib().pmcs_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_finfet_stl(path, substrate=(200.0, 100.0, 50.0), fin=(10.0, 40.0, 100.0), step=(5.0, 10.0, 5.0)):
    """Procedural FinFET surface as a binary STL (nm); returns the triangle count."""
    arrs = [(C.c_double * 3)(*v) for v in (substrate, fin, step)]
    err = C.create_string_buffer(512)
    n = lib().pmcs_write_finfet_stl(path.encode(), arrs[0], arrs[1], arrs[2], err, 512)
    if n < 0:
        raise RuntimeError(err.value.decode())
    return n


def config_text(name, **overrides):
    """Reads a bundled workload TOML and applies key = value overrides (section.key form);
    keys the file lacks are added to their section."""
    path = os.path.join(CONFIGS, name)
    lines = open(path).read().splitlines()
    section = None
    out = []
    done = set()
    depth = 0  # bracket depth of a multi-line array value
    for line in lines:
        s = line.strip()
        if depth == 0 and s.startswith("[") and s.endswith("]") and "=" not in s:
            section = s[1:-1]
            out.append(line)
            continue
        if depth == 0 and "=" in s and section:
            key = s.split("=", 1)[0].strip()
            full = f"{section}.{key}"
            depth = s.count("[") - s.count("]")
            if full in overrides:
                if depth != 0:
                    raise KeyError(f"override of multi-line value {full} is not supported")
                out.append(f"{key} = {overrides[full]}")
                done.add(full)
                continue
        elif depth:
            depth += s.count("[") - s.count("]")
        out.append(line)
    # Keys the file does not set are added to their section (appended as a new
    # section when the file has none of that name).
    missing = sorted(set(overrides) - done)
    text = "\n".join(out) + "\n"
    for full in missing:
        sec, key = full.split(".", 1)
        header = f"[{sec}]"
        if header in text.splitlines():
            idx = text.splitlines().index(header)
            rows = text.splitlines()
            rows.insert(idx + 1, f"{key} = {overrides[full]}")
            text = "\n".join(rows) + "\n"
        else:
            text += f"\n{header}\n{key} = {overrides[full]}\n"
    return text
