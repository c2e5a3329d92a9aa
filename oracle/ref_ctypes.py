"""ctypes binding of oracle/_ref/libphonomc_ref.so (see oracle/ref_harness.h).

Checker infrastructure: used by tests/golden/make_golden.py to write known-answer
vectors and by the CPU-baseline leg of bench.py.  Never imported by the product.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_LIB = os.path.join(HERE, "_ref", "libphonomc_ref.so")


class MaterialArrays(C.Structure):
    _fields_ = [("volume_a3", C.c_double), ("reciprocal", C.c_double * 9), ("qpoints", C.c_int),
                ("branches", C.c_int), ("mesh", C.c_int * 3), ("frequency_thz", C.POINTER(C.c_double)),
                ("qpoint_fractions", C.POINTER(C.c_double)), ("group_velocity", C.POINTER(C.c_double)),
                ("ntemps", C.c_int), ("temperatures", C.POINTER(C.c_double)), ("gamma", C.POINTER(C.c_double)),
                ("gamma_normal", C.POINTER(C.c_double)), ("gamma_umklapp", C.POINTER(C.c_double)),
                ("gamma_isotope", C.POINTER(C.c_double)), ("temperature_lookup_dt", C.c_double)]


def available():
    return os.path.exists(REF_LIB)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError(f"{REF_LIB} missing: build with make -f oracle/ref.mk")
        _lib = C.CDLL(REF_LIB, mode=os.RTLD_LAZY)
        d, P = C.c_double, C.c_void_p
        _lib.ref_bose_occupation.restype = d
        _lib.ref_bose_occupation.argtypes = [P, d, C.c_int, C.c_int]
        _lib.ref_relaxation_weight.restype = d
        _lib.ref_relaxation_weight.argtypes = [P, d, C.c_int, C.c_int, d]
        for name in ("ref_temperature_from_energy_density", "ref_energy_density_from_temperature",
                     "ref_crystal_energy_density"):
            getattr(_lib, name).restype = d
            getattr(_lib, name).argtypes = [P, d]
        for name in ("ref_zero_point_energy_density", "ref_energy_density_normalization"):
            getattr(_lib, name).restype = d
            getattr(_lib, name).argtypes = [P]
        _lib.ref_material_active_count.argtypes = [P]
        _lib.ref_domain_facet_count.argtypes = [P]
        _lib.ref_domain_grid_count.argtypes = [P]
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def material_arrays(view, lookup_dt=0.1):
    """The ref_material_arrays struct of a material view (pointers into the view's arrays)."""
    m = MaterialArrays()
    m.volume_a3 = view.volume_a3
    for k in range(9):
        m.reciprocal[k] = view.reciprocal[k]
    m.qpoints, m.branches = view.qpoints, view.branches
    for k in range(3):
        m.mesh[k] = view.mesh[k]
    m.frequency_thz = C.cast(view.frequency_thz, C.POINTER(C.c_double))
    m.qpoint_fractions = C.cast(view.qpoint_fractions, C.POINTER(C.c_double))
    m.group_velocity = C.cast(view.group_velocity, C.POINTER(C.c_double))
    m.ntemps = view.ntemps
    m.temperatures = C.cast(view.temperatures, C.POINTER(C.c_double))
    m.gamma = C.cast(view.gamma, C.POINTER(C.c_double))
    m.gamma_normal = C.cast(view.gamma_normal, C.POINTER(C.c_double)) if view.gamma_normal else None
    m.gamma_umklapp = C.cast(view.gamma_umklapp, C.POINTER(C.c_double)) if view.gamma_umklapp else None
    m.gamma_isotope = C.cast(view.gamma_isotope, C.POINTER(C.c_double)) if view.gamma_isotope else None
    m.temperature_lookup_dt = lookup_dt
    return m


class RefMaterial:
    def __init__(self, view, lookup_dt=0.1, include_stationary=False):
        """view: a pmcs_material_view (phonomc_amd.runtime.MaterialView) or compatible struct.
        include_stationary: the constructor flag of conservative transport (Callaway)."""
        m = material_arrays(view, lookup_dt)
        self.arrays = m
        err = C.create_string_buffer(512)
        h = C.c_void_p()
        if include_stationary:
            rc = lib().ref_material_create_stationary(C.byref(m), 1, C.byref(h), err, 512)
        else:
            rc = lib().ref_material_create(C.byref(m), C.byref(h), err, 512)
        if rc != 0:
            raise RuntimeError(err.value.decode())
        self.h = h

    def active_modes(self):
        n = lib().ref_material_active_count(self.h)
        q, b = np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32)
        lib().ref_material_active_modes(self.h, _ip(q), _ip(b))
        return q, b

    def bose(self, t, q, b):
        return lib().ref_bose_occupation(self.h, t, int(q), int(b))

    def relaxation_weight(self, t, q, b, dt):
        return lib().ref_relaxation_weight(self.h, t, int(q), int(b), dt)

    def temperature_from_energy(self, e):
        return lib().ref_temperature_from_energy_density(self.h, e)

    def energy_from_temperature(self, t):
        return lib().ref_energy_density_from_temperature(self.h, t)

    def crystal_energy(self, t):
        return lib().ref_crystal_energy_density(self.h, t)

    def zero_point(self):
        return lib().ref_zero_point_energy_density(self.h)

    def normalization(self):
        return lib().ref_energy_density_normalization(self.h)

    def close(self):
        if self.h:
            lib().ref_material_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RefDomain:
    def __init__(self, toml_path):
        err = C.create_string_buffer(512)
        h = C.c_void_p()
        if lib().ref_domain_create(toml_path.encode(), C.byref(h), err, 512) != 0:
            raise RuntimeError(err.value.decode())
        self.h = h

    def facets(self):
        n = lib().ref_domain_facet_count(self.h)
        normals, cents, shifts = np.zeros(3 * n), np.zeros(3 * n), np.zeros(3 * n)
        areas, rough_vals = np.zeros(n), np.zeros(n)
        cond = C.create_string_buffer(n + 1)
        rough, faces = np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32)
        lib().ref_domain_facets(self.h, _dp(normals), _dp(cents), _dp(areas), cond, _dp(shifts), _dp(rough_vals),
                                _ip(rough), _ip(faces))
        return dict(normals=normals.reshape(n, 3), centroids=cents.reshape(n, 3), areas=areas,
                    conditions=cond.raw[:n].decode(), shifts=shifts.reshape(n, 3), roughness=rough_vals,
                    rough=rough, faces=faces)

    def grid(self):
        n = lib().ref_domain_grid_count(self.h)
        c, v = np.zeros(3 * n), np.zeros(n)
        lib().ref_domain_grid(self.h, _dp(c), _dp(v))
        return c.reshape(n, 3), v

    def bounds(self):
        bmin, bmax, vol = np.zeros(3), np.zeros(3), C.c_double()
        lib().ref_domain_bounds(self.h, _dp(bmin), _dp(bmax), C.byref(vol))
        return bmin, bmax, vol.value

    def trace(self, pos, vel):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        n = len(pos)
        pt, t, f = np.zeros(3 * n), np.zeros(n), np.zeros(n, dtype=np.int32)
        lib().ref_domain_trace(self.h, n, _dp(pos), _dp(vel), _dp(pt), _dp(t), _ip(f))
        return pt.reshape(n, 3), t, f

    def grid_index(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        n = len(pos)
        c = np.zeros(n, dtype=np.int32)
        lib().ref_domain_grid_index(self.h, n, _dp(pos), _ip(c))
        return c

    def contains(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        out = np.zeros(len(pos), dtype=np.int32)
        lib().ref_domain_contains(self.h, len(pos), _dp(pos), _ip(out))
        return out.astype(bool)

    def nearest_facet(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        out = np.zeros(len(pos), dtype=np.int32)
        lib().ref_domain_nearest_facet(self.h, len(pos), _dp(pos), _ip(out))
        return out

    def sample_surface(self, facet, n, seed):
        pts, u = np.zeros(3 * n), np.zeros(3 * n)
        err = C.create_string_buffer(512)
        if lib().ref_domain_sample_surface(self.h, facet, n, C.c_uint64(seed), _dp(pts), _dp(u), err, 512) != 0:
            raise RuntimeError(err.value.decode())
        return pts.reshape(n, 3), u.reshape(n, 3)

    def rough_sample(self, material, facet, tref, in_active, in_occ, seeds):
        in_active = np.ascontiguousarray(in_active, dtype=np.int32)
        in_occ = np.ascontiguousarray(in_occ, dtype=np.float64)
        seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
        n = len(in_active)
        out_a, out_o, u = np.zeros(n, dtype=np.int32), np.zeros(n), np.zeros(2 * n)
        err = C.create_string_buffer(512)
        rc = lib().ref_rough_sample(self.h, material.h, facet, C.c_double(tref), n, _ip(in_active), _dp(in_occ),
                                    seeds.ctypes.data_as(C.POINTER(C.c_uint64)), _ip(out_a), _dp(out_o), _dp(u),
                                    err, 512)
        if rc != 0:
            raise RuntimeError(err.value.decode())
        return out_a, out_o, u.reshape(n, 2)

    def rough_sample_strict(self, material, facet, t_bg, in_active, in_occ, seeds):
        """Strict-elastic table of conservative transport, signed weights."""
        in_active = np.ascontiguousarray(in_active, dtype=np.int32)
        in_occ = np.ascontiguousarray(in_occ, dtype=np.float64)
        seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
        n = len(in_active)
        out_a, out_o, u = np.zeros(n, dtype=np.int32), np.zeros(n), np.zeros(2 * n)
        err = C.create_string_buffer(512)
        rc = lib().ref_rough_sample_strict(self.h, material.h, facet, C.c_double(t_bg), n, _ip(in_active),
                                           _dp(in_occ), seeds.ctypes.data_as(C.POINTER(C.c_uint64)), _ip(out_a),
                                           _dp(out_o), _dp(u), err, 512)
        if rc != 0:
            raise RuntimeError(err.value.decode())
        return out_a, out_o, u.reshape(n, 2)

    def interact(self, material, facet, tref, weight, in_active, in_vel, in_occ, seeds):
        in_active = np.ascontiguousarray(in_active, dtype=np.int32)
        in_vel = np.ascontiguousarray(in_vel, dtype=np.float64)
        in_occ = np.ascontiguousarray(in_occ, dtype=np.float64)
        seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
        n = len(in_active)
        action, out_a = np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32)
        out_v, out_o, absorbed, shift = np.zeros(3 * n), np.zeros(n), np.zeros(n), np.zeros(3 * n)
        err = C.create_string_buffer(512)
        rc = lib().ref_interact_boundary(self.h, material.h, facet, C.c_double(tref), C.c_double(weight), n,
                                         _ip(in_active), _dp(in_vel), _dp(in_occ),
                                         seeds.ctypes.data_as(C.POINTER(C.c_uint64)), _ip(action), _ip(out_a),
                                         _dp(out_v), _dp(out_o), _dp(absorbed), _dp(shift), err, 512)
        if rc != 0:
            raise RuntimeError(err.value.decode())
        return action, out_a, out_v.reshape(n, 3), out_o, absorbed, shift.reshape(n, 3)

    def reservoir(self, material, facet, density, dt, seed, n):
        init = C.c_int()
        out_a, out_t, u = np.zeros(n, dtype=np.int32), np.zeros(n), np.zeros(2 * n)
        err = C.create_string_buffer(512)
        rc = lib().ref_reservoir(self.h, material.h, facet, C.c_double(density), C.c_double(dt), C.c_uint64(seed), n,
                                 C.byref(init), _ip(out_a), _dp(out_t), _dp(u), err, 512)
        if rc != 0:
            raise RuntimeError(err.value.decode())
        return init.value, out_a, out_t, u.reshape(n, 2)

    def close(self):
        if self.h:
            lib().ref_domain_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def refresh_temperatures(material, volumes, t_bg, particle_volume, active, cell, occupation):
    volumes = np.ascontiguousarray(volumes, dtype=np.float64)
    active = np.ascontiguousarray(active, dtype=np.int32)
    cell = np.ascontiguousarray(cell, dtype=np.int32)
    occupation = np.ascontiguousarray(occupation, dtype=np.float64)
    nc = len(volumes)
    e, t = np.zeros(nc), np.zeros(nc)
    err = C.create_string_buffer(512)
    rc = lib().ref_refresh_temperatures(material.h, nc, _dp(volumes), C.c_double(t_bg), C.c_double(particle_volume),
                                        len(active), _ip(active), _ip(cell), _dp(occupation), _dp(e), _dp(t), err, 512)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return e, t


def rta_apply(material, cell_temperature, active, cell, carrier_temperature, occupation, dt, particle_volume, local,
              lifetime_temperature):
    cell_temperature = np.ascontiguousarray(cell_temperature, dtype=np.float64)
    active = np.ascontiguousarray(active, dtype=np.int32)
    cell = np.ascontiguousarray(cell, dtype=np.int32)
    carrier_temperature = np.ascontiguousarray(carrier_temperature, dtype=np.float64)
    occ = np.array(occupation, dtype=np.float64)
    res = C.c_double()
    err = C.create_string_buffer(512)
    rc = lib().ref_rta_apply(material.h, len(cell_temperature), _dp(cell_temperature), len(active), _ip(active),
                             _ip(cell), _dp(carrier_temperature), _dp(occ), C.c_double(dt),
                             C.c_double(particle_volume), int(local), C.c_double(lifetime_temperature), C.byref(res),
                             err, 512)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return occ, res.value


def heat_source_rates(domain, ncells):
    """PrescribedHeatSource::configure cell rates (eV/ps) of the domain's input."""
    rates = np.zeros(ncells)
    err = C.create_string_buffer(512)
    if lib().ref_heat_source_rates(domain.h, _dp(rates), err, 512) != 0:
        raise RuntimeError(err.value.decode())
    return rates


def source_apply(domain, material, integrated_time, cell_temperature, active, cell, occupation, particle_volume, t_bg):
    """accrue(integrated_time) + PhononSourceOperator::apply on one carrier set."""
    cell_temperature = np.ascontiguousarray(cell_temperature, dtype=np.float64)
    active = np.ascontiguousarray(active, dtype=np.int32)
    cell = np.ascontiguousarray(cell, dtype=np.int32)
    occ = np.array(occupation, dtype=np.float64)
    dep, upd = C.c_double(), C.c_longlong()
    err = C.create_string_buffer(512)
    rc = lib().ref_source_apply(domain.h, material.h, C.c_double(integrated_time), _dp(cell_temperature), len(active),
                                _ip(active), _ip(cell), _dp(occ), C.c_double(particle_volume), C.c_double(t_bg),
                                C.byref(dep), C.byref(upd), err, 512)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return occ, dep.value, upd.value


def domain_summary(toml_path, folder):
    """ref_domain_summary: the reference's summary.txt and grid_centers.csv in `folder`."""
    L = lib()
    L.ref_domain_summary.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int]
    err = C.create_string_buffer(512)
    if L.ref_domain_summary(toml_path.encode(), folder.encode(), err, 512) != 0:
        raise RuntimeError(err.value.decode())


def run(toml_path, material_view, steps, threads=0, lookup_dt=0.1, warmup=0):
    mat = RefMaterial(material_view, lookup_dt)  # keeps the arrays struct alive
    secs, ps = C.c_double(), C.c_longlong()
    err = C.create_string_buffer(1024)
    rc = lib().ref_run(toml_path.encode(), C.byref(mat.arrays), warmup, steps, threads, C.byref(secs), C.byref(ps),
                       err, 1024)
    mat.close()
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return secs.value, ps.value


def run_materials(toml_path, material_views, steps, threads=0, lookup_dt=0.1, warmup=0):
    """ref_run_materials: (seconds, particle_steps, interface events, interface transmissions,
    run summary text)."""
    arr = (MaterialArrays * len(material_views))(*[material_arrays(v, lookup_dt) for v in material_views])
    secs, ps = C.c_double(), C.c_longlong()
    iface = (C.c_longlong * 2)()
    err = C.create_string_buffer(1024)
    summary = C.create_string_buffer(1 << 16)
    rc = lib().ref_run_materials(toml_path.encode(), arr, len(material_views), warmup, steps, threads, C.byref(secs),
                                 C.byref(ps), iface, summary, 1 << 16, err, 1024)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return secs.value, ps.value, int(iface[0]), int(iface[1]), summary.value.decode()


def interface_dmm(material_views, axis, bin_thz, t_ref, weights, material, active, velocity, occupation,
                  next_material, seeds, lookup_dt=0.1):
    """ref_interface_dmm on n carriers: dict of action, material, active, occupation,
    velocity (n x 3) and the consumed uniforms (n x 3)."""
    arr = (MaterialArrays * len(material_views))(*[material_arrays(v, lookup_dt) for v in material_views])
    n = len(material)
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    mat, act, vel, occ, nxt = i32(material), i32(active), f64(velocity).reshape(-1), f64(occupation), i32(next_material)
    sd = np.ascontiguousarray(seeds, dtype=np.uint64)
    w = f64(weights)
    out = dict(action=np.zeros(n, np.int32), material=np.zeros(n, np.int32), active=np.zeros(n, np.int32),
               occupation=np.zeros(n), velocity=np.zeros(3 * n), uniforms=np.zeros(3 * n))
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
    err = C.create_string_buffer(1024)
    rc = lib().ref_interface_dmm(arr, len(material_views), int(axis), C.c_double(bin_thz), C.c_double(t_ref), _dp(w),
                                 n, ip(mat), ip(act), _dp(vel), _dp(occ), ip(nxt),
                                 sd.ctypes.data_as(C.POINTER(C.c_uint64)), ip(out["action"]), ip(out["material"]),
                                 ip(out["active"]), _dp(out["occupation"]), _dp(out["velocity"]), _dp(out["uniforms"]),
                                 err, 1024)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    out["velocity"] = out["velocity"].reshape(n, 3)
    out["uniforms"] = out["uniforms"].reshape(n, 3)
    return out


def callaway_data(material, temperature):
    """PhononMaterial::callaway_data(T) by active index."""
    n = lib().ref_material_active_count(material.h)
    out = {k: np.zeros(n) for k in ("omega", "capacity", "normal_rate", "resistive_rate")}
    q = np.zeros(3 * n)
    err = C.create_string_buffer(512)
    rc = lib().ref_callaway_data(material.h, C.c_double(temperature), _dp(out["omega"]), _dp(out["capacity"]), _dp(q),
                                 _dp(out["normal_rate"]), _dp(out["resistive_rate"]), err, 512)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    out["wavevector"] = q.reshape(n, 3)
    return out


def callaway_apply(material, t_bg, ncells, active, cell, occupation, dt, particle_volume):
    """CarrierCollisionOperator::configure(T_bg) + apply: (new occupations, residual eV)."""
    active = np.ascontiguousarray(active, dtype=np.int32)
    cell = np.ascontiguousarray(cell, dtype=np.int32)
    occ = np.array(occupation, dtype=np.float64)
    res = C.c_double()
    err = C.create_string_buffer(512)
    rc = lib().ref_callaway_apply(material.h, C.c_double(t_bg), int(ncells), len(active), _ip(active), _ip(cell),
                                  _dp(occ), C.c_double(dt), C.c_double(particle_volume), C.byref(res), err, 512)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return occ, res.value


def nl_callaway_evolve(omega, wavevector, normal_rate, resistive_rate, occupation, temperature, dt):
    """nonlinear_callaway_evolve on one cell: the new occupations, or RuntimeError
    with the reference's message."""
    n = len(omega)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (omega, np.reshape(wavevector, -1), normal_rate,
                                                                resistive_rate, occupation)]
    out = np.zeros(n)
    err = C.create_string_buffer(512)
    rc = lib().ref_nl_callaway_evolve(n, *[_dp(a) for a in arrs], C.c_double(temperature), C.c_double(dt), _dp(out),
                                      err, 512)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return out


def nl_callaway_rates(material, temperature, active):
    """PhononMaterial::mode_callaway_rates(T) of active modes: (normal, resistive) 1/ps."""
    active = np.ascontiguousarray(active, dtype=np.int32)
    rn, rr = np.zeros(len(active)), np.zeros(len(active))
    err = C.create_string_buffer(512)
    rc = lib().ref_nl_callaway_rates(material.h, C.c_double(temperature), len(active), _ip(active), _dp(rn), _dp(rr),
                                     err, 512)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return rn, rr


def nl_callaway_apply(material, t_bg, cell_temperature, local, lifetime_temperature, active, cell, occupation, dt,
                      particle_volume):
    """NonlinearCallawayCollision::configure(T_bg) + apply: (new occupations, residual eV)."""
    active = np.ascontiguousarray(active, dtype=np.int32)
    cell = np.ascontiguousarray(cell, dtype=np.int32)
    ct = np.ascontiguousarray(cell_temperature, dtype=np.float64)
    occ = np.array(occupation, dtype=np.float64)
    res = C.c_double()
    err = C.create_string_buffer(512)
    rc = lib().ref_nl_callaway_apply(material.h, C.c_double(t_bg), len(ct), _dp(ct), int(local),
                                     C.c_double(lifetime_temperature), len(active), _ip(active), _ip(cell), _dp(occ),
                                     C.c_double(dt), C.c_double(particle_volume), C.byref(res), err, 512)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return occ, res.value


def amm_rows(material_views, axis, densities, bin_thz, parallel_bin, face, side, n_source, edge_capacity=2000000):
    """AcousticInterfaceSampler rows of one interface side: (reflected [n], edge_begin [n+1],
    edge_active, edge_probability, audit[9])."""
    arrays = (MaterialArrays * len(material_views))(*[material_arrays(v) for v in material_views])
    dens = np.ascontiguousarray(densities, dtype=np.float64)
    refl = np.zeros(n_source, dtype=np.int32)
    eb = np.zeros(n_source + 1, dtype=np.int32)
    ea = np.zeros(edge_capacity, dtype=np.int32)
    ep = np.zeros(edge_capacity)
    audit = np.zeros(9)
    err = C.create_string_buffer(512)
    rc = lib().ref_amm_rows(arrays, len(material_views), int(axis), _dp(dens), C.c_double(bin_thz),
                            C.c_double(parallel_bin), int(face), int(side), _ip(refl), _ip(eb), _ip(ea), _dp(ep),
                            int(edge_capacity), _dp(audit), err, 512)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return refl, eb, ea[:eb[-1]].copy(), ep[:eb[-1]].copy(), audit


def interface_model(material_views, axis, model, amm_fraction, densities, parallel_bin, bin_thz, t_ref, weights,
                    material, active, velocity, occupation, next_material, seeds, lookup_dt=0.1):
    """ref_interface_model (model 0 DMM, 1 AMM, 2 MMM) on n carriers: dict of action,
    material, active, occupation, velocity (n x 3) and the draws made (n x 4)."""
    arr = (MaterialArrays * len(material_views))(*[material_arrays(v, lookup_dt) for v in material_views])
    n = len(material)
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    mat, act, vel, occ, nxt = i32(material), i32(active), f64(velocity).reshape(-1), f64(occupation), i32(next_material)
    sd = np.ascontiguousarray(seeds, dtype=np.uint64)
    w, dens = f64(weights), f64(densities)
    out = dict(action=np.zeros(n, np.int32), material=np.zeros(n, np.int32), active=np.zeros(n, np.int32),
               occupation=np.zeros(n), velocity=np.zeros(3 * n), uniforms=np.zeros(4 * n))
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
    err = C.create_string_buffer(1024)
    rc = lib().ref_interface_model(arr, len(material_views), int(axis), int(model), C.c_double(amm_fraction),
                                   _dp(dens), C.c_double(parallel_bin), C.c_double(bin_thz), C.c_double(t_ref), _dp(w),
                                   n, ip(mat), ip(act), _dp(vel), _dp(occ), ip(nxt),
                                   sd.ctypes.data_as(C.POINTER(C.c_uint64)), ip(out["action"]), ip(out["material"]),
                                   ip(out["active"]), _dp(out["occupation"]), _dp(out["velocity"]),
                                   _dp(out["uniforms"]), err, 1024)
    if rc != 0:
        raise RuntimeError(err.value.decode())
    out["velocity"] = out["velocity"].reshape(n, 3)
    out["uniforms"] = out["uniforms"].reshape(n, 4)
    return out
