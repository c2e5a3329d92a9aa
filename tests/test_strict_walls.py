"""Strict-elastic wall reflections of conservative (linearized Callaway)
transport against the reference's RoughBoundarySampler built for that transport
(src/solver/RoughBoundaryInitialization.cpp:9-25, RoughBoundarySampler.cpp:77-368,
:459-533): specularity from |q.n|, mirror partners within the equal-frequency
shell, capacity-limited specular columns, residual-flux windows over the shell
and signed occupation remaps.  (Heat-capacity weights are one value per shell,
so within a window they cancel.)

The fixtures (tests/golden/rough_c2c.npz with the 1 nm walls of the film,
rough_c2c_smooth.npz with 0.05 nm walls so that most reflections are specular,
rough_finfet_cw.npz on all eleven FinFET facets that reflect; written by
make_golden.py --strict; rough_c2r.npz, the RTA film with
simulation.conservative_boundaries, written by make_golden.py --nonlinear) hold the reference's outputs on the uniforms it drew.
The product's host tables (csrc/host/boundary_tables.cpp) fed to the oracle's
sampler with the same uniforms must reproduce them bit for bit: outgoing mode,
occupation and its sign.  Planted mistakes in the tables or in the sampler's
signed-weight rule fail the same vectors."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle_ctypes as O
from phonomc_amd import runtime as rt

SMOOTH = {"boundary.boundary_values": "[302, 298, 0.0, 0.0, 0.05, 0.05]"}
CASES = [("c2c_callaway.toml", "rough_c2c.npz", {}), ("c2c_callaway.toml", "rough_c2c_smooth.npz", SMOOTH),
         ("finfet_callaway.toml", "rough_finfet_cw.npz", {}),
         # conservative RTA walls (simulation.conservative_boundaries, no N/U split)
         ("c2r_conservative.toml", "rough_c2r.npz", {})]
_problems = {}


def problem(cfg, overrides=None):
    key = (cfg, tuple(sorted((overrides or {}).items())))
    if key not in _problems:
        _problems[key] = rt.Problem(rt.config_text(cfg, **(overrides or {})), input_directory=rt.CONFIGS,
                                    mode_sampling=1)
    return _problems[key]


def table_of(p, facet):
    _, _, g, _, rough, _ = p.views()
    t = g.facets[facet].rough_table
    assert t >= 0, f"facet {facet} has no reflection table"
    assert rough[t].strict_elastic == 1
    return t


def outcomes(p, table, g, facet, k_range, signed=1):
    ins, occ, u = g[f"f{facet}_in"], g[f"f{facet}_occ"], g[f"f{facet}_u"]
    modes, occs = [], []
    for k in k_range:
        a, o, _, _ = O.rough_sample_signed(p, table, int(ins[k]), float(occ[k]), [u[k, 0], u[k, 1], 0.5, 0.5],
                                           signed)
        modes.append(a)
        occs.append(o)
    return np.array(modes), np.array(occs)


@pytest.mark.parametrize("cfg,fixture,overrides", CASES)
def test_strict_reflections_bit_exact(golden, cfg, fixture, overrides):
    p, g = problem(cfg, overrides), golden(fixture)
    assert float(g["t_bg"]) == p.views()[0].background_temperature
    facets = [int(f) for f in g["facets"]]
    _, _, geo, _, _, _ = p.views()
    # The same facets reflect through tables: everything but reservoirs and paired periodic faces.
    ours = [f for f in range(geo.n_facets) if geo.facets[f].rough_table >= 0]
    assert ours == facets
    negative = 0
    for f in facets:
        n = len(g[f"f{f}_in"])
        modes, occs = outcomes(p, table_of(p, f), g, f, range(n))
        assert np.array_equal(modes, g[f"f{f}_out"]), f"facet {f}: outgoing modes differ"
        assert np.array_equal(occs, g[f"f{f}_out_occ"]), f"facet {f}: occupations differ"
        assert np.array_equal(np.signbit(occs), np.signbit(g[f"f{f}_out_occ"]))
        negative += int(np.sum(occs < 0))
    # The vectors carry signed (negative) remapped occupations.
    assert negative > 10


def _edited_table(p, table, **arrays):
    """A copy of a RoughTable with some arrays replaced (kept alive on the copy)."""
    _, _, _, _, rough, _ = p.views()
    r = rough[table]
    copy = rt.RoughTable()
    C.pointer(copy)[0] = r
    keep = []
    for name, values in arrays.items():
        buf = np.ascontiguousarray(values)
        keep.append(buf)
        ctype = C.c_double if buf.dtype == np.float64 else C.c_int32
        setattr(copy, name, buf.ctypes.data_as(C.POINTER(ctype)))
    copy._keep = keep
    return copy


def _last_best_partners(p, table, normal):
    """Specular partners chosen with the wrong tie-break: the last outgoing mode
    of the shell window at the smallest mirror mismatch instead of the first
    (the reference keeps the first, RoughBoundarySampler.cpp:262-280)."""
    _, m, _, _, rough, _ = p.views()
    r = rough[table]
    na = m.n_active
    v = rt.arr(m.velocity, 3 * na).reshape(na, 3)
    begin, end = rt.arr(r.diffuse_begin, na, np.int32), rt.arr(r.diffuse_end, na, np.int32)
    order = rt.arr(r.outgoing_sorted, r.n_outgoing, np.int32)
    match = rt.arr(r.spec_match, na, np.int32).copy()
    n_in = -np.asarray(normal)
    ties = 0
    for a in np.nonzero((v @ n_in < 0) & (end > begin))[0]:
        vtry = v[a] - 2.0 * (v[a] @ n_in) * n_in
        cand = order[begin[a]:end[a]]
        ref = np.maximum(np.maximum(np.linalg.norm(vtry), np.linalg.norm(v[cand], axis=1)), 1e-12)
        metric = np.linalg.norm(vtry - v[cand], axis=1) / ref
        best = np.nonzero(metric == metric.min())[0]
        ties += len(best) > 1
        match[a] = cand[best[-1]]
    return match, ties


def test_planted_table_mistakes_are_caught(golden):
    """Plausible slips in the strict table builder or the sampler fail the
    reference's vectors: a shell window one mode short (diffuse channel), the
    last instead of the first best mirror partner (specular channel, exact ties
    between the degenerate transverse branches), and unsigned occupation
    weights."""
    cfg, fixture, overrides = CASES[1]
    p, g = problem(cfg, overrides), golden(fixture)
    f = int(g["facets"][0])
    t = table_of(p, f)
    _, m, geo, _, rough, _ = p.views()
    r = rough[t]
    na = m.n_active
    begin = rt.arr(r.diffuse_begin, na, np.int32)
    end = rt.arr(r.diffuse_end, na, np.int32)
    short = np.where(end - begin > 1, end - 1, end).astype(np.int32)
    last_best, ties = _last_best_partners(p, t, geo.facets[f].normal[:3])
    assert ties > 100
    k = range(2000)
    ref_modes, ref_occ = g[f"f{f}_out"][:2000], g[f"f{f}_out_occ"][:2000]
    # the smooth wall takes both channels
    u0 = g[f"f{f}_u"][:2000, 0]
    spec = rt.arr(r.specularity, na)[g[f"f{f}_in"][:2000]]
    assert np.sum(u0 <= spec) > 300 and np.sum(u0 > spec) > 300
    for edited in (_edited_table(p, t, diffuse_end=short), _edited_table(p, t, spec_match=last_best)):
        modes, _ = outcomes(p, edited, g, f, k)
        assert np.sum(modes != ref_modes) > 0
    modes, occs = outcomes(p, t, g, f, k, signed=0)
    assert np.sum((modes != ref_modes) | (occs != ref_occ)) > 0
