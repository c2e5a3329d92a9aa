"""pmc_import_material: the reference's phono3py material folders (POSCAR and
kappa-fbz.hdf5) converted to the raw material table that the GPU hosts read.
Runs where the importer is built (HDF5 C++ library present) and the reference's
material folders exist; skipped elsewhere.

The reference's own reader (src/material/MaterialDataReader.cpp) needs the HDF5
library, so it is not part of the oracle build: the import is checked against
what the files fix independently -- the POSCAR lattice (reciprocal vectors
2 pi A^-1, volume |det A|), the 31^3 Gamma-centred full mesh, unit shapes and
finite values -- and then used end to end (a problem builds from it, its
tables pass the solver's own validation and the Callaway first-BZ checks)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from phonomc_amd import runtime as rt

TOOL = os.path.join(rt.BUILD, "pmc_import_material")
MATERIALS = "/root/reference/material"
pytestmark = pytest.mark.skipif(not (os.path.exists(TOOL) and os.path.isdir(MATERIALS)),
                                reason="importer not built or reference materials absent")


def poscar(path):
    lines = [l.strip() for l in open(path) if l.strip()]
    scale = float(lines[1])
    return np.array([[float(x) for x in lines[2 + i].split()[:3]] for i in range(3)]) * scale


@pytest.fixture(scope="module")
def imported():
    out = tempfile.mkdtemp()
    tables = {}
    for name in ("Si", "SiC"):
        dst = os.path.join(out, name.lower())
        os.makedirs(dst)
        r = subprocess.run([TOOL, os.path.join(MATERIALS, name), os.path.join(dst, "material.pmcmat")],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        tables[name] = dst
    return tables


@pytest.mark.parametrize("name", ["Si", "SiC"])
def test_import_matches_the_files(imported, name):
    text = rt.config_text("c1_cross_plane.toml", **{"io.material_folder": f'"{imported[name]}"',
                                                    "simulation.particle_count": 2000})
    problem = rt.Problem(text, input_directory=rt.CONFIGS)
    mv = problem.material()
    a = poscar(os.path.join(MATERIALS, name, "POSCAR"))
    assert mv.volume_a3 == pytest.approx(abs(np.linalg.det(a)), rel=1e-12)
    recip = np.array(mv.reciprocal[:]).reshape(3, 3)
    np.testing.assert_allclose(recip, 2 * np.pi * np.linalg.inv(a), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(a @ recip, 2 * np.pi * np.eye(3), atol=1e-9)  # A B = 2 pi I
    assert tuple(mv.mesh[:]) == (31, 31, 31) and mv.qpoints == 31 ** 3 and mv.branches == 6
    q = rt.arr(mv.qpoint_fractions, 3 * mv.qpoints).reshape(-1, 3)
    idx = np.rint((q - np.floor(q)) * 31).astype(int) % 31
    distinct = len(np.unique(idx[:, 0] * 961 + idx[:, 1] * 31 + idx[:, 2]))
    # every q point within 2e-6 mesh units of the mesh; SiC (N/U data) holds each
    # mesh point once, Si lists six zone-boundary points twice (as two images)
    assert np.abs((q - np.floor(q)) * 31 - np.rint((q - np.floor(q)) * 31)).max() <= 2e-6 * 31
    assert distinct == (31 ** 3 if name == "SiC" else 31 ** 3 - 6)
    f = rt.arr(mv.frequency_thz, mv.qpoints * mv.branches)
    assert np.all(np.isfinite(f)) and f.max() > 10.0
    _, m, _, _, _, _ = problem.views()
    assert m.n_active > 0.99 * mv.qpoints * mv.branches


def test_imported_nu_split_runs_callaway(imported):
    """SiC carries gamma_N / gamma_U (and an isotope linewidth): a Callaway input
    builds on it, through the completed-mesh and first-BZ checks."""
    text = rt.config_text("c1c_callaway.toml", **{"io.material_folder": f'"{imported["SiC"]}"',
                                                  "simulation.particle_count": 2000})
    problem = rt.Problem(text, input_directory=rt.CONFIGS)
    cw = problem.callaway()
    assert cw is not None
    _, m, _, _, _, _ = problem.views()
    gn = rt.arr(cw.normal_rate, m.n_active)
    gr = rt.arr(cw.resistive_rate, m.n_active)
    assert np.all(gn >= 0) and np.all(gr > 0) and gn.max() > 0
