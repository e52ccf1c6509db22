"""The oracle against the reference itself.

The reference's Boost-free units (Code/ProcessMLFMM.h, Code/GMRES.h,
Code/CGS.h, Code/GaussPoints.cpp, Code/EMSource.cpp) are compiled from
/root/reference into oracle/_ref/ref_harness (oracle/Makefile `ref`), fed the
oracle's tree and tables, and compared with the oracle's own results.  Where
the harness has not been built (no reference sources on this machine), the
committed fixtures in tests/golden/ carry the reference's outputs instead.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import pyoracle
from mlfma_pmchwt_amd import geometry

GOLD = Path(__file__).resolve().parent / "golden"
needs_harness = pytest.mark.skipif(not pyoracle.HARNESS.exists(), reason="reference harness not built here")


@pytest.fixture(scope="module")
def nas_s2(tmp_path_factory):
    p = tmp_path_factory.mktemp("nas") / "s2.nas"
    geometry.write_nas(p, geometry.icosphere(2, 0.5))
    return p


@pytest.fixture(scope="module")
def nas_s2_ml(tmp_path_factory):
    p = tmp_path_factory.mktemp("nas") / "s2ml.nas"
    geometry.write_nas(p, geometry.icosphere(2, 1.2))
    return p


@needs_harness
def test_gauss_tables_bit_exact():
    assert pyoracle.harness("gauss")["mismatches"] == 0


@needs_harness
def test_emsource_bit_exact():
    assert pyoracle.harness("emsource")["mismatches"] == 0


@needs_harness
@pytest.mark.parametrize("ie,eps", [(0, (1.0, 0.0)), (1, (1.0, 0.0)), (2, (4.0, 0.0)), (2, (2.5, -0.1))])
def test_stages_match_reference(nas_s2_ml, ie, eps):
    r = pyoracle.harness("stages", nas_s2_ml, ie, eps[0], eps[1], 11)
    assert r["compared"] > 0
    assert r["max_rel"] <= 1e-15


@needs_harness
@pytest.mark.parametrize("ie", [0, 1])
def test_solvers_match_reference(nas_s2, ie):
    r = pyoracle.harness("solvers", nas_s2, ie, 1.0, 0.0)
    assert r["gmres_ref_conv"] == r["gmres_orc_conv"] == 1
    assert r["cgs_ref_conv"] == r["cgs_orc_conv"] == 1
    assert r["gmres_x_rel"] <= 1e-12 and r["cgs_x_rel"] <= 1e-12


def test_golden_reference_fixtures():
    """Fixtures written by tests/golden/make_golden.py from the reference harness."""
    f = GOLD / "reference_checks.json"
    if not f.exists():
        pytest.skip("fixture not generated")
    gold = json.loads(f.read_text())
    assert gold["gauss"]["mismatches"] == 0
    assert gold["emsource"]["mismatches"] == 0
    for s in gold["stages"]:
        assert s["max_rel"] <= 1e-15
    for s in gold["solvers"]:
        assert s["gmres_x_rel"] <= 1e-12 and s["cgs_x_rel"] <= 1e-12


def test_golden_solution_vectors():
    """Reference GMRES (Code/GMRES.h) solution on the oracle operator, stored as a fixture."""
    f = GOLD / "gmres_efie_s2.npz"
    if not f.exists():
        pytest.skip("fixture not generated")
    g = np.load(f)
    mesh = geometry.icosphere(2, 0.5)
    orc = pyoracle.Oracle(mesh, ie=0)
    b = orc.rhs(0.0, 0.0, 0.0)
    assert np.max(np.abs(b - g["rhs"])) <= 1e-12 * np.max(np.abs(b))
    x, st = orc.gmres(b, restart=int(g["restart"]), tol_rel=float(g["tol"]), tol_abs=1e-12)
    assert np.linalg.norm(x - g["x_ref"]) <= 1e-12 * np.linalg.norm(g["x_ref"])


def _rwg_meshes():
    import sys
    sys.path.insert(0, str(GOLD))
    from make_golden import rwg_meshes
    return rwg_meshes()


@needs_harness
def test_rwg_numbering_matches_reference(tmp_path):
    """Code/RWGGenerator.cpp compiled from the reference (oracle/Makefile `ref`) against the
    oracle's build_rwg on the same NAS files: count and every field of every function."""
    for name, m in _rwg_meshes().items():
        p = tmp_path / f"{name}.nas"
        geometry.write_nas(p, m)
        r = pyoracle.harness("rwg", p)
        assert r["ref_count"] == r["orc_count"] > 0, name
        assert r["mismatches"] == 0, name


def test_golden_rwg_numbering(tmp_path):
    """The reference's RWG numbering (tests/golden/rwg_reference.npz, written by the compiled
    Code/RWGGenerator.cpp): the oracle reproduces faces, edge and free vertices and edge
    lengths bit for bit, also with shuffled non-contiguous ids and shuffled face order."""
    gold = np.load(GOLD / "rwg_reference.npz")
    for name, m in _rwg_meshes().items():
        p = tmp_path / f"{name}.nas"
        geometry.write_nas(p, m)
        o = pyoracle.Oracle(p, ie=0, mode=0)
        ints, ln, _ = o.rwg()
        assert np.array_equal(ints, gold[f"{name}_ints"]), name
        assert np.array_equal(ln, gold[f"{name}_len"]), name
