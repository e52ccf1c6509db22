"""The reference-facing entry points on the GPU.

* A NAS file read by the product (mlfma_create_nas) and by the oracle: RWG numbering equal
  to the compiled reference's RWGGenerator (tests/golden/rwg_reference.npz) bit for bit, also
  with shuffled non-contiguous ids and face order; every tree level identical; matvec within
  1e-10 of the oracle.
* The C++ drop-in (include/mlfma_dropin.hpp) driven by a C++ program the way
  MGMRES::mgmres drives the reference's ax (Code/GMRES.h:50-66): its one-shot y = A * 1 dump in
  the reference's proposed format equals the oracle's A * 1 within 1e-10; its host GMRES,
  calling applyAx once per Arnoldi step, reaches a true residual of 1e-6 and agrees with the
  device GMRES of MLFMM::matrix_solver; applyAx(0, null, null) throws std::invalid_argument.
"""
import json
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from conftest import relerr
from mlfma_pmchwt_amd import api, geometry
import pyoracle

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
GOLD = Path(__file__).resolve().parent / "golden"
EXE = ROOT / "mlfma_pmchwt_amd" / "_build" / "dropin_gmres"


def _rwg_meshes():
    sys.path.insert(0, str(GOLD))
    from make_golden import rwg_meshes
    return rwg_meshes()


@pytest.mark.parametrize("name", ["shuffled_ids", "cube", "ellipsoid"])
def test_nas_through_product_and_oracle(tmp_path, name):
    m = _rwg_meshes()[name]
    p = tmp_path / f"{name}.nas"
    geometry.write_nas(p, m)
    print(f"[nas] {name}: {m.n_faces} faces", flush=True)
    op = api.Operator(str(p), integral_equation=api.PMCHWT, eps_r_re=3.0, eps_r_im=-0.2)
    orc = pyoracle.Oracle(p, ie=2, eps=(3.0, -0.2))
    gold = np.load(GOLD / "rwg_reference.npz")
    ints, ln, _ = op.rwg()
    assert np.array_equal(ints, gold[f"{name}_ints"]) and np.array_equal(ln, gold[f"{name}_len"])
    assert op.max_level == orc.max_level
    for lvl in range(orc.max_level + 1):
        a, b = op.level(lvl), orc.level(lvl)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), lvl
    x = np.random.default_rng(3).standard_normal(op.n) + 0.5j
    assert relerr(op.apply(x), orc.apply(x)) < 1e-10
    op.close()


def test_cpp_dropin_caller(tmp_path):
    assert EXE.exists(), "C++ caller not built (python -m mlfma_pmchwt_amd.build)"
    mesh = geometry.icosphere(2, 0.8)
    p = tmp_path / "sphere.nas"
    geometry.write_nas(p, mesh)
    dump = tmp_path / "ax_ones.txt"
    print("[dropin] PMCHWT eps 4-0.1j icosphere 2 r=0.8", flush=True)
    r = subprocess.run([str(EXE), str(p), "2", "4.0", "-0.1", str(dump)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    res = json.loads(r.stdout.strip().splitlines()[-1])
    print(f"[dropin] {res}")
    assert res["invalid_argument_thrown"]
    assert res["host_gmres_converged"] and res["host_true_residual"] <= 1.001e-6
    assert res["device_converged"] and res["device_true_residual"] <= 1.001e-6
    assert res["solution_rel_diff"] < 1e-5

    lines = dump.read_text().splitlines()
    head = [ln for ln in lines if ln.startswith("#")]
    assert head == ["# MatVec one-shot result", "# algorithm = MLFMM", "# selectAlgorithm = 2",
                    "# selectIntegralEqu = 2", f"# size = {2 * mesh.n_edges_interior}",
                    "# x is fixed to complex<double>(1.0, 0.0)", "# columns: index x_real x_imag y_real y_imag abs_y"]
    rows = np.array([[float(t) for t in ln.split("\t")] for ln in lines if not ln.startswith("#")])
    assert np.array_equal(rows[:, 0], np.arange(rows.shape[0]))
    assert np.all(rows[:, 1] == 1.0) and np.all(rows[:, 2] == 0.0)
    y = rows[:, 3] + 1j * rows[:, 4]
    assert np.allclose(rows[:, 5], np.abs(y), rtol=1e-15, atol=0)
    orc = pyoracle.Oracle(p, ie=2, eps=(4.0, -0.1))
    assert relerr(y, orc.apply(np.ones(orc.n, dtype=np.complex128))) < 1e-10
