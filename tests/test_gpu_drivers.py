"""Multi-incidence drivers on the GPU, checked against the single-vector entry
points they batch (each of which is pinned against the oracle elsewhere):

* batched excitation = one mlfma_excitation per column (bit-identical);
* monostatic sweep (all directions in one batched solve) = one solve and one
  backscatter RCS per direction;
* the batch runner (reference configuration format, one operator per mesh,
  frequency and material) = building each task's operator and solving it
  directly, in both dual (bistatic) and mono mode.
"""
import numpy as np
import pytest

from mlfma_pmchwt_amd import api, geometry

pytestmark = pytest.mark.gpu

TOL = 1e-7   # solver tolerance of the comparisons: RCS agree to far below 0.01 dB


def test_excitation_batch_equals_single():
    op = api.Operator(geometry.icosphere(2, 0.5), integral_equation=api.PMCHWT, eps_r_re=4.0)
    inc = [(0.0, 0.0, 0.0), (90.0, 30.0, 45.0), (0.0, 120.0, 300.0)]
    V = op.excitation_batch([a[0] for a in inc], [a[1] for a in inc], [a[2] for a in inc])
    for j, a in enumerate(inc):
        assert np.array_equal(V[:, j], op.excitation(*a))
    op.close()


@pytest.mark.parametrize("ie", [api.CFIE, api.PMCHWT])
def test_monostatic_equals_per_direction(ie):
    op = api.Operator(geometry.icosphere(2, 0.5), integral_equation=ie, eps_r_re=4.0, eps_r_im=-0.1)
    th, ph = api.scatter_grid(0, 180, 0, 0, 45)
    for pol in (0.0, 90.0):
        rcs, st = op.monostatic(th, ph, pol=pol, tol_rel=TOL, tol_abs=0.0)
        assert all(s["converged"] for s in st)
        for a in range(th.size):
            x, s1 = op.solve(op.excitation(pol, th[a], ph[a]), restart=50, tol_rel=TOL, tol_abs=0.0)
            r1 = op.rcs(x, th[a], ph[a], pol=pol)[0]
            print(f"[mono] ie={ie} pol={pol} theta={th[a]}: batch {rcs[a, 0]:.5f} single {r1[0]:.5f} dBsm")
            assert np.max(np.abs(rcs[a] - r1)) < 1e-3
    op.close()


def test_monostatic_graphs_fp32_after_smaller_batch():
    """A monostatic sweep with recorded matvecs in FP32 mode after an earlier call sized the
    FP32 workspaces for fewer right-hand sides: the workspaces are sized before any recording,
    and the sweep equals the same sweep run eagerly."""
    mesh = geometry.icosphere(2, 0.5)
    th, ph = api.scatter_grid(0, 180, 0, 0, 30)
    out = []
    for graphs in (False, True):
        op = api.Operator(mesh, integral_equation=api.PMCHWT, eps_r_re=4.0, eps_r_im=-0.1, far_precision=api.FP32,
                          near_precision=api.FP32, max_rhs=8)
        op.set_apply_precision(api.FP32)
        op.apply(np.ones((op.n, 2), dtype=np.complex128))   # FP32 workspaces sized for 2
        op.set_graphs(graphs)
        rcs, st = op.monostatic(th, ph, pol=0.0, tol_rel=1e-6, tol_abs=0.0)
        assert all(s["converged"] for s in st)
        out.append(rcs)
        op.close()
    print(f"[mono] graphs vs eager, FP32 mode: max {np.max(np.abs(out[0] - out[1])):.2e} dB")
    assert np.max(np.abs(out[0] - out[1])) < 1e-6


def _write_meshes(tmp_path):
    paths = []
    for k, (s, r) in enumerate([(2, 0.4), (2, 0.6)]):
        p = tmp_path / f"sphere{k}.nas"
        geometry.write_nas(p, geometry.icosphere(s, r))
        paths.append(str(p))
    return paths


def test_batch_dual_equals_direct(tmp_path):
    nas = _write_meshes(tmp_path)
    cfg = f"""nas_files = {nas[0]}; {nas[1]}
frequencies = 299792458
incident_angles = 0:0, 60:30
material_list = 4.0,-0.1;2.5,0.0
sca_theta = 0:180
sca_phi = 0:0
step = 60
integral_equation = 2
algorithm = 2, 0
"""
    b = api.Batch(text=cfg)
    assert b.n_tasks == 2 * 1 * 2 * 2 * 2
    rc = b.run(api.SolveParams(api.GMRES, 2000, 50, TOL, 0.0, 0))
    assert rc == 0 or rc == 1   # MoM tasks report MLFMA_ERR_ARG
    th, ph = b.grid()
    checked = 0
    for i in range(b.n_tasks):
        t = b.task(i)
        if t["algorithm"] != 2:
            assert b.result(i, 0)[0] == 1 and b.result(i, 1)[0] == 1   # not provided here
            continue
        if i % 3:   # a sample of the tasks is re-solved directly
            continue
        op = api.Operator(t["nas"], integral_equation=api.PMCHWT, frequency_hz=t["frequency"],
                          eps_r_re=t["eps_r"].real, eps_r_im=t["eps_r"].imag)
        for pol, pdeg in ((0, 90.0), (1, 0.0)):   # "h" then "v"
            status, rcs, st = b.result(i, pol)
            assert status == 0 and st["converged"]
            x, _ = op.solve(op.excitation(pdeg, t["theta"], t["phi"]), restart=50, tol_rel=TOL, tol_abs=0.0)
            ref = op.rcs(x, th, ph, pol=pdeg)
            assert np.max(np.abs(rcs - ref)) < 1e-3
            checked += 1
        op.close()
    assert checked >= 4
    b.close()


def test_batch_mono_equals_sweep(tmp_path):
    nas = _write_meshes(tmp_path)[:1]
    b = api.Batch(text=f"""nas_files = {nas[0]}
frequencies = 299792458
incident_angles = 0:0, 30:0
sca_theta = 0:90
sca_phi = 0:0
step = 45
integral_equation = 1
mono_dual = mono
""")
    b.run(api.SolveParams(api.GMRES, 2000, 50, TOL, 0.0, 0))
    th, ph = b.grid()
    op = api.Operator(nas[0], integral_equation=api.CFIE)
    for pol, pdeg in ((0, 90.0), (1, 0.0)):
        ref, _ = op.monostatic(th, ph, pol=pdeg, tol_rel=TOL, tol_abs=0.0)
        for i in range(b.n_tasks):   # mono ignores the incidence: every task has the sweep
            status, rcs, st = b.result(i, pol)
            assert status == 0
            assert np.max(np.abs(rcs - ref)) < 1e-6
    op.close()
    b.close()


def test_batch_concurrent_objects_equal_sequential(tmp_path):
    """Objects of uneven size (four spheres, two materials: eight operators)
    processed by four host threads at once give the one-after-another results:
    each object has its own operator, stream and recordings, so the threads
    share nothing but the GPU."""
    paths = []
    for k, (s, r) in enumerate([(1, 0.3), (2, 0.4), (2, 0.6), (3, 0.5)]):
        p = tmp_path / f"obj{k}.nas"
        geometry.write_nas(p, geometry.icosphere(s, r))
        paths.append(str(p))
    cfg = f"""nas_files = {'; '.join(paths)}
frequencies = 299792458
incident_angles = 0:0, 60:30, 120:90
material_list = 4.0,-0.1;2.5,0.0
sca_theta = 0:180
sca_phi = 0:90
step = 30
integral_equation = 2
"""
    out = {}
    for conc in (1, 4):
        b = api.Batch(text=cfg)
        print(f"[batch concurrency] {conc} threads over {b.n_tasks} tasks", flush=True)
        assert b.run(api.SolveParams(api.GMRES, 2000, 50, 1e-5, 0.0, 0), concurrency=conc) == 0
        out[conc] = [[b.result(i, pol) for pol in (0, 1)] for i in range(b.n_tasks)]
        b.close()
    for i, (seq, par) in enumerate(zip(out[1], out[4])):
        for (s1, r1, t1), (s4, r4, t4) in zip(seq, par):
            assert s1 == 0 and s4 == 0
            assert t1["iterations"] == t4["iterations"]
            assert np.max(np.abs(r1 - r4)) < 1e-9, i
    with pytest.raises(api.MlfmaError):
        api.Batch(text=cfg).run(concurrency=0)
