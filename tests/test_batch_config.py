"""Batch configuration format (Code/batch_config.txt, parsed as in
Code/BatchRunner.h:103-302 and expanded as in :55-100) and the scattering grid
(Code/RCS.h:46-51, 239-246).  Host-only: the parser and the grid touch no GPU.

The expected task lists are written out by hand from the format's rules:
incident_angles replaces the incidence list, incident_theta_range appends
(theta, 0) with round((end - start) / step) + 1 points, incident_phi_range
crosses every existing theta with its phi values (or uses theta = 90 when
the list is empty), material_list items carry their own mu unless only
eps is given and mu_r was set earlier, and tasks run mesh x frequency x
incidence x algorithm x material.
"""
import numpy as np
import pytest

from mlfma_pmchwt_amd import api


def tasks(b):
    return [b.task(i) for i in range(b.n_tasks)]


def test_expansion_order_and_fields():
    b = api.Batch(text="""# comment
nas_files = a.nas; b.nas
frequencies = 1e9, 2e9
incident_angles = 90:0, 45:30
algorithm = 2, 1
material_list = 2.25,-0.0225,1.0,0.0;4.0,-0.08
integral_equation = 2
matrix_solver = 1
gauss_points = 4
mono_dual = dual
omp_threads = 14
""")
    t = tasks(b)
    assert b.n_tasks == 2 * 2 * 2 * 2 * 2
    # innermost: material, then algorithm, then incidence, then frequency, then mesh
    assert [x["eps_r"] for x in t[:2]] == [2.25 - 0.0225j, 4.0 - 0.08j]
    assert [x["algorithm"] for x in t[:4]] == [2, 2, 1, 1]
    assert [(x["theta"], x["phi"]) for x in t[:8:4]] == [(90.0, 0.0), (45.0, 30.0)]
    assert t[8]["frequency"] == 2e9 and t[16]["nas"] == "b.nas"
    assert all(x["integral_equation"] == 2 and x["solver"] == 1 and x["gauss_points"] == 4 and not x["mono"]
               for x in t)


def test_incidence_ranges():
    b = api.Batch(text="""nas_files = a.nas
frequencies = 1e9
incident_angles = 10:5
incident_theta_range = 0:20:10
""")
    assert [(x["theta"], x["phi"]) for x in tasks(b)] == [(10, 5), (0, 0), (10, 0), (20, 0)]
    b = api.Batch(text="""nas_files = a.nas
frequencies = 1e9
incident_theta_range = 0:90:45
incident_phi_range = 0:180:90
""")
    assert [(x["theta"], x["phi"]) for x in tasks(b)] == [(th, ph) for th in (0, 45, 90) for ph in (0, 90, 180)]
    b = api.Batch(text="""nas_files = a.nas
frequencies = 1e9
incident_phi_range = 0:30:15
""")
    assert [(x["theta"], x["phi"]) for x in tasks(b)] == [(90, 0), (90, 15), (90, 30)]


def test_materials_and_mu():
    b = api.Batch(text="""nas_files = a.nas
frequencies = 1e9
incident_angles = 0:0
mu_r = 2.0, -0.5
epsilon_r_list = 3.0; 5.0, -1.0
""")
    assert [(x["eps_r"], x["mu_r"]) for x in tasks(b)] == [(3.0, 2.0 - 0.5j), (5.0 - 1.0j, 2.0 - 0.5j)]
    b = api.Batch(text="""nas_files = a.nas
frequencies = 1e9
incident_angles = 0:0
mu_r = 2.0, 0.0
material_list = 4.0,-0.1;6.0,0.0,1.5,0.0
""")
    assert [(x["eps_r"], x["mu_r"]) for x in tasks(b)] == [(4.0 - 0.1j, 2.0), (6.0, 1.5)]
    b = api.Batch(text="""nas_files = a.nas
frequencies = 1e9
incident_angles = 0:0
epsilon_r = 2.5, -0.1
mu_r = 1.2
""")
    assert [(x["eps_r"], x["mu_r"]) for x in tasks(b)] == [(2.5 - 0.1j, 1.2)]


def test_defaults_and_errors():
    b = api.Batch(text="nas_files = a.nas\nfrequencies = 1e9\nincident_angles = 0:0\n")
    t = b.task(0)
    assert (t["algorithm"], t["integral_equation"], t["solver"], t["gauss_points"], t["mono"]) == (2, 1, 0, 7, False)
    assert b.n_points == 181   # theta 0..180 step 1 at phi 0
    assert api.Batch(text="nas_files = a.nas\n").n_tasks == 0   # no frequencies, no incidences
    with pytest.raises(api.MlfmaError):
        api.Batch(text="nas_files = a.nas\nfrequencies = abc\n")   # std::stod rejects it, as in the reference


@pytest.mark.parametrize("grid,expect_th,expect_ph", [
    ((0, 180, 0, 360, 30), np.arange(0, 181, 30), np.zeros(7)),
    ((90, 90, 0, 360, 90), np.full(5, 90.0), np.arange(0, 361, 90)),
    ((10, 35, 20, 20, 10), np.array([10, 20, 30]), np.full(3, 20.0)),   # floor(25 / 10) + 1 points
])
def test_scatter_grid(grid, expect_th, expect_ph):
    th, ph = api.scatter_grid(*grid)
    assert np.array_equal(th, expect_th) and np.array_equal(ph, expect_ph)
    b = api.Batch(text="nas_files = a.nas\nfrequencies = 1e9\nincident_angles = 0:0\n"
                       f"sca_theta = {grid[0]}:{grid[1]}\nsca_phi = {grid[2]}:{grid[3]}\nstep = {grid[4]}\n")
    bt, bp = b.grid()
    assert np.array_equal(bt, expect_th) and np.array_equal(bp, expect_ph)
