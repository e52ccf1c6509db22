"""The C ABI of libmlfma.so on a machine without a GPU: it loads, exports every
entry point declared in include/mlfma.h, and reports errors through codes and
messages instead of crashing."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from conftest import gpu_available
from mlfma_pmchwt_amd import api, geometry

HEADER = Path(__file__).resolve().parents[1] / "include" / "mlfma.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(mlfma_[a-z_0-9]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = api.library()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_default_params_are_the_reference_constants():
    p = api.default_params()
    assert p.integral_equation == api.PMCHWT and p.gauss_points == 7
    assert p.cfie_alpha == 0.2 and p.leaf_coeff == 0.3          # Code/PhyConstant.h:14-15
    assert p.frequency_hz == 299792458.0 and p.eps_r_re == 1.0 and p.far_precision == api.FP64


def _create(**kw):
    lib = api.library()
    m = geometry.icosphere(1, 0.5)
    p = api.default_params(**kw)
    h = C.c_void_p()
    rc = lib.mlfma_create(C.byref(p), m.n_points, api._p(m.point_ids), api._p(m.xyz), m.n_faces,
                          api._p(m.face_ids), api._p(m.faces), C.byref(h))
    return rc, lib.mlfma_last_error(None).decode(), h


@pytest.mark.parametrize("kw,needle", [({"gauss_points": 5}, "Gauss point count: 5"),
                                       ({"integral_equation": 3}, "integral_equation"),
                                       ({"frequency_hz": -1.0}, "frequency"),
                                       ({"max_rhs": 0}, "max_rhs")])
def test_invalid_parameters_rejected(kw, needle):
    rc, msg, h = _create(**kw)
    assert rc == 1 and needle in msg and not h.value


def test_bad_nas_file(tmp_path):
    lib = api.library()
    bad = tmp_path / "bad.nas"
    bad.write_text("GRID*   1 0.0 0.0\nCTRIA3 1 1 1 2 3\n")
    p = api.default_params()
    h = C.c_void_p()
    rc = lib.mlfma_create_nas(C.byref(p), str(bad).encode(), C.byref(h))
    assert rc == 6 and "GRID*" in lib.mlfma_last_error(None).decode()


@pytest.mark.skipif(gpu_available(), reason="checks the no-device error path")
def test_no_device_is_an_error_not_a_crash():
    rc, msg, h = _create()
    assert rc == 2 and "hip" in msg.lower() and not h.value


def test_null_handles():
    lib = api.library()
    assert lib.mlfma_apply(None, 1, None, None) == 1
    assert lib.mlfma_stage(None, 0, None, None) == 1
    assert lib.mlfma_dims(None, None) == 1
