"""The C ABI library loads and exports every entry point include/rhg.h
declares; host-only entry points behave.  No GPU calls."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "rhg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rhg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(R):
    names = declared_functions()
    assert len(names) >= 25
    lib = ctypes.CDLL(R.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version(R):
    assert "gfx950" in R.version()


def test_estimators_host_only(R, O):
    """rhg_widom_estimators (src/statistics.c:2539,4747-4800) matches the
    oracle's restatement on arbitrary accumulators."""
    rng = np.random.default_rng(1)
    b = R.WidomBlocks()
    arr = np.zeros((5, 4))
    for k in range(5):
        arr[k] = [1000 + k, rng.uniform(1, 2), rng.uniform(-5, 5), 1000.0]
        b.count[k], b.sum_w[k], b.sum_wdu[k], b.ideal_gas[k] = arr[k]
    g = R.widom_estimators(b, 298.0, 593.395728552254)
    o = O.widom_estimators(arr, 298.0, 593.395728552254)
    for key in ("kh_blocks", "u_blocks"):
        assert np.array_equal(g[key], o[key])
    for key in ("kh", "kh_err", "u", "u_err", "kh_overall", "w_overall"):
        assert g[key] == o[key], key


def test_estimator_rejects_bad_input(R):
    b = R.WidomBlocks()
    out = np.zeros(16)
    rc = R.lib.rhg_widom_estimators(ctypes.byref(b), -1.0, 1.0, 1.0,
                                    out.ctypes.data_as(ctypes.c_void_p))
    assert rc == -1


def test_synthetic_positions_min_separation(W):
    box = W.cell_box(20.0, 22.0, 24.0, 85.0, 95.0, 80.0)
    xyz = W.synthetic_positions(3, box, 300, 2.2)
    m = np.asarray(box).reshape(3, 3)
    inv = np.linalg.inv(m)
    s = xyz @ inv
    d = s[:, None, :] - s[None, :, :]
    d -= np.round(d)
    r = np.linalg.norm(d @ m, axis=-1)
    np.fill_diagonal(r, 99.0)
    assert r.min() >= 2.2 - 1e-9
    assert np.all((s >= -1e-12) & (s < 1.0 + 1e-12))


def test_workloads_shapes(W):
    c1 = W.c1_framework()
    assert c1.n == 4608
    assert abs(c1.charge.sum()) < 1e-9
    assert np.allclose(W.perpendicular_widths(c1.box), 51.6)
    c2 = W.c2_framework()
    assert c2.n == 1200 and c2.bc == 2
    assert W.perpendicular_widths(c2.box).min() >= 24.0
    for i in (0, 7, 255):
        job = W.c3_job(i)
        assert 100 <= job.framework.n <= 8000
        assert W.perpendicular_widths(job.framework.box).min() >= 24.0 - 1e-9
        assert 1e4 <= job.n_trials <= 1e6
