"""The C-ABI library loads, exports every symbol include/*.h declares, and its host-only logic
(grids, argument validation with the reference's error texts) works without a GPU."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "hrweno_capi.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(hrw_[a-z0-9_]+)\s*\(", text))
    return sorted(n for n in names if not n.endswith("_fn"))


def test_library_exports_every_declared_symbol(hh):
    from hrweno_hip import _lib
    lib = _lib.lib()
    syms = declared_symbols()
    assert len(syms) >= 40
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (hrw_\w+)", out))
    missing = [s for s in syms if s not in exported]
    assert not missing, missing
    for s in syms:
        getattr(lib, s)
    assert set(_lib.SIGNATURES) == set(syms)


def test_version(hh):
    assert "gfx950" in hh.version()


def test_grid_linear_matches_oracle(hh, oracle):
    g = hh.Grid1().linear(-1.0, 1.0, 1000, name="x")
    assert np.array_equal(g.edges, oracle.grid_linear(-1.0, 1.0, 1000))
    c, w = oracle.grid_compute(g.edges)
    assert np.array_equal(g.center, c) and np.array_equal(g.width, w)
    assert np.array_equal(g.left, g.edges[:-1]) and np.array_equal(g.right, g.edges[1:])
    assert g.scale == "linear" and g.name == "x" and g.ncells == 1000


def test_grid_kinds_match_oracle(hh, oracle):
    assert np.array_equal(hh.Grid1().bilinear(0.0, 10.0, 1e3, [124, 365]).edges,
                          oracle.grid_bilinear(0.0, 10.0, 1e3, 124, 365))
    assert np.array_equal(hh.Grid1().geometric(10.0, 1e3, 1.1, 100).edges,
                          oracle.grid_geometric(10.0, 1e3, 1.1, 100))
    assert np.array_equal(hh.Grid1().log(0.1, 1e3, 10 ** 4).edges,
                          oracle.grid_log(0.1, 1e3, 10 ** 4))


def test_reference_grid_tests(hh):
    """test/test_grid.f90:31-162 on the library's grids."""
    rtol = 1e-5
    for g, xmin, xmax, n in [
        (hh.Grid1().linear(1.0, 1e3, 10 ** 6, name="F [N]"), 1.0, 1e3, 10 ** 6),
        (hh.Grid1().log(1e-1, 1e3, 10 ** 4, name="T [K]"), 1e-1, 1e3, 10 ** 4),
        (hh.Grid1().geometric(1e1, 1e3, 1.1, 10 ** 2, name="P [W]"), 1e1, 1e3, 10 ** 2),
    ]:
        assert g.ncells == n
        assert abs(g.edges[0] - xmin) <= rtol * abs(xmin)
        assert abs(g.edges[n] - xmax) <= rtol * abs(xmax)
        assert np.array_equal(g.width, g.right - g.left)
        assert np.array_equal(g.center, (g.left + g.right) / 2)
    g = hh.Grid1().bilinear(0.0, 1e1, 1e3, [124, 365], name="W [J]")
    assert g.ncells == 489 and g.edges[0] == 0.0
    assert abs(g.edges[124] - 1e1) <= rtol * 1e1 and abs(g.edges[489] - 1e3) <= rtol * 1e3


@pytest.mark.parametrize("call,msg", [
    (lambda hh: hh.Grid1().linear(1.0, 1.0, 4), "Invalid input 'xmin', 'xmax'. Valid range: xmax > xmin."),
    (lambda hh: hh.Grid1().linear(0.0, 1.0, 0), "Invalid input 'ncells'. Valid range: ncells > 1."),
    (lambda hh: hh.Grid1().log(0.0, 1.0, 4), "Invalid input 'xmin'. Valid range: xmin > 0."),
    (lambda hh: hh.Grid1().geometric(0.0, 1.0, 0.0, 4), "Invalid input 'ratio'. Valid range: ratio > 0"),
    (lambda hh: hh.Grid1().bilinear(0.0, 0.0, 1.0, [2, 2]),
     "Invalid input 'xmin', 'xcross'. Valid range: xcross > xmin."),
])
def test_grid_errors_use_reference_texts(hh, call, msg):
    with pytest.raises(hh.HrwError) as ei:
        call(hh)
    assert str(ei.value) == msg and ei.value.code == 1


def test_weno_tables_match_oracle(hh, oracle):
    for k in (1, 2, 3):
        d, c = hh.tables(k)
        od, oc = oracle.tables(k)
        assert np.array_equal(d, od) and np.array_equal(c.T, oc)
    with pytest.raises(hh.HrwError, match="Invalid input 'k'. Valid range: 1 <= k <= 3."):
        hh.tables(4)


def test_nonuniform_problem_argument_errors(hh):
    """The non-uniform-grid problem checks its arguments before touching the device."""
    xe = np.linspace(-1.0, 1.0, 11)
    with pytest.raises(hh.HrwError, match="needs HRW_BC_ZERO_GRADIENT"):
        hh.Burgers1D(10, bc=hh.BC_PERIODIC, xedges=xe)
    with pytest.raises(hh.HrwError, match=r"size\(xedges\) /= ncells \+ 1"):
        hh.Burgers1D(12, bc=hh.BC_ZERO_GRADIENT, xedges=xe)


def test_batch_line_far_from_origin_rejected(hh):
    """The batched stage kernel takes 1/width from 1/rx: a line whose edges lie more than 1e8
    cell widths from the origin is refused before anything reaches the device."""
    with pytest.raises(hh.HrwError, match="1e8 cell widths"):
        hh.BurgersBatch([100, 100], [3, 3], [1e-3, 1e-3], xmin=[-1.0, 1e9], xmax=[1.0, 1e9 + 1.0])


def test_batch_form_and_config_need_a_batch(hh):
    """The batch form setters check their handle before anything else."""
    from hrweno_hip import _lib
    lib = _lib.lib()
    r, f = C.c_int(-1), C.c_int(-1)
    assert lib.hrw_batch_set_form(None, 1) != 0
    assert b"null batch" in lib.hrw_last_error()
    assert lib.hrw_batch_config(None, C.byref(r), C.byref(f)) != 0
    assert r.value == -1 and f.value == -1
