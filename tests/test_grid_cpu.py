"""The grid front end (phd_amd/grid.py): the reference's grid.dat reader and
metric (module_tools.f90:223-442, module_deriv.f90:80-201).  The stencils
are fourth order at the ends and sixth order inside, so they differentiate
polynomials of degree <= 4 exactly: the metric of such a grid is known in
closed form.  The reference's own grid fixture (testcases/sod/grid.dat, a
uniform 400 x 11 x 1 grid) reads back with the reference's layout."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_metric_exact_on_polynomial_grids():
    from phd_amd.grid import metric
    i = np.arange(1, 41, dtype=np.float64)
    for x, dxdi in ((0.01 * i, 0.01 + 0 * i),
                    (1e-3 * i ** 2, 2e-3 * i),
                    (1e-4 * i ** 3 + 1e-2 * i, 3e-4 * i ** 2 + 1e-2),
                    (1e-6 * i ** 4 + 1e-2 * i, 4e-6 * i ** 3 + 1e-2)):
        m = metric(x)
        assert m.size == x.size + 6
        assert np.all(m[:3] == 1.0) and np.all(m[-3:] == 1.0)
        assert np.max(np.abs(m[3:-3] * dxdi - 1.0)) <= 1e-12


def test_metric_orders():
    """On x = i^6 the sixth-order interior stencil is still exact while the
    fourth-order closures are not: only the six end points carry error."""
    from phd_amd.grid import metric
    i = np.arange(1, 31, dtype=np.float64)
    m = metric(1e-9 * i ** 6)
    exact = 6e-9 * i ** 5
    err = np.abs(m[3:-3] * exact - 1.0)
    assert np.max(err[3:-3]) <= 1e-12
    assert np.min(err[[0, 1, 2, -3, -2, -1]]) > 1e-6


def test_reference_grid_fixture():
    from phd_amd import chemdata as cd
    from phd_amd.grid import box_from_grid, read_grid
    path = os.path.join(HERE, "golden", "sod_grid.dat")
    xt, yt, zt = read_grid(path, 400, 11, 1)
    assert xt.size == 400 and yt.size == 11 and zt.size == 1
    assert np.all(np.diff(xt) > 0)
    adi = cd.Adim()
    box = box_from_grid(path, 400, 11, 1, 1, adi)
    dx_i, dy_i, dz_i = box.inv_spacing(adi)
    h = xt[1] - xt[0]
    assert np.max(np.abs(dx_i[3:-3] * h / adi.L_ref - 1.0)) <= 1e-10
    assert np.all(dx_i[:3] == adi.L_ref) and np.all(dy_i == adi.L_ref)
    with pytest.raises(ValueError):
        read_grid(path, 400, 12, 1)   # one y point too many: Z-DIRECTION not found


def test_uniform_grid_matches_box(tmp_path):
    """A grid.dat of the uniform periodic box gives, in the interior, the
    metric the uniform Box computes (dx_i = L_ref / dx) to rounding."""
    from phd_amd import chemdata as cd
    from phd_amd.grid import box_from_grid
    from phd_amd.ics import Box
    n, L = 24, 0.01
    x = np.arange(n) * (L / n)
    p = tmp_path / "grid.dat"
    rec = ["X-DIRECTION"] + [repr(float(v)) for v in x] + ["Y-DIRECTION"] + [repr(float(v)) for v in x] + \
          ["Z-DIRECTION", "0.0", "END"]
    p.write_text("\n".join(rec) + "\n")
    adi = cd.Adim()
    g = box_from_grid(str(p), n, n, 1, 2, adi).inv_spacing(adi)
    u = Box(n, n, 1, 3, 2, (L, L, L)).inv_spacing(adi)
    for a, b in zip(g[:2], u[:2]):
        assert np.max(np.abs(a[3:-3] / b[3:-3] - 1.0)) <= 1e-12
