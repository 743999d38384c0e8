"""Test infrastructure only (the oracle): a numpy restatement of MISO's
uniform-table interpolation (interpolate_uniform_table1d/2d and the
interpolator constructors, miso/include/miso/table_interpolator.hpp:11-20,
27-36, 75-101, 114-125), evaluated in the table's own precision with the
reference's operation order.  Used only by tests/ to check the device
kernels; nothing in the product imports it."""
import numpy as np


def _clamp(v, lo, hi):
    # util::clamp (utility.hpp:68-75): v < lo ? lo : (v > hi ? hi : v)
    return np.where(v < lo, lo, np.where(v > hi, hi, v))


def _index(x, x_min, dxi, n, dt):
    f = (x - dt(x_min)) * dt(dxi)
    f = _clamp(f, dt(0), dt(n - 1)).astype(dt)
    i = _clamp(f.astype(np.int64), 0, n - 2)   # static_cast<int> truncates toward zero
    t = (f - i.astype(dt)).astype(dt)
    return i, t


def table1d(table, x_min, x_max, x):
    """UniformTableInterpolator1D(table, x_min, x_max).interpolate(x)."""
    table = np.asarray(table)
    dt = table.dtype.type
    n = table.shape[0]
    dx = dt((dt(x_max) - dt(x_min)) / dt(n - 1))
    dxi = dt(dt(1) / dx)
    i, t = _index(np.asarray(x, dt), x_min, dxi, n, dt)
    return ((dt(1) - t) * table[i] + t * table[i + 1]).astype(dt)


def table2d(table, x0_min, x0_max, x1_min, x1_max, x0, x1):
    """UniformTableInterpolator2D(table, ...).interpolate(x0, x1); table[i0][i1]."""
    table = np.asarray(table)
    dt = table.dtype.type
    n0, n1 = table.shape
    d0 = dt((dt(x0_max) - dt(x0_min)) / dt(n0 - 1))
    d1 = dt((dt(x1_max) - dt(x1_min)) / dt(n1 - 1))
    i, t = _index(np.asarray(x0, dt), x0_min, dt(dt(1) / d0), n0, dt)
    j, u = _index(np.asarray(x1, dt), x1_min, dt(dt(1) / d1), n1, dt)
    ti, ui = (dt(1) - t).astype(dt), (dt(1) - u).astype(dt)
    flat = table.ravel()
    ij00 = i * n1 + j
    return (ti * ui * flat[ij00] + ti * u * flat[ij00 + 1] + t * ui * flat[ij00 + n1]
            + t * u * flat[ij00 + n1 + 1]).astype(dt)
