#!/usr/bin/env python3
"""Write tests/golden/calc_golden.npz: golden vectors of the reference's
per-dial math, computed by the reference's own header-only Calculate*.h
(src/Utils/include) compiled in this container into oracle/_ref
(`make -C oracle ref`).  The fixture holds only inputs and outputs.

Also tests/golden/calc2d_golden.npz: the 2-D Bilinear/Bicubic math of
CalculateBilinearInterpolation.h / CalculateBicubicSpline.h on seeded grids.

Cases: the knot arrays and x sweeps of the reference's unit tests
(tests/GTests/unitTest_Calculate{Monotonic,Compact,Graph}*.cpp), plus seeded
random records of every 1-D type with x inside the knots, exactly on knots,
just below the first knot (the (xx<0) index branch) and far outside on both
sides (extrapolation), with wide and tight output clamps.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import pyoracle  # noqa: E402

MONO, UNIF, GEN, COMP, GRAPH = 1, 2, 3, 4, 5


def unit_test_records():
    """Knot arrays of the reference's unit tests (data, dim, x sweep, lo, hi)."""
    out = []
    xs = np.arange(-1.5, 1.5 + 1e-9, 0.01)
    for t in (MONO, COMP):
        out.append((t, [0.0, 1.0, 0.0, 1.0], 2, np.arange(-1.0, 2.0 + 1e-9, 0.1), -10.0, 10.0))
        out.append((t, [-1.0, 2.0 / 2, 0.0, 1.0, 0.0], 3, xs, -10.0, 10.0))
        out.append((t, [-1.0, 2.0 / 5, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0], 6, xs, -10.0, 10.0))
        out.append((t, [-1.0, 2.0 / 12, 0.5, 1.5, 1.0, 1.0, 1.0, 1.0, 0.5, 1.0, 1.0,
                        1.0, 1.0, 1.5, 0.5], 13, np.arange(-1.1, 1.1 + 1e-9, 0.01),
                    -10.0, 10.0))
        n = 17
        d = [-1.0, 2.0 / (n - 1)] + [float(np.cos(4.0 * (-1.0 + p * 2.0 / (n - 1))))
                                     for p in range(n)]
        out.append((t, d, n, xs, -10.0, 10.0))
    out.append((GRAPH, [0.0, 0.0, 1.0, 1.0], 4, np.arange(-2.0, 2.0 + 1e-9, 0.1), -10.0, 10.0))
    out.append((GRAPH, [1.0, -1.0, 0.0, 0.0, 1.0, 1.0], 6, np.arange(-3.0, 3.0 + 1e-9, 0.1), -10.0, 10.0))
    out.append((GRAPH, [2.0, 0.5], 2, np.array([-1.0, 0.5, 3.0]), -10.0, 10.0))  # 1 point
    return out


def catmull_rom(x, y):
    s = np.empty_like(y)
    s[0] = (y[1] - y[0]) / (x[1] - x[0])
    s[-1] = (y[-1] - y[-2]) / (x[-1] - x[-2])
    s[1:-1] = (y[2:] - y[:-2]) / (x[2:] - x[:-2])
    return s


def random_records(rng):
    out = []
    for rep in range(40):
        for t in (MONO, COMP, UNIF, GEN, GRAPH):
            n = int(rng.integers(2 if t in (MONO, COMP, GRAPH) else 3, 16))
            if t == GEN:
                n = int(rng.integers(2, 33))
            lo_x, hi_x = -3.0, 3.0
            if t in (MONO, COMP, UNIF):
                x = lo_x + np.arange(n) * ((hi_x - lo_x) / (n - 1.0))
            else:
                x = np.sort(np.concatenate([[lo_x, hi_x], rng.uniform(lo_x, hi_x, n - 2)]))
            y = np.maximum(1.0 + 0.3 * rng.standard_normal(n) * x, 1e-3)
            if rep % 5 == 0:
                y = rng.standard_normal(n)  # non-positive and oscillating knots
            if t in (MONO, COMP):
                d = [x[0], (x[-1] - x[0]) / (n - 1.0)] + list(y)
                dim = n
            elif t == UNIF:
                s = catmull_rom(x, y) if rep % 2 else rng.standard_normal(n)
                d = [x[0], (x[-1] - x[0]) / (n - 1.0)]
                for k in range(n):
                    d += [y[k], s[k]]
                dim = len(d)
            elif t == GEN:
                s = catmull_rom(x, y)
                d = [x[0], (x[-1] - x[0]) / (n - 1.0)]
                for k in range(n):
                    d += [y[k], s[k], x[k]]
                dim = len(d)
            else:
                d = []
                for k in range(n):
                    d += [y[k], x[k]]
                dim = len(d)
            xs = np.concatenate([
                rng.uniform(-3.0, 3.0, 24), x, x[:1] - 1e-9, x[:1] - 0.3,
                rng.uniform(-8.0, -3.0, 4), rng.uniform(3.0, 8.0, 4), [0.0, -0.0]])
            lo, hi = (-1e20, 1e20) if rep % 3 else (0.9, 1.1)
            out.append((t, d, dim, xs, lo, hi))
    return out


BILIN, BICUB = 9, 10


def grid_records(rng):
    """2-D grids (Bilinear/Bicubic payloads): nx, ny from 2 up to 20 (past
    the 16 knots the reference's four-step index search reaches), sorted
    knots with uneven spacing, smooth and oscillating values; points inside,
    on knots, just outside and far outside, with wide and tight clamps."""
    out = []
    shapes = [(2, 2), (3, 2), (2, 5), (5, 4), (9, 3), (16, 4), (17, 3), (20, 7), (4, 17)]
    for rep in range(6):
        for t in (BILIN, BICUB):
            for nx, ny in shapes:
                xk = np.sort(np.concatenate([[-3.0, 3.0], rng.uniform(-3, 3, nx - 2)]))
                yk = np.sort(np.concatenate([[-2.0, 2.0], rng.uniform(-2, 2, ny - 2)]))
                if rep % 3 == 0:
                    v = rng.standard_normal(nx * ny)
                else:
                    v = 1.0 + 0.2 * np.add.outer(xk, 0.5 * yk).ravel() \
                        + 0.05 * rng.standard_normal(nx * ny)
                d = np.concatenate([[nx, ny], xk, yk, v])
                kx = rng.integers(0, nx, 6)
                ky = rng.integers(0, ny, 6)
                px = np.concatenate([rng.uniform(-3, 3, 20), xk[kx], [-3.5, 3.5, -9.0, 9.0,
                                     xk[0] - 1e-9, xk[-1] + 1e-9]])
                py = np.concatenate([rng.uniform(-2, 2, 20), yk[ky], [-2.5, 2.5, 6.0, -6.0,
                                     yk[0] - 1e-9, yk[-1] + 1e-9]])
                lo, hi = (-1e20, 1e20) if rep % 2 == 0 else (0.9, 1.1)
                out.append((t, d, px, py, lo, hi))
    return out


def write_grid_golden(ref, rng):
    recs = grid_records(rng)
    types, lo_all, hi_all, data_all, doff, xs, ys, poff, exp = [], [], [], [], [0], [], [], [0], []
    for t, d, px, py, lo, hi in recs:
        padded = np.concatenate([d, np.zeros(16)])  # the reference's index reads past small grids
        types.append(t)
        lo_all.append(lo)
        hi_all.append(hi)
        data_all.append(d)
        doff.append(doff[-1] + d.size)
        for x, y in zip(px, py):
            xs.append(float(x))
            ys.append(float(y))
            exp.append(ref.ref_calc2d(t, float(x), float(y), lo, hi, padded.ctypes.data))
        poff.append(len(xs))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "calc2d_golden.npz")
    np.savez_compressed(out, type=np.array(types, np.int32), lo=np.array(lo_all),
                        hi=np.array(hi_all), data=np.concatenate(data_all),
                        data_off=np.array(doff, np.int64), x=np.array(xs), y=np.array(ys),
                        p_off=np.array(poff, np.int64), expected=np.array(exp))
    print(f"wrote {out}: {len(recs)} grids, {len(xs)} evaluations")


def main():
    ref = pyoracle.ref_lib()
    if ref is None:
        sys.exit("oracle/_ref/libref_calc.so is missing: run `make -C oracle ref` "
                 "where the reference tree is present")
    rng = np.random.default_rng(20261203)
    recs = unit_test_records() + random_records(rng)
    types, dims, xoff, doff, xs_all, lo_all, hi_all, data_all, exp = \
        [], [], [0], [0], [], [], [], [], []
    for t, d, dim, xs, lo, hi in recs:
        d = np.asarray(d, dtype=np.float64)
        types.append(t)
        dims.append(dim)
        lo_all.append(lo)
        hi_all.append(hi)
        data_all.append(d)
        doff.append(doff[-1] + d.size)
        for x in xs:
            xs_all.append(float(x))
            exp.append(ref.ref_calc(t, float(x), lo, hi, d.ctypes.data, dim))
        xoff.append(len(xs_all))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "calc_golden.npz")
    np.savez_compressed(
        out, type=np.array(types, np.int32), dim=np.array(dims, np.int32),
        lo=np.array(lo_all), hi=np.array(hi_all),
        data=np.concatenate(data_all), data_off=np.array(doff, np.int64),
        x=np.array(xs_all), x_off=np.array(xoff, np.int64),
        expected=np.array(exp))
    print(f"wrote {out}: {len(recs)} records, {len(xs_all)} evaluations")
    write_grid_golden(ref, np.random.default_rng(20261204))


if __name__ == "__main__":
    main()
