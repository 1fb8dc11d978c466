"""Generates the polynomial coefficients of m_cispi_m1 in cp3_math.hpp.

sqrt(2) sin(pi f) = f * S(f^2) and sqrt(2) cos(pi f) = C(f^2) on |f| <= 1/2,
each the minimax polynomial for absolute error, found by the Remez exchange
algorithm in 40-digit arithmetic (mpmath), then rounded to the target type.
(The factor sqrt(2) makes e^{2 pi i f} - 1 = (-a^2, a b) for a = f S(f^2),
b = C(f^2).)  The script prints the coefficients (highest power first, as
Horner's scheme uses them) and the maximum error of the rounded polynomials
on a dense grid, evaluated in the target type with one rounding per fused
multiply-add.

    python3 cudaprob3_amd/csrc/fit_cispi.py
"""
import mpmath as mp
import numpy as np

mp.mp.dps = 40
HALF = mp.mpf(1) / 2


def basis(x, j, odd):
    return x * (x * x) ** j if odd else (x * x) ** j


def remez(func, deg, odd, grid_n=4000, iters=40):
    """Minimax coefficients c[0..deg] of sum_j c_j basis(x, j) for func on
    [0, 1/2] (both target functions have the parity of their basis, so one
    side of the symmetric interval suffices)."""
    m = deg + 2
    xs = [HALF / 2 - HALF / 2 * mp.cos(mp.pi * k / (m - 1)) for k in range(m)]
    if odd:  # the odd error vanishes at 0: start inside the interval
        xs[0] = HALF / (4 * m)
    grid = [HALF * k / grid_n for k in range(grid_n + 1)]
    c = None
    for _ in range(iters):
        A = mp.matrix(m, m)
        b = mp.matrix(m, 1)
        for i, x in enumerate(xs):
            for j in range(deg + 1):
                A[i, j] = basis(x, j, odd)
            A[i, m - 1] = (-1) ** i
            b[i] = func(x)
        sol = mp.lu_solve(A, b)
        c = [sol[j] for j in range(deg + 1)]
        err = [sum(basis(x, j, odd) * c[j] for j in range(deg + 1)) - func(x)
               for x in grid]
        # alternating extrema of the error curve
        idx = [k for k in range(1, grid_n)
               if (err[k] - err[k - 1]) * (err[k + 1] - err[k]) <= 0]
        idx = ([] if odd else [0]) + idx + [grid_n]
        sel = []
        for k in idx:
            if sel and mp.sign(err[k]) == mp.sign(err[sel[-1]]):
                if abs(err[k]) > abs(err[sel[-1]]):
                    sel[-1] = k
            else:
                sel.append(k)
        while len(sel) > m:
            sel = sel[1:] if abs(err[sel[0]]) < abs(err[sel[-1]]) else sel[:-1]
        if len(sel) < m:
            break
        xs = [grid[k] for k in sel]
    return c


def evaluate(coefs_hi_first, f, odd, dtype):
    """The polynomial in dtype: z = f*f, Horner in z (one rounding per step,
    as an FMA), then a product with f for the odd one."""
    wide = np.float64 if dtype == np.float32 else np.longdouble
    z = (f.astype(wide) * f).astype(dtype)
    acc = np.full_like(f, coefs_hi_first[0])
    for c in coefs_hi_first[1:]:
        acc = (acc.astype(wide) * z + wide(c)).astype(dtype)
    if odd:
        acc = (acc.astype(wide) * f).astype(dtype)
    return acc


def report(name, func, deg, odd, dtype):
    c = remez(func, deg, odd)
    r = [dtype(float(v)) for v in c][::-1]
    f = np.linspace(-0.5, 0.5, 20001).astype(dtype)
    p = evaluate(r, f, odd, dtype)
    ref = np.array([float(func(mp.mpf(float(x)))) for x in f])
    e = np.abs(p.astype(np.float64) - ref).max()
    print(f"{name}, degree {2 * deg + (1 if odd else 0)}, "
          f"{np.dtype(dtype).name}: max error {e:.3g}")
    print("  " + ", ".join(repr(float(v)) for v in r))


if __name__ == "__main__":
    sin = lambda x: mp.sqrt(2) * mp.sin(mp.pi * x)
    cos = lambda x: mp.sqrt(2) * mp.cos(mp.pi * x)
    report("sqrt(2) sin(pi f)", sin, 4, True, np.float32)
    report("sqrt(2) cos(pi f)", cos, 5, False, np.float32)
    report("sqrt(2) sin(pi f)", sin, 7, True, np.float64)
    report("sqrt(2) cos(pi f)", cos, 8, False, np.float64)
