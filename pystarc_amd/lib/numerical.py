"""Romberg quadrature for k_b (pystarc/lib/numerical.py:74-97): trapezoid
refinement with Richardson extrapolation, stopping when two diagonal entries
differ by less than ``tol`` (absolute), at most ``max_order`` halvings."""
from __future__ import annotations


def romberg_integrate(f, a, b, tol=1e-8, max_order=12):
    prev_row = [0.5 * (b - a) * (f(a) + f(b))]
    h = b - a
    for i in range(1, max_order + 1):
        h *= 0.5
        mids = sum(f(a + (2 * k - 1) * h) for k in range(1, 2 ** (i - 1) + 1))
        row = [0.5 * prev_row[0] + h * mids]
        for j in range(1, i + 1):
            row.append(row[j - 1] + (row[j - 1] - prev_row[j - 1]) / (4 ** j - 1))
        if i >= 2 and abs(row[i] - prev_row[i - 1]) < tol:
            return row[i]
        prev_row = row
    return prev_row[-1]
