"""Velocity sets in the reference's order (gpu/misc.jl:195-272).

The comprehensions `[[p,q] for p in [-1,1], q in [-1,1]] |> vec` (and
`[[p,q,r] ...]`) run p fastest (Julia is column-major), which fixes the
order of the diagonal velocities; q indices here are 0-based.
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np


def _pairs(axes):
    out = []
    for q in (-1, 1):          # outer (second comprehension variable)
        for p in (-1, 1):      # inner: p runs fastest
            v = [0, 0, 0]
            v[axes[0]], v[axes[1]] = p, q
            out.append(v)
    return out


def d2q9():
    cs = [[0, 0], [-1, 0], [1, 0], [0, -1], [0, 1]] + [v[:2] for v in _pairs((0, 1))]
    ws = [Fraction(4, 9)] + [Fraction(1, 9)] * 4 + [Fraction(1, 36)] * 4
    return np.array(cs, dtype=np.int64), ws


def _corners():
    out = []
    for r in (-1, 1):              # outermost
        for q in (-1, 1):
            for p in (-1, 1):      # p runs fastest
                out.append([p, q, r])
    return out


_AXES3 = [[0, 0, 0], [-1, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0], [0, 0, -1], [0, 0, 1]]


def d3q15():
    cs = _AXES3 + _corners()
    ws = [Fraction(2, 9)] + [Fraction(1, 9)] * 6 + [Fraction(1, 72)] * 8
    return np.array(cs, dtype=np.int64), ws


def d3q19():
    cs = _AXES3 + _pairs((1, 2)) + _pairs((0, 2)) + _pairs((0, 1))
    ws = [Fraction(1, 3)] + [Fraction(1, 18)] * 6 + [Fraction(1, 36)] * 12
    return np.array(cs, dtype=np.int64), ws


def d3q27():
    cs = _AXES3 + _pairs((1, 2)) + _pairs((0, 2)) + _pairs((0, 1)) + _corners()
    ws = [Fraction(8, 27)] + [Fraction(2, 27)] * 6 + [Fraction(1, 54)] * 12 + [Fraction(1, 216)] * 8
    return np.array(cs, dtype=np.int64), ws


VELOCITY_SETS = {"D2Q9": d2q9, "D3Q15": d3q15, "D3Q19": d3q19, "D3Q27": d3q27}


def velocity_set(name: str):
    """(cs int array (Q, d), ws float64 array (Q,), opposite index array (Q,))."""
    key = name.lstrip(":")
    if key not in VELOCITY_SETS:
        raise NotImplementedError(
            f"velocity set {name} is not available in the hip backend (D2Q9, D3Q15, D3Q19 and D3Q27 are)")
    cs, ws = VELOCITY_SETS[key]()
    w = np.array([float(x) for x in ws])
    opp = np.array([next(j for j in range(len(cs)) if (cs[j] == -cs[i]).all()) for i in range(len(cs))])
    return cs, w, opp
