"""Velocity sets against the reference's tables (gpu/misc.jl:195-248) and
against the isotropy conditions any D2Q9/D3Q19 must satisfy."""
from fractions import Fraction

import numpy as np
import pytest

from lbmengine_hip.lattice import d2q9, d3q15, d3q19, d3q27, velocity_set


def test_d2q9_order_and_opposites():
    cs, w, opp = velocity_set("D2Q9")
    # gpu/misc.jl:195-209: rest, [±1,0], [0,±1], then p fastest in [[p,q] ...]
    assert cs.tolist() == [[0, 0], [-1, 0], [1, 0], [0, -1], [0, 1], [-1, -1], [1, -1], [-1, 1], [1, 1]]
    # oppositeVectorId, 1-based in the reference: [1,3,2,5,4,9,8,7,6]
    assert (opp + 1).tolist() == [1, 3, 2, 5, 4, 9, 8, 7, 6]
    assert np.allclose(w, [4 / 9] + [1 / 9] * 4 + [1 / 36] * 4, rtol=0, atol=0)


def test_d3q19_order_and_opposites():
    cs, w, opp = velocity_set(":D3Q19")
    assert cs[:7].tolist() == [[0, 0, 0], [-1, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0], [0, 0, -1], [0, 0, 1]]
    assert cs[7:11].tolist() == [[0, -1, -1], [0, 1, -1], [0, -1, 1], [0, 1, 1]]
    assert cs[11:15].tolist() == [[-1, 0, -1], [1, 0, -1], [-1, 0, 1], [1, 0, 1]]
    assert cs[15:].tolist() == [[-1, -1, 0], [1, -1, 0], [-1, 1, 0], [1, 1, 0]]
    assert (opp + 1).tolist() == [1, 3, 2, 5, 4, 7, 6, 11, 10, 9, 8, 15, 14, 13, 12, 19, 18, 17, 16]


def test_d3q15_d3q27_order_and_opposites():
    # gpu/misc.jl:212-272: rest, axes, [edges as D3Q19], then [[p,q,r] ...]
    # with p fastest
    corners = [[p, q, r] for r in (-1, 1) for q in (-1, 1) for p in (-1, 1)]
    cs15, w15, opp15 = velocity_set("D3Q15")
    assert cs15[7:].tolist() == corners
    assert np.allclose(w15, [2 / 9] + [1 / 9] * 6 + [1 / 72] * 8, rtol=0, atol=0)
    assert (opp15 + 1).tolist() == [1, 3, 2, 5, 4, 7, 6, 15, 14, 13, 12, 11, 10, 9, 8]
    cs27, w27, opp27 = velocity_set("D3Q27")
    cs19, _, _ = velocity_set("D3Q19")
    assert cs27[:19].tolist() == cs19.tolist() and cs27[19:].tolist() == corners
    assert np.allclose(w27, [8 / 27] + [2 / 27] * 6 + [1 / 54] * 12 + [1 / 216] * 8, rtol=0, atol=0)
    assert all((cs27[opp27[i]] == -cs27[i]).all() for i in range(27))


@pytest.mark.parametrize("make", [d2q9, d3q15, d3q19, d3q27])
def test_isotropy_exact(make):
    cs, ws = make()
    d = cs.shape[1]
    assert sum(ws) == 1
    for a in range(d):
        assert sum(w * int(c[a]) for c, w in zip(cs, ws)) == 0
        for b in range(d):
            m2 = sum(w * int(c[a]) * int(c[b]) for c, w in zip(cs, ws))
            assert m2 == (Fraction(1, 3) if a == b else 0)
            # third moments vanish
            for g in range(d):
                assert sum(w * int(c[a]) * int(c[b]) * int(c[g]) for c, w in zip(cs, ws)) == 0
    # fourth moment of the isotropic lattices: c_s^4 (d_ab d_gd + d_ag d_bd + d_ad d_bg)
    delta = np.eye(d, dtype=int)
    for a in range(d):
        for b in range(d):
            for g in range(d):
                for e in range(d):
                    m4 = sum(w * int(c[a]) * int(c[b]) * int(c[g]) * int(c[e]) for c, w in zip(cs, ws))
                    iso = Fraction(1, 9) * (delta[a, b] * delta[g, e] + delta[a, g] * delta[b, e] + delta[a, e] * delta[b, g])
                    assert m4 == iso


def test_unsupported_sets_raise():
    with pytest.raises(NotImplementedError):
        velocity_set("D1Q3")
