"""The dual-number restatement of the oracle's tick (oracle/lbm_oracle_dual.cpp),
pinned on the CPU: its value parts reproduce the FP64 oracle's swimmer
speeds bit for bit, and its tangent parts equal central finite differences
of the oracle's loss along B1, B2 and a force-field direction."""
import math

import numpy as np

KW3 = dict(dims=3, velocities="D3Q19", latticeParameter=1.0, relaxationTimeRatio=0.8, kineticParameter=3 / 16)


def case(oracle, n=12, r=3.2, seed=1, beta=2.0, walls=False):
    from lbmengine_hip.model import resolve_model, resolve_particle
    rng = np.random.default_rng(seed)
    F = rng.normal(0.0, 1e-6, (n, n, n, 3))
    kw = dict(**KW3, xlims=(-n // 2, n // 2 - 1), forceDensity=F)
    if walls:
        kw["walledDimensions"] = [3]
    pk = dict(radius=r, position=rng.uniform(-0.5, 0.5, 3), swimmingSpeed=2e-3, beta=beta,
              swimmingDirection=rng.standard_normal(3), coupleForces=True, coupleTorques=True)
    rm = resolve_model(**kw)
    p = resolve_particle(rm, kind="squirmer", **pk)
    return rm, p, F, kw, pk


def oracle_speeds(oracle, kw, pk, ticks, window, f0=None):
    o = oracle.oracle_model(**kw)
    if f0 is not None:
        o.distributions = f0
    o.add("squirmer", **pk)
    sp = []
    for t in range(ticks):
        o.run(1)
        if t >= ticks - window:
            v = o.particles[0].velocity
            s2 = 0.0
            for vk in v:  # the dual's order: sqrt(((v0 v0) + v1 v1) + v2 v2)
                s2 = s2 + float(vk) * float(vk)
            sp.append(math.sqrt(s2))
    return np.array(sp), o


def test_dual_values_reproduce_the_oracle(oracle):
    for walls in (False, True):
        rm, p, F, kw, pk = case(oracle, walls=walls)
        o0 = oracle.oracle_model(**kw)
        f0 = o0.distributions
        sp_o, _ = oracle_speeds(oracle, kw, pk, 12, 6)
        Fj = np.asfortranarray(F).reshape(-1, order="F")
        loss, dl, sp_d = oracle.dual_loss_jvp(rm, [p], f0, ticks=12, window=6, F=Fj)
        assert np.array_equal(sp_d, sp_o)
        assert dl == 0.0


def test_dual_tangents_match_finite_differences(oracle):
    from lbmengine_hip.model import resolve_particle
    rm, p, F, kw, pk = case(oracle, n=12, r=3.2, seed=4)
    f0 = oracle.oracle_model(**kw).distributions
    Fj = np.asfortranarray(F).reshape(-1, order="F")
    rng = np.random.default_rng(9)
    dFd = rng.standard_normal(Fj.shape)
    T, W, target, pen = 14, 7, 1e-4, 1e3

    def loss_at(b1, b2, Fv):
        q = resolve_particle(rm, kind="squirmer", **pk)
        q.B1, q.B2 = b1, b2
        return oracle.dual_loss_jvp(rm, [q], f0, ticks=T, window=W, F=Fv, target=target, penalty=pen)[0]

    for dB1, dB2, dF in ((1.0, 0.0, None), (0.0, 1.0, None), (0.0, 0.0, dFd), (0.3, -0.7, dFd)):
        _, dl, _ = oracle.dual_loss_jvp(rm, [p], f0, ticks=T, window=W, F=Fj, dB1=[dB1 * p.B1], dB2=[dB2 * p.B2],
                                        dF=None if dF is None else dF * 1e-6, target=target, penalty=pen)
        h = 1e-6
        Fp = Fj if dF is None else Fj + h * 1e-6 * dF
        Fm = Fj if dF is None else Fj - h * 1e-6 * dF
        fd = (loss_at(p.B1 * (1 + h * dB1), p.B2 * (1 + h * dB2), Fp)
              - loss_at(p.B1 * (1 - h * dB1), p.B2 * (1 - h * dB2), Fm)) / (2 * h)
        assert abs(dl - fd) <= 1e-5 * abs(fd), (dB1, dB2, dl, fd)
