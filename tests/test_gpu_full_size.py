"""The batched and the adjoint workloads at their full problem sizes, with
shortened runs:

* 64 D3Q19 TRT boxes of edges {48, 64, 96}, one squirmer each (radius 6-10,
  beta uniform in [-5, 5]), FP32 storage, advanced together 200 ticks in one
  launch per tick: each box's velocity field within 1e-3 relative L2 of the
  same box run alone in FP64, its swimming speed within 1e-3.
* The 96^3 adjoint (squirmer r = 10, Guo force field N(0, 1e-7)), FP32, 200
  ticks with checkpoints every 50: the gradient along B1, B2 and two seeded
  field directions against FP64 central differences, within 1e-4 relative.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

import lbmengine_hip as lh

sys.path.insert(0, str(Path(__file__).resolve().parent.parent))
import bench  # noqa: E402

pytestmark = pytest.mark.gpu


def test_batch_64_boxes_against_fp64(gpu):
    print("64-box FP32 batch, 200 ticks, against each box in FP64", flush=True)
    boxes = bench.batch_boxes(lh, "f32", 0)
    b = lh.LBMbatch(boxes)
    b.autotune(3)
    b.run(200)
    b.sync()
    eu, es = [], []
    for m, (n, r, beta, pos, e) in zip(boxes, bench.batch_specs(0)):
        ref = lh.modelInit(**bench.workload_kwargs("f64", n))
        lh.addSquirmer(ref, radius=r, position=pos, swimmingSpeed=2e-3 / 3, beta=beta, swimmingDirection=e,
                       coupleForces=True, coupleTorques=True)
        ref.run(200)
        ref.sync()
        u64, v64 = ref.fluidVelocity, ref.particles[0].velocity
        eu.append(float(np.linalg.norm(m.fluidVelocity - u64) / np.linalg.norm(u64)))
        es.append(float(abs(np.linalg.norm(m.particles[0].velocity) - np.linalg.norm(v64)) / np.linalg.norm(v64)))
        del ref
    print(f"  u-field rel L2 max {max(eu):.3e} median {np.median(eu):.3e}; speed max {max(es):.3e}", flush=True)
    assert max(eu) < 1e-3 and max(es) < 1e-3
    assert all(m.tick == 200 for m in boxes)


def test_adjoint_96_against_fp64_differences(gpu):
    from lbmengine_hip import adjoint
    n, r, ticks, K, window = 96, 10.0, 200, 50, 100
    rng = np.random.default_rng(5)
    F = rng.normal(0.0, 1e-7, (n, n, n, 3))
    lo, hi = -n // 2 + r + 6, n // 2 - 1 - r - 6
    pos = rng.uniform(lo, hi, 3)
    e = rng.standard_normal(3)
    e /= np.linalg.norm(e)
    B1, B2 = 1e-3, 2e-3
    opts = dict(ticks=ticks, window=window, target_speed=1e-4, force_penalty=1.0)
    print("96^3 adjoint, FP32, 200 ticks, against FP64 central differences", flush=True)
    m = bench.adjoint_model(lh, "f32", n, r, B1, B2, F, pos, e)
    g = adjoint.swimming_loss_gradient(m, checkpoint_every=K, **opts)
    del m
    f0 = bench.adjoint_model(lh, "f64", n, r, B1, B2, F, pos, e).distributions
    dirs = [(1.0, 0.0, None), (0.0, 1.0, None),
            (rng.standard_normal(), rng.standard_normal(), rng.standard_normal(F.shape)),
            (0.0, 0.0, rng.standard_normal(F.shape))]
    for d1, d2, dF in dirs:
        dF = np.zeros_like(F) if dF is None else dF
        h = 1e-6
        while True:
            s1, s2, sF = h * B1 * d1, h * B2 * d2, h * 1e-7 * dF
            ta, tb = [], []
            mp = bench.adjoint_model(lh, "f64", n, r, B1 + s1, B2 + s2, F + sF, pos, e)
            mm = bench.adjoint_model(lh, "f64", n, r, B1 - s1, B2 - s2, F - sF, pos, e)
            mp.distributions = f0
            mm.distributions = f0
            Lp = adjoint.swimming_loss(mp, force_field=F + sF, trajectory=ta, **opts)
            Lm = adjoint.swimming_loss(mm, force_field=F - sF, trajectory=tb, **opts)
            if adjoint.link_switches(mp, ta, tb) == 0:
                break
            h /= 10.0
            assert h > 1e-10, "link switches at every step size"
        fd = (Lp - Lm) / 2.0
        ad = g.dB1[0] * s1 + g.dB2[0] * s2 + float(np.sum(g.dF * sF))
        err = abs(ad - fd) / abs(fd)
        print(f"  direction ({d1:+.2f}, {d2:+.2f}, field): rel err {err:.2e}", flush=True)
        assert err < 1e-4
