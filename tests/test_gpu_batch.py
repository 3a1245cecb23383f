"""Batched boxes: every box of an LBMbatch must end bit-identical to the same
box advanced alone with LBMpropagate (same kernels, same arithmetic), over
runs that exercise graph replays, odd and even run lengths, boxes run alone
between batch runs, block shapes and the autotuner; one box is also checked
against the oracle, and an FP32 batch against FP64 boxes run alone.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KW3 = dict(dims=3, velocities="D3Q19", latticeParameter=1.0, relaxationTimeRatio=0.8, kineticParameter=3 / 16)
KW2 = dict(dims=2, velocities="D2Q9", latticeParameter=1.0, relaxationTimeRatio=0.8, collisionModel="trt")


def rel(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    scale = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / scale)


def twins(gpu, kw, particles, seed):
    """Two identical models (same perturbed populations, same particles)."""
    a = gpu.modelInit(**kw)
    b = gpu.modelInit(**kw)
    rng = np.random.default_rng(seed)
    f = a.distributions
    f = np.asfortranarray(f * (1 + 0.01 * rng.standard_normal(f.shape)))
    f[a.distributions == 0] = 0.0
    a.distributions = f
    b.distributions = f
    for kind, pk in particles:
        add = gpu.addSquirmer if kind == "squirmer" else gpu.addBead
        add(a, **pk)
        add(b, **pk)
    return a, b


def sq(r, pos, beta, e, B1=2e-3):
    return ("squirmer", dict(radius=r, position=pos, swimmingSpeed=2 * B1 / 3, beta=beta, swimmingDirection=e,
                             coupleForces=True, coupleTorques=True))


def same(a, b):
    assert np.array_equal(a.distributions, b.distributions)
    assert a.tick == b.tick and a.time == b.time
    for pa, pb in zip(a.particles, b.particles):
        for attr in ("position", "velocity", "angularVelocity", "swimmingDirection"):
            assert np.array_equal(getattr(pa, attr), getattr(pb, attr)), attr
        fa, ta = pa.load()
        fb, tb = pb.load()
        assert np.array_equal(fa, fb) and np.array_equal(ta, tb)


def uneven_boxes(gpu, storage="f64"):
    """Five D3Q19 TRT boxes of different shapes; four carry particles."""
    specs = [
        (dict(xlims=(-10, 9)), [sq(4.0, [0.3, -0.2, 0.1], -3.0, [1, 0, 0])]),
        (dict(xlims=(-12, 11), ylims=(-9, 10), zlims=(-8, 9)), [sq(5.0, [0.1, 0.4, -0.3], 2.0, [0, 1, 1])]),
        (dict(xlims=(0, 15), ylims=(0, 11), zlims=(0, 9)), []),
        (dict(xlims=(-13, 12), ylims=(-9, 10), zlims=(-8, 9)),
         [sq(4.0, [-4.6, 0.3, 0.1], 1.0, [1, 0.2, 0]),
          ("bead", dict(radius=4.0, position=[4.6, -0.2, 0.3], coupleForces=True, coupleTorques=True))]),
        (dict(xlims=(-9, 8), walledDimensions=[3]), [sq(3.5, [0.2, 0.1, -0.4], 5.0, [0, 0, 1])]),
    ]
    out = []
    for i, (lims, parts) in enumerate(specs):
        out.append(twins(gpu, dict(**KW3, storage=storage, **lims), parts, seed=40 + i))
    return out


@pytest.mark.parametrize("storage", ["f64", "f32", "f16"])
def test_batch_matches_boxes_run_alone(gpu, storage):
    pairs = uneven_boxes(gpu, storage)
    batch = gpu.LBMbatch([a for a, _ in pairs])
    batch.set_graph(4)
    for steps in (1, 9, 12, 4):  # graph replays, odd and even run lengths
        gpu.LBMpropagate(batch, ticks=steps)
        for _, b in pairs:
            gpu.LBMpropagate(b, ticks=steps)
        for a, b in pairs:
            same(a, b)


def test_batch_against_oracle(gpu, oracle):
    kw = dict(**KW3, xlims=(-10, 9))
    pk = dict(radius=4.0, position=[0.3, -0.2, 0.1], swimmingSpeed=2e-3, beta=-3.0, swimmingDirection=[1, 1, 0],
              coupleForces=True, coupleTorques=True)
    g = gpu.modelInit(**kw)
    other = gpu.modelInit(**dict(KW3, xlims=(0, 15), ylims=(0, 11), zlims=(0, 9)))
    o = oracle.oracle_model(**kw)
    gpu.addSquirmer(g, **pk)
    o.add("squirmer", **pk)
    batch = gpu.LBMbatch([other, g])
    gpu.LBMpropagate(batch, ticks=30)
    o.run(30)
    assert rel(g.distributions, o.distributions) <= 1e-12
    for attr in ("position", "velocity", "angularVelocity", "swimmingDirection"):
        assert rel(getattr(g.particles[0], attr), getattr(o.particles[0], attr)) <= 1e-12
    ng, dg = g.particles[0].links()
    no, do = o.particles[0].links()
    assert np.array_equal(ng, no) and np.array_equal(dg, do)


def test_boxes_run_alone_between_batch_runs(gpu):
    """A box advanced on its own (by an odd count) between batch runs leaves
    the boxes at different buffer parities; the batch follows each box."""
    pairs = uneven_boxes(gpu)[:3]
    batch = gpu.LBMbatch([a for a, _ in pairs])
    batch.set_graph(2)
    gpu.LBMpropagate(batch, ticks=5)
    pairs[1][0].run(3)
    gpu.LBMpropagate(batch, ticks=6)
    for i, (_, b) in enumerate(pairs):
        gpu.LBMpropagate(b, ticks=14 if i == 1 else 11)
    for a, b in pairs:
        same(a, b)


@pytest.mark.parametrize("block", [(64, 1, 1), (16, 4, 1), (8, 4, 2)])
def test_batch_block_shapes(gpu, block):
    pairs = uneven_boxes(gpu)[1:4]
    batch = gpu.LBMbatch([a for a, _ in pairs])
    batch.set_block(*block)
    assert batch.get_block() == block
    gpu.LBMpropagate(batch, ticks=7)
    for _, b in pairs:
        gpu.LBMpropagate(b, ticks=7)
    for a, b in pairs:
        same(a, b)


@pytest.mark.parametrize("threads", [64, 192, 320])
def test_batch_row_tiles(gpu, threads):
    """Every box its own tile shape of `threads` threads spanning its rows
    (x extents 16..26 nodes: tiles of 16 to 64 threads wide, some with
    several z layers or overhanging the box), then back to one shared shape."""
    pairs = uneven_boxes(gpu)
    batch = gpu.LBMbatch([a for a, _ in pairs])
    batch.set_graph(4)
    batch.set_row_tiles(threads)
    assert batch.get_row_tiles() == threads
    for steps in (7, 8):
        gpu.LBMpropagate(batch, ticks=steps)
        for _, b in pairs:
            gpu.LBMpropagate(b, ticks=steps)
        for a, b in pairs:
            same(a, b)
    batch.set_block(64, 1, 1)
    assert batch.get_row_tiles() == 0 and batch.get_block() == (64, 1, 1)
    gpu.LBMpropagate(batch, ticks=5)
    for _, b in pairs:
        gpu.LBMpropagate(b, ticks=5)
    for a, b in pairs:
        same(a, b)


def test_batch_autotune_keeps_state(gpu):
    pairs = uneven_boxes(gpu, "f32")
    batch = gpu.LBMbatch([a for a, _ in pairs])
    gpu.LBMpropagate(batch, ticks=3)
    for _, b in pairs:
        gpu.LBMpropagate(b, ticks=3)
    f0 = [a.distributions for a, _ in pairs]
    shape, ms = batch.autotune(2)
    if shape[1] == 0:  # row tiles
        assert ms > 0 and shape[0] >= 64 and batch.get_row_tiles() == shape[0]
    else:
        assert ms > 0 and np.prod(shape) >= 64 and batch.get_block() == tuple(shape)
    for (a, _), f in zip(pairs, f0):
        assert np.array_equal(a.distributions, f)
    gpu.LBMpropagate(batch, ticks=8)
    for _, b in pairs:
        gpu.LBMpropagate(b, ticks=8)
    for a, b in pairs:
        same(a, b)


def test_d2q9_batch(gpu):
    specs = [
        (dict(xlims=(-24, 23)), [sq(6.0, [0.3, 0.2], -1.0, [1, 0.5])]),
        (dict(xlims=(-16, 15), ylims=(-20, 19)), [sq(5.0, [0.1, -0.4], 3.0, [0, 1])]),
        (dict(xlims=(0, 31), ylims=(0, 17), walledDimensions=[2]), []),
    ]
    pairs = [twins(gpu, dict(**KW2, **lims), parts, seed=60 + i) for i, (lims, parts) in enumerate(specs)]
    batch = gpu.LBMbatch([a for a, _ in pairs])
    gpu.LBMpropagate(batch, ticks=41)
    for _, b in pairs:
        gpu.LBMpropagate(b, ticks=41)
    for a, b in pairs:
        same(a, b)


def test_batch_rejects_mixed_kinds(gpu):
    a = gpu.modelInit(**KW3, xlims=(-8, 7))
    b = gpu.modelInit(**KW3, xlims=(-8, 7), storage="f32")
    with pytest.raises(gpu.LBMError):
        gpu.LBMbatch([a, b])
    with pytest.raises(gpu.LBMError):
        gpu.LBMbatch([a, a])


def test_fp32_batch_against_fp64_boxes(gpu):
    """A small beta sweep: FP32 boxes in one batch against FP64 boxes run
    alone (velocity field and swimming speed within 1e-3 relative)."""
    rng = np.random.default_rng(7)
    edges = [32, 40, 48]
    lo, hi = [], []
    for i, beta in enumerate([-5.0, 0.0, 5.0]):
        n = edges[i]
        e = rng.standard_normal(3)
        pk = dict(radius=6.0, position=rng.uniform(-1, 1, 3), swimmingSpeed=2e-3 / 3, beta=beta,
                  swimmingDirection=e, coupleForces=True, coupleTorques=True)
        for store, acc in (("f32", lo), ("f64", hi)):
            m = gpu.modelInit(**KW3, xlims=(-n // 2, n // 2 - 1), storage=store)
            gpu.addSquirmer(m, **pk)
            acc.append(m)
    batch = gpu.LBMbatch(lo)
    batch.autotune(2)
    gpu.LBMpropagate(batch, ticks=400)
    for m in hi:
        gpu.LBMpropagate(m, ticks=400)
    for a, b in zip(lo, hi):
        eu = float(np.linalg.norm(a.fluidVelocity - b.fluidVelocity) / np.linalg.norm(b.fluidVelocity))
        va, vb = np.linalg.norm(a.particles[0].velocity), np.linalg.norm(b.particles[0].velocity)
        assert eu < 1e-3 and abs(va - vb) / vb < 1e-3, (eu, va, vb)


def test_batch_with_interactions_matches_boxes_run_alone(gpu):
    """Interaction terms ride in each box's arguments: a batch of bonded
    boxes ends bit-identical to the boxes advanced alone."""
    from lbmengine_hip import interactions as it
    kw = dict(**KW3, xlims=(-16, 15))
    bead = dict(radius=3.0, coupleForces=True, coupleTorques=True)
    pairs = []
    for i, k in enumerate((0.05, 0.02)):
        a, b = twins(gpu, kw, [("bead", dict(bead, position=[-4.2, 0.3 * i, 0.1])),
                               ("bead", dict(bead, position=[5.1, -1.2, 0.7]))], seed=80 + i)
        for m in (a, b):
            it.addLinearBond(m, 1, 2, equilibriumDisplacement=7.0, stiffness=k)
            it.addDipoles(m, magneticField=[0.0, 0.3, 0.4], permeability=1.0)
        pairs.append((a, b))
    batch = gpu.LBMbatch([a for a, _ in pairs])
    for steps in (3, 40):
        gpu.LBMpropagate(batch, ticks=steps)
        for _, b in pairs:
            gpu.LBMpropagate(b, ticks=steps)
        for a, b in pairs:
            same(a, b)


def test_box_segments_ahead_of_their_tiles(gpu):
    """Thirty-three boxes: each box's particle blocks are launched two boxes
    ahead of its tiles (a sixteenth of the batch), and a box without
    particles has no such blocks; every box still ends bit-identical to the
    box advanced alone, for FP32 and FP16 storage and two block shapes."""
    rng = np.random.default_rng(7)
    for storage, block in (("f32", (32, 2, 1)), ("f16", (16, 4, 2))):
        pairs = []
        for i in range(33):
            n = int(rng.choice([16, 20, 24]))
            lims = dict(xlims=(-n // 2, n // 2 - 1), ylims=(-10, 9), zlims=(-9, 10))
            parts = [] if i == 5 else [sq(float(rng.uniform(2.5, 3.5)), list(rng.uniform(-1, 1, 3)),
                                          float(rng.uniform(-5, 5)), list(rng.standard_normal(3)))]
            pairs.append(twins(gpu, dict(**KW3, storage=storage, **lims), parts, seed=100 + i))
        batch = gpu.LBMbatch([a for a, _ in pairs])
        batch.set_block(*block)
        batch.set_graph(4)
        for steps in (2, 9):
            gpu.LBMpropagate(batch, ticks=steps)
            for _, b in pairs:
                b.set_block(*block)
                gpu.LBMpropagate(b, ticks=steps)
            for a, b in pairs:
                same(a, b)
