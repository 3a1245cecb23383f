"""Discrete adjoint of the step against central finite differences of the
forward step (FP64 runs of the same loss on perturbed parameters):

    L = (mean |V| over the last `window` ticks - target)^2 + penalty * sum F^2

with respect to the squirmer's B1, B2 and the Guo force field, along seeded
directions, and against the dual-number restatement of the oracle's tick.
All hold the starting populations fixed (the model's initial state is an
input; its Guo-shifted equilibrium is not differentiated).  Steps are
relative (1e-6 of each parameter's scale).  The link
set is piecewise constant in the positions: where the +/- runs' link sets
differ at some tick (adjoint.link_switches), the difference quotient holds
a jump the derivative does not, and the step is divided by 10."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KW3 = dict(dims=3, velocities="D3Q19", latticeParameter=1.0, relaxationTimeRatio=0.8, kineticParameter=3 / 16)


def setup(n, r, seed, f_sigma=1e-7):
    rng = np.random.default_rng(seed)
    F = rng.normal(0.0, f_sigma, (n, n, n, 3))
    e = rng.standard_normal(3)
    pos = rng.uniform(-1.0, 1.0, 3)
    return F, pos, e / np.linalg.norm(e)


def make(gpu, storage, n, r, B1, B2, F, pos, e):
    m = gpu.modelInit(**KW3, xlims=(-n // 2, n // 2 - 1), storage=storage, forceDensity=np.asfortranarray(F))
    gpu.addSquirmer(m, radius=r, position=pos, swimmingSpeed=2 * B1 / 3, beta=B2 / B1, swimmingDirection=e,
                    coupleForces=True, coupleTorques=True)
    return m


def directional_checks(gpu, storage, n, r, ticks, K, window, seed, ndirs, target=1e-4, penalty=1e3):
    from lbmengine_hip import adjoint
    F, pos, e = setup(n, r, seed)
    B1, B2 = 1e-3, 2e-3
    m = make(gpu, storage, n, r, B1, B2, F, pos, e)
    # the gradient holds the starting populations fixed: so do the runs below
    f0 = make(gpu, "f64", n, r, B1, B2, F, pos, e).distributions
    g = adjoint.swimming_loss_gradient(m, ticks=ticks, checkpoint_every=K, window=window, target_speed=target,
                                       force_penalty=penalty)
    assert m.tick == ticks
    opts = dict(ticks=ticks, window=window, target_speed=target, force_penalty=penalty)
    rng = np.random.default_rng(seed + 100)
    h = 1e-6
    out = []
    for d in range(ndirs):
        if d == 0:
            dB1, dB2, dFd = 1.0, 0.0, np.zeros_like(F)
        elif d == 1:
            dB1, dB2, dFd = 0.0, 1.0, np.zeros_like(F)
        else:
            dB1, dB2, dFd = rng.standard_normal(), rng.standard_normal(), rng.standard_normal(F.shape)
        # steps relative to each parameter's scale, shrunk while the two
        # runs' link sets differ at some tick (a jump, not a derivative)
        hd = h
        while True:
            sB1, sB2, sF = hd * B1 * dB1, hd * B2 * dB2, hd * 1e-7 * dFd
            ta, tb = [], []
            mp = make(gpu, "f64", n, r, B1 + sB1, B2 + sB2, F + sF, pos, e)
            mm = make(gpu, "f64", n, r, B1 - sB1, B2 - sB2, F - sF, pos, e)
            mp.distributions = f0
            mm.distributions = f0
            Lp = adjoint.swimming_loss(mp, force_field=F + sF, trajectory=ta, **opts)
            Lm = adjoint.swimming_loss(mm, force_field=F - sF, trajectory=tb, **opts)
            if adjoint.link_switches(mp, ta, tb) == 0:
                break
            hd /= 10.0
            assert hd > 1e-10, "link switches at every step size"
        fd = (Lp - Lm) / 2.0
        ad = g.dB1[0] * sB1 + g.dB2[0] * sB2 + float(np.sum(g.dF * sF))
        out.append((ad, fd))
    return g, out


def test_fp64_adjoint_against_finite_differences(gpu):
    g, pairs = directional_checks(gpu, "f64", n=24, r=5.0, ticks=40, K=16, window=10, seed=3, ndirs=5)
    assert np.isfinite(g.loss) and g.loss > 0
    for ad, fd in pairs:
        assert abs(ad - fd) <= 1e-5 * abs(fd), (ad, fd)


def test_fp32_adjoint_against_fp64_finite_differences(gpu):
    g, pairs = directional_checks(gpu, "f32", n=40, r=7.0, ticks=300, K=64, window=100, seed=5, ndirs=4)
    for ad, fd in pairs:
        assert abs(ad - fd) <= 1e-3 * abs(fd), (ad, fd)


def test_adjoint_against_the_dual_number_oracle(gpu, oracle):
    """FP64 adjoint against the forward-mode derivative of the oracle's tick
    (dual numbers) from the same populations, along B1, B2, a force-field
    direction and a mixed one."""
    from lbmengine_hip import adjoint
    from lbmengine_hip.model import resolve_particle
    n, r, T, W, target, pen = 16, 3.6, 16, 6, 1e-4, 1e3
    F, pos, e = setup(n, r, 21, f_sigma=1e-6)
    B1, B2 = 1e-3, -2e-3
    m = make(gpu, "f64", n, r, B1, B2, F, pos, e)
    f0 = m.distributions
    p = resolve_particle(m.resolved, kind="squirmer", radius=r, position=pos, swimmingSpeed=2 * B1 / 3,
                         beta=B2 / B1, swimmingDirection=e, coupleForces=True, coupleTorques=True)
    g = adjoint.swimming_loss_gradient(m, ticks=T, checkpoint_every=5, window=W, target_speed=target,
                                       force_penalty=pen)
    Fj = np.asfortranarray(F).reshape(-1, order="F")
    gF = np.asfortranarray(g.dF).reshape(-1, order="F")
    rng = np.random.default_rng(4)
    dFd = rng.standard_normal(Fj.shape) * 1e-6
    for a, b, dF in ((B1, 0.0, None), (0.0, B2, None), (0.0, 0.0, dFd), (0.4 * B1, -0.3 * B2, dFd)):
        L, dl, _ = oracle.dual_loss_jvp(m.resolved, [p], f0, ticks=T, window=W, F=Fj, dB1=[a], dB2=[b], dF=dF,
                                        target=target, penalty=pen)
        ad = g.dB1[0] * a + g.dB2[0] * b + (0.0 if dF is None else float(np.dot(gF, dF)))
        assert abs(L - g.loss) <= 1e-12 * L
        assert abs(ad - dl) <= 1e-8 * abs(dl), (a, b, ad, dl)


@pytest.mark.parametrize("sep", [7.5, 6.4], ids=["gap", "overlap"])
def test_two_particles_against_the_dual_number_oracle(gpu, oracle, sep):
    """Two squirmers (radii 3.6 and 3.4) whose centres lie `sep` apart: 0.5
    node apart at their closest (nodes in the gap hold link slots of both)
    or overlapping by 0.6 (slots with two owners: the Ladd terms' owner chain
    runs backwards through both particles).  The second swimmer's speed is
    the loss; FP64 adjoint against the dual-number oracle along each
    particle's B1 and B2, a force-field direction and a mixed one."""
    from lbmengine_hip import adjoint
    from lbmengine_hip.model import resolve_particle
    n, T, W, target, pen = 20, 14, 6, 1e-4, 1e3
    rng = np.random.default_rng(31)
    F = rng.normal(0.0, 1e-6, (n, n, n, 3))
    ax = np.array([0.99, -0.1, 0.05])
    ax /= np.linalg.norm(ax)
    c0 = np.array([-0.2, 0.1, 0.15])
    specs = [dict(radius=3.6, position=c0 - 0.5 * sep * ax, swimmingSpeed=2e-3 / 3, beta=-2.0,
                  swimmingDirection=[0.6, 0.8, 0.0]),
             dict(radius=3.4, position=c0 + 0.5 * sep * ax, swimmingSpeed=1.5e-3 / 3, beta=3.0,
                  swimmingDirection=[-0.48, 0.6, 0.64])]
    m = gpu.modelInit(**KW3, xlims=(-n // 2, n // 2 - 1), storage="f64", forceDensity=np.asfortranarray(F))
    for sp in specs:
        gpu.addSquirmer(m, **sp, coupleForces=True, coupleTorques=True)
    f0 = m.distributions
    ps = [resolve_particle(m.resolved, kind="squirmer", **sp, coupleForces=True, coupleTorques=True)
          for sp in specs]
    g = adjoint.swimming_loss_gradient(m, ticks=T, checkpoint_every=5, window=W, particle=1, target_speed=target,
                                       force_penalty=pen)
    assert g.dB1.shape == (2,) and g.dB2.shape == (2,)
    Fj = np.asfortranarray(F).reshape(-1, order="F")
    gF = np.asfortranarray(g.dF).reshape(-1, order="F")
    dFd = rng.standard_normal(Fj.shape) * 1e-6
    B1 = [p.B1 for p in ps]
    B2 = [p.B2 for p in ps]
    dirs = ([B1[0], 0.0], [0.0, 0.0], None), ([0.0, B1[1]], [0.0, 0.0], None), ([0.0, 0.0], [B2[0], 0.0], None), \
           ([0.0, 0.0], [0.0, B2[1]], None), ([0.0, 0.0], [0.0, 0.0], dFd), \
           ([0.3 * B1[0], -0.5 * B1[1]], [0.2 * B2[0], 0.7 * B2[1]], dFd)
    for a, b, dF in dirs:
        L, dl, _ = oracle.dual_loss_jvp(m.resolved, ps, f0, ticks=T, window=W, F=Fj, dB1=a, dB2=b, dF=dF,
                                        swimmer=1, target=target, penalty=pen)
        ad = float(np.dot(g.dB1, a) + np.dot(g.dB2, b)) + (0.0 if dF is None else float(np.dot(gF, dF)))
        assert abs(L - g.loss) <= 1e-12 * L
        assert abs(ad - dl) <= 1e-8 * abs(dl), (a, b, ad, dl)


def test_walled_lattice_against_the_dual_number_oracle(gpu, oracle):
    """A lattice with wall layers at z = 0 and z = n - 1 (the bulk kernel's
    general path: wall tests, bounce-back of the adjoint populations): FP64
    adjoint against the dual-number oracle along B1, B2, a force-field
    direction and a mixed one; FP32 storage (dL/dF summed per segment in
    FP32, folded into FP64) against the FP64 adjoint's directional
    derivatives within 1e-3."""
    from lbmengine_hip import adjoint
    from lbmengine_hip.model import resolve_particle
    n, r, T, W, target, pen = 18, 3.4, 14, 6, 1e-4, 1e3
    rng = np.random.default_rng(41)
    F = rng.normal(0.0, 1e-6, (n, n, n, 3))
    pos, e = [0.3, -0.4, 0.2], [0.36, 0.48, 0.8]
    B1, B2 = 1e-3, -1.5e-3

    def walled(storage):
        m = gpu.modelInit(**KW3, xlims=(-n // 2, n // 2 - 1), walledDimensions=[3], storage=storage,
                          forceDensity=np.asfortranarray(F))
        gpu.addSquirmer(m, radius=r, position=pos, swimmingSpeed=2 * B1 / 3, beta=B2 / B1, swimmingDirection=e,
                        coupleForces=True, coupleTorques=True)
        return m

    m = walled("f64")
    f0 = m.distributions
    p = resolve_particle(m.resolved, kind="squirmer", radius=r, position=pos, swimmingSpeed=2 * B1 / 3,
                         beta=B2 / B1, swimmingDirection=e, coupleForces=True, coupleTorques=True)
    g = adjoint.swimming_loss_gradient(m, ticks=T, checkpoint_every=5, window=W, target_speed=target,
                                       force_penalty=pen)
    Fj = np.asfortranarray(F).reshape(-1, order="F")
    gF = np.asfortranarray(g.dF).reshape(-1, order="F")
    dFd = rng.standard_normal(Fj.shape) * 1e-6
    dirs = ((B1, 0.0, None), (0.0, B2, None), (0.0, 0.0, dFd), (0.4 * B1, -0.3 * B2, dFd))
    for a, b, dF in dirs:
        L, dl, _ = oracle.dual_loss_jvp(m.resolved, [p], f0, ticks=T, window=W, F=Fj, dB1=[a], dB2=[b], dF=dF,
                                        target=target, penalty=pen)
        ad = g.dB1[0] * a + g.dB2[0] * b + (0.0 if dF is None else float(np.dot(gF, dF)))
        assert abs(L - g.loss) <= 1e-12 * L
        assert abs(ad - dl) <= 1e-8 * abs(dl), (a, b, ad, dl)
    m32 = walled("f32")
    m32.distributions = f0
    g32 = adjoint.swimming_loss_gradient(m32, ticks=T, checkpoint_every=5, window=W, target_speed=target,
                                         force_penalty=pen)
    gF32 = np.asfortranarray(g32.dF).reshape(-1, order="F")
    for a, b, dF in dirs:
        ad = g.dB1[0] * a + g.dB2[0] * b + (0.0 if dF is None else float(np.dot(gF, dF)))
        ad32 = g32.dB1[0] * a + g32.dB2[0] * b + (0.0 if dF is None else float(np.dot(gF32, dF)))
        assert abs(ad32 - ad) <= 1e-3 * abs(ad), (a, b, ad32, ad)


def test_checkpoint_interval_does_not_change_the_gradient(gpu):
    from lbmengine_hip import adjoint
    F, pos, e = setup(20, 4.5, 9)
    res = []
    for K in (7, 30):
        m = make(gpu, "f64", 20, 4.5, 1e-3, -1e-3, F, pos, e)
        res.append(adjoint.swimming_loss_gradient(m, ticks=30, checkpoint_every=K, window=8))
    a, b = res
    assert a.loss == b.loss
    assert np.array_equal(a.dB1, b.dB1) and np.array_equal(a.dB2, b.dB2)
    assert np.array_equal(a.dF, b.dF)


def test_autograd_wrapper_matches_the_adjoint(gpu):
    import torch
    from lbmengine_hip import adjoint
    n, r = 20, 4.5
    F, pos, e = setup(n, r, 11)

    def make_model(B1, B2, Fa):
        return make(gpu, "f64", n, r, float(B1[0]), float(B2[0]), Fa, pos, e)

    B1 = torch.tensor([1e-3], dtype=torch.float64, requires_grad=True)
    B2 = torch.tensor([5e-4], dtype=torch.float64, requires_grad=True)
    Ft = torch.tensor(F, requires_grad=True)
    L = adjoint.swimming_speed_loss(B1, B2, Ft, make_model, ticks=20, checkpoint_every=8, window=5,
                                    force_penalty=10.0)
    (3.0 * L).backward()
    ref = adjoint.swimming_loss_gradient(make_model(B1.detach(), B2.detach(), F), ticks=20, checkpoint_every=8,
                                         window=5, force_penalty=10.0)
    assert L.item() == ref.loss
    assert B1.grad.item() == 3.0 * ref.dB1[0] and B2.grad.item() == 3.0 * ref.dB2[0]
    assert np.array_equal(Ft.grad.numpy(), 3.0 * ref.dF)


def test_trajectory_log_after_adjoint(gpu):
    """The adjoint's forward run logs into its own ring: afterwards the
    handle's log refuses the ticks it did not see and serves later ones."""
    import ctypes as C

    from lbmengine_hip import _capi
    n, r = 24, 4.0
    F, pos, e = setup(n, r, 5)
    m = make(gpu, "f32", n, r, 1e-3, 0.0, F, pos, e)
    fn, h = m._fn, m._h
    _capi.check(fn["lbm_set_trajectory_log"](h, 64))
    m.run(10)
    from lbmengine_hip import adjoint
    adjoint.swimming_loss_gradient(m, ticks=20, checkpoint_every=10, window=5, target_speed=1e-4, force_penalty=0.0)
    assert m.tick == 30
    buf = np.zeros((5, 9))
    rc = fn["lbm_get_trajectory"](h, 0, 25, 5, _capi.as_dptr(buf))  # ticks inside the adjoint's span
    assert rc == _capi.LBM_EINVAL
    m.run(4)
    m.sync()
    out = np.zeros((4, 9))
    _capi.check(fn["lbm_get_trajectory"](h, 0, 31, 4, _capi.as_dptr(out)))
    assert np.allclose(out[-1, :3], m.particles[0].position, rtol=0, atol=0)
    assert np.allclose(out[-1, 3:6], m.particles[0].velocity, rtol=0, atol=0)
