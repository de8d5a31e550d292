"""Structural pins of the oracle's sampler pieces against the reference's own tests and
against brute force: the checkpoint U-turn test against the recursive tree build, the
window-end step search, the accept-probability branch of the pooled mass weights, the
MEADS slice drift and the conjugate-Gaussian NUTS gate."""
import math

import numpy as np
import pytest

import uncertaintea_amd as ut
from oracle import binding as B


def _gauss2():
    # x ~ normal(0, 1), y ~ normal(0, 3): the model of nuts_uturn_turning.jl:4-8
    b = ut.PlanBuilder()
    b.latent("normal", 0.0, 1.0, name="x")
    b.latent("normal", 0.0, 3.0, name="y")
    return B.OracleModel(b.build("dyadic_model"))


def _is_turning(lq, rq, lp, rp, im):
    # state_init.jl:309-320 with a diagonal inverse mass
    d = rq - lq
    return float(np.sum(d * im * lp)) <= 0.0 or float(np.sum(d * im * rp)) <= 0.0


def _recursive_tree(tq, tp, im, lo, depth, direction):
    """nuts_uturn_turning.jl:25-55 over precomputed leapfrog leaves: leaves lo .. lo + 2^depth - 1
    in integration order. Returns (first, last, turning, steps)."""
    if depth == 0:
        return lo, lo, False, 1
    l0, l1, lt, ls = _recursive_tree(tq, tp, im, lo, depth - 1, direction)
    if lt:
        return l0, l1, True, ls
    r0, r1, rt, rs = _recursive_tree(tq, tp, im, l1 + 1, depth - 1, direction)
    steps = ls + rs
    if rt:
        return l0, r1, True, steps
    if direction > 0:
        turned = _is_turning(tq[l0], tq[r1], tp[l0], tp[r1], im)
    else:
        turned = _is_turning(tq[r1], tq[l0], tp[r1], tp[l0], im)
    return l0, r1, turned, steps


def _old_criterion(q0, p0, tq, tp, im, direction):
    # nuts_uturn_turning.jl:57-82: start point against the current leaf
    for i in range(len(tq)):
        if direction > 0:
            t = _is_turning(q0, tq[i], p0, tp[i], im)
        else:
            t = _is_turning(tq[i], q0, tp[i], p0, im)
        if t:
            return True, i + 1
    return False, len(tq)


def test_checkpoint_uturn_matches_recursive_tree():
    """nuts_uturn_turning.jl:84-150: over step sizes 0.3..1.5 and depths 2..4 the checkpoint
    subtree and the recursive build agree on turning and integration steps, at least four
    subtrees turn, and at least one differs from the start-vs-current criterion."""
    om = _gauss2()
    q0, p0, im = np.array([1.6, -1.2]), np.array([0.5, 0.9]), np.ones(2)
    turning_count = diff_from_old = 0
    for step in (0.3, 0.6, 0.9, 1.2, 1.5):
        for depth in (2, 3, 4):
            steps, turning, div, tq, tp = om.nuts_subtree(q0, p0, im, step, +1, depth)
            assert not div
            _, _, rt, rs = _recursive_tree(tq, tp, im, 0, depth, +1)
            assert (turning, steps) == (rt, rs), (step, depth)
            turning_count += turning
            diff_from_old += _old_criterion(q0, p0, tq, tp, im, +1) != (turning, steps)
    assert turning_count >= 4
    assert diff_from_old >= 1


@pytest.mark.parametrize("direction", [+1, -1])
def test_checkpoint_uturn_brute_force_random(direction):
    """Random starts, metrics, directions and depths 1..6 on the 2-D model: the checkpoint
    slots (popcount of the block start) and the orientation of the test reproduce the
    recursive build exactly, including subtrees whose inner sub-subtrees turn first."""
    om = _gauss2()
    rng = np.random.default_rng(20261204 + direction)
    inner_turns = 0
    for _ in range(300):
        q0 = rng.normal(size=2) * np.array([1.0, 3.0])
        p0 = rng.normal(size=2)
        im = np.exp(rng.uniform(-1.0, 1.0, size=2))
        step = rng.uniform(0.05, 1.2)
        depth = int(rng.integers(1, 7))
        steps, turning, div, tq, tp = om.nuts_subtree(q0, p0, im, step, direction, depth)
        if div:
            continue
        _, _, rt, rs = _recursive_tree(tq, tp, im, 0, depth, direction)
        assert (turning, steps) == (rt, rs)
        inner_turns += turning and steps < (1 << depth)
    assert inner_turns >= 20


# ------------------------------------------------------------------ step-size search
def _conj(y=0.3):
    # mu ~ normal(0, 1), y ~ normal(mu, 1): warmup_driver_regression.jl:2-8
    b = ut.PlanBuilder()
    mu = b.latent("normal", 0.0, 1.0, name="mu")
    b.observe("normal", mu, 1.0, value=y, name="y")
    return b.build("warmup_driver_model")


def _two_sum(y=0.4):
    # a, b ~ normal(0, 1), y ~ normal(a + b, 1): warmup_driver_regression.jl:148-153
    b = ut.PlanBuilder()
    a = b.latent("normal", 0.0, 1.0, name="a")
    c = b.latent("normal", 0.0, 1.0, name="b")
    b.observe("normal", a + c, 1.0, value=y, name="y")
    return b.build("perchain_stepsearch_model")


def _scalar_search(om, q, im, chain, itc, s0):
    """adaptation.jl:516-571 restated for one chain: momentum N(0, 1) / sqrt(im) from the
    search stream, one leapfrog step per trial, first crossing of log 0.5."""
    P = len(q)
    _, _, z, _ = B.philox_fill(chain, itc, 4, 0, P)  # stream 4: step search
    p0 = z.astype(np.float64) / np.sqrt(im)

    def lj_grad(x):
        lj, g = om.logjoint_grad_f64(x.reshape(P, 1))
        return float(lj[0]), g[:, 0]

    lj0, g0 = lj_grad(q)
    H0 = 0.5 * float(np.sum(p0 * p0 * im)) - lj0

    def lar(h):
        p = p0 + 0.5 * h * g0
        x = q + h * im * p
        lj1, g1 = lj_grad(x)
        p = p + 0.5 * h * g1
        return H0 - (0.5 * float(np.sum(p * p * im)) - lj1)

    s = s0
    direction = 1.0 if lar(s) > math.log(0.5) else -1.0
    for _ in range(20):
        nxt = s * 2.0 ** direction
        if nxt < 1e-8 or nxt > 1e3:
            break
        s = nxt
        a = lar(s)
        if (direction > 0 and a <= math.log(0.5)) or (direction < 0 and a >= math.log(0.5)):
            break
    return min(max(s, 1e-8), 1e3)


def test_step_search_first_crossing():
    """warmup_driver_regression.jl:68-136 per chain: searches from s0 and 2 s0 (same momentum)
    stop at the same first crossing, by halving from above and by doubling from below, and the
    two directions bracket each other within the doubling grid."""
    om = B.OracleModel(_conj())
    rng = np.random.default_rng(1)
    q = rng.normal(size=(1, 8))
    im = np.ones(1)
    down = om.step_search(q, im, 10.0, 44)
    down2 = om.step_search(q, im, 20.0, 44)
    up = om.step_search(q, im, 1e-4, 44)
    up2 = om.step_search(q, im, 2e-4, 44)
    assert np.array_equal(down, down2) and np.array_equal(up, up2)
    for d, u in zip(down, up):
        assert d < 10.0 and u > 1e-4
        assert float(np.log2(10.0 / d)).is_integer()
        assert float(np.log2(20.0 / d)).is_integer()
        assert float(np.log2(u / 1e-4)).is_integer()
        assert u / 4 <= d <= u * 4


def test_step_search_per_chain_equals_scalar_search():
    """warmup_driver_regression.jl:139-240: the per-chain search over a batch reproduces C
    independent scalar searches, from one anchor and from per-chain anchors (the window-end
    re-search form), at several Philox coordinates and non-unit metrics."""
    om = B.OracleModel(_two_sum())
    q = np.random.default_rng(7).normal(size=(2, 4))
    for itc, anchors, im in ((2158, np.full(4, 0.7), np.ones(2)),
                             (9931, np.array([0.3, 0.5, 0.9, 1.4]), np.ones(2)),
                             (77, np.array([5.0, 1e-3, 0.05, 40.0]), np.array([0.4, 2.5]))):
        batched = om.step_search(q, im, anchors, itc)
        scalar = [_scalar_search(om, q[:, c].copy(), im, c, itc, anchors[c]) for c in range(4)]
        assert np.array_equal(batched, np.array(scalar)), (itc, batched, scalar)
    # the bounds: an anchor at the top of the range cannot double past 1e3
    s = om.step_search(q[:, :1], np.ones(2), 1e3, 5)
    assert s[0] <= 1e3


# ------------------------------------------------------------ pooled mass weights
def test_mass_weight_accept_probability_ramp_hand_computed():
    """adaptation.jl:225-240 with data_types.jl:189-192: a rejected NUTS draw weighs
    1 + (clamp(accept, 0, 1) - 1) * progress, progress = max(count - 4, 0) / (window - 4)
    counting only earlier non-zero-weight updates; divergent draws weigh 0."""
    x = np.array([0.3, -0.2, 0.5, 0.1, -0.4, 0.2, 0.0, 0.35, -0.15, 0.25])
    acc = np.full(10, 0.4)
    flags = np.zeros(10, np.uint8)
    flags[6] = 1  # divergent: weight 0, count does not advance
    # counts before each update: 0 1 2 3 4 5 | 6 | 6 7 8 -> progress 0 0 0 0 0 1/6 | - | 2/6 3/6 4/6
    w = np.array([1.0, 1.0, 1.0, 1.0, 1.0, 0.9, 0.0, 0.8, 0.7, 0.6])
    im, eff = B.running_variance(x.reshape(10, 1, 1), acc.reshape(10, 1), flags.reshape(10, 1),
                                 window_length=10, use_accepted=False)
    ws, wq = w.sum(), (w * w).sum()
    mean = (w * x).sum() / ws
    m2 = (w * (x - mean) ** 2).sum()
    e = ws * ws / wq
    shrink = e / (e + 5.0)
    assert eff == pytest.approx(e, rel=1e-13)
    assert im[0] == pytest.approx(shrink * m2 / (ws - wq / ws) + (1 - shrink), rel=1e-12)
    # accept probabilities are clamped to [0, 1] and non-finite ones weigh 0
    acc2 = acc.copy()
    acc2[7], acc2[8], acc2[9] = 1.7, -0.3, np.nan
    w2 = np.array([1.0, 1.0, 1.0, 1.0, 1.0, 0.9, 0.0, 1.0, 1.0 - 3 / 6, 0.0])
    im2, eff2 = B.running_variance(x.reshape(10, 1, 1), acc2.reshape(10, 1), flags.reshape(10, 1),
                                   window_length=10, use_accepted=False)
    assert eff2 == pytest.approx(w2.sum() ** 2 / (w2 * w2).sum(), rel=1e-13)
    # a window of at most 4 draws has progress 1: the weight is the accept probability
    im3, eff3 = B.running_variance(x[:4].reshape(4, 1, 1), acc[:4].reshape(4, 1), flags[:4].reshape(4, 1),
                                   window_length=4, use_accepted=False)
    assert eff3 == pytest.approx((4 * 0.4) ** 2 / (4 * 0.16), rel=1e-13)


# --------------------------------------------------------- conjugate-Gaussian gate
def _two_param(y=0.7):
    # mu ~ normal(0, 1), log_sigma ~ normal(0, 0.5), y ~ normal(mu, exp(log_sigma))
    # (device_persistent_nuts.jl:44-49)
    b = ut.PlanBuilder()
    mu = b.latent("normal", 0.0, 1.0, name="mu")
    ls = b.latent("normal", 0.0, 0.5, name="log_sigma")
    b.observe("normal", mu, ut.plan.exp(ls), value=y, name="y")
    return b.build("dpnuts_two_param")


def test_conjugate_gaussian_gate():
    """device_persistent_nuts.jl:57-72: 64 chains, 500 warmup + 1000 draws of the conjugate
    Gaussian at y = 0.3: posterior mean 0.15 and sd sqrt(0.5) within 0.03, R-hat < 1.01, no
    divergences; the binary32 kernel-order twin passes the same gate."""
    from uncertaintea_amd import diagnostics as D
    om = B.OracleModel(_conj())
    for mode in ("f64", "twin"):
        r = om.run("nuts", mode, 64, num_samples=1000, num_warmup=500, seed=511, threads=8)
        d = r["samples"][..., 0].astype(np.float64)
        assert np.all(np.isfinite(d))
        assert abs(d.mean() - 0.15) < 0.03, mode
        assert abs(d.std(ddof=1) - math.sqrt(0.5)) < 0.03, mode
        assert float(np.max(np.asarray(D.rhat(r["samples"])))) < 1.01
        assert not r["divergent"].any()


def test_two_parameter_deep_trees():
    """device_persistent_nuts.jl:99-120: the location / log-scale model reaches depth >= 3,
    stays (nearly) divergence-free with R-hat < 1.01, and the twin agrees with binary64."""
    from uncertaintea_amd import diagnostics as D
    om = B.OracleModel(_two_param())
    a = om.run("nuts", "f64", 64, num_samples=1000, num_warmup=500, seed=2, threads=8)
    b = om.run("nuts", "twin", 64, num_samples=1000, num_warmup=500, seed=3, threads=8)
    for r in (a, b):
        assert np.all(np.isfinite(r["samples"]))
        assert float(np.max(np.asarray(D.rhat(r["samples"])))) < 1.01
        assert r["divergent"].mean() < 0.01
        assert r["tree_depth"].max() >= 3
    for k in range(2):
        da, db = a["samples"][..., k].astype(np.float64), b["samples"][..., k].astype(np.float64)
        assert abs(da.mean() - db.mean()) < 0.03
        assert abs(da.std() - db.std()) < 0.03


# ----------------------------------------------------------- float32 saturation edges
GUARDED = {  # device_float32_saturation.jl:36-66: the seven positive families with log(x) / 1 / x terms
    "gamma": (2.0, 1.0), "lognormal": (0.0, 1.0), "weibull": (2.0, 1.5), "inversegamma": (3.0, 2.0),
    "frechet": (2.0, 1.5), "rayleigh": (1.5,), "inversegaussian": (1.0, 2.0),
}


def _single(family, args):
    b = ut.PlanBuilder()
    b.latent(family, *args, transform="log", name="x")
    return B.OracleModel(b.build(f"dfs_{family}"))


def _beta_bern():
    b = ut.PlanBuilder()
    p = b.latent("beta", 2.0, 2.0, transform="logit", name="p")
    b.observe("bernoulli", p, value=1.0, name="y")
    return B.OracleModel(b.build("dfs_beta_bern"))


def _point(om, theta, precision):
    q = np.array([[theta]], dtype=precision)
    v, g = om.logjoint_grad_twin(q, precision=precision)
    return float(v[0]), float(g[0, 0])


def _truth(om, theta):
    v, g = om.logjoint_grad_f64(np.array([[theta]]))
    return float(v[0]), float(g[0, 0])


@pytest.mark.parametrize("family", sorted(GUARDED))
def test_float32_exp_subnormal_band_rejects_on_the_value(family):
    """device_float32_saturation.jl:78-86: in the binary32 subnormal band of exp(theta) and past
    underflow the guarded families score -Inf with a finite (Jacobian-only) gradient."""
    om = _single(family, GUARDED[family])
    for theta in (-90.0, -100.0, -120.0):
        v, g = _point(om, theta, np.float32)
        assert v == -np.inf and not np.isnan(g), (family, theta, v, g)
        assert g == 1.0  # d theta / d theta of the log transform's log-abs-det


def test_float32_just_outside_the_band_tracks_binary64():
    """device_float32_saturation.jl:88-98: at theta = -80 exp(theta) is a normal binary32 number
    and the values and gradients track the binary64 truth to 1e-4; the binary64 evaluator itself
    equals scipy's log density plus the log-abs-det."""
    from scipy import stats
    sp = {"gamma": lambda x: stats.gamma.logpdf(x, 2.0, scale=1.0),
          "lognormal": lambda x: stats.lognorm.logpdf(x, 1.0),
          "weibull": lambda x: stats.weibull_min.logpdf(x, 2.0, scale=1.5),
          "rayleigh": lambda x: stats.rayleigh.logpdf(x, scale=1.5)}
    for family, f in sp.items():
        om = _single(family, GUARDED[family])
        v, g = _point(om, -80.0, np.float32)
        vr, gr = _truth(om, -80.0)
        assert vr == pytest.approx(f(math.exp(-80.0)) - 80.0, rel=1e-12)
        assert np.isfinite(v)
        assert v == pytest.approx(vr, rel=1e-4) and g == pytest.approx(gr, rel=1e-4), family


def test_exponential_is_exact_through_the_band():
    """device_float32_saturation.jl:100-106: exponential has no log(x) term and stays finite
    through the band and past underflow, gradient 1 - rate exp(theta) = 1."""
    om = _single("exponential", (1.5,))
    for theta in (-90.0, -120.0):
        v, g = _point(om, theta, np.float32)
        assert np.isfinite(v) and abs(g - 1.0) <= 1e-6


def test_binary64_boundary_follows_the_value_precision():
    """device_float32_saturation.jl:108-117: in binary64 theta = -100 is far from the boundary
    (finite, gradient to 1e-9 of the reference), while the binary64 subnormal band rejects."""
    for family in ("gamma", "lognormal"):
        om = _single(family, GUARDED[family])
        v, g = _point(om, -100.0, np.float64)
        _, gr = _truth(om, -100.0)
        assert np.isfinite(v) and g == pytest.approx(gr, rel=1e-9)
        vb, gb = _point(om, -720.0, np.float64)
        assert vb == -np.inf and not np.isnan(gb)
        assert _truth(om, -720.0)[0] == -np.inf


def test_float32_logit_cliff():
    """device_float32_saturation.jl:120-160: past the binary32 logistic cliff (theta >= 18) the
    beta latent's value rounds to 1 and scores -Inf with a finite gradient; the pre-cliff band
    stays within 4 eps32 exp(theta) + 1e-5 of binary64; the normal range within 1e-5."""
    om = _beta_bern()
    for theta in (18.0, 20.0, 30.0):
        v, g = _point(om, theta, np.float32)
        assert v == -np.inf and not np.isnan(g)
    for theta in (13.0, 15.0, 16.0):
        v, g = _point(om, theta, np.float32)
        vr, gr = _truth(om, theta)
        tol = 4 * float(np.finfo(np.float32).eps) * math.exp(theta) + 1e-5
        assert np.isfinite(v) and abs(v - vr) <= tol
        assert np.isfinite(g) and abs(g - gr) <= tol
    for theta in (-1.5, 0.3, 2.0):
        v, g = _point(om, theta, np.float32)
        vr, gr = _truth(om, theta)
        assert abs(v - vr) <= 1e-5 and abs(g - gr) <= 1e-5
