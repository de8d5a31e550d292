"""Plan descriptions shared by the plan tests: the configuration models re-expressed
as general static execution plans, and non-configuration models that exercise every
family, transform and expression op of the plan format."""
import numpy as np

import uncertaintea_amd as ut
from uncertaintea_amd import plan as PL
from uncertaintea_amd import synthetic as S


def hier_plan(num_groups=8, per_group=8, seed=3, generic=False):
    y, off, sig, _ = S.hierarchical_data(num_groups, per_group, seed=seed)
    J = len(sig)
    grp = np.repeat(np.arange(J), np.diff(off))
    b = ut.PlanBuilder()
    mu = b.latent("normal", 0.0, 5.0, name="mu")
    tau = b.latent("halfcauchy", 5.0, transform="log", name="tau")
    th = b.latent_vector(J, "normal", 0.0, 1.0, name="theta_raw")
    with b.loop(np.column_stack([grp, sig[grp], y]), name="y") as it:
        it.observe("normal", mu + tau * th[it.col(0)], it.col(1), value=it.col(2))
    return b.build("hier_plan", generic=generic), ut.HierarchicalNormalNC(y, off, sig)


def logistic_plan(n=1000, d=25, seed=20261203, generic=False, cfg=1, x_scale=1.0, matrix_cores=False):
    x, y, _ = S.logistic_data(n, d, seed=seed, cfg=cfg, x_scale=x_scale)
    b = ut.PlanBuilder()
    beta = b.latent_vector(d, "normal", 0.0, 1.0, name="beta")
    with b.loop(np.column_stack([x.astype(np.float64), y]), name="rows") as it:
        it.glm(beta, 0, d)
    return b.build("logistic_plan", generic=generic, matrix_cores=matrix_cores), (x, y)


def funnel_plan(dim=100, v_scale=3.0, generic=False):
    b = ut.PlanBuilder()
    v = b.latent("normal", 0.0, v_scale, name="v")
    b.latent_vector(dim, "normal", 0.0, PL.exp(v * 0.5), name="x")
    return b.build("funnel_plan", generic=generic), ut.NealsFunnel(dim, v_scale)


def eight_schools_centered():
    """Centred eight schools: theta_j ~ normal(mu, tau) as a vector latent whose
    scale is a latent (DIAG_NORMAL with slot arguments), y_j read through a data
    index (PARAM with an OBS offset)."""
    y = np.array([28.0, 8.0, -3.0, 7.0, -1.0, 1.0, 18.0, 12.0])
    sigma = np.array([15.0, 10.0, 16.0, 11.0, 9.0, 11.0, 10.0, 18.0])
    b = ut.PlanBuilder()
    mu = b.latent("normal", 0.0, 5.0, name="mu")
    tau = b.latent("halfcauchy", 5.0, transform="log", name="tau")
    theta = b.latent_vector(8, "normal", mu, tau, name="theta")
    with b.loop(np.column_stack([np.arange(8), sigma, y]), name="y") as it:
        it.observe("normal", theta[it.col(0)], it.col(1), value=it.col(2))
    return b.build("eight_schools_centered"), (y, sigma)


def linear_regression(n=200, d=3, seed=7):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, d))
    yv = 0.5 + x @ np.array([1.0, -0.5, 0.25])[:d] + 0.3 * rng.normal(size=n)
    b = ut.PlanBuilder()
    alpha = b.latent("normal", 0.0, 10.0, name="alpha")
    beta = b.latent_vector(d, "normal", 0.0, 1.0, name="beta")
    sigma = b.latent("halfnormal", 1.0, transform="log", name="sigma")
    with b.loop(np.column_stack([x, yv]), name="y") as it:
        mu = alpha
        for k in range(d):
            mu = mu + beta[k] * it.col(k)
        it.observe("normal", mu, sigma, value=it.col(d))
    return b.build("linear_regression"), (x, yv)


def kitchen_sink(n=40, seed=11):
    """Every family, transform and op of the plan format."""
    rng = np.random.default_rng(seed)
    x = rng.normal(size=n)
    bern = (rng.uniform(size=n) < 0.4).astype(float)
    cnt = rng.poisson(2.0, size=n).astype(float)
    pos = rng.gamma(2.0, 1.0, size=n)
    heavy = rng.standard_t(4.0, size=n)
    b = ut.PlanBuilder()
    s1 = b.latent("lognormal", 0.0, 0.5, transform="log", name="s1")
    s2 = b.latent("exponential", 2.0, transform="log", name="s2")
    c = b.latent("cauchy", 0.0, 2.5, name="c")
    lap = b.latent("laplace", 0.2, 1.0, name="lap")
    g = b.latent("logistic", 0.0, 1.0, name="g")
    q = b.latent("halfnormal", 1.0, transform="logit", name="q")
    nu = b.latent("gamma", 4.0, 0.5, transform="log", name="nu")
    hc = b.latent("halfcauchy", 1.0, transform="log", name="hc")
    w = b.latent_vector(3, "normal", 0.0, 1.0, name="w")
    det = b.deterministic(PL.sqrt(s1 * s1 + 1.0) - PL.log1p(s2))
    b.observe("normal", det, 2.0, value=0.3)
    b.observe("halfcauchy", hc, value=0.7)
    b.observe("bernoulli", q, value=1.0)
    b.observe("normal", w[0], PL.exp(w[1] * 0.5), value=np.array([0.1, -0.2, 0.4]))
    with b.loop(np.column_stack([x, bern, cnt, pos, heavy]), name="obs") as it:
        eta = it.deterministic(g + lap * it.col(0) - PL.tanh(c) / 4.0)
        it.observe("bernoullilogit", eta, value=it.col(1))
        it.observe("bernoulli", PL.logistic(c * 0.1 + w[2] * it.col(0)), value=it.col(1))
        it.observe("poisson", PL.exp(0.5 + 0.2 * it.col(0) - PL.abs_(w[0]) * 0.1), value=it.col(2))
        it.observe("gamma", s1 + 1.0, s2 ** 0.5, value=it.col(3))
        it.observe("studentt", nu + 1.0, -w[1], hc + 0.5, value=it.col(4))
        it.observe("laplace", lap, 1.5, value=it.col(0))
        it.observe("cauchy", c, s1, value=it.col(4))
        it.observe("logistic", g, 2.0, value=it.col(0))
        it.observe("lognormal", -0.1, s2, value=it.col(3))
        it.observe("exponential", s2, value=it.col(3))
        it.observe("halfnormal", hc, value=it.col(3))
    return b.build("kitchen_sink"), dict(x=x, bern=bern, cnt=cnt, pos=pos, heavy=heavy)


def kitchen_sink_more(n=40, seed=12):
    """The families beyond the first kitchen sink: gumbel, frechet, rayleigh,
    inversegaussian, beta, inversegamma and weibull as latents (identity, log and logit
    transforms) and observed, and the count families binomial, geometric and
    negativebinomial observed in a loop and at top level."""
    rng = np.random.default_rng(seed)
    x = rng.normal(size=n)
    pos = rng.gamma(2.0, 1.0, size=n)
    unit = rng.beta(2.0, 2.0, size=n)
    trials = rng.integers(1, 11, size=n).astype(float)
    succ = rng.binomial(trials.astype(int), 0.4).astype(float)
    cnt = rng.poisson(2.0, size=n).astype(float)
    b = ut.PlanBuilder()
    gb = b.latent("gumbel", 0.3, 1.2, name="gb")
    fr = b.latent("frechet", 3.0, 1.5, transform="log", name="fr")
    ry = b.latent("rayleigh", 1.5, transform="log", name="ry")
    ig = b.latent("inversegaussian", 1.0, 2.0, transform="log", name="ig")
    bt = b.latent("beta", 2.0, 3.0, transform="logit", name="bt")
    ivg = b.latent("inversegamma", 3.0, 2.0, transform="log", name="ivg")
    wb = b.latent("weibull", 2.0, 1.5, transform="log", name="wb")
    b.observe("binomial", 12.0, bt, value=5.0)
    b.observe("geometric", PL.logistic(gb), value=0.0)
    b.observe("negativebinomial", 3.0, bt, value=4.0)
    with b.loop(np.column_stack([x, pos, unit, trials, succ, cnt]), name="obs") as it:
        it.observe("gumbel", gb, ry + 1.0, value=it.col(0))
        it.observe("frechet", fr + 1.0, ig, value=it.col(1))
        it.observe("rayleigh", wb, value=it.col(1))
        it.observe("inversegaussian", ivg, fr * 2.0, value=it.col(1))
        it.observe("beta", bt * 4.0 + 0.5, ry + 1.0, value=it.col(2))
        it.observe("inversegamma", ig + 1.0, wb, value=it.col(1))
        it.observe("weibull", fr, ivg, value=it.col(1))
        it.observe("binomial", it.col(3), PL.logistic(gb * 0.3 + it.col(0) * 0.2), value=it.col(4))
        it.observe("geometric", bt, value=it.col(5))
        it.observe("negativebinomial", ry + 0.5, bt, value=it.col(5))
    return b.build("kitchen_sink_more"), dict(x=x, pos=pos, unit=unit, trials=trials, succ=succ, cnt=cnt)


def glm_large(n=8192, d=8, seed=5):
    """A logistic regression with an intercept (not the fused family): N rows for the
    observation-strided interpreter under NUTS."""
    x, y, _ = S.logistic_data(n, d, seed=seed, cfg=1)
    b = ut.PlanBuilder()
    a = b.latent("normal", 0.0, 2.0, name="alpha")
    beta = b.latent_vector(d, "normal", 0.0, 1.0, name="beta")
    with b.loop(np.column_stack([x.astype(np.float64), y]), name="rows") as it:
        it.glm(beta, 0, d, intercept=a)
    return b.build("glm_intercept"), (x, y)
