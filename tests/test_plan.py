"""CPU: the general static execution plan (ut_plan_desc) -- its host audit through the
C ABI, and the oracle's two restatements of the plan interpreter.

Pins of the oracle (test infrastructure):
* the binary64 plan evaluator (reference semantics, forward-mode dual gradients)
  equals the dedicated binary64 oracles of the configuration models it re-expresses
  (hierarchical, logistic, funnel) to 1e-12, and on non-configuration models equals
  scipy.stats log densities (every family, transform and op), with gradients
  against central finite differences;
* the kernel-order twin (reverse-mode, the device's partials) equals that evaluator
  in binary64 to 1e-11 (1e-13 without Student-t / gamma shapes, whose dual
  derivative of log Gamma is a finite difference) and in binary32 to 1e-6;
* the prior draw of a plan equals the dedicated models' initial columns.
Audit: every rule of the reference's lowering the plan format keeps is rejected with
the reference's error kind."""
import ctypes as ct
import sys
from pathlib import Path

import numpy as np
import pytest
from scipy import stats

sys.path.insert(0, str(Path(__file__).resolve().parent))
import plan_models as PMS  # noqa: E402

import uncertaintea_amd as ut  # noqa: E402
from oracle import binding as B  # noqa: E402
from uncertaintea_amd import _lib as L  # noqa: E402
from uncertaintea_amd import plan as PL  # noqa: E402


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(1e-300, float(np.max(np.abs(b)))))


def _fd(o, q, h=1e-6):
    fd = np.zeros_like(q)
    for r in range(q.shape[0]):
        e = np.zeros_like(q)
        e[r] = h
        fd[r] = (o.logjoint_grad_f64(q + e)[0] - o.logjoint_grad_f64(q - e)[0]) / (2 * h)
    return fd


@pytest.mark.parametrize("which", ["hier", "logistic", "funnel"])
def test_plan_oracle_equals_dedicated_model_oracles(which):
    if which == "hier":
        pm, dm = PMS.hier_plan()
    elif which == "logistic":
        pm, (x, y) = PMS.logistic_plan(n=300, d=7)
        dm = ut.LogisticRegression(x, y)
    else:
        pm, dm = PMS.funnel_plan(dim=20)
    op, od = B.OracleModel(pm), B.OracleModel(dm)
    q = np.random.default_rng(4).normal(size=(pm.num_params, 12)) * 0.5
    l1, g1 = op.logjoint_grad_f64(q)
    l2, g2 = od.logjoint_grad_f64(q)
    assert _rel(l1, l2) <= 1e-12 and _rel(g1, g2) <= 1e-12
    lt, gt = op.logjoint_grad_twin(q, precision=np.float64)
    assert _rel(lt, l1) <= 1e-13 and _rel(gt, g1) <= 1e-13
    l32, g32 = op.logjoint_grad_twin(q.astype(np.float32))
    assert _rel(l32, l1) <= 1e-6 and _rel(g32, g1) <= 1e-6


def _kitchen_sink_scipy(q, d):
    """The kitchen-sink plan's logjoint written with scipy.stats (unconstrained q)."""
    s1, s2, c, lap, g, qq, nu, hc = np.exp(q[0]), np.exp(q[1]), q[2], q[3], q[4], 1 / (1 + np.exp(-q[5])), \
        np.exp(q[6]), np.exp(q[7])
    w = q[8:11]
    lp = stats.lognorm.logpdf(s1, 0.5, scale=1.0) + q[0]
    lp += stats.expon.logpdf(s2, scale=0.5) + q[1]
    lp += stats.cauchy.logpdf(c, 0.0, 2.5) + stats.laplace.logpdf(lap, 0.2, 1.0) + stats.logistic.logpdf(g, 0, 1)
    lp += stats.halfnorm.logpdf(qq, scale=1.0) + np.log(qq * (1 - qq))
    lp += stats.gamma.logpdf(nu, 4.0, scale=2.0) + q[6]
    lp += stats.halfcauchy.logpdf(hc, scale=1.0) + q[7]
    lp += stats.norm.logpdf(w, 0, 1).sum()
    det = np.sqrt(s1 * s1 + 1.0) - np.log1p(s2)
    lp += stats.norm.logpdf(0.3, det, 2.0) + stats.halfcauchy.logpdf(0.7, scale=hc)
    lp += np.log(qq)
    lp += stats.norm.logpdf([0.1, -0.2, 0.4], w[0], np.exp(w[1] * 0.5)).sum()
    x, bern, cnt, pos, heavy = d["x"], d["bern"], d["cnt"], d["pos"], d["heavy"]
    eta = g + lap * x - np.tanh(c) / 4.0
    lp += np.sum(bern * eta - np.logaddexp(0, eta))
    p = 1 / (1 + np.exp(-(c * 0.1 + w[2] * x)))
    lp += stats.bernoulli.logpmf(bern, p).sum()
    lp += stats.poisson.logpmf(cnt, np.exp(0.5 + 0.2 * x - np.abs(w[0]) * 0.1)).sum()
    lp += stats.gamma.logpdf(pos, s1 + 1.0, scale=1.0 / np.sqrt(s2)).sum()
    lp += stats.t.logpdf(heavy, nu + 1.0, -w[1], hc + 0.5).sum()
    lp += stats.laplace.logpdf(x, lap, 1.5).sum() + stats.cauchy.logpdf(heavy, c, s1).sum()
    lp += stats.logistic.logpdf(x, g, 2.0).sum() + stats.lognorm.logpdf(pos, s2, scale=np.exp(-0.1)).sum()
    lp += stats.expon.logpdf(pos, scale=1.0 / s2).sum() + stats.halfnorm.logpdf(pos, scale=hc).sum()
    return lp


def test_plan_oracle_against_scipy_every_family_transform_and_op():
    pm, d = PMS.kitchen_sink()
    o = B.OracleModel(pm)
    q = np.random.default_rng(2).normal(size=(pm.num_params, 6)) * 0.4
    lj, g = o.logjoint_grad_f64(q)
    ref = np.array([_kitchen_sink_scipy(q[:, c], d) for c in range(q.shape[1])])
    assert _rel(lj, ref) <= 1e-12, (lj, ref)
    assert _rel(g, _fd(o, q)) <= 1e-7
    lt, gt = o.logjoint_grad_twin(q, precision=np.float64)
    assert _rel(lt, lj) <= 1e-12 and _rel(gt, g) <= 1e-10
    l32, g32 = o.logjoint_grad_twin(q.astype(np.float32))
    assert _rel(l32, lj) <= 1e-6 and _rel(g32, g) <= 1e-5


def _kitchen_sink_more_scipy(q, d):
    """kitchen_sink_more's logjoint with scipy.stats (unconstrained q)."""
    gb, fr, ry, ig, ivg, wb = q[0], np.exp(q[1]), np.exp(q[2]), np.exp(q[3]), np.exp(q[5]), np.exp(q[6])
    bt = 1 / (1 + np.exp(-q[4]))
    lp = stats.gumbel_r.logpdf(gb, 0.3, 1.2)
    lp += stats.invweibull.logpdf(fr, 3.0, scale=1.5) + q[1]
    lp += stats.rayleigh.logpdf(ry, scale=1.5) + q[2]
    lp += stats.invgauss.logpdf(ig, 1.0 / 2.0, scale=2.0) + q[3]
    lp += stats.beta.logpdf(bt, 2.0, 3.0) + np.log(bt * (1 - bt))
    lp += stats.invgamma.logpdf(ivg, 3.0, scale=2.0) + q[5]
    lp += stats.weibull_min.logpdf(wb, 2.0, scale=1.5) + q[6]
    lp += stats.binom.logpmf(5, 12, bt) + stats.geom.logpmf(1, 1 / (1 + np.exp(-gb)))
    lp += stats.nbinom.logpmf(4, 3.0, bt)
    x, pos, unit, trials, succ, cnt = (d[k] for k in ("x", "pos", "unit", "trials", "succ", "cnt"))
    lp += stats.gumbel_r.logpdf(x, gb, ry + 1.0).sum()
    lp += stats.invweibull.logpdf(pos, fr + 1.0, scale=ig).sum()
    lp += stats.rayleigh.logpdf(pos, scale=wb).sum()
    lp += stats.invgauss.logpdf(pos, ivg / (fr * 2.0), scale=fr * 2.0).sum()
    lp += stats.beta.logpdf(unit, bt * 4.0 + 0.5, ry + 1.0).sum()
    lp += stats.invgamma.logpdf(pos, ig + 1.0, scale=wb).sum()
    lp += stats.weibull_min.logpdf(pos, fr, scale=ivg).sum()
    lp += stats.binom.logpmf(succ, trials, 1 / (1 + np.exp(-(gb * 0.3 + x * 0.2)))).sum()
    lp += stats.geom.logpmf(cnt + 1, bt).sum() + stats.nbinom.logpmf(cnt, ry + 0.5, bt).sum()
    return lp


def test_plan_oracle_against_scipy_more_families():
    pm, d = PMS.kitchen_sink_more()
    o = B.OracleModel(pm)
    q = np.random.default_rng(5).normal(size=(pm.num_params, 6)) * 0.4
    lj, g = o.logjoint_grad_f64(q)
    ref = np.array([_kitchen_sink_more_scipy(q[:, c], d) for c in range(q.shape[1])])
    assert _rel(lj, ref) <= 1e-11, (lj, ref)
    assert _rel(g, _fd(o, q)) <= 1e-6
    lt, gt = o.logjoint_grad_twin(q, precision=np.float64)
    assert _rel(lt, lj) <= 1e-12 and _rel(gt, g) <= 1e-9
    l32, g32 = o.logjoint_grad_twin(q.astype(np.float32))
    assert _rel(l32, lj) <= 1e-6 and _rel(g32, g) <= 1e-5


@pytest.mark.parametrize("name", ["eight_schools_centered", "linear_regression"])
def test_plan_oracle_against_scipy_regressions(name):
    pm, data = getattr(PMS, name)()
    o = B.OracleModel(pm)
    q = np.random.default_rng(3).normal(size=(pm.num_params, 5)) * 0.4
    lj, g = o.logjoint_grad_f64(q)
    ref = []
    for c in range(q.shape[1]):
        u = q[:, c]
        if name == "eight_schools_centered":
            y, sigma = data
            mu, tau, th = u[0], np.exp(u[1]), u[2:]
            v = stats.norm.logpdf(mu, 0, 5) + stats.halfcauchy.logpdf(tau, scale=5) + u[1]
            v += stats.norm.logpdf(th, mu, tau).sum() + stats.norm.logpdf(y, th, sigma).sum()
        else:
            x, yv = data
            a, beta, s = u[0], u[1:4], np.exp(u[4])
            v = stats.norm.logpdf(a, 0, 10) + stats.norm.logpdf(beta, 0, 1).sum() + stats.halfnorm.logpdf(s) + u[4]
            v += stats.norm.logpdf(yv, a + x @ beta, s).sum()
        ref.append(v)
    assert _rel(lj, ref) <= 1e-12
    assert _rel(g, _fd(o, q)) <= 1e-7
    lt, gt = o.logjoint_grad_twin(q, precision=np.float64)
    assert _rel(lt, lj) <= 1e-13 and _rel(gt, g) <= 1e-13


def test_plan_prior_draw_matches_the_dedicated_models():
    """The plan's ancestral prior draw gives the dedicated models' initial columns:
    mu = 5 z_0, tau = 5 |z_1 / z_{P+1}|, theta_raw_j = z_{2+j} (hierarchical) and
    v = 3 z_0, x_i = e^{v/2} z_{1+i} (funnel), from the same Philox normals."""
    pm, dm = PMS.hier_plan()
    o = B.OracleModel(pm)
    P = pm.num_params
    for chain in (0, 5):
        col = o.prior_column(L.UT_INIT_PRIOR, chain, 77)
        z = np.array([B.philox_fill(chain, 77, 5, 0, 2 * P)[2]]).reshape(-1).astype(np.float64)
        assert col[0] == 5.0 * z[0]
        assert col[1] == np.log(5.0 * abs(z[1] / z[P + 1]))
        assert np.array_equal(col[2:], z[2:P])
    fm, _ = PMS.funnel_plan(dim=6)
    col = B.OracleModel(fm).prior_column(L.UT_INIT_PRIOR, 3, 9)
    z = B.philox_fill(3, 9, 5, 0, 7)[2].astype(np.float64)
    assert col[0] == 3.0 * z[0] and np.allclose(col[1:], np.exp(col[0] / 2) * z[1:7], rtol=1e-15, atol=0)
    u = B.OracleModel(fm).prior_column(L.UT_INIT_UNIFORM, 3, 9)
    assert np.all(np.abs(u) <= 2.0)


# ---------------------------------------------------------------- audit
def _desc(build):
    b = ut.PlanBuilder()
    build(b)
    return b.build("audit")


def _rc(model):
    rc = L.lib().ut_plan_validate(ct.byref(model.desc))
    return rc, L.lib().ut_last_error().decode()


def test_audit_accepts_the_models():
    for f in ("hier_plan", "logistic_plan", "funnel_plan", "eight_schools_centered", "linear_regression",
              "kitchen_sink"):
        m = getattr(PMS, f)()[0]
        assert _rc(m)[0] == L.UT_OK, f
        assert ut.device_lowering_report(m)["lowerable"]


def test_audit_rejections():
    """The reference's lowering rules: unsupported shapes are UT_E_UNSUPPORTED with a
    message that points at device_lowering_report; inconsistent descriptions UT_E_ARG
    or UT_E_DIM."""
    # a slot read before it is written: build a plan then swap its two steps
    m = _desc(lambda b: (b.observe("normal", b.latent("normal", 0.0, 1.0), 1.0, value=0.5)))
    st = m._steps
    st[0], st[1] = L.PlanStep.from_buffer_copy(st[1]), L.PlanStep.from_buffer_copy(st[0])
    rc, msg = _rc(m)
    assert rc == L.UT_E_UNSUPPORTED and "before it is written" in msg and "device_lowering_report" in msg
    # a latent inside a loop body
    with_loop = ut.PlanBuilder()
    a = with_loop.latent("normal", 0.0, 1.0)
    with with_loop.loop(np.ones((3, 1))) as it:
        it.observe("normal", a, 1.0, value=it.col(0))
    lm = with_loop.build("loop")
    assert _rc(lm)[0] == L.UT_OK
    lm._steps[2].value_source = 1
    rc, msg = _rc(lm)
    assert rc == L.UT_E_UNSUPPORTED and "latent choices inside loops" in msg
    # a loop-local value read after its loop
    b = ut.PlanBuilder()
    a = b.latent("normal", 0.0, 1.0)
    with b.loop(np.ones((2, 1))) as it:
        t = it.deterministic(a * it.col(0))
    b.observe("normal", t, 1.0, value=0.0)
    rc, msg = _rc(b.build("escape"))
    assert rc == L.UT_E_UNSUPPORTED and "do not escape" in msg
    # staged rows that the steps do not consume
    m = PMS.linear_regression(n=10)[0]
    m.desc.num_observed = m.desc.num_observed - 1
    rc, msg = _rc(m)
    assert rc == L.UT_E_DIM and "staged rows" in msg
    # a parameter row owned by no latent
    m = PMS.linear_regression(n=10)[0]
    m.desc.num_params = m.desc.num_params + 1
    rc, msg = _rc(m)
    assert rc == L.UT_E_ARG and "belongs to no latent" in msg
    # a data-indexed row out of range at some iteration
    b = ut.PlanBuilder()
    v = b.latent_vector(3, "normal", 0.0, 1.0)
    with b.loop(np.array([[0.0, 1.0], [3.0, 0.5]])) as it:
        it.observe("normal", v[it.col(0)], 1.0, value=it.col(1))
    rc, msg = _rc(b.build("oob"))
    assert rc == L.UT_E_ARG and "not a valid row" in msg
    # an OBS row past the iteration's block
    b = ut.PlanBuilder()
    a = b.latent("normal", 0.0, 1.0)
    with b.loop(np.ones((3, 2))) as it:
        it.observe("normal", a + it.col(1), 1.0, value=it.col(0))
    m = b.build("obs_row")
    assert _rc(m)[0] == L.UT_OK
    obs = [i for i in range(m.desc.num_exprs) if m._exprs[i].op == L.UT_OP_OBS and m._exprs[i].index == 1]
    m._exprs[obs[0]].index = 2
    rc, msg = _rc(m)
    assert rc == L.UT_E_ARG and "outside the iteration's block" in msg
    # nested loops
    m = PMS.hier_plan()[0]
    m._steps[4].kind = L.UT_STEP_LOOP
    rc, msg = _rc(m)
    assert rc == L.UT_E_UNSUPPORTED and "nested loops" in msg
    # a discrete latent
    b = ut.PlanBuilder()
    b.latent("poisson", 2.0)
    rc, msg = _rc(b.build("discrete"))
    assert rc == L.UT_E_UNSUPPORTED and "discrete latents" in msg
    # OBS outside a loop body, and through the Python layer's report
    b = ut.PlanBuilder()
    b.observe("normal", PL.Expr(L.UT_OP_OBS, index=0), 1.0, value=0.0)
    b.latent("normal", 0.0, 1.0)
    m = b.build("obs")
    rc, msg = _rc(m)
    assert rc == L.UT_E_UNSUPPORTED and "OBS nodes" in msg
    rep = ut.device_lowering_report(m)
    assert not rep["lowerable"] and "OBS nodes" in rep["issues"][0]
    with pytest.raises(ValueError, match="device_lowering_report"):
        m.validate()


def test_plan_conditioning_replaces_named_data():
    pm, _ = PMS.eight_schools_centered()
    y2 = np.arange(24, dtype=float).reshape(8, 3)
    y2[:, 0] = np.arange(8)
    y2[:, 1] = 10.0
    m2 = pm.condition(constraints={"y": y2})
    assert m2 is not pm and _rc(m2)[0] == L.UT_OK
    assert np.array_equal(m2._observed[-24:], y2.reshape(-1))
    with pytest.raises(ValueError, match="device_lowering_report"):
        pm.condition(constraints={"mu": 1.0})
