"""General static execution plans: the host side of `ut_plan_create`.

The reference lowers a `@tea (static)` model to an isbits DeviceExecutionPlan --
a tuple of choice, deterministic and loop steps whose arguments are small
expression trees over slots (src/device/plan/steps.jl:1-539, exprs.jl:1-31) -- and
stages the observed values into one dense stream in the kernel's pre-order, with a
trip count per loop (src/device/staging.jl:499-550).  `PlanBuilder` writes the same
structure directly: latent and observed choices of the supported families,
deterministic bindings, and loops over the rows of a data block, with arguments
built from ordinary Python arithmetic on `Expr` values.  `build()` serialises it
into the flat `ut_plan_desc` of the C ABI (steps in pre-order, expression nodes in
post-order, the staged rows) and returns a `PlanModel`, a TeaModel that every
device entry point and sampler accepts.

    b = PlanBuilder()
    mu = b.latent("normal", 0.0, 5.0)
    tau = b.latent("halfcauchy", 5.0, transform="log")
    theta_raw = b.latent_vector(J, "normal", 0.0, 1.0)
    with b.loop(np.column_stack([group, sigma, y])) as it:
        it.observe("normal", mu + tau * theta_raw[it.col(0)], it.col(1), value=it.col(2))
    model = b.build("hierarchical")

Plans of the fused families (the iid-normal logistic GLM, the non-centred
hierarchical normal, Neal's funnel) run their dedicated kernels unless
`generic=True`; everything else runs on the plan interpreter.
"""
from __future__ import annotations

import ctypes as ct
from contextlib import contextmanager

import numpy as np

from . import _lib as L
from .models import TeaModel

FAMILIES = {
    "normal": (L.UT_FAMILY_NORMAL, 2), "noncentered_normal": (L.UT_FAMILY_NONCENTERED_NORMAL, 2),
    "lognormal": (L.UT_FAMILY_LOGNORMAL, 2), "halfnormal": (L.UT_FAMILY_HALFNORMAL, 1),
    "halfcauchy": (L.UT_FAMILY_HALFCAUCHY, 1), "exponential": (L.UT_FAMILY_EXPONENTIAL, 1),
    "cauchy": (L.UT_FAMILY_CAUCHY, 2), "laplace": (L.UT_FAMILY_LAPLACE, 2), "logistic": (L.UT_FAMILY_LOGISTIC, 2),
    "studentt": (L.UT_FAMILY_STUDENTT, 3), "gamma": (L.UT_FAMILY_GAMMA, 2),
    "bernoulli": (L.UT_FAMILY_BERNOULLI, 1), "bernoullilogit": (L.UT_FAMILY_BERNOULLI_LOGIT, 1),
    "poisson": (L.UT_FAMILY_POISSON, 1), "gumbel": (L.UT_FAMILY_GUMBEL, 2),
    "frechet": (L.UT_FAMILY_FRECHET, 2), "rayleigh": (L.UT_FAMILY_RAYLEIGH, 1),
    "inversegaussian": (L.UT_FAMILY_INVERSEGAUSSIAN, 2), "beta": (L.UT_FAMILY_BETA, 2),
    "inversegamma": (L.UT_FAMILY_INVERSEGAMMA, 2), "weibull": (L.UT_FAMILY_WEIBULL, 2),
    "binomial": (L.UT_FAMILY_BINOMIAL, 2), "geometric": (L.UT_FAMILY_GEOMETRIC, 1),
    "negativebinomial": (L.UT_FAMILY_NEGATIVEBINOMIAL, 2),
}
TRANSFORMS = {"identity": L.UT_TRANSFORM_IDENTITY, "log": L.UT_TRANSFORM_LOG, "logit": L.UT_TRANSFORM_LOGIT}
ENGINES = {L.UT_ENGINE_INTERPRETER: "interpreter", L.UT_ENGINE_LOGISTIC: "logistic_glm",
           L.UT_ENGINE_HIER: "hierarchical_normal_nc", L.UT_ENGINE_FUNNEL: "funnel",
           L.UT_ENGINE_GLM_MATRIX_CORES: "logistic_glm_matrix_cores", L.UT_ENGINE_GP: "gp_poisson"}


# ---------------------------------------------------------------- expressions
class Expr:
    """A node of an argument expression (the reference's DeviceLiteralExpr /
    DeviceSlotExpr / DevicePrimitiveExpr, plus PARAM and OBS leaves)."""

    __slots__ = ("op", "a", "b", "index", "value")

    def __init__(self, op, a=None, b=None, index=0, value=0.0):
        self.op, self.a, self.b, self.index, self.value = op, a, b, int(index), float(value)

    def _bin(self, op, other, swap=False):
        o = as_expr(other)
        return Expr(op, o, self) if swap else Expr(op, self, o)

    def __add__(self, o): return self._bin(L.UT_OP_ADD, o)
    def __radd__(self, o): return self._bin(L.UT_OP_ADD, o, True)
    def __sub__(self, o): return self._bin(L.UT_OP_SUB, o)
    def __rsub__(self, o): return self._bin(L.UT_OP_SUB, o, True)
    def __mul__(self, o): return self._bin(L.UT_OP_MUL, o)
    def __rmul__(self, o): return self._bin(L.UT_OP_MUL, o, True)
    def __truediv__(self, o): return self._bin(L.UT_OP_DIV, o)
    def __rtruediv__(self, o): return self._bin(L.UT_OP_DIV, o, True)
    def __pow__(self, o): return self._bin(L.UT_OP_POW, o)
    def __neg__(self): return Expr(L.UT_OP_NEG, self)

    def _nodes(self, out: list) -> int:
        """Append this tree's nodes in post-order (operands first); returns its root."""
        a = self.a._nodes(out) if self.a is not None else -1
        b = self.b._nodes(out) if self.b is not None else -1
        out.append((self.op, a, b, self.index, self.value))
        return len(out) - 1

    def _uses(self, op):
        return self.op == op or (self.a is not None and self.a._uses(op)) or (self.b is not None and self.b._uses(op))


def as_expr(x) -> Expr:
    if isinstance(x, Expr):
        return x
    if isinstance(x, (int, float, np.integer, np.floating)):
        return Expr(L.UT_OP_LITERAL, value=float(x))
    raise TypeError(f"cannot use {type(x).__name__} in a plan expression")


def _unary(op):
    return lambda x: Expr(op, as_expr(x))


exp = _unary(L.UT_OP_EXP)
log = _unary(L.UT_OP_LOG)
log1p = _unary(L.UT_OP_LOG1P)
sqrt = _unary(L.UT_OP_SQRT)
abs_ = _unary(L.UT_OP_ABS)
tanh = _unary(L.UT_OP_TANH)
logistic = _unary(L.UT_OP_LOGISTIC)


class VectorLatent:
    """Rows first .. first + n - 1 of a DIAG_NORMAL latent: `v[k]` with an int reads
    row first + k; `v[expr]` with a data expression (a loop's column) reads the row
    the data names (the reference cannot lower this: vector bindings are never
    materialised there)."""

    def __init__(self, first: int, n: int):
        self.first, self.n = first, n

    def __len__(self):
        return self.n

    def __getitem__(self, k):
        if isinstance(k, (int, np.integer)):
            if not 0 <= k < self.n:
                raise IndexError("vector latent index out of range")
            return Expr(L.UT_OP_PARAM, index=self.first + int(k))
        return Expr(L.UT_OP_PARAM, a=as_expr(k), index=self.first)


# ---------------------------------------------------------------- builder
class _Loop:
    def __init__(self, builder, data, start, name):
        self.b = builder
        self.data = np.ascontiguousarray(np.atleast_2d(np.asarray(data, dtype=np.float64)))
        if self.data.ndim != 2:
            raise ValueError("a loop's data block must be n x stride")
        self.start, self.name = int(start), name
        self.body = []
        self.iterator_slot = 0

    def col(self, k: int) -> Expr:
        """Column k of the iteration's row of the data block."""
        if not 0 <= k < self.data.shape[1]:
            raise IndexError("loop column out of range")
        return Expr(L.UT_OP_OBS, index=int(k))

    @property
    def index(self) -> Expr:
        """The iteration's index start + t (the loop's iterator slot)."""
        if not self.iterator_slot:
            self.iterator_slot = self.b._new_slot()
        return Expr(L.UT_OP_SLOT, index=self.iterator_slot)

    def _value_col(self, value) -> int:
        if not (isinstance(value, Expr) and value.op == L.UT_OP_OBS):
            raise ValueError("an observed value inside a loop is a column of its data block: value=it.col(k)")
        return value.index

    def observe(self, family: str, *args, value, bind: bool = False):
        code, arity = _family(family)
        if code in (L.UT_FAMILY_NONCENTERED_NORMAL,):
            raise ValueError(f"{family} cannot be observed")
        if len(args) != arity:
            raise ValueError(f"{family} takes {arity} arguments, got {len(args)}")
        slot = self.b._new_slot() if bind else 0
        self.body.append(dict(kind=L.UT_STEP_CHOICE, family=code, args=[as_expr(a) for a in args],
                              value_offset=self._value_col(value), binding_slot=slot))
        return Expr(L.UT_OP_SLOT, index=slot) if bind else None

    def glm(self, coefficients: VectorLatent, first_col: int, outcome_col: int, intercept=0.0):
        """y ~ bernoullilogit(intercept + sum_d coefficients[d] x_d) with covariates
        x_d in columns first_col .. first_col + D - 1 and y in outcome_col =
        first_col + D (the reference's DeviceBernoulliLogitGLMChoiceStepDyn)."""
        D = len(coefficients)
        if outcome_col != first_col + D:
            raise ValueError("the GLM outcome column follows its D covariate columns")
        self.body.append(dict(kind=L.UT_STEP_CHOICE, family=L.UT_FAMILY_BERNOULLI_LOGIT_GLM,
                              args=[as_expr(intercept)], value_offset=int(first_col), dimension=D,
                              coef_source=coefficients.first))

    def deterministic(self, expr) -> Expr:
        slot = self.b._new_slot()
        self.body.append(dict(kind=L.UT_STEP_DETERMINISTIC, args=[as_expr(expr)], binding_slot=slot))
        return Expr(L.UT_OP_SLOT, index=slot)


def _family(name):
    if name not in FAMILIES:
        raise ValueError(f"unknown family {name!r}; the plan supports {sorted(FAMILIES)}")
    return FAMILIES[name]


class PlanBuilder:
    """Builds a general static execution plan step by step (reference pre-order)."""

    def __init__(self):
        self.items = []   # ("step", dict) | ("loop", _Loop) | ("observe", dict, value)
        self.num_params = 0
        self.num_slots = 0
        self.names = []        # per parameter row
        self.transforms = []   # per parameter row
        self._open_loop = None

    def _new_slot(self) -> int:
        self.num_slots += 1
        return self.num_slots

    def _rows(self, n, name, transform):
        first = self.num_params + 1
        self.num_params += n
        self.names += [name if n == 1 else f"{name}[{k + 1}]" for k in range(n)]
        self.transforms += [transform] * n
        return first

    def _check_top(self):
        if self._open_loop is not None:
            raise ValueError("latent and top-level steps cannot be added inside a loop body")

    def latent(self, family: str, *args, transform: str = "identity", name: str | None = None) -> Expr:
        """A scalar latent choice; returns its (constrained) value."""
        self._check_top()
        code, arity = _family(family)
        if len(args) != arity:
            raise ValueError(f"{family} takes {arity} arguments, got {len(args)}")
        if transform not in TRANSFORMS:
            raise ValueError(f"transform must be one of {sorted(TRANSFORMS)}")
        row = self._rows(1, name or f"p{self.num_params + 1}", transform)
        slot = self._new_slot()
        self.items.append(("step", dict(kind=L.UT_STEP_CHOICE, family=code, args=[as_expr(a) for a in args],
                                        value_source=row, transform=TRANSFORMS[transform], binding_slot=slot)))
        return Expr(L.UT_OP_SLOT, index=slot)

    def latent_vector(self, n: int, family: str, mu, sigma, name: str | None = None) -> VectorLatent:
        """n iid normal(mu, sigma) latents (DIAG_NORMAL; identity transform)."""
        self._check_top()
        if family != "normal":
            raise ValueError("vector latents are iid normal (the reference's DeviceDiagNormalChoiceStepDyn)")
        first = self._rows(int(n), name or f"v{self.num_params + 1}", "identity")
        self.items.append(("step", dict(kind=L.UT_STEP_CHOICE, family=L.UT_FAMILY_DIAG_NORMAL,
                                        args=[as_expr(mu), as_expr(sigma)], value_source=first,
                                        dimension=int(n))))
        return VectorLatent(first, int(n))

    def observe(self, family: str, *args, value, name: str | None = None, bind: bool = False):
        """A top-level observed choice (a scalar, or a vector for family 'normal' with
        an array value: n iid normal(mu, sigma) observations)."""
        self._check_top()
        v = np.atleast_1d(np.asarray(value, dtype=np.float64))
        if v.size > 1:
            if family != "normal" or bind:
                raise ValueError("vector observations are iid normal and bind nothing")
            self.items.append(("observe", dict(kind=L.UT_STEP_CHOICE, family=L.UT_FAMILY_DIAG_NORMAL,
                                               args=[as_expr(a) for a in args], dimension=int(v.size)), v, name))
            return None
        code, arity = _family(family)
        if len(args) != arity:
            raise ValueError(f"{family} takes {arity} arguments, got {len(args)}")
        slot = self._new_slot() if bind else 0
        self.items.append(("observe", dict(kind=L.UT_STEP_CHOICE, family=code, args=[as_expr(a) for a in args],
                                           binding_slot=slot), v, name))
        return Expr(L.UT_OP_SLOT, index=slot) if bind else None

    def deterministic(self, expr) -> Expr:
        self._check_top()
        slot = self._new_slot()
        self.items.append(("step", dict(kind=L.UT_STEP_DETERMINISTIC, args=[as_expr(expr)],
                                        binding_slot=slot)))
        return Expr(L.UT_OP_SLOT, index=slot)

    @contextmanager
    def loop(self, data, start: int = 1, name: str | None = None):
        """A loop over the rows of `data` (n x stride): one iteration per row, whose
        columns its body reads with it.col(k)."""
        self._check_top()
        lp = _Loop(self, data, start, name)
        self._open_loop = lp
        try:
            yield lp
        finally:
            self._open_loop = None
        self.items.append(("loop", lp))

    def build(self, name: str = "plan", generic: bool = False, matrix_cores: bool = False) -> "PlanModel":
        return PlanModel(self, name, generic=generic, matrix_cores=matrix_cores)


def _serialise(builder: PlanBuilder, data_override: dict | None = None):
    """(steps, exprs, observed, trip_counts, loop_starts, observed_names) arrays."""
    steps, nodes, observed, trips, starts, names = [], [], [], [], [], {}
    data_override = data_override or {}

    def step_rec(d):
        args = [-1, -1, -1]
        for j, e in enumerate(d.get("args", [])):
            args[j] = e._nodes(nodes)
        return (d["kind"], d.get("family", 0), d.get("value_source", 0), d.get("transform", 0),
                d.get("binding_slot", 0), args, d.get("dimension", 0), d.get("coef_source", 0),
                d.get("value_offset", 0), 0, 0, 0, 0)

    for item in builder.items:
        if item[0] == "step":
            steps.append(step_rec(item[1]))
        elif item[0] == "observe":
            _, d, v, nm = item
            if nm is not None:
                names[nm] = ("observe", len(observed), v.size)
                if nm in data_override:
                    v = np.atleast_1d(np.asarray(data_override[nm], dtype=np.float64)).reshape(v.shape)
            steps.append(step_rec(d))
            observed.extend(v.tolist())
        else:
            lp = item[1]
            data = lp.data
            if lp.name is not None:
                names[lp.name] = ("loop", len(observed), data.shape)
                if lp.name in data_override:
                    data = np.ascontiguousarray(np.asarray(data_override[lp.name], dtype=np.float64))
                    if data.ndim != 2 or data.shape[1] != lp.data.shape[1]:
                        raise ValueError(f"DimensionMismatch: loop data {lp.name!r} must keep {lp.data.shape[1]} "
                                         f"columns")
            trips.append(data.shape[0])
            starts.append(lp.start)
            rec = list(step_rec(dict(kind=L.UT_STEP_LOOP)))
            rec[9], rec[10], rec[11], rec[12] = len(trips), lp.iterator_slot, len(lp.body), data.shape[1]
            steps.append(tuple(rec))
            for d in lp.body:
                steps.append(step_rec(d))
            observed.extend(data.reshape(-1).tolist())
    return steps, nodes, observed, trips, starts, names


class PlanModel(TeaModel):
    """A model given as a general static execution plan (see PlanBuilder)."""

    def __init__(self, builder: PlanBuilder, name: str = "plan", generic: bool = False, matrix_cores: bool = False,
                 _data: dict | None = None):
        super().__init__()
        self.name = name
        self._builder = builder
        self.generic, self.matrix_cores = bool(generic), bool(matrix_cores)
        self._data = dict(_data or {})
        steps, nodes, observed, trips, starts, self._names = _serialise(builder, self._data)
        self._steps = (L.PlanStep * max(1, len(steps)))()
        for i, r in enumerate(steps):
            st = self._steps[i]
            (st.kind, st.family, st.value_source, st.transform, st.binding_slot, args, st.dimension,
             st.coef_source, st.value_offset, st.loop_id, st.iterator_slot, st.body_length, st.stride) = r
            for j in range(3):
                st.args[j] = args[j]
        self._exprs = (L.PlanExpr * max(1, len(nodes)))()
        for i, (op, a, b, index, value) in enumerate(nodes):
            e = self._exprs[i]
            e.op, e.a, e.b, e.index, e.value = op, a, b, index, value
        self._observed = np.ascontiguousarray(observed if observed else [0.0], dtype=np.float64)
        self._trips = np.ascontiguousarray(trips if trips else [0], dtype=np.int32)
        self._starts = np.ascontiguousarray(starts if starts else [0], dtype=np.int32)
        d = L.PlanDesc()
        d.num_params, d.num_slots, d.num_loops = builder.num_params, builder.num_slots, len(trips)
        d.num_steps, d.steps = len(steps), ct.cast(self._steps, ct.POINTER(L.PlanStep))
        d.num_exprs, d.exprs = len(nodes), ct.cast(self._exprs, ct.POINTER(L.PlanExpr))
        d.num_observed = len(observed)
        d.observed = self._observed.ctypes.data_as(ct.POINTER(ct.c_double))
        d.trip_counts = self._trips.ctypes.data_as(ct.POINTER(ct.c_int32))
        d.loop_starts = self._starts.ctypes.data_as(ct.POINTER(ct.c_int32))
        d.flags = (L.UT_PLAN_GENERIC if self.generic else 0) | (L.UT_PLAN_MATRIX_CORES if self.matrix_cores else 0)
        self.desc = d
        self.observed_names = tuple(self._names)

    @property
    def num_params(self):
        return self._builder.num_params

    @property
    def param_names(self):
        return list(self._builder.names)

    def constrain(self, u):
        c = np.array(u, dtype=np.float64, copy=True)
        for r, t in enumerate(self._builder.transforms):
            if t == "log":
                c[..., r] = np.exp(c[..., r])
            elif t == "logit":
                c[..., r] = 1.0 / (1.0 + np.exp(-c[..., r]))
        return c

    def _create_plan(self, out):
        return L.lib().ut_plan_create(ct.byref(self.desc), out)

    def validate(self) -> None:
        """Host-only audit (no device); raises the reference's error kinds."""
        L.check(L.lib().ut_plan_validate(ct.byref(self.desc)))

    @property
    def engine(self) -> str:
        e = ct.c_int32()
        L.check(L.lib().ut_plan_engine(self.plan, ct.byref(e)))
        return ENGINES.get(int(e.value), str(e.value))

    def with_flags(self, generic: bool | None = None, matrix_cores: bool | None = None) -> "PlanModel":
        return PlanModel(self._builder, self.name, self.generic if generic is None else generic,
                         self.matrix_cores if matrix_cores is None else matrix_cores, self._data)

    def _conditioned(self, args, constraints):
        if args:
            raise ValueError(f"model {self.name}: a plan's arguments are its data blocks; condition by name through "
                             f"constraints.\nSee device_lowering_report(model) for details.")
        data = dict(self._data)
        data.update(constraints)
        return PlanModel(self._builder, self.name, self.generic, self.matrix_cores, data)

    def lowering_report(self) -> dict:
        rc = L.lib().ut_plan_validate(ct.byref(self.desc))
        issues = [] if rc == L.UT_OK else [L.lib().ut_last_error().decode(errors="replace")]
        return {"model": self.name, "lowerable": rc == L.UT_OK, "plan": "PlanModel", "num_params": self.num_params,
                "arguments": [], "observed": list(self.observed_names), "issues": issues,
                "num_steps": int(self.desc.num_steps), "num_slots": int(self.desc.num_slots),
                "staged_rows": int(self.desc.num_observed)}
