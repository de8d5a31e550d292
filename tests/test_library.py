"""CPU checks of the C ABI: the library loads, exports every symbol the header
declares, and validates arguments before touching a device."""
import ctypes as ct
import re
from pathlib import Path

import numpy as np
import pytest

from uncertaintea_amd import _lib as L

ROOT = Path(__file__).resolve().parent.parent


def header_symbols():
    text = (ROOT / "include" / "uncertaintea_hip.h").read_text()
    return sorted(set(re.findall(r"\b(ut_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header_symbols():
    lib = L.lib()
    syms = header_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), s
    assert set(L.EXPORTED) == set(syms)


def test_abi_version_and_host_philox(golden):
    assert L.lib().ut_abi_version() == 2
    for key, ctr, expected in golden["philox_kat"]:
        assert L.philox4x32(key, ctr).tolist() == expected


def test_plan_argument_validation_without_device():
    h = ct.c_void_p()
    x = np.zeros((4, 2), np.float32)
    y = np.array([0, 1, 2, 0], np.float32)  # 2 is not a Bernoulli outcome
    rc = L.lib().ut_plan_logistic_glm(L.ptr(x), L.ptr(y), 4, 2, 0.0, 1.0, ct.byref(h))
    assert rc == L.UT_E_ARG and b"0/1" in L.lib().ut_last_error()
    rc = L.lib().ut_plan_logistic_glm(L.ptr(x), L.ptr(y), 4, 300, 0.0, 1.0, ct.byref(h))
    assert rc == L.UT_E_UNSUPPORTED
    # the reference's unsupported-model message points at the lowering report
    # (src/device/workspace.jl:24-35)
    assert b"See device_lowering_report(model) for details." in L.lib().ut_last_error()
    rc = L.lib().ut_plan_funnel(0, 3.0, ct.byref(h))
    assert rc == L.UT_E_DIM
    off = np.array([0, 2], np.int32)
    sig = np.array([-1.0])
    rc = L.lib().ut_plan_hier_normal_nc(L.ptr(np.zeros(2, np.float32)), L.ptr(off), L.ptr(sig), 1, 0.0, 5.0, 5.0,
                                        ct.byref(h))
    assert rc == L.UT_E_ARG
    with pytest.raises(ValueError):
        L.check(rc)


def test_sampler_defaults_match_reference():
    o = L.SamplerOpts()
    L.lib().ut_sampler_opts_default(ct.byref(o))
    assert (o.max_tree_depth, o.num_leapfrog_steps, o.step_size, o.target_accept) == (10, 10, 0.1, 0.8)
    assert (o.divergence_threshold, o.mass_matrix_regularization, o.mass_matrix_min_samples) == (1000.0, 1e-3, 10)
    assert (o.adapt_step_size, o.adapt_mass_matrix, o.init_max_retries) == (1, 1, 100)
    # progress callback: none by default, reported every 10 iterations when set
    # (reference callback = nothing, callback_every = 10, api_batched_nuts.jl:42-43)
    assert not o.callback and o.callback_every == 10
    mo = L.MeadsOpts()
    L.lib().ut_meads_opts_default(ct.byref(mo))
    assert not mo.callback and mo.callback_every == 10 and mo.num_folds == 4


def test_product_modules_do_not_use_the_oracle():
    """The oracle is test infrastructure: no module of the package imports or loads it
    (the build step compiles it beside the library, which is not using it), and no
    device source includes it."""
    pkg = ROOT / "uncertaintea_amd"
    for f in sorted(pkg.glob("*.py")):
        text = f.read_text()
        assert not re.search(r"^\s*(from|import)\s+oracle\b", text, re.M), f.name
        if f.name != "_build.py":
            assert "liboracle" not in text and "ORACLE" not in text, f.name
    for f in sorted((pkg / "csrc").glob("*")):
        if f.suffix in (".hip", ".h", ".cpp"):
            assert not re.search(r"#include\s*[<\"][^>\"]*oracle", f.read_text()), f.name


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_conditioning_replaces_arguments_and_observations():
    """args / constraints condition the model (reference device_batched_logjoint(model,
    params, args, constraints), src/device/api.jl:34-88) on the host, before any plan
    exists; unknown constraint names and wrong argument counts raise the reference's
    error kinds, the former naming device_lowering_report."""
    import uncertaintea_amd as ut
    x = np.arange(12, dtype=np.float32).reshape(4, 3)
    y = np.array([0, 1, 1, 0], np.float32)
    m = ut.LogisticRegression(x, y)
    assert m.condition() is m and m.condition((), None) is m and m.condition((), {}) is m
    y2 = np.array([1, 1, 0, 0], np.float32)
    m2 = m.condition(constraints={"y": y2})
    assert m2 is not m and np.array_equal(m2.y, y2) and np.array_equal(m2.x, m.x)
    m3 = m.condition((2 * x,), {"y": y2})
    assert np.array_equal(m3.x, 2 * x)
    with pytest.raises(ValueError, match="device_lowering_report"):
        m.condition(constraints={"beta": np.zeros(3)})
    with pytest.raises(ValueError, match="DimensionMismatch"):
        m.condition((x, y))
    h = ut.HierarchicalNormalNC(np.zeros(4, np.float32), np.array([0, 2, 4], np.int32), np.ones(2))
    h2 = h.condition((np.array([0, 1, 4], np.int32), np.array([1.0, 2.0])), {"y": np.ones(4, np.float32)})
    assert np.array_equal(h2.offsets, [0, 1, 4]) and np.array_equal(h2.y, np.ones(4))
    f = ut.NealsFunnel(5)
    assert f.condition() is f
    with pytest.raises(ValueError, match="device_lowering_report"):
        f.condition(constraints={"v": 0.0})
    rep = ut.device_lowering_report(h)
    assert rep["lowerable"] and rep["num_params"] == 4 and rep["observed"] == ["y"] and "batched_nuts" in rep["samplers"]
