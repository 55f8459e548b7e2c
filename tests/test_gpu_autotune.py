"""Autotuner (tgx_set_autotune): the tuned launch shapes give the same
energies, forces and virial as the built-in ones to the precision mode's
rounding; the call that tunes already returns the chosen shapes' results;
decisions go to the cache file, and a second context that reads the cache
with tuning off makes the same launches (results equal bit for bit)."""
import numpy as np
import pytest

import tgx
import workloads as W

pytestmark = pytest.mark.gpu

# fused kernel step vs separate GEMMs (other k order in FP32), SOAP waves
# per workgroup (same arithmetic per site)
TOL = {"fp64": (1e-10, 1e-9, 1e-8), "fp32": (1e-5, 1e-4, 1e-3), "mixed": (1e-5, 1e-4, 1e-3)}


@pytest.mark.parametrize("prec", ["mixed", "fp32", "fp64"])
def test_autotune_agrees_and_caches(tmp_path, prec):
    pos, sp, lat, model = W.carbon216()
    ref = tgx.Engine(0, prec)
    ref.set_model(model)
    e0, f0, v0 = ref.energy_forces(pos, sp, lat)
    ref.close()

    path = tmp_path / "tune.txt"
    a = tgx.Engine(0, prec)
    a.set_model(model)
    a.set_autotune(True, str(path))
    e1, f1, v1 = a.energy_forces(pos, sp, lat)  # times the candidates
    e2, f2, v2 = a.energy_forces(pos, sp, lat)  # uses the decisions
    a.close()
    keys = [ln.split()[0].split("/")[0] for ln in path.read_text().splitlines()]
    print(prec, path.read_text(), flush=True)
    assert {"fwd", "step", "bwd"} <= set(keys)
    assert np.array_equal(e1, e2) and np.array_equal(f1, f2) and np.array_equal(v1, v2)

    b = tgx.Engine(0, prec)
    b.set_model(model)
    b.set_autotune(False, str(path))
    e3, f3, v3 = b.energy_forces(pos, sp, lat)
    b.close()
    assert np.array_equal(e3, e2) and np.array_equal(f3, f2) and np.array_equal(v3, v2)

    te, tf, tv = TOL[prec]
    assert np.abs(e2 - e0).max() <= te
    assert np.abs(f2 - f0).max() <= tf
    assert np.abs(v2 - v0).max() <= tv


def test_autotune_rejects_bad_mode():
    e = tgx.Engine(0, "fp64")
    with pytest.raises(tgx.TgxError):
        e.set_autotune(2)
    e.close()


@pytest.mark.parametrize("prec", ["mixed", "fp32", "fp64"])
def test_every_kernel_step_plan(monkeypatch, prec):
    """Each launch plan the tuner may choose (TGX_GEMM_PLAN="path t0 t1"
    forces one): the separate GEMMs in every tile shape, against the
    built-in plan.  Tile shapes change neither the products nor their k
    order per element; the single BF16 kernel and the separate GEMMs sum
    the site energies in another order (FP32 rounding)."""
    pos, sp, lat, model = W.carbon216()
    eng = tgx.Engine(0, prec)
    eng.set_model(model)
    e0, f0, v0 = eng.energy_forces(pos, sp, lat)
    te, tf, tv = TOL[prec]
    plans = ["2 0 0", "2 1 0", "2 2 0", "2 0 1", "2 1 1", "2 2 1"]
    try:
        for plan in plans:
            monkeypatch.setenv("TGX_GEMM_PLAN", plan)
            e, f, v = eng.energy_forces(pos, sp, lat)
            print(prec, plan, np.abs(f - f0).max(), flush=True)
            assert np.abs(e - e0).max() <= te, plan
            assert np.abs(f - f0).max() <= tf, plan
            assert np.abs(v - v0).max() <= tv, plan
    finally:
        monkeypatch.delenv("TGX_GEMM_PLAN", raising=False)
        eng.close()


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_soap_waves_per_workgroup(monkeypatch, prec):
    """The SOAP kernels with 1, 2 and 4 waves per workgroup (the tuner's
    candidates; TGX_SOAP_WPB forces one) give the same results bit for bit:
    the shape changes only which sites share a workgroup."""
    pos, sp, lat, model = W.cho(120)  # N_R = 24: one site per wave
    pos2, sp2, lat2, model2 = W.carbon216()  # N_R = 8: two sites per wave
    eng = tgx.Engine(0, prec)
    try:
        for (p_, s_, l_, m_) in ((pos, sp, lat, model), (pos2, sp2, lat2, model2)):
            eng.set_model(m_)
            out = []
            for w in ("1", "2", "4"):
                monkeypatch.setenv("TGX_SOAP_WPB", w)
                out.append(eng.energy_forces(p_, s_, l_))
            for e, f, v in out[1:]:
                assert np.array_equal(e, out[0][0]) and np.array_equal(f, out[0][1])
                assert np.array_equal(v, out[0][2])
    finally:
        monkeypatch.delenv("TGX_SOAP_WPB", raising=False)
        eng.close()
