"""Local-property models (Hirshfeld-volume-like) on the GPU against the
oracle's restatement of local_property_predict (local_properties.f90:46-146)
and get_local_properties' scatter onto the full list
(gap_interface.f90:342-351): FP64 V within 1e-10, dV_i/dr_j within 1e-8;
mixed precision within the descriptor's FP32 and the GEMM's BF16 rounding."""
import numpy as np
import pytest

import oracle
import structures as S
import tgx
import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = tgx.Engine(0, "fp64")
    yield e
    e.close()


def _prop(h, pos, sp, lat, seed=5, n_sparse=60):
    rng = np.random.default_rng(seed)
    q = oracle.soap(h, pos, sp, lat)["soap"]
    Qs = q[rng.choice(len(q), n_sparse, replace=len(q) < n_sparse)]
    al = rng.normal(0.0, 0.5, n_sparse)
    e, _ = oracle.gap_kernel_step(q, Qs, al, 0.9, 3.0)
    return Qs, al, -float(np.median(e))


@pytest.mark.parametrize("case", ["carbon216", "small24"])
def test_local_property_fp64(eng, case):
    if case == "carbon216":
        pos, sp, lat, model = W.carbon216()
    else:
        pos, sp, lat = S.carbon_cell(24, 2.6, 1.2, 61)
        _, _, _, model = W.carbon216()
    eng.set_model(model)
    eng.clear_local_properties()
    Qs, al, V0 = _prop(model.hypers, pos, sp, lat)
    k = eng.add_local_property(0, Qs, al, 0.9, 3.0, V0)
    V, der = eng.local_property(k, pos, sp, lat, 4.2, derivatives=True)
    V0r, der0 = oracle.local_property(model.hypers, Qs, al, 0.9, 3.0, V0, pos, sp, lat, 4.2,
                                      derivatives=True)
    print(case, "dV", np.abs(V - V0r).max(), "dder", np.abs(der - der0).max(),
          "clipped", int(np.sum(V0r == 0)), flush=True)
    assert np.sum(V0r == 0) > 0
    assert np.array_equal(V == 0, V0r == 0)
    assert np.abs(V - V0r).max() < 1e-10
    assert der.shape == der0.shape and np.abs(der - der0).max() < 1e-8
    V2, _ = eng.local_property(k, pos, sp, lat)
    assert np.abs(V2 - V).max() < 1e-12


def test_local_property_central_species_block_and_mixed(eng):
    """A property on the second block of a two-block C/H/O model, that block
    restricted to central species 2; then mixed precision."""
    pos, sp, lat, model = W.cho()
    h2 = S.cho_hypers()
    h2.central_species = 2
    q2 = oracle.soap(h2, pos, sp, lat)["soap"]
    m2 = S.model_from_soaps(h2, q2, 50, seed=4)
    full = tgx.GapModel(hypers=model.hypers, Qs=model.Qs, alphas=model.alphas, delta=model.delta,
                        zeta=model.zeta, e0=model.e0, more_soap=[m2])
    eng.set_model(full)
    eng.clear_local_properties()
    Qs, al, V0 = _prop(h2, pos, sp, lat, seed=9, n_sparse=40)
    k = eng.add_local_property(1, Qs, al, 0.9, 3.0, V0)
    V, der = eng.local_property(k, pos, sp, lat, 4.5, derivatives=True)
    V0r, der0 = oracle.local_property(h2, Qs, al, 0.9, 3.0, V0, pos, sp, lat, 4.5,
                                      derivatives=True)
    assert np.all(V[sp != 2] == 0.0)
    assert np.abs(V - V0r).max() < 1e-10 and np.abs(der - der0).max() < 1e-8
    eng.set_precision("mixed")
    try:
        Vm, derm = eng.local_property(k, pos, sp, lat, 4.5, derivatives=True)
    finally:
        eng.set_precision("fp64")
    # sites clipped in one precision but not the other are left out
    scale = np.abs(V0r).max()
    live = (Vm > 0) & (V0r > 0)
    assert np.abs(Vm - V0r)[live].max() < 2e-2 * scale
    rows = np.repeat(live, eng.neighbors(pos, lat, 4.5)["n_neigh"])
    assert np.abs(derm - der0)[rows].max() < 2e-2 * max(1.0, np.abs(der0).max())


def test_local_property_triclinic_values(eng):
    pos, sp, lat = S.triclinic_cell(40, 2.4, 1.1, 31, skew=(0.3, -0.2, 0.4), rotate=True)
    _, _, _, model = W.carbon216()
    eng.set_model(model)
    eng.clear_local_properties()
    Qs, al, V0 = _prop(model.hypers, pos, sp, lat)
    k = eng.add_local_property(0, Qs, al, 0.9, 3.0, V0)
    V, _ = eng.local_property(k, pos, sp, lat)
    V0r, _ = oracle.local_property(model.hypers, Qs, al, 0.9, 3.0, V0, pos, sp, lat)
    assert np.abs(V - V0r).max() < 1e-10
    with pytest.raises(tgx.TgxError) as ei:
        eng.local_property(k, pos, sp, lat, 4.2, derivatives=True)
    assert ei.value.code == tgx.TGX_ERR_UNSUPPORTED
