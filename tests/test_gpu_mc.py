"""Device Monte Carlo (tgx_mc_run) against the oracle's sequential
Metropolis restatement, which recomputes the whole system per trial as the
reference does.  FP64: identical accept/reject sequence, positions bit for
bit, energy trace within 1e-8 eV.  Mixed: the acceptance rate and the energy
of the final positions checked against the oracle's recomputation."""
import os

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
    _, _, _, model = W.carbon216()
    e.set_model(model)
    yield e
    e.close()


def _system():
    pos, sp, lat = S.carbon_cell(128, 2.0, 1.2, 1004)  # 10.9 A box >= 2 (3.7 + 1.2)
    _, _, _, model = W.carbon216()
    return pos, sp, lat, model


@pytest.mark.parametrize("graph", [True, False])
def test_mc_fp64_matches_oracle(eng, graph, monkeypatch):
    pos, sp, lat, model = _system()
    if not graph:
        monkeypatch.setenv("TGX_NO_GRAPH", "1")
    n_trials = 150
    print(f"graph={graph} n={len(pos)} trials={n_trials}", flush=True)
    g = eng.mc_run(pos, sp, lat, n_trials, 1500.0, 0.6, 1.2, 2024)
    p_o, tr_o, acc_o = oracle.mc_run(model, pos, sp, lat, n_trials, 1500.0, 0.6, 2024)
    assert g["n_accepted"] == acc_o
    assert np.array_equal(g["pos"], p_o)
    assert np.abs(g["e_trace"] - tr_o).max() < 1e-8


def test_mc_rebuilds_and_energy_consistency(eng):
    """A long hot run forces list rebuilds; the accumulated energy must equal
    a fresh evaluation of the final positions."""
    pos, sp, lat, model = _system()
    g = eng.mc_run(pos, sp, lat, 1500, 6000.0, 0.6, 1.1, 77)
    assert g["n_rebuilds"] >= 1 and g["n_accepted"] > 300
    e, _, _ = oracle.energy_forces(model, g["pos"], sp, lat, do_forces=False)
    assert abs(e.sum() - g["e_trace"][-1]) < 1e-8


def test_mc_mixed_precision(eng):
    pos, sp, lat, model = _system()
    eng.set_precision("mixed")
    try:
        g = eng.mc_run(pos, sp, lat, 400, 1500.0, 0.6, 1.2, 31)
    finally:
        eng.set_precision("fp64")
    _, tr_o, acc_o = oracle.mc_run(model, pos, sp, lat, 400, 1500.0, 0.6, 31)
    assert abs(g["n_accepted"] - acc_o) <= 0.1 * 400
    e, _, _ = oracle.energy_forces(model, g["pos"], sp, lat, do_forces=False)
    # the trace accumulates mixed-precision site-energy changes
    assert abs(e.sum() - g["e_trace"][-1]) < 5e-3 * len(pos)


def test_mc_mixed_partial_energies_match_row_sums(eng):
    """Mixed precision: a plain trial's decision kernel adds the few-site
    energy kernel's FP32 partials itself; with sample collection on (the
    training phase of delayed acceptance, same random streams) the row-sum
    kernel adds them first.  The same arithmetic in the same order, so the
    two chains agree bit for bit."""
    pos, sp, lat, model = _system()
    eng.set_precision("mixed")
    eng.set_mc_batch(1)  # the one-trial sequence on both sides
    try:
        a = eng.mc_run(pos, sp, lat, 300, 1500.0, 0.6, 1.2, 41)
        eng.surrogate_init(model.hypers.dim, seed=4)
        b = eng.mc_run_da(pos, sp, lat, 300, 1500.0, 0.6, 1.2, 41, n_train_trials=300,
                          train_every=1 << 20, epochs=1, batch=64, lr=1e-3, max_rmse=1.0)
    finally:
        eng.set_precision("fp64")
        eng.set_mc_batch(0)
    print("accepted", a["n_accepted"], b["n_accepted"], flush=True)
    assert a["n_accepted"] == b["n_accepted"] and a["n_accepted"] > 50
    assert np.array_equal(a["pos"], b["pos"])
    assert np.array_equal(a["e_trace"], b["e_trace"])


def test_mc_rejects_bad_arguments(eng):
    pos, sp, lat, _ = _system()
    with pytest.raises(tgx.TgxError):  # skin too small for the move
        eng.mc_run(pos, sp, lat, 10, 1500.0, 0.6, 0.5, 1)
    small = S.carbon_cell(40, 2.0, 1.2, 3)
    with pytest.raises(tgx.TgxError) as ei:  # box below 2 (rcut + skin)
        eng.mc_run(small[0], small[1], small[2], 10, 1500.0, 0.6, 1.2, 1)
    assert ei.value.code == tgx.TGX_ERR_UNSUPPORTED
