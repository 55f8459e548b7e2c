"""Speculative Monte Carlo batches (tgx_set_mc_batch): trials proposed and
evaluated together against the current state and decided in order, a trial
that an earlier accepted move of its batch may have changed evaluated again.
The chain must be that of the one-trial sequence: the same decisions and
the same positions bit for bit, the same energy trace to rounding (the
batch's kernel step may add a site's sparse-point terms in another order
than the one-trial kernel), in FP64, mixed precision and delayed acceptance, on the 512-atom configuration (17.2 A box) where
batches carry several trials and on the small 128-atom box where nearly
every accepted move invalidates the rest of its batch."""
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
    e.set_mc_batch(0)
    e.close()


def _run(eng, pos, sp, lat, nb, n, T, seed, prec="fp64", skin=1.2):
    eng.set_mc_batch(nb)
    eng.set_precision(prec)
    try:
        return eng.mc_run(pos, sp, lat, n, T, 0.6, skin, seed)
    finally:
        eng.set_precision("fp64")
        eng.set_mc_batch(0)


@pytest.mark.parametrize("case", [(512, 17.22, 1500.0, 600), (128, 10.9, 1500.0, 300),
                                  (512, 17.22, 6000.0, 400)],
                         ids=["cfg512", "box128", "hot512"])
@pytest.mark.parametrize("nb", [8, 32, 64])
def test_batches_equal_one_trial_sequence_fp64(eng, case, nb):
    n_atoms, L, T, n = case
    pos, sp, lat = S.carbon_cell(n_atoms, 2.0, 1.2, 1004)
    print(f"n={n_atoms} T={T} trials={n} nb={nb}", flush=True)
    seq = _run(eng, pos, sp, lat, 1, n, T, 7)
    bat = _run(eng, pos, sp, lat, nb, n, T, 7)
    print("accepted", seq["n_accepted"], bat["n_accepted"], "rebuilds", seq["n_rebuilds"],
          bat["n_rebuilds"], flush=True)
    assert bat["n_accepted"] == seq["n_accepted"]
    assert np.array_equal(bat["pos"], seq["pos"])
    assert np.abs(bat["e_trace"] - seq["e_trace"]).max() < 1e-10


def test_batches_match_oracle_chain(eng):
    pos, sp, lat = S.carbon_cell(512, 2.0, 1.2, 1004)
    _, _, _, model = W.carbon216()
    g = _run(eng, pos, sp, lat, 32, 120, 1500.0, 2024)
    p_o, tr_o, acc_o = oracle.mc_run(model, pos, sp, lat, 120, 1500.0, 0.6, 2024)
    assert g["n_accepted"] == acc_o
    assert np.array_equal(g["pos"], p_o)
    assert np.abs(g["e_trace"] - tr_o).max() < 1e-8


def test_batches_mixed_precision(eng):
    """Mixed precision: the batch's site energies come from the BF16 kernel
    GEMM, the one-trial sequence's from the few-site FP32 kernel, so the
    traces differ at the mixed energies' rounding (the decisions of this
    seeded run do not)."""
    pos, sp, lat = S.carbon_cell(512, 2.0, 1.2, 1004)
    seq = _run(eng, pos, sp, lat, 1, 500, 1500.0, 31, "mixed")
    bat = _run(eng, pos, sp, lat, 32, 500, 1500.0, 31, "mixed")
    assert bat["n_accepted"] == seq["n_accepted"]
    assert np.array_equal(bat["pos"], seq["pos"])
    assert np.abs(bat["e_trace"] - seq["e_trace"]).max() < 5e-4


def _fitted_surrogate(eng, model):
    qs, ys = [], []
    for k in range(4):
        p, s, l = S.carbon_cell(64, 1.8 + 0.2 * k, 1.2, 40 + k)
        q = oracle.soap(model.hypers, p, s, l)["soap"]
        e, _ = oracle.gap_kernel_step(q, model.Qs, model.alphas, model.delta, model.zeta)
        qs.append(q)
        ys.append(e)
    eng.surrogate_init(model.hypers.dim, seed=3)
    eng.surrogate_fit(np.concatenate(qs), np.concatenate(ys), epochs=40, batch=128, lr=2e-3,
                      seed=2)
    return eng.surrogate_get()


@pytest.mark.parametrize("max_rmse", [0.5, 2e-3])
def test_batches_delayed_acceptance(eng, max_rmse):
    """Two-stage chain with a fitted surrogate (and with a drift bound that
    trips part way): batched equals one-trial, stage counters included."""
    pos, sp, lat = S.carbon_cell(512, 2.0, 1.2, 1004)
    _, _, _, model = W.carbon216()
    P, mu, sigma = _fitted_surrogate(eng, model)
    out = []
    for nb in (1, 32):
        eng.surrogate_set(P, mu, sigma)
        eng.set_mc_batch(nb)
        try:
            out.append(eng.mc_run_da(pos, sp, lat, 400, 1500.0, 0.6, 1.2, 99, n_train_trials=0,
                                     max_rmse=max_rmse))
        finally:
            eng.set_mc_batch(0)
    seq, bat = out
    print({k: (seq[k], bat[k]) for k in ("n_accepted", "n_stage1_rejected", "n_gap_evaluations",
                                         "fallback")}, flush=True)
    for k in ("n_accepted", "n_stage1_rejected", "n_gap_evaluations", "fallback"):
        assert bat[k] == seq[k], k
    assert np.array_equal(bat["pos"], seq["pos"])
    assert np.abs(bat["e_trace"] - seq["e_trace"]).max() < 1e-10
