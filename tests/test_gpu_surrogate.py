"""The surrogate network's device kernels (inference, gradient, Adam
training) against the oracle's restatement, what training gains, and
delayed-acceptance Monte Carlo against the oracle's two-stage chain.

Tolerances: the device accumulates in FP32 over BF16 inputs that the oracle
rounds identically, so energies agree to 1e-5 of their scale and gradients
to 1e-3 of the largest component."""
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


def _samples(n_cells=6, seed=40):
    """SOAP descriptors and GAP site energies of carbon cells (the training
    data delayed acceptance collects along a run)."""
    _, _, _, model = W.carbon216()
    qs, ys = [], []
    for k in range(n_cells):
        pos, sp, lat = S.carbon_cell(64, 1.8 + 0.2 * k, 1.2, seed + k)
        q = oracle.soap(model.hypers, pos, sp, lat)["soap"]
        e, _ = oracle.gap_kernel_step(q, model.Qs, model.alphas, model.delta, model.zeta)
        qs.append(q)
        ys.append(e)
    return np.concatenate(qs), np.concatenate(ys)


def test_surrogate_forward_matches_oracle(eng):
    q, _ = _samples(2)
    eng.surrogate_init(q.shape[1], seed=5)
    P, _, _ = eng.surrogate_get()
    eng.surrogate_set(P, 0.1, 2.0)
    e = eng.surrogate_predict(q)
    e_o = oracle.mlp(P, q, 0.1, 2.0, round_bf16=True)
    print(f"n={len(q)} max|de|={np.abs(e - e_o).max():.3e}", flush=True)
    assert np.abs(e - e_o).max() <= 1e-5 * (np.abs(e_o).max() + 2.0)


def test_surrogate_gradient_matches_oracle(eng):
    q, y = _samples(2)
    eng.surrogate_init(q.shape[1], seed=6)
    P, _, _ = eng.surrogate_get()
    eng.surrogate_set(P, float(y.mean()), float(y.std()))
    g, loss = eng.surrogate_gradient(q, y)
    _, loss_o, g_o = oracle.mlp(P, q, float(y.mean()), float(y.std()), round_bf16=True, y=y,
                                grad=True)
    assert loss == pytest.approx(loss_o, rel=1e-4)
    assert np.abs(g - g_o).max() <= 1e-3 * np.abs(g_o).max()


def test_surrogate_training_gain(eng):
    """Adam on the device kernels: held-out RMSE of the trained network
    against the untrained one and against predicting the mean."""
    q, y = _samples(8)
    rng = np.random.default_rng(0)
    idx = rng.permutation(len(q))
    tr, te = idx[: 3 * len(q) // 4], idx[3 * len(q) // 4:]
    eng.surrogate_init(q.shape[1], seed=7)
    P, _, _ = eng.surrogate_get()
    eng.surrogate_set(P, float(y[tr].mean()), float(y[tr].std()))
    rmse0 = np.sqrt(np.mean((eng.surrogate_predict(q[te]) - y[te]) ** 2))
    eng.surrogate_fit(q[tr], y[tr], epochs=80, batch=128, lr=2e-3, seed=1)
    rmse1 = np.sqrt(np.mean((eng.surrogate_predict(q[te]) - y[te]) ** 2))
    spread = y[te].std()
    print(f"spread {spread:.4g} untrained {rmse0:.4g} trained {rmse1:.4g}", flush=True)
    assert rmse1 < 0.35 * spread and rmse1 < 0.35 * rmse0


def _mc_system():
    pos, sp, lat = S.carbon_cell(128, 2.0, 1.2, 1004)
    _, _, _, model = W.carbon216()
    return pos, sp, lat, model


def _fitted(eng, model, seed=3):
    q, y = _samples(6, seed=70)
    eng.surrogate_init(model.hypers.dim, seed=seed)
    eng.surrogate_fit(q, y, epochs=40, batch=128, lr=2e-3, seed=2)
    return eng.surrogate_get()


def test_delayed_acceptance_fp64_matches_oracle(eng):
    pos, sp, lat, model = _mc_system()
    eng.set_model(model)
    P, mu, sigma = _fitted(eng, model)
    n_trials = 120
    g = eng.mc_run_da(pos, sp, lat, n_trials, 1500.0, 0.6, 1.2, 99, n_train_trials=0)
    p_o, tr_o, acc_o, r1_o = oracle.mc_run_da(model, pos, sp, lat, n_trials, 1500.0, 0.6, 99,
                                              P, mu, sigma)
    print(f"acc {g['n_accepted']}/{acc_o} stage-1 rejections {g['n_stage1_rejected']}/{r1_o}",
          flush=True)
    assert g["n_accepted"] == acc_o and g["n_stage1_rejected"] == r1_o
    assert g["n_gap_evaluations"] == n_trials - r1_o
    assert np.array_equal(g["pos"], p_o)
    assert np.abs(g["e_trace"] - tr_o).max() < 1e-8


def test_delayed_acceptance_online_training(eng):
    """Trained on the run's own first 1,000 trials, the surrogate then
    rejects trials without a GAP evaluation; the chain stays exact (energy
    trace equals a fresh evaluation) and its acceptance matches plain
    Metropolis within sampling noise."""
    pos, sp, lat, model = _mc_system()
    eng.set_model(model)
    eng.surrogate_init(model.hypers.dim, seed=11)
    n = 3000
    g = eng.mc_run_da(pos, sp, lat, n, 1500.0, 0.6, 1.2, 5, n_train_trials=1000,
                      train_every=256, epochs=4, batch=256, lr=2e-3, max_rmse=0.3)
    plain = eng.mc_run(pos, sp, lat, n, 1500.0, 0.6, 1.2, 5)
    print({k: v for k, v in g.items() if k not in ("pos", "e_trace")},
          "plain", plain["n_accepted"], flush=True)
    assert g["fallback"] == 0 and g["n_stage1_rejected"] > 0
    assert g["n_gap_evaluations"] < n
    e, _, _ = oracle.energy_forces(model, g["pos"], sp, lat, do_forces=False)
    assert abs(e.sum() - g["e_trace"][-1]) < 1e-8
    a_da, a_pl = g["n_accepted"] / n, plain["n_accepted"] / n
    assert abs(a_da - a_pl) < 0.06


@pytest.mark.parametrize("max_rmse", [1e-9, 2e-3])
def test_delayed_acceptance_fallback_matches_oracle(eng, max_rmse):
    """The drift fallback: once the running mean square of the stage-2
    energy difference exceeds max_rmse^2 the run continues as plain
    Metropolis.  With max_rmse ~ 0 it switches after the first GAP-evaluated
    trial; the chain (before and after the switch) equals the oracle's
    two-stage chain with the same rule, move for move."""
    pos, sp, lat, model = _mc_system()
    eng.set_model(model)
    P, mu, sigma = _fitted(eng, model)
    n_trials = 150
    print(f"max_rmse={max_rmse}", flush=True)
    g = eng.mc_run_da(pos, sp, lat, n_trials, 1500.0, 0.6, 1.2, 77, n_train_trials=0,
                      max_rmse=max_rmse)
    p_o, tr_o, acc_o, r1_o, fb = oracle.mc_run_da(model, pos, sp, lat, n_trials, 1500.0, 0.6,
                                                  77, P, mu, sigma, max_rmse=max_rmse,
                                                  with_fallback=True)
    print(f"fallback at {fb}: gpu {g['fallback']} acc {g['n_accepted']}/{acc_o} "
          f"stage-1 rejections {g['n_stage1_rejected']}/{r1_o}", flush=True)
    if max_rmse < 1e-6:
        assert 0 < fb < n_trials
    assert g["fallback"] == (1 if fb >= 0 else 0)
    assert g["n_accepted"] == acc_o and g["n_stage1_rejected"] == r1_o
    assert g["n_gap_evaluations"] == n_trials - r1_o
    assert np.array_equal(g["pos"], p_o)
    assert np.abs(g["e_trace"] - tr_o).max() < 1e-8
