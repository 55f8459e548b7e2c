"""Every benchmark configuration at its own size, through the C-ABI, against
the FP64 oracle (SOAP + kernel step gap/gap.f90:39-244, 2-body
gap/gap.f90:329-460, velocity Verlet md/md_utils.f90:486-532, Metropolis
mc/mc_interface.f90:245-299).

Tolerances (written per check below): FP64 mode 1e-8 eV/atom, 1e-6 eV/A and
a virial within 1e-9 eV per atom; FP32 mode 1e-5 eV/atom energy,
force MAE 1e-5 eV/A; mixed mode (FP32 SOAP, BF16 kernel GEMM) force MAE
<= 1 meV/A, energy <= 10 meV/atom, virial <= 2 meV per atom."""
import numpy as np
import pytest

import oracle
import structures as S
import tgx
import workloads as W

pytestmark = pytest.mark.gpu

AMU = 103.6426965268
TOL = {
    "fp64": dict(e=1e-8, f_max=1e-6, f_mae=1e-6, v_atom=1e-9),
    "fp32": dict(e=1e-5, f_max=1e-3, f_mae=1e-5, v_atom=1e-5),
    "mixed": dict(e=1e-2, f_max=3e-2, f_mae=1e-3, v_atom=2e-3),
}


def _check(prec, got, want, n, label):
    (e, f, v), (e0, f0, v0) = got, want
    de, dfm, dfa = np.abs(e - e0).max(), np.abs(f - f0).max(), np.abs(f - f0).mean()
    dv = np.abs(v - v0).max() / n
    print(f"{label} {prec}: |dE| {de:.3g} eV/atom, |dF| max {dfm:.3g} MAE {dfa:.3g} eV/A, "
          f"|dV| {dv:.3g} eV/atom", flush=True)
    t = TOL[prec]
    assert de <= t["e"] and dfm <= t["f_max"] and dfa <= t["f_mae"] and dv <= t["v_atom"]


# ---- config 1: C/H/O 4,096 atoms, D = 2,700, S = 4,000 ---------------------

@pytest.fixture(scope="module")
def cho4096():
    pos, sp, lat = S.cho_cell(4096, 1.5, 1.0, 1001)
    h = S.cho_hypers()
    q = oracle.soap(h, pos, sp, lat)["soap"]
    model = S.model_from_soaps(h, q, 4000, seed=1001)
    want = oracle.energy_forces(model, pos, sp, lat, chunk_sites=512)
    return pos, sp, lat, model, want


@pytest.mark.parametrize("prec", ["fp64", "fp32", "mixed"])
def test_cfg1_cho4096(cho4096, prec):
    pos, sp, lat, model, want = cho4096
    assert model.hypers.dim == 2700 and len(model.alphas) == 4000
    eng = tgx.Engine(0, prec)
    eng.set_model(model)
    print(f"cfg1 {prec}: {len(pos)} atoms", flush=True)
    _check(prec, eng.energy_forces(pos, sp, lat), want, len(pos), "cfg1")
    eng.close()


# ---- config 2: 256 structures of 32-1,024 atoms, S = 2,000 ------------------

@pytest.fixture(scope="module")
def batch256():
    structs = S.batch_cells(256, 1002)
    h = S.carbon_hypers()
    qs = [oracle.soap(h, p, s, l)["soap"] for p, s, l in structs[:64]]
    model = S.model_from_soaps(h, np.concatenate(qs), 2000, seed=1002,
                               two_body=[S.carbon_2b(1002)])
    want = [oracle.energy_forces(model, p, s, l) for p, s, l in structs]
    return structs, model, want


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_cfg2_batch256(batch256, prec):
    structs, model, want = batch256
    eng = tgx.Engine(0, prec)
    eng.set_model(model)
    n = sum(len(p) for p, _, _ in structs)
    print(f"cfg2 {prec}: {len(structs)} structures, {n} atoms", flush=True)
    offs, e, f, v = eng.energy_forces_batched(structs)
    e0 = np.concatenate([w[0] for w in want])
    f0 = np.concatenate([w[1] for w in want])
    v0 = np.stack([w[2] for w in want])
    natoms = np.diff(offs)
    dv = np.max(np.abs(v - v0).reshape(len(structs), -1).max(1) / natoms)
    de, dfm, dfa = np.abs(e - e0).max(), np.abs(f - f0).max(), np.abs(f - f0).mean()
    print(f"cfg2 {prec}: |dE| {de:.3g} eV/atom, |dF| max {dfm:.3g} MAE {dfa:.3g}, "
          f"|dV| {dv:.3g} eV/atom", flush=True)
    t = TOL[prec]
    assert de <= t["e"] and dfm <= t["f_max"] and dfa <= t["f_mae"] and dv <= t["v_atom"]
    eng.close()


# ---- config 3: 32,768 C atoms, E+F and NVE MD -------------------------------

@pytest.fixture(scope="module")
def carbon32k():
    pos, sp, lat = S.carbon_cell(32768, 2.0, 1.2, 1003)
    _, _, _, model = W.carbon216()
    want = oracle.energy_forces(model, pos, sp, lat, chunk_sites=4096)
    return pos, sp, lat, model, want


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_cfg3_energy_forces_32k(carbon32k, prec):
    pos, sp, lat, model, want = carbon32k
    eng = tgx.Engine(0, prec)
    eng.set_model(model)
    print(f"cfg3 {prec}: {len(pos)} atoms", flush=True)
    _check(prec, eng.energy_forces(pos, sp, lat), want, len(pos), "cfg3")
    eng.close()


def test_cfg3_md_fp64_two_steps_vs_oracle(carbon32k):
    """Two FP64 velocity-Verlet steps of the 32,768-atom cell at 3,000 K
    (captured step graph) against the oracle's VV on oracle forces."""
    pos, sp, lat, model, want = carbon32k
    n = len(pos)
    masses = np.full(n, 12.011)
    vel = S.maxwell_boltzmann(n, masses, 3000.0, seed=1003)
    eng = tgx.Engine(0, "fp64")
    eng.set_model(model)
    print("cfg3 md fp64: 2 steps", flush=True)
    out = eng.md_run(pos, vel, sp, lat, masses, 2, 1.0, skin=0.5)
    eng.close()
    m = masses * AMU
    p, v = pos.copy(), vel.copy()
    pprev, fprev = np.zeros_like(p), np.zeros_like(p)
    f = want[1].copy()
    dtp = 1.0
    for s in range(2):
        dtp = oracle.velocity_verlet(p, pprev, v, f, fprev, m, 1.0, dtp, s == 0)
        e, f, _ = oracle.energy_forces(model, p, sp, lat, chunk_sites=4096)
    print("dpos", np.abs(out["pos"] - p).max(), "dF", np.abs(out["forces"] - f).max(),
          "dEp", abs(out["e_pot"][-1] - e.sum()), flush=True)
    assert np.abs(out["pos"] - p).max() < 1e-9
    assert np.abs(out["forces"] - f).max() < 1e-6
    assert abs(out["e_pot"][-1] - e.sum()) < 1e-8 * n


def test_cfg3_md_mixed_drift(carbon32k):
    """200 mixed-precision steps of the 32,768-atom cell at 3,000 K with the
    0.5 A skin: the list is rebuilt and the total energy drifts by less than
    2e-4 eV/atom (the 1,000-step benchmark run drifts -6e-5 eV/atom)."""
    pos, sp, lat, model, _ = carbon32k
    n = len(pos)
    masses = np.full(n, 12.011)
    vel = S.maxwell_boltzmann(n, masses, 3000.0, seed=1003)
    eng = tgx.Engine(0, "mixed")
    eng.set_model(model)
    print("cfg3 md mixed: 200 steps", flush=True)
    eng.md_begin(pos, vel, sp, lat, masses, 1.0, 0.5, history=256)
    eng.md_steps(200)
    out = eng.md_read()
    eng.md_end()
    eng.close()
    et = out["e_pot"] + out["e_kin"]
    drift = (et[-1] - et[0]) / n
    print(f"rebuilds {out['n_rebuilds']} drift {drift:.3g} eV/atom", flush=True)
    assert out["steps"] == 200 and out["n_rebuilds"] >= 10
    assert abs(drift) < 2e-4


# ---- config 4: Metropolis MC on 512 C atoms at 1,500 K ----------------------

@pytest.fixture(scope="module")
def carbon512():
    pos, sp, lat = S.carbon_cell(512, 2.0, 1.2, 1004)
    _, _, _, model = W.carbon216()
    return pos, sp, lat, model


def test_cfg4_mc_fp64_chain_vs_oracle(carbon512):
    """300 FP64 trials (move_max 0.6: per-component bound 0.3 A, skin 1.2 A)
    equal the oracle's full-recomputation chain move for move."""
    pos, sp, lat, model = carbon512
    eng = tgx.Engine(0, "fp64")
    eng.set_model(model)
    print("cfg4 fp64 chain: 300 trials", flush=True)
    g = eng.mc_run(pos, sp, lat, 300, 1500.0, 0.6, 1.2, 4004)
    eng.close()
    p_o, tr_o, acc_o = oracle.mc_run(model, pos, sp, lat, 300, 1500.0, 0.6, 4004)
    print(f"accepted {g['n_accepted']}/{acc_o}", flush=True)
    assert g["n_accepted"] == acc_o
    assert np.array_equal(g["pos"], p_o)
    assert np.abs(g["e_trace"] - tr_o).max() < 1e-8 * len(pos)


def test_cfg4_delayed_acceptance_20k_training(carbon512):
    """Mixed precision: the surrogate trained on the chain's own first 20,000
    trials, then 6,000 delayed-acceptance trials against 6,000 plain ones from
    the same state.  Acceptance rates agree within sampling noise, the
    surrogate rejects trials without a GAP evaluation, and the final energy
    of the chain equals a fresh evaluation of its final positions."""
    pos, sp, lat, model = carbon512
    eng = tgx.Engine(0, "mixed")
    eng.set_model(model)
    eng.surrogate_init(model.hypers.dim, seed=4)
    print("cfg4 DA: 20000 training trials", flush=True)
    tr = eng.mc_run_da(pos, sp, lat, 20000, 1500.0, 0.6, 1.2, 1, n_train_trials=20000,
                       train_every=2048, epochs=3, batch=512, lr=2e-3, max_rmse=0.2)
    start = tr["pos"]
    k = 6000
    pl = eng.mc_run(start, sp, lat, k, 1500.0, 0.6, 1.2, 3)
    da = eng.mc_run_da(start, sp, lat, k, 1500.0, 0.6, 1.2, 3, n_train_trials=0, max_rmse=0.2)
    a_pl, a_da = pl["n_accepted"] / k, da["n_accepted"] / k
    print(f"train rmse {tr['train_rmse']:.3g} eV/site, acceptance plain {a_pl:.4f} DA {a_da:.4f}, "
          f"stage-1 rejections {da['n_stage1_rejected']}, fallback {da['fallback']}", flush=True)
    assert da["fallback"] == 0 and da["n_stage1_rejected"] > 0
    # two chains of 6,000 correlated trials (binomial standard error 0.009)
    assert abs(a_da - a_pl) < 0.05
    # the chain's running energy (site energies of the trials, FP32 partial
    # sums) against a fresh full mixed evaluation: the descriptors of a site
    # may round to a neighbouring BF16 value in one path and not the other
    e_fresh, _, _ = eng.energy_forces(da["pos"], sp, lat, do_forces=False)
    print(f"trace - fresh {da['e_trace'][-1] - e_fresh.sum():.3g} eV", flush=True)
    assert abs(da["e_trace"][-1] - e_fresh.sum()) < 1e-4 * len(pos)
    e_o, _, _ = oracle.energy_forces(model, da["pos"], sp, lat, do_forces=False)
    assert abs(e_fresh.sum() - e_o.sum()) / len(pos) < 2e-3
    eng.close()
