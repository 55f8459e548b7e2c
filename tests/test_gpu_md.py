"""GPU molecular dynamics against the oracle's velocity Verlet
(md/md_utils.f90:486-532) driven by oracle forces, and NVE energy checks."""
import numpy as np
import pytest

import oracle
import structures as S
import tgx
import workloads as W

pytestmark = pytest.mark.gpu

AMU = 103.6426965268


def _oracle_md(model, pos, vel, sp, lat, masses_amu, n_steps, dt):
    pos = pos.copy()
    vel = vel.copy()
    m = masses_amu * AMU
    pprev = np.zeros_like(pos)
    fprev = np.zeros_like(pos)
    e, f, _ = oracle.energy_forces(model, pos, sp, lat)
    ep = [e.sum()]
    ek = []
    dtp = dt
    for s in range(n_steps):
        dtp = oracle.velocity_verlet(pos, pprev, vel, f, fprev, m, dt, dtp, s == 0)
        ek.append(0.5 * np.sum(m[:, None] * vel * vel))
        e, f, _ = oracle.energy_forces(model, pos, sp, lat)
        ep.append(e.sum())
    vfin = vel + 0.5 * (f + fprev) / m[:, None] * dt if n_steps else vel
    ek.append(0.5 * np.sum(m[:, None] * vfin * vfin))
    return pos, vfin, f, np.array(ep), np.array(ek)


@pytest.mark.parametrize("graph", ["1", "0"])
def test_md_fp64_matches_oracle(graph, monkeypatch):
    monkeypatch.setenv("TGX_NO_GRAPH", "0" if graph == "1" else "1")
    pos, sp, lat, model = W.carbon216()
    masses = np.full(len(pos), 12.011)
    vel = S.maxwell_boltzmann(len(pos), masses, 3000.0, seed=5)
    eng = tgx.Engine(0, "fp64")
    eng.set_model(model)
    n_steps, dt = 12, 1.0
    out = eng.md_run(pos, vel, sp, lat, masses, n_steps, dt, skin=0.5)
    p0, v0, f0, ep0, ek0 = _oracle_md(model, pos, vel, sp, lat, masses, n_steps, dt)
    print("dpos", np.abs(out["pos"] - p0).max(), "dvel", np.abs(out["vel"] - v0).max(),
          "dEp", np.abs(out["e_pot"] - ep0).max(), "rebuilds", out["n_rebuilds"])
    assert np.abs(out["pos"] - p0).max() < 1e-9
    assert np.abs(out["vel"] - v0).max() < 1e-9
    assert np.abs(out["forces"] - f0).max() < 1e-6
    assert np.abs(out["e_pot"] - ep0).max() < 1e-8 * len(pos)
    assert np.abs(out["e_kin"] - ek0).max() < 1e-8 * len(pos)
    eng.close()


def test_md_skin_rebuilds_and_nve_mixed():
    """Mixed precision, 300 steps at 3000 K with skin 0.3 A: several
    rebuilds, total energy drift small relative to kinetic energy
    fluctuations."""
    pos, sp, lat, model = W.carbon216()
    masses = np.full(len(pos), 12.011)
    vel = S.maxwell_boltzmann(len(pos), masses, 3000.0, seed=6)
    eng = tgx.Engine(0, "mixed")
    eng.set_model(model)
    eng.md_begin(pos, vel, sp, lat, masses, 0.5, 0.3, history=512)
    eng.md_steps(300)
    out = eng.md_read()
    eng.md_end()
    et = out["e_pot"] + out["e_kin"]
    print("rebuilds", out["n_rebuilds"], "drift", et[-1] - et[0], "std", et.std(),
          "ek std", out["e_kin"].std())
    assert out["steps"] == 300
    assert out["n_rebuilds"] >= 2
    assert abs(et[-1] - et[0]) < 0.05 * out["e_kin"].std() + 1e-3
    eng.close()


def test_md_steps_resume_equals_single_run():
    pos, sp, lat, model = W.carbon216()
    masses = np.full(len(pos), 12.011)
    vel = S.maxwell_boltzmann(len(pos), masses, 1000.0, seed=8)
    eng = tgx.Engine(0, "fp64")
    eng.set_model(model)
    eng.md_begin(pos, vel, sp, lat, masses, 1.0, 0.5, history=64)
    eng.md_steps(4)
    a = eng.md_read()
    eng.md_steps(6)
    b = eng.md_read()
    eng.md_end()
    eng.md_begin(pos, vel, sp, lat, masses, 1.0, 0.5, history=64)
    eng.md_steps(10)
    c = eng.md_read()
    eng.md_end()
    assert a["steps"] == 4 and b["steps"] == 10
    assert np.array_equal(b["pos"], c["pos"])
    assert np.array_equal(b["e_pot"], c["e_pot"])
    eng.close()
