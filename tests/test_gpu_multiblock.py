"""Models of several soap_turbo blocks (calculate_gap_soap.f90:324-447: each
block evaluated in turn, energies, forces and virials added) through the
C-ABI against the FP64 oracle: two C/H/O blocks with different central
species, radial and angular sizes, cutoffs, zeta and delta, plus a 2-body
term.

Tolerances: FP64 1e-8 eV/atom, 1e-6 eV/A, virial 1e-9 eV per atom; mixed
precision force MAE <= 1 meV/A, energy <= 10 meV/atom."""
import dataclasses

import numpy as np
import pytest

import oracle
import structures as S
import tgx

pytestmark = pytest.mark.gpu


def _model(n=300, seed=61):
    pos, sp, lat = S.cho_cell(n, 1.5, 1.0, seed)
    h1 = dataclasses.replace(S.cho_hypers(), central_species=1)
    h2 = dataclasses.replace(S.cho_hypers(), central_species=2, l_max=6, alpha_max=(6, 6, 6),
                             rcut_hard=(3.5, 3.5, 3.5))
    m1 = S.model_from_soaps(h1, oracle.soap(h1, pos, sp, lat)["soap"], 150, seed=1)
    m2 = S.model_from_soaps(h2, oracle.soap(h2, pos, sp, lat)["soap"], 80, seed=2, zeta=3.0,
                            delta=0.8)
    two_b = S.carbon_2b(5, rcut=3.0)
    model = dataclasses.replace(m1, e0=np.array([-0.3, -0.1, -0.6]), more_soap=[m2],
                                two_body=[two_b])
    return pos, sp, lat, model


def _check(got, want, n, e_tol, f_tol, v_tol, label):
    """Energy and virial bounds asserted here; returns the force max and MAE."""
    (e, f, v), (e0, f0, v0) = got, want
    de, df, dv = np.abs(e - e0).max(), np.abs(f - f0).max(), np.abs(v - v0).max() / n
    print(f"{label}: |dE| {de:.3g} |dF| {df:.3g} MAE {np.abs(f - f0).mean():.3g} |dV| {dv:.3g}",
          flush=True)
    assert de <= e_tol and dv <= v_tol
    return df, np.abs(f - f0).mean()


def test_two_blocks_fp64_and_mixed():
    pos, sp, lat, model = _model()
    want = oracle.energy_forces(model, pos, sp, lat)
    eng = tgx.Engine(0, "fp64")
    eng.set_model(model)
    assert eng.lib.tgx_n_soap_models(eng.ctx) == 2
    df, _ = _check(eng.energy_forces(pos, sp, lat), want, len(pos), 1e-8, 1e-6, 1e-9, "fp64")
    assert df <= 1e-6
    e_only, _, _ = eng.energy_forces(pos, sp, lat, do_forces=False)
    assert np.abs(e_only - want[0]).max() <= 1e-8
    e_na, f_na, _ = eng.energy_forces_noalloc(pos, sp, lat)
    assert np.abs(f_na - want[1]).max() <= 1e-6
    eng.set_precision("mixed")
    _, mae = _check(eng.energy_forces(pos, sp, lat), want, len(pos), 1e-2, 3e-2, 2e-3, "mixed")
    assert mae <= 1e-3
    eng.close()


def test_two_blocks_batched_and_md():
    pos, sp, lat, model = _model()
    other = S.cho_cell(120, 1.5, 1.0, 62)
    eng = tgx.Engine(0, "fp64")
    eng.set_model(model)
    offs, e, f, v = eng.energy_forces_batched([(pos, sp, lat), other])
    for k, (p, s, l) in enumerate([(pos, sp, lat), other]):
        e0, f0, v0 = oracle.energy_forces(model, p, s, l)
        assert np.abs(e[offs[k]:offs[k + 1]] - e0).max() <= 1e-8
        assert np.abs(f[offs[k]:offs[k + 1]] - f0).max() <= 1e-6
        assert np.abs(v[k] - v0).max() <= 1e-9 * len(p)
    # three FP64 velocity-Verlet steps against the oracle's VV
    masses = np.array([12.011, 1.008, 15.999])[sp - 1]
    vel = S.maxwell_boltzmann(len(pos), masses, 500.0, seed=3)
    out = eng.md_run(pos, vel, sp, lat, masses, 3, 0.5, skin=0.3)
    m = masses * 103.6426965268
    p, vv = pos.copy(), vel.copy()
    pprev, fprev = np.zeros_like(p), np.zeros_like(p)
    _, f, _ = oracle.energy_forces(model, p, sp, lat)
    dtp = 0.5
    for s in range(3):
        dtp = oracle.velocity_verlet(p, pprev, vv, f, fprev, m, 0.5, dtp, s == 0)
        e, f, _ = oracle.energy_forces(model, p, sp, lat)
    assert np.abs(out["pos"] - p).max() < 1e-9
    assert np.abs(out["forces"] - f).max() < 1e-6
    assert abs(out["e_pot"][-1] - e.sum()) < 1e-8 * len(p)
    eng.close()


def test_block_rules():
    pos, sp, lat, model = _model(n=80, seed=63)
    eng = tgx.Engine(0, "fp64")
    eng.set_model(model)
    # Monte Carlo takes one block only
    with pytest.raises(tgx.TgxError) as e:
        eng.mc_run(pos, sp, lat, 4, 1000.0, 0.2, 0.5, 1)
    assert e.value.code == tgx.TGX_ERR_UNSUPPORTED
    # a block on another species list is refused
    h = dataclasses.replace(S.carbon_hypers())
    q = np.zeros((4, h.dim))
    hc = h.to_c()
    rc = eng.lib.tgx_add_soap_model(eng.ctx, hc, h.dim, 4, q.ctypes.data_as(
        tgx.C.POINTER(tgx.C.c_double)), np.zeros(4).ctypes.data_as(tgx.C.POINTER(tgx.C.c_double)),
        1.0, 2.0)
    assert rc == tgx.TGX_ERR_ARG
    assert eng.lib.tgx_n_soap_models(eng.ctx) == 2
    # clearing leaves a usable context: one block again
    assert eng.lib.tgx_clear_soap_models(eng.ctx) == 0
    assert eng.lib.tgx_n_soap_models(eng.ctx) == 0
    eng.set_model(dataclasses.replace(model, more_soap=[]))
    e1, f1, _ = eng.energy_forces(pos, sp, lat)
    e0, f0, _ = oracle.energy_forces(dataclasses.replace(model, more_soap=[]), pos, sp, lat)
    assert np.abs(f1 - f0).max() <= 1e-6
    eng.close()
