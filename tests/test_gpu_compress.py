"""compress_soap models on the GPU (descriptor q = P p / |P p|, the P
transformation read by read_gap.f90:485-549) against the oracle: FP64
descriptors within 5e-10, energies within 1e-8 eV/atom, forces within 1e-6
eV/A and the virial within 1e-6 eV; FP32 and mixed within their usual
bounds; a compressed block beside an uncompressed one; the no-allocation
entry point and velocity-Verlet MD on a compressed model."""
import dataclasses

import numpy as np
import pytest

import oracle
import structures as S
import tgx
import workloads as W

pytestmark = pytest.mark.gpu


def _random_P(h, dim, seed, per_row=3):
    rng = np.random.default_rng(seed)
    pi, pj, pe = [], [], []
    for i in range(1, dim + 1):
        for j in rng.choice(h.full_dim, per_row, replace=False):
            pi.append(i)
            pj.append(j + 1)
            pe.append(rng.normal())
    return np.array(pi), np.array(pj), np.array(pe)


def _compressed_model(h, pos, sp, lat, dim=96, seed=2, n_sparse=120, two_body=None):
    hc = dataclasses.replace(h, compress_P=_random_P(h, dim, seed), compress_dim=dim)
    q = oracle.soap(hc, pos, sp, lat)["soap"]
    return hc, S.model_from_soaps(hc, q, n_sparse, seed=seed, two_body=two_body or ())


@pytest.fixture(scope="module")
def carbon():
    pos, sp, lat, base = W.carbon216()
    hc, model = _compressed_model(base.hypers, pos, sp, lat, two_body=[S.carbon_2b(1000)])
    return pos, sp, lat, hc, model


def _check(e, f, v, e0, f0, v0, te=1e-8, tf=1e-6, tv=1e-6):
    de, df, dv = np.abs(e - e0).max(), np.abs(f - f0).max(), np.abs(v - v0).max()
    print(f"|dE| {de:.2e} |dF| {df:.2e} |dV| {dv:.2e}", flush=True)
    assert de < te and df < tf and dv < tv


def test_compressed_descriptors_and_energy_forces_fp64(carbon):
    pos, sp, lat, hc, model = carbon
    eng = tgx.Engine(0, "fp64")
    eng.set_model(model)
    q, _ = eng.soap(pos, sp, lat)
    q0 = oracle.soap(hc, pos, sp, lat)["soap"]
    assert q.shape == (len(pos), 96)
    print("|dq|", np.abs(q - q0).max())
    # the uncompressed FP64 descriptors agree to about 1e-10; P's random-sign
    # weights and the renormalisation by |P q_f| < |q_f| amplify that
    assert np.abs(q - q0).max() < 5e-10
    e, f, v = eng.energy_forces(pos, sp, lat)
    _check(e, f, v, *oracle.energy_forces(model, pos, sp, lat))
    # a small cell (supercell images) and the no-allocation entry point
    p2, s2, l2 = S.carbon_cell(24, 2.6, 1.2, 61)
    e0, f0, v0 = oracle.energy_forces(model, p2, s2, l2)
    _check(*eng.energy_forces(p2, s2, l2), e0, f0, v0)
    _check(*eng.energy_forces_noalloc(p2, s2, l2), e0, f0, v0)
    eng.close()


@pytest.mark.parametrize("prec", ["fp32", "mixed"])
def test_compressed_reduced_precision(carbon, prec):
    pos, sp, lat, hc, model = carbon
    eng = tgx.Engine(0, prec)
    eng.set_model(model)
    e, f, v = eng.energy_forces(pos, sp, lat)
    e0, f0, v0 = oracle.energy_forces(model, pos, sp, lat)
    mae = np.abs(f - f0).mean()
    print(prec, "|dE|", np.abs(e - e0).max(), "force MAE", mae, "max", np.abs(f - f0).max())
    if prec == "fp32":
        assert np.abs(e - e0).max() < 1e-5 and np.abs(f - f0).max() < 1e-4
    else:
        assert mae < 1e-3 and np.abs(e - e0).max() < 5e-3
    eng.close()


def test_compressed_block_beside_uncompressed_cho():
    """Block 0 uncompressed C/H/O; block 1 compressed with central species 2."""
    pos, sp, lat, model = W.cho()
    h2 = S.cho_hypers()
    h2.central_species = 2
    hc, m2 = _compressed_model(h2, pos, sp, lat, dim=150, seed=7, n_sparse=60)
    full = tgx.GapModel(hypers=model.hypers, Qs=model.Qs, alphas=model.alphas, delta=model.delta,
                        zeta=model.zeta, e0=model.e0, more_soap=[m2])
    eng = tgx.Engine(0, "fp64")
    eng.set_model(full)
    e, f, v = eng.energy_forces(pos, sp, lat)
    _check(e, f, v, *oracle.energy_forces(full, pos, sp, lat))
    eng.close()


def test_compressed_md_matches_oracle_vv(carbon):
    pos, sp, lat, hc, model = carbon
    masses = np.full(len(pos), 12.011)
    vel = S.maxwell_boltzmann(len(pos), masses, 3000.0, seed=5)
    eng = tgx.Engine(0, "fp64")
    eng.set_model(model)
    out = eng.md_run(pos, vel, sp, lat, masses, 6, 1.0, skin=0.5)
    m = masses * 103.6426965268
    p, vv = pos.copy(), vel.copy()
    pprev, fprev = np.zeros_like(p), np.zeros_like(p)
    _, f, _ = oracle.energy_forces(model, p, sp, lat)
    dtp = 1.0
    for s in range(6):
        dtp = oracle.velocity_verlet(p, pprev, vv, f, fprev, m, 1.0, dtp, s == 0)
        _, f, _ = oracle.energy_forces(model, p, sp, lat)
    print("dpos", np.abs(out["pos"] - p).max())
    assert np.abs(out["pos"] - p).max() < 1e-9
    assert np.abs(out["forces"] - f).max() < 1e-6
    eng.close()


def test_compressed_refused_by_monte_carlo(carbon):
    pos, sp, lat, hc, model = carbon
    eng = tgx.Engine(0, "fp64")
    eng.set_model(model)
    with pytest.raises(tgx.TgxError) as ei:
        eng.mc_run(pos, sp, lat, 4, 1500.0, 0.6, 1.2, 1)
    assert ei.value.code == tgx.TGX_ERR_UNSUPPORTED
    eng.close()
