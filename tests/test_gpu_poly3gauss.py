"""The poly3gauss radial basis on the GPU (a normalised Gaussian of width
atom_sigma_r at the origin, then n_max - 1 poly3 functions, orthonormalised
by W = S^(-1/2)) against the oracle: FP64 descriptors within 1e-10 and
FP32 within 2e-5; energies within 1e-8 eV/atom, forces within 1e-6 eV/A and
the virial within 1e-6 eV in FP64; the FP32 table path, C/H/O, radial
enhancement with r-dependent widths, and the FP64 Monte Carlo chain move
for move."""
import dataclasses

import numpy as np
import pytest

import oracle
import structures as S
import tgx
import workloads as W

pytestmark = pytest.mark.gpu


def _gauss(h, **kw):
    return dataclasses.replace(h, basis="poly3gauss", **kw)


@pytest.fixture(scope="module")
def carbon():
    pos, sp, lat, base = W.carbon216()
    h = _gauss(base.hypers)
    q = oracle.soap(h, pos, sp, lat)["soap"]
    model = S.model_from_soaps(h, q, 200, seed=3, two_body=[S.carbon_2b(1000)])
    return pos, sp, lat, h, model


def _check(got, ref, te=1e-8, tf=1e-6, tv=1e-6):
    (e, f, v), (e0, f0, v0) = got, ref
    de, df, dv = np.abs(e - e0).max(), np.abs(f - f0).max(), np.abs(v - v0).max()
    print(f"|dE| {de:.2e} |dF| {df:.2e} |dV| {dv:.2e}", flush=True)
    assert de < te and df < tf and dv < tv


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("fp32", 2e-5)])
def test_poly3gauss_descriptors(carbon, prec, tol):
    pos, sp, lat, h, model = carbon
    eng = tgx.Engine(0, prec)
    eng.set_model(model)
    q, _ = eng.soap(pos, sp, lat)
    q0 = oracle.soap(h, pos, sp, lat)["soap"]
    print(prec, "|dq|", np.abs(q - q0).max())
    assert np.abs(q - q0).max() < tol
    eng.close()


def test_poly3gauss_energy_forces_fp64_and_mixed(carbon):
    pos, sp, lat, h, model = carbon
    eng = tgx.Engine(0, "fp64")
    eng.set_model(model)
    ref = oracle.energy_forces(model, pos, sp, lat)
    _check(eng.energy_forces(pos, sp, lat), ref)
    eng.set_precision("fp32")
    e, f, _ = eng.energy_forces(pos, sp, lat)
    print("fp32 |dE|", np.abs(e - ref[0]).max(), "|dF|", np.abs(f - ref[1]).max())
    assert np.abs(e - ref[0]).max() < 1e-5 and np.abs(f - ref[1]).max() < 1e-4
    eng.set_precision("mixed")
    e, f, _ = eng.energy_forces(pos, sp, lat)
    mae = np.abs(f - ref[1]).mean()
    print("mixed force MAE", mae, "|dE|", np.abs(e - ref[0]).max())
    assert mae < 1e-3 and np.abs(e - ref[0]).max() < 5e-3
    eng.close()


def test_poly3gauss_cho_enhancement_and_sigma_scaling():
    """C/H/O, radial_enhancement 1, r-dependent radial and angular widths."""
    pos, sp, lat = S.cho_cell(200, 1.5, 1.0, 91)
    h = _gauss(S.cho_hypers(), radial_enhancement=1, atom_sigma_r_scaling=(0.05,) * 3,
               atom_sigma_t_scaling=(0.04,) * 3)
    q = oracle.soap(h, pos, sp, lat)["soap"]
    model = S.model_from_soaps(h, q, 100, seed=5)
    eng = tgx.Engine(0, "fp64")
    eng.set_model(model)
    qg, _ = eng.soap(pos, sp, lat)
    print("|dq|", np.abs(qg - q).max())
    assert np.abs(qg - q).max() < 1e-10
    _check(eng.energy_forces(pos, sp, lat), oracle.energy_forces(model, pos, sp, lat))
    eng.close()


def test_poly3gauss_monte_carlo_fp64_chain():
    pos, sp, lat = S.carbon_cell(128, 2.0, 1.2, 1004)
    _, _, _, base = W.carbon216()
    h = _gauss(base.hypers)
    model = S.model_from_soaps(h, oracle.soap(h, pos, sp, lat)["soap"], 100, seed=8)
    eng = tgx.Engine(0, "fp64")
    eng.set_model(model)
    g = eng.mc_run(pos, sp, lat, 120, 1500.0, 0.6, 1.2, 31)
    p_o, tr_o, acc_o = oracle.mc_run(model, pos, sp, lat, 120, 1500.0, 0.6, 31)
    print("accepted", g["n_accepted"], acc_o)
    assert g["n_accepted"] == acc_o
    assert np.array_equal(g["pos"], p_o)
    assert np.abs(g["e_trace"] - tr_o).max() < 1e-8
    eng.close()
