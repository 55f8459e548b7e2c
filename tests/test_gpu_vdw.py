"""Tkatchenko-Scheffler dispersion on the GPU against the oracle's
restatement of get_ts_energy_and_forces (vdw/vdw_interface.f90:127-385) fed
with the oracle's Hirshfeld volumes and their derivatives
(local_properties.f90:46-146): FP64 energies within 1e-12 eV, forces
within 1e-10 eV/A and the virial within 1e-9 eV; mixed precision within the
volumes' FP32/BF16 rounding."""
import numpy as np
import pytest

import oracle
import structures as S
import tgx
import workloads as W

pytestmark = pytest.mark.gpu

HARTREE, BOHR = 27.211386024367243, 0.5291772105638411
# (C6 eV A^6, R0 A, alpha0 A^3) per species: C and H as get_vdw_ref_params
# sets them, O from the TS tables
REF = np.array([[46.6 * HARTREE * BOHR ** 6, 1.900, 12.0 * BOHR ** 3],
                [6.5 * HARTREE * BOHR ** 6, 1.64, 4.5 * BOHR ** 3],
                [15.6 * HARTREE * BOHR ** 6, 1.71, 5.4 * BOHR ** 3]])


@pytest.fixture(scope="module")
def eng():
    e = tgx.Engine(0, "fp64")
    yield e
    e.close()


def _volume_model(h, pos, sp, lat, seed=5, n_sparse=60, scale=0.05, v0=1.0):
    """A Hirshfeld-volume-like model: V = 1 + small SOAP-kernel variation."""
    rng = np.random.default_rng(seed)
    q = oracle.soap(h, pos, sp, lat)["soap"]
    Qs = q[rng.choice(len(q), n_sparse, replace=len(q) < n_sparse)]
    return Qs, scale * rng.normal(0.0, 0.5, n_sparse), v0


def _run(eng, model, block_hypers, pos, sp, lat, kw, block=0, lp=None, do_forces=True):
    eng.set_model(model)
    eng.clear_local_properties()
    Qs, al, v0 = lp if lp is not None else _volume_model(block_hypers, pos, sp, lat)
    k = eng.add_local_property(block, Qs, al, 0.9, 3.0, v0)
    ns = int(sp.max())
    eng.set_vdw_ts(k, REF[:ns, 0], REF[:ns, 1], REF[:ns, 2], **kw)
    e, f, v = eng.vdw_energy_forces(pos, sp, lat, do_forces=do_forces)
    rl = max(kw.get("rcut", 10.0), block_hypers.rcut_max)
    V, der = oracle.local_property(block_hypers, Qs, al, 0.9, 3.0, v0, pos, sp, lat, rl,
                                   derivatives=do_forces)
    e0, f0, v0_ = oracle.ts_vdw(REF[:ns, 0], REF[:ns, 1], REF[:ns, 2], pos, sp, lat, V, der,
                                rcut_list=rl, do_forces=do_forces, **kw)
    return (e, f, v), (e0, f0, v0_), V


def _check(got, ref, tol_e=1e-12, tol_f=1e-10, tol_v=1e-9):
    (e, f, v), (e0, f0, v0) = got, ref
    de = np.abs(e - e0).max()
    print(f"TS |dE| {de:.2e} |E| {np.abs(e0).max():.3e}", end="")
    assert de < tol_e * max(1.0, np.abs(e0).max())
    if f is not None:
        df, dv = np.abs(f - f0).max(), np.abs(v - v0).max()
        print(f" |dF| {df:.2e} |F| {np.abs(f0).max():.3e} |dV| {dv:.2e}", flush=True)
        assert df < tol_f * max(1.0, np.abs(f0).max())
        assert dv < tol_v * max(1.0, np.abs(v0).max())


def test_ts_carbon216_default_cutoffs_fp64(eng):
    """216 C at 2.0 g/cm^3 with the default cutoffs (rcut 10 A: a supercell
    of the 12.9 A box, buffers at both ends populated)."""
    pos, sp, lat, model = W.carbon216()
    got, ref, V = _run(eng, model, model.hypers, pos, sp, lat, {})
    assert np.all(V > 0)
    _check(got, ref)
    # energies only
    e, f, _ = eng.vdw_energy_forces(pos, sp, lat, do_forces=False)
    assert f is None and np.abs(e - got[0]).max() < 1e-13


def test_ts_cho_clipped_volumes_and_central_species_block(eng):
    """C/H/O with volumes on the second block (central species 2, so C and
    O atoms have V = 0 and no dispersion) and some H volumes clipped."""
    pos, sp, lat, model = W.cho()
    h2 = S.cho_hypers()
    h2.central_species = 2
    q2 = oracle.soap(h2, pos, sp, lat)["soap"]
    m2 = S.model_from_soaps(h2, q2, 50, seed=4)
    full = tgx.GapModel(hypers=model.hypers, Qs=model.Qs, alphas=model.alphas, delta=model.delta,
                        zeta=model.zeta, e0=model.e0, more_soap=[m2])
    Qs, al, _ = _volume_model(h2, pos, sp, lat, seed=9, n_sparse=40, scale=3.0)
    e_k, _ = oracle.gap_kernel_step(q2, Qs, al, 0.9, 3.0)
    v0 = -float(np.quantile(e_k, 0.2))  # a fifth of the H volumes clip to 0
    kw = dict(rcut=6.0, buffer=1.0, rcut_inner=0.5, buffer_inner=0.5)
    got, ref, V = _run(eng, full, h2, pos, sp, lat, kw, block=1, lp=(Qs, al, v0))
    assert np.all(V[sp != 2] == 0) and np.sum((sp == 2) & (V == 0)) > 0
    assert np.all(got[0][V == 0] == 0)
    _check(got, ref)


def test_ts_small_and_triclinic_cells(eng):
    _, _, _, model = W.carbon216()
    kw = dict(rcut=7.0, buffer=1.0, rcut_inner=0.5, buffer_inner=0.5)
    pos, sp, lat = S.carbon_cell(24, 2.6, 1.2, 61)
    got, ref, _ = _run(eng, model, model.hypers, pos, sp, lat, kw)
    _check(got, ref)
    pos, sp, lat = S.triclinic_cell(40, 2.4, 1.1, 31, skew=(0.3, -0.2, 0.4), rotate=True)
    got, ref, _ = _run(eng, model, model.hypers, pos, sp, lat, kw)
    _check(got, ref)


def test_ts_mixed_precision(eng):
    pos, sp, lat, model = W.carbon216()
    eng.set_precision("mixed")
    try:
        got, ref, _ = _run(eng, model, model.hypers, pos, sp, lat, {})
    finally:
        eng.set_precision("fp64")
    (e, f, v), (e0, f0, v0) = got, ref
    print("mixed TS", np.abs(e - e0).max(), np.abs(f - f0).max(), np.abs(v - v0).max())
    # the volumes carry the BF16 rounding of q.Qs (relative ~2^-9, raised to
    # the power zeta = 3) and C6 goes as V^2: about 0.5 % of the largest
    # force, so the bound is 1 % (energies 0.2 %)
    assert np.abs(e - e0).max() < 2e-3 * np.abs(e0).max()
    assert np.abs(f - f0).max() < 1e-2 * np.abs(f0).max()
    assert np.abs(v - v0).max() < 1e-2 * np.abs(v0).max()


def test_ts_errors(eng):
    pos, sp, lat, model = W.carbon216()
    eng.set_model(model)
    eng.clear_local_properties()
    with pytest.raises(tgx.TgxError) as ei:
        eng.vdw_energy_forces(pos, sp, lat)
    assert ei.value.code == tgx.TGX_ERR_STATE
    with pytest.raises(tgx.TgxError):
        eng.set_vdw_ts(0, REF[:1, 0], REF[:1, 1], REF[:1, 2])
