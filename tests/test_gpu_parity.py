"""GPU parity: the HIP path through the C-ABI against the FP64 CPU oracle.

Tolerances: neighbour lists bit-exact (ids, rjs, xyz); FP64 mode energies
1e-8 eV/atom, forces 1e-6 eV/A, virial 1e-6 eV; FP32 / mixed modes at the
stated looser bounds.
"""
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


def _nb_equal(a, b, ulps_angles=True):
    assert np.array_equal(a["n_neigh"], b["n_neigh"])
    assert np.array_equal(a["list"], b["list"])
    assert np.array_equal(a["rjs"], b["rjs"])  # bit-exact
    assert np.array_equal(a["xyz"], b["xyz"])
    assert np.allclose(a["thetas"], b["thetas"], rtol=0, atol=4e-16 * np.pi)
    assert np.allclose(a["phis"], b["phis"], rtol=0, atol=4e-16 * np.pi)
    assert np.array_equal(a["indices"], b["indices"])


@pytest.mark.parametrize("case", ["carbon216", "small32", "cho300", "carbon2000", "dense84",
                                  "dense145", "carbon6000", "vacuum1100"])
def test_neighbors_bit_exact(eng, case):
    """Lists of up to 64, 65-128 and over 128 entries (dense84, dense145:
    1,500 C at 3.2 g/cm^3 with rcut 5.0 / 6.0 A) take the three sort paths.
    The one-workgroup build of captured Monte Carlo rebuilds is checked by
    the Monte Carlo chains against the oracle."""
    if case == "carbon216":
        pos, sp, lat, _ = W.carbon216()
        rc = 3.7
    elif case == "small32":
        pos, sp, lat, _ = W.small_box()
        rc = 3.7
    elif case == "cho300":
        pos, sp, lat, _ = W.cho()
        rc = 4.0
    elif case == "carbon2000":
        pos, sp, lat = S.carbon_cell(2000, 2.0, 1.2, 55)
        rc = 4.2
    elif case == "carbon6000":
        pos, sp, lat = S.carbon_cell(6000, 2.0, 1.2, 57)
        rc = 4.2
    elif case == "vacuum1100":
        # a 300-atom cluster straddling the faces of a 1,100 A box (the
        # build for boxes beyond the FP32 screen's range)
        pos, sp, _ = S.carbon_cell(300, 2.0, 1.2, 58)
        lat = np.diag([1100.0] * 3).ravel()
        pos = pos - 7.0
        rc = 3.7
    else:
        pos, sp, lat = S.carbon_cell(1500, 3.2, 1.0, 56)
        rc = 5.0 if case == "dense84" else 6.0
    print(f"case={case} n={len(pos)} rcut={rc}", flush=True)
    g = eng.neighbors(pos, lat, rc)
    o = oracle.neighbors(pos, lat, rc)
    if case == "dense84":
        assert 64 < o["n_neigh"].max() - 1 <= 128
    if case == "dense145":
        assert o["n_neigh"].max() - 1 > 128
    _nb_equal(g, o)


def test_neighbors_unwrapped_positions(eng):
    """Positions outside the box (MD without wrapping) give the same list."""
    pos, sp, lat, _ = W.carbon216()
    L = lat[0]
    rng = np.random.default_rng(3)
    shifted = pos + L * rng.integers(-2, 3, size=pos.shape)
    g = eng.neighbors(shifted, lat, 3.7)
    o = oracle.neighbors(shifted, lat, 3.7)
    _nb_equal(g, o)


def test_soap_fp64_and_fp32(eng):
    pos, sp, lat, model = W.carbon216()
    eng.set_model(model)
    ref = oracle.soap(model.hypers, pos, sp, lat)["soap"]
    q64, site = eng.soap(pos, sp, lat)
    assert np.array_equal(site, np.arange(len(pos)))
    assert np.abs(q64 - ref).max() < 1e-10
    eng.set_precision("fp32")
    try:
        q32, _ = eng.soap(pos, sp, lat)
    finally:
        eng.set_precision("fp64")
    assert np.abs(q32 - ref).max() < 2e-5


@pytest.mark.parametrize("case", ["carbon216", "small32"])
def test_soap_fp32_few_sites_split(eng, case, monkeypatch):
    """The workgroup-per-site FP32 forward kernel of Monte Carlo trials (the
    list split over four waves, partials added in wave order; TGX_FWD_SPLIT=1
    selects it for any call of up to 256 sites) against the one-wave-per-site
    kernel, which sums the pairs in another order, and both against the FP64
    oracle."""
    pos, sp, lat, model = W.carbon216() if case == "carbon216" else W.small_box()
    eng.set_model(model)
    ref = oracle.soap(model.hypers, pos, sp, lat)["soap"]
    eng.set_precision("fp32")
    try:
        q_wave, _ = eng.soap(pos, sp, lat)
        monkeypatch.setenv("TGX_FWD_SPLIT", "1")
        q_split, _ = eng.soap(pos, sp, lat)
        monkeypatch.delenv("TGX_FWD_SPLIT")
    finally:
        eng.set_precision("fp64")
    print(case, len(pos), np.abs(q_split - ref).max(), np.abs(q_split - q_wave).max())
    assert len(pos) <= 256
    assert np.abs(q_split - ref).max() < 2e-5
    assert np.abs(q_wave - ref).max() < 2e-5
    assert np.abs(q_split - q_wave).max() < 2e-6


def test_soap_multispecies(eng):
    pos, sp, lat, model = W.cho()
    eng.set_model(model)
    ref = oracle.soap(model.hypers, pos, sp, lat)["soap"]
    q, _ = eng.soap(pos, sp, lat)
    assert np.abs(q - ref).max() < 1e-10


@pytest.mark.parametrize("prec,etol,gtol", [("fp64", 1e-11, 1e-11), ("fp32", 2e-5, 2e-5),
                                            ("mixed", 5e-2, 5e-2)])
def test_kernel_step(eng, prec, etol, gtol):
    pos, sp, lat, model = W.carbon216()
    eng.set_model(model)
    q = oracle.soap(model.hypers, pos, sp, lat)["soap"]
    e_ref, qss_ref = oracle.gap_kernel_step(q, model.Qs, model.alphas, model.delta, model.zeta)
    eng.set_precision(prec)
    try:
        e, qss = eng.kernel_step(q)
    finally:
        eng.set_precision("fp64")
    scale_e = np.abs(e_ref).max()
    scale_g = np.abs(qss_ref).max()
    print(prec, np.abs(e - e_ref).max() / scale_e, np.abs(qss - qss_ref).max() / scale_g)
    assert np.abs(e - e_ref).max() <= etol * scale_e
    assert np.abs(qss - qss_ref).max() <= gtol * scale_g


@pytest.mark.parametrize("zeta,n_sparse", [(4.0, 500), (2.0, 300), (3.0, 64), (4.0, 1100),
                                           (2.0, 600)])
def test_fused_bf16_kernel_step(eng, zeta, n_sparse, monkeypatch):
    """The single-kernel BF16 step (both GEMMs in one workgroup per 128
    sites; sparse sets above 512 in chunks of 256) against the two separate
    GEMMs.  Both form the same BF16 products
    in the same k order with FP32 accumulation; the single kernel multiplies
    with the operands swapped and sums the site energies in another order,
    so dE/dq and E may differ at FP32 rounding level (1e-6 relative)."""
    pos, sp, lat, model = W.carbon216()
    m2 = S.model_from_soaps(model.hypers, model.Qs, n_sparse, seed=5, zeta=zeta)
    eng.set_model(m2)
    q = oracle.soap(model.hypers, pos, sp, lat)["soap"]
    eng.set_precision("mixed")
    try:
        e1, g1 = eng.kernel_step(q)
        e4, _ = eng.kernel_step(q, want_qss=False)  # energies-only variant
        monkeypatch.setenv("TGX_GEMM_UNFUSED", "1")
        e2, g2 = eng.kernel_step(q)
        e5, _ = eng.kernel_step(q, want_qss=False)
        monkeypatch.delenv("TGX_GEMM_UNFUSED")
        e3, g3 = eng.kernel_step(q[:77])  # a partial block of rows
        # energies of at most 128 sites: one workgroup per 32 sparse points,
        # then the row sums (a repeated call must agree)
        few = [eng.kernel_step(q[:m], want_qss=False)[0] for m in (40, 40, 100, 7)]
    finally:
        eng.set_precision("fp64")
    print("dE/dq bit-identical:", np.array_equal(g1, g2),
          "max rel diff", np.abs(g1 - g2).max() / np.abs(g2).max())
    # E_i = delta^2 sum_s alpha_s K_is^zeta with |K| <= 1: FP32 sums in another
    # order differ by a few ulp of sum_s |alpha_s|
    etol = 1e-6 * m2.delta ** 2 * np.abs(m2.alphas).sum()
    assert np.abs(g1 - g2).max() <= 1e-6 * np.abs(g2).max()
    assert np.abs(e1 - e2).max() <= etol
    assert np.array_equal(e4, e1)
    assert np.abs(e5 - e2).max() <= etol
    assert np.array_equal(g3, g1[:77])
    assert np.abs(e3 - e1[:77]).max() <= etol
    assert np.array_equal(few[0], few[1])
    for m, ef in zip((40, 40, 100, 7), few):
        assert ef.shape == (m,) and np.abs(ef - e1[:m]).max() <= etol, m


def test_kernel_step_fractional_zeta(eng):
    pos, sp, lat, model = W.carbon216()
    m2 = S.model_from_soaps(model.hypers, model.Qs, 300, seed=4, zeta=2.5)
    eng.set_model(m2)
    q = oracle.soap(model.hypers, pos, sp, lat)["soap"]
    e_ref, qss_ref = oracle.gap_kernel_step(q, m2.Qs, m2.alphas, m2.delta, m2.zeta)
    e, qss = eng.kernel_step(q)
    assert np.abs(e - e_ref).max() < 1e-11
    assert np.abs(qss - qss_ref).max() < 1e-11


def _ef_check(eng, model, pos, sp, lat, etol, ftol, vtol, prec="fp64"):
    eng.set_model(model)
    eng.set_precision(prec)
    try:
        e, f, v = eng.energy_forces(pos, sp, lat)
    finally:
        eng.set_precision("fp64")
    e0, f0, v0 = oracle.energy_forces(model, pos, sp, lat)
    de, df, dv = np.abs(e - e0).max(), np.abs(f - f0).max(), np.abs(v - v0).max()
    print(f"{prec} n={len(pos)} dE={de:.3e} dF={df:.3e} dV={dv:.3e} "
          f"F_mae={np.abs(f - f0).mean():.3e}")
    assert de <= etol and df <= ftol and dv <= vtol, (de, df, dv)
    return e, f, v, e0, f0, v0


def test_energy_forces_fp64_carbon216(eng):
    pos, sp, lat, model = W.carbon216()
    _ef_check(eng, model, pos, sp, lat, 1e-8, 1e-6, 1e-6)


def test_energy_forces_fp64_small_box(eng):
    pos, sp, lat, model = W.small_box()
    _ef_check(eng, model, pos, sp, lat, 1e-8, 1e-6, 1e-6)


def test_energy_forces_fp64_cho(eng):
    pos, sp, lat, model = W.cho()
    _ef_check(eng, model, pos, sp, lat, 1e-8, 1e-6, 1e-6)


def test_energy_forces_central_species(eng):
    pos, sp, lat, model = W.cho()
    h = S.cho_hypers()
    h.central_species = 1
    q = oracle.soap(h, pos, sp, lat)["soap"]
    m = S.model_from_soaps(h, q, 100, seed=9)
    _ef_check(eng, m, pos, sp, lat, 1e-8, 1e-6, 1e-6)


def test_energy_forces_mixed_and_fp32(eng):
    """FP32: SOAP and GEMM in FP32.  Mixed: FP32 SOAP, BF16 GEMM (its error is
    the BF16 rounding of q and Qs, reproduced by a numpy emulation): force
    MAE <= 1 meV/A, energy <= 3 meV/atom, virial <= 2 meV per atom."""
    pos, sp, lat, model = W.carbon216()
    n = len(pos)
    _ef_check(eng, model, pos, sp, lat, 1e-5, 1e-4, 1e-4 * n, prec="fp32")
    _, f, _, _, f0, _ = _ef_check(eng, model, pos, sp, lat, 3e-3, 2e-2, 2e-3 * n, prec="mixed")
    assert np.abs(f - f0).mean() < 1e-3


def test_energies_only(eng):
    pos, sp, lat, model = W.carbon216()
    eng.set_model(model)
    e, _, _ = eng.energy_forces(pos, sp, lat, do_forces=False)
    e0, _, _ = oracle.energy_forces(model, pos, sp, lat, do_forces=False)
    assert np.abs(e - e0).max() < 1e-8


def test_single_atom(eng):
    pos = np.array([[1.0, 2.0, 3.0]])
    sp = np.array([1], np.int32)
    lat = S.cubic_lattice(20.0)
    _, _, _, model = W.carbon216()
    eng.set_model(model)
    e, f, v = eng.energy_forces(pos, sp, lat)
    e0, f0, v0 = oracle.energy_forces(model, pos, sp, lat)
    assert np.abs(e - e0).max() < 1e-10
    assert np.abs(f).max() == 0.0 and np.abs(v).max() == 0.0


def test_bad_arguments_raise(eng):
    _, _, _, model = W.carbon216()
    eng.set_model(model)
    pos = np.zeros((2, 3))
    with pytest.raises(tgx.TgxError):
        eng.energy_forces(pos, np.array([1, 5], np.int32), S.cubic_lattice(10.0))
    with pytest.raises(tgx.TgxError) as ei:
        eng.energy_forces(pos + [[0, 0, 0], [np.nan, 0, 0]], np.array([1, 1], np.int32),
                          S.cubic_lattice(10.0))
    assert ei.value.code == tgx.TGX_ERR_ARG
    flat = np.array([10.0, 0, 0, 1.0, 10.0, 0, 11.0, 1.0, 0.0])  # no volume
    with pytest.raises(tgx.TgxError) as ei:
        eng.energy_forces(pos + [[0, 0, 0], [1.5, 0, 0]], np.array([1, 1], np.int32), flat)
    assert ei.value.code == tgx.TGX_ERR_ARG


def test_list_buffer_does_not_change_results(eng):
    """Lists built with a buffer hold extra pairs beyond the cutoff; they add
    exact zeros, but shift which lane sums which pair, so results agree to
    rounding of the lane reductions."""
    pos, sp, lat, model = W.carbon216()
    eng.set_model(model)
    e0, f0, v0 = eng.energy_forces(pos, sp, lat)
    eng.set_list_buffer(0.7)
    try:
        e1, f1, v1 = eng.energy_forces(pos, sp, lat)
    finally:
        eng.set_list_buffer(0.0)
    assert np.abs(e0 - e1).max() <= 1e-13
    assert np.abs(f0 - f1).max() <= 1e-12
    assert np.abs(v0 - v1).max() <= 1e-12 * np.abs(v0).max()



@pytest.mark.parametrize("rows", [20000, 28000])
def test_chunked_kernel_step_tile_heights(eng, rows):
    """The chunked single-kernel step (600 sparse points) takes 96-row
    workgroups where 128-row ones would leave the last round of workgroups
    mostly empty: on 256 CUs the 216- and 20,000-site calls use 96 rows,
    the 28,000-site call 128.  A site's E and dE/dq depend on its own row
    only, so tiled copies of the 216 descriptor rows must reproduce the
    216-row call bit for bit."""
    pos, sp, lat, model = W.carbon216()
    m2 = S.model_from_soaps(model.hypers, model.Qs, 600, seed=5, zeta=4.0)
    eng.set_model(m2)
    q = oracle.soap(model.hypers, pos, sp, lat)["soap"]
    big = np.ascontiguousarray(np.tile(q, (rows // len(q) + 1, 1))[:rows])
    print(f"rows={rows}", flush=True)
    eng.set_precision("mixed")
    try:
        e1, g1 = eng.kernel_step(q)
        e2, g2 = eng.kernel_step(big)
    finally:
        eng.set_precision("fp64")
    idx = np.arange(rows) % len(q)
    assert np.array_equal(e2, e1[idx])
    assert np.array_equal(g2, g1[idx])

@pytest.mark.parametrize("rcut2b", [3.2, 4.6])
def test_two_body_fused_and_separate(eng, rcut2b):
    """The SOAP adjoint adds the 2-body terms itself when their cutoff lies
    within the SOAP cutoff (3.2 here); a longer 2-body cutoff (4.6) takes the
    separate pair kernel.  Both against the oracle in FP64."""
    pos, sp, lat, model = W.carbon216()
    m = tgx.GapModel(model.hypers, model.Qs, model.alphas, model.delta, model.zeta,
                     two_body=[S.carbon_2b(1000, rcut=rcut2b)])
    _ef_check(eng, m, pos, sp, lat, 1e-8, 1e-6, 1e-6)


def test_timing_events_are_opt_in(eng):
    """Stage and kernel timing events only with tgx_set_timing on (0 figures
    otherwise); the results do not depend on them."""
    pos, sp, lat, model = W.carbon216()
    eng.set_model(model)
    e0, f0, v0 = eng.energy_forces(pos, sp, lat)
    assert all(t == 0.0 for t in eng.kernel_times().values())
    assert all(t == 0.0 for t in eng.stage_times().values())
    eng.set_timing(True)
    try:
        e1, f1, v1 = eng.energy_forces(pos, sp, lat)
        kt, st = eng.kernel_times(), eng.stage_times()
    finally:
        eng.set_timing(False)
    print(kt, st, flush=True)
    assert kt["soap_fwd"] > 0 and kt["soap_bwd"] > 0 and kt["force_gather"] > 0
    assert st["total"] >= st["soap_fwd"] > 0
    assert np.array_equal(e0, e1) and np.array_equal(f0, f1) and np.array_equal(v0, v1)
    eng.energy_forces(pos, sp, lat)
    assert all(t == 0.0 for t in eng.kernel_times().values())
