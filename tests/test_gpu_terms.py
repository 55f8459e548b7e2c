"""angle_3b and core_pot terms on the GPU (one wave per atom, FP64 in every
precision mode) against the oracle's restatement of gap.f90:552-1015:
FP64 within 1e-8 eV/atom, 1e-6 eV/A and 1e-6 eV, alone and together with
SOAP and distance_2b terms, on cubic, triclinic and batched structures,
and through velocity Verlet."""
import numpy as np
import pytest

import oracle
import structures as S
import tgx
import workloads as W

pytestmark = pytest.mark.gpu

AMU = 103.6426965268


@pytest.fixture(scope="module")
def eng():
    e = tgx.Engine(0, "fp64")
    yield e
    e.close()


def _core_table(n=120, lo=0.3, hi=2.1):
    x = lo + (hi - lo) * np.linspace(0, 1, n) ** 1.3
    V = 30.0 * np.exp(-x / 0.35) / x
    dV0 = -30.0 * np.exp(-lo / 0.35) * (1 / (0.35 * lo) + 1 / lo ** 2)
    return x, V, dV0


def _a3(kernel, sp=(1, 1, 1), seed=3, rcut=2.6):
    rng = np.random.default_rng(seed)
    S3 = 40
    Q = np.stack([rng.uniform(2.4, 5.0, S3), rng.uniform(0.0, 1.0, S3),
                  rng.uniform(1.2, 4.5, S3)], 1)
    sig = (0.5, 0.4, 0.5) if kernel == "exp" else (2.0, 1.2, 2.0)
    return tgx.Angle3b(sp0=sp[0], sp1=sp[1], sp2=sp[2], rcut=rcut, delta=0.8, sigma=sig, Qs=Q,
                       alphas=rng.normal(0, 0.05, S3), cutoff=rng.uniform(0.5, 1.0, S3),
                       kernel=kernel)


def _core(sp1=1, sp2=1, natural=False):
    x, V, dV0 = _core_table()
    return tgx.CorePot(sp1=sp1, sp2=sp2, x=x, V=V, yp1=1e31 if natural else dV0, ypn=1e31)


def _with_terms(model, a3=(), cp=()):
    return tgx.GapModel(hypers=model.hypers, Qs=model.Qs, alphas=model.alphas,
                        delta=model.delta, zeta=model.zeta, e0=model.e0,
                        two_body=list(model.two_body), more_soap=list(model.more_soap),
                        angle_3b=list(a3), core_pot=list(cp))


def _check(eng, model, pos, sp, lat, prec="fp64", soap=True, tol=(1e-8, 1e-6, 1e-6)):
    eng.set_model(model)
    eng.set_precision(prec)
    try:
        e, f, v = eng.energy_forces(pos, sp, lat)
    finally:
        eng.set_precision("fp64")
    e0, f0, v0 = oracle.energy_forces(model, pos, sp, lat, soap=soap)[:3]
    de, df, dv = np.abs(e - e0).max(), np.abs(f - f0).max(), np.abs(v - v0).max()
    print(f"{prec} n={len(pos)} dE={de:.3e} dF={df:.3e} dV={dv:.3e} |F|={np.abs(f0).max():.2f}",
          flush=True)
    assert de <= tol[0] and df <= tol[1] and dv <= tol[2], (de, df, dv)
    return e, f, v


@pytest.mark.parametrize("kernel", ["exp", "pol"])
def test_terms_alone_fp64(eng, kernel):
    """angle_3b + core_pot only (no SOAP block is evaluated: its sparse set
    is a single zero-weight point)."""
    pos, sp, lat = S.carbon_cell(48, 2.2, 1.05, 41)
    _, _, _, base = W.carbon216()
    m = _with_terms(base, [_a3(kernel)], [_core()])
    m.alphas = np.zeros_like(m.alphas)
    m.two_body = []
    _check(eng, m, pos, sp, lat)


@pytest.mark.parametrize("kernel", ["exp", "pol"])
def test_terms_with_soap_and_2b(eng, kernel):
    pos, sp, lat, base = W.carbon216()
    m = _with_terms(base, [_a3(kernel), _a3("exp", seed=4, rcut=2.2)], [_core(natural=True)])
    _check(eng, m, pos, sp, lat)
    n = len(pos)
    _check(eng, m, pos, sp, lat, "mixed", tol=(3e-3, 2e-2, 2e-3 * n))
    # energies only
    eng.set_model(m)
    e = eng.energy_forces(pos, sp, lat, do_forces=False)[0]
    e0 = oracle.energy_forces(m, pos, sp, lat, do_forces=False)[0]
    assert np.abs(e - e0).max() < 1e-8


def test_terms_species_filters_cho(eng):
    """C-centred C-H angle_3b and a C-H core_pot on the C/H/O cell with a
    central-species SOAP block."""
    pos, sp, lat, base = W.cho()
    m = _with_terms(base, [_a3("exp", sp=(1, 1, 2), rcut=2.8)], [_core(1, 2)])
    _check(eng, m, pos, sp, lat)


def test_terms_triclinic_and_small_cell(eng):
    _, _, _, base = W.carbon216()
    m = _with_terms(base, [_a3("pol")], [_core()])
    pos, sp, lat = S.triclinic_cell(40, 2.4, 1.05, 77, skew=(0.3, -0.4, 0.2), rotate=True)
    _check(eng, m, pos, sp, lat)
    pos, sp, lat = S.carbon_cell(12, 2.6, 1.05, 78)
    _check(eng, m, pos, sp, lat)


def test_terms_batched_equals_single(eng):
    _, _, _, base = W.carbon216()
    m = _with_terms(base, [_a3("exp")], [_core()])
    eng.set_model(m)
    structs = [S.carbon_cell(n, rho, 1.05, 300 + n) for n, rho in ((24, 2.0), (60, 2.6), (33, 3.0))]
    _, e, f, v = eng.energy_forces_batched(structs)
    o = 0
    for k, (p, s, l) in enumerate(structs):
        e0, f0, v0 = oracle.energy_forces(m, p, s, l)[:3]
        assert np.abs(e[o:o + len(p)] - e0).max() < 1e-8
        assert np.abs(f[o:o + len(p)] - f0).max() < 1e-6
        assert np.abs(v[k] - v0).max() < 1e-6
        o += len(p)


def test_terms_md_and_mc(eng):
    """Velocity Verlet with the terms (graph-captured step) against the
    oracle's; Monte Carlo refuses such models."""
    pos, sp, lat, base = W.carbon216()
    m = _with_terms(base, [_a3("exp")], [_core()])
    eng.set_model(m)
    masses = np.full(len(pos), 12.011)
    vel = S.maxwell_boltzmann(len(pos), masses, 2000.0, seed=6)
    out = eng.md_run(pos, vel, sp, lat, masses, 6, 1.0, skin=0.5)
    p, v = pos.copy(), vel.copy()
    mm = masses * AMU
    pprev, fprev = np.zeros_like(p), np.zeros_like(p)
    e, f, _ = oracle.energy_forces(m, p, sp, lat)[:3]
    dtp = 1.0
    for s in range(6):
        dtp = oracle.velocity_verlet(p, pprev, v, f, fprev, mm, 1.0, dtp, s == 0)
        e, f, _ = oracle.energy_forces(m, p, sp, lat)[:3]
    assert np.abs(out["pos"] - p).max() < 1e-9
    assert np.abs(out["forces"] - f).max() < 1e-6
    with pytest.raises(tgx.TgxError) as ei:
        eng.mc_run(pos, sp, lat, 10, 1000.0, 0.2, 0.5, 1)
    assert ei.value.code == tgx.TGX_ERR_UNSUPPORTED
