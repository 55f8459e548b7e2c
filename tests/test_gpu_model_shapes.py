"""The FP32 SOAP forward and adjoint against the FP64 oracle over the model
shapes that change the kernels' channel layout: one species (n_max 8 and
12, so radial blocks of 8 and 16), several species with even and odd radial
offsets, l_max below the compile-time bound, a central-species filter, and
batched structures against single calls.

Tolerances: FP32 mode energies 1e-5 eV/atom, forces 1e-4 eV/A, virial
1e-4 eV per atom."""
import dataclasses

import numpy as np
import pytest

import oracle
import structures as S
import tgx
import workloads as W

pytestmark = pytest.mark.gpu


def _model(case):
    if case == "carbon216":
        return W.carbon216()
    if case == "cho300":
        return W.cho()
    if case == "carbon_l6":
        pos, sp, lat, _ = W.carbon216()
        h = dataclasses.replace(S.carbon_hypers(), l_max=6)
    elif case == "carbon_n12":
        pos, sp, lat, _ = W.carbon216()
        h = dataclasses.replace(S.carbon_hypers(), alpha_max=(12,))
    elif case == "cho_odd":  # radial offsets 0, 5, 8
        pos, sp, lat = S.cho_cell(200, 1.5, 1.0, 77)
        h = dataclasses.replace(S.cho_hypers(), alpha_max=(5, 3, 4))
    elif case == "cho_central":
        pos, sp, lat = S.cho_cell(200, 1.5, 1.0, 78)
        h = dataclasses.replace(S.cho_hypers(), central_species=2)
    else:
        raise ValueError(case)
    q = oracle.soap(h, pos, sp, lat)["soap"]
    return pos, sp, lat, S.model_from_soaps(h, q, 150, seed=5)


CASES = ["carbon216", "cho300", "carbon_l6", "carbon_n12", "cho_odd", "cho_central"]


@pytest.mark.parametrize("case", CASES)
def test_fp32_model_shapes(case):
    pos, sp, lat, model = _model(case)
    print(f"case={case} n={len(pos)} nr={model.hypers.n_r} l_max={model.hypers.l_max}",
          flush=True)
    eng = tgx.Engine(0, "fp32")
    eng.set_model(model)
    e1, f1, v1 = eng.energy_forces(pos, sp, lat)
    eng.close()
    e0, f0, v0 = oracle.energy_forces(model, pos, sp, lat)
    n = len(pos)
    print(f"|dE| {np.abs(e1 - e0).max():.3g} |dF| {np.abs(f1 - f0).max():.3g} "
          f"|dV| {np.abs(v1 - v0).max():.3g}", flush=True)
    assert np.abs(e1 - e0).max() < 1e-5
    assert np.abs(f1 - f0).max() < 1e-4
    assert np.abs(v1 - v0).max() < 1e-4 * n


def test_mixed_batched_equals_single():
    """Batched structures (a cell per structure) take the same kernels:
    each structure's energies and forces equal its own single call bit for
    bit, the virial to the rounding of its per-structure sum."""
    _, _, _, model = W.carbon216()
    structs = [S.carbon_cell(n, rho, 1.2, 900 + n) for n, rho in ((40, 2.6), (150, 2.0),
                                                                   (333, 1.7))]
    eng = tgx.Engine(0, "mixed")
    eng.set_model(model)
    offs, e, f, v = eng.energy_forces_batched(structs)
    for k, (p, s, l) in enumerate(structs):
        e1, f1, v1 = eng.energy_forces(p, s, l)
        assert np.array_equal(e[offs[k]:offs[k + 1]], e1)
        assert np.array_equal(f[offs[k]:offs[k + 1]], f1)
        assert np.abs(v[k] - v1).max() <= 1e-12 * max(1.0, np.abs(v1).max())
    eng.close()
