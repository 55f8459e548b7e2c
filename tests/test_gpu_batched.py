"""Batched predict: many independent periodic structures in one launch
sequence (tgx_energy_forces_batched).  Each structure's energies, forces and
virial must equal the single-structure path (same arithmetic per site and
pair, so to rounding of the virial sums) and the FP64 oracle."""
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


def _mixed_sizes():
    # 8 atoms at 2 g/cm^3 is a 4.3 A box (27 supercell images), 200 atoms a
    # linked-cell box: both branches of the batched list build in one batch
    return S.batch_cells(7, seed=77, n_lo=8, n_hi=200) + [S.carbon_cell(64, 3.0, 1.2, 5)]


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_batched_equals_single(eng, prec):
    _, _, _, model = W.carbon216()
    eng.set_model(model)
    eng.set_precision(prec)
    try:
        structs = _mixed_sizes()
        print(f"prec={prec} sizes={[len(p) for p, _, _ in structs]}", flush=True)
        offs, e, f, v = eng.energy_forces_batched(structs)
        for s, (p, sp, lat) in enumerate(structs):
            a, b = offs[s], offs[s + 1]
            e1, f1, v1 = eng.energy_forces(p, sp, lat)
            assert np.array_equal(e[a:b], e1), s
            assert np.array_equal(f[a:b], f1), s
            assert np.abs(v[s] - v1).max() <= 1e-12 * max(1.0, np.abs(v1).max()), s
    finally:
        eng.set_precision("fp64")


def test_batched_fp64_vs_oracle(eng):
    _, _, _, model = W.carbon216()
    eng.set_model(model)
    structs = _mixed_sizes()
    offs, e, f, v = eng.energy_forces_batched(structs)
    for s, (p, sp, lat) in enumerate(structs):
        a, b = offs[s], offs[s + 1]
        eo, fo, vo = oracle.energy_forces(model, p, sp, lat)
        assert np.abs(e[a:b] - eo).max() <= 1e-8, s
        assert np.abs(f[a:b] - fo).max() <= 1e-6, s
        assert np.abs(v[s] - vo).max() <= 1e-6, s


def test_batched_multispecies_energies_only(eng):
    pos, sp, lat, model = W.cho()
    eng.set_model(model)
    structs = [S.cho_cell(40, 1.5, 1.0, 31), (pos, sp, lat), S.cho_cell(90, 1.5, 1.0, 32)]
    offs, e, _, _ = eng.energy_forces_batched(structs, do_forces=False)
    for s, (p, q, l) in enumerate(structs):
        eo, _, _ = oracle.energy_forces(model, p, q, l, do_forces=False)
        assert np.abs(e[offs[s]:offs[s + 1]] - eo).max() <= 1e-8, s


def test_batched_bad_offsets(eng):
    _, _, _, model = W.carbon216()
    eng.set_model(model)
    p, sp, lat = S.carbon_cell(20, 2.0, 1.2, 3)
    with pytest.raises(tgx.TgxError):
        eng.energy_forces_batched([(p, sp, lat), (p[:0], sp[:0], lat)])
