"""Non-orthorhombic cells through the C-ABI.

Neighbour lists bit for bit against the oracle's restatement of the
reference's brute-force path (neighbors_interface.f90:390-421) with
get_distance's general minimum image (:582-629): 24 cells of different skew,
orientation, size and cutoff, many of them thinner than 2 rcut (supercell
images, read_xyz.f90:323-373).  Bit-exact: neighbour counts, supercell ids
and order, rjs, xyz, supercell repetitions; theta and phi within 4e-16 pi
(device acos/atan2).

Energies, forces, virials and descriptors of such cells (image atoms on the
device) against the oracle, which runs them on the reference's list: FP64
within 1e-8 eV/atom, 1e-6 eV/A, 1e-6 eV; descriptors within 1e-10."""
import numpy as np
import pytest

import oracle
import structures as S
import tgx
import workloads as W

pytestmark = pytest.mark.gpu

_rng = np.random.default_rng(2024)
CASES = []
for k in range(24):
    n = int(_rng.choice([6, 10, 17, 30, 48, 80]))
    rho = float(_rng.uniform(1.4, 3.0))
    skew = tuple(float(x) for x in _rng.uniform(-0.5, 0.5, 3))
    rot = bool(k % 2)
    rc = float(_rng.uniform(2.2, 4.5))
    CASES.append((k, n, rho, skew, rot, rc))


@pytest.fixture(scope="module")
def eng():
    e = tgx.Engine(0, "fp64")
    yield e
    e.close()


@pytest.mark.parametrize("case", CASES, ids=[f"c{c[0]}" for c in CASES])
def test_general_cell_lists_bit_exact(eng, case):
    k, n, rho, skew, rot, rc = case
    pos, sp, lat = S.triclinic_cell(n, rho, 1.0, 500 + k, skew=skew, rotate=rot)
    print(f"case {k}: n={n} rho={rho:.2f} skew={skew} rot={rot} rcut={rc:.2f}", flush=True)
    g = eng.neighbors(pos, lat, rc)
    o = oracle.neighbors(pos, lat, rc)
    print("indices", o["indices"], "pairs", len(o["list"]), flush=True)
    assert np.array_equal(g["indices"], o["indices"])
    assert np.array_equal(g["n_neigh"], o["n_neigh"])
    assert np.array_equal(g["list"], o["list"])
    assert np.array_equal(g["rjs"], o["rjs"])
    assert np.array_equal(g["xyz"], o["xyz"])
    assert np.allclose(g["thetas"], o["thetas"], rtol=0, atol=4e-16 * np.pi)
    assert np.allclose(g["phis"], o["phis"], rtol=0, atol=4e-16 * np.pi)


def test_general_cell_supercells_present():
    """The case list covers cells thinner than 2 rcut."""
    small = 0
    for k, n, rho, skew, rot, rc in CASES:
        lat = S.triclinic_cell(n, rho, 1.0, 500 + k, skew=skew, rotate=rot)[2]
        small += int(np.prod(oracle.neighbors(np.zeros((1, 3)), lat, rc)["indices"]) > 1)
    assert small >= 8


def _ef(eng, model, pos, sp, lat, prec="fp64"):
    eng.set_model(model)
    eng.set_precision(prec)
    try:
        return eng.energy_forces(pos, sp, lat)
    finally:
        eng.set_precision("fp64")


def _ef_vs_oracle(eng, model, pos, sp, lat, etol=1e-8, ftol=1e-6, vtol=1e-6, prec="fp64"):
    e, f, v = _ef(eng, model, pos, sp, lat, prec)
    e0, f0, v0 = oracle.energy_forces(model, pos, sp, lat)[:3]
    de, df, dv = np.abs(e - e0).max(), np.abs(f - f0).max(), np.abs(v - v0).max()
    print(f"{prec} n={len(pos)} dE={de:.3e} dF={df:.3e} dV={dv:.3e}", flush=True)
    assert de <= etol and df <= ftol and dv <= vtol, (de, df, dv)
    return e, f, v


def _sheared(lat, rot_seed=None):
    """Another basis of the same lattice (a, a + b, c - 2b), optionally
    rotated together with the positions."""
    L = np.asarray(lat, float).reshape(3, 3)
    M = np.array([L[0], L[1] + L[0], L[2] - 2 * L[1]])
    A = np.eye(3)
    if rot_seed is not None:
        A = np.linalg.qr(np.random.default_rng(rot_seed).normal(size=(3, 3)))[0]
    return M @ A.T, A


@pytest.mark.parametrize("rot", [None, 3])
def test_general_cell_ef_same_crystal(eng, rot):
    """carbon216 in a sheared (and rotated) basis: the oracle on the
    reference's general-cell list, and the orthorhombic GPU result rotated."""
    pos, sp, lat, model = W.carbon216()
    M, A = _sheared(lat, rot)
    e, f, v = _ef_vs_oracle(eng, model, pos @ A.T, sp, M.reshape(9))
    e1, f1, v1 = _ef(eng, model, pos, sp, lat)
    assert np.abs(e - e1).max() < 1e-9
    assert np.abs(f - f1 @ A.T).max() < 1e-8
    assert np.abs(v - A @ v1 @ A.T).max() < 1e-8


TRI = [(30, 2.0, (0.3, -0.2, 0.25), False, 1), (17, 2.6, (-0.45, 0.35, 0.1), True, 2),
       (80, 1.8, (0.5, 0.5, -0.4), True, 3), (6, 2.2, (0.2, 0.1, -0.3), False, 4),
       (48, 3.0, (-0.3, -0.3, 0.45), True, 5), (120, 2.0, (0.1, -0.45, 0.3), False, 6)]


@pytest.mark.parametrize("case", TRI, ids=[f"n{c[0]}" for c in TRI])
def test_general_cell_ef_fp64(eng, case):
    """Random triclinic carbon cells (several thinner than 2 rcut), atoms
    also moved out of the cell by lattice translations and small offsets."""
    n, rho, skew, rot, seed = case
    pos, sp, lat = S.triclinic_cell(n, rho, 1.1, 900 + seed, skew=skew, rotate=rot)
    M = lat.reshape(3, 3)
    rng = np.random.default_rng(seed)
    pos = pos + rng.integers(-2, 3, (n, 3)) @ M + rng.normal(0, 0.05, (n, 3))
    print(f"n={n} rho={rho} skew={skew} rot={rot}", flush=True)
    _, _, _, model = W.carbon216()
    _ef_vs_oracle(eng, model, pos, sp, lat)


def test_general_cell_ef_cho_central_species(eng):
    """C/H/O model, one block restricted to central species 1, and the
    2-body-free multi-species path, on a sheared basis of the CHO cell."""
    pos, sp, lat, model = W.cho()
    M, A = _sheared(lat, 11)
    _ef_vs_oracle(eng, model, pos @ A.T, sp, M.reshape(9))
    h = S.cho_hypers()
    h.central_species = 1
    q = oracle.soap(h, pos, sp, lat)["soap"]
    m = S.model_from_soaps(h, q, 100, seed=9)
    _ef_vs_oracle(eng, m, pos @ A.T, sp, M.reshape(9))


def test_general_cell_ef_mixed(eng):
    """Mixed precision on a triclinic cell at the orthorhombic bounds."""
    pos, sp, lat = S.triclinic_cell(120, 2.0, 1.1, 905, skew=(0.1, -0.45, 0.3), rotate=True)
    _, _, _, model = W.carbon216()
    n = len(pos)
    e, f, v = _ef(eng, model, pos, sp, lat, "mixed")
    e0, f0, v0 = oracle.energy_forces(model, pos, sp, lat)[:3]
    assert np.abs(e - e0).max() < 3e-3 and np.abs(f - f0).max() < 2e-2
    assert np.abs(f - f0).mean() < 1e-3 and np.abs(v - v0).max() < 2e-3 * n


def test_general_cell_noalloc_and_descriptors(eng):
    pos, sp, lat = S.triclinic_cell(48, 3.0, 1.1, 905, skew=(-0.3, -0.3, 0.45), rotate=True)
    _, _, _, model = W.carbon216()
    e, f, v = _ef(eng, model, pos, sp, lat)
    e2, f2, v2 = eng.energy_forces_noalloc(pos, sp, lat)
    assert np.array_equal(e, e2) and np.array_equal(f, f2) and np.array_equal(v, v2)
    q, site = eng.soap(pos, sp, lat)
    q0 = oracle.soap(model.hypers, pos, sp, lat)["soap"]
    assert np.abs(q - q0).max() < 1e-10


def test_general_cell_md_mc_refused(eng):
    pos, sp, lat = S.triclinic_cell(20, 2.0, 1.1, 9)
    _, _, _, model = W.carbon216()
    eng.set_model(model)
    with pytest.raises(tgx.TgxError) as e:
        eng.md_run(pos, np.zeros_like(pos), sp, lat, np.full(len(pos), 12.0), 2, 1.0, 0.5)
    assert e.value.code == tgx.TGX_ERR_UNSUPPORTED
    with pytest.raises(tgx.TgxError) as e:
        eng.mc_run(pos, sp, lat, 10, 1000.0, 0.2, 0.5, 1)
    assert e.value.code == tgx.TGX_ERR_UNSUPPORTED
    # the context still works afterwards
    e1 = eng.energy_forces(pos, sp, lat)[0]
    assert np.abs(e1 - oracle.energy_forces(model, pos, sp, lat)[0]).max() < 1e-8
