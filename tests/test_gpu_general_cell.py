"""Neighbour lists of non-orthorhombic cells through the C-ABI, bit for bit
against the oracle's restatement of the reference's brute-force path
(neighbors_interface.f90:390-421) with get_distance's general minimum image
(:582-629): 24 cells of different skew, orientation, size and cutoff, many
of them thinner than 2 rcut (supercell images, read_xyz.f90:323-373).

Bit-exact: neighbour counts, supercell ids and order, rjs, xyz, supercell
repetitions; theta and phi within 4e-16 pi (device acos/atan2)."""
import numpy as np
import pytest

import oracle
import structures as S
import tgx

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


def test_general_cell_energy_forces_refused(eng):
    import workloads as W

    _, _, _, model = W.carbon216()
    eng.set_model(model)
    pos, sp, lat = S.triclinic_cell(20, 2.0, 1.1, 9)
    with pytest.raises(tgx.TgxError) as e:
        eng.energy_forces(pos, sp, lat)
    assert e.value.code == tgx.TGX_ERR_UNSUPPORTED
