"""Context state rules through the C-ABI: calls that would overwrite the
buffers an active MD run's captured step reads are refused, non-default nf
is refused, and a context destroyed in the middle of an MD run releases it."""
import dataclasses

import numpy as np
import pytest

import oracle
import structures as S
import tgx
import workloads as W

pytestmark = pytest.mark.gpu


def test_calls_during_md_are_refused_and_md_continues():
    pos, sp, lat, model = W.carbon216()
    masses = np.full(len(pos), 12.011)
    vel = S.maxwell_boltzmann(len(pos), masses, 1000.0, seed=11)
    ref = tgx.Engine(0, "fp64")
    ref.set_model(model)
    want = ref.md_run(pos, vel, sp, lat, masses, 6, 1.0, skin=0.5)
    ref.close()

    eng = tgx.Engine(0, "fp64")
    eng.set_model(model)
    eng.md_begin(pos, vel, sp, lat, masses, 1.0, 0.5, history=16)
    eng.md_steps(3)
    refused = {
        "energy_forces": lambda: eng.energy_forces(pos, sp, lat),
        "energy_forces_batched": lambda: eng.energy_forces_batched([(pos, sp, lat)]),
        "soap": lambda: eng.soap(pos, sp, lat),
        "neighbors": lambda: eng.neighbors(pos, lat, 3.7),
        "kernel_step": lambda: eng.kernel_step(np.zeros((2, model.hypers.dim))),
        "set_model": lambda: eng.set_model(model),
        "set_precision": lambda: eng.set_precision("mixed"),
        "set_list_buffer": lambda: eng.set_list_buffer(0.1),
    }
    for name, call in refused.items():
        print("call", name, flush=True)
        with pytest.raises(tgx.TgxError) as e:
            call()
        assert e.value.code == tgx.TGX_ERR_STATE, name
    eng.md_steps(3)
    got = eng.md_read()
    eng.md_end()
    assert got["steps"] == 6
    assert np.array_equal(got["pos"], want["pos"])
    assert np.array_equal(got["forces"], want["forces"])
    # after md_end the context takes ordinary calls again
    e, f, v = eng.energy_forces(got["pos"], sp, lat)
    e0, f0, v0 = oracle.energy_forces(model, got["pos"], sp, lat)
    assert np.abs(e - e0).max() < 1e-8 and np.abs(f - f0).max() < 1e-6
    eng.close()


def test_destroy_during_md_run():
    pos, sp, lat, model = W.carbon216()
    masses = np.full(len(pos), 12.011)
    vel = S.maxwell_boltzmann(len(pos), masses, 1000.0, seed=12)
    for _ in range(3):
        eng = tgx.Engine(0, "mixed")
        eng.set_model(model)
        eng.md_begin(pos, vel, sp, lat, masses, 1.0, 0.5, history=8)
        eng.md_steps(2)
        eng.close()  # no tgx_md_end: destroy releases the graph and the buffers


def test_non_default_nf_is_refused():
    _, _, _, model = W.carbon216()
    h = dataclasses.replace(model.hypers, nf=(3.0,))
    eng = tgx.Engine(0, "fp64")
    with pytest.raises(tgx.TgxError) as e:
        eng.set_model(dataclasses.replace(model, hypers=h))
    assert e.value.code == tgx.TGX_ERR_UNSUPPORTED
    eng.close()
