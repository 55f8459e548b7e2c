"""tgx_energy_forces_noalloc through the C-ABI: the same results as
tgx_energy_forces (bit for bit: same kernels, same launch shapes) on two
calls in a row, the second with a smaller structure; no device memory is
allocated by the calls; a structure too large for the scratch array is
refused with TGX_ERR_CAPACITY and leaves the context usable."""
import ctypes as C

import numpy as np
import pytest

import oracle
import structures as S
import tgx
import workloads as W

pytestmark = pytest.mark.gpu


def _hip_runtime():
    """The HIP runtime libtgx.so is linked to (already mapped)."""
    tgx.load_library()
    with open("/proc/self/maps") as fh:
        paths = {ln.split()[-1] for ln in fh if "libamdhip64" in ln}
    assert paths, "libamdhip64 is not loaded"
    return C.CDLL(sorted(paths)[0])


def _free_device_bytes():
    hip = _hip_runtime()
    free, total = C.c_size_t(0), C.c_size_t(0)
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_noalloc_equals_energy_forces(prec):
    _, _, _, model = W.carbon216()
    big = S.carbon_cell(512, 2.0, 1.2, 31)
    small = S.carbon_cell(100, 2.2, 1.2, 32)
    ref = tgx.Engine(0, prec)
    ref.set_model(model)
    want = [ref.energy_forces(*big), ref.energy_forces(*small)]
    ref.close()
    eng = tgx.Engine(0, prec)
    eng.set_model(model)
    eng.energy_forces_noalloc(*small)  # first launches of every kernel
    free0 = _free_device_bytes()
    got = []
    for k, s in enumerate((big, small)):
        print(f"{prec} call {k}: {len(s[0])} atoms", flush=True)
        got.append(eng.energy_forces_noalloc(*s))
    free1 = _free_device_bytes()
    for (e, f, v), (e0, f0, v0) in zip(got, want):
        assert np.array_equal(e, e0)
        assert np.array_equal(f, f0)
        assert np.array_equal(v, v0)
    assert free1 >= free0, (free0, free1)  # nothing allocated by the calls
    if prec == "fp64":  # and against the oracle, to the FP64 bar
        e_o, f_o, v_o = oracle.energy_forces(model, *small)
        e, f, v = got[1]
        assert np.abs(e - e_o).max() < 1e-8 and np.abs(f - f_o).max() < 1e-6
        assert np.abs(v - v_o).max() < 1e-6
    eng.close()


def test_noalloc_too_large_is_refused_and_context_stays_usable():
    _, _, _, model = W.carbon216()
    eng = tgx.Engine(0, "fp64")
    eng.set_model(model)
    huge = S.carbon_cell(6000, 2.0, 1.2, 33)
    with pytest.raises(tgx.TgxError) as e:
        eng.energy_forces_noalloc(*huge)
    assert e.value.code == tgx.TGX_ERR_CAPACITY
    pos, sp, lat = S.carbon_cell(64, 2.0, 1.2, 34)
    e1, f1, v1 = eng.energy_forces_noalloc(pos, sp, lat)
    e2, f2, v2 = eng.energy_forces(pos, sp, lat)
    assert np.array_equal(e1, e2) and np.array_equal(f1, f2)
    eng.close()
