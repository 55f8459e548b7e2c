"""CPU checks of the C-ABI boundary: the library loads, exports every entry
point include/tgx.h declares, the ctypes/Fortran mirror of tgx_soap_hypers
has the C layout, and the host-only orthonormalisation matches the oracle
bit for bit.  No device compute is called."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
import structures as S
import tgx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tgx.h")


def _declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(tgx_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    lib = tgx.load_library()
    names = _declared()
    assert len(names) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", tgx.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (tgx_\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        getattr(lib, n)


def test_soap_hypers_layout_matches_c(tmp_path):
    src = tmp_path / "layout.c"
    fields = [f[0] for f in tgx.CSoapHypers._fields_]
    body = "\n".join(f'  printf("{f} %zu\\n", offsetof(tgx_soap_hypers, {f}));' for f in fields)
    src.write_text(f"""#include <stdio.h>
#include <stddef.h>
#include "tgx.h"
int main(void) {{
  printf("size %zu\\n", sizeof(tgx_soap_hypers));
{body}
  return 0;
}}
""")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True,
                                                         text=True, check=True).stdout.split("\n")
               if line)
    assert int(got["size"]) == C.sizeof(tgx.CSoapHypers)
    for f in fields:
        assert int(got[f]) == getattr(tgx.CSoapHypers, f).offset, f


@pytest.mark.parametrize("nmax", [1, 2, 5, 8, 12, 16])
def test_orthonormalization_matches_oracle_bitwise(nmax):
    S1, W1 = tgx.poly3_orthonormalization(nmax)
    S2, W2 = oracle.poly3_overlap(nmax)
    assert np.array_equal(W1, W2)
    assert np.abs(S1 - S2).max() <= 2.3e-16


@pytest.mark.parametrize("nmax,sg,rh", [(1, 0.5, 3.7), (4, 0.5, 3.7), (8, 0.5, 3.7),
                                         (8, 0.3, 5.0), (12, 0.45, 4.0), (16, 0.5, 4.5)])
def test_poly3gauss_orthonormalization_matches_oracle(nmax, sg, rh):
    """Library (binary128, libquadmath) and oracle (its own binary128) W of
    the poly3gauss basis agree to FP64 rounding."""
    S1, W1 = tgx.poly3gauss_orthonormalization(nmax, sg, rh)
    S2, W2 = oracle.poly3gauss_overlap(nmax, rh, sg)
    assert np.abs(S1 - S2).max() <= 4.5e-16
    # at n_max = 16 cond(S) exceeds 1e20, so the two binary128 computations
    # (their own square roots and Jacobi sweeps) part at about 1e-12
    tol = 4e-16 if nmax <= 14 else 1e-11
    assert np.abs(W1 - W2).max() <= tol * np.abs(W2).max()


def test_orthonormalization_rejects_bad_size():
    lib = tgx.load_library()
    W = np.zeros(4)
    assert lib.tgx_poly3_orthonormalization(0, None, W.ctypes.data_as(C.POINTER(C.c_double))) \
        == tgx.TGX_ERR_ARG
    assert lib.tgx_poly3_orthonormalization(8, None, None) == tgx.TGX_ERR_ARG


def test_null_context_is_an_error_not_a_crash():
    lib = tgx.load_library()
    assert lib.tgx_destroy(None) == tgx.TGX_OK
    assert lib.tgx_last_error(None) == b"null context"
    assert lib.tgx_set_precision(None, 0) == tgx.TGX_ERR_ARG
    assert lib.tgx_md_steps(None, 1) == tgx.TGX_ERR_ARG
    assert lib.tgx_sync(None) == tgx.TGX_ERR_ARG


def test_hypers_host_logic():
    h = S.carbon_hypers()
    assert h.dim == 324 and h.n_r == 8
    c = h.to_c()
    assert c.rcut_soft[0] == pytest.approx(3.2) and c.rcut_hard[0] == pytest.approx(3.7)
    h0 = tgx.SoapHypers.uniform(1, buffer=0.0)
    assert h0.to_c().rcut_soft[0] == h0.to_c().rcut_hard[0]  # buffer 0 -> no soft zone
    h3 = S.cho_hypers()
    assert h3.n_r == 24 and h3.dim == 2700  # read_gap.f90:556-558


def test_structure_generator_density_and_min_distance():
    pos, sp, lat = S.carbon_cell(300, 2.0, 1.2, 3)
    L = lat[0]
    rho = 300 * 12.011 / (L ** 3 * 1e-24 * 6.02214076e23)
    assert rho == pytest.approx(2.0, rel=1e-12)
    d = pos[:, None, :] - pos[None, :, :]
    d -= L * np.round(d / L)
    r = np.sqrt((d ** 2).sum(-1)) + np.eye(300) * 10
    assert r.min() >= 1.2
    v = S.maxwell_boltzmann(300, np.full(300, 12.011), 3000.0, 5)
    m = 12.011 * tgx.AMU
    assert np.abs(v.sum(0)).max() < 1e-12
    T = 2 * 0.5 * m * (v ** 2).sum() / (3 * 299 * 8.6173303e-5)
    assert T == pytest.approx(3000.0, rel=1e-12)


def test_create_without_device_fails_cleanly():
    lib = tgx.load_library()
    ctx = C.c_void_p()
    rc = lib.tgx_create(0, 0, C.byref(ctx))
    if rc == tgx.TGX_OK:  # a GPU is present: nothing to check here
        lib.tgx_destroy(ctx)
        pytest.skip("device present")
    assert rc == tgx.TGX_ERR_HIP and not ctx.value


@pytest.mark.parametrize("cls,cname", [(tgx.DaOptions, "tgx_da_options"),
                                       (tgx.McStats, "tgx_mc_stats")])
def test_mc_structs_layout_matches_c(tmp_path, cls, cname):
    fields = [f[0] for f in cls._fields_]
    body = "\n".join(f'  printf("{f} %zu\\n", offsetof({cname}, {f}));' for f in fields)
    src = tmp_path / "layout.c"
    src.write_text(f"""#include <stdio.h>
#include <stddef.h>
#include "tgx.h"
int main(void) {{
  printf("size %zu\\n", sizeof({cname}));
{body}
  return 0;
}}
""")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True,
                                                         text=True, check=True).stdout.split("\n")
               if line)
    assert int(got["size"]) == C.sizeof(cls)
    for f in fields:
        assert int(got[f]) == getattr(cls, f).offset, f
