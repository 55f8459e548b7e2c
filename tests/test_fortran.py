"""The Fortran host side (iso_c_binding module and drivers): struct layout
on the CPU, and the predict driver on the GPU reading model files in the
reference's .gap / alphas / desc formats (gap/read_gap.f90:45-875)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
import tgx
import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "turbogap_2.0_amd", "fortran")
FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"


def test_fortran_struct_layout_matches_c(tmp_path):
    if not os.path.exists(FC):
        pytest.skip("no Fortran compiler")
    exe = tmp_path / "layout"
    libdir = os.path.dirname(tgx.LIB_PATH)
    subprocess.run([FC, "-J", str(tmp_path), os.path.join(FDIR, "tgx_mod.f90"),
                    os.path.join(ROOT, "tests", "fortran", "layout_check.f90"), "-L", libdir,
                    "-ltgx", f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True, cwd=tmp_path)
    out = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                 check=True).stdout.splitlines() if l.strip())
    import ctypes as C
    assert int(out["size"]) == C.sizeof(tgx.CSoapHypers)
    for f in ("n_species", "l_max", "alpha_max", "rcut_hard", "rcut_soft", "central_weight", "nf"):
        assert int(out[f]) == getattr(tgx.CSoapHypers, f).offset, f
    assert int(out["da.size"]) == C.sizeof(tgx.DaOptions)
    for f, _ in tgx.DaOptions._fields_:
        assert int(out["da." + f]) == getattr(tgx.DaOptions, f).offset, f
    assert int(out["st.size"]) == C.sizeof(tgx.McStats)
    for f, _ in tgx.McStats._fields_:
        if f != "pad_":
            assert int(out["st." + f]) == getattr(tgx.McStats, f).offset, f


def write_model_files(d, model, species=("C",)):
    """The reference's formats: .gap blocks, alphas (one per line), desc
    (dim*n_sparse values, column-major), 2b alphas ("alpha cutoff")."""
    h = model.hypers
    np.savetxt(os.path.join(d, "alphas.dat"), model.alphas, fmt="%.17e")
    np.savetxt(os.path.join(d, "desc.dat"), model.Qs.reshape(-1), fmt="%.17e")
    lines = ["gap_beg soap_turbo", f"n_species = {h.n_species}",
             "species = " + " ".join(species),
             "rcut = " + " ".join(str(x) for x in h.rcut_hard),
             "buffer = " + " ".join(str(x) for x in h.buffer),
             "atom_sigma_r = " + " ".join(str(x) for x in h.atom_sigma_r),
             "atom_sigma_t = " + " ".join(str(x) for x in h.atom_sigma_t),
             "n_max = " + " ".join(str(x) for x in h.alpha_max), f"l_max = {h.l_max}",
             f"zeta = {model.zeta}", f"delta = {model.delta}",
             "alphas_sparse = alphas.dat", "desc_sparse = desc.dat", "gap_end"]
    for k, b in enumerate(model.two_body):
        np.savetxt(os.path.join(d, f"a2b_{k}.dat"), np.stack([b.alphas, b.cutoff], 1), fmt="%.17e")
        np.savetxt(os.path.join(d, f"d2b_{k}.dat"), b.Qs, fmt="%.17e")
        lines += ["gap_beg distance_2b", f"species1 = {species[b.sp1 - 1]}",
                  f"species2 = {species[b.sp2 - 1]}", f"rcut = {b.rcut}", f"delta = {b.delta}",
                  f"sigma = {b.sigma}", f"alphas_sparse = a2b_{k}.dat",
                  f"desc_sparse = d2b_{k}.dat", "gap_end"]
    open(os.path.join(d, "model.gap"), "w").write("\n".join(lines) + "\n")


def write_xyz(path, pos, sp, lat, symbols=("C",)):
    with open(path, "w") as f:
        f.write(f"{len(pos)}\n")
        f.write('Lattice="' + " ".join(f"{x:.17e}" for x in lat) +
                '" Properties=species:S:1:pos:R:3\n')
        for p, s in zip(pos, sp):
            f.write(f"{symbols[s - 1]} {p[0]:.17e} {p[1]:.17e} {p[2]:.17e}\n")


@pytest.mark.gpu
def test_fortran_predict_driver(tmp_path):
    exe = os.path.join(FDIR, "tgx_predict")
    if not os.path.exists(exe):
        pytest.fail("tgx_predict is not built (run __graft_entry__.build())")
    pos, sp, lat, model = W.carbon216()
    write_model_files(str(tmp_path), model)
    write_xyz(str(tmp_path / "s.xyz"), pos, sp, lat)
    out = subprocess.run([exe, "model.gap", "s.xyz", "fp64"], capture_output=True, text=True,
                         cwd=tmp_path, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = out.stdout.splitlines()
    etot = float(lines[0].split()[1])
    vir = np.array(lines[1].split()[1:], float).reshape(3, 3).T
    rows = np.array([l.split()[1:] for l in lines[2:2 + len(pos)]], float)
    e0, f0, v0 = oracle.energy_forces(model, pos, sp, lat)
    assert abs(etot - e0.sum()) < 1e-8 * len(pos)
    assert np.abs(rows[:, 3] - e0).max() < 1e-8
    assert np.abs(rows[:, 4:7] - f0).max() < 1e-6
    assert np.abs(vir - v0).max() < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["plain", "da"])
def test_fortran_mc_driver(tmp_path, kind):
    """The Fortran Monte Carlo driver (reference model files, xyz input):
    plain FP64 Metropolis equals the oracle's chain move for move; delayed
    acceptance with a surrogate trained on the first 256 trials reports its
    stages consistently (every trial is either rejected by the surrogate or
    evaluated with the GAP) and the same final energy as a fresh evaluation."""
    exe = os.path.join(FDIR, "tgx_mc")
    if not os.path.exists(exe):
        pytest.fail("tgx_mc is not built (run __graft_entry__.build())")
    pos, sp, lat, model = W.carbon216()
    write_model_files(str(tmp_path), model)
    write_xyz(str(tmp_path / "s.xyz"), pos, sp, lat)
    n_trials, T, mm, skin, seed = 400, 1500.0, 0.6, 1.2, 11
    args = [exe, "model.gap", "s.xyz", str(n_trials), str(T), str(mm), str(skin), str(seed), kind]
    if kind == "da":
        args += ["256"]
    out = subprocess.run(args, capture_output=True, text=True, cwd=tmp_path, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    kv = {}
    fpos = np.zeros_like(pos)
    for l in out.stdout.splitlines():
        w = l.split()
        if w[0] == "pos":
            fpos[int(w[1]) - 1] = [float(x) for x in w[2:5]]
        else:
            kv[w[0]] = float(w[1])
    e_new, _, _ = oracle.energy_forces(model, fpos, sp, lat)
    assert abs(kv["e_final"] - e_new.sum()) < 1e-8 * len(pos)
    if kind == "plain":
        p_o, tr_o, acc_o = oracle.mc_run(model, pos, sp, lat, n_trials, T, mm, seed)
        assert int(kv["accepted"]) == acc_o
        assert np.array_equal(fpos, p_o)
    else:
        assert int(kv["stage1_rejected"]) + int(kv["gap_evaluations"]) == n_trials
        assert int(kv["fallback"]) in (0, 1)
        assert 0 < kv["accepted"] <= kv["gap_evaluations"]
