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
    """The reference's formats: .gap blocks (one soap_turbo block per SOAP
    block of the model), alphas (one per line), desc (dim*n_sparse values,
    column-major), 2b alphas ("alpha cutoff")."""
    lines = []
    for k, blk in enumerate([model] + list(model.more_soap)):
        h = blk.hypers
        np.savetxt(os.path.join(d, f"alphas{k}.dat"), blk.alphas, fmt="%.17e")
        np.savetxt(os.path.join(d, f"desc{k}.dat"), blk.Qs.reshape(-1), fmt="%.17e")
        extra = []
        if h.compress_P is not None:
            pi, pj, pe = h.compress_P
            with open(os.path.join(d, f"compress{k}.dat"), "w") as f:
                f.write(" ".join(str(x) for x in h.alpha_max) + f" {h.l_max} {h.compress_dim}\n")
                if np.array_equal(pi, np.arange(1, len(pi) + 1)) and np.all(pe == 1.0):
                    f.writelines(f"{j}\n" for j in pj)  # the older format: kept indices
                else:
                    f.write("P_transformation\n")
                    f.writelines(f"{i} {j} {e:.17e}\n" for i, j, e in zip(pi, pj, pe))
                    f.write("end_transformation\n")
            extra += ["compress_soap = .true.", f"file_compress_soap = compress{k}.dat"]
        if h.basis != "poly3":
            extra.append(f"basis = {h.basis}")
        lines += ["gap_beg soap_turbo", f"n_species = {h.n_species}",
                  "species = " + " ".join(species),
                  "rcut = " + " ".join(str(x) for x in h.rcut_hard),
                  "buffer = " + " ".join(str(x) for x in h.buffer),
                  "atom_sigma_r = " + " ".join(str(x) for x in h.atom_sigma_r),
                  "atom_sigma_t = " + " ".join(str(x) for x in h.atom_sigma_t),
                  "n_max = " + " ".join(str(x) for x in h.alpha_max), f"l_max = {h.l_max}",
                  f"central_species = {h.central_species}",
                  f"zeta = {blk.zeta}", f"delta = {blk.delta}",
                  f"alphas_sparse = alphas{k}.dat", f"desc_sparse = desc{k}.dat"] + extra + \
            ["gap_end"]
    for k, b in enumerate(model.two_body):
        np.savetxt(os.path.join(d, f"a2b_{k}.dat"), np.stack([b.alphas, b.cutoff], 1), fmt="%.17e")
        np.savetxt(os.path.join(d, f"d2b_{k}.dat"), b.Qs, fmt="%.17e")
        lines += ["gap_beg distance_2b", f"species1 = {species[b.sp1 - 1]}",
                  f"species2 = {species[b.sp2 - 1]}", f"rcut = {b.rcut}", f"delta = {b.delta}",
                  f"sigma = {b.sigma}", f"alphas_sparse = a2b_{k}.dat",
                  f"desc_sparse = d2b_{k}.dat", "gap_end"]
    for k, b in enumerate(getattr(model, "angle_3b", [])):
        np.savetxt(os.path.join(d, f"a3b_{k}.dat"), np.stack([b.alphas, b.cutoff], 1),
                   fmt="%.17e")
        np.savetxt(os.path.join(d, f"d3b_{k}.dat"), np.asarray(b.Qs).reshape(-1), fmt="%.17e")
        lines += ["gap_beg angle_3b", f"species_center = {species[b.sp0 - 1]}",
                  f"species1 = {species[b.sp1 - 1]}", f"species2 = {species[b.sp2 - 1]}",
                  f"rcut = {b.rcut}", f"delta = {b.delta}",
                  "sigma = " + " ".join(repr(float(x)) for x in b.sigma),
                  f"kernel_type = {b.kernel}", f"alphas_sparse = a3b_{k}.dat",
                  f"desc_sparse = d3b_{k}.dat", "gap_end"]
    for k, b in enumerate(getattr(model, "core_pot", [])):
        with open(os.path.join(d, f"core_{k}.dat"), "w") as f:
            f.write(f"{len(b.x)} {float(b.yp1):.17e} {float(b.ypn):.17e}\n")
            for x, v in zip(b.x, b.V):
                f.write(f"{x:.17e} {v:.17e}\n")
        lines += ["gap_beg core_pot", f"species1 = {species[b.sp1 - 1]}",
                  f"species2 = {species[b.sp2 - 1]}", f"core_pot_file = core_{k}.dat", "gap_end"]
    open(os.path.join(d, "model.gap"), "w").write("\n".join(lines) + "\n")


def write_xyz(path, pos, sp, lat, symbols=("C",), mode="w"):
    with open(path, mode) as f:
        f.write(f"{len(pos)}\n")
        f.write('Lattice="' + " ".join(f"{x:.17e}" for x in lat) +
                '" Properties=species:S:1:pos:R:3\n')
        for p, s in zip(pos, sp):
            f.write(f"{symbols[s - 1]} {p[0]:.17e} {p[1]:.17e} {p[2]:.17e}\n")


def read_extxyz(path):
    """Frames of an extended-XYZ file: (comment key/values, columns)."""
    frames = []
    lines = open(path).read().splitlines()
    k = 0
    while k < len(lines):
        if not lines[k].strip():
            k += 1
            continue
        n = int(lines[k])
        head = lines[k + 1]
        kv = {}
        import shlex
        for tok in shlex.split(head):
            if "=" in tok:
                a, b = tok.split("=", 1)
                kv[a] = b
        rows = [lines[k + 2 + i].split() for i in range(n)]
        frames.append((kv, rows))
        k += 2 + n
    return frames


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


@pytest.mark.gpu
def test_fortran_predict_multiframe_two_blocks(tmp_path):
    """A model of two soap_turbo blocks (C/H/O, central species C and H) and
    three frames of different sizes in one xyz file: stdout per frame against
    the oracle at 1e-8 eV/atom, 1e-6 eV/A; the extended-XYZ output carries
    the same values at its printed precision (F16.8 columns, F16.6 energy)
    with the reference's layout."""
    import dataclasses

    import structures as S

    exe = os.path.join(FDIR, "tgx_predict")
    if not os.path.exists(exe):
        pytest.fail("tgx_predict is not built (run __graft_entry__.build())")
    pos, sp, lat = S.cho_cell(150, 1.5, 1.0, 71)
    h1 = dataclasses.replace(S.cho_hypers(), central_species=1)
    h2 = dataclasses.replace(S.cho_hypers(), central_species=2, l_max=6, alpha_max=(6, 6, 6),
                             rcut_hard=(3.5, 3.5, 3.5))
    m1 = S.model_from_soaps(h1, oracle.soap(h1, pos, sp, lat)["soap"], 60, seed=1)
    m2 = S.model_from_soaps(h2, oracle.soap(h2, pos, sp, lat)["soap"], 40, seed=2, zeta=3.0)
    model = dataclasses.replace(m1, more_soap=[m2])
    names = ("C", "H", "O")
    write_model_files(str(tmp_path), model, species=names)
    frames = [(pos, sp, lat), S.cho_cell(90, 1.5, 1.0, 72), S.cho_cell(120, 1.7, 1.0, 73)]
    for k, (p, q, l) in enumerate(frames):
        write_xyz(str(tmp_path / "s.xyz"), p, q, l, symbols=names, mode="w" if k == 0 else "a")
    out = subprocess.run([exe, "model.gap", "s.xyz", "fp64", "out.xyz"], capture_output=True,
                         text=True, cwd=tmp_path, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = out.stdout.splitlines()
    ext = read_extxyz(str(tmp_path / "out.xyz"))
    assert len(ext) == len(frames)
    k = 0
    for fi, (p, q, l) in enumerate(frames):
        n = len(p)
        etot = float(lines[k].split()[1])
        vir = np.array(lines[k + 1].split()[1:], float).reshape(3, 3).T
        rows = np.array([r.split()[1:] for r in lines[k + 2:k + 2 + n]], float)
        k += 2 + n
        e0, f0, v0 = oracle.energy_forces(model, p, q, l)
        assert abs(etot - e0.sum()) < 1e-8 * n
        assert np.abs(rows[:, 3] - e0).max() < 1e-8
        assert np.abs(rows[:, 4:7] - f0).max() < 1e-6
        assert np.abs(vir - v0).max() < 1e-9 * n
        kv, cols = ext[fi]
        assert kv["Properties"] == "species:S:1:pos:R:3:forces:R:3:local_energy:R:1"
        assert int(kv["i_config"]) == fi + 1
        assert np.allclose(np.array(kv["Lattice"].split(), float), l, atol=1e-9)
        assert abs(float(kv["energy"]) - e0.sum()) < 1e-6
        assert np.abs(np.array(kv["virial"].split(), float).reshape(3, 3).T - v0).max() < 1e-7
        assert [c[0] for c in cols] == [names[s - 1] for s in q]
        arr = np.array([c[1:] for c in cols], float)
        assert np.abs(arr[:, 0:3] - p).max() < 1e-7
        assert np.abs(arr[:, 3:6] - f0).max() < 1e-7
        assert np.abs(arr[:, 6] - e0).max() < 1e-7


@pytest.mark.gpu
def test_fortran_predict_angle_3b_core_pot(tmp_path):
    """angle_3b ("exp" and "pol") and core_pot blocks read from .gap files in
    the reference's formats, evaluated with SOAP and 2b, against the
    oracle."""
    import test_gpu_terms as T
    exe = os.path.join(FDIR, "tgx_predict")
    if not os.path.exists(exe):
        pytest.fail("tgx_predict is not built (run __graft_entry__.build())")
    pos, sp, lat, base = W.carbon216()
    model = T._with_terms(base, [T._a3("exp"), T._a3("pol", seed=5, rcut=2.3)], [T._core()])
    write_model_files(str(tmp_path), model)
    write_xyz(str(tmp_path / "s.xyz"), pos, sp, lat)
    out = subprocess.run([exe, "model.gap", "s.xyz", "fp64"], capture_output=True, text=True,
                         cwd=tmp_path, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = out.stdout.splitlines()
    etot = float(lines[0].split()[1])
    vir = np.array(lines[1].split()[1:], float).reshape(3, 3).T
    rows = np.array([l.split()[1:] for l in lines[2:2 + len(pos)]], float)
    e0, f0, v0 = oracle.energy_forces(model, pos, sp, lat)[:3]
    assert abs(etot - e0.sum()) < 1e-8 * len(pos)
    assert np.abs(rows[:, 3] - e0).max() < 1e-8
    assert np.abs(rows[:, 4:7] - f0).max() < 1e-6
    assert np.abs(vir - v0).max() < 1e-6


@pytest.mark.gpu
def test_fortran_predict_compressed_blocks(tmp_path):
    """Two compressed soap_turbo blocks read from .gap files: a general P in
    the P_transformation format (block 1) and a selection of components in
    the older one-index-per-line format (block 2, central species H, basis
    poly3gauss); the predict driver against the oracle."""
    import dataclasses

    import structures as S
    import test_gpu_compress as TC

    exe = os.path.join(FDIR, "tgx_predict")
    if not os.path.exists(exe):
        pytest.fail("tgx_predict is not built (run __graft_entry__.build())")
    pos, sp, lat = S.cho_cell(120, 1.5, 1.0, 81)
    h = S.cho_hypers()
    h1 = dataclasses.replace(h, compress_P=TC._random_P(h, 200, 3), compress_dim=200)
    sel = np.sort(np.random.default_rng(5).choice(h.full_dim, 300, replace=False)) + 1
    h2 = dataclasses.replace(h, central_species=2, compress_dim=300, basis="poly3gauss",
                             compress_P=(np.arange(1, 301), sel, np.ones(300)))
    m1 = S.model_from_soaps(h1, oracle.soap(h1, pos, sp, lat)["soap"], 50, seed=1)
    m2 = S.model_from_soaps(h2, oracle.soap(h2, pos, sp, lat)["soap"], 30, seed=2)
    model = dataclasses.replace(m1, more_soap=[m2])
    names = ("C", "H", "O")
    write_model_files(str(tmp_path), model, species=names)
    write_xyz(str(tmp_path / "s.xyz"), pos, sp, lat, symbols=names)
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
