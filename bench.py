"""Benchmark: NVE molecular dynamics of 32,768 carbon atoms (E+F every step).

Workload (headline metric "atom-steps/s (E+F) at 32k C atoms, mixed vs FP64;
force MAE vs FP64 oracle"): 32,768 C at 2.0 g/cm^3 (min distance 1.2 A,
seed 1003), Maxwell-Boltzmann velocities at 3,000 K, velocity Verlet with
dt = 1 fs, neighbour list with 0.5 A skin rebuilt on displacement.  GAP
model: SOAP rcut 3.7 (buffer 0.5), n_max 8, l_max 8, sigma 0.5, zeta 4,
delta 1, 500 sparse vectors drawn from the normalised SOAPs of a seeded
216-atom carbon cell, alphas ~ N(0, 0.01), plus a C-C distance_2b term.

A step = one hipGraph replay: integrator, skin test, conditional list
rebuild, SOAP expansion, kernel GEMMs, SOAP adjoint and force gather, 2-body
term, FP64 reductions.  Inputs are resident in HBM before the timed region.
`value` is the mixed path (FP32 SOAP + BF16 MFMA kernel GEMM); the FP64 GPU
path is timed the same way and reported beside it.
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "turbogap_2.0_amd"))

import structures as S  # noqa: E402
import tgx  # noqa: E402

METRIC = "atom-steps/s (E+F) at 32k C atoms, mixed vs FP64; force MAE vs FP64 oracle"
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def recorded_traffic(kernel, precision, atoms):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    measurement (FETCH_SIZE + WRITE_SIZE, scaled by a copy-kernel
    calibration), or None when no measurement matches this run."""
    try:
        with open(TRAFFIC_FILE) as fh:
            rec = json.load(fh)
    except (OSError, ValueError):
        return None
    for r in rec.get("kernels", []):
        if r.get("kernel") == kernel and r.get("precision") == precision \
                and r.get("atoms") == atoms:
            return r.get("bytes_per_launch")
    return None


def copy_probe(mbytes):
    """A plain device copy of known size, for calibrating the PMC byte
    counters (FETCH_SIZE/WRITE_SIZE) against algorithmic bytes."""
    import torch

    n = mbytes * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n,
                      "launches": 3}))


def cpu_model():
    """The host CPU's model name and logical CPU count (as lscpu reports)."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    name = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{name}, {os.cpu_count()} logical CPUs"


def read_tune(path):
    """The autotuner's decisions of this run (shape key -> index of the
    chosen candidate launch shape), and the file removed."""
    try:
        with open(path) as fh:
            dec = dict(ln.rsplit(" ", 1) for ln in fh.read().splitlines() if " " in ln)
        os.remove(path)
    except OSError:
        return None
    return {k: int(v) for k, v in dec.items()}


def build_model(eng):
    """Sparse set from the library's own FP64 SOAPs of a 216-atom cell."""
    pos, sp, lat = S.carbon_cell(216, 2.0, 1.2, 1000)
    h = S.carbon_hypers()
    tmp = S.model_from_soaps(h, np.eye(h.dim)[:1], 1, seed=0)
    eng.set_model(tmp)
    q, _ = eng.soap(pos, sp, lat)
    return S.model_from_soaps(h, q, 500, seed=1000, two_body=[S.carbon_2b(1000)])


def zorder(pos, lat):
    """Atom order along a Z-order (Morton) curve of the wrapped positions,
    10 bits per axis of the orthorhombic box: the order the library's MD
    run (tgx_md_begin) keeps its atoms in, so that single-call kernel timings see the MD step's
    memory locality."""
    L = np.array([lat[0], lat[4], lat[8]])
    f = np.mod(pos / L, 1.0)
    g = np.minimum((f * 1024).astype(np.int64), 1023)
    key = np.zeros(len(pos), dtype=np.int64)
    for bit in range(10):
        for d in range(3):
            key |= ((g[:, d] >> bit) & 1) << (3 * bit + d)
    return np.argsort(key, kind="stable")


def soap_costs(h, n_sites, n_pairs):
    """Algorithmic floating-point operations of the SOAP expansion and its
    adjoint (multiply-add = 2), L2 = (l_max+1)^2 real harmonics, n_max of
    the neighbour's species.  Per pair within the cutoff: radial lookup
    (n_max x 16), harmonics and angular factors (L2 x 6), c accumulation
    (n_max x L2 x 2); adjoint: h and h' (2 x n_max x L2 x 2), harmonic
    gradients and contraction (L2 x 16).  Per site: power spectrum
    (sum over l of N_R(N_R+1)/2 x (2l+1) x 2) plus normalisation (3 x dim);
    adjoint: dE/dc = M_l c (N_R^2 x L2 x 2) plus projection (4 x dim)."""
    L2 = (h.l_max + 1) ** 2
    nr = h.n_r
    nm = max(h.alpha_max)
    npair_nn = nr * (nr + 1) // 2
    ps = npair_nn * sum(2 * (2 * l + 1) for l in range(h.l_max + 1))
    fwd = n_pairs * (nm * 16 + L2 * 6 + nm * L2 * 2) + n_sites * (ps + 3 * h.dim)
    bwd = n_pairs * (nm * 16 + 2 * nm * L2 * 2 + L2 * 16) + n_sites * (nr * nr * L2 * 2 + 4 * h.dim)
    return fwd, bwd


def bench_mc512(args):
    """Metropolis MC, 512 C at 2.0 g/cm^3, 1,500 K, single-atom moves with
    per-component bound 0.3 A (move_max 0.6): plain GAP Metropolis against
    delayed acceptance with the surrogate trained on the chain's own first
    20,000 trials.  Accepted moves/s of the timed trials of each mode from
    the same start state, in speculative batches of 8 trials; plain
    Metropolis also one trial per launch sequence (the same chain)."""
    n = 512
    pos, sp, lat = S.carbon_cell(n, 2.0, 1.2, 1004)
    eng = tgx.Engine(0, "fp64")
    model = build_model(eng)
    eng.set_model(model)
    eng.set_precision("mixed")
    T, mm, skin = 1500.0, 0.6, 1.2
    # surrogate trained online on plain-MC trials (GAP site energies)
    eng.surrogate_init(model.hypers.dim, seed=4)
    t0 = time.perf_counter()
    tr = eng.mc_run_da(pos, sp, lat, 20000, T, mm, skin, 1, n_train_trials=20000,
                       train_every=2048, epochs=3, batch=512, lr=2e-3, max_rmse=0.2)
    t_train = time.perf_counter() - t0
    start = tr["pos"]
    # the one-trial launch sequence first, then speculative batches (the
    # default): the same chain, decided move for move
    eng.set_mc_batch(1)
    eng.mc_run(start, sp, lat, 64, T, mm, skin, 2)  # warm-up
    t0 = time.perf_counter()
    one = eng.mc_run(start, sp, lat, args.trials, T, mm, skin, 3)
    t_one = time.perf_counter() - t0
    eng.set_mc_batch(0)
    eng.mc_run(start, sp, lat, 64, T, mm, skin, 2)  # warm-up
    t0 = time.perf_counter()
    pl = eng.mc_run(start, sp, lat, args.trials, T, mm, skin, 3)
    t_plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    da = eng.mc_run_da(start, sp, lat, args.trials, T, mm, skin, 3, n_train_trials=0,
                       max_rmse=0.2)
    t_da = time.perf_counter() - t0
    acc_pl, acc_da = pl["n_accepted"], da["n_accepted"]
    v_pl, v_da = acc_pl / t_plain, acc_da / t_da
    print(f"[bench] mc512 plain {v_pl:.4g} acc/s ({acc_pl}/{args.trials}), DA {v_da:.4g} acc/s "
          f"({acc_da}/{args.trials}, stage-1 rejections {da['n_stage1_rejected']})",
          file=sys.stderr, flush=True)
    line = {"metric": "accepted Monte Carlo moves/s, 512 C at 1500 K, delayed acceptance",
            "value": round(v_da, 1), "unit": "accepted moves/s", "n_gpus": 1,
            "steps": args.trials, "warmup": 64, "higher_is_better": True, "scaling": "strong",
            "vs_baseline": None, "dtype": "f32+bf16", "data": "synthetic",
            "config": {"workload": "Metropolis MC, 512 C at 2.0 g/cm3, 1500 K, single-atom "
                                   "moves, per-component bound 0.3 A, skin 1.2 A",
                       "atoms": n, "trials": args.trials},
            "plain_value": round(v_pl, 1), "speedup_over_plain": round(v_da / v_pl, 3),
            "plain_one_trial_value": round(one["n_accepted"] / t_one, 1),
            "batch_speedup_plain": round(v_pl / (one["n_accepted"] / t_one), 3),
            "acceptance_plain": acc_pl / args.trials, "acceptance_da": acc_da / args.trials,
            "stage1_rejected": da["n_stage1_rejected"],
            "gap_evaluations": da["n_gap_evaluations"],
            "surrogate_train_rmse_eV_per_site": tr["train_rmse"],
            "surrogate_dE_rmse_eV": da["da_rmse"], "fallback": da["fallback"],
            "training_phase_s": round(t_train, 3), "training_trials": 20000,
            "training_accepted": tr["n_accepted"]}
    print(json.dumps(line))
    eng.close()


def bench_batch256(args):
    """Batched predict: 256 periodic carbon cells of 32-1,024 atoms
    (log-uniform), 1.5-3.2 g/cm^3, carbon SOAP, 2,000 sparse from the batch's
    own SOAPs; all structures in one launch sequence (host buffers in and
    out, lists built per call)."""
    structs = S.batch_cells(256, 1002)
    n_atoms = sum(len(p) for p, _, _ in structs)
    eng = tgx.Engine(0, "fp64")
    tune_file = os.path.join(tempfile.gettempdir(), f"tgx_autotune_{os.getpid()}.txt")
    if not args.no_autotune:  # decided in the untimed first call of each mode
        eng.set_autotune(True, tune_file)
    h = S.carbon_hypers()
    tmp = S.model_from_soaps(h, np.eye(h.dim)[:1], 1, seed=0)
    eng.set_model(tmp)
    qs = [eng.soap(p, s_, l)[0] for p, s_, l in structs[:64]]
    model = S.model_from_soaps(h, np.concatenate(qs), 2000, seed=1002,
                               two_body=[S.carbon_2b(1002)])
    eng.set_model(model)
    out = {}
    for prec in ("mixed", "fp64"):
        eng.set_precision(prec)
        eng.energy_forces_batched(structs)
        reps = 5
        t0 = time.perf_counter()
        for _ in range(reps):
            res = eng.energy_forces_batched(structs)
        dt = (time.perf_counter() - t0) / reps
        eng.set_timing(True)  # one more call, with the kernel timing events
        eng.energy_forces_batched(structs)
        out[prec] = (dt, res, eng.kernel_times())
        eng.set_timing(False)
    dt, (offs, e_m, f_m, v_m), kt = out["mixed"]
    _, (_, e_d, f_d, v_d), _ = out["fp64"]
    value = n_atoms / dt
    print(f"[bench] batch256: {1e3 * dt:.3f} ms per batch, {value:.4g} atom-evals/s",
          file=sys.stderr, flush=True)
    line = {"metric": "atom-evaluations/s (E+F), 256-structure batched predict, mixed",
            "value": round(value, 1), "unit": "atom-evals/s", "n_gpus": 1, "steps": 5,
            "warmup": 1, "ms_per_step": round(1e3 * dt, 4), "higher_is_better": True,
            "scaling": "strong", "vs_baseline": None, "dtype": "f32+bf16", "data": "synthetic",
            "config": {"workload": "256 carbon cells, 32-1024 atoms log-uniform, 1.5-3.2 g/cm3, "
                                   "carbon SOAP, 2000 sparse, + C-C 2b", "atoms": n_atoms,
                       "structures": 256},
            "structures_per_s": round(256 / dt, 1),
            "fp64_ms_per_batch": round(1e3 * out["fp64"][0], 4),
            "mixed_vs_fp64": {"force_mae_eV_A": float(np.abs(f_m - f_d).mean()),
                              "energy_max_eV_atom": float(np.abs(e_m - e_d).max())},
            "kernel_ms": {k: round(v, 4) for k, v in kt.items()},
            "autotune": read_tune(tune_file)}
    print(json.dumps(line))
    eng.close()


def bench_cho4096(args):
    """C/H/O 60/30/10 at 1.5 g/cm^3, 4,096 atoms, rcut 4.0, n_max 8 per
    species (dim 2,700), 4,000 sparse from the cell's own SOAPs, zeta 4:
    E+F with the kernel GEMM in FP64, FP32 and BF16 MFMA, errors against the
    FP64 path (itself parity-tested against the oracle)."""
    pos, sp, lat = S.cho_cell(4096, 1.5, 1.0, 1001)
    eng = tgx.Engine(0, "fp64")
    tune_file = os.path.join(tempfile.gettempdir(), f"tgx_autotune_{os.getpid()}.txt")
    if not args.no_autotune:  # decided in the untimed first call of each mode
        eng.set_autotune(True, tune_file)
    h = S.cho_hypers()
    tmp = S.model_from_soaps(h, np.eye(h.dim)[:1], 1, seed=0)
    eng.set_model(tmp)
    q, _ = eng.soap(pos, sp, lat)
    model = S.model_from_soaps(h, q, 4000, seed=1001)
    eng.set_model(model)
    res = {}
    for prec in ("fp64", "fp32", "mixed"):
        eng.set_precision(prec)
        eng.energy_forces(pos, sp, lat)
        reps = 5
        t0 = time.perf_counter()
        for _ in range(reps):
            e, f, v = eng.energy_forces(pos, sp, lat)
        dt = (time.perf_counter() - t0) / reps
        eng.set_timing(True)  # one more call, with the kernel timing events
        eng.energy_forces(pos, sp, lat)
        res[prec] = dict(dt=dt, e=e, f=f, v=v, kt=eng.kernel_times())
        eng.set_timing(False)
    ref = res["fp64"]
    modes = {}
    for prec, r in res.items():
        modes[prec] = {"ms_per_eval": round(1e3 * r["dt"], 4),
                       "atom_evals_per_s": round(4096 / r["dt"], 1),
                       "kernel_ms": {k: round(v, 4) for k, v in r["kt"].items()},
                       "energy_max_eV_atom": float(np.abs(r["e"] - ref["e"]).max()),
                       "force_mae_eV_A": float(np.abs(r["f"] - ref["f"]).mean()),
                       "virial_max_eV": float(np.abs(r["v"] - ref["v"]).max()),
                       "virial_max_eV_per_atom": float(np.abs(r["v"] - ref["v"]).max() / 4096)}
    print(f"[bench] cho4096: " + ", ".join(f"{k} {v['ms_per_eval']} ms" for k, v in modes.items()),
          file=sys.stderr, flush=True)
    line = {"metric": "atom-evaluations/s (E+F), 4096-atom C/H/O, 4000 sparse, mixed",
            "value": modes["mixed"]["atom_evals_per_s"], "unit": "atom-evals/s", "n_gpus": 1,
            "steps": 5, "warmup": 1, "higher_is_better": True, "scaling": "strong",
            "vs_baseline": None, "dtype": "f32+bf16", "data": "synthetic",
            "config": {"workload": "C/H/O 60/30/10 4096 atoms 1.5 g/cm3, rcut 4.0, n_max 8 per "
                                   "species, l_max 8, dim 2700, 4000 sparse, zeta 4",
                       "atoms": 4096},
            "modes": modes,
            "autotune": read_tune(tune_file)}
    print(json.dumps(line))
    eng.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # the workload's own length: 1,000 velocity-Verlet steps of 1 fs
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--atoms", type=int, default=32768)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp32", "fp64"])
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cpu-1core", action="store_true",
                    help="skip the one-core oracle evaluation of the CPU baseline")
    ap.add_argument("--no-autotune", action="store_true",
                    help="md32k: built-in launch shapes instead of the autotuner's")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MB",
                    help="only run a plain device copy of MB MiB (PMC calibration)")
    ap.add_argument("--workload", default="md32k", choices=["md32k", "mc512", "batch256",
                                                            "cho4096"],
                    help="md32k: the headline NVE MD line; the others print their own line")
    ap.add_argument("--trials", type=int, default=4096, help="mc512: timed trials per mode")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(a run over several GPUs is not part of this benchmark).", file=sys.stderr)
        sys.exit(2)
    if args.workload == "mc512":
        bench_mc512(args)
        return
    if args.workload == "batch256":
        bench_batch256(args)
        return
    if args.workload == "cho4096":
        bench_cho4096(args)
        return
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return

    n = args.atoms
    pos, sp, lat = S.carbon_cell(n, 2.0, 1.2, 1003)
    masses = np.full(n, 12.011)
    vel = S.maxwell_boltzmann(n, masses, 3000.0, seed=1003)

    eng = tgx.Engine(0, "fp64")
    model = build_model(eng)
    eng.set_model(model)
    # launch shapes from the library's autotuner (timed on the first, eager
    # evaluation of each shape, before the MD step graph is captured)
    tune_file = os.path.join(tempfile.gettempdir(), f"tgx_autotune_{os.getpid()}.txt")
    if not args.no_autotune:
        eng.set_autotune(True, tune_file)

    def run_md(prec, steps, warmup):
        eng.set_precision(prec)
        eng.md_begin(pos, vel, sp, lat, masses, 1.0, 0.5, history=steps + warmup + 1)
        eng.md_steps(warmup)
        eng.sync()
        t0 = time.perf_counter()
        eng.md_steps(steps)
        eng.sync()
        dt = time.perf_counter() - t0
        out = eng.md_read()
        eng.md_end()
        return dt, out

    t_mixed, out_mixed = run_md(args.precision, args.steps, args.warmup)
    ms_per_step = 1e3 * t_mixed / args.steps
    value = n * args.steps / t_mixed
    print(f"[bench] {args.precision}: {ms_per_step:.3f} ms/step, {value:.4g} atom-steps/s, "
          f"rebuilds {out_mixed['n_rebuilds']}", file=sys.stderr, flush=True)
    et = out_mixed["e_pot"] + out_mixed["e_kin"]
    drift = float((et[-1] - et[0]) / n)

    fp64 = None
    if not args.no_fp64 and args.precision != "fp64":
        k64 = max(20, args.steps // 5)
        t64, _ = run_md("fp64", k64, 2)
        fp64 = n * k64 / t64
        print(f"[bench] fp64: {1e3 * t64 / k64:.3f} ms/step, {fp64:.4g} atom-steps/s",
              file=sys.stderr, flush=True)

    # Dominant kernel: single-kernel device times (HIP events on the library
    # stream, bracketing only that kernel) averaged over instrumented E+F calls
    # on the atoms in the MD run's Z-order.
    eng.set_precision(args.precision)
    zo = zorder(pos, lat)
    pos_z, sp_z = np.ascontiguousarray(pos[zo]), np.ascontiguousarray(sp[zo])
    # lists with the MD skin, so that each kernel sees the MD step's pair
    # slots and the profile's per-kernel averages cover one workload
    eng.set_list_buffer(0.5)
    kt = {"soap_fwd": 0.0, "gemm": 0.0, "soap_bwd": 0.0, "force_gather": 0.0}
    eng.set_timing(True)
    n_inst = 5
    for _ in range(n_inst):
        e_m, f_m, v_m = eng.energy_forces(pos_z, sp_z, lat)
        for k, v in eng.kernel_times().items():
            kt[k] += v / n_inst
    st = eng.stage_times()
    eng.set_timing(False)
    # per-atom results back in the input's order
    e_m, f_m = e_m[np.argsort(zo)], f_m[np.argsort(zo)]
    # host buffers in and out (PCIe, list rebuilt from scratch every call),
    # without the timing events
    t0 = time.perf_counter()
    for _ in range(n_inst):
        eng.energy_forces(pos_z, sp_z, lat)
    host_rate = n * n_inst / (time.perf_counter() - t0)
    eng.set_list_buffer(0.0)
    nb = eng.neighbors(pos, lat, model.hypers.rcut_max)
    n_pairs = int(np.sum(nb["rjs"] > 0))
    fwd_flop, bwd_flop = soap_costs(model.hypers, n, n_pairs)
    dom = max(("soap_fwd", "gemm", "soap_bwd"), key=kt.get)
    D, Sn = model.hypers.dim, len(model.alphas)
    gemm_flop = 4.0 * n * Sn * D
    flops = {"soap_fwd": fwd_flop, "gemm": gemm_flop, "soap_bwd": bwd_flop}[dom]
    ach = flops / (kt[dom] * 1e-3) / 1e12
    peak = PEAK_FP32_TFLOPS if dom != "gemm" or args.precision != "mixed" else 2500.0
    if args.precision == "fp64":
        peak = PEAK_FP64_TFLOPS
    # the dominant kernel's arithmetic unit: the kernel GEMM runs on MFMA in
    # every mode; the SOAP kernels on the FP32 (or FP64) vector ALUs
    if dom == "gemm":
        bound = "mfma"
    else:
        bound = "valu_fp64" if args.precision == "fp64" else "valu_fp32"
    roofline = {"bound": bound, "kernel": dom, "achieved": round(ach, 3), "peak": peak,
                "unit": "TFLOP/s", "frac": round(ach / peak, 5),
                "traffic": recorded_traffic(dom, args.precision, n),
                "kernel_ms": {k: round(v, 4) for k, v in kt.items()},
                "flop_per_launch": flops,
                "stage_ms": {k: round(v, 4) for k, v in st.items()}}

    cpu = None
    mae = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        # repeat full FP64 E+F evaluations of the whole cell for >= 10 s of CPU
        # work on every thread the host gives, then one evaluation on one core
        cores = oracle.threads()
        reps, t_cpu = 0, 0.0
        while reps == 0 or (t_cpu < 10.0 and reps < 20):
            t0 = time.perf_counter()
            e_o, f_o, v_o = oracle.energy_forces(model, pos, sp, lat, chunk_sites=2048)
            t_cpu += time.perf_counter() - t0
            reps += 1
        one = None
        if not args.no_cpu_1core:
            oracle.set_threads(1)
            t0 = time.perf_counter()
            oracle.energy_forces(model, pos, sp, lat, chunk_sites=2048)
            t1 = time.perf_counter() - t0
            oracle.set_threads(cores)
            one = {"value": round(n / t1, 2), "cores": 1,
                   "sample": f"1 FP64 E+F evaluation of the full {n}-atom cell, {t1:.1f} s"}
        cpu = {"value": round(n * reps / t_cpu, 2), "unit": "atom-steps/s", "cores": cores,
               "kind": "port", "cpu_model": cpu_model(),
               "sample": f"{reps} FP64 E+F evaluations of the full {n}-atom cell by the "
                         f"oracle (linked cells, SOAP with materialised dq/dr, kernel GEMMs, "
                         f"force loop), {t_cpu:.1f} s",
               "one_core": one}
        mae = {"force_mae_eV_A": float(np.abs(f_m - f_o).mean()),
               "force_max_eV_A": float(np.abs(f_m - f_o).max()),
               "energy_max_eV_atom": float(np.abs(e_m - e_o).max()),
               "virial_max_eV": float(np.abs(v_m - v_o).max()),
               "virial_max_eV_per_atom": float(np.abs(v_m - v_o).max() / n)}

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"mixed": "f32+bf16", "fp32": "f32", "fp64": "f64"}[args.precision],
        "data": "synthetic",
        "config": {"workload": f"NVE MD, {n} C atoms at 2.0 g/cm3, SOAP n_max 8 l_max 8 "
                               f"rcut 3.7, 500 sparse, zeta 4, + C-C 2b, dt 1 fs, skin 0.5 A",
                   "atoms": n, "precision": args.precision, "n_pairs_within_rcut": n_pairs},
        "fp64_value": None if fp64 is None else round(fp64, 1),
        "mixed_over_fp64": None if fp64 is None else round(value / fp64, 3),
        "accuracy_vs_fp64_oracle": mae,
        "energy_drift_eV_per_atom": drift,
        "host_buffer_atom_evals_per_s": round(host_rate, 1),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "autotune": read_tune(tune_file),
    }
    print(json.dumps(line))
    eng.close()


if __name__ == "__main__":
    main()
