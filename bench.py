"""Benchmark: NVE molecular dynamics of 32,768 carbon atoms (E+F every step).

Workload (BASELINE metric "atom-steps/s (E+F) at 32k C atoms, mixed vs FP64;
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


def build_model(eng):
    """Sparse set from the library's own FP64 SOAPs of a 216-atom cell."""
    pos, sp, lat = S.carbon_cell(216, 2.0, 1.2, 1000)
    h = S.carbon_hypers()
    tmp = S.model_from_soaps(h, np.eye(h.dim)[:1], 1, seed=0)
    eng.set_model(tmp)
    q, _ = eng.soap(pos, sp, lat)
    return S.model_from_soaps(h, q, 500, seed=1000, two_body=[S.carbon_2b(1000)])


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--atoms", type=int, default=32768)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp32", "fp64"])
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MB",
                    help="only run a plain device copy of MB MiB (PMC calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is described in DESIGN.md, not run).", file=sys.stderr)
        sys.exit(2)

    n = args.atoms
    pos, sp, lat = S.carbon_cell(n, 2.0, 1.2, 1003)
    masses = np.full(n, 12.011)
    vel = S.maxwell_boltzmann(n, masses, 3000.0, seed=1003)

    eng = tgx.Engine(0, "fp64")
    model = build_model(eng)
    eng.set_model(model)

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
        k64 = max(5, args.steps // 5)
        t64, _ = run_md("fp64", k64, 2)
        fp64 = n * k64 / t64
        print(f"[bench] fp64: {1e3 * t64 / k64:.3f} ms/step, {fp64:.4g} atom-steps/s",
              file=sys.stderr, flush=True)

    # Dominant kernel: single-kernel device times (HIP events on the library
    # stream, bracketing only that kernel) averaged over instrumented E+F calls.
    eng.set_precision(args.precision)
    # lists with the MD skin, so that each kernel sees the MD step's pair
    # slots and the profile's per-kernel averages cover one workload
    eng.set_list_buffer(0.5)
    kt = {"soap_fwd": 0.0, "gemm": 0.0, "soap_bwd": 0.0, "force_gather": 0.0}
    n_inst = 5
    t_host = 0.0
    for _ in range(n_inst):
        t0 = time.perf_counter()
        e_m, f_m, v_m = eng.energy_forces(pos, sp, lat)
        t_host += time.perf_counter() - t0
        for k, v in eng.kernel_times().items():
            kt[k] += v / n_inst
    # host buffers in and out (PCIe, list rebuilt from scratch every call)
    host_rate = n * n_inst / t_host
    st = eng.stage_times()
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
    roofline = {"bound": "mfma", "kernel": dom, "achieved": round(ach, 3), "peak": peak,
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

        # repeat full FP64 E+F evaluations of the whole cell for >= 10 s of CPU work
        reps, t_cpu = 0, 0.0
        while reps == 0 or (t_cpu < 10.0 and reps < 20):
            t0 = time.perf_counter()
            e_o, f_o, v_o = oracle.energy_forces(model, pos, sp, lat, chunk_sites=2048)
            t_cpu += time.perf_counter() - t0
            reps += 1
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = {"value": round(n * reps / t_cpu, 2), "unit": "atom-steps/s", "cores": cores,
               "kind": "port",
               "sample": f"{reps} FP64 E+F evaluations of the full {n}-atom cell by the "
                         f"oracle (linked cells, SOAP with materialised dq/dr, kernel GEMMs, "
                         f"force loop), {t_cpu:.1f} s"}
        mae = {"force_mae_eV_A": float(np.abs(f_m - f_o).mean()),
               "force_max_eV_A": float(np.abs(f_m - f_o).max()),
               "energy_max_eV_atom": float(np.abs(e_m - e_o).max()),
               "virial_max_eV": float(np.abs(v_m - v_o).max())}

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
    }
    print(json.dumps(line))
    eng.close()


if __name__ == "__main__":
    main()
