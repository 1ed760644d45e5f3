"""Benchmark: particle-bead-steps/s of the PIMD step on one MI355X.

Workload (the configuration the headline metric is quoted on): spin-polarised
uniform electron gas, N = 1024 electrons, P = 32 beads, r_s = 2, theta = 0.5,
xi = 1 (bosonic exchange), Coulomb by 3-D Ewald with alpha = 5/L and the
reference's |m_d| <= 8 reciprocal cube, massive Nose-Hoover chains (M = 4),
dt = 0.05.  Synthetic initial state (seeded uniform particle centres, beads
scattered by +-0.4 bohr, Maxwell velocities); inputs resident on the device
before timing.  One step = the full reference driver step (NHC half-kick,
drift, wrap, forces: trap/pair/Ewald, cycle energies, V_B forward/backward,
exchange forces, NHC half-kick), replayed as a captured hipGraph.

Prints ONE JSON line.  `roofline` is for the kernel of the step that takes
the most device time, chosen live between the real-space Ewald kernel
k_ewald_real (FP64 vector ALU) and the persistent exchange-recursion solve
k_rec_persist (FP64 vector ALU, limited by its serial dependency chain),
each timed alone with HIP events around it on its stream in the serial
stage profile: algorithmic FP64 work per launch / that average duration.  `kernels` gives the same figures for the other large
kernels, and `step_rate` the algorithmic work of the whole step over
ms_per_step.  `cpu_baseline` times the reference's own single-thread nocl
build (oracle/_ref) on a bounded sample of the same step.
"""
import argparse
import ctypes as C
import json
import math
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "md-lab_amd"))

import numpy as np  # noqa: E402

import pyxmd  # noqa: E402

METRIC = "particle-bead-steps/s, N=1024 P=32 UEG; surrogate vs Ewald at ΔE≤1e-4 Ha"
FP64_PEAK_TFLOPS = 78.6
# Algorithmic FP64 flop-equivalents (add/mul 1, FMA 2, exp/erfc/log/sincos 20):
# per real-space image term (r^2 5, sqrt 1, erfc 20, exp 20, prefactors 9, force 9)
FLOP_PER_IMAGE = 65.0
# per log-sum-exp term of the V_B / backward recursions (coefficient + value,
# max, exp, accumulate)
FLOP_PER_LSE_TERM = 23.0
# per (bead, half-space k vector) of the reciprocal sum: phase product and
# structure-factor accumulation (S), weight and force projection (F)
FLOP_PER_BEAD_K = 37.0
# per degree of freedom of the massive NHC update, both halves (M = 4)
FLOP_PER_DOF_NHC = 470.0


def recursion_work(N, xi):
    """Algorithmic flop-eq of both recursion directions (N(N+1)/2 log-sum-exp
    terms each; xi = 0 keeps only the N k = 1 terms)."""
    terms = N * (N + 1) / 2 if xi != 0 else N
    return 2 * terms * FLOP_PER_LSE_TERM


def step_work(N, P, xi, alpha, L, rcut, nk):
    """Algorithmic flop-eq of one full step at this workload, by stage."""
    pairs = N * (N - 1) / 2 * P
    w = {
        "ewald_real": pairs * expected_images(alpha, L, alpha * rcut) * FLOP_PER_IMAGE,
        "ewald_recip": N * P * nk * FLOP_PER_BEAD_K,
        "recursions": recursion_work(N, xi),
        # cycle energies (21 NP + 13.5 N^2) and exchange forces (one exp and
        # 2 x 7 flops per G entry, N^2 entries per end)
        "cycle_energies_exchange_forces": 21 * N * P + 13.5 * N * N + 2 * N * N * 34,
        "nhc_bead_forces": 3 * N * P * (FLOP_PER_DOF_NHC + 10),
    }
    w["total"] = sum(w.values())
    return w


def expected_images(alpha, L, xcut):
    """Mean number of periodic images of a random pair within r_c = xcut/alpha."""
    rc = xcut / alpha
    return 4.0 / 3.0 * math.pi * rc ** 3 / L ** 3


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE summary (profiles/hbm_traffic.json),
    or None when that file holds no entry for this kernel and workload."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None
    for e in table.get("entries", []):
        if e.get("kernel") == kernel and e.get("workload") == workload:
            return e.get("hbm_bytes_per_launch")
    return None


def counter_flops(kernel, workload):
    """FP64 FLOPs per launch of `kernel` counted by the hardware
    (SQ_INSTS_VALU_FLOPS_FP64[_TRANS] x 64 + SQ_INSTS_VALU_MFMA_MOPS_F64 x 512)
    from the committed rocprofv3 --pmc summary (profiles/fp64_counters.json),
    or None when it holds no entry for this workload."""
    path = os.path.join(ROOT, "profiles", "fp64_counters.json")
    try:
        with open(path) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None
    if table.get("workload") != workload:
        return None
    for name, e in table.get("kernels", {}).items():
        if kernel in name:
            return e["valu_fp64_flop"] + e["mfma_fp64_flop"]
    return None


def cpu_baseline(N, P, rs, theta, xi, nsum, pairs):
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_bench")
    if not os.path.exists(exe):
        return {"unavailable": "oracle/_ref/ref_bench is not built (it is compiled by __graft_entry__.build() "
                               "from the reference sources, which this machine does not have)"}
    out = subprocess.run(["taskset", "-c", "0", exe, str(N), str(P), str(rs), str(theta), str(xi), str(nsum),
                          str(pairs), "1"], capture_output=True, text=True, timeout=600)
    if out.returncode != 0:
        return {"unavailable": f"ref_bench exited with status {out.returncode}: {out.stderr.strip()[-200:]}"}
    kv = dict(t.split("=") for t in out.stdout.split())
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": float(kv["pbs_per_s"]), "unit": "particle-bead-steps/s", "cores": 1, "kind": "reference",
        "sample": (f"reference nocl build (gcc -O2), 1 thread on {cpu}: md_coulomb_3d_ewald_force_x timed on "
                   f"{kv['pairs_timed']} pairs ({float(kv['t_pairs']):.1f} s, {float(kv['per_pair']) * 1e3:.3f} ms/pair)"
                   f" scaled to N(N-1)/2*P pairs, plus one full-size exchange stage ({float(kv['t_exchange']):.2f} s)"
                   f" and integrator pass ({float(kv['t_integrator']):.3f} s): {float(kv['t_step_est']):.0f} s/step"),
    }


def replica_table(n=64, seed=2024):
    """Batched-replica workload: n systems, N log-uniform over {8, 16, ..., 512},
    P in {16, 32, 64}, half harmonic-trap Gaussian bosons (beta = 1), half UEG
    at r_s in {1, 2, 4} (theta = 0.5), xi uniform over {0, 0.25, 0.5, 0.75, 1}."""
    rng = np.random.default_rng(seed)
    table = []
    for _ in range(n):
        N = int(8 * 2 ** rng.integers(0, 7))
        P = int(rng.choice([16, 32, 64]))
        xi = float(rng.choice([0.0, 0.25, 0.5, 0.75, 1.0]))
        if rng.random() < 0.5:
            table.append(dict(kind="trap", N=N, P=P, xi=xi, T=1.0, L=10.0, rs=None))
        else:
            rs = float(rng.choice([1.0, 2.0, 4.0]))
            L = pyxmd.ueg_box(N, rs)
            table.append(dict(kind="ueg", N=N, P=P, xi=xi, T=pyxmd.ueg_temperature(N, L, 0.5), L=L, rs=rs))
    return table


def batch_bench(lib, steps, dt=0.01, one_by_one=True):
    """pbs/s of the replica table advanced by md_pimd_run_batch_x, and of the
    same replicas advanced one after another (each its own captured graph)."""
    table = replica_table()
    # One converged Ewald parameter set for every UEG box of the table (the
    # interaction parameters are process globals): alpha = 2 / bohr, |m| <= 200;
    # each box is then evaluated with the engine's own split.
    def build():
        out = []
        for i, r in enumerate(table):
            s = pyxmd.System(lib, r["kind"], r["N"], r["P"], r["T"], r["xi"], seed=100 + i, L=r["L"], nsum=200,
                             alpha_L=2.0 * r["L"])
            out.append(s)
        lib.md_set_coulomb_truc_x(2.0)
        lib.md_set_coulomb_n_sum_x(200)
        return out
    pbs = sum(r["N"] * r["P"] for r in table)
    res = {}
    for mode in ("batched", "one_by_one") if one_by_one else ("batched",):
        systems = build()
        cfgs = (pyxmd.StepConfig * len(systems))(*[s.step_config(dt, use_graph=1) for s in systems])
        ptrs = (pyxmd.PIMD_P * len(systems))(*[s.pimd for s in systems])
        run = (lambda k: lib.md_pimd_run_batch_x(ptrs, cfgs, len(systems), k)) if mode == "batched" else \
            (lambda k: sum(lib.md_pimd_run_steps_x(systems[i].pimd, C.byref(cfgs[i]), k)
                           for i in range(len(systems))))
        assert run(2) == 0 and lib.md_hip_synchronize_x() == 0
        t0 = time.perf_counter()
        assert run(steps) == 0 and lib.md_hip_synchronize_x() == 0
        sec = time.perf_counter() - t0
        res[mode] = pbs * steps / sec
        res[mode + "_ms_per_step"] = 1e3 * sec / steps
    kinds = {k: sum(1 for r in table if r["kind"] == k) for k in ("trap", "ueg")}
    return {"workload": f"{len(table)} replicas (seeded table: N log-uniform 8..512, P in {{16,32,64}}, "
                        f"{kinds['trap']} trap Gaussian beta=1, {kinds['ueg']} UEG r_s in {{1,2,4}} theta=0.5 "
                        f"converged Ewald, xi in {{0,.25,.5,.75,1}}), NHC M=4, dt={dt}",
            "particle_beads": pbs, "steps": steps, "unit": "particle-bead-steps/s",
            "value": res["batched"], "ms_per_step": res["batched_ms_per_step"],
            "one_by_one_value": res.get("one_by_one"), "one_by_one_ms_per_step": res.get("one_by_one_ms_per_step"),
            "speedup": res["batched"] / res["one_by_one"] if one_by_one else None}


def surrogate_bench(lib, s, cfg, steps, train_steps, nlow=5, K=50, f_thr=0.05, e_thr=1e-4, r_cut=8.0, lr=1e-3):
    """The learned stand-in for the reciprocal Ewald term on the headline
    workload: the low shells |m| <= nlow stay exact (FP64) and the network
    stands in for the rest of the reciprocal sum.  Online training on
    train_steps exact MD steps of this run (each bead slice is one
    FP64-labelled configuration); then `steps` steps with the surrogate alone
    (timed: the surrogate step's own rate), and `steps` steps with an exact
    judge every K steps (exact fallback with online training above the
    thresholds); errors against the FP64 reciprocal term on the final
    configuration, beside those of the exact low shells without the
    network."""
    sur = lib.md_alloc_ewald_surrogate_x(2024, r_cut)
    if not sur:
        return {"error": "surrogate allocation failed"}
    assert lib.md_ewald_surrogate_set_exact_shells_x(sur, nlow) == 0
    rep = np.zeros(2)
    t0 = time.perf_counter()
    assert lib.md_pimd_train_surrogate_steps_x(s.pimd, C.byref(cfg), sur, train_steps, lr, 1, pyxmd.dptr(rep)) == 0
    t_train = time.perf_counter() - t0
    chk0 = np.zeros(10)
    assert lib.md_ewald_surrogate_check_x(sur, s.pimd, pyxmd.dptr(chk0)) == 0
    run = np.zeros(8)
    # surrogate steps only (no judge): warm-up (captures the step graph), then timed
    assert lib.md_pimd_run_steps_surrogate_x(s.pimd, C.byref(cfg), sur, 40, 41, f_thr, e_thr, lr,
                                             pyxmd.dptr(run)) == 0
    assert lib.md_hip_synchronize_x() == 0
    t0 = time.perf_counter()
    assert lib.md_pimd_run_steps_surrogate_x(s.pimd, C.byref(cfg), sur, steps, steps + 1, f_thr, e_thr, lr,
                                             pyxmd.dptr(run)) == 0
    assert lib.md_hip_synchronize_x() == 0
    sec = time.perf_counter() - t0
    assert int(run[0]) == steps
    # judged run
    t0 = time.perf_counter()
    assert lib.md_pimd_run_steps_surrogate_x(s.pimd, C.byref(cfg), sur, steps, K, f_thr, e_thr, lr,
                                             pyxmd.dptr(run)) == 0
    assert lib.md_hip_synchronize_x() == 0
    sec_j = time.perf_counter() - t0
    chk = np.zeros(10)
    assert lib.md_ewald_surrogate_check_x(sur, s.pimd, pyxmd.dptr(chk)) == 0
    # the exact low shells alone: their errors are the network's targets
    nb = s.N * s.P
    et, ft = np.zeros(nb), np.zeros(3 * nb)
    assert lib.md_ewald_surrogate_labels_x(sur, s.pimd, pyxmd.dptr(et), pyxmd.dptr(ft)) == 0
    e_slice = et.reshape(s.N, s.P).sum(0) / s.N
    f_exact2 = float(np.sum(np.square(chk[5:8]))) * nb
    low_frel = float(np.sqrt(np.sum(ft ** 2) / f_exact2)) if f_exact2 > 0 else None
    return {
        "workload": (f"UEG N={s.N} P={s.P}: reciprocal shells |m| <= {nlow} exact (FP64), the rest by a network on "
                     f"64 radial + 32 angular descriptors (r_cut={r_cut} bohr), MLP 96-128-128-1 tanh, BF16 "
                     f"matrix-core inference with FP32 accumulation; online training on {train_steps} exact MD "
                     f"steps x {s.P} slices = {train_steps * s.P} FP64-labelled configurations (Adam, lr={lr}); "
                     f"judged run: exact judge every K={K} steps, fallback above a relative RMS force error "
                     f"{f_thr} or {e_thr} Ha/particle"),
        "value": nb * steps / sec, "unit": "particle-bead-steps/s", "ms_per_step": 1e3 * sec / steps,
        "steps": steps, "judged_ms_per_step": 1e3 * sec_j / steps, "train_seconds": t_train,
        "train_loss_first_last": [float(rep[0]), float(rep[1])],
        "judged_run": {"surrogate_steps": int(run[0]), "judge_steps": int(run[1]), "fallback_steps": int(run[2]),
                       "fallbacks": int(run[3]), "force_rel_rms_mean": float(run[4]),
                       "force_rel_rms_max": float(run[5]), "energy_error_per_particle_mean": float(run[6]),
                       "energy_error_per_particle_max": float(run[7])},
        "final_energy_error_per_particle": float(chk[0]), "final_energy_error_per_particle_worst_slice": float(chk[1]),
        "final_force_rms_error_per_component": [float(v) for v in chk[2:5]],
        "final_exact_recip_force_rms_per_component": [float(v) for v in chk[5:8]],
        "final_force_rel_rms_error": float(chk[8]), "after_training_force_rel_rms_error": float(chk0[8]),
        "after_training_energy_error_per_particle": float(chk0[0]),
        "exact_recip_energy_per_particle": float(chk[9]),
        "without_network": {"energy_error_per_particle": float(-e_slice.mean()),
                            "energy_error_per_particle_worst_slice": float(np.abs(e_slice).max()),
                            "force_rel_rms_error": low_frel},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--N", type=int, default=1024)
    ap.add_argument("--P", type=int, default=32)
    ap.add_argument("--cpu-pairs", type=int, default=25000, help="pairs in the reference CPU sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=10,
                    help="steps of the batched-replica measurement (0 skips it)")
    ap.add_argument("--only-batch", action="store_true",
                    help="run only the batched-replica measurement (for profiling it)")
    ap.add_argument("--lib", default=None,
                    help="engine library to load instead of the in-tree build (comparisons between builds)")
    ap.add_argument("--serial-only", action="store_true",
                    help="only run --steps steps with every stage alone on one stream (for per-kernel counters)")
    ap.add_argument("--sur-steps", type=int, default=200, help="timed surrogate steps (0 skips the surrogate leg)")
    ap.add_argument("--sur-train", type=int, default=625, help="exact MD steps of online surrogate training")
    ap.add_argument("--sur-shells", type=int, default=5, help="reciprocal shells |m| <= n kept exact beside the network")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a plain device-to-device copy of this many bytes (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch

        n = args.copy_probe // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": 8 * n, "dispatches": 3}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not supported "
                                   "(independent replicas, or the bead slices of one system, would shard "
                                   "across GPUs)"}))
        sys.exit(2)

    N, P, rs, theta, xi, nsum, dt = args.N, args.P, 2.0, 0.5, 1.0, 8, 0.05
    lib = pyxmd.Lib(args.lib) if args.lib else pyxmd.Lib()
    lib.md_set_nhc_length_x(4)
    if args.only_batch:
        print(json.dumps({"batched_replicas": batch_bench(lib, args.batch_steps, one_by_one=False)}))
        return
    L = pyxmd.ueg_box(N, rs)
    T = pyxmd.ueg_temperature(N, L, theta)
    s = pyxmd.System(lib, "ueg", N, P, T, xi, seed=1, L=L, nsum=nsum)
    cfg = s.step_config(dt, use_graph=1)
    names = ["nhc1_wrap", "cycle_energies", "vb_recursions", "bead_forces", "ewald_real", "ewald_recip",
             "exchange_forces", "nhc2", "step", "rec_persist", "ewald_real_kernel"]
    if args.serial_only:
        stages = np.zeros(11)
        assert lib.md_pimd_profile_steps_x(s.pimd, C.byref(cfg), args.steps, 1, pyxmd.dptr(stages)) == 0
        print(json.dumps({"serial_only": True, "steps": args.steps,
                          "stages_ms": {k: round(float(v), 5) for k, v in zip(names, stages)}}))
        return
    assert lib.md_pimd_compute_forces_x(s.pimd, C.byref(cfg)) == 0
    assert lib.md_pimd_run_steps_x(s.pimd, C.byref(cfg), args.warmup) == 0
    assert lib.md_hip_synchronize_x() == 0

    t0 = time.perf_counter()
    assert lib.md_pimd_run_steps_x(s.pimd, C.byref(cfg), args.steps) == 0
    assert lib.md_hip_synchronize_x() == 0
    t1 = time.perf_counter()
    sec = t1 - t0
    value = N * P * args.steps / sec

    # per-stage device times, every stage alone on one stream (HIP events)
    stages = np.zeros(11)
    assert lib.md_pimd_profile_steps_x(s.pimd, C.byref(cfg), 5, 1, pyxmd.dptr(stages)) == 0
    a_, rc_, nk_ = C.c_double(), C.c_double(), C.c_int()
    assert lib.md_pimd_get_ewald_plan_x(s.pimd, C.byref(a_), C.byref(rc_), C.byref(nk_)) == 0
    alpha, rcut, nk = a_.value, rc_.value, nk_.value
    work = step_work(N, P, xi, alpha, L, rcut, nk)
    rec_ms = stages[9]
    nblk = (N + 63) // 64
    # the real-space kernel alone (stage 10: HIP events around it on its
    # stream in the serial profile, without its partial-force reduction)
    real_ms = float(stages[10]) if stages[10] > 0 else float(stages[4])
    kernels = {}
    for name, key, ms in (("k_ewald_real", "ewald_real", real_ms),
                          ("k_rec_persist (V_B forward + backward recursion)", "recursions", rec_ms),
                          ("reciprocal stage (k_phase+k_sk_mfma+k_recip_force_mfma)", "ewald_recip", stages[5])):
        tf = work[key] / (ms * 1e-3) / 1e12
        kernels[name] = {"avg_ms": float(ms), "work_flop_eq": work[key], "achieved_tflops": tf,
                         "frac": tf / FP64_PEAK_TFLOPS}

    b = s.host_beads()
    finite = bool(np.all(np.isfinite(b["x"])) and np.all(np.isfinite(b["f"])))

    batched = batch_bench(lib, args.batch_steps) if args.batch_steps > 0 else None
    surrogate = (surrogate_bench(lib, s, cfg, args.sur_steps, args.sur_train, nlow=args.sur_shells)
                 if args.sur_steps > 0 else None)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(N, P, rs, theta, xi, nsum, args.cpu_pairs)

    traffic_key = f"UEG N={N} P={P} alpha'L={alpha * L:.2f}"
    step_tf = work["total"] / (sec / args.steps) / 1e12
    if real_ms >= rec_ms:
        kname, kwork, kms = "k_ewald_real", work["ewald_real"], real_ms
        roof_extra = {"kernel": "k_ewald_real (real-space Ewald, 64x64 bead tile pairs, both images of each pair)",
                      "limiter": "FP64 VALU issue and LDS table lookups",
                      "work_unit": f"pairs x mean images in r_c x {FLOP_PER_IMAGE:.0f} flop-eq"}
    else:
        kname, kwork, kms = "k_rec_persist", work["recursions"], rec_ms
        roof_extra = {"kernel": "k_rec_persist (V_B forward + backward recursion, both directions in one launch)",
                      "limiter": "2N-long serial dependency chain (latency), not throughput",
                      "work_unit": f"2 directions x N(N+1)/2 log-sum-exp terms x {FLOP_PER_LSE_TERM:.0f} flop-eq",
                      "blocks_per_direction": nblk, "us_per_block": 1e3 * float(rec_ms) / nblk}
    traffic = measured_traffic(kname, traffic_key)
    hw_flop = counter_flops(kname, traffic_key)
    achieved = kwork / (kms * 1e-3) / 1e12
    line = {
        "metric": METRIC, "value": value, "unit": "particle-bead-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * sec / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"UEG N={N} P={P} r_s={rs} theta={theta} xi={xi}, FP64 Ewald with the reference "
                               f"parameters alpha=5/L, |m|<=8 (converged; evaluated with the engine's split "
                               f"alpha'L={alpha * L:.2f}, r_c={rcut:.2f}, {nk} half-space k), NHC M=4, dt={dt}",
                   "path": "plain FP64 Ewald (no surrogate)", "energy_error_vs_fp64_ewald": None,
                   "hipgraph": True, "finite_state": finite},
        "roofline": dict({"bound": "fp64_valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                          "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                          "avg_ms": float(kms), "work_flop_eq": kwork, "counted_fp64_flop": hw_flop,
                          "counted_tflops": hw_flop / (kms * 1e-3) / 1e12 if hw_flop else None}, **roof_extra),
        "kernels": kernels,
        "step_rate": {"work_flop_eq_per_step": work, "achieved_tflops": step_tf,
                      "frac": step_tf / FP64_PEAK_TFLOPS},
        "stages_ms": {k: round(float(v), 5) for k, v in zip(names, stages)},
        "cpu_baseline": cpu,
        "batched_replicas": batched,
        "surrogate": surrogate,
    }
    if cpu and "value" in cpu:
        line["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
