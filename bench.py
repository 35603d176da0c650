#!/usr/bin/env python3
"""CCSD / (T) benchmark on one MI355X.

Workload (a "step" is one CCSD iteration): plain FP64 CCSD with No=40
occupied, Nv=400 virtual, NG=1600 auxiliary orbitals, the vertex-built ladder
in virtual slices of 64, DIIS(6), MP2 start, on a seeded synthetic vertex.
All tensors are resident in HBM when the timed region starts.

Printed as one JSON line:
  value        seconds per FP64 CCSD iteration (lower is better)
  roofline     the FP64 MFMA GEMM kernel family, timed with HIP events on
               the engine's stream (algorithmic 2*M*N*K per launch)
  cpu_baseline the FP64 CPU oracle timed on a sample of the same workload
  extra        iteration TFLOP/s from the reference's algorithmic flop
               count, split-BF16x3 iteration time, (T) block sample
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("CCSD iter time (s) + (T) time, FP64 vs mixed at |ΔE|≤1µEh; "
          "TFLOP/s % peak")
PEAK_FP64_TFLOPS = 78.6


def slice_pairs(Nv, s):
    n = [min(s, Nv - k) for k in range(0, Nv, s)]
    return (sum(n) ** 2 + sum(x * x for x in n)) / 2.0


def ccsd_iteration_flops(No, Nv, NG, s):
    """Algorithmic flops of one CCSD iteration (real arithmetic, 2*M*N*K per
    term): ladder 2 P(s) Nv^2 (NG + No^2), seven O(No^3 Nv^3) ring products,
    five K=NG products and the O(No^4 Nv^2) hole ladder."""
    return (2.0 * slice_pairs(Nv, s) * Nv ** 2 * (NG + No ** 2) + 14.0 * Nv ** 3 * No ** 3
            + 10.0 * NG * Nv ** 2 * No ** 2 + 6.0 * Nv ** 2 * No ** 4)


def triples_flops(No, Nv, NG):
    return No ** 3 * (2.0 * Nv ** 4 + 2.0 * Nv ** 3 * No) + 2.0 * NG * Nv ** 3 * No


def cpu_sample(g, No, Nv, NG, s, seconds_budget=30.0):
    """FP64 CPU oracle on one full-size ladder slice pair (s x s virtual
    block at the workload's Nv, NG, No): the dominant term of the iteration.
    Returns (seconds, flops) of the sample."""
    from oracle import ccsd as oc
    rng = np.random.default_rng(0)
    Gab = g[:, No:, No:]
    X = 0.01 * rng.standard_normal((Nv, Nv, No, No))
    R2 = np.zeros((Nv, Nv, No, No))
    t = time.perf_counter()
    Vxycd = oc.ein("Gxc,Gyd->xycd", Gab[:, s:2 * s, :], Gab[:, 0:s, :])
    Rxyij = oc.ein("xycd,cdij->xyij", Vxycd, X)
    oc.slice_into_residuum(Rxyij, s, 0, R2)
    dt = time.perf_counter() - t
    return dt, 2.0 * s * s * Nv * Nv * (NG + No * No)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--No", type=int, default=40)
    ap.add_argument("--Nv", type=int, default=400)
    ap.add_argument("--NG", type=int, default=1600)
    ap.add_argument("--slice", type=int, default=64)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--mixed-steps", type=int, default=2)
    ap.add_argument("--factors", type=int, default=0,
                    help="iteration mode: NR > 0 runs the factor-based ladder with NR random "
                         "complex CoulombFactors / FactorOrbitals (factor slices of Nv)")
    ap.add_argument("--triples-blocks", type=int, default=64,
                    help="(T) blocks timed in FP64 and compared with split-BF16x3")
    ap.add_argument("--triples-full", type=int, default=1,
                    help="time one full split-BF16x3 (T) pass over every i<=j<=k block")
    ap.add_argument("--refine-steps", type=int, default=4,
                    help="refinement iterations timed after the FP64 anchor")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--uccsd-no", type=int, default=10,
                    help="consumers: occupied spatial orbitals of the spin-orbital CCSD system")
    ap.add_argument("--uccsd-nv", type=int, default=60,
                    help="consumers: virtual spatial orbitals of the spin-orbital CCSD system")
    ap.add_argument("--eom", type=int, default=0,
                    help="consumers: also time the IP- and EA-EOM-CCSD right-applies (needs PPPP)")
    ap.add_argument("--mode", choices=["iteration", "converge", "batched", "complex", "consumers",
                                       "small", "gemm"],
                    default="iteration",
                    help="iteration: timed FP64 iterations (default line); converge: FP64 and "
                         "mixed runs to |dT|/|T| <= 1e-8 with (T); batched: 64 small systems")
    ap.add_argument("--gemm-shape", default="16000,16000,16000",
                    help="gemm mode: M,N,K of the timed product")
    ap.add_argument("--gemm-precision", default="bf16x3", choices=["fp64", "fp32", "bf16x3", "bf16"])
    ap.add_argument("--gemm-variant", type=int, default=-1,
                    help="gemm mode: kernel variant of the bf16 / split-BF16x3 products (-1: "
                         "automatic)")
    ap.add_argument("--gemm-reps", type=int, default=2)
    ap.add_argument("--systems", type=int, default=64)
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--batch-solver", choices=["lockstep", "threads"], default="lockstep",
                    help="batched mode: grouped launches over all systems, or one engine per "
                         "system on --workers host threads")
    ap.add_argument("--batched-slice", type=int, default=0,
                    help="batched mode: integralsSliceSize (0: No, the reference default)")
    ap.add_argument("--batched-refine-interval", type=int, default=0,
                    help="batched: FP64 anchor interval of the mixed runs (0: plain FP64 after the switch)")
    ap.add_argument("--graphs", type=int, default=0,
                    help="batched: replay each system's residual from a captured hipGraph")
    ap.add_argument("--level-shift", type=float, default=2.0,
                    help="levelShift of every CCSD run (reference argument; 2.0 keeps the "
                         "No=40 Nv=400 NG=1600 trajectory converging)")
    ap.add_argument("--max-iterations", type=int, default=400)
    ap.add_argument("--precisions", default="fp64,mixed")
    ap.add_argument("--switch-threshold", type=float, default=1e-5,
                    help="mixed runs: |dT|/|T| at which the bulk precision ends")
    ap.add_argument("--refine-precision", default="bf16", choices=["fp32", "bf16x3", "bf16"],
                    help="mixed runs: precision of the refinement residuals between anchors")
    ap.add_argument("--triples-precision", default="bf16x3", choices=["fp32", "bf16x3"],
                    help="(T) precision of the mixed run")
    ap.add_argument("--coarse-threshold", type=float, default=0.0,
                    help="mixed runs: plain-bf16 coarse stage until |dT|/|T| reaches this value")
    ap.add_argument("--refine-interval", type=int, default=128,
                    help="mixed runs: FP64 anchor every k iterations after the switch (0: plain "
                         "FP64 iterations)")
    ap.add_argument("--bulk-refine-interval", type=int, default=16,
                    help="mixed runs: split-BF16x3 anchor every k bulk iterations and "
                         "refinement residuals between (0: plain bulk iterations)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path "
              "(requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from sisi4s_amd import engine as eng

    if args.mode == "converge":
        return converge(args)
    if args.mode == "batched":
        return batched(args)
    if args.mode == "gemm":
        return gemm_shape(args)
    if args.mode == "consumers":
        return consumers(args)
    if args.mode == "small":
        return small_case(args)
    if args.mode == "complex":
        return complex_path(args)

    No, Nv, NG, s = args.No, args.Nv, args.NG, args.slice
    log = lambda m: print("[bench] " + m, file=sys.stderr, flush=True)  # noqa: E731
    t0 = time.perf_counter()
    g, ei, ea = eng.generate_system(args.seed, No, Nv, NG)
    log("generated system No=%d Nv=%d NG=%d in %.1fs" % (No, Nv, NG, time.perf_counter() - t0))

    e = eng.Engine(No, Nv, NG)
    e.set_vertex(g)
    e.set_eigenenergies(ei, ea)
    t0 = time.perf_counter()
    e.integrals_from_vertex()
    t_int = time.perf_counter() - t0
    log("integrals %.2fs" % t_int)
    if args.factors > 0:
        rng = np.random.default_rng(args.seed + 2)
        NR = args.factors
        lam = (rng.standard_normal((NG, NR)) + 1j * rng.standard_normal((NG, NR))) / np.sqrt(NR * NG)
        pi = (rng.standard_normal((No + Nv, NR)) + 1j * rng.standard_normal((No + Nv, NR))) / 2
        e.set_coulomb_factors(lam, pi, slice_size=Nv)
        log("factor-based ladder with NR=%d complex factors" % NR)

    # MP2 closed form sum V (2V - V') / D from the engine's PPHH, to check the
    # engine's MP2 energy and the driver's first iteration before anything is
    # reported: the MP2 start with a level shift s gives T = -V / (D + s), so
    # iteration 1 is the same sum over D + s
    V = e.get_integrals(eng.PPHH)
    D2 = (ea[:, None, None, None] + ea[None, :, None, None]
          - ei[None, None, :, None] - ei[None, None, None, :])
    W = V * (2.0 * V - V.transpose(1, 0, 2, 3))
    e_mp2_closed = float(np.sum(W / (-D2)))
    e_it1_closed = float(np.sum(W / (-(D2 + args.level_shift))))
    del V, D2, W
    e_mp2 = e.mp2_energy()
    log("MP2 %.10f vs closed form %.10f" % (e_mp2, e_mp2_closed))
    if abs(e_mp2 - e_mp2_closed) > 1e-9 * abs(e_mp2_closed):
        raise SystemExit("MP2 energy differs from the closed form by %.3e" % abs(e_mp2 - e_mp2_closed))

    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=10000, slice_size=s,
              amplitudes_convergence=1e-8, energy_convergence=1e-10, level_shift=args.level_shift)
    e.ccsd_begin(eng.default_opts(**kw))
    trajectory = []
    for k in range(max(args.warmup, 1)):
        t0 = time.perf_counter()
        rep = e.ccsd_iterate(1)
        trajectory.append((rep["energy"], rep["delta"]))
        log("warmup iteration %d: %.2fs E=%.10f dT/T=%.2e" % (k + 1, time.perf_counter() - t0,
                                                             rep["energy"], rep["delta"]))
        if k == 0:
            d1 = abs(rep["energy"] - e_it1_closed)
            log("iteration 1 (MP2 start, levelShift %g) %.10f vs closed form %.10f: |d| %.1e"
                % (args.level_shift, rep["energy"], e_it1_closed, d1))
            if d1 > 1e-9 * abs(e_it1_closed):
                raise SystemExit("iteration-1 energy differs from the closed form by %.3e" % d1)
    e.enable_stats(True)
    e.reset_stats()
    times = []
    for k in range(args.steps):
        t0 = time.perf_counter()
        rep = e.ccsd_iterate(1)
        times.append(time.perf_counter() - t0)
        trajectory.append((rep["energy"], rep["delta"]))
        log("timed iteration %d: %.3fs E=%.10f dT/T=%.2e" % (k + 1, times[-1], rep["energy"],
                                                             rep["delta"]))
    st = e.stats()
    e.enable_stats(False)
    t_iter = float(np.sum(times)) / max(len(times), 1)
    T1, T2 = e.amplitudes()  # FP64 amplitudes after the timed iterations, for (T)
    flops_it = ccsd_iteration_flops(No, Nv, NG, s)
    if args.factors > 0:
        # the vertex ladder's term replaced by the factor ladder's products per
        # factor slice pair, complex (four real products each)
        NR = args.factors
        flops_it += (4.0 * (4.0 * NR * Nv ** 2 * No ** 2 + 4.0 * NR ** 2 * Nv * No ** 2)
                     - 2.0 * slice_pairs(Nv, s) * Nv ** 2 * (NG + No ** 2))
    g_sec, g_flops, g_n = st["gemm_seconds"][0], st["gemm_flops"][0], st["gemm_launches"][0]
    achieved = g_flops / g_sec / 1e12 if g_sec > 0 else 0.0
    other_bw = st["other_bytes"] / st["other_seconds"] / 1e9 if st["other_seconds"] > 0 else 0.0

    # mixed schedule: split-BF16x3 bulk iterations, then refinement iterations
    # (an FP64 anchor followed by plain-bf16 residual changes), timed the same
    # way on the same trajectory
    mixed = {}
    if args.mixed_steps > 0:
        om = eng.default_opts(precision="bf16x3", switch_threshold=0.0, **kw)
        e.ccsd_begin(om)
        e.ccsd_iterate(2)
        e.enable_stats(True)
        e.reset_stats()
        mt = []
        for k in range(args.mixed_steps):
            t0 = time.perf_counter()
            e.ccsd_iterate(1)
            mt.append(time.perf_counter() - t0)
        sm = e.stats()
        e.enable_stats(False)
        mixed = dict(seconds_per_iteration=float(np.mean(mt)),
                     gemm_tflops_fp32=(sm["gemm_flops"][1] / sm["gemm_seconds"][1] / 1e12
                                       if sm["gemm_seconds"][1] > 0 else None),
                     gemm_tflops_bf16x3=(sm["gemm_flops"][2] / sm["gemm_seconds"][2] / 1e12
                                         if sm["gemm_seconds"][2] > 0 else None),
                     speedup_per_iteration=t_iter / float(np.mean(mt)))
        log("mixed bulk iteration %.3fs" % mixed["seconds_per_iteration"])
    refine = {}
    if args.refine_steps > 0:
        # switch early (threshold 1): bulk iterations, an FP64 anchor, then
        # plain-bf16 refinement iterations until the next anchor
        orf = eng.default_opts(precision="bf16x3", switch_threshold=1.0,
                               refine_interval=args.refine_interval,
                               refine_precision=args.refine_precision,
                               bulk_refine_interval=args.bulk_refine_interval, **kw)
        e.ccsd_begin(orf)
        rep_r = e.ccsd_iterate(2)
        # the switch waits for |dT|/|T| <= 1: bulk iterations until the first
        # FP64 anchor, timed as bulk-refinement iterations (split-BF16x3
        # anchors every bulk_refine_interval, plain-bf16 refinements between;
        # the anchors are the slow ones)
        bulk_t = []
        for k in range(30):
            if rep_r["fp64_iterations"] > 0:
                break
            t0 = time.perf_counter()
            rep_r = e.ccsd_iterate(1)
            bulk_t.append(time.perf_counter() - t0)
        if bulk_t and rep_r["fp64_iterations"] > 0:
            bulk_t = bulk_t[:-1]   # the last one was the FP64 anchor
        if args.bulk_refine_interval > 0 and len(bulk_t) >= 3:
            med = float(np.median(bulk_t))
            slow = [t for t in bulk_t if t > 1.5 * med]
            fast = [t for t in bulk_t if t <= 1.5 * med]
            refine_bulk = dict(bulk_refinement_seconds_per_iteration=float(np.mean(fast)),
                               bulk_anchor_seconds=float(np.mean(slow)) if slow else None,
                               iterations_timed=len(bulk_t),
                               bulk_refine_interval=args.bulk_refine_interval)
            log("bulk refinement iteration %.3f s, bulk anchor %s s"
                % (refine_bulk["bulk_refinement_seconds_per_iteration"],
                   refine_bulk["bulk_anchor_seconds"]))
        else:
            refine_bulk = None
        anchor_t, ref_t = [], []
        for k in range(args.refine_steps + 1):
            n64 = rep_r["fp64_iterations"]
            t0 = time.perf_counter()
            rep_r = e.ccsd_iterate(1)
            dt = time.perf_counter() - t0
            (anchor_t if rep_r["fp64_iterations"] > n64 else ref_t).append(dt)
        refine = dict(refinement_seconds_per_iteration=float(np.mean(ref_t)) if ref_t else None,
                      anchor_seconds=float(np.mean(anchor_t)) if anchor_t else None,
                      refinement_iterations_timed=len(ref_t), anchors_timed=len(anchor_t),
                      refine_precision=args.refine_precision,
                      refine_interval=args.refine_interval, bulk=refine_bulk)
        log("refinement iteration %s s, anchor %s s" % (refine["refinement_seconds_per_iteration"],
                                                        refine["anchor_seconds"]))

    # (T) on the FP64 amplitudes: FP64 and split-BF16x3 on a sample of blocks
    # (their energy difference), then one full split-BF16x3 pass
    triples = {}
    nblk = No * (No + 1) * (No + 2) // 6
    if args.triples_blocks > 0:
        rng = np.random.default_rng(1)
        allb = [(a, b, c) for a in range(No) for b in range(a, No) for c in range(b, No)]
        pick = rng.choice(len(allb), min(args.triples_blocks, len(allb)), replace=False)
        blocks = [allb[k] for k in sorted(pick)]
        e.triples(T1, T2, precision="fp64", blocks=[blocks[0]])  # builds the (T) integrals
        per = {}
        for tp in ("fp64", "bf16x3"):
            e.triples(T1, T2, precision=tp, blocks=[blocks[0]])
            t0 = time.perf_counter()
            et = e.triples(T1, T2, precision=tp, blocks=blocks)
            per[tp] = ((time.perf_counter() - t0) / len(blocks), et)
            log("(T) %s %.4fs per block on %d blocks" % (tp, per[tp][0], len(blocks)))
        triples = dict(sample_blocks=len(blocks),
                       fp64_seconds_per_block=per["fp64"][0],
                       bf16x3_seconds_per_block=per["bf16x3"][0],
                       fp64_tflops=(triples_flops(No, Nv, NG) / nblk) / per["fp64"][0] / 1e12,
                       sample_energy_fp64=per["fp64"][1],
                       sample_energy_bf16x3=per["bf16x3"][1],
                       sample_abs_difference=abs(per["fp64"][1] - per["bf16x3"][1]))
        if args.triples_full:
            t0 = time.perf_counter()
            et = e.triples(T1, T2, precision="bf16x3")
            tt = time.perf_counter() - t0
            triples.update(full_pass_bf16x3_seconds=tt, full_pass_blocks=nblk,
                           full_pass_energy_bf16x3=et,
                           full_pass_tflops_bf16x3=triples_flops(No, Nv, NG) / tt / 1e12)
            log("(T) full split-BF16x3 pass %.1fs over %d blocks" % (tt, nblk))

    cpu = None
    if not args.no_cpu:
        dt, fl = cpu_sample(g, No, Nv, NG, s)
        rate = fl / dt
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = dict(value=flops_it / rate, unit="s/iteration (extrapolated)", cores=cores,
                   kind="port", cpu_model=cpu_model(),
                   sample=("oracle ladder slice pair %dx%d at Nv=%d NG=%d No=%d: %.1f s, "
                           "%.1f GFLOP/s; iteration = algorithmic flops / rate"
                           % (s, s, Nv, NG, No, dt, rate / 1e9)))
        log("cpu sample %.1fs" % dt)

    traffic = None
    tj = os.path.join(ROOT, "profiles", "ccsd_fp64_No40_Nv400_NG1600_traffic.json")
    if (No, Nv, NG, s) == (40, 400, 1600, 64) and os.path.exists(tj):
        # HBM bytes per launch of the FP64 GEMM family from the committed
        # rocprofv3 FETCH_SIZE / WRITE_SIZE passes (calibrated; see profiles/)
        traffic = json.load(open(tj))["hbm_bytes_per_launch"]

    out = {
        "metric": METRIC,
        "value": t_iter,
        "unit": "s/iteration (CCSD, FP64)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_iter,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Philox vertex, No=%d Nv=%d NG=%d)" % (No, Nv, NG),
        "config": {"workload": "CCSD FP64 iteration, No=%d Nv=%d NG=%d, ladder slices of %d, "
                               "DIIS(6), MP2 start, levelShift %g%s"
                               % (No, Nv, NG, s, args.level_shift,
                                  ", factor-based ladder NR=%d" % args.factors
                                  if args.factors > 0 else ""),
                   "No": No, "Nv": Nv, "NG": NG, "slice": s, "mixer": "DiisMixer(6)",
                   "level_shift": args.level_shift},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic,
                     "kernel": "gemm_k4_kernel<FP64>", "launches": g_n,
                     "avg_launch_ms": 1e3 * g_sec / max(g_n, 1)},
        "cpu_baseline": cpu,
        "extra": {
            "iteration_tflops": flops_it / t_iter / 1e12,
            "iteration_pct_fp64_peak": 100.0 * flops_it / t_iter / 1e12 / PEAK_FP64_TFLOPS,
            "iteration_algorithmic_flops": flops_it,
            "gemm_seconds_per_iteration": g_sec / args.steps,
            "bandwidth_kernels_seconds_per_iteration": st["other_seconds"] / args.steps,
            "bandwidth_kernels_GBps": other_bw,
            "integrals_seconds": t_int,
            "mp2_energy": e_mp2,
            "mp2_closed_form": e_mp2_closed,
            "iteration1_energy": trajectory[0][0],
            "iteration1_closed_form": e_it1_closed,
            "trajectory_energy_delta": trajectory,
            "mixed_bf16x3_bulk": mixed,
            "refinement": refine,
            "triples": triples,
            "gemm_variants": {"bf16x3": eng.gemm_variants(3), "bf16": eng.gemm_variants(1)},
            "tuning_records": [l for l in eng.tune_dump().splitlines() if l and l[0] != "#"],
        },
    }
    print(json.dumps(out), flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def complex_vertex(g, No, seed):
    """A genuinely complex vertex with the physics of g: a random phase per
    auxiliary function and per orbital, Gamma[G,p,q] e^{i(f_G - t_p + t_q)};
    every Coulomb integral only picks up orbital phases."""
    rng = np.random.default_rng(seed)
    NG, Np, _ = g.shape
    f = np.exp(1j * rng.uniform(0, 2 * np.pi, NG))
    t = np.exp(1j * rng.uniform(0, 2 * np.pi, Np))
    z = g.astype(np.complex128)
    z *= f[:, None, None]
    z *= t.conj()[None, :, None]
    z *= t[None, None, :]
    return z


def small_case(args):
    """The reference's CPU-runnable case: No=10, Nv=50, NG=300, MP2 start,
    DIIS(6), 30 iterations, slices of No; the whole run on the GPU (FP64)
    beside the whole run of the FP64 CPU oracle on the host cores, and the
    largest energy difference over the 30 iterations."""
    from sisi4s_amd import engine as eng
    from oracle import ccsd as oc
    No, Nv, NG = 10, 50, 300
    g, ei, ea = eng.generate_system(args.seed, No, Nv, NG)
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=30, amplitudes_convergence=0.0,
              energy_convergence=0.0, slice_size=No)
    e = eng.Engine(No, Nv, NG)
    e.set_system(g, ei, ea)
    e.ccsd_solve(eng.default_opts(**kw))  # warm-up run (allocations, split caches)
    times = []
    for _ in range(max(args.steps, 1)):
        t0 = time.perf_counter()
        rep = e.ccsd_solve(eng.default_opts(**kw))
        times.append(time.perf_counter() - t0)
    gpu = float(np.median(times))
    t0 = time.perf_counter()
    ref = oc.solve(g, ei, ea, **kw)
    cpu = time.perf_counter() - t0
    dE = float(np.abs(e.energies() - np.array(ref["energies"])).max())
    print(json.dumps({"mode": "small", "workload": "No=10 Nv=50 NG=300, MP2 start, DIIS(6), "
                      "30 iterations, slices of No", "gpu_seconds": gpu,
                      "gpu_ms_per_iteration": 1e3 * gpu / 30, "cpu_seconds": cpu,
                      "cpu_kind": "port (numpy FP64 oracle)", "cpu_threads": os.cpu_count(),
                      "energy": rep["energy"], "max_energy_difference": dE}), flush=True)


def spin_orbital_classes(g, No, Nv):
    """Antisymmetrized spin-orbital classes <pq||rs> (PPHH, HHHH, PPPP, HHHP,
    PPPH, PHPH) of a closed-shell real vertex; spin orbital 2 P + s inside
    the hole and particle blocks."""
    h = np.repeat(np.arange(No), 2)
    p = np.repeat(np.arange(No, No + Nv), 2)
    sh, sp = np.tile([0, 1], No), np.tile([0, 1], Nv)
    idx = {"H": (h, sh), "P": (p, sp)}

    def block(cls):
        (a, sa), (b, sb), (c, sc), (d, sd) = (idx[x] for x in cls)

        def coulomb(i, j, k, l):  # <ij|kl> = (ik|jl) on spatial indices
            ui, uj, uk, ul = (np.unique(x) for x in (i, j, k, l))
            S = np.einsum("Gpr,Gqs->pqrs", g[:, ui][:, :, uk], g[:, uj][:, :, ul], optimize=True)
            return S[np.ix_(np.searchsorted(ui, i), np.searchsorted(uj, j),
                            np.searchsorted(uk, k), np.searchsorted(ul, l))]
        direct = coulomb(a, b, c, d) * ((sa[:, None, None, None] == sc[None, None, :, None])
                                        & (sb[None, :, None, None] == sd[None, None, None, :]))
        exch = coulomb(a, b, d, c).transpose(0, 1, 3, 2) * (
            (sa[:, None, None, None] == sd[None, None, None, :])
            & (sb[None, :, None, None] == sc[None, None, :, None]))
        return direct - exch
    return {c: block(c) for c in ("PPHH", "HHHH", "PPPP", "HHHP", "PPPH", "PHPH")}


def consumers(args):
    """Direct-ring CCD (DrccdEnergyFromCoulombIntegrals) at the given size and
    spin-orbital CCSD (UccsdAmplitudesFromCoulombIntegrals) on a closed-shell
    system in spin orbitals: FP64 iteration time and the GEMM family's rate."""
    from sisi4s_amd import engine as eng
    log = lambda m: print("[bench] " + m, file=sys.stderr, flush=True)  # noqa: E731
    out = {"metric": "iteration time (s), FP64: direct-ring CCD and spin-orbital CCSD",
           "mode": "consumers"}

    def timed(e, label, flops):
        opts = eng.default_opts(mixer="DiisMixer", max_residua=6, max_iterations=10000,
                                amplitudes_convergence=1e-8, energy_convergence=1e-10)
        e.ccsd_begin(opts)
        for k in range(args.warmup):
            e.ccsd_iterate(1)
        e.enable_stats(True)
        e.reset_stats()
        times = []
        for k in range(args.steps):
            t0 = time.perf_counter()
            rep = e.ccsd_iterate(1)
            times.append(time.perf_counter() - t0)
            log("%s iteration %d: %.3fs E=%.10f" % (label, k + 1, times[-1], rep["energy"]))
        st = e.stats()
        e.enable_stats(False)
        t = float(np.mean(times))
        gs, gf, gn = st["gemm_seconds"][0], st["gemm_flops"][0], st["gemm_launches"][0]
        return dict(seconds_per_iteration=t, algorithmic_flops=flops,
                    tflops=flops / t / 1e12 if flops else None,
                    gemm_tflops=gf / gs / 1e12 if gs > 0 else None,
                    gemm_frac_fp64_peak=(gf / gs / 1e12) / PEAK_FP64_TFLOPS if gs > 0 else None,
                    gemm_launches_per_iteration=gn / args.steps,
                    gemm_avg_launch_ms=1e3 * gs / max(gn, 1),
                    energy=rep["energy"])

    No, Nv, NG = args.No, args.Nv, args.NG
    g, ei, ea = eng.generate_system(args.seed, No, Nv, NG)
    e = eng.Engine(No, Nv, NG)
    e.set_system(g, ei, ea)
    del g
    e.set_drccd()
    nvo = float(Nv * No)
    out["drccd"] = dict(config="No=%d Nv=%d NG=%d, vertex integrals" % (No, Nv, NG),
                        **timed(e, "drccd", 4 * 2.0 * nvo ** 3))
    del e
    o, v = args.uccsd_no, args.uccsd_nv
    gs_, ei_, ea_ = eng.generate_system(args.seed + 1, o, v, 4 * v)
    t0 = time.perf_counter()
    ints = spin_orbital_classes(gs_, o, v)
    log("spin-orbital integrals %.1fs" % (time.perf_counter() - t0))
    u = eng.Engine(2 * o, 2 * v, 1)
    u.set_eigenenergies(np.repeat(ei_, 2), np.repeat(ea_, 2))
    for c, V in ints.items():
        u.set_integrals(getattr(eng, c), V)
    del ints
    u.set_formulation(eng.FORMULATION_UCCSD)
    O, W = 2 * o, 2 * v
    # dominant products: particle ladder 2 V^4 O^2, ring 2 O^3 V^3 (W_mbej T2),
    # and the tau <am||ef> product 2 V^3 O^3
    out["uccsd"] = dict(config="spin orbitals No=%d Nv=%d (closed shell No=%d Nv=%d, NG=%d)"
                        % (O, W, o, v, 4 * v),
                        **timed(u, "uccsd", 2.0 * W ** 4 * O ** 2 + 4.0 * O ** 3 * W ** 3))
    if args.eom:
        # EOM-CCSD right-applies on the same engine and amplitudes: wall time
        # per call through the C ABI (host buffers in and out, Hbar rebuilt
        # each call)
        T1, T2 = u.amplitudes()
        rng = np.random.default_rng(args.seed)
        r1, r2 = rng.normal(size=O), rng.normal(size=(O, O, W))
        r2 = r2 - r2.transpose(1, 0, 2)
        q1, q2 = rng.normal(size=W), rng.normal(size=(O, W, W))
        q2 = q2 - q2.transpose(0, 2, 1)
        for label, call, x1, x2 in (("ip", u.eom_ip_sigma, r1, r2), ("ea", u.eom_ea_sigma, q1, q2)):
            call(T1, T2, x1, x2)
            times = []
            for k in range(args.steps):
                t0 = time.perf_counter()
                call(T1, T2, x1, x2)
                times.append(time.perf_counter() - t0)
                log("%s-eom right-apply %d: %.3fs" % (label, k + 1, times[-1]))
            out["%s_eom_right_apply_seconds" % label] = float(np.mean(times))
    print(json.dumps(out), flush=True)


def complex_path(args):
    """Complex CCSD iterations (CcsdEnergyFromCoulombIntegrals.cxx:496-891)
    and complex (T) blocks at the given size, FP64; flops counted as four
    real products per complex product."""
    from sisi4s_amd import engine as eng
    No, Nv, NG, s = args.No, args.Nv, args.NG, args.slice
    log = lambda m: print("[bench] " + m, file=sys.stderr, flush=True)  # noqa: E731
    g, ei, ea = eng.generate_system(args.seed, No, Nv, NG)
    z = complex_vertex(g, No, args.seed + 1)
    del g
    e = eng.Engine(No, Nv, NG)
    e.set_vertex_complex(z)
    e.set_eigenenergies(ei, ea)
    del z
    t0 = time.perf_counter()
    e.integrals_from_vertex()
    t_int = time.perf_counter() - t0
    log("complex integrals %.2fs" % t_int)
    opts = eng.default_opts(mixer="DiisMixer", max_residua=6, max_iterations=10000,
                            slice_size=s, amplitudes_convergence=1e-8, energy_convergence=1e-10)
    e.ccsd_begin(opts)
    for k in range(args.warmup):
        t0 = time.perf_counter()
        rep = e.ccsd_iterate(1)
        log("warmup iteration %d: %.2fs E=%.10f" % (k + 1, time.perf_counter() - t0, rep["energy"]))
    e.enable_stats(True)
    e.reset_stats()
    times = []
    for k in range(args.steps):
        t0 = time.perf_counter()
        rep = e.ccsd_iterate(1)
        times.append(time.perf_counter() - t0)
        log("timed iteration %d: %.3fs E=%.10f Im E=%.1e" % (k + 1, times[-1], rep["energy"],
                                                              rep["energy_imag"]))
    st = e.stats()
    e.enable_stats(False)
    t_iter = float(np.mean(times))
    flops_it = 4.0 * ccsd_iteration_flops(No, Nv, NG, s)
    g_sec, g_flops, g_n = st["gemm_seconds"][0], st["gemm_flops"][0], st["gemm_launches"][0]
    achieved = g_flops / g_sec / 1e12 if g_sec > 0 else 0.0
    triples = {}
    if args.triples_blocks > 0:
        rng = np.random.default_rng(1)
        blocks = []
        while len(blocks) < args.triples_blocks:
            b = sorted(rng.choice(No, 3, replace=False).tolist())
            if b not in blocks:
                blocks.append(b)
        e.triples(precision="fp64", blocks=[blocks[0]])  # builds the (T) integrals
        nblk = No * (No + 1) * (No + 2) // 6
        t0 = time.perf_counter()
        e.triples(precision="fp64", blocks=blocks)
        tb = (time.perf_counter() - t0) / len(blocks)
        log("complex (T) %.4fs per block, %.0fs extrapolated" % (tb, tb * nblk))
        triples = dict(seconds_per_block=tb, blocks=nblk, seconds_extrapolated=tb * nblk,
                       tflops=4.0 * triples_flops(No, Nv, NG) / (tb * nblk) / 1e12,
                       sample="%d random i<j<k blocks, FP64" % len(blocks))
    out = {
        "metric": "complex CCSD iteration time (s), FP64; complex (T) time",
        "value": t_iter,
        "unit": "s/iteration (complex CCSD, FP64)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * t_iter,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "c128",
        "data": "synthetic (seeded Philox vertex with random auxiliary and orbital phases, "
                "No=%d Nv=%d NG=%d)" % (No, Nv, NG),
        "config": {"workload": "complex CCSD FP64 iteration, No=%d Nv=%d NG=%d, ladder slices of "
                               "%d, DIIS(6), MP2 start" % (No, Nv, NG, s),
                   "No": No, "Nv": Nv, "NG": NG, "slice": s},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                     "kernel": "gemm_k4_kernel<FP64>", "launches": g_n,
                     "avg_launch_ms": 1e3 * g_sec / max(g_n, 1)},
        "cpu_baseline": None,
        "extra": {"iteration_tflops": flops_it / t_iter / 1e12,
                  "iteration_pct_fp64_peak": 100.0 * flops_it / t_iter / 1e12 / PEAK_FP64_TFLOPS,
                  "iteration_algorithmic_flops": flops_it, "integrals_seconds": t_int,
                  "gemm_seconds_per_iteration": g_sec / args.steps,
                  "bandwidth_kernels_seconds_per_iteration": st["other_seconds"] / args.steps,
                  "energy_after_timed_iterations": rep["energy"],
                  "energy_imag": rep["energy_imag"], "triples": triples},
    }
    print(json.dumps(out), flush=True)


def converge(args):
    """FP64 and mixed CCSD to |dT|/|T| <= 1e-8 on the same system, then (T)
    on each run's amplitudes: time to the converged energy and |dE|."""
    from sisi4s_amd import engine as eng
    No, Nv, NG, s = args.No, args.Nv, args.NG, args.slice
    log = lambda m: print("[converge] " + m, file=sys.stderr, flush=True)  # noqa: E731
    g, ei, ea = eng.generate_system(args.seed, No, Nv, NG)
    e = eng.Engine(No, Nv, NG)
    e.set_system(g, ei, ea)
    runs = {}
    todo = [r for r in (("fp64", "fp64", "fp64"), ("mixed", "bf16x3", args.triples_precision))
            if r[0] in args.precisions.split(",")]
    for name, prec, tprec in todo:
        o = eng.default_opts(mixer="DiisMixer", max_residua=6, max_iterations=args.max_iterations,
                             slice_size=s, amplitudes_convergence=1e-8, energy_convergence=1e-10,
                             precision=prec, switch_threshold=args.switch_threshold,
                             level_shift=args.level_shift,
                             refine_interval=args.refine_interval if prec != "fp64" else 0,
                             refine_precision=args.refine_precision,
                             bulk_refine_interval=(args.bulk_refine_interval
                                                   if prec != "fp64" else 0),
                             coarse_threshold=args.coarse_threshold if prec != "fp64" else 0.0)
        t0 = time.perf_counter()
        e.ccsd_begin(o)
        rep = None
        while True:
            try:
                rep = e.ccsd_iterate(5)
            except eng.EngineError as ex:
                log("%s failed: %s" % (name, ex))
                rep = dict(rep or {}, converged=0, failed=str(ex))
                break
            log("%s it=%d E=%.10f dT/T=%.2e prec=%d (%.0fs)" % (
                name, rep["iterations"], rep["energy"], rep["delta"], rep["precision"],
                time.perf_counter() - t0))
            if rep["converged"] or rep["iterations"] >= o.max_iterations:
                break
        t_ccsd = time.perf_counter() - t0
        if rep.get("failed"):
            runs[name] = dict(failed=rep["failed"], ccsd_seconds=t_ccsd)
            continue
        blocks = None
        if args.triples_blocks > 0:
            rng = np.random.default_rng(1)
            blocks = []
            while len(blocks) < args.triples_blocks:
                b = sorted(rng.choice(No, 3, replace=False).tolist())
                if b not in blocks:
                    blocks.append(b)
        # one-time (T) integrals (and planes) outside the timed sample, once
        # per run: they are per system, not per block
        t0 = time.perf_counter()
        e.triples(precision=tprec, blocks=[[0, 1, 2]])
        t_setup = time.perf_counter() - t0
        t0 = time.perf_counter()
        et = e.triples(precision=tprec, blocks=blocks)
        t_t = time.perf_counter() - t0
        nblk = No * (No + 1) * (No + 2) // 6
        runs[name] = dict(ccsd_seconds=t_ccsd, iterations=rep["iterations"],
                          fp64_iterations=rep["fp64_iterations"], energy=rep["energy"],
                          converged=bool(rep["converged"]), triples_precision=tprec,
                          triples_seconds=t_setup + (t_t if blocks is None
                                                     else t_t * nblk / len(blocks)),
                          triples_sampled=blocks is not None, triples_energy_sample=et)
        if prec != "fp64":
            runs[name]["schedule"] = dict(switch_threshold=o.switch_threshold,
                                          refine_interval=o.refine_interval,
                                          refine_precision=args.refine_precision,
                                          bulk_refine_interval=o.bulk_refine_interval,
                                          coarse_threshold=o.coarse_threshold)
        log("%s done: %s" % (name, runs[name]))
    if len(runs) < 2:
        print(json.dumps({"mode": "converge", "level_shift": args.level_shift, "runs": runs}))
        return
    f, m = runs["fp64"], runs["mixed"]
    out = {"metric": METRIC, "mode": "converge", "level_shift": args.level_shift,
           "config": {"workload": "CCSD to |dT|/|T|<=1e-8 + (T), No=%d Nv=%d NG=%d slice=%d DIIS(6)"
                                  % (No, Nv, NG, s)},
           "fp64": f, "mixed": m,
           "ccsd_dE": abs(m["energy"] - f["energy"]),
           "triples_dE_sample": abs(m["triples_energy_sample"] - f["triples_energy_sample"]),
           "speedup_ccsd": f["ccsd_seconds"] / m["ccsd_seconds"],
           "speedup_total": (f["ccsd_seconds"] + f["triples_seconds"])
           / (m["ccsd_seconds"] + m["triples_seconds"])}
    print(json.dumps(out), flush=True)


def gemm_shape(args):
    """One product shape on device-resident operands (a warm-up call, then
    --gemm-reps timed calls): the GEMM microbenchmark the kernel profiles use."""
    from sisi4s_amd import engine as eng
    M, N, K = (int(x) for x in args.gemm_shape.split(","))
    passes = {"bf16": 1, "bf16x3": 3}.get(args.gemm_precision)
    name = None
    if passes is not None:
        eng.set_gemm_variant(passes, args.gemm_variant)
        if args.gemm_variant >= 0:
            name = eng.gemm_variants(passes)[args.gemm_variant]
    k_us, t_us = eng.gemm_bench(args.gemm_precision, M, N, K, reps=args.gemm_reps)
    out = {"mode": "gemm", "precision": args.gemm_precision, "M": M, "N": N, "K": K,
           "variant": name, "kernel_us": k_us, "call_us": t_us,
           "tflops": 2.0 * M * N * K / k_us / 1e6}
    print(json.dumps(out), flush=True)


def batched(args):
    """Independent seeded systems (64 by default), No in [4,16], Nv in [30,120],
    NG = 4 Nv, FP64 and mixed runs to 1e-8; per-system |dE|."""
    from sisi4s_amd import engine as eng
    rng = np.random.default_rng(args.seed)
    specs = []
    for k in range(args.systems):
        No = int(rng.integers(4, 17))
        Nv = int(rng.integers(30, 121))
        specs.append((args.seed + k, No, Nv, 4 * Nv))
    systems = [eng.generate_system(*sp) for sp in specs]
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=300, level_shift=args.level_shift,
              amplitudes_convergence=1e-8, energy_convergence=1e-10)
    res = {}
    lockstep = args.batch_solver == "lockstep"
    for name, prec in (("fp64", "fp64"), ("mixed", "fp32" if lockstep else "bf16x3")):
        t0 = time.perf_counter()
        # FP32 bulk then plain FP64 iterations: at these sizes no product is
        # large enough for the bf16 kernels, so refinement residuals (more
        # products per iteration at FP32 speed) would cost more than FP64
        ri = args.batched_refine_interval if prec != "fp64" else 0
        opts = eng.default_opts(precision=prec, refine_interval=ri, graphs=args.graphs,
                                slice_size=args.batched_slice, **kw)
        st = None
        if lockstep:
            E, reps, st = eng.ccsd_solve_lockstep(systems, opts, strict=False)
        else:
            E, reps = eng.ccsd_solve_batched(systems, opts, workers=args.workers, strict=False)
        res[name] = dict(seconds=time.perf_counter() - t0, energies=E.tolist(), lockstep=st,
                         failed=[k for k, r in enumerate(reps) if r["status"] != 0],
                         conv=[bool(r["converged"]) and r["status"] == 0 for r in reps],
                         converged=int(sum(r["converged"] for r in reps)),
                         iterations=int(sum(r["iterations"] for r in reps)),
                         fp64_iterations=int(sum(r["fp64_iterations"] for r in reps)))
        print("[batched] %s %.1fs" % (name, res[name]["seconds"]), file=sys.stderr, flush=True)
    # compare the systems that converged in both runs
    ok = [k for k in range(len(specs)) if res["fp64"]["conv"][k] and res["mixed"]["conv"][k]]
    dE = np.abs(np.array(res["fp64"]["energies"])[ok] - np.array(res["mixed"]["energies"])[ok])
    out = {"metric": METRIC, "mode": "batched",
           "config": {"workload": "%d seeded systems, No U[4,16], Nv U[30,120], NG=4Nv, DIIS(6), "
                                  "to 1e-8" % args.systems, "solver": args.batch_solver,
                      "workers": args.workers, "slice": args.batched_slice,
                      "mixed_refine_interval": args.batched_refine_interval,
                      "graphs": args.graphs},
           "fp64_seconds": res["fp64"]["seconds"], "mixed_seconds": res["mixed"]["seconds"],
           "speedup": res["fp64"]["seconds"] / res["mixed"]["seconds"],
           "max_dE": float(np.nanmax(dE)) if len(dE) else None,
           "systems_compared": len(ok),
           "systems_within_1uEh": int((dE <= 1e-6).sum()),
           "fp64_diverged": res["fp64"]["failed"],
           "level_shift": args.level_shift,
           "fp64": {k: v for k, v in res["fp64"].items() if k not in ("energies", "conv")},
           "mixed": {k: v for k, v in res["mixed"].items() if k not in ("energies", "conv")}}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
