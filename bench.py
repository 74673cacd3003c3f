#!/usr/bin/env python3
"""Benchmark: admixture-MCMC sweeps/s on one MI355X at N=2000, L=100000, K=5.

A step is one full sweep of the reference hot loop (structure.c:497-638) on
synthetic Balding-Nichols SNP data (K_true=5, F in [0.01, 0.1], Q ~ Dir(0.3),
1% missing) under the F model with alpha and lambda inference and
METROFREQ=10: UpdateP, UpdateQ (Gibbs, Metropolis every 10th sweep), the fused
UpdateZ, alpha, lambda, epsilon, Fst, and the post-burn-in data collection with
lnP(D).  W warm-up sweeps are burn-in; the K timed sweeps are all post-burn-in
(the costlier half of a run).  Genotypes, Q and P are resident in HBM before
timing starts.

Prints one JSON line (see the keys below); rank-0 only, one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
VALU_SIMDS = 1024          # MI355X: 256 CUs x 4 SIMDs
CLOCK_HZ = 2.4e9           # MI355X peak engine clock


def z_bytes_per_launch(N, L, K, A):
    """Algorithmic HBM bytes of one fused UpdateZ launch: 2-bit genotypes
    (N*L/4), P read (L*K*A doubles), allele counts written (L*K*A int32), Q read
    (N*K doubles), ancestry counts written (N*K int32)."""
    return N * L / 4 + L * K * A * 8 + L * K * A * 4 + N * K * 8 + N * K * 4


def sweep_bytes(N, L, K, A):
    """Algorithmic bytes per sweep: the UpdateZ traffic plus UpdateP's count read
    and P write and one pass over epsilon."""
    return z_bytes_per_launch(N, L, K, A) + L * K * A * (4 + 8) + L * A * 8


def cpu_baseline(N, L, K, sample_loci, sweeps, seed, ktrue=5, flo=0.01, fhi=0.1, qalpha=0.3):
    import numpy as np

    import oracle
    import structure_amd as sa
    d = sa.synth_snp(N, sample_loci, Ktrue=ktrue, flo=flo, fhi=fhi, qalpha=qalpha, miss=0.01,
                     seed=seed, want_geno=True)
    m = oracle.make_model(N, sample_loci, K, 2, freqscorr=1, inferalpha=1, inferlambda=1,
                          metrofreq=10, burnin=0, seed=seed, chain=0)
    ch = oracle.Chain(m, d["geno"], d["nall"])
    ch.init()
    ch.sweep(0)   # untimed first sweep
    t0 = time.perf_counter()
    for r in range(1, 1 + sweeps):   # reps 1..10 include one Metropolis sweep (rep 10)
        ch.sweep(r)
    dt = time.perf_counter() - t0
    del np
    per_sweep_sample = dt / sweeps
    # every phase of a sweep is linear in L at fixed N and K
    per_sweep_full = per_sweep_sample * (L / sample_loci)
    return {"value": 1.0 / per_sweep_full, "unit": "sweeps/s", "cores": 1, "kind": "port",
            "sample": (f"oracle (FP64 C restatement, -O2) on N={N}, L={sample_loci} "
                       f"(1/{L // sample_loci} of the loci), {sweeps} post-burn-in sweeps in "
                       f"{dt:.1f} s, scaled linearly to L={L}")}


def batched_ksweep(d, N, L, seed, sweeps, warmup, kmax=10, reps=10):
    """The K-sweep users run (K = 1..kmax, reps Philox seeds each) on one data
    set: chain-sweeps/s with all chains in one batch (one launch per kernel and
    phase, HIP-graph replay) against the same chains run one at a time on the
    GPU (each alone in its own context, graph replay)."""
    import structure_amd as sa
    models = [sa.make_model(N, L, K, 2, freqscorr=1, inferalpha=1, inferlambda=1, metrofreq=10,
                            burnin=warmup, numreps=sweeps, seed=seed + r, chain=K)
              for K in range(1, kmax + 1) for r in range(reps)]
    data = [(sa.GENO_SNP2, d["words"], d["nall"])]
    b = sa.Batch(models, data)
    b.init()
    b.run(0, warmup)
    b.synchronize()
    t0 = time.perf_counter()
    b.run(warmup, sweeps)
    b.synchronize()
    t_batch = time.perf_counter() - t0
    b.close()
    t_plain = 0.0
    for m in models:
        ch = sa.Chain(m, sa.GENO_SNP2, d["words"], d["nall"])
        ch.init()
        ch.run(0, warmup)
        ch.synchronize()
        t0 = time.perf_counter()
        ch.run(warmup, sweeps)
        ch.synchronize()
        t_plain += time.perf_counter() - t0
        ch.close()
    n = len(models)
    return {"chains": n, "K": f"1..{kmax} x {reps} seeds", "sweeps_timed": sweeps,
            "warmup": warmup,
            "batched_chain_sweeps_per_s": round(n * sweeps / t_batch, 2),
            "plain_chain_sweeps_per_s": round(n * sweeps / t_plain, 2),
            "speedup": round(t_plain / t_batch, 3),
            "batched_site_updates_per_s": round(
                sum(N * L * m.K for m in models) * sweeps / t_batch, 1)}


def _timed_run(ch, warmup, steps, direct=False):
    """Warm-up sweeps (burn-in), then `steps` timed sweeps from graph replay (or,
    with `direct`, launched one kernel at a time); returns seconds."""
    if direct:
        ch.set_timing(True)   # timing mode launches every phase directly
    ch.init()
    ch.run(0, warmup)
    ch.synchronize()
    t0 = time.perf_counter()
    ch.run(warmup, steps)
    ch.synchronize()
    return time.perf_counter() - t0


def _z_roofline(ch, start, n, zbytes):
    """Average UpdateZ launch time over n directly launched sweeps (HIP events
    on the context stream) and the algorithmic-byte HBM fraction."""
    ch.set_timing(True)
    ch.get_timing()
    ch.run(start, n)
    t = ch.get_timing()
    ch.set_timing(False)
    zms, zn = t["Z"]
    z_avg = zms / max(zn, 1) / 1e3
    ach = zbytes / z_avg / 1e9
    return t, {"bound": "hbm", "achieved": round(ach, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
               "frac": round(ach / PEAK_HBM_GBS, 5), "traffic": None,
               "kernel": "k_update_z_* (fused UpdateZ + counts + lnP)",
               "algorithmic_bytes_per_launch": zbytes, "avg_launch_ms": round(z_avg * 1e3, 4),
               "launches_timed": zn}


def workload_c3(args):
    """Large single chain: N=20000, L=500000 SNPs (2.5 GB of 2-bit genotypes),
    K=8, F model; Z never materialised."""
    import structure_amd as sa
    N, L, K = 20000, 500000, 8
    t0 = time.perf_counter()
    d = sa.synth_snp(N, L, Ktrue=8, flo=0.005, fhi=0.05, qalpha=0.2, miss=0.01, seed=args.seed)
    print(f"bench: synthetic data N={N} L={L} in {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)
    warm, steps = args.warmup, args.steps
    model = sa.make_model(N, L, K, 2, freqscorr=1, inferalpha=1, inferlambda=1, metrofreq=10,
                          burnin=warm, numreps=steps, seed=args.seed, chain=0)
    ch = sa.Chain(model, sa.GENO_SNP2, d["words"], d["nall"])
    dt = _timed_run(ch, warm, steps)
    zb = z_bytes_per_launch(N, L, K, 2)
    timing, roof = _z_roofline(ch, warm + steps, min(steps, 5), zb)
    ch.close()
    sps = steps / dt
    res = {"metric": "sweeps/s at N=20k,L=500k,K=8 (%HBM peak)", "value": round(sps, 3),
           "unit": "sweeps/s", "n_gpus": 1, "steps": steps, "warmup": warm,
           "ms_per_step": round(1e3 * dt / steps, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": f"synthetic (Balding-Nichols SNP, K_true=8, seed {args.seed})",
           "config": {"workload": "config 3: N=20000 L=500000 K=8 biallelic 2-bit, F model + "
                                  "alpha + lambda, METROFREQ=10, post-burn-in sweeps",
                      "N": N, "L": L, "K": K, "genotype_format": "SNP2"},
           "roofline": roof, "site_updates_per_s": round(N * L * K * sps, 1),
           "phase_ms_per_sweep": {p: round(v[0] / max(min(steps, 5), 1), 4)
                                  for p, v in timing.items()}}
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(N, L, K, 2000, 5, args.seed, ktrue=8, flo=0.005,
                                           fhi=0.05, qalpha=0.2)
    return res


def _ragged_datasets(seed, n=64):
    """Config 5: n independent data sets, N ~ U[100,5000], L log-uniform in
    [1e3, 2e5], half SNP (Balding-Nichols) and half microsatellite (2-20
    alleles), K_true and run K in [2, 8], Q ~ Dir(U[0.1,1]), missing U[0,5%]."""
    import concurrent.futures as cf

    import numpy as np

    import structure_amd as sa
    rng = np.random.default_rng(seed)
    specs = []
    for j in range(n):
        specs.append(dict(N=int(rng.integers(100, 5001)),
                          L=int(round(10 ** rng.uniform(3, np.log10(2e5)))),
                          snp=j % 2 == 0, ktrue=int(rng.integers(2, 9)), K=int(rng.integers(2, 9)),
                          qalpha=float(rng.uniform(0.1, 1.0)), miss=float(rng.uniform(0, 0.05)),
                          seed=int(rng.integers(1, 2 ** 31))))

    def make(sp):
        if sp["snp"]:
            d = sa.synth_snp(sp["N"], sp["L"], Ktrue=sp["ktrue"], qalpha=sp["qalpha"],
                             miss=sp["miss"], seed=sp["seed"], threads=16)
            return sa.GENO_SNP2, d["words"], d["nall"], 2
        d = sa.synth_microsat(sp["N"], sp["L"], amin=2, amax=20, Ktrue=sp["ktrue"],
                              qalpha=sp["qalpha"], miss=sp["miss"], seed=sp["seed"])
        fmt, words = sa.pack(d["geno"], sa.GENO_BYTE)
        return fmt, words, d["nall"], d["A"]

    # the generators are multithreaded themselves; two at a time overlaps packing
    with cf.ThreadPoolExecutor(max_workers=2) as ex:
        data = list(ex.map(make, specs))
    return specs, data


def _r3(x):
    return None if x is None else round(x, 3)


def _batch_vs_plain(models, datasets, doc, warm, sweeps, plain_sweeps=None, direct=False):
    """Chain-sweeps/s of all chains in one batch against the same chains each
    run alone (one after the other) on the GPU.  With `direct` (for kernel
    tracers) the batch's sweeps are launched without graphs and the chains are
    not run alone."""
    import structure_amd as sa
    b = sa.Batch(models, datasets, doc)
    t_batch = _timed_run(b, warm, sweeps, direct)
    # per-phase device time of a batch sweep (events around each phase, direct launches)
    nt = min(sweeps, 20)
    b.set_timing(True)
    b.get_timing()
    b.run(warm + sweeps, nt)
    tm = b.get_timing()
    b.set_timing(False)
    phases = {p: round(v[0] / nt, 4) for p, v in tm.items()}
    b.close()
    n = len(models)
    if direct:
        return n * sweeps / t_batch, None, phases
    ps = plain_sweeps or sweeps
    t_plain = 0.0
    for m, j in zip(models, doc):
        fmt, words, nall = datasets[j]
        ch = sa.Chain(m, fmt, words, nall)
        t_plain += _timed_run(ch, warm, ps)
        ch.close()
    return n * sweeps / t_batch, n * ps / t_plain, phases


def workload_c5(args):
    """64 ragged data sets, one chain each, all in one batch."""
    import structure_amd as sa
    t0 = time.perf_counter()
    specs, data = _ragged_datasets(args.seed)
    print(f"bench: 64 ragged data sets in {time.perf_counter() - t0:.1f}s", file=sys.stderr,
          flush=True)
    warm, sweeps = args.warmup, args.steps
    models = [sa.make_model(sp["N"], sp["L"], sp["K"], A, freqscorr=1, inferalpha=1,
                            metrofreq=10, burnin=warm, numreps=sweeps, seed=args.seed + j,
                            chain=j)
              for j, (sp, (_, _, _, A)) in enumerate(zip(specs, data))]
    dsets = [(f, w, na) for f, w, na, _ in data]
    batched, plain, phases = _batch_vs_plain(models, dsets, list(range(len(models))), warm,
                                             sweeps, direct=args.direct)
    sites = sum(sp["N"] * sp["L"] * sp["K"] for sp in specs)
    return {"metric": "chain-sweeps/s, 64 ragged data sets in one batch",
            "value": round(batched, 3), "unit": "chain-sweeps/s", "n_gpus": 1, "steps": sweeps,
            "warmup": warm, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64", "data": f"synthetic ragged (seed {args.seed})",
            "config": {"workload": "config 5: 64 data sets, N~U[100,5000], L log-uniform "
                                   "[1e3,2e5], half 2-bit SNP / half byte microsatellite "
                                   "(2-20 alleles), K in [2,8], F model + alpha",
                       "datasets": len(specs), "total_sites": sum(sp["N"] * sp["L"] for sp in specs)},
            "plain_chain_sweeps_per_s": _r3(plain), "speedup": _r3(plain and batched / plain),
            "site_updates_per_s": round(sites * batched / len(specs), 1),
            "phase_ms_per_batch_sweep": phases,
            "launch": "direct" if args.direct else "graph replay"}


def workload_c1batch(args):
    """Config 1 replicated: 100 chains of N=200, L=500 microsatellites, K=3,
    no F model, alpha inferred; batched against one at a time."""
    import structure_amd as sa
    d = sa.synth_microsat(200, 500, amin=2, amax=10, Ktrue=3, qalpha=0.5, miss=0.02,
                          seed=args.seed)
    fmt, words = sa.pack(d["geno"])
    A = d["A"]
    warm, sweeps = args.warmup, args.steps
    models = [sa.make_model(200, 500, 3, A, freqscorr=0, inferalpha=1, lambda0=1.0,
                            metrofreq=10, burnin=warm, numreps=sweeps, seed=args.seed + r,
                            chain=3) for r in range(100)]
    batched, plain, phases = _batch_vs_plain(models, [(fmt, words, d["nall"])], [0] * 100,
                                             warm, sweeps, plain_sweeps=max(sweeps // 10, 20),
                                             direct=args.direct)
    res = {"metric": "chain-sweeps/s, 100 replicate chains of config 1 in one batch",
           "value": round(batched, 3), "unit": "chain-sweeps/s", "n_gpus": 1, "steps": sweeps,
           "warmup": warm, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": f"synthetic microsatellites (seed {args.seed})",
           "config": {"workload": "config 1 x 100 seeds: N=200 L=500 A in [2,10] K=3, "
                                  "no F model, alpha inferred", "chains": 100},
           "plain_chain_sweeps_per_s": _r3(plain), "speedup": _r3(plain and batched / plain),
           "phase_ms_per_batch_sweep": phases,
           "launch": "direct" if args.direct else "graph replay"}
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_c1(d, A, args.seed)
    return res


def cpu_baseline_c1(d, A, seed):
    """The oracle on one config-1 chain (one core), full-size data."""
    import oracle
    m = oracle.make_model(200, 500, 3, A, freqscorr=0, inferalpha=1, lambda0=1.0, metrofreq=10,
                          burnin=0, seed=seed, chain=3)
    ch = oracle.Chain(m, d["geno"], d["nall"])
    ch.init()
    ch.sweep(0)
    n = 2500   # about 10 s of one core
    t0 = time.perf_counter()
    for r in range(1, 1 + n):
        ch.sweep(r)
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "chain-sweeps/s", "cores": 1, "kind": "port",
            "sample": f"oracle (FP64 C restatement) on one config-1 chain, {n} sweeps in "
                      f"{dt:.1f} s (one chain per core: multiply by the cores used)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--N", type=int, default=2000)
    ap.add_argument("--L", type=int, default=100000)
    ap.add_argument("--K", type=int, default=5)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-sample-loci", type=int, default=10000)
    ap.add_argument("--cpu-sweeps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--direct", action="store_true",
                    help="c5 / c1batch: launch the timed sweeps one kernel at a time instead of "
                         "replaying HIP graphs, and skip the one-chain-at-a-time runs (for "
                         "kernel tracers)")
    ap.add_argument("--batch-sweeps", type=int, default=10)
    ap.add_argument("--tune", action="store_true",
                    help="search the tile shapes (structure_amd.tune) before timing")
    ap.add_argument("--workload", default="c2", choices=["c2", "c3", "c5", "c1batch"],
                    help="c2 (default, the headline line), c3, c5 or c1batch")
    args, _ = ap.parse_known_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; multi-GPU runs are not part of it "
              f"(asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)
    if args.workload != "c2":
        defaults = {"c3": (3, 10), "c5": (10, 20), "c1batch": (20, 200)}
        w, k = defaults[args.workload]
        if "--warmup" not in sys.argv:
            args.warmup = w
        if "--steps" not in sys.argv:
            args.steps = k
        fn = {"c3": workload_c3, "c5": workload_c5, "c1batch": workload_c1batch}
        print(json.dumps(fn[args.workload](args)), flush=True)
        return

    import structure_amd as sa
    N, L, K = args.N, args.L, args.K
    t0 = time.perf_counter()
    d = sa.synth_snp(N, L, Ktrue=5, flo=0.01, fhi=0.1, qalpha=0.3, miss=0.01, seed=args.seed)
    print(f"bench: synthetic data N={N} L={L} in {time.perf_counter() - t0:.1f}s", file=sys.stderr,
          flush=True)
    model = sa.make_model(N, L, K, 2, freqscorr=1, inferalpha=1, inferlambda=1, metrofreq=10,
                          burnin=args.warmup, numreps=args.steps, seed=args.seed, chain=0)
    tuning = None
    if args.tune:
        from structure_amd import tune
        tuning, _ = tune.autotune([model], [(sa.GENO_SNP2, d["words"], d["nall"])], cache=None,
                                  log=lambda s: print("bench: " + s, file=sys.stderr, flush=True))
    ch = sa.Batch([model], [(sa.GENO_SNP2, d["words"], d["nall"])], tuning=tuning)
    ch.init()
    ch.run(0, args.warmup)
    ch.synchronize()
    print(f"bench: {args.warmup} warm-up sweeps done", file=sys.stderr, flush=True)

    # timed region: K sweeps replayed from captured HIP graphs
    ch.synchronize()
    t0 = time.perf_counter()
    ch.run(args.warmup, args.steps)
    ch.synchronize()
    dt = time.perf_counter() - t0
    like = ch.get_chain(0, sa.LIKE)
    # per-phase kernel durations: the same sweeps launched directly, with HIP
    # events around every phase on the chain's stream
    nt = min(args.steps, 100)
    ch.set_timing(True)
    ch.get_timing()   # reset
    ch.run(args.warmup + args.steps, nt)
    timing = ch.get_timing()
    ch.set_timing(False)
    ch.close()

    sweeps_per_s = args.steps / dt
    zms, zn = timing["Z"]
    z_avg_s = zms / max(zn, 1) / 1e3
    zb = z_bytes_per_launch(N, L, K, 2)
    achieved = zb / z_avg_s / 1e9
    traffic, valu_instr = None, None
    tpath = os.path.join(ROOT, "profiles", "updatez_c2_traffic.json")
    if os.path.exists(tpath):
        try:
            with open(tpath) as f:
                prof = json.load(f)
            traffic = prof.get("bytes_per_launch")
            valu_instr = prof.get("valu_instr_per_launch")
        except (OSError, ValueError):
            traffic, valu_instr = None, None
    # the binding bound: VALU issue, one wave64 instruction per 4 cycles per SIMD
    valu = None
    if valu_instr:
        peak_ips = VALU_SIMDS * CLOCK_HZ / 4.0
        valu = {"instr_per_launch": valu_instr, "source": "rocprofv3 SQ_INSTS_VALU of this kernel",
                "achieved_instr_per_s": round(valu_instr / z_avg_s, 1),
                "peak_instr_per_s": peak_ips,
                "issue_frac": round(valu_instr / z_avg_s / peak_ips, 4)}
    result = {
        "metric": "sweeps/s at N=2k,L=100k,K=5 (%HBM peak); batched vs plain chain-sweeps/s",
        "value": round(sweeps_per_s, 3),
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * dt / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": f"synthetic (Balding-Nichols SNP, seed {args.seed}); random initial state",
        "config": {"workload": (f"config 2: N={N} L={L} K={K} biallelic 2-bit, F model + "
                                f"alpha + lambda, METROFREQ=10, post-burn-in sweeps with lnP(D)"),
                   "N": N, "L": L, "K": K, "genotype_format": "SNP2"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 5),
                     "traffic": traffic, "kernel": "k_update_z_snp2 (fused UpdateZ + counts + lnP)",
                     "algorithmic_bytes_per_launch": zb, "avg_launch_ms": round(z_avg_s * 1e3, 4),
                     "launches_timed": zn, "valu": valu,
                     "traffic_frac": (round(traffic / z_avg_s / 1e9 / PEAK_HBM_GBS, 5)
                                      if traffic else None)},
        "site_updates_per_s": round(N * L * K * sweeps_per_s, 1),
        "sweep_hbm_frac": round(sweep_bytes(N, L, K, 2) * sweeps_per_s / 1e9 / PEAK_HBM_GBS, 5),
        "phase_ms_per_sweep": {p: round(v[0] / nt, 4) for p, v in timing.items()},
        "phase_timing_sweeps": nt,
        "tuning": ({n: getattr(tuning, n) for n, _ in sa.Tuning._fields_} if tuning else
                   "library defaults"),
        "lnP_last": like[0],
    }
    if not args.no_batch:
        print("bench: batched K-sweep (100 chains) ...", file=sys.stderr, flush=True)
        result["batched"] = batched_ksweep(d, N, L, args.seed, args.batch_sweeps, 10)
    if not args.no_cpu:
        print("bench: CPU baseline (oracle) ...", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(N, L, K, args.cpu_sample_loci, args.cpu_sweeps,
                                              args.seed)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
