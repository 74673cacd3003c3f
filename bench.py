#!/usr/bin/env python3
"""Benchmark: admixture-MCMC sweeps/s on one MI355X at N=2000, L=100000, K=5.

A step is one full sweep of the reference hot loop (structure.c:497-638) on
synthetic Balding-Nichols SNP data (K_true=5, F in [0.01, 0.1], Q ~ Dir(0.3),
1% missing) under the F model with alpha and lambda inference and
METROFREQ=10: UpdateP, UpdateQ (Gibbs, Metropolis every 10th sweep), the fused
UpdateZ, alpha, lambda, epsilon, Fst, and the post-burn-in data collection with
lnP(D).  Every sweep collects (BURNIN=0: the costlier half of a run); the W
warm-up sweeps (at least two) capture the HIP graphs of both sweep types, so
the K timed sweeps are replays only.  Genotypes, Q and P are resident in HBM
before timing starts.  Outside the timed region the line also carries the
lnP(D) error against the FP64 oracle and the CPU baseline (the oracle run as one
pinned chain per host core).

Prints one JSON line (see the keys below); rank-0 only, one GPU.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
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


CPU_SHARE = 16   # host cores one GPU's runs may use on the benchmark host
# Warm-up sweeps, at least.  With METROFREQ=10 the library replays graphs of 20
# sweeps (two Metropolis periods) and single-sweep graphs around them: 30 warm-up
# sweeps from rep 0 capture a 20-sweep graph (reps 0-19) and both single-sweep
# graphs (20-29), so the timed sweeps, from rep 30 on, capture nothing.
MIN_WARM = 30


def cpu_info():
    """(model name, cores visible to this process, cores the baseline uses)."""
    model = "unknown"
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.lower().startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    cpus = sorted(os.sched_getaffinity(0))
    return model, cpus, cpus[:CPU_SHARE]


def _save_sample(geno, nall):
    import numpy as np
    d = tempfile.mkdtemp(prefix="sgpu_cpu_")
    np.save(os.path.join(d, "geno.npy"), np.ascontiguousarray(geno, dtype=np.int8))
    np.save(os.path.join(d, "nall.npy"), np.ascontiguousarray(nall, dtype=np.int32))
    return d


def _run_pinned(jobs, cpus):
    """Run the oracle worker (oracle/cpu_bench.py) once per job, each process
    pinned to one core with taskset, at most one process per core at a time.
    jobs: list of (sample dir, worker arguments).  Returns the parsed lines."""
    worker = os.path.join(ROOT, "oracle", "cpu_bench.py")
    pending = list(enumerate(jobs))
    running = {}   # cpu -> (job index, Popen)
    results = [None] * len(jobs)
    free = list(cpus)
    while pending or running:
        while pending and free:
            cpu = free.pop(0)
            j, (d, extra) = pending.pop(0)
            cmd = ["taskset", "-c", str(cpu), sys.executable, worker, d] + [str(x) for x in extra]
            running[cpu] = (j, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                                stderr=subprocess.PIPE, text=True))
        time.sleep(0.05)
        for cpu in list(running):
            j, proc = running[cpu]
            if proc.poll() is None:
                continue
            out, err = proc.communicate()
            if proc.returncode != 0:
                raise RuntimeError(f"CPU baseline worker failed ({proc.returncode}): {err[-500:]}")
            results[j] = json.loads(out.strip().splitlines()[-1])
            del running[cpu]
            free.append(cpu)
    return results


def _rm_sample(d):
    for f in ("geno.npy", "nall.npy"):
        try:
            os.remove(os.path.join(d, f))
        except OSError:
            pass
    try:
        os.rmdir(d)
    except OSError:
        pass


def cpu_baseline_replicates(geno, nall, L_full, Ks, sweeps, seed, model_args, what):
    """The oracle (FP64 C restatement of the reference CPU sweep, -O3) run as one
    chain per core, pinned, on a locus sample of the workload: chain c takes
    K = Ks[c % len(Ks)] and seed + c.  Each chain's sweeps/s on the sample is
    scaled to the full locus count (every phase is linear in L at fixed N, K);
    the baseline is the sum over the cores (replicate chains are independent)."""
    model, visible, cpus = cpu_info()
    N, _, Ls = geno.shape
    d = _save_sample(geno, nall)
    try:
        jobs = [(d, ["--K", Ks[c % len(Ks)], "--sweeps", sweeps, "--seed", seed + c,
                     "--chain", c] + model_args) for c in range(len(cpus))]
        t0 = time.perf_counter()
        res = _run_pinned(jobs, cpus)
        wall = time.perf_counter() - t0
    finally:
        _rm_sample(d)
    rates = [r["sweeps"] / r["seconds"] * (Ls / L_full) for r in res]
    ks_used = sorted({Ks[c % len(Ks)] for c in range(len(cpus))})
    return {"value": sum(rates), "unit": "chain-sweeps/s", "cores": len(cpus), "kind": "port",
            "per_chain_sweeps_per_s": sum(rates) / len(rates),
            "cpu_model": model, "nproc": len(visible),
            "sample": (f"{what}: oracle (FP64 C restatement of the reference CPU sweep, -O3) as "
                       f"one chain per core on {len(cpus)} cores (taskset-pinned, K in "
                       f"{ks_used}), each on N={N}, L={Ls} loci of the same generator "
                       f"({Ls / L_full:.4g} of the loci), {sweeps} sweeps after one untimed "
                       f"sweep, rates scaled linearly to L={L_full} and summed over the cores "
                       f"({wall:.1f} s wall); cores capped at {CPU_SHARE}, the host share of "
                       f"one GPU")}


def batched_ksweep(d, N, L, seed, sweeps, warmup, kmax=10, reps=10):
    """The K-sweep users run (K = 1..kmax, reps Philox seeds each) on one data
    set: chain-sweeps/s with all chains in one batch (one launch per kernel and
    phase, HIP-graph replay) against the same chains run one at a time on the
    GPU (each alone in its own context, graph replay).  Every sweep collects
    (BURNIN=0), so the warm-up (at least two sweeps: a Metropolis and a Gibbs
    one) captures both sweep graphs that the timed sweeps replay, on both sides.
    Returns the line and, per K, the final (Q, P, lnP) of replicate 0 for the
    error check against the oracle."""
    import structure_amd as sa
    warmup = max(warmup, 10)
    models = [sa.make_model(N, L, K, 2, freqscorr=1, inferalpha=1, inferlambda=1, metrofreq=10,
                            burnin=0, numreps=sweeps, seed=seed + r, chain=K)
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
    finals = {}
    for ci, m in enumerate(models):
        if ci % reps < 3:   # three replicates per K for the error against the oracle
            finals[ci] = (m.K, b.get_chain(ci, sa.Q), b.get_chain(ci, sa.P),
                          b.get_chain(ci, sa.LIKE)[0])
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
            "warmup": warmup, "launch": "graph replay, graphs captured in the warm-up",
            "batched_chain_sweeps_per_s": round(n * sweeps / t_batch, 2),
            "plain_chain_sweeps_per_s": round(n * sweeps / t_plain, 2),
            "speedup": round(t_plain / t_batch, 3),
            "batched_site_updates_per_s": round(
                sum(N * L * m.K for m in models) * sweeps / t_batch, 1)}, finals


def lnp_error_per_k(words, N, L, finals):
    """Relative error of each chain's last lnP(D) (FP64 sum of the UpdateZ tile
    partials) against the FP64 oracle's CalcLike (CalcLike.c:387-426) on the same
    (Q, P), computed on the host from the same 2-bit words; mean and max per K."""
    import oracle
    per_k = {}
    for ci, (K, Q, P, like) in sorted(finals.items()):
        ref = oracle.calc_like_snp2(oracle.make_model(N, L, K, 2), words, Q, P)
        per_k.setdefault(K, []).append(abs(like - ref) / abs(ref))
    return {str(K): {"mean_rel_err": sum(v) / len(v), "max_rel_err": max(v), "chains": len(v)}
            for K, v in sorted(per_k.items())}


def _timed_run(ch, warmup, steps, direct=False):
    """Warm-up sweeps, then `steps` timed sweeps from graph replay (or, with
    `direct`, launched one kernel at a time); returns seconds.  The chains
    collect from sweep 0 (BURNIN=0); a warm-up of MIN_WARM sweeps captures every
    graph the timed sweeps replay, so no capture falls inside the timed region
    (shorter warm-ups, as the config-3 line's, start the timed region off a
    Metropolis sweep and replay only single-sweep graphs, both captured by then)."""
    if direct:
        ch.set_timing(True)   # timing mode launches every phase directly
    ch.init()
    ch.run(0, max(warmup, 2))
    ch.synchronize()
    t0 = time.perf_counter()
    ch.run(max(warmup, 2), steps)
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


def _unpack_snp2(words, N, Ls):
    """Genotypes [N][2][Ls] int8 (-9 missing) of the first Ls loci of 2-bit words
    [ceil(N/16)][Lpad] (code 0 = (0,0), 1 = (0,1), 2 = (1,1), 3 = missing)."""
    import numpy as np
    G = (N + 15) // 16
    w = words.reshape(G, -1)[:, :Ls]
    codes = np.stack([(w >> np.uint32(2 * t)) & np.uint32(3) for t in range(16)], axis=1)
    codes = codes.reshape(G * 16, Ls)[:N]
    geno = np.empty((N, 2, Ls), np.int8)
    geno[:, 0, :] = codes == 2
    geno[:, 1, :] = codes != 0
    mi, ml = np.nonzero(codes == 3)
    geno[mi, :, ml] = -9
    return geno


def workload_c3(args):
    """Large single chain: N=20000, L=500000 SNPs (2.5 GB of 2-bit genotypes),
    K=8, F model; Z never materialised."""
    import structure_amd as sa
    N, L, K = 20000, 500000, 8
    t0 = time.perf_counter()
    d = sa.synth_snp(N, L, Ktrue=8, flo=0.005, fhi=0.05, qalpha=0.2, miss=0.01, seed=args.seed)
    print(f"bench: synthetic data N={N} L={L} in {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)
    warm, steps = max(args.warmup, 2), args.steps
    model = sa.make_model(N, L, K, 2, freqscorr=1, inferalpha=1, inferlambda=1, metrofreq=10,
                          burnin=0, numreps=steps, seed=args.seed, chain=0)
    ch = sa.Chain(model, sa.GENO_SNP2, d["words"], d["nall"])
    dt = _timed_run(ch, warm, steps)
    zb = z_bytes_per_launch(N, L, K, 2)
    timing, roof = _z_roofline(ch, warm + steps, min(steps, 5), zb)
    Q, P, like = ch.get(sa.Q), ch.get(sa.P), ch.get(sa.LIKE)[0]
    ch.close()
    sps = steps / dt
    res = {"metric": "sweeps/s at N=20k,L=500k,K=8 (%HBM peak)", "value": round(sps, 3),
           "unit": "sweeps/s", "n_gpus": 1, "steps": steps, "warmup": warm,
           "ms_per_step": round(1e3 * dt / steps, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": f"synthetic (Balding-Nichols SNP, K_true=8, seed {args.seed})",
           "config": {"workload": "config 3: N=20000 L=500000 K=8 biallelic 2-bit, F model + "
                                  "alpha + lambda, METROFREQ=10, collecting sweeps",
                      "N": N, "L": L, "K": K, "genotype_format": "SNP2"},
           "roofline": roof, "site_updates_per_s": round(N * L * K * sps, 1),
           "phase_ms_per_sweep": {p: round(v[0] / max(min(steps, 5), 1), 4)
                                  for p, v in timing.items()}}
    if not args.no_check:
        import oracle
        ref = oracle.calc_like_snp2(oracle.make_model(N, L, K, 2), d["words"], Q, P)
        res["lnP_rel_err_vs_fp64_oracle"] = abs(like - ref) / abs(ref)
    if not args.no_cpu:
        print("bench: CPU baseline (oracle, one chain per core) ...", file=sys.stderr, flush=True)
        res["cpu_baseline"] = cpu_baseline_replicates(
            _unpack_snp2(d["words"], N, 1000), d["nall"][:1000], L, [K], 3, args.seed,
            ["--freqscorr", 1, "--inferlambda", 1], "config 3")
    return res


def _ragged_datasets(seed, n=64, keep_geno=False):
    """Config 5: n independent data sets, N ~ U[100,5000], L log-uniform in
    [1e3, 2e5], half SNP (Balding-Nichols) and half microsatellite (2-20
    alleles), K_true and run K in [2, 8], Q ~ Dir(U[0.1,1]), missing U[0,5%].
    Returns the specs and (format, words, NumAlleles, A, genotypes or None)."""
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
            return sa.GENO_SNP2, d["words"], d["nall"], 2, None
        d = sa.synth_microsat(sp["N"], sp["L"], amin=2, amax=20, Ktrue=sp["ktrue"],
                              qalpha=sp["qalpha"], miss=sp["miss"], seed=sp["seed"])
        fmt, words = sa.pack(d["geno"], sa.GENO_BYTE)
        return fmt, words, d["nall"], d["A"], (d["geno"] if keep_geno else None)

    # the generators are multithreaded themselves; two at a time overlaps packing
    with cf.ThreadPoolExecutor(max_workers=2) as ex:
        data = list(ex.map(make, specs))
    return specs, data


def _r3(x):
    return None if x is None else round(x, 3)


def _batch_vs_plain(models, datasets, doc, warm, sweeps, plain_sweeps=None, direct=False,
                    keep=False):
    """Chain-sweeps/s of all chains in one batch against the same chains each
    run alone (one after the other) on the GPU.  With `direct` (for kernel
    tracers) the batch's sweeps are launched without graphs and the chains are
    not run alone.  With `keep`, also each chain's final (Q, P, lnP)."""
    import structure_amd as sa
    b = sa.Batch(models, datasets, doc)
    t_batch = _timed_run(b, warm, sweeps, direct)
    finals = None
    if keep:
        finals = [(b.get_chain(c, sa.Q), b.get_chain(c, sa.P), b.get_chain(c, sa.LIKE)[0])
                  for c in range(len(models))]
    # per-phase device time of a batch sweep (events around each phase, direct launches)
    nt = min(sweeps, 20)
    b.set_timing(True)
    b.get_timing()
    b.run(max(warm, 2) + sweeps, nt)
    tm = b.get_timing()
    b.set_timing(False)
    phases = {p: round(v[0] / nt, 4) for p, v in tm.items()}
    b.close()
    n = len(models)
    if direct:
        return n * sweeps / t_batch, None, phases, finals
    ps = plain_sweeps or sweeps
    t_plain = 0.0
    for m, j in zip(models, doc):
        fmt, words, nall = datasets[j]
        ch = sa.Chain(m, fmt, words, nall)
        t_plain += _timed_run(ch, warm, ps)
        ch.close()
    return n * sweeps / t_batch, n * ps / t_plain, phases, finals


def _c5_errors(specs, data, finals):
    """Per data set: relative error of the last lnP(D) against the FP64 oracle's
    CalcLike on the same (Q, P), host threads over the data sets."""
    import concurrent.futures as cf

    import oracle

    def one(j):
        sp, (fmt, words, nall, A, geno) = specs[j], data[j]
        Q, P, like = finals[j]
        m = oracle.make_model(sp["N"], sp["L"], sp["K"], A)
        if geno is None:
            ref = oracle.calc_like_snp2(m, words, Q, P, threads=1)
        else:
            ref = oracle.Chain(m, geno, nall).calc_like(Q, P)
        return abs(like - ref) / abs(ref)

    with cf.ThreadPoolExecutor(max_workers=min(CPU_SHARE, len(os.sched_getaffinity(0)))) as ex:
        return list(ex.map(one, range(len(specs))))


def _c5_cpu_baseline(specs, data, seed):
    """The oracle per data set on one pinned core each (a pool over the cores):
    each chain's sweep time on its first 1000 loci scaled to its full L; the
    batch of 64 chains then takes sum(t) / cores per sweep on the pool."""
    model, visible, cpus = cpu_info()
    Ls_max = 1000
    dirs, jobs = [], []
    try:
        for j, (sp, (fmt, words, nall, A, geno)) in enumerate(zip(specs, data)):
            Ls = min(sp["L"], Ls_max)
            g = _unpack_snp2(words, sp["N"], Ls) if geno is None else geno[:, :, :Ls]
            d = _save_sample(g, nall[:Ls])
            dirs.append(d)
            jobs.append((d, ["--K", sp["K"], "--sweeps", 10, "--seed", seed + j, "--chain", j,
                             "--freqscorr", 1, "--inferlambda", 0]))
        t0 = time.perf_counter()
        res = _run_pinned(jobs, cpus)
        wall = time.perf_counter() - t0
    finally:
        for d in dirs:
            _rm_sample(d)
    t_sweep = [r["seconds"] / r["sweeps"] * sp["L"] / min(sp["L"], Ls_max)
               for r, sp in zip(res, specs)]
    per_batch_sweep = sum(t_sweep) / len(cpus)
    return {"value": len(specs) / per_batch_sweep, "unit": "chain-sweeps/s", "cores": len(cpus),
            "kind": "port", "cpu_model": model, "nproc": len(visible),
            "sample": (f"config 5: oracle (FP64 C restatement, -O3), each data set's chain on "
                       f"one taskset-pinned core (pool of {len(cpus)} cores), 10 sweeps on its "
                       f"first min(L, {Ls_max}) loci after one untimed sweep, sweep time scaled "
                       f"linearly to its L; 64 chain-sweeps per (sum of the chains' sweep times "
                       f"/ cores) ({wall:.1f} s wall)")}


def workload_c5(args):
    """64 ragged data sets, one chain each, all in one batch."""
    import structure_amd as sa
    t0 = time.perf_counter()
    specs, data = _ragged_datasets(args.seed, keep_geno=not args.no_check or not args.no_cpu)
    print(f"bench: 64 ragged data sets in {time.perf_counter() - t0:.1f}s", file=sys.stderr,
          flush=True)
    warm, sweeps = max(args.warmup, MIN_WARM), args.steps
    models = [sa.make_model(sp["N"], sp["L"], sp["K"], dd[3], freqscorr=1, inferalpha=1,
                            metrofreq=10, burnin=0, numreps=sweeps, seed=args.seed + j, chain=j)
              for j, (sp, dd) in enumerate(zip(specs, data))]
    dsets = [(f, w, na) for f, w, na, _, _ in data]
    batched, plain, phases, finals = _batch_vs_plain(
        models, dsets, list(range(len(models))), warm, sweeps, direct=args.direct,
        keep=not args.no_check)
    sites = sum(sp["N"] * sp["L"] * sp["K"] for sp in specs)
    res = {"metric": "chain-sweeps/s, 64 ragged data sets in one batch",
           "value": round(batched, 3), "unit": "chain-sweeps/s", "n_gpus": 1, "steps": sweeps,
           "warmup": warm, "warmup_requested": args.warmup, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": f"synthetic ragged (seed {args.seed})",
           "config": {"workload": "config 5: 64 data sets, N~U[100,5000], L log-uniform "
                                  "[1e3,2e5], half 2-bit SNP / half byte microsatellite "
                                  "(2-20 alleles), K in [2,8], F model + alpha",
                      "datasets": len(specs),
                      "total_sites": sum(sp["N"] * sp["L"] for sp in specs)},
           "plain_chain_sweeps_per_s": _r3(plain), "speedup": _r3(plain and batched / plain),
           "site_updates_per_s": round(sites * batched / len(specs), 1),
           "phase_ms_per_batch_sweep": phases,
           "launch": "direct" if args.direct else "graph replay"}
    if finals is not None:
        print("bench: lnP(D) per data set against the FP64 oracle ...", file=sys.stderr,
              flush=True)
        errs = _c5_errors(specs, data, finals)
        res["lnP_rel_err_vs_fp64_oracle"] = {
            "per_dataset": [float(f"{e:.3e}") for e in errs], "max": max(errs),
            "mean": sum(errs) / len(errs)}
    if not args.no_cpu:
        print("bench: CPU baseline (oracle, pinned pool) ...", file=sys.stderr, flush=True)
        res["cpu_baseline"] = _c5_cpu_baseline(specs, data, args.seed)
    return res


def workload_c1batch(args):
    """Config 1 replicated: 100 chains of N=200, L=500 microsatellites, K=3,
    no F model, alpha inferred; batched against one at a time."""
    import structure_amd as sa
    d = sa.synth_microsat(200, 500, amin=2, amax=10, Ktrue=3, qalpha=0.5, miss=0.02,
                          seed=args.seed)
    fmt, words = sa.pack(d["geno"])
    A = d["A"]
    warm, sweeps = max(args.warmup, MIN_WARM), args.steps
    models = [sa.make_model(200, 500, 3, A, freqscorr=0, inferalpha=1, lambda0=1.0,
                            metrofreq=10, burnin=0, numreps=sweeps, seed=args.seed + r,
                            chain=3) for r in range(100)]
    batched, plain, phases, _ = _batch_vs_plain(models, [(fmt, words, d["nall"])], [0] * 100,
                                                warm, sweeps, plain_sweeps=max(sweeps // 10, 20),
                                                direct=args.direct)
    res = {"metric": "chain-sweeps/s, 100 replicate chains of config 1 in one batch",
           "value": round(batched, 3), "unit": "chain-sweeps/s", "n_gpus": 1, "steps": sweeps,
           "warmup": warm, "warmup_requested": args.warmup, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": f"synthetic microsatellites (seed {args.seed})",
           "config": {"workload": "config 1 x 100 seeds: N=200 L=500 A in [2,10] K=3, "
                                  "no F model, alpha inferred", "chains": 100},
           "plain_chain_sweeps_per_s": _r3(plain), "speedup": _r3(plain and batched / plain),
           "phase_ms_per_batch_sweep": phases,
           "launch": "direct" if args.direct else "graph replay"}
    if not args.no_cpu:
        print("bench: CPU baseline (oracle, one chain per core) ...", file=sys.stderr, flush=True)
        res["cpu_baseline"] = cpu_baseline_replicates(
            d["geno"], d["nall"], 500, [3], 2000, args.seed,
            ["--freqscorr", 0, "--inferlambda", 0], "config 1 replicates (full data)")
    return res


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
    ap.add_argument("--no-check", action="store_true",
                    help="skip the lnP(D) error columns against the FP64 oracle")
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
    # every sweep collects (BURNIN=0): the warm-up, at least MIN_WARM sweeps,
    # captures every graph that the timed sweeps replay; right after start-up the
    # first 20 sweeps also run 6% slower than the next (clocks and caches warming),
    # so fewer would time the ramp.
    warm = max(args.warmup, MIN_WARM)
    model = sa.make_model(N, L, K, 2, freqscorr=1, inferalpha=1, inferlambda=1, metrofreq=10,
                          burnin=0, numreps=args.steps, seed=args.seed, chain=0)
    tuning = None
    if args.tune:
        from structure_amd import tune
        tuning, _ = tune.autotune([model], [(sa.GENO_SNP2, d["words"], d["nall"])], cache=None,
                                  log=lambda s: print("bench: " + s, file=sys.stderr, flush=True))
    ch = sa.Batch([model], [(sa.GENO_SNP2, d["words"], d["nall"])], tuning=tuning)
    ch.init()
    ch.run(0, warm)
    ch.synchronize()
    print(f"bench: {warm} warm-up sweeps done", file=sys.stderr, flush=True)

    # timed region: K sweeps replayed from HIP graphs captured in the warm-up
    ch.synchronize()
    t0 = time.perf_counter()
    ch.run(warm, args.steps)
    ch.synchronize()
    dt = time.perf_counter() - t0
    like = ch.get_chain(0, sa.LIKE)
    Qf, Pf = ch.get_chain(0, sa.Q), ch.get_chain(0, sa.P)
    # per-phase kernel durations: the same sweeps launched directly, with HIP
    # events around every phase on the chain's stream
    nt = min(args.steps, 100)
    ch.set_timing(True)
    ch.get_timing()   # reset
    ch.run(warm + args.steps, nt)
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
        "warmup": warm,
        "warmup_requested": args.warmup,
        "ms_per_step": round(1e3 * dt / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": f"synthetic (Balding-Nichols SNP, seed {args.seed}); random initial state",
        "config": {"workload": (f"config 2: N={N} L={L} K={K} biallelic 2-bit, F model + "
                                f"alpha + lambda, METROFREQ=10, collecting sweeps with lnP(D)"),
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
    if not args.no_check:
        import oracle
        ref = oracle.calc_like_snp2(oracle.make_model(N, L, K, 2), d["words"], Qf, Pf)
        result["lnP_rel_err_vs_fp64_oracle"] = abs(like[0] - ref) / abs(ref)
    if not args.no_batch:
        print("bench: batched K-sweep (100 chains) ...", file=sys.stderr, flush=True)
        result["batched"], finals = batched_ksweep(d, N, L, args.seed, args.batch_sweeps,
                                                   args.warmup)
        if not args.no_check:
            print("bench: lnP(D) per K against the FP64 oracle ...", file=sys.stderr, flush=True)
            result["batched"]["lnP_rel_err_vs_fp64_oracle_per_K"] = lnp_error_per_k(
                d["words"], N, L, finals)
    if not args.no_cpu:
        print("bench: CPU baseline (oracle, one chain per core) ...", file=sys.stderr, flush=True)
        ds = sa.synth_snp(N, args.cpu_sample_loci, Ktrue=5, flo=0.01, fhi=0.1, qalpha=0.3,
                          miss=0.01, seed=args.seed, want_geno=True)
        margs = ["--freqscorr", 1, "--inferlambda", 1]
        result["cpu_baseline"] = cpu_baseline_replicates(
            ds["geno"], ds["nall"], L, [K], args.cpu_sweeps, args.seed, margs,
            f"config 2 (K={K} replicate chains)")
        if "batched" in result:
            result["batched"]["cpu_baseline"] = cpu_baseline_replicates(
                ds["geno"], ds["nall"], L, list(range(1, 11)), args.cpu_sweeps, args.seed, margs,
                "config 4 (K-sweep chains, K = 1 + core % 10)")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
