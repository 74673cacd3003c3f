"""Chain-level comparison cases: data, model options and the statistics
compared across seeds (posterior-mean Q after label alignment, estimated
ln P(D), mean alpha / F / lambda)."""
import numpy as np

from structure_amd import host

SEEDS = list(range(1, 9))

CASES = {
    # config-1 shape scaled down: microsatellites, admixture without the F model
    "microsat": dict(N=100, L=60, K=3, data=dict(kind="microsat", amin=2, amax=8, Ktrue=3,
                                                   qalpha=0.5, miss=0.02, seed=5),
                     opts=dict(freqscorr=0, inferalpha=1, inferlambda=0, metrofreq=10,
                               alpha=1.0, lambda_=1.0),
                     burnin=300, numreps=700),
    # config-2 options scaled down: SNPs, F model, alpha and lambda inferred
    "snp_fmodel": dict(N=100, L=200, K=3, data=dict(kind="snp", Ktrue=3, flo=0.02, fhi=0.1,
                                                     qalpha=0.4, miss=0.01, seed=6),
                       opts=dict(freqscorr=1, inferalpha=1, inferlambda=1, metrofreq=10,
                                 alpha=1.0, lambda_=1.0, fpriormean=0.01, fpriorsd=0.05),
                       burnin=300, numreps=700),
}


def make_data(case):
    import structure_amd as sa
    c = CASES[case]
    d = dict(c["data"])
    kind = d.pop("kind")
    if kind == "microsat":
        r = sa.synth_microsat(c["N"], c["L"], **d)
        return r["geno"], r["nall"], r["A"]
    r = sa.synth_snp(c["N"], c["L"], want_geno=True, **d)
    return r["geno"], r["nall"], 2


def compare(ref_stats, got_stats, K):
    """Chain-level agreement: returns a dict of the largest standardised
    differences (|mean difference| / combined standard error)."""
    qr, qg = np.asarray(ref_stats["qbar"]), np.asarray(got_stats["qbar"])
    R = qr.shape[0]
    anchor = qr[0]
    aligned = []
    for q in list(qr) + list(qg):
        perm = host.align_q(anchor, q)
        aligned.append(q[:, perm])
    aligned = np.array(aligned)
    ar, ag = aligned[:R], aligned[R:]
    se = np.sqrt(ar.var(axis=0, ddof=1) / R + ag.var(axis=0, ddof=1) / ag.shape[0])
    dq = np.abs(ar.mean(axis=0) - ag.mean(axis=0))
    out = dict(q_max_abs=float(dq.max()),
               q_frac_outside=float(np.mean(dq > 4 * se + 0.01)))
    for key in ("lnp", "alpha", "fst", "lam"):
        if key not in ref_stats:
            continue
        a = np.asarray(ref_stats[key], dtype=float).reshape(R, -1)
        b = np.asarray(got_stats[key], dtype=float).reshape(len(got_stats[key]), -1)
        if key != "lnp":   # per-cluster means are label dependent: compare their sums
            a, b = a.sum(axis=1, keepdims=True), b.sum(axis=1, keepdims=True)
        s = np.sqrt(a.var(axis=0, ddof=1) / len(a) + b.var(axis=0, ddof=1) / len(b))
        d = np.abs(a.mean(axis=0) - b.mean(axis=0))
        out[key + "_z"] = float(np.max(d / np.maximum(s, 1e-12 * np.abs(a.mean(axis=0)) + 1e-300)))
    return out
