"""Write tests/golden/ref_vectors.npz from the reference's own C code.

Runs only where the reference sources were built into oracle/_ref (see
oracle/Makefile).  Every array in the file is either an input or the output
the reference produced for it; tests/test_golden.py checks the oracle against
them without needing the reference.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from oracle import ref  # noqa: E402
from helpers import geno_ref_layout, random_geno, random_p, random_q, z_uniforms  # noqa: E402

TWO31 = 2.0 ** 31


def main():
    if not ref.available():
        sys.exit("oracle/_ref/libstructref.so is not built")
    L_ = ref.lib()
    rng = np.random.default_rng(20240521)
    out = {}

    # --- UpdateZ + counts + CalcLike + CalcLikeIndDiff (UpdateZ.c, UpdateP.c, UpdateQ.c, CalcLike.c)
    N, L, K, A = 12, 17, 3, 4
    geno, nall = random_geno(rng, N, L, A, miss=0.08)
    Q = random_q(rng, N, K, a=0.6)
    P = random_p(rng, L, K, A, nall, a=0.7)
    seed, chain, rep = 777, 2, 5
    ref.configure(N, L, K, A)
    g32 = geno_ref_layout(geno)
    ra = z_uniforms(seed, chain, rep, N, L)
    Z = np.zeros((N, 2, L), np.int32)
    L_.UpdateZ(ref.ip(Z), ref.dp(Q), ref.dp(P), ref.ip(g32), ref.dp(ra))
    nA = np.zeros((L, K, A), np.int32)
    L_.GetNumFromPops(ref.ip(nA), ref.ip(g32), ref.ip(Z), ref.ip(nall), ref.individuals(N))
    nQ = np.zeros((N, K), np.int32)
    for i in range(N):
        row = np.zeros(K, np.int32)
        L_.GetNumLociPop(ref.ip(row), ref.ip(Z), i)
        nQ[i] = row
    like = L_.CalcLike(ref.ip(g32), None, ref.dp(Q), ref.dp(P), None, None, None)
    Q2 = random_q(rng, N, K, a=0.6)
    diffs = np.array([L_.CalcLikeIndDiff(ref.ip(g32), None, ref.dp(np.ascontiguousarray(Q2[i])),
                                         ref.dp(np.ascontiguousarray(Q[i])), ref.dp(P), i, None)
                      for i in range(N)])
    out.update(z_geno=geno, z_nall=nall, z_Q=Q, z_P=P, z_Q2=Q2,
               z_key=np.array([seed, chain, rep], np.uint32), z_Z=Z, z_nA=nA, z_nQ=nQ,
               z_like=np.array(like), z_diffs=diffs)

    # --- mylgamma (mymath.c:382-429)
    zs = np.concatenate([rng.uniform(1e-3, 1, 20), rng.uniform(1, 30, 20),
                         10 ** rng.uniform(1, 6, 20)])
    out["lg_z"] = zs
    out["lg_v"] = np.array([L_.mylgamma(float(z)) for z in zs])

    # --- LogProbQ, LogProbQonepop, AlphaPriorDiff (UpdateAlpha.c:12-186)
    Nq, Kq = 15, 4
    ref.configure(Nq, 3, Kq, 2, ALPHAPRIORA=1.7, ALPHAPRIORB=2.5)
    Qa = random_q(rng, Nq, Kq, a=0.4)
    Qa[0, 1] = 1e-70
    alphas = rng.uniform(0.05, 4, 6)
    inds = ref.individuals(Nq)
    out.update(lp_Q=Qa, lp_alphas=alphas,
               lp_logprobq=np.array([L_.LogProbQ(ref.dp(Qa), a, inds) for a in alphas]),
               lp_onepop=np.array([[L_.LogProbQonepop(ref.dp(Qa), a, 3.3, inds, p)
                                    for p in range(Kq)] for a in alphas]),
               lp_prior=np.array([L_.AlphaPriorDiff(a, 0.9) for a in alphas]))

    # --- FlikeFreqs, FPriorDiff (UpdateFst.c:10-51)
    Nf, Lf, Kf, Af = 10, 20, 3, 4
    _, nallf = random_geno(rng, Nf, Lf, Af)
    nallf[4] = 0
    ref.configure(Nf, Lf, Kf, Af, FPRIORMEAN=0.05, FPRIORSD=0.07)
    Pf = random_p(rng, Lf, Kf, Af, nallf)
    epsf = np.zeros((Lf, Af))
    for l in range(Lf):
        if nallf[l]:
            epsf[l, :nallf[l]] = rng.dirichlet(np.ones(nallf[l]))
    fs = rng.uniform(0.01, 0.9, 5)
    out.update(fl_nall=nallf, fl_P=Pf, fl_eps=epsf, fl_f=fs,
               fl_v=np.array([[L_.FlikeFreqs(f, ref.dp(epsf), ref.dp(Pf), ref.ip(nallf), p)
                               for p in range(Kf)] for f in fs]),
               fl_prior=np.array([L_.FPriorDiff(f, 0.04) for f in fs]))

    # --- full Metropolis updates with a known rand() queue
    def queue(n):
        return rng.integers(1, 2 ** 31, size=n).astype(np.int32)

    # UpdateAlpha, single alpha (UpdateAlpha.c:189-253)
    ref.configure(Nq, 3, Kq, 2, ALPHAPROPSD=0.4, ALPHAMAX=3.0)
    al_in, al_out, al_z, al_u = [], [], [], []
    for t in range(12):
        a0 = np.full(Kq, rng.uniform(0.1, 2.5))
        q = queue(100)
        ref.set_rand(q)
        z = L_.snorm()
        pos = L_.sref_rand_consumed()
        u = q[pos] / TWO31
        a = a0.copy()
        ref.set_rand(q)
        L_.UpdateAlpha(ref.dp(Qa), ref.dp(a), inds, 1)
        al_in.append(a0)
        al_out.append(a)
        al_z.append(z)
        al_u.append(u)
    out.update(al_in=np.array(al_in), al_out=np.array(al_out), al_z=np.array(al_z),
               al_u=np.array(al_u))

    # NonIndependenceUpdateEpsilon (UpdateEpsilon.c:71-145)
    ref.configure(Nf, Lf, Kf, Af)
    Ff = rng.uniform(0.02, 0.3, Kf)
    q = queue(5 * Lf)
    e = epsf.copy()
    ref.set_rand(q)
    L_.NonIndependenceUpdateEpsilon(ref.dp(Pf), ref.dp(e), ref.dp(Ff), ref.ip(nallf), 0.8)
    out.update(ep_F=Ff, ep_queue=q, ep_out=e, ep_consumed=np.array(L_.sref_rand_consumed()))

    # InitFreqPriors (Init.c:172-240)
    ref.configure(N, L, K, A, FREQSCORR=1, LAMBDA=0.9, FPRIORMEAN=0.02)
    ei = np.zeros((L, A))
    Fi = np.zeros(K)
    L_.InitFreqPriors(ref.dp(ei), ref.dp(Fi), ref.ip(g32), ref.ip(nall))
    out.update(ifp_eps=ei, ifp_F=Fi)

    path = os.path.join(HERE, "ref_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, sum(v.nbytes for v in out.values()), "bytes")


if __name__ == "__main__":
    main()
