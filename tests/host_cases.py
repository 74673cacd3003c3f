"""Inputs for the host-file tests (parameter files, genotype files and result
states), shared by tests/test_host_files.py and tests/golden/make_golden_host.py."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TESTDATA1 = os.path.join(HERE, "golden", "host", "testdata1.str")

MAIN_T1 = """Parameter file with commentary around the definitions; a stray # here
is not a definition, nor is #defined or # define.
#define MAXPOPS    3      // (int) number of populations assumed
#define BURNIN    500   // (int) length of burnin period
#define NUMREPS   1000
#define INFILE   {infile}   // (str) name with spaces
#define OUTFILE  {outfile}  //(str)
#define NUMINDS    20
#define NUMLOCI    7
#define PLOIDY       2
#define MISSING     -999
#define ONEROWPERIND 0
#define LABEL     1
#define POPDATA   1
#define POPFLAG   0
#define LOCDATA   0
#define PHENOTYPE 0
#define EXTRACOLS 2
#define MARKERNAMES      1
#define RECESSIVEALLELES 0
#define MAPDISTANCES     0
#define NOSUCHKEY 5
#define SEED 0x10
#define ALPHA 0.5
#define LAMBDA 0.75
#define FREQSCORR 0
#define ANCESTPINT .8
#define UPDATEFREQ 0
"""

EXTRA_T1 = """#define METROFREQ 7
#define INFERLAMBDA 1
#define POPSPECIFICLAMBDA 1
#define FPRIORSD 0.025
#define LOCISPOP 010
#define ALPHAPROPSD   0.1
#define PRINTQHAT 1
#define NOQS 3
"""


def write_param_files(d, infile="my data/file one.str", outfile="out/run1"):
    m = os.path.join(d, "mainparams_t1")
    e = os.path.join(d, "extraparams_t1")
    with open(m, "w") as f:
        f.write(MAIN_T1.format(infile=infile, outfile=outfile))
    with open(e, "w") as f:
        f.write(EXTRA_T1)
    return m, e


def write_genotype_file(path, N, L, seed, onerow=False, missing=-9, extracols=2,
                        markernames=True, popflag=True):
    """A data file with labels, population, pop flag, extra columns and marker
    names; alleles are arbitrary integers (e.g. fragment lengths)."""
    rng = np.random.default_rng(seed)
    alle = [rng.choice([97, 101, 103, 250, 7, -1, 0, 12], size=rng.integers(1, 6), replace=False)
            for _ in range(L)]
    lines = []
    if markernames:
        lines.append(" ".join(f"loc_{j}" if j % 3 else f"marker{j}xyz" for j in range(L)))
    geno = np.zeros((N, 2, L), np.int64)
    for i in range(N):
        for c in range(2):
            for j in range(L):
                geno[i, c, j] = missing if rng.random() < 0.1 else rng.choice(alle[j])
    for i in range(N):
        lead = [f"ind{i:03d}", str(int(rng.integers(1, 4)))]
        if popflag:
            lead.append(str(int(rng.integers(0, 2))))
        lead += [str(int(rng.integers(0, 50))) for _ in range(extracols)]
        if onerow:
            vals = [str(geno[i, c, j]) for j in range(L) for c in range(2)]
            lines.append("\t".join(lead + vals))
        else:
            for c in range(2):
                lines.append(" ".join(lead + [str(v) for v in geno[i, c]]))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def input_params(datafile, N, L, onerow=False, missing=-9, extracols=2, markernames=True,
                 popflag=True):
    return dict(datafile=datafile, numinds=N, numloci=L, missing=missing, label=1, popdata=1,
                popflag=int(popflag), phenotype=0, extracols=extracols,
                markernames=int(markernames), mapdistances=0, onerowperind=int(onerow),
                recessivealleles=0, ploidy=2, phaseinfo=0, locdata=0, locispop=0)


TESTDATA1_PARAMS = dict(numinds=200, numloci=5, missing=-999, label=1, popdata=1, popflag=1,
                        phenotype=0, extracols=0, markernames=0, mapdistances=0,
                        onerowperind=0, recessivealleles=0, ploidy=2, phaseinfo=0, locdata=0,
                        locispop=0)

OUTPUT_VARIANTS = [
    dict(maxpops=3, popdata=1, freqscorr=1, onefst=0, popalphas=0, inferlambda=0,
         printqhat=1),
    dict(maxpops=2, popdata=0, freqscorr=0, onefst=0, popalphas=1, inferlambda=1,
         popspecificlambda=0, printqhat=0),
    dict(maxpops=5, popdata=1, freqscorr=1, onefst=1, popalphas=0, inferlambda=1,
         popspecificlambda=1, printqhat=1, randomize=0, computeprob=1),
    dict(maxpops=1, popdata=1, freqscorr=1, onefst=0, popalphas=0, inferlambda=0,
         printqhat=0, computeprob=0),
    # ANCESTDIST probability intervals (PrintAncestDist, output.c:762-797) and the
    # PFROMPOPFLAGONLY line of PrintSums (output.c:1115-1117)
    dict(maxpops=3, popdata=1, freqscorr=1, onefst=0, popalphas=0, inferlambda=0,
         printqhat=1, ancestdist=1, numboxes=50, ancestpint=0.8),
    dict(maxpops=4, popdata=0, freqscorr=0, onefst=0, popalphas=1, inferlambda=1,
         printqhat=0, ancestdist=1, numboxes=1000, ancestpint=0.95, pfrompopflagonly=1),
]


def result_state(N, L, K, A, nall, seed, nrep=37, numboxes=None):
    """Posterior sums of nrep collected sweeps with plausible values; with
    numboxes, also the ANCESTDIST histogram of the same Q draws."""
    rng = np.random.default_rng(seed)
    qs = rng.dirichlet(np.full(K, 0.7), size=(nrep, N))
    q = qs.sum(axis=0)
    P = np.zeros((L, K, A))
    E = np.zeros((L, A))
    for l in range(L):
        if nall[l]:
            P[l, :, :nall[l]] = rng.dirichlet(np.ones(nall[l]), size=(nrep, K)).sum(axis=0)
            E[l, :nall[l]] = rng.dirichlet(np.ones(nall[l]), size=nrep).sum(axis=0)
    likes = -1000.0 * (1 + rng.random(nrep))
    out = dict(qsum=q, psum=P, epssum=E, fstsum=rng.uniform(0, 0.3, K) * nrep,
               alphasum=rng.uniform(0.2, 2, K) * nrep, lambdasum=rng.uniform(0.5, 2, K) * nrep,
               sumlikes=float(likes.sum()), sumsqlikes=float((likes ** 2).sum()))
    if numboxes:
        box = np.minimum((qs * numboxes).astype(np.int64), numboxes - 1)
        ad = np.zeros((N, K, numboxes), np.int32)
        for r in range(nrep):
            for i in range(N):
                for k in range(K):
                    ad[i, k, box[r, i, k]] += 1
        out["ancestdist"] = ad
    return out
