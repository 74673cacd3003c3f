"""The `structure` command end to end on the GPU (structure_amd/bin/structure):
the reference's parameter files and data file in, the reference's result file
out.  The run is checked against the oracle chain with the same key, written
through the same result writer, and a -B K-sweep batch against single runs."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from host_cases import TESTDATA1
from structure_amd import host

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "structure_amd", "bin", "structure")

MAIN = """#define MAXPOPS    2
#define BURNIN    40
#define NUMREPS   60
#define INFILE   {infile}
#define OUTFILE  {outfile}
#define NUMINDS    200
#define NUMLOCI    5
#define PLOIDY       2
#define MISSING     -999
#define ONEROWPERIND 0
#define LABEL     1
#define POPDATA   1
#define POPFLAG   1
#define LOCDATA   0
#define PHENOTYPE 0
#define EXTRACOLS 0
#define MARKERNAMES      0
#define RECESSIVEALLELES 0
#define MAPDISTANCES     0
"""

EXTRA = """#define FREQSCORR   1
#define INFERALPHA  1
#define ALPHA     1.0
#define INFERLAMBDA 1
#define LAMBDA    1.0
#define FPRIORMEAN 0.01
#define FPRIORSD   0.05
#define ALPHAPROPSD 0.025
#define METROFREQ 10
#define UPDATEFREQ 20
#define RANDOMIZE 0
#define SEED 4242
#define PRINTQHAT 1
"""


def _files(tmp_path, outfile="run"):
    (tmp_path / "mainparams").write_text(MAIN.format(infile=TESTDATA1, outfile=outfile))
    (tmp_path / "extraparams").write_text(EXTRA)


def _run(tmp_path, *args):
    print("structure", *args, flush=True)
    r = subprocess.run([BIN, *args], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    return r.stdout


def _numbers(text):
    return [float(x) for x in re.findall(r"-?\d+\.\d+", text)]


def _oracle_file(tmp_path, K, seed, name):
    """The oracle chain with the command's key, written by the same writer."""
    p = host.get_params(["structure", "-m", str(tmp_path / "mainparams"), "-e",
                         str(tmp_path / "extraparams"), "-K", str(K)])
    p.seed = seed
    inp = host.InputFile(p)
    m = oracle.make_model(inp.N, inp.L, K, inp.A, burnin=p.burnin, numreps=p.numreps,
                          freqscorr=p.freqscorr, onefst=p.onefst, inferalpha=p.inferalpha,
                          popalphas=p.popalphas, unifprioralpha=p.unifprioralpha,
                          inferlambda=p.inferlambda, popspecificlambda=p.popspecificlambda,
                          metrofreq=p.metrofreq, computeprob=p.computeprob, alpha0=p.alpha,
                          lambda0=p.lambda_, alphapropsd=p.alphapropsd, alphamax=p.alphamax,
                          alphapriora=p.alphapriora, alphapriorb=p.alphapriorb,
                          fpriormean=p.fpriormean, fpriorsd=p.fpriorsd, seed=seed, chain=K)
    ch = oracle.Chain(m, inp.geno16(), inp.nall)
    ch.init()
    total = p.burnin + p.numreps
    for rep in range(total):
        ch.sweep(rep)
    p.outfile = str(tmp_path / name).encode()
    inp.write_results(total, ch.QSum, ch.PSum, ch.epsSum, ch.FstSum, ch.alphaSum, ch.lambdaSum,
                      ch.sumlikes, ch.sumsqlikes, final=2, params=p)
    return (tmp_path / (name + "_f")).read_text()


def test_structure_command_matches_oracle_chain(tmp_path):
    _files(tmp_path)
    out = _run(tmp_path)
    assert "Final results printed to file run_f" in out
    assert "Structure seed: 4242" in out
    got = (tmp_path / "run_f").read_text()
    assert "Estimated Ln Prob of Data" in got and "Inferred ancestry of individuals:" in got
    want = _oracle_file(tmp_path, 2, 4242, "oracle")
    a = got.replace("run", "X").splitlines()
    b = want.replace(str(tmp_path / "oracle"), "X").splitlines()
    assert len(a) == len(b)
    for la, lb in zip(a, b):
        if la != lb:   # FP64 trajectories agree to ~1e-9: printed digits may differ by 1 ulp
            na, nb = _numbers(la), _numbers(lb)
            assert len(na) == len(nb) and np.allclose(na, nb, rtol=0, atol=0.11), (la, lb)
    q = np.loadtxt(tmp_path / "run_q", usecols=(2, 3))
    assert q.shape == (200, 2) and np.allclose(q.sum(axis=1), 1.0, atol=2e-4)


def test_ksweep_batch_equals_single_runs(tmp_path):
    _files(tmp_path, outfile="sweep")
    out = _run(tmp_path, "-B", "1", "3", "2", "-D", "900")
    assert "K= 3" in out
    for K in (1, 2, 3):
        summ = (tmp_path / f"sweep_K{K}_summary").read_text()
        assert f"K = {K}, 2 replicates" in summ
    # replicate 2 of K=2 is the chain (seed 901, K 2): run it alone
    single = tmp_path / "single"
    single.mkdir()
    _files(single, outfile="one")
    _run(single, "-K", "2", "-D", "901")
    a = (tmp_path / "sweep_K2_R2_f").read_text()
    b = (single / "one_f").read_text()

    def body(t):   # drop the command line and file-name lines
        return [ln for ln in t.splitlines()
                if not ln.startswith(("Command line", "DATAFILE", "Output File"))]
    assert body(a) == body(b)


OPTIONS = """#define ANCESTDIST 1
#define NUMBOXES 100
#define ANCESTPINT 0.9
#define INTERMEDSAVE 2
#define PRINTLIKES 1
#define PFROMPOPFLAGONLY 1
"""


def test_structure_options_match_oracle_chain(tmp_path):
    """ANCESTDIST (intervals in the ancestry table), INTERMEDSAVE (result files
    <OUTFILE>_1, _2 at (rep - BURNIN) / savefreq, structure.c:640-661),
    PRINTLIKES (one lnP(D) line per sweep, structure.c:663-667) and
    PFROMPOPFLAGONLY, against the oracle chain with the same key."""
    _files(tmp_path)
    (tmp_path / "extraparams").write_text(EXTRA + OPTIONS)
    out = _run(tmp_path)
    p = host.get_params(["structure", "-m", str(tmp_path / "mainparams"), "-e",
                         str(tmp_path / "extraparams")])
    seed, K = 4242, 2
    p.seed = seed
    inp = host.InputFile(p)
    assert inp.popflag.min() >= 0 and 0 < (inp.popflag == 1).sum() < inp.N
    m = oracle.make_model(inp.N, inp.L, K, inp.A, burnin=p.burnin, numreps=p.numreps,
                          freqscorr=p.freqscorr, inferalpha=p.inferalpha,
                          inferlambda=p.inferlambda, metrofreq=p.metrofreq, alpha0=p.alpha,
                          lambda0=p.lambda_, alphapropsd=p.alphapropsd,
                          fpriormean=p.fpriormean, fpriorsd=p.fpriorsd, seed=seed, chain=K,
                          ancestdist=1, numboxes=100, pfrompopflagonly=1)
    ch = oracle.Chain(m, inp.geno_int8(), inp.nall, popflag=inp.popflag)
    ch.init()
    total = p.burnin + p.numreps
    savefreq = host.save_freq(p)
    assert savefreq == 20
    likes = []
    for rep in range(total):
        ch.sweep(rep)
        likes.append(ch.like if rep + 1 > p.burnin else ch.calc_like(ch.Q, ch.P))
        n = rep + 1 - p.burnin
        if n > 0 and n % savefreq == 0 and rep + 1 != total:
            pi = host.get_params(["structure", "-m", str(tmp_path / "mainparams"), "-e",
                                  str(tmp_path / "extraparams")])
            pi.seed = seed
            pi.outfile = str(tmp_path / "oracle").encode()
            inp.write_results(rep + 1, ch.QSum, ch.PSum, ch.epsSum, ch.FstSum, ch.alphaSum,
                              ch.lambdaSum, ch.sumlikes, ch.sumsqlikes, final=0, params=pi,
                              ancestdist=ch.ancestDist, saveindex=n // savefreq)
    pi.outfile = str(tmp_path / "oracle").encode()
    inp.write_results(total, ch.QSum, ch.PSum, ch.epsSum, ch.FstSum, ch.alphaSum, ch.lambdaSum,
                      ch.sumlikes, ch.sumsqlikes, final=2, params=pi, ancestdist=ch.ancestDist)
    printed = [int(x) for x in re.findall(r"^ *(-?\d+)$", out, flags=re.M)]
    assert len(printed) == total
    assert np.all(np.abs(np.array(printed) - np.round(likes)) <= 1), "PRINTLIKES lines"
    for ours, theirs in [("run_1", "oracle_1"), ("run_2", "oracle_2"), ("run_f", "oracle_f")]:
        a = (tmp_path / ours).read_text().replace(ours[:3], "X").splitlines()
        b = (tmp_path / theirs).read_text().replace(str(tmp_path / "oracle"), "X").splitlines()
        assert len(a) == len(b), ours
        assert any("probability intervals" in ln for ln in a)
        assert any("POPFLAG=1 ONLY" in ln for ln in a)
        for la, lb in zip(a, b):
            if la != lb:   # FP64 trajectories agree to ~1e-9: printed digits may differ
                na, nb = _numbers(la), _numbers(lb)
                assert len(na) == len(nb) and np.allclose(na, nb, rtol=0, atol=0.11), (la, lb)
    assert not (tmp_path / "run_3").exists()


def test_structure_with_200_alleles_at_a_locus(tmp_path):
    """A locus with 200 distinct alleles (recoded codes 0..199, beyond the int8
    range): the command packs the recoded int32 genotypes into bytes
    (sgpu_pack_byte32) and its result file matches the oracle chain's."""
    rng = np.random.default_rng(12)
    N, L = 120, 4
    lines = []
    for i in range(N):
        for c in range(2):
            al = [1000 + (2 * i + c) % 200,
                  int(rng.choice([5, 7])) if rng.random() > 0.02 else -9,
                  int(rng.choice([11, 12, 13])), int(rng.choice([1, 2]))]
            lines.append(f"ind{i} " + " ".join(str(a) for a in al))
    (tmp_path / "many.str").write_text("\n".join(lines) + "\n")
    main = MAIN.replace("NUMINDS    200", f"NUMINDS    {N}").replace(
        "NUMLOCI    5", f"NUMLOCI    {L}").replace("POPDATA   1", "POPDATA   0").replace(
        "POPFLAG   1", "POPFLAG   0").replace("MISSING     -999", "MISSING     -9")
    (tmp_path / "mainparams").write_text(main.format(infile=str(tmp_path / "many.str"),
                                                     outfile="many"))
    (tmp_path / "extraparams").write_text(EXTRA.replace("PRINTQHAT 1", "PRINTQHAT 0"))
    out = _run(tmp_path)
    assert "Genotypes packed as bytes" in out
    got = (tmp_path / "many_f").read_text()
    assert "200 alleles" in got
    want = _oracle_file(tmp_path, 2, 4242, "oracle")
    a = got.replace("many", "X").splitlines()
    b = want.replace(str(tmp_path / "oracle"), "X").splitlines()
    assert len(a) == len(b)
    for la, lb in zip(a, b):
        if la != lb and not la.startswith(("Input File", "DATAFILE")):
            na, nb = _numbers(la), _numbers(lb)
            assert len(na) == len(nb) and np.allclose(na, nb, rtol=0, atol=0.11), (la, lb)
