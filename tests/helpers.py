"""Shared test helpers: small random states and Philox-derived uniforms."""
from __future__ import annotations

import numpy as np

import oracle

RAND_MAX = 2 ** 31 - 1


def random_geno(rng, N, L, A_max, miss=0.05, biallelic=False):
    """Genotypes [N][2][L] int8 with recoded alleles and -9 for missing copies.

    nall[l] = number of distinct alleles seen at l (first-appearance coding is
    not needed here: any value < nall[l] is a valid recoded allele)."""
    if biallelic:
        nall = np.full(L, 2, np.int32)
    else:
        nall = rng.integers(1, A_max + 1, size=L).astype(np.int32)
    geno = np.empty((N, 2, L), np.int8 if A_max <= 127 else np.int16)
    for l in range(L):
        geno[:, :, l] = rng.integers(0, nall[l], size=(N, 2))
    geno[rng.random((N, 2, L)) < miss] = -9
    return geno, nall


def random_q(rng, N, K, a=1.0):
    return rng.dirichlet(np.full(K, a), size=N)


def random_p(rng, L, K, A, nall, a=1.0):
    P = np.zeros((L, K, A))
    for l in range(L):
        if nall[l] > 0:
            P[l, :, : nall[l]] = rng.dirichlet(np.full(nall[l], a), size=K)
    return P


def z_uniforms(seed, chain, rep, N, L):
    """randomArr for the reference UpdateZ: [(ind*L + loc)*2 + line] (UpdateZ.c:37)."""
    r = np.zeros(N * L * 2)
    for i in range(N):
        for l in range(L):
            w = oracle.philox_block(seed, chain, rep, l, i >> 1, oracle.PH_Z << 24)
            for c in range(2):
                r[(i * L + l) * 2 + c] = w[2 * (i & 1) + c] * 2.0 ** -32
    return r


def geno_ref_layout(geno):
    """[N][2][L] int8 -> int32 in the reference GenPos layout (structure.h:22)."""
    return np.ascontiguousarray(geno.astype(np.int32))


def random_snp_codes(rng, N, L, miss=0.05):
    """2-bit genotype codes [N][L] (0 = (0,0), 1 = (0,1), 2 = (1,1), 3 = missing)
    and the same genotypes as [N][2][L] int8 (-9 missing)."""
    codes = rng.integers(0, 3, size=(N, L)).astype(np.uint32)
    codes[rng.random((N, L)) < miss] = 3
    geno = np.empty((N, 2, L), np.int8)
    geno[:, 0, :] = (codes == 2)
    geno[:, 1, :] = (codes != 0)
    miss_i, miss_l = np.nonzero(codes == 3)
    geno[miss_i, :, miss_l] = -9
    return codes, geno


def pack_snp2_words(codes, Lpad=None):
    """Words [ceil(N/16)][Lpad] of 2-bit codes [N][L]: individual 16g+t at bits
    2t..2t+1 of word (g, l); padding individuals and loci read 3 (missing)."""
    N, L = codes.shape
    if Lpad is None:
        Lpad = (L + 255) // 256 * 256
    G = (N + 15) // 16
    full = np.full((G * 16, Lpad), 3, np.uint32)
    full[:N, :L] = codes
    full = full.reshape(G, 16, Lpad)
    words = np.zeros((G, Lpad), np.uint32)
    for t in range(16):
        words |= full[:, t, :] << np.uint32(2 * t)
    return words
