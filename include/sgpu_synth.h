/*
 * sgpu_synth.h -- seeded synthetic genotype generators (host only).
 *
 * No counterpart in the reference, which reads real data only (datain.c).
 * They produce the benchmark and test datasets of the admixture sweep:
 * already recoded by first appearance (datain.c:824-904), heterozygotes listed
 * as (0,1), whole genotypes missing.  Every individual draws from its own
 * xoshiro256** stream seeded from (seed, individual), so the output does not
 * depend on the thread count.
 */
#ifndef SGPU_SYNTH_H
#define SGPU_SYNTH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Balding-Nichols SNP panel: ancestral p ~ U(0.05, 0.95) per locus; per
 * population F_k ~ U(flo, fhi) and p_k ~ Beta(p(1-F)/F, (1-p)(1-F)/F);
 * Q_i ~ Dirichlet(qalpha); each copy draws z ~ Q_i then allele ~ Bernoulli(p_z);
 * a genotype is missing with probability miss.  Writes SGPU_GENO_SNP2 words
 * into packed (sgpu_packed_words(SGPU_GENO_SNP2, N, L) of them), optionally the
 * unpacked geno[N][2][L] (int8, may be NULL) and nall[L] (observed alleles),
 * and optionally the true Q [N][Ktrue] (may be NULL). */
int sgpu_synth_snp(int32_t N, int32_t L, int32_t Ktrue, double flo, double fhi, double qalpha,
                   double miss, uint64_t seed, int32_t nthreads, uint32_t *packed,
                   size_t packed_words, int8_t *geno, int32_t *nall, double *qtrue);

/* Microsatellite-like panel: per locus an allele count U{amin..amax}; per
 * population frequencies ~ Dirichlet(1,...,1); Q_i ~ Dirichlet(qalpha); each
 * copy draws z ~ Q_i then an allele from that population; genotypes missing
 * with probability miss; alleles recoded by first appearance.  Writes
 * geno[N][2][L] and nall[L]; returns MAXALLELES (>= 1) or a negative status. */
int sgpu_synth_microsat(int32_t N, int32_t L, int32_t amin, int32_t amax, int32_t Ktrue,
                        double qalpha, double miss, uint64_t seed, int8_t *geno, int32_t *nall,
                        double *qtrue);

#ifdef __cplusplus
}
#endif
#endif
