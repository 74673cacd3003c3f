/*
 * Genotype packing for the sgpu C-ABI (host side).
 *
 * Input is the reference's recoded genotype array Geno[ind][line][loc]
 * (structure.h:22 GenPos; alleles recoded 0..k-1 by first appearance,
 * datain.c:824-904; -9 = missing).  Two device formats, each laid out so that
 * the lanes of a wavefront read consecutive words (256-byte loads):
 *   SGPU_GENO_SNP2  2 bits per (ind, locus), 16 individuals per uint32 word,
 *                   [word row][Lpad]: its UpdateZ kernel maps one locus to a lane
 *   SGPU_GENO_BYTE  1 byte per allele copy, 2 individuals per uint32 word,
 *                   [locus][Gpad]: its kernels map one individual pair (or one
 *                   individual) to a lane
 */
#include <stdint.h>
#include <string.h>

#include "../../include/sgpu.h"

#define ZLOCI 256   /* loci per UpdateZ block; rows are padded to a multiple */

#define WAVE 64    /* byte format: word rows (individual pairs) padded to a wavefront */

static size_t lpad_of(int32_t L) { return ((size_t)L + ZLOCI - 1) / ZLOCI * ZLOCI; }
static size_t gpad_of(int32_t N) { return (((size_t)N + 1) / 2 + WAVE - 1) / WAVE * WAVE; }

size_t sgpu_packed_words(int32_t fmt, int32_t N, int32_t L)
{
    if (N < 1 || L < 1) {
        return 0;
    }
    if (fmt == SGPU_GENO_SNP2) {
        return ((size_t)N + 15) / 16 * lpad_of(L);
    }
    if (fmt == SGPU_GENO_BYTE) {
        return (size_t)L * gpad_of(N);
    }
    return 0;
}

int sgpu_pack_snp2(const int8_t *geno, int32_t N, int32_t L, uint32_t *out, size_t out_words)
{
    const size_t lp = lpad_of(L);
    if (!geno || !out || out_words != sgpu_packed_words(SGPU_GENO_SNP2, N, L)) {
        return SGPU_ERR_ARG;
    }
    memset(out, 0xff, out_words * sizeof(uint32_t));   /* padding = code 3 (missing) */
    for (int32_t i = 0; i < N; i++) {
        const int8_t *c0 = geno + (size_t)i * 2 * L;
        const int8_t *c1 = c0 + L;
        uint32_t *row = out + (size_t)(i >> 4) * lp;
        const int sh = 2 * (i & 15);
        for (int32_t l = 0; l < L; l++) {
            const int a = c0[l], b = c1[l];
            uint32_t code;
            if (a == -9 && b == -9) {
                code = 3u;
            } else if (a == -9 || b == -9) {
                return SGPU_ERR_UNSUPPORTED;   /* one copy missing */
            } else if (a == 0 && b == 0) {
                code = 0u;
            } else if (a == 0 && b == 1) {
                code = 1u;
            } else if (a == 1 && b == 1) {
                code = 2u;
            } else {
                return SGPU_ERR_UNSUPPORTED;   /* allele > 1 or heterozygote stored as (1,0) */
            }
            row[l] = (row[l] & ~(3u << sh)) | (code << sh);
        }
    }
    return SGPU_OK;
}

/* Byte format from any integer genotype type: G is the element type, miss the
 * missing code.  Alleles 0..254; 255 marks a missing copy. */
#define PACK_BYTE_BODY(G)                                                                  \
    const size_t gp = gpad_of(N);                                                          \
    if (!geno || !out || out_words != sgpu_packed_words(SGPU_GENO_BYTE, N, L)) {          \
        return SGPU_ERR_ARG;                                                               \
    }                                                                                      \
    memset(out, 0xff, out_words * sizeof(uint32_t)); /* padding rows: copies missing */    \
    /* loci in blocks of 64, so that both the reads along a copy and the writes       */  \
    /* down a locus column stay within a few KiB                                       */  \
    for (int32_t l0 = 0; l0 < L; l0 += 64) {                                               \
        const int32_t l1 = l0 + 64 < L ? l0 + 64 : L;                                      \
        for (int32_t i = 0; i < N; i++) {                                                  \
            const int sh = 16 * (i & 1);                                                   \
            for (int c = 0; c < 2; c++) {                                                  \
                const G *g = geno + ((size_t)i * 2 + c) * L;                               \
                const int s = sh + 8 * c;                                                  \
                for (int32_t l = l0; l < l1; l++) {                                        \
                    const int32_t x = (int32_t)g[l];                                       \
                    uint32_t v;                                                            \
                    if (x == miss) {                                                       \
                        v = 0xffu;                                                         \
                    } else if (x >= 0 && x < 255) {                                        \
                        v = (uint32_t)x;                                                   \
                    } else {                                                               \
                        return SGPU_ERR_ARG;                                               \
                    }                                                                      \
                    uint32_t *w = out + (size_t)l * gp + (size_t)(i >> 1);                 \
                    *w = (*w & ~(0xffu << s)) | (v << s);                                  \
                }                                                                          \
            }                                                                              \
        }                                                                                  \
    }                                                                                      \
    return SGPU_OK;

int sgpu_pack_byte(const int8_t *geno, int32_t N, int32_t L, uint32_t *out, size_t out_words)
{
    const int32_t miss = -9;
    PACK_BYTE_BODY(int8_t)
}

int sgpu_pack_byte32(const int32_t *geno, int32_t N, int32_t L, int32_t missing, uint32_t *out,
                     size_t out_words)
{
    const int32_t miss = missing;
    PACK_BYTE_BODY(int32_t)
}
