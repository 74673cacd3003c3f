/*
 * Test harness linked into oracle/_ref/libstructref.so together with the
 * reference CPU sources (compiled where they lie under /root/reference by
 * oracle/Makefile).  The reference is built with -Drand=sref_rand, so every
 * rand() it makes (ran.c:128-147 RandomReal/RandomInteger, hence rnd(),
 * snorm, RGamma) reads the queue below instead of libc.  Tests fill the queue
 * with known integers in [0, RAND_MAX] and so know every uniform the
 * reference consumed.  TEST INFRASTRUCTURE ONLY.
 */
#include <stdlib.h>
#include <string.h>

static int *sref_queue = NULL;
static size_t sref_qlen = 0;
static size_t sref_qpos = 0;
static size_t sref_overrun = 0;

void sref_set_rand(const int *vals, size_t n)
{
    free(sref_queue);
    sref_queue = NULL;
    sref_qlen = 0;
    sref_qpos = 0;
    sref_overrun = 0;
    if (n) {
        sref_queue = (int *)malloc(n * sizeof(int));
        memcpy(sref_queue, vals, n * sizeof(int));
        sref_qlen = n;
    }
}

int sref_rand(void)
{
    if (sref_qpos < sref_qlen) {
        return sref_queue[sref_qpos++];
    }
    sref_overrun++;
    return 0;
}

size_t sref_rand_consumed(void) { return sref_qpos; }
size_t sref_rand_overrun(void) { return sref_overrun; }
int sref_rand_max(void) { return RAND_MAX; }
