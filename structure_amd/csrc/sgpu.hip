// C-ABI of the MI355X admixture sweep (include/sgpu.h): context, device
// state, and the per-phase entry points that replace the reference's *CL
// wrappers and its OpenCL dispatch layer (Kernels.c).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <vector>

#include "../../include/sgpu.h"
#include "kernel_update.h"
#include "kernel_z.h"

using namespace sgpu;

namespace {

thread_local char g_err[1024] = "";

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define HIP_OK(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            return fail(SGPU_ERR_DEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                \
        }                                                                                   \
    } while (0)

#define LAUNCH_OK()                                                                         \
    do {                                                                                    \
        hipError_t e_ = hipGetLastError();                                                  \
        if (e_ != hipSuccess) {                                                             \
            return fail(SGPU_ERR_DEVICE, "kernel launch failed: %s (%s:%d)",                \
                        hipGetErrorString(e_), __FILE__, __LINE__);                         \
        }                                                                                   \
    } while (0)

size_t cdiv(size_t a, size_t b) { return (a + b - 1) / b; }

}  // namespace

struct sgpu_ctx {
    sgpu_model m{};
    ModelArgs ma{};
    int fmt = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    int Lpad = 0, G = 0, inds_per_row = 0;
    // UpdateZ tiling
    int n_lt = 0, n_it = 0, rows_per_tile = 0;
    size_t z_lds = 0;
    // Metropolis tiling
    int n_ranges = 0, loci_per_range = 0;
    // partial-buffer block counts
    int q_blocks = 0, l_blocks = 0;
    int logq_rows = 0;   // rows of logq_part written by the last Q update
    // device state
    uint32_t *geno = nullptr;
    int32_t *nall = nullptr, *ahist = nullptr;
    double *Q = nullptr, *P = nullptr, *eps = nullptr, *F = nullptr, *alpha = nullptr,
           *lambda = nullptr;
    int32_t *nApart = nullptr, *nQpart = nullptr, *nA = nullptr, *nQ = nullptr, *nQfold = nullptr;
    double *likepart = nullptr, *logq_part = nullptr, *delta_part = nullptr,
           *lam_part = nullptr, *fst_part = nullptr;
    double *QSum = nullptr, *PSum = nullptr, *epsSum = nullptr, *FstSum = nullptr,
           *alphaSum = nullptr, *lambdaSum = nullptr, *like3 = nullptr;
    int8_t *zdbg = nullptr;
    uint32_t *repd = nullptr;     // device copy of the sweep index read by every kernel
    uint32_t rep_shadow = 0;      // host view of *repd
    bool rep_known = false;
    bool capturing = false;
    bool use_graphs = true;
    hipGraphExec_t graph[4] = {nullptr, nullptr, nullptr, nullptr};   // [metro | 2*collect]
    int nhist = 0, nloci_nonzero = 0;
    int64_t nsum = 0;
    bool initialized = false;
    bool like_ready = false;
    std::vector<void *> allocs;
    // phase timing (sgpu_set_timing)
    struct Pending {
        int phase;
        hipEvent_t a, b;
    };
    bool timing = false;
    std::vector<hipEvent_t> free_events;
    std::vector<Pending> pending;
    double t_ms[SGPU_NPHASE] = {};
    int64_t t_n[SGPU_NPHASE] = {};
};

namespace {

template <typename T>
int dalloc(sgpu_ctx *c, T **p, size_t count)
{
    void *q = nullptr;
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) {
        return fail(SGPU_ERR_NOMEM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    }
    c->allocs.push_back(q);
    *p = (T *)q;
    return SGPU_OK;
}

ModelArgs model_args(const sgpu_model &m)
{
    ModelArgs a{};
    a.N = m.N;
    a.L = m.L;
    a.K = m.K;
    a.A = m.A;
    a.freqscorr = m.freqscorr;
    a.onefst = m.onefst;
    a.popalphas = m.popalphas;
    a.unifprioralpha = m.unifprioralpha;
    a.popspecificlambda = m.popspecificlambda;
    a.alphapropsd = m.alphapropsd;
    a.alphamax = m.alphamax;
    a.alphapriora = m.alphapriora;
    a.alphapriorb = m.alphapriorb;
    a.fpriormean = m.fpriormean;
    a.fpriorsd = m.fpriorsd;
    a.seed = m.seed;
    a.chain = m.chain;
    return a;
}

ZArgs z_args(sgpu_ctx *c, uint32_t rep)
{
    ZArgs a{};
    a.geno = c->geno;
    a.Q = c->Q;
    a.P = c->P;
    a.nApart = c->nApart;
    a.nQpart = c->nQpart;
    a.likepart = c->likepart;
    a.zout = nullptr;
    a.N = c->m.N;
    a.L = c->m.L;
    a.K = c->m.K;
    a.A = c->m.A;
    a.Lpad = c->Lpad;
    a.G = c->G;
    a.rows_per_tile = c->rows_per_tile;
    a.inds_per_row = c->inds_per_row;
    a.seed = c->m.seed;
    a.chain = c->m.chain;
    return a;
}

typedef void (*zkern_t)(const uint32_t *, const double *, const double *, int32_t *, int32_t *,
                        double *, int8_t *, ZDims, const uint32_t *);

template <int K>
zkern_t pick_snp2(int mode)
{
    switch (mode) {
    case 0: return k_update_z_snp2<K, true, false, false>;
    case 1: return k_update_z_snp2<K, false, false, false>;
    case 2: return k_update_z_snp2<K, false, true, false>;
    case 3: return k_update_z_snp2<K, false, false, true>;
    default: return k_update_z_snp2<K, false, true, true>;
    }
}

zkern_t snp2_kernel(int K, int mode)
{
    switch (K) {
    case 1: return pick_snp2<1>(mode);
    case 2: return pick_snp2<2>(mode);
    case 3: return pick_snp2<3>(mode);
    case 4: return pick_snp2<4>(mode);
    case 5: return pick_snp2<5>(mode);
    case 6: return pick_snp2<6>(mode);
    case 7: return pick_snp2<7>(mode);
    case 8: return pick_snp2<8>(mode);
    case 9: return pick_snp2<9>(mode);
    case 10: return pick_snp2<10>(mode);
    case 11: return pick_snp2<11>(mode);
    case 12: return pick_snp2<12>(mode);
    case 13: return pick_snp2<13>(mode);
    case 14: return pick_snp2<14>(mode);
    case 15: return pick_snp2<15>(mode);
    case 16: return pick_snp2<16>(mode);
    default: return nullptr;
    }
}

typedef void (*mkern_t)(ModelArgs, const uint32_t *, int, const double *, const double *,
                        const double *, int, double *, const uint32_t *);

mkern_t metro_snp2_kernel(int K)
{
    switch (K) {
    case 1: return k_metro_ratio_snp2<1>;
    case 2: return k_metro_ratio_snp2<2>;
    case 3: return k_metro_ratio_snp2<3>;
    case 4: return k_metro_ratio_snp2<4>;
    case 5: return k_metro_ratio_snp2<5>;
    case 6: return k_metro_ratio_snp2<6>;
    case 7: return k_metro_ratio_snp2<7>;
    case 8: return k_metro_ratio_snp2<8>;
    case 9: return k_metro_ratio_snp2<9>;
    case 10: return k_metro_ratio_snp2<10>;
    case 11: return k_metro_ratio_snp2<11>;
    case 12: return k_metro_ratio_snp2<12>;
    case 13: return k_metro_ratio_snp2<13>;
    case 14: return k_metro_ratio_snp2<14>;
    case 15: return k_metro_ratio_snp2<15>;
    case 16: return k_metro_ratio_snp2<16>;
    default: return nullptr;
    }
}

// mode 0 = initial Z, 1 = update, 2 = update + lnP terms (+2 with a debug copy of Z)
int launch_z(sgpu_ctx *c, uint32_t rep, int mode, int8_t *zdev)
{
    ZArgs a = z_args(c, rep);
    a.zout = zdev;
    dim3 grid(c->n_lt, c->n_it), block(kZLoci);
    if (c->fmt == SGPU_GENO_SNP2) {
        zkern_t k = snp2_kernel(c->m.K, (zdev && mode > 0) ? mode + 2 : mode);
        if (!k) {
            return fail(SGPU_ERR_UNSUPPORTED, "no SNP2 UpdateZ kernel for K=%d", c->m.K);
        }
        ZDims zd{};
        zd.N = c->m.N;
        zd.L = c->m.L;
        zd.A = c->m.A;
        zd.Lpad = c->Lpad;
        zd.G = c->G;
        zd.rows_per_tile = c->rows_per_tile;
        zd.seed = c->m.seed;
        zd.chain = c->m.chain;
            hipLaunchKernelGGL(k, grid, block, c->z_lds, c->stream, c->geno, c->Q, c->P, c->nApart,
                           c->nQpart, c->likepart, zdev, zd, (const uint32_t *)c->repd);
    } else {
        if (mode == 0) {
            hipLaunchKernelGGL((k_update_z_byte<true, false>), grid, block, c->z_lds, c->stream, a,
                               (const uint32_t *)c->repd);
        } else if (mode == 1) {
            hipLaunchKernelGGL((k_update_z_byte<false, false>), grid, block, c->z_lds, c->stream, a,
                               (const uint32_t *)c->repd);
        } else {
            hipLaunchKernelGGL((k_update_z_byte<false, true>), grid, block, c->z_lds, c->stream, a,
                               (const uint32_t *)c->repd);
        }
    }
    LAUNCH_OK();
    return SGPU_OK;
}

size_t state_elems(const sgpu_ctx *c, int which, size_t *elem)
{
    const size_t N = c->m.N, L = c->m.L, K = c->m.K, A = c->m.A;
    *elem = sizeof(double);
    switch (which) {
    case SGPU_Q: case SGPU_QSUM: return N * K;
    case SGPU_P: case SGPU_PSUM: return L * K * A;
    case SGPU_EPS: case SGPU_EPSSUM: return L * A;
    case SGPU_FST: case SGPU_ALPHA: case SGPU_LAMBDA: case SGPU_FSTSUM: case SGPU_ALPHASUM:
    case SGPU_LAMBDASUM: return K;
    case SGPU_NA: *elem = sizeof(int32_t); return L * K * A;
    case SGPU_NQ: *elem = sizeof(int32_t); return N * K;
    case SGPU_LIKE: return 3;
    case SGPU_NSUM: *elem = sizeof(int64_t); return 1;
    default: return 0;
    }
}

double *state_ptr(sgpu_ctx *c, int which)
{
    switch (which) {
    case SGPU_Q: return c->Q;
    case SGPU_P: return c->P;
    case SGPU_EPS: return c->eps;
    case SGPU_FST: return c->F;
    case SGPU_ALPHA: return c->alpha;
    case SGPU_LAMBDA: return c->lambda;
    case SGPU_QSUM: return c->QSum;
    case SGPU_PSUM: return c->PSum;
    case SGPU_FSTSUM: return c->FstSum;
    case SGPU_EPSSUM: return c->epsSum;
    case SGPU_ALPHASUM: return c->alphaSum;
    case SGPU_LAMBDASUM: return c->lambdaSum;
    case SGPU_LIKE: return c->like3;
    default: return nullptr;
    }
}

int launch_logq(sgpu_ctx *c)
{
    hipLaunchKernelGGL(k_logq_stats, dim3(c->q_blocks), dim3(kNT), 0, c->stream, c->ma, c->Q,
                       c->logq_part);
    LAUNCH_OK();
    c->logq_rows = c->q_blocks;
    return SGPU_OK;
}

hipEvent_t take_event(sgpu_ctx *c)
{
    if (!c->free_events.empty()) {
        hipEvent_t e = c->free_events.back();
        c->free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) {
        return nullptr;
    }
    return e;
}

int drain_timing(sgpu_ctx *c)
{
    if (c->pending.empty()) {
        return SGPU_OK;
    }
    HIP_OK(hipStreamSynchronize(c->stream));
    for (int t = 0; t < 4; t++) {
        if (c->graph[t]) {
            (void)hipGraphExecDestroy(c->graph[t]);
        }
    }
    for (auto &p : c->pending) {
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, p.a, p.b));
        c->t_ms[p.phase] += ms;
        c->t_n[p.phase] += 1;
        c->free_events.push_back(p.a);
        c->free_events.push_back(p.b);
    }
    c->pending.clear();
    return SGPU_OK;
}

// Brackets one phase's launches with events on the context stream.
struct PhaseTimer {
    sgpu_ctx *c;
    int phase;
    hipEvent_t a = nullptr;
    PhaseTimer(sgpu_ctx *ctx, int ph) : c(ctx), phase(ph)
    {
        if (c && c->timing && (a = take_event(c))) {
            (void)hipEventRecord(a, c->stream);
        }
    }
    ~PhaseTimer()
    {
        if (!a) {
            return;
        }
        hipEvent_t b = take_event(c);
        if (!b) {
            c->free_events.push_back(a);
            return;
        }
        (void)hipEventRecord(b, c->stream);
        c->pending.push_back({phase, a, b});
        if (c->pending.size() > 8192) {
            (void)drain_timing(c);
        }
    }
};

__global__ void k_set_rep(uint32_t *repd, uint32_t v) { repd[0] = v; }
__global__ void k_advance_rep(uint32_t *repd) { repd[0] += 1u; }

// Make the device sweep index equal rep (no-op while capturing a graph, whose
// kernels read the index that the previous replay advanced).
int use_rep(sgpu_ctx *c, uint32_t rep)
{
    if (c->capturing || (c->rep_known && c->rep_shadow == rep)) {
        return SGPU_OK;
    }
    hipLaunchKernelGGL(k_set_rep, dim3(1), dim3(1), 0, c->stream, c->repd, rep);
    LAUNCH_OK();
    c->rep_shadow = rep;
    c->rep_known = true;
    return SGPU_OK;
}

int check_ready(sgpu_ctx *c)
{
    if (!c) {
        return fail(SGPU_ERR_ARG, "null context");
    }
    if (!c->initialized) {
        return fail(SGPU_ERR_STATE, "sgpu_init_state has not been called");
    }
    return SGPU_OK;
}

}  // namespace

extern "C" {

const char *sgpu_last_error(void) { return g_err; }

const char *sgpu_version(void)
{
    return "sgpu 0.1 gfx950: fused UpdateZ (SNP2 and byte formats, K<=16), FP64 draws";
}

int sgpu_create(const sgpu_model *m, const sgpu_data *d, sgpu_ctx **out)
{
    if (!m || !d || !out || !d->geno || !d->nall) {
        return fail(SGPU_ERR_ARG, "null argument");
    }
    *out = nullptr;
    if (m->N < 1 || m->L < 1 || m->K < 1 || m->A < 1) {
        return fail(SGPU_ERR_ARG, "bad shape N=%d L=%d K=%d A=%d", m->N, m->L, m->K, m->A);
    }
    if (m->K > kMaxK) {
        return fail(SGPU_ERR_UNSUPPORTED, "K=%d exceeds the supported maximum %d", m->K, kMaxK);
    }
    if (m->A > kMaxAlleles) {
        return fail(SGPU_ERR_UNSUPPORTED, "MAXALLELES=%d exceeds %d", m->A, kMaxAlleles);
    }
    if (d->format == SGPU_GENO_SNP2 && m->A != 2) {
        return fail(SGPU_ERR_ARG, "SNP2 genotypes need MAXALLELES == 2 (got %d)", m->A);
    }
    if (d->format != SGPU_GENO_SNP2 && d->format != SGPU_GENO_BYTE) {
        return fail(SGPU_ERR_ARG, "unknown genotype format %d", d->format);
    }
    const size_t need = sgpu_packed_words(d->format, m->N, m->L);
    if (d->geno_words != need) {
        return fail(SGPU_ERR_ARG, "genotype buffer has %zu words, format needs %zu",
                    d->geno_words, need);
    }
    for (int l = 0; l < m->L; l++) {
        if (d->nall[l] < 0 || d->nall[l] > m->A) {
            return fail(SGPU_ERR_ARG, "NumAlleles[%d]=%d outside [0, %d]", l, d->nall[l], m->A);
        }
    }
    // every allele index must lie below NumAlleles of its locus (the kernels index
    // P and the count rows with it)
    {
        const size_t lp = need / (size_t)cdiv(m->N, d->format == SGPU_GENO_SNP2 ? 16 : 2);
        for (size_t row = 0; row * lp < need; row++) {
            const uint32_t *w = d->geno + row * lp;
            for (int l = 0; l < m->L; l++) {
                const uint32_t x = w[l];
                if (d->format == SGPU_GENO_SNP2) {
                    for (int j = 0; j < 16; j++) {
                        const uint32_t code = (x >> (2 * j)) & 3u;
                        if (code != 3u && (int)(code != 0u) >= d->nall[l]) {
                            return fail(SGPU_ERR_ARG, "allele 1 at locus %d exceeds NumAlleles=%d", l,
                                        d->nall[l]);
                        }
                        if (code == 0u && d->nall[l] < 1) {
                            return fail(SGPU_ERR_ARG, "allele 0 at locus %d with NumAlleles=0", l);
                        }
                    }
                } else {
                    for (int b = 0; b < 4; b++) {
                        const uint32_t a = (x >> (8 * b)) & 0xffu;
                        if (a != 0xffu && (int)a >= d->nall[l]) {
                            return fail(SGPU_ERR_ARG, "allele %u at locus %d exceeds NumAlleles=%d",
                                        a, l, d->nall[l]);
                        }
                    }
                }
            }
        }
    }
    if (m->freqscorr && (m->fpriormean <= 0.0 || m->fpriorsd <= 0.0)) {
        return fail(SGPU_ERR_ARG, "FPRIORMEAN and FPRIORSD must be positive");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        return fail(SGPU_ERR_DEVICE, "no HIP device available");
    }

    sgpu_ctx *c = new sgpu_ctx();
    c->m = *m;
    c->ma = model_args(*m);
    c->fmt = d->format;
    if (hipGetDevice(&c->device) != hipSuccess) {
        delete c;
        return fail(SGPU_ERR_DEVICE, "hipGetDevice failed");
    }
    int rc = SGPU_OK;
    auto cleanup = [&](int code) {
        sgpu_destroy(c);
        return code;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(SGPU_ERR_DEVICE, "hipStreamCreate failed");
    }

    const int N = m->N, L = m->L, K = m->K, A = m->A;
    c->inds_per_row = d->format == SGPU_GENO_SNP2 ? 16 : 2;
    c->G = (int)cdiv(N, c->inds_per_row);
    c->n_lt = (int)cdiv(L, kZLoci);
    c->Lpad = c->n_lt * kZLoci;
    // about 2048 UpdateZ blocks (8 per CU) with per-tile LDS counts <= 48 KiB
    int n_it = (int)std::min<size_t>(std::max<size_t>(cdiv(2048, c->n_lt), 1), c->G);
    int rows = (int)cdiv(c->G, n_it);
    while ((size_t)rows * c->inds_per_row * K * 4 > 48 * 1024 && rows > 1) {
        rows = (rows + 1) / 2;
    }
    c->rows_per_tile = rows;
    c->n_it = (int)cdiv(c->G, rows);
    c->z_lds = (size_t)rows * c->inds_per_row * K * sizeof(uint32_t);
    c->q_blocks = (int)cdiv(N, kNT);
    c->l_blocks = (int)cdiv(L, kNT);
    c->n_ranges = (int)std::max<size_t>(1, std::min<size_t>(cdiv(2048, c->q_blocks), L));
    c->loci_per_range = (int)(cdiv(cdiv(L, c->n_ranges), 4) * 4);   // multiple of 4 loci
    c->n_ranges = (int)cdiv(L, c->loci_per_range);

    if ((rc = dalloc(c, &c->geno, need)) || (rc = dalloc(c, &c->nall, L)) ||
        (rc = dalloc(c, &c->ahist, 2 * (A + 1))) || (rc = dalloc(c, &c->Q, (size_t)N * K)) ||
        (rc = dalloc(c, &c->P, (size_t)L * K * A)) || (rc = dalloc(c, &c->eps, (size_t)L * A)) ||
        (rc = dalloc(c, &c->F, K)) || (rc = dalloc(c, &c->alpha, K)) ||
        (rc = dalloc(c, &c->lambda, K)) ||
        (rc = dalloc(c, &c->nApart, (size_t)c->n_it * L * K * A)) ||
        (rc = dalloc(c, &c->nQpart, (size_t)c->n_lt * N * K)) ||
        (rc = dalloc(c, &c->nA, (size_t)L * K * A)) || (rc = dalloc(c, &c->nQ, (size_t)N * K)) ||
        (rc = dalloc(c, &c->nQfold, (size_t)kRowSplit * N * K)) ||
        (rc = dalloc(c, &c->likepart, (size_t)c->n_it * c->n_lt)) ||
        (rc = dalloc(c, &c->logq_part, cdiv(N, kGibbsInds) * (K + 1))) ||
        (rc = dalloc(c, &c->delta_part, (size_t)c->n_ranges * N)) ||
        (rc = dalloc(c, &c->lam_part, (size_t)c->l_blocks * K)) ||
        (rc = dalloc(c, &c->fst_part, (size_t)c->l_blocks * K)) ||
        (rc = dalloc(c, &c->QSum, (size_t)N * K)) || (rc = dalloc(c, &c->PSum, (size_t)L * K * A)) ||
        (rc = dalloc(c, &c->epsSum, (size_t)L * A)) || (rc = dalloc(c, &c->FstSum, K)) ||
        (rc = dalloc(c, &c->alphaSum, K)) || (rc = dalloc(c, &c->lambdaSum, K)) ||
        (rc = dalloc(c, &c->like3, 3)) || (rc = dalloc(c, &c->repd, 2))) {
        return cleanup(rc);
    }
    // allele-count histogram over loci with more than one allele (lambda update)
    std::map<int, int> hist;
    int nzero = 0;
    for (int l = 0; l < L; l++) {
        if (d->nall[l] > 1) {
            hist[d->nall[l]]++;
        }
        if (d->nall[l] == 0) {
            nzero++;
        }
    }
    std::vector<int32_t> hv;
    for (auto &kv : hist) {
        hv.push_back(kv.first);
        hv.push_back(kv.second);
    }
    c->nhist = (int)hist.size();
    c->nloci_nonzero = L - nzero;
    hipError_t e = hipMemcpyAsync(c->geno, d->geno, need * sizeof(uint32_t), hipMemcpyHostToDevice,
                                  c->stream);
    if (e == hipSuccess) {
        e = hipMemcpyAsync(c->nall, d->nall, (size_t)L * sizeof(int32_t), hipMemcpyHostToDevice,
                           c->stream);
    }
    if (e == hipSuccess && !hv.empty()) {
        e = hipMemcpyAsync(c->ahist, hv.data(), hv.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                           c->stream);
    }
    if (e == hipSuccess) {
        e = hipStreamSynchronize(c->stream);
    }
    if (e != hipSuccess) {
        fail(SGPU_ERR_DEVICE, "upload failed: %s", hipGetErrorString(e));
        return cleanup(SGPU_ERR_DEVICE);
    }
    *out = c;
    return SGPU_OK;
}

int sgpu_destroy(sgpu_ctx *c)
{
    if (!c) {
        return SGPU_OK;
    }
    if (c->stream) {
        (void)hipStreamSynchronize(c->stream);
    }
    for (void *p : c->allocs) {
        (void)hipFree(p);
    }
    if (c->zdbg) {
        (void)hipFree(c->zdbg);
    }
    for (int t = 0; t < 4; t++) {
        if (c->graph[t]) {
            (void)hipGraphExecDestroy(c->graph[t]);
        }
    }
    for (auto &p : c->pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (hipEvent_t e : c->free_events) {
        (void)hipEventDestroy(e);
    }
    if (c->stream) {
        (void)hipStreamDestroy(c->stream);
    }
    delete c;
    return SGPU_OK;
}

int sgpu_init_state(sgpu_ctx *c)
{
    if (!c) {
        return fail(SGPU_ERR_ARG, "null context");
    }
    const sgpu_model &m = c->m;
    const int N = m.N, L = m.L, K = m.K, A = m.A;
    std::vector<double> kv(K), q((size_t)N * K, 1.0 / K);
    for (int k = 0; k < K; k++) {
        kv[k] = m.alpha0;
    }
    HIP_OK(hipMemcpyAsync(c->alpha, kv.data(), K * sizeof(double), hipMemcpyHostToDevice, c->stream));
    for (int k = 0; k < K; k++) {
        kv[k] = m.lambda0;
    }
    HIP_OK(hipMemcpyAsync(c->lambda, kv.data(), K * sizeof(double), hipMemcpyHostToDevice, c->stream));
    for (int k = 0; k < K; k++) {
        kv[k] = m.freqscorr ? m.fpriormean : 0.5;
    }
    HIP_OK(hipMemcpyAsync(c->F, kv.data(), K * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipMemcpyAsync(c->Q, q.data(), q.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipMemsetAsync(c->P, 0, (size_t)L * K * A * sizeof(double), c->stream));
    HIP_OK(hipMemsetAsync(c->QSum, 0, (size_t)N * K * sizeof(double), c->stream));
    HIP_OK(hipMemsetAsync(c->PSum, 0, (size_t)L * K * A * sizeof(double), c->stream));
    HIP_OK(hipMemsetAsync(c->epsSum, 0, (size_t)L * A * sizeof(double), c->stream));
    HIP_OK(hipMemsetAsync(c->FstSum, 0, K * sizeof(double), c->stream));
    HIP_OK(hipMemsetAsync(c->alphaSum, 0, K * sizeof(double), c->stream));
    HIP_OK(hipMemsetAsync(c->lambdaSum, 0, K * sizeof(double), c->stream));
    HIP_OK(hipMemsetAsync(c->like3, 0, 3 * sizeof(double), c->stream));
    hipLaunchKernelGGL(k_init_eps, dim3(cdiv(L, 256)), dim3(256), 0, c->stream, c->ma, c->fmt,
                       c->geno, c->Lpad, c->nall, m.lambda0, c->eps);
    LAUNCH_OK();
    c->rep_known = false;
    int rc = use_rep(c, 0);
    if (rc || (rc = launch_z(c, 0, 0, nullptr))) {
        return rc;
    }
    if ((rc = launch_logq(c))) {
        return rc;
    }
    c->nsum = 0;
    c->like_ready = false;
    c->initialized = true;
    return SGPU_OK;
}

int sgpu_get_state(sgpu_ctx *c, int32_t which, void *host, size_t bytes)
{
    if (!c || !host) {
        return fail(SGPU_ERR_ARG, "null argument");
    }
    size_t elem;
    const size_t n = state_elems(c, which, &elem);
    if (n == 0) {
        return fail(SGPU_ERR_ARG, "unknown state selector %d", which);
    }
    if (bytes != n * elem) {
        return fail(SGPU_ERR_ARG, "state %d needs %zu bytes, got %zu", which, n * elem, bytes);
    }
    if (which == SGPU_NSUM) {
        HIP_OK(hipStreamSynchronize(c->stream));
        memcpy(host, &c->nsum, sizeof(int64_t));
        return SGPU_OK;
    }
    const void *src = state_ptr(c, which);
    if (which == SGPU_NA || which == SGPU_NQ) {
        const bool isA = which == SGPU_NA;
        int32_t *dst = isA ? c->nA : c->nQ;
        hipLaunchKernelGGL(k_reduce_counts, dim3(cdiv(n, 256)), dim3(256), 0, c->stream,
                           isA ? c->nApart : c->nQpart, isA ? c->n_it : c->n_lt, n, dst);
        LAUNCH_OK();
        src = dst;
    }
    HIP_OK(hipMemcpyAsync(host, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return SGPU_OK;
}

int sgpu_set_state(sgpu_ctx *c, int32_t which, const void *host, size_t bytes)
{
    if (!c || !host) {
        return fail(SGPU_ERR_ARG, "null argument");
    }
    size_t elem;
    const size_t n = state_elems(c, which, &elem);
    if (n == 0) {
        return fail(SGPU_ERR_ARG, "unknown state selector %d", which);
    }
    if (bytes != n * elem) {
        return fail(SGPU_ERR_ARG, "state %d needs %zu bytes, got %zu", which, n * elem, bytes);
    }
    if (which == SGPU_NSUM) {
        memcpy(&c->nsum, host, sizeof(int64_t));
        return SGPU_OK;
    }
    if (which == SGPU_NA || which == SGPU_NQ) {
        const bool isA = which == SGPU_NA;
        int32_t *part = isA ? c->nApart : c->nQpart;
        const int nparts = isA ? c->n_it : c->n_lt;
        HIP_OK(hipMemsetAsync(part, 0, (size_t)nparts * n * sizeof(int32_t), c->stream));
        HIP_OK(hipMemcpyAsync(part, host, bytes, hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        return SGPU_OK;
    }
    HIP_OK(hipMemcpyAsync(state_ptr(c, which), host, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    if (which == SGPU_Q) {
        return launch_logq(c);
    }
    return SGPU_OK;
}

int sgpu_update_p(sgpu_ctx *c, uint32_t rep, int32_t collect)
{
    int rc = check_ready(c);
    if (rc || (rc = use_rep(c, rep))) {
        return rc;
    }
    PhaseTimer timer(c, SGPU_PH_P);
    const size_t nthreads = (size_t)c->m.L * c->m.K;
    if (c->m.A <= 8) {
        hipLaunchKernelGGL(k_update_p<8>, dim3(cdiv(nthreads, kNT)), dim3(kNT), 0, c->stream, c->ma,
                           c->nall, c->nApart, c->n_it, c->eps, c->F, c->lambda, c->P, c->PSum,
                           (int)collect, (const uint32_t *)c->repd);
    } else {
        hipLaunchKernelGGL(k_update_p<0>, dim3(cdiv(nthreads, kNT)), dim3(kNT), 0, c->stream, c->ma,
                           c->nall, c->nApart, c->n_it, c->eps, c->F, c->lambda, c->P, c->PSum,
                           (int)collect, (const uint32_t *)c->repd);
    }
    LAUNCH_OK();
    return SGPU_OK;
}

int sgpu_update_q(sgpu_ctx *c, uint32_t rep, int32_t collect)
{
    int rc = check_ready(c);
    if (rc || (rc = use_rep(c, rep))) {
        return rc;
    }
    PhaseTimer timer(c, SGPU_PH_Q);
    const bool metro = c->m.metrofreq > 0 && rep % (uint32_t)c->m.metrofreq == 0;
    if (metro) {
        dim3 grid(c->q_blocks, c->n_ranges);
        if (c->fmt == SGPU_GENO_SNP2) {
            mkern_t k = metro_snp2_kernel(c->m.K);
            if (!k) {
                return fail(SGPU_ERR_UNSUPPORTED, "no SNP2 Metropolis kernel for K=%d", c->m.K);
            }
            hipLaunchKernelGGL(k, grid, dim3(kNT), 0, c->stream, c->ma, c->geno, c->Lpad, c->Q,
                               c->P, c->alpha, c->loci_per_range, c->delta_part, (const uint32_t *)c->repd);
        } else {
            hipLaunchKernelGGL(k_metro_ratio_byte, grid, dim3(kNT), 0, c->stream, c->ma, c->geno,
                               c->Lpad, c->Q, c->P, c->alpha, c->loci_per_range, c->delta_part, (const uint32_t *)c->repd);
        }
        LAUNCH_OK();
        hipLaunchKernelGGL(k_metro_accept, dim3(c->q_blocks), dim3(kNT), 0, c->stream, c->ma,
                           c->n_ranges, c->delta_part, c->alpha, c->Q, c->QSum, (int)collect,
                           c->logq_part, (int32_t *)nullptr, (const uint32_t *)c->repd);
        c->logq_rows = c->q_blocks;
    } else {
        const size_t nk = (size_t)c->m.N * c->m.K;
        hipLaunchKernelGGL(k_fold_rows, dim3(cdiv(nk, kNT), kRowSplit), dim3(kNT), 0, c->stream,
                           c->nQpart, c->n_lt, nk, c->nQfold);
        LAUNCH_OK();
        hipLaunchKernelGGL(k_update_q_gibbs, dim3(cdiv(c->m.N, kGibbsInds)), dim3(kGibbsInds), 0,
                           c->stream, c->ma, c->nQfold, c->alpha, c->Q, c->QSum, (int)collect,
                           c->logq_part, (const uint32_t *)c->repd);
        c->logq_rows = (int)cdiv(c->m.N, kGibbsInds);
    }
    LAUNCH_OK();
    return SGPU_OK;
}

int sgpu_update_z(sgpu_ctx *c, uint32_t rep, int32_t flags, int8_t *zout)
{
    int rc = check_ready(c);
    if (rc || (rc = use_rep(c, rep))) {
        return rc;
    }
    PhaseTimer timer(c, SGPU_PH_Z);
    int8_t *zdev = nullptr;
    const size_t zn = (size_t)c->m.N * 2 * c->m.L;
    if (zout) {
        if (!c->zdbg) {
            HIP_OK(hipMalloc(&c->zdbg, zn));
        }
        zdev = c->zdbg;
    }
    const bool like = (flags & SGPU_Z_LIKE) != 0;
    if ((rc = launch_z(c, rep, like ? 2 : 1, zdev))) {
        return rc;
    }
    c->like_ready = like;
    if (zout) {
        HIP_OK(hipMemcpyAsync(zout, zdev, zn, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
    }
    return SGPU_OK;
}

int sgpu_update_alpha(sgpu_ctx *c, uint32_t rep)
{
    int rc = check_ready(c);
    if (rc || (rc = use_rep(c, rep))) {
        return rc;
    }
    PhaseTimer timer(c, SGPU_PH_ALPHA);
    hipLaunchKernelGGL(k_update_alpha, dim3(1), dim3(kNT), 0, c->stream, c->ma, c->logq_part,
                       c->logq_rows, c->alpha, (const uint32_t *)c->repd);
    LAUNCH_OK();
    return SGPU_OK;
}

int sgpu_update_lambda(sgpu_ctx *c, uint32_t rep)
{
    int rc = check_ready(c);
    if (rc || (rc = use_rep(c, rep))) {
        return rc;
    }
    PhaseTimer timer(c, SGPU_PH_LAMBDA);
    hipLaunchKernelGGL(k_lambda_partial, dim3(c->l_blocks), dim3(kNT), 0, c->stream, c->ma, c->nall,
                       c->P, c->eps, c->lam_part);
    LAUNCH_OK();
    hipLaunchKernelGGL(k_lambda_finalize, dim3(1), dim3(kNT), 0, c->stream, c->ma, c->lam_part,
                       c->l_blocks, c->ahist, c->nhist, c->lambda, (const uint32_t *)c->repd);
    LAUNCH_OK();
    return SGPU_OK;
}

int sgpu_update_eps(sgpu_ctx *c, uint32_t rep, int32_t collect)
{
    int rc = check_ready(c);
    if (rc || (rc = use_rep(c, rep))) {
        return rc;
    }
    PhaseTimer timer(c, SGPU_PH_EPS);
    hipLaunchKernelGGL(k_update_eps, dim3(cdiv(c->m.L, kNT / c->m.K)), dim3(kNT), 0, c->stream,
                       c->ma, c->nall, c->P,
                       c->F, c->lambda, c->eps, c->epsSum, (int)collect, (const uint32_t *)c->repd);
    LAUNCH_OK();
    return SGPU_OK;
}

int sgpu_update_fst(sgpu_ctx *c, uint32_t rep)
{
    int rc = check_ready(c);
    if (rc || (rc = use_rep(c, rep))) {
        return rc;
    }
    PhaseTimer timer(c, SGPU_PH_FST);
    hipLaunchKernelGGL(k_fst_partial, dim3(c->l_blocks), dim3(kNT), 0, c->stream, c->ma, c->nall,
                       c->eps, c->P, c->F, c->fst_part, (const uint32_t *)c->repd);
    LAUNCH_OK();
    hipLaunchKernelGGL(k_fst_finalize, dim3(1), dim3(kNT), 0, c->stream, c->ma, c->fst_part,
                       c->l_blocks, c->nloci_nonzero, c->F, (const uint32_t *)c->repd);
    LAUNCH_OK();
    return SGPU_OK;
}

int sgpu_collect(sgpu_ctx *c, uint32_t rep)
{
    (void)rep;
    int rc = check_ready(c);
    if (rc) {
        return rc;
    }
    PhaseTimer timer(c, SGPU_PH_COLLECT);
    const int with_like = (c->m.computeprob && c->like_ready) ? 1 : 0;
    hipLaunchKernelGGL(k_collect, dim3(1), dim3(kNT), 0, c->stream, c->ma, c->likepart,
                       c->n_it * c->n_lt, with_like, c->alpha, c->lambda, c->F, c->alphaSum,
                       c->lambdaSum, c->FstSum, c->like3);
    LAUNCH_OK();
    c->nsum++;
    return SGPU_OK;
}

int sgpu_sweep(sgpu_ctx *c, uint32_t rep)
{
    int rc = check_ready(c);
    if (rc) {
        return rc;
    }
    const sgpu_model &m = c->m;
    const int collect = ((int64_t)rep + 1 > m.burnin) ? 1 : 0;
    if ((rc = sgpu_update_p(c, rep, collect))) return rc;
    if ((rc = sgpu_update_q(c, rep, collect))) return rc;
    if ((rc = sgpu_update_z(c, rep, (collect && m.computeprob) ? SGPU_Z_LIKE : 0, nullptr)))
        return rc;
    if (m.inferalpha && (rc = sgpu_update_alpha(c, rep))) return rc;
    if (m.inferlambda && (rc = sgpu_update_lambda(c, rep))) return rc;
    if (m.freqscorr) {
        if ((rc = sgpu_update_eps(c, rep, collect))) return rc;
        if ((rc = sgpu_update_fst(c, rep))) return rc;
    }
    if (collect && (rc = sgpu_collect(c, rep))) return rc;
    return SGPU_OK;
}

int sgpu_run_sweeps(sgpu_ctx *c, uint32_t rep0, uint32_t n)
{
    int rc = check_ready(c);
    if (rc) {
        return rc;
    }
    if (!c->use_graphs || c->timing) {
        for (uint32_t r = 0; r < n; r++) {
            if ((rc = sgpu_sweep(c, rep0 + r))) {
                return rc;
            }
        }
        return SGPU_OK;
    }
    // One graph per sweep type (Metropolis or Gibbs Q, burn-in or collecting),
    // captured on first use; each replay ends by advancing the device index.
    const sgpu_model &m = c->m;
    if ((rc = use_rep(c, rep0))) {
        return rc;
    }
    for (uint32_t r = 0; r < n; r++) {
        const uint32_t rep = rep0 + r;
        const bool metro = m.metrofreq > 0 && rep % (uint32_t)m.metrofreq == 0;
        const bool collect = (int64_t)rep + 1 > m.burnin;
        const int type = (metro ? 1 : 0) | (collect ? 2 : 0);
        if (!c->graph[type]) {
            hipGraph_t g = nullptr;
            HIP_OK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            c->capturing = true;
            rc = sgpu_sweep(c, rep);
            if (rc == SGPU_OK) {
                hipLaunchKernelGGL(k_advance_rep, dim3(1), dim3(1), 0, c->stream, c->repd);
            }
            c->capturing = false;
            hipError_t e = hipStreamEndCapture(c->stream, &g);
            if (rc) {
                if (g) {
                    (void)hipGraphDestroy(g);
                }
                return rc;
            }
            if (e != hipSuccess) {
                return fail(SGPU_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
            }
            e = hipGraphInstantiate(&c->graph[type], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                c->graph[type] = nullptr;
                return fail(SGPU_ERR_DEVICE, "graph instantiate failed: %s", hipGetErrorString(e));
            }
            // the captured sweep counted itself; the replay below counts again
            if (collect) {
                c->nsum--;
            }
        }
        HIP_OK(hipGraphLaunch(c->graph[type], c->stream));
        c->rep_shadow = rep + 1;
        if (collect) {
            c->nsum++;
        }
        c->like_ready = collect && m.computeprob;
        c->logq_rows = metro ? c->q_blocks : (int)cdiv(m.N, kGibbsInds);
    }
    return SGPU_OK;
}

int sgpu_set_timing(sgpu_ctx *c, int32_t enable)
{
    if (!c) {
        return fail(SGPU_ERR_ARG, "null context");
    }
    int rc = drain_timing(c);
    c->timing = enable != 0;
    return rc;
}

int sgpu_get_timing(sgpu_ctx *c, double *ms, int64_t *launches)
{
    if (!c || !ms || !launches) {
        return fail(SGPU_ERR_ARG, "null argument");
    }
    int rc = drain_timing(c);
    if (rc) {
        return rc;
    }
    for (int p = 0; p < SGPU_NPHASE; p++) {
        ms[p] = c->t_ms[p];
        launches[p] = c->t_n[p];
        c->t_ms[p] = 0.0;
        c->t_n[p] = 0;
    }
    return SGPU_OK;
}

int sgpu_synchronize(sgpu_ctx *c)
{
    if (!c) {
        return fail(SGPU_ERR_ARG, "null context");
    }
    HIP_OK(hipStreamSynchronize(c->stream));
    return SGPU_OK;
}

}  // extern "C"
