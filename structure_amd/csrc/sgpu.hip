// C-ABI of the MI355X admixture sweep (include/sgpu.h): context, device
// state, and the per-phase entry points that replace the reference's *CL
// wrappers and its OpenCL dispatch layer (Kernels.c).
//
// A context holds a batch of chains (one chain for sgpu_create).  Each phase
// is one launch per kernel over every chain of the batch: block b of a launch
// reads its chain and its block coordinates from a block map built once at
// creation, and the chain's shapes and state addresses from a device array of
// chain descriptors.  UpdateZ and the Metropolis ratio are compiled per K, so
// they launch once per distinct (genotype format, K) among the chains.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
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

// Default block targets of the tiled phases (16 and 4 blocks per CU on 256 CUs;
// the 2-bit UpdateZ is measured fastest near 4096 at config 2, where 2048
// leaves the second round of blocks on the CUs a third full).
constexpr int kZBlocksDefault = 4096;
// Sweeps per span graph: the whole Metropolis periods that make at least
// kPeriodSweeps (for periods up to kMaxPeriodSweeps: one graph holds a period of
// the reference's default METROFREQ = 20), or kPeriodSweeps Gibbs sweeps when
// there is no Metropolis sweep or the period is longer.
constexpr int kPeriodSweeps = 16;
constexpr int kMaxPeriodSweeps = 64;
// Metropolis-ratio blocks per launch over the batch: at N=2000, L=100000, K=5, 2048
// blocks (256 locus ranges) took the three Metropolis passes from 275 to 260 us per
// Metropolis sweep against 1024; the batch workloads measured the same either way.
constexpr int kMetroBlocksDefault = 2048;
// LDS per UpdateZ block for the per-individual counts (keeps 3 blocks per CU).
constexpr size_t kZLdsMax = 48 * 1024;
// Most loci one Metropolis-ratio block walks.
constexpr long long kMetroMaxLoci = 2048;
// LDS per byte-format UpdateZ block for the staged frequencies and counts, and
// the most loci staged at once: small enough that five or more blocks share a
// CU (measured on 32 microsatellite data sets with K x A up to 160: chunks of
// up to 60 KiB ran 1.15x slower than chunks of at most 16 loci and 32 KiB).
constexpr size_t kByteLdsMax = 32 * 1024;
constexpr int kByteChunkMax = 16;

// One launch over a set of chains: the block map and its LDS need.
struct Plan {
    BlockRef *map = nullptr;
    int nblocks = 0;
    size_t lds = 0;
};

// A launch of a K-templated kernel: the chains of one (format, K) and, for
// UpdateZ, one PFROMPOPFLAGONLY setting.
struct KGroup {
    int fmt = 0, K = 0;
    bool pf = false;
    Plan plan;
};

struct Dataset {
    int fmt = 0, N = 0, L = 0, A = 0, Lpad = 0, G = 0, Gpad = 0;
    std::vector<int32_t> nall_host;   // NumAlleles, for the UpdateP work lists
    uint32_t *geno = nullptr;
    int32_t *nall = nullptr, *ahist = nullptr;
    // POPFLAG of the data file (null without one): per individual, and as the
    // per-word-row masks of the 2-bit path (bit 2t: individual 16g+t has POPFLAG 1)
    int32_t *popflag = nullptr;
    uint32_t *pfmask = nullptr;
    int nhist = 0, nloci_nonzero = 0;
};

struct HostChain {
    sgpu_model m{};
    int dataset = 0;
    size_t z_lds = 0;
    ChainDesc d{};   // host copy of the device descriptor
    int32_t *nA = nullptr, *nQ = nullptr;   // reduced counts for state reads
};

}  // namespace

struct sgpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<Dataset> data;
    std::vector<HostChain> chains;
    ChainDesc *cds = nullptr;       // device descriptors [nchains]
    double *lambda0 = nullptr;      // device [nchains], for InitFreqPriors
    // batch-wide cadence (equal for every chain)
    int32_t burnin = 0, metrofreq = 0, computeprob = 0;
    bool any_alpha = false, any_lambda = false, any_freqscorr = false;
    // launch plans
    std::vector<KGroup> z_groups, metro_groups;
    Plan p_update_p, p_fold, p_gibbs, p_logq, p_accept, p_alpha, p_lam_part,
        p_lam_fin, p_eps, p_fst_part, p_fst_fin, p_collect, p_init_eps;
    // the fused launches of a sweep: UpdateP with the lambda decision blocks of
    // F-model chains, in Gibbs sweeps also with the Gibbs Q blocks
    // (p_update_pq_sw), in Metropolis sweeps with the Q proposal blocks
    // (p_update_pm_sw); in Metropolis sweeps the proposal launch keeps only the
    // lambda log-sum blocks of chains without the F model (none: no launch), which
    // in Gibbs sweeps run in a launch of their own (p_lam_part_p); the epsilon launch with the scalar
    // blocks, epsilon + Fst blocks and count-fold blocks; the Fst decision
    // (F-model chains); separate lambda launches for pop-specific lambda with F
    Plan p_update_pq_sw, p_update_pm_sw, p_propose_sw, p_lam_part_p, p_postz, p_tail,
        p_lam_part_x, p_lam_fin_x;
    Plan p_eps_logsum;   // the log-eps sums of F-model lambda chains, outside a sweep
    // nQfold holds the fold of the current nQpart (false after UpdateZ or a host
    // write of the ancestry counts, until a fold runs)
    bool fold_fresh = false;
    // lam_epart holds the log-eps sums of the current eps (false after
    // initialisation or a host write of eps, until k_eps_logsum runs)
    bool lam_fresh = false;
    std::vector<Plan> p_logq_chain;   // per chain, after a host write of Q
    // device sweep index read by every kernel
    uint32_t *repd = nullptr;
    uint32_t rep_shadow = 0;          // host view of *repd
    bool rep_known = false;
    bool capturing = false;
    bool use_graphs = true;
    hipGraphExec_t graph[4] = {nullptr, nullptr, nullptr, nullptr};   // [metro | 2*collect]
    hipGraphExec_t period_graph[2] = {nullptr, nullptr};              // [collect]
    uint32_t period_sweeps = kPeriodSweeps;   // sweeps per period graph
    int64_t nsum = 0;
    bool initialized = false;
    bool like_ready = false;
    bool tail_advances = false;       // the last sweep's closing launch advances *repd
    int byte_rows = kByteRows;        // block size of the byte-format UpdateZ launches
    int8_t *zdbg = nullptr;
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
    a.inferalpha = m.inferalpha;
    a.inferlambda = m.inferlambda;
    a.alphapropsd = m.alphapropsd;
    a.alphamax = m.alphamax;
    a.alphapriora = m.alphapriora;
    a.alphapriorb = m.alphapriorb;
    a.fpriormean = m.fpriormean;
    a.fpriorsd = m.fpriorsd;
    a.invsqrtn = pow((double)m.N, -0.5);
    a.seed = m.seed;
    a.chain = m.chain;
    a.ancestdist = m.ancestdist ? 1 : 0;
    a.numboxes = m.ancestdist ? m.numboxes : 0;
    a.pfonly = m.pfrompopflagonly ? 1 : 0;
    return a;
}

// The dispatcher hands a launch's workgroups to the XCDs in turn (workgroup w
// runs on XCD w % kXcds), and each XCD has an L2 of its own.
constexpr int kXcds = 8;

// Block orders of a plan.  kXFast lists a chain's grid row by row (x fastest).
// kXcdRuns then cuts the list into kXcds contiguous runs and places run x at
// workgroups x, x + kXcds, x + 2 kXcds, ...: each run executes on one XCD in
// list order, so blocks near each other in the list share its L2.
enum PlanOrder { kXFast = 0, kXcdRuns = 2 };

// Appends the blocks of the chains selected by sel to v, chain by chain: dims(chain)
// gives the chain's own grid (nx, ny), listed row by row (x fastest), each block
// with the given role.
template <class Sel, class Dims>
void add_blocks(const sgpu_ctx *c, std::vector<BlockRef> &v, Sel sel, Dims dims,
                uint32_t role = kRoleMain)
{
    for (size_t ci = 0; ci < c->chains.size(); ci++) {
        if (!sel(c->chains[ci])) {
            continue;
        }
        const std::pair<int, int> g = dims(c->chains[ci]);
        for (int by = 0; by < g.second; by++) {
            for (int bx = 0; bx < g.first; bx++) {
                v.push_back(BlockRef{(uint32_t)ci, (uint32_t)bx, (uint32_t)by, role});
            }
        }
    }
}

int upload_plan(sgpu_ctx *c, Plan &p, std::vector<BlockRef> &v, int order);

// Block map over the chains selected by sel; dims(chain) gives the chain's
// own grid (nx, ny).  Blocks are listed chain by chain, in the given order.
template <class Sel, class Dims>
int build_plan(sgpu_ctx *c, Plan &p, Sel sel, Dims dims, int order = kXFast)
{
    std::vector<BlockRef> v;
    add_blocks(c, v, sel, dims);
    return upload_plan(c, p, v, order);
}

int upload_plan(sgpu_ctx *c, Plan &p, std::vector<BlockRef> &v, int order)
{
    if (order & kXcdRuns) {
        const size_t run = v.size() / kXcds;   // the last v.size() % kXcds blocks stay at the end
        std::vector<BlockRef> w(v);
        for (size_t x = 0; x < (size_t)kXcds; x++) {
            for (size_t s = 0; s < run; s++) {
                w[s * kXcds + x] = v[x * run + s];
            }
        }
        v.swap(w);
    }
    p.nblocks = (int)v.size();
    if (v.empty()) {
        return SGPU_OK;
    }
    int rc = dalloc(c, &p.map, v.size());
    if (rc) {
        return rc;
    }
    HIP_OK(hipMemcpyAsync(p.map, v.data(), v.size() * sizeof(BlockRef), hipMemcpyHostToDevice,
                          c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));   // v is freed on return
    return SGPU_OK;
}

auto all_chains = [](const HostChain &) { return true; };

typedef void (*zkern_t)(const ChainDesc *, const BlockRef *, const uint32_t *);

template <int K, bool PF>
zkern_t pick_z_pf(int fmt, int mode)
{
    if (fmt == SGPU_GENO_BYTE) {
        switch (mode) {
        case 0: return k_update_z_byte<K, true, false, false, PF>;
        case 1: return k_update_z_byte<K, false, false, false, PF>;
        case 2: return k_update_z_byte<K, false, true, false, PF>;
        case 3: return k_update_z_byte<K, false, false, true, PF>;
        default: return k_update_z_byte<K, false, true, true, PF>;
        }
    }
    switch (mode) {
    case 0: return k_update_z_snp2<K, true, false, false, PF>;
    case 1: return k_update_z_snp2<K, false, false, false, PF>;
    case 2: return k_update_z_snp2<K, false, true, false, PF>;
    case 3: return k_update_z_snp2<K, false, false, true, PF>;
    default: return k_update_z_snp2<K, false, true, true, PF>;
    }
}

// PF: the variant that filters the allele counts by POPFLAG (PFROMPOPFLAGONLY)
template <int K>
zkern_t pick_z(int fmt, int mode, bool pf)
{
    return pf ? pick_z_pf<K, true>(fmt, mode) : pick_z_pf<K, false>(fmt, mode);
}

// mode 0 = initial Z, 1 = update, 2 = update + lnP terms (+2 with a debug copy of Z)
zkern_t z_kernel(int fmt, int K, int mode, bool pf)
{
    switch (K) {
    case 1: return pick_z<1>(fmt, mode, pf);
    case 2: return pick_z<2>(fmt, mode, pf);
    case 3: return pick_z<3>(fmt, mode, pf);
    case 4: return pick_z<4>(fmt, mode, pf);
    case 5: return pick_z<5>(fmt, mode, pf);
    case 6: return pick_z<6>(fmt, mode, pf);
    case 7: return pick_z<7>(fmt, mode, pf);
    case 8: return pick_z<8>(fmt, mode, pf);
    case 9: return pick_z<9>(fmt, mode, pf);
    case 10: return pick_z<10>(fmt, mode, pf);
    case 11: return pick_z<11>(fmt, mode, pf);
    case 12: return pick_z<12>(fmt, mode, pf);
    case 13: return pick_z<13>(fmt, mode, pf);
    case 14: return pick_z<14>(fmt, mode, pf);
    case 15: return pick_z<15>(fmt, mode, pf);
    case 16: return pick_z<16>(fmt, mode, pf);
    case 17: return pick_z<17>(fmt, mode, pf);
    case 18: return pick_z<18>(fmt, mode, pf);
    case 19: return pick_z<19>(fmt, mode, pf);
    case 20: return pick_z<20>(fmt, mode, pf);
    default: return nullptr;
    }
}

typedef void (*mkern_t)(const ChainDesc *, const BlockRef *, const uint32_t *);

template <int K>
mkern_t pick_metro(int fmt)
{
    return fmt == SGPU_GENO_BYTE ? k_metro_ratio_byte<K> : k_metro_ratio_snp2<K>;
}

mkern_t metro_kernel(int fmt, int K)
{
    switch (K) {
    case 1: return pick_metro<1>(fmt);
    case 2: return pick_metro<2>(fmt);
    case 3: return pick_metro<3>(fmt);
    case 4: return pick_metro<4>(fmt);
    case 5: return pick_metro<5>(fmt);
    case 6: return pick_metro<6>(fmt);
    case 7: return pick_metro<7>(fmt);
    case 8: return pick_metro<8>(fmt);
    case 9: return pick_metro<9>(fmt);
    case 10: return pick_metro<10>(fmt);
    case 11: return pick_metro<11>(fmt);
    case 12: return pick_metro<12>(fmt);
    case 13: return pick_metro<13>(fmt);
    case 14: return pick_metro<14>(fmt);
    case 15: return pick_metro<15>(fmt);
    case 16: return pick_metro<16>(fmt);
    case 17: return pick_metro<17>(fmt);
    case 18: return pick_metro<18>(fmt);
    case 19: return pick_metro<19>(fmt);
    case 20: return pick_metro<20>(fmt);
    default: return nullptr;
    }
}

int launch_z(sgpu_ctx *c, int mode)
{
    c->fold_fresh = false;   // new ancestry-count partials
    for (const KGroup &g : c->z_groups) {
        zkern_t k = z_kernel(g.fmt, g.K, mode, g.pf);
        if (!k) {
            return fail(SGPU_ERR_UNSUPPORTED, "no UpdateZ kernel for K=%d", g.K);
        }
        const int threads = g.fmt == SGPU_GENO_BYTE ? c->byte_rows : kZLoci;
        hipLaunchKernelGGL(k, dim3(g.plan.nblocks), dim3(threads), g.plan.lds, c->stream, c->cds,
                           g.plan.map, (const uint32_t *)c->repd);
        LAUNCH_OK();
    }
    return SGPU_OK;
}

// Launch a batched kernel over a plan (nothing when the plan is empty).
template <typename... Args>
int launch(sgpu_ctx *c, void (*k)(const ChainDesc *, const BlockRef *, Args...), const Plan &p,
           int nthreads, Args... args)
{
    if (p.nblocks == 0) {
        return SGPU_OK;
    }
    hipLaunchKernelGGL(k, dim3(p.nblocks), dim3(nthreads), p.lds, c->stream, c->cds, p.map,
                       args...);
    LAUNCH_OK();
    return SGPU_OK;
}

size_t state_elems(const sgpu_model &m, int which, size_t *elem)
{
    const size_t N = m.N, L = m.L, K = m.K, A = m.A;
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
    case SGPU_ANCESTDIST:
        *elem = sizeof(int32_t);
        return m.ancestdist ? N * K * (size_t)m.numboxes : 0;
    default: return 0;
    }
}

void *state_ptr(const ChainDesc &d, int which)
{
    switch (which) {
    case SGPU_ANCESTDIST: return d.adist;
    case SGPU_Q: return d.Q;
    case SGPU_P: return d.P;
    case SGPU_EPS: return d.eps;
    case SGPU_FST: return d.F;
    case SGPU_ALPHA: return d.alpha;
    case SGPU_LAMBDA: return d.lambda;
    case SGPU_QSUM: return d.QSum;
    case SGPU_PSUM: return d.PSum;
    case SGPU_FSTSUM: return d.FstSum;
    case SGPU_EPSSUM: return d.epsSum;
    case SGPU_ALPHASUM: return d.alphaSum;
    case SGPU_LAMBDASUM: return d.lambdaSum;
    case SGPU_LIKE: return d.like3;
    default: return nullptr;
    }
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
        if (c && c->timing && !c->capturing && (a = take_event(c))) {
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

// The sampler's device special functions over an array (sgpu_eval_math).
__global__ void k_eval_math(int fn, const double *__restrict__ x, double *__restrict__ y, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    const double v = x[i];
    double sn, cs;
    sincos_d(v, &sn, &cs);
    y[i] = fn == SGPU_FN_LOG ? log_d(v)
                             : (fn == SGPU_FN_COS ? cs : (fn == SGPU_FN_SIN ? sn : lgamma_lanczos(v)));
}

}  // namespace

// Device staging of sgpu_eval_math_noalloc (x, then y), allocated with the code
// object: each GPU holds its own copy until the process ends.  Host code never
// uses this name as an address, only the pointer hipGetSymbolAddress returns
// for the current device (g_stage_ptr).
static __device__ __align__(256) double scratch[2 * SGPU_EVAL_MATH_NOALLOC_MAX];

namespace {

constexpr int kStageDevices = 64;
std::mutex g_stage_mu;
double *g_stage_ptr[kStageDevices] = {};   // address of `scratch` on each device

struct DevBuf {
    void *p = nullptr;
    ~DevBuf()
    {
        if (p) {
            (void)hipFree(p);
        }
    }
};

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

int check_model(const sgpu_model *m, int ci)
{
    if (m->N < 1 || m->L < 1 || m->K < 1 || m->A < 1) {
        return fail(SGPU_ERR_ARG, "chain %d: bad shape N=%d L=%d K=%d A=%d", ci, m->N, m->L, m->K,
                    m->A);
    }
    if (m->K > kMaxK) {
        return fail(SGPU_ERR_UNSUPPORTED, "chain %d: K=%d exceeds the supported maximum %d", ci,
                    m->K, kMaxK);
    }
    if (m->A > kMaxAlleles) {
        return fail(SGPU_ERR_UNSUPPORTED, "chain %d: MAXALLELES=%d exceeds %d", ci, m->A,
                    kMaxAlleles);
    }
    if (m->freqscorr && (m->fpriormean <= 0.0 || m->fpriorsd <= 0.0)) {
        return fail(SGPU_ERR_ARG, "chain %d: FPRIORMEAN and FPRIORSD must be positive", ci);
    }
    if (m->ancestdist && (m->numboxes < 1 || (size_t)m->N * m->K * m->numboxes > ((size_t)1 << 34))) {
        return fail(SGPU_ERR_ARG, "chain %d: NUMBOXES=%d outside [1, 2^34 / (N K)]", ci,
                    m->numboxes);
    }
    // the kernels index [L][K][A] and [N][K] with 32-bit products
    if ((size_t)m->L * m->K * m->A >= ((size_t)1 << 31) || (size_t)m->N * m->K >= ((size_t)1 << 31)) {
        return fail(SGPU_ERR_UNSUPPORTED, "chain %d: L*K*A or N*K reaches 2^31", ci);
    }
    return SGPU_OK;
}

// Every allele index must lie below NumAlleles of its locus (the kernels index
// P and the count rows with it).
int check_data(const sgpu_data *d, int N, int L, int A, int di)
{
    if (!d->geno || !d->nall) {
        return fail(SGPU_ERR_ARG, "data set %d: null genotype or NumAlleles pointer", di);
    }
    if (d->format != SGPU_GENO_SNP2 && d->format != SGPU_GENO_BYTE) {
        return fail(SGPU_ERR_ARG, "data set %d: unknown genotype format %d", di, d->format);
    }
    if (d->format == SGPU_GENO_SNP2 && A != 2) {
        return fail(SGPU_ERR_ARG, "data set %d: SNP2 genotypes need MAXALLELES == 2 (got %d)", di,
                    A);
    }
    const size_t need = sgpu_packed_words(d->format, N, L);
    if (d->geno_words != need) {
        return fail(SGPU_ERR_ARG, "data set %d: genotype buffer has %zu words, format needs %zu",
                    di, d->geno_words, need);
    }
    for (int l = 0; l < L; l++) {
        if (d->nall[l] < 0 || d->nall[l] > A) {
            return fail(SGPU_ERR_ARG, "data set %d: NumAlleles[%d]=%d outside [0, %d]", di, l,
                        d->nall[l], A);
        }
    }
    // 2-bit words are [rows][Lpad], byte words [L][Gpad]: (row, l) at
    // row * row_stride + l * locus_stride
    const size_t rows = d->format == SGPU_GENO_SNP2 ? cdiv(N, 16) : need / L;
    const size_t row_stride = d->format == SGPU_GENO_SNP2 ? need / rows : 1;
    const size_t locus_stride = d->format == SGPU_GENO_SNP2 ? 1 : rows;
    for (size_t row = 0; row < rows; row++) {
        for (int l = 0; l < L; l++) {
            const uint32_t x = d->geno[row * row_stride + (size_t)l * locus_stride];
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
                        return fail(SGPU_ERR_ARG, "allele %u at locus %d exceeds NumAlleles=%d", a,
                                    l, d->nall[l]);
                    }
                }
            }
        }
    }
    return SGPU_OK;
}

int upload_dataset(sgpu_ctx *c, Dataset &ds, const sgpu_data *d)
{
    const size_t need = sgpu_packed_words(d->format, ds.N, ds.L);
    int rc;
    if ((rc = dalloc(c, &ds.geno, need)) || (rc = dalloc(c, &ds.nall, ds.L)) ||
        (rc = dalloc(c, &ds.ahist, 2 * (ds.A + 1)))) {
        return rc;
    }
    // allele-count histogram over loci with more than one allele (lambda update)
    std::map<int, int> hist;
    int nzero = 0;
    for (int l = 0; l < ds.L; l++) {
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
    ds.nhist = (int)hist.size();
    ds.nloci_nonzero = ds.L - nzero;
    ds.nall_host.assign(d->nall, d->nall + ds.L);
    if (d->popflag) {
        const int G2 = (int)cdiv(ds.N, 16);
        std::vector<uint32_t> masks((size_t)G2, 0u);
        for (int i = 0; i < ds.N; i++) {
            if (d->popflag[i] == 1) {
                masks[(size_t)(i / 16)] |= 1u << (2 * (i % 16));
            }
        }
        if ((rc = dalloc(c, &ds.popflag, ds.N)) || (rc = dalloc(c, &ds.pfmask, masks.size()))) {
            return rc;
        }
        HIP_OK(hipMemcpyAsync(ds.popflag, d->popflag, (size_t)ds.N * sizeof(int32_t),
                              hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipMemcpyAsync(ds.pfmask, masks.data(), masks.size() * sizeof(uint32_t),
                              hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));   // masks is freed on return
    }
    HIP_OK(hipMemcpyAsync(ds.geno, d->geno, need * sizeof(uint32_t), hipMemcpyHostToDevice,
                          c->stream));
    HIP_OK(hipMemcpyAsync(ds.nall, d->nall, (size_t)ds.L * sizeof(int32_t), hipMemcpyHostToDevice,
                          c->stream));
    if (!hv.empty()) {
        HIP_OK(hipMemcpyAsync(ds.ahist, hv.data(), hv.size() * sizeof(int32_t),
                              hipMemcpyHostToDevice, c->stream));
    }
    HIP_OK(hipStreamSynchronize(c->stream));
    return SGPU_OK;
}

// Tiles of the chain's UpdateZ and Metropolis launches for the batch's block
// targets, then its device state.
// Tile sizes shared by every chain of a batch, so that blocks of one launch
// carry about the same work whatever the chain: word rows per 2-bit UpdateZ
// block and loci per Metropolis block, each set so that the whole batch comes
// to about the target number of blocks.
struct BatchTiles {
    int snp_rows;   // 16-individual word rows per 2-bit UpdateZ block (before the LDS cap)
    int metro_lpr;  // loci per Metropolis block (multiple of 4)
    int byte_loci;  // loci per byte-format UpdateZ block (0 = byte_zt)
    int byte_chunk; // loci staged in LDS at once by the byte path (0 = as many as fit)
    int byte_rows;  // word rows (individual pairs) per byte-format UpdateZ block: 64, 128, 256
    int byte_zt;    // default loci per byte-format UpdateZ block
};

BatchTiles batch_tiles(const sgpu_model *models, const sgpu_data *data,
                       const int32_t *data_of_chain, int nchains, int z_target,
                       int metro_target)
{
    long long snp_work = 0, metro_work = 0;   // in (row x locus tile) and (256 inds x locus)
    for (int ci = 0; ci < nchains; ci++) {
        const sgpu_model &m = models[ci];
        if (data[data_of_chain[ci]].format == SGPU_GENO_SNP2) {
            snp_work += (long long)cdiv(m.N, 16) * (long long)cdiv(m.L, kZLoci);
        }
        metro_work += (long long)cdiv(m.N, kNT) * m.L;
    }
    BatchTiles t{};
    t.snp_rows = (int)std::max<long long>(1, snp_work / std::max(z_target, 1));
    // byte format: the largest block (word rows) that launches at most 15% more
    // lanes than 64-row blocks would, then the longest locus tile (<= 64) that still
    // gives z_target blocks; failing that, the shortest tile whose per-tile
    // ancestry-count partials (N K int32 per tile) stay within half of the
    // genotype stream (2 N L bytes), i.e. at least 4 K loci
    auto lanes = [&](int R) {
        long long s = 0;
        for (int ci = 0; ci < nchains; ci++) {
            if (data[data_of_chain[ci]].format == SGPU_GENO_BYTE) {
                s += (long long)cdiv(cdiv(models[ci].N, 2), R) * R * models[ci].L;
            }
        }
        return s;
    };
    t.byte_rows = 64;
    for (int R : {256, 128}) {
        if (lanes(R) * 100 <= lanes(64) * 115) {
            t.byte_rows = R;
            break;
        }
    }
    auto blocks = [&](int zt) {
        long long s = 0;
        for (int ci = 0; ci < nchains; ci++) {
            if (data[data_of_chain[ci]].format == SGPU_GENO_BYTE) {
                s += (long long)cdiv(models[ci].L, zt) * cdiv(cdiv(models[ci].N, 2), t.byte_rows);
            }
        }
        return s;
    };
    int kmax = 1;
    for (int ci = 0; ci < nchains; ci++) {
        if (data[data_of_chain[ci]].format == SGPU_GENO_BYTE) {
            kmax = std::max(kmax, models[ci].K);
        }
    }
    t.byte_zt = 0;
    for (int zt : {64, 32, 16}) {
        if (blocks(zt) >= z_target) {
            t.byte_zt = zt;
            break;
        }
    }
    if (t.byte_zt == 0) {
        t.byte_zt = 64;
        while (t.byte_zt > 16 && t.byte_zt / 2 >= 4 * kmax) {
            t.byte_zt /= 2;
        }
    }
    // at most kMetroMaxLoci loci per block, so a large batch keeps many blocks
    const long long lpr = std::min<long long>(
        kMetroMaxLoci,
        std::max<long long>(4, (metro_work + metro_target - 1) / std::max(metro_target, 1)));
    t.metro_lpr = (int)((lpr + 3) / 4 * 4);
    return t;
}

// UpdateP work list of a chain: the dense grid gives every (locus, pop) group A
// threads and leaves a - nall[l] of them idle, which with microsatellites
// (A the largest allele count) is most of a launch.  The list gives a group
// exactly nall[l] threads, groups packed whole into blocks of kNT.  Used when it
// saves at least an eighth of the threads.
int build_p_slots(sgpu_ctx *c, HostChain &h, const Dataset &ds)
{
    ChainDesc &d = h.d;
    const int L = h.m.L, K = h.m.K, A = h.m.A;
    const int gpb = kNT / A;
    d.p_blocks = (int)cdiv((size_t)L * K, (size_t)gpb);
    d.pslots = nullptr;
    if (L >= (1 << 19) || A > 255 || K > 31) {
        return SGPU_OK;   // the fields of a list entry
    }
    std::vector<uint32_t> v;
    int fill = 0;
    for (int l = 0; l < L; l++) {
        const int na = ds.nall_host[l];
        for (int k = 0; k < K && na > 0; k++) {
            if (fill + na > kNT) {
                v.insert(v.end(), (size_t)(kNT - fill), 0xFFFFFFFFu);
                fill = 0;
            }
            for (int a = 0; a < na; a++) {
                v.push_back((uint32_t)l << 13 | (uint32_t)k << 8 | (uint32_t)a);
            }
            fill += na;
        }
    }
    if (v.empty() || fill > 0) {
        v.insert(v.end(), (size_t)(kNT - fill), 0xFFFFFFFFu);
    }
    const size_t blocks = v.size() / kNT;
    if (blocks * 8 > (size_t)d.p_blocks * 7) {
        return SGPU_OK;   // the dense grid is nearly as small
    }
    uint32_t *dev = nullptr;
    int rc = dalloc(c, &dev, v.size());
    if (rc) {
        return rc;
    }
    HIP_OK(hipMemcpyAsync(dev, v.data(), v.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                          c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));   // v is freed on return
    d.pslots = dev;
    d.p_blocks = (int)blocks;
    return SGPU_OK;
}

int setup_chain(sgpu_ctx *c, HostChain &h, const Dataset &ds, const BatchTiles &bt)
{
    const sgpu_model &m = h.m;
    ChainDesc &d = h.d;
    const int N = m.N, L = m.L, K = m.K, A = m.A;
    d.m = model_args(m);
    d.fmt = ds.fmt;
    d.Lpad = ds.Lpad;
    d.G = ds.G;
    d.Gpad = ds.Gpad;
    d.geno = ds.geno;
    d.nall = ds.nall;
    d.ahist = ds.ahist;
    d.nhist = ds.nhist;
    d.nloci_nonzero = ds.nloci_nonzero;
    if (ds.fmt == SGPU_GENO_SNP2) {
        // one locus per lane, 256 loci per block; individual tiles sized so that
        // one launch over the batch has about z_target blocks, LDS counts <= 48 KiB
        d.zt_loci = kZLoci;
        d.z_chunk = 0;
        d.n_lt = (int)cdiv(L, kZLoci);
        int rows = std::min(bt.snp_rows, d.G);
        // LDS counts <= kZLdsMax; that also keeps a tile's copies (rows x 32) within
        // the 16-bit count partials
        while (((size_t)rows * 16 * K * 4 > kZLdsMax || rows * 32 > 65535) && rows > 1) {
            rows = (rows + 1) / 2;
        }
        d.n_it = (int)cdiv(d.G, rows);
        d.rows_per_tile = (int)cdiv(d.G, d.n_it);   // even tiles
        h.z_lds = (size_t)rows * 8 * K * sizeof(uint32_t);   // counts of individual pairs
    } else {
        // one individual pair per lane, byte_rows (64, 128 or 256) pairs per block,
        // locus tiles from batch_tiles; a chunk of loci's K x A frequencies (FP64)
        // and counts (u32) sit in LDS
        d.zt_loci = bt.byte_loci > 0 ? std::min(bt.byte_loci, kZLoci) : bt.byte_zt;
        d.n_lt = (int)cdiv(L, d.zt_loci);
        d.rows_per_tile = bt.byte_rows;
        d.n_it = (int)cdiv(d.G, bt.byte_rows);
        const size_t per_locus = (size_t)K * A * (sizeof(double) + sizeof(uint32_t));
        size_t chunk = std::max<size_t>(
            1, std::min<size_t>({kByteLdsMax / per_locus, (size_t)d.zt_loci, (size_t)kByteChunkMax}));
        if (bt.byte_chunk > 0) {
            chunk = std::min<size_t>(chunk, (size_t)bt.byte_chunk);
        }
        d.z_chunk = (int)chunk;
        h.z_lds = (size_t)d.z_chunk * per_locus;
    }
    // the Gibbs update sums up to kQDirectRows ancestry partials itself; more are
    // first folded to kRowSplit rows by their own pass
    d.q_direct = d.n_lt <= kQDirectRows ? 1 : 0;
    d.q_blocks = (int)cdiv(N, kNT);
    d.l_blocks = (int)cdiv(L, kNT);
    d.e_blocks = (int)cdiv(L, kNT / K);
    {
        const int n_ranges = (int)cdiv(L, std::min<size_t>(bt.metro_lpr, cdiv(L, 4) * 4));
        d.loci_per_range = (int)(cdiv(cdiv(L, n_ranges), 4) * 4);   // even, multiple of 4 loci
        d.n_ranges = (int)cdiv(L, d.loci_per_range);
    }
    d.zout = nullptr;
    d.adist = nullptr;
    d.pfmask = nullptr;
    d.popflag = nullptr;
    if (m.pfrompopflagonly) {
        if (!ds.popflag) {
            // CheckIfMissing's rule (Util.c:75-77)
            return fail(SGPU_ERR_ARG, "PFROMPOPFLAGONLY can only be turned on when the data "
                                      "set has POPFLAG data");
        }
        if (ds.fmt == SGPU_GENO_SNP2) {
            d.pfmask = ds.pfmask;
        } else {
            d.popflag = ds.popflag;
        }
    }

    int rc;
    if (m.ancestdist && (rc = dalloc(c, &d.adist, (size_t)N * K * m.numboxes))) {
        return rc;
    }
    if ((rc = dalloc(c, &d.Q, (size_t)N * K)) || (rc = dalloc(c, &d.P, (size_t)L * K * A)) ||
        (rc = dalloc(c, &d.eps, (size_t)L * A)) || (rc = dalloc(c, &d.F, K)) ||
        (rc = dalloc(c, &d.alpha, K)) || (rc = dalloc(c, &d.lambda, K)) ||
        (rc = dalloc(c, &d.nApart, (size_t)d.n_it * L * K * A)) ||
        (rc = dalloc(c, &d.nQpart, (size_t)d.n_lt * N * K)) ||
        (rc = dalloc(c, &h.nA, (size_t)L * K * A)) || (rc = dalloc(c, &h.nQ, (size_t)N * K)) ||
        (rc = dalloc(c, &d.nQfold, (size_t)kRowSplit * N * K)) ||
        (rc = dalloc(c, &d.logq_rows, 1)) ||
        (rc = dalloc(c, &d.likepart, (size_t)d.n_it * d.n_lt)) ||
        (rc = dalloc(c, &d.logq_part,
                     std::max(cdiv(N, kNT / K), cdiv(N, kAcceptInds)) * (K + 1))) ||
        (rc = dalloc(c, &d.delta_part, (size_t)d.n_ranges * N)) ||
        (rc = dalloc(c, &d.qprop, (size_t)N * K)) ||
        (rc = dalloc(c, &d.fst_part, (size_t)std::max(d.l_blocks, d.e_blocks) * K)) ||
        (rc = dalloc(c, &d.QSum, (size_t)N * K)) || (rc = dalloc(c, &d.PSum, (size_t)L * K * A)) ||
        (rc = dalloc(c, &d.epsSum, (size_t)L * A)) || (rc = dalloc(c, &d.FstSum, K)) ||
        (rc = dalloc(c, &d.alphaSum, K)) || (rc = dalloc(c, &d.lambdaSum, K)) ||
        (rc = dalloc(c, &d.like3, 3))) {
        return rc;
    }
    d.lam_mode = !m.inferlambda ? kLamNone
                 : !m.freqscorr    ? kLamP
                 : (m.popspecificlambda ? kLamSplit : kLamEps);
    if ((rc = build_p_slots(c, h, ds))) {
        return rc;
    }
    if (d.lam_mode == kLamEps && (rc = dalloc(c, &d.lam_epart, (size_t)d.e_blocks))) {
        return rc;
    }
    return dalloc(c, &d.lam_part, (size_t)d.l_blocks * K);
}

int upload_descriptors(sgpu_ctx *c)
{
    std::vector<ChainDesc> v;
    for (const HostChain &h : c->chains) {
        v.push_back(h.d);
    }
    HIP_OK(hipMemcpyAsync(c->cds, v.data(), v.size() * sizeof(ChainDesc), hipMemcpyHostToDevice,
                          c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return SGPU_OK;
}

int build_plans(sgpu_ctx *c)
{
    int rc;
    auto one = [](const HostChain &) { return std::make_pair(1, 1); };
    auto qb = [](const HostChain &h) { return std::make_pair(h.d.q_blocks, 1); };
    auto lb = [](const HostChain &h) { return std::make_pair(h.d.l_blocks, 1); };
    if ((rc = build_plan(c, c->p_update_p, all_chains,
                         [](const HostChain &h) { return std::make_pair(h.d.p_blocks, 1); })) ||
        (rc = build_plan(c, c->p_fold, [](const HostChain &h) { return h.d.q_direct == 0; },
                         [](const HostChain &h) {
                             return std::make_pair((int)cdiv((size_t)h.m.N * h.m.K, kNT), kRowSplit);
                         })) ||
        (rc = build_plan(c, c->p_gibbs, all_chains,
                         [](const HostChain &h) { return std::make_pair((int)cdiv(h.m.N, kNT / h.m.K), 1); })) ||
        (rc = build_plan(c, c->p_logq, all_chains, qb)) ||
        (rc = build_plan(c, c->p_accept, all_chains,
                         [](const HostChain &h) {
                             return std::make_pair((int)cdiv(h.m.N, kAcceptInds), 1);
                         })) ||
        (rc = build_plan(c, c->p_alpha, [](const HostChain &h) { return h.m.inferalpha != 0; }, one)) ||
        (rc = build_plan(c, c->p_lam_part, [](const HostChain &h) { return h.m.inferlambda != 0; }, lb)) ||
        (rc = build_plan(c, c->p_lam_fin, [](const HostChain &h) { return h.m.inferlambda != 0; }, one)) ||
        (rc = build_plan(c, c->p_eps, [](const HostChain &h) { return h.m.freqscorr != 0; },
                         [](const HostChain &h) { return std::make_pair(h.d.e_blocks, 1); })) ||
        (rc = build_plan(c, c->p_fst_part, [](const HostChain &h) { return h.m.freqscorr != 0; },
                         [](const HostChain &h) {
                             return std::make_pair(h.d.l_blocks, h.m.onefst ? 1 : h.m.K);
                         })) ||
        (rc = build_plan(c, c->p_fst_fin, [](const HostChain &h) { return h.m.freqscorr != 0; }, one)) ||
        (rc = build_plan(c, c->p_collect, all_chains, one)) ||
        (rc = build_plan(c, c->p_tail, [](const HostChain &h) { return h.m.freqscorr != 0; }, one)) ||
        (rc = build_plan(c, c->p_init_eps, all_chains, [](const HostChain &h) {
             return std::make_pair((int)cdiv(h.m.L, kNT / 64), 1);   // a wavefront per locus
         }))) {
        return rc;
    }
    {
        // sweep launches mixing roles; the single blocks of a chain come first so
        // that they start with the launch
        auto gibbs_dims = [](const HostChain &h) {
            return std::make_pair((int)cdiv(h.m.N, kNT / h.m.K), 1);
        };
        auto mode = [](int lm) { return [lm](const HostChain &h) { return h.d.lam_mode == lm; }; };
        auto p_dims = [](const HostChain &h) { return std::make_pair(h.d.p_blocks, 1); };
        std::vector<BlockRef> v;
        add_blocks(c, v, mode(kLamEps), one, kRoleLambda);
        add_blocks(c, v, all_chains, gibbs_dims, kRoleQGibbs);
        add_blocks(c, v, all_chains, p_dims);
        if ((rc = upload_plan(c, c->p_update_pq_sw, v, kXFast))) {
            return rc;
        }
        v.clear();
        add_blocks(c, v, mode(kLamEps), one, kRoleLambda);
        add_blocks(c, v, all_chains, gibbs_dims, kRoleMetroProp);
        add_blocks(c, v, all_chains, p_dims);
        if ((rc = upload_plan(c, c->p_update_pm_sw, v, kXFast))) {
            return rc;
        }
        v.clear();
        add_blocks(c, v, mode(kLamP), lb, kRoleLamPart);
        if ((rc = upload_plan(c, c->p_propose_sw, v, kXFast))) {
            return rc;
        }
        if ((rc = build_plan(c, c->p_lam_part_p, mode(kLamP), lb)) ||
            (rc = build_plan(c, c->p_eps_logsum, mode(kLamEps),
                             [](const HostChain &h) { return std::make_pair(h.d.e_blocks, 1); })) ||
            (rc = build_plan(c, c->p_lam_part_x, mode(kLamSplit), lb)) ||
            (rc = build_plan(c, c->p_lam_fin_x, mode(kLamSplit), one))) {
            return rc;
        }
        v.clear();
        add_blocks(c, v, all_chains, one, kRoleScalars);
        add_blocks(c, v, [](const HostChain &h) { return h.m.freqscorr != 0; },
                   [](const HostChain &h) { return std::make_pair(h.d.e_blocks, 1); });
        add_blocks(c, v, [](const HostChain &h) { return h.d.q_direct == 0; },
                   [](const HostChain &h) {
                       return std::make_pair((int)cdiv((size_t)h.m.N * h.m.K, kNT), kRowSplit);
                   },
                   kRoleFold);
        if ((rc = upload_plan(c, c->p_postz, v, kXFast))) {
            return rc;
        }
    }
    c->p_logq_chain.resize(c->chains.size());
    for (size_t ci = 0; ci < c->chains.size(); ci++) {
        const HostChain *target = &c->chains[ci];
        if ((rc = build_plan(c, c->p_logq_chain[ci],
                             [target](const HostChain &h) { return &h == target; }, qb))) {
            return rc;
        }
    }
    // K-templated launches, heavier K first: UpdateZ once per (format, K,
    // PFROMPOPFLAGONLY), the Metropolis ratio once per (format, K)
    std::map<std::tuple<int, int, int>, int> zkeys;
    std::map<std::pair<int, int>, int> mkeys;
    for (const HostChain &h : c->chains) {
        zkeys[{-h.d.fmt, -h.m.K, h.m.pfrompopflagonly ? 1 : 0}] = 1;
        mkeys[{-h.d.fmt, -h.m.K}] = 1;
    }
    for (auto &kv : zkeys) {
        const int fmt = -std::get<0>(kv.first), K = -std::get<1>(kv.first);
        const bool pf = std::get<2>(kv.first) != 0;
        auto sel = [fmt, K, pf](const HostChain &h) {
            return h.d.fmt == fmt && h.m.K == K && (h.m.pfrompopflagonly != 0) == pf;
        };
        KGroup zg;
        zg.fmt = fmt;
        zg.K = K;
        zg.pf = pf;
        // x fastest, so the blocks running at one time share their individual
        // tile.  Orders that keep a locus tile's blocks on one XCD (y fastest, or
        // XCD bands of locus tiles) fetch each frequency tile once per XCD
        // (L2 fetches 148 -> 59 MB per launch) but measure 9% slower at N=2000,
        // L=100000, K=5, where the kernel is bound by instruction issue.
        if ((rc = build_plan(c, zg.plan, sel,
                             [](const HostChain &h) { return std::make_pair(h.d.n_lt, h.d.n_it); }))) {
            return rc;
        }
        for (const HostChain &h : c->chains) {
            if (sel(h)) {
                zg.plan.lds = std::max(zg.plan.lds, h.z_lds);
            }
        }
        c->z_groups.push_back(zg);
    }
    for (auto &kv : mkeys) {
        const int fmt = -kv.first.first, K = -kv.first.second;
        auto sel = [fmt, K](const HostChain &h) { return h.d.fmt == fmt && h.m.K == K; };
        KGroup mg;
        mg.fmt = fmt;
        mg.K = K;
        // the blocks of one locus range read the same frequencies: one XCD runs them
        if ((rc = build_plan(c, mg.plan, sel,
                             [](const HostChain &h) { return std::make_pair(h.d.q_blocks, h.d.n_ranges); },
                             kXFast | kXcdRuns))) {
            return rc;
        }
        c->metro_groups.push_back(mg);
    }
    return SGPU_OK;
}

int chain_index(sgpu_ctx *c, int32_t chain)
{
    if (!c) {
        return fail(SGPU_ERR_ARG, "null context");
    }
    if (chain < 0 || (size_t)chain >= c->chains.size()) {
        return fail(SGPU_ERR_ARG, "chain %d outside [0, %zu)", chain, c->chains.size());
    }
    return SGPU_OK;
}

}  // namespace

extern "C" {

const char *sgpu_last_error(void) { return g_err; }

const char *sgpu_version(void)
{
    return "sgpu 0.3 gfx950: chain-batched launches, fused UpdateZ (SNP2 and byte formats, K<=20), "
           "FP64 draws";
}

int sgpu_create(const sgpu_model *m, const sgpu_data *d, sgpu_ctx **out)
{
    const int32_t zero = 0;
    return sgpu_create_batch(m, 1, d, 1, &zero, nullptr, out);
}

int sgpu_create_batch(const sgpu_model *models, int32_t nchains, const sgpu_data *data,
                      int32_t ndata, const int32_t *data_of_chain, const sgpu_tuning *tuning,
                      sgpu_ctx **out)
{
    if (!out) {
        return fail(SGPU_ERR_ARG, "null argument");
    }
    *out = nullptr;
    if (!models || !data || !data_of_chain || nchains < 1 || ndata < 1) {
        return fail(SGPU_ERR_ARG, "null argument or empty batch");
    }
    for (int ci = 0; ci < nchains; ci++) {
        int rc = check_model(&models[ci], ci);
        if (rc) {
            return rc;
        }
        const int di = data_of_chain[ci];
        if (di < 0 || di >= ndata) {
            return fail(SGPU_ERR_ARG, "chain %d names data set %d of %d", ci, di, ndata);
        }
        const sgpu_model &m0 = models[0];
        const sgpu_model &mc = models[ci];
        if (mc.burnin != m0.burnin || mc.metrofreq != m0.metrofreq ||
            mc.computeprob != m0.computeprob) {
            return fail(SGPU_ERR_ARG,
                        "chain %d: BURNIN, METROFREQ and COMPUTEPROB must match across a batch", ci);
        }
    }
    // each data set's shape comes from its first chain; later chains must agree
    std::vector<int> first(ndata, -1);
    for (int ci = 0; ci < nchains; ci++) {
        const int di = data_of_chain[ci];
        if (first[di] < 0) {
            first[di] = ci;
            int rc = check_data(&data[di], models[ci].N, models[ci].L, models[ci].A, di);
            if (rc) {
                return rc;
            }
        } else {
            const sgpu_model &a = models[first[di]], &b = models[ci];
            if (a.N != b.N || a.L != b.L || a.A != b.A) {
                return fail(SGPU_ERR_ARG, "chain %d: N, L, A differ from chain %d on data set %d",
                            ci, first[di], di);
            }
        }
    }
    if (tuning && (tuning->byte_loci < 0 || tuning->byte_loci > kZLoci || tuning->byte_chunk < 0 ||
                   tuning->z_blocks < 0 || tuning->metro_blocks < 0)) {
        return fail(SGPU_ERR_ARG, "tuning: byte_loci must lie in [0, %d], the rest be >= 0",
                    kZLoci);
    }
    const int z_target = (tuning && tuning->z_blocks > 0) ? tuning->z_blocks : kZBlocksDefault;
    const int metro_target =
        (tuning && tuning->metro_blocks > 0) ? tuning->metro_blocks : kMetroBlocksDefault;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        return fail(SGPU_ERR_DEVICE, "no HIP device available");
    }

    sgpu_ctx *c = new sgpu_ctx();
    if (hipGetDevice(&c->device) != hipSuccess) {
        delete c;
        return fail(SGPU_ERR_DEVICE, "hipGetDevice failed");
    }
    auto cleanup = [&](int code) {
        sgpu_destroy(c);
        return code;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(SGPU_ERR_DEVICE, "hipStreamCreate failed");
    }
    c->burnin = models[0].burnin;
    c->metrofreq = models[0].metrofreq;
    c->period_sweeps = (uint32_t)(c->metrofreq > 0 && c->metrofreq <= kMaxPeriodSweeps
                                      ? c->metrofreq * ((kPeriodSweeps + c->metrofreq - 1) / c->metrofreq)
                                      : kPeriodSweeps);
    c->computeprob = models[0].computeprob;

    int rc;
    c->data.resize(ndata);
    for (int di = 0; di < ndata; di++) {
        if (first[di] < 0) {
            continue;   // no chain uses it
        }
        const sgpu_model &m = models[first[di]];
        Dataset &ds = c->data[di];
        ds.fmt = data[di].format;
        ds.N = m.N;
        ds.L = m.L;
        ds.A = m.A;
        ds.G = (int)cdiv(m.N, ds.fmt == SGPU_GENO_SNP2 ? 16 : 2);
        ds.Lpad = (int)cdiv(m.L, kZLoci) * kZLoci;
        ds.Gpad = (int)(sgpu_packed_words(ds.fmt, m.N, m.L) / (size_t)m.L);   // byte format
        if ((rc = upload_dataset(c, ds, &data[di]))) {
            return cleanup(rc);
        }
    }
    BatchTiles bt = batch_tiles(models, data, data_of_chain, nchains, z_target, metro_target);
    bt.byte_loci = tuning ? tuning->byte_loci : 0;
    bt.byte_chunk = tuning ? tuning->byte_chunk : 0;
    c->byte_rows = bt.byte_rows;
    c->chains.resize(nchains);
    std::vector<double> lam0(nchains);
    for (int ci = 0; ci < nchains; ci++) {
        HostChain &h = c->chains[ci];
        h.m = models[ci];
        h.dataset = data_of_chain[ci];
        lam0[ci] = h.m.lambda0;
        c->any_alpha |= h.m.inferalpha != 0;
        c->any_lambda |= h.m.inferlambda != 0;
        c->any_freqscorr |= h.m.freqscorr != 0;
        if ((rc = setup_chain(c, h, c->data[h.dataset], bt))) {
            return cleanup(rc);
        }
    }
    if ((rc = dalloc(c, &c->cds, nchains)) || (rc = dalloc(c, &c->lambda0, nchains)) ||
        (rc = dalloc(c, &c->repd, 2)) || (rc = upload_descriptors(c)) || (rc = build_plans(c))) {
        return cleanup(rc);
    }
    // repd[0]: the device sweep index; repd[1]: the tail pass's completion count
    hipError_t e = hipMemsetAsync(c->repd, 0, 2 * sizeof(uint32_t), c->stream);
    if (e == hipSuccess) {
        e = hipMemcpyAsync(c->lambda0, lam0.data(), nchains * sizeof(double),
                           hipMemcpyHostToDevice, c->stream);
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

int sgpu_num_chains(const sgpu_ctx *c) { return c ? (int)c->chains.size() : 0; }

int sgpu_destroy(sgpu_ctx *c)
{
    if (!c) {
        return SGPU_OK;
    }
    if (c->stream) {
        (void)hipStreamSynchronize(c->stream);
    }
    // the executable graphs' kernel nodes hold addresses of the buffers below:
    // release the graphs first, then the memory they refer to (the reference
    // likewise releases its kernels before its buffers, Kernels.c:177-193)
    for (int t = 0; t < 4; t++) {
        if (c->graph[t]) {
            (void)hipGraphExecDestroy(c->graph[t]);
            c->graph[t] = nullptr;
        }
    }
    for (int t = 0; t < 2; t++) {
        if (c->period_graph[t]) {
            (void)hipGraphExecDestroy(c->period_graph[t]);
            c->period_graph[t] = nullptr;
        }
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
    for (HostChain &h : c->chains) {
        const sgpu_model &m = h.m;
        const ChainDesc &d = h.d;
        const int N = m.N, L = m.L, K = m.K, A = m.A;
        std::vector<double> alpha(K, m.alpha0), lambda(K, m.lambda0),
            F(K, m.freqscorr ? m.fpriormean : 0.5), q((size_t)N * K, 1.0 / K);
        HIP_OK(hipMemcpyAsync(d.alpha, alpha.data(), K * sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
        HIP_OK(hipMemcpyAsync(d.lambda, lambda.data(), K * sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
        HIP_OK(hipMemcpyAsync(d.F, F.data(), K * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipMemcpyAsync(d.Q, q.data(), q.size() * sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
        HIP_OK(hipMemsetAsync(d.P, 0, (size_t)L * K * A * sizeof(double), c->stream));
        HIP_OK(hipMemsetAsync(d.QSum, 0, (size_t)N * K * sizeof(double), c->stream));
        HIP_OK(hipMemsetAsync(d.PSum, 0, (size_t)L * K * A * sizeof(double), c->stream));
        HIP_OK(hipMemsetAsync(d.epsSum, 0, (size_t)L * A * sizeof(double), c->stream));
        HIP_OK(hipMemsetAsync(d.FstSum, 0, K * sizeof(double), c->stream));
        HIP_OK(hipMemsetAsync(d.alphaSum, 0, K * sizeof(double), c->stream));
        HIP_OK(hipMemsetAsync(d.lambdaSum, 0, K * sizeof(double), c->stream));
        HIP_OK(hipMemsetAsync(d.like3, 0, 3 * sizeof(double), c->stream));
        if (d.adist) {
            HIP_OK(hipMemsetAsync(d.adist, 0, (size_t)N * K * m.numboxes * sizeof(int32_t),
                                  c->stream));
        }
        // the host vectors die at the end of this iteration
        HIP_OK(hipStreamSynchronize(c->stream));
    }
    int rc = launch(c, k_init_eps, c->p_init_eps, kNT, (const double *)c->lambda0);
    if (rc) {
        return rc;
    }
    c->lam_fresh = false;
    c->rep_known = false;
    if ((rc = use_rep(c, 0)) || (rc = launch_z(c, 0)) || (rc = launch(c, k_logq_stats, c->p_logq, kNT))) {
        return rc;
    }
    c->nsum = 0;
    c->like_ready = false;
    c->initialized = true;
    return SGPU_OK;
}

int sgpu_get_chain_state(sgpu_ctx *c, int32_t chain, int32_t which, void *host, size_t bytes)
{
    int rc = chain_index(c, chain);
    if (rc) {
        return rc;
    }
    if (!host) {
        return fail(SGPU_ERR_ARG, "null argument");
    }
    HostChain &h = c->chains[chain];
    size_t elem;
    const size_t n = state_elems(h.m, which, &elem);
    if (n == 0) {
        return fail(SGPU_ERR_ARG, which == SGPU_ANCESTDIST
                                      ? "chain %d runs without ANCESTDIST"
                                      : "unknown state selector %d",
                    which == SGPU_ANCESTDIST ? chain : which);
    }
    if (bytes != n * elem) {
        return fail(SGPU_ERR_ARG, "state %d needs %zu bytes, got %zu", which, n * elem, bytes);
    }
    if (which == SGPU_NSUM) {
        HIP_OK(hipStreamSynchronize(c->stream));
        memcpy(host, &c->nsum, sizeof(int64_t));
        return SGPU_OK;
    }
    const void *src = state_ptr(h.d, which);
    if (which == SGPU_NA || which == SGPU_NQ) {
        const bool isA = which == SGPU_NA;
        int32_t *dst = isA ? h.nA : h.nQ;
        hipLaunchKernelGGL(k_reduce_counts, dim3(cdiv(n, 256)), dim3(256), 0, c->stream,
                           isA ? h.d.nApart : h.d.nQpart, isA ? h.d.n_it : h.d.n_lt, n, dst);
        LAUNCH_OK();
        src = dst;
    }
    HIP_OK(hipMemcpyAsync(host, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return SGPU_OK;
}

int sgpu_set_chain_state(sgpu_ctx *c, int32_t chain, int32_t which, const void *host,
                         size_t bytes)
{
    int rc = chain_index(c, chain);
    if (rc) {
        return rc;
    }
    if (!host) {
        return fail(SGPU_ERR_ARG, "null argument");
    }
    HostChain &h = c->chains[chain];
    size_t elem;
    const size_t n = state_elems(h.m, which, &elem);
    if (n == 0) {
        return fail(SGPU_ERR_ARG, which == SGPU_ANCESTDIST
                                      ? "chain %d runs without ANCESTDIST"
                                      : "unknown state selector %d",
                    which == SGPU_ANCESTDIST ? chain : which);
    }
    if (bytes != n * elem) {
        return fail(SGPU_ERR_ARG, "state %d needs %zu bytes, got %zu", which, n * elem, bytes);
    }
    if (which == SGPU_NSUM) {
        memcpy(&c->nsum, host, sizeof(int64_t));
        return SGPU_OK;
    }
    if (which == SGPU_NA || which == SGPU_NQ) {
        // the 16-bit per-tile partials take the totals, up to 65535 per partial
        const bool isA = which == SGPU_NA;
        uint16_t *part = isA ? h.d.nApart : h.d.nQpart;
        const int nparts = isA ? h.d.n_it : h.d.n_lt;
        const int32_t *tot = (const int32_t *)host;
        std::vector<uint16_t> hp((size_t)nparts * n, 0);
        for (size_t j = 0; j < n; j++) {
            if (tot[j] < 0 || (int64_t)tot[j] > (int64_t)65535 * nparts) {
                return fail(SGPU_ERR_ARG, "count %d at %zu outside [0, %lld]", tot[j], j,
                            (long long)65535 * nparts);
            }
            int32_t left = tot[j];
            for (int t = 0; left > 0; t++) {
                const int32_t v = std::min(left, 65535);
                hp[(size_t)t * n + j] = (uint16_t)v;
                left -= v;
            }
        }
        HIP_OK(hipMemcpyAsync(part, hp.data(), hp.size() * sizeof(uint16_t),
                              hipMemcpyHostToDevice, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
        if (!isA) {
            c->fold_fresh = false;
        }
        return SGPU_OK;
    }
    HIP_OK(hipMemcpyAsync(state_ptr(h.d, which), host, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    if (which == SGPU_EPS) {
        c->lam_fresh = false;
    }
    if (which == SGPU_Q) {
        return launch(c, k_logq_stats, c->p_logq_chain[chain], kNT);
    }
    return SGPU_OK;
}

int sgpu_get_state(sgpu_ctx *c, int32_t which, void *host, size_t bytes)
{
    return sgpu_get_chain_state(c, 0, which, host, bytes);
}

int sgpu_set_state(sgpu_ctx *c, int32_t which, const void *host, size_t bytes)
{
    return sgpu_set_chain_state(c, 0, which, host, bytes);
}

}  // extern "C"

namespace {

// UpdateP over the plan pp: p_update_p, or a sweep's p_update_pm_sw / p_update_pq_sw
// with lambda decision and Gibbs Q blocks.
int update_p(sgpu_ctx *c, uint32_t rep, int32_t collect, const Plan &pp)
{
    int rc = check_ready(c);
    if (rc || (rc = use_rep(c, rep))) {
        return rc;
    }
    PhaseTimer timer(c, SGPU_PH_P);
    return launch(c, k_update_p, pp, kNT, (const uint32_t *)c->repd, (int)collect);
}

// The fold of the ancestry-count partials the Gibbs update reads, unless nQfold
// already holds the fold of the current partials.
int ensure_fold(sgpu_ctx *c)
{
    if (c->fold_fresh) {
        return SGPU_OK;
    }
    int rc = launch(c, k_fold_rows, c->p_fold, kNT);
    if (rc) {
        return rc;
    }
    c->fold_fresh = true;
    return SGPU_OK;
}

// The log-eps sums that the lambda decision of F-model chains reads, unless
// lam_epart already holds those of the current eps.
int ensure_lam(sgpu_ctx *c)
{
    if (c->lam_fresh) {
        return SGPU_OK;
    }
    if (c->p_eps_logsum.nblocks > 0) {
        int rc = launch(c, k_eps_logsum, c->p_eps_logsum, kNT);
        if (rc) {
            return rc;
        }
    }
    c->lam_fresh = true;
    return SGPU_OK;
}

// UpdateQ over the (individual, pop) plan qp: p_gibbs, or in a Metropolis sweep
// p_propose_sw with the lambda log-sum blocks.
int update_q(sgpu_ctx *c, uint32_t rep, int32_t collect, const Plan &qp)
{
    int rc = check_ready(c);
    if (rc || (rc = use_rep(c, rep))) {
        return rc;
    }
    PhaseTimer timer(c, SGPU_PH_Q);
    const uint32_t *r = c->repd;
    const bool metro = c->metrofreq > 0 && rep % (uint32_t)c->metrofreq == 0;
    if (metro) {
        // (individual, pop) grid; in a sweep the proposals ran in the UpdateP
        // launch and qp holds only lambda log-sum blocks, if any
        if (qp.nblocks > 0 && (rc = launch(c, k_metro_propose, qp, kNT, r))) {
            return rc;
        }
        for (const KGroup &g : c->metro_groups) {
            mkern_t k = metro_kernel(g.fmt, g.K);
            if (!k) {
                return fail(SGPU_ERR_UNSUPPORTED, "no Metropolis kernel for K=%d", g.K);
            }
            if ((rc = launch(c, k, g.plan, kNT, r))) {
                return rc;
            }
        }
        return launch(c, k_metro_accept, c->p_accept, kNT, r, (int)collect);
    }
    if ((rc = ensure_fold(c))) {
        return rc;
    }
    return launch(c, k_update_q_gibbs, qp, kNT, r, (int)collect);
}

}  // namespace

extern "C" {

int sgpu_update_p(sgpu_ctx *c, uint32_t rep, int32_t collect)
{
    if (!c) {
        return fail(SGPU_ERR_ARG, "null context");
    }
    return update_p(c, rep, collect, c->p_update_p);
}

int sgpu_update_q(sgpu_ctx *c, uint32_t rep, int32_t collect)
{
    if (!c) {
        return fail(SGPU_ERR_ARG, "null context");
    }
    return update_q(c, rep, collect, c->p_gibbs);
}

int sgpu_update_z(sgpu_ctx *c, uint32_t rep, int32_t flags, int8_t *zout)
{
    int rc = check_ready(c);
    if (rc || (rc = use_rep(c, rep))) {
        return rc;
    }
    if (zout && c->chains.size() != 1) {
        return fail(SGPU_ERR_ARG, "a debug copy of Z needs a single-chain context");
    }
    PhaseTimer timer(c, SGPU_PH_Z);
    const bool like = (flags & SGPU_Z_LIKE) != 0;
    if (zout) {
        HostChain &h = c->chains[0];
        const size_t zn = (size_t)h.m.N * 2 * h.m.L;
        if (!c->zdbg) {
            HIP_OK(hipMalloc(&c->zdbg, zn));
        }
        h.d.zout = c->zdbg;
        if ((rc = upload_descriptors(c)) || (rc = launch_z(c, like ? 4 : 3))) {
            return rc;
        }
        h.d.zout = nullptr;
        if ((rc = upload_descriptors(c))) {
            return rc;
        }
        HIP_OK(hipMemcpyAsync(zout, c->zdbg, zn, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipStreamSynchronize(c->stream));
    } else if ((rc = launch_z(c, like ? 2 : 1))) {
        return rc;
    }
    c->like_ready = like;
    return SGPU_OK;
}

int sgpu_update_alpha(sgpu_ctx *c, uint32_t rep)
{
    int rc = check_ready(c);
    if (rc || (rc = use_rep(c, rep))) {
        return rc;
    }
    PhaseTimer timer(c, SGPU_PH_ALPHA);
    return launch(c, k_update_alpha, c->p_alpha, kNT, (const uint32_t *)c->repd);
}

int sgpu_update_lambda(sgpu_ctx *c, uint32_t rep)
{
    int rc = check_ready(c);
    if (rc || (rc = use_rep(c, rep))) {
        return rc;
    }
    PhaseTimer timer(c, SGPU_PH_LAMBDA);
    if ((rc = launch(c, k_lambda_partial, c->p_lam_part, kNT))) {
        return rc;
    }
    return launch(c, k_lambda_finalize, c->p_lam_fin, kNT, (const uint32_t *)c->repd);
}

int sgpu_update_eps(sgpu_ctx *c, uint32_t rep, int32_t collect)
{
    int rc = check_ready(c);
    if (rc || (rc = use_rep(c, rep))) {
        return rc;
    }
    PhaseTimer timer(c, SGPU_PH_EPS);
    return launch(c, k_update_eps, c->p_eps, kNT, c->repd, c->repd + 1,
                  collect ? kEpsCollect : 0);
}

int sgpu_update_fst(sgpu_ctx *c, uint32_t rep)
{
    int rc = check_ready(c);
    if (rc || (rc = use_rep(c, rep))) {
        return rc;
    }
    PhaseTimer timer(c, SGPU_PH_FST);
    const uint32_t *r = c->repd;
    if ((rc = launch(c, k_fst_partial, c->p_fst_part, kNT, r))) {
        return rc;
    }
    return launch(c, k_fst_finalize, c->p_fst_fin, kNT, r);
}

int sgpu_collect(sgpu_ctx *c, uint32_t rep)
{
    (void)rep;
    int rc = check_ready(c);
    if (rc) {
        return rc;
    }
    PhaseTimer timer(c, SGPU_PH_COLLECT);
    const int with_like = (c->computeprob && c->like_ready) ? 1 : 0;
    if ((rc = launch(c, k_collect, c->p_collect, 128, with_like))) {
        return rc;
    }
    c->nsum++;
    return SGPU_OK;
}

int sgpu_sweep(sgpu_ctx *c, uint32_t rep)
{
    int rc = check_ready(c);
    if (rc) {
        return rc;
    }
    // The reference's phase order (structure.c:497-638) in four launches (Gibbs
    // sweeps) or six (Metropolis).  UpdateP and the Gibbs Q update share a launch:
    // the Gibbs update reads the counts of the last UpdateZ and alpha, neither of
    // which UpdateP reads or writes.  The lambda update moves, as nothing between
    // reads or writes what it reads and writes: F-model chains decide in the
    // UpdateP launch from the log-eps sums of the last epsilon pass (lambda is read
    // next by the epsilon update, eps written first by it); other chains sum log P
    // after UpdateP and decide in the epsilon launch (lambda is read next by
    // UpdateP).  Then UpdateZ; epsilon with the Fst partials and the log-eps sums,
    // the count fold for the next sweep, and per chain alpha (it reads only Q and
    // alpha), the lnP(D) sum and the scalar sums; the Fst decision with FstSum,
    // which may close the sweep.
    const int collect = ((int64_t)rep + 1 > c->burnin) ? 1 : 0;
    const bool like = collect && c->computeprob;
    const bool metro = c->metrofreq > 0 && rep % (uint32_t)c->metrofreq == 0;
    if ((rc = ensure_lam(c))) return rc;
    if (metro) {
        if ((rc = update_p(c, rep, collect, c->p_update_pm_sw))) return rc;
        if ((rc = update_q(c, rep, collect, c->p_propose_sw))) return rc;
    } else {
        if ((rc = ensure_fold(c))) return rc;
        if ((rc = update_p(c, rep, collect, c->p_update_pq_sw))) return rc;
        if (c->p_lam_part_p.nblocks > 0) {   // lambda log sums of the new P (no F model)
            PhaseTimer timer(c, SGPU_PH_LAMBDA);
            if ((rc = launch(c, k_lambda_partial, c->p_lam_part_p, kNT))) return rc;
        }
    }
    if ((rc = sgpu_update_z(c, rep, like ? SGPU_Z_LIKE : 0, nullptr))) return rc;
    if (c->p_lam_fin_x.nblocks > 0) {   // pop-specific lambda with the F model
        PhaseTimer timer(c, SGPU_PH_LAMBDA);
        if ((rc = launch(c, k_lambda_partial, c->p_lam_part_x, kNT)) ||
            (rc = launch(c, k_lambda_finalize, c->p_lam_fin_x, kNT, (const uint32_t *)c->repd))) {
            return rc;
        }
    }
    // in a captured sweep the last launch (Fst decision, else epsilon) advances
    // the device sweep index
    c->tail_advances = false;
    int ef = (collect ? kEpsCollect : 0) | (c->any_freqscorr ? kEpsFst : 0) |
             (c->any_alpha ? kEpsAlpha : 0) | ((like && c->like_ready) ? kEpsLike : 0);
    if (c->any_freqscorr || c->p_fold.nblocks > 0 || (ef & (kEpsAlpha | kEpsCollect))) {
        if (c->capturing && c->p_tail.nblocks == 0) {
            ef |= kEpsAdvance;
        }
        PhaseTimer timer(c, SGPU_PH_EPS);
        if ((rc = launch(c, k_update_eps, c->p_postz, kNT, c->repd, c->repd + 1, ef))) {
            return rc;
        }
        c->fold_fresh = true;
        c->tail_advances = (ef & kEpsAdvance) != 0;
    }
    if (c->p_tail.nblocks > 0) {
        const int tf = kTailFst | (collect ? kTailCollect : 0) | (c->capturing ? kTailAdvance : 0);
        PhaseTimer timer(c, SGPU_PH_COLLECT);
        if ((rc = launch(c, k_sweep_tail, c->p_tail, kTailNT, c->repd, c->repd + 1, tf))) {
            return rc;
        }
        c->tail_advances = (tf & kTailAdvance) != 0;
    }
    if (collect) {
        c->nsum++;
    }
    return SGPU_OK;
}

}  // extern "C"

namespace {

// Captures sweeps rep .. rep + span - 1 into one executable graph, each closed
// by the advance of the device sweep index.  The host-side sweep count is left
// as it was: the replays count.
int capture_sweeps(sgpu_ctx *c, hipGraphExec_t *exec, uint32_t rep, uint32_t span)
{
    hipGraph_t g = nullptr;
    const int64_t nsum0 = c->nsum;
    HIP_OK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    int rc = SGPU_OK;
    for (uint32_t s = 0; s < span && rc == SGPU_OK; s++) {
        rc = sgpu_sweep(c, rep + s);
        if (rc == SGPU_OK && !c->tail_advances) {
            hipLaunchKernelGGL(k_advance_rep, dim3(1), dim3(1), 0, c->stream, c->repd);
        }
    }
    c->capturing = false;
    c->nsum = nsum0;
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
    e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        *exec = nullptr;
        return fail(SGPU_ERR_DEVICE, "graph instantiate failed: %s", hipGetErrorString(e));
    }
    return SGPU_OK;
}

}  // namespace

extern "C" {

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
    // Graphs captured on first use and replayed: one per span of period_sweeps
    // sweeps that lies wholly in burn-in or in the collecting sweeps and starts
    // at a Metropolis sweep (periods up to kMaxPeriodSweeps) or holds none (longer
    // periods, or none) -- one replay for the lot, as a graph launch costs
    // about 10 us of idle GPU between replays -- and one per single sweep type
    // (Metropolis or Gibbs Q, burn-in or collecting) for the sweeps around
    // them.  Each captured sweep ends by advancing the device index.  Every
    // sweep leaves the count fold and the log-eps sums current for the next, so
    // the graphs are captured (and replayed) with them current: make them now
    // if they are not.
    if ((rc = use_rep(c, rep0)) || (rc = ensure_fold(c)) || (rc = ensure_lam(c))) {
        return rc;
    }
    const uint32_t S = c->period_sweeps;
    for (uint32_t r = 0; r < n;) {
        const uint32_t rep = rep0 + r;
        const bool collect = (int64_t)rep + 1 > c->burnin;
        const uint32_t mf = c->metrofreq > 0 ? (uint32_t)c->metrofreq : 0u;
        const bool aligned = mf == 0 || (mf <= (uint32_t)kMaxPeriodSweeps
                                             ? rep % mf == 0
                                             : rep % mf != 0 && rep % mf + S <= mf);
        const bool one_side = collect || (int64_t)rep + S <= c->burnin;
        hipGraphExec_t *exec;
        uint32_t span;
        if (n - r >= S && aligned && one_side) {
            exec = &c->period_graph[collect ? 1 : 0];
            span = S;
        } else {
            const bool metro = c->metrofreq > 0 && rep % (uint32_t)c->metrofreq == 0;
            exec = &c->graph[(metro ? 1 : 0) | (collect ? 2 : 0)];
            span = 1;
        }
        if (!*exec && (rc = capture_sweeps(c, exec, rep, span))) {
            return rc;
        }
        HIP_OK(hipGraphLaunch(*exec, c->stream));
        c->rep_shadow = rep + span;
        if (collect) {
            c->nsum += span;
        }
        c->like_ready = collect && c->computeprob;
        r += span;
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

int sgpu_eval_math(int32_t fn, const double *x, double *y, int64_t n)
{
    if (fn < SGPU_FN_LOG || fn > SGPU_FN_SIN || n < 0 || (n > 0 && (!x || !y))) {
        return fail(SGPU_ERR_ARG, "sgpu_eval_math: fn %d, n %lld", fn, (long long)n);
    }
    if (n == 0) {
        return SGPU_OK;
    }
    const size_t bytes = (size_t)n * sizeof(double);
    DevBuf dx, dy;
    HIP_OK(hipMalloc(&dx.p, bytes));
    HIP_OK(hipMalloc(&dy.p, bytes));
    HIP_OK(hipMemcpy(dx.p, x, bytes, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_eval_math, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, fn,
                       (const double *)dx.p, (double *)dy.p, n);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpy(y, dy.p, bytes, hipMemcpyDeviceToHost));
    return SGPU_OK;
}

int sgpu_eval_math_noalloc(int32_t fn, const double *x, double *y, int64_t n)
{
    if (fn < SGPU_FN_LOG || fn > SGPU_FN_SIN || n < 0 || (n > 0 && (!x || !y))) {
        return fail(SGPU_ERR_ARG, "sgpu_eval_math_noalloc: fn %d, n %lld", fn, (long long)n);
    }
    if (n > SGPU_EVAL_MATH_NOALLOC_MAX) {
        return fail(SGPU_ERR_ARG, "sgpu_eval_math_noalloc: n = %lld exceeds the %d values of its "
                                  "static staging", (long long)n, SGPU_EVAL_MATH_NOALLOC_MAX);
    }
    if (n == 0) {
        return SGPU_OK;
    }
    // one call at a time per process: the staging is shared by every call on a GPU
    std::lock_guard<std::mutex> lock(g_stage_mu);
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kStageDevices) {
        return fail(SGPU_ERR_DEVICE, "sgpu_eval_math_noalloc: device %d beyond %d", dev,
                    kStageDevices);
    }
    if (!g_stage_ptr[dev]) {
        // once per device, that device current: the array's address on it
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            return fail(SGPU_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed: %s",
                        hipGetErrorString(e));
        }
        g_stage_ptr[dev] = (double *)p;
    }
    double *sx = g_stage_ptr[dev];                  // x staged in the first half
    double *sy = sx + SGPU_EVAL_MATH_NOALLOC_MAX;   // y in the second
    const size_t bytes = (size_t)n * sizeof(double);
    HIP_OK(hipMemcpy(sx, x, bytes, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_eval_math, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, fn,
                       (const double *)sx, sy, n);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpy(y, sy, bytes, hipMemcpyDeviceToHost));
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
