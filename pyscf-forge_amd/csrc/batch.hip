// Many independent CAS problems in one persistent launch.
//
// sigma_p,v = H_p c_p,v for every problem p of a batch and every active
// vector slot v of p (Davidson roots), with problems of different size
// (norb, nelec) packed into one pair of flat vectors.  One kernel,
// k_batch_row, does the whole product: workgroups take work items
// (problem, range of alpha strings) from a queue (one atomic counter,
// items sorted by decreasing cost), and for every string I and active slot
// v of the item run the fused-row scheme of sigma.hip:
//   * builder wavefronts gather t1[pq][b] for a chunk of FUSED_NB beta
//     strings into LDS (double-buffered, maps two chunks ahead);
//   * one MFMA wavefront per 16-row tile of h_p (A fragments in registers,
//     reloaded when the problem changes) contracts the chunk;
//   * the beta part of the scatter goes to the LDS copy of sigma row I,
//     the alpha part (sigma[J][b] += s vt1[pq][I][b] for the alpha link
//     I -> J of pair pq, the reverse of the gather link) goes to HBM with
//     FP64 atomic adds, so no workspace and no second kernel is needed;
//     the row is flushed with atomic adds as well.
// Accumulation order therefore varies between runs (~1e-16 relative).
//
// Work item size and grid size come from an autotuner (fcis_batch_tune)
// that times candidate schedules on the batch itself.  The zero fill of
// the output, the queue reset and the kernel are captured once into a
// hipGraph and replayed on every call with the same buffers.
//
// Problems are the reference's direct_spin1 CAS spaces (libfci_gas tasks
// own alpha destination ranges the same way, fci_contract_gas.c:1934-2036);
// the batching itself has no counterpart in the reference, whose solvers
// handle one problem per call.
#include <algorithm>
#include <new>
#include <stdlib.h>
#include <string.h>
#include <vector>

#include "fused.h"

namespace {

constexpr int BATCH_SMALL_MAXW = 5;      // npair <= 80 (norb <= 12)
constexpr int BATCH_SMALL_KS = 20;       // 4-wide k-steps for npair <= 80
constexpr int BATCH_SMALL_KS_BF16 = 3;   // 32-wide k-steps for npair <= 96

// Per-problem tables live in three shared device arrays (kernel arguments,
// so that every table load is a global load): offsets are in elements.
struct BatchProb {
        int64_t pa;                // pmapa [npair][nstra]: (target << 1 | negative) or -1
        int64_t pb;                // beta pair map [npair][nbp] (int16)
        int64_t hf;                // A fragments [ks][mt][lane] (and bf16 low halves)
        int64_t voff;              // packed offset of vector slot 0
        int64_t ndet;
        int npair, mtiles, ksteps, nvec;
        int nb, nstra, nbp, aoff;  // aoff: index of slot 0 in the active mask
};

struct BatchItem {
        int prob, r0, r1, pad;
};

// Builder wavefronts of the batched kernel (FP64: 3, enough at <= 12
// orbitals, which leaves room for two workgroups per CU; FP32/bf16: 7).
template <int PREC>
constexpr int batch_nbld()
{
        return PREC == FCIS_PREC_FP64 ? 3 : 7;
}

template <int PREC, int KS, int MAXW>
constexpr size_t batch_lds_bytes(int64_t nbpmax)
{
        typedef typename FusedTypes<PREC>::B BT;
        return (size_t)nbpmax * sizeof(double) +
               2 * sizeof(BT) * KS * FusedTypes<PREC>::KW * FUSED_NB +
               sizeof(int) * 16 * MAXW + 3 * sizeof(int16_t) * 16 * MAXW * FUSED_NB;
}

__device__ __forceinline__ void atomic_add_f64(double *p, double v)
{
        unsafeAtomicAdd(p, v);   // hardware FP64 add (coarse-grained device memory)
}

template <int PREC, int KS, int MAXW>
__global__ __launch_bounds__(64 * (MAXW + batch_nbld<PREC>())) void k_batch_row(
        const BatchProb *__restrict__ probs, const BatchItem *__restrict__ items, int nitems,
        int *__restrict__ counter, const uint8_t *__restrict__ active, int64_t nbpmax,
        const int32_t *__restrict__ pmapa_all, const int16_t *__restrict__ pmapb_all,
        const typename FusedTypes<PREC>::A *__restrict__ hfrag_all,
        const typename FusedTypes<PREC>::A *__restrict__ hfrag_lo_all,
        const double *__restrict__ cin, double *__restrict__ cout)
{
        typedef FusedTypes<PREC> Ty;
        typedef typename Ty::B BT;
        constexpr int KW = Ty::KW, NB = FUSED_NB, NBLD = batch_nbld<PREC>();
        constexpr int KPMAX = KS * KW;
        constexpr int MAXP = 16 * MAXW;
        extern __shared__ __align__(16) char smem[];
        __shared__ int s_item;
        double *srow = reinterpret_cast<double *>(smem);
        BT *buf0 = reinterpret_cast<BT *>(smem + (size_t)nbpmax * sizeof(double));
        BT *buf1 = buf0 + KPMAX * NB;
        int *arow = reinterpret_cast<int *>(buf1 + KPMAX * NB);
        int16_t *maps = reinterpret_cast<int16_t *>(arow + MAXP);

        const int tid = threadIdx.x, nthr = blockDim.x;
        const int wave = tid >> 6, lane = tid & 63, li = lane & 15, lk = lane >> 4;
        for (int64_t i = tid; i < nbpmax; i += nthr) srow[i] = 0.0;

        // Barrier schedule per item (all roles): (A) item fetched; per active
        // (string, slot): (1) row start, (2) arow, (3) maps of chunks 0,1,
        // (4) chunk 0 built, one per chunk; (B) item done.
        if (wave >= MAXW) {
                // ---- builder wavefronts (issue priority as in the fused row
                // kernel: builders first for FP64/FP32, MFMA wavefronts first
                // for bf16x3)
                if constexpr (PREC != FCIS_PREC_BF16X3) __builtin_amdgcn_s_setprio(1);
                const int bt = tid - MAXW * 64;
                for (;;) {
                        __syncthreads();   // (A)
                        const int it = s_item;
                        if (it >= nitems) break;
                        FusedBuild<NBLD, MAXW> st;
                        const BatchItem W = items[it];
                        const BatchProb P = probs[W.prob];
                        const int32_t *pmapa = pmapa_all + P.pa;
                        const int16_t *pmapb = pmapb_all + P.pb;
                        for (int e = P.npair * NB + tid; e < KPMAX * NB; e += nthr) {
                                buf0[e] = (BT)0;
                                buf1[e] = (BT)0;
                        }
                        const int64_t nch = (P.nb + NB - 1) / NB;
                        for (int r = W.r0; r < W.r1; r++) {
                                for (int v = 0; v < P.nvec; v++) {
                                        if (!active[P.aoff + v]) continue;
                                        const double *c = cin + P.voff + (int64_t)v * P.ndet;
                                        __syncthreads();   // (1)
                                        for (int q = bt; q < P.npair; q += NBLD * 64)
                                                arow[q] = pmapa[(int64_t)q * P.nstra + r];
                                        __syncthreads();   // (2)
                                        fb_issue_maps(st, P.npair, P.nbp, 0, bt, pmapb);
                                        fb_commit_maps(st, P.npair, bt, maps);
                                        if (nch > 1) {
                                                fb_issue_maps(st, P.npair, P.nbp, 1, bt, pmapb);
                                                fb_commit_maps(st, P.npair, bt, maps + MAXP * NB);
                                        }
                                        __syncthreads();   // (3)
                                        fb_issue_c(st, P.npair, P.nb, r, 0, bt, arow, maps, c);
                                        fb_commit_c(st, P.npair, bt, buf0);
                                        __syncthreads();   // (4)
                                        for (int64_t j = 0; j < nch; j++) {
                                                if (j + 1 < nch)
                                                        fb_issue_c(st, P.npair, P.nb, r, j + 1, bt, arow,
                                                                   maps + ((j + 1) % 3) * MAXP * NB, c);
                                                if (j + 2 < nch)
                                                        fb_issue_maps(st, P.npair, P.nbp, j + 2, bt, pmapb);
                                                if (j + 1 < nch)
                                                        fb_commit_c(st, P.npair, bt, ((j + 1) & 1) ? buf1 : buf0);
                                                if (j + 2 < nch)
                                                        fb_commit_maps(st, P.npair, bt,
                                                                       maps + ((j + 2) % 3) * MAXP * NB);
                                                __syncthreads();
                                        }
                                        double *dst = cout + P.voff + (int64_t)v * P.ndet + (int64_t)r * P.nb;
                                        for (int i = tid; i < P.nb; i += nthr) {
                                                double x = srow[i];
                                                if (x != 0.0) atomic_add_f64(dst + i, x);
                                                srow[i] = 0.0;
                                        }
                                }
                        }
                        __syncthreads();   // (B)
                }
                return;
        }

        // ---- MFMA wavefronts (waves >= the problem's mtiles only keep the
        // barrier schedule)
        if constexpr (PREC == FCIS_PREC_BF16X3) __builtin_amdgcn_s_setprio(1);
        typename Ty::A ah[KS];
        typename Ty::A al[PREC == FCIS_PREC_BF16X3 ? KS : 1];
        int cur = -1;
        for (;;) {
                if (tid == 0) s_item = atomicAdd(counter, 1);
                __syncthreads();   // (A)
                const int it = s_item;
                if (it >= nitems) break;
                const BatchItem W = items[it];
                const BatchProb P = probs[W.prob];
                const bool mma = wave < P.mtiles;
                if (W.prob != cur) {
                        cur = W.prob;
                        if (mma) {
                                // fragments past the problem's k-steps are zero, so
                                // that mma_steps may run past them
#pragma unroll
                                for (int ks = 0; ks < KS; ks++) {
                                        const size_t o = ((size_t)ks * P.mtiles + wave) * 64 + lane;
                                        ah[ks] = ks < P.ksteps ? hfrag_all[P.hf + o] : typename Ty::A{};
                                        if constexpr (PREC == FCIS_PREC_BF16X3)
                                                al[ks] = ks < P.ksteps ? hfrag_lo_all[P.hf + o] : typename Ty::A{};
                                }
                        }
                }
                for (int e = P.npair * NB + tid; e < KPMAX * NB; e += nthr) {
                        buf0[e] = (BT)0;
                        buf1[e] = (BT)0;
                }
                const int64_t nch = (P.nb + NB - 1) / NB;
                for (int r = W.r0; r < W.r1; r++) {
                        for (int v = 0; v < P.nvec; v++) {
                                if (!active[P.aoff + v]) continue;
                                double *sig = cout + P.voff + (int64_t)v * P.ndet;
                                __syncthreads();   // (1)
                                __syncthreads();   // (2)
                                __syncthreads();   // (3)
                                __syncthreads();   // (4)
                                for (int64_t j = 0; j < nch; j++) {
                                        if (mma) {
                                                const BT *bb = (j & 1) ? buf1 : buf0;
                                                const int16_t *mb = maps + (j % 3) * MAXP * NB;
                                                // two unconditional depths: half or all of KS
                                                const typename Ty::Acc acc =
                                                        P.ksteps <= KS / 2
                                                                ? mma_steps<PREC, KS / 2, KS>(ah, al, bb, lk, li)
                                                                : mma_steps<PREC, KS, KS>(ah, al, bb, lk, li);
                                                const int64_t b = j * NB + li;
                                                if (b < P.nb) {
#pragma unroll
                                                        for (int q = 0; q < 4; q++) {
                                                                // f64 MFMA: D row lk + 4q; f32/bf16: 4lk + q
                                                                int pq = wave * 16 + (PREC == FCIS_PREC_FP64
                                                                                              ? lk + 4 * q
                                                                                              : 4 * lk + q);
                                                                if (pq < P.npair) {
                                                                        double x = (double)acc[q];
                                                                        int e = mb[pq * NB + li];
                                                                        if (e >= 0)
                                                                                atomicAdd(&srow[e >> 1], (e & 1) ? -x : x);
                                                                        int a = arow[pq];
                                                                        if (a >= 0) {
                                                                                double y = (a & 1) ? -x : x;
                                                                                int J = a >> 1;
                                                                                if (J == r)
                                                                                        atomicAdd(&srow[b], y);
                                                                                else
                                                                                        atomic_add_f64(sig + (int64_t)J * P.nb + b, y);
                                                                        }
                                                                }
                                                        }
                                                }
                                        }
                                        __syncthreads();
                                }
                                double *dst = sig + (int64_t)r * P.nb;
                                for (int i = tid; i < P.nb; i += nthr) {
                                        double x = srow[i];
                                        if (x != 0.0) atomic_add_f64(dst + i, x);
                                        srow[i] = 0.0;
                                }
                        }
                }
                __syncthreads();   // (B)
        }
}

// Zero the packed output and reset the work queue (graph-capturable: no
// memset nodes).
__global__ void k_batch_reset(int64_t n, double *__restrict__ out, int *__restrict__ counter)
{
        int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (i == 0) *counter = 0;
        if (i < n) out[i] = 0.0;
}

}  // namespace

// ---------------------------------------------------------------- batch object

struct fcis_batch {
        int nprob, precision;
        int variant;                         // 0: norb <= 12 kernels, 1: norb <= 16 kernels
        std::vector<BatchProb> probs;
        int32_t *d_pmapa;                    // all problems' tables (BatchProb offsets)
        int16_t *d_pmapb;
        void *d_hfrag, *d_hfrag_lo;
        std::vector<int64_t> offs;           // packed offset of each problem's slot 0
        std::vector<double> cost;            // per-string cost of each problem (all slots)
        int64_t total;                       // packed vector length
        int nslots;
        int64_t nbpmax;
        size_t lds;
        int ncu, occ;
        int rows_per_item, grid;             // schedule
        BatchProb *d_probs;
        BatchItem *d_items;
        int nitems, items_cap;
        int *d_counter;
        uint8_t *d_active;
        hipStream_t cap_stream;
        hipGraph_t graph;
        hipGraphExec_t exec;
        const double *g_in;
        double *g_out;
        int use_graph;
        double tuned_ms;
};

namespace {

template <int PREC, int KS, int MAXW>
int launch_row(fcis_batch *b, const double *cin, double *cout, hipStream_t st)
{
        typedef typename FusedTypes<PREC>::A A;
        hipLaunchKernelGGL((k_batch_row<PREC, KS, MAXW>), dim3(b->grid),
                           dim3(64 * (MAXW + batch_nbld<PREC>())), b->lds, st, b->d_probs,
                           b->d_items, b->nitems, b->d_counter, b->d_active, b->nbpmax, b->d_pmapa,
                           b->d_pmapb, (const A *)b->d_hfrag, (const A *)b->d_hfrag_lo, cin, cout);
        FCIS_CHECK_LAUNCH();
        return FCIS_SUCCESS;
}

template <int PREC, int KS, int MAXW>
const void *row_kernel()
{
        return (const void *)k_batch_row<PREC, KS, MAXW>;
}

const void *batch_kernel(int precision, int variant)
{
        if (precision == FCIS_PREC_FP64)
                return variant ? row_kernel<FCIS_PREC_FP64, FUSED_KS, FUSED_MAXW>()
                               : row_kernel<FCIS_PREC_FP64, BATCH_SMALL_KS, BATCH_SMALL_MAXW>();
        if (precision == FCIS_PREC_FP32)
                return variant ? row_kernel<FCIS_PREC_FP32, FUSED_KS, FUSED_MAXW>()
                               : row_kernel<FCIS_PREC_FP32, BATCH_SMALL_KS, BATCH_SMALL_MAXW>();
        return variant ? row_kernel<FCIS_PREC_BF16X3, FUSED_KS_BF16, FUSED_MAXW>()
                       : row_kernel<FCIS_PREC_BF16X3, BATCH_SMALL_KS_BF16, BATCH_SMALL_MAXW>();
}

size_t batch_lds(int precision, int variant, int64_t nbpmax)
{
        if (precision == FCIS_PREC_FP64)
                return variant ? batch_lds_bytes<FCIS_PREC_FP64, FUSED_KS, FUSED_MAXW>(nbpmax)
                               : batch_lds_bytes<FCIS_PREC_FP64, BATCH_SMALL_KS, BATCH_SMALL_MAXW>(nbpmax);
        if (precision == FCIS_PREC_FP32)
                return variant ? batch_lds_bytes<FCIS_PREC_FP32, FUSED_KS, FUSED_MAXW>(nbpmax)
                               : batch_lds_bytes<FCIS_PREC_FP32, BATCH_SMALL_KS, BATCH_SMALL_MAXW>(nbpmax);
        return variant ? batch_lds_bytes<FCIS_PREC_BF16X3, FUSED_KS_BF16, FUSED_MAXW>(nbpmax)
                       : batch_lds_bytes<FCIS_PREC_BF16X3, BATCH_SMALL_KS_BF16, BATCH_SMALL_MAXW>(nbpmax);
}

int batch_threads(int precision, int variant)
{
        const int maxw = variant ? FUSED_MAXW : BATCH_SMALL_MAXW;
        const int nbld = precision == FCIS_PREC_FP64 ? batch_nbld<FCIS_PREC_FP64>()
                       : precision == FCIS_PREC_FP32 ? batch_nbld<FCIS_PREC_FP32>()
                                                      : batch_nbld<FCIS_PREC_BF16X3>();
        return 64 * (maxw + nbld);
}

int launch_batch_kernel(fcis_batch *b, const double *cin, double *cout, hipStream_t st)
{
        const int pr = b->precision, v = b->variant;
        if (pr == FCIS_PREC_FP64)
                return v ? launch_row<FCIS_PREC_FP64, FUSED_KS, FUSED_MAXW>(b, cin, cout, st)
                         : launch_row<FCIS_PREC_FP64, BATCH_SMALL_KS, BATCH_SMALL_MAXW>(b, cin, cout, st);
        if (pr == FCIS_PREC_FP32)
                return v ? launch_row<FCIS_PREC_FP32, FUSED_KS, FUSED_MAXW>(b, cin, cout, st)
                         : launch_row<FCIS_PREC_FP32, BATCH_SMALL_KS, BATCH_SMALL_MAXW>(b, cin, cout, st);
        return v ? launch_row<FCIS_PREC_BF16X3, FUSED_KS_BF16, FUSED_MAXW>(b, cin, cout, st)
                 : launch_row<FCIS_PREC_BF16X3, BATCH_SMALL_KS_BF16, BATCH_SMALL_MAXW>(b, cin, cout, st);
}

// Work items of rows_per_item alpha strings, largest cost first (ties keep
// problem order so that consecutive items of one problem reuse its h).
int build_items(fcis_batch *b, int rows_per_item)
{
        std::vector<BatchItem> items;
        std::vector<double> cost;
        for (int p = 0; p < b->nprob; p++) {
                const BatchProb &P = b->probs[p];
                for (int r = 0; r < P.nstra; r += rows_per_item) {
                        int r1 = std::min(P.nstra, r + rows_per_item);
                        items.push_back(BatchItem{p, r, r1, 0});
                        cost.push_back(b->cost[p] * (r1 - r));
                }
        }
        std::vector<int> order(items.size());
        for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
        std::vector<BatchItem> sorted(items.size());
        for (size_t i = 0; i < order.size(); i++) sorted[i] = items[order[i]];
        if ((int)sorted.size() > b->items_cap) {
                (void)hipFree(b->d_items);
                b->d_items = nullptr;
                b->items_cap = 0;
                if (fcis::dev_malloc(&b->d_items, sizeof(BatchItem) * sorted.size()) != hipSuccess) {
                        fcis::set_error("fcis_batch: item table allocation failed");
                        return FCIS_ERR_MEMORY;
                }
                b->items_cap = (int)sorted.size();
        }
        FCIS_HIP(hipMemcpy(b->d_items, sorted.data(), sizeof(BatchItem) * sorted.size(),
                           hipMemcpyHostToDevice));
        b->nitems = (int)sorted.size();
        b->rows_per_item = rows_per_item;
        return FCIS_SUCCESS;
}

void drop_graph(fcis_batch *b)
{
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->exec = nullptr;
        b->graph = nullptr;
        b->g_in = nullptr;
        b->g_out = nullptr;
}

int enqueue(fcis_batch *b, const double *cin, double *cout, hipStream_t st)
{
        hipLaunchKernelGGL(k_batch_reset, dim3((b->total + 255) / 256), dim3(256), 0, st, b->total,
                           cout, b->d_counter);
        FCIS_CHECK_LAUNCH();
        return launch_batch_kernel(b, cin, cout, st);
}

int capture(fcis_batch *b, const double *cin, double *cout)
{
        drop_graph(b);
        FCIS_HIP(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
        int st = enqueue(b, cin, cout, b->cap_stream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
        if (st != FCIS_SUCCESS) {
                if (g) (void)hipGraphDestroy(g);
                return st;
        }
        if (e != hipSuccess) {
                fcis::set_error("fcis_batch: graph capture failed: %s", hipGetErrorString(e));
                return FCIS_ERR_DEVICE;
        }
        b->graph = g;
        FCIS_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
        b->g_in = cin;
        b->g_out = cout;
        return FCIS_SUCCESS;
}


// Bytes of one A fragment element (bf16x3: 8 bf16 per lane).
size_t frag_elem_bytes(int precision)
{
        return precision == FCIS_PREC_FP64 ? 8 : precision == FCIS_PREC_FP32 ? 4 : 16;
}

// Build every problem's tables with a (minimal-workspace) sigma plan, copy
// them into the batch's shared arrays and drop the plan.
int batch_init(fcis_batch *b, int nprob, const fcis_space *const *spaces, const double *const *h2e,
               const int *nvec, int precision, hipStream_t stream)
{
        b->nprob = nprob;
        b->precision = precision;
        b->use_graph = 1;
        const int kstep = precision == FCIS_PREC_BF16X3 ? 32 : 4;
        const size_t fb = frag_elem_bytes(precision);
        int64_t off = 0, npa = 0, npb = 0, nhf = 0;
        int aoff = 0, maxmt = 0, maxks = 0;
        for (int p = 0; p < nprob; p++) {
                const fcis_space *s = spaces[p];
                if (s == nullptr || h2e[p] == nullptr || nvec[p] < 1 || nvec[p] > 64) {
                        fcis::set_error("fcis_batch_create: problem %d: invalid space, h2e or nvec", p);
                        return FCIS_ERR_INVALID;
                }
                if (s->nstra * s->nstrb == 0 || s->norb > 16) {
                        fcis::set_error("fcis_batch_create: problem %d (norb %d) outside the batched "
                                        "kernel's range (norb <= 16, nonempty)", p, s->norb);
                        return FCIS_ERR_INVALID;
                }
                BatchProb P;
                P.npair = s->npair;
                P.mtiles = (s->npair + 15) / 16;
                P.ksteps = (s->npair + kstep - 1) / kstep;
                P.nb = (int)s->nstrb;
                P.nstra = (int)s->nstra;
                P.nbp = (int)((s->nstrb + FUSED_NB - 1) / FUSED_NB * FUSED_NB);
                P.ndet = s->nstra * s->nstrb;
                P.nvec = nvec[p];
                P.voff = off;
                P.aoff = aoff;
                P.pa = npa;
                P.pb = npb;
                P.hf = nhf;
                npa += (int64_t)P.npair * P.nstra;
                npb += (int64_t)P.npair * P.nbp;
                nhf += (int64_t)P.ksteps * P.mtiles * 64;
                b->probs.push_back(P);
                b->offs.push_back(off);
                b->cost.push_back((double)P.nb * P.mtiles * P.ksteps * P.nvec + 64.0 * P.nvec);
                off += P.ndet * P.nvec;
                aoff += P.nvec;
                maxmt = std::max(maxmt, P.mtiles);
                maxks = std::max(maxks, P.ksteps);
                b->nbpmax = std::max(b->nbpmax, (int64_t)P.nbp);
        }
        b->total = off;
        b->nslots = aoff;
        if (fcis::dev_malloc(&b->d_pmapa, 4 * (size_t)npa + 16) != hipSuccess ||
            fcis::dev_malloc(&b->d_pmapb, 2 * (size_t)npb + 16) != hipSuccess ||
            fcis::dev_malloc(&b->d_hfrag, fb * (size_t)nhf + 16) != hipSuccess ||
            fcis::dev_malloc(&b->d_hfrag_lo, precision == FCIS_PREC_BF16X3 ? fb * (size_t)nhf + 16 : 16) != hipSuccess ||
            fcis::dev_malloc(&b->d_probs, sizeof(BatchProb) * nprob) != hipSuccess ||
            fcis::dev_malloc(&b->d_counter, 256) != hipSuccess ||
            fcis::dev_malloc(&b->d_active, (size_t)b->nslots + 16) != hipSuccess) {
                fcis::set_error("fcis_batch_create: device allocation failed");
                return FCIS_ERR_MEMORY;
        }
        for (int p = 0; p < nprob; p++) {
                const BatchProb &P = b->probs[p];
                fcis_sigma_plan *pl = nullptr;
                int st = fcis::sigma_plan_create(&pl, spaces[p], h2e[p], precision, 1, false);
                if (st != FCIS_SUCCESS) return st;
                const bool ok = pl->fused && pl->mtiles == P.mtiles && pl->ksteps == P.ksteps &&
                                pl->nbp == P.nbp;
                hipError_t e = hipSuccess;
                if (ok) {
                        const size_t nf = (size_t)P.ksteps * P.mtiles * 64;
                        e = hipMemcpy(b->d_pmapa + P.pa, pl->d_pmapa, 4 * (size_t)P.npair * P.nstra,
                                      hipMemcpyDeviceToDevice);
                        if (e == hipSuccess)
                                e = hipMemcpy(b->d_pmapb + P.pb, pl->d_pmapb16, 2 * (size_t)P.npair * P.nbp,
                                              hipMemcpyDeviceToDevice);
                        if (e == hipSuccess)
                                e = hipMemcpy((char *)b->d_hfrag + fb * P.hf, pl->d_hfrag, fb * nf,
                                              hipMemcpyDeviceToDevice);
                        if (e == hipSuccess && precision == FCIS_PREC_BF16X3)
                                e = hipMemcpy((char *)b->d_hfrag_lo + fb * P.hf, pl->d_hfrag_lo, fb * nf,
                                              hipMemcpyDeviceToDevice);
                }
                fcis_sigma_plan_free(pl);
                if (!ok) {
                        fcis::set_error("fcis_batch_create: problem %d has no fused sigma tables", p);
                        return FCIS_ERR_INVALID;
                }
                FCIS_HIP(e);
        }
        const int small_ks = precision == FCIS_PREC_BF16X3 ? BATCH_SMALL_KS_BF16 : BATCH_SMALL_KS;
        b->variant = (maxmt <= BATCH_SMALL_MAXW && maxks <= small_ks) ? 0 : 1;
        b->lds = batch_lds(precision, b->variant, b->nbpmax);
        if (b->lds > 160 * 1024) {
                fcis::set_error("fcis_batch_create: LDS need %zu exceeds 160 KiB", b->lds);
                return FCIS_ERR_INVALID;
        }
        int dev = 0;
        FCIS_HIP(hipGetDevice(&dev));
        FCIS_HIP(hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, dev));
        const void *k = batch_kernel(precision, b->variant);
        FCIS_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)b->lds));
        FCIS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b->occ, k,
                                                              batch_threads(precision, b->variant), b->lds));
        if (b->occ < 1) b->occ = 1;
        b->grid = b->ncu * b->occ;
        FCIS_HIP(hipMemcpy(b->d_probs, b->probs.data(), sizeof(BatchProb) * nprob, hipMemcpyHostToDevice));
        std::vector<uint8_t> ones(b->nslots, 1);
        FCIS_HIP(hipMemcpy(b->d_active, ones.data(), b->nslots, hipMemcpyHostToDevice));
        FCIS_HIP(hipStreamCreate(&b->cap_stream));
        (void)stream;
        return build_items(b, 1);
}
}  // namespace

extern "C" {

int fcis_batch_create(fcis_batch_t **out, int nprob, const fcis_space_t *const *spaces,
                      const double *const *h2e, const int *nvec, int precision, void *stream)
{
        if (out == nullptr || nprob < 1 || spaces == nullptr || h2e == nullptr || nvec == nullptr) {
                fcis::set_error("fcis_batch_create: invalid argument");
                return FCIS_ERR_INVALID;
        }
        if (precision != FCIS_PREC_FP64 && precision != FCIS_PREC_FP32 &&
            precision != FCIS_PREC_BF16X3) {
                fcis::set_error("fcis_batch_create: unknown precision %d", precision);
                return FCIS_ERR_INVALID;
        }
        *out = nullptr;
        fcis_batch *b = new (std::nothrow) fcis_batch();
        if (b == nullptr) return FCIS_ERR_MEMORY;
        int st = batch_init(b, nprob, spaces, h2e, nvec, precision, (hipStream_t)stream);
        if (st != FCIS_SUCCESS) {
                fcis_batch_free(b);
                return st;
        }
        *out = b;
        return FCIS_SUCCESS;
}

void fcis_batch_free(fcis_batch_t *b)
{
        if (b == nullptr) return;
        drop_graph(b);
        (void)hipFree(b->d_pmapa);
        (void)hipFree(b->d_pmapb);
        (void)hipFree(b->d_hfrag);
        (void)hipFree(b->d_hfrag_lo);
        (void)hipFree(b->d_probs);
        (void)hipFree(b->d_items);
        (void)hipFree(b->d_counter);
        (void)hipFree(b->d_active);
        if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
        delete b;
}

int fcis_batch_info(const fcis_batch_t *b, int64_t *total, int64_t *offsets, int *rows_per_item,
                    int *grid)
{
        if (b == nullptr) return FCIS_ERR_INVALID;
        if (total) *total = b->total;
        if (offsets)
                for (int p = 0; p < b->nprob; p++) offsets[p] = b->offs[p];
        if (rows_per_item) *rows_per_item = b->rows_per_item;
        if (grid) *grid = b->grid;
        return FCIS_SUCCESS;
}

int fcis_batch_set_active(fcis_batch_t *b, const uint8_t *mask, void *stream)
{
        if (b == nullptr || mask == nullptr) {
                fcis::set_error("fcis_batch_set_active: invalid argument");
                return FCIS_ERR_INVALID;
        }
        hipStream_t st = (hipStream_t)stream;
        FCIS_HIP(hipStreamSynchronize(st));
        FCIS_HIP(hipMemcpy(b->d_active, mask, b->nslots, hipMemcpyHostToDevice));
        return FCIS_SUCCESS;
}

int fcis_batch_set_graph(fcis_batch_t *b, int enable)
{
        if (b == nullptr) return FCIS_ERR_INVALID;
        b->use_graph = enable ? 1 : 0;
        if (!enable) drop_graph(b);
        return FCIS_SUCCESS;
}

int fcis_batch_sigma(fcis_batch_t *b, const double *d_in, double *d_out, void *stream)
{
        if (b == nullptr || d_in == nullptr || d_out == nullptr || d_in == d_out) {
                fcis::set_error("fcis_batch_sigma: invalid or aliased buffers");
                return FCIS_ERR_INVALID;
        }
        hipStream_t st = (hipStream_t)stream;
        if (!b->use_graph) return enqueue(b, d_in, d_out, st);
        if (b->exec == nullptr || b->g_in != d_in || b->g_out != d_out) {
                int s = capture(b, d_in, d_out);
                if (s != FCIS_SUCCESS) return s;
        }
        FCIS_HIP(hipGraphLaunch(b->exec, st));
        return FCIS_SUCCESS;
}

int fcis_batch_tune(fcis_batch_t *b, const double *d_in, double *d_out, int reps, double *best_ms,
                    void *stream)
{
        if (b == nullptr || d_in == nullptr || d_out == nullptr || d_in == d_out) {
                fcis::set_error("fcis_batch_tune: invalid or aliased buffers");
                return FCIS_ERR_INVALID;
        }
        if (reps < 1) reps = 1;
        hipStream_t st = (hipStream_t)stream;
        drop_graph(b);
        hipEvent_t e0, e1;
        FCIS_HIP(hipEventCreate(&e0));
        FCIS_HIP(hipEventCreate(&e1));
        double best = 1e300;
        int brow = 1, bgrid = b->occ * b->ncu;
        int status = FCIS_SUCCESS;
        // time one schedule: min over reps
        auto trial = [&](int rows, int grid) {
                if (status != FCIS_SUCCESS) return;
                status = build_items(b, rows);
                b->grid = grid;
                float ms = 0.0f;
                double tmin = 1e300;
                for (int r = 0; r < reps && status == FCIS_SUCCESS; r++) {
                        if (hipEventRecord(e0, st) != hipSuccess) status = FCIS_ERR_DEVICE;
                        if (status == FCIS_SUCCESS) status = enqueue(b, d_in, d_out, st);
                        if (status == FCIS_SUCCESS && (hipEventRecord(e1, st) != hipSuccess ||
                                                       hipEventSynchronize(e1) != hipSuccess ||
                                                       hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                                status = FCIS_ERR_DEVICE;
                        tmin = std::min(tmin, (double)ms);
                }
                if (status == FCIS_SUCCESS && tmin < best) {
                        best = tmin;
                        brow = rows;
                        bgrid = grid;
                }
        };
        // warm-up, then alpha strings per item at full residency, then the
        // grid size at the best item size
        status = enqueue(b, d_in, d_out, st);
        if (status == FCIS_SUCCESS && hipStreamSynchronize(st) != hipSuccess) status = FCIS_ERR_DEVICE;
        for (int rows : {1, 2, 4, 8}) trial(rows, b->occ * b->ncu);
        const int rbest = brow;
        for (int g : {b->ncu, 2 * b->ncu})
                if (g < b->occ * b->ncu) trial(rbest, g);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (status != FCIS_SUCCESS) {
                if (status == FCIS_ERR_DEVICE) fcis::set_error("fcis_batch_tune: device error while timing");
                return status;
        }
        b->grid = bgrid;
        b->tuned_ms = best;
        if (best_ms) *best_ms = best;
        return build_items(b, brow);
}

}  // extern "C"
