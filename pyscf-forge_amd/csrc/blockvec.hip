// Block vector algebra over many device vectors of different lengths, one
// launch per call: the Davidson work of many solves at once (subspace
// projections, Ritz vectors, residuals, Gram-Schmidt, preconditioning;
// pyscf.lib.davidson1's dot / axpy / precond, fci_gas.py:1337-1350).
//
// An entry names a block of nx vectors X and a block of ny vectors Y of
// one length n through pointer pools.  gram computes X^T Y, gemm computes
// Y = beta Y + X C, both streaming every vector once per (16-wide or
// 4-wide) output tile instead of once per output.  Work is split into
// tasks of LIST_CHUNK elements so that long and short vectors balance;
// gram partial sums are added on the host in task order (reproducible).
#include <algorithm>
#include <math.h>
#include <mutex>
#include <string.h>
#include <vector>

#include "common.h"

namespace {

constexpr int LIST_THREADS = 256;
constexpr int64_t LIST_CHUNK = 8192;    // elements per task (~2k tasks for 13M elements)
constexpr int GRAM_TI = 8;              // X rows per gram task
constexpr int GRAM_TJ = 4;              // Y columns per gram task
constexpr int GEMM_TJ = 16;             // Y outputs per gemm task (in-place safe below this)
constexpr int BLOCK_MAXX = 128;         // nx limit (gemm coefficients staged in LDS)

struct BlockTask {
        int32_t entry, i0, j0, pad;
        int64_t start, len;
};

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        return v;
}

// part[task][ii][jj] = sum_{t in chunk} X[i0+ii][t] Y[j0+jj][t]
__global__ __launch_bounds__(LIST_THREADS) void k_gram_list(const BlockTask *__restrict__ tasks,
                                                            const fcis_block_entry *__restrict__ ent,
                                                            const double *const *__restrict__ xs,
                                                            const double *const *__restrict__ ys,
                                                            double *__restrict__ part)
{
        __shared__ double red[LIST_THREADS / 64][GRAM_TI * GRAM_TJ];
        const BlockTask T = tasks[blockIdx.x];
        const fcis_block_entry E = ent[T.entry];
        const int ni = min(GRAM_TI, E.nx - T.i0), nj = min(GRAM_TJ, E.ny - T.j0);
        const double *xp[GRAM_TI], *yp[GRAM_TJ];
#pragma unroll
        for (int a = 0; a < GRAM_TI; a++) xp[a] = a < ni ? xs[E.xoff + T.i0 + a] + T.start : nullptr;
#pragma unroll
        for (int b = 0; b < GRAM_TJ; b++) yp[b] = b < nj ? ys[E.yoff + T.j0 + b] + T.start : nullptr;
        double acc[GRAM_TI][GRAM_TJ];
#pragma unroll
        for (int a = 0; a < GRAM_TI; a++)
#pragma unroll
                for (int b = 0; b < GRAM_TJ; b++) acc[a][b] = 0.0;
        for (int64_t k = threadIdx.x; k < T.len; k += LIST_THREADS) {
                double xv[GRAM_TI], yv[GRAM_TJ];
#pragma unroll
                for (int a = 0; a < GRAM_TI; a++) xv[a] = a < ni ? xp[a][k] : 0.0;
#pragma unroll
                for (int b = 0; b < GRAM_TJ; b++) yv[b] = b < nj ? yp[b][k] : 0.0;
#pragma unroll
                for (int a = 0; a < GRAM_TI; a++)
#pragma unroll
                        for (int b = 0; b < GRAM_TJ; b++) acc[a][b] += xv[a] * yv[b];
        }
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
        for (int a = 0; a < GRAM_TI; a++)
#pragma unroll
                for (int b = 0; b < GRAM_TJ; b++) {
                        double v = wave_sum(acc[a][b]);
                        if (lane == 0) red[wave][a * GRAM_TJ + b] = v;
                }
        __syncthreads();
        if (threadIdx.x < GRAM_TI * GRAM_TJ) {
                double v = 0.0;
                for (int w = 0; w < LIST_THREADS / 64; w++) v += red[w][threadIdx.x];
                part[(int64_t)blockIdx.x * GRAM_TI * GRAM_TJ + threadIdx.x] = v;
        }
}

// One output tile of the block product, NJ outputs wide; every thread
// handles UNR elements per pass so that UNR independent loads of each X_i
// are in flight.
template <int NJ, int UNR>
__device__ __forceinline__ void gemm_tile(const BlockTask &T, const fcis_block_entry &E, int nj,
                                          const double *const *xb, double *const *yb,
                                          const double *cs)
{
        double *yp[NJ];
#pragma unroll
        for (int b = 0; b < NJ; b++) yp[b] = b < nj ? yb[b] + T.start : nullptr;
        const int nx = E.nx;
        for (int64_t k0 = threadIdx.x; k0 < T.len; k0 += (int64_t)LIST_THREADS * UNR) {
                bool in[UNR];
#pragma unroll
                for (int u = 0; u < UNR; u++) in[u] = k0 + u * LIST_THREADS < T.len;
                double acc[UNR][NJ];
#pragma unroll
                for (int u = 0; u < UNR; u++)
#pragma unroll
                        for (int b = 0; b < NJ; b++)
                                acc[u][b] = (b < nj && in[u] && E.beta != 0.0)
                                                    ? E.beta * yp[b][k0 + u * LIST_THREADS] : 0.0;
#pragma unroll 2
                for (int i = 0; i < nx; i++) {
                        const double *x = xb[i] + T.start + k0;
                        double xv[UNR];
#pragma unroll
                        for (int u = 0; u < UNR; u++) xv[u] = in[u] ? x[u * LIST_THREADS] : 0.0;
#pragma unroll
                        for (int b = 0; b < NJ; b++) {
                                const double c = cs[i * GEMM_TJ + b];
#pragma unroll
                                for (int u = 0; u < UNR; u++) acc[u][b] += c * xv[u];
                        }
                }
#pragma unroll
                for (int u = 0; u < UNR; u++)
#pragma unroll
                        for (int b = 0; b < NJ; b++)
                                if (b < nj && in[u]) yp[b][k0 + u * LIST_THREADS] = acc[u][b];
        }
}

// Y[j0+jj][t] = beta Y[j0+jj][t] + sum_i C[i][j0+jj] X[i][t] for the
// chunk; every thread reads all X[i][t] of its elements before writing its
// Y[.][t], so Y may alias X when ny <= GEMM_TJ (one output tile per chunk).
__global__ __launch_bounds__(LIST_THREADS) void k_gemm_list(const BlockTask *__restrict__ tasks,
                                                            const fcis_block_entry *__restrict__ ent,
                                                            const double *const *__restrict__ xs,
                                                            double *const *__restrict__ ys,
                                                            const double *__restrict__ coef)
{
        __shared__ double cs[BLOCK_MAXX * GEMM_TJ];
        const BlockTask T = tasks[blockIdx.x];
        const fcis_block_entry E = ent[T.entry];
        const int nj = min(GEMM_TJ, E.ny - T.j0), nx = E.nx;
        for (int e = threadIdx.x; e < nx * GEMM_TJ; e += LIST_THREADS) {
                int i = e / GEMM_TJ, jj = e - i * GEMM_TJ;
                cs[e] = jj < nj ? coef[E.ooff + (int64_t)i * E.ny + T.j0 + jj] : 0.0;
        }
        __syncthreads();
        const double *const *xb = xs + E.xoff;
        double *const *yb = ys + E.yoff + T.j0;
        if (nj <= 1)
                gemm_tile<1, 8>(T, E, nj, xb, yb, cs);
        else if (nj <= 2)
                gemm_tile<2, 8>(T, E, nj, xb, yb, cs);
        else if (nj <= 4)
                gemm_tile<4, 4>(T, E, nj, xb, yb, cs);
        else if (nj <= 8)
                gemm_tile<8, 4>(T, E, nj, xb, yb, cs);
        else
                gemm_tile<GEMM_TJ, 2>(T, E, nj, xb, yb, cs);
}

__global__ __launch_bounds__(LIST_THREADS) void k_precond_list(const BlockTask *__restrict__ tasks,
                                                               const fcis_precond_entry *__restrict__ ent,
                                                               double shift)
{
        const BlockTask T = tasks[blockIdx.x];
        const fcis_precond_entry E = ent[T.entry];
        for (int64_t k = threadIdx.x; k < T.len; k += LIST_THREADS) {
                const int64_t i = T.start + k;
                double d = E.hdiag[i] - E.e;
                if (fabs(d) < shift) d = shift;
                E.out[i] = E.r[i] / d;
        }
}

// Pinned host staging + device table buffers of the list calls: a ring of
// STAGE_SLOTS per device (grown on demand; calls are serialised per
// process).  A call packs its tables into the next slot, copies them to the
// device and launches on the caller's stream; gemm and preconditioner calls
// then record the slot's event and return without waiting, so that the
// host prepares the next call while the device runs this one.  A slot is
// written again only after its event has completed (the copy and the
// kernel that read it are done).  gram calls wait for their results.
constexpr int STAGE_SLOTS = 8;
struct Stage {
        char *host = nullptr, *dev = nullptr;
        size_t cap = 0;
        hipEvent_t done = nullptr;   // recorded after the slot's last kernel
        bool pending = false;
};
Stage g_stage[16][STAGE_SLOTS];
int g_next[16];
std::mutex g_mu;

// The next slot of device dev, its previous use finished.
int stage_acquire(int dev, Stage **out)
{
        Stage &S = g_stage[dev & 15][g_next[dev & 15]];
        g_next[dev & 15] = (g_next[dev & 15] + 1) % STAGE_SLOTS;
        if (S.done == nullptr) FCIS_HIP(hipEventCreateWithFlags(&S.done, hipEventDisableTiming));
        if (S.pending) FCIS_HIP(hipEventSynchronize(S.done));
        S.pending = false;
        *out = &S;
        return FCIS_SUCCESS;
}

// Mark the slot in use by the work just queued on stream st.
int stage_release_async(Stage &S, hipStream_t st)
{
        FCIS_HIP(hipEventRecord(S.done, st));
        S.pending = true;
        return FCIS_SUCCESS;
}

int stage_reserve(Stage &s, size_t bytes)
{
        if (bytes <= s.cap) return FCIS_SUCCESS;
        size_t cap = std::max(bytes + bytes / 2, (size_t)1 << 20);
        if (s.host) (void)hipHostFree(s.host);
        if (s.dev) (void)hipFree(s.dev);
        s.host = s.dev = nullptr;
        s.cap = 0;
        if (hipHostMalloc((void **)&s.host, cap) != hipSuccess ||
            fcis::dev_malloc((void **)&s.dev, cap) != hipSuccess) {
                fcis::set_error("block vector ops: staging allocation of %zu bytes failed", cap);
                return FCIS_ERR_MEMORY;
        }
        s.cap = cap;
        return FCIS_SUCCESS;
}

size_t align_up(size_t x)
{
        return (x + 255) & ~(size_t)255;
}

// Packs consecutive host arrays into the staging buffer.
struct Packer {
        std::vector<std::pair<const void *, size_t>> parts;
        std::vector<size_t> offs;
        size_t total = 0;
        size_t add(const void *p, size_t bytes)
        {
                size_t o = total;
                parts.push_back({p, bytes});
                offs.push_back(o);
                total = align_up(o + bytes);
                return o;
        }
        void copy(char *dst) const
        {
                for (size_t k = 0; k < parts.size(); k++)
                        if (parts[k].first && parts[k].second)   // nullptr: output-only area
                                memcpy(dst + offs[k], parts[k].first, parts[k].second);
        }
};

int check_entries(const char *who, int nent, const fcis_block_entry *ent, int nxs, int nys,
                  int nout, bool gemm)
{
        for (int e = 0; e < nent; e++) {
                const fcis_block_entry &E = ent[e];
                if (E.n < 0 || E.nx < 0 || E.ny < 0 || E.xoff < 0 || E.yoff < 0 || E.ooff < 0 ||
                    E.xoff + E.nx > nxs || E.yoff + E.ny > nys ||
                    (int64_t)E.ooff + (int64_t)E.nx * E.ny > nout || (gemm && E.nx > BLOCK_MAXX)) {
                        fcis::set_error("%s: entry %d out of range", who, e);
                        return FCIS_ERR_INVALID;
                }
        }
        return FCIS_SUCCESS;
}

int null_pointers(const char *who, int n, const void *const *p)
{
        for (int k = 0; k < n; k++)
                if (p[k] == nullptr) {
                        fcis::set_error("%s: pointer %d is NULL", who, k);
                        return FCIS_ERR_INVALID;
                }
        return FCIS_SUCCESS;
}

}  // namespace

extern "C" {

int fcis_gram_list(int nent, const fcis_block_entry *ent, int nxs, const double *const *xs, int nys,
                   const double *const *ys, int nout, double *out, void *stream)
{
        if (nent < 0 || (nent > 0 && (ent == nullptr || xs == nullptr || ys == nullptr ||
                                      out == nullptr))) {
                fcis::set_error("fcis_gram_list: invalid argument");
                return FCIS_ERR_INVALID;
        }
        int st = check_entries("fcis_gram_list", nent, ent, nxs, nys, nout, false);
        if (st == FCIS_SUCCESS) st = null_pointers("fcis_gram_list", nxs, (const void *const *)xs);
        if (st == FCIS_SUCCESS) st = null_pointers("fcis_gram_list", nys, (const void *const *)ys);
        if (st != FCIS_SUCCESS) return st;
        for (int k = 0; k < nout; k++) out[k] = 0.0;
        std::vector<BlockTask> tasks;
        for (int e = 0; e < nent; e++)
                for (int i0 = 0; i0 < ent[e].nx; i0 += GRAM_TI)
                        for (int j0 = 0; j0 < ent[e].ny; j0 += GRAM_TJ)
                                for (int64_t s = 0; s < ent[e].n; s += LIST_CHUNK)
                                        tasks.push_back(BlockTask{e, i0, j0, 0, s,
                                                                  std::min(LIST_CHUNK, ent[e].n - s)});
        if (tasks.empty()) return FCIS_SUCCESS;
        int dev = 0;
        FCIS_HIP(hipGetDevice(&dev));
        std::lock_guard<std::mutex> lock(g_mu);
        Stage *Sp = nullptr;
        st = stage_acquire(dev, &Sp);
        if (st != FCIS_SUCCESS) return st;
        Stage &S = *Sp;
        Packer pk;
        const size_t o_t = pk.add(tasks.data(), tasks.size() * sizeof(BlockTask));
        const size_t o_e = pk.add(ent, nent * sizeof(fcis_block_entry));
        const size_t o_x = pk.add(xs, nxs * sizeof(double *));
        const size_t o_y = pk.add(ys, nys * sizeof(double *));
        const size_t npart = tasks.size() * GRAM_TI * GRAM_TJ;
        const size_t o_p = pk.add(nullptr, npart * sizeof(double));
        st = stage_reserve(S, pk.total);
        if (st != FCIS_SUCCESS) return st;
        hipStream_t s = (hipStream_t)stream;
        pk.copy(S.host);
        FCIS_HIP(hipMemcpyAsync(S.dev, S.host, o_p, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_gram_list, dim3(tasks.size()), dim3(LIST_THREADS), 0, s,
                           (const BlockTask *)(S.dev + o_t), (const fcis_block_entry *)(S.dev + o_e),
                           (const double *const *)(S.dev + o_x), (const double *const *)(S.dev + o_y),
                           (double *)(S.dev + o_p));
        FCIS_CHECK_LAUNCH();
        FCIS_HIP(hipMemcpyAsync(S.host + o_p, S.dev + o_p, npart * sizeof(double),
                                hipMemcpyDeviceToHost, s));
        FCIS_HIP(hipStreamSynchronize(s));
        const double *part = (const double *)(S.host + o_p);
        for (size_t t = 0; t < tasks.size(); t++) {
                const BlockTask &T = tasks[t];
                const fcis_block_entry &E = ent[T.entry];
                const int ni = std::min(GRAM_TI, E.nx - T.i0), nj = std::min(GRAM_TJ, E.ny - T.j0);
                for (int a = 0; a < ni; a++)
                        for (int b = 0; b < nj; b++)
                                out[E.ooff + (int64_t)(T.i0 + a) * E.ny + T.j0 + b] +=
                                        part[t * GRAM_TI * GRAM_TJ + a * GRAM_TJ + b];
        }
        return FCIS_SUCCESS;
}

int fcis_gemm_list(int nent, const fcis_block_entry *ent, int nxs, const double *const *xs, int nys,
                   double *const *ys, int ncoef, const double *coef, void *stream)
{
        if (nent < 0 || (nent > 0 && (ent == nullptr || xs == nullptr || ys == nullptr ||
                                      (ncoef > 0 && coef == nullptr)))) {
                fcis::set_error("fcis_gemm_list: invalid argument");
                return FCIS_ERR_INVALID;
        }
        int st = check_entries("fcis_gemm_list", nent, ent, nxs, nys, ncoef, true);
        if (st == FCIS_SUCCESS) st = null_pointers("fcis_gemm_list", nxs, (const void *const *)xs);
        if (st == FCIS_SUCCESS) st = null_pointers("fcis_gemm_list", nys, (const void *const *)ys);
        if (st != FCIS_SUCCESS) return st;
        std::vector<BlockTask> tasks;
        for (int e = 0; e < nent; e++)
                for (int j0 = 0; j0 < ent[e].ny; j0 += GEMM_TJ)
                        for (int64_t s = 0; s < ent[e].n; s += LIST_CHUNK)
                                tasks.push_back(BlockTask{e, 0, j0, 0, s, std::min(LIST_CHUNK, ent[e].n - s)});
        if (tasks.empty()) return FCIS_SUCCESS;
        int dev = 0;
        FCIS_HIP(hipGetDevice(&dev));
        std::lock_guard<std::mutex> lock(g_mu);
        Stage *Sp = nullptr;
        st = stage_acquire(dev, &Sp);
        if (st != FCIS_SUCCESS) return st;
        Stage &S = *Sp;
        Packer pk;
        const size_t o_t = pk.add(tasks.data(), tasks.size() * sizeof(BlockTask));
        const size_t o_e = pk.add(ent, nent * sizeof(fcis_block_entry));
        const size_t o_x = pk.add(xs, nxs * sizeof(double *));
        const size_t o_y = pk.add(ys, nys * sizeof(double *));
        const size_t o_c = pk.add(coef, ncoef * sizeof(double));
        st = stage_reserve(S, pk.total);
        if (st != FCIS_SUCCESS) return st;
        hipStream_t s = (hipStream_t)stream;
        pk.copy(S.host);
        FCIS_HIP(hipMemcpyAsync(S.dev, S.host, pk.total, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_gemm_list, dim3(tasks.size()), dim3(LIST_THREADS), 0, s,
                           (const BlockTask *)(S.dev + o_t), (const fcis_block_entry *)(S.dev + o_e),
                           (const double *const *)(S.dev + o_x), (double *const *)(S.dev + o_y),
                           (const double *)(S.dev + o_c));
        FCIS_CHECK_LAUNCH();
        return stage_release_async(S, s);
}

int fcis_precond_list(int nent, const fcis_precond_entry *ent, double level_shift, void *stream)
{
        if (nent < 0 || (nent > 0 && ent == nullptr)) {
                fcis::set_error("fcis_precond_list: invalid argument");
                return FCIS_ERR_INVALID;
        }
        std::vector<BlockTask> tasks;
        for (int e = 0; e < nent; e++) {
                const fcis_precond_entry &E = ent[e];
                if (E.n < 0 || (E.n > 0 && (E.hdiag == nullptr || E.r == nullptr || E.out == nullptr))) {
                        fcis::set_error("fcis_precond_list: entry %d invalid", e);
                        return FCIS_ERR_INVALID;
                }
                for (int64_t s = 0; s < E.n; s += LIST_CHUNK)
                        tasks.push_back(BlockTask{e, 0, 0, 0, s, std::min(LIST_CHUNK, E.n - s)});
        }
        if (tasks.empty()) return FCIS_SUCCESS;
        int dev = 0;
        FCIS_HIP(hipGetDevice(&dev));
        std::lock_guard<std::mutex> lock(g_mu);
        Stage *Sp = nullptr;
        int st = stage_acquire(dev, &Sp);
        if (st != FCIS_SUCCESS) return st;
        Stage &S = *Sp;
        Packer pk;
        const size_t o_t = pk.add(tasks.data(), tasks.size() * sizeof(BlockTask));
        const size_t o_e = pk.add(ent, nent * sizeof(fcis_precond_entry));
        st = stage_reserve(S, pk.total);
        if (st != FCIS_SUCCESS) return st;
        hipStream_t s = (hipStream_t)stream;
        pk.copy(S.host);
        FCIS_HIP(hipMemcpyAsync(S.dev, S.host, pk.total, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_precond_list, dim3(tasks.size()), dim3(LIST_THREADS), 0, s,
                           (const BlockTask *)(S.dev + o_t), (const fcis_precond_entry *)(S.dev + o_e),
                           level_shift);
        FCIS_CHECK_LAUNCH();
        return stage_release_async(S, s);
}

}  // extern "C"
