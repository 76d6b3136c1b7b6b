// sigma = H c on the device: gather -> MFMA contraction -> pull scatter.
//
// For the absorbed pair matrix h (npair x npair) and the symmetric pair
// operators F_pq = E_pq + E_qp (p != q), F_pp = E_pp, E = E^alpha + E^beta:
//
//   t1[pq, I, b] = sum_{alpha links (pq, J, s) of I} s c[J, b]
//                + sum_{beta links (pq, J, s) of b}  s c[I, J]      (gather)
//   vt1[rs, d]   = sum_pq h[rs, pq] t1[pq, d]                       (MFMA)
//   sigma[J, Jb] = sum_{alpha links (pq, I, s) of J} s vt1[pq, I, Jb]
//                + sum_{beta links (pq, b, s) of Jb} s vt1[pq, J, b]  (scatter)
//
// This is the t1 / dgemm / spread formulation of pyscf's
// direct_spin1.contract_2e (libfci FCIcontract_2e_spin1, called by the
// reference at pyscf/csf_fci/csf.py:449), reorganised for the GPU: the
// scatter is written destination-owned ("pull": every sigma element is
// produced by exactly one thread, as libfci_gas's tasks own destination
// alpha ranges, pyscf/lib/mcscf/fci_contract_gas.c:2051-2069, :2360-2386),
// so sigma needs no atomics and is bitwise reproducible.  Pulling through a
// link of the destination uses a_i^+ a_a |J> = s |I>  <=>  a_a^+ a_i |I> = s |J>.
//
// t1 and vt1 live in HBM, tiled over alpha strings so that a tile fits the
// plan's workspace budget (the reference tiles its t1 under a byte budget
// too, fci_contract_gas.c:55-71).
#include <stdlib.h>
#include <string.h>
#include <vector>

#include "common.h"

namespace {

constexpr int GEMM_WAVES = 4;           // wavefronts per workgroup
constexpr int GEMM_NSUB = 2;            // 16-column subtiles per wavefront
constexpr int GEMM_MT = 9;              // 16-row tiles per workgroup (144 rows)
constexpr int GEMM_NCOL = GEMM_WAVES * GEMM_NSUB * 16;  // 128 dets per workgroup

typedef double v4d __attribute__((ext_vector_type(4)));

// Gather: one thread per determinant (I, b), loop over all pairs pq so that
// every store to t1[pq, :] is coalesced across the wavefront.  pmap[pq][str]
// holds (target << 1 | negative) for the (unique) link of str with pair pq,
// or -1.  Alpha lookups are uniform across the workgroup (one I per block).
template <typename T>
__global__ __launch_bounds__(256) void k_gather(int npair, int64_t nb, int64_t a0,
                                                int64_t nstra, int64_t ld,
                                                const int32_t *__restrict__ pmapa,
                                                const int32_t *__restrict__ pmapb,
                                                const double *__restrict__ c,
                                                T *__restrict__ t1)
{
        const int64_t nchunk = (nb + 255) / 256;
        const int64_t row = blockIdx.x / nchunk;
        int64_t b = (blockIdx.x - row * nchunk) * 256 + threadIdx.x;
        int64_t I = a0 + row;
        if (b >= nb) return;
        const double *crow = c + I * nb;
        int64_t d = row * nb + b;
        for (int pq = 0; pq < npair; pq++) {
                double v = 0.0;
                int32_t ea = pmapa[(int64_t)pq * nstra + I];
                if (ea >= 0) {
                        double x = c[(int64_t)(ea >> 1) * nb + b];
                        v = (ea & 1) ? -x : x;
                }
                int32_t eb = pmapb[(int64_t)pq * nb + b];
                if (eb >= 0) {
                        double x = crow[eb >> 1];
                        v += (eb & 1) ? -x : x;
                }
                t1[(int64_t)pq * ld + d] = (T)v;
        }
}

// FP64 MFMA contraction vt1 = h . t1 with v_mfma_f64_16x16x4_f64.
// hfrag holds h pre-arranged per (k-step, 16-row tile, lane) so that every
// A-fragment load is one contiguous 512-byte read:
//   hfrag[((ks * mtiles + mt) * 64) + lane] = h[mt*16 + lane%16][4*ks + lane/16]
// Lane layout of the 16x16x4 f64 MFMA (measured on gfx950): A[i][k] at lane
// i + 16k, B[k][j] at lane j + 16k, D[i][j] in lane j + 16*(i%4), element
// i/4 (the f32/bf16 forms use lane j + 16*(i/4), element i%4).
__global__ __launch_bounds__(256) void k_contract_f64(int npair, int mtiles, int ksteps,
                                                      int64_t ndet, int64_t ld,
                                                      const double *__restrict__ hfrag,
                                                      const double *__restrict__ t1,
                                                      double *__restrict__ vt1)
{
        const int lane = threadIdx.x & 63;
        const int wave = threadIdx.x >> 6;
        const int mt0 = blockIdx.y * GEMM_MT;
        const int64_t col0 = (int64_t)blockIdx.x * GEMM_NCOL + wave * (GEMM_NSUB * 16);
        const int li = lane & 15, lk = lane >> 4;

        v4d acc[GEMM_MT][GEMM_NSUB];
#pragma unroll
        for (int m = 0; m < GEMM_MT; m++)
#pragma unroll
                for (int j = 0; j < GEMM_NSUB; j++) acc[m][j] = v4d{0, 0, 0, 0};

        for (int ks = 0; ks < ksteps; ks++) {
                int k = 4 * ks + lk;
                double bfr[GEMM_NSUB];
#pragma unroll
                for (int j = 0; j < GEMM_NSUB; j++) {
                        int64_t col = col0 + j * 16 + li;
                        bfr[j] = (k < npair && col < ndet) ? t1[(int64_t)k * ld + col] : 0.0;
                }
                const double *hf = hfrag + ((int64_t)ks * mtiles + mt0) * 64 + lane;
#pragma unroll
                for (int m = 0; m < GEMM_MT; m++) {
                        if (mt0 + m < mtiles) {
                                double a = hf[m * 64];
#pragma unroll
                                for (int j = 0; j < GEMM_NSUB; j++)
                                        acc[m][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                                                a, bfr[j], acc[m][j], 0, 0, 0);
                        }
                }
        }
#pragma unroll
        for (int m = 0; m < GEMM_MT; m++) {
                if (mt0 + m >= mtiles) continue;
#pragma unroll
                for (int j = 0; j < GEMM_NSUB; j++) {
                        int64_t col = col0 + j * 16 + li;
                        if (col >= ndet) continue;
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                                int row = (mt0 + m) * 16 + lk + 4 * r;
                                if (row < npair) vt1[(int64_t)row * ld + col] = acc[m][j][r];
                        }
                }
        }
}

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// FP32 MFMA contraction (v_mfma_f32_16x16x4_f32, same lane layout as the
// f64 form); t1 and vt1 are float, h fragments are float.
__global__ __launch_bounds__(256) void k_contract_f32(int npair, int mtiles, int ksteps,
                                                      int64_t ndet, int64_t ld,
                                                      const float *__restrict__ hfrag,
                                                      const float *__restrict__ t1,
                                                      float *__restrict__ vt1)
{
        const int lane = threadIdx.x & 63;
        const int wave = threadIdx.x >> 6;
        const int mt0 = blockIdx.y * GEMM_MT;
        const int64_t col0 = (int64_t)blockIdx.x * GEMM_NCOL + wave * (GEMM_NSUB * 16);
        const int li = lane & 15, lk = lane >> 4;

        v4f acc[GEMM_MT][GEMM_NSUB];
#pragma unroll
        for (int m = 0; m < GEMM_MT; m++)
#pragma unroll
                for (int j = 0; j < GEMM_NSUB; j++) acc[m][j] = v4f{0, 0, 0, 0};

        for (int ks = 0; ks < ksteps; ks++) {
                int k = 4 * ks + lk;
                float bfr[GEMM_NSUB];
#pragma unroll
                for (int j = 0; j < GEMM_NSUB; j++) {
                        int64_t col = col0 + j * 16 + li;
                        bfr[j] = (k < npair && col < ndet) ? t1[(int64_t)k * ld + col] : 0.0f;
                }
                const float *hf = hfrag + ((int64_t)ks * mtiles + mt0) * 64 + lane;
#pragma unroll
                for (int m = 0; m < GEMM_MT; m++) {
                        if (mt0 + m < mtiles) {
                                float a = hf[m * 64];
#pragma unroll
                                for (int j = 0; j < GEMM_NSUB; j++)
                                        acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                                                a, bfr[j], acc[m][j], 0, 0, 0);
                        }
                }
        }
#pragma unroll
        for (int m = 0; m < GEMM_MT; m++) {
                if (mt0 + m >= mtiles) continue;
#pragma unroll
                for (int j = 0; j < GEMM_NSUB; j++) {
                        int64_t col = col0 + j * 16 + li;
                        if (col >= ndet) continue;
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                                int row = (mt0 + m) * 16 + 4 * lk + r;
                                if (row < npair) vt1[(int64_t)row * ld + col] = acc[m][j][r];
                        }
                }
        }
}

// 3-term bf16 split contraction with FP32 accumulation:
//   h = h_hi + h_lo, t = t_hi + t_lo (each term rounded to bf16),
//   h t ~= h_hi t_hi + h_hi t_lo + h_lo t_hi
// on v_mfma_f32_16x16x32_bf16 (gfx950).  Lane layout: A[i][k] for
// k = 8*(lane/16) + e (e = 0..7) at lane i + 16*(lane/16), B[k][j] likewise,
// D as in the 16x16x4 forms.  t1 is read as float and split in registers;
// h fragments are pre-split on the host:
//   hhi/hlo[((ks * mtiles + mt) * 64) + lane][e] = h[mt*16 + lane%16][32*ks + 8*(lane/16) + e]
__global__ __launch_bounds__(256) void k_contract_bf16x3(int npair, int mtiles, int ksteps,
                                                         int64_t ndet, int64_t ld,
                                                         const bf16x8 *__restrict__ hhi,
                                                         const bf16x8 *__restrict__ hlo,
                                                         const float *__restrict__ t1,
                                                         float *__restrict__ vt1)
{
        const int lane = threadIdx.x & 63;
        const int wave = threadIdx.x >> 6;
        const int mt0 = blockIdx.y * GEMM_MT;
        const int64_t col0 = (int64_t)blockIdx.x * GEMM_NCOL + wave * (GEMM_NSUB * 16);
        const int li = lane & 15, lk = lane >> 4;

        v4f acc[GEMM_MT][GEMM_NSUB];
#pragma unroll
        for (int m = 0; m < GEMM_MT; m++)
#pragma unroll
                for (int j = 0; j < GEMM_NSUB; j++) acc[m][j] = v4f{0, 0, 0, 0};

        for (int ks = 0; ks < ksteps; ks++) {
                bf16x8 bh[GEMM_NSUB], bl[GEMM_NSUB];
#pragma unroll
                for (int j = 0; j < GEMM_NSUB; j++) {
                        int64_t col = col0 + j * 16 + li;
#pragma unroll
                        for (int e = 0; e < 8; e++) {
                                int k = 32 * ks + 8 * lk + e;
                                float x = (k < npair && col < ndet) ? t1[(int64_t)k * ld + col] : 0.0f;
                                __bf16 hi = (__bf16)x;
                                bh[j][e] = hi;
                                bl[j][e] = (__bf16)(x - (float)hi);
                        }
                }
                const int64_t fo = ((int64_t)ks * mtiles + mt0) * 64 + lane;
#pragma unroll
                for (int m = 0; m < GEMM_MT; m++) {
                        if (mt0 + m < mtiles) {
                                bf16x8 ah = hhi[fo + m * 64];
                                bf16x8 al = hlo[fo + m * 64];
#pragma unroll
                                for (int j = 0; j < GEMM_NSUB; j++) {
                                        acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                                al, bh[j], acc[m][j], 0, 0, 0);
                                        acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                                ah, bl[j], acc[m][j], 0, 0, 0);
                                        acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                                ah, bh[j], acc[m][j], 0, 0, 0);
                                }
                        }
                }
        }
#pragma unroll
        for (int m = 0; m < GEMM_MT; m++) {
                if (mt0 + m >= mtiles) continue;
#pragma unroll
                for (int j = 0; j < GEMM_NSUB; j++) {
                        int64_t col = col0 + j * 16 + li;
                        if (col >= ndet) continue;
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                                int row = (mt0 + m) * 16 + 4 * lk + r;
                                if (row < npair) vt1[(int64_t)row * ld + col] = acc[m][j][r];
                        }
                }
        }
}

// Pull scatter for one alpha tile [a0, a0+ta): one thread per destination
// determinant (J, Jb).  accumulate = 0 on the first tile (sigma overwritten).
template <typename T>
__global__ __launch_bounds__(256) void k_scatter(int64_t nb, int64_t nstra, int64_t a0,
                                                 int64_t ta, int64_t ld, int nlinka,
                                                 int nlinkb, int accumulate,
                                                 const fcis_link *__restrict__ linka,
                                                 const fcis_link *__restrict__ linkb,
                                                 const T *__restrict__ vt1,
                                                 double *__restrict__ sigma)
{
        const int64_t nchunk = (nb + 255) / 256;
        const int64_t J = blockIdx.x / nchunk;
        int64_t Jb = (blockIdx.x - J * nchunk) * 256 + threadIdx.x;
        if (Jb >= nb) return;
        double acc = accumulate ? sigma[J * nb + Jb] : 0.0;
        for (int k = 0; k < nlinka; k++) {
                fcis_link L = linka[(int64_t)k * nstra + J];
                int64_t I = L.addr;
                if (I < a0 || I >= a0 + ta) continue;
                double x = (double)vt1[(int64_t)link_pair(L.info) * ld + (I - a0) * nb + Jb];
                acc += link_neg(L.info) ? -x : x;
        }
        if (J >= a0 && J < a0 + ta) {
                const T *row = vt1 + (J - a0) * nb;
                for (int k = 0; k < nlinkb; k++) {
                        fcis_link L = linkb[(int64_t)k * nb + Jb];
                        double x = (double)row[(int64_t)link_pair(L.info) * ld + L.addr];
                        acc += link_neg(L.info) ? -x : x;
                }
        }
        sigma[J * nb + Jb] = acc;
}

__global__ void k_build_pmap(uint32_t nstr, int nlink, const fcis_link *__restrict__ link,
                             int32_t *__restrict__ pmap)
{
        uint32_t str = blockIdx.x * blockDim.x + threadIdx.x;
        int slot = blockIdx.y;
        if (str >= nstr) return;
        fcis_link L = link[(size_t)slot * nstr + str];
        pmap[(size_t)link_pair(L.info) * nstr + str] = (L.addr << 1) | link_neg(L.info);
}

}  // namespace

struct fcis_sigma_plan {
        const fcis_space *space;
        int precision;
        int npair, mtiles, ksteps;   // ksteps: MFMA k-steps (4 for f64/f32, 32 for bf16)
        void *d_hfrag;               // FP64 / FP32 fragments, or bf16 hi fragments
        void *d_hfrag_lo;            // bf16 lo fragments (BF16X3 only)
        int32_t *d_pmapa, *d_pmapb;
        int64_t tile_a;
        int ntiles;
        void *d_t1, *d_vt1;
        uint64_t ws_bytes;
        std::vector<double> h;
        int timing;
        hipEvent_t ev[4];
        double ms[3];
        int64_t launches[3];
};

namespace {

// float -> bf16 bits, round to nearest even (NaN kept quiet).
uint16_t bf16_bits(float x)
{
        uint32_t u;
        memcpy(&u, &x, 4);
        if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
        u += 0x7fffu + ((u >> 16) & 1u);
        return (uint16_t)(u >> 16);
}

float bf16_value(uint16_t b)
{
        uint32_t u = (uint32_t)b << 16;
        float x;
        memcpy(&x, &u, 4);
        return x;
}

size_t frag_count(const fcis_sigma_plan *p)
{
        return (size_t)p->ksteps * p->mtiles * 64;
}

int upload_h(fcis_sigma_plan *p, const double *h2e)
{
        const int np = p->npair;
        p->h.assign(h2e, h2e + (size_t)np * np);
        const size_t nf = frag_count(p);
        if (p->precision == FCIS_PREC_BF16X3) {
                // 8 consecutive k per lane; k-step of 32
                std::vector<uint16_t> hi(nf * 8), lo(nf * 8);
                for (int ks = 0; ks < p->ksteps; ks++)
                        for (int mt = 0; mt < p->mtiles; mt++)
                                for (int l = 0; l < 64; l++)
                                        for (int e = 0; e < 8; e++) {
                                                int i = mt * 16 + (l & 15);
                                                int k = 32 * ks + 8 * (l >> 4) + e;
                                                float x = (i < np && k < np)
                                                                  ? (float)h2e[(size_t)i * np + k]
                                                                  : 0.0f;
                                                uint16_t xh = bf16_bits(x);
                                                size_t o = (((size_t)ks * p->mtiles + mt) * 64 + l) * 8 + e;
                                                hi[o] = xh;
                                                lo[o] = bf16_bits(x - bf16_value(xh));
                                        }
                FCIS_HIP(hipMemcpy(p->d_hfrag, hi.data(), hi.size() * 2, hipMemcpyHostToDevice));
                FCIS_HIP(hipMemcpy(p->d_hfrag_lo, lo.data(), lo.size() * 2, hipMemcpyHostToDevice));
                return FCIS_SUCCESS;
        }
        std::vector<double> frag(nf, 0.0);
        for (int ks = 0; ks < p->ksteps; ks++)
                for (int mt = 0; mt < p->mtiles; mt++)
                        for (int l = 0; l < 64; l++) {
                                int i = mt * 16 + (l & 15), k = 4 * ks + (l >> 4);
                                if (i < np && k < np)
                                        frag[((size_t)ks * p->mtiles + mt) * 64 + l] =
                                                h2e[(size_t)i * np + k];
                        }
        if (p->precision == FCIS_PREC_FP64) {
                FCIS_HIP(hipMemcpy(p->d_hfrag, frag.data(), nf * 8, hipMemcpyHostToDevice));
        } else {
                std::vector<float> f(frag.begin(), frag.end());
                FCIS_HIP(hipMemcpy(p->d_hfrag, f.data(), nf * 4, hipMemcpyHostToDevice));
        }
        return FCIS_SUCCESS;
}

}  // namespace

extern "C" {

int fcis_sigma_plan_create(fcis_sigma_plan_t **out, const fcis_space_t *space,
                           const double *h2e, int precision, uint64_t workspace_bytes)
{
        if (out == nullptr || space == nullptr || h2e == nullptr) {
                fcis::set_error("fcis_sigma_plan_create: null argument");
                return FCIS_ERR_INVALID;
        }
        if (precision != FCIS_PREC_FP64 && precision != FCIS_PREC_FP32 &&
            precision != FCIS_PREC_BF16X3) {
                fcis::set_error("fcis_sigma_plan_create: unknown precision %d", precision);
                return FCIS_ERR_INVALID;
        }
        *out = nullptr;
        fcis_sigma_plan *p = new (std::nothrow) fcis_sigma_plan();
        if (p == nullptr) return FCIS_ERR_MEMORY;
        p->space = space;
        p->precision = precision;
        p->npair = space->npair;
        p->mtiles = (space->npair + 15) / 16;
        const int kstep = precision == FCIS_PREC_BF16X3 ? 32 : 4;
        p->ksteps = (space->npair + kstep - 1) / kstep;
        const size_t elem = precision == FCIS_PREC_FP64 ? 8 : 4;
        const int64_t nb = space->nstrb;
        if (workspace_bytes == 0) workspace_bytes = (uint64_t)16 << 30;
        // two buffers (t1, vt1) of npair x tile_a x nb elements
        uint64_t per_alpha = 2ull * p->npair * (uint64_t)nb * elem;
        int64_t ta = per_alpha ? (int64_t)(workspace_bytes / per_alpha) : space->nstra;
        if (ta < 1) ta = 1;
        if (ta > space->nstra) ta = space->nstra;
        p->tile_a = ta;
        p->ntiles = (int)((space->nstra + ta - 1) / ta);
        size_t buf = (size_t)p->npair * ta * nb * elem;
        p->ws_bytes = 2 * buf;
        const size_t nf = frag_count(p);
        size_t hbytes = precision == FCIS_PREC_BF16X3 ? nf * 16 : nf * elem;
        size_t pa = (size_t)p->npair * space->nstra * 4, pb = (size_t)p->npair * nb * 4;
        if (hipMalloc(&p->d_hfrag, hbytes) != hipSuccess ||
            hipMalloc(&p->d_hfrag_lo, precision == FCIS_PREC_BF16X3 ? hbytes : 16) != hipSuccess ||
            hipMalloc(&p->d_pmapa, pa ? pa : 4) != hipSuccess ||
            hipMalloc(&p->d_pmapb, pb ? pb : 4) != hipSuccess ||
            hipMalloc(&p->d_t1, buf ? buf : 8) != hipSuccess ||
            hipMalloc(&p->d_vt1, buf ? buf : 8) != hipSuccess) {
                fcis_sigma_plan_free(p);
                fcis::set_error("fcis_sigma_plan_create: device allocation of %zu bytes failed",
                                2 * buf + 2 * hbytes + pa + pb);
                return FCIS_ERR_MEMORY;
        }
        int st = upload_h(p, h2e);
        if (st != FCIS_SUCCESS) {
                fcis_sigma_plan_free(p);
                return st;
        }
        FCIS_HIP(hipMemset(p->d_pmapa, 0xff, pa ? pa : 4));
        FCIS_HIP(hipMemset(p->d_pmapb, 0xff, pb ? pb : 4));
        if (space->nlinka > 0) {
                hipLaunchKernelGGL(k_build_pmap, dim3((space->nstra + 255) / 256, space->nlinka),
                                   dim3(256), 0, 0, (uint32_t)space->nstra, space->nlinka,
                                   space->d_linka, p->d_pmapa);
                FCIS_CHECK_LAUNCH();
        }
        if (space->nlinkb > 0) {
                hipLaunchKernelGGL(k_build_pmap, dim3((nb + 255) / 256, space->nlinkb),
                                   dim3(256), 0, 0, (uint32_t)nb, space->nlinkb,
                                   space->d_linkb, p->d_pmapb);
                FCIS_CHECK_LAUNCH();
        }
        FCIS_HIP(hipDeviceSynchronize());
        *out = p;
        return FCIS_SUCCESS;
}

int fcis_sigma_plan_set_h2e(fcis_sigma_plan_t *p, const double *h2e)
{
        if (p == nullptr || h2e == nullptr) return FCIS_ERR_INVALID;
        return upload_h(p, h2e);
}

int fcis_sigma_plan_info(const fcis_sigma_plan_t *p, int *ntiles, int64_t *tile_alpha,
                         uint64_t *ws)
{
        if (p == nullptr) return FCIS_ERR_INVALID;
        if (ntiles) *ntiles = p->ntiles;
        if (tile_alpha) *tile_alpha = p->tile_a;
        if (ws) *ws = p->ws_bytes;
        return FCIS_SUCCESS;
}

int fcis_sigma_plan_execute(fcis_sigma_plan_t *p, const double *ci0, double *ci1, void *stream)
{
        if (p == nullptr || ci0 == nullptr || ci1 == nullptr) {
                fcis::set_error("fcis_sigma_plan_execute: null argument");
                return FCIS_ERR_INVALID;
        }
        if (ci0 == ci1) {
                fcis::set_error("fcis_sigma_plan_execute: ci0 and ci1 must not alias");
                return FCIS_ERR_INVALID;
        }
        hipStream_t st = (hipStream_t)stream;
        const fcis_space *s = p->space;
        const int64_t nb = s->nstrb, na = s->nstra;
        if (na * nb == 0) return FCIS_SUCCESS;
        const dim3 gblk(64 * GEMM_WAVES);
        for (int t = 0; t < p->ntiles; t++) {
                const int64_t a0 = (int64_t)t * p->tile_a;
                const int64_t ta = na - a0 < p->tile_a ? na - a0 : p->tile_a;
                const int64_t ndet = ta * nb;
                const int64_t ld = ndet;
                const dim3 ggrid((ndet + GEMM_NCOL - 1) / GEMM_NCOL, (p->mtiles + GEMM_MT - 1) / GEMM_MT);
                const dim3 rgrid(ta * ((nb + 255) / 256)), sgrid(na * ((nb + 255) / 256));
                if (p->timing) FCIS_HIP(hipEventRecord(p->ev[0], st));
                if (p->precision == FCIS_PREC_FP64) {
                        double *t1 = (double *)p->d_t1, *vt1 = (double *)p->d_vt1;
                        hipLaunchKernelGGL(k_gather<double>, rgrid, dim3(256), 0, st, p->npair, nb, a0,
                                           na, ld, p->d_pmapa, p->d_pmapb, ci0, t1);
                        if (p->timing) FCIS_HIP(hipEventRecord(p->ev[1], st));
                        hipLaunchKernelGGL(k_contract_f64, ggrid, gblk, 0, st, p->npair, p->mtiles,
                                           p->ksteps, ndet, ld, (const double *)p->d_hfrag, t1, vt1);
                        if (p->timing) FCIS_HIP(hipEventRecord(p->ev[2], st));
                        hipLaunchKernelGGL(k_scatter<double>, sgrid, dim3(256), 0, st, nb, na, a0, ta,
                                           ld, s->nlinka, s->nlinkb, t > 0 ? 1 : 0, s->d_linka,
                                           s->d_linkb, vt1, ci1);
                } else {
                        float *t1 = (float *)p->d_t1, *vt1 = (float *)p->d_vt1;
                        hipLaunchKernelGGL(k_gather<float>, rgrid, dim3(256), 0, st, p->npair, nb, a0,
                                           na, ld, p->d_pmapa, p->d_pmapb, ci0, t1);
                        if (p->timing) FCIS_HIP(hipEventRecord(p->ev[1], st));
                        if (p->precision == FCIS_PREC_FP32)
                                hipLaunchKernelGGL(k_contract_f32, ggrid, gblk, 0, st, p->npair,
                                                   p->mtiles, p->ksteps, ndet, ld,
                                                   (const float *)p->d_hfrag, t1, vt1);
                        else
                                hipLaunchKernelGGL(k_contract_bf16x3, ggrid, gblk, 0, st, p->npair,
                                                   p->mtiles, p->ksteps, ndet, ld,
                                                   (const bf16x8 *)p->d_hfrag,
                                                   (const bf16x8 *)p->d_hfrag_lo, t1, vt1);
                        if (p->timing) FCIS_HIP(hipEventRecord(p->ev[2], st));
                        hipLaunchKernelGGL(k_scatter<float>, sgrid, dim3(256), 0, st, nb, na, a0, ta,
                                           ld, s->nlinka, s->nlinkb, t > 0 ? 1 : 0, s->d_linka,
                                           s->d_linkb, vt1, ci1);
                }
                FCIS_CHECK_LAUNCH();
                if (p->timing) {
                        FCIS_HIP(hipEventRecord(p->ev[3], st));
                        FCIS_HIP(hipEventSynchronize(p->ev[3]));
                        for (int k = 0; k < 3; k++) {
                                float ms = 0.0f;
                                FCIS_HIP(hipEventElapsedTime(&ms, p->ev[k], p->ev[k + 1]));
                                p->ms[k] += ms;
                                p->launches[k] += 1;
                        }
                }
        }
        return FCIS_SUCCESS;
}

int fcis_sigma_plan_set_timing(fcis_sigma_plan_t *p, int enable)
{
        if (p == nullptr) return FCIS_ERR_INVALID;
        if (enable && !p->timing) {
                for (int k = 0; k < 4; k++) FCIS_HIP(hipEventCreate(&p->ev[k]));
        }
        if (!enable && p->timing) {
                for (int k = 0; k < 4; k++) (void)hipEventDestroy(p->ev[k]);
        }
        p->timing = enable ? 1 : 0;
        for (int k = 0; k < 3; k++) {
                p->ms[k] = 0.0;
                p->launches[k] = 0;
        }
        return FCIS_SUCCESS;
}

int fcis_sigma_plan_get_timing(const fcis_sigma_plan_t *p, double *ms, int64_t *launches)
{
        if (p == nullptr) return FCIS_ERR_INVALID;
        for (int k = 0; k < 3; k++) {
                if (ms) ms[k] = p->ms[k];
                if (launches) launches[k] = p->launches[k];
        }
        return FCIS_SUCCESS;
}

void fcis_sigma_plan_free(fcis_sigma_plan_t *p)
{
        if (p == nullptr) return;
        if (p->timing) fcis_sigma_plan_set_timing(p, 0);
        (void)hipFree(p->d_hfrag);
        (void)hipFree(p->d_hfrag_lo);
        (void)hipFree(p->d_pmapa);
        (void)hipFree(p->d_pmapb);
        (void)hipFree(p->d_t1);
        (void)hipFree(p->d_vt1);
        delete p;
}

}  // extern "C"
