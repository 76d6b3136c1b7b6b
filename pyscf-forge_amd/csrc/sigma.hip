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
#include <mutex>
#include <algorithm>
#include <string>
#include <vector>

#include "common.h"
#include "fused.h"

namespace {

constexpr int GEMM_WAVES = 4;           // wavefronts per workgroup
constexpr int GEMM_NSUB = 2;            // 16-column subtiles per wavefront
constexpr int GEMM_MT = 9;              // 16-row tiles per workgroup (144 rows)
constexpr int GEMM_NCOL = GEMM_WAVES * GEMM_NSUB * 16;  // 128 dets per workgroup

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

// FP32 MFMA contraction (v_mfma_f32_16x16x4_f32: A and B as in the f64
// form, D[i][j] in lane j + 16*(i/4), element i%4); t1, vt1, h are float.
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

// ---------------------------------------------------------------------------
// Fused row kernel: gather + MFMA contraction + beta scatter for whole
// alpha strings.
//
// A workgroup owns one alpha string I at a time (persistent over rows) and
// walks its nb determinants in chunks of FUSED_NB.  It has one wavefront per
// 16-row tile of h (mtiles <= FUSED_MAXW); each wavefront keeps its tile's A
// fragments for every k-step in registers, so the MFMA loop reads only B
// fragments, from LDS.  The B tile (npair x FUSED_NB) is gathered from c
// straight into LDS,
//   t1[pq][b] = sa c[Ja, b] + sb c[I, Jb]   (pmapa[pq][I] -> (Ja, sa),
//                                             pmapb[pq][b] -> (Jb, sb)),
// double-buffered and software-pipelined (pair maps two chunks ahead, c one
// chunk ahead of the MFMAs), so t1 never touches HBM.  Of the product
// vt1 = h t1:
//   * the beta part of the scatter stays on chip: sigma[I, Jb] +=
//     sb vt1[pq][b] for the beta link (pq -> Jb) of b, accumulated into the
//     LDS copy of sigma row I with LDS float atomics (the only
//     non-deterministic summation order in the engine; see the
//     deterministic three-kernel path);
//   * the alpha part is written to HBM in alpha-link-slot layout,
//     walpha[k][I][b] for the k-th alpha link of I (72 of the 136 pair
//     rows at norb = 16), and pulled by k_alpha_pull.
// Independent accumulators per k-step phase (FP64/FP32) or per split term
// (bf16x3) keep several MFMA chains in flight per wavefront.
//
// ROWS: restricted rows (fcis_rowset).  Position a0 + r processes alpha
// string rows[a0 + r] over its listed beta chunks only; walpha is indexed by
// position.  Otherwise position r of the tile processes alpha string
// a0 + rows[r] (a processing order of the tile's strings, see
// fused_row_order) or a0 + r when rows is null, all chunks; walpha is
// indexed by string.
template <int PREC, bool ROWS>
__global__ __launch_bounds__(64 * (FUSED_MTILES + FUSED_NBLD)) void k_fused_row(
        int npair, int mtiles, int mrow_f, int na, int nlinka, int64_t nb, int64_t nstra, int64_t a0,
        int64_t ta, int64_t ld, int accumulate, const fcis_link *__restrict__ linka,
        const int32_t *__restrict__ pmapa, const int16_t *__restrict__ pmapb, int64_t nbp,
        const typename FusedTypes<PREC>::G *__restrict__ c, const void *__restrict__ hfrag_v,
        const void *__restrict__ hfrag_lo_v, const int16_t *__restrict__ mpair,
        const double *__restrict__ falpha, const typename FusedTypes<PREC>::G *__restrict__ fbeta,
        typename FusedTypes<PREC>::Out *__restrict__ walpha,
        double *__restrict__ sigma, const int32_t *__restrict__ rows,
        const int2 *__restrict__ rowcl, const uint16_t *__restrict__ clist,
        const bf16x8 *__restrict__ dtile, const uint64_t *__restrict__ strb)
{
        typedef FusedTypes<PREC> Ty;
        typedef typename Ty::B BT;
        typedef typename Ty::G G;
        constexpr int KS = Ty::KS, KW = Ty::KW, NB = FUSED_NB;
        constexpr int KPMAX = KS * KW;
        constexpr int MAXP = 16 * FUSED_MAXW;
        constexpr int NBLD = FUSED_NBLD;
        extern __shared__ __align__(16) char smem[];
        double *srow = reinterpret_cast<double *>(smem);   // nbp doubles: 16-byte aligned end
        // B tiles: [k][col] in BT, or for bf16x3 bf16 high and low parts [col][k]
        constexpr size_t BB = fused_bbytes<PREC>();
        constexpr bool PRESPLIT = PREC == FCIS_PREC_BF16X3;
        char *bbase = smem + (size_t)nbp * sizeof(double);
        BT *buf0 = reinterpret_cast<BT *>(bbase);
        BT *buf1 = reinterpret_cast<BT *>(bbase + BB);
        // abyte[pq]: byte offset of c row Ja of I's alpha link with pair pq,
        // its sign in bit 0, or FUSED_OOB (pq >= npair too)
        uint32_t *abyte = reinterpret_cast<uint32_t *>(bbase + 2 * BB);
        int *aslot = reinterpret_cast<int *>(abyte + MAXP);        // walpha slot of pq, or -1
        int16_t *maps = reinterpret_cast<int16_t *>(aslot + MAXP); // FUSED_NMAP x [MAXP][NB] beta maps
        // bf16x3: the beta-occupation part of the diagonal pairs runs on the
        // matrix pipe (which has room in this kernel): the pair-map wavefront
        // multiplies the 16 rows h[pp][k] (A fragments from LDS) into each
        // chunk's tile and adds sum_{p in b} (h[pp] t1)[p][b] to sigma row I;
        // the other kernels form sum_k fBeta[k][b] t1[k][b] in the builders
        constexpr bool DTILE = PREC == FCIS_PREC_BF16X3;
        constexpr bool BDIAG = !DTILE;
        bf16x8 *dlds = reinterpret_cast<bf16x8 *>(maps + FUSED_NMAP * MAXP * NB);

        const int tid = threadIdx.x, nthr = blockDim.x;
        const __amdgpu_buffer_rsrc_t rc = fused_c_rsrc(c, (uint32_t)(nstra * nb * (int64_t)sizeof(G)));
        const int wave = tid >> 6, lane = tid & 63, li = lane & 15, lk = lane >> 4;
        const bool mfma_wave = wave < mtiles;
        const int64_t nch = (nb + NB - 1) / NB;
        const int nelem = npair * NB;
        // ROWS: entry c0 + jj of the chunk lists, read as its 32-bit word (the
        // list is padded to whole words): a uniform address, so a scalar load
        // that never makes the vector-memory counter wait for loads in flight
        auto list_chunk = [&](int64_t c0, int64_t jj) -> int64_t {
                const int64_t q = c0 + jj;
                return (int64_t)((reinterpret_cast<const uint32_t *>(clist)[q >> 1] >> (16 * (int)(q & 1))) &
                                 0xffffu);
        };
        auto row_of = [&](int64_t r) -> int64_t {
                return ROWS ? (int64_t)rows[a0 + r] : a0 + (rows != nullptr ? (int64_t)rows[r] : r);
        };

        if constexpr (PRESPLIT) {
                // rows k >= npair (and the padding) stay zero; commits write the rest
                for (int e = tid; e < (int)(2 * BB / 4); e += nthr) reinterpret_cast<uint32_t *>(bbase)[e] = 0u;
        } else {
                for (int e = nelem + tid; e < KPMAX * NB; e += nthr) {
                        buf0[e] = (BT)0;
                        buf1[e] = (BT)0;
                }
        }
        for (int64_t i = tid; i < nb; i += nthr) srow[i] = 0.0;
        if constexpr (DTILE)
                for (int e = tid; e < 2 * KS * 64; e += nthr) dlds[e] = dtile[e];

        // Barrier schedule per row (every role): (1) row start, (2) abyte,
        // (3) aslot + maps of chunks 0..2, (4) chunk 0 built, then one per chunk.
        if (!mfma_wave) {
                // ---- builder wavefronts: seven gather c into the B tiles, the
                // eighth loads the beta pair maps and forms the diagonal term.
                // Builders go first at issue (their commits set the chunk time
                // when they straggle).
                __builtin_amdgcn_s_setprio(1);
                const int bi = wave - mtiles;
                constexpr int NBA = NBLD - 1;   // gathering builders
                if (bi == NBA) {
                        // pair-map wavefront.  Phase j: the beta pair-map slice of
                        // chunk j + 4 is loaded (16-byte loads) and that of chunk
                        // j + 3, loaded one phase earlier, stored to LDS slice
                        // (j + 3) % 4 (for the builders' gathers of phase j + 1 and
                        // the MFMA scatter of phase j + 3).
                        constexpr int EMW = (2 * MAXP + 63) / 64;
                        v4i m0[EMW], m1[EMW];
                        for (int64_t r = blockIdx.x; r < ta; r += gridDim.x) {
                                const int64_t I = row_of(r);
                                int64_t nl = nch, c0 = 0;
                                if constexpr (ROWS) {
                                        const int2 e = rowcl[a0 + r];
                                        c0 = e.x;
                                        nl = e.y;
                                }
                                auto cc = [&](int64_t jj) -> int64_t {
                                        const int64_t q = jj < nl ? jj : nl - 1;
                                        return ROWS ? list_chunk(c0, q) : (int64_t)__builtin_amdgcn_readfirstlane((int)q);
                                };
                                auto issue = [&](v4i(&m)[EMW], int64_t jj) {
                                        const int64_t b0 = cc(jj) * NB;
#pragma unroll
                                        for (int s = 0; s < EMW; s++) {
                                                const int t = lane + s * 64, tc = t < 2 * npair ? t : 2 * npair - 1;
                                                m[s] = *reinterpret_cast<const v4i *>(pmapb + (int64_t)(tc >> 1) * nbp + b0 +
                                                                                      (tc & 1) * 8);
                                        }
                                };
                                auto store = [&](const v4i(&m)[EMW], int64_t jj) {
                                        int16_t *mb = maps + (jj % FUSED_NMAP) * MAXP * NB;
#pragma unroll
                                        for (int s = 0; s < EMW; s++) {
                                                const int t = lane + s * 64;
                                                if (t < 2 * npair)
                                                        *reinterpret_cast<v4i *>(mb + (t >> 1) * NB + (t & 1) * 8) = m[s];
                                        }
                                };
                                __syncthreads();   // (1)
                                __syncthreads();   // (2)
                                issue(m0, 0);
                                store(m0, 0);
                                issue(m0, 1);
                                store(m0, 1);
                                issue(m0, 2);
                                store(m0, 2);
                                issue(m0, 3);
                                __syncthreads();   // (3)
                                __syncthreads();   // (4)
                                // bf16x3: the diagonal tile of chunk j (in B buffer j & 1
                                // this phase), rows p = 4 lk + q, column li
                                auto diag_tile = [&](int64_t j) {
                                        if constexpr (DTILE) {
                                                const int64_t b = cc(j) * NB + li;
                                                const uint64_t occ = strb[b < nb ? b : nb - 1];
                                                const v4f dv = mma_presplit_lds<KS>(
                                                        dlds, reinterpret_cast<const __bf16 *>((j & 1) ? buf1 : buf0), lane,
                                                        lk, li);
                                                float part = 0.0f;
#pragma unroll
                                                for (int q = 0; q < 4; q++)
                                                        part += ((occ >> (4 * lk + q)) & 1u) ? dv[q] : 0.0f;
                                                part += __shfl_xor(part, 16);
                                                part += __shfl_xor(part, 32);
                                                if (lk == 0 && b < nb) atomicAdd(&srow[b], (double)part);
                                        }
                                };
                                auto mphase = [&](int64_t j, v4i(&ld_)[EMW], v4i(&st_)[EMW]) {
                                        issue(ld_, j + 4);
                                        store(st_, j + 3);
                                        diag_tile(j);
                                        __syncthreads();
                                };
                                int64_t j = 0;
                                for (; j + 1 < nl; j += 2) {
                                        mphase(j, m1, m0);
                                        mphase(j + 1, m0, m1);
                                }
                                if (j < nl) mphase(j, m1, m0);
                                double *dst = sigma + I * nb;
                                for (int64_t i = tid; i < nb; i += nthr) {
                                        double v = srow[i];
                                        dst[i] = accumulate ? dst[i] + v : v;
                                        srow[i] = 0.0;
                                }
                        }
                        return;
                }
                const int bt = bi * 64 + lane;
                // the beta occupations' part of the diagonal term,
                //   d[b] = sum_k fBeta[k][b] t1[k][b],
                // formed at commit from the tile elements in registers (a thread
                // keeps one column; lanes of a column summed, then one LDS
                // atomic per column and wavefront)
                const __amdgpu_buffer_rsrc_t fbr = fused_c_rsrc(fbeta, (uint32_t)(npair * nbp * (int64_t)sizeof(G)));
                auto add_diag = [&](double d, int64_t b0) {
                        d += __shfl_xor(d, 16);
                        d += __shfl_xor(d, 32);
                        if (lane < 16 && b0 + (bt & 15) < nb) atomicAdd(&srow[b0 + (bt & 15)], d);
                };
                typedef FusedGather<NBA, FUSED_MAXW, G> Gather;
                Gather s0, s1;   // c of even / odd chunks in flight
                for (int64_t r = blockIdx.x; r < ta; r += gridDim.x) {
                        const int64_t I = row_of(r);
                        int64_t nl = nch;
                        int64_t c0 = 0;
                        if constexpr (ROWS) {
                                const int2 e = rowcl[a0 + r];
                                c0 = e.x;
                                nl = e.y;
                        }
                        // the chunk index through readfirstlane: opaque to strength
                        // reduction, which otherwise keeps one address register pair per
                        // gathered element live across the chunk loop
                        auto chunk = [&](int64_t jj) -> int64_t {
                                return ROWS ? list_chunk(c0, jj) : (int64_t)__builtin_amdgcn_readfirstlane((int)jj);
                        };
                        __syncthreads();   // (1)
                        for (int q = bt; q < MAXP; q += NBA * 64) {
                                const int a = q < npair ? pmapa[(int64_t)q * nstra + I] : -1;
                                abyte[q] = a >= 0 ? (uint32_t)(a >> 1) * (uint32_t)nb * (uint32_t)sizeof(G) |
                                                            (uint32_t)(a & 1)
                                                  : FUSED_OOB;
                                aslot[q] = -1;
                        }
                        const uint32_t rowI_byte = (uint32_t)(I * nb) * (uint32_t)sizeof(G);
                        __syncthreads();   // (2)
                        // walpha slots: the off-diagonal alpha links (slots na ..);
                        // the diagonal ones act within row I (the diagonal term)
                        for (int k = na + bt; k < nlinka; k += NBA * 64)
                                aslot[link_pair(linka[(int64_t)k * nstra + I].info)] = k - na;
                        // chunk jj clamped to the row (loads past the last chunk
                        // repeat it and are never committed)
                        auto cc = [&](int64_t jj) -> int64_t { return chunk(jj < nl ? jj : nl - 1); };
                        auto mbuf = [&](int64_t jj) -> int16_t * { return maps + (jj % FUSED_NMAP) * MAXP * NB; };
                        __syncthreads();   // (3)
                        // loads left in flight here, oldest first: c of chunk 1 -- the
                        // order the loop body leaves them in, so that the loop head sees
                        // one order and every wait stays partial
                        fg_issue<NBA, FUSED_MAXW, BDIAG>(s0, rowI_byte, cc(0), bt, abyte, mbuf(0), rc, fbr, nbp);
                        fg_issue<NBA, FUSED_MAXW, BDIAG>(s1, rowI_byte, cc(1), bt, abyte, mbuf(1), rc, fbr, nbp);
                        if constexpr (PRESPLIT)
                                fg_commit_split<NBA, FUSED_MAXW, BDIAG>(s0, npair, bt, reinterpret_cast<__bf16 *>(buf0));
                        else
                                add_diag(fg_commit<NBA, FUSED_MAXW, BDIAG>(s0, npair, bt, buf0), cc(0) * NB);
                        __syncthreads();   // (4)
                        // phase j: `issue` issues the c loads of chunk j+2 (its map
                        // slice is in LDS); `commit` holds the c values of chunk j+1
                        // (issued one phase earlier), committed now.  The loads of
                        // this phase stay in flight across the barrier.
                        auto phase = [&](int64_t j, Gather &issue, Gather &commit) {
                                fg_issue<NBA, FUSED_MAXW, BDIAG>(issue, rowI_byte, cc(j + 2), bt, abyte, mbuf(j + 2), rc,
                                                                 fbr, nbp);
                                // unconditional: past the row's last chunk the commit goes
                                // to the B buffer that no wavefront reads in this phase
                                // (clamped chunk data); its diagonal term is dropped
                                if constexpr (PRESPLIT) {
                                        fg_commit_split<NBA, FUSED_MAXW, BDIAG>(
                                                commit, npair, bt,
                                                reinterpret_cast<__bf16 *>(((j + 1) & 1) ? buf1 : buf0));
                                } else {
                                        const double d = fg_commit<NBA, FUSED_MAXW, BDIAG>(commit, npair, bt,
                                                                                           ((j + 1) & 1) ? buf1 : buf0);
                                        if (j + 1 < nl) add_diag(d, cc(j + 1) * NB);
                                }
                                __syncthreads();
                        };
                        int64_t j = 0;
                        for (; j + 1 < nl; j += 2) {
                                phase(j, s0, s1);
                                phase(j + 1, s1, s0);
                        }
                        if (j < nl) phase(j, s0, s1);
                        double *dst = sigma + I * nb;
                        for (int64_t i = tid; i < nb; i += nthr) {
                                double v = srow[i];
                                dst[i] = accumulate ? dst[i] + v : v;
                                srow[i] = 0.0;
                        }
                }
                return;
        }

        // ---- MFMA wavefronts: this wavefront's 16-row tile of the fused h
        // (rows: the off-diagonal pairs mpair[m], then row mrow_f whose A
        // entries are fAlpha(I), set per alpha string) in registers
        typename Ty::A ah[KS];
        typename Ty::A al[PREC == FCIS_PREC_BF16X3 ? KS : 1];
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
                // zero past the problem's k-steps: mma_steps runs all KS
                const size_t o = ((size_t)ks * mtiles + wave) * 64 + lane;
                const int kstep = PREC == FCIS_PREC_BF16X3 ? 32 : 4;
                const bool in = ks * kstep < npair;
                ah[ks] = in ? ((const typename Ty::A *)hfrag_v)[o] : typename Ty::A{};
                if constexpr (PREC == FCIS_PREC_BF16X3)
                        al[ks] = in ? ((const typename Ty::A *)hfrag_lo_v)[o] : typename Ty::A{};
        }
        // this lane's four output rows (f64 MFMA: D row lk + 4q; f32/bf16 MFMA:
        // D row 4lk + q) and their pairs (-1: the diagonal row or padding)
        int qpair[4];
        bool qdiag[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
                const int m = wave * 16 + (PREC == FCIS_PREC_FP64 ? lk + 4 * q : 4 * lk + q);
                qpair[q] = m < mrow_f ? (int)mpair[m] : -1;
                qdiag[q] = m == mrow_f;
        }
        const bool frow = wave == mrow_f / 16 && li == mrow_f % 16;   // holds fAlpha(I)
        for (int64_t r = blockIdx.x; r < ta; r += gridDim.x) {
                const int64_t I = row_of(r);
                int64_t nl = nch;
                int64_t c0 = 0;
                if constexpr (ROWS) {
                        const int2 e = rowcl[a0 + r];
                        c0 = e.x;
                        nl = e.y;
                }
                __syncthreads();   // (1)
                if (frow) {
                        // A row mrow_f = fAlpha(I): A[i][k] at lane i + 16 (k % 4) for
                        // k-step k / 4 (f64 / f32), lane i + 16 (k / 8 % 4) element k % 8
                        // for k-step k / 32 (bf16)
                        const double *fa = falpha + I * npair;
#pragma unroll
                        for (int ks = 0; ks < KS; ks++) {
                                if constexpr (PREC == FCIS_PREC_BF16X3) {
#pragma unroll
                                        for (int e = 0; e < 8; e++) {
                                                const int k = 32 * ks + 8 * lk + e;
                                                const float x = k < npair ? (float)fa[k] : 0.0f;
                                                const __bf16 hi = (__bf16)x;
                                                ah[ks][e] = hi;
                                                al[ks][e] = (__bf16)(x - (float)hi);
                                        }
                                } else {
                                        const int k = 4 * ks + lk;
                                        ah[ks] = k < npair ? (typename Ty::A)fa[k] : (typename Ty::A)0;
                                }
                        }
                }
                __syncthreads();   // (2)
                __syncthreads();   // (3)
                int slot[4];
#pragma unroll
                for (int q = 0; q < 4; q++) slot[q] = qpair[q] >= 0 ? aslot[qpair[q]] : -1;
                __syncthreads();   // (4)
                typename Ty::Out *wrow = walpha + (ROWS ? r : I - a0) * nbp;
                for (int64_t j = 0; j < nl; j++) {
                        const BT *bb = (j & 1) ? buf1 : buf0;
                        const int16_t *mb = maps + (j % FUSED_NMAP) * MAXP * NB;
                        const int64_t b = (ROWS ? list_chunk(c0, j) : j) * NB + li;
                        // the beta-scatter targets, read before the MFMA chain so that
                        // the scatter does not wait on LDS (the diagonal row adds to
                        // its own determinant)
                        int em[4];
#pragma unroll
                        for (int q = 0; q < 4; q++)
                                em[q] = qpair[q] >= 0 ? (int)mb[qpair[q] * NB + li] : (qdiag[q] ? (int)(b << 1) : -1);
                        typename Ty::Acc acc;
                        if constexpr (PRESPLIT)
                                acc = mma_presplit<KS>(ah, al, reinterpret_cast<const __bf16 *>(bb), lk, li);
                        else
                                acc = mma_steps<PREC, KS, KS>(ah, al, bb, lk, li);
                        if (b < nb) {
#pragma unroll
                                for (int q = 0; q < 4; q++) {
                                        const double v = (double)acc[q];
                                        if (em[q] >= 0) atomicAdd(&srow[em[q] >> 1], fused_flip(v, em[q] & 1));
                                        if (slot[q] >= 0) wrow[(int64_t)slot[q] * ld + b] = acc[q];
                                }
                        }
                        __syncthreads();
                }
                double *dst = sigma + I * nb;
                for (int64_t i = tid; i < nb; i += nthr) {
                        double v = srow[i];
                        dst[i] = accumulate ? dst[i] + v : v;
                        srow[i] = 0.0;
                }
        }
}

// Alpha pull for one alpha tile [a0, a0+ta): one thread per destination
// determinant (J, Jb), adding sum_k s_k walpha[rslot_k - kfirst][I_k - a0][Jb]
// over the off-diagonal alpha links k >= kfirst (= na; the diagonal links
// come first) of J whose target I_k lies in the tile; the reverse slot
// rslot_k is the slot of the link I_k -> J in I_k's list.  walpha rows
// are wrs (nb rounded up to whole 16-determinant chunks) apart, so that the
// fused kernel's 16-wide stores fill whole 128-byte lines.
template <typename T>
__global__ __launch_bounds__(256) void k_alpha_pull(int64_t nb, int64_t wrs, int64_t nstra, int64_t a0,
                                                    int64_t ta, int64_t ld, int kfirst, int nlinka,
                                                    const fcis_link *__restrict__ linka,
                                                    const T *__restrict__ walpha,
                                                    double *__restrict__ sigma)
{
        // two adjacent determinants per thread, one 2-wide load per link: the
        // walpha rows and slots have even lengths, so a pair never crosses one
        typedef typename std::conditional<std::is_same<T, double>::value, double2, float2>::type T2;
        const int64_t nchunk = ((nb + 1) / 2 + 255) / 256;
        const int64_t J = blockIdx.x / nchunk;
        const int64_t Jb = 2 * ((blockIdx.x - J * nchunk) * 256 + threadIdx.x);
        // the links of J, all loaded at once (nlinka <= npair <= 16 * FUSED_MAXW
        // on the fused path)
        __shared__ fcis_link lks[16 * FUSED_MAXW];
        for (int k = threadIdx.x; k < nlinka; k += 256) lks[k] = linka[(int64_t)k * nstra + J];
        __syncthreads();
        if (Jb >= nb) return;
        // links in groups of U with all U loads issued before any is used
        // (links outside the tile read walpha[Jb] and add zero); the sums
        // run in link order either way
        constexpr int U = 12;
        double acc0 = 0.0, acc1 = 0.0;
        int k = kfirst;
        for (; k + U <= nlinka; k += U) {
                T2 v[U];
                bool in[U], neg[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                        const fcis_link L = lks[k + u];
                        const int64_t I = L.addr;
                        in[u] = I >= a0 && I < a0 + ta;
                        neg[u] = link_neg(L.info);
                        // each walpha element is read once: non-temporal loads
                        typedef T NT2 __attribute__((ext_vector_type(2)));
                        const NT2 w = __builtin_nontemporal_load(reinterpret_cast<const NT2 *>(
                                walpha + (in[u] ? (int64_t)(link_rslot(L.info) - kfirst) * ld + (I - a0) * wrs + Jb
                                                : Jb)));
                        v[u].x = w.x;
                        v[u].y = w.y;
                }
#pragma unroll
                for (int u = 0; u < U; u++) {
                        if (in[u]) {
                                const double x0 = (double)v[u].x, x1 = (double)v[u].y;
                                acc0 += neg[u] ? -x0 : x0;
                                acc1 += neg[u] ? -x1 : x1;
                        }
                }
        }
        for (; k < nlinka; k++) {
                const fcis_link L = lks[k];
                const int64_t I = L.addr;
                if (I < a0 || I >= a0 + ta) continue;
                const T2 v = *reinterpret_cast<const T2 *>(walpha + (int64_t)(link_rslot(L.info) - kfirst) * ld +
                                                           (I - a0) * wrs + Jb);
                const double x0 = (double)v.x, x1 = (double)v.y;
                acc0 += link_neg(L.info) ? -x0 : x0;
                acc1 += link_neg(L.info) ? -x1 : x1;
        }
        sigma[J * nb + Jb] += acc0;
        if (Jb + 1 < nb) sigma[J * nb + Jb + 1] += acc1;
}

__global__ void k_zero(int64_t n, double *__restrict__ x)
{
        int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (i < n) x[i] = 0.0;
}

// int16 copy of pmapb with rows padded to nbp entries (-1 padding).
__global__ void k_pmap16(int npair, int64_t nb, int64_t nbp, const int32_t *__restrict__ in,
                         int16_t *__restrict__ out)
{
        int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= (int64_t)npair * nbp) return;
        int64_t k = i / nbp, b = i - k * nbp;
        out[i] = b < nb ? (int16_t)in[k * nb + b] : (int16_t)-1;
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
        // the α links of J in LDS (nlinka entries of dynamic shared memory)
        extern __shared__ fcis_link lks[];
        for (int k = threadIdx.x; k < nlinka; k += 256) lks[k] = linka[(int64_t)k * nstra + J];
        __syncthreads();
        if (Jb >= nb) return;
        // U loads in flight per thread; every sum runs in link order
        constexpr int U = 8;
        double acc = accumulate ? sigma[J * nb + Jb] : 0.0;
        int k = 0;
        for (; k + U <= nlinka; k += U) {
                T v[U];
                bool in[U], neg[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                        const fcis_link L = lks[k + u];
                        const int64_t I = L.addr;
                        in[u] = I >= a0 && I < a0 + ta;
                        neg[u] = link_neg(L.info);
                        v[u] = vt1[in[u] ? (int64_t)link_pair(L.info) * ld + (I - a0) * nb + Jb : Jb];
                }
#pragma unroll
                for (int u = 0; u < U; u++) {
                        if (in[u]) {
                                const double x = (double)v[u];
                                acc += neg[u] ? -x : x;
                        }
                }
        }
        for (; k < nlinka; k++) {
                const fcis_link L = lks[k];
                const int64_t I = L.addr;
                if (I < a0 || I >= a0 + ta) continue;
                const double x = (double)vt1[(int64_t)link_pair(L.info) * ld + (I - a0) * nb + Jb];
                acc += link_neg(L.info) ? -x : x;
        }
        if (J >= a0 && J < a0 + ta) {
                const T *row = vt1 + (J - a0) * nb;
                k = 0;
                for (; k + U <= nlinkb; k += U) {
                        fcis_link L[U];
                        T v[U];
#pragma unroll
                        for (int u = 0; u < U; u++) L[u] = linkb[(int64_t)(k + u) * nb + Jb];
#pragma unroll
                        for (int u = 0; u < U; u++) v[u] = row[(int64_t)link_pair(L[u].info) * ld + L[u].addr];
#pragma unroll
                        for (int u = 0; u < U; u++) {
                                const double x = (double)v[u];
                                acc += link_neg(L[u].info) ? -x : x;
                        }
                }
                for (; k < nlinkb; k++) {
                        const fcis_link L = linkb[(int64_t)k * nb + Jb];
                        const double x = (double)row[(int64_t)link_pair(L.info) * ld + L.addr];
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

// A fragments of the MFMA tiles whose rows are the pairs rowpair[m] (-1: a
// zero row), laid out so that each fragment load is one contiguous read:
//   FP64 / FP32: frag[(ks * mtiles + mt) * 64 + lane]
//                  = h[rowpair[mt*16 + lane%16]][4 ks + lane/16]
//   bf16x3:      hi / lo[((ks * mtiles + mt) * 64 + lane) * 8 + e]
//                  = h[rowpair[mt*16 + lane%16]][32 ks + 8 (lane/16) + e]
int upload_frags(const fcis_sigma_plan *p, const double *h2e, const std::vector<int> &rowpair, int mtiles,
                 void *dst, void *dst_lo)
{
        const int np = p->npair;
        const size_t nf = (size_t)p->ksteps * mtiles * 64;
        auto hval = [&](int m, int k) -> double {
                const int row = m < (int)rowpair.size() ? rowpair[m] : -1;
                return (row >= 0 && k < np) ? h2e[(size_t)row * np + k] : 0.0;
        };
        if (p->precision == FCIS_PREC_BF16X3) {
                std::vector<uint16_t> hi(nf * 8), lo(nf * 8);
                for (int ks = 0; ks < p->ksteps; ks++)
                        for (int mt = 0; mt < mtiles; mt++)
                                for (int l = 0; l < 64; l++)
                                        for (int e = 0; e < 8; e++) {
                                                const float x = (float)hval(mt * 16 + (l & 15), 32 * ks + 8 * (l >> 4) + e);
                                                const uint16_t xh = bf16_bits(x);
                                                const size_t o = (((size_t)ks * mtiles + mt) * 64 + l) * 8 + e;
                                                hi[o] = xh;
                                                lo[o] = bf16_bits(x - bf16_value(xh));
                                        }
                FCIS_HIP(hipMemcpy(dst, hi.data(), hi.size() * 2, hipMemcpyHostToDevice));
                FCIS_HIP(hipMemcpy(dst_lo, lo.data(), lo.size() * 2, hipMemcpyHostToDevice));
                return FCIS_SUCCESS;
        }
        std::vector<double> frag(nf, 0.0);
        for (int ks = 0; ks < p->ksteps; ks++)
                for (int mt = 0; mt < mtiles; mt++)
                        for (int l = 0; l < 64; l++)
                                frag[((size_t)ks * mtiles + mt) * 64 + l] = hval(mt * 16 + (l & 15), 4 * ks + (l >> 4));
        if (p->precision == FCIS_PREC_FP64) {
                FCIS_HIP(hipMemcpy(dst, frag.data(), nf * 8, hipMemcpyHostToDevice));
        } else {
                std::vector<float> f(frag.begin(), frag.end());
                FCIS_HIP(hipMemcpy(dst, f.data(), nf * 4, hipMemcpyHostToDevice));
        }
        return FCIS_SUCCESS;
}

// The off-diagonal pairs in pair order: the rows of the fused kernel's h.
std::vector<int> offdiag_pairs(int norb)
{
        std::vector<int> out;
        for (int pp = 0; pp < norb; pp++)
                for (int q = 0; q < pp; q++) out.push_back(fcis::pair_index(pp, q));
        std::sort(out.begin(), out.end());
        return out;
}

int upload_h(fcis_sigma_plan *p, const double *h2e)
{
        const int np = p->npair;
        p->h.assign(h2e, h2e + (size_t)np * np);
        std::vector<int> natural(p->mtiles * 16);
        for (int m = 0; m < (int)natural.size(); m++) natural[m] = m < np ? m : -1;
        int st = upload_frags(p, h2e, natural, p->mtiles, p->d_hfrag, p->d_hfrag_lo);
        if (st != FCIS_SUCCESS || !p->fused) return st;
        // fused kernel: the off-diagonal rows; the diagonal pairs pp act on the
        // determinant itself, sigma[D] += sum_p occ_p(D) vt1[pp][D] =
        // sum_k (fAlpha[I][k] + fBeta[k][Ib]) t1[k][D], fAlpha / fBeta the
        // sums of h[pp][k] over the occupied orbitals of the alpha / beta string
        std::vector<int> rows = offdiag_pairs(p->space->norb);
        st = upload_frags(p, h2e, rows, p->mtiles_f, p->d_hfrag_f, p->d_hfrag_f_lo);
        if (st != FCIS_SUCCESS) return st;
        const fcis_space *s = p->space;
        const int n = s->norb;
        if (p->precision == FCIS_PREC_BF16X3) {
                // the diagonal pairs' rows for the bf16x3 kernel's diagonal tile
                // (k-steps past the problem's stay zero)
                std::vector<int> diag(n);
                for (int q = 0; q < n; q++) diag[q] = fcis::pair_index(q, q);
                char *d = static_cast<char *>(p->d_dtile);
                FCIS_HIP(hipMemset(d, 0, fused_dtile_bytes()));
                st = upload_frags(p, h2e, diag, 1, d, d + fused_dtile_bytes() / 2);
                if (st != FCIS_SUCCESS) return st;
        }
        std::vector<double> fa((size_t)s->nstra * np), fb((size_t)np * p->nbp, 0.0);
        for (int64_t a = 0; a < s->nstra; a++) {
                const uint64_t str = fcis::space_string(s, 0, a);
                for (int k = 0; k < np; k++) {
                        double v = 0.0;
                        for (int q = 0; q < n; q++)
                                if ((str >> q) & 1) v += h2e[(size_t)fcis::pair_index(q, q) * np + k];
                        fa[(size_t)a * np + k] = v;
                }
        }
        for (int64_t b = 0; b < s->nstrb; b++) {
                const uint64_t str = fcis::space_string(s, 1, b);
                for (int k = 0; k < np; k++) {
                        double v = 0.0;
                        for (int q = 0; q < n; q++)
                                if ((str >> q) & 1) v += h2e[(size_t)fcis::pair_index(q, q) * np + k];
                        fb[(size_t)k * p->nbp + b] = v;
                }
        }
        FCIS_HIP(hipMemcpy(p->d_falpha, fa.data(), 8 * fa.size(), hipMemcpyHostToDevice));
        if (p->precision == FCIS_PREC_FP64) {
                FCIS_HIP(hipMemcpy(p->d_fbeta, fb.data(), 8 * fb.size(), hipMemcpyHostToDevice));
        } else {
                std::vector<float> fb32(fb.begin(), fb.end());
                FCIS_HIP(hipMemcpy(p->d_fbeta, fb32.data(), 4 * fb32.size(), hipMemcpyHostToDevice));
        }
        return FCIS_SUCCESS;
}

}  // namespace

namespace {

// Processing order of the alpha strings of a single-tile fused plan.  The
// persistent workgroups take positions r = blockIdx.x + k * grid, and
// workgroups are handed to the 8 XCDs round robin, so the 32 workgroups an
// XCD runs together in round k hold positions with one residue mod 8.  A
// row gathers c from its alpha neighbours (strings one excitation away), and
// rows that run together on an XCD share that XCD's L2, so each XCD gets a
// block of strings with few distinct neighbours:
//   "greedy"   blocks grown from the lowest unplaced string by repeatedly
//              adding the string with most neighbours already in the
//              block's neighbourhood (about 26 distinct neighbours per row
//              at (16o,8e) against 72 unshared),
//   "xcd"      consecutive colex ranks per XCD (about 30),
//   "colex"    no reordering (position r = string r; about 44 per row).
// FCIS_ROW_ORDER selects one (default greedy).
std::vector<int32_t> greedy_blocks(const fcis_space *s, int block)
{
        const int64_t na = s->nstra;
        const int n = s->norb;
        // neighbour lists (every string one alpha excitation away, itself
        // included) from the strings
        std::vector<int32_t> nb_off(na + 1, 0), nb;
        nb.reserve((size_t)na * (s->nlinka ? s->nlinka : 1));
        for (int64_t a = 0; a < na; a++) {
                const uint64_t str = fcis::space_string(s, 0, a);
                nb.push_back((int32_t)a);
                for (int q = 0; q < n; q++) {
                        if (!((str >> q) & 1)) continue;
                        for (int pp = 0; pp < n; pp++) {
                                if ((str >> pp) & 1) continue;
                                nb.push_back((int32_t)fcis::space_addr(s, 0, (str & ~(1ull << q)) | (1ull << pp)));
                        }
                }
                nb_off[a + 1] = (int32_t)nb.size();
        }
        std::vector<int32_t> order;
        order.reserve(na);
        std::vector<char> placed(na, 0), in_u(na, 0);
        std::vector<int32_t> cnt(na, 0), touched, ulist;
        int64_t next = 0;
        while ((int64_t)order.size() < na) {
                while (placed[next]) next++;
                auto add = [&](int64_t x) {
                        placed[x] = 1;
                        order.push_back((int32_t)x);
                        for (int32_t i = nb_off[x]; i < nb_off[x + 1]; i++) {
                                const int32_t y = nb[i];
                                if (in_u[y]) continue;
                                in_u[y] = 1;
                                ulist.push_back(y);
                                for (int32_t k = nb_off[y]; k < nb_off[y + 1]; k++) {
                                        if (cnt[nb[k]]++ == 0) touched.push_back(nb[k]);
                                }
                        }
                };
                add(next);
                for (int b = 1; b < block && (int64_t)order.size() < na; b++) {
                        int64_t best = -1;
                        for (int32_t z : touched)
                                if (!placed[z] && (best < 0 || cnt[z] > cnt[best] || (cnt[z] == cnt[best] && z < best)))
                                        best = z;
                        if (best < 0) {
                                while (placed[next]) next++;
                                best = next;
                        }
                        add(best);
                }
                for (int32_t z : touched) cnt[z] = 0;
                for (int32_t y : ulist) in_u[y] = 0;
                touched.clear();
                ulist.clear();
        }
        return order;
}

int build_row_order(fcis_sigma_plan *p)
{
        const char *env = getenv("FCIS_ROW_ORDER");
        const std::string mode = env ? env : "greedy";
        const int64_t na = p->space->nstra;
        const int grid = p->fused_grid, nx = 8;
        if (mode == "colex" || grid % nx != 0 || na < grid) return FCIS_SUCCESS;
        const int per = grid / nx;
        std::vector<int32_t> order;
        if (mode == "greedy") {
                // depends on the space and the block size only: kept on the space
                static std::mutex order_mu;
                std::lock_guard<std::mutex> guard(order_mu);
                fcis_space *sm = const_cast<fcis_space *>(p->space);
                if (sm->h_order != nullptr && sm->order_block == per) {
                        order.assign(sm->h_order, sm->h_order + na);
                } else {
                        order = greedy_blocks(p->space, per);
                        int32_t *keep = (int32_t *)malloc(4 * (size_t)na);
                        if (keep != nullptr) {
                                memcpy(keep, order.data(), 4 * (size_t)na);
                                free(sm->h_order);
                                sm->h_order = keep;
                                sm->order_block = per;
                        }
                }
        } else {
                order.resize(na);
                for (int64_t a = 0; a < na; a++) order[a] = (int32_t)a;
        }
        // position r = k * grid + w runs on workgroup w (XCD w % 8) in round
        // k: full rounds give XCD x the entries x * per .. x * per + per - 1
        // of that round's slice of the order
        std::vector<int32_t> pos(na);
        for (int64_t r = 0; r < na; r++) {
                const int64_t k = r / grid, w = r - k * grid;
                const bool full = (k + 1) * grid <= na;
                pos[r] = full ? order[k * grid + (w % nx) * per + w / nx] : order[r];
        }
        if (fcis::dev_malloc(&p->d_order, 4 * (size_t)na) != hipSuccess) {
                fcis::set_error("fcis_sigma_plan_create: device allocation failed (row order)");
                return FCIS_ERR_MEMORY;
        }
        FCIS_HIP(hipMemcpy(p->d_order, pos.data(), 4 * (size_t)na, hipMemcpyHostToDevice));
        return FCIS_SUCCESS;
}

// One launch of the fused row kernel over positions [a0, a0 + ta) (alpha
// strings, or the row set's positions when ROWS), walpha in p->d_vt1.
template <bool ROWS>
void launch_fused(const fcis_sigma_plan *p, hipStream_t st, int64_t a0, int64_t ta, int64_t lda,
                  int acc, const double *ci0, double *sig, const fcis_rowset *rs)
{
        const fcis_space *s = p->space;
        const int64_t nb = s->nstrb, na = s->nstra;
        const dim3 fgrid(p->fused_grid), fblk(64 * (p->mtiles_f + p->nbld));
        const int32_t *rows = ROWS ? rs->d_rows : p->d_order;
        const int2 *rowcl = ROWS ? rs->d_rowcl : nullptr;
        const uint16_t *clist = ROWS ? rs->d_clist : nullptr;
        // the FP32 / bf16x3 kernels gather the float copy of c (p->d_cg)
#define FCIS_LAUNCH_FUSED(PR, OUT, CG)                                                                \
        hipLaunchKernelGGL((k_fused_row<PR, ROWS>), fgrid, fblk, p->fused_lds, st, p->npair,      \
                           p->mtiles_f, p->mrow_f, s->na, s->nlinka, nb, na, a0, ta, lda, acc,        \
                           s->d_linka, p->d_pmapa, p->d_pmapb16, p->nbp, CG, p->d_hfrag_f,            \
                           p->d_hfrag_f_lo, p->d_mpair, p->d_falpha,                                  \
                           (const FusedTypes<PR>::G *)p->d_fbeta, (OUT *)p->d_vt1, sig, rows, rowcl,  \
                           clist, (const bf16x8 *)p->d_dtile, s->d_strb)
        if (p->precision == FCIS_PREC_FP64)
                FCIS_LAUNCH_FUSED(FCIS_PREC_FP64, double, ci0);
        else if (p->precision == FCIS_PREC_FP32)
                FCIS_LAUNCH_FUSED(FCIS_PREC_FP32, float, p->d_cg);
        else
                FCIS_LAUNCH_FUSED(FCIS_PREC_BF16X3, float, p->d_cg);
#undef FCIS_LAUNCH_FUSED
}

// The float copy of c that the FP32 / bf16x3 fused kernels gather.
__global__ __launch_bounds__(256) void k_to_float(int64_t n, const double *__restrict__ x, float *__restrict__ y)
{
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) y[i] = (float)x[i];
}

void fused_gather_copy(const fcis_sigma_plan *p, hipStream_t st, const double *ci0)
{
        if (!p->fused || p->precision == FCIS_PREC_FP64) return;
        const int64_t n = p->space->nstra * p->space->nstrb;
        const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
        hipLaunchKernelGGL(k_to_float, dim3(blocks), dim3(256), 0, st, n, ci0, p->d_cg);
}

// The fused row kernel of a precision, ROWS or not.
template <bool ROWS>
const void *fused_kernel(int precision)
{
        if (precision == FCIS_PREC_FP64) return (const void *)k_fused_row<FCIS_PREC_FP64, ROWS>;
        if (precision == FCIS_PREC_FP32) return (const void *)k_fused_row<FCIS_PREC_FP32, ROWS>;
        return (const void *)k_fused_row<FCIS_PREC_BF16X3, ROWS>;
}

// Alpha pull onto the wanted determinants of a row set, for the positions
// [r0, r0 + tr) of one launch: one thread per wanted determinant m =
// dets[d] = (J, Jb) adds sum_k s_k walpha[rslot_k][rpos[I_k] - r0][Jb] over
// the alpha links k of J whose source I_k has its position in the range.
// The last launch (out != nullptr) writes out[d] = sigma[m] + that sum.
template <typename T>
__global__ __launch_bounds__(256) void k_alpha_pull_rows(
        int64_t ndets, const uint32_t *__restrict__ dets, int64_t nb, int64_t wrs, int64_t nstra, int64_t r0,
        int64_t tr, int64_t ld, int kfirst, int nlinka, const fcis_link *__restrict__ linka,
        const int32_t *__restrict__ rpos, const T *__restrict__ walpha, double *__restrict__ sigma,
        double *__restrict__ out)
{
        for (int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; d < ndets;
             d += (int64_t)gridDim.x * blockDim.x) {
                const int64_t m = dets[d];
                const int64_t J = m / nb, Jb = m - J * nb;
                // U links, their positions and U walpha loads in flight at a
                // time; the sum runs in link order
                constexpr int U = 8;
                double acc = 0.0;
                int k = kfirst;
                for (; k + U <= nlinka; k += U) {
                        fcis_link L[U];
                        int32_t pos[U];
                        T v[U];
#pragma unroll
                        for (int u = 0; u < U; u++) L[u] = linka[(int64_t)(k + u) * nstra + J];
#pragma unroll
                        for (int u = 0; u < U; u++) pos[u] = rpos[L[u].addr];
#pragma unroll
                        for (int u = 0; u < U; u++) {
                                const bool in = pos[u] >= r0 && pos[u] < r0 + tr;
                                v[u] = walpha[in ? (int64_t)(link_rslot(L[u].info) - kfirst) * ld +
                                                           (pos[u] - r0) * wrs + Jb
                                                 : Jb];
                        }
#pragma unroll
                        for (int u = 0; u < U; u++) {
                                if (pos[u] >= r0 && pos[u] < r0 + tr) {
                                        const double x = (double)v[u];
                                        acc += link_neg(L[u].info) ? -x : x;
                                }
                        }
                }
                for (; k < nlinka; k++) {
                        const fcis_link L = linka[(int64_t)k * nstra + J];
                        const int64_t pos = rpos[L.addr];
                        if (pos < r0 || pos >= r0 + tr) continue;
                        const double v = (double)walpha[(int64_t)(link_rslot(L.info) - kfirst) * ld +
                                                        (pos - r0) * wrs + Jb];
                        acc += link_neg(L.info) ? -v : v;
                }
                if (out)
                        out[d] = sigma[m] + acc;
                else
                        sigma[m] += acc;
        }
}

}  // namespace

int fcis::sigma_plan_execute_rows(fcis_sigma_plan *p, const fcis_rowset *rs, const double *ci0,
                                  double *sig, double *out, hipStream_t st)
{
        if (p == nullptr || rs == nullptr || ci0 == nullptr || sig == nullptr || out == nullptr ||
            ci0 == sig || !p->fused) {
                fcis::set_error("sigma_plan_execute_rows: invalid argument or unfused plan");
                return FCIS_ERR_INVALID;
        }
        const fcis_space *s = p->space;
        const int64_t nb = s->nstrb, na = s->nstra;
        if (rs->nrows == 0 || rs->ndets == 0) {
                if (rs->ndets > 0)
                        hipLaunchKernelGGL(k_zero, dim3((rs->ndets + 255) / 256), dim3(256), 0, st,
                                           rs->ndets, out);
                FCIS_CHECK_LAUNCH();
                return FCIS_SUCCESS;
        }
        const int64_t tile = p->tile_a;
        const int ntiles = (int)((rs->nrows + tile - 1) / tile);
        const int acc = ntiles > 1 ? 1 : 0;
        fused_gather_copy(p, st, ci0);
        if (acc)
                hipLaunchKernelGGL(k_zero, dim3((na * nb + 255) / 256), dim3(256), 0, st, na * nb, sig);
        const int64_t pgrid = std::min<int64_t>((rs->ndets + 255) / 256, 16384);
        for (int t = 0; t < ntiles; t++) {
                const int64_t r0 = (int64_t)t * tile;
                const int64_t tr = std::min<int64_t>(tile, rs->nrows - r0);
                const int64_t lda = tr * p->nbp;   // walpha rows nbp apart
                double *o = t == ntiles - 1 ? out : nullptr;
                if (p->timing) FCIS_HIP(hipEventRecord(p->ev[0], st));
                if (p->timing) FCIS_HIP(hipEventRecord(p->ev[1], st));
                launch_fused<true>(p, st, r0, tr, lda, acc, ci0, sig, rs);
                if (p->timing) FCIS_HIP(hipEventRecord(p->ev[2], st));
                if (p->precision == FCIS_PREC_FP64)
                        hipLaunchKernelGGL(k_alpha_pull_rows<double>, dim3(pgrid), dim3(256), 0, st,
                                           rs->ndets, rs->d_dets, nb, p->nbp, na, r0, tr, lda, s->na, s->nlinka,
                                           s->d_linka, rs->d_rpos, (const double *)p->d_vt1, sig, o);
                else
                        hipLaunchKernelGGL(k_alpha_pull_rows<float>, dim3(pgrid), dim3(256), 0, st,
                                           rs->ndets, rs->d_dets, nb, p->nbp, na, r0, tr, lda, s->na, s->nlinka,
                                           s->d_linka, rs->d_rpos, (const float *)p->d_vt1, sig, o);
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

int fcis::sigma_plan_create(fcis_sigma_plan **out, const fcis_space *space, const double *h2e,
                            int precision, uint64_t workspace_bytes, bool honour_env)
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
        // The fused row path keeps h in registers (npair <= 136), sigma row I
        // in LDS (nb <= 16384, int16 beta pair map) and needs reverse link
        // slots (nlink <= 255).  FCIS_DETERMINISTIC selects the bitwise
        // reproducible three-kernel path instead.
        const size_t lds = precision == FCIS_PREC_FP64 ? fused_lds_bytes<FCIS_PREC_FP64>(nb)
                         : precision == FCIS_PREC_FP32 ? fused_lds_bytes<FCIS_PREC_FP32>(nb)
                                                        : fused_lds_bytes<FCIS_PREC_BF16X3>(nb);
        // fused rows: the npair - norb off-diagonal pairs and the diagonal row
        p->mrow_f = space->npair - space->norb;
        p->mtiles_f = (p->mrow_f + 1 + 15) / 16;
        p->fused = p->mtiles <= FUSED_MAXW && p->mtiles_f <= FUSED_MTILES &&
                   p->ksteps <= (precision == FCIS_PREC_BF16X3 ? FUSED_KS_BF16 : FUSED_KS) &&
                   nb < FUSED_MAXNB && space->nlinka <= 255 && lds <= 160 * 1024 &&
                   (uint64_t)space->nstra * (uint64_t)nb * 8ull < (uint64_t)FUSED_OOB &&
                   !(honour_env && getenv("FCIS_DETERMINISTIC"));
        if (p->fused) {
                int dev = 0, ncu = 0, occ = 0;
                FCIS_HIP(hipGetDevice(&dev));
                FCIS_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
                const void *fk = fused_kernel<false>(precision);
                const void *fkr = fused_kernel<true>(precision);
                FCIS_HIP(hipFuncSetAttribute(fk, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                FCIS_HIP(hipFuncSetAttribute(fkr, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                p->nbld = FUSED_NBLD;
                p->nbp = (nb + FUSED_NB - 1) / FUSED_NB * FUSED_NB;
                FCIS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fk, 64 * (p->mtiles_f + p->nbld), lds));
                if (occ < 1) occ = 1;
                p->fused_grid = ncu * occ;
                p->fused_lds = lds;
        }
        // workspace per alpha string: walpha (the nlinka - na off-diagonal
        // alpha links, fused path) or t1 + vt1 (npair rows each)
        const int nwslot = space->nlinka - space->na;
        const uint64_t rows_per_alpha = p->fused ? (uint64_t)nwslot : 2ull * p->npair;
        const int64_t wrs = p->fused ? p->nbp : nb;   // walpha rows padded to whole chunks
        uint64_t per_alpha = rows_per_alpha * (uint64_t)wrs * elem;
        int64_t ta = per_alpha ? (int64_t)(workspace_bytes / per_alpha) : space->nstra;
        if (ta < 1) ta = 1;
        if (ta > space->nstra) ta = space->nstra;
        p->tile_a = ta;
        p->ntiles = (int)((space->nstra + ta - 1) / ta);
        size_t buf = (size_t)(p->fused ? nwslot : p->npair) * ta * wrs * elem;
        p->ws_bytes = p->fused ? buf : 2 * buf;
        const size_t nf = frag_count(p);
        size_t hbytes = precision == FCIS_PREC_BF16X3 ? nf * 16 : nf * elem;
        const size_t nff = (size_t)p->ksteps * p->mtiles_f * 64;
        const size_t hfbytes = precision == FCIS_PREC_BF16X3 ? nff * 16 : nff * elem;
        size_t pa = (size_t)p->npair * space->nstra * 4, pb = (size_t)p->npair * nb * 4;
        if (fcis::dev_malloc(&p->d_hfrag, hbytes) != hipSuccess ||
            fcis::dev_malloc(&p->d_hfrag_lo, precision == FCIS_PREC_BF16X3 ? hbytes : 16) != hipSuccess ||
            fcis::dev_malloc(&p->d_pmapa, pa ? pa : 4) != hipSuccess ||
            fcis::dev_malloc(&p->d_pmapb, pb ? pb : 4) != hipSuccess ||
            fcis::dev_malloc(&p->d_pmapb16, p->fused ? (size_t)p->npair * p->nbp * 2 + 16 : 16) != hipSuccess ||
            fcis::big_malloc(&p->d_t1, (buf && !p->fused) ? buf : 8) != hipSuccess ||
            fcis::big_malloc(&p->d_vt1, buf ? buf : 8) != hipSuccess ||
            (p->fused && (fcis::dev_malloc(&p->d_hfrag_f, hfbytes) != hipSuccess ||
                          fcis::dev_malloc(&p->d_hfrag_f_lo, precision == FCIS_PREC_BF16X3 ? hfbytes : 16) !=
                                  hipSuccess ||
                          fcis::dev_malloc(&p->d_mpair, 2 * 16 * FUSED_MTILES) != hipSuccess ||
                          fcis::dev_malloc(&p->d_falpha, 8 * (size_t)space->nstra * p->npair) != hipSuccess ||
                          fcis::dev_malloc(&p->d_fbeta, elem * (size_t)p->npair * p->nbp) != hipSuccess ||
                          fcis::big_malloc(&p->d_cg, precision == FCIS_PREC_FP64
                                                             ? 16
                                                             : 4 * (size_t)space->nstra * nb) != hipSuccess ||
                          fcis::dev_malloc(&p->d_dtile, fused_dtile_bytes()) != hipSuccess))) {
                fcis_sigma_plan_free(p);
                fcis::set_error("fcis_sigma_plan_create: device allocation of %zu bytes failed",
                                2 * buf + 2 * hbytes + pa + pb);
                return FCIS_ERR_MEMORY;
        }
        if (p->fused) {
                std::vector<int> rows = offdiag_pairs(space->norb);
                std::vector<int16_t> mp(16 * FUSED_MTILES, -1);
                for (size_t m = 0; m < rows.size(); m++) mp[m] = (int16_t)rows[m];
                FCIS_HIP(hipMemcpy(p->d_mpair, mp.data(), 2 * mp.size(), hipMemcpyHostToDevice));
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
        if (p->fused && pb > 0) {
                int64_t n16 = (int64_t)p->npair * p->nbp;
                hipLaunchKernelGGL(k_pmap16, dim3((n16 + 255) / 256), dim3(256), 0, 0, p->npair, nb,
                                   p->nbp, p->d_pmapb, p->d_pmapb16);
                FCIS_CHECK_LAUNCH();
        }
        if (p->fused && p->ntiles == 1) {
                int st2 = build_row_order(p);
                if (st2 != FCIS_SUCCESS) {
                        fcis_sigma_plan_free(p);
                        return st2;
                }
        }
        FCIS_HIP(hipDeviceSynchronize());
        *out = p;
        return FCIS_SUCCESS;
}

extern "C" {

int fcis_sigma_plan_create(fcis_sigma_plan_t **out, const fcis_space_t *space,
                           const double *h2e, int precision, uint64_t workspace_bytes)
{
        return fcis::sigma_plan_create(out, space, h2e, precision, workspace_bytes, true);
}

int fcis_sigma_plan_path(const fcis_sigma_plan_t *p)
{
        if (p == nullptr) return FCIS_ERR_INVALID;
        return p->fused ? 1 : 0;
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
        fused_gather_copy(p, st, ci0);
        for (int t = 0; t < p->ntiles; t++) {
                const int64_t a0 = (int64_t)t * p->tile_a;
                const int64_t ta = na - a0 < p->tile_a ? na - a0 : p->tile_a;
                const int64_t ndet = ta * nb;
                const int64_t ld = ndet;
                const dim3 ggrid((ndet + GEMM_NCOL - 1) / GEMM_NCOL, (p->mtiles + GEMM_MT - 1) / GEMM_MT);
                const dim3 rgrid(ta * ((nb + 255) / 256)), sgrid(na * ((nb + 255) / 256));
                if (p->timing) FCIS_HIP(hipEventRecord(p->ev[0], st));
                if (p->fused) {
                        const int acc = p->ntiles > 1 ? 1 : 0;
                        const int64_t lda = ta * p->nbp;   // walpha slot stride (rows nbp apart)
                        if (t == 0 && acc)
                                hipLaunchKernelGGL(k_zero, dim3((na * nb + 255) / 256), dim3(256), 0, st,
                                                   na * nb, ci1);
                        if (p->timing) FCIS_HIP(hipEventRecord(p->ev[1], st));
                        launch_fused<false>(p, st, a0, ta, lda, acc, ci0, ci1, nullptr);
                        if (p->timing) FCIS_HIP(hipEventRecord(p->ev[2], st));
                        const dim3 pgrid(na * (((nb + 1) / 2 + 255) / 256));   // two determinants per thread
                        if (p->precision == FCIS_PREC_FP64)
                                hipLaunchKernelGGL(k_alpha_pull<double>, pgrid, dim3(256), 0, st, nb, p->nbp, na, a0,
                                                   ta, lda, s->na, s->nlinka, s->d_linka,
                                                   (const double *)p->d_vt1, ci1);
                        else
                                hipLaunchKernelGGL(k_alpha_pull<float>, pgrid, dim3(256), 0, st, nb, p->nbp, na, a0,
                                                   ta, lda, s->na, s->nlinka, s->d_linka,
                                                   (const float *)p->d_vt1, ci1);
                } else if (p->precision == FCIS_PREC_FP64) {
                        double *t1 = (double *)p->d_t1, *vt1 = (double *)p->d_vt1;
                        hipLaunchKernelGGL(k_gather<double>, rgrid, dim3(256), 0, st, p->npair, nb, a0,
                                           na, ld, p->d_pmapa, p->d_pmapb, ci0, t1);
                        if (p->timing) FCIS_HIP(hipEventRecord(p->ev[1], st));
                        hipLaunchKernelGGL(k_contract_f64, ggrid, gblk, 0, st, p->npair, p->mtiles,
                                           p->ksteps, ndet, ld, (const double *)p->d_hfrag, t1, vt1);
                        if (p->timing) FCIS_HIP(hipEventRecord(p->ev[2], st));
                        hipLaunchKernelGGL(k_scatter<double>, sgrid, dim3(256), s->nlinka * sizeof(fcis_link), st, nb, na, a0, ta,
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
                        hipLaunchKernelGGL(k_scatter<float>, sgrid, dim3(256), s->nlinka * sizeof(fcis_link), st, nb, na, a0, ta,
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
        (void)hipFree(p->d_pmapb16);
        fcis::big_free(p->d_t1);
        fcis::big_free(p->d_vt1);
        (void)hipFree(p->d_order);
        (void)hipFree(p->d_hfrag_f);
        (void)hipFree(p->d_hfrag_f_lo);
        (void)hipFree(p->d_mpair);
        (void)hipFree(p->d_falpha);
        (void)hipFree(p->d_fbeta);
        fcis::big_free(p->d_cg);
        (void)hipFree(p->d_dtile);
        delete p;
}

}  // extern "C"
