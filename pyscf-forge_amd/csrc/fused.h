// Building blocks of the fused row kernels (single problem: sigma.hip,
// batched problems: batch.hip).  gfx950 only.
#pragma once

#include <type_traits>
#include <vector>

#include "common.h"

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef int v4i __attribute__((ext_vector_type(4)));

namespace {

constexpr int FUSED_NB = 16;      // determinants per chunk (one MFMA column block)
constexpr int FUSED_MAXW = 9;     // 16-row tiles of npair (norb <= 16): B tile and LDS sizing
constexpr int FUSED_MTILES = 8;   // MFMA wavefronts of the single-problem fused kernel
constexpr int FUSED_NBLD = 8;     // its builder wavefronts (7 gathering + 1 pair-map / diagonal)
constexpr int FUSED_KS = 34;      // k-steps held in registers (4-wide MFMA): npair <= 136
constexpr int FUSED_KS_BF16 = 5;  // k-steps held in registers (32-wide bf16 MFMA)
constexpr int FUSED_NACC = 2;     // independent accumulator chains (FP64/FP32)
constexpr int FUSED_MAXNB = 16384;  // sigma row kept in LDS; pmapb fits int16
constexpr int FUSED_NMAP = 4;       // beta pair-map slices in LDS (chunks j .. j+3)

template <int PREC>
struct FusedTypes;
template <>
struct FusedTypes<FCIS_PREC_FP64> {
        typedef double G;    // element type the builders gather from c
        typedef double B;
        typedef double A;
        typedef v4d Acc;
        typedef double Out;
        static constexpr int KS = FUSED_KS;
        static constexpr int KW = 4;
        static constexpr int NBLD = 7;    // 9 + 7 wavefronts: 4 per SIMD at <= 128 VGPRs
};
template <>
struct FusedTypes<FCIS_PREC_FP32> {
        typedef float G;    // element type the builders gather from c
        typedef float B;
        typedef float A;
        typedef v4f Acc;
        typedef float Out;
        static constexpr int KS = FUSED_KS;
        static constexpr int KW = 4;
        static constexpr int NBLD = 7;    // 9 + 7 wavefronts: 4 per SIMD at <= 128 VGPRs
};
template <>
struct FusedTypes<FCIS_PREC_BF16X3> {
        typedef float G;    // element type the builders gather from c
        typedef float B;
        typedef bf16x8 A;
        typedef v4f Acc;
        typedef float Out;
        static constexpr int KS = FUSED_KS_BF16;
        static constexpr int KW = 32;
        static constexpr int NBLD = 7;
};

// bf16x3 fused kernel: the builders split each B element into bf16 high and
// low parts once and store them column-major, [col][k] with this k stride
// (padded against LDS bank conflicts), so that an MFMA wavefront reads its
// 8 k-values of a column as one 16-byte load instead of converting every
// element itself.
constexpr int FUSED_KPS_BF16 = FUSED_KS_BF16 * 32 + 8;

// The bf16x3 fused kernel's diagonal tile: A fragments (bf16 high, then low
// parts, [k-step][lane]) of the 16 rows h[pp][k], p < norb, kept in LDS.
constexpr size_t fused_dtile_bytes()
{
        return 2 * sizeof(bf16x8) * FUSED_KS_BF16 * 64;
}

// Bytes of one B buffer of the fused row kernel.
template <int PREC>
constexpr size_t fused_bbytes()
{
        return PREC == FCIS_PREC_BF16X3
                       ? 2 * sizeof(__bf16) * FUSED_NB * FUSED_KPS_BF16
                       : sizeof(typename FusedTypes<PREC>::B) * FusedTypes<PREC>::KS * FusedTypes<PREC>::KW *
                                 FUSED_NB;
}

template <int PREC>
constexpr size_t fused_lds_bytes(int64_t nb)
{
        return (size_t)((nb + FUSED_NB - 1) / FUSED_NB * FUSED_NB) * sizeof(double) + 2 * fused_bbytes<PREC>() +
               2 * sizeof(int) * 16 * FUSED_MAXW +
               FUSED_NMAP * sizeof(int16_t) * 16 * FUSED_MAXW * FUSED_NB +
               (PREC == FCIS_PREC_BF16X3 ? fused_dtile_bytes() : 0);
}

// Builder wavefronts: gather the B tile of chunk j+1 (pair maps, then c)
// while the MFMA wavefronts consume chunk j.
// Builder state of one chunk: the B-tile element values in flight and the
// next-but-one chunk's beta pair-map slice in flight.
template <int NBLD, int MAXW = FUSED_MAXW>
struct FusedBuild {
        // tile rows: 16 per MFMA wavefront, at most npair <= 136 (norb <= 16)
        static constexpr int ROWS = 16 * MAXW < 136 ? 16 * MAXW : 136;
        static constexpr int EB = (ROWS * FUSED_NB + 64 * NBLD - 1) / (64 * NBLD);
        static constexpr int EM = (2 * 16 * MAXW + 64 * NBLD - 1) / (64 * NBLD);
        static_assert(EB <= 16, "element flags hold 16 elements");
        typedef typename std::conditional<(EB <= 8), uint32_t, uint64_t>::type Flags;
        Flags fl;             // element s: bit 4s alpha term, 4s+1 its sign, 4s+2 beta term, 4s+3 sign
        double va[EB], vb[EB];
        v4i mv[EM];           // 8 int16 map entries per 16-byte load
        int mr[EB];           // beta pair-map entry of each element (int16 bits)
};

// Beta pair-map entries of this thread's elements of chunk j straight from
// pmapb into registers (element (k, col) -> pmapb[k][b0 + col]; loads
// clamped to the tile, no branches).
template <int NBLD, int MAXW>
__device__ __forceinline__ void fb_issue_mapv(FusedBuild<NBLD, MAXW> &st, int npair, int64_t nbp, int64_t j,
                                              int bt, const int16_t *__restrict__ pmapb)
{
        constexpr int NB = FUSED_NB, EB = FusedBuild<NBLD, MAXW>::EB;
        const int nbt = NBLD * 64, nelem = npair * NB;
        const uint16_t *pm = reinterpret_cast<const uint16_t *>(pmapb) + j * NB;
#pragma unroll
        for (int s = 0; s < EB; s++) {
                const int e = bt + s * nbt, ec = e < nelem ? e : nelem - 1;
                const int k = ec / NB;
                st.mr[s] = pm[(int64_t)k * nbp + (ec - k * NB)];
        }
}

// The entries of fb_issue_mapv into the LDS pair-map slice of the chunk
// (read by the scatter of the MFMA wavefronts).
template <int NBLD, int MAXW>
__device__ __forceinline__ void fb_store_mapv(const FusedBuild<NBLD, MAXW> &st, int npair, int bt,
                                              int16_t *__restrict__ mb)
{
        constexpr int EB = FusedBuild<NBLD, MAXW>::EB;
        const int nbt = NBLD * 64, nelem = npair * FUSED_NB;
#pragma unroll
        for (int s = 0; s < EB; s++) {
                const int e = bt + s * nbt;
                if (e < nelem) mb[e] = (int16_t)st.mr[s];
        }
}

// Issue the c loads of chunk j (element (k, col) of the tile is
//   sa c[Ja, b0+col] + sb c[I, Jb],  arow[k] -> (Ja, sa), mb[k][col] -> (Jb, sb)),
// addressed with 32-bit byte offsets (the fused path needs c below 4 GiB).
template <int NBLD, int MAXW, bool REGMAP = false>
__device__ __forceinline__ void fb_issue_c(FusedBuild<NBLD, MAXW> &st, int npair, int64_t nb, int64_t I,
                                           int64_t j, int bt, const int *__restrict__ arow,
                                           const int16_t *__restrict__ mb,
                                           const double *__restrict__ c)
{
        // REGMAP: the beta map entries come from st.mr (fb_issue_mapv), not
        // from the LDS slice mb
        constexpr int NB = FUSED_NB, EB = FusedBuild<NBLD, MAXW>::EB;
        const int nbt = NBLD * 64, nelem = npair * NB;
        const uint32_t b0 = (uint32_t)(j * NB);
        const char *cb = reinterpret_cast<const char *>(c);
        const uint32_t rowI = (uint32_t)(I * nb);
        // No branches around the loads: every lane loads (invalid elements
        // from c[0], discarded at commit), so that the number of loads in
        // flight is static and a commit waits only for its own chunk's loads
        // while the next chunk's are outstanding.
        typedef typename FusedBuild<NBLD, MAXW>::Flags Flags;
        Flags fl = 0;
        uint32_t oa[EB], ob[EB];
#pragma unroll
        for (int s = 0; s < EB; s++) {
                const int e = bt + s * nbt;
                const int ec = e < nelem ? e : nelem - 1;
                const int k = ec / NB;
                const int col = ec - k * NB;
                const bool in = e < nelem && b0 + (uint32_t)col < (uint32_t)nb;
                // m = -1 where b has no link with pair k
                const int a = arow[k], m = REGMAP ? (int)(int16_t)st.mr[s] : (int)mb[ec];
                const bool ia = in && a >= 0, ib = in && m >= 0;
                fl |= (Flags)((ia ? 1u : 0u) | ((uint32_t)(a & 1) << 1) | (ib ? 4u : 0u) |
                              ((uint32_t)(m & 1) << 3)) << (4 * s);
                oa[s] = ia ? ((uint32_t)(a >> 1) * (uint32_t)nb + b0 + col) * 8u : 0u;
                ob[s] = ib ? (rowI + (uint32_t)(m >> 1)) * 8u : 0u;
        }
        st.fl = fl;
#pragma unroll
        for (int s = 0; s < EB; s++) {
                st.va[s] = *reinterpret_cast<const double *>(cb + oa[s]);
                st.vb[s] = *reinterpret_cast<const double *>(cb + ob[s]);
        }
}

template <int NBLD, int MAXW, typename BT>
__device__ __forceinline__ void fb_commit_c(FusedBuild<NBLD, MAXW> &st, int npair, int bt, BT *bb)
{
        constexpr int EB = FusedBuild<NBLD, MAXW>::EB;
        const int nbt = NBLD * 64, nelem = npair * FUSED_NB;
#pragma unroll
        for (int s = 0; s < EB; s++) {
                int e = bt + s * nbt;
                if (e < nelem) {
                        const uint32_t f = (uint32_t)(st.fl >> (4 * s));
                        const double va = (f & 1) ? ((f & 2) ? -st.va[s] : st.va[s]) : 0.0;
                        const double vb = (f & 4) ? ((f & 8) ? -st.vb[s] : st.vb[s]) : 0.0;
                        bb[e] = (BT)(va + vb);
                }
        }
}

// fb_commit_c for the bf16x3 kernel: element (k, col) split into bf16 high
// and low parts, bh[col][k] and bl[col][k] (k stride FUSED_KPS_BF16).
template <int NBLD, int MAXW>
__device__ __forceinline__ void fb_commit_split(FusedBuild<NBLD, MAXW> &st, int npair, int bt,
                                                __bf16 *__restrict__ bh)
{
        constexpr int EB = FusedBuild<NBLD, MAXW>::EB, NB = FUSED_NB, KP = FUSED_KPS_BF16;
        __bf16 *bl = bh + NB * KP;
        const int nbt = NBLD * 64, nelem = npair * NB;
#pragma unroll
        for (int s = 0; s < EB; s++) {
                const int e = bt + s * nbt;
                if (e < nelem) {
                        const uint32_t f = (uint32_t)(st.fl >> (4 * s));
                        const double va = (f & 1) ? ((f & 2) ? -st.va[s] : st.va[s]) : 0.0;
                        const double vb = (f & 4) ? ((f & 8) ? -st.vb[s] : st.vb[s]) : 0.0;
                        const float x = (float)(va + vb);
                        const __bf16 hi = (__bf16)x;
                        const int k = e / NB, col = e - k * NB;
                        bh[col * KP + k] = hi;
                        bl[col * KP + k] = (__bf16)(x - (float)hi);
                }
        }
}

// Beta pair-map slice of chunk j, pmapb[k][b0 .. b0+NB) for k < npair (16
// int16 = two 16-byte loads per pair; pmapb rows are padded to a multiple of
// FUSED_NB entries with -1, so no load crosses a row).
template <int NBLD, int MAXW>
__device__ __forceinline__ void fb_issue_maps(FusedBuild<NBLD, MAXW> &st, int npair, int64_t nbp,
                                              int64_t j, int bt, const int16_t *__restrict__ pmapb)
{
        constexpr int EM = FusedBuild<NBLD, MAXW>::EM;
        const int64_t b0 = j * FUSED_NB;
#pragma unroll
        for (int s = 0; s < EM; s++) {
                // unconditional (clamped) load, as in fb_issue_c
                const int t = bt + s * NBLD * 64, tc = t < 2 * npair ? t : 2 * npair - 1;
                const int k = tc >> 1, h = (tc & 1) * 8;
                st.mv[s] = *reinterpret_cast<const v4i *>(pmapb + (int64_t)k * nbp + b0 + h);
        }
}

template <int NBLD, int MAXW>
__device__ __forceinline__ void fb_commit_maps(FusedBuild<NBLD, MAXW> &st, int npair, int bt,
                                               int16_t *__restrict__ mb)
{
        constexpr int EM = FusedBuild<NBLD, MAXW>::EM;
#pragma unroll
        for (int s = 0; s < EM; s++) {
                int t = bt + s * NBLD * 64;
                if (t < 2 * npair) {
                        int k = t >> 1, h = (t & 1) * 8;
                        *reinterpret_cast<v4i *>(mb + k * FUSED_NB + h) = st.mv[s];
                }
        }
}

// ---- Builders of the single-problem fused kernel: c read through a raw
// buffer resource, so that absent links load zero instead of being selected
// away.  A raw buffer load at a byte offset >= num_records returns 0; the
// builders give every absent link the offset FUSED_OOB, which stays out of
// range for any c of the fused path (below 4 GiB - 1 MiB) even after a
// column offset (< 128 KiB) is added.  The FP64 kernel gathers c itself; the
// FP32 and bf16x3 kernels, whose tile elements are rounded to float anyway,
// gather a float copy of c (half the bytes per lane through the vector
// memory pipeline, the builders' bound).
constexpr uint32_t FUSED_OOB = 0xFFF00000u;

template <typename V>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t fused_c_rsrc(const V *c, uint32_t bytes)
{
        return __builtin_amdgcn_make_buffer_rsrc(const_cast<V *>(c), (short)0, (int)bytes, 0x00020000);
}

template <typename V>
__device__ __forceinline__ V fused_c_load(__amdgpu_buffer_rsrc_t rc, uint32_t off)
{
        if constexpr (sizeof(V) == 8)
                return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b64(rc, off, 0, 0));
        else
                return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b32(rc, off, 0, 0));
}

// -x for bit = 1, x for bit = 0 (sign flip without selects)
__device__ __forceinline__ double fused_flip(double x, uint32_t bit)
{
        return __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, x) ^ ((uint64_t)bit << 63));
}

__device__ __forceinline__ float fused_flip(float x, uint32_t bit)
{
        return __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, x) ^ (bit << 31));
}

// Builder state of one chunk: c values in flight (element type V) and their
// two signs.
template <int NBLD, int MAXW, typename V>
struct FusedGather {
        // tile rows: 16 per MFMA wavefront, at most npair <= 136 (norb <= 16)
        static constexpr int ROWS = 16 * MAXW < 136 ? 16 * MAXW : 136;
        static constexpr int EB = (ROWS * FUSED_NB + 64 * NBLD - 1) / (64 * NBLD);
        static_assert(EB <= 16, "sign flags hold 16 elements");
        typedef uint32_t Flags;
        Flags sg;             // element s: bit 2s alpha sign, 2s+1 beta sign
        V va[EB], vb[EB];
        V vf[EB];             // fBeta[k][b] of each element (diagonal term)
};

// Issue the loads of chunk j: element e = bt + s * NBLD * 64 = (k, col),
// k = e / 16, of the tile is sa c[Ja, b0 + col] + sb c[I, Jb], with
// abyte[k] = byte offset of c row Ja (sa in bit 0; FUSED_OOB if I has no
// alpha link with pair k, also for k >= npair) and mb[e] = Jb << 1 | sb or
// -1 (the chunk's beta pair-map slice).  With DIAG, also fBeta[k][b0 + col]
// of every element (rows nbp apart, through fr: rows past npair read zero)
// for the diagonal term.  Elements past the tile or past the row's end load
// values that are never used (or zeros).
template <int NBLD, int MAXW, bool DIAG, typename V>
__device__ __forceinline__ void fg_issue(FusedGather<NBLD, MAXW, V> &st, uint32_t rowI_byte, int64_t j, int bt,
                                         const uint32_t *__restrict__ abyte, const int16_t *__restrict__ mb,
                                         __amdgpu_buffer_rsrc_t rc, __amdgpu_buffer_rsrc_t fr, int64_t nbp)
{
        constexpr int NB = FUSED_NB, EB = FusedGather<NBLD, MAXW, V>::EB;
        constexpr int nbt = NBLD * 64;
        constexpr uint32_t W = sizeof(V);
        static_assert(EB * nbt <= 16 * MAXW * FUSED_NB, "elements stay in the map slice");
        const uint32_t cbyte = (uint32_t)(j * NB) * W;
        typedef typename FusedGather<NBLD, MAXW, V>::Flags Flags;
        Flags sg = 0;
        uint32_t oa[EB], ob[EB];
#pragma unroll
        for (int s = 0; s < EB; s++) {
                const int e = bt + s * nbt;
                const uint32_t av = abyte[e >> 4];
                const int m = (int)mb[e];
                oa[s] = (av & ~1u) + cbyte + (uint32_t)(e & 15) * W;
                ob[s] = m >= 0 ? rowI_byte + (uint32_t)(m >> 1) * W : FUSED_OOB;
                sg |= (Flags)((av & 1u) | ((uint32_t)(m & 1) << 1)) << (2 * s);
        }
        st.sg = sg;
#pragma unroll
        for (int s = 0; s < EB; s++) {
                st.va[s] = fused_c_load<V>(rc, oa[s]);
                st.vb[s] = fused_c_load<V>(rc, ob[s]);
        }
        if constexpr (DIAG) {
                // element s of this thread: row (bt >> 4) + s * nbt / 16, column bt & 15
                static_assert(nbt % 16 == 0, "a thread keeps its column");
                const uint32_t voff = (uint32_t)(((int64_t)(bt >> 4) * nbp + j * NB + (bt & 15)) * W);
#pragma unroll
                for (int s = 0; s < EB; s++)
                        st.vf[s] = fused_c_load<V>(fr, voff + (uint32_t)(s * (nbt / 16) * nbp * W));
        }
}

// Commit the tile elements, [k][col] in BT; with DIAG, return sum_s fBeta *
// t1 over this thread's elements (its column's part of the diagonal term).
template <int NBLD, int MAXW, bool DIAG = false, typename V, typename BT>
__device__ __forceinline__ double fg_commit(const FusedGather<NBLD, MAXW, V> &st, int npair, int bt, BT *bb)
{
        constexpr int EB = FusedGather<NBLD, MAXW, V>::EB;
        constexpr int nbt = NBLD * 64;
        const int nelem = npair * FUSED_NB;
        V d = 0;
#pragma unroll
        for (int s = 0; s < EB; s++) {
                const int e = bt + s * nbt;
                if (e < nelem) {
                        const uint32_t f = (uint32_t)(st.sg >> (2 * s));
                        const V t = fused_flip(st.va[s], f & 1u) + fused_flip(st.vb[s], (f >> 1) & 1u);
                        bb[e] = (BT)t;
                        if constexpr (DIAG) d += st.vf[s] * t;
                }
        }
        return (double)d;
}

// fg_commit for the bf16x3 kernel (column-major high / low parts, as
// fb_commit_split).
template <int NBLD, int MAXW, bool DIAG = false, typename V>
__device__ __forceinline__ double fg_commit_split(const FusedGather<NBLD, MAXW, V> &st, int npair, int bt,
                                                  __bf16 *__restrict__ bh)
{
        constexpr int EB = FusedGather<NBLD, MAXW, V>::EB, NB = FUSED_NB, KP = FUSED_KPS_BF16;
        constexpr int nbt = NBLD * 64;
        __bf16 *bl = bh + NB * KP;
        const int nelem = npair * NB;
        V d = 0;
#pragma unroll
        for (int s = 0; s < EB; s++) {
                const int e = bt + s * nbt;
                if (e < nelem) {
                        const uint32_t f = (uint32_t)(st.sg >> (2 * s));
                        const V t = fused_flip(st.va[s], f & 1u) + fused_flip(st.vb[s], (f >> 1) & 1u);
                        const float x = (float)t;
                        const __bf16 hi = (__bf16)x;
                        const int k = e >> 4, col = e & 15;
                        bh[col * KP + k] = hi;
                        bl[col * KP + k] = (__bf16)(x - (float)hi);
                        if constexpr (DIAG) d += st.vf[s] * t;
                }
        }
        return (double)d;
}

// One chunk of the contraction on MFMA: the 16 x 16 block
//   acc[i][j] = sum_k h[16 w + i][k] t1[k][j]
// over exactly KSP k-steps (4 k per step; 32 for bf16x3), without
// conditions: A fragments past the problem's k-steps must be zero in
// registers and B rows past npair zero in LDS.  ah/al: this wavefront's A
// fragments (al: bf16 low halves), bb: the B tile in LDS ([k][FUSED_NB]).
template <int PREC, int KSP, int KS>
__device__ __forceinline__ typename FusedTypes<PREC>::Acc mma_steps(
        const typename FusedTypes<PREC>::A (&ah)[KS],
        const typename FusedTypes<PREC>::A (&al)[PREC == FCIS_PREC_BF16X3 ? KS : 1],
        const typename FusedTypes<PREC>::B *bb, int lk, int li)
{
        static_assert(KSP <= KS, "more k-steps than A fragments");
        typedef FusedTypes<PREC> Ty;
        constexpr int NB = FUSED_NB;
        if constexpr (PREC == FCIS_PREC_BF16X3) {
                typename Ty::Acc a0c = {0, 0, 0, 0}, a1c = a0c, a2c = a0c;
#pragma unroll
                for (int ks = 0; ks < KSP; ks++) {
                        bf16x8 bh, bl;
#pragma unroll
                        for (int e = 0; e < 8; e++) {
                                float xv = bb[(32 * ks + 8 * lk + e) * NB + li];
                                __bf16 hi = (__bf16)xv;
                                bh[e] = hi;
                                bl[e] = (__bf16)(xv - (float)hi);
                        }
                        a0c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[ks], bh, a0c, 0, 0, 0);
                        a1c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[ks], bl, a1c, 0, 0, 0);
                        a2c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[ks], bh, a2c, 0, 0, 0);
                }
                return (a0c + a1c) + a2c;
        } else {
                typename Ty::Acc ac[FUSED_NACC];
#pragma unroll
                for (int z = 0; z < FUSED_NACC; z++) ac[z] = typename Ty::Acc{0, 0, 0, 0};
#pragma unroll
                for (int ks = 0; ks < KSP; ks++) {
                        typename Ty::B bv = bb[(4 * ks + lk) * NB + li];
                        if constexpr (PREC == FCIS_PREC_FP64)
                                ac[ks % FUSED_NACC] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                                        ah[ks], bv, ac[ks % FUSED_NACC], 0, 0, 0);
                        else
                                ac[ks % FUSED_NACC] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                                        ah[ks], bv, ac[ks % FUSED_NACC], 0, 0, 0);
                }
                return ac[0] + ac[1];
        }
}

// One 16 x 16 bf16x3 block whose A fragments come from LDS (af: KS high
// parts then KS low parts, [k-step][lane]) against the pre-split B tile, the
// same three products and accumulation order as mma_presplit.
template <int KS>
__device__ __forceinline__ v4f mma_presplit_lds(const bf16x8 *__restrict__ af, const __bf16 *bh, int lane,
                                                int lk, int li)
{
        constexpr int KP = FUSED_KPS_BF16;
        const __bf16 *bl = bh + FUSED_NB * KP;
        v4f a0c = {0, 0, 0, 0}, a1c = a0c, a2c = a0c;
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
                const bf16x8 ahk = af[ks * 64 + lane], alk = af[(KS + ks) * 64 + lane];
                const bf16x8 h = *reinterpret_cast<const bf16x8 *>(bh + li * KP + 32 * ks + 8 * lk);
                const bf16x8 l = *reinterpret_cast<const bf16x8 *>(bl + li * KP + 32 * ks + 8 * lk);
                a0c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alk, h, a0c, 0, 0, 0);
                a1c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahk, l, a1c, 0, 0, 0);
                a2c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahk, h, a2c, 0, 0, 0);
        }
        return (a0c + a1c) + a2c;
}

// mma_steps for the bf16x3 fused kernel, whose B tile the builders stored
// pre-split and column-major (fb_commit_split): same three products, same
// accumulation order, no conversions on the MFMA wavefronts.
template <int KS>
__device__ __forceinline__ v4f mma_presplit(const bf16x8 (&ah)[KS], const bf16x8 (&al)[KS],
                                            const __bf16 *bh, int lk, int li)
{
        constexpr int KP = FUSED_KPS_BF16;
        const __bf16 *bl = bh + FUSED_NB * KP;
        v4f a0c = {0, 0, 0, 0}, a1c = a0c, a2c = a0c;
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
                const bf16x8 h = *reinterpret_cast<const bf16x8 *>(bh + li * KP + 32 * ks + 8 * lk);
                const bf16x8 l = *reinterpret_cast<const bf16x8 *>(bl + li * KP + 32 * ks + 8 * lk);
                a0c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[ks], h, a0c, 0, 0, 0);
                a1c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[ks], l, a1c, 0, 0, 0);
                a2c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[ks], h, a2c, 0, 0, 0);
        }
        return (a0c + a1c) + a2c;
}

}  // namespace

// State behind fcis_sigma_plan_t (sigma.hip); the batched kernels of
// batch.hip reuse its h fragments and pair maps.
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
        int fused;                   // fused row kernel + alpha pull
        int fused_grid;              // persistent workgroups of the fused kernel
        size_t fused_lds;            // its dynamic LDS bytes
        int16_t *d_pmapb16;          // beta pair map as int16, rows padded to nbp (fused path)
        int64_t nbp;
        int nbld;                    // builder wavefronts of the fused kernel
        // fused kernel's h: rows = the off-diagonal pairs (mpair), then row
        // mrow_f for the alpha occupations' diagonal term, zero padding
        int mtiles_f, mrow_f;
        void *d_hfrag_f, *d_hfrag_f_lo;
        void *d_dtile;               // bf16x3: A fragments of the diagonal pairs' rows (hi, lo)
        int16_t *d_mpair;            // [16 * FUSED_MTILES] pair of each fused row
        double *d_falpha;            // [nstr_a][npair]: sum_{p in I} h[pp][k]
        void *d_fbeta;               // [npair][nbp]:    sum_{p in b} h[pp][k] (double, float for
                                     // the FP32 / bf16x3 kernels)
        float *d_cg;                 // float copy of c gathered by the FP32 / bf16x3 fused kernels
        int32_t *d_order;            // alpha-string processing order (single-tile fused plans) or null
        int timing;
        hipEvent_t ev[4];
        double ms[3];
        int64_t launches[3];
};

// Restricted rows of the fused kernel (restricted determinant spaces): the
// alpha strings to process, in processing order, each with its list of
// FUSED_NB-wide beta chunks, and the determinants whose sigma is wanted.
// Exact when every determinant reachable by one excitation from a wanted
// determinant or from the support of c lies in a listed chunk of a listed row
// (t1 and vt1 vanish elsewhere).
struct fcis_rowset {
        int64_t nrows;
        const int32_t *d_rows;     // [nrows] alpha string of each position
        const int2 *d_rowcl;       // [nrows] (offset into d_clist, chunk count >= 1)
        const uint16_t *d_clist;   // chunk indices, ascending per row; allocated to whole 32-bit words
        const int32_t *d_rpos;     // [nstra] position of each alpha string, -1 if absent
        int64_t ndets;
        const uint32_t *d_dets;    // [ndets] wanted determinants (I * nstrb + Ib)
};

namespace fcis {
// fcis_sigma_plan_create; honour_env = false ignores FCIS_DETERMINISTIC
// (the batch always needs the fused tables).
int sigma_plan_create(fcis_sigma_plan **out, const fcis_space *space, const double *h2e,
                      int precision, uint64_t workspace_bytes, bool honour_env);
// sigma on the wanted determinants of a row set (fused plans only):
// out[d] = (H ci0)[dets[d]]; ci0 is a full (nstra, nstrb) array, sigma_ws a
// full-size work array.
int sigma_plan_execute_rows(fcis_sigma_plan *p, const fcis_rowset *rs, const double *ci0,
                            double *sigma_ws, double *out, hipStream_t stream);
}  // namespace fcis
