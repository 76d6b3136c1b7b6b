// Spin-resolved transition 1- and 2-RDMs on the device (the adjoint of the
// sigma product; the backward of E = c^T H c).
//
// pyscf conventions (direct_spin1.trans_rdm12s; pyscf/lib/mcscf/fci_gas.h:
// 255-268): dm1[p,q] = <bra|q^+ p|ket>, dm2[p,q,r,s] = <bra|p^+ r^+ s q|ket>.
// With T^s_x[d][(r,s)] = (E^s_rs x)[d] and an extra column T^s_x[d][n^2] =
// x[d] (the identity operator):
//   G^{st}[u][v] = sum_d T^s_bra[d][u] T^t_ket[d][v]       (MFMA, long K)
//   dm2^{st}[p,q,r,s] = G^{st}[(q,p)][(r,s)] - delta_{st} delta_qr dm1^s[s,p]
//   dm1^s[p,q] = G^{ss}[n^2][(q,p)]
// because <bra|E_pq E_rs|ket> = <E_qp bra|E_rs ket>.  The reference computes
// the same quantities path by path (fci_rdm_gas.c:318-604, :1334-1465).
#include <stdlib.h>
#include <string.h>
#include <mutex>
#include <vector>

#include "common.h"

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

// T[d][w] for one alpha tile [a0, a0+ta), w in [0, wp): zero, then fill.
// One thread per (determinant, link slot); the slot loop runs over alpha
// links then beta links, and slot == nlinka + nlinkb writes the identity.
// link (cre a, des i, J, s) of string I gives (E_{ia} x)[I] += s x[J].
__global__ __launch_bounds__(256) void k_rdm_build_t(int norb, int64_t nb, int64_t nstra,
                                                     int64_t a0, int64_t ndet, int wp,
                                                     int nlinka, int nlinkb,
                                                     const fcis_link *__restrict__ linka,
                                                     const fcis_link *__restrict__ linkb,
                                                     const double *__restrict__ x,
                                                     double *__restrict__ ta_out,
                                                     double *__restrict__ tb_out)
{
        int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (d >= ndet) return;
        int64_t I = a0 + d / nb, b = d % nb;
        double *ra = ta_out + d * wp;
        double *rb = tb_out + d * wp;
        for (int w = 0; w < wp; w++) {
                ra[w] = 0.0;
                rb[w] = 0.0;
        }
        for (int k = 0; k < nlinka; k++) {
                fcis_link L = linka[(int64_t)k * nstra + I];
                double v = x[(int64_t)L.addr * nb + b];
                ra[link_des(L.info) * norb + link_cre(L.info)] = link_neg(L.info) ? -v : v;
        }
        const double *row = x + I * nb;
        for (int k = 0; k < nlinkb; k++) {
                fcis_link L = linkb[(int64_t)k * nb + b];
                double v = row[L.addr];
                rb[link_des(L.info) * norb + link_cre(L.info)] = link_neg(L.info) ? -v : v;
        }
        ra[norb * norb] = x[I * nb + b];
        rb[norb * norb] = x[I * nb + b];
}

constexpr int RDM_KCHUNK = 4096;  // determinants per workgroup

// part[chunk][u][v] = sum_{d in chunk} A[d][u] B[d][v] for one 16x16 (u,v)
// block per wavefront, v_mfma_f64_16x16x4_f64 with the long determinant
// index as the MFMA k dimension: A-fragment lane (i=u, k) reads A[d0+k][u0+i].
__global__ __launch_bounds__(256) void k_rdm_gemm(int64_t ndet, int wp, int nblk,
                                                  const double *__restrict__ A,
                                                  const double *__restrict__ B,
                                                  double *__restrict__ part)
{
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int li = lane & 15, lk = lane >> 4;
        const int blk = blockIdx.y * 4 + wave;  // (u-block, v-block) pair index
        if (blk >= nblk * nblk) return;
        const int ub = blk / nblk, vb = blk % nblk;
        const int64_t d0 = (int64_t)blockIdx.x * RDM_KCHUNK;
        const int64_t d1 = d0 + RDM_KCHUNK < ndet ? d0 + RDM_KCHUNK : ndet;
        v4d acc = {0, 0, 0, 0};
        for (int64_t d = d0; d < d1; d += 4) {
                int64_t dd = d + lk;
                double a = 0.0, bv = 0.0;
                if (dd < d1) {
                        a = A[dd * wp + ub * 16 + li];
                        bv = B[dd * wp + vb * 16 + li];
                }
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, acc, 0, 0, 0);
        }
        double *out = part + (size_t)blockIdx.x * wp * wp;
        for (int r = 0; r < 4; r++) {
                int u = ub * 16 + lk + 4 * r, v = vb * 16 + li;  // f64 MFMA D layout
                out[(size_t)u * wp + v] = acc[r];
        }
}

// out[e] (+)= sum_c part[c][e], chunks summed in order (deterministic).
__global__ void k_rdm_reduce(int64_t nelem, int nchunk, int accumulate,
                             const double *__restrict__ part, double *__restrict__ out)
{
        int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (e >= nelem) return;
        double s = accumulate ? out[e] : 0.0;
        for (int c = 0; c < nchunk; c++) s += part[(size_t)c * nelem + e];
        out[e] = s;
}


// Buffers of one RDM evaluation: T (four [nd][wp] arrays for the bra and
// ket, alpha and beta), part (per-chunk partial products) and G (the three
// [wp][wp] Gram blocks aa, ab, bb), for alpha tiles of ta strings.
struct RdmBuffers {
        double *T, *part, *G;
        int64_t ta;
};

struct RdmShape {
        int n, wp, nblk;
        int64_t nb, na;
        size_t W2;
};

RdmShape rdm_shape(const fcis_space *s)
{
        RdmShape r;
        r.n = s->norb;
        r.wp = (r.n * r.n + 1 + 15) / 16 * 16;
        r.nblk = r.wp / 16;
        r.nb = s->nstrb;
        r.na = s->nstra;
        r.W2 = (size_t)r.wp * r.wp;
        return r;
}

// Bytes of T and part for a tile of ta alpha strings.
size_t rdm_tile_bytes(const RdmShape &r, int64_t ta)
{
        const int64_t nd = ta * r.nb;
        const int64_t nchunk = (nd + RDM_KCHUNK - 1) / RDM_KCHUNK;
        return 4 * (size_t)nd * r.wp * 8 + (size_t)nchunk * r.W2 * 8;
}

// G = the three Gram blocks of bra and ket over all alpha tiles (launches
// only; stream-ordered).
int rdm_gram(const fcis_space *s, const RdmShape &r, const RdmBuffers &b, const double *d_bra,
             const double *d_ket, hipStream_t st)
{
        const int64_t ndt = b.ta * r.nb;
        double *Tab = b.T, *Tbb = b.T + (size_t)ndt * r.wp;                          // bra alpha / beta
        double *Tak = b.T + 2 * (size_t)ndt * r.wp, *Tbk = b.T + 3 * (size_t)ndt * r.wp;  // ket
        for (int64_t a0 = 0; a0 < r.na; a0 += b.ta) {
                int64_t t = r.na - a0 < b.ta ? r.na - a0 : b.ta;
                int64_t nd = t * r.nb;
                int64_t nchunk = (nd + RDM_KCHUNK - 1) / RDM_KCHUNK;
                hipLaunchKernelGGL(k_rdm_build_t, dim3((nd + 255) / 256), dim3(256), 0, st, r.n, r.nb,
                                   r.na, a0, nd, r.wp, s->nlinka, s->nlinkb, s->d_linka, s->d_linkb,
                                   d_bra, Tab, Tbb);
                hipLaunchKernelGGL(k_rdm_build_t, dim3((nd + 255) / 256), dim3(256), 0, st, r.n, r.nb,
                                   r.na, a0, nd, r.wp, s->nlinka, s->nlinkb, s->d_linka, s->d_linkb,
                                   d_ket, Tak, Tbk);
                const double *As[3] = {Tab, Tab, Tbb};
                const double *Bs[3] = {Tak, Tbk, Tbk};
                for (int g = 0; g < 3; g++) {
                        hipLaunchKernelGGL(k_rdm_gemm, dim3(nchunk, (r.nblk * r.nblk + 3) / 4),
                                           dim3(256), 0, st, nd, r.wp, r.nblk, As[g], Bs[g], b.part);
                        hipLaunchKernelGGL(k_rdm_reduce, dim3((r.W2 + 255) / 256), dim3(256), 0, st,
                                           (int64_t)r.W2, (int)nchunk, a0 > 0 ? 1 : 0, b.part,
                                           b.G + g * r.W2);
                }
                if (hipGetLastError() != hipSuccess) {
                        fcis::set_error("trans_rdm12s: kernel launch failed");
                        return FCIS_ERR_DEVICE;
                }
        }
        return FCIS_SUCCESS;
}

// The RDMs in their final layout from G, on the device:
//   out = [dm1a n^2][dm1b n^2][dm2aa n^4][dm2ab n^4][dm2bb n^4]
//   dm1^s[p][q] = G^ss[n^2][(q,p)],
//   dm2^st[p][q][r][t] = G^st[(q,p)][(r,t)] - delta_st delta_qr dm1^s[t][p].
__global__ void k_rdm_assemble(int n, int wp, const double *__restrict__ G, double *__restrict__ out)
{
        const int n2 = n * n;
        const int64_t n4 = (int64_t)n2 * n2, W2 = (int64_t)wp * wp;
        const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (e >= 2 * n2 + 3 * n4) return;
        if (e < 2 * n2) {
                const int sp = (int)(e / n2), idx = (int)(e - (int64_t)sp * n2);
                const int p = idx / n, q = idx - p * n;
                out[e] = G[(sp ? 2 : 0) * W2 + (int64_t)n2 * wp + q * n + p];
                return;
        }
        const int64_t f = e - 2 * n2;
        const int g = (int)(f / n4);
        const int64_t idx = f - g * n4;
        const int p = (int)(idx / ((int64_t)n * n2)), q = (int)(idx / n2 % n), r = (int)(idx / n % n),
                  t = (int)(idx % n);
        double v = G[g * W2 + (int64_t)(q * n + p) * wp + r * n + t];
        if (g != 1 && q == r) v -= G[g * W2 + (int64_t)n2 * wp + p * n + t];
        out[e] = v;
}

// Scratch of the non-allocating RDM entry point, one copy per device:
// G, the per-chunk partial products and the T arrays of an alpha tile, then
// the assembled RDMs.  The library holds it on every GPU it runs on until
// the process ends.
constexpr size_t RDM_SCRATCH_BYTES = (size_t)128 << 20;
static __device__ __align__(256) char scratch[RDM_SCRATCH_BYTES];

// Device address of `scratch` per device, from hipGetSymbolAddress on the
// first call on that device (with that device current).
constexpr int RDM_MAXDEV = 64;
void *g_scratch_ptr[RDM_MAXDEV];
std::mutex g_scratch_mu;

int scratch_ptr(char **out)
{
        int dev = 0;
        FCIS_HIP(hipGetDevice(&dev));
        if (dev < 0 || dev >= RDM_MAXDEV) {
                fcis::set_error("fcis_trans_rdm12s_noalloc: device %d out of range", dev);
                return FCIS_ERR_INVALID;
        }
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (g_scratch_ptr[dev] == nullptr) {
                void *p = nullptr;
                hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
                if (e != hipSuccess || p == nullptr) {
                        fcis::set_error("fcis_trans_rdm12s_noalloc: hipGetSymbolAddress: %s",
                                        hipGetErrorString(e));
                        return FCIS_ERR_DEVICE;
                }
                g_scratch_ptr[dev] = p;
        }
        *out = static_cast<char *>(g_scratch_ptr[dev]);
        return FCIS_SUCCESS;
}

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

}  // namespace

extern "C" int fcis_trans_rdm12s(const fcis_space_t *s, const double *d_bra, const double *d_ket,
                                 double *dm1a, double *dm1b, double *dm2aa, double *dm2ab,
                                 double *dm2bb, void *stream)
{
        if (s == nullptr || d_bra == nullptr || d_ket == nullptr) {
                fcis::set_error("fcis_trans_rdm12s: null argument");
                return FCIS_ERR_INVALID;
        }
        hipStream_t st = (hipStream_t)stream;
        const RdmShape r = rdm_shape(s);
        const int n = r.n, n2 = n * n, wp = r.wp;
        const size_t W2 = r.W2;
        // alpha tile sized so that four T arrays stay within ~2 GiB
        int64_t per_alpha = 4 * r.nb * wp * 8;
        int64_t ta = per_alpha ? ((int64_t)2 << 30) / per_alpha : r.na;
        if (ta < 1) ta = 1;
        if (ta > r.na) ta = r.na;
        int64_t ndt = ta * r.nb;
        int64_t maxchunk = (ndt + RDM_KCHUNK - 1) / RDM_KCHUNK;
        double *T = nullptr, *part = nullptr, *G = nullptr;
        size_t tbytes = 4 * (size_t)ndt * wp * 8;
        if (fcis::dev_malloc(&T, tbytes ? tbytes : 8) != hipSuccess ||
            fcis::dev_malloc(&part, maxchunk * W2 * 8 + 8) != hipSuccess ||
            fcis::dev_malloc(&G, 3 * W2 * 8) != hipSuccess) {
                (void)hipFree(T);
                (void)hipFree(part);
                fcis::set_error("fcis_trans_rdm12s: device allocation failed");
                return FCIS_ERR_MEMORY;
        }
        int ret = rdm_gram(s, r, RdmBuffers{T, part, G, ta}, d_bra, d_ket, st);
        std::vector<double> h(3 * W2);
        if (ret == FCIS_SUCCESS &&
            (hipMemcpyAsync(h.data(), G, 3 * W2 * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
             hipStreamSynchronize(st) != hipSuccess)) {
                fcis::set_error("fcis_trans_rdm12s: copy failed");
                ret = FCIS_ERR_DEVICE;
        }
        (void)hipFree(T);
        (void)hipFree(part);
        (void)hipFree(G);
        if (ret != FCIS_SUCCESS) return ret;
        const double *Gaa = h.data(), *Gab = Gaa + W2, *Gbb = Gab + W2;
        std::vector<double> d1a(n2), d1b(n2);
        for (int p = 0; p < n; p++)
                for (int q = 0; q < n; q++) {
                        d1a[p * n + q] = Gaa[(size_t)n2 * wp + q * n + p];
                        d1b[p * n + q] = Gbb[(size_t)n2 * wp + q * n + p];
                }
        if (dm1a) memcpy(dm1a, d1a.data(), 8 * n2);
        if (dm1b) memcpy(dm1b, d1b.data(), 8 * n2);
        double *outs[3] = {dm2aa, dm2ab, dm2bb};
        const double *Gs[3] = {Gaa, Gab, Gbb};
        for (int g = 0; g < 3; g++) {
                double *o = outs[g];
                if (o == nullptr) continue;
                for (int p = 0; p < n; p++)
                        for (int q = 0; q < n; q++)
                                for (int r2 = 0; r2 < n; r2++)
                                        for (int t = 0; t < n; t++) {
                                                double v = Gs[g][(size_t)(q * n + p) * wp + r2 * n + t];
                                                if (g != 1 && q == r2)
                                                        v -= (g == 0 ? d1a : d1b)[t * n + p];
                                                o[(((size_t)p * n + q) * n + r2) * n + t] = v;
                                        }
        }
        return FCIS_SUCCESS;
}

extern "C" int fcis_trans_rdm12s_noalloc(const fcis_space_t *s, const double *d_bra,
                                         const double *d_ket, double *dm1a, double *dm1b,
                                         double *dm2aa, double *dm2ab, double *dm2bb, void *stream)
{
        if (s == nullptr || d_bra == nullptr || d_ket == nullptr) {
                fcis::set_error("fcis_trans_rdm12s_noalloc: null argument");
                return FCIS_ERR_INVALID;
        }
        hipStream_t st = (hipStream_t)stream;
        const RdmShape r = rdm_shape(s);
        const int n2 = r.n * r.n;
        const int64_t n4 = (int64_t)n2 * n2;
        // scratch layout: G | assembled RDMs | part | T
        const size_t gbytes = align256(3 * r.W2 * 8);
        const size_t obytes = align256((size_t)(2 * n2 + 3 * n4) * 8);
        const size_t fixed = gbytes + obytes;
        int64_t ta = r.na;
        while (ta > 1 && fixed + rdm_tile_bytes(r, ta) + 512 > RDM_SCRATCH_BYTES) ta = (ta + 1) / 2;
        if (r.na > 0 && fixed + rdm_tile_bytes(r, ta) + 512 > RDM_SCRATCH_BYTES) {
                fcis::set_error("fcis_trans_rdm12s_noalloc: one alpha string of this space needs %zu "
                                "bytes of scratch, above the fixed %zu (norb %d, %lld beta strings); "
                                "use fcis_trans_rdm12s",
                                fixed + rdm_tile_bytes(r, 1), RDM_SCRATCH_BYTES, r.n, (long long)r.nb);
                return FCIS_ERR_INVALID;
        }
        char *base = nullptr;
        int ret = scratch_ptr(&base);
        if (ret != FCIS_SUCCESS) return ret;
        double *G = reinterpret_cast<double *>(base);
        double *rd = reinterpret_cast<double *>(base + gbytes);
        const int64_t ndt = ta * r.nb;
        const int64_t nchunk = (ndt + RDM_KCHUNK - 1) / RDM_KCHUNK;
        double *part = reinterpret_cast<double *>(base + fixed);
        double *T = reinterpret_cast<double *>(base + fixed + align256((size_t)nchunk * r.W2 * 8));
        if (r.na > 0) {
                ret = rdm_gram(s, r, RdmBuffers{T, part, G, ta}, d_bra, d_ket, st);
                if (ret != FCIS_SUCCESS) return ret;
        } else {
                hipLaunchKernelGGL(k_rdm_reduce, dim3((3 * r.W2 + 255) / 256), dim3(256), 0, st,
                                   (int64_t)(3 * r.W2), 0, 0, G, G);
        }
        const int64_t ntot = 2 * n2 + 3 * n4;
        hipLaunchKernelGGL(k_rdm_assemble, dim3((ntot + 255) / 256), dim3(256), 0, st, r.n, r.wp, G, rd);
        FCIS_CHECK_LAUNCH();
        double *outs[5] = {dm1a, dm1b, dm2aa, dm2ab, dm2bb};
        for (int k = 0; k < 5; k++) {
                if (outs[k] == nullptr) continue;
                const size_t off = k < 2 ? (size_t)k * n2 : 2 * (size_t)n2 + (size_t)(k - 2) * n4;
                const size_t cnt = k < 2 ? (size_t)n2 : (size_t)n4;
                FCIS_HIP(hipMemcpyAsync(outs[k], rd + off, cnt * 8, hipMemcpyDeviceToHost, st));
        }
        FCIS_HIP(hipStreamSynchronize(st));
        return FCIS_SUCCESS;
}
