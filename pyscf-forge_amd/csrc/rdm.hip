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
        const int n = s->norb, n2 = n * n;
        const int wp = (n2 + 1 + 15) / 16 * 16;
        const int nblk = wp / 16;
        const int64_t nb = s->nstrb, na = s->nstra;
        const size_t W2 = (size_t)wp * wp;
        // alpha tile sized so that four T arrays stay within ~2 GiB
        int64_t per_alpha = 4 * nb * wp * 8;
        int64_t ta = per_alpha ? ((int64_t)2 << 30) / per_alpha : na;
        if (ta < 1) ta = 1;
        if (ta > na) ta = na;
        int64_t ndt = ta * nb;
        int64_t maxchunk = (ndt + RDM_KCHUNK - 1) / RDM_KCHUNK;
        double *T = nullptr, *part = nullptr, *G = nullptr;
        size_t tbytes = 4 * (size_t)ndt * wp * 8;
        if (hipMalloc(&T, tbytes ? tbytes : 8) != hipSuccess ||
            hipMalloc(&part, maxchunk * W2 * 8 + 8) != hipSuccess ||
            hipMalloc(&G, 3 * W2 * 8) != hipSuccess) {
                (void)hipFree(T);
                (void)hipFree(part);
                fcis::set_error("fcis_trans_rdm12s: device allocation failed");
                return FCIS_ERR_MEMORY;
        }
        double *Tab = T, *Tbb = T + (size_t)ndt * wp;        // bra alpha / beta
        double *Tak = T + 2 * (size_t)ndt * wp, *Tbk = T + 3 * (size_t)ndt * wp;  // ket
        int ret = FCIS_SUCCESS;
        for (int64_t a0 = 0; a0 < na && ret == FCIS_SUCCESS; a0 += ta) {
                int64_t t = na - a0 < ta ? na - a0 : ta;
                int64_t nd = t * nb;
                int64_t nchunk = (nd + RDM_KCHUNK - 1) / RDM_KCHUNK;
                hipLaunchKernelGGL(k_rdm_build_t, dim3((nd + 255) / 256), dim3(256), 0, st, n, nb,
                                   na, a0, nd, wp, s->nlinka, s->nlinkb, s->d_linka, s->d_linkb,
                                   d_bra, Tab, Tbb);
                hipLaunchKernelGGL(k_rdm_build_t, dim3((nd + 255) / 256), dim3(256), 0, st, n, nb,
                                   na, a0, nd, wp, s->nlinka, s->nlinkb, s->d_linka, s->d_linkb,
                                   d_ket, Tak, Tbk);
                const double *As[3] = {Tab, Tab, Tbb};
                const double *Bs[3] = {Tak, Tbk, Tbk};
                for (int g = 0; g < 3; g++) {
                        hipLaunchKernelGGL(k_rdm_gemm, dim3(nchunk, (nblk * nblk + 3) / 4),
                                           dim3(256), 0, st, nd, wp, nblk, As[g], Bs[g], part);
                        hipLaunchKernelGGL(k_rdm_reduce, dim3((W2 + 255) / 256), dim3(256), 0, st,
                                           (int64_t)W2, (int)nchunk, a0 > 0 ? 1 : 0, part,
                                           G + g * W2);
                }
                if (hipGetLastError() != hipSuccess) {
                        fcis::set_error("fcis_trans_rdm12s: kernel launch failed");
                        ret = FCIS_ERR_DEVICE;
                }
        }
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
                                for (int r = 0; r < n; r++)
                                        for (int t = 0; t < n; t++) {
                                                double v = Gs[g][(size_t)(q * n + p) * wp + r * n + t];
                                                if (g != 1 && q == r)
                                                        v -= (g == 0 ? d1a : d1b)[t * n + p];
                                                o[(((size_t)p * n + q) * n + r) * n + t] = v;
                                        }
        }
        return FCIS_SUCCESS;
}
