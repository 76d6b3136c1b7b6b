// Hamiltonian block over selected determinants (pspace), by the
// Slater-Condon rules from the un-absorbed integrals.
//
// The reference builds small-space Hamiltonians one column at a time with a
// full sigma per column (fci_gas.py:1130-1148, capped at 64 columns,
// :36-39) or, for CSF pspaces, with pyscf's determinant pspace builder
// (csf.py:246-366).  Here one thread computes H[i][j] = <D_i|H|D_j> for a
// pair of determinants D = (alpha string, beta string):
//   no difference   sum_{p in a,b} h_pp + 1/2 sum_{p,q same spin} [(pp|qq) - (pq|qp)]
//                   + sum_{p in a, q in b} (pp|qq)
//   one p <- q      s [h_pq + sum_{k same spin, both} ((pq|kk) - (pk|kq))
//                      + sum_{k other spin} (pq|kk)]
//   two, same spin  s [(p1 q1|p2 q2) - (p1 q2|p2 q1)]
//   two, a and b    s_a s_b (pa qa|pb qb)
//   more than two   0
// with the signs of the excitation operators a+_p a_q acting on the ket's
// strings (for two same-spin moves: the first on the ket, the second on its
// result).  h1e and the (npair, npair) pair matrix are read from global
// memory (at most 150 KB at 16 orbitals, L2-resident).  The one-electron
// part may differ between the spins (h_a for alpha moves and occupations,
// h_b for beta; the UHF pspace of csf.py:305-320); fcis_pspace is h_a = h_b.
#include <vector>

#include "gas.h"

namespace {

__device__ __forceinline__ double pe(const double *__restrict__ eri, int np, int p, int q, int r,
                                     int s)
{
        return eri[(size_t)fcis::pair_index(p, q) * np + fcis::pair_index(r, s)];
}

// sign of a+_p a_q on s (q occupied, p empty or p == q)
__device__ __forceinline__ double exc_sign(uint64_t s, int p, int q)
{
        const int lo = p < q ? p : q, hi = p < q ? q : p;
        const uint64_t m = hi > lo + 1 ? ((1ull << hi) - (1ull << (lo + 1))) : 0ull;
        return (__popcll(s & m) & 1) ? -1.0 : 1.0;
}

__device__ double same_spin_diag(uint64_t s, const double *__restrict__ h1e,
                                 const double *__restrict__ eri, int n, int np)
{
        double v = 0.0;
        for (uint64_t a = s; a; a &= a - 1) {
                const int p = __ffsll((long long)a) - 1;
                v += h1e[p * n + p];
                for (uint64_t b = s & ((1ull << p) - 1); b; b &= b - 1) {
                        const int q = __ffsll((long long)b) - 1;
                        v += pe(eri, np, p, p, q, q) - pe(eri, np, p, q, q, p);
                }
        }
        return v;
}

// <I|H|J> for a single move p <- q within one spin; `same` are the other
// electrons of that spin (common to I and J), `other` the other spin's string.
__device__ double single_elem(int p, int q, uint64_t same, uint64_t other,
                              const double *__restrict__ h1e, const double *__restrict__ eri,
                              int n, int np)
{
        double v = h1e[p * n + q];
        for (uint64_t a = same; a; a &= a - 1) {
                const int k = __ffsll((long long)a) - 1;
                v += pe(eri, np, p, q, k, k) - pe(eri, np, p, k, k, q);
        }
        for (uint64_t a = other; a; a &= a - 1) {
                const int k = __ffsll((long long)a) - 1;
                v += pe(eri, np, p, q, k, k);
        }
        return v;
}

__global__ __launch_bounds__(256) void k_pspace(int n, int np, int m, const int64_t *__restrict__ addr,
                                                int64_t nstrb, const uint64_t *__restrict__ stra,
                                                const uint64_t *__restrict__ strb,
                                                const double *__restrict__ h1a,
                                                const double *__restrict__ h1b,
                                                const double *__restrict__ eri, double *__restrict__ h0)
{
        const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (t >= (int64_t)m * m) return;
        const int i = (int)(t / m), j = (int)(t - (int64_t)i * m);
        if (j > i) return;
        const int64_t ai = addr[i] / nstrb, bi = addr[i] - ai * nstrb;
        const int64_t aj = addr[j] / nstrb, bj = addr[j] - aj * nstrb;
        const uint64_t Ia = stra[ai], Ib = strb[bi], Ja = stra[aj], Jb = strb[bj];
        const uint64_t da = Ia ^ Ja, db = Ib ^ Jb;
        const int na = __popcll(da) >> 1, nb = __popcll(db) >> 1;
        double v = 0.0;
        if (na + nb == 0) {
                v = same_spin_diag(Ia, h1a, eri, n, np) + same_spin_diag(Ib, h1b, eri, n, np);
                for (uint64_t a = Ia; a; a &= a - 1) {
                        const int p = __ffsll((long long)a) - 1;
                        for (uint64_t b = Ib; b; b &= b - 1) {
                                const int q = __ffsll((long long)b) - 1;
                                v += pe(eri, np, p, p, q, q);
                        }
                }
        } else if (na + nb == 1) {
                const bool alpha = na == 1;
                const uint64_t I = alpha ? Ia : Ib, J = alpha ? Ja : Jb, d = alpha ? da : db;
                const int p = __ffsll((long long)(d & I)) - 1;   // in I only
                const int q = __ffsll((long long)(d & J)) - 1;   // in J only
                v = exc_sign(J, p, q) *
                    single_elem(p, q, I & J, alpha ? Ib : Ia, alpha ? h1a : h1b, eri, n, np);
        } else if (na + nb == 2 && (na == 2 || nb == 2)) {
                const bool alpha = na == 2;
                const uint64_t I = alpha ? Ia : Ib, J = alpha ? Ja : Jb, d = alpha ? da : db;
                uint64_t pI = d & I, qJ = d & J;
                const int p1 = __ffsll((long long)pI) - 1;
                const int p2 = __ffsll((long long)(pI & (pI - 1))) - 1;
                const int q1 = __ffsll((long long)qJ) - 1;
                const int q2 = __ffsll((long long)(qJ & (qJ - 1))) - 1;
                const uint64_t K = (J & ~(1ull << q1)) | (1ull << p1);
                const double s = exc_sign(J, p1, q1) * exc_sign(K, p2, q2);
                v = s * (pe(eri, np, p1, q1, p2, q2) - pe(eri, np, p1, q2, p2, q1));
        } else if (na + nb == 2) {
                const int pa = __ffsll((long long)(da & Ia)) - 1, qa = __ffsll((long long)(da & Ja)) - 1;
                const int pb = __ffsll((long long)(db & Ib)) - 1, qb = __ffsll((long long)(db & Jb)) - 1;
                v = exc_sign(Ja, pa, qa) * exc_sign(Jb, pb, qb) * pe(eri, np, pa, qa, pb, qb);
        }
        h0[(int64_t)i * m + j] = v;
        h0[(int64_t)j * m + i] = v;
}

}  // namespace

int fcis::pspace_strings(int n, int m, const uint64_t *d_stra, const uint64_t *d_strb,
                         const double *h1e, const double *eri, double *d_h0, hipStream_t st)
{
        // determinant i is (stra[i], strb[i]): address i * m + i over an
        // "alpha" list and a "beta" list of m strings each
        const int np = n * (n + 1) / 2;
        const size_t nh = (size_t)n * n, ne = (size_t)np * np;
        std::vector<int64_t> addr(m);
        for (int i = 0; i < m; i++) addr[i] = (int64_t)i * m + i;
        double *d_h = nullptr;
        FCIS_HIP(fcis::dev_malloc(&d_h, 8 * (nh + ne) + 8 * (size_t)m));
        int64_t *d_addr = reinterpret_cast<int64_t *>(d_h + nh + ne);
        int ret = FCIS_SUCCESS;
        if (hipMemcpyAsync(d_h, h1e, 8 * nh, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(d_h + nh, eri, 8 * ne, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(d_addr, addr.data(), 8 * (size_t)m, hipMemcpyHostToDevice, st) != hipSuccess) {
                fcis::set_error("pspace: upload failed");
                ret = FCIS_ERR_DEVICE;
        } else {
                const int64_t nt = (int64_t)m * m;
                hipLaunchKernelGGL(k_pspace, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, st, n, np, m,
                                   d_addr, (int64_t)m, d_stra, d_strb, d_h, d_h, d_h + nh, d_h0);
                hipError_t e = hipGetLastError();
                if (e == hipSuccess) e = hipStreamSynchronize(st);
                if (e != hipSuccess) {
                        fcis::set_error("pspace: %s", hipGetErrorString(e));
                        ret = FCIS_ERR_DEVICE;
                }
        }
        (void)hipFree(d_h);
        return ret;
}

extern "C" {

int fcis_pspace_uhf(const fcis_space_t *s, const double *h1a, const double *h1b, const double *eri,
                    int m, const int64_t *d_addr, double *d_h0, void *stream)
{
        if (s == nullptr || h1a == nullptr || h1b == nullptr || eri == nullptr || m < 0 ||
            (m > 0 && (d_addr == nullptr || d_h0 == nullptr))) {
                fcis::set_error("fcis_pspace: invalid argument");
                return FCIS_ERR_INVALID;
        }
        if ((int64_t)m * m > ((int64_t)1 << 31)) {
                fcis::set_error("fcis_pspace: %d determinants exceed the 46340 limit", m);
                return FCIS_ERR_INVALID;
        }
        if (m == 0) return FCIS_SUCCESS;
        hipStream_t st = (hipStream_t)stream;
        const int n = s->norb, np = s->npair;
        double *d_h = nullptr;
        const size_t nh = (size_t)n * n, ne = (size_t)np * np;
        FCIS_HIP(fcis::dev_malloc(&d_h, 8 * (2 * nh + ne)));
        int ret = FCIS_SUCCESS;
        if (hipMemcpyAsync(d_h, h1a, 8 * nh, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(d_h + nh, h1b, 8 * nh, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(d_h + 2 * nh, eri, 8 * ne, hipMemcpyHostToDevice, st) != hipSuccess) {
                fcis::set_error("fcis_pspace: upload failed");
                ret = FCIS_ERR_DEVICE;
        } else {
                const int64_t nt = (int64_t)m * m;
                hipLaunchKernelGGL(k_pspace, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, st, n,
                                   np, m, d_addr, s->nstrb, s->d_stra, s->d_strb, d_h, d_h + nh,
                                   d_h + 2 * nh, d_h0);
                hipError_t e = hipGetLastError();
                if (e == hipSuccess) e = hipStreamSynchronize(st);
                if (e != hipSuccess) {
                        fcis::set_error("fcis_pspace: %s", hipGetErrorString(e));
                        ret = FCIS_ERR_DEVICE;
                }
        }
        (void)hipFree(d_h);
        return ret;
}

int fcis_pspace(const fcis_space_t *s, const double *h1e, const double *eri, int m,
                const int64_t *d_addr, double *d_h0, void *stream)
{
        return fcis_pspace_uhf(s, h1e, h1e, eri, m, d_addr, d_h0, stream);
}

}  // extern "C"
