// Sector-native restricted (GAS) spaces: sigma, diagonal, RDMs and
// H blocks on the GAS determinants without a CAS embedding, so that the
// orbital limit is the reference's (63 in all subspaces, 31 per subspace,
// pyscf/lib/mcscf/fci_gas.h:24-27) and device memory scales with the GAS
// space, not with the CAS space of sum(norb) orbitals.
//
// The restricted Hamiltonian is P H P (the identity the reference's test
// checks libfci_gas against, pyscf/mcscf/test/test_gasci.py:522-575).  With
// the absorbed pair matrix h and the symmetric pair operators
// F_pq = E_pq + E_qp (p != q), F_pp = E_pp, E = E^alpha + E^beta,
//   sigma_I = sum_{pq,rs} h[rs,pq] sum_K <I|F_rs|K> <K|F_pq|c>,
// where K runs over every determinant one excitation (or none) from a GAS
// determinant: the closure S1 = D plus the blocks (sa', sb) and (sa, sb')
// with sa' (sb') one inter-subspace move from a block's alpha (beta) sector
// (the reference's sector pool, fci_string_gas.c:588-791).  The device path
// is the t1 / GEMM / pull scheme of the CAS three-kernel sigma on S1:
//   t1[pq, K]  = sum over the links (pq, J, s) of K's alpha and beta
//                strings with J in D: s c[J]                     (gather)
//   vt1        = h . t1                                          (rocBLAS GEMM)
//   sigma[I]   = sum over the links (pq, K, s) of I's strings: s vt1[pq, K]
//                                                                (pull)
// Every sigma element is produced by one thread (destination-owned, no
// atomics, bitwise reproducible).  Two direct forms avoid the closure where
// npair^2 per closure column would far exceed the excitations (many
// orbitals, few electrons): the operator form (one-spin operator matrices
// for the same-spin parts, sorted link lists for the opposite-spin part;
// described with its kernels below) and the link-pair loops.  S1 is laid out row-major over alpha
// strings: the strings of the alpha sectors of S1 in sector order are the
// rows, a row holds the strings of its sector's S1 beta sectors one after
// the other, and t1 / vt1 are tiled over whole rows to fit the workspace.
//
// String tables per spin: every string of the spin's S1 sectors, in sector
// order, gets a global index g (sector base + address in the sector); the
// pair map pmap[pq][g] is the global index of a_p^+ a_q |g> (or a_q^+ a_p,
// whichever is nonzero; the string itself for p == q in the string) with
// the sign in bit 31, or -1 when the result is zero or leaves the spin's
// S1 sectors.  The link lists hold the same entries compactly, k(n-k+1)
// slots per string of k electrons (diagonal (q,q) first for each occupied
// q, then (p,q) for every empty p): lk[slot][g] as in pmap and lkp[slot][g]
// the pair index, -1 where the result leaves the S1 sectors.  The pull
// walks them (npair may exceed the link count many times over at large
// orbital counts); the gather writes every t1 row and reads the pair map.
// RDMs use the same closure with ordered pairs:
//   D^s_v[d n + c, K] = <K| a_d^+ a_c |v> (spin s),
//   <bra| a_p^+ a_q a_r^+ a_s |ket> = sum_K D_bra[q n + p, K] D_ket[r n + s, K],
//   <bra| a_p^+ a_q |ket>           = sum_{K in D} bra_K D_ket[p n + q, K],
// assembled into dm1[p,q] = <q^+ p> and dm2[p,q,r,s] = <p^+ r^+ s q>
// (fci_gas.h:255-268) with rocBLAS GEMMs over K.
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <map>
#include <new>
#include <set>
#include <vector>

#include <rocblas/rocblas.h>

#include "gas.h"

namespace {

constexpr int NT = 256;
// Elements a rocBLAS GEMM operand may span from its base pointer
// ((columns - 1) * ld + rows): DGEMMs of npair 136 returned wrong results,
// without an error, for ld = 47.4M and 63.2M (spans 6.4e9, 8.5e9) and were
// exact for ld = 31.6M (span 4.29e9), so t1 tiles keep below 2^32.
constexpr uint64_t GEMM_MAX_SPAN = 0xffffffffull;

__constant__ fcis::BinomTable c_binom = fcis::make_binom_table();

struct SubInfo {
        int ngas;
        int norb[GAS_MAXNGAS], start[GAS_MAXNGAS];
        int8_t sub_of[64];              // subspace of each orbital
};

// Device view of the native tables (kernel argument).
struct NatView {
        int n, npair, nsecA, nsecB, nblock;
        int64_t nstrA, nstrB, ndet;
        const uint64_t *strA, *strB;
        const uint16_t *ssecA, *ssecB;
        const uint32_t *saddrA, *saddrB;
        const int32_t *pmapA, *pmapB;
        const int32_t *lkA, *lkB;       // [nlink][nstr] target | sign << 31, or -1
        const int16_t *lkpA, *lkpB;     // [nlink][nstr] pair index
        int nlinkA, nlinkB;
        const int64_t *sbaseA, *sbaseB;
        const int64_t *cbase, *rowlen;
        const int32_t *segptr, *segsec;
        const int64_t *segoff;
        const int32_t *dblk;
        const int64_t *s1bo;
        const int64_t *boff;
        const int32_t *bsa, *bsb;
        const int64_t *bnb;
};

__device__ __forceinline__ int upper_index(const int64_t *__restrict__ a, int n, int64_t x)
{
        int lo = 0, hi = n - 1;
        while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (a[mid] <= x) lo = mid;
                else hi = mid - 1;
        }
        return lo;
}

// Position of S1 column col: row r (global alpha string), its local alpha
// sector A and address a, the local beta sector B of the column and the
// beta address b.
struct S1Pos {
        int64_t r, a, b;
        int A, B;
};

__device__ __forceinline__ S1Pos s1_locate(const NatView &v, int64_t col)
{
        S1Pos p;
        p.A = upper_index(v.cbase, v.nsecA, col);
        int64_t j = col - v.cbase[p.A];
        const int64_t L = v.rowlen[p.A];
        p.a = j / L;
        j -= p.a * L;
        const int s0 = v.segptr[p.A], s1 = v.segptr[p.A + 1];
        const int seg = s0 + upper_index(v.segoff + s0, s1 - s0, j);
        p.B = v.segsec[seg];
        p.b = j - v.segoff[seg];
        p.r = v.sbaseA[p.A] + p.a;
        return p;
}

struct DPos {
        int64_t a, b;
        int A, B, blk;
};

__device__ __forceinline__ DPos d_locate(const NatView &v, int64_t d)
{
        DPos q;
        q.blk = upper_index(v.boff, v.nblock, d);
        const int64_t nb = v.bnb[q.blk], off = d - v.boff[q.blk];
        q.a = off / nb;
        q.b = off - q.a * nb;
        q.A = v.bsa[q.blk];
        q.B = v.bsb[q.blk];
        return q;
}

// Index of GAS determinant (alpha local sector A, address a; beta local
// sector B, address b), or -1 when (A, B) is not a block of D.
__device__ __forceinline__ int64_t d_index(const NatView &v, int A, int64_t a, int B, int64_t b)
{
        const int blk = v.dblk[(int64_t)A * v.nsecB + B];
        return blk < 0 ? -1 : v.boff[blk] + a * v.bnb[blk] + b;
}

__device__ __forceinline__ uint64_t unrank_dev(int n, int k, int64_t addr)
{
        uint64_t s = 0;
        int e = k;
        for (int p = n - 1; p >= 0 && e > 0; p--) {
                const int64_t c = c_binom.v[p][e];
                if (c <= addr) {
                        s |= 1ull << p;
                        addr -= c;
                        e--;
                }
        }
        return s;
}

__device__ __forceinline__ int64_t rank_dev(uint64_t s)
{
        int64_t a = 0;
        int i = 1;
        while (s) {
                const int p = __ffsll((long long)s) - 1;
                a += c_binom.v[p][i++];
                s &= s - 1;
        }
        return a;
}

// Strings of one spin: g -> (string, local sector, address in sector).
__global__ __launch_bounds__(NT) void k_nat_strings(SubInfo si, int nsec, int64_t nstr,
                                                    const int64_t *__restrict__ sbase,
                                                    const uint8_t *__restrict__ occ,
                                                    const int64_t *__restrict__ stride,
                                                    uint64_t *__restrict__ str,
                                                    uint16_t *__restrict__ ssec,
                                                    uint32_t *__restrict__ saddr)
{
        const int64_t g = (int64_t)blockIdx.x * NT + threadIdx.x;
        if (g >= nstr) return;
        const int s = upper_index(sbase, nsec, g);
        int64_t addr = g - sbase[s];
        saddr[g] = (uint32_t)addr;
        ssec[g] = (uint16_t)s;
        uint64_t x = 0;
        for (int k = 0; k < si.ngas; k++) {
                const int64_t st = stride[(int64_t)s * si.ngas + k];
                const int64_t q = addr / st;
                addr -= q * st;
                x |= unrank_dev(si.norb[k], occ[(int64_t)s * si.ngas + k], q) << si.start[k];
        }
        str[g] = x;
}

// Pair map of one spin (see the file comment).  mv[(s * ngas + r) * ngas + t]
// is the local sector of sector s with one electron moved from subspace r
// to subspace t, or -1.
__global__ __launch_bounds__(NT) void k_nat_pmap(SubInfo si, int n, int64_t nstr,
                                                 const uint64_t *__restrict__ str,
                                                 const uint16_t *__restrict__ ssec,
                                                 const int64_t *__restrict__ sbase,
                                                 const int64_t *__restrict__ stride,
                                                 const int32_t *__restrict__ mv,
                                                 int32_t *__restrict__ pmap,
                                                 int32_t *__restrict__ lk,
                                                 int16_t *__restrict__ lkp)
{
        const int64_t g = (int64_t)blockIdx.x * NT + threadIdx.x;
        if (g >= nstr) return;
        const uint64_t S = str[g];
        const int s = ssec[g];
        const int ng = si.ngas;
        int slot = 0;
        for (uint64_t occ = S; occ; occ &= occ - 1) {
                const int q = __ffsll((long long)occ) - 1;
                pmap[(int64_t)fcis::pair_index(q, q) * nstr + g] = (int32_t)g;
                lk[(int64_t)slot * nstr + g] = (int32_t)g;
                lkp[(int64_t)slot * nstr + g] = (int16_t)fcis::pair_index(q, q);
                slot++;
                const int kq = si.sub_of[q];
                for (int p = 0; p < n; p++) {
                        if ((S >> p) & 1) continue;
                        const int kp = si.sub_of[p];
                        const int ts = kq == kp ? s : mv[((int64_t)s * ng + kq) * ng + kp];
                        const int this_slot = slot++;
                        lkp[(int64_t)this_slot * nstr + g] = (int16_t)fcis::pair_index(p, q);
                        if (ts < 0) {
                                lk[(int64_t)this_slot * nstr + g] = -1;
                                continue;
                        }
                        const uint64_t T = (S & ~(1ull << q)) | (1ull << p);
                        int64_t addr = 0;
                        for (int k = 0; k < ng; k++) {
                                const uint64_t sub = (T >> si.start[k]) & ((1ull << si.norb[k]) - 1);
                                addr += rank_dev(sub) * stride[(int64_t)ts * ng + k];
                        }
                        const int lo = p < q ? p : q, hi = p < q ? q : p;
                        const uint64_t m = hi > lo + 1 ? ((1ull << hi) - (1ull << (lo + 1))) : 0ull;
                        const uint32_t neg = (uint32_t)(__popcll(S & m) & 1);
                        const uint32_t t = (uint32_t)(sbase[ts] + addr) | (neg << 31);
                        pmap[(int64_t)fcis::pair_index(p, q) * nstr + g] = (int32_t)t;
                        lk[(int64_t)this_slot * nstr + g] = (int32_t)t;
                }
        }
}

__device__ __forceinline__ int32_t pm_index(int32_t e) { return e & 0x7fffffff; }

// t1[pq][k] for the S1 columns col0 + k, k < ncol.
template <typename T>
__global__ __launch_bounds__(NT) void k_nat_gather(NatView v, int64_t col0, int64_t ncol, int64_t ld,
                                                   const double *__restrict__ c, T *__restrict__ t1)
{
        const int64_t k = (int64_t)blockIdx.x * NT + threadIdx.x;
        if (k >= ncol) return;
        const S1Pos p = s1_locate(v, col0 + k);
        const int64_t gb = v.sbaseB[p.B] + p.b;
        const int32_t *dA = v.dblk + (int64_t)p.A * v.nsecB;
        for (int pq = 0; pq < v.npair; pq++) {
                double x = 0.0;
                const int32_t ea = v.pmapA[(int64_t)pq * v.nstrA + p.r];
                if (ea != -1) {
                        const int32_t g = pm_index(ea);
                        const int blk = v.dblk[(int64_t)v.ssecA[g] * v.nsecB + p.B];
                        if (blk >= 0) {
                                const double y = c[v.boff[blk] + (int64_t)v.saddrA[g] * v.bnb[blk] + p.b];
                                x = ea < 0 ? -y : y;
                        }
                }
                const int32_t eb = v.pmapB[(int64_t)pq * v.nstrB + gb];
                if (eb != -1) {
                        const int32_t g = pm_index(eb);
                        const int blk = dA[v.ssecB[g]];
                        if (blk >= 0) {
                                const double y = c[v.boff[blk] + p.a * v.bnb[blk] + v.saddrB[g]];
                                x += eb < 0 ? -y : y;
                        }
                }
                t1[(int64_t)pq * ld + k] = (T)x;
        }
}

// sigma[d] (+)= pulls from vt1 of the S1 rows [r0, r1) (columns from col0).
template <typename T>
__global__ __launch_bounds__(NT) void k_nat_pull(NatView v, int64_t r0, int64_t r1, int64_t col0,
                                                 int64_t ld, int accumulate,
                                                 const T *__restrict__ vt1, double *__restrict__ sigma)
{
        const int64_t d = (int64_t)blockIdx.x * NT + threadIdx.x;
        if (d >= v.ndet) return;
        const DPos q = d_locate(v, d);
        const int64_t r = v.sbaseA[q.A] + q.a, gb = v.sbaseB[q.B] + q.b;
        double acc = accumulate ? sigma[d] : 0.0;
        for (int k = 0; k < v.nlinkA; k++) {
                const int32_t ea = v.lkA[(int64_t)k * v.nstrA + r];
                if (ea == -1) continue;
                const int64_t g = pm_index(ea);
                if (g < r0 || g >= r1) continue;
                const int A2 = v.ssecA[g];
                const int64_t bo = v.s1bo[(int64_t)A2 * v.nsecB + q.B];
                if (bo < 0) continue;
                const int pq = v.lkpA[(int64_t)k * v.nstrA + r];
                const int64_t col = v.cbase[A2] + (int64_t)v.saddrA[g] * v.rowlen[A2] + bo + q.b;
                const double x = (double)vt1[(int64_t)pq * ld + (col - col0)];
                acc += ea < 0 ? -x : x;
        }
        if (r >= r0 && r < r1) {
                const int64_t rowc = v.cbase[q.A] + q.a * v.rowlen[q.A] - col0;
                const int64_t *bo = v.s1bo + (int64_t)q.A * v.nsecB;
                for (int k = 0; k < v.nlinkB; k++) {
                        const int32_t eb = v.lkB[(int64_t)k * v.nstrB + gb];
                        if (eb == -1) continue;
                        const int32_t g = pm_index(eb);
                        const int64_t o = bo[v.ssecB[g]];
                        if (o < 0) continue;
                        const int pq = v.lkpB[(int64_t)k * v.nstrB + gb];
                        const double x = (double)vt1[(int64_t)pq * ld + rowc + o + v.saddrB[g]];
                        acc += eb < 0 ? -x : x;
                }
        }
        sigma[d] = acc;
}

// Diagonal of H: sum_p h_pp over both strings, (J - K) over same-spin
// pairs, J over opposite-spin pairs (fci_hdiag_gas.c:144-164, :678-709),
// with J[p][q] = (pp|qq), K[p][q] = (pq|qp) from the un-absorbed integrals.
__global__ __launch_bounds__(NT) void k_nat_hdiag(NatView v, const double *__restrict__ hjk,
                                                  double *__restrict__ out)
{
        const int64_t d = (int64_t)blockIdx.x * NT + threadIdx.x;
        if (d >= v.ndet) return;
        const DPos q = d_locate(v, d);
        const uint64_t sa = v.strA[v.sbaseA[q.A] + q.a], sb = v.strB[v.sbaseB[q.B] + q.b];
        const int n = v.n;
        const double *h = hjk, *J = hjk + n, *K = hjk + n + n * n;
        double e = 0.0;
        for (uint64_t x = sa; x; x &= x - 1) {
                const int p = __ffsll((long long)x) - 1;
                e += h[p];
                for (uint64_t y = sa & ((1ull << p) - 1); y; y &= y - 1) {
                        const int r = __ffsll((long long)y) - 1;
                        e += J[p * n + r] - K[p * n + r];
                }
                for (uint64_t y = sb; y; y &= y - 1) e += J[p * n + __ffsll((long long)y) - 1];
        }
        for (uint64_t x = sb; x; x &= x - 1) {
                const int p = __ffsll((long long)x) - 1;
                e += h[p];
                for (uint64_t y = sb & ((1ull << p) - 1); y; y &= y - 1) {
                        const int r = __ffsll((long long)y) - 1;
                        e += J[p * n + r] - K[p * n + r];
                }
        }
        out[d] = e;
}

// D^SPIN_x[d n + c][k] = <K| a_d^+ a_c |x> for the S1 columns col0 + k.
template <int SPIN>
__global__ __launch_bounds__(NT) void k_nat_rdm_gather(NatView v, int64_t col0, int64_t ncol, int64_t ld,
                                                       const double *__restrict__ x,
                                                       double *__restrict__ D)
{
        const int64_t k = (int64_t)blockIdx.x * NT + threadIdx.x;
        if (k >= ncol) return;
        const S1Pos p = s1_locate(v, col0 + k);
        const int64_t gb = v.sbaseB[p.B] + p.b;
        const uint64_t S = SPIN == 0 ? v.strA[p.r] : v.strB[gb];
        const int64_t self = d_index(v, p.A, p.a, p.B, p.b);
        const double xs = self >= 0 ? x[self] : 0.0;
        const int n = v.n;
        for (int dd = 0; dd < n; dd++) {
                const bool occd = (S >> dd) & 1;
                for (int cc = 0; cc < n; cc++) {
                        double val = 0.0;
                        if (occd) {
                                if (cc == dd) {
                                        val = xs;
                                } else if (!((S >> cc) & 1)) {
                                        const int pq = fcis::pair_index(cc, dd);
                                        const int32_t e = SPIN == 0 ? v.pmapA[(int64_t)pq * v.nstrA + p.r]
                                                                    : v.pmapB[(int64_t)pq * v.nstrB + gb];
                                        if (e != -1) {
                                                const int32_t g = pm_index(e);
                                                const int64_t j = SPIN == 0
                                                        ? d_index(v, v.ssecA[g], v.saddrA[g], p.B, p.b)
                                                        : d_index(v, p.A, p.a, v.ssecB[g], v.saddrB[g]);
                                                if (j >= 0) val = e < 0 ? -x[j] : x[j];
                                        }
                                }
                        }
                        D[(int64_t)(dd * n + cc) * ld + k] = val;
                }
        }
}

// e[k] = x at S1 column col0 + k if that determinant is in D, else 0.
__global__ __launch_bounds__(NT) void k_nat_s1_embed(NatView v, int64_t col0, int64_t ncol,
                                                     const double *__restrict__ x, double *__restrict__ e)
{
        const int64_t k = (int64_t)blockIdx.x * NT + threadIdx.x;
        if (k >= ncol) return;
        const S1Pos p = s1_locate(v, col0 + k);
        const int64_t j = d_index(v, p.A, p.a, p.B, p.b);
        e[k] = j >= 0 ? x[j] : 0.0;
}

// S^2 = S_z (S_z + 1) + N_beta - sum_pq E^alpha_pq E^beta_qp on the GAS
// determinants (P S^2 P; exact when D is spin complete).  For determinant I
// the double sum is |I_a & I_b| c_I (p = q, doubly occupied) plus, for every
// alpha-only orbital p and beta-only orbital q, the spin exchange
// s_a s_b c_J with J_a = a_q^+ a_p I_a (sign s_a), J_b = a_p^+ a_q I_b (s_b),
// both from the pair maps.  shift = S_z (S_z + 1) + N_beta.
__global__ __launch_bounds__(NT) void k_nat_ss(NatView v, double shift, const double *__restrict__ c,
                                               double *__restrict__ out)
{
        const int64_t d = (int64_t)blockIdx.x * NT + threadIdx.x;
        if (d >= v.ndet) return;
        const DPos q = d_locate(v, d);
        const int64_t r = v.sbaseA[q.A] + q.a, gb = v.sbaseB[q.B] + q.b;
        const uint64_t sa = v.strA[r], sb = v.strB[gb];
        double acc = (shift - (double)__popcll(sa & sb)) * c[d];
        for (uint64_t oa = sa & ~sb; oa; oa &= oa - 1) {
                const int p = __ffsll((long long)oa) - 1;
                for (uint64_t ob = sb & ~sa; ob; ob &= ob - 1) {
                        const int pq = fcis::pair_index(p, __ffsll((long long)ob) - 1);
                        const int32_t ea = v.pmapA[(int64_t)pq * v.nstrA + r];
                        const int32_t eb = v.pmapB[(int64_t)pq * v.nstrB + gb];
                        if (ea == -1 || eb == -1) continue;
                        const int32_t ja = pm_index(ea), jb = pm_index(eb);
                        const int64_t j = d_index(v, v.ssecA[ja], v.saddrA[ja], v.ssecB[jb], v.saddrB[jb]);
                        if (j < 0) continue;
                        acc -= ((ea < 0) != (eb < 0)) ? -c[j] : c[j];
                }
        }
        out[d] = acc;
}

__global__ __launch_bounds__(NT) void k_nat_ss_diag(NatView v, double shift, double *__restrict__ out)
{
        const int64_t d = (int64_t)blockIdx.x * NT + threadIdx.x;
        if (d >= v.ndet) return;
        const DPos q = d_locate(v, d);
        const uint64_t sa = v.strA[v.sbaseA[q.A] + q.a], sb = v.strB[v.sbaseB[q.B] + q.b];
        out[d] = shift - (double)__popcll(sa & sb);
}

// Direct sigma (no closure): one wavefront per (block, alpha address, 64
// beta addresses) of D, destination-owned.  With the link lists of the
// strings (pair P, target, sign) and h symmetric-summed hs = h + h^T,
//   sigma_I = sum over alpha links (P, K) of I_a, links (R, J) of K:
//                 h[P][R] s s' c[J, I_b]                      (alpha-alpha)
//           + the same over beta links                         (beta-beta)
//           + sum over alpha links (P, J_a) of I_a and beta links (R, J_b)
//             of I_b: hs[P][R] s s' c[J_a, J_b]                (alpha-beta)
// (the t1 / GEMM form's sum over the closure, with K restricted to one spin
// at a time; J must lie in D).  The alpha loops are uniform across the
// wavefront (scalar loads of links and h), the c loads of the alpha-alpha
// part coalesced; cost per determinant ~ (nlink_a + nlink_b)^2, against
// npair^2 per closure column for the GEMM form.
// Inner link loop of the direct form, wavefront-uniform string K: sum over
// the links (R, J) of K (slots of lk / lkp, ns strings) whose J pairs with
// a block of D through dtab[ssec[J] * dstride] of h[P][R] s s' x[boff +
// saddr[J] * rowlen + off] (rowlen the block's beta count, or rowlen_c when
// given), U links per step with all their loads issued together.
template <bool ROWLEN_BLOCK>
__device__ __forceinline__ double uniform_links(const NatView &v, const int32_t *__restrict__ lk,
                                                const int16_t *__restrict__ lkp, int64_t ns, int nl, int32_t K,
                                                const uint16_t *__restrict__ ssec,
                                                const uint32_t *__restrict__ saddr,
                                                const int32_t *__restrict__ dtab, int64_t dstride,
                                                const double *__restrict__ hP, const double *__restrict__ x,
                                                int64_t off, int64_t rowlen_c, bool neg1)
{
        constexpr int U = 8;
        double acc = 0.0;
        for (int k2 = 0; k2 < nl; k2 += U) {
                int32_t e[U], J[U], blk2[U];
#pragma unroll
                for (int u = 0; u < U; u++) e[u] = k2 + u < nl ? lk[(int64_t)(k2 + u) * ns + K] : -1;
#pragma unroll
                for (int u = 0; u < U; u++) J[u] = e[u] == -1 ? 0 : pm_index(e[u]);
#pragma unroll
                for (int u = 0; u < U; u++) blk2[u] = e[u] == -1 ? -1 : dtab[(int64_t)ssec[J[u]] * dstride];
#pragma unroll
                for (int u = 0; u < U; u++) {
                        if (blk2[u] < 0) continue;
                        const double coef = hP[lkp[(int64_t)(k2 + u) * ns + K]];
                        const int64_t rl = ROWLEN_BLOCK ? v.bnb[blk2[u]] : rowlen_c;
                        const double xv = x[v.boff[blk2[u]] + (int64_t)saddr[J[u]] * rl + off];
                        acc += ((e[u] < 0) != neg1) ? -coef * xv : coef * xv;
                }
        }
        return acc;
}

// The beta-beta part runs in k_nat_direct_bb on the transposed vector (the
// same uniform form with the spins' roles swapped); k_nat_direct adds its
// result st (transposed blocks) to the alpha-alpha and alpha-beta parts.
__global__ __launch_bounds__(NT) void k_nat_direct(NatView v, const int64_t *__restrict__ wbase,
                                                   const double *__restrict__ h,
                                                   const double *__restrict__ hs,
                                                   const double *__restrict__ c,
                                                   const double *__restrict__ st, double *__restrict__ sigma)
{
        const int64_t w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (NT / 64) + (threadIdx.x >> 6))) +
                          (int64_t)0;
        const int lane = threadIdx.x & 63;
        if (w >= wbase[v.nblock]) return;
        const int blk = upper_index(wbase, v.nblock, w);
        const int64_t nb = v.bnb[blk];
        const int64_t nch = (nb + 63) >> 6;
        const int64_t local = w - wbase[blk];
        const int64_t a = local / nch;
        const int64_t b = (local - a * nch) * 64 + lane;
        const bool valid = b < nb;
        const int64_t bl = valid ? b : nb - 1;
        const int A = v.bsa[blk], B = v.bsb[blk];
        const int64_t r = v.sbaseA[A] + a, gb = v.sbaseB[B] + bl;
        const int npair = v.npair;
        const int64_t nsa = v.nstrA, nsb = v.nstrB;
        double acc = 0.0;
        // alpha-alpha: uniform
        for (int k1 = 0; k1 < v.nlinkA; k1++) {
                const int32_t e1 = v.lkA[(int64_t)k1 * nsa + r];
                if (e1 == -1) continue;
                const int32_t K = pm_index(e1);
                const double *hP = h + (int64_t)v.lkpA[(int64_t)k1 * nsa + r] * npair;
                acc += uniform_links<true>(v, v.lkA, v.lkpA, nsa, v.nlinkA, K, v.ssecA, v.saddrA, v.dblk + B,
                                           v.nsecB, hP, c, bl, 0, e1 < 0);
        }
        // alpha-beta: alpha links uniform, beta links per lane
        for (int k1 = 0; k1 < v.nlinkA; k1++) {
                const int32_t e1 = v.lkA[(int64_t)k1 * nsa + r];
                if (e1 == -1) continue;
                const int32_t Ja = pm_index(e1);
                const int32_t *dA2 = v.dblk + (int64_t)v.ssecA[Ja] * v.nsecB;
                const int64_t aJ = v.saddrA[Ja];
                const double *hP = hs + (int64_t)v.lkpA[(int64_t)k1 * nsa + r] * npair;
                constexpr int U = 8;
                for (int k2 = 0; k2 < v.nlinkB; k2 += U) {
                        int32_t e2[U], Jb[U], blk2[U];
#pragma unroll
                        for (int u = 0; u < U; u++)
                                e2[u] = k2 + u < v.nlinkB ? v.lkB[(int64_t)(k2 + u) * nsb + gb] : -1;
#pragma unroll
                        for (int u = 0; u < U; u++) Jb[u] = e2[u] == -1 ? 0 : pm_index(e2[u]);
#pragma unroll
                        for (int u = 0; u < U; u++) blk2[u] = e2[u] == -1 ? -1 : dA2[v.ssecB[Jb[u]]];
#pragma unroll
                        for (int u = 0; u < U; u++) {
                                if (blk2[u] < 0) continue;
                                const double coef = hP[v.lkpB[(int64_t)(k2 + u) * nsb + gb]];
                                const double x = c[v.boff[blk2[u]] + aJ * v.bnb[blk2[u]] + v.saddrB[Jb[u]]];
                                acc += ((e1 < 0) != (e2[u] < 0)) ? -coef * x : coef * x;
                        }
                }
        }
        if (valid) {
                const int64_t na = (v.boff[blk + 1] - v.boff[blk]) / nb;
                sigma[v.boff[blk] + a * nb + b] = acc + st[v.boff[blk] + b * na + a];
        }
}

// Beta-beta part of the direct sigma on transposed blocks (block b of D
// stored beta-major, [nstr(sb)][nstr(sa)] at the same offset): one
// wavefront per (block, beta address, 64 alpha addresses), the beta link
// loops uniform and the loads of ct coalesced:
//   st[I] = sum over beta links (P, K) of I_b, links (R, J) of K:
//           h[P][R] s s' ct[I_a, J]
__global__ __launch_bounds__(NT) void k_nat_direct_bb(NatView v, const int64_t *__restrict__ wbase,
                                                      const double *__restrict__ h,
                                                      const double *__restrict__ ct, double *__restrict__ st)
{
        const int64_t w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (NT / 64) + (threadIdx.x >> 6))) +
                          (int64_t)0;
        const int lane = threadIdx.x & 63;
        if (w >= wbase[v.nblock]) return;
        const int blk = upper_index(wbase, v.nblock, w);
        const int64_t nb = v.bnb[blk];
        const int64_t na = (v.boff[blk + 1] - v.boff[blk]) / nb;
        const int64_t nch = (na + 63) >> 6;
        const int64_t local = w - wbase[blk];
        const int64_t b = local / nch;
        const int64_t a = (local - b * nch) * 64 + lane;
        const bool valid = a < na;
        const int64_t al = valid ? a : na - 1;
        const int A = v.bsa[blk], B = v.bsb[blk];
        const int64_t gb = v.sbaseB[B] + b;
        const int npair = v.npair;
        const int64_t nsb = v.nstrB;
        const int32_t *dA = v.dblk + (int64_t)A * v.nsecB;
        double acc = 0.0;
        for (int k1 = 0; k1 < v.nlinkB; k1++) {
                const int32_t e1 = v.lkB[(int64_t)k1 * nsb + gb];
                if (e1 == -1) continue;
                const int32_t K = pm_index(e1);
                const double *hP = h + (int64_t)v.lkpB[(int64_t)k1 * nsb + gb] * npair;
                acc += uniform_links<false>(v, v.lkB, v.lkpB, nsb, v.nlinkB, K, v.ssecB, v.saddrB, dA, 1, hP, ct,
                                            al, na, e1 < 0);
        }
        if (valid) st[v.boff[blk] + b * na + a] = acc;
}

// ct[block b] = c[block b] transposed (alpha-major -> beta-major).
__global__ __launch_bounds__(NT) void k_nat_transpose(NatView v, const double *__restrict__ c,
                                                      double *__restrict__ ct)
{
        const int64_t d = (int64_t)blockIdx.x * NT + threadIdx.x;
        if (d >= v.ndet) return;
        const DPos q = d_locate(v, d);
        const int64_t nb = v.bnb[q.blk], na = (v.boff[q.blk + 1] - v.boff[q.blk]) / nb;
        ct[v.boff[q.blk] + q.b * na + q.a] = c[d];
}

// ---- Operator form of the direct sigma ----
// Same-spin part as a sparse one-spin operator: for the strings I of a
// spin's D sectors (sectors that are that spin's sector of some block),
//   A[I][J] = sum_{P,R} h[P][R] <I|F_P|K> <K|F_R|J>   (K over all strings)
// over the strings J of the spin's D sectors at most two excitations from I
// (I itself, single and double replacements), so that
//   sigma_aa[I_a, I_b] = sum_J A_a[I_a][J] c[J, I_b]
// (J's block with I_b's sector must be in D).  A row's entries are grouped
// by J's sector in ascending order (segments); the number of strings of a
// sector T within two excitations of a string of sector S depends only on
// the two sectors' occupations, so every row of S has the same length and
// the same segment boundaries.  Within a segment the entries follow the
// enumeration order of the replacements of I: I, then (q, p) for the k
// occupied q and n - k empty p, then pairs q < s of occupied and p < r of
// empty orbitals.  The opposite-spin part walks link lists sorted the same
// way (targets in the spin's D sectors only, diagonal slots first), so a
// wavefront skips whole segments whose sector pair is not a block of D.
constexpr int OP_MAXSEG = 1024;    // segments per source sector (LDS counters)

struct OpView {
        int ndsec;
        int64_t nrows;                  // strings of the spin's D sectors
        const int32_t *dsec;            // [ndsec] local sector
        const int32_t *dindex;          // [nsec] position in dsec, or -1
        const int64_t *rowbase;         // [ndsec + 1] first row of each D sector
        const int64_t *opbase;          // [ndsec] first operator entry
        const int32_t *oplen;           // [ndsec] row length
        const int32_t *opsegptr;        // [ndsec + 1]
        const int32_t *opsegT, *opsegoff;   // segment target sector and start in the row
        const int16_t *opsegk;          // [ndsec][nsec] segment (relative) or -1
        const int64_t *lsbase;          // [ndsec] first link entry
        const int32_t *lslen;           // [ndsec] links per string
        const int32_t *lssegptr;        // [ndsec + 1]
        const int32_t *lssegT, *lssegoff;
        const int16_t *lssegk;          // [ndsec][nsec]
        const uint32_t *opaddr;         // target address within its sector
        const uint64_t *ls;             // address | pair << 32 | sign << 63
        const double *opcoef;           // A entries (per plan)
};

__device__ __forceinline__ int cre_des_neg(uint64_t S, int cre, int des)
{
        const int lo = cre < des ? cre : des, hi = cre < des ? des : cre;
        const uint64_t m = hi > lo + 1 ? ((1ull << hi) - (1ull << (lo + 1))) : 0ull;
        return __popcll(S & m) & 1;
}

__device__ __forceinline__ uint64_t lane_below_mask()
{
        const int lane = threadIdx.x & 63;
        return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// (i, j), i < j, of combination index t over pairs ordered by j, then i.
__device__ __forceinline__ void unpair(int64_t t, int &i, int &j)
{
        int jj = (int)((1.0 + sqrt(1.0 + 8.0 * (double)t)) * 0.5);
        while ((int64_t)jj * (jj - 1) / 2 > t) jj--;
        while ((int64_t)(jj + 1) * jj / 2 <= t) jj++;
        j = jj;
        i = (int)(t - (int64_t)jj * (jj - 1) / 2);
}

// Local sector of the string reached from sector S by moving one electron
// from orbital q to orbital p (same sector inside one subspace), or -1.
__device__ __forceinline__ int move_sector(const SubInfo &si, const int32_t *__restrict__ mv, int S, int q, int p)
{
        if (S < 0) return -1;
        const int kq = si.sub_of[q], kp = si.sub_of[p];
        return kq == kp ? S : mv[((int64_t)S * si.ngas + kq) * si.ngas + kp];
}

__device__ __forceinline__ int64_t sector_address(const SubInfo &si, const int64_t *__restrict__ stride, int T,
                                                  uint64_t J)
{
        int64_t addr = 0;
        for (int k = 0; k < si.ngas; k++) {
                const uint64_t sub = (J >> si.start[k]) & ((1ull << si.norb[k]) - 1);
                addr += rank_dev(sub) * stride[(int64_t)T * si.ngas + k];
        }
        return addr;
}

// Structure of one spin's operator rows (OPS) or sorted link lists (!OPS):
// one wavefront per string of the spin's D sectors enumerates its
// replacements, keeps those whose sector is a segment of its row and writes
// them in enumeration order within each segment (ranks by ballot; per-
// segment counters in LDS).  Link lists are row-major (ROWMAJOR: the links
// of one string contiguous) or slot-major within each D sector (slot s of
// string a of sector i at lsbase[i] + s * nstr + a).
template <bool OPS>
__global__ __launch_bounds__(NT) void k_op_struct(NatView v, OpView o, int spin, int rowmajor, SubInfo si,
                                                  const int64_t *__restrict__ stride,
                                                  const int32_t *__restrict__ mv, uint32_t *__restrict__ opaddr,
                                                  uint64_t *__restrict__ ls, int *__restrict__ err)
{
        __shared__ int32_t cnt[NT / 64][OP_MAXSEG];
        __shared__ int8_t orb[NT / 64][64];
        const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int64_t row = (int64_t)blockIdx.x * (NT / 64) + wv;
        const bool active = row < o.nrows;
        const int i = active ? upper_index(o.rowbase, o.ndsec, row) : 0;
        const int64_t a = active ? row - o.rowbase[i] : 0;
        const int S = o.dsec[i];
        const int64_t *sbase = spin ? v.sbaseB : v.sbaseA;
        const uint64_t I = active ? (spin ? v.strB : v.strA)[sbase[S] + a] : 0ull;
        const int n = v.n, k = __popcll(I), nv = n - k;
        const bool occb = lane < n && ((I >> lane) & 1);
        const uint64_t mo = __ballot(occb), me = __ballot(lane < n && !occb);
        if (occb) orb[wv][__popcll(mo & lane_below_mask())] = (int8_t)lane;
        else if (lane < n) orb[wv][k + __popcll(me & lane_below_mask())] = (int8_t)lane;
        const int sp0 = OPS ? o.opsegptr[i] : o.lssegptr[i];
        const int nseg = (OPS ? o.opsegptr[i + 1] : o.lssegptr[i + 1]) - sp0;
        for (int j = lane; j < nseg && j < OP_MAXSEG; j += 64) cnt[wv][j] = 0;
        __syncthreads();
        if (!active) return;
        const int16_t *segk = (OPS ? o.opsegk : o.lssegk) + (int64_t)i * (spin ? v.nsecB : v.nsecA);
        const int32_t *segoff = (OPS ? o.opsegoff : o.lssegoff) + sp0;
        const int64_t nsS = o.rowbase[i + 1] - o.rowbase[i];
        const int32_t rowlen = OPS ? o.oplen[i] : o.lslen[i];
        const int64_t total = OPS ? 1 + (int64_t)k * nv + (int64_t)(k * (k - 1) / 2) * (nv * (nv - 1) / 2)
                                  : (int64_t)k + (int64_t)k * nv;
        const int64_t c2v = (int64_t)nv * (nv - 1) / 2;
        for (int64_t x0 = 0; x0 < total; x0 += 64) {
                const int64_t x = x0 + lane;
                int seg = -1;
                uint64_t val = 0;
                if (x < total) {
                        uint64_t J = I;
                        int T = S;
                        int cre = -1, des = -1;
                        if (OPS) {
                                if (x == 0) {
                                } else if (x < 1 + (int64_t)k * nv) {
                                        const int64_t y = x - 1;
                                        des = orb[wv][y / nv];
                                        cre = orb[wv][k + y % nv];
                                        T = move_sector(si, mv, S, des, cre);
                                        J = I ^ (1ull << des) ^ (1ull << cre);
                                } else {
                                        const int64_t y = x - 1 - (int64_t)k * nv;
                                        int oi, oj, vi, vj;
                                        unpair(y / c2v, oi, oj);
                                        unpair(y % c2v, vi, vj);
                                        const int q = orb[wv][oi], s2 = orb[wv][oj];
                                        const int p = orb[wv][k + vi], r = orb[wv][k + vj];
                                        T = move_sector(si, mv, move_sector(si, mv, S, q, p), s2, r);
                                        J = I ^ (1ull << q) ^ (1ull << s2) ^ (1ull << p) ^ (1ull << r);
                                }
                        } else {
                                if (x < k) {
                                        des = cre = orb[wv][x];
                                } else {
                                        const int64_t y = x - k;
                                        des = orb[wv][y / nv];
                                        cre = orb[wv][k + y % nv];
                                        T = move_sector(si, mv, S, des, cre);
                                        J = I ^ (1ull << des) ^ (1ull << cre);
                                }
                        }
                        if (T >= 0) seg = segk[T];
                        if (seg >= 0) {
                                const uint64_t addr = (uint64_t)sector_address(si, stride, T, J);
                                val = OPS ? addr
                                          : addr | ((uint64_t)fcis::pair_index(cre, des) << 32) |
                                                    ((uint64_t)(cre != des ? cre_des_neg(I, cre, des) : 0) << 63);
                        }
                }
                uint64_t pend = __ballot(seg >= 0);
                while (pend) {
                        const int l = __ffsll((long long)pend) - 1;
                        const int k0 = __builtin_amdgcn_readlane(seg, l);
                        const uint64_t m = __ballot(seg == k0);
                        const int base = cnt[wv][k0];
                        const int32_t send = k0 + 1 < nseg ? segoff[k0 + 1] : rowlen;
                        if (seg == k0) {
                                const int64_t slot = segoff[k0] + base + __popcll(m & lane_below_mask());
                                if (slot >= send)
                                        atomicAdd(err, 1);
                                else if (OPS)
                                        opaddr[o.opbase[i] + a * o.oplen[i] + slot] = (uint32_t)val;
                                else
                                        ls[o.lsbase[i] + (rowmajor ? a * o.lslen[i] + slot : slot * nsS + a)] = val;
                        }
                        __builtin_amdgcn_wave_barrier();
                        if (lane == 0) cnt[wv][k0] = base + __popcll(m);
                        __builtin_amdgcn_wave_barrier();
                        pend &= ~m;
                }
        }
        // every segment filled exactly
        for (int j = lane; j < nseg; j += 64) {
                const int32_t send = j + 1 < nseg ? segoff[j + 1] : rowlen;
                if (cnt[wv][j] != send - segoff[j]) atomicAdd(err, 1);
        }
}

// <I| sum_{P,R} h[P][R] F_P F_R |J> over one spin's strings (I and J at
// most two replacements apart), summed over the paths I -> K -> J in a
// fixed order.
__device__ double op_element(uint64_t I, uint64_t J, int n, const double *__restrict__ h, int np)
{
        const uint64_t d = I ^ J;
        const int nd = __popcll(d) >> 1;
        const uint64_t full = n == 64 ? ~0ull : ((1ull << n) - 1);
        double acc = 0.0;
        if (nd == 0) {
                for (uint64_t x = I; x; x &= x - 1) {
                        const int p = __ffsll((long long)x) - 1;
                        const double *hp = h + (int64_t)fcis::pair_index(p, p) * np;
                        for (uint64_t y = I; y; y &= y - 1) {
                                const int r = __ffsll((long long)y) - 1;
                                acc += hp[fcis::pair_index(r, r)];
                        }
                }
                for (uint64_t x = I; x; x &= x - 1) {
                        const int q = __ffsll((long long)x) - 1;
                        for (uint64_t y = ~I & full; y; y &= y - 1) {
                                const int e = __ffsll((long long)y) - 1;
                                const int P = fcis::pair_index(e, q);
                                acc += h[(int64_t)P * np + P];
                        }
                }
        } else if (nd == 1) {
                const int q = __ffsll((long long)(I & d)) - 1, p = __ffsll((long long)(J & d)) - 1;
                const int Ppq = fcis::pair_index(p, q);
                const double sJ = cre_des_neg(I, p, q) ? -1.0 : 1.0;
                double t = 0.0;
                for (uint64_t x = I; x; x &= x - 1) {
                        const int r = __ffsll((long long)x) - 1;
                        t += h[(int64_t)fcis::pair_index(r, r) * np + Ppq];
                }
                for (uint64_t x = J; x; x &= x - 1) {
                        const int r = __ffsll((long long)x) - 1;
                        t += h[(int64_t)Ppq * np + fcis::pair_index(r, r)];
                }
                acc = sJ * t;
                for (uint64_t y = ~I & full & ~(1ull << p); y; y &= y - 1) {
                        const int e = __ffsll((long long)y) - 1;
                        const uint64_t K = I ^ (1ull << q) ^ (1ull << e);
                        const int ng = cre_des_neg(I, e, q) ^ cre_des_neg(K, p, e);
                        const double v = h[(int64_t)fcis::pair_index(e, q) * np + fcis::pair_index(p, e)];
                        acc += ng ? -v : v;
                }
                for (uint64_t y = I & ~(1ull << q); y; y &= y - 1) {
                        const int e = __ffsll((long long)y) - 1;
                        const uint64_t K = I ^ (1ull << e) ^ (1ull << p);
                        const int ng = cre_des_neg(I, p, e) ^ cre_des_neg(K, e, q);
                        const double v = h[(int64_t)fcis::pair_index(p, e) * np + fcis::pair_index(e, q)];
                        acc += ng ? -v : v;
                }
        } else if (nd == 2) {
                const uint64_t rm = I & d, ad = J & d;
                const int q0 = __ffsll((long long)rm) - 1, q1 = 63 - __clzll((long long)rm);
                const int p0 = __ffsll((long long)ad) - 1, p1 = 63 - __clzll((long long)ad);
                const int qs[2] = {q0, q1}, ps[2] = {p0, p1};
#pragma unroll
                for (int u = 0; u < 2; u++)
#pragma unroll
                        for (int w = 0; w < 2; w++) {
                                const int a1 = qs[u], b1 = ps[w], a2 = qs[1 - u], b2 = ps[1 - w];
                                const uint64_t K = I ^ (1ull << a1) ^ (1ull << b1);
                                const int ng = cre_des_neg(I, b1, a1) ^ cre_des_neg(K, b2, a2);
                                const double v =
                                        h[(int64_t)fcis::pair_index(b1, a1) * np + fcis::pair_index(b2, a2)];
                                acc += ng ? -v : v;
                        }
        }
        return acc;
}

// Operator entries of one spin for the integrals h: one wavefront per row.
__global__ __launch_bounds__(NT) void k_op_coef(NatView v, OpView o, int spin, const double *__restrict__ h,
                                                double *__restrict__ coef)
{
        const int64_t row = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (NT / 64) + (threadIdx.x >> 6))) +
                            (int64_t)0;
        const int lane = threadIdx.x & 63;
        if (row >= o.nrows) return;
        const int i = upper_index(o.rowbase, o.ndsec, row);
        const int64_t a = row - o.rowbase[i];
        const int64_t *sbase = spin ? v.sbaseB : v.sbaseA;
        const uint64_t *str = spin ? v.strB : v.strA;
        const uint64_t I = str[sbase[o.dsec[i]] + a];
        const int64_t r0 = o.opbase[i] + a * o.oplen[i];
        const int s0 = o.opsegptr[i], s1 = o.opsegptr[i + 1];
        for (int sg = s0; sg < s1; sg++) {
                const int64_t e0 = o.opsegoff[sg], e1 = sg + 1 < s1 ? o.opsegoff[sg + 1] : o.oplen[i];
                const int64_t jb = sbase[o.opsegT[sg]];
                for (int64_t e = e0 + lane; e < e1; e += 64)
                        coef[r0 + e] = op_element(I, str[jb + o.opaddr[r0 + e]], v.n, h, v.npair);
        }
}

// Same-spin part through the operator of one spin: the sum over J of a row
// of A times x[base(J's block) + addr(J) * ld]; the block of J's sector T
// with the other spin's sector is dblk[T * tstride + other] (alpha: T * nsecB
// + B; beta: A * nsecB + T).
__device__ __forceinline__ double op_row(const OpView &o, int i, int64_t row0, const NatView &v,
                                         const int32_t *__restrict__ dtab, int64_t tstride,
                                         const double *__restrict__ x, int64_t lane_off, int64_t ld)
{
        double acc = 0.0;
        const int s0 = o.opsegptr[i], s1 = o.opsegptr[i + 1];
        for (int sg = s0; sg < s1; sg++) {
                const int blk2 = dtab[(int64_t)o.opsegT[sg] * tstride];
                if (blk2 < 0) continue;
                const int64_t e0 = row0 + o.opsegoff[sg], e1 = row0 + (sg + 1 < s1 ? o.opsegoff[sg + 1] : o.oplen[i]);
                const double *xb = x + v.boff[blk2] + lane_off;
                constexpr int U = 8;
                int64_t e = e0;
                for (; e + U <= e1; e += U) {
                        double cf[U], xv[U];
#pragma unroll
                        for (int u = 0; u < U; u++) {
                                cf[u] = o.opcoef[e + u];
                                xv[u] = xb[(int64_t)o.opaddr[e + u] * ld];
                        }
#pragma unroll
                        for (int u = 0; u < U; u++) acc += cf[u] * xv[u];
                }
                for (; e < e1; e++) acc += o.opcoef[e] * xb[(int64_t)o.opaddr[e] * ld];
        }
        return acc;
}

// Wavefront geometry of the operator-form kernels on a block with `nr`
// rows (the wavefront-uniform string) and `ncol` columns (the lane string):
// ncol > 32: one row and 64 columns per wavefront; ncol <= 32: g = 64 / ncol
// rows of all columns per wavefront (rows of one sector share their segment
// structure, so the loops stay uniform; only the row's tables differ per
// lane).  Wavefronts per block: op_waves.
__host__ __device__ __forceinline__ int64_t op_waves(int64_t nr, int64_t ncol, int packmax)
{
        return ncol > packmax ? nr * ((ncol + 63) / 64) : (nr + 64 / ncol - 1) / (64 / ncol);
}

struct OpLane {
        int64_t row, col;     // clamped into the block
        bool valid;
};

__device__ __forceinline__ OpLane op_lane(int64_t local, int lane, int64_t nr, int64_t ncol, int packmax)
{
        OpLane q;
        if (ncol > packmax) {
                const int64_t nch = (ncol + 63) >> 6;
                q.row = local / nch;
                const int64_t b = (local - q.row * nch) * 64 + lane;
                q.valid = b < ncol;
                q.col = q.valid ? b : ncol - 1;
        } else {
                const int g = 64 / (int)ncol;
                const int r = lane / (int)ncol;
                const int64_t a = local * g + r;
                q.valid = r < g && a < nr;
                q.row = q.valid ? a : nr - 1;
                q.col = lane - r * ncol < ncol ? lane - r * ncol : ncol - 1;
        }
        return q;
}

// Beta-beta part on the transposed vector (block b stored beta-major):
// wavefronts over (block, beta addresses, alpha addresses) as op_lane.
__global__ __launch_bounds__(NT) void k_nat_op_bb(NatView v, OpView ob, const int64_t *__restrict__ wbase,
                                                  int packmax, const double *__restrict__ ct,
                                                  double *__restrict__ st)
{
        const int64_t w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (NT / 64) + (threadIdx.x >> 6))) +
                          (int64_t)0;
        const int lane = threadIdx.x & 63;
        if (w >= wbase[v.nblock]) return;
        const int blk = upper_index(wbase, v.nblock, w);
        const int64_t nb = v.bnb[blk];
        const int64_t na = (v.boff[blk + 1] - v.boff[blk]) / nb;
        const OpLane q = op_lane(w - wbase[blk], lane, nb, na, packmax);
        const int A = v.bsa[blk], B = v.bsb[blk];
        const int ib = ob.dindex[B];
        const double acc = op_row(ob, ib, ob.opbase[ib] + q.row * ob.oplen[ib], v, v.dblk + (int64_t)A * v.nsecB, 1,
                                  ct, q.col, na);
        if (q.valid) st[v.boff[blk] + q.row * na + q.col] = acc;
}

// alpha-alpha and alpha-beta parts of the operator form at alpha address a
// (wavefront-uniform when each wavefront holds one alpha string) and beta
// address bl of block blk.
__device__ __forceinline__ double op_direct_at(const NatView &v, const OpView &oa, const OpView &ob, int blk,
                                               int64_t a, int64_t bl, int64_t nb, const double *__restrict__ hs,
                                               const double *__restrict__ c)
{
        const int A = v.bsa[blk], B = v.bsb[blk];
        const int ia = oa.dindex[A], ib = ob.dindex[B];
        const int npair = v.npair;
        // alpha-alpha
        double acc = op_row(oa, ia, oa.opbase[ia] + a * oa.oplen[ia], v, v.dblk + B, v.nsecB, c, bl, nb);
        // alpha-beta
        const uint64_t *la = oa.ls + oa.lsbase[ia] + a * oa.lslen[ia];
        const int64_t nsB = ob.rowbase[ib + 1] - ob.rowbase[ib];
        const uint64_t *lb = ob.ls + ob.lsbase[ib] + bl;
        const int ka0 = oa.lssegptr[ia], ka1 = oa.lssegptr[ia + 1];
        const int kb0 = ob.lssegptr[ib], kb1 = ob.lssegptr[ib + 1];
        for (int ka = ka0; ka < ka1; ka++) {
                const int32_t *dA2 = v.dblk + (int64_t)oa.lssegT[ka] * v.nsecB;
                const int sa0 = oa.lssegoff[ka], sa1 = ka + 1 < ka1 ? oa.lssegoff[ka + 1] : oa.lslen[ia];
                for (int kb = kb0; kb < kb1; kb++) {
                        const int blk2 = dA2[ob.lssegT[kb]];
                        if (blk2 < 0) continue;
                        const int64_t nb2 = v.bnb[blk2];
                        const double *cb = c + v.boff[blk2];
                        const int sb0 = ob.lssegoff[kb], sb1 = kb + 1 < kb1 ? ob.lssegoff[kb + 1] : ob.lslen[ib];
                        constexpr int U = 8;
                        for (int j0 = sb0; j0 < sb1; j0 += U) {
                                int64_t bj[U];
                                int rr[U];
                                bool ng[U];
#pragma unroll
                                for (int u = 0; u < U; u++) {
                                        const uint64_t e = j0 + u < sb1 ? lb[(int64_t)(j0 + u) * nsB] : 0ull;
                                        bj[u] = (int64_t)(uint32_t)e;
                                        rr[u] = (int)((e >> 32) & 0xffff);
                                        ng[u] = (e >> 63) != 0;
                                }
                                const int nu = sb1 - j0 < U ? sb1 - j0 : U;
                                for (int s = sa0; s < sa1; s++) {
                                        const uint64_t e = la[s];
                                        const double *hrow = hs + (int64_t)((e >> 32) & 0xffff) * npair;
                                        const double *xrow = cb + (int64_t)(uint32_t)e * nb2;
                                        const bool na_ = (e >> 63) != 0;
#pragma unroll
                                        for (int u = 0; u < U; u++) {
                                                if (u < nu) {
                                                        const double t = hrow[rr[u]] * xrow[bj[u]];
                                                        acc += (na_ != ng[u]) ? -t : t;
                                                }
                                        }
                                }
                        }
                }
        }
        return acc;
}

// Direct sigma in the operator form: alpha-alpha through A_alpha on c,
// alpha-beta over the sorted link lists (alpha links of I_a row-major,
// beta links of I_b per lane, slot-major, eight held in registers while the
// alpha links of a segment pass), plus the beta-beta part st (transposed
// blocks).  Wavefronts over (block, alpha addresses, beta addresses) as
// op_lane: one alpha string per wavefront (uniform loads) where the block
// has 64 beta strings or more, several otherwise.
__global__ __launch_bounds__(NT) void k_nat_op_direct(NatView v, OpView oa, OpView ob,
                                                      const int64_t *__restrict__ wbase, int packmax,
                                                      const double *__restrict__ hs, const double *__restrict__ c,
                                                      const double *__restrict__ st, double *__restrict__ sigma)
{
        const int64_t w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (NT / 64) + (threadIdx.x >> 6))) +
                          (int64_t)0;
        const int lane = threadIdx.x & 63;
        if (w >= wbase[v.nblock]) return;
        const int blk = upper_index(wbase, v.nblock, w);
        const int64_t nb = v.bnb[blk];
        const int64_t na = (v.boff[blk + 1] - v.boff[blk]) / nb;
        const int64_t local = w - wbase[blk];
        const OpLane q = op_lane(local, lane, na, nb, packmax);
        double acc;
        if (nb > packmax) {
                // one alpha string per wavefront: its address is uniform
                const int64_t a = __builtin_amdgcn_readfirstlane((int)q.row) + (int64_t)0;
                acc = op_direct_at(v, oa, ob, blk, a, q.col, nb, hs, c);
        } else {
                acc = op_direct_at(v, oa, ob, blk, q.row, q.col, nb, hs, c);
        }
        if (q.valid) sigma[v.boff[blk] + q.row * nb + q.col] = acc + st[v.boff[blk] + q.col * na + q.row];
}

__global__ void k_nat_fill(int32_t *__restrict__ x, int64_t n, int32_t val)
{
        for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT)
                x[i] = val;
}

unsigned blocks_for(int64_t n)
{
        return (unsigned)std::max<int64_t>(1, (n + NT - 1) / NT);
}

}  // namespace

struct gas_native {
        int n = 0, npair = 0;
        int nsec[2] = {0, 0};
        std::vector<int> glob[2];       // local sector -> pool sector
        std::vector<int64_t> sbase[2];  // [nsec + 1] first global string index
        int64_t nstr[2] = {0, 0};
        std::vector<int64_t> rowlen;    // [nsecA] S1 row length
        std::vector<int64_t> cbase;     // [nsecA + 1] first S1 column of the sector's rows
        int64_t n1 = 0;                 // S1 size
        int64_t maxrow = 0;
        SubInfo si{};
        NatView v{};
        std::vector<void *> dev;        // device allocations
        // per spin: local sector occupations, strides and moves (host and
        // device), the D sectors (local ids) of the spin
        std::vector<uint8_t> occ[2];
        std::vector<int64_t> stride[2];
        std::vector<int32_t> mv[2];
        const int64_t *d_stride[2] = {nullptr, nullptr};
        const int32_t *d_mv[2] = {nullptr, nullptr};
        std::vector<int32_t> dsec[2];
        // operator form: host segment tables (op_host), device structure
        // (op_device: 0 not built, 1 built, -1 failed)
        bool op_host_done = false;
        int op_dev_state = 0;
        struct OpHost {
                std::vector<int32_t> dindex;
                std::vector<int64_t> rowbase, opbase, lsbase;
                std::vector<int32_t> oplen, opsegptr, opsegT, opsegoff;
                std::vector<int32_t> lslen, lssegptr, lssegT, lssegoff;
                std::vector<int16_t> opsegk, lssegk;
                int64_t nnz = 0, nls = 0;
                int maxseg = 0;
        } oph[2];
        OpView op[2] = {};              // device tables (opcoef unset)
        double op_same = 0.0, op_cross = 0.0;   // estimated wavefront iterations per sigma
        std::vector<int32_t> bsa_loc, bsb_loc, dblk_loc;
};

struct gas_native_plan {
        const fcis_gas_space *g = nullptr;
        int precision = FCIS_PREC_FP64;
        int direct = 0;                 // 0: t1 / GEMM / pull, 1: link-pair loops, 2: operator form
        double *d_opcoef = nullptr;     // operator form: A_alpha then A_beta entries
        // operator form: several strings per wavefront on blocks of at most
        // packmax columns (32: measured best of 32 / 16 / 8 / 4 / none on
        // spaces of 16-62 orbitals)
        int packmax = 32;
        OpView oa{}, ob{};
        rocblas_handle blas = nullptr;
        void *d_h = nullptr;            // npair x npair, double or float (direct: h, then h + h^T)
        void *d_t1 = nullptr, *d_vt1 = nullptr;
        int64_t ld = 0;                 // columns per tile
        std::vector<int64_t> r0, c0;    // tile row and column starts (+ ends)
        int64_t *d_wbase = nullptr;     // direct: wavefronts before each block (+ total),
        int64_t nwave = 0, nwave_bb = 0;   // then the same for the transposed beta-beta kernel
        double *d_ct = nullptr;         // direct: transposed c and beta-beta part (2 ndet)
};

namespace {

template <typename T>
int dev_upload(gas_native *nt, T **out, const std::vector<T> &h, hipStream_t st)
{
        *out = nullptr;
        if (fcis::dev_malloc(out, sizeof(T) * std::max<size_t>(h.size(), 1)) != hipSuccess) {
                fcis::set_error("sector-native GAS space: device allocation failed");
                return FCIS_ERR_MEMORY;
        }
        nt->dev.push_back(*out);
        if (!h.empty())
                FCIS_HIP(hipMemcpyAsync(*out, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, st));
        return FCIS_SUCCESS;
}

template <typename T>
int dev_alloc(gas_native *nt, T **out, size_t count)
{
        *out = nullptr;
        if (fcis::dev_malloc(out, sizeof(T) * std::max<size_t>(count, 1)) != hipSuccess) {
                fcis::set_error("sector-native GAS space: device allocation of %zu bytes failed",
                                sizeof(T) * count);
                return FCIS_ERR_MEMORY;
        }
        nt->dev.push_back(*out);
        return FCIS_SUCCESS;
}

#define NAT_TRY(x)                                \
        do {                                      \
                int r_ = (x);                     \
                if (r_ != FCIS_SUCCESS) return r_; \
        } while (0)

int build_native(fcis_gas_space *g, gas_native *nt, const fcis::SectorPool &pool, hipStream_t st)
{
        const int ng = g->ngas;
        int n = 0;
        for (int k = 0; k < ng; k++) n += g->norb[k];
        nt->n = n;
        nt->npair = n * (n + 1) / 2;
        nt->si.ngas = ng;
        for (int k = 0; k < ng; k++) {
                nt->si.norb[k] = g->norb[k];
                nt->si.start[k] = g->start[k];
                for (int p = g->start[k]; p < g->start[k] + g->norb[k]; p++) nt->si.sub_of[p] = (int8_t)k;
        }
        // S1 sector pairs (pool ids)
        std::set<std::pair<int, int>> s1;
        for (int b = 0; b < g->nblock; b++) {
                const int sa = g->bsa[b], sb = g->bsb[b];
                for (int x : fcis::gas_moves(g, pool, sa)) s1.insert({x, sb});
                for (int y : fcis::gas_moves(g, pool, sb)) s1.insert({sa, y});
        }
        std::vector<int> loc[2] = {std::vector<int>(g->nsector, -1), std::vector<int>(g->nsector, -1)};
        {
                std::set<int> sa, sb;
                for (auto &e : s1) {
                        sa.insert(e.first);
                        sb.insert(e.second);
                }
                nt->glob[0].assign(sa.begin(), sa.end());
                nt->glob[1].assign(sb.begin(), sb.end());
        }
        for (int spin = 0; spin < 2; spin++) {
                nt->nsec[spin] = (int)nt->glob[spin].size();
                nt->sbase[spin].assign(nt->nsec[spin] + 1, 0);
                for (int i = 0; i < nt->nsec[spin]; i++) {
                        loc[spin][nt->glob[spin][i]] = i;
                        nt->sbase[spin][i + 1] = nt->sbase[spin][i] + g->nstr[nt->glob[spin][i]];
                }
                nt->nstr[spin] = nt->sbase[spin][nt->nsec[spin]];
                if (nt->nstr[spin] >= 0x7fffffffLL) {
                        fcis::set_error("sector-native GAS space: 2^31 or more %s strings",
                                        spin ? "beta" : "alpha");
                        return FCIS_ERR_INVALID;
                }
        }
        const int nsa = nt->nsec[0], nsb = nt->nsec[1];
        if ((int64_t)nsa * nsb > ((int64_t)1 << 24)) {
                fcis::set_error("sector-native GAS space: %d x %d sector pairs exceed 2^24", nsa, nsb);
                return FCIS_ERR_INVALID;
        }
        // S1 rows: per alpha sector the beta sectors of S1, in order
        std::vector<int32_t> segptr(nsa + 1, 0), segsec;
        std::vector<int64_t> segoff, s1bo((size_t)nsa * nsb, -1);
        nt->rowlen.assign(nsa, 0);
        nt->cbase.assign(nsa + 1, 0);
        {
                std::vector<std::vector<int>> lists(nsa);
                for (auto &e : s1) lists[loc[0][e.first]].push_back(loc[1][e.second]);
                for (int A = 0; A < nsa; A++) {
                        std::sort(lists[A].begin(), lists[A].end());
                        segptr[A] = (int32_t)segsec.size();
                        int64_t off = 0;
                        for (int B : lists[A]) {
                                segsec.push_back(B);
                                segoff.push_back(off);
                                s1bo[(size_t)A * nsb + B] = off;
                                off += g->nstr[nt->glob[1][B]];
                        }
                        nt->rowlen[A] = off;
                        nt->maxrow = std::max(nt->maxrow, off);
                        nt->cbase[A + 1] = nt->cbase[A] + off * g->nstr[nt->glob[0][A]];
                }
                segptr[nsa] = (int32_t)segsec.size();
        }
        nt->n1 = nt->cbase[nsa];
        // D blocks on local sectors
        std::vector<int32_t> dblk((size_t)nsa * nsb, -1), bsa(g->nblock), bsb(g->nblock);
        std::vector<int64_t> boff(g->nblock + 1), bnb(g->nblock);
        for (int b = 0; b < g->nblock; b++) {
                bsa[b] = loc[0][g->bsa[b]];
                bsb[b] = loc[1][g->bsb[b]];
                dblk[(size_t)bsa[b] * nsb + bsb[b]] = b;
                boff[b] = g->boff[b];
                bnb[b] = g->nstr[g->bsb[b]];
        }
        boff[g->nblock] = g->ndet;
        nt->bsa_loc = bsa;
        nt->bsb_loc = bsb;
        nt->dblk_loc = dblk;
        {
                std::set<int> da(bsa.begin(), bsa.end()), db(bsb.begin(), bsb.end());
                nt->dsec[0].assign(da.begin(), da.end());
                nt->dsec[1].assign(db.begin(), db.end());
        }
        // per spin: sector occupations, strides and the move table
        NatView &v = nt->v;
        v.n = n;
        v.npair = nt->npair;
        v.nsecA = nsa;
        v.nsecB = nsb;
        v.nblock = g->nblock;
        v.nstrA = nt->nstr[0];
        v.nstrB = nt->nstr[1];
        v.ndet = g->ndet;
        for (int spin = 0; spin < 2; spin++) {
                const int ns = nt->nsec[spin];
                std::vector<uint8_t> occ((size_t)ns * ng);
                std::vector<int64_t> stride((size_t)ns * ng);
                std::vector<int32_t> mv((size_t)ns * ng * ng, -1);
                for (int i = 0; i < ns; i++) {
                        const int s = nt->glob[spin][i];
                        for (int k = 0; k < ng; k++) {
                                occ[(size_t)i * ng + k] = g->occ[(size_t)s * ng + k];
                                stride[(size_t)i * ng + k] = g->stride[(size_t)s * ng + k];
                        }
                        for (int r = 0; r < ng; r++)
                                for (int t = 0; t < ng; t++) {
                                        if (r == t || pool[s][r] == 0 || pool[s][t] >= g->norb[t]) continue;
                                        std::vector<uint8_t> o = pool[s];
                                        o[r]--;
                                        o[t]++;
                                        auto it = std::lower_bound(pool.begin(), pool.end(), o);
                                        if (it != pool.end() && *it == o)
                                                mv[((size_t)i * ng + r) * ng + t] = loc[spin][(int)(it - pool.begin())];
                                }
                }
                const int64_t ns_str = nt->nstr[spin];
                int64_t *d_sbase;
                uint8_t *d_occ;
                int64_t *d_stride;
                int32_t *d_mv;
                uint64_t *d_str;
                uint16_t *d_ssec;
                uint32_t *d_saddr;
                int32_t *d_pmap, *d_lk;
                int16_t *d_lkp;
                const int ke = spin ? g->nb : g->na;
                const int nlink = ke * (n - ke + 1);
                NAT_TRY(dev_upload(nt, &d_sbase, nt->sbase[spin], st));
                NAT_TRY(dev_upload(nt, &d_occ, occ, st));
                NAT_TRY(dev_upload(nt, &d_stride, stride, st));
                NAT_TRY(dev_upload(nt, &d_mv, mv, st));
                nt->occ[spin] = occ;
                nt->stride[spin] = stride;
                nt->mv[spin] = mv;
                nt->d_stride[spin] = d_stride;
                nt->d_mv[spin] = d_mv;
                NAT_TRY(dev_alloc(nt, &d_str, ns_str));
                NAT_TRY(dev_alloc(nt, &d_ssec, ns_str));
                NAT_TRY(dev_alloc(nt, &d_saddr, ns_str));
                NAT_TRY(dev_alloc(nt, &d_pmap, (size_t)nt->npair * ns_str));
                NAT_TRY(dev_alloc(nt, &d_lk, (size_t)nlink * ns_str));
                NAT_TRY(dev_alloc(nt, &d_lkp, (size_t)nlink * ns_str));
                hipLaunchKernelGGL(k_nat_strings, dim3(blocks_for(ns_str)), dim3(NT), 0, st, nt->si, ns,
                                   ns_str, d_sbase, d_occ, d_stride, d_str, d_ssec, d_saddr);
                FCIS_CHECK_LAUNCH();
                hipLaunchKernelGGL(k_nat_fill, dim3(4096), dim3(NT), 0, st, d_pmap,
                                   (int64_t)nt->npair * ns_str, (int32_t)-1);
                FCIS_CHECK_LAUNCH();
                hipLaunchKernelGGL(k_nat_pmap, dim3(blocks_for(ns_str)), dim3(NT), 0, st, nt->si, n, ns_str,
                                   d_str, d_ssec, d_sbase, d_stride, d_mv, d_pmap, d_lk, d_lkp);
                FCIS_CHECK_LAUNCH();
                if (spin == 0) {
                        v.strA = d_str;
                        v.ssecA = d_ssec;
                        v.saddrA = d_saddr;
                        v.pmapA = d_pmap;
                        v.lkA = d_lk;
                        v.lkpA = d_lkp;
                        v.nlinkA = nlink;
                        v.sbaseA = d_sbase;
                } else {
                        v.strB = d_str;
                        v.ssecB = d_ssec;
                        v.saddrB = d_saddr;
                        v.pmapB = d_pmap;
                        v.lkB = d_lk;
                        v.lkpB = d_lkp;
                        v.nlinkB = nlink;
                        v.sbaseB = d_sbase;
                }
        }
        int64_t *d_cbase, *d_rowlen, *d_segoff, *d_s1bo, *d_boff, *d_bnb;
        int32_t *d_segptr, *d_segsec, *d_dblk, *d_bsa, *d_bsb;
        NAT_TRY(dev_upload(nt, &d_cbase, nt->cbase, st));
        NAT_TRY(dev_upload(nt, &d_rowlen, nt->rowlen, st));
        NAT_TRY(dev_upload(nt, &d_segptr, segptr, st));
        NAT_TRY(dev_upload(nt, &d_segsec, segsec, st));
        NAT_TRY(dev_upload(nt, &d_segoff, segoff, st));
        NAT_TRY(dev_upload(nt, &d_s1bo, s1bo, st));
        NAT_TRY(dev_upload(nt, &d_dblk, dblk, st));
        NAT_TRY(dev_upload(nt, &d_boff, boff, st));
        NAT_TRY(dev_upload(nt, &d_bsa, bsa, st));
        NAT_TRY(dev_upload(nt, &d_bsb, bsb, st));
        NAT_TRY(dev_upload(nt, &d_bnb, bnb, st));
        v.cbase = d_cbase;
        v.rowlen = d_rowlen;
        v.segptr = d_segptr;
        v.segsec = d_segsec;
        v.segoff = d_segoff;
        v.s1bo = d_s1bo;
        v.dblk = d_dblk;
        v.boff = d_boff;
        v.bsa = d_bsa;
        v.bsb = d_bsb;
        v.bnb = d_bnb;
        FCIS_HIP(hipStreamSynchronize(st));
        return FCIS_SUCCESS;
}

// Row tiles: consecutive S1 rows whose columns fit `cap`.
void make_tiles(const gas_native *nt, int64_t cap, std::vector<int64_t> &r0, std::vector<int64_t> &c0)
{
        r0.clear();
        c0.clear();
        r0.push_back(0);
        c0.push_back(0);
        int64_t cur = 0;
        for (int A = 0; A < nt->nsec[0]; A++) {
                const int64_t L = nt->rowlen[A];
                const int64_t nrow = nt->sbase[0][A + 1] - nt->sbase[0][A];
                int64_t done = 0;
                while (done < nrow) {
                        const int64_t room = (cap - (nt->cbase[A] + done * L - c0.back())) / L;
                        if (room <= 0) {
                                r0.push_back(nt->sbase[0][A] + done);
                                c0.push_back(nt->cbase[A] + done * L);
                                continue;
                        }
                        done += std::min(room, nrow - done);
                }
                cur = nt->cbase[A + 1];
        }
        (void)cur;
        r0.push_back(nt->nstr[0]);
        c0.push_back(nt->n1);
}

// Local sector reached from local sector S (spin `spin`) by moving one
// electron from subspace x to subspace y (S itself for x == y), or -1.
int host_move(const gas_native *nt, int spin, int S, int x, int y)
{
        if (S < 0) return -1;
        const int ng = nt->si.ngas;
        return x == y ? S : nt->mv[spin][((size_t)S * ng + x) * ng + y];
}

// Segment tables of the operator form for both spins (see the device
// comment): per D sector S the number of strings of each D sector T within
// two replacements of a string of S (operator rows) and within one
// (link lists, diagonal slots included), counted from the occupations:
// o[x] electrons and w[x] = norb[x] - o[x] holes in subspace x.
void op_host(const fcis_gas_space *g, gas_native *nt)
{
        if (nt->op_host_done) return;
        const int ng = nt->si.ngas;
        for (int spin = 0; spin < 2; spin++) {
                auto &H = nt->oph[spin];
                const int ns = nt->nsec[spin];
                const std::vector<int32_t> &ds = nt->dsec[spin];
                const int nd = (int)ds.size();
                H.dindex.assign(ns, -1);
                for (int i = 0; i < nd; i++) H.dindex[ds[i]] = i;
                H.rowbase.assign(nd + 1, 0);
                H.opbase.assign(nd, 0);
                H.lsbase.assign(nd, 0);
                H.oplen.assign(nd, 0);
                H.lslen.assign(nd, 0);
                H.opsegptr.assign(nd + 1, 0);
                H.lssegptr.assign(nd + 1, 0);
                H.opsegk.assign((size_t)nd * ns, -1);
                H.lssegk.assign((size_t)nd * ns, -1);
                H.opsegT.clear();
                H.opsegoff.clear();
                H.lssegT.clear();
                H.lssegoff.clear();
                H.nnz = H.nls = 0;
                H.maxseg = 0;
                for (int i = 0; i < nd; i++) {
                        const int S = ds[i];
                        const int64_t nstr = nt->sbase[spin][S + 1] - nt->sbase[spin][S];
                        H.rowbase[i + 1] = H.rowbase[i] + nstr;
                        std::vector<int64_t> o(ng), w(ng);
                        int k = 0;
                        for (int x = 0; x < ng; x++) {
                                o[x] = nt->occ[spin][(size_t)S * ng + x];
                                w[x] = nt->si.norb[x] - o[x];
                                k += (int)o[x];
                        }
                        std::map<int, int64_t> cop, cls;
                        cop[S] += 1;
                        cls[S] += k;
                        for (int x = 0; x < ng; x++)
                                for (int y = 0; y < ng; y++) {
                                        const int64_t m = o[x] * w[y];
                                        const int T = host_move(nt, spin, S, x, y);
                                        if (m == 0 || T < 0 || H.dindex[T] < 0) continue;
                                        cop[T] += m;
                                        cls[T] += m;
                                }
                        // doubles: removed electrons from x1 <= x2, added to y1 <= y2
                        for (int x1 = 0; x1 < ng; x1++)
                                for (int x2 = x1; x2 < ng; x2++) {
                                        const int64_t mo = x1 == x2 ? o[x1] * (o[x1] - 1) / 2 : o[x1] * o[x2];
                                        if (mo == 0) continue;
                                        for (int y1 = 0; y1 < ng; y1++)
                                                for (int y2 = y1; y2 < ng; y2++) {
                                                        const int64_t mw = y1 == y2 ? w[y1] * (w[y1] - 1) / 2
                                                                                    : w[y1] * w[y2];
                                                        if (mw == 0) continue;
                                                        const int T = host_move(nt, spin, host_move(nt, spin, S, x1, y1),
                                                                                x2, y2);
                                                        if (T < 0 || H.dindex[T] < 0) continue;
                                                        cop[T] += mo * mw;
                                                }
                                }
                        H.opsegptr[i] = (int32_t)H.opsegT.size();
                        int64_t off = 0;
                        for (auto &e : cop) {
                                H.opsegk[(size_t)i * ns + e.first] = (int16_t)(H.opsegT.size() - H.opsegptr[i]);
                                H.opsegT.push_back(e.first);
                                H.opsegoff.push_back((int32_t)off);
                                off += e.second;
                        }
                        H.oplen[i] = (int32_t)off;
                        H.opbase[i] = H.nnz;
                        H.nnz += nstr * off;
                        H.maxseg = std::max<int>(H.maxseg, (int)cop.size());
                        H.lssegptr[i] = (int32_t)H.lssegT.size();
                        off = 0;
                        for (auto &e : cls) {
                                H.lssegk[(size_t)i * ns + e.first] = (int16_t)(H.lssegT.size() - H.lssegptr[i]);
                                H.lssegT.push_back(e.first);
                                H.lssegoff.push_back((int32_t)off);
                                off += e.second;
                        }
                        H.lslen[i] = (int32_t)off;
                        H.lsbase[i] = H.nls;
                        H.nls += nstr * off;
                        H.maxseg = std::max<int>(H.maxseg, (int)cls.size());
                }
                H.opsegptr[nd] = (int32_t)H.opsegT.size();
                H.lssegptr[nd] = (int32_t)H.lssegT.size();
        }
        // work estimate: wavefront iterations of the sigma kernels, same-spin
        // entries and opposite-spin link pairs
        const auto &HA = nt->oph[0], &HB = nt->oph[1];
        const int nsb = nt->nsec[1];
        double same = 0.0, cross = 0.0;
        auto seglen = [](const std::vector<int32_t> &off, int k, int k1, int32_t len) {
                return (double)((k + 1 < k1 ? off[k + 1] : len) - off[k]);
        };
        for (int b = 0; b < g->nblock; b++) {
                const int A = nt->bsa_loc[b], B = nt->bsb_loc[b];
                const int ia = HA.dindex[A], ib = HB.dindex[B];
                const double na = (double)(nt->sbase[0][A + 1] - nt->sbase[0][A]);
                const double nb = (double)(nt->sbase[1][B + 1] - nt->sbase[1][B]);
                double aa = 0.0, ab = 0.0, bb = 0.0;
                for (int k = HA.opsegptr[ia]; k < HA.opsegptr[ia + 1]; k++)
                        if (nt->dblk_loc[(size_t)HA.opsegT[k] * nsb + B] >= 0)
                                aa += seglen(HA.opsegoff, k, HA.opsegptr[ia + 1], HA.oplen[ia]);
                for (int k = HB.opsegptr[ib]; k < HB.opsegptr[ib + 1]; k++)
                        if (nt->dblk_loc[(size_t)A * nsb + HB.opsegT[k]] >= 0)
                                bb += seglen(HB.opsegoff, k, HB.opsegptr[ib + 1], HB.oplen[ib]);
                for (int ka = HA.lssegptr[ia]; ka < HA.lssegptr[ia + 1]; ka++)
                        for (int kb = HB.lssegptr[ib]; kb < HB.lssegptr[ib + 1]; kb++)
                                if (nt->dblk_loc[(size_t)HA.lssegT[ka] * nsb + HB.lssegT[kb]] >= 0)
                                        ab += seglen(HA.lssegoff, ka, HA.lssegptr[ia + 1], HA.lslen[ia]) *
                                              seglen(HB.lssegoff, kb, HB.lssegptr[ib + 1], HB.lslen[ib]);
                const double wa = na * std::ceil(nb / 64.0), wb = nb * std::ceil(na / 64.0);
                same += wa * aa + wb * bb;
                cross += wa * ab;
        }
        nt->op_same = same;
        nt->op_cross = cross;
        nt->op_host_done = true;
}

// Device tables of the operator form for both spins (structure only; the
// entries of A come with each plan).  On failure the operator form is
// marked unusable and the status returned.
int op_device(const fcis_gas_space *g, gas_native *nt, hipStream_t st)
{
        if (nt->op_dev_state == 1) return FCIS_SUCCESS;
        if (nt->op_dev_state < 0) {
                fcis::set_error("sector-native GAS space: operator tables unavailable");
                return FCIS_ERR_INVALID;
        }
        op_host(g, nt);
        nt->op_dev_state = -1;
        for (int spin = 0; spin < 2; spin++)
                if (nt->oph[spin].maxseg > OP_MAXSEG) {
                        fcis::set_error("sector-native GAS space: %d target sectors per sector exceed %d",
                                        nt->oph[spin].maxseg, OP_MAXSEG);
                        return FCIS_ERR_INVALID;
                }
        int *d_err;
        NAT_TRY(dev_alloc(nt, &d_err, 1));
        FCIS_HIP(hipMemsetAsync(d_err, 0, sizeof(int), st));
        for (int spin = 0; spin < 2; spin++) {
                const auto &H = nt->oph[spin];
                OpView &o = nt->op[spin];
                o = OpView{};
                o.ndsec = (int)nt->dsec[spin].size();
                o.nrows = H.rowbase[o.ndsec];
                int32_t *p32;
                int64_t *p64;
                int16_t *p16;
                NAT_TRY(dev_upload(nt, &p32, nt->dsec[spin], st));
                o.dsec = p32;
                NAT_TRY(dev_upload(nt, &p32, H.dindex, st));
                o.dindex = p32;
                NAT_TRY(dev_upload(nt, &p64, H.rowbase, st));
                o.rowbase = p64;
                NAT_TRY(dev_upload(nt, &p64, H.opbase, st));
                o.opbase = p64;
                NAT_TRY(dev_upload(nt, &p32, H.oplen, st));
                o.oplen = p32;
                NAT_TRY(dev_upload(nt, &p32, H.opsegptr, st));
                o.opsegptr = p32;
                NAT_TRY(dev_upload(nt, &p32, H.opsegT, st));
                o.opsegT = p32;
                NAT_TRY(dev_upload(nt, &p32, H.opsegoff, st));
                o.opsegoff = p32;
                NAT_TRY(dev_upload(nt, &p16, H.opsegk, st));
                o.opsegk = p16;
                NAT_TRY(dev_upload(nt, &p64, H.lsbase, st));
                o.lsbase = p64;
                NAT_TRY(dev_upload(nt, &p32, H.lslen, st));
                o.lslen = p32;
                NAT_TRY(dev_upload(nt, &p32, H.lssegptr, st));
                o.lssegptr = p32;
                NAT_TRY(dev_upload(nt, &p32, H.lssegT, st));
                o.lssegT = p32;
                NAT_TRY(dev_upload(nt, &p32, H.lssegoff, st));
                o.lssegoff = p32;
                NAT_TRY(dev_upload(nt, &p16, H.lssegk, st));
                o.lssegk = p16;
                uint32_t *addr;
                uint64_t *ls;
                NAT_TRY(dev_alloc(nt, &addr, (size_t)H.nnz));
                NAT_TRY(dev_alloc(nt, &ls, (size_t)H.nls));
                o.opaddr = addr;
                o.ls = ls;
                const dim3 grid((unsigned)((o.nrows + NT / 64 - 1) / (NT / 64)));
                hipLaunchKernelGGL(k_op_struct<true>, grid, dim3(NT), 0, st, nt->v, o, spin, 0, nt->si,
                                   nt->d_stride[spin], nt->d_mv[spin], addr, ls, d_err);
                FCIS_CHECK_LAUNCH();
                hipLaunchKernelGGL(k_op_struct<false>, grid, dim3(NT), 0, st, nt->v, o, spin, spin == 0 ? 1 : 0,
                                   nt->si, nt->d_stride[spin], nt->d_mv[spin], addr, ls, d_err);
                FCIS_CHECK_LAUNCH();
        }
        int herr = 0;
        FCIS_HIP(hipMemcpyAsync(&herr, d_err, sizeof(int), hipMemcpyDeviceToHost, st));
        FCIS_HIP(hipStreamSynchronize(st));
        if (herr != 0) {
                fcis::set_error("sector-native GAS space: operator tables inconsistent (%d segments)", herr);
                return FCIS_ERR_DEVICE;
        }
        nt->op_dev_state = 1;
        return FCIS_SUCCESS;
}

// Bytes of the operator form's device tables (structure, both spins) and
// of a plan's entries.
uint64_t op_struct_bytes(const gas_native *nt)
{
        return (uint64_t)(nt->oph[0].nnz + nt->oph[1].nnz) * 4 + (uint64_t)(nt->oph[0].nls + nt->oph[1].nls) * 8;
}

uint64_t op_coef_bytes(const gas_native *nt)
{
        return (uint64_t)(nt->oph[0].nnz + nt->oph[1].nnz) * 8;
}

// Column capacity of a tile for `per_col` bytes per column within `budget`.
int64_t tile_columns(const gas_native *nt, uint64_t budget, uint64_t per_col)
{
        int64_t cols = (int64_t)(budget / per_col);
        return std::min(cols, nt->n1);
}

struct Blas {
        rocblas_handle h = nullptr;
        ~Blas()
        {
                if (h) rocblas_destroy_handle(h);
        }
};

struct DevBuf {
        double *p = nullptr;
        hipStream_t st = nullptr;
        ~DevBuf()
        {
                if (p) {
                        (void)hipStreamSynchronize(st);
                        (void)hipFree(p);
                }
        }
};

}  // namespace

int fcis::gas_native_create(fcis_gas_space *g, const fcis::SectorPool &pool, hipStream_t st)
{
        gas_native *nt = new (std::nothrow) gas_native();
        if (nt == nullptr) return FCIS_ERR_MEMORY;
        g->nat = nt;
        g->ncas = 0;
        g->d_map = nullptr;
        g->d_rows = g->d_rpos = nullptr;
        g->d_rowcl = nullptr;
        g->d_clist = nullptr;
        return build_native(g, nt, pool, st);
}

void fcis::gas_native_free(gas_native *nt)
{
        if (nt == nullptr) return;
        for (void *p : nt->dev) (void)hipFree(p);
        delete nt;
}

int fcis::gas_native_plan_path(const gas_native_plan *p)
{
        return p->direct == 2 ? 3 : p->direct == 1 ? 4 : 2;
}

int fcis::gas_native_info(const fcis_gas_space *g, int64_t *nclosure, int64_t *maxrow, int64_t *nstra,
                          int64_t *nstrb)
{
        const gas_native *nt = g->nat;
        if (nclosure) *nclosure = nt ? nt->n1 : 0;
        if (maxrow) *maxrow = nt ? nt->maxrow : 0;
        if (nstra) *nstra = nt ? nt->nstr[0] : 0;
        if (nstrb) *nstrb = nt ? nt->nstr[1] : 0;
        return FCIS_SUCCESS;
}

int fcis::gas_native_hdiag(const fcis_gas_space *g, const double *h1e, const double *eri,
                           double *d_hdiag, hipStream_t st)
{
        const gas_native *nt = g->nat;
        const int n = nt->n, np = nt->npair;
        std::vector<double> hjk((size_t)n + 2 * (size_t)n * n);
        for (int p = 0; p < n; p++) {
                hjk[p] = h1e[p * n + p];
                for (int q = 0; q < n; q++) {
                        hjk[n + p * n + q] = eri[(size_t)fcis::pair_index(p, p) * np + fcis::pair_index(q, q)];
                        hjk[n + n * n + p * n + q] =
                                eri[(size_t)fcis::pair_index(p, q) * np + fcis::pair_index(q, p)];
                }
        }
        DevBuf d;
        d.st = st;
        if (fcis::dev_malloc(&d.p, 8 * hjk.size()) != hipSuccess) {
                fcis::set_error("fcis_gas_make_hdiag: device allocation failed");
                return FCIS_ERR_MEMORY;
        }
        FCIS_HIP(hipMemcpyAsync(d.p, hjk.data(), 8 * hjk.size(), hipMemcpyHostToDevice, st));
        if (g->ndet > 0) {
                hipLaunchKernelGGL(k_nat_hdiag, dim3(blocks_for(g->ndet)), dim3(NT), 0, st, nt->v, d.p,
                                   d_hdiag);
                FCIS_CHECK_LAUNCH();
        }
        return FCIS_SUCCESS;
}

int fcis::gas_native_ss(const fcis_gas_space *g, const double *d_ci0, double *d_ci1, hipStream_t st)
{
        const double sz = 0.5 * (g->na - g->nb);
        if (g->ndet > 0) {
                if (d_ci0 != nullptr)
                        hipLaunchKernelGGL(k_nat_ss, dim3(blocks_for(g->ndet)), dim3(NT), 0, st, g->nat->v,
                                           sz * (sz + 1.0) + g->nb, d_ci0, d_ci1);
                else
                        hipLaunchKernelGGL(k_nat_ss_diag, dim3(blocks_for(g->ndet)), dim3(NT), 0, st,
                                           g->nat->v, sz * (sz + 1.0) + g->nb, d_ci1);
                FCIS_CHECK_LAUNCH();
        }
        return FCIS_SUCCESS;
}

void fcis::gas_native_plan_free(gas_native_plan *p)
{
        if (p == nullptr) return;
        if (p->blas) rocblas_destroy_handle(p->blas);
        if (p->d_h) (void)hipFree(p->d_h);
        if (p->d_t1) fcis::big_free(p->d_t1);
        if (p->d_wbase) (void)hipFree(p->d_wbase);
        if (p->d_ct) (void)hipFree(p->d_ct);
        if (p->d_opcoef) (void)hipFree(p->d_opcoef);
        delete p;
}

int fcis::gas_native_plan_create(gas_native_plan **out, const fcis_gas_space *g, const double *h2e,
                                 int precision, uint64_t workspace_bytes)
{
        *out = nullptr;
        if (precision != FCIS_PREC_FP64 && precision != FCIS_PREC_FP32) {
                fcis::set_error("fcis_gas_sigma_plan_create: a sector-native space computes in FP64 "
                                "or FP32");
                return FCIS_ERR_INVALID;
        }
        const gas_native *nt = g->nat;
        gas_native_plan *p = new (std::nothrow) gas_native_plan();
        if (p == nullptr) return FCIS_ERR_MEMORY;
        p->g = g;
        p->precision = precision;
        const size_t es = precision == FCIS_PREC_FP64 ? 8 : 4;
        const int np = nt->npair;
        if (workspace_bytes == 0) workspace_bytes = (uint64_t)8 << 30;
        // Path, by estimated time in ns (FP64 sigma; the direct forms compute
        // in FP64 for either precision), each model fitted to measured sigma
        // times at 16-62 orbitals: the t1 / GEMM / pull form by its GEMM,
        // npair^2 + 60 npair per closure column at about 28 per ns; the
        // link-pair form by its k(n-k+1)-squared slot visits per
        // determinant; the operator form by its wavefront iterations (one per
        // same-spin entry and one per opposite-spin link pair), about 0.08
        // ns each, plus launches.  The operator form needs its tables (about
        // 12 bytes per entry and 8 per sorted link) within the workspace.
        // Measured (ms, operator / link-pair / t1): RAS(2,8,30) (4,4) 1.5M
        // determinants 11.3 / 142 / 1134; GAS 62 orbitals 1.35M 18.1 / 368 /
        // 10284; RAS(4,12,4) (6,6) 7.0M 26.6 / 251 / 164.
        // FCIS_GAS_NATIVE_PATH=direct / links / t1 forces one.
        {
                gas_native *nm = g->nat;
                op_host(g, nm);
                const double la = nt->v.nlinkA, lb = nt->v.nlinkB;
                const double t1_ns = (double)nt->n1 * ((double)np * np + 60.0 * np) * 1.0e-3 / 28.0;
                const double links_ns = (double)g->ndet * (la * la + lb * lb + 2.0 * la * lb) * 2.0 * 16.0 * 1.0e-3 / 28.0;
                const double op_ns = 0.08 * (nm->op_same + nm->op_cross) + 5.0e4;
                const bool op_fits = op_struct_bytes(nm) + op_coef_bytes(nm) <= workspace_bytes &&
                                     nm->oph[0].maxseg <= OP_MAXSEG && nm->oph[1].maxseg <= OP_MAXSEG &&
                                     nm->op_dev_state >= 0;
                p->direct = 0;
                double best = t1_ns;
                if (links_ns < best) {
                        p->direct = 1;
                        best = links_ns;
                }
                if (op_fits && op_ns < best) p->direct = 2;
                const char *env = getenv("FCIS_GAS_NATIVE_PATH");
                if (env != nullptr && strcmp(env, "direct") == 0) p->direct = 2;
                if (env != nullptr && strcmp(env, "links") == 0) p->direct = 1;
                if (env != nullptr && strcmp(env, "t1") == 0) p->direct = 0;
                if (p->direct == 2) {
                        const int r = op_device(g, nm, nullptr);
                        if (r != FCIS_SUCCESS) {
                                gas_native_plan_free(p);
                                return r;
                        }
                }
        }
        if (p->direct) {
                std::vector<int64_t> wb(2 * (size_t)(g->nblock + 1), 0);
                for (int b = 0; b < g->nblock; b++) {
                        const int64_t nas = g->nstr[g->bsa[b]], nbs = g->nstr[g->bsb[b]];
                        // operator form: several strings per wavefront on narrow blocks
                        wb[b + 1] = wb[b] + (p->direct == 2 ? op_waves(nas, nbs, p->packmax) : nas * ((nbs + 63) / 64));
                        wb[g->nblock + 1 + b + 1] = wb[g->nblock + 1 + b] +
                                                    (p->direct == 2 ? op_waves(nbs, nas, p->packmax) : nbs * ((nas + 63) / 64));
                }
                p->nwave = wb[g->nblock];
                p->nwave_bb = wb[2 * (size_t)g->nblock + 1];
                std::vector<double> hh((size_t)2 * np * np);
                for (int i = 0; i < np; i++)
                        for (int j = 0; j < np; j++) {
                                hh[(size_t)i * np + j] = h2e[(size_t)i * np + j];
                                hh[(size_t)np * np + (size_t)i * np + j] = h2e[(size_t)i * np + j] + h2e[(size_t)j * np + i];
                        }
                if (fcis::dev_malloc(&p->d_h, 8 * hh.size()) != hipSuccess ||
                    fcis::dev_malloc(&p->d_wbase, 8 * wb.size()) != hipSuccess ||
                    fcis::dev_malloc(&p->d_ct, 16 * (size_t)std::max<int64_t>(g->ndet, 1)) != hipSuccess) {
                        gas_native_plan_free(p);
                        fcis::set_error("fcis_gas_sigma_plan_create: device allocation failed");
                        return FCIS_ERR_MEMORY;
                }
                if (hipMemcpy(p->d_h, hh.data(), 8 * hh.size(), hipMemcpyHostToDevice) != hipSuccess ||
                    hipMemcpy(p->d_wbase, wb.data(), 8 * wb.size(), hipMemcpyHostToDevice) != hipSuccess) {
                        gas_native_plan_free(p);
                        fcis::set_error("fcis_gas_sigma_plan_create: upload failed");
                        return FCIS_ERR_DEVICE;
                }
                if (p->direct == 2) {
                        const gas_native *nm = g->nat;
                        const int64_t nza = nm->oph[0].nnz, nzb = nm->oph[1].nnz;
                        if (fcis::dev_malloc(&p->d_opcoef, 8 * (size_t)std::max<int64_t>(nza + nzb, 1)) != hipSuccess) {
                                gas_native_plan_free(p);
                                fcis::set_error("fcis_gas_sigma_plan_create: device allocation of the operator "
                                                "entries failed");
                                return FCIS_ERR_MEMORY;
                        }
                        p->oa = nm->op[0];
                        p->ob = nm->op[1];
                        p->oa.opcoef = p->d_opcoef;
                        p->ob.opcoef = p->d_opcoef + nza;
                        for (int spin = 0; spin < 2; spin++) {
                                const OpView &o = spin ? p->ob : p->oa;
                                hipLaunchKernelGGL(k_op_coef, dim3((unsigned)((o.nrows + NT / 64 - 1) / (NT / 64))),
                                                   dim3(NT), 0, nullptr, nm->v, o, spin,
                                                   static_cast<const double *>(p->d_h), const_cast<double *>(o.opcoef));
                        }
                        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(nullptr) != hipSuccess) {
                                gas_native_plan_free(p);
                                fcis::set_error("fcis_gas_sigma_plan_create: operator entries failed");
                                return FCIS_ERR_DEVICE;
                        }
                }
                *out = p;
                return FCIS_SUCCESS;
        }
        int64_t cols = tile_columns(nt, workspace_bytes, 2 * es * (uint64_t)np);
        if (cols < nt->maxrow) {
                gas_native_plan_free(p);
                fcis::set_error("fcis_gas_sigma_plan_create: a workspace of %llu bytes holds %lld "
                                "columns, one closure row has %lld",
                                (unsigned long long)workspace_bytes, (long long)cols,
                                (long long)nt->maxrow);
                return FCIS_ERR_MEMORY;
        }
        cols = std::min<int64_t>(cols, (int64_t)(GEMM_MAX_SPAN / (uint64_t)np));
        if (cols < nt->maxrow) {
                gas_native_plan_free(p);
                fcis::set_error("fcis_gas_sigma_plan_create: one closure row (%lld columns) exceeds the "
                                "GEMM tile limit of %lld columns", (long long)nt->maxrow,
                                (long long)(GEMM_MAX_SPAN / (uint64_t)np));
                return FCIS_ERR_INVALID;
        }
        cols = std::max<int64_t>(cols, 1);
        p->ld = cols;
        make_tiles(nt, cols, p->r0, p->c0);
        if (fcis::big_malloc(&p->d_t1, 2 * es * (size_t)np * cols) != hipSuccess ||
            fcis::dev_malloc(&p->d_h, es * (size_t)np * np) != hipSuccess) {
                gas_native_plan_free(p);
                fcis::set_error("fcis_gas_sigma_plan_create: device allocation failed");
                return FCIS_ERR_MEMORY;
        }
        p->d_vt1 = static_cast<char *>(p->d_t1) + es * (size_t)np * cols;
        if (precision == FCIS_PREC_FP64) {
                if (hipMemcpy(p->d_h, h2e, 8 * (size_t)np * np, hipMemcpyHostToDevice) != hipSuccess) {
                        gas_native_plan_free(p);
                        fcis::set_error("fcis_gas_sigma_plan_create: upload failed");
                        return FCIS_ERR_DEVICE;
                }
        } else {
                std::vector<float> hf(h2e, h2e + (size_t)np * np);
                if (hipMemcpy(p->d_h, hf.data(), 4 * hf.size(), hipMemcpyHostToDevice) != hipSuccess) {
                        gas_native_plan_free(p);
                        fcis::set_error("fcis_gas_sigma_plan_create: upload failed");
                        return FCIS_ERR_DEVICE;
                }
        }
        if (rocblas_create_handle(&p->blas) != rocblas_status_success) {
                p->blas = nullptr;
                gas_native_plan_free(p);
                fcis::set_error("fcis_gas_sigma_plan_create: rocblas_create_handle failed");
                return FCIS_ERR_DEVICE;
        }
        *out = p;
        return FCIS_SUCCESS;
}

namespace {

template <typename T>
int native_execute(gas_native_plan *p, const double *ci0, double *ci1, hipStream_t st)
{
        const fcis_gas_space *g = p->g;
        const gas_native *nt = g->nat;
        const int np = nt->npair;
        T *t1 = static_cast<T *>(p->d_t1), *vt1 = static_cast<T *>(p->d_vt1);
        const T one = 1, zero = 0;
        if (rocblas_set_stream(p->blas, st) != rocblas_status_success) {
                fcis::set_error("fcis_gas_sigma_plan_execute: rocblas_set_stream failed");
                return FCIS_ERR_DEVICE;
        }
        const int ntile = (int)p->r0.size() - 1;
        for (int t = 0; t < ntile; t++) {
                const int64_t col0 = p->c0[t], ncol = p->c0[t + 1] - col0;
                if (ncol > 0) {
                        hipLaunchKernelGGL(k_nat_gather<T>, dim3(blocks_for(ncol)), dim3(NT), 0, st, nt->v,
                                           col0, ncol, p->ld, ci0, t1);
                        FCIS_CHECK_LAUNCH();
                        rocblas_status r;
                        if (sizeof(T) == 8)
                                r = rocblas_dgemm(p->blas, rocblas_operation_none, rocblas_operation_none,
                                                  (rocblas_int)ncol, np, np, (const double *)&one,
                                                  (const double *)t1, (rocblas_int)p->ld,
                                                  (const double *)p->d_h, np, (const double *)&zero,
                                                  (double *)vt1, (rocblas_int)p->ld);
                        else
                                r = rocblas_sgemm(p->blas, rocblas_operation_none, rocblas_operation_none,
                                                  (rocblas_int)ncol, np, np, (const float *)&one,
                                                  (const float *)t1, (rocblas_int)p->ld,
                                                  (const float *)p->d_h, np, (const float *)&zero,
                                                  (float *)vt1, (rocblas_int)p->ld);
                        if (r != rocblas_status_success) {
                                fcis::set_error("fcis_gas_sigma_plan_execute: rocBLAS GEMM failed (status %d)",
                                                (int)r);
                                return FCIS_ERR_DEVICE;
                        }
                }
                if (g->ndet > 0) {
                        hipLaunchKernelGGL(k_nat_pull<T>, dim3(blocks_for(g->ndet)), dim3(NT), 0, st, nt->v,
                                           p->r0[t], p->r0[t + 1], col0, p->ld, t > 0 ? 1 : 0, vt1, ci1);
                        FCIS_CHECK_LAUNCH();
                }
        }
        return FCIS_SUCCESS;
}

}  // namespace

int fcis::gas_native_plan_execute(gas_native_plan *p, const double *d_ci0, double *d_ci1, hipStream_t st)
{
        if (p->direct) {
                if (p->nwave > 0) {
                        const fcis_gas_space *g = p->g;
                        const int64_t np = g->nat->npair;
                        const double *h = static_cast<const double *>(p->d_h);
                        double *ct = p->d_ct, *sbb = p->d_ct + g->ndet;
                        const int64_t *wbt = p->d_wbase + g->nblock + 1;
                        hipLaunchKernelGGL(k_nat_transpose, dim3(blocks_for(g->ndet)), dim3(NT), 0, st, g->nat->v,
                                           d_ci0, ct);
                        if (p->direct == 2) {
                                hipLaunchKernelGGL(k_nat_op_bb,
                                                   dim3((unsigned)((p->nwave_bb + NT / 64 - 1) / (NT / 64))), dim3(NT),
                                                   0, st, g->nat->v, p->ob, wbt, p->packmax, ct, sbb);
                                hipLaunchKernelGGL(k_nat_op_direct,
                                                   dim3((unsigned)((p->nwave + NT / 64 - 1) / (NT / 64))), dim3(NT), 0,
                                                   st, g->nat->v, p->oa, p->ob, p->d_wbase, p->packmax, h + np * np,
                                                   d_ci0, sbb, d_ci1);
                                FCIS_CHECK_LAUNCH();
                                return FCIS_SUCCESS;
                        }
                        hipLaunchKernelGGL(k_nat_direct_bb,
                                           dim3((unsigned)((p->nwave_bb + NT / 64 - 1) / (NT / 64))), dim3(NT), 0,
                                           st, g->nat->v, wbt, h, ct, sbb);
                        hipLaunchKernelGGL(k_nat_direct, dim3((unsigned)((p->nwave + NT / 64 - 1) / (NT / 64))),
                                           dim3(NT), 0, st, g->nat->v, p->d_wbase, h, h + np * np, d_ci0, sbb,
                                           d_ci1);
                        FCIS_CHECK_LAUNCH();
                }
                return FCIS_SUCCESS;
        }
        if (p->ld > INT32_MAX) {
                fcis::set_error("fcis_gas_sigma_plan_execute: tile of more than 2^31 columns");
                return FCIS_ERR_INVALID;
        }
        return p->precision == FCIS_PREC_FP64 ? native_execute<double>(p, d_ci0, d_ci1, st)
                                              : native_execute<float>(p, d_ci0, d_ci1, st);
}

int fcis::gas_native_rdm(const fcis_gas_space *g, const double *d_bra, const double *d_ket,
                         double *dm1a, double *dm1b, double *dm2aa, double *dm2ab, double *dm2bb,
                         hipStream_t st)
{
        const gas_native *nt = g->nat;
        const int n = nt->n, n2 = n * n;
        const bool two = dm2aa != nullptr || dm2ab != nullptr || dm2bb != nullptr;
        const bool same = d_bra == d_ket;
        // device: y[2][n2] (dm1 rows), M[3][n2][n2] (aa, ab, bb), then the
        // tile: D matrices (ket a, ket b, bra a, bra b) and the embedded bra
        const int nmat = two ? (same ? 2 : 4) : 2;
        const size_t fixed = 2 * (size_t)n2 + (two ? 3 * (size_t)n2 * n2 : 0);
        size_t freeb = 0, totb = 0;
        FCIS_HIP(hipMemGetInfo(&freeb, &totb));
        const uint64_t budget = std::min<uint64_t>((uint64_t)freeb / 2, (uint64_t)16 << 30);
        const uint64_t per_col = 8 * ((uint64_t)nmat * n2 + 1);
        int64_t cols = tile_columns(nt, budget > 8 * fixed ? budget - 8 * fixed : 0, per_col);
        if (cols < nt->maxrow) cols = nt->maxrow;
        cols = std::max<int64_t>(cols, 1);
        if (cols > INT32_MAX) cols = INT32_MAX;
        std::vector<int64_t> r0, c0;
        make_tiles(nt, cols, r0, c0);
        DevBuf buf;
        buf.st = st;
        if (fcis::dev_malloc(&buf.p, 8 * (fixed + (size_t)cols * ((size_t)nmat * n2 + 1))) != hipSuccess) {
                fcis::set_error("fcis_gas_trans_rdm12s: device allocation failed");
                return FCIS_ERR_MEMORY;
        }
        double *y = buf.p, *M = y + 2 * (size_t)n2;
        double *Dk[2], *Db[2], *e;
        double *tile = buf.p + fixed;
        Dk[0] = tile;
        Dk[1] = tile + (size_t)n2 * cols;
        if (two && !same) {
                Db[0] = tile + 2 * (size_t)n2 * cols;
                Db[1] = tile + 3 * (size_t)n2 * cols;
        } else {
                Db[0] = Dk[0];
                Db[1] = Dk[1];
        }
        e = tile + (size_t)nmat * n2 * cols;
        FCIS_HIP(hipMemsetAsync(buf.p, 0, 8 * fixed, st));
        Blas blas;
        if (rocblas_create_handle(&blas.h) != rocblas_status_success) {
                blas.h = nullptr;
                fcis::set_error("fcis_gas_trans_rdm12s: rocblas_create_handle failed");
                return FCIS_ERR_DEVICE;
        }
        if (rocblas_set_stream(blas.h, st) != rocblas_status_success) {
                fcis::set_error("fcis_gas_trans_rdm12s: rocblas_set_stream failed");
                return FCIS_ERR_DEVICE;
        }
        const double one = 1.0;
        for (size_t t = 0; t + 1 < r0.size(); t++) {
                const int64_t col0 = c0[t], ncol = c0[t + 1] - col0;
                if (ncol <= 0) continue;
                const unsigned gb = blocks_for(ncol);
                hipLaunchKernelGGL(k_nat_rdm_gather<0>, dim3(gb), dim3(NT), 0, st, nt->v, col0, ncol, cols,
                                   d_ket, Dk[0]);
                hipLaunchKernelGGL(k_nat_rdm_gather<1>, dim3(gb), dim3(NT), 0, st, nt->v, col0, ncol, cols,
                                   d_ket, Dk[1]);
                if (two && !same) {
                        hipLaunchKernelGGL(k_nat_rdm_gather<0>, dim3(gb), dim3(NT), 0, st, nt->v, col0, ncol,
                                           cols, d_bra, Db[0]);
                        hipLaunchKernelGGL(k_nat_rdm_gather<1>, dim3(gb), dim3(NT), 0, st, nt->v, col0, ncol,
                                           cols, d_bra, Db[1]);
                }
                hipLaunchKernelGGL(k_nat_s1_embed, dim3(gb), dim3(NT), 0, st, nt->v, col0, ncol, d_bra, e);
                FCIS_CHECK_LAUNCH();
                rocblas_status r = rocblas_status_success;
                for (int s = 0; s < 2 && r == rocblas_status_success; s++)
                        r = rocblas_dgemv(blas.h, rocblas_operation_transpose, (rocblas_int)ncol, n2, &one,
                                          Dk[s], (rocblas_int)cols, e, 1, &one, y + (size_t)s * n2, 1);
                if (two) {
                        const double *A[3] = {Db[0], Db[0], Db[1]}, *B[3] = {Dk[0], Dk[1], Dk[1]};
                        for (int m = 0; m < 3 && r == rocblas_status_success; m++)
                                r = rocblas_dgemm(blas.h, rocblas_operation_transpose, rocblas_operation_none,
                                                  n2, n2, (rocblas_int)ncol, &one, A[m], (rocblas_int)cols, B[m],
                                                  (rocblas_int)cols, &one, M + (size_t)m * n2 * n2, n2);
                }
                if (r != rocblas_status_success) {
                        fcis::set_error("fcis_gas_trans_rdm12s: rocBLAS failed (status %d)", (int)r);
                        return FCIS_ERR_DEVICE;
                }
        }
        std::vector<double> hy(2 * (size_t)n2), hM(two ? 3 * (size_t)n2 * n2 : 0);
        FCIS_HIP(hipMemcpyAsync(hy.data(), y, 8 * hy.size(), hipMemcpyDeviceToHost, st));
        if (two) FCIS_HIP(hipMemcpyAsync(hM.data(), M, 8 * hM.size(), hipMemcpyDeviceToHost, st));
        FCIS_HIP(hipStreamSynchronize(st));
        // y[s][d n + c] = <bra| a_d^+ a_c |ket>;  M[m] column-major:
        // M(i, j) = hM[j n2 + i] = sum_K D_bra[i, K] D_ket[j, K]
        double *dm1[2] = {dm1a, dm1b};
        for (int s = 0; s < 2; s++) {
                if (dm1[s] == nullptr) continue;
                for (int p = 0; p < n; p++)
                        for (int q = 0; q < n; q++) dm1[s][p * n + q] = hy[(size_t)s * n2 + q * n + p];
        }
        double *dm2[3] = {dm2aa, dm2ab, dm2bb};
        for (int m = 0; m < 3; m++) {
                if (dm2[m] == nullptr) continue;
                const double *Mm = hM.data() + (size_t)m * n2 * n2;
                const double *ys = hy.data() + (size_t)(m == 2 ? 1 : 0) * n2;
                for (int p = 0; p < n; p++)
                        for (int q = 0; q < n; q++)
                                for (int r = 0; r < n; r++)
                                        for (int s = 0; s < n; s++) {
                                                double v = Mm[(size_t)(r * n + s) * n2 + (q * n + p)];
                                                if (m != 1 && q == r) v -= ys[p * n + s];
                                                dm2[m][((size_t)(p * n + q) * n + r) * n + s] = v;
                                        }
        }
        return FCIS_SUCCESS;
}
