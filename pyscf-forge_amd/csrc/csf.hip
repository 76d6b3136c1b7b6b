// CSF <-> determinant transform and the CSF-basis Hamiltonian diagonal
// (the spin-adapted basis of pyscf-forge's CSF-FCI solver,
// pyscf/csf_fci/csfstring.py:569-674, pyscf/lib/csf/csfstring.c).
//
// Determinants are grouped in CSD order (csdstring.py:10-14): blocks by the
// number of doubly occupied orbitals npair, then the pair configuration,
// the unpaired configuration (nspin electrons in the norb-npair other
// orbitals), then the spin configuration (nup up-spins among the nspin
// singles), each in colex string order.  Within a configuration the CSFs
// are U^T (spin dets), with U (ndet_spin x ncsf) the genealogical spin
// eigenvectors (Clebsch-Gordan products, csfstring.c:86-141) times the
// block phase (-1)^(npair(npair-1)/2 + npair*max(0, nb-npair))
// (csfstring.py:623-627).  The device kernels are batched over vectors:
//   det2csf: x[v][coff + conf*ncsf + k] = sum_d U[d][k] c[v][mask[off + conf*ndet + d]]
//   csf2det: c[v][mask[off + conf*ndet + d]] = sum_k U[d][k] x[v][coff + conf*ncsf + k]
// where mask (the CSD -> determinant address map) is built on the device.
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <new>
#include <vector>

#include <rocblas/rocblas.h>

#include "common.h"

namespace {

constexpr int CSF_MAXBLK = 33;

struct CsfBlock {
        int npair, nspin, nup;
        int64_t off;        // CSD offset of the block
        int64_t coff;       // CSF offset of the block
        int64_t nconf;      // pair configurations x unpaired configurations
        int64_t sconf;      // unpaired configurations per pair configuration
        int ndet;           // spin determinants per configuration
        int ncsf;           // CSFs per configuration
        int64_t uoff;       // offset of U in the concatenated table
};

struct CsfTable {
        int nblk;
        int norb, na, nb;
        int64_t nstrb;
        CsfBlock b[CSF_MAXBLK];
};

__constant__ fcis::BinomTable c_binom = fcis::make_binom_table();

__device__ inline uint32_t cbinom(int n, int k)
{
        return (k < 0 || k > n) ? 0u : c_binom.v[n][k];
}

__device__ inline uint64_t cunrank(int n, int k, uint64_t addr)
{
        uint64_t s = 0;
        int e = k;
        for (int p = n - 1; p >= 0 && e > 0; p--) {
                uint64_t c = cbinom(p, e);
                if (c <= addr) {
                        s |= 1ull << p;
                        addr -= c;
                        e--;
                }
        }
        return s;
}

__device__ inline uint64_t crank(uint64_t s)
{
        uint64_t a = 0;
        int i = 1;
        while (s) {
                int p = __ffsll((unsigned long long)s) - 1;
                a += cbinom(p, i);
                i++;
                s &= s - 1;
        }
        return a;
}

// bit i of `bits` -> bit (i-th set bit of `slots`)
__device__ inline uint64_t deposit(uint64_t bits, uint64_t slots)
{
        uint64_t out = 0;
        while (slots) {
                uint64_t low = slots & (~slots + 1);
                if (bits & 1) out |= low;
                bits >>= 1;
                slots ^= low;
        }
        return out;
}

__device__ inline int find_block_csd(const CsfTable &t, int64_t pos)
{
        int b = 0;
        while (b + 1 < t.nblk && t.b[b + 1].off <= pos) b++;
        return b;
}

__device__ inline int find_block_csf(const CsfTable &t, int64_t i)
{
        int b = 0;
        while (b + 1 < t.nblk && t.b[b + 1].coff <= i) b++;
        return b;
}

// Determinant address of every CSD position.
__global__ void k_csd_mask(CsfTable t, int64_t ndet_all, uint32_t *__restrict__ mask)
{
        int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (pos >= ndet_all) return;
        const CsfBlock &B = t.b[find_block_csd(t, pos)];
        int64_t local = pos - B.off;
        int64_t per_d = B.sconf * B.ndet;
        int64_t dconf = local / per_d;
        int64_t rem = local - dconf * per_d;
        int64_t sconf = rem / B.ndet;
        int64_t spins = rem - sconf * B.ndet;
        uint64_t full = (t.norb == 64) ? ~0ull : ((1ull << t.norb) - 1);
        uint64_t dstr = cunrank(t.norb, B.npair, (uint64_t)dconf);
        uint64_t free_orbs = ~dstr & full;
        uint64_t sbits = deposit(cunrank(t.norb - B.npair, B.nspin, (uint64_t)sconf), free_orbs);
        uint64_t up = deposit(cunrank(B.nspin, B.nup, (uint64_t)spins), sbits);
        uint64_t astr = dstr | up;
        uint64_t bstr = dstr | (sbits & ~up);
        mask[pos] = (uint32_t)(crank(astr) * (uint64_t)t.nstrb + crank(bstr));
}

// det <-> CSF as one GEMM per spin block over the determinant vector in
// CSD order: a block is an (nconf x ndet) row-major matrix of
// determinants (configuration-major, spin determinants within), its U an
// (ndet x ncsf) row-major matrix, its CSFs an (nconf x ncsf) row-major
// matrix, so CSF = D U and D = CSF U^T (rocBLAS, column-major views).  The
// two kernels below move determinants between address order and CSD order
// (mask[pos] = address of CSD position pos; positions of blocks without
// CSFs receive zero).
__global__ void k_csd_gather(int64_t ndet_all, const uint32_t *__restrict__ mask,
                             const double *__restrict__ det, double *__restrict__ dcsd)
{
        const int v = blockIdx.y;
        const double *src = det + (int64_t)v * ndet_all;
        double *dst = dcsd + (int64_t)v * ndet_all;
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        for (int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; pos < ndet_all; pos += stride)
                dst[pos] = src[mask[pos]];
}

__global__ void k_csd_scatter(CsfTable t, int64_t ndet_all, const uint32_t *__restrict__ mask,
                              const double *__restrict__ dcsd, double *__restrict__ det)
{
        const int v = blockIdx.y;
        const double *src = dcsd + (int64_t)v * ndet_all;
        double *dst = det + (int64_t)v * ndet_all;
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        for (int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; pos < ndet_all; pos += stride)
                dst[mask[pos]] = t.b[find_block_csd(t, pos)].ncsf > 0 ? src[pos] : 0.0;
}

// Determinant supports of m CSFs (full-space indices): for CSF i the
// determinant addresses and coefficients of its nonzero U column entries,
// in CSD order, compacted into row i of [m][maxdet]; count[i] entries, -1
// for an index out of range.  One thread per CSF.
__global__ void k_csf_support(CsfTable t, int m, int64_t ncsf_all, const int64_t *__restrict__ csf,
                              const uint32_t *__restrict__ mask, const double *__restrict__ U,
                              int maxdet, int64_t *__restrict__ addr, double *__restrict__ coef,
                              int *__restrict__ count)
{
        const int i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= m) return;
        const int64_t c = csf[i];
        if (c < 0 || c >= ncsf_all) {
                count[i] = -1;
                return;
        }
        const CsfBlock &B = t.b[find_block_csf(t, c)];
        const int64_t li = c - B.coff;
        const int64_t conf = li / B.ncsf;
        const int k = (int)(li - conf * B.ncsf);
        const uint32_t *mp = mask + B.off + conf * B.ndet;
        const double *u = U + B.uoff + k;
        int n = 0;
        for (int d = 0; d < B.ndet && n < maxdet; d++) {
                const double v = u[(int64_t)d * B.ncsf];
                if (v != 0.0) {
                        addr[(int64_t)i * maxdet + n] = mp[d];
                        coef[(int64_t)i * maxdet + n] = v;
                        n++;
                }
        }
        count[i] = n;
}

// diag(U^T H_conf U): H_conf diagonal = hdiag_det, off-diagonal between spin
// dets that swap the spins of singles i (up in d) and j (down in d):
// sgn * (o_i o_j | o_j o_i), sgn = -(-1)^(i+j) (csfstring.c:295-361, with the
// singles counted in orbital order).  kex[p*norb+q] = (pq|qp).
__global__ void k_hdiag_csf(CsfTable t, int64_t ncsf_all, const uint32_t *__restrict__ mask,
                            const double *__restrict__ U, const double *__restrict__ hdet,
                            const double *__restrict__ kex, double *__restrict__ out)
{
        int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= ncsf_all) return;
        const CsfBlock &B = t.b[find_block_csf(t, i)];
        int64_t li = i - B.coff;
        int64_t conf = li / B.ncsf;
        int k = (int)(li - conf * B.ncsf);
        const uint32_t *m = mask + B.off + conf * B.ndet;
        const double *u = U + B.uoff + k;
        // orbitals of the singles of this configuration
        uint64_t full = (t.norb == 64) ? ~0ull : ((1ull << t.norb) - 1);
        int64_t dconf = conf / B.sconf, sconf = conf - dconf * B.sconf;
        uint64_t dstr = cunrank(t.norb, B.npair, (uint64_t)dconf);
        uint64_t sbits = deposit(cunrank(t.norb - B.npair, B.nspin, (uint64_t)sconf), ~dstr & full);
        int orb[FCIS_MAXORB];
        {
                uint64_t s = sbits;
                int n = 0;
                while (s) {
                        orb[n++] = __ffsll((unsigned long long)s) - 1;
                        s &= s - 1;
                }
        }
        double acc = 0.0;
        for (int d = 0; d < B.ndet; d++) {
                double ud = u[(int64_t)d * B.ncsf];
                if (ud == 0.0) continue;
                acc += ud * ud * hdet[m[d]];
                uint64_t sp = cunrank(B.nspin, B.nup, (uint64_t)d);
                for (int a = 0; a < B.nspin; a++) {
                        if (!((sp >> a) & 1)) continue;
                        for (int b = 0; b < B.nspin; b++) {
                                if ((sp >> b) & 1) continue;
                                uint64_t e = sp ^ (1ull << a) ^ (1ull << b);
                                int64_t ea = (int64_t)crank(e);
                                double v = ((a + b) & 1 ? 1.0 : -1.0) * kex[orb[a] * t.norb + orb[b]];
                                acc += ud * v * u[ea * B.ncsf];
                        }
                }
        }
        out[i] = acc;
}

// ---- host-side genealogical couplings and Clebsch-Gordan coefficients ----

std::vector<std::vector<int64_t>> gentable(int nspin, int smult)
{
        int n0 = (nspin - (smult - 1)) / 2, n1 = (nspin + (smult - 1)) / 2;
        std::vector<std::vector<int64_t>> g(n0 + 1, std::vector<int64_t>(n1 + 1, 0));
        for (int i1 = n0; i1 <= n1; i1++) g[n0][i1] = 1;
        for (int i0 = n0; i0 > 0; i0--) {
                std::vector<int64_t> row(g[i0].begin() + i0, g[i0].end());
                std::vector<int64_t> cum(row.size());
                for (size_t a = 0; a < row.size(); a++) {
                        int64_t sum = 0;
                        for (size_t b = a; b < row.size(); b++) sum += row[b];
                        cum[a] = sum;
                }
                g[i0 - 1][i0 - 1] = cum[0];
                for (size_t a = 0; a < cum.size(); a++) g[i0 - 1][i0 + a] = cum[a];
        }
        return g;
}

int64_t count_csfs(int nspin, int smult)
{
        if (smult < 1 || (smult - 1) % 2 != nspin % 2 || smult > nspin + 1) return 0;
        return gentable(nspin, smult)[0][0];
}

std::vector<uint64_t> coupling_strings(int nspin, int smult)
{
        int64_t ncsf = count_csfs(nspin, smult);
        std::vector<uint64_t> out(ncsf);
        if (ncsf == 0) return out;
        auto g = gentable(nspin, smult);
        int n1 = (nspin + smult - 1) / 2;
        for (int64_t a = 0; a < ncsf; a++) {
                uint64_t s = 1;
                int64_t rem = a;
                int i0 = 0, i1 = 0;
                for (int ispin = 0; ispin < nspin; ispin++) {
                        if (i1 == n1) break;
                        if (g[i0][i1 + 1] <= rem) {
                                rem -= g[i0][i1 + 1];
                                i0++;
                        } else {
                                s |= 1ull << ispin;
                                i1++;
                        }
                }
                out[a] = s;
        }
        return out;
}

// U[d][k] for spin det d (nup of nspin, colex) and coupling k.
void spin_evecs(int nspin, int nup, int two_ms_total, int smult, std::vector<double> &U)
{
        (void)two_ms_total;
        std::vector<uint64_t> coups = coupling_strings(nspin, smult);
        int64_t ncsf = coups.size();
        int64_t ndet = fcis::binom_host(nspin, nup);
        U.assign((size_t)(ndet * ncsf), 0.0);
        for (int64_t d = 0; d < ndet; d++) {
                uint64_t det = fcis_addr2str(nspin, nup, d);
                for (int64_t k = 0; k < ncsf; k++) {
                        uint64_t cp = coups[k];
                        int two_s = 1, two_ms = (det & 1) ? 1 : -1;
                        double num = 1, den = 1;
                        int sgn = 1, osgn = -two_ms;
                        bool ok = true;
                        for (int ispin = 1; ispin < nspin; ispin++) {
                                int sup = (cp >> ispin) & 1, msup = (det >> ispin) & 1;
                                if (msup) {
                                        two_ms++;
                                        osgn = -osgn;
                                } else {
                                        two_ms--;
                                        sgn *= osgn;
                                }
                                num *= (sup == msup) ? two_s + two_ms + 1 : two_s + 1 - two_ms;
                                if (num == 0) {
                                        ok = false;
                                        break;
                                }
                                den *= two_s + 1;
                                if (msup && !sup) sgn = -sgn;
                                num /= 2;
                                two_s += sup ? 1 : -1;
                        }
                        U[(size_t)(d * ncsf + k)] = ok ? sgn * sqrt(num / den) : 0.0;
                }
        }
}

}  // namespace

struct fcis_csf {
        const fcis_space *space;
        int smult;
        CsfTable t;
        int64_t ndet, ncsf;
        uint32_t *d_mask;
        double *d_U;
        rocblas_handle blas;      // the per-block GEMMs of det2csf / csf2det
        double *d_work;           // determinant vectors in CSD order (det2csf / csf2det)
        int64_t work_n;           // its capacity in doubles
        double *d_kex;
        // point-group packing: the CSFs whose configuration has irrep wfnsym
        int64_t nsym;
        uint32_t *d_sym;        // [nsym] CSF indices, ascending
};

namespace {

// Packed CSF vectors: packed[v][i] = full[v][sym[i]] and back (zero
// elsewhere in the full vector).
__global__ void k_csf_pack(int64_t nsym, int64_t ncsf, const uint32_t *__restrict__ sym,
                           const double *__restrict__ full, double *__restrict__ packed)
{
        const int v = blockIdx.y;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nsym;
             i += (int64_t)gridDim.x * blockDim.x)
                packed[v * nsym + i] = full[v * ncsf + sym[i]];
}

__global__ void k_csf_zero(int64_t n, double *__restrict__ x)
{
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
             i += (int64_t)gridDim.x * blockDim.x)
                x[i] = 0.0;
}

__global__ void k_csf_unpack(int64_t nsym, int64_t ncsf, const uint32_t *__restrict__ sym,
                             const double *__restrict__ packed, double *__restrict__ full)
{
        const int v = blockIdx.y;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nsym;
             i += (int64_t)gridDim.x * blockDim.x)
                full[v * ncsf + sym[i]] = packed[v * nsym + i];
}

uint64_t host_unrank(int n, int k, int64_t addr)
{
        uint64_t s = 0;
        int e = k;
        for (int p = n - 1; p >= 0 && e > 0; p--) {
                int64_t c = fcis::binom_host(p, e);
                if (c <= addr) {
                        s |= 1ull << p;
                        addr -= c;
                        e--;
                }
        }
        return s;
}

uint64_t host_deposit(uint64_t bits, uint64_t slots)
{
        uint64_t out = 0;
        while (slots) {
                uint64_t low = slots & (~slots + 1);
                if (bits & 1) out |= low;
                bits >>= 1;
                slots ^= low;
        }
        return out;
}

// Irrep of configuration (block B, pair configuration dconf, singles
// configuration sconf): XOR of the irreps of the singly occupied orbitals
// (pairs contribute the totally symmetric irrep; D2h-subgroup irrep ids
// multiply by XOR).
int conf_irrep(const CsfTable &t, const CsfBlock &B, int64_t dconf, int64_t sconf,
               const int *orbsym)
{
        const uint64_t full = (t.norb == 64) ? ~0ull : ((1ull << t.norb) - 1);
        const uint64_t dstr = host_unrank(t.norb, B.npair, dconf);
        uint64_t sbits = host_deposit(host_unrank(t.norb - B.npair, B.nspin, sconf), ~dstr & full);
        int ir = 0;
        while (sbits) {
                ir ^= orbsym[__builtin_ctzll(sbits)];
                sbits &= sbits - 1;
        }
        return ir;
}

}  // namespace

static int csf_upload(fcis_csf *c, const std::vector<double> &Uall, hipStream_t st)
{
        if (!Uall.empty()) FCIS_HIP(hipMemcpy(c->d_U, Uall.data(), 8 * Uall.size(), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_csd_mask, dim3((c->ndet + 255) / 256), dim3(256), 0, st, c->t, c->ndet,
                           c->d_mask);
        FCIS_CHECK_LAUNCH();
        FCIS_HIP(hipStreamSynchronize(st));
        return FCIS_SUCCESS;
}

extern "C" {

int fcis_csf_create(fcis_csf_t **out, const fcis_space_t *s, int smult, void *stream)
{
        if (out == nullptr || s == nullptr || smult < 1) {
                fcis::set_error("fcis_csf_create: invalid argument");
                return FCIS_ERR_INVALID;
        }
        *out = nullptr;
        const int norb = s->norb, na = s->na, nb = s->nb;
        if (abs(na - nb) > smult - 1 || (smult - 1) % 2 != abs(na - nb) % 2) {
                fcis::set_error("fcis_csf_create: 2S+1 = %d impossible for (na, nb) = (%d, %d)",
                                smult, na, nb);
                return FCIS_ERR_INVALID;
        }
        if (s->nstra * s->nstrb >= ((int64_t)1 << 32)) {
                fcis::set_error("fcis_csf_create: more than 2^32 determinants");
                return FCIS_ERR_INVALID;
        }
        if (s->h_stra != nullptr) {
                fcis::set_error("fcis_csf_create: the CSF transform needs a colex-ordered space");
                return FCIS_ERR_INVALID;
        }
        fcis_csf *c = new (std::nothrow) fcis_csf();
        if (c == nullptr) return FCIS_ERR_MEMORY;
        c->space = s;
        c->smult = smult;
        CsfTable &t = c->t;
        t.norb = norb;
        t.na = na;
        t.nb = nb;
        t.nstrb = s->nstrb;
        t.nblk = 0;
        int64_t off = 0, coff = 0, uoff = 0;
        std::vector<double> Uall;
        int min_npair = na + nb - norb > 0 ? na + nb - norb : 0;
        int nless = na < nb ? na : nb;
        for (int npair = min_npair; npair <= nless; npair++) {
                if (t.nblk >= CSF_MAXBLK) {
                        delete c;
                        fcis::set_error("fcis_csf_create: too many blocks");
                        return FCIS_ERR_INVALID;
                }
                CsfBlock &B = t.b[t.nblk++];
                B.npair = npair;
                B.nspin = na + nb - 2 * npair;
                B.nup = (B.nspin + na - nb) / 2;
                B.off = off;
                B.coff = coff;
                B.sconf = fcis::binom_host(norb - npair, B.nspin);
                B.nconf = fcis::binom_host(norb, npair) * B.sconf;
                B.ndet = (int)fcis::binom_host(B.nspin, B.nup);
                B.ncsf = (int)count_csfs(B.nspin, smult);
                B.uoff = uoff;
                if ((double)B.ndet * B.ncsf > (double)(1 << 27)) {
                        delete c;
                        fcis::set_error("fcis_csf_create: spin block of %d singles exceeds the "
                                        "2^27-coefficient umat limit", B.nspin);
                        return FCIS_ERR_INVALID;
                }
                if (B.ncsf > 0) {
                        std::vector<double> U;
                        spin_evecs(B.nspin, B.nup, na - nb, smult, U);
                        int64_t ncomm = (int64_t)npair * (npair - 1) / 2 +
                                        (int64_t)npair * (nb - npair > 0 ? nb - npair : 0);
                        double sg = (ncomm % 2) ? -1.0 : 1.0;
                        for (double &x : U) x *= sg;
                        Uall.insert(Uall.end(), U.begin(), U.end());
                        uoff += (int64_t)U.size();
                }
                off += B.nconf * B.ndet;
                coff += B.nconf * B.ncsf;
        }
        c->ndet = off;
        c->ncsf = coff;
        if (off != s->nstra * s->nstrb) {
                delete c;
                fcis::set_error("fcis_csf_create: CSD blocks do not cover the determinant space");
                return FCIS_ERR_INVALID;
        }
        c->d_mask = nullptr;
        c->d_U = nullptr;
        c->blas = nullptr;
        c->d_work = nullptr;
        c->work_n = 0;
        c->d_kex = nullptr;
        c->d_sym = nullptr;
        c->nsym = 0;
        if (fcis::dev_malloc(&c->d_mask, 4 * (size_t)off + 4) != hipSuccess ||
            fcis::dev_malloc(&c->d_U, 8 * (Uall.size() + 1)) != hipSuccess ||

            fcis::dev_malloc(&c->d_kex, 8 * (size_t)norb * norb + 8) != hipSuccess) {
                fcis_csf_free(c);
                fcis::set_error("fcis_csf_create: device allocation failed");
                return FCIS_ERR_MEMORY;
        }
        if (rocblas_create_handle(&c->blas) != rocblas_status_success) {
                c->blas = nullptr;
                fcis_csf_free(c);
                fcis::set_error("fcis_csf_create: rocBLAS handle creation failed");
                return FCIS_ERR_DEVICE;
        }
        int st = csf_upload(c, Uall, (hipStream_t)stream);
        if (st != FCIS_SUCCESS) {
                fcis_csf_free(c);
                return st;
        }
        *out = c;
        return FCIS_SUCCESS;
}

void fcis_csf_free(fcis_csf_t *c)
{
        if (c == nullptr) return;
        (void)hipFree(c->d_mask);
        (void)hipFree(c->d_U);
        (void)hipFree(c->d_work);
        if (c->blas) rocblas_destroy_handle(c->blas);
        (void)hipFree(c->d_kex);
        if (c->d_sym) (void)hipFree(c->d_sym);
        delete c;
}

int fcis_csf_symmetry(fcis_csf_t *c, const int *orbsym, int wfnsym, int64_t *nsym, void *stream)
{
        if (c == nullptr) {
                fcis::set_error("fcis_csf_symmetry: null transformer");
                return FCIS_ERR_INVALID;
        }
        if (c->d_sym) (void)hipFree(c->d_sym);
        c->d_sym = nullptr;
        c->nsym = c->ncsf;
        if (orbsym == nullptr) {
                if (nsym) *nsym = c->ncsf;
                return FCIS_SUCCESS;
        }
        for (int p = 0; p < c->t.norb; p++) {
                if (orbsym[p] < 0 || orbsym[p] > 7) {
                        fcis::set_error("fcis_csf_symmetry: orbsym[%d] = %d is not a D2h-subgroup "
                                        "irrep id (0..7)", p, orbsym[p]);
                        return FCIS_ERR_INVALID;
                }
        }
        std::vector<uint32_t> idx;
        for (int b = 0; b < c->t.nblk; b++) {
                const CsfBlock &B = c->t.b[b];
                if (B.ncsf == 0) continue;
                for (int64_t conf = 0; conf < B.nconf; conf++) {
                        const int64_t dconf = conf / B.sconf, sconf = conf - dconf * B.sconf;
                        if (conf_irrep(c->t, B, dconf, sconf, orbsym) != wfnsym) continue;
                        for (int k = 0; k < B.ncsf; k++)
                                idx.push_back((uint32_t)(B.coff + conf * B.ncsf + k));
                }
        }
        c->nsym = (int64_t)idx.size();
        if (nsym) *nsym = c->nsym;
        if (fcis::dev_malloc(&c->d_sym, 4 * idx.size() + 4) != hipSuccess) {
                c->d_sym = nullptr;
                c->nsym = c->ncsf;
                fcis::set_error("fcis_csf_symmetry: device allocation failed");
                return FCIS_ERR_MEMORY;
        }
        hipStream_t st = (hipStream_t)stream;
        if (!idx.empty())
                FCIS_HIP(hipMemcpyAsync(c->d_sym, idx.data(), 4 * idx.size(), hipMemcpyHostToDevice, st));
        FCIS_HIP(hipStreamSynchronize(st));
        return FCIS_SUCCESS;
}

int fcis_csf_irrep(const fcis_csf_t *c, const int *orbsym, int64_t csf, int *irrep)
{
        if (c == nullptr || orbsym == nullptr || irrep == nullptr || csf < 0 || csf >= c->ncsf) {
                fcis::set_error("fcis_csf_irrep: invalid argument");
                return FCIS_ERR_INVALID;
        }
        int b = 0;
        while (b + 1 < c->t.nblk && c->t.b[b + 1].coff <= csf) b++;
        const CsfBlock &B = c->t.b[b];
        const int64_t conf = (csf - B.coff) / B.ncsf;
        const int64_t dconf = conf / B.sconf, sconf = conf - dconf * B.sconf;
        *irrep = conf_irrep(c->t, B, dconf, sconf, orbsym);
        return FCIS_SUCCESS;
}

int fcis_csf_max_support(const fcis_csf_t *c, int *maxdet)
{
        if (c == nullptr || maxdet == nullptr) {
                fcis::set_error("fcis_csf_max_support: invalid argument");
                return FCIS_ERR_INVALID;
        }
        int mx = 1;
        for (int b = 0; b < c->t.nblk; b++)
                if (c->t.b[b].ncsf > 0 && c->t.b[b].ndet > mx) mx = c->t.b[b].ndet;
        *maxdet = mx;
        return FCIS_SUCCESS;
}

int fcis_csf_support(const fcis_csf_t *c, int m, const int64_t *d_csf, int maxdet, int64_t *d_addr,
                     double *d_coef, int *d_count, void *stream)
{
        int need = 0;
        if (c == nullptr || m < 0 || (m > 0 && (d_csf == nullptr || d_addr == nullptr ||
                                                d_coef == nullptr || d_count == nullptr)) ||
            fcis_csf_max_support(c, &need) != FCIS_SUCCESS || maxdet < need) {
                fcis::set_error("fcis_csf_support: invalid argument (maxdet below the largest "
                                "configuration's determinant count?)");
                return FCIS_ERR_INVALID;
        }
        if (m == 0) return FCIS_SUCCESS;
        hipLaunchKernelGGL(k_csf_support, dim3((m + 63) / 64), dim3(64), 0, (hipStream_t)stream, c->t, m,
                           c->ncsf, d_csf, c->d_mask, c->d_U, maxdet, d_addr, d_coef, d_count);
        FCIS_CHECK_LAUNCH();
        return FCIS_SUCCESS;
}

int fcis_csf_describe(const fcis_csf_t *c, int64_t csf, uint64_t *docc, uint64_t *socc,
                      uint64_t *coupling)
{
        if (c == nullptr || docc == nullptr || socc == nullptr || coupling == nullptr || csf < 0 ||
            csf >= c->ncsf) {
                fcis::set_error("fcis_csf_describe: invalid argument");
                return FCIS_ERR_INVALID;
        }
        int b = 0;
        while (b + 1 < c->t.nblk && c->t.b[b + 1].coff <= csf) b++;
        const CsfBlock &B = c->t.b[b];
        const int64_t rel = csf - B.coff;
        const int64_t conf = rel / B.ncsf, k = rel - conf * B.ncsf;
        const int64_t dconf = conf / B.sconf, sconf = conf - dconf * B.sconf;
        const uint64_t full = (c->t.norb == 64) ? ~0ull : ((1ull << c->t.norb) - 1);
        *docc = host_unrank(c->t.norb, B.npair, dconf);
        *socc = host_deposit(host_unrank(c->t.norb - B.npair, B.nspin, sconf), ~*docc & full);
        *coupling = coupling_strings(B.nspin, c->smult)[k];
        return FCIS_SUCCESS;
}

int fcis_csf_pack(const fcis_csf_t *c, const double *d_full, double *d_packed, int nvec,
                  void *stream)
{
        if (c == nullptr || c->d_sym == nullptr || d_full == nullptr || d_packed == nullptr ||
            nvec < 1 || nvec > 65535) {
                fcis::set_error("fcis_csf_pack: invalid argument (no symmetry set?)");
                return FCIS_ERR_INVALID;
        }
        if (c->nsym == 0) return FCIS_SUCCESS;
        const unsigned g = (unsigned)std::min<int64_t>((c->nsym + 255) / 256, 8192);
        hipLaunchKernelGGL(k_csf_pack, dim3(g, nvec), dim3(256), 0, (hipStream_t)stream, c->nsym,
                           c->ncsf, c->d_sym, d_full, d_packed);
        FCIS_CHECK_LAUNCH();
        return FCIS_SUCCESS;
}

int fcis_csf_unpack(const fcis_csf_t *c, const double *d_packed, double *d_full, int nvec,
                    void *stream)
{
        if (c == nullptr || c->d_sym == nullptr || d_full == nullptr || d_packed == nullptr ||
            nvec < 1 || nvec > 65535) {
                fcis::set_error("fcis_csf_unpack: invalid argument (no symmetry set?)");
                return FCIS_ERR_INVALID;
        }
        hipStream_t st = (hipStream_t)stream;
        const int64_t n = c->ncsf * nvec;
        hipLaunchKernelGGL(k_csf_zero, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 8192)),
                           dim3(256), 0, st, n, d_full);
        FCIS_CHECK_LAUNCH();
        if (c->nsym == 0) return FCIS_SUCCESS;
        const unsigned g = (unsigned)std::min<int64_t>((c->nsym + 255) / 256, 8192);
        hipLaunchKernelGGL(k_csf_unpack, dim3(g, nvec), dim3(256), 0, st, c->nsym, c->ncsf,
                           c->d_sym, d_packed, d_full);
        FCIS_CHECK_LAUNCH();
        return FCIS_SUCCESS;
}

int fcis_csf_info(const fcis_csf_t *c, int64_t *ncsf, int64_t *ndet)
{
        if (c == nullptr) return FCIS_ERR_INVALID;
        if (ncsf) *ncsf = c->ncsf;
        if (ndet) *ndet = c->ndet;
        return FCIS_SUCCESS;
}

// The CSD-order work array for nvec vectors (grown on demand).
static int csf_work(fcis_csf_t *c, int nvec)
{
        const int64_t need = c->ndet * (int64_t)nvec;
        if (need <= c->work_n) return FCIS_SUCCESS;
        (void)hipFree(c->d_work);
        c->d_work = nullptr;
        c->work_n = 0;
        if (fcis::dev_malloc(&c->d_work, 8 * (size_t)need) != hipSuccess) {
                fcis::set_error("CSF transform: device allocation of %zu bytes failed", 8 * (size_t)need);
                return FCIS_ERR_MEMORY;
        }
        c->work_n = need;
        return FCIS_SUCCESS;
}

// CSF = D U (det2csf) or D = CSF U^T (csf2det) for every block with CSFs,
// all nvec vectors in one strided-batched call per block.
static int csf_gemms(fcis_csf_t *c, bool to_csf, const double *d_csf_in, double *d_csf_out, int nvec,
                     hipStream_t st)
{
        if (rocblas_set_stream(c->blas, st) != rocblas_status_success) {
                fcis::set_error("CSF transform: rocblas_set_stream failed");
                return FCIS_ERR_DEVICE;
        }
        const double one = 1.0, zero = 0.0;
        // vectors per call: every operand within 2^32 elements of its base
        // pointer (rocBLAS GEMMs whose operands spanned more returned wrong
        // results without an error)
        const int64_t span = std::max<int64_t>(std::max<int64_t>(c->ndet, c->ncsf), 1);
        const int vchunk = (int)std::max<int64_t>(1, std::min<int64_t>(nvec, (int64_t)0xffffffffLL / span));
        for (int v0 = 0; v0 < nvec; v0 += vchunk) {
                const int nv = std::min(vchunk, nvec - v0);
                double *work = c->d_work + (size_t)v0 * c->ndet;
                for (int b = 0; b < c->t.nblk; b++) {
                        const CsfBlock &B = c->t.b[b];
                        if (B.ncsf == 0 || B.nconf == 0) continue;
                        const double *U = c->d_U + B.uoff;
                        double *D = work + B.off;
                        rocblas_status r;
                        // column-major views: U is (ncsf x ndet), D (ndet x nconf) and
                        // the CSFs (ncsf x nconf); det2csf: CSF = U D, csf2det: D = U^T CSF
                        if (to_csf)
                                r = rocblas_dgemm_strided_batched(
                                        c->blas, rocblas_operation_none, rocblas_operation_none, B.ncsf,
                                        (rocblas_int)B.nconf, B.ndet, &one, U, B.ncsf, 0, D, B.ndet, c->ndet, &zero,
                                        d_csf_out + (size_t)v0 * c->ncsf + B.coff, B.ncsf, c->ncsf, nv);
                        else
                                r = rocblas_dgemm_strided_batched(
                                        c->blas, rocblas_operation_transpose, rocblas_operation_none, B.ndet,
                                        (rocblas_int)B.nconf, B.ncsf, &one, U, B.ncsf, 0,
                                        d_csf_in + (size_t)v0 * c->ncsf + B.coff, B.ncsf, c->ncsf, &zero, D, B.ndet,
                                        c->ndet, nv);
                        if (r != rocblas_status_success) {
                                fcis::set_error("CSF transform: rocblas_dgemm_strided_batched failed (status %d)",
                                                (int)r);
                                return FCIS_ERR_DEVICE;
                        }
                }
        }
        return FCIS_SUCCESS;
}

int fcis_csf_det2csf(const fcis_csf_t *c, const double *d_det, double *d_csf, int nvec, void *stream)
{
        if (c == nullptr || d_det == nullptr || (d_csf == nullptr && c->ncsf > 0) || nvec < 1 || nvec > 65535) {
                fcis::set_error("fcis_csf_det2csf: invalid argument");
                return FCIS_ERR_INVALID;
        }
        if (c->ncsf == 0) return FCIS_SUCCESS;
        fcis_csf_t *cm = const_cast<fcis_csf_t *>(c);
        int r = csf_work(cm, nvec);
        if (r != FCIS_SUCCESS) return r;
        hipStream_t st = (hipStream_t)stream;
        const unsigned g = (unsigned)std::min<int64_t>((c->ndet + 255) / 256, 16384);
        hipLaunchKernelGGL(k_csd_gather, dim3(g, nvec), dim3(256), 0, st, c->ndet, c->d_mask, d_det, c->d_work);
        FCIS_CHECK_LAUNCH();
        return csf_gemms(cm, true, nullptr, d_csf, nvec, st);
}

int fcis_csf_csf2det(const fcis_csf_t *c, const double *d_csf, double *d_det, int nvec, void *stream)
{
        if (c == nullptr || d_det == nullptr || (d_csf == nullptr && c->ncsf > 0) || nvec < 1 || nvec > 65535) {
                fcis::set_error("fcis_csf_csf2det: invalid argument");
                return FCIS_ERR_INVALID;
        }
        fcis_csf_t *cm = const_cast<fcis_csf_t *>(c);
        int r = csf_work(cm, nvec);
        if (r != FCIS_SUCCESS) return r;
        hipStream_t st = (hipStream_t)stream;
        r = csf_gemms(cm, false, d_csf, nullptr, nvec, st);
        if (r != FCIS_SUCCESS) return r;
        const unsigned g = (unsigned)std::min<int64_t>((c->ndet + 255) / 256, 16384);
        hipLaunchKernelGGL(k_csd_scatter, dim3(g, nvec), dim3(256), 0, st, c->t, c->ndet, c->d_mask, c->d_work,
                           d_det);
        FCIS_CHECK_LAUNCH();
        return FCIS_SUCCESS;
}

int fcis_csf_hdiag(fcis_csf_t *c, const double *eri, const double *d_hdiag_det,
                   double *d_hdiag_csf, void *stream)
{
        if (c == nullptr || eri == nullptr || d_hdiag_det == nullptr || d_hdiag_csf == nullptr) {
                fcis::set_error("fcis_csf_hdiag: invalid argument");
                return FCIS_ERR_INVALID;
        }
        const int n = c->t.norb, np = n * (n + 1) / 2;
        std::vector<double> kex((size_t)n * n);
        for (int p = 0; p < n; p++)
                for (int q = 0; q < n; q++) {
                        size_t pq = fcis::pair_index(p, q);
                        kex[(size_t)p * n + q] = eri[pq * np + pq];   // (pq|qp) = pair[pq][qp]
                }
        hipStream_t st = (hipStream_t)stream;
        FCIS_HIP(hipMemcpyAsync(c->d_kex, kex.data(), 8 * kex.size(), hipMemcpyHostToDevice, st));
        if (c->ncsf > 0) {
                hipLaunchKernelGGL(k_hdiag_csf, dim3((c->ncsf + 127) / 128), dim3(128), 0, st, c->t,
                                   c->ncsf, c->d_mask, c->d_U, d_hdiag_det, c->d_kex, d_hdiag_csf);
                FCIS_CHECK_LAUNCH();
        }
        FCIS_HIP(hipStreamSynchronize(st));
        return FCIS_SUCCESS;
}

}  // extern "C"
