// Restricted (GAS) determinant spaces on the device
// (pyscf-forge's GAS-CI, pyscf/mcscf/fci_gas.py, pyscf/lib/mcscf/).
//
// A GAS space splits the norb active orbitals into ngas consecutive
// subspaces and keeps the determinants of a block set D: rows (alpha
// occupation per subspace, beta occupation per subspace), unique and
// lexicographically sorted.  Layout, as the reference's gas_space_t
// (fci_gas.h:52-60, fci_string_gas.c:588-791):
//   sectors  the alpha and beta occupation tuples of D and their single
//            inter-subspace excitations, unique, sorted lexicographically;
//            a sector's strings are addressed row-major over subspaces with
//            the last subspace fastest, each subspace by its combinatorial
//            address;
//   blocks   in the order of D; block b is an alpha-major rectangle
//            (nstr(sa) x nstr(sb)) at offset = running determinant count.
// The restricted Hamiltonian is the projection P H P of the CAS one
// (pyscf/mcscf/test/test_gasci.py:522-575 checks libfci_gas that way), so
// the device path embeds a GAS vector in the CAS (nstr_a, nstr_b) array
// (fci_gas.py gas2fci :141-182), runs the CAS sigma and projects back
// (fci2gas :185-208).  d_map[d] is the CAS determinant index of GAS
// determinant d; it is built on the device from per-sector CAS string
// addresses.  A sector-native space (cas == null, gas_native.hip) has the
// same layout and no CAS: its sigma, diagonal and RDMs run on the GAS
// determinants and their single-excitation closure directly.
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <new>
#include <vector>

#include "gas.h"

namespace {

constexpr int GAS_THREADS = 256;

uint64_t unrank_local(int n, int k, int64_t addr)
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

int64_t rank_colex(uint64_t s)
{
        int64_t a = 0;
        int i = 1;
        while (s) {
                int p = __builtin_ctzll(s);
                a += fcis::binom_host(p, i++);
                s &= s - 1;
        }
        return a;
}


// Grid-stride: d_map[d] for every GAS determinant.  Block b of a
// determinant is found by binary search over the block offsets.
__global__ void k_gas_map(const int64_t *__restrict__ boff, const int64_t *__restrict__ bnb,
                          const int64_t *__restrict__ abase, const int64_t *__restrict__ bbase,
                          int nblock, const uint32_t *__restrict__ secaddr, int64_t nstrb,
                          int64_t ndet, uint32_t *__restrict__ map)
{
        for (int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; d < ndet;
             d += (int64_t)gridDim.x * blockDim.x) {
                int lo = 0, hi = nblock - 1;
                while (lo < hi) {
                        int mid = (lo + hi + 1) >> 1;
                        if (boff[mid] <= d) lo = mid;
                        else hi = mid - 1;
                }
                int64_t r = d - boff[lo];
                int64_t nb = bnb[lo];
                int64_t ia = r / nb, ib = r - ia * nb;
                map[d] = (uint32_t)((int64_t)secaddr[abase[lo] + ia] * nstrb +
                                    secaddr[bbase[lo] + ib]);
        }
}

__global__ void k_gas_zero(double *__restrict__ x, int64_t n)
{
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
             i += (int64_t)gridDim.x * blockDim.x) {
                x[i] = 0.0;
        }
}

// cas[v][map[d]] = gas[v][d]
__global__ void k_gas_scatter(const uint32_t *__restrict__ map, const double *__restrict__ gas,
                              double *__restrict__ cas, int64_t ndet, int64_t ncas)
{
        const int v = blockIdx.y;
        const double *src = gas + (int64_t)v * ndet;
        double *dst = cas + (int64_t)v * ncas;
        for (int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; d < ndet;
             d += (int64_t)gridDim.x * blockDim.x) {
                dst[map[d]] = src[d];
        }
}

// gas[v][d] = cas[v][map[d]]
__global__ void k_gas_gather(const uint32_t *__restrict__ map, const double *__restrict__ cas,
                             double *__restrict__ gas, int64_t ndet, int64_t ncas)
{
        const int v = blockIdx.y;
        const double *src = cas + (int64_t)v * ncas;
        double *dst = gas + (int64_t)v * ndet;
        for (int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; d < ndet;
             d += (int64_t)gridDim.x * blockDim.x) {
                dst[d] = src[map[d]];
        }
}

__global__ void k_gas_det_index(const uint32_t *__restrict__ map, int64_t ndet, int64_t n,
                                const int64_t *__restrict__ in, int64_t *__restrict__ out)
{
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
             i += (int64_t)gridDim.x * blockDim.x) {
                const int64_t a = in[i];
                out[i] = (a >= 0 && a < ndet) ? (int64_t)map[a] : -1;
        }
}

int no_cas(const char *who)
{
        fcis::set_error("%s: the space is sector-native (no CAS embedding)", who);
        return FCIS_ERR_INVALID;
}

unsigned grid_for(int64_t n)
{
        int64_t g = (n + GAS_THREADS - 1) / GAS_THREADS;
        return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, 8192));
}

int embed(const fcis_gas_space *g, const double *in, double *out, int nvec, hipStream_t st)
{
        hipLaunchKernelGGL(k_gas_zero, dim3(grid_for(g->ncas * nvec)), dim3(GAS_THREADS), 0, st,
                           out, g->ncas * nvec);
        FCIS_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_gas_scatter, dim3(grid_for(g->ndet), nvec), dim3(GAS_THREADS), 0, st,
                           g->d_map, in, out, g->ndet, g->ncas);
        FCIS_CHECK_LAUNCH();
        return FCIS_SUCCESS;
}

int project(const fcis_gas_space *g, const double *in, double *out, int nvec, hipStream_t st)
{
        hipLaunchKernelGGL(k_gas_gather, dim3(grid_for(g->ndet), nvec), dim3(GAS_THREADS), 0, st,
                           g->d_map, in, out, g->ndet, g->ncas);
        FCIS_CHECK_LAUNCH();
        return FCIS_SUCCESS;
}

int lex_cmp(const uint8_t *a, const uint8_t *b, int n)
{
        for (int i = 0; i < n; i++) {
                if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
        }
        return 0;
}

// Device scratch of n doubles, freed after the stream drains.
struct Scratch {
        double *p = nullptr;
        hipStream_t st;
        explicit Scratch(hipStream_t s) : st(s) {}
        ~Scratch()
        {
                if (p) {
                        (void)hipStreamSynchronize(st);
                        (void)hipFree(p);
                }
        }
};

int sector_of(const fcis_gas_space *g, const std::vector<std::vector<uint8_t>> &pool,
              uint64_t str)
{
        std::vector<uint8_t> o(g->ngas);
        for (int k = 0; k < g->ngas; k++)
                o[k] = (uint8_t)__builtin_popcountll((str >> g->start[k]) & ((1ull << g->norb[k]) - 1));
        auto it = std::lower_bound(pool.begin(), pool.end(), o);
        return (it != pool.end() && *it == o) ? (int)(it - pool.begin()) : -1;
}

}  // namespace

uint64_t fcis::gas_sector_string(const fcis_gas_space *g, int s, int64_t addr)
{
        uint64_t str = 0;
        const uint8_t *occ = &g->occ[(size_t)s * g->ngas];
        const int64_t *st = &g->stride[(size_t)s * g->ngas];
        for (int k = 0; k < g->ngas; k++) {
                int64_t q = addr / st[k];
                addr -= q * st[k];
                str |= unrank_local(g->norb[k], occ[k], q) << g->start[k];
        }
        return str;
}

std::vector<int> fcis::gas_moves(const fcis_gas_space *g, const SectorPool &pool, int s)
{
        std::vector<int> out{s};
        for (int r = 0; r < g->ngas; r++) {
                for (int t = 0; t < g->ngas; t++) {
                        if (r == t || pool[s][r] == 0 || pool[s][t] >= g->norb[t]) continue;
                        std::vector<uint8_t> o = pool[s];
                        o[r]--;
                        o[t]++;
                        auto it = std::lower_bound(pool.begin(), pool.end(), o);
                        if (it != pool.end() && *it == o) out.push_back((int)(it - pool.begin()));
                }
        }
        return out;
}

int fcis::gas_layout(fcis_gas_space *g, int ngas, const int *norb, int na, int nb, int nblock,
                     const int *block_occ, SectorPool &pool, const char *who)
{
        g->ngas = ngas;
        g->na = na;
        g->nb = nb;
        int tot = 0;
        for (int k = 0; k < ngas; k++) {
                if (norb[k] < 1 || norb[k] > GAS_MAXLOCAL) {
                        fcis::set_error("%s: subspace %d has %d orbitals", who, k, norb[k]);
                        return FCIS_ERR_INVALID;
                }
                g->norb[k] = norb[k];
                g->start[k] = tot;
                tot += norb[k];
        }
        if (tot > GAS_MAXORB) {
                fcis::set_error("%s: %d orbitals (at most %d)", who, tot, GAS_MAXORB);
                return FCIS_ERR_INVALID;
        }
        // blocks: valid occupations, electron counts, strictly increasing
        const int w = 2 * ngas;
        std::vector<uint8_t> rows((size_t)nblock * w);
        for (int b = 0; b < nblock; b++) {
                int sa = 0, sb = 0;
                for (int k = 0; k < w; k++) {
                        int v = block_occ[(size_t)b * w + k];
                        if (v < 0 || v > norb[k % ngas]) {
                                fcis::set_error("%s: block %d occupation out of range", who, b);
                                return FCIS_ERR_INVALID;
                        }
                        rows[(size_t)b * w + k] = (uint8_t)v;
                        (k < ngas ? sa : sb) += v;
                }
                if (sa != na || sb != nb) {
                        fcis::set_error("%s: block %d holds (%d, %d) electrons, the space (%d, %d)",
                                        who, b, sa, sb, na, nb);
                        return FCIS_ERR_INVALID;
                }
                if (b > 0 && lex_cmp(&rows[(size_t)(b - 1) * w], &rows[(size_t)b * w], w) >= 0) {
                        fcis::set_error("%s: blocks must be unique and sorted (block %d)", who, b);
                        return FCIS_ERR_INVALID;
                }
        }
        // sector pool: block sectors and their single inter-subspace moves
        pool.clear();
        std::vector<uint8_t> t(ngas);
        for (int b = 0; b < nblock; b++) {
                for (int spin = 0; spin < 2; spin++) {
                        const uint8_t *o = &rows[(size_t)b * w + spin * ngas];
                        pool.emplace_back(o, o + ngas);
                        for (int r = 0; r < ngas; r++) {
                                for (int s = 0; s < ngas; s++) {
                                        if (r == s || o[r] == 0 || o[s] >= norb[s]) continue;
                                        std::copy(o, o + ngas, t.begin());
                                        t[r]--;
                                        t[s]++;
                                        pool.push_back(t);
                                }
                        }
                }
        }
        std::sort(pool.begin(), pool.end());
        pool.erase(std::unique(pool.begin(), pool.end()), pool.end());
        if (pool.size() > 65534) {
                fcis::set_error("%s: more than 65534 sectors", who);
                return FCIS_ERR_INVALID;
        }
        g->nsector = (int)pool.size();
        g->occ.resize((size_t)g->nsector * ngas);
        g->stride.resize((size_t)g->nsector * ngas);
        g->nstr.resize(g->nsector);
        for (int s = 0; s < g->nsector; s++) {
                int64_t acc = 1;
                for (int k = ngas - 1; k >= 0; k--) {
                        g->occ[(size_t)s * ngas + k] = pool[s][k];
                        g->stride[(size_t)s * ngas + k] = acc;
                        acc *= fcis::binom_host(norb[k], pool[s][k]);
                }
                g->nstr[s] = acc;
        }
        auto find = [&](const uint8_t *o) {
                std::vector<uint8_t> key(o, o + ngas);
                return (int)(std::lower_bound(pool.begin(), pool.end(), key) - pool.begin());
        };
        g->nblock = nblock;
        g->boff.resize(nblock);
        g->bsa.resize(nblock);
        g->bsb.resize(nblock);
        int64_t off = 0;
        for (int b = 0; b < nblock; b++) {
                g->bsa[b] = find(&rows[(size_t)b * w]);
                g->bsb[b] = find(&rows[(size_t)b * w + ngas]);
                g->boff[b] = off;
                off += g->nstr[g->bsa[b]] * g->nstr[g->bsb[b]];
        }
        g->ndet = off;
        return FCIS_SUCCESS;
}

namespace {

// Tables of fcis_rowset for the fused CAS sigma restricted to S1 (see the
// struct); rows sorted by chunk count, longest first, for the persistent
// grid's round-robin.
int build_rowset(fcis_gas_space *g, const std::vector<std::vector<uint8_t>> &pool,
                 const std::vector<uint8_t> &blk, hipStream_t st)
{
        const int ns = g->nsector;
        const int64_t na = g->cas->nstra, nb = g->cas->nstrb;
        const int64_t nch = (nb + FUSED_NB - 1) / FUSED_NB;
        // S1 sector pairs
        std::vector<std::vector<char>> allow(ns, std::vector<char>(ns, 0));
        auto moves = [&](int s) { return fcis::gas_moves(g, pool, s); };
        for (int b = 0; b < g->nblock; b++) {
                (void)blk;
                const int sa = g->bsa[b], sb = g->bsb[b];
                for (int x : moves(sa)) allow[x][sb] = 1;
                for (int y : moves(sb)) allow[sa][y] = 1;
        }
        // sector of every CAS beta string; chunk lists per alpha sector
        std::vector<int> bsec(nb);
        for (int64_t j = 0; j < nb; j++) bsec[j] = sector_of(g, pool, fcis::space_string(g->cas, 1, j));
        std::vector<int64_t> soff(ns, 0), scnt(ns, 0);
        std::vector<uint16_t> clist;
        std::vector<char> mark(nch);
        for (int s = 0; s < ns; s++) {
                std::fill(mark.begin(), mark.end(), 0);
                for (int64_t j = 0; j < nb; j++)
                        if (bsec[j] >= 0 && allow[s][bsec[j]]) mark[j / FUSED_NB] = 1;
                soff[s] = (int64_t)clist.size();
                for (int64_t c = 0; c < nch; c++)
                        if (mark[c]) clist.push_back((uint16_t)c);
                scnt[s] = (int64_t)clist.size() - soff[s];
        }
        std::vector<std::pair<int64_t, int32_t>> order;   // (-count, alpha string)
        std::vector<int> asec(na);
        for (int64_t i = 0; i < na; i++) {
                asec[i] = sector_of(g, pool, fcis::space_string(g->cas, 0, i));
                if (asec[i] >= 0 && scnt[asec[i]] > 0) order.emplace_back(-scnt[asec[i]], (int32_t)i);
        }
        std::sort(order.begin(), order.end());
        const int64_t nrows = (int64_t)order.size();
        std::vector<int32_t> rows(nrows), rpos(na, -1);
        std::vector<int2> rowcl(nrows);
        int64_t tot = 0;
        for (int64_t r = 0; r < nrows; r++) {
                const int32_t I = order[r].second;
                rows[r] = I;
                rpos[I] = (int32_t)r;
                rowcl[r] = make_int2((int)soff[asec[I]], (int)scnt[asec[I]]);
                tot += scnt[asec[I]];
        }
        g->nchunk_rows = tot;
        if (fcis::dev_malloc(&g->d_rows, 4 * (size_t)nrows + 4) != hipSuccess ||
            fcis::dev_malloc(&g->d_rpos, 4 * (size_t)na + 4) != hipSuccess ||
            fcis::dev_malloc(&g->d_rowcl, 8 * (size_t)nrows + 8) != hipSuccess ||
            fcis::dev_malloc(&g->d_clist, 2 * clist.size() + 2) != hipSuccess) {
                fcis::set_error("fcis_gas_space_create: device allocation failed");
                return FCIS_ERR_MEMORY;
        }
        FCIS_HIP(hipMemcpyAsync(g->d_rows, rows.data(), 4 * (size_t)nrows, hipMemcpyHostToDevice, st));
        FCIS_HIP(hipMemcpyAsync(g->d_rpos, rpos.data(), 4 * (size_t)na, hipMemcpyHostToDevice, st));
        FCIS_HIP(hipMemcpyAsync(g->d_rowcl, rowcl.data(), 8 * (size_t)nrows, hipMemcpyHostToDevice, st));
        FCIS_HIP(hipMemcpyAsync(g->d_clist, clist.data(), 2 * clist.size(), hipMemcpyHostToDevice, st));
        FCIS_HIP(hipStreamSynchronize(st));
        g->rs.nrows = nrows;
        g->rs.d_rows = g->d_rows;
        g->rs.d_rowcl = g->d_rowcl;
        g->rs.d_clist = g->d_clist;
        g->rs.d_rpos = g->d_rpos;
        g->rs.ndets = g->ndet;
        g->rs.d_dets = g->d_map;
        return FCIS_SUCCESS;
}

// Adjacent-orbital Givens rotations on the string axis of a GAS tensor
// (transform_ci_within_gas): one workgroup per fiber (a line of `extent`
// elements at base + i * stride), staged in LDS; rotation r mixes the
// disjoint string pairs (left, right) of its list in parallel, rotations in
// order with a barrier between them, then the reflection signs.
__global__ __launch_bounds__(256) void k_rotate_fibers(double *__restrict__ x, int64_t nfiber,
                                                       const int64_t *__restrict__ base,
                                                       const int64_t *__restrict__ stride, int extent,
                                                       int nrot, const int32_t *__restrict__ roff,
                                                       const int32_t *__restrict__ left,
                                                       const int32_t *__restrict__ right,
                                                       const double *__restrict__ cs,
                                                       const double *__restrict__ sign)
{
        extern __shared__ double fib[];
        for (int64_t f = blockIdx.x; f < nfiber; f += gridDim.x) {
                const int64_t b0 = base[f], st = stride[f];
                for (int i = threadIdx.x; i < extent; i += blockDim.x) fib[i] = x[b0 + i * st];
                __syncthreads();
                for (int r = 0; r < nrot; r++) {
                        const double c = cs[2 * r], s = cs[2 * r + 1];
                        for (int k = roff[r] + threadIdx.x; k < roff[r + 1]; k += blockDim.x) {
                                const int l = left[k], q = right[k];
                                const double a = fib[l], b = fib[q];
                                fib[l] = c * a + s * b;
                                fib[q] = -s * a + c * b;
                        }
                        __syncthreads();
                }
                for (int i = threadIdx.x; i < extent; i += blockDim.x) x[b0 + i * st] = fib[i] * sign[i];
                __syncthreads();
        }
}

}  // namespace

extern "C" {

int fcis_rotate_fibers(double *d_x, int64_t nfiber, const int64_t *d_base, const int64_t *d_stride,
                       int extent, int nrot, const int32_t *d_roff, const int32_t *d_left,
                       const int32_t *d_right, const double *d_cs, const double *d_sign, void *stream)
{
        if (d_x == nullptr || nfiber < 0 || extent < 1 || nrot < 0 || extent > 16384 ||
            (nfiber > 0 && (d_base == nullptr || d_stride == nullptr || d_roff == nullptr ||
                            d_sign == nullptr || (nrot > 0 && (d_cs == nullptr || d_left == nullptr ||
                                                               d_right == nullptr))))) {
                fcis::set_error("fcis_rotate_fibers: invalid argument (extent <= 16384)");
                return FCIS_ERR_INVALID;
        }
        if (nfiber == 0) return FCIS_SUCCESS;
        const size_t lds = 8 * (size_t)extent;
        if (lds > 64 * 1024)
                FCIS_HIP(hipFuncSetAttribute((const void *)k_rotate_fibers,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        const unsigned grid = (unsigned)std::min<int64_t>(nfiber, 65536);
        hipLaunchKernelGGL(k_rotate_fibers, dim3(grid), dim3(256), lds, (hipStream_t)stream, d_x, nfiber,
                           d_base, d_stride, extent, nrot, d_roff, d_left, d_right, d_cs, d_sign);
        FCIS_CHECK_LAUNCH();
        return FCIS_SUCCESS;
}

void fcis_gas_space_free(fcis_gas_space_t *g)
{
        if (g == nullptr) return;
        if (g->d_map) (void)hipFree(g->d_map);
        if (g->d_rows) (void)hipFree(g->d_rows);
        if (g->d_rpos) (void)hipFree(g->d_rpos);
        if (g->d_rowcl) (void)hipFree(g->d_rowcl);
        if (g->d_clist) (void)hipFree(g->d_clist);
        if (g->nat) fcis::gas_native_free(g->nat);
        delete g;
}

int fcis_gas_space_create(fcis_gas_space_t **out, const fcis_space_t *cas, int ngas,
                          const int *norb, int nblock, const int *block_occ, void *stream)
{
        if (out == nullptr || cas == nullptr || norb == nullptr || block_occ == nullptr ||
            ngas < 1 || ngas > GAS_MAXNGAS || nblock < 1) {
                fcis::set_error("fcis_gas_space_create: invalid argument");
                return FCIS_ERR_INVALID;
        }
        *out = nullptr;
        hipStream_t st = (hipStream_t)stream;
        fcis_gas_space *g = new (std::nothrow) fcis_gas_space();
        if (g == nullptr) return FCIS_ERR_MEMORY;
        g->cas = cas;
        fcis::SectorPool pool;
        int r0 = fcis::gas_layout(g, ngas, norb, cas->na, cas->nb, nblock, block_occ, pool,
                                  "fcis_gas_space_create");
        if (r0 != FCIS_SUCCESS) {
                fcis_gas_space_free(g);
                return r0;
        }
        int tot = 0;
        for (int k = 0; k < ngas; k++) tot += norb[k];
        if (tot != cas->norb) {
                fcis_gas_space_free(g);
                fcis::set_error("fcis_gas_space_create: subspaces hold %d orbitals, the CAS "
                                "space %d", tot, cas->norb);
                return FCIS_ERR_INVALID;
        }
        g->ncas = cas->nstra * cas->nstrb;
        if (g->ncas >= ((int64_t)1 << 32)) {
                fcis_gas_space_free(g);
                fcis::set_error("fcis_gas_space_create: CAS space of 2^32 or more determinants");
                return FCIS_ERR_INVALID;
        }
        if (g->ndet >= ((int64_t)1 << 32)) {
                fcis_gas_space_free(g);
                fcis::set_error("fcis_gas_space_create: 2^32 or more determinants");
                return FCIS_ERR_INVALID;
        }
        std::vector<uint8_t> rows;      // unused by the CAS row tables
        // CAS string addresses (in the CAS space's own order) of every
        // sector used by a block, per spin
        std::vector<int64_t> base[2] = {std::vector<int64_t>(g->nsector, -1),
                                        std::vector<int64_t>(g->nsector, -1)};
        std::vector<uint32_t> secaddr;
        for (int b = 0; b < nblock; b++) {
                for (int spin = 0; spin < 2; spin++) {
                        const int s = spin ? g->bsb[b] : g->bsa[b];
                        if (base[spin][s] >= 0) continue;
                        base[spin][s] = (int64_t)secaddr.size();
                        for (int64_t a = 0; a < g->nstr[s]; a++) {
                                secaddr.push_back(
                                        (uint32_t)fcis::space_addr(cas, spin, fcis::gas_sector_string(g, s, a)));
                        }
                }
        }
        std::vector<int64_t> tab(4 * (size_t)nblock);
        for (int b = 0; b < nblock; b++) {
                tab[b] = g->boff[b];
                tab[nblock + b] = g->nstr[g->bsb[b]];
                tab[2 * nblock + b] = base[0][g->bsa[b]];
                tab[3 * nblock + b] = base[1][g->bsb[b]];
        }
        int64_t *d_tab = nullptr;
        uint32_t *d_sec = nullptr;
        int st_ok = FCIS_SUCCESS;
        if (fcis::dev_malloc(&g->d_map, 4 * (size_t)g->ndet + 4) != hipSuccess ||
            fcis::dev_malloc(&d_tab, 8 * tab.size()) != hipSuccess ||
            fcis::dev_malloc(&d_sec, 4 * secaddr.size() + 4) != hipSuccess) {
                fcis::set_error("fcis_gas_space_create: device allocation failed");
                st_ok = FCIS_ERR_MEMORY;
        } else if (hipMemcpyAsync(d_tab, tab.data(), 8 * tab.size(), hipMemcpyHostToDevice, st) !=
                           hipSuccess ||
                   hipMemcpyAsync(d_sec, secaddr.data(), 4 * secaddr.size(),
                                  hipMemcpyHostToDevice, st) != hipSuccess) {
                fcis::set_error("fcis_gas_space_create: upload failed");
                st_ok = FCIS_ERR_DEVICE;
        } else {
                hipLaunchKernelGGL(k_gas_map, dim3(grid_for(g->ndet)), dim3(GAS_THREADS), 0, st,
                                   d_tab, d_tab + nblock, d_tab + 2 * nblock, d_tab + 3 * nblock,
                                   nblock, d_sec, cas->nstrb, g->ndet, g->d_map);
                hipError_t e = hipGetLastError();
                if (e == hipSuccess) e = hipStreamSynchronize(st);
                if (e != hipSuccess) {
                        fcis::set_error("fcis_gas_space_create: %s", hipGetErrorString(e));
                        st_ok = FCIS_ERR_DEVICE;
                }
        }
        if (d_tab) (void)hipFree(d_tab);
        if (d_sec) (void)hipFree(d_sec);
        if (st_ok == FCIS_SUCCESS) st_ok = build_rowset(g, pool, rows, st);
        if (st_ok != FCIS_SUCCESS) {
                fcis_gas_space_free(g);
                return st_ok;
        }
        *out = g;
        return FCIS_SUCCESS;
}

int fcis_gas_string_order(int ngas, const int *norb, int nelec, uint64_t *out)
{
        if (ngas < 1 || ngas > GAS_MAXNGAS || norb == nullptr || out == nullptr || nelec < 0) {
                fcis::set_error("fcis_gas_string_order: invalid argument");
                return FCIS_ERR_INVALID;
        }
        int tot = 0, start[GAS_MAXNGAS];
        for (int k = 0; k < ngas; k++) {
                if (norb[k] < 1) {
                        fcis::set_error("fcis_gas_string_order: empty subspace");
                        return FCIS_ERR_INVALID;
                }
                start[k] = tot;
                tot += norb[k];
        }
        if (tot > FCIS_MAXORB || nelec > tot) {
                fcis::set_error("fcis_gas_string_order: invalid orbital or electron count");
                return FCIS_ERR_INVALID;
        }
        const int64_t ns = fcis::binom_host(tot, nelec);
        std::vector<std::pair<std::vector<uint8_t>, int64_t>> key(ns);
        for (int64_t a = 0; a < ns; a++) {
                const uint64_t x = unrank_local(tot, nelec, a);
                key[a].first.resize(ngas);
                for (int k = 0; k < ngas; k++)
                        key[a].first[k] = (uint8_t)__builtin_popcountll(
                                (x >> start[k]) & ((1ull << norb[k]) - 1));
                key[a].second = a;
        }
        std::stable_sort(key.begin(), key.end(),
                         [](const std::pair<std::vector<uint8_t>, int64_t> &u,
                            const std::pair<std::vector<uint8_t>, int64_t> &v) {
                                 return u.first < v.first;
                         });
        for (int64_t a = 0; a < ns; a++) out[a] = unrank_local(tot, nelec, key[a].second);
        return FCIS_SUCCESS;
}

int fcis_gas_space_info(const fcis_gas_space_t *g, int64_t *ndet, int *nblock, int *nsector)
{
        if (g == nullptr) {
                fcis::set_error("fcis_gas_space_info: null space");
                return FCIS_ERR_INVALID;
        }
        if (ndet) *ndet = g->ndet;
        if (nblock) *nblock = g->nblock;
        if (nsector) *nsector = g->nsector;
        return FCIS_SUCCESS;
}

int fcis_gas_block_info(const fcis_gas_space_t *g, int64_t *offset, int *sa, int *sb,
                        int64_t *na, int64_t *nb)
{
        if (g == nullptr) {
                fcis::set_error("fcis_gas_block_info: null space");
                return FCIS_ERR_INVALID;
        }
        for (int b = 0; b < g->nblock; b++) {
                if (offset) offset[b] = g->boff[b];
                if (sa) sa[b] = g->bsa[b];
                if (sb) sb[b] = g->bsb[b];
                if (na) na[b] = g->nstr[g->bsa[b]];
                if (nb) nb[b] = g->nstr[g->bsb[b]];
        }
        return FCIS_SUCCESS;
}

int fcis_gas_sector_info(const fcis_gas_space_t *g, int *occ, int64_t *nstr)
{
        if (g == nullptr) {
                fcis::set_error("fcis_gas_sector_info: null space");
                return FCIS_ERR_INVALID;
        }
        for (int s = 0; s < g->nsector; s++) {
                if (occ) {
                        for (int k = 0; k < g->ngas; k++) {
                                occ[(size_t)s * g->ngas + k] = g->occ[(size_t)s * g->ngas + k];
                        }
                }
                if (nstr) nstr[s] = g->nstr[s];
        }
        return FCIS_SUCCESS;
}

int fcis_gas_addr2str(const fcis_gas_space_t *g, int sector, int64_t addr, uint64_t *str)
{
        if (g == nullptr || str == nullptr || sector < 0 || sector >= g->nsector || addr < 0 ||
            addr >= g->nstr[sector]) {
                fcis::set_error("fcis_gas_addr2str: invalid sector or address");
                return FCIS_ERR_INVALID;
        }
        *str = fcis::gas_sector_string(g, sector, addr);
        return FCIS_SUCCESS;
}

int fcis_gas_str2addr(const fcis_gas_space_t *g, int sector, uint64_t str, int64_t *addr)
{
        if (g == nullptr || addr == nullptr || sector < 0 || sector >= g->nsector) {
                fcis::set_error("fcis_gas_str2addr: invalid sector");
                return FCIS_ERR_INVALID;
        }
        int64_t a = 0;
        for (int k = 0; k < g->ngas; k++) {
                uint64_t sub = (str >> g->start[k]) & ((1ull << g->norb[k]) - 1);
                if (__builtin_popcountll(sub) != g->occ[(size_t)sector * g->ngas + k]) {
                        fcis::set_error("fcis_gas_str2addr: string not in sector %d", sector);
                        return FCIS_ERR_INVALID;
                }
                a += rank_colex(sub) * g->stride[(size_t)sector * g->ngas + k];
        }
        if (str >> (g->start[g->ngas - 1] + g->norb[g->ngas - 1])) {
                fcis::set_error("fcis_gas_str2addr: string has orbitals outside the space");
                return FCIS_ERR_INVALID;
        }
        *addr = a;
        return FCIS_SUCCESS;
}

int fcis_gas_det_index(const fcis_gas_space_t *g, int64_t n, const int64_t *d_addr,
                       int64_t *d_cas_addr, void *stream)
{
        if (g == nullptr || n < 0 || (n > 0 && (d_addr == nullptr || d_cas_addr == nullptr))) {
                fcis::set_error("fcis_gas_det_index: invalid argument");
                return FCIS_ERR_INVALID;
        }
        if (g->cas == nullptr) return no_cas("fcis_gas_det_index");
        if (n == 0) return FCIS_SUCCESS;
        hipLaunchKernelGGL(k_gas_det_index, dim3(grid_for(n)), dim3(GAS_THREADS), 0,
                           (hipStream_t)stream, g->d_map, g->ndet, n, d_addr, d_cas_addr);
        FCIS_CHECK_LAUNCH();
        return FCIS_SUCCESS;
}

int fcis_gas_embed(const fcis_gas_space_t *g, const double *d_gas, double *d_cas, int nvec,
                   void *stream)
{
        if (g == nullptr || d_gas == nullptr || d_cas == nullptr || nvec < 1 || nvec > 65535) {
                fcis::set_error("fcis_gas_embed: invalid argument");
                return FCIS_ERR_INVALID;
        }
        if (g->cas == nullptr) return no_cas("fcis_gas_embed");
        return embed(g, d_gas, d_cas, nvec, (hipStream_t)stream);
}

int fcis_gas_project(const fcis_gas_space_t *g, const double *d_cas, double *d_gas, int nvec,
                     void *stream)
{
        if (g == nullptr || d_gas == nullptr || d_cas == nullptr || nvec < 1 || nvec > 65535) {
                fcis::set_error("fcis_gas_project: invalid argument");
                return FCIS_ERR_INVALID;
        }
        if (g->cas == nullptr) return no_cas("fcis_gas_project");
        return project(g, d_cas, d_gas, nvec, (hipStream_t)stream);
}

int fcis_gas_make_hdiag(const fcis_gas_space_t *g, const double *h1e, const double *eri,
                        double *d_hdiag, void *stream)
{
        if (g == nullptr || h1e == nullptr || eri == nullptr || d_hdiag == nullptr) {
                fcis::set_error("fcis_gas_make_hdiag: invalid argument");
                return FCIS_ERR_INVALID;
        }
        hipStream_t st = (hipStream_t)stream;
        if (g->cas == nullptr) return fcis::gas_native_hdiag(g, h1e, eri, d_hdiag, st);
        Scratch tmp(st);
        FCIS_HIP(fcis::dev_malloc(&tmp.p, 8 * (size_t)g->ncas));
        int r = fcis_make_hdiag(g->cas, h1e, eri, tmp.p, stream);
        if (r != FCIS_SUCCESS) return r;
        return project(g, tmp.p, d_hdiag, 1, st);
}

int fcis_gas_trans_rdm12s(const fcis_gas_space_t *g, const double *d_bra, const double *d_ket,
                          double *dm1a, double *dm1b, double *dm2aa, double *dm2ab,
                          double *dm2bb, void *stream)
{
        if (g == nullptr || d_bra == nullptr || d_ket == nullptr) {
                fcis::set_error("fcis_gas_trans_rdm12s: invalid argument");
                return FCIS_ERR_INVALID;
        }
        hipStream_t st = (hipStream_t)stream;
        if (g->cas == nullptr)
                return fcis::gas_native_rdm(g, d_bra, d_ket, dm1a, dm1b, dm2aa, dm2ab, dm2bb, st);
        Scratch tmp(st);
        FCIS_HIP(fcis::dev_malloc(&tmp.p, 16 * (size_t)g->ncas));
        int r = embed(g, d_bra, tmp.p, 1, st);
        if (r == FCIS_SUCCESS) r = embed(g, d_ket, tmp.p + g->ncas, 1, st);
        if (r == FCIS_SUCCESS) {
                r = fcis_trans_rdm12s(g->cas, tmp.p, tmp.p + g->ncas, dm1a, dm1b, dm2aa, dm2ab,
                                      dm2bb, stream);
        }
        return r;
}

void fcis_gas_sigma_plan_free(fcis_gas_sigma_plan_t *p)
{
        if (p == nullptr) return;
        if (p->cas_plan) fcis_sigma_plan_free(p->cas_plan);
        if (p->d_a) (void)hipFree(p->d_a);
        if (p->nat) fcis::gas_native_plan_free(p->nat);
        delete p;
}

int fcis_gas_sigma_plan_create(fcis_gas_sigma_plan_t **out, const fcis_gas_space_t *g,
                               const double *h2e, int precision, uint64_t workspace_bytes)
{
        if (out == nullptr || g == nullptr || h2e == nullptr) {
                fcis::set_error("fcis_gas_sigma_plan_create: invalid argument");
                return FCIS_ERR_INVALID;
        }
        *out = nullptr;
        fcis_gas_sigma_plan *p = new (std::nothrow) fcis_gas_sigma_plan();
        if (p == nullptr) return FCIS_ERR_MEMORY;
        p->gas = g;
        p->cas_plan = nullptr;
        p->d_a = p->d_b = nullptr;
        p->nat = nullptr;
        if (g->cas == nullptr) {
                int r = fcis::gas_native_plan_create(&p->nat, g, h2e, precision, workspace_bytes);
                if (r != FCIS_SUCCESS) {
                        fcis_gas_sigma_plan_free(p);
                        return r;
                }
                *out = p;
                return FCIS_SUCCESS;
        }
        if (fcis::dev_malloc(&p->d_a, 16 * (size_t)g->ncas) != hipSuccess) {
                fcis_gas_sigma_plan_free(p);
                fcis::set_error("fcis_gas_sigma_plan_create: device allocation failed");
                return FCIS_ERR_MEMORY;
        }
        p->d_b = p->d_a + g->ncas;
        int r = fcis_sigma_plan_create(&p->cas_plan, g->cas, h2e, precision, workspace_bytes);
        if (r != FCIS_SUCCESS) {
                fcis_gas_sigma_plan_free(p);
                return r;
        }
        *out = p;
        return FCIS_SUCCESS;
}

int fcis_gas_sigma_execute(const fcis_gas_space_t *g, fcis_sigma_plan_t *cas_plan,
                           const double *d_ci0, double *d_ci1, double *d_work, void *stream)
{
        if (g == nullptr || cas_plan == nullptr || d_ci0 == nullptr || d_ci1 == nullptr ||
            d_work == nullptr || d_ci0 == d_ci1 || cas_plan->space != g->cas) {
                fcis::set_error("fcis_gas_sigma_execute: invalid argument (distinct device "
                                "arrays and a plan on the space's CAS space are required)");
                return FCIS_ERR_INVALID;
        }
        if (g->cas == nullptr) return no_cas("fcis_gas_sigma_execute");
        hipStream_t st = (hipStream_t)stream;
        double *a = d_work, *b = d_work + g->ncas;
        int r = embed(g, d_ci0, a, 1, st);
        if (r != FCIS_SUCCESS) return r;
        if (cas_plan->fused) return fcis::sigma_plan_execute_rows(cas_plan, &g->rs, a, b, d_ci1, st);
        r = fcis_sigma_plan_execute(cas_plan, a, b, stream);
        if (r == FCIS_SUCCESS) r = project(g, b, d_ci1, 1, st);
        return r;
}

int fcis_gas_contract_ss(const fcis_gas_space_t *g, const fcis_spin_plan_t *spin,
                         const double *d_ci0, double *d_ci1, double *d_work, void *stream)
{
        if (g != nullptr && g->cas == nullptr && d_ci0 != nullptr && d_ci1 != nullptr &&
            d_ci0 != d_ci1)
                return fcis::gas_native_ss(g, d_ci0, d_ci1, (hipStream_t)stream);
        if (g == nullptr || spin == nullptr || d_ci0 == nullptr || d_ci1 == nullptr ||
            d_work == nullptr || d_ci0 == d_ci1) {
                fcis::set_error("fcis_gas_contract_ss: invalid argument");
                return FCIS_ERR_INVALID;
        }
        hipStream_t st = (hipStream_t)stream;
        double *a = d_work, *b = d_work + g->ncas;
        int r = embed(g, d_ci0, a, 1, st);
        if (r == FCIS_SUCCESS) r = fcis_spin_plan_contract(spin, a, b, stream);
        if (r == FCIS_SUCCESS) r = project(g, b, d_ci1, 1, st);
        return r;
}

int fcis_gas_rows_info(const fcis_gas_space_t *g, int64_t *nrows, int64_t *nchunks,
                       int64_t *nchunks_cas)
{
        if (g == nullptr) {
                fcis::set_error("fcis_gas_rows_info: null space");
                return FCIS_ERR_INVALID;
        }
        if (g->cas == nullptr) return no_cas("fcis_gas_rows_info");
        if (nrows) *nrows = g->rs.nrows;
        if (nchunks) *nchunks = g->nchunk_rows;
        if (nchunks_cas) *nchunks_cas = g->cas->nstra * ((g->cas->nstrb + FUSED_NB - 1) / FUSED_NB);
        return FCIS_SUCCESS;
}

int fcis_gas_sigma_plan_execute(fcis_gas_sigma_plan_t *p, const double *d_ci0, double *d_ci1,
                                void *stream)
{
        if (p == nullptr || d_ci0 == nullptr || d_ci1 == nullptr || d_ci0 == d_ci1) {
                fcis::set_error("fcis_gas_sigma_plan_execute: invalid argument (input and "
                                "output must be distinct device arrays)");
                return FCIS_ERR_INVALID;
        }
        if (p->nat) return fcis::gas_native_plan_execute(p->nat, d_ci0, d_ci1, (hipStream_t)stream);
        return fcis_gas_sigma_execute(p->gas, p->cas_plan, d_ci0, d_ci1, p->d_a, stream);
}

int fcis_gas_space_create_native(fcis_gas_space_t **out, int ngas, const int *norb, int na,
                                 int nb, int nblock, const int *block_occ, void *stream)
{
        if (out == nullptr || norb == nullptr || block_occ == nullptr || ngas < 1 ||
            ngas > GAS_MAXNGAS || nblock < 1 || na < 0 || nb < 0) {
                fcis::set_error("fcis_gas_space_create_native: invalid argument");
                return FCIS_ERR_INVALID;
        }
        *out = nullptr;
        fcis_gas_space *g = new (std::nothrow) fcis_gas_space();
        if (g == nullptr) return FCIS_ERR_MEMORY;
        g->cas = nullptr;
        fcis::SectorPool pool;
        int r = fcis::gas_layout(g, ngas, norb, na, nb, nblock, block_occ, pool,
                                 "fcis_gas_space_create_native");
        if (r == FCIS_SUCCESS) r = fcis::gas_native_create(g, pool, (hipStream_t)stream);
        if (r != FCIS_SUCCESS) {
                fcis_gas_space_free(g);
                return r;
        }
        *out = g;
        return FCIS_SUCCESS;
}

int fcis_gas_det_strings(const fcis_gas_space_t *g, int64_t n, const int64_t *addr,
                         uint64_t *stra, uint64_t *strb)
{
        if (g == nullptr || n < 0 || (n > 0 && (addr == nullptr || stra == nullptr ||
                                                strb == nullptr))) {
                fcis::set_error("fcis_gas_det_strings: invalid argument");
                return FCIS_ERR_INVALID;
        }
        for (int64_t i = 0; i < n; i++) {
                const int64_t d = addr[i];
                if (d < 0 || d >= g->ndet) {
                        fcis::set_error("fcis_gas_det_strings: determinant %lld outside the space",
                                        (long long)d);
                        return FCIS_ERR_INVALID;
                }
                const int b = (int)(std::upper_bound(g->boff.begin(), g->boff.end(), d) -
                                    g->boff.begin()) - 1;
                const int64_t nbs = g->nstr[g->bsb[b]];
                const int64_t ia = (d - g->boff[b]) / nbs, ib = d - g->boff[b] - ia * nbs;
                stra[i] = fcis::gas_sector_string(g, g->bsa[b], ia);
                strb[i] = fcis::gas_sector_string(g, g->bsb[b], ib);
        }
        return FCIS_SUCCESS;
}

int fcis_gas_pspace(const fcis_gas_space_t *g, const double *h1e, const double *eri, int m,
                    const int64_t *addr, double *d_h0, void *stream)
{
        if (g == nullptr || h1e == nullptr || eri == nullptr || m < 0 ||
            (m > 0 && (addr == nullptr || d_h0 == nullptr))) {
                fcis::set_error("fcis_gas_pspace: invalid argument");
                return FCIS_ERR_INVALID;
        }
        if ((int64_t)m * m > ((int64_t)1 << 31)) {
                fcis::set_error("fcis_gas_pspace: %d determinants exceed the 46340 limit", m);
                return FCIS_ERR_INVALID;
        }
        if (m == 0) return FCIS_SUCCESS;
        std::vector<uint64_t> s(2 * (size_t)m);
        int r = fcis_gas_det_strings(g, m, addr, s.data(), s.data() + m);
        if (r != FCIS_SUCCESS) return r;
        hipStream_t st = (hipStream_t)stream;
        uint64_t *d_s = nullptr;
        if (fcis::dev_malloc(&d_s, 16 * (size_t)m) != hipSuccess) {
                fcis::set_error("fcis_gas_pspace: device allocation failed");
                return FCIS_ERR_MEMORY;
        }
        int n = 0;
        for (int k = 0; k < g->ngas; k++) n += g->norb[k];
        if (hipMemcpyAsync(d_s, s.data(), 16 * (size_t)m, hipMemcpyHostToDevice, st) != hipSuccess) {
                fcis::set_error("fcis_gas_pspace: upload failed");
                r = FCIS_ERR_DEVICE;
        } else {
                r = fcis::pspace_strings(n, m, d_s, d_s + m, h1e, eri, d_h0, st);
        }
        (void)hipStreamSynchronize(st);
        (void)hipFree(d_s);
        return r;
}

int fcis_gas_sigma_plan_path(const fcis_gas_sigma_plan_t *p)
{
        if (p == nullptr) return FCIS_ERR_INVALID;
        if (p->nat) return fcis::gas_native_plan_path(p->nat);
        return fcis_sigma_plan_path(p->cas_plan);
}

int fcis_gas_ss_diag(const fcis_gas_space_t *g, double *d_out, void *stream)
{
        if (g == nullptr || d_out == nullptr) {
                fcis::set_error("fcis_gas_ss_diag: invalid argument");
                return FCIS_ERR_INVALID;
        }
        if (g->cas != nullptr) {
                fcis::set_error("fcis_gas_ss_diag: a CAS-embedded space takes the diagonal of its "
                                "CAS spin plan (fcis_spin_plan_diag) projected");
                return FCIS_ERR_INVALID;
        }
        return fcis::gas_native_ss(g, nullptr, d_out, (hipStream_t)stream);
}

int fcis_gas_native_info(const fcis_gas_space_t *g, int *native, int64_t *nclosure,
                         int64_t *maxrow, int64_t *nstra, int64_t *nstrb)
{
        if (g == nullptr) {
                fcis::set_error("fcis_gas_native_info: null space");
                return FCIS_ERR_INVALID;
        }
        if (native) *native = g->cas == nullptr;
        return fcis::gas_native_info(g, nclosure, maxrow, nstra, nstrb);
}

}  // extern "C"
