// String addressing and single-excitation link tables, built on the device.
//
// Strings of k electrons in n orbitals are addressed by their colex rank
// addr = sum_i C(p_i, i) over ascending occupied orbitals p_1 < ... < p_k,
// the order of pyscf's cistring (make_strings / str2addr) and of libfci_gas
// (pyscf/lib/mcscf/fci_string_gas.c:263-298).
//
// Each string owns nlink = k + k(n-k) links, in pyscf's gen_linkstr_index
// row order: the diagonal (i,i) entries for occupied i ascending, then for
// every occupied i ascending and every virtual a ascending the excitation
// a_a^+ a_i |str> = sign |addr>, sign = (-1)^{#occupied strictly between a,i}
// (cistring.cre_des_sign; fci_string_gas.c:252-258 cd_sign).  One thread
// computes one (string, slot) entry, independently of every other entry.
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#include "common.h"

namespace {

__constant__ fcis::BinomTable c_binom = fcis::make_binom_table();

__device__ inline uint32_t dbinom(int n, int k)
{
        return (k < 0 || k > n) ? 0u : c_binom.v[n][k];
}

__device__ inline uint64_t unrank(int n, int k, uint32_t addr)
{
        uint64_t s = 0;
        int e = k;
        for (int p = n - 1; p >= 0 && e > 0; p--) {
                uint32_t c = dbinom(p, e);
                if (c <= addr) {
                        s |= 1ull << p;
                        addr -= c;
                        e--;
                }
        }
        return s;
}

__device__ inline uint32_t rank(uint64_t s)
{
        uint32_t addr = 0;
        int i = 1;
        while (s) {
                int p = __ffsll((unsigned long long)s) - 1;
                addr += dbinom(p, i);
                i++;
                s &= s - 1;
        }
        return addr;
}

// Position of the j-th (0-based) set bit of x.
__device__ inline int nth_set(uint64_t x, int j)
{
        for (int t = 0; t < j; t++) x &= x - 1;
        return __ffsll((unsigned long long)x) - 1;
}

struct LinkEntry {
        int cre, des;
        uint32_t addr;
        int sign;
        int rslot;      // slot of the reverse excitation in addr's list
};

// perm: address of each string by colex rank (ordered spaces) or NULL.
__device__ inline LinkEntry link_entry(int n, int k, uint64_t s, uint32_t str, int slot,
                                       const uint32_t *__restrict__ perm)
{
        LinkEntry e;
        if (slot < k) {
                int p = nth_set(s, slot);
                e.cre = e.des = p;
                e.addr = str;
                e.sign = 1;
                e.rslot = slot;
                return e;
        }
        int nvir = n - k;
        int j = slot - k;
        int i = j / nvir, v = j - i * nvir;
        uint64_t full = (n == 64) ? ~0ull : ((1ull << n) - 1);
        int q = nth_set(s, i);
        int p = nth_set(~s & full, v);
        uint64_t s1 = (s ^ (1ull << q)) | (1ull << p);
        int hi = p > q ? p : q, lo = p > q ? q : p;
        uint64_t mask = (1ull << hi) - (1ull << (lo + 1));
        e.cre = p;
        e.des = q;
        e.addr = perm ? perm[rank(s1)] : rank(s1);
        e.sign = (__popcll(s & mask) & 1) ? -1 : 1;
        // reverse excitation a_q^+ a_p |s1>: p is occupied in s1 (occupied
        // index = #occupied below p), q is empty (virtual index = #empty below q)
        int io = __popcll(s1 & ((1ull << p) - 1));
        int iv = __popcll(~s1 & full & ((1ull << q) - 1));
        e.rslot = k + io * nvir + iv;
        return e;
}

__global__ void k_make_strings(int n, int k, uint32_t nstr, uint64_t *out)
{
        uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
        if (a < nstr) out[a] = unrank(n, k, a);
}

// grid.x covers strings, grid.y covers link slots.
__global__ void k_build_links(int n, int k, uint32_t nstr, int nlink,
                              const uint64_t *__restrict__ strs,
                              const uint32_t *__restrict__ perm,
                              fcis_link *__restrict__ link)
{
        uint32_t str = blockIdx.x * blockDim.x + threadIdx.x;
        int slot = blockIdx.y;
        if (str >= nstr || slot >= nlink) return;
        LinkEntry e = link_entry(n, k, strs[str], str, slot, perm);
        fcis_link out;
        out.addr = (int32_t)e.addr;
        out.info = link_pack(e.cre, e.des, fcis::pair_index(e.cre, e.des), e.sign < 0 ? 1 : 0,
                             nlink <= 255 ? e.rslot : 0);
        link[(size_t)slot * nstr + str] = out;
}

__global__ void k_export_links(int n, int k, uint32_t nstr, int nlink, int tril,
                               const uint64_t *__restrict__ strs,
                               const uint32_t *__restrict__ perm,
                               int32_t *__restrict__ out)
{
        uint32_t str = blockIdx.x * blockDim.x + threadIdx.x;
        int slot = blockIdx.y;
        if (str >= nstr || slot >= nlink) return;
        LinkEntry e = link_entry(n, k, strs[str], str, slot, perm);
        int4 v;
        if (tril) {
                v.x = fcis::pair_index(e.cre, e.des);
                v.y = 0;
        } else {
                v.x = e.cre;
                v.y = e.des;
        }
        v.z = (int)e.addr;
        v.w = e.sign;
        reinterpret_cast<int4 *>(out)[(size_t)str * nlink + slot] = v;
}

thread_local char g_err[512] = "";

}  // namespace

namespace fcis {

void set_error(const char *fmt, ...)
{
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(g_err, sizeof(g_err), fmt, ap);
        va_end(ap);
}

uint64_t space_string(const fcis_space *s, int spin, int64_t addr)
{
        const uint64_t *h = spin ? s->h_strb : s->h_stra;
        return h ? h[addr] : fcis_addr2str(s->norb, spin ? s->nb : s->na, addr);
}

int64_t space_addr(const fcis_space *s, int spin, uint64_t str)
{
        const int64_t r = fcis_str2addr(s->norb, spin ? s->nb : s->na, str);
        const uint32_t *p = spin ? s->h_permb : s->h_perma;
        return p ? (int64_t)p[r] : r;
}

int build_links(fcis_space *s, hipStream_t stream)
{
        for (int spin = 0; spin < 2; spin++) {
                int k = spin ? s->nb : s->na;
                uint32_t nstr = (uint32_t)(spin ? s->nstrb : s->nstra);
                int nlink = spin ? s->nlinkb : s->nlinka;
                uint64_t *strs = spin ? s->d_strb : s->d_stra;
                const uint32_t *perm = spin ? s->d_permb : s->d_perma;
                fcis_link *link = spin ? s->d_linkb : s->d_linka;
                if (perm == nullptr) {   // colex: strings from their addresses
                        hipLaunchKernelGGL(k_make_strings, dim3((nstr + 255) / 256), dim3(256), 0,
                                           stream, s->norb, k, nstr, strs);
                        FCIS_CHECK_LAUNCH();
                }
                if (nlink > 0) {
                        hipLaunchKernelGGL(k_build_links, dim3((nstr + 255) / 256, nlink),
                                           dim3(256), 0, stream, s->norb, k, nstr, nlink,
                                           strs, perm, link);
                        FCIS_CHECK_LAUNCH();
                }
        }
        return FCIS_SUCCESS;
}

}  // namespace fcis

extern "C" {

const char *fcis_last_error(void)
{
        return g_err;
}

int64_t fcis_str2addr(int norb, int nelec, uint64_t string)
{
        (void)norb;
        int64_t addr = 0;
        int i = 1;
        for (int p = 0; p < 64 && i <= nelec; p++) {
                if (string & (1ull << p)) {
                        addr += fcis::binom_host(p, i);
                        i++;
                }
        }
        return addr;
}

uint64_t fcis_addr2str(int norb, int nelec, int64_t addr)
{
        uint64_t s = 0;
        int e = nelec;
        for (int p = norb - 1; p >= 0 && e > 0; p--) {
                int64_t c = fcis::binom_host(p, e);
                if (c <= addr) {
                        s |= 1ull << p;
                        addr -= c;
                        e--;
                }
        }
        return s;
}

int fcis_space_create(fcis_space_t **out, int norb, int na, int nb, void *stream)
{
        if (out == nullptr || norb < 1 || norb > FCIS_MAXORB || na < 0 || nb < 0 ||
            na > norb || nb > norb) {
                fcis::set_error("fcis_space_create: invalid (norb=%d, na=%d, nb=%d)",
                                norb, na, nb);
                return FCIS_ERR_INVALID;
        }
        *out = nullptr;
        fcis_space *s = (fcis_space *)calloc(1, sizeof(fcis_space));
        if (s == nullptr) return FCIS_ERR_MEMORY;
        s->norb = norb;
        s->na = na;
        s->nb = nb;
        s->npair = norb * (norb + 1) / 2;
        s->nstra = fcis::binom_host(norb, na);
        s->nstrb = fcis::binom_host(norb, nb);
        s->nlinka = na + na * (norb - na);
        s->nlinkb = nb + nb * (norb - nb);
        if (s->nstra * s->nstrb > (int64_t)1 << 40) {
                free(s);
                fcis::set_error("fcis_space_create: determinant space too large");
                return FCIS_ERR_INVALID;
        }
        (void)hipGetDevice(&s->device);
        size_t la = sizeof(fcis_link) * (size_t)s->nstra * (s->nlinka ? s->nlinka : 1);
        size_t lb = sizeof(fcis_link) * (size_t)s->nstrb * (s->nlinkb ? s->nlinkb : 1);
        if (fcis::dev_malloc(&s->d_linka, la) != hipSuccess ||
            fcis::dev_malloc(&s->d_linkb, lb) != hipSuccess ||
            fcis::dev_malloc(&s->d_stra, 8 * (size_t)s->nstra) != hipSuccess ||
            fcis::dev_malloc(&s->d_strb, 8 * (size_t)s->nstrb) != hipSuccess) {
                fcis_space_free(s);
                fcis::set_error("fcis_space_create: device allocation failed");
                return FCIS_ERR_MEMORY;
        }
        int st = fcis::build_links(s, (hipStream_t)stream);
        if (st != FCIS_SUCCESS) {
                fcis_space_free(s);
                return st;
        }
        *out = s;
        return FCIS_SUCCESS;
}

void fcis_space_free(fcis_space_t *s)
{
        if (s == nullptr) return;
        (void)hipFree(s->d_linka);
        (void)hipFree(s->d_linkb);
        (void)hipFree(s->d_stra);
        (void)hipFree(s->d_strb);
        (void)hipFree(s->d_perma);
        (void)hipFree(s->d_permb);
        free(s->h_stra);
        free(s->h_strb);
        free(s->h_perma);
        free(s->h_permb);
        free(s->h_order);
        free(s);
}

int fcis_space_create_ordered(fcis_space_t **out, int norb, int na, int nb,
                              const uint64_t *stra, const uint64_t *strb, void *stream)
{
        if (out == nullptr || stra == nullptr || strb == nullptr || norb < 1 ||
            norb > FCIS_MAXORB || na < 0 || nb < 0 || na > norb || nb > norb) {
                fcis::set_error("fcis_space_create_ordered: invalid argument");
                return FCIS_ERR_INVALID;
        }
        *out = nullptr;
        fcis_space *s = (fcis_space *)calloc(1, sizeof(fcis_space));
        if (s == nullptr) return FCIS_ERR_MEMORY;
        s->norb = norb;
        s->na = na;
        s->nb = nb;
        s->npair = norb * (norb + 1) / 2;
        s->nstra = fcis::binom_host(norb, na);
        s->nstrb = fcis::binom_host(norb, nb);
        s->nlinka = na + na * (norb - na);
        s->nlinkb = nb + nb * (norb - nb);
        if (s->nstra * s->nstrb > (int64_t)1 << 40 || s->nstra >= ((int64_t)1 << 32) ||
            s->nstrb >= ((int64_t)1 << 32)) {
                free(s);
                fcis::set_error("fcis_space_create_ordered: determinant space too large");
                return FCIS_ERR_INVALID;
        }
        (void)hipGetDevice(&s->device);
        // strings: each a k-subset of norb orbitals, every one exactly once
        for (int spin = 0; spin < 2; spin++) {
                const int k = spin ? nb : na;
                const int64_t ns = spin ? s->nstrb : s->nstra;
                const uint64_t *in = spin ? strb : stra;
                uint64_t *h = (uint64_t *)malloc(8 * (size_t)ns + 8);
                uint32_t *perm = (uint32_t *)malloc(4 * (size_t)ns + 4);
                (spin ? s->h_strb : s->h_stra) = h;
                (spin ? s->h_permb : s->h_perma) = perm;
                if (h == nullptr || perm == nullptr) {
                        fcis_space_free(s);
                        return FCIS_ERR_MEMORY;
                }
                for (int64_t r = 0; r < ns; r++) perm[r] = UINT32_MAX;
                for (int64_t a = 0; a < ns; a++) {
                        const uint64_t x = in[a];
                        const int64_t r = fcis_str2addr(norb, k, x);
                        if ((norb < 64 && (x >> norb)) || __builtin_popcountll(x) != k || r >= ns ||
                            perm[r] != UINT32_MAX) {
                                fcis_space_free(s);
                                fcis::set_error("fcis_space_create_ordered: %s string %lld is "
                                                "invalid or repeated", spin ? "beta" : "alpha",
                                                (long long)a);
                                return FCIS_ERR_INVALID;
                        }
                        h[a] = x;
                        perm[r] = (uint32_t)a;
                }
        }
        hipStream_t st = (hipStream_t)stream;
        size_t la = sizeof(fcis_link) * (size_t)s->nstra * (s->nlinka ? s->nlinka : 1);
        size_t lb = sizeof(fcis_link) * (size_t)s->nstrb * (s->nlinkb ? s->nlinkb : 1);
        if (fcis::dev_malloc(&s->d_linka, la) != hipSuccess || fcis::dev_malloc(&s->d_linkb, lb) != hipSuccess ||
            fcis::dev_malloc(&s->d_stra, 8 * (size_t)s->nstra) != hipSuccess ||
            fcis::dev_malloc(&s->d_strb, 8 * (size_t)s->nstrb) != hipSuccess ||
            fcis::dev_malloc(&s->d_perma, 4 * (size_t)s->nstra) != hipSuccess ||
            fcis::dev_malloc(&s->d_permb, 4 * (size_t)s->nstrb) != hipSuccess) {
                fcis_space_free(s);
                fcis::set_error("fcis_space_create_ordered: device allocation failed");
                return FCIS_ERR_MEMORY;
        }
        if (hipMemcpyAsync(s->d_stra, s->h_stra, 8 * (size_t)s->nstra, hipMemcpyHostToDevice, st) !=
                    hipSuccess ||
            hipMemcpyAsync(s->d_strb, s->h_strb, 8 * (size_t)s->nstrb, hipMemcpyHostToDevice, st) !=
                    hipSuccess ||
            hipMemcpyAsync(s->d_perma, s->h_perma, 4 * (size_t)s->nstra, hipMemcpyHostToDevice, st) !=
                    hipSuccess ||
            hipMemcpyAsync(s->d_permb, s->h_permb, 4 * (size_t)s->nstrb, hipMemcpyHostToDevice, st) !=
                    hipSuccess) {
                fcis_space_free(s);
                fcis::set_error("fcis_space_create_ordered: upload failed");
                return FCIS_ERR_DEVICE;
        }
        int r = fcis::build_links(s, st);
        if (r == FCIS_SUCCESS && hipStreamSynchronize(st) != hipSuccess) r = FCIS_ERR_DEVICE;
        if (r != FCIS_SUCCESS) {
                fcis_space_free(s);
                return r;
        }
        *out = s;
        return FCIS_SUCCESS;
}

int fcis_space_info(const fcis_space_t *s, int64_t *nstra, int64_t *nstrb,
                    int *nlinka, int *nlinkb)
{
        if (s == nullptr) return FCIS_ERR_INVALID;
        if (nstra) *nstra = s->nstra;
        if (nstrb) *nstrb = s->nstrb;
        if (nlinka) *nlinka = s->nlinka;
        if (nlinkb) *nlinkb = s->nlinkb;
        return FCIS_SUCCESS;
}

int fcis_gen_linkstr_index(const fcis_space_t *s, int spin, int trilidx,
                           int32_t *d_out, void *stream)
{
        if (s == nullptr || (spin != 0 && spin != 1)) {
                fcis::set_error("fcis_gen_linkstr_index: invalid argument");
                return FCIS_ERR_INVALID;
        }
        int k = spin ? s->nb : s->na;
        uint32_t nstr = (uint32_t)(spin ? s->nstrb : s->nstra);
        int nlink = spin ? s->nlinkb : s->nlinka;
        if (nlink == 0) return FCIS_SUCCESS;  // empty table: nothing to write
        if (d_out == nullptr) {
                fcis::set_error("fcis_gen_linkstr_index: null output");
                return FCIS_ERR_INVALID;
        }
        hipLaunchKernelGGL(k_export_links, dim3((nstr + 255) / 256, nlink), dim3(256), 0,
                           (hipStream_t)stream, s->norb, k, nstr, nlink, trilidx ? 1 : 0,
                           spin ? s->d_strb : s->d_stra, spin ? s->d_permb : s->d_perma, d_out);
        FCIS_CHECK_LAUNCH();
        return FCIS_SUCCESS;
}

int fcis_device_info(char *buf, int buflen)
{
        hipDeviceProp_t prop;
        int dev = 0;
        FCIS_HIP(hipGetDevice(&dev));
        FCIS_HIP(hipGetDeviceProperties(&prop, dev));
        snprintf(buf, (size_t)buflen, "%s CUs=%d LDS/block=%zu mem=%zu", prop.gcnArchName,
                 prop.multiProcessorCount, (size_t)prop.sharedMemPerBlock,
                 (size_t)prop.totalGlobalMem);
        return FCIS_SUCCESS;
}

}  // extern "C"
