// Internal definitions shared by the fcisigma HIP sources (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "../../include/fcisigma.h"

#define FCIS_MAXORB 32

namespace fcis {

void set_error(const char *fmt, ...);

// Check a HIP call; on failure record the error and return FCIS_ERR_DEVICE.
#define FCIS_HIP(call)                                                        \
        do {                                                                  \
                hipError_t e_ = (call);                                       \
                if (e_ != hipSuccess) {                                       \
                        ::fcis::set_error("%s:%d %s: %s", __FILE__, __LINE__, \
                                          #call, hipGetErrorString(e_));      \
                        return FCIS_ERR_DEVICE;                               \
                }                                                             \
        } while (0)

#define FCIS_CHECK_LAUNCH()                                                   \
        do {                                                                  \
                hipError_t e_ = hipGetLastError();                            \
                if (e_ != hipSuccess) {                                       \
                        ::fcis::set_error("%s:%d launch: %s", __FILE__,       \
                                          __LINE__, hipGetErrorString(e_));   \
                        return FCIS_ERR_DEVICE;                               \
                }                                                             \
        } while (0)

// hipMalloc that leaves no error state behind when it fails: HIP keeps the
// failure as the thread's last error, which the next hipGetLastError (the
// launch check after the next kernel) would otherwise report.
template <typename T>
static inline hipError_t dev_malloc(T **p, size_t bytes)
{
        hipError_t e = hipMalloc(p, bytes);
        if (e != hipSuccess) (void)hipGetLastError();
        return e;
}

// Plan workspaces (devmem.hip): blocks of 256 MiB and more freed through
// big_free stay allocated for the next big_alloc of about their size on the
// same device (at most four per process); a failed allocation releases that
// device's kept blocks and retries.  big_release(-1) frees every kept block.
hipError_t big_alloc(void **p, size_t bytes);
void big_free(void *p);
size_t big_kept_bytes(int device);
void big_release(int device);

template <typename T>
static inline hipError_t big_malloc(T **p, size_t bytes)
{
        void *q = nullptr;
        const hipError_t e = big_alloc(&q, bytes);
        *p = static_cast<T *>(q);
        return e;
}

static inline int64_t binom_host(int n, int k)
{
        if (k < 0 || k > n) return 0;
        if (k > n - k) k = n - k;
        int64_t r = 1;
        for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
        return r;
}

// Binomial coefficients C(n, k), n <= FCIS_MAXORB, for a __constant__
// table in each source that needs them on the device.
struct BinomTable {
        uint32_t v[FCIS_MAXORB + 1][FCIS_MAXORB + 1];
};

constexpr BinomTable make_binom_table()
{
        BinomTable t{};
        for (int n = 0; n <= FCIS_MAXORB; n++) {
                t.v[n][0] = 1;
                for (int k = 1; k <= n; k++) {
                        t.v[n][k] = t.v[n - 1][k - 1] + (k <= n - 1 ? t.v[n - 1][k] : 0u);
                }
        }
        return t;
}

__host__ __device__ static inline int pair_index(int p, int q)
{
        int hi = p > q ? p : q, lo = p > q ? q : p;
        return hi * (hi + 1) / 2 + lo;
}

}  // namespace fcis

// A compact link: target string address and a packed descriptor
//   info = neg | pair << 1 | des << 12 | cre << 18 | rslot << 24
// for a_cre^+ a_des |str> = (-1)^neg |addr>, pair = pair_index(cre, des);
// rslot is the slot of the reverse link (a_des^+ a_cre |addr> = same sign
// |str>) in addr's own list (valid when nlink <= 255).
// Stored transposed, link[k * nstr + str], so that a wavefront reading the
// k-th link of 64 consecutive strings issues one coalesced load.
struct fcis_link {
        int32_t addr;
        int32_t info;
};

__host__ __device__ static inline int link_neg(int32_t info) { return info & 1; }
__host__ __device__ static inline int link_pair(int32_t info) { return (info >> 1) & 0x7ff; }
__host__ __device__ static inline int link_des(int32_t info) { return (info >> 12) & 0x3f; }
__host__ __device__ static inline int link_cre(int32_t info) { return (info >> 18) & 0x3f; }
__host__ __device__ static inline int link_rslot(int32_t info) { return (info >> 24) & 0xff; }
__host__ __device__ static inline int32_t link_pack(int cre, int des, int pair, int neg, int rslot)
{
        return neg | (pair << 1) | (des << 12) | (cre << 18) | ((rslot & 0xff) << 24);
}

struct fcis_space {
        int norb, na, nb;
        int npair;
        int64_t nstra, nstrb;
        int nlinka, nlinkb;
        fcis_link *d_linka;     // [nlinka][nstra]
        fcis_link *d_linkb;     // [nlinkb][nstrb]
        uint64_t *d_stra;       // strings in address order
        uint64_t *d_strb;
        int device;
        // Spaces with a caller-given string order (fcis_space_create_ordered):
        // host copies of the strings in address order and the address of
        // each string by colex rank (host and device); NULL for colex spaces.
        uint64_t *h_stra, *h_strb;
        uint32_t *h_perma, *h_permb;
        uint32_t *d_perma, *d_permb;
        // processing order of the alpha strings for fused plans with
        // XCD blocks of order_block strings (host, malloc; built by the
        // first such plan, reused by later ones)
        int32_t *h_order;
        int order_block;
};

namespace fcis {
int build_links(fcis_space *s, hipStream_t stream);

// Host views of a space's addressing (colex or caller-ordered).
uint64_t space_string(const fcis_space *s, int spin, int64_t addr);
int64_t space_addr(const fcis_space *s, int spin, uint64_t str);
}
