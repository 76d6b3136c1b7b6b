// Hamiltonian diagonal and the Davidson vector primitives.
#include <stdlib.h>
#include <string.h>
#include <vector>

#include "common.h"

namespace {

// hdiag[I, b] = ea[I] + eb[b] + sum_{i in I} jb[b][i]
//   ea/eb: one-spin string energies  sum_i h_ii + sum_{i<j} (J_ij - K_ij)
//   jb[b][i] = sum_{j in b} J_ij   (opposite-spin Coulomb)
// (direct_spin1.make_hdiag; fci_hdiag_gas.c:144-164 spin_diagonal and the
// alpha-beta Coulomb term).  One thread per determinant.
__global__ __launch_bounds__(256) void k_hdiag(int norb, int64_t nb, int64_t ndet,
                                               const uint64_t *__restrict__ stra,
                                               const double *__restrict__ ea,
                                               const double *__restrict__ eb,
                                               const double *__restrict__ jb,
                                               double *__restrict__ hdiag)
{
        int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (d >= ndet) return;
        int64_t I = d / nb, b = d - I * nb;
        uint64_t s = stra[I];
        double v = ea[I] + eb[b];
        const double *row = jb + b * norb;
        while (s) {
                int i = __ffsll((unsigned long long)s) - 1;
                v += row[i];
                s &= s - 1;
        }
        hdiag[d] = v;
}

constexpr int DOT_THREADS = 256;
constexpr int DOT_BLOCKS = 512;
constexpr int MAX_COEF = 64;

// Partial dot products: part[i * DOT_BLOCKS + blk] = sum over this block's
// grid-stride share of X_i . v.  Fixed grid and fixed tree order make the
// result bitwise reproducible.
__global__ __launch_bounds__(DOT_THREADS) void k_dots(int64_t n, const double *__restrict__ x,
                                                      int64_t ldx, const double *__restrict__ v,
                                                      double *__restrict__ part)
{
        __shared__ double red[DOT_THREADS];
        const int i = blockIdx.y;
        const double *xi = x + (int64_t)i * ldx;
        double s = 0.0;
        for (int64_t k = (int64_t)blockIdx.x * DOT_THREADS + threadIdx.x; k < n;
             k += (int64_t)DOT_BLOCKS * DOT_THREADS)
                s += xi[k] * v[k];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int w = DOT_THREADS / 2; w > 0; w >>= 1) {
                if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
                __syncthreads();
        }
        if (threadIdx.x == 0) part[(int64_t)i * DOT_BLOCKS + blockIdx.x] = red[0];
}

struct Coefs {
        double c[MAX_COEF];
};

__global__ __launch_bounds__(256) void k_lincomb(int m, int64_t n, const double *__restrict__ x,
                                                 int64_t ldx, Coefs coef, double beta,
                                                 double *__restrict__ y)
{
        int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (k >= n) return;
        double acc = beta == 0.0 ? 0.0 : beta * y[k];
        for (int i = 0; i < m; i++) acc += coef.c[i] * x[(int64_t)i * ldx + k];
        y[k] = acc;
}

__global__ __launch_bounds__(256) void k_precond(int64_t n, const double *__restrict__ hdiag,
                                                 double e, double shift,
                                                 const double *r, double *out)
{
        int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (k >= n) return;
        double den = hdiag[k] - e;
        if (fabs(den) < shift) den = shift;
        out[k] = r[k] / den;
}

// out[k] = delta_{k,addr} + scale * sign(k), sign from the splitmix64
// finaliser of k + seed, seed = (root+1) * 0x9E3779B97F4A7C15: bit 0 set -> +1.
// (fci_gas.py:61-88 _add_distributed_noise, restated for the device.)
__global__ __launch_bounds__(256) void k_init_guess(int64_t n, int64_t addr, uint64_t seed,
                                                    double scale, double *__restrict__ out)
{
        int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if (k >= n) return;
        uint64_t v = (uint64_t)k + seed;
        v ^= v >> 30;
        v *= 0xBF58476D1CE4E5B9ull;
        v ^= v >> 27;
        v *= 0x94D049BB133111EBull;
        v ^= v >> 31;
        double x = (v & 1ull) ? scale : -scale;
        out[k] = (k == addr ? 1.0 : 0.0) + x;
}

// Per-device scratch for the two-stage reductions (grown on demand).
struct Scratch {
        double *d = nullptr;
        size_t bytes = 0;
};
Scratch g_scratch[16];

int scratch(size_t bytes, double **out)
{
        int dev = 0;
        FCIS_HIP(hipGetDevice(&dev));
        if (dev < 0 || dev >= 16) return FCIS_ERR_INVALID;
        Scratch &s = g_scratch[dev];
        if (s.bytes < bytes) {
                if (s.d) (void)hipFree(s.d);
                s.d = nullptr;
                s.bytes = 0;
                FCIS_HIP(fcis::dev_malloc(&s.d, bytes));
                s.bytes = bytes;
        }
        *out = s.d;
        return FCIS_SUCCESS;
}

}  // namespace

extern "C" {

int fcis_make_hdiag(const fcis_space_t *s, const double *h1e, const double *eri,
                    double *d_hdiag, void *stream)
{
        if (s == nullptr || h1e == nullptr || eri == nullptr || d_hdiag == nullptr) {
                fcis::set_error("fcis_make_hdiag: null argument");
                return FCIS_ERR_INVALID;
        }
        int n = s->norb, np = s->npair;
        std::vector<double> J(n * n), K(n * n);
        for (int i = 0; i < n; i++)
                for (int j = 0; j < n; j++) {
                        J[i * n + j] = eri[(size_t)fcis::pair_index(i, i) * np + fcis::pair_index(j, j)];
                        K[i * n + j] = eri[(size_t)fcis::pair_index(i, j) * np + fcis::pair_index(j, i)];
                }
        std::vector<double> ea(s->nstra), eb(s->nstrb), jb((size_t)s->nstrb * n);
        for (int spin = 0; spin < 2; spin++) {
                int64_t ns = spin ? s->nstrb : s->nstra;
                double *e = spin ? eb.data() : ea.data();
                for (int64_t a = 0; a < ns; a++) {
                        uint64_t str = fcis::space_string(s, spin, a);
                        double v = 0;
                        for (int i = 0; i < n; i++) {
                                if (!((str >> i) & 1)) continue;
                                v += h1e[i * n + i];
                                for (int j = 0; j < i; j++)
                                        if ((str >> j) & 1) v += J[i * n + j] - K[i * n + j];
                        }
                        e[a] = v;
                        if (spin) {
                                for (int i = 0; i < n; i++) {
                                        double c = 0;
                                        for (int j = 0; j < n; j++)
                                                if ((str >> j) & 1) c += J[i * n + j];
                                        jb[(size_t)a * n + i] = c;
                                }
                        }
                }
        }
        double *d_ea = nullptr, *d_eb = nullptr, *d_jb = nullptr;
        hipStream_t st = (hipStream_t)stream;
        int64_t ndet = s->nstra * s->nstrb;
        int ret = FCIS_SUCCESS;
        if (fcis::dev_malloc(&d_ea, 8 * (ea.size() + 1)) != hipSuccess ||
            fcis::dev_malloc(&d_eb, 8 * (eb.size() + 1)) != hipSuccess ||
            fcis::dev_malloc(&d_jb, 8 * (jb.size() + 1)) != hipSuccess) {
                fcis::set_error("fcis_make_hdiag: device allocation failed");
                ret = FCIS_ERR_MEMORY;
        } else if (hipMemcpyAsync(d_ea, ea.data(), 8 * ea.size(), hipMemcpyHostToDevice, st) != hipSuccess ||
                   hipMemcpyAsync(d_eb, eb.data(), 8 * eb.size(), hipMemcpyHostToDevice, st) != hipSuccess ||
                   hipMemcpyAsync(d_jb, jb.data(), 8 * jb.size(), hipMemcpyHostToDevice, st) != hipSuccess) {
                fcis::set_error("fcis_make_hdiag: copy failed");
                ret = FCIS_ERR_DEVICE;
        } else if (ndet > 0) {
                hipLaunchKernelGGL(k_hdiag, dim3((ndet + 255) / 256), dim3(256), 0, st, n,
                                   s->nstrb, ndet, s->d_stra, d_ea, d_eb, d_jb, d_hdiag);
                if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
                        fcis::set_error("fcis_make_hdiag: kernel failed");
                        ret = FCIS_ERR_DEVICE;
                }
        }
        (void)hipStreamSynchronize(st);
        (void)hipFree(d_ea);
        (void)hipFree(d_eb);
        (void)hipFree(d_jb);
        return ret;
}

int fcis_dots(int m, int64_t n, const double *d_x, int64_t ldx, const double *d_v, double *out,
              void *stream)
{
        if (m < 0 || n < 0 || (m > 0 && (d_x == nullptr || d_v == nullptr || out == nullptr))) {
                fcis::set_error("fcis_dots: invalid argument");
                return FCIS_ERR_INVALID;
        }
        if (m == 0) return FCIS_SUCCESS;
        if (m > 65535) return FCIS_ERR_INVALID;
        double *part = nullptr;
        int st = scratch(sizeof(double) * (size_t)m * DOT_BLOCKS, &part);
        if (st != FCIS_SUCCESS) return st;
        hipStream_t hs = (hipStream_t)stream;
        hipLaunchKernelGGL(k_dots, dim3(DOT_BLOCKS, m), dim3(DOT_THREADS), 0, hs, n, d_x, ldx, d_v,
                           part);
        FCIS_CHECK_LAUNCH();
        std::vector<double> h((size_t)m * DOT_BLOCKS);
        FCIS_HIP(hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, hs));
        FCIS_HIP(hipStreamSynchronize(hs));
        for (int i = 0; i < m; i++) {
                double s = 0.0;
                for (int b = 0; b < DOT_BLOCKS; b++) s += h[(size_t)i * DOT_BLOCKS + b];
                out[i] = s;
        }
        return FCIS_SUCCESS;
}

int fcis_lincomb(int m, int64_t n, const double *d_x, int64_t ldx, const double *coef,
                 double beta, double *d_y, void *stream)
{
        if (m < 0 || m > MAX_COEF || n < 0 || d_y == nullptr ||
            (m > 0 && (d_x == nullptr || coef == nullptr))) {
                fcis::set_error("fcis_lincomb: invalid argument (m=%d, max %d)", m, MAX_COEF);
                return FCIS_ERR_INVALID;
        }
        if (n == 0) return FCIS_SUCCESS;
        Coefs c;
        memset(&c, 0, sizeof(c));
        for (int i = 0; i < m; i++) c.c[i] = coef[i];
        hipLaunchKernelGGL(k_lincomb, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, m,
                           n, d_x, ldx, c, beta, d_y);
        FCIS_CHECK_LAUNCH();
        return FCIS_SUCCESS;
}

int fcis_precond(int64_t n, const double *d_hdiag, double e, double level_shift,
                 const double *d_r, double *d_out, void *stream)
{
        if (n < 0 || d_hdiag == nullptr || d_r == nullptr || d_out == nullptr) {
                fcis::set_error("fcis_precond: invalid argument");
                return FCIS_ERR_INVALID;
        }
        if (n == 0) return FCIS_SUCCESS;
        hipLaunchKernelGGL(k_precond, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n,
                           d_hdiag, e, level_shift, d_r, d_out);
        FCIS_CHECK_LAUNCH();
        return FCIS_SUCCESS;
}

int fcis_init_guess(int64_t n, int64_t addr, int root, double noise_norm, double *d_out,
                    void *stream)
{
        if (n <= 0 || addr < 0 || addr >= n || root < 0 || d_out == nullptr ||
            !(noise_norm >= 0.0)) {
                fcis::set_error("fcis_init_guess: invalid argument");
                return FCIS_ERR_INVALID;
        }
        uint64_t seed = (uint64_t)(root + 1) * 0x9E3779B97F4A7C15ull;
        double scale = noise_norm / sqrt((double)n);
        hipLaunchKernelGGL(k_init_guess, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                           n, addr, seed, scale, d_out);
        FCIS_CHECK_LAUNCH();
        return FCIS_SUCCESS;
}

int fcis_absorb_h1e(int norb, int nelec, double fac, const double *h1e, const double *eri,
                    double *h2e)
{
        if (norb < 1 || h1e == nullptr || eri == nullptr || h2e == nullptr) {
                fcis::set_error("fcis_absorb_h1e: invalid argument");
                return FCIS_ERR_INVALID;
        }
        int n = norb, np = n * (n + 1) / 2;
        std::vector<double> f((size_t)n * n);
        for (int i = 0; i < n; i++)
                for (int k = 0; k < n; k++) {
                        double s = 0;
                        for (int j = 0; j < n; j++)
                                s += eri[(size_t)fcis::pair_index(i, j) * np + fcis::pair_index(j, k)];
                        f[i * n + k] = (h1e[i * n + k] - 0.5 * s) * (1.0 / (nelec + 1e-100));
                }
        for (int i = 0; i < n; i++)
                for (int j = 0; j <= i; j++) {
                        size_t ij = fcis::pair_index(i, j);
                        for (int k = 0; k < n; k++)
                                for (int l = 0; l <= k; l++) {
                                        size_t kl = fcis::pair_index(k, l);
                                        double v = eri[ij * np + kl];
                                        if (k == l) v += f[i * n + j];
                                        if (i == j) v += f[k * n + l];
                                        h2e[ij * np + kl] = v * fac;
                                }
                }
        return FCIS_SUCCESS;
}

}  // extern "C"
