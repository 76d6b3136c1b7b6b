// Total spin S^2 on determinant vectors (the reference's S^2 contraction,
// _GasSpinPlan, pyscf/mcscf/fci_gas.py:266-419, used by contract_ss
// :993-1001 and the spin penalty of its GAS solver).
//
//   S^2 = S_z (S_z + 1) + S_- S_+,   S_- S_+ = N_b - sum_pq E^a_qp E^b_pq
//
// (E^s_pq = a+_ps a_qs), so for a destination determinant (Ia, Ib)
//   (S^2 c)[Ia, Ib] = [S_z(S_z+1) + N_b - |Ia & Ib|] c[Ia, Ib]
//                     - sum_{p != q} s_a s_b c[Ja, Jb],
// over the alpha links a+_q a_p Ia = s_a Ja of Ia (so <Ia|E^a_pq|Ja> = s_a)
// whose beta partner a+_p a_q Ib = s_b Jb exists (q occupied in Ib, p empty;
// <Ib|E^b_qp|Jb> = s_b).  The p = q terms are the |Ia & Ib| count.  One
// workgroup per alpha string, threads over beta strings: the alpha link is
// wave-uniform, the beta pair map row and the c row Ja are read coalesced /
// row-local.
//
// The spin ladder operators S+ / S- (fcis_spin_ladder) share the file.
#include <algorithm>
#include <new>

#include "common.h"

struct fcis_spin_plan {
        const fcis_space *space;
        int32_t *d_pmapb;               // [npair][nstrb]: (addr << 1) | neg, or -1
};

namespace {

__global__ void k_spin_pmap(uint32_t nstr, int nlink, const fcis_link *__restrict__ link,
                            int32_t *__restrict__ pmap)
{
        const uint32_t str = blockIdx.x * blockDim.x + threadIdx.x;
        if (str >= nstr) return;
        const fcis_link L = link[(size_t)blockIdx.y * nstr + str];
        pmap[(size_t)link_pair(L.info) * nstr + str] = (L.addr << 1) | link_neg(L.info);
}

__global__ void k_fill_i32(int32_t *__restrict__ x, int64_t n, int32_t v)
{
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
             i += (int64_t)gridDim.x * blockDim.x)
                x[i] = v;
}

__global__ __launch_bounds__(256) void k_contract_ss(int nlinka, int64_t nstra, int64_t nb,
                                                     double d0, const fcis_link *__restrict__ linka,
                                                     const int32_t *__restrict__ pmapb,
                                                     const uint64_t *__restrict__ stra,
                                                     const uint64_t *__restrict__ strb,
                                                     const double *__restrict__ c,
                                                     double *__restrict__ out)
{
        for (int64_t I = blockIdx.x; I < nstra; I += gridDim.x) {
                const uint64_t sa = stra[I];
                for (int64_t jb = threadIdx.x; jb < nb; jb += blockDim.x) {
                        const uint64_t sb = strb[jb];
                        double acc = (d0 - (double)__popcll(sa & sb)) * c[I * nb + jb];
                        for (int k = 0; k < nlinka; k++) {
                                const fcis_link L = linka[(int64_t)k * nstra + I];
                                const int q = link_cre(L.info), p = link_des(L.info);
                                if (p == q || !((sb >> q) & 1) || ((sb >> p) & 1)) continue;
                                const int32_t e = pmapb[(int64_t)link_pair(L.info) * nb + jb];
                                const double v = c[(int64_t)L.addr * nb + (e >> 1)];
                                acc -= ((link_neg(L.info) ^ (e & 1)) ? -v : v);
                        }
                        out[I * nb + jb] = acc;
                }
        }
}

__global__ void k_ss_diag(int64_t nstra, int64_t nb, double d0, const uint64_t *__restrict__ stra,
                          const uint64_t *__restrict__ strb, double *__restrict__ out)
{
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nstra * nb;
             i += (int64_t)gridDim.x * blockDim.x) {
                const int64_t a = i / nb;
                out[i] = d0 - (double)__popcll(stra[a] & strb[i - a * nb]);
        }
}

__constant__ fcis::BinomTable c_binom_spin = fcis::make_binom_table();

__device__ inline uint32_t colex_rank(uint64_t s)
{
        uint32_t addr = 0;
        for (int i = 1; s; i++, s &= s - 1) {
                const int p = __ffsll((unsigned long long)s) - 1;
                addr += p >= i ? c_binom_spin.v[p][i] : 0u;
        }
        return addr;
}

// S- (op = -1) or S+ (op = +1) from the (na, nb) vector c (row length
// nbs) onto destination determinant (I, jb) of the (na + op, nb - op)
// space: one workgroup per destination alpha string, threads over beta.
// The source of orbital p is (I ^ p, Ib ^ p) with p empty in I and
// occupied in Ib (S-) or the reverse (S+); its sign counts the occupied
// orbitals above p in both source strings, which equal those of I and Ib.
__global__ __launch_bounds__(256) void k_sladder(int op, int64_t nstra, int64_t nb, int64_t nbs,
                                                 double fac, const uint64_t *__restrict__ stra,
                                                 const uint64_t *__restrict__ strb,
                                                 const double *__restrict__ c,
                                                 double *__restrict__ out)
{
        for (int64_t I = blockIdx.x; I < nstra; I += gridDim.x) {
                const uint64_t sa = stra[I];
                for (int64_t jb = threadIdx.x; jb < nb; jb += blockDim.x) {
                        const uint64_t sb = strb[jb];
                        uint64_t cand = op < 0 ? (~sa & sb) : (sa & ~sb);
                        double acc = 0.0;
                        while (cand) {
                                const int p = __ffsll((unsigned long long)cand) - 1;
                                cand &= cand - 1;
                                const uint64_t bit = 1ull << p;
                                const double v = c[(int64_t)colex_rank(sa ^ bit) * nbs + colex_rank(sb ^ bit)];
                                const int par = (__popcll(sa >> (p + 1)) + __popcll(sb >> (p + 1))) & 1;
                                acc += par ? -v : v;
                        }
                        out[I * nb + jb] = fac * acc;
                }
        }
}

// Spin-dependent one-electron term, destination-owned: one workgroup per
// alpha string I, threads over beta strings.  The link (cre, des) of a
// string S gives <S'|E_{cre,des}|S> = sign, i.e. <S|E_{des,cre}|S'> = sign,
// so the destination's own links carry h[des][cre].
__global__ __launch_bounds__(256) void k_contract_1e_ab(int n, int nlinka, int nlinkb, int64_t nstra,
                                                        int64_t nb, int accumulate,
                                                        const fcis_link *__restrict__ linka,
                                                        const fcis_link *__restrict__ linkb,
                                                        const double *__restrict__ ha,
                                                        const double *__restrict__ hb,
                                                        const double *__restrict__ c,
                                                        double *__restrict__ out)
{
        for (int64_t I = blockIdx.x; I < nstra; I += gridDim.x) {
                for (int64_t jb = threadIdx.x; jb < nb; jb += blockDim.x) {
                        double acc = accumulate ? out[I * nb + jb] : 0.0;
                        for (int k = 0; k < nlinka; k++) {
                                const fcis_link L = linka[(int64_t)k * nstra + I];
                                const double v = ha[link_des(L.info) * n + link_cre(L.info)] *
                                                 c[(int64_t)L.addr * nb + jb];
                                acc += link_neg(L.info) ? -v : v;
                        }
                        for (int k = 0; k < nlinkb; k++) {
                                const fcis_link L = linkb[(int64_t)k * nb + jb];
                                const double v = hb[link_des(L.info) * n + link_cre(L.info)] *
                                                 c[I * nb + L.addr];
                                acc += link_neg(L.info) ? -v : v;
                        }
                        out[I * nb + jb] = acc;
                }
        }
}

double ss_base(const fcis_space *s)
{
        const double sz = 0.5 * (s->na - s->nb);
        return sz * (sz + 1.0) + s->nb;
}

}  // namespace

extern "C" {

void fcis_spin_plan_free(fcis_spin_plan_t *p)
{
        if (p == nullptr) return;
        if (p->d_pmapb) (void)hipFree(p->d_pmapb);
        delete p;
}

int fcis_spin_plan_create(fcis_spin_plan_t **out, const fcis_space_t *s, void *stream)
{
        if (out == nullptr || s == nullptr) {
                fcis::set_error("fcis_spin_plan_create: invalid argument");
                return FCIS_ERR_INVALID;
        }
        *out = nullptr;
        fcis_spin_plan *p = new (std::nothrow) fcis_spin_plan();
        if (p == nullptr) return FCIS_ERR_MEMORY;
        p->space = s;
        p->d_pmapb = nullptr;
        const int64_t n = (int64_t)s->npair * s->nstrb;
        if (fcis::dev_malloc(&p->d_pmapb, 4 * (size_t)n + 4) != hipSuccess) {
                fcis_spin_plan_free(p);
                fcis::set_error("fcis_spin_plan_create: device allocation failed");
                return FCIS_ERR_MEMORY;
        }
        hipStream_t st = (hipStream_t)stream;
        hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 8192) + 1),
                           dim3(256), 0, st, p->d_pmapb, n, -1);
        if (s->nlinkb > 0 && s->nstrb > 0)
                hipLaunchKernelGGL(k_spin_pmap, dim3((unsigned)((s->nstrb + 255) / 256), s->nlinkb),
                                   dim3(256), 0, st, (uint32_t)s->nstrb, s->nlinkb, s->d_linkb,
                                   p->d_pmapb);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
                fcis_spin_plan_free(p);
                fcis::set_error("fcis_spin_plan_create: %s", hipGetErrorString(e));
                return FCIS_ERR_DEVICE;
        }
        *out = p;
        return FCIS_SUCCESS;
}

int fcis_spin_plan_contract(const fcis_spin_plan_t *p, const double *d_ci0, double *d_ci1,
                            void *stream)
{
        if (p == nullptr || d_ci0 == nullptr || d_ci1 == nullptr || d_ci0 == d_ci1) {
                fcis::set_error("fcis_spin_plan_contract: invalid argument (distinct device "
                                "arrays required)");
                return FCIS_ERR_INVALID;
        }
        const fcis_space *s = p->space;
        if (s->nstra * s->nstrb == 0) return FCIS_SUCCESS;
        const unsigned grid = (unsigned)std::min<int64_t>(s->nstra, 65536);
        hipLaunchKernelGGL(k_contract_ss, dim3(grid), dim3(256), 0, (hipStream_t)stream, s->nlinka,
                           s->nstra, s->nstrb, ss_base(s), s->d_linka, p->d_pmapb, s->d_stra,
                           s->d_strb, d_ci0, d_ci1);
        FCIS_CHECK_LAUNCH();
        return FCIS_SUCCESS;
}

int fcis_spin_plan_diag(const fcis_spin_plan_t *p, double *d_out, void *stream)
{
        if (p == nullptr || d_out == nullptr) {
                fcis::set_error("fcis_spin_plan_diag: invalid argument");
                return FCIS_ERR_INVALID;
        }
        const fcis_space *s = p->space;
        const int64_t n = s->nstra * s->nstrb;
        if (n == 0) return FCIS_SUCCESS;
        hipLaunchKernelGGL(k_ss_diag, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 16384)),
                           dim3(256), 0, (hipStream_t)stream, s->nstra, s->nstrb, ss_base(s),
                           s->d_stra, s->d_strb, d_out);
        FCIS_CHECK_LAUNCH();
        return FCIS_SUCCESS;
}

int fcis_contract_1e_ab(const fcis_space_t *s, const double *d_h1a, const double *d_h1b,
                        const double *d_ci0, double *d_ci1, int accumulate, void *stream)
{
        if (s == nullptr || d_h1a == nullptr || d_h1b == nullptr || d_ci0 == nullptr ||
            d_ci1 == nullptr || d_ci0 == d_ci1) {
                fcis::set_error("fcis_contract_1e_ab: invalid argument (distinct device arrays)");
                return FCIS_ERR_INVALID;
        }
        if (s->nstra * s->nstrb == 0) return FCIS_SUCCESS;
        const unsigned grid = (unsigned)std::min<int64_t>(s->nstra, 65536);
        hipLaunchKernelGGL(k_contract_1e_ab, dim3(grid), dim3(256), 0, (hipStream_t)stream, s->norb,
                           s->nlinka, s->nlinkb, s->nstra, s->nstrb, accumulate ? 1 : 0, s->d_linka,
                           s->d_linkb, d_h1a, d_h1b, d_ci0, d_ci1);
        FCIS_CHECK_LAUNCH();
        return FCIS_SUCCESS;
}

int fcis_spin_ladder(const fcis_space_t *src, const fcis_space_t *dst, int op,
                     const double *d_ci0, double *d_ci1, void *stream)
{
        if (src == nullptr || dst == nullptr || d_ci0 == nullptr || d_ci1 == nullptr ||
            d_ci0 == d_ci1 || (op != -1 && op != 1)) {
                fcis::set_error("fcis_spin_ladder: invalid argument (op = -1 or +1, distinct "
                                "device arrays)");
                return FCIS_ERR_INVALID;
        }
        if (dst->norb != src->norb || dst->na != src->na + op || dst->nb != src->nb - op) {
                fcis::set_error("fcis_spin_ladder: destination space must be (na %+d, nb %+d) "
                                "of the source in the same orbitals", op, -op);
                return FCIS_ERR_INVALID;
        }
        if (src->d_perma || src->d_permb || dst->d_perma || dst->d_permb) {
                fcis::set_error("fcis_spin_ladder: colex-addressed spaces only");
                return FCIS_ERR_INVALID;
        }
        const int64_t n = dst->nstra * dst->nstrb;
        if (n == 0) return FCIS_SUCCESS;
        // passing the alpha string: na - 1 operators for S-, na for S+
        const double fac = ((op < 0 ? src->na - 1 : src->na) & 1) ? -1.0 : 1.0;
        const unsigned grid = (unsigned)std::min<int64_t>(dst->nstra, 65536);
        hipLaunchKernelGGL(k_sladder, dim3(grid), dim3(256), 0, (hipStream_t)stream, op, dst->nstra,
                           dst->nstrb, src->nstrb, fac, dst->d_stra, dst->d_strb, d_ci0, d_ci1);
        FCIS_CHECK_LAUNCH();
        return FCIS_SUCCESS;
}

}  // extern "C"
