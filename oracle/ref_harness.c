/*
 * Thin C driver around the reference's own GAS-FCI library (libfci_gas),
 * used ONLY as a checker and as the CPU baseline.  It is compiled together
 * with the reference sources where they lie (pyscf/lib/mcscf/*.c); no
 * reference source is copied.  For a CAS problem the GAS space has one
 * subspace and one determinant block {na, nb}; the reference's tests assert
 * that this equals pyscf.fci.direct_spin1 (pyscf/mcscf/test/test_gasci.py:
 * 272-293 for sigma/hdiag, :765-786 for RDMs), and for one block the GAS
 * determinant order is the (nstr_alpha, nstr_beta) C-order of direct_spin1.
 *
 * Entry points follow the reference ABI in pyscf/lib/mcscf/fci_gas.h:
 *   gas_space_from_blocks          fci_gas.h:220, fci_string_gas.c:1683
 *   gas_space_compress_links       fci_gas.h:222
 *   fci_contract_gas_plan_*        fci_gas.h:242-249
 *   fci_make_hdiag_gas             fci_gas.h:205
 *   fci_rdm_gas_plan_*             fci_gas.h:283-291
 */
#include <stdlib.h>
#include <string.h>
#include "fci_gas.h"

typedef struct {
        gas_space_t gas;
        gas_contract_plan_t *plan;
        double *h2e;
        double *gos;
} refh_plan_t;

static int cas_space(gas_space_t *gas, int norb, int na, int nb, int compress)
{
        int norbs[1] = {norb};
        int occ[2] = {na, nb};
        memset(gas, 0, sizeof(*gas));
        if (gas_space_from_blocks(gas, 1, norbs, na, nb, 1, occ) != 0) {
                return -1;
        }
        if (compress && gas_space_compress_links(gas) != 0) {
                gas_space_free(gas);
                return -1;
        }
        return 0;
}

/* Reusable sigma plan; h2e is the absorbed (npair,npair) pair matrix.  The
 * plan keeps its own copies of h2e and gos = h2e + h2e^T (fci_gas.py:588). */
void *refh_plan_create(int norb, int na, int nb, const double *h2e)
{
        size_t npair = (size_t)norb * (norb + 1) / 2;
        refh_plan_t *p = calloc(1, sizeof(*p));
        if (p == 0) {
                return 0;
        }
        p->h2e = malloc(sizeof(double) * npair * npair);
        p->gos = malloc(sizeof(double) * npair * npair);
        if (p->h2e == 0 || p->gos == 0 || cas_space(&p->gas, norb, na, nb, 1) != 0) {
                free(p->h2e);
                free(p->gos);
                free(p);
                return 0;
        }
        memcpy(p->h2e, h2e, sizeof(double) * npair * npair);
        for (size_t i = 0; i < npair; i++) {
                for (size_t j = 0; j < npair; j++) {
                        p->gos[i * npair + j] = h2e[i * npair + j] + h2e[j * npair + i];
                }
        }
        if (fci_contract_gas_plan_create(&p->plan, &p->gas, p->h2e, p->gos) != 0) {
                gas_space_free(&p->gas);
                free(p->h2e);
                free(p->gos);
                free(p);
                return 0;
        }
        return p;
}

int refh_plan_execute(void *handle, const double *ci0, double *ci1)
{
        refh_plan_t *p = handle;
        return fci_contract_gas_plan_execute(p->plan, ci0, ci1);
}

void refh_plan_free(void *handle)
{
        refh_plan_t *p = handle;
        if (p == 0) {
                return;
        }
        fci_contract_gas_plan_free(p->plan);
        gas_space_free(&p->gas);
        free(p->h2e);
        free(p->gos);
        free(p);
}

int refh_contract_2e(int norb, int na, int nb, const double *h2e,
                     const double *ci0, double *ci1)
{
        void *p = refh_plan_create(norb, na, nb, h2e);
        int ret;
        if (p == 0) {
                return -1;
        }
        ret = refh_plan_execute(p, ci0, ci1);
        refh_plan_free(p);
        return ret;
}

int refh_make_hdiag(int norb, int na, int nb, const double *h1e,
                    const double *eri, double *hdiag)
{
        gas_space_t gas;
        int ret;
        if (cas_space(&gas, norb, na, nb, 0) != 0) {
                return -1;
        }
        ret = fci_make_hdiag_gas(&gas, h1e, eri, hdiag);
        gas_space_free(&gas);
        return ret;
}

int refh_trans_rdm12s(int norb, int na, int nb, const double *bra,
                      const double *ket, double *dm1a, double *dm1b,
                      double *dm2aa, double *dm2ab, double *dm2bb)
{
        gas_space_t gas;
        gas_rdm_plan_t *plan = 0;
        int ret;
        if (cas_space(&gas, norb, na, nb, 0) != 0) {
                return -1;
        }
        ret = fci_rdm_gas_plan_create(&plan, &gas);
        if (ret == 0) {
                ret = fci_rdm_gas_plan_make_rdm12s(plan, bra, ket, dm1a, dm1b,
                                                   dm2aa, dm2ab, dm2bb);
                fci_rdm_gas_plan_free(plan);
        }
        gas_space_free(&gas);
        return ret;
}

/* Raw link table of the alpha sector (spin = 0) or beta sector (spin = 1):
 * out[(src * nlink + k) * 4 + {0,1,2,3}] = (cre p, des q, dst addr, sign)
 * with a_p^+ a_q |src> = sign |dst>.  Returns nlink, or -1. */
int refh_link_table(int norb, int na, int nb, int spin, int *out)
{
        gas_space_t gas;
        int nlink;
        if (cas_space(&gas, norb, na, nb, 0) != 0) {
                return -1;
        }
        gas_sid_t s = spin == 0 ? gas.block[0].sa : gas.block[0].sb;
        const gas_link_table_t *t = gas_get_link_table(&gas, s, s);
        if (t == 0) {
                gas_space_free(&gas);
                return -1;
        }
        nlink = (int)t->nlink;
        for (uint32_t i = 0; i < t->nsrc; i++) {
                for (uint32_t k = 0; k < t->nlink; k++) {
                        const gas_link_entry_t *e = t->link + (size_t)i * t->nlink + k;
                        int *o = out + ((size_t)i * t->nlink + k) * 4;
                        o[0] = gas_link_cre(e);
                        o[1] = gas_link_des(e);
                        o[2] = (int)e->addr;
                        o[3] = e->sign;
                }
        }
        gas_space_free(&gas);
        return nlink;
}

/* Address of a string in the single CAS sector (spin 0 alpha, 1 beta). */
long refh_str2addr(int norb, int na, int nb, int spin, unsigned long long str)
{
        gas_space_t gas;
        long addr;
        if (cas_space(&gas, norb, na, nb, 0) != 0) {
                return -1;
        }
        gas_sid_t s = spin == 0 ? gas.block[0].sa : gas.block[0].sb;
        addr = (long)gas_str2addr_sector(&gas, s, str);
        gas_space_free(&gas);
        return addr;
}

/* ------------------------------------------------------------------------ */
/* Restricted (multi-space) GAS spaces, gas_space_from_blocks with the      */
/* caller's canonical block set D (fci_gas.h:220, fci_string_gas.c:1683).   */
/* ------------------------------------------------------------------------ */

typedef struct {
        int ngas, na, nb, nblock;
        int norb[GAS_MAX_NGAS];
        int *occ;                       /* [nblock][2*ngas] */
        gas_space_t raw;                /* raw links: layout, hdiag, RDMs */
} refh_gas_t;

static int gas_build(const refh_gas_t *h, gas_space_t *gas, int compress)
{
        memset(gas, 0, sizeof(*gas));
        if (gas_space_from_blocks(gas, h->ngas, h->norb, h->na, h->nb,
                                  h->nblock, h->occ) != 0) {
                return -1;
        }
        if (compress && gas_space_compress_links(gas) != 0) {
                gas_space_free(gas);
                return -1;
        }
        return 0;
}

void refh_gas_free(void *handle)
{
        refh_gas_t *h = handle;
        if (h == 0) {
                return;
        }
        gas_space_free(&h->raw);
        free(h->occ);
        free(h);
}

void *refh_gas_create(int ngas, const int *norb, int na, int nb,
                      int nblock, const int *block_occ)
{
        if (ngas <= 0 || ngas > GAS_MAX_NGAS || nblock <= 0) {
                return 0;
        }
        refh_gas_t *h = calloc(1, sizeof(*h));
        if (h == 0) {
                return 0;
        }
        h->ngas = ngas;
        h->na = na;
        h->nb = nb;
        h->nblock = nblock;
        memcpy(h->norb, norb, sizeof(int) * (size_t)ngas);
        h->occ = malloc(sizeof(int) * (size_t)nblock * 2 * ngas);
        if (h->occ == 0) {
                free(h);
                return 0;
        }
        memcpy(h->occ, block_occ, sizeof(int) * (size_t)nblock * 2 * ngas);
        if (gas_build(h, &h->raw, 0) != 0) {
                free(h->occ);
                free(h);
                return 0;
        }
        return h;
}

/* ndet, nblock, nsector of the built space. */
void refh_gas_info(void *handle, long *ndet, int *nblock, int *nsector)
{
        refh_gas_t *h = handle;
        *ndet = (long)h->raw.ndet;
        *nblock = (int)h->raw.nblock;
        *nsector = (int)h->raw.nsector;
}

/* Per block: offset, alpha/beta sector ids and string counts. */
void refh_gas_blocks(void *handle, long *offset, int *sa, int *sb,
                     long *na, long *nb)
{
        refh_gas_t *h = handle;
        for (uint32_t b = 0; b < h->raw.nblock; b++) {
                const gas_block_t *blk = h->raw.block + b;
                offset[b] = (long)blk->offset;
                sa[b] = blk->sa;
                sb[b] = blk->sb;
                na[b] = (long)h->raw.sector_nstr[blk->sa];
                nb[b] = (long)h->raw.sector_nstr[blk->sb];
        }
}

/* Per sector: occupations [nsector][ngas] and string count. */
void refh_gas_sectors(void *handle, int *occ, long *nstr)
{
        refh_gas_t *h = handle;
        for (uint32_t s = 0; s < h->raw.nsector; s++) {
                for (int g = 0; g < h->ngas; g++) {
                        occ[s * h->ngas + g] = h->raw.sector_occ[s * h->ngas + g];
                }
                nstr[s] = (long)h->raw.sector_nstr[s];
        }
}

/* All strings of one sector in address order. */
void refh_gas_sector_strings(void *handle, int sector, unsigned long long *out)
{
        refh_gas_t *h = handle;
        for (uint32_t a = 0; a < h->raw.sector_nstr[sector]; a++) {
                out[a] = gas_addr2str_sector(&h->raw, (gas_sid_t)sector, a);
        }
}

/* H c with the absorbed pair matrix h2e; gos = h2e + h2e^T (fci_gas.py:588). */
int refh_gas_contract(void *handle, const double *h2e, const double *ci0,
                      double *ci1)
{
        refh_gas_t *h = handle;
        gas_space_t gas;
        gas_contract_plan_t *plan = 0;
        size_t norb = 0, npair;
        double *gos;
        int ret;
        for (int g = 0; g < h->ngas; g++) {
                norb += (size_t)h->norb[g];
        }
        npair = norb * (norb + 1) / 2;
        gos = malloc(sizeof(double) * npair * npair);
        if (gos == 0 || gas_build(h, &gas, 1) != 0) {
                free(gos);
                return -1;
        }
        for (size_t i = 0; i < npair; i++) {
                for (size_t j = 0; j < npair; j++) {
                        gos[i * npair + j] = h2e[i * npair + j] + h2e[j * npair + i];
                }
        }
        ret = fci_contract_gas_plan_create(&plan, &gas, h2e, gos);
        if (ret == 0) {
                ret = fci_contract_gas_plan_execute(plan, ci0, ci1);
                fci_contract_gas_plan_free(plan);
        }
        gas_space_free(&gas);
        free(gos);
        return ret;
}

int refh_gas_hdiag(void *handle, const double *h1e, const double *eri,
                   double *hdiag)
{
        refh_gas_t *h = handle;
        return fci_make_hdiag_gas(&h->raw, h1e, eri, hdiag);
}

int refh_gas_rdm12s(void *handle, const double *bra, const double *ket,
                    double *dm1a, double *dm1b, double *dm2aa, double *dm2ab,
                    double *dm2bb)
{
        refh_gas_t *h = handle;
        gas_rdm_plan_t *plan = 0;
        int ret = fci_rdm_gas_plan_create(&plan, &h->raw);
        if (ret == 0) {
                ret = fci_rdm_gas_plan_make_rdm12s(plan, bra, ket, dm1a, dm1b,
                                                   dm2aa, dm2ab, dm2bb);
                fci_rdm_gas_plan_free(plan);
        }
        return ret;
}

/* Reusable restricted sigma plan (compressed-link space + contract plan with
 * its own h2e and gos copies), for timing the reference on a GAS space. */
typedef struct {
        gas_space_t gas;
        gas_contract_plan_t *plan;
        double *h2e, *gos;
} refh_gas_plan_t;

void refh_gas_plan_free(void *handle)
{
        refh_gas_plan_t *p = handle;
        if (p == 0) {
                return;
        }
        fci_contract_gas_plan_free(p->plan);
        gas_space_free(&p->gas);
        free(p->h2e);
        free(p->gos);
        free(p);
}

void *refh_gas_plan_create(void *space, const double *h2e)
{
        refh_gas_t *h = space;
        size_t norb = 0, npair;
        for (int g = 0; g < h->ngas; g++) {
                norb += (size_t)h->norb[g];
        }
        npair = norb * (norb + 1) / 2;
        refh_gas_plan_t *p = calloc(1, sizeof(*p));
        if (p == 0) {
                return 0;
        }
        p->h2e = malloc(sizeof(double) * npair * npair);
        p->gos = malloc(sizeof(double) * npair * npair);
        if (p->h2e == 0 || p->gos == 0 || gas_build(h, &p->gas, 1) != 0) {
                free(p->h2e);
                free(p->gos);
                free(p);
                return 0;
        }
        memcpy(p->h2e, h2e, sizeof(double) * npair * npair);
        for (size_t i = 0; i < npair; i++) {
                for (size_t j = 0; j < npair; j++) {
                        p->gos[i * npair + j] = h2e[i * npair + j] + h2e[j * npair + i];
                }
        }
        if (fci_contract_gas_plan_create(&p->plan, &p->gas, p->h2e, p->gos) != 0) {
                p->plan = 0;
                refh_gas_plan_free(p);
                return 0;
        }
        return p;
}

int refh_gas_plan_execute(void *handle, const double *ci0, double *ci1)
{
        refh_gas_plan_t *p = handle;
        return fci_contract_gas_plan_execute(p->plan, ci0, ci1);
}
