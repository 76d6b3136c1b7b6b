// Restricted (GAS) spaces: the layout shared by the CAS-embedded path
// (gas.hip) and the sector-native path (gas_native.hip).
#pragma once

#include <vector>

#include "common.h"
#include "fused.h"

#define GAS_MAXNGAS 15
#define GAS_MAXLOCAL 31         // orbitals per subspace (fci_gas.h:24-27)
#define GAS_MAXORB 63           // orbitals in all subspaces (fci_gas.h:24-27)

struct gas_native;
struct gas_native_plan;

struct fcis_gas_space {
        const fcis_space *cas;          // null for a sector-native space
        int ngas, na, nb;
        int norb[GAS_MAXNGAS], start[GAS_MAXNGAS];
        int nsector;
        std::vector<uint8_t> occ;       // [nsector][ngas]
        std::vector<int64_t> stride;    // [nsector][ngas]
        std::vector<int64_t> nstr;      // [nsector]
        int nblock;
        std::vector<int64_t> boff;      // [nblock]
        std::vector<int> bsa, bsb;      // [nblock]
        int64_t ndet;
        int64_t ncas;
        uint32_t *d_map;                // [ndet] CAS determinant index
        // Restricted rows of the fused CAS sigma (fcis_rowset): every alpha
        // string whose sector pairs with some beta sector in S1, the GAS
        // blocks and their single-excitation neighbours (sa', sb), (sa, sb')
        // with sa' (sb') one inter-subspace move from sa (sb) or equal, each
        // with the FUSED_NB-wide beta chunks holding strings of those beta
        // sectors.  t1 and vt1 vanish outside S1 when c lives on the GAS
        // determinants, so sigma on them is exact.
        int32_t *d_rows, *d_rpos;
        int2 *d_rowcl;
        uint16_t *d_clist;
        fcis_rowset rs;
        int64_t nchunk_rows;            // sum of chunk counts over rows
        gas_native *nat;                // sector-native tables (cas == null)
};

struct fcis_gas_sigma_plan {
        const fcis_gas_space *gas;
        fcis_sigma_plan_t *cas_plan;
        double *d_a, *d_b;              // CAS-sized embed / sigma buffers
        gas_native_plan *nat;           // sector-native plan (cas == null)
};

namespace fcis {

using SectorPool = std::vector<std::vector<uint8_t>>;

// Validated block rows, the sector pool (block sectors and their single
// inter-subspace moves, unique and sorted) and the block layout of g
// (occ, stride, nstr, boff, bsa, bsb, ndet).  Sets the error and returns a
// status on invalid input.
int gas_layout(fcis_gas_space *g, int ngas, const int *norb, int na, int nb, int nblock,
               const int *block_occ, SectorPool &pool, const char *who);

// Sector ids (in the pool) one inter-subspace move from sector s, s first.
std::vector<int> gas_moves(const fcis_gas_space *g, const SectorPool &pool, int s);

uint64_t gas_sector_string(const fcis_gas_space *g, int s, int64_t addr);

// Sector-native path (gas_native.hip).
int gas_native_create(fcis_gas_space *g, const SectorPool &pool, hipStream_t st);
void gas_native_free(gas_native *n);
int gas_native_hdiag(const fcis_gas_space *g, const double *h1e, const double *eri,
                     double *d_hdiag, hipStream_t st);
int gas_native_rdm(const fcis_gas_space *g, const double *d_bra, const double *d_ket,
                   double *dm1a, double *dm1b, double *dm2aa, double *dm2ab, double *dm2bb,
                   hipStream_t st);
int gas_native_plan_create(gas_native_plan **out, const fcis_gas_space *g, const double *h2e,
                           int precision, uint64_t workspace_bytes);
int gas_native_plan_execute(gas_native_plan *p, const double *d_ci0, double *d_ci1,
                            hipStream_t st);
void gas_native_plan_free(gas_native_plan *p);
// 2: closure t1 / GEMM / pull, 3: direct link-pair kernel.
int gas_native_plan_path(const gas_native_plan *p);
// d_ci1 = P S^2 P d_ci0, or the diagonal of S^2 when d_ci0 is null.
int gas_native_ss(const fcis_gas_space *g, const double *d_ci0, double *d_ci1, hipStream_t st);
// S1 size and the string counts of the native tables (zeros for a
// CAS-embedded space).
int gas_native_info(const fcis_gas_space *g, int64_t *nclosure, int64_t *maxrow, int64_t *nstra,
                    int64_t *nstrb);

// Slater-Condon H block over m determinants given by their strings
// (pspace.hip; h1e n x n, eri the (npair, npair) pair matrix, host).
int pspace_strings(int n, int m, const uint64_t *d_stra, const uint64_t *d_strb,
                   const double *h1e, const double *eri, double *d_h0, hipStream_t st);

}  // namespace fcis
