/*
 * fcisigma -- MI355X (gfx950) direct-CI sigma engine, C ABI.
 *
 * Plain C types only.  Device arrays are HIP device pointers; `stream` is a
 * hipStream_t passed as void* (NULL = default stream).  Every entry point
 * returns an int status:
 *     FCIS_SUCCESS 0, FCIS_ERR_INVALID -1, FCIS_ERR_MEMORY -2,
 *     FCIS_ERR_DEVICE -3
 * mirroring gas_status_t of the reference (pyscf/lib/mcscf/fci_gas.h:38-42)
 * plus a device-error code; fcis_last_error() describes the last failure.
 *
 * CI vectors are (nstr_alpha, nstr_beta) row-major float64 on the device,
 * the layout of pyscf.fci.direct_spin1 CAS vectors.  Integrals are row-major
 * host arrays: h1e (norb,norb), two-electron tensors as (npair,npair) 4-fold
 * pair matrices, npair = norb(norb+1)/2, pair index max*(max+1)/2+min
 * (fci_gas.h:171-176 gas_link_pair_index).
 */
#ifndef FCISIGMA_H
#define FCISIGMA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
        FCIS_SUCCESS = 0,
        FCIS_ERR_INVALID = -1,
        FCIS_ERR_MEMORY = -2,
        FCIS_ERR_DEVICE = -3
};

/* Arithmetic of the t1 contraction (sigma plan). */
enum {
        FCIS_PREC_FP64 = 0,     /* FP64 MFMA                                  */
        FCIS_PREC_FP32 = 1,     /* FP32 MFMA, FP32 t1/vt1, FP64 gather/scatter */
        FCIS_PREC_BF16X3 = 2    /* 3-term bf16 split on MFMA, FP32 accumulate  */
};

typedef struct fcis_space fcis_space_t;
typedef struct fcis_sigma_plan fcis_sigma_plan_t;

/* Text of the last error on the calling thread. */
const char *fcis_last_error(void);

/* Library / device version info: writes "gfx950 ..." into buf. */
int fcis_device_info(char *buf, int buflen);

/* Plan workspaces of 256 MiB and more stay allocated after their plan is
 * freed, for the next plan of about the same size on that device (at most
 * four blocks per process; freeing and re-allocating blocks of tens of GB
 * costs up to about a second in the driver).  A failed allocation releases
 * them and retries.  fcis_kept_memory: bytes kept on `device` (-1: all);
 * fcis_release_kept_memory frees them.  No counterpart in the reference
 * (memory management of this library). */
int fcis_kept_memory(int device, uint64_t *bytes);
int fcis_release_kept_memory(int device);

/*
 * CAS string space (norb, na, nb): alpha/beta string addressing and link
 * tables, built ON THE DEVICE.  Replaces the string space and link-table
 * construction that pyscf.fci.direct_spin1 obtains from
 * cistring.gen_linkstr_index (called for the reference's solvers through
 * direct_spin1.contract_2e, pyscf/csf_fci/csf.py:449) and that libfci_gas
 * builds in gas_space_from_blocks / fill_link_table
 * (pyscf/lib/mcscf/fci_string_gas.c:1683, :1304-1427; ABI fci_gas.h:220).
 * norb <= 32, 0 <= na, nb <= norb.
 */
int fcis_space_create(fcis_space_t **out, int norb, int na, int nb, void *stream);
/* The same space with the caller's address order: stra / strb (host, nstr
 * each) list every alpha / beta string once, string a at address a.  Link
 * tables, sigma, diagonal, RDM, pspace and S^2 kernels work in that order
 * (the CSF transform needs the colex order).  Used for restricted spaces,
 * whose sectors are contiguous address ranges in sector-major order
 * (fcis_gas_string_order).  No counterpart in the reference (its restricted
 * spaces address strings per sector, fci_string_gas.c:462-489). */
int fcis_space_create_ordered(fcis_space_t **out, int norb, int na, int nb,
                              const uint64_t *stra, const uint64_t *strb, void *stream);
void fcis_space_free(fcis_space_t *space);
int fcis_space_info(const fcis_space_t *space, int64_t *nstr_alpha,
                    int64_t *nstr_beta, int *nlink_alpha, int *nlink_beta);

/*
 * Export a link table in pyscf's cistring.gen_linkstr_index layout:
 * int32 [nstr, nlink, 4] rows (cre, des, addr, sign), or with trilidx != 0
 * (pair, 0, addr, sign) as gen_linkstr_index_trilidx.  Row order: the
 * diagonal (i,i) entries over occupied i ascending, then occupied i
 * ascending x virtual a ascending.  spin 0 = alpha, 1 = beta.  d_out is a
 * device array.  (Replaces cistring.gen_linkstr_index; libfci_gas analogue
 * fill_link_table, fci_string_gas.c:1304.)
 */
int fcis_gen_linkstr_index(const fcis_space_t *space, int spin, int trilidx,
                           int32_t *d_out, void *stream);

/* String <-> address (colex rank, cistring.str2addr/addr2str;
 * fci_string_gas.c:263-298).  Host-side helpers. */
int64_t fcis_str2addr(int norb, int nelec, uint64_t string);
uint64_t fcis_addr2str(int norb, int nelec, int64_t addr);

/*
 * absorb_h1e: h2e = fac * [eri + delta_kl f_ij + delta_ij f_kl],
 * f = (h1e - 1/2 sum_j (ij|jk)) / (nelec + 1e-100).  Host arrays.
 * Replaces direct_spin1.absorb_h1e as wrapped at
 * pyscf/mcscf/fci_gas.py:132-138 (called by the reference at
 * fci_gas.py:1166 and csf.py:445 with fac = 0.5).
 */
int fcis_absorb_h1e(int norb, int nelec, double fac, const double *h1e,
                    const double *eri, double *h2e_out);

/*
 * Sigma plan: sigma = H c for an absorbed pair matrix h2e (host array,
 * copied).  Replaces fci_contract_gas_plan_create / _execute / _free
 * (fci_gas.h:242-249; fci_contract_gas.c:2388-2454) and the
 * direct_spin1.contract_2e call at csf.py:449 / fci_gas.py:919-939.
 * workspace_bytes bounds the t1/vt1 device workspace (0 = library default);
 * the plan tiles alpha strings so that the workspace fits.
 * precision: FCIS_PREC_*.  The plan must not be executed concurrently.
 */
int fcis_sigma_plan_create(fcis_sigma_plan_t **out, const fcis_space_t *space,
                           const double *h2e, int precision,
                           uint64_t workspace_bytes);
/* d_ci0 and d_ci1 are device (nstr_a, nstr_b) float64 arrays, must differ
 * (the reference rejects aliasing, fci_contract_gas.c:2426).  d_ci1 is fully
 * overwritten. */
int fcis_sigma_plan_execute(fcis_sigma_plan_t *plan, const double *d_ci0,
                            double *d_ci1, void *stream);
void fcis_sigma_plan_free(fcis_sigma_plan_t *plan);
/* Tiles per execute, alpha strings per tile, workspace bytes. */
int fcis_sigma_plan_info(const fcis_sigma_plan_t *plan, int *ntiles,
                         int64_t *tile_alpha, uint64_t *workspace_bytes);
/* 1 if the plan runs the fused row path (norb <= 16: gather, MFMA
 * contraction and beta scatter in one kernel, then an alpha pull; its beta
 * scatter accumulates with LDS atomics, so results are reproducible to
 * rounding only), 0 for the bitwise-reproducible three-kernel path
 * (gather, contraction, pull scatter; selected for larger spaces or with
 * the environment variable FCIS_DETERMINISTIC set at plan creation). */
int fcis_sigma_plan_path(const fcis_sigma_plan_t *plan);
/* Replace the Hamiltonian of an existing plan (same space). */
int fcis_sigma_plan_set_h2e(fcis_sigma_plan_t *plan, const double *h2e);
/* Per-kernel timing with HIP events recorded on the execute stream
 * (enable != 0 turns it on and clears the counters; execute then waits for
 * its last event).  ms[0..2] = accumulated gather / contraction / scatter
 * milliseconds, launches[0..2] = launch counts.  (Measurement hook, no
 * counterpart in the reference.) */
int fcis_sigma_plan_set_timing(fcis_sigma_plan_t *plan, int enable);
int fcis_sigma_plan_get_timing(const fcis_sigma_plan_t *plan, double *ms,
                               int64_t *launches);

/*
 * Hamiltonian diagonal from the un-absorbed integrals (host h1e, eri pair
 * matrix) into device d_hdiag (nstr_a*nstr_b).  Replaces fci_make_hdiag_gas
 * (fci_gas.h:205; fci_hdiag_gas.c:678) / direct_spin1.make_hdiag.
 */
int fcis_make_hdiag(const fcis_space_t *space, const double *h1e,
                    const double *eri, double *d_hdiag, void *stream);

/*
 * Spin-resolved transition 1- and 2-RDMs (pyscf conventions:
 * dm1[p,q] = <bra|q^+ p|ket>, dm2[p,q,r,s] = <bra|p^+ r^+ s q|ket>,
 * dm2ab alpha on (p,q), beta on (r,s)).  Device bra/ket, HOST outputs
 * (norb^2 and norb^4 float64).  Replaces fci_rdm_gas_plan_make_rdm12s
 * (fci_gas.h:287; fci_rdm_gas.c:2207) / direct_spin1.trans_rdm12s.
 * Any output pointer may be NULL to skip it (dm2 NULLs skip the 2-RDM).
 */
int fcis_trans_rdm12s(const fcis_space_t *space, const double *d_bra,
                      const double *d_ket, double *dm1a, double *dm1b,
                      double *dm2aa, double *dm2ab, double *dm2bb,
                      void *stream);
/* fcis_trans_rdm12s without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving or optimisation loops, harnesses that
 * forbid hipMalloc).  Same arguments, conventions and outputs; its
 * intermediates (the T arrays of an alpha tile, per-chunk partial products,
 * the Gram blocks and the assembled RDMs) live in one fixed per-device
 * scratch array of 128 MiB that the library holds on every GPU it runs on
 * until the process ends, tiled over alpha strings to fit.  Spaces whose
 * single alpha string needs more than that (nstr_beta * (norb^2 + 1)
 * roughly above 4 M, e.g. norb = 18 with 9 beta electrons) are rejected with
 * FCIS_ERR_INVALID.  Calls on one GPU share the array: calls that overlap
 * on that GPU (other streams or threads) race.  The call waits for its
 * stream before returning (host outputs). */
int fcis_trans_rdm12s_noalloc(const fcis_space_t *space, const double *d_bra,
                              const double *d_ket, double *dm1a, double *dm1b,
                              double *dm2aa, double *dm2ab, double *dm2bb,
                              void *stream);

/*
 * Hamiltonian block over m selected determinants (pspace): d_h0[i][j] =
 * <D_i|H|D_j> (device, m x m, row-major) for the CAS addresses d_addr[i] =
 * alpha_i * nstr_b + beta_i (device int64), by the Slater-Condon rules from
 * the HOST un-absorbed h1e and (npair, npair) eri.  m <= 46340.  Replaces
 * the sigma-per-column small-space builder (fci_gas.py:1130-1148) and the
 * determinant pspace under csf.pspace (csf.py:246-366).
 */
int fcis_pspace(const fcis_space_t *space, const double *h1e, const double *eri, int m,
                const int64_t *d_addr, double *d_h0, void *stream);
/* The same block with a spin-dependent one-electron part: h1a for the alpha
 * occupations and moves, h1b for the beta ones (host, norb x norb each).
 * Replaces FCIpspace_h0tril_uhf under csf.pspace (csf.py:305-320). */
int fcis_pspace_uhf(const fcis_space_t *space, const double *h1a, const double *h1b,
                    const double *eri, int m, const int64_t *d_addr, double *d_h0, void *stream);

/*
 * Total spin S^2 on CAS determinant vectors: a plan holding the beta pair
 * map (4 * npair * nstr_b bytes) on the space, which must outlive it.
 * contract: d_ci1 = S^2 d_ci0 (distinct device arrays); diag: the diagonal
 * of S^2.  Replaces the S^2 plan of the reference's GAS solver
 * (_GasSpinPlan.contract / diagonal_vector, fci_gas.py:266-419) and
 * pyscf.fci.spin_op.contract_ss for CAS vectors.
 */
typedef struct fcis_spin_plan fcis_spin_plan_t;
int fcis_spin_plan_create(fcis_spin_plan_t **out, const fcis_space_t *space, void *stream);
int fcis_spin_plan_contract(const fcis_spin_plan_t *plan, const double *d_ci0, double *d_ci1,
                            void *stream);
int fcis_spin_plan_diag(const fcis_spin_plan_t *plan, double *d_out, void *stream);
void fcis_spin_plan_free(fcis_spin_plan_t *plan);

/*
 * Spin ladder operators on CAS determinant vectors: op = -1 applies
 * S- = sum_p a+_{p beta} a_{p alpha}, from the (na, nb) space src to the
 * (na - 1, nb + 1) space dst; op = +1 applies S+ = sum_p a+_{p alpha}
 * a_{p beta}, to (na + 1, nb - 1).  Both spaces colex-addressed, same norb;
 * d_ci1 (dst-sized) is overwritten.  Signs follow pyscf's string
 * convention (a single creation or annihilation at p counts the occupied
 * orbitals above p), with (-1)^(na - 1) (S-) or (-1)^na (S+) for passing
 * the alpha string.  Replaces spin_op.contract_sladder
 * (pyscf/csf_fci/spin_op.py:22-71).
 */
int fcis_spin_ladder(const fcis_space_t *src, const fcis_space_t *dst, int op,
                     const double *d_ci0, double *d_ci1, void *stream);

/*
 * Spin-dependent one-electron contraction on CAS determinant vectors:
 * d_ci1 (+)= sum_pq h1a[p][q] E^alpha_pq ci0 + h1b[p][q] E^beta_pq ci0, with
 * h1a, h1b DEVICE norb x norb arrays; accumulate = 0 overwrites d_ci1.
 * Replaces direct_uhf.contract_1e as the CSF solver adds the spin part of
 * a two-component h1e to sigma (csf.py:518-529).
 */
int fcis_contract_1e_ab(const fcis_space_t *space, const double *d_h1a, const double *d_h1b,
                        const double *d_ci0, double *d_ci1, int accumulate, void *stream);

/* ---- Configuration state functions (spin-adapted basis of the CSF-FCI
 *      solver, pyscf/csf_fci) ----------------------------------------- */

/*
 * CSF transformer for spin multiplicity smult = 2S+1 on a CAS space.
 * Built on the device: the CSD -> determinant address map (make_csd_mask,
 * csdstring.py:28-47) and the per-block spin eigenvectors
 * sgn * umat (get_spin_evecs csfstring.py:913-960 -> FCICSFmakecsf
 * csfstring.c:86-141; phase csfstring.py:623-627).  CSF order is the
 * reference's: npair blocks, pair configuration, unpaired configuration,
 * coupling string (get_csfvec_shape csfstring.py:858-911).  The space must
 * outlive the transformer.  Fails with FCIS_ERR_INVALID if no state of that
 * spin exists for (na, nb).
 */
typedef struct fcis_csf fcis_csf_t;
int fcis_csf_create(fcis_csf_t **out, const fcis_space_t *space, int smult,
                    void *stream);
void fcis_csf_free(fcis_csf_t *csf);
int fcis_csf_info(const fcis_csf_t *csf, int64_t *ncsf, int64_t *ndet);
/* nvec determinant vectors (rows of ndet) -> nvec CSF vectors (rows of
 * ncsf), no normalisation (_transform_det2csf csfstring.py:569-674 /
 * vec_det2csf(normalize=False) :45-52). */
int fcis_csf_det2csf(const fcis_csf_t *csf, const double *d_det, double *d_csf,
                     int nvec, void *stream);
/* nvec CSF vectors -> determinant vectors; determinants of spin blocks with
 * no CSF of this multiplicity are set to zero (transform_civec_csf2det
 * csfstring.py:416-487 / vec_csf2det(normalize=False) :54-60).
 * Both transforms run one GEMM per spin block on a CSD-ordered work array
 * the transformer owns (grown to nvec vectors on first use, nvec <= 65535):
 * calls on one transformer must not overlap. */
int fcis_csf_csf2det(const fcis_csf_t *csf, const double *d_csf, double *d_det,
                     int nvec, void *stream);
/* Diagonal of H in the CSF basis (make_hdiag_csf csf.py:70-144 ->
 * FCICSFhdiag csfstring.c:295-361): diag(U^T H_conf U) per configuration,
 * from the determinant diagonal d_hdiag_det (fcis_make_hdiag) and the
 * exchange integrals of the HOST (npair, npair) pair matrix eri. */
int fcis_csf_hdiag(fcis_csf_t *csf, const double *eri, const double *d_hdiag_det,
                   double *d_hdiag_csf, void *stream);

/*
 * Point-group packing of CSF vectors (csf_symm / CSFTransformer.pack_csf,
 * pyscf/csf_fci/csfstring.py:74-103): with orbital irreps orbsym[norb]
 * (D2h-subgroup ids 0..7, products by XOR), keep the CSFs whose
 * configuration irrep (XOR over its singly occupied orbitals) is wfnsym;
 * *nsym receives their count (orbsym NULL clears the packing, nsym = ncsf).
 * pack: d_packed[v] (nsym) = the kept entries of d_full[v] (ncsf); unpack:
 * the reverse, zero elsewhere.  irrep: the irrep of CSF `csf`.
 */
int fcis_csf_symmetry(fcis_csf_t *csf, const int *orbsym, int wfnsym, int64_t *nsym,
                      void *stream);
int fcis_csf_pack(const fcis_csf_t *csf, const double *d_full, double *d_packed, int nvec,
                  void *stream);
int fcis_csf_unpack(const fcis_csf_t *csf, const double *d_packed, double *d_full, int nvec,
                    void *stream);
int fcis_csf_irrep(const fcis_csf_t *csf, const int *orbsym, int64_t csf_index, int *irrep);

/*
 * Configuration and spin coupling of full-space CSF csf: the bit masks of
 * its doubly and singly occupied orbitals and its genealogical coupling
 * string (bit i = 1: the i-th singly occupied orbital, counted from the
 * lowest, couples up).  Replaces csfaddrs2str (csfstring.py:1169-1181),
 * behind printable_csfstring (:1183-1207).
 */
int fcis_csf_describe(const fcis_csf_t *t, int64_t csf, uint64_t *docc, uint64_t *socc,
                      uint64_t *coupling);

/*
 * Determinant supports of m full-space CSFs (device d_csf[m]): row i of the
 * device arrays d_addr / d_coef [m][maxdet] holds the determinant addresses
 * and coefficients of CSF i (nonzero entries, CSD order), d_count[i] their
 * number (-1: index out of range).  maxdet >= fcis_csf_max_support (the
 * largest configuration's determinant count).  The configurations' rows of
 * U that the reference gathers for its CSF pspace (csf.py:282-366,
 * transform_opmat_det2csf_pspace csfstring.py:676-787).
 */
int fcis_csf_max_support(const fcis_csf_t *t, int *maxdet);
int fcis_csf_support(const fcis_csf_t *t, int m, const int64_t *d_csf, int maxdet, int64_t *d_addr,
                     double *d_coef, int *d_count, void *stream);

/* ---- Many CAS problems per launch ---------------------------------------
 *
 * A batch of nprob independent CAS problems (spaces[p], absorbed pair
 * matrices h2e[p] as in fcis_sigma_plan_create), each with nvec[p] <= 64
 * vector slots (Davidson roots).  Vectors of all problems are packed into
 * one flat device array: slot v of problem p starts at offsets[p] +
 * v * ndet_p (fcis_batch_info).  fcis_batch_sigma computes, for every
 * active slot, d_out = H_p d_in in one persistent kernel launch (replayed
 * from a hipGraph captured on the first call with a given buffer pair) and
 * zeroes inactive slots.  norb <= 16 per problem.  The spaces must outlive
 * the batch.  The reference's solvers contract one problem per call
 * (direct_spin1.contract_2e, pyscf/csf_fci/csf.py:449); this entry point
 * has no single counterpart: it is the contract_2e of many fcisolver.kernel
 * calls at once.
 */
typedef struct fcis_batch fcis_batch_t;
int fcis_batch_create(fcis_batch_t **out, int nprob, const fcis_space_t *const *spaces,
                      const double *const *h2e, const int *nvec, int precision,
                      void *stream);
void fcis_batch_free(fcis_batch_t *batch);
/* total packed length; offsets (host array of nprob, may be NULL); the
 * current schedule (alpha strings per work item, persistent workgroups). */
int fcis_batch_info(const fcis_batch_t *batch, int64_t *total, int64_t *offsets,
                    int *rows_per_item, int *grid);
/* Host mask of sum(nvec) bytes, problem-major: nonzero = compute the slot. */
int fcis_batch_set_active(fcis_batch_t *batch, const uint8_t *mask, void *stream);
/* enable = 0 launches eagerly instead of through the captured graph. */
int fcis_batch_set_graph(fcis_batch_t *batch, int enable);
int fcis_batch_sigma(fcis_batch_t *batch, const double *d_in, double *d_out,
                     void *stream);
/* Autotuner: times candidate schedules (alpha strings per work item x grid
 * size) on these buffers (d_out is overwritten), keeps the fastest and
 * returns its time in *best_ms.  (Tile scheduling of the reference's
 * OpenMP tasks is a cost model instead, fci_contract_gas.c:1845-2036.) */
int fcis_batch_tune(fcis_batch_t *batch, const double *d_in, double *d_out, int reps,
                    double *best_ms, void *stream);

/* Block vector algebra on many device vectors of different length, one
 * launch per call (the Davidson work of many solves at once:
 * pyscf.lib.davidson1's projections, axpy and precond, called at
 * fci_gas.py:1337-1350).  An entry names nx vectors X = xs[xoff ..
 * xoff+nx) and ny vectors Y = ys[yoff .. yoff+ny), all of length n
 * (xs, ys: HOST arrays of device pointers). */
typedef struct {
        int64_t n;
        int32_t nx, ny;        /* gemm: nx <= 128 */
        int32_t xoff, yoff;
        int32_t ooff;          /* gram: first output; gemm: first coefficient */
        int32_t pad;
        double beta;           /* gemm only; beta == 0: Y is not read */
} fcis_block_entry;
typedef struct {
        const double *hdiag, *r;
        double *out;           /* out = r / (hdiag - e), |.| < level_shift -> level_shift */
        int64_t n;
        double e;
} fcis_precond_entry;
/* Ordering: every list call runs on `stream` after the work queued there
 * before it.  fcis_gram_list returns when its results are in `out`;
 * fcis_gemm_list and fcis_precond_list return once their work is queued
 * (the host arrays they take are copied before they return), so the caller
 * reads their outputs through work on the same stream or after
 * synchronising it. */
/* out[ooff + i*ny + j] = <X_i, Y_j> (out: HOST array of nout; partial sums
 * added in a fixed order). */
int fcis_gram_list(int nent, const fcis_block_entry *ent, int nxs, const double *const *xs,
                   int nys, const double *const *ys, int nout, double *out, void *stream);
/* Y_j = beta Y_j + sum_i coef[ooff + i*ny + j] X_i (coef: HOST array of
 * ncoef).  Y may alias X (in-place block transforms) when ny <= 16. */
int fcis_gemm_list(int nent, const fcis_block_entry *ent, int nxs, const double *const *xs,
                   int nys, double *const *ys, int ncoef, const double *coef, void *stream);
int fcis_precond_list(int nent, const fcis_precond_entry *ent, double level_shift,
                      void *stream);

/* ---- Davidson primitives on device vectors (pyscf.lib.davidson1 work,
 *      called at fci_gas.py:1344-1350; vectors stay in HBM) ----------- */

/* out[i] = <X_i, v>, i < m; X rows of length n with row stride ldx.
 * Deterministic two-stage reduction.  out is a HOST array. */
int fcis_dots(int m, int64_t n, const double *d_x, int64_t ldx,
              const double *d_v, double *out, void *stream);
/* y = beta*y + sum_i coef[i] X_i  (coef: host array of m <= 64 values). */
int fcis_lincomb(int m, int64_t n, const double *d_x, int64_t ldx,
                 const double *coef, double beta, double *d_y, void *stream);
/* out = r / (hdiag - e), |hdiag - e| < level_shift -> level_shift
 * (precond, fci_gas.py:1337-1342).  out may alias r. */
int fcis_precond(int64_t n, const double *d_hdiag, double e,
                 double level_shift, const double *d_r, double *d_out,
                 void *stream);
/* Initial guess of root `root`: unit vector at `addr` plus the reproducible
 * splitmix64 sign noise of total norm noise_norm over all n entries
 * (get_init_guess / _add_distributed_noise, fci_gas.py:61-88, :881-902). */
int fcis_init_guess(int64_t n, int64_t addr, int root, double noise_norm,
                    double *d_out, void *stream);

/* ---- Restricted (GAS) determinant spaces (pyscf/mcscf/fci_gas.py,
 *      libfci_gas pyscf/lib/mcscf/fci_gas.h) ----------------------------
 *
 * GAS vectors are flat float64 device arrays in the reference's canonical
 * block order: blocks in the order of the block set D, block b an
 * alpha-major (nstr(sa), nstr(sb)) rectangle at offset[b] (fci_gas.h:52-60).
 * The restricted Hamiltonian is P H P with P the projection on the GAS
 * determinants; the device path embeds into the CAS array of the borrowed
 * CAS space, applies the CAS kernels and projects back. */
typedef struct fcis_gas_space fcis_gas_space_t;
typedef struct fcis_gas_sigma_plan fcis_gas_sigma_plan_t;

/* Space of the block set D (nblock rows of 2*ngas ints: alpha occupation
 * per subspace, then beta), unique and lexicographically sorted, on the CAS
 * space `cas` (norb = sum(norb[]), na, nb), which must outlive it.
 * ngas <= 15, 1 <= norb[k] <= 31; the CAS space must hold fewer than 2^32
 * determinants.  Sector ids, block offsets and sector string addresses
 * equal the reference's.  Replaces gas_space_from_blocks (fci_gas.h:220,
 * fci_string_gas.c:1683). */
int fcis_gas_space_create(fcis_gas_space_t **out, const fcis_space_t *cas, int ngas,
                          const int *norb, int nblock, const int *block_occ, void *stream);
void fcis_gas_space_free(fcis_gas_space_t *gas);       /* gas_space_free */
/* All strings of nelec electrons in sum(norb[]) orbitals, sorted by their
 * occupation tuple over the ngas subspaces (the reference's sector order),
 * colex within a tuple: the sector-major address order for
 * fcis_space_create_ordered.  out: host, C(sum(norb), nelec) entries. */
int fcis_gas_string_order(int ngas, const int *norb, int nelec, uint64_t *out);
/* Sequences of two-element rotations on fibers of a device array, the string
 * axes of GAS blocks under an orbital rotation within each subspace
 * (transform_ci_within_gas, fci_gas.py:1003-1091): for every fiber f (the
 * elements d_x[base[f] + i * stride[f]], i < extent <= 16384) and rotation
 * r in order, each pair k in [roff[r], roff[r+1]) becomes
 *   x[left[k]] <- c x[left[k]] + s x[right[k]],
 *   x[right[k]] <- -s x[left[k]] + c x[right[k]],   (c, s) = cs[2r], cs[2r+1],
 * then x[i] *= sign[i].  All arrays device-resident. */
int fcis_rotate_fibers(double *d_x, int64_t nfiber, const int64_t *d_base, const int64_t *d_stride,
                       int extent, int nrot, const int32_t *d_roff, const int32_t *d_left,
                       const int32_t *d_right, const double *d_cs, const double *d_sign, void *stream);
/* Determinants, blocks and sectors (gas_space_t ndet/nblock/nsector). */
int fcis_gas_space_info(const fcis_gas_space_t *gas, int64_t *ndet, int *nblock,
                        int *nsector);
/* Per block (host arrays of nblock): offset, alpha/beta sector ids, string
 * counts (GasSpace.block_descriptors, fci_gas.py:533-542). */
int fcis_gas_block_info(const fcis_gas_space_t *gas, int64_t *offset, int *sa, int *sb,
                        int64_t *na, int64_t *nb);
/* Per sector: occupations (host [nsector][ngas]) and string counts. */
int fcis_gas_sector_info(const fcis_gas_space_t *gas, int *occ, int64_t *nstr);
/* gas_addr2str_sector / gas_str2addr_sector (fci_string_gas.c:462-489). */
int fcis_gas_addr2str(const fcis_gas_space_t *gas, int sector, int64_t addr, uint64_t *str);
int fcis_gas_str2addr(const fcis_gas_space_t *gas, int sector, uint64_t str, int64_t *addr);
/* d_cas[v] = gas2fci(d_gas[v]) and d_gas[v] = fci2gas(d_cas[v]) for nvec
 * vectors (fci_gas.py:141-208); d_cas is fully overwritten. */
int fcis_gas_embed(const fcis_gas_space_t *gas, const double *d_gas, double *d_cas, int nvec,
                   void *stream);
int fcis_gas_project(const fcis_gas_space_t *gas, const double *d_cas, double *d_gas,
                     int nvec, void *stream);
/* d_cas_addr[i] = CAS determinant index (alpha * nstr_b + beta) of GAS
 * determinant d_addr[i] (device int64 arrays; -1 for an address outside
 * the space): GAS pspaces through fcis_pspace. */
int fcis_gas_det_index(const fcis_gas_space_t *gas, int64_t n, const int64_t *d_addr,
                       int64_t *d_cas_addr, void *stream);
/* Diagonal of H in GAS order (fci_make_hdiag_gas, fci_gas.h:205). */
int fcis_gas_make_hdiag(const fcis_gas_space_t *gas, const double *h1e, const double *eri,
                        double *d_hdiag, void *stream);
/* Spin-resolved transition RDMs of GAS vectors, conventions and host
 * outputs as fcis_trans_rdm12s (fci_rdm_gas_plan_make_rdm12s, fci_gas.h:287). */
int fcis_gas_trans_rdm12s(const fcis_gas_space_t *gas, const double *d_bra,
                          const double *d_ket, double *dm1a, double *dm1b, double *dm2aa,
                          double *dm2ab, double *dm2bb, void *stream);
/* Reusable restricted sigma plan: d_ci1 = P H P d_ci0 with the absorbed
 * pair matrix h2e, precision and workspace as fcis_sigma_plan_create.  The
 * space must outlive the plan; d_ci0 != d_ci1.  Replaces
 * fci_contract_gas_plan_create / _execute / _free (fci_gas.h:242-249). */
int fcis_gas_sigma_plan_create(fcis_gas_sigma_plan_t **out, const fcis_gas_space_t *gas,
                               const double *h2e, int precision, uint64_t workspace_bytes);
int fcis_gas_sigma_plan_execute(fcis_gas_sigma_plan_t *plan, const double *d_ci0,
                                double *d_ci1, void *stream);
void fcis_gas_sigma_plan_free(fcis_gas_sigma_plan_t *plan);
/* d_ci1 = P H P d_ci0 with a caller's sigma plan on the space's CAS space
 * (any precision; the Davidson driver's plans), d_work a device array of
 * 2 * nstr_a * nstr_b doubles.  With a fused plan (norb <= 16) the fused
 * kernel runs only over the alpha strings and 16-determinant beta chunks
 * within one excitation of the GAS space (fcis_gas_rows_info), where t1
 * and vt1 can be nonzero; otherwise the full CAS sigma is projected.  The
 * work of fci_contract_gas_plan_execute (fci_gas.h:245). */
int fcis_gas_sigma_execute(const fcis_gas_space_t *gas, fcis_sigma_plan_t *cas_plan,
                           const double *d_ci0, double *d_ci1, double *d_work, void *stream);
/* d_ci1 = P S^2 P d_ci0 on GAS vectors with a spin plan on the space's CAS
 * space (exact S^2 when D is spin complete, as the reference requires,
 * fci_gas.py:993-1001), d_work 2 * nstr_a * nstr_b doubles.  A sector-native
 * space applies S^2 = S_z(S_z+1) + N_b - sum_pq E^a_pq E^b_qp on its own
 * tables; spin and d_work are then ignored and may be NULL. */
int fcis_gas_contract_ss(const fcis_gas_space_t *gas, const fcis_spin_plan_t *spin,
                         const double *d_ci0, double *d_ci1, double *d_work, void *stream);
/* Restricted work of the fused path: alpha rows, beta chunks over those
 * rows, and the chunk count of the full CAS (nstr_a * ceil(nstr_b / 16)). */
int fcis_gas_rows_info(const fcis_gas_space_t *gas, int64_t *nrows, int64_t *nchunks,
                       int64_t *nchunks_cas);

/* Sector-native space of the block set D (arguments as fcis_gas_space_create,
 * with the electron counts na, nb instead of a CAS space): the same sector
 * ids, block offsets and sector string addresses, and no CAS embedding.
 * sum(norb[]) <= 63, norb[k] <= 31, ngas <= 15, as libfci_gas allows
 * (fci_gas.h:24-27); fewer than 2^31 strings per spin in the sectors one
 * move from D.  fcis_gas_sigma_plan_* (FP64 or FP32), fcis_gas_make_hdiag,
 * fcis_gas_trans_rdm12s, fcis_gas_pspace and the descriptor queries accept
 * it; the CAS-embedding entry points (embed, project, det_index,
 * sigma_execute, contract_ss, rows_info) reject it.  Replaces
 * gas_space_from_blocks (fci_gas.h:220, fci_string_gas.c:1683) and with it
 * the sector-addressed contraction, diagonal and RDM plans of libfci_gas
 * (fci_contract_gas.c:2388-2454, fci_hdiag_gas.c:678-709,
 * fci_rdm_gas.c:2207-2243). */
int fcis_gas_space_create_native(fcis_gas_space_t **out, int ngas, const int *norb, int na,
                                 int nb, int nblock, const int *block_occ, void *stream);
/* Path of a restricted sigma plan: 0 / 1 the CAS plan's three-kernel /
 * fused path (CAS-embedded space); on a sector-native space 2 = t1 over
 * the single-excitation closure, GEMM and pull, 3 = the operator form
 * (same-spin one-spin operators with their two-replacement entries,
 * opposite-spin sorted link lists; tables kept with the space, entries
 * with the plan), 4 = the link-pair loops without tables.  Chosen by
 * estimated time (3 and 4 compute in FP64 for either precision); the
 * environment variable FCIS_GAS_NATIVE_PATH=t1 / direct / links at plan
 * creation forces one. */
int fcis_gas_sigma_plan_path(const fcis_gas_sigma_plan_t *plan);
/* Diagonal of S^2 on the determinants of a sector-native space (device
 * d_out, ndet doubles; _GasSpinPlan.diagonal_vector, fci_gas.py:374-398). */
int fcis_gas_ss_diag(const fcis_gas_space_t *gas, double *d_out, void *stream);
/* native = 1 for a sector-native space; nclosure = determinants within one
 * excitation of D (the t1 columns of its sigma), maxrow = the longest
 * closure row (the fewest t1 columns a workspace must hold), nstra / nstrb =
 * strings of the sectors one move from D per spin (zeros for a CAS-embedded
 * space). */
int fcis_gas_native_info(const fcis_gas_space_t *gas, int *native, int64_t *nclosure,
                         int64_t *maxrow, int64_t *nstra, int64_t *nstrb);
/* Alpha and beta strings of the GAS determinants addr[i], i < n (host
 * arrays; gas_addr2str_sector over the block layout, fci_string_gas.c:462-489). */
int fcis_gas_det_strings(const fcis_gas_space_t *gas, int64_t n, const int64_t *addr,
                         uint64_t *stra, uint64_t *strb);
/* H block over the m GAS determinants addr[] (host int64) into device d_h0
 * (m x m), Slater-Condon rules on the un-absorbed h1e and pair matrix, for
 * either kind of space (the reference's small-space builder,
 * fci_gas.py:955-988, 1130-1148). */
int fcis_gas_pspace(const fcis_gas_space_t *gas, const double *h1e, const double *eri, int m,
                    const int64_t *addr, double *d_h0, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FCISIGMA_H */
