/*
 * mjk.h -- C ABI of the MI355X direct Coulomb/exchange (J/K) engine.
 *
 * Drop-in boundary for PySCF's direct J/K hot path.  The reference crosses
 * from Python into C once per J/K build through ctypes:
 *
 *   void CVHFnr_direct_drv(int (*intor)(), void (*fdot)(), JKOperator **jkop,
 *                          double **dms, double **vjk, int n_dm, int ncomp,
 *                          int *shls_slice, int *ao_loc, CINTOpt *cintopt,
 *                          CVHFOpt *vhfopt, int *atm, int natm, int *bas,
 *                          int nbas, double *env);
 *   (pyscf/lib/vhf/nr_direct.h:74-78, called from pyscf/scf/_vhf.py:540-639)
 *
 * with the screening context CVHFOpt built by _VHFOpt
 * (pyscf/lib/vhf/optimizer.h:23-34, pyscf/scf/_vhf.py:151-264).  Function
 * pointers to libcint and to the digestion operators do not cross a device
 * boundary, so the same work is exposed as a handle-based API: the basis
 * (the reference's atm/bas/env arrays, unchanged) is uploaded once and
 * keeps shell pairs, Schwarz bounds and work buffers resident in HBM; each
 * J/K build passes densities and receives J and K.
 *
 * Conventions (all arrays C-contiguous):
 *   atm  int32 [natm][6], bas int32 [nbas][8], env float64 -- PySCF layout
 *        (pyscf/gto/mole.py:58-88), spherical AOs, ao_loc = cumsum((2l+1)*nctr).
 *   dms  float64 [n_dm][nao][nao]; vj, vk float64 [n_dm][nao][nao].
 *   For a batch of molecules, dms/vj/vk hold the molecules' [nao_m][nao_m]
 *   blocks back to back inside every density slot.
 *   hermi: 0 any dm, 1 symmetric dm (K mirrored from the lower triangle),
 *          2 anti-symmetric (upper triangle of K = -lower).  J is always the
 *          symmetric Coulomb matrix (pyscf/scf/_vhf.py:405-464).
 *   direct_scf_tol <= 0: no screening (module-level get_jk without an
 *          optimizer, pyscf/scf/hf.py:948-1022); > 0: the CVHFnrs8_prescreen
 *          predicate with Schwarz q_cond and dm_cond
 *          (pyscf/lib/vhf/optimizer.c:90-117, 408-454, 494-518).
 * Return codes: 0 success, < 0 error (mjk_last_error gives the message);
 * the library never exits the process.
 */
#ifndef MJK_H
#define MJK_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MJK_OK 0
#define MJK_EINVAL (-1)       /* bad argument or shape */
#define MJK_ENOMEM (-2)       /* request exceeds the context's HBM budget */
#define MJK_EDEVICE (-3)      /* HIP runtime or kernel failure */
#define MJK_EUNSUPPORTED (-4) /* angular momentum or feature not built */
#define MJK_EINTERNAL (-5)    /* internal consistency check failed */

typedef struct mjk_ctx mjk_ctx;
typedef struct mjk_basis mjk_basis;

typedef struct {
    double ms_total;            /* host wall time of the call */
    double ms_quartets;         /* GPU time of the ERI/J/K kernels (HIP events) */
    double ms_transform;        /* GPU time of dm_cond and sph<->cart transforms */
    long long quartets;         /* evaluated (post-screening) shell quartets */
    long long prim_quartets;    /* evaluated primitive quartets */
    double flops;               /* algorithmic FLOPs of the evaluated quartets */
    long long quartets_fp32;    /* of which evaluated in FP32 (mixed precision) */
    double flops_fp32;          /* algorithmic FLOPs of the FP32 quartets */
} mjk_stats;

/* Context on one HIP device; hbm_budget = 0 means 80% of free memory.
 * Replaces the process-wide state the reference keeps in OpenMP/ctypes
 * (pyscf/lib/misc.py:95-107 load_library, pyscf/lib/misc.py:125-154). */
int mjk_ctx_create(mjk_ctx** ctx, int device, size_t hbm_budget);
void mjk_ctx_destroy(mjk_ctx* ctx);
const char* mjk_last_error(const mjk_ctx* ctx);
/* Bytes of HBM currently held by the context. */
size_t mjk_ctx_hbm_used(const mjk_ctx* ctx);
/* Kernel selection knob (no counterpart in the reference, whose loop has a
 * single CPU code path): the largest LDS allocation, in bytes, that a
 * row-kernel block may take (one block per bra pair accumulating its K'
 * rows in LDS).  Classes whose rows do not fit this limit, or with the
 * kernel's static LDS (up to 12 KB: the Rys table copy of the one- and
 * two-root classes) the workgroup's 160 KB, run their fallback kernel (the
 * tile kernel up to 81 Cartesian components per quartet, the cooperative
 * kernel above); 0 runs the fallback kernels only.  Default and maximum
 * 162816 (160 KB less 1 KB).  Results agree either way to rounding. */
int mjk_ctx_set_row_lds(mjk_ctx* ctx, size_t bytes);

/* No reference counterpart (a launch-overhead option): a build without
 * statistics or profiling is captured once into a HIP graph and replayed
 * while its inputs, flags and buffers are unchanged (one graph launch per
 * build instead of one launch per kernel).  On by default; 0 runs every
 * build eagerly.  Device memory that any thread of the process frees while
 * a capture is in flight is released after the capture ends, so captures
 * and frees in other threads may overlap. */
int mjk_ctx_set_graphs(mjk_ctx* ctx, int enable);

/* Upload one molecule's basis and build its shell pairs, Schwarz bounds
 * (the q_cond of CVHFnr_int2e_q_cond, pyscf/lib/vhf/optimizer.c:408-454)
 * and quartet tile lists for screening at `cutoff` (<= 0: none).
 * The two-electron operator is env[8] (PTR_RANGE_OMEGA, as libcint reads it;
 * set by Mole.with_range_coulomb, pyscf/gto/mole.py:2930-2979): 0 the Coulomb
 * 1/r, > 0 erf(omega r)/r, < 0 erfc(|omega| r)/r.  It holds for the bounds
 * and for every J/K build of the handle (one handle per omega, like the
 * per-omega optimizers of pyscf/scf/hf.py:2062-2067).
 * Replaces _VHFOpt.__init__/init_cvhf_direct (pyscf/scf/_vhf.py:151-206). */
int mjk_basis_upload(mjk_ctx* ctx, const int* atm, int natm, const int* bas, int nbas,
                     const double* env, double cutoff, mjk_basis** out);

/* Batch of nmol independent molecules in one handle (one launch per quartet
 * class covers all of them).  No reference counterpart: the reference
 * builds one molecule at a time. */
int mjk_basis_upload_batch(mjk_ctx* ctx, int nmol, const int* natm, const int* nbas,
                           const int* const* atm, const int* const* bas,
                           const double* const* env, double cutoff, mjk_basis** out);
void mjk_basis_destroy(mjk_basis* basis);

/* Batched SCF (no reference counterpart): active[m] = 0 drops molecule m's
 * quartets from the following builds of the handle; 1 keeps them.  All
 * molecules start active.  The row kernels' tasks of the active molecules
 * are compacted on the device before each build and run as persistent
 * launches (blocks claim tasks from a counter), so an iteration's cost
 * follows its active molecules.  mjk_get_jk then copies only the active
 * molecules' blocks: the inactive ones' blocks of dms are not read and their
 * blocks of vj/vk are not written (mjk_get_jk_device writes them as zero).
 * The reference converges one molecule per SCF loop (pyscf/scf/hf.py:171-208). */
int mjk_basis_set_active(mjk_ctx* ctx, mjk_basis* basis, const int* active);

/* nao of molecule `mol` (spherical AOs). */
int mjk_basis_nao(const mjk_basis* basis, int mol);
/* Total spherical elements of one density slot: sum_m nao_m^2. */
long long mjk_basis_slot_size(const mjk_basis* basis);
/* Schwarz bound q_cond [nbas][nbas] of molecule `mol` over the reference's
 * shells (max over the contracted functions), for comparison with
 * CVHFnr_int2e_q_cond (pyscf/lib/vhf/optimizer.c:408-454). */
int mjk_basis_q_cond(mjk_ctx* ctx, const mjk_basis* basis, int mol, double* q_cond);

/* J/K build from host arrays (synchronous).  Replaces the
 * _vhf.direct -> CVHFnr_direct_drv(CVHFdot_nrs8, CVHFnrs8_ji_s2kl,
 * CVHFnrs8_li_s2kj|s1kj) call chain (pyscf/scf/_vhf.py:405-464,
 * pyscf/lib/vhf/nr_direct.c:355-476), including hermi_triu.
 * fp32_tol > 0 enables mixed precision: a wavefront of 64 quartets whose
 * bounds q_ij q_kl max|D| (the prescreen's density factor) are all below
 * fp32_tol evaluates its recurrences in FP32; Rys roots, the contracted
 * block's digestion and all J/K accumulation stay FP64.  0: all FP64. */
int mjk_get_jk(mjk_ctx* ctx, mjk_basis* basis, const double* dms, int n_dm, int hermi,
               int with_j, int with_k, double direct_scf_tol, double fp32_tol,
               double* vj, double* vk, mjk_stats* stats);

/* Same, with dms/vj/vk already in device memory (HBM-resident inputs).
 * Asynchronous on the context's stream unless stats != NULL. */
int mjk_get_jk_device(mjk_ctx* ctx, mjk_basis* basis, const double* d_dms, int n_dm, int hermi,
                      int with_j, int with_k, double direct_scf_tol, double fp32_tol,
                      double* d_vj, double* d_vk, mjk_stats* stats);

/* Measurement and batching helpers with no reference counterpart.
 *
 * mjk_get_jk_resident: the J/K build of the densities that the last
 * mjk_get_jk call on this basis uploaded (still resident in the basis's
 * device buffers), with J and K left in device memory -- the build alone,
 * without the host<->device copies of mjk_get_jk, replayed as a HIP graph
 * like any plain build.  Synchronous.  MJK_EINVAL when no density of n_dm
 * slots is resident.
 * mjk_basis_fetch_jk: copies the resident J and/or K (either may be NULL)
 * of n_dm slots to the host.
 * mjk_basis_set_class_filter: cls >= 0 restricts the following builds of
 * the handle to quartet class cls (its index in the basis's class list for
 * the current cutoff), -1 restores all classes.  J and K are linear in the
 * quartets, so the filtered builds of all classes add up to the full build.
 * mjk_basis_nclasses / mjk_basis_class: the class list; l4 = (la lb lc ld),
 * ctr4 = contraction counts of the four shells (2 for a generally contracted
 * s shell kept whole, 1 otherwise), nblocks = launch size (0: no surviving
 * task), kernel = the kernel a build of n_dm densities with this hermi runs
 * for it (0 tile, 1 row, 2 cooperative, 3 merged row kernel, 4 row kernel
 * with only the A rows of K' in LDS; -1 the class does not run: the other
 * representation of its quartets does). */
int mjk_get_jk_resident(mjk_ctx* ctx, mjk_basis* basis, int n_dm, int hermi, int with_j, int with_k,
                        double direct_scf_tol, double fp32_tol, mjk_stats* stats);
int mjk_basis_fetch_jk(mjk_ctx* ctx, mjk_basis* basis, int n_dm, double* vj, double* vk);
int mjk_basis_set_class_filter(mjk_ctx* ctx, mjk_basis* basis, int cls);
int mjk_basis_nclasses(const mjk_basis* basis);
/* Tuning key of class k: key9 = (la lb lc ld ca cb cc cd bucket), bucket =
 * floor(log2) of the class's Schwarz-prescreened primitive quartets. */
int mjk_basis_class_key(mjk_ctx* ctx, const mjk_basis* basis, int k, int* key9);
/* Per-class tuning table (no reference counterpart; the reference has one CPU
 * code path with fixed 64-AO blocking, pyscf/lib/vhf/nr_direct.c:316-340).
 * n entries of 11 ints (la lb lc ld ca cb cc cd bucket kernel rprim) and one
 * double (scale): a class takes the entry of its key with the nearest bucket
 * (at most 2 away) -- kernel 0 the default choice, 1 the row kernel, 2 its
 * fallback (a merged class then runs its split pairs); rprim > 0 bra
 * primitive pairs per row task; scale > 0 multiplies the row tasks' size.
 * Replaces the table; every basis rebuilds its task lists at its next build.
 * Results agree with the untuned build to rounding. */
int mjk_ctx_set_tuning(mjk_ctx* ctx, int n, const int* keys, const double* scale);
/* Entries of the context's tuning table.  mjk_ctx_create loads the table
 * from the file MJK_TUNING names, else tuning/mi355x.txt beside the
 * library's folder (lines "la lb lc ld ca cb cc cd bucket kernel rprim
 * scale ...", '#' comments); MJK_TUNE=0 starts with none. */
int mjk_ctx_tuning_entries(mjk_ctx* ctx);
int mjk_basis_class(mjk_ctx* ctx, const mjk_basis* basis, int k, int n_dm, int hermi, int* l4, int* ctr4,
                    long long* nblocks, int* kernel);

/* Device-resident closed-shell SCF: the RHF loop of pyscf/scf/hf.py:47-242
 * (kernel; get_fock/CDIIS pyscf/scf/diis.py:40-66; eig :1340-1347; make_rdm1
 * :840; incremental get_veff :2416-2427) with the density, Fock matrix, DIIS
 * history and MO coefficients in HBM, J/K from this engine and the dense
 * algebra on rocBLAS/rocSOLVER on the context's stream.  The reference runs
 * this loop in numpy with D crossing to the C driver every cycle.
 *   mjk_rhf_create: basis = one molecule; s1e/h1e = overlap and core
 *     Hamiltonian (host, nao x nao); nocc doubly occupied orbitals; enuc =
 *     nuclear repulsion; diis_space vectors (8 in the reference).
 *   mjk_rhf_run: dm0 = initial density (host) or NULL for the core
 *     Hamiltonian guess; converged when |dE| < conv_tol and the orbital
 *     gradient norm < conv_tol_grad; level_shift shifts the virtual space
 *     after DIIS (0: none).  Synchronous; fills res.
 *   mjk_rhf_fetch: any of dm, fock (nao x nao), mo_energy (nao), mo_coeff
 *     (row-major nao x nmo) may be NULL.
 *   mjk_rhf_device_density: the converged density in device memory (valid
 *     until mjk_rhf_destroy). */
typedef struct mjk_scf mjk_scf;
typedef struct {
    double e_tot;
    double gnorm;
    int cycles;
    int converged;
} mjk_rhf_result;
int mjk_rhf_create(mjk_ctx* ctx, mjk_basis* basis, const double* s1e, const double* h1e, int nocc, double enuc,
                   int diis_space, mjk_scf** scf);
int mjk_rhf_run(mjk_ctx* ctx, mjk_scf* scf, const double* dm0, int max_cycle, double conv_tol, double conv_tol_grad,
                double direct_scf_tol, double fp32_tol, double level_shift, mjk_rhf_result* res);
int mjk_rhf_fetch(mjk_ctx* ctx, mjk_scf* scf, double* dm, double* mo_energy, double* mo_coeff, double* fock);
const double* mjk_rhf_device_density(const mjk_scf* scf);
/* Density-fitted SCF (pyscf/df/df_jk.py:32-125, the SCF of density_fit()):
 * the following runs build J and K of the full density from the fitted
 * tensor every cycle, K from the occupied orbitals on the device; df = NULL
 * returns to exact J/K.  The fit must belong to this context and molecule
 * (mjk_df_build); a DF run starts from the core Hamiltonian guess.  The
 * mjk_df declarations follow below. */
struct mjk_df;
int mjk_rhf_set_df(mjk_ctx* ctx, mjk_scf* scf, struct mjk_df* df);
void mjk_rhf_destroy(mjk_scf* scf);

/* One-electron integrals on the device (mol.intor('int1e_ovlp'),
 * 'int1e_kin', 'int1e_nuc' of pyscf/gto/moleintor.py; the core Hamiltonian
 * of pyscf/scf/hf.py:315-345): any of s, t, v (nao x nao, spherical) may be
 * NULL; device_out = 1 when they are device pointers (asynchronous then on the
 * context's stream), 0 for host arrays.  Point nuclei; shells up to f. */
int mjk_int1e_device(mjk_ctx* ctx, const int* atm, int natm, const int* bas, int nbas, const double* env, double* s,
                     double* t, double* v, int device_out);

/* Density-fitted J/K (pyscf/df/incore.py:129-214 cholesky_eri; pyscf/df/df_jk.py:278-418
 * get_jk), Coulomb operator.
 *   mjk_df_build: the orbital basis (atm/bas/env, shells up to f) and the
 *     auxiliary basis (its own atm/bas/env, shells up to g, as auxmol holds
 *     them).  Computes (P|Q) and (ij|P) on the device, then B = L^-1 (P|ij)
 *     with (P|Q) = L L^T; when the factorisation fails, B = (V w^-1/2)^T (P|ij)
 *     over the eigenpairs with w > lindep.  B stays in HBM (naux x nao^2, all
 *     orbital pairs).  lindep < 0 keeps the raw (ij|P) (no fit; for tests).
 *   mjk_df_naux: fitted functions (rows of B).
 *   mjk_df_get_jk: n_dm densities (host, nao x nao each, any symmetry);
 *     J = sum_P B_P tr(B_P D); K = sum_P B_P D B_P, or, when orbo != NULL
 *     (n_dm x nao x nocc, D = X X^T per density, the reference's mo_coeff
 *     path), K = sum_P (B_P X)(B_P X)^T.  vj/vk host, n_dm x nao x nao.
 *   mjk_df_get_jk_device: the same with device arrays, asynchronous.
 *   mjk_df_fetch: which 0 = the raw (P|Q) (naux_raw^2), 1 = B as the device
 *     holds it, C order (nao, nao, naux): the reference's _cderi unpacked and
 *     transposed. */
typedef struct mjk_df mjk_df;
int mjk_df_build(mjk_ctx* ctx, const int* atm, int natm, const int* bas, int nbas, const double* env,
                 const int* aux_atm, int aux_natm, const int* aux_bas, int aux_nbas, const double* aux_env,
                 double lindep, mjk_df** df);
int mjk_df_naux(const mjk_df* df);
int mjk_df_get_jk(mjk_ctx* ctx, mjk_df* df, const double* dms, int n_dm, const double* orbo, int nocc, int with_j,
                  int with_k, double* vj, double* vk);
int mjk_df_get_jk_device(mjk_ctx* ctx, mjk_df* df, const double* d_dms, int n_dm, const double* d_orbo, int nocc,
                         int with_j, int with_k, double* d_vj, double* d_vk);
int mjk_df_fetch(mjk_ctx* ctx, mjk_df* df, int which, double* out);
void mjk_df_destroy(mjk_df* df);

/* Request server (continuous batching on one device).  The reference serves
 * one caller at a time: each SCF.get_jk call crosses into CVHFnr_direct_drv
 * once (pyscf/scf/_vhf.py:540-639) and returns.  Here requests -- a J/K build
 * of given densities, or a whole RHF -- are submitted, run concurrently on
 * nslots worker contexts (one stream each), and are collected by ticket.
 *   mjk_server_create: hbm_budget = bytes the running requests may hold
 *     together (0: 80% of the free device memory).
 *   mjk_submit: copies the request (molecule arrays, densities); returns a
 *     ticket at once.  MJK_ENOMEM when the request's estimate alone exceeds
 *     the budget.  Admission is in arrival order: the head starts when a
 *     slot is free and the bytes the running requests' contexts have tracked
 *     (each at least its estimate until its buffers exist) plus the head's
 *     estimate fit the budget.  An RHF request runs the device-resident SCF
 *     (mjk_rhf_*: core guess, CDIIS 8, converged at conv_tol and
 *     sqrt(conv_tol) on the orbital gradient).
 *   mjk_poll: state and timings of a ticket, non-blocking.
 *   mjk_wait_any: waits up to timeout_s for a request to finish; *ticket =
 *     the next finished ticket not yet returned (0 on time-out).
 *   mjk_result: copies the outputs (J/K: n_dm x nao x nao each; RHF: the
 *     density and orbital energies; any pointer may be NULL) and releases the
 *     request; a failed request returns its error.
 *   mjk_server_hbm_held: the bytes admission counts now.
 *   mjk_server_error: the calling thread's last server error.
 *   mjk_server_destroy: runs the queued requests to the end, then frees. */
typedef struct mjk_server mjk_server;
#define MJK_REQ_JK 0
#define MJK_REQ_RHF 1
#define MJK_REQ_QUEUED 0
#define MJK_REQ_RUNNING 1
#define MJK_REQ_DONE 2
#define MJK_REQ_FAILED 3
typedef struct {
    int kind;                 /* MJK_REQ_JK or MJK_REQ_RHF */
    const int* atm;           /* natm x 6, pyscf layout */
    int natm;
    const int* bas;           /* nbas x 8 */
    int nbas;
    const double* env;
    int env_size;
    double direct_scf_tol;
    double fp32_tol;
    /* MJK_REQ_JK */
    const double* dm;         /* n_dm x nao x nao */
    int n_dm;
    int hermi;
    int with_j;
    int with_k;
    /* MJK_REQ_RHF */
    int nelectron;
    int max_cycle;
    double conv_tol;
} mjk_request;
typedef struct {
    int state;                /* MJK_REQ_* */
    int nao;
    int n_dm;
    int cycles;
    int converged;
    int slot;
    double e_tot;
    double t_submit, t_start, t_done;   /* seconds since the server's creation */
    size_t hbm_estimate;
    size_t hbm_peak;          /* tracked device bytes the request held at most */
    char error[256];
} mjk_request_info;
int mjk_server_create(int device, int nslots, size_t hbm_budget, mjk_server** server);
int mjk_submit(mjk_server* server, const mjk_request* request, long long* ticket);
int mjk_poll(mjk_server* server, long long ticket, mjk_request_info* info);
int mjk_wait_any(mjk_server* server, double timeout_s, long long* ticket);
int mjk_result(mjk_server* server, long long ticket, double* vj, double* vk, double* dm, double* mo_energy);
size_t mjk_server_hbm_held(mjk_server* server);
const char* mjk_server_error(const mjk_server* server);
void mjk_server_destroy(mjk_server* server);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The build of mjk_get_jk_device (same kernels, results
 * equal to rounding) with every work buffer -- Cartesian densities, the
 * private J'/K' accumulator copies, dm_cond -- carved out of a static device
 * array of MJK_NOALLOC_SCRATCH_BYTES per device, so that the call allocates
 * and frees no device memory.  The caller passes the inputs (a basis handle,
 * the densities in device memory) and the outputs (device memory).
 *   - The basis's quartet task lists must exist for direct_scf_tol: upload it
 *     with that cutoff (or run one mjk_get_jk_device build at it); otherwise
 *     MJK_EINVAL.
 *   - As many accumulator copies as fit are used (at most the context's);
 *     a build whose densities, one J'/K' copy and dm_cond do not fit returns
 *     MJK_ENOMEM (e.g. one density of about 2900 Cartesian functions).
 *   - No statistics, no profiling, no graph replay; asynchronous on the
 *     context's stream.
 *   - Every call on one device shares that device's array, which keeps the
 *     last call's contents (each call clears what must start at zero): calls
 *     that overlap on one GPU -- from two contexts, or streams not ordered
 *     with each other -- race.  Calls on one context are ordered. */
#define MJK_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int mjk_get_jk_device_noalloc(mjk_ctx* ctx, mjk_basis* basis, const double* d_dms, int n_dm, int hermi,
                              int with_j, int with_k, double direct_scf_tol, double fp32_tol,
                              double* d_vj, double* d_vk);

/* Per-quartet-class kernel profiling with HIP events on the launch stream
 * (the classes of a build run concurrently on up to four streams, so the
 * per-class times of one build overlap).
 * mjk_profile(…, 1) resets and starts accumulating times and work tallies
 * (each J/K build then synchronises at its end); mjk_profile(…, 2) does the
 * same with times only (the kernels then keep no tallies, as in an
 * unprofiled build); 0 stops.  mjk_profile_read returns the number of
 * classes written: cls[4k..4k+3] = (la lb lc ld), accumulated GPU ms,
 * evaluated quartets, primitive quartets, algorithmic FLOPs and launches.
 * No reference counterpart (the reference's timer is logger.timer,
 * pyscf/lib/logger.py:167-180). */
int mjk_profile(mjk_ctx* ctx, mjk_basis* basis, int enable);
int mjk_profile_read(mjk_ctx* ctx, const mjk_basis* basis, int max_classes, int* cls, double* ms,
                     long long* nq, long long* nprimq, double* flops, long long* launches);
/* Kernel each profiled class last ran on, in mjk_profile_read's order:
 * 0 tile kernel (k_jk), 1 row kernel (k_jk_row), 2 cooperative kernel
 * (k_jk_coop), 3 merged row kernel (k_jk_row_gc), 4 row kernel with the B
 * rows of K' in memory (k_jk_row_m).  Returns the number of classes
 * written. */
int mjk_profile_kinds(mjk_ctx* ctx, const mjk_basis* basis, int max_classes, int* kind);
/* Accumulated GPU time of the quartet phase of the profiled builds (from the
 * end of the density transform to the end of the last class kernel, over all
 * the context's streams) and the number of builds. */
int mjk_profile_span(mjk_ctx* ctx, const mjk_basis* basis, double* ms, long long* builds);

/* Device memory helpers (for callers without another GPU runtime). */
int mjk_device_alloc(mjk_ctx* ctx, size_t bytes, void** ptr);
int mjk_device_free(mjk_ctx* ctx, void* ptr);
int mjk_memcpy_h2d(mjk_ctx* ctx, void* dst, const void* src, size_t bytes);
int mjk_memcpy_d2h(mjk_ctx* ctx, void* dst, const void* src, size_t bytes);
int mjk_synchronize(mjk_ctx* ctx);
/* Device-to-device copy of `bytes` (a multiple of 8) with a plain HIP kernel. */
int mjk_device_copy(mjk_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Free and total HBM of a device (admission control of the request loop). */
int mjk_mem_info(int device, size_t* free_bytes, size_t* total_bytes);

/* One-electron integrals for the SCF caller, computed on the host:
 * which = 0 overlap (int1e_ovlp), 1 kinetic (int1e_kin), 2 nuclear
 * attraction (int1e_nuc); out float64 [nao][nao] spherical.  Replaces
 * mol.intor('int1e_ovlp'/'int1e_kin'/'int1e_nuc') as used by
 * get_ovlp/get_hcore (pyscf/scf/hf.py:315-345, pyscf/gto/mole.py:3474). */
int mjk_int1e(int which, const int* atm, int natm, const int* bas, int nbas, const double* env,
              double* out);

/* Cartesian -> real solid harmonic coefficients the engine uses for shells of
 * angular momentum l <= 3: out[m * ncart(l) + c], m = 0..2l in PySCF's AO
 * order (p: x, y, z; l >= 2: m = -l..l), Cartesian components c in the order
 * lx = l..0, ly = l-lx..0, including the s and p angular factors
 * 1/(2 sqrt(pi)) and sqrt(3/(4 pi)).  Replaces libcint's CINTc2s_ket_sph as
 * used by pyscf.gto.cart2sph (pyscf/gto/mole.py:157-187); host only. */
int mjk_c2s(int l, double* out);

/* Rys roots x_i (= t_i^2) and weights of the interpolation table, evaluated
 * on the host (table validation).  n <= 9.  The kernels' root cursor is also
 * run over the table in memory and over the padded per-root-count layout that
 * the row kernels copy to LDS; MJK_EINTERNAL if the two differ in any bit.  No reference counterpart (the
 * reference's roots live inside libcint). */
int mjk_rys_roots(int n, double T, double* x, double* w);

#ifdef __cplusplus
}
#endif

#endif /* MJK_H */
