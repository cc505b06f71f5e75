/*
 * vmps_hxv.h -- C ABI of the MI355X (gfx950) effective-Hamiltonian engine.
 *
 * The reference (TNWorkers/VMPS) binds its Krylov solvers to the pivot
 * operators by C++ overload resolution, not by an FFI: LanczosSolver calls
 * the free functions HxV / dot / norm on PivotMatrix2 / PivotVector
 * (pivot/DmrgPivotMatrix2.h:78-83, pivot/DmrgPivotVector.h:121-424).  This
 * header is the flat boundary those calls are routed through: plain pointers,
 * sizes and status codes, no C++ or GPU types.  Every entry point below names
 * the reference interface it stands in for.
 *
 * Conventions
 *  - Quantum numbers are int arrays of length nq (1 or 2).  A component of kind
 *    VMPS_SYM_U1 is an additive charge, one of kind VMPS_SYM_SU2 is q = 2S+1
 *    (symmetry/U1.h, symmetry/SU2.h).
 *  - Dense blocks are column-major FP64, exactly Eigen's default storage, so a
 *    caller passes block.data() without copies.
 *  - A PivotVector is one packed FP64 array holding its blocks in the order they
 *    were declared with vmps_pivot2_add_V_block (component s, then block q).
 *  - Calls return VMPS_OK or an error status; vmps_last_error() gives the
 *    message of the calling thread's last failure.  Nothing throws across the
 *    boundary.
 *  - Objects are independent: distinct plans may be used from distinct threads
 *    concurrently (one HIP stream per plan); a single plan is not reentrant.
 */
#ifndef VMPS_HXV_H
#define VMPS_HXV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VMPS_ABI_VERSION 4

typedef int vmps_status;
#define VMPS_OK 0
#define VMPS_ERR_ARG 1        /* invalid argument */
#define VMPS_ERR_LAYOUT 2     /* block layout inconsistent with the pivot */
#define VMPS_ERR_OOM 3        /* device allocation failed */
#define VMPS_ERR_HIP 4        /* HIP runtime error */
#define VMPS_ERR_NONFINITE 5  /* NaN/Inf found in an input vector */
#define VMPS_ERR_STATE 6      /* call out of order (e.g. no block structure yet) */
#define VMPS_ERR_INTERNAL 7
#define VMPS_ERR_NOCONV 8     /* an iterative method stopped at its iteration limit */

#define VMPS_SYM_U1 1
#define VMPS_SYM_SU2 2

typedef enum {
  VMPS_PREC_FP64 = 0,   /* plain FP64 drop-in (FP64 MFMA) */
  VMPS_PREC_FP32 = 1,   /* FP32 MFMA, FP32 intermediates, FP64 vectors */
  VMPS_PREC_BF16X3 = 2  /* split-BF16 x3 MFMA with FP32 accumulation, FP64 vectors */
} vmps_precision;

/* A pivot: the host image of PivotMatrix2, PivotMatrix1 or PivotMatrix0 plus
 * the PivotVector layout.  The vmps_pivot2_* builder calls below apply to all
 * three kinds (the name is kept from the two-site API). */
typedef struct vmps_pivot2 vmps_pivot2;
typedef struct vmps_plan vmps_plan;     /* device-resident staged block-triple plan */

int vmps_abi_version(void);
const char* vmps_last_error(void);

/* ---- pivot description (PivotMatrix2, pivot/DmrgPivotMatrix2.h:15-76) ---- */

/* PivotMatrix2(L, R, W12, W34, qloc12, qloc34, qOp12, qOp34)
 * (DmrgPivotMatrix2.h:34-52): local bases and operator bases of both sites.
 * kinds[nq] = VMPS_SYM_U1 / VMPS_SYM_SU2 per component. */
vmps_status vmps_pivot2_create(int nq, const int* kinds, int nloc12, const int* qloc12, int nloc34,
                               const int* qloc34, int nop12, const int* qop12, int nop34, const int* qop34,
                               vmps_pivot2** out);
/* PivotMatrix1 (pivot/DmrgPivotMatrix1.h:11-66; Terms[0].qloc / .qOp): a
 * one-site pivot.  Its W goes into slot 12 of vmps_pivot2_add_W; PivotVector
 * components are the local states s (DmrgPivotVector.h:28-31). */
vmps_status vmps_pivot1_create(int nq, const int* kinds, int nloc, const int* qloc, int nop, const int* qop,
                               vmps_pivot2** out);
/* PivotMatrix0(L, R) (pivot/DmrgPivotMatrix0.h:10-27): a zero-site pivot (the
 * bond matrix C); the vector has one component, s = 0, with blocks (q, q). */
vmps_status vmps_pivot0_create(int nq, const int* kinds, vmps_pivot2** out);
/* PivotMatrix2::Terms / PivotMatrix1::Terms (DmrgPivotMatrix2.h:54,
 * DmrgPivotMatrix1.h:56): HxV sums over terms (DmrgPivotMatrix2.h:233-245).
 * A pivot starts with one term.  add_term appends a term with its own
 * operator bases (qop34 unused for one-site, both unused for zero-site) and
 * makes it current; add_env, add_W and the block-structure calls act on the
 * current term.  All terms share the local bases and the vector layout. */
vmps_status vmps_pivot_add_term(vmps_pivot2* p, int nop12, const int* qop12, int nop34, const int* qop34,
                                int* term);
vmps_status vmps_pivot_select_term(vmps_pivot2* p, int term);
/* nsite (0, 1, 2) and number of terms; either pointer may be NULL. */
vmps_status vmps_pivot_info(const vmps_pivot2* p, int* nsite, int* nterms);
void vmps_pivot2_destroy(vmps_pivot2* p);

/* Index s of PivotVector::data[s]: for two sites the combined basis
 * loc12.combine(loc34) (tensors/Qbasis.h:686-798; DmrgPivotVector.h:33-48),
 * for one site qloc itself, for zero sites the single vacuum component.
 * ns receives its size; qloc (ns*nq ints) may be NULL. */
vmps_status vmps_pivot2_local_basis(const vmps_pivot2* p, int* ns, int* qloc);

/* One sector of the environment Tripod L (which=0) or R (which=1):
 * Tripod::push_back({in,out,mid}, block[a][0] for a < adim)
 * (tensors/Multipede.h:77-95,152-154).  rows/cols/data describe the adim
 * blocks packed back to back; a block with rows*cols == 0 is absent
 * (DmrgPivotMatrix2.h:207-208 skip such blocks). */
vmps_status vmps_pivot2_add_env(vmps_pivot2* p, int which, const int* qin, const int* qout, const int* qmid,
                                int adim, const int64_t* rows, const int64_t* cols, const double* data);

/* One sparse block of W12[s1][s2][k] (which=12) or W34[s3][s4][k] (which=34):
 * Biped<Sym,SparseMatrix<double,ColMajor>>::push_back(qin, qout, mat)
 * (MpoTerms.h:1175-1233).  nnz triplets (row, col, val) in any order.
 * One-site pivots take their W[s1][s2][k] in slot 12. */
vmps_status vmps_pivot2_add_W(vmps_pivot2* p, int which, int s1, int s2, int k, const int* qin, const int* qout,
                              int rows, int cols, int64_t nnz, const int* row, const int* col, const double* val);

/* Declare block (qin, qout) of PivotVector::data[s] (Biped::push_back,
 * tensors/Biped.h:82-104; extend_A materialises every allowed block,
 * DmrgContractions.h:1832-1867).  *offset receives its element offset in the
 * packed vector. */
vmps_status vmps_pivot2_add_V_block(vmps_pivot2* p, int s, const int* qin, const int* qout, int64_t rows,
                                    int64_t cols, int64_t* offset);
vmps_status vmps_pivot2_vsize(const vmps_pivot2* p, int64_t* size);

/* Bulk forms of the three builder calls above, for hosts that hold a whole
 * Tripod, W slot or PivotVector layout at once (same checks, item by item;
 * the first failing item's status is returned):
 *  - add_env_sectors: nsec sectors; qin/qout/qmid nsec*nq ints, adim[nsec],
 *    rows/cols/blocks sum(adim) entries in sector order, blocks[i] the
 *    column-major data of block i (ignored when rows*cols == 0), so a host
 *    holding one allocation per block (Eigen MatrixXd) passes them as they are;
 *  - add_W_blocks: nblk blocks of slot `which`; idx = nblk x {s1, s2, k},
 *    qin/qout nblk*nq, dims = nblk x {rows, cols}, nnz[nblk], and the
 *    triplets of all blocks back to back in row/col/val;
 *  - add_V_blocks: nblk blocks, component s[b]; offsets (may be NULL)
 *    receives each block's element offset. */
vmps_status vmps_pivot2_add_env_sectors(vmps_pivot2* p, int which, int nsec, const int* qin, const int* qout,
                                        const int* qmid, const int* adim, const int64_t* rows, const int64_t* cols,
                                        const double* const* blocks);
vmps_status vmps_pivot2_add_W_blocks(vmps_pivot2* p, int which, int nblk, const int* idx, const int* qin,
                                     const int* qout, const int* dims, const int64_t* nnz, const int* row,
                                     const int* col, const double* val);
vmps_status vmps_pivot2_add_V_blocks(vmps_pivot2* p, int nblk, const int* s, const int* qin, const int* qout,
                                     const int64_t* rows, const int64_t* cols, int64_t* offsets);

/* precalc_blockStructure(L, Vin, W12, W34, Vin, R, qloc12, qloc34, qOp12,
 * qOp34, qlhs, qrhs, factor_cgcs) (tensors/DmrgIndexGymnastics.h:698-856),
 * or the one-site form (L, Vin, W, Vin, R, qloc, qOp, ...) (:485-590), for
 * every term: fills the pivot's block structure, including the SU(2) 6j/9j
 * factors.  Zero-site pivots need none (the call is accepted). */
vmps_status vmps_pivot2_precalc(vmps_pivot2* p);

/* Use the caller's own qlhs / qrhs / factor_cgcs of the current term
 * (DmrgPivotMatrix2.h:66-70, DmrgPivotMatrix1.h:58-62) instead of
 * vmps_pivot2_precalc.  qlhs[2*i] = {s1s3, q13}; entries
 * rhs_begin[i] .. rhs_begin[i+1]-1 of qrhs (12 values each:
 * s2s4,q24,qL,qR,s1,s2,k12,qW12,s3,s4,k34,qW34; one-site: s2,q2,qL,qR,k,qW
 * followed by six zeros) and of factors belong to lhs entry i. */
vmps_status vmps_pivot2_set_blockstructure(vmps_pivot2* p, int64_t nlhs, const int64_t* qlhs,
                                           const int64_t* rhs_begin, const int64_t* qrhs, const double* factors);
/* Read the block structure back.  Call with NULL arrays to get nlhs / nrhs. */
vmps_status vmps_pivot2_get_blockstructure(const vmps_pivot2* p, int64_t* nlhs, int64_t* nrhs, int64_t* qlhs,
                                           int64_t* rhs_begin, int64_t* qrhs, double* factors);

/* Excited-state penalty: PivotMatrix2::A0proj / Epenalty
 * (DmrgPivotMatrix2.h:72-75, applied at :249-272).  A0 holds n packed vectors. */
vmps_status vmps_pivot2_set_projectors(vmps_pivot2* p, double Epenalty, int n, const double* A0);

/* ---- device plan and HxV ---------------------------------------------- */

typedef struct {
  int64_t n_terms;  /* contributing (out, L, Vin, R) block quadruples */
  int64_t n_T;      /* distinct (L block, Vin block) products */
  int64_t n_U;      /* distinct (out block, R block) sums */
  int64_t n_out;    /* output blocks */
  int64_t n_L, n_R, n_V;
  double flops_alg; /* sum over distinct (L,V) of 2 m'mn + over distinct (out,R) of 2 m'nn' */
  double flops_ref; /* per-term (L.V).R work of the reference loop */
  double bytes_alg_fp64;
  int64_t L_elems, R_elems, V_elems, T_elems, U_elems;
  int64_t max_block_dim;
  double flops_A;   /* stage A share of flops_alg: T = L . Vin */
  double flops_C;   /* stage C share of flops_alg: Vout += U . R */
} vmps_plan_stats;

/* Work counts of the staged plan of a pivot, computed on the host only (no
 * device needed): the numbers vmps_plan_stats_get reports for its plan. */
vmps_status vmps_pivot2_stats(const vmps_pivot2* p, vmps_plan_stats* st);

/* Flatten the pivot into the staged block-triple tables and upload L, R and
 * the tables to the current HIP device.  The pivot may be destroyed after. */
vmps_status vmps_plan_create(const vmps_pivot2* p, vmps_precision prec, vmps_plan** out);
/* Several independent pivots in one plan, one launch per stage for all of
 * them (the batched HxV of VUMPS cell sites, VumpsSolver.h:1168-1211, run
 * there one pivot per OpenMP thread).  The plan's vectors are the pivots'
 * vectors concatenated: pivot i owns elements [offsets[i], offsets[i+1]);
 * offsets (n+1 entries) may be NULL.  vmps_hxv / vmps_hxv_device apply every
 * pivot at once; vmps_dot sums over all of them.  Projectors are not
 * supported here. */
vmps_status vmps_plan_create_batch(const vmps_pivot2* const* pivots, int n, vmps_precision prec, vmps_plan** out,
                                   int64_t* offsets);
void vmps_plan_destroy(vmps_plan* plan);
vmps_status vmps_plan_stats_get(const vmps_plan* plan, vmps_plan_stats* st);
vmps_status vmps_plan_set_precision(vmps_plan* plan, vmps_precision prec);

/* HxV(H, Vin, Vout) (DmrgPivotMatrix2.h:78-83 -> OxV :85-273).
 * Host pointers, synchronous.  Vout has Vin's layout. */
vmps_status vmps_hxv(vmps_plan* plan, const double* vin, double* vout);
/* Same on device-resident vectors, enqueued on `stream` (hipStream_t; NULL =
 * the plan's own stream).  Does not synchronise. */
vmps_status vmps_hxv_device(vmps_plan* plan, const double* d_vin, double* d_vout, void* stream);

/* dot(V1, V2) = sum_s sum_q tr(V1^T V2) * coeff_dot(out[q])
 * (DmrgPivotVector.h:296-339).  Host pointers. */
vmps_status vmps_dot(vmps_plan* plan, const double* x, const double* y, double* result);

/* SU(2) recoupling symbols with the reference's argument convention q = 2S+1:
 * coupling_6j / coupling_9j (symmetry/SU2Wrappers.h:71-94, which pass q-1 =
 * 2j to GSL).  Host-only, evaluated by the same Racah routine the block
 * structure uses; for callers that recouple MPS tensors (contract_AA2,
 * split_AA2). */
double vmps_coupling_6j(int q1, int q2, int q3, int q4, int q5, int q6);
double vmps_coupling_9j(int q1, int q2, int q3, int q4, int q5, int q6, int q7, int q8, int q9);

/* Timing helper: mean device time of one HxV on the plan's stream, measured
 * with HIP events around `iters` back-to-back calls on resident vectors. */
vmps_status vmps_plan_time_hxv(vmps_plan* plan, int warmup, int iters, double* ms_per_hxv);

/* Tile configuration of a plan.  No counterpart in the reference (its Eigen
 * GEMMs need none); it serves the GPU path's tuning.
 *   wave_edge: output tile edge of the FP64 / FP32 wave kernels, 16 or 32;
 *   big_min:   split-BF16 x3 only: problems with both dimensions >= big_min
 *              use the LDS-staged 128 x 128 workgroup tiles, the rest
 *              32 x 32 wave tiles (0 = none use the big tiles);
 *   big_min64: FP64 / FP32 only: problems with both dimensions >= big_min64
 *              use LDS-staged 128 x 128 workgroup tiles (0 = none, the default);
 *   explicit_u: 0 = the combine U = sum c T is done on stage C's operand
 *              load (U never stored); 1 = a separate kernel forms U once per
 *              (output block, R block) and stage C reads it (pays off when
 *              many terms share an output block that spans many tiles). */
typedef struct {
  int wave_edge;
  int big_min;
  int big_min64;
  int explicit_u;
  double ms;       /* HxV time of the chosen configuration (autotune; 0 if taken from the cache) */
  int from_cache;  /* 1: a plan with the same block-size histogram was tuned before in this process */
  int n_tried;     /* configurations timed */
} vmps_tile_config;

vmps_status vmps_plan_get_tiles(const vmps_plan* plan, vmps_tile_config* cfg);
/* Results do not depend on the configuration beyond FP rounding of the
 * split-BF16 path (FP64: bit-identical sums in the same order per tile). */
vmps_status vmps_plan_set_tiles(vmps_plan* plan, const vmps_tile_config* cfg);  /* ms .. n_tried ignored */
/* Autotuner: times every candidate configuration of the plan's precision
 * (`iters` HxV each on the plan's resident vectors) and keeps the fastest.
 * The choice is cached per block-size histogram (the share of the work in
 * classes of block rows x columns) and per precision, so later plans of the
 * same shape take it without timing.  Thread-safe. */
vmps_status vmps_plan_autotune(vmps_plan* plan, int iters, vmps_tile_config* out);
/* Forgets every cached choice of vmps_plan_autotune in this process, so that
 * the next plan of each shape is timed again (for instance, to time two
 * solver configurations that should each pay their own tuning).  No
 * counterpart in the reference, which has no tile configuration.  Thread-safe. */
vmps_status vmps_autotune_clear_cache(void);

/* ---- environment updates ---------------------------------------------------
 * contract_L(Lold, Abra, W, Aket, qloc, qOp, Lnew)  (tensors/DmrgContractions.h:34-158)
 * contract_R(Rold, Abra, W, Aket, qloc, qOp, Rnew)  (tensors/DmrgContractions.h:179-311)
 * (MODE = FULL, RANDOMIZE = false).  The new environment is returned with its
 * sectors in creation order and, within a sector, blocks a = 0 .. adim-1. */
typedef struct vmps_env vmps_env;

/* which = 0: contract_L, which = 1: contract_R */
vmps_status vmps_env_create(int which, int nq, const int* kinds, int nloc, const int* qloc, int nop, const int* qop,
                            vmps_env** out);
void vmps_env_destroy(vmps_env* e);

/* ---- two-site split -------------------------------------------------------
 * Batched thin SVD of dense FP64 blocks on the device: the per-sector SVD of
 * Biped::truncateSVD (tensors/Biped.h:754-775) that split_AA2
 * (tensors/DmrgContractions.h:2460-2630) applies to every middle sector of a
 * two-site wavefunction.  Block b is rows[b] x cols[b], column-major, packed
 * back to back in A.  With r_b = min(rows[b], cols[b]) the outputs are packed
 * per block: U (rows x r), S (r values, descending), Vt (r x cols), all
 * column-major.  One-sided Jacobi in FP64 (every singular value to rounding-
 * level relative accuracy); max_sweeps (may be NULL) receives the largest
 * number of Jacobi sweeps a block needed.  A block still rotating after 60
 * sweeps makes the call return VMPS_ERR_NOCONV (outputs written as they
 * are).  The singular vector of a zero singular value (U column when
 * rows >= cols, else Vt row) is zero, not completed to an orthonormal basis;
 * the split discards such values.  Dimensions up to 2^20.  Host buffers,
 * synchronous. */
vmps_status vmps_block_svd(int nblocks, const int64_t* rows, const int64_t* cols, const double* A, double* U,
                           double* S, double* Vt, int* max_sweeps);
/* The split's factorisation through the reduced density matrix of one
 * side, on the device (stands in for the dense SVD of Biped::truncateSVD,
 * tensors/Biped.h:754-775, inside split_AA2, tensors/DmrgContractions.h:
 * 2460-2630).  For every block Psi (rows x cols, column-major, packed as in
 * vmps_block_svd):
 *   side 0: rho = Psi Psi^T (k = rows), C = W^T Psi   (k x cols)
 *   side 1: rho = Psi^T Psi (k = cols), C = Psi W     (rows x k)
 * with rho = W diag(lambda) W^T: lambda (k values, ascending; lambda =
 * sigma^2 of Psi, absolute accuracy eps ||Psi||^2), W (k x k, orthonormal
 * columns, the new site tensor of the sweep direction) and C (the centre
 * that moves on).  Outputs packed per block like the inputs.  Dense GEMMs by
 * rocBLAS, the eigensolver rocSOLVER's dsyevd; host buffers, synchronous;
 * VMPS_ERR_NOCONV if the eigensolver fails. */
vmps_status vmps_block_eigh_split(int nblocks, const int64_t* rows, const int64_t* cols, const double* A, int side,
                                  double* lambda, double* W, double* C);
/* The same factorisation with the sector matrices built on the device from
 * the blocks of a device vector d_v (vsize doubles; the two-site
 * wavefunction, e.g. the eigenvector vmps_lanczos_device left in a plan
 * vector): the recoupling of split_AA2 (tensors/DmrgContractions.h:2460-2630)
 * as a gather.  Destination d is the r x c block at (dst_row[d], dst_col[d])
 * of sector dst_block[d]; it receives the sum, in order, of
 * part_factor[p] * v[part_src[p] .. part_src[p] + r c) (column-major r x c)
 * over p in [part_begin[d], part_begin[d+1]); entries no destination covers
 * are zero.  Outputs as vmps_block_eigh_split. */
vmps_status vmps_block_eigh_split_gather(const double* d_v, int64_t vsize, int nblocks, const int64_t* rows,
                                         const int64_t* cols, int ndst, const int* dst_block, const int64_t* dst_row,
                                         const int64_t* dst_col, const int64_t* dst_r, const int64_t* dst_c,
                                         const int64_t* part_begin, const int64_t* part_src,
                                         const double* part_factor, int side, double* lambda, double* W, double* C);
/* vmps_block_svd that allocates no device memory: no counterpart in the
 * reference; for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc).  Arguments,
 * results and errors as vmps_block_svd (same kernels, same results).  The
 * work arrays live in a fixed 64 MiB device array of the library, one per
 * GPU, held until the process ends: the input, the Jacobi work copies, the
 * singular values and the packed U, S, Vt together (8 bytes x (2 rows cols
 * + (rows + cols) k + k^2 + 2 k) per block with k = min(rows, cols), plus a
 * few words of descriptors, each region rounded up to 256 bytes) must fit,
 * and k <= 8192 per block; larger calls
 * return VMPS_ERR_ARG before touching the device.  Calls on one GPU share
 * that array, so calls that overlap on one GPU (from several threads) race;
 * every call overwrites what it uses, nothing carries over.  Runs on the null
 * stream and returns after the results are copied back. */
vmps_status vmps_block_svd_noalloc(int nblocks, const int64_t* rows, const int64_t* cols, const double* A,
                                   double* U, double* S, double* Vt, int* max_sweeps);
/* one sector {in, out, mid} of Lold / Rold, packed as in vmps_pivot2_add_env */
vmps_status vmps_env_add_old(vmps_env* e, const int* qin, const int* qout, const int* qmid, int adim,
                             const int64_t* rows, const int64_t* cols, const double* data);
/* block (qin, qout) of A[s] (bra = 1: Abra, bra = 0: Aket), column-major rows x cols */
vmps_status vmps_env_add_A(vmps_env* e, int bra, int s, const int* qin, const int* qout, int64_t rows, int64_t cols,
                           const double* data);
/* sparse block of W[s1][s2][k] */
vmps_status vmps_env_add_W(vmps_env* e, int s1, int s2, int k, const int* qin, const int* qout, int rows, int cols,
                           int64_t nnz, const int* row, const int* col, const double* val);
/* bulk forms of the three calls above, laid out as the vmps_pivot2_add_*
 * bulk forms (A blocks: s[nblk], rows/cols[nblk], blocks[nblk] pointers) */
vmps_status vmps_env_add_old_sectors(vmps_env* e, int nsec, const int* qin, const int* qout, const int* qmid,
                                     const int* adim, const int64_t* rows, const int64_t* cols,
                                     const double* const* blocks);
vmps_status vmps_env_add_A_blocks(vmps_env* e, int bra, int nblk, const int* s, const int* qin, const int* qout,
                                  const int64_t* rows, const int64_t* cols, const double* const* blocks);
vmps_status vmps_env_add_W_blocks(vmps_env* e, int nblk, const int* idx, const int* qin, const int* qout,
                                  const int* dims, const int64_t* nnz, const int* row, const int* col,
                                  const double* val);
/* plan and run the contraction on the current HIP device (synchronous) */
vmps_status vmps_env_contract(vmps_env* e, vmps_precision prec);
/* result: nsec sectors, nslots = sum of adim, size = total elements */
vmps_status vmps_env_result_size(const vmps_env* e, int64_t* nsec, int64_t* nslots, int64_t* size);
/* qin/qout/qmid: nsec*nq ints each, adim: nsec, rows/cols: nslots (0 = absent
 * block), data: the blocks packed in slot order */
vmps_status vmps_env_result_get(const vmps_env* e, int* qin, int* qout, int* qmid, int* adim, int64_t* rows,
                                int64_t* cols, double* data);
/* work counts of the last contraction's plan */
vmps_status vmps_env_stats(const vmps_env* e, vmps_plan_stats* st);

/* ---- environments resident on the device ----------------------------------
 * The solvers keep every environment Heff[l].Terms[t].L / R in host memory
 * and hand the new one from contract_L / contract_R straight to the next
 * pivot (solvers/DmrgSolver.h:1945-1953).  A vmps_denv keeps such an
 * environment on the device it was made on: the sector layout on the host,
 * the blocks in one device allocation, so a sweep need not copy environments
 * between host and device.  It is never written once made.  A pivot or
 * contraction that takes it (vmps_pivot2_add_env_device,
 * vmps_env_add_old_device) shares the ownership of its device blocks: the
 * denv may be destroyed at any point afterwards, the blocks stay allocated
 * until the last of the denv, the pivot and the contraction is destroyed. */
typedef struct vmps_denv vmps_denv;

/* a device environment from host blocks, laid out as vmps_env_add_old_sectors */
vmps_status vmps_denv_create(int nq, int nsec, const int* qin, const int* qout, const int* qmid, const int* adim,
                             const int64_t* rows, const int64_t* cols, const double* const* blocks,
                             vmps_denv** out);
void vmps_denv_destroy(vmps_denv* d);
/* layout and (data != NULL) the blocks, as vmps_env_result_size / _get */
vmps_status vmps_denv_size(const vmps_denv* d, int64_t* nsec, int64_t* nslots, int64_t* size);
vmps_status vmps_denv_get(const vmps_denv* d, int* qin, int* qout, int* qmid, int* adim, int64_t* rows,
                          int64_t* cols, double* data);
/* vmps_env_contract whose result stays on the device as a new vmps_denv
 * (vmps_env_result_get then reports VMPS_ERR_STATE; result_size and stats
 * still answer).  With the old environment on the device (or none), it
 * returns once the contraction's kernels are queued: every reader of the new
 * environment -- a plan that gathers it, a contraction that takes it as its
 * old environment, vmps_denv_get, vmps_denv_destroy -- waits for them first, so the
 * caller can build the next plan on the host meanwhile.  A contraction from
 * a host environment returns when it has finished. */
vmps_status vmps_env_contract_device(vmps_env* e, vmps_precision prec, vmps_denv** out);
/* the old environment of a contraction: the whole Tripod Lold / Rold at once
 * (instead of vmps_env_add_old*) */
vmps_status vmps_env_add_old_device(vmps_env* e, const vmps_denv* E);
/* the whole environment L (which = 0) or R (which = 1) of the current term
 * (instead of vmps_pivot2_add_env*); the L of all terms must be all on the
 * host or all on the device, and likewise the R */
vmps_status vmps_pivot2_add_env_device(vmps_pivot2* p, int which, const vmps_denv* E);

/* ---- PivotVector algebra on device vectors (DmrgPivotVector.h:121-424) ----
 * All vectors are packed device arrays of the plan's layout; work is enqueued
 * on the plan's stream and scalar results are returned after a synchronise. */

/* Device vector of the plan's layout (PivotVector storage), zero-filled;
 * upload / download copy whole vectors between host and device. */
vmps_status vmps_vec_alloc(vmps_plan* plan, double** d_v);
vmps_status vmps_vec_free(vmps_plan* plan, double* d_v);
vmps_status vmps_vec_upload(vmps_plan* plan, double* d_v, const double* h_v);
vmps_status vmps_vec_download(vmps_plan* plan, const double* d_v, double* h_v);

/* dot(V1, V2) with coeff_dot weights (DmrgPivotVector.h:296-339) */
vmps_status vmps_vec_dot(vmps_plan* plan, const double* d_x, const double* d_y, double* result);
/* norm(V) = sqrt(dot(V, V)) (DmrgPivotVector.h:341-352) */
vmps_status vmps_vec_norm(vmps_plan* plan, const double* d_x, double* result);
/* y = a*x + b*y: addScale(a, x, y) is (a, 1) (DmrgPivotVector.h:289-293),
 * y *= a is (0, a), y /= a is (0, 1/a) (:223-253), y += x / y -= x are (+-1, 1) */
vmps_status vmps_vec_axpby(vmps_plan* plan, double a, const double* d_x, double b, double* d_y);
/* GaussianRandomVector<PivotVector>::fill: entries uniform in (-1, 1), then
 * normalised (DmrgPivotVector.h:410-424); counter-based, reproducible per seed */
vmps_status vmps_vec_fill_random(vmps_plan* plan, double* d_v, uint64_t seed);

/* contract_AA2 (tensors/DmrgContractions.h:1869-1927) into a device vector
 * of the plan's layout: d_out is zeroed, then for every term t
 *   d_out[out_off[t] ..] (r x c, column-major) += factor[t] * A1[i1[t]] . A2[i2[t]]
 * with A1[b] (rows1[b] x cols1[b]) and A2[b] (rows2[b] x cols2[b]) host
 * blocks, column-major float64, copied to the device (A1 blocks: the
 * Biped blocks of A[l][s1], A2 blocks those of A[l+1][s2]; factor: the
 * coeff_twoSiteGate factor of the term; out_off: the offset of the Apair
 * block (in1, q2) of s1s2 in the pivot vector).  The products are rocBLAS
 * GEMMs on the plan's stream; returns when d_out is complete. */
vmps_status vmps_contract_aa2_device(vmps_plan* plan, int n1, const int64_t* rows1, const int64_t* cols1,
                                     const double* const* A1, int n2, const int64_t* rows2, const int64_t* cols2,
                                     const double* const* A2, int nterms, const int64_t* out_off, const int* i1,
                                     const int* i2, const double* factor, double* d_out);

/* ---- Krylov eigensolver (stands in for the external LanczosSolver that
 * DmrgSolver::iteration_two calls, solvers/DmrgSolver.h:1466-1471) ---- */
typedef struct {
  int max_iter;             /* Krylov dimension limit (dimK, DmrgTypedefs.h:373) */
  int min_iter;
  double tol_eigval;        /* |E_k - E_{k-1}| (DmrgTypedefs.h:371) */
  double tol_state;         /* beta_k |y_k|  (DmrgTypedefs.h:372) */
  int precision;            /* vmps_precision of the HxV inside the iterations */
  int refine_fp64;          /* mixed solves: restart in FP64 from the mixed eigenvector */
  double refine_tol_eigval;
  double refine_tol_state;
  uint64_t seed;            /* start vector when none is given */
  int graph;                /* 1: each iteration is one replay of a captured HIP graph (default);
                               0: the same kernels launched one by one */
  int refine_method;        /* mixed solves with refine_fp64: VMPS_REFINE_CORRECTION (default) or
                               VMPS_REFINE_RESTART */
} vmps_lanczos_opts;

/* Refinement of a mixed-precision eigenpair to the FP64 tolerances:
 * CORRECTION: FP64 residual, correction from conjugate gradients on the
 *   projected equation (1 - psi psi^T)(H' - E)(1 - psi psi^T) t = -r with the
 *   mixed operator H', FP64 Rayleigh-Ritz on {psi, t}; one FP64 HxV per step
 *   (at most 4 steps, then RESTART takes over).
 * RESTART: an FP64 Lanczos solve started from the mixed eigenvector. */
#define VMPS_REFINE_CORRECTION 0
#define VMPS_REFINE_RESTART 1

typedef struct {
  double energy;            /* <psi|H|psi> with the FP64 operator when refined / FP64 */
  double residual;          /* ||H psi - E psi||, same operator */
  int iterations;           /* iterations of the main solve */
  int iterations_refine;    /* iterations of the FP64 refinement */
  int hxv_calls;
  int converged;
  double energy_mixed;      /* Rayleigh quotient of the mixed solve's vector: with the FP64 operator when
                               refine_method is VMPS_REFINE_CORRECTION (the refinement forms it anyway),
                               else with the mixed operator; residual_mixed likewise */
  double residual_mixed;
  double ms;                /* wall time of the call */
  int hxv_fp64;             /* HxV calls with the FP64 operator (all of hxv_calls for FP64 solves) */
  int refine_steps;         /* correction steps of VMPS_REFINE_CORRECTION */
  int refine_cg;            /* mixed HxV calls inside those corrections */
} vmps_lanczos_result;

void vmps_lanczos_opts_default(vmps_lanczos_opts* o);
/* Lowest eigenpair of the pivot.  v: host vector, on entry the start vector
 * (NULL: random start from opts->seed), on exit the normalised eigenvector. */
vmps_status vmps_lanczos(vmps_plan* plan, double* v, const vmps_lanczos_opts* opts, vmps_lanczos_result* res);
/* Same with a device-resident start / result vector. */
vmps_status vmps_lanczos_device(vmps_plan* plan, double* d_v, const vmps_lanczos_opts* opts,
                                vmps_lanczos_result* res);

/* One lowest eigenpair per pivot of a batch plan (vmps_plan_create_batch):
 * the independent solves of VumpsSolver's loop over the unit cell
 * (VUMPS/VumpsSolver.h:1168-1211) run together.  Each pivot keeps its own
 * Gram-Schmidt coefficients, tridiagonal matrix and convergence test (opts
 * as for vmps_lanczos); every iteration of the whole batch is one replay of a
 * captured HIP graph (opts->graph).  v: host vector of the batch layout
 * (pivot i at offsets[i] of vmps_plan_create_batch), start vectors on entry
 * (NULL: uniform random, seed opts->seed), eigenvectors on exit.  res: one
 * result per pivot.  A mixed precision with refine_fp64 is refined by an FP64
 * batch solve started from the mixed eigenvectors (the restart form).
 * VMPS_ERR_STATE for a plan that is not a batch. */
vmps_status vmps_lanczos_batch(vmps_plan* plan, double* v, const vmps_lanczos_opts* opts, vmps_lanczos_result* res);

/* Per-stage device time (ms per call, HIP events on the plan's stream, mean
 * over `iters` launches of each stage alone on resident data):
 * ms[0] stage A (T = L.V), ms[1] stage B (U = sum c T), ms[2] stage C
 * (Vout = sum U.R), ms[3] the whole HxV. */
vmps_status vmps_plan_time_stages(vmps_plan* plan, int warmup, int iters, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* VMPS_HXV_H */
