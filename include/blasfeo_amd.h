/*
 * blasfeo_amd.h -- C ABI of the MI355X (gfx950) batched BLASFEO level-3 hot path.
 *
 * Two layers are declared here:
 *
 *  1. The BLASFEO descriptor types and the BLASFEO struct API for the hot-path
 *     routines, with the reference's names, argument order and meaning.  The
 *     struct layouts are byte-identical to the panel-major build of the
 *     reference (include/blasfeo_common.h:67-103), so a descriptor built by
 *     BLASFEO code can be handed over unchanged.  In this library the memory a
 *     descriptor points to (pA, dA) lives in GPU memory; the single-instance
 *     routines run one GPU instance and return when it is done (the reference
 *     routines are synchronous and return void as well).
 *
 *  2. The batched entry points, one per hot-path routine.  Each takes a
 *     device-resident array of argument records; record i holds exactly the
 *     arguments of one reference call (sizes, scalars, pointers to
 *     device-resident descriptors and their row/column offsets).  They run
 *     asynchronously on the given HIP stream and perform no host
 *     synchronisation.  The Riccati sweep, the trapezoid factorisations above
 *     128 rows and the mixed-precision solve keep their per-instance plan in a
 *     per-(device, stream) workspace that the library allocates on the first
 *     call (and enlarges for a larger batch); every other entry allocates
 *     nothing.  After one eager call of the same size on a stream, a call can
 *     be captured into a HIP graph; blasfeo_amd_d_back_ric_trf_batched_noalloc
 *     allocates nothing at all.  They return an
 *     int status for argument/launch errors only; numerical failures follow
 *     the reference's conventions (zero pivot => zero reciprocal in dA).
 *
 * Panel-major addressing (include/blasfeo_common.h:112): element (i,j) of a
 * matrix with panel size ps and panel stride cn sits at
 *     pA[(i - i%ps)*cn + j*ps + i%ps].
 * The panel size is not stored in the descriptor (the reference fixes it at
 * build time: 4 for the Haswell build, 8 for Skylake-X/Zen5); here every
 * batched call names it explicitly, and the single-instance routines use the
 * process-wide value set by blasfeo_amd_set_ps() (default 4).
 */
#ifndef BLASFEO_AMD_H
#define BLASFEO_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* descriptor types: include/blasfeo_common.h:67-103 (PANELMAJ branch)       */
/* ------------------------------------------------------------------------ */

struct blasfeo_dmat
	{
	double *mem;  /* start of the caller's memory chunk */
	double *pA;   /* pm*cn panel-major doubles */
	double *dA;   /* reciprocal diagonal cache, min(m,n) rounded up */
	int m;        /* rows */
	int n;        /* columns */
	int pm;       /* rows rounded up to the panel size */
	int cn;       /* columns rounded up to the panel leading dimension (panel stride) */
	int use_dA;   /* >0: dA holds valid reciprocals of the first use_dA diagonal entries */
	int memsize;  /* bytes needed */
	};

struct blasfeo_smat
	{
	float *mem;
	float *pA;
	float *dA;
	int m;
	int n;
	int pm;
	int cn;
	int use_dA;
	int memsize;
	};

struct blasfeo_dvec
	{
	double *mem;
	double *pa;
	int m;
	int pm;
	int memsize;
	};

struct blasfeo_svec
	{
	float *mem;
	float *pa;
	int m;
	int pm;
	int memsize;
	};

/* include/blasfeo_common.h:172-185: explicitly panel-major, runtime panel size */
struct blasfeo_pm_dmat
	{
	double *mem;
	double *pA;
	double *dA;
	int m;
	int n;
	int pm;
	int cn;
	int use_dA;
	int ps;
	int memsize;
	};

/* ------------------------------------------------------------------------ */
/* status codes and library control                                         */
/* ------------------------------------------------------------------------ */

#define BLASFEO_AMD_OK          0
#define BLASFEO_AMD_EINVAL     -1  /* bad argument (batch<0, unsupported ps, NULL array with batch>0) */
#define BLASFEO_AMD_ELAUNCH    -2  /* HIP runtime reported an error at launch */
#define BLASFEO_AMD_ENODEV     -3  /* no usable gfx950 device */
#define BLASFEO_AMD_ETOOBIG    -4  /* a size exceeds what the selected kernel supports */
#define BLASFEO_AMD_ENOMEM     -5  /* device allocation failed (single-instance helpers only) */

/* HIP stream handle without pulling in HIP headers (hipStream_t is a pointer). */
typedef void *blasfeo_amd_stream_t;

/* Library version string and the last error message of the calling thread. */
const char *blasfeo_amd_version(void);
const char *blasfeo_amd_last_error(void);

/* Launch-shape knobs read by the batched entries (process-wide; no counterpart
 * in the reference, whose block sizes are fixed at build time).  Set by the
 * autotuner (blasfeo_amd/autotune.py) from measurements on the caller's
 * workload; 0 restores the built-in choice.
 *   "ric_per_cu",  index = Riccati LDS class 0..7: resident instances per CU
 *                  of the class's persistent launch (1..64);
 *   "ric_threads", index = class 0..7: threads per instance (64, 128 or 256;
 *                  64 only for classes 0..4);
 *   "gemm_grid_y", index 0: most workgroups per gemm/syrk/trmm instance (1..64);
 *   "potrf_bw",    index 0: tile columns per block panel of potrf_l for
 *                  128 < n <= 256 (1, 2 or 4: 16-, 32- or 64-column panels,
 *                  four, two or one instance per CU; built-in 1);
 *   "trsm_thin",   index 0: left-side solves (llnn, lltn, llnu, lunn) with
 *                  hint->max_n <= 4 and m <= 256 on one wavefront per instance:
 *                  1 always, 2 never; built-in: when the batch has at least
 *                  16 instances per CU;
 *   "gemm_wave",   index 0: gemm_nt/nn and syrk_ln(_mn) with size bounds
 *                  up to 256 on one wavefront per 32 x 32 output tile
 *                  (operands from global memory): 1 or 0 (built-in) when the
 *                  hint bounds m and n by 256, 2 never (the LDS-staged
 *                  64 x 64 engine);
 *   "lds_threads", index = 16-row bucket 2..7 (rows 16 index + 1 .. 16 index + 16)
 *                  of the LDS-resident potrf_l(_mn) and dsyrk_dpotrf_ln launches
 *                  (size bound 33..128): threads per workgroup, 256 or 512;
 *                  built-in 256.
 * blasfeo_amd_tune_set returns BLASFEO_AMD_OK or BLASFEO_AMD_EINVAL;
 * blasfeo_amd_tune_get returns the value (0 = built-in) or BLASFEO_AMD_EINVAL. */
int blasfeo_amd_tune_set(const char *name, int index, int value);
int blasfeo_amd_tune_get(const char *name, int index);

/* Panel size used by the single-instance API and by blasfeo_memsize_dmat /
 * blasfeo_create_dmat (double: 4 or 8, float: panel size is 8 for ps=4 and
 * 16 for ps=8, following include/blasfeo_block_size.h:99-108 and :53-62).
 * Returns BLASFEO_AMD_EINVAL for other values. */
int blasfeo_amd_set_ps(int ps);
int blasfeo_amd_get_ps(void);

/* ------------------------------------------------------------------------ */
/* memory and layout helpers (auxiliary/d_aux_lib4.c, d_aux_lib8.c)          */
/* ------------------------------------------------------------------------ */

/* auxiliary/d_aux_lib4.c:1519-1530 (ps=4, pld=4) and d_aux_lib8.c:64-75 (ps=8, pld=8) */
size_t blasfeo_memsize_dmat(int m, int n);
size_t blasfeo_memsize_diag_dmat(int m, int n);
size_t blasfeo_memsize_dvec(int m);
/* auxiliary/d_aux_lib4.c:1548-1570 */
void blasfeo_create_dmat(int m, int n, struct blasfeo_dmat *sA, void *memory);
void blasfeo_create_dvec(int m, struct blasfeo_dvec *sa, void *memory);
/* auxiliary/d_aux_common.c:49-83 (runtime panel size, pld=4) */
size_t blasfeo_pm_memsize_dmat(int ps, int m, int n);
void blasfeo_pm_create_dmat(int ps, int m, int n, struct blasfeo_pm_dmat *sA, void *memory);
/* the same layout rules with an explicit panel size for struct blasfeo_dmat */
size_t blasfeo_amd_memsize_dmat(int ps, int m, int n);
void blasfeo_amd_create_dmat(int ps, int m, int n, struct blasfeo_dmat *sA, void *memory);
size_t blasfeo_amd_memsize_smat(int ps, int m, int n);
void blasfeo_amd_create_smat(int ps, int m, int n, struct blasfeo_smat *sA, void *memory);

/* Column-major <-> panel-major conversion (auxiliary/d_aux_lib4.c:1607, :2238,
 * :2086).  Work on host or device memory: when sA's memory is device memory the
 * conversion is staged through host memory and copied synchronously. */
void blasfeo_pack_dmat(int m, int n, double *A, int lda, struct blasfeo_dmat *sA, int ai, int aj);
void blasfeo_pack_tran_dmat(int m, int n, double *A, int lda, struct blasfeo_dmat *sA, int ai, int aj);
void blasfeo_unpack_dmat(int m, int n, struct blasfeo_dmat *sA, int ai, int aj, double *A, int lda);
void blasfeo_pack_dvec(int m, double *x, int incx, struct blasfeo_dvec *sy, int yi);
void blasfeo_unpack_dvec(int m, struct blasfeo_dvec *sx, int xi, double *y, int incy);

/* auxiliary/x_aux_ext_dep.c:40-59: zeroed allocation; here in GPU memory. */
void blasfeo_allocate_dmat(int m, int n, struct blasfeo_dmat *sA);
void blasfeo_free_dmat(struct blasfeo_dmat *sA);
void blasfeo_allocate_dvec(int m, struct blasfeo_dvec *sa);
void blasfeo_free_dvec(struct blasfeo_dvec *sa);

/* ------------------------------------------------------------------------ */
/* single-instance struct API (include/blasfeo_d_blasfeo_api.h)              */
/* Descriptors are host structs whose pA/dA point to GPU memory.             */
/* ------------------------------------------------------------------------ */

/* include/blasfeo_d_blasfeo_api.h:141 */
void blasfeo_dgemm_nn(int m, int n, int k, double alpha, struct blasfeo_dmat *sA, int ai, int aj, struct blasfeo_dmat *sB, int bi, int bj, double beta, struct blasfeo_dmat *sC, int ci, int cj, struct blasfeo_dmat *sD, int di, int dj);
/* include/blasfeo_d_blasfeo_api.h:143 */
void blasfeo_dgemm_nt(int m, int n, int k, double alpha, struct blasfeo_dmat *sA, int ai, int aj, struct blasfeo_dmat *sB, int bi, int bj, double beta, struct blasfeo_dmat *sC, int ci, int cj, struct blasfeo_dmat *sD, int di, int dj);
/* include/blasfeo_d_blasfeo_api.h:149 */
void blasfeo_dsyrk_ln(int m, int k, double alpha, struct blasfeo_dmat *sA, int ai, int aj, struct blasfeo_dmat *sB, int bi, int bj, double beta, struct blasfeo_dmat *sC, int ci, int cj, struct blasfeo_dmat *sD, int di, int dj);
/* include/blasfeo_d_blasfeo_api.h:150 */
void blasfeo_dsyrk_ln_mn(int m, int n, int k, double alpha, struct blasfeo_dmat *sA, int ai, int aj, struct blasfeo_dmat *sB, int bi, int bj, double beta, struct blasfeo_dmat *sC, int ci, int cj, struct blasfeo_dmat *sD, int di, int dj);
/* include/blasfeo_d_blasfeo_api.h:195 */
void blasfeo_dtrmm_rlnn(int m, int n, double alpha, struct blasfeo_dmat *sA, int ai, int aj, struct blasfeo_dmat *sB, int bi, int bj, struct blasfeo_dmat *sD, int di, int dj);
/* include/blasfeo_d_blasfeo_api.h:222 */
void blasfeo_dtrsm_rltn(int m, int n, double alpha, struct blasfeo_dmat *sA, int ai, int aj, struct blasfeo_dmat *sB, int bi, int bj, struct blasfeo_dmat *sD, int di, int dj);
/* include/blasfeo_d_blasfeo_api.h:280-281 */
void blasfeo_dpotrf_l(int m, struct blasfeo_dmat *sC, int ci, int cj, struct blasfeo_dmat *sD, int di, int dj);
void blasfeo_dpotrf_l_mn(int m, int n, struct blasfeo_dmat *sC, int ci, int cj, struct blasfeo_dmat *sD, int di, int dj);
/* include/blasfeo_d_blasfeo_api.h:284-285 */
void blasfeo_dsyrk_dpotrf_ln(int m, int k, struct blasfeo_dmat *sA, int ai, int aj, struct blasfeo_dmat *sB, int bi, int bj, struct blasfeo_dmat *sC, int ci, int cj, struct blasfeo_dmat *sD, int di, int dj);
/* include/blasfeo_d_blasfeo_api.h:291 */
void blasfeo_dgetrf_rp(int m, int n, struct blasfeo_dmat *sC, int ci, int cj, struct blasfeo_dmat *sD, int di, int dj, int *ipiv);

/* ------------------------------------------------------------------------ */
/* batched argument records (device-resident arrays, one record per call)    */
/* ------------------------------------------------------------------------ */

/* one blasfeo_dgemm_nt / _nn / dsyrk_ln(_mn) call (include/blasfeo_d_blasfeo_api.h:141-150).
 * For syrk_ln, n is ignored (the square m x m lower triangle is written);
 * for syrk_ln_mn, n columns of the m x n lower trapezoid are written. */
struct blasfeo_amd_dgemm_arg
	{
	int m, n, k;
	int ai, aj, bi, bj, ci, cj, di, dj;
	double alpha, beta;
	struct blasfeo_dmat *sA, *sB, *sC, *sD;
	};

/* one blasfeo_dtrsm_rltn / dtrmm_rlnn call (include/blasfeo_d_blasfeo_api.h:195,222) */
struct blasfeo_amd_dtrsm_arg
	{
	int m, n;
	int ai, aj, bi, bj, di, dj;
	double alpha;
	struct blasfeo_dmat *sA, *sB, *sD;
	};

/* one blasfeo_dpotrf_l(_mn) / dgetrf_rp call (include/blasfeo_d_blasfeo_api.h:280-291).
 * n is used by potrf_l_mn and getrf_rp; ipiv (device, >= min(m,n) ints) by getrf_rp. */
struct blasfeo_amd_dfact_arg
	{
	int m, n;
	int ci, cj, di, dj;
	struct blasfeo_dmat *sC, *sD;
	int *ipiv;
	};

/* one blasfeo_dsyrk_dpotrf_ln(_mn) call (include/blasfeo_d_blasfeo_api.h:284-285) */
struct blasfeo_amd_dsyrk_dpotrf_arg
	{
	int m, n, k;
	int ai, aj, bi, bj, ci, cj, di, dj;
	struct blasfeo_dmat *sA, *sB, *sC, *sD;
	};

/* FP32 twins (include/blasfeo_s_blasfeo_api.h) */
struct blasfeo_amd_sgemm_arg
	{
	int m, n, k;
	int ai, aj, bi, bj, ci, cj, di, dj;
	float alpha, beta;
	struct blasfeo_smat *sA, *sB, *sC, *sD;
	};
struct blasfeo_amd_strsm_arg
	{
	int m, n;
	int ai, aj, bi, bj, di, dj;
	float alpha;
	struct blasfeo_smat *sA, *sB, *sD;
	};
struct blasfeo_amd_sfact_arg
	{
	int m, n;
	int ci, cj, di, dj;
	struct blasfeo_smat *sC, *sD;
	int *ipiv;
	};

/* Optional host-side hints: upper bounds of the sizes in the batch, used to
 * pick the kernel variant.  Zero means unknown.  May be NULL. */
struct blasfeo_amd_hint
	{
	int max_m, max_n, max_k;
	};

/* ------------------------------------------------------------------------ */
/* batched entry points (device-resident args; asynchronous on `stream`)     */
/* ps: panel size of every matrix in the batch (double: 4 or 8).             */
/* ------------------------------------------------------------------------ */

/* replaces batch x blasfeo_dgemm_nt (blasfeo_hp_pm/d_blas3_lib4.c:7215,727) */
int blasfeo_amd_dgemm_nt_batched(int ps, int batch, const struct blasfeo_amd_dgemm_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);
/* replaces batch x blasfeo_dgemm_nn (blasfeo_hp_pm/d_blas3_lib4.c:7208,315) */
int blasfeo_amd_dgemm_nn_batched(int ps, int batch, const struct blasfeo_amd_dgemm_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);
/* replaces batch x blasfeo_dsyrk_ln (blasfeo_hp_pm/d_blas3_lib4.c:7460,5400) */
int blasfeo_amd_dsyrk_ln_batched(int ps, int batch, const struct blasfeo_amd_dgemm_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);
/* replaces batch x blasfeo_dsyrk_ln_mn (blasfeo_hp_pm/d_blas3_lib4.c:6459) */
int blasfeo_amd_dsyrk_ln_mn_batched(int ps, int batch, const struct blasfeo_amd_dgemm_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);
/* replaces batch x blasfeo_dtrsm_rltn (blasfeo_hp_pm/d_blas3_lib4.c:7306,4123) */
int blasfeo_amd_dtrsm_rltn_batched(int ps, int batch, const struct blasfeo_amd_dtrsm_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);
/* replaces batch x blasfeo_dtrsm_llnn (blasfeo_hp_pm/d_blas3_lib4.c:7236,3481;
 * blasfeo_ref/x_blas3_ref.c:445): D = alpha * A^{-1} * B, A lower m x m, B m x n */
int blasfeo_amd_dtrsm_llnn_batched(int ps, int batch, const struct blasfeo_amd_dtrsm_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);
/* replaces batch x blasfeo_dtrsm_lltn (blasfeo_hp_pm/d_blas3_lib4.c:3782;
 * blasfeo_ref/x_blas3_ref.c:674): D = alpha * A^{-T} * B, A lower m x m, B m x n */
int blasfeo_amd_dtrsm_lltn_batched(int ps, int batch, const struct blasfeo_amd_dtrsm_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);
/* replaces batch x blasfeo_dtrmm_rlnn (blasfeo_hp_pm/d_blas3_lib4.c:4893) */
int blasfeo_amd_dtrmm_rlnn_batched(int ps, int batch, const struct blasfeo_amd_dtrsm_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);
/* LU solve pieces (examples/example_d_lu_factorization.c:163-171: drowpe, then
 * trsm_llnu with the unit-lower L, then trsm_lunn with the upper U of one
 * getrf_rp result).
 * replaces batch x blasfeo_dtrsm_llnu (blasfeo_ref/x_blas3_ref.c:572): D = alpha * L^{-1} * B,
 * L unit lower m x m (its diagonal is not read); sD->use_dA = 0, sA untouched */
int blasfeo_amd_dtrsm_llnu_batched(int ps, int batch, const struct blasfeo_amd_dtrsm_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);
/* replaces batch x blasfeo_dtrsm_lunn (blasfeo_ref/x_blas3_ref.c:920): D = alpha * U^{-1} * B,
 * U the upper triangle of sA (m x m, non-unit); the dA / use_dA rule of trsm_lltn */
int blasfeo_amd_dtrsm_lunn_batched(int ps, int batch, const struct blasfeo_amd_dtrsm_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);
/* One blasfeo_drowpe(kmax, ipiv, sA) (auxiliary/d_aux_lib4.c:2767): rows ii and
 * ipiv[ii] swapped over all sA->n columns for ii = 0 .. kmax-1 in order;
 * sA->use_dA = 0.  ipiv is a device array; kmax is capped at sA->m and entries
 * outside 0 .. sA->m-1 are skipped (the reference would write out of bounds). */
struct blasfeo_amd_drowpe_arg
	{
	int kmax;
	int *ipiv;
	struct blasfeo_dmat *sA;
	};
int blasfeo_amd_drowpe_batched(int ps, int batch, const struct blasfeo_amd_drowpe_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);
/* replaces batch x blasfeo_dpotrf_l (blasfeo_hp_pm/d_lapack_lib4.c:4269,918) */
int blasfeo_amd_dpotrf_l_batched(int ps, int batch, const struct blasfeo_amd_dfact_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);
/* replaces batch x blasfeo_dpotrf_l_mn (blasfeo_hp_pm/d_lapack_lib4.c:1155) */
int blasfeo_amd_dpotrf_l_mn_batched(int ps, int batch, const struct blasfeo_amd_dfact_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);
/* replaces batch x blasfeo_dsyrk_dpotrf_ln(_mn) (blasfeo_hp_pm/d_lapack_lib4.c:1848) */
int blasfeo_amd_dsyrk_dpotrf_ln_batched(int ps, int batch, const struct blasfeo_amd_dsyrk_dpotrf_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);
/* replaces batch x blasfeo_dgetrf_np (blasfeo_hp_pm/d_lapack_lib4.c:2078; REF
 * blasfeo_ref/x_lapack_ref.c:880): LU without pivoting, unit-lower L and U in D,
 * dA = 1/U(j,j) (a zero pivot gives inf, as the reference divides unconditionally),
 * use_dA = 1 at offset (0,0); the record's ipiv is ignored */
int blasfeo_amd_dgetrf_np_batched(int ps, int batch, const struct blasfeo_amd_dfact_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);
/* replaces batch x blasfeo_dgetrf_rp (blasfeo_hp_pm/d_lapack_lib4.c:4311,2396) */
int blasfeo_amd_dgetrf_rp_batched(int ps, int batch, const struct blasfeo_amd_dfact_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);

/* FP32 twins: blasfeo_hp_pm/s_blas3_lib8.c:48,316,751,972; s_lapack_lib8.c:49,931.
 * ps is the float panel size (4, 8 or 16). */
int blasfeo_amd_sgemm_nt_batched(int ps, int batch, const struct blasfeo_amd_sgemm_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);
int blasfeo_amd_sgemm_nn_batched(int ps, int batch, const struct blasfeo_amd_sgemm_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);
int blasfeo_amd_ssyrk_ln_batched(int ps, int batch, const struct blasfeo_amd_sgemm_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);
int blasfeo_amd_strsm_rltn_batched(int ps, int batch, const struct blasfeo_amd_strsm_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);
int blasfeo_amd_spotrf_l_batched(int ps, int batch, const struct blasfeo_amd_sfact_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);
int blasfeo_amd_sgetrf_rp_batched(int ps, int batch, const struct blasfeo_amd_sfact_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);

/* ------------------------------------------------------------------------ */
/* column-major <-> panel-major conversion on the device                    */
/* ------------------------------------------------------------------------ */

/* One blasfeo_pack_dmat(m, n, A, lda, sA, ai, aj) (auxiliary/d_aux_lib4.c:1607),
 * blasfeo_pack_tran_dmat (:2086; the transpose of the m x n A into sA, an
 * n x m block: sA(ai+r, aj+c) = A[c + r*lda]) or blasfeo_unpack_dmat(m, n, sA,
 * ai, aj, A, lda) (:2238), with A a column-major device buffer (lib8 twins in
 * auxiliary/d_aux_lib8.c).  pack and pack_tran set sA->use_dA = 0, as the
 * reference's do. */
struct blasfeo_amd_dpack_arg
	{
	int m, n, lda, ai, aj;
	double *A;
	struct blasfeo_dmat *sA;
	};

/* Batched conversions, one workgroup per instance (more for wide blocks:
 * hint->max_n, the largest n, may be given), each 64 x 16 tile staged through
 * LDS so that the column-major and the panel-major sides are both read and
 * written in contiguous runs. */
int blasfeo_amd_dpack_batched(int ps, int batch, const struct blasfeo_amd_dpack_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);
int blasfeo_amd_dpack_tran_batched(int ps, int batch, const struct blasfeo_amd_dpack_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);
int blasfeo_amd_dunpack_batched(int ps, int batch, const struct blasfeo_amd_dpack_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);

/* Panel size conversion on the device.  No counterpart in the reference, whose
 * blasfeo_pm_dmat API (include/blasfeo_d_aux.h:222-226) only sizes, creates
 * and prints: one blasfeo_dgecp(m, n, sA, ai, aj, sB, bi, bj)
 * (auxiliary/d_aux_lib4.c:3006; B(bi+i, bj+j) = A(ai+i, aj+j), sB->use_dA = 0)
 * between blasfeo_pm_dmat descriptors whose panel sizes (4 or 8, from their ps
 * fields) may differ, so that a batch can move between ps = 4 and ps = 8
 * without a host round trip.  Instances with another panel size are skipped. */
struct blasfeo_amd_pm_dgecp_arg
	{
	int m, n, ai, aj, bi, bj;
	struct blasfeo_pm_dmat *sA;
	struct blasfeo_pm_dmat *sB;
	};
int blasfeo_amd_pm_dgecp_batched(int batch, const struct blasfeo_amd_pm_dgecp_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Riccati recursion (examples/example_d_riccati_recursion.c:115-140)        */
/* ------------------------------------------------------------------------ */

/* One instance of d_back_ric_trf_libstr(N, nx, nu, hsBAbt, hsRSQrq, hsL, hswork_mat).
 * nx, nu: device arrays of N+1 ints.  BAbt: device array of N descriptors,
 * BAbt[k] holds [B_k; A_k]^T, (nu[k]+nx[k]) x nx[k+1].  RSQrq: N+1 descriptors,
 * lower triangle of [R S; S^T Q], (nu[k]+nx[k])^2.  L: N+1 output descriptors.
 * work: one descriptor of at least max_k (nu[k]+nx[k]) x nx[k+1], used only
 * by the unfused (graph) form. */
struct blasfeo_amd_riccati_arg
	{
	int N;
	const int *nx;
	const int *nu;
	struct blasfeo_dmat *BAbt;
	struct blasfeo_dmat *RSQrq;
	struct blasfeo_dmat *L;
	struct blasfeo_dmat *work;
	};

/* Fused backward factorisation sweep, one workgroup per instance, all stages in
 * one launch (the per-stage trmm_rlnn + syrk_dpotrf_ln chain of
 * examples/example_d_riccati_recursion.c:120-138).  hint->max_m bounds
 * nu[k]+nx[k] (at most 1024) and hint->max_n bounds nx[k] (0: max_m). */
int blasfeo_amd_d_back_ric_trf_batched(int ps, int batch, const struct blasfeo_amd_riccati_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);

/* The same sweep for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc); it has no
 * counterpart in the reference.  It allocates and creates nothing: the plan's
 * class counts, queue heads and lists live in a fixed device array of the
 * library (2.4 MB per GPU, held until the process ends), and the class
 * launches run in order on `stream` instead of side by side on helper
 * streams.  At most 65536 instances per call (more: BLASFEO_AMD_ETOOBIG).
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (other streams, or other host threads) race and must not be issued. */
int blasfeo_amd_d_back_ric_trf_batched_noalloc(int ps, int batch, const struct blasfeo_amd_riccati_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);

/* One instance of d_back_ric_trs_libstr(N, nx, nu, hsBAbt, hsb, hsrq, hsL, hsPb,
 * hsux, hspi, hswork_vec) (examples/example_d_riccati_recursion.c:203-267):
 * the backward and forward substitutions on the factors L[0..N] of
 * blasfeo_amd_d_back_ric_trf_batched.  Device arrays of descriptors:
 * BAbt[N] as above, b[N] (nx[k+1]), rq[N+1] (nu[k]+nx[k]), L[N+1], and the
 * outputs Pb[N] (nx[k+1]), ux[N+1] (nu[k]+nx[k]) and pi[N] (nx[k+1]).
 * L[k]->dA / use_dA follow blasfeo_dtrsv_lnn(_mn) and blasfeo_dtrsv_ltn(_mn)
 * (blasfeo_hp_pm/d_blas2_lib4.c:1275-1330): the cached reciprocal diagonal is
 * used when use_dA == 1, else it is formed, stored and use_dA set to 1.
 * work (one vector of max nx) is accepted for the reference's argument list
 * and not written: the kernel keeps that vector on chip. */
struct blasfeo_amd_riccati_trs_arg
	{
	int N;
	const int *nx;
	const int *nu;
	struct blasfeo_dmat *BAbt;
	struct blasfeo_dvec *b;
	struct blasfeo_dvec *rq;
	struct blasfeo_dmat *L;
	struct blasfeo_dvec *Pb;
	struct blasfeo_dvec *ux;
	struct blasfeo_dvec *pi;
	struct blasfeo_dvec *work;
	};

/* Batched Riccati solve, one wavefront per instance, both sweeps in one launch
 * (the trmv_ltn / trmv_lnn / gemv_n / trsv_lnn_mn backward chain and the
 * trsv_ltn_mn / gemv_t / trmv forward chain of
 * examples/example_d_riccati_recursion.c:211-264).  hint->max_m bounds
 * nu[k]+nx[k] and hint->max_n bounds nx[k], both <= 128. */
int blasfeo_amd_d_back_ric_trs_batched(int ps, int batch, const struct blasfeo_amd_riccati_trs_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);

/* One instance of d_back_ric_sv_libstr(N, nx, nu, hsBAbt, hsRSQrq, hsL, hsux,
 * hspi, hswork_mat, hswork_vec) (examples/example_d_riccati_recursion.c:58-111):
 * the factorisation and the solve in one recursion, on matrices augmented by
 * one row that carries the linear terms (nk = nu[k] + nx[k]):
 *   BAbt[k]  (nk+1) x nx[k+1]: [B_k^T; A_k^T; b_k^T];
 *   RSQrq[k] (nk+1) x nk: the lower triangle of [R S; S^T Q] and the row [r; q]^T
 *            (RSQrq[N]: (nx[N]+1) x nx[N], [Q_N; q_N^T]);
 *   L[k]     (nk+1) x nk outputs (the factor and the row of the linear term);
 *   ux[0..N] (nk) and pi[0..N-1] (nx[k+1]) outputs.
 * work_mat and work_vec are accepted for the reference's argument list and not
 * written: the sweep keeps those intermediates on chip. */
struct blasfeo_amd_riccati_sv_arg
	{
	int N;
	const int *nx;
	const int *nu;
	struct blasfeo_dmat *BAbt;
	struct blasfeo_dmat *RSQrq;
	struct blasfeo_dmat *L;
	struct blasfeo_dvec *ux;
	struct blasfeo_dvec *pi;
	struct blasfeo_dmat *work_mat;
	struct blasfeo_dvec *work_vec;
	};

/* Batched d_back_ric_sv_libstr: the backward sweep on the augmented matrices
 * (the LDS-resident sweep of blasfeo_amd_d_back_ric_trf_batched, the factor's
 * last row carrying the linear terms), then one wavefront per instance for the
 * forward substitution.  hint->max_m bounds nu[k]+nx[k] (at most 127) and
 * hint->max_n bounds nx[k] (at most 64); larger bounds return
 * BLASFEO_AMD_ETOOBIG. */
int blasfeo_amd_d_back_ric_sv_batched(int ps, int batch, const struct blasfeo_amd_riccati_sv_arg *args, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);

/* Mixed-precision Riccati solve: no single counterpart in the reference; it
 * computes what d_back_ric_trf_libstr followed by d_back_ric_trs_libstr
 * (examples/example_d_riccati_recursion.c:115-140, 203-267) compute, the
 * solution of each instance's KKT system, to normwise backward error <= tol.
 * The factorisation runs in FP32 (stage matrices and factors held in FP32 in
 * LDS, half the bytes of the FP64 sweep; FP32 matrix-core trmm/syrk and
 * trailing updates; the factors written to L in FP64 storage), the solve on
 * those factors, then iterative refinement: FP64 residuals of the KKT
 * equations
 *   r_k = RSQ_k ux_k + rq_k + BAbt_k pi_k - [0; pi_{k-1}],
 *   s_k = BAbt_k^T ux_k + b_k - x_{k+1},
 * and a correction solved on the FP32 factors (the same solve with b = s,
 * rq = r).  The backward error is max_k of ||r_k|| / (||RSQ_k|| ||ux_k|| +
 * ||rq_k|| + ||BAbt_k|| ||pi_k|| + ||pi_{k-1}||) and ||s_k|| / (||BAbt_k^T||
 * ||ux_k|| + ||b_k|| + ||x_{k+1}||) (infinity norms).  An instance still
 * above tol after maxit corrections is factored and solved again in FP64
 * (its L and vectors overwritten).  Arguments as blasfeo_amd_riccati_arg and
 * blasfeo_amd_riccati_trs_arg (BAbt, RSQrq, b and rq are read-only inputs;
 * L, Pb, ux, pi outputs), plus correction vectors of the same shapes
 * (res_b as b, res_rq as rq, d_Pb as Pb, d_ux as ux, d_pi as pi) and one
 * info record per instance: iters = corrections applied, path 0 = mixed,
 * 1 = FP64, berr = final backward error.  hint->max_m bounds nu[k]+nx[k] (at
 * most 128) and hint->max_n bounds nx[k] (at most 64). */
struct blasfeo_amd_riccati_mixed_arg
	{
	int N;
	const int *nx;
	const int *nu;
	struct blasfeo_dmat *BAbt;
	struct blasfeo_dmat *RSQrq;
	struct blasfeo_dmat *L;
	struct blasfeo_dmat *work_mat;
	struct blasfeo_dvec *b;
	struct blasfeo_dvec *rq;
	struct blasfeo_dvec *Pb;
	struct blasfeo_dvec *ux;
	struct blasfeo_dvec *pi;
	struct blasfeo_dvec *work_vec;
	struct blasfeo_dvec *res_b;
	struct blasfeo_dvec *res_rq;
	struct blasfeo_dvec *d_Pb;
	struct blasfeo_dvec *d_ux;
	struct blasfeo_dvec *d_pi;
	struct blasfeo_amd_solve_info *info;
	};
int blasfeo_amd_d_ric_solve_mixed_batched(int ps, int batch, const struct blasfeo_amd_riccati_mixed_arg *args, double tol, int maxit, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Cholesky solve A X = B with A symmetric positive definite                */
/* ------------------------------------------------------------------------ */

/* Outcome of one mixed-precision solve.
 * iters: conjugate-gradient iterations (products A p) of the refinement;
 * path:  0 = FP32 factor + FP64 refinement converged,
 *        1 = solved by the FP64 factorisation (the FP32 factor broke down or
 *            the refinement did not reach tol within max_iter);
 * berr:  normwise backward error max_c |b_c - A x_c|_inf / (|A|_inf |x_c|_inf + |b_c|_inf)
 *        of the returned solution when path = 0 (-1 when path = 1). */
struct blasfeo_amd_solve_info
	{
	int iters;
	int path;
	double berr;
	};

/* One solve: n x n SPD A (lower triangle read) at (ai,aj), right-hand sides
 * B (n x nrhs) at (bi,bj), solution X (n x nrhs) at (xi,xj).  sL (n x n at
 * (li,lj)) receives the FP64 Cholesky factor of instances that take path 1
 * (it may alias sA; then A is overwritten for those instances only).
 * info may be NULL. */
struct blasfeo_amd_dposv_arg
	{
	int n, nrhs;
	int ai, aj, bi, bj, xi, xj, li, lj;
	struct blasfeo_dmat *sA, *sB, *sX, *sL;
	struct blasfeo_amd_solve_info *info;
	};

/* No single counterpart in the reference: it computes what the sequence
 * blasfeo_dpotrf_l + blasfeo_dtrsm_llnn + blasfeo_dtrsm_lltn
 * (blasfeo_hp_pm/d_lapack_lib4.c:4269; d_blas3_lib4.c:7236,3782) computes, to
 * normwise backward error <= tol.  The factor is formed in FP32 and kept in
 * LDS (FP32 matrix-core trailing updates, FP64 16x16 diagonal blocks); the
 * refinement is conjugate gradients in FP64 (products with the FP64 A)
 * preconditioned by that factor, at most max_iter iterations, stopped when the
 * true residual meets berr <= tol.  Instances that do not converge are
 * re-solved in FP64 into sL on the device; n > 256 always takes that path. */
int blasfeo_amd_dposv_mixed_batched(int ps, int batch, const struct blasfeo_amd_dposv_arg *args, double tol, int max_iter, const struct blasfeo_amd_hint *hint, blasfeo_amd_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* BLASFEO_AMD_H */
