/*
 * oracle.c -- CPU oracle for the batched BLASFEO hot path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg, always as the checker or the reported CPU
 * baseline, never as the thing measured or shipped.  The GPU library
 * (blasfeo_amd/) neither links nor calls it.
 *
 * The arithmetic restates the reference's REFERENCE backend
 * (blasfeo_ref/x_blas3_ref.c, blasfeo_ref/x_lapack_ref.c); see oracle_impl.h
 * for the per-routine citations.  Built with -ffp-contract=off so that every
 * multiply-add rounds twice, as the reference C does on a non-FMA build.
 *
 * Pinning (tests/test_oracle.py): closed forms (A*I = A, trsm with A = I,
 * diagonal potrf = sqrt), the reference's integer-valued test data
 * (tests/test_x.c:144-169) against exact rational arithmetic, backward-error
 * identities (A = L L^T, P A = L U), permutation matrices with known pivots,
 * exact ties (first index wins) and zero pivots, and the Riccati closed form
 * P_k = (N-k+1) I for A = I, B = 0, Q = R = I.
 */
#include <math.h>
#include <stddef.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "../include/blasfeo_amd.h"

#define REAL double
#define MAT struct blasfeo_dmat
#define VEC struct blasfeo_dvec
#define PFX(name) or_d##name
#define SQRT sqrt
#include "oracle_impl.h"
#undef REAL
#undef MAT
#undef VEC
#undef PFX
#undef SQRT

#define REAL float
#define MAT struct blasfeo_smat
#define VEC struct blasfeo_svec
#define PFX(name) or_s##name
#define SQRT sqrtf
#include "oracle_impl.h"
#undef REAL
#undef MAT
#undef VEC
#undef PFX
#undef SQRT

/* ---- batched drivers over the same argument records as the GPU entries ---- */
/* (host pointers inside the records; instances are independent, so they are
 * spread over threads with a static schedule when OpenMP is enabled) */

void or_dgemm_nt_batch(int ps, int batch, const struct blasfeo_amd_dgemm_arg *a)
	{
#pragma omp parallel for schedule(dynamic, 16)
	for(int i = 0; i < batch; i++)
		or_dgemm_nt(ps, a[i].m, a[i].n, a[i].k, a[i].alpha, a[i].sA, a[i].ai, a[i].aj, a[i].sB, a[i].bi, a[i].bj, a[i].beta, a[i].sC, a[i].ci, a[i].cj, a[i].sD, a[i].di, a[i].dj);
	}

void or_dgemm_nn_batch(int ps, int batch, const struct blasfeo_amd_dgemm_arg *a)
	{
#pragma omp parallel for schedule(dynamic, 16)
	for(int i = 0; i < batch; i++)
		or_dgemm_nn(ps, a[i].m, a[i].n, a[i].k, a[i].alpha, a[i].sA, a[i].ai, a[i].aj, a[i].sB, a[i].bi, a[i].bj, a[i].beta, a[i].sC, a[i].ci, a[i].cj, a[i].sD, a[i].di, a[i].dj);
	}

void or_dsyrk_ln_batch(int ps, int batch, const struct blasfeo_amd_dgemm_arg *a)
	{
#pragma omp parallel for schedule(dynamic, 16)
	for(int i = 0; i < batch; i++)
		or_dsyrk_ln(ps, a[i].m, a[i].k, a[i].alpha, a[i].sA, a[i].ai, a[i].aj, a[i].sB, a[i].bi, a[i].bj, a[i].beta, a[i].sC, a[i].ci, a[i].cj, a[i].sD, a[i].di, a[i].dj);
	}

void or_dpotrf_l_batch(int ps, int batch, const struct blasfeo_amd_dfact_arg *a)
	{
#pragma omp parallel for schedule(dynamic, 16)
	for(int i = 0; i < batch; i++)
		or_dpotrf_l(ps, a[i].m, a[i].sC, a[i].ci, a[i].cj, a[i].sD, a[i].di, a[i].dj);
	}

void or_dgetrf_rp_batch(int ps, int batch, const struct blasfeo_amd_dfact_arg *a)
	{
#pragma omp parallel for schedule(dynamic, 16)
	for(int i = 0; i < batch; i++)
		or_dgetrf_rp(ps, a[i].m, a[i].n, a[i].sC, a[i].ci, a[i].cj, a[i].sD, a[i].di, a[i].dj, a[i].ipiv);
	}

void or_sgetrf_rp_batch(int ps, int batch, const struct blasfeo_amd_sfact_arg *a)
	{
#pragma omp parallel for schedule(dynamic, 16)
	for(int i = 0; i < batch; i++)
		or_sgetrf_rp(ps, a[i].m, a[i].n, a[i].sC, a[i].ci, a[i].cj, a[i].sD, a[i].di, a[i].dj, a[i].ipiv);
	}

void or_dric_trf_batch(int ps, int batch, const struct blasfeo_amd_riccati_arg *a)
	{
#pragma omp parallel for schedule(dynamic, 4)
	for(int i = 0; i < batch; i++)
		or_dric_trf(ps, a[i].N, a[i].nx, a[i].nu, a[i].BAbt, a[i].RSQrq, a[i].L, a[i].work);
	}

void or_dpack_batch(int ps, int batch, const struct blasfeo_amd_dpack_arg *a)
	{
#pragma omp parallel for schedule(dynamic, 16)
	for(int i = 0; i < batch; i++)
		or_dpack(ps, a[i].m, a[i].n, a[i].A, a[i].lda, a[i].sA, a[i].ai, a[i].aj);
	}

void or_dpack_tran_batch(int ps, int batch, const struct blasfeo_amd_dpack_arg *a)
	{
#pragma omp parallel for schedule(dynamic, 16)
	for(int i = 0; i < batch; i++)
		or_dpack_tran(ps, a[i].m, a[i].n, a[i].A, a[i].lda, a[i].sA, a[i].ai, a[i].aj);
	}

void or_dunpack_batch(int ps, int batch, const struct blasfeo_amd_dpack_arg *a)
	{
#pragma omp parallel for schedule(dynamic, 16)
	for(int i = 0; i < batch; i++)
		or_dunpack(ps, a[i].m, a[i].n, a[i].sA, a[i].ai, a[i].aj, a[i].A, a[i].lda);
	}

/* blasfeo_dgecp between blasfeo_pm_dmat descriptors of any panel sizes:
 * B(bi+i, bj+j) = A(ai+i, aj+j), sB->use_dA = 0 (auxiliary/d_aux_lib4.c:3006) */
#define PMEL(s, i, j) ((s)->pA[((i) - (i) % (s)->ps) * (s)->cn + (j) * (s)->ps + (i) % (s)->ps])
void or_pm_dgecp(int m, int n, struct blasfeo_pm_dmat *sA, int ai, int aj, struct blasfeo_pm_dmat *sB, int bi, int bj)
	{
	sB->use_dA = 0;
	for(int j = 0; j < n; j++)
		for(int i = 0; i < m; i++)
			PMEL(sB, bi+i, bj+j) = PMEL(sA, ai+i, aj+j);
	}
#undef PMEL

void or_pm_dgecp_batch(int batch, const struct blasfeo_amd_pm_dgecp_arg *a)
	{
#pragma omp parallel for schedule(dynamic, 16)
	for(int i = 0; i < batch; i++)
		or_pm_dgecp(a[i].m, a[i].n, a[i].sA, a[i].ai, a[i].aj, a[i].sB, a[i].bi, a[i].bj);
	}

void or_dric_trs_batch(int ps, int batch, const struct blasfeo_amd_riccati_trs_arg *a)
	{
#pragma omp parallel for schedule(dynamic, 4)
	for(int i = 0; i < batch; i++)
		or_dric_trs(ps, a[i].N, a[i].nx, a[i].nu, a[i].BAbt, a[i].b, a[i].rq, a[i].L, a[i].Pb, a[i].ux, a[i].pi, a[i].work);
	}

void or_dric_sv_batch(int ps, int batch, const struct blasfeo_amd_riccati_sv_arg *a)
	{
#pragma omp parallel for schedule(dynamic, 4)
	for(int i = 0; i < batch; i++)
		or_dric_sv(ps, a[i].N, a[i].nx, a[i].nu, a[i].BAbt, a[i].RSQrq, a[i].L, a[i].ux, a[i].pi, a[i].work_mat, a[i].work_vec);
	}

/* The FP64 Cholesky solve a reference caller writes:
 * blasfeo_dpotrf_l(A -> L); blasfeo_dtrsm_llnn(L, B -> X); blasfeo_dtrsm_lltn(L, X -> X). */
void or_dposv_batch(int ps, int batch, const struct blasfeo_amd_dposv_arg *a)
	{
#pragma omp parallel for schedule(dynamic, 4)
	for(int i = 0; i < batch; i++)
		{
		or_dpotrf_l(ps, a[i].n, a[i].sA, a[i].ai, a[i].aj, a[i].sL, a[i].li, a[i].lj);
		or_dtrsm_llnn(ps, a[i].n, a[i].nrhs, 1.0, a[i].sL, a[i].li, a[i].lj, a[i].sB, a[i].bi, a[i].bj, a[i].sX, a[i].xi, a[i].xj);
		or_dtrsm_lltn(ps, a[i].n, a[i].nrhs, 1.0, a[i].sL, a[i].li, a[i].lj, a[i].sX, a[i].xi, a[i].xj, a[i].sX, a[i].xi, a[i].xj);
		}
	}

int or_num_threads(void)
	{
	int t = 1;
#ifdef _OPENMP
#pragma omp parallel
#pragma omp single
	t = omp_get_num_threads();
#endif
	return t;
	}
