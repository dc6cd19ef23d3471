/*
 * drop_in.c -- a C host program that uses the library the way a BLASFEO
 * caller does (TEST: run by tests/test_gpu_c_host.py on the GPU).
 *
 * It allocates matrices with blasfeo_allocate_dmat (GPU memory), fills them
 * with blasfeo_pack_dmat, and calls the reference-named single-instance
 * routines declared in include/blasfeo_amd.h with the reference's argument
 * lists (include/blasfeo_d_blasfeo_api.h:143,222,280,284,291):
 * blasfeo_dgemm_nt, blasfeo_dpotrf_l, blasfeo_dtrsm_rltn (the use_dA cache
 * of its input A, both used and formed), blasfeo_dsyrk_dpotrf_ln,
 * blasfeo_dgetrf_rp (pivots returned to the host), plus one batched entry
 * (blasfeo_amd_dpotrf_l_batched over device-resident descriptors and
 * records).  Every call is repeated by the CPU oracle on host copies created
 * with blasfeo_create_dmat on host memory, and every byte of each GPU matrix
 * (panels and dA) is compared with the reference's rule
 * (tests/test_x_common.c:38-42, tol 1e-11), use_dA exactly, pivots exactly.
 * Both panel sizes (blasfeo_amd_set_ps 4 and 8).  Prints one line per check;
 * exit status 0 when all pass.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "blasfeo_amd.h"

/* the CPU oracle (oracle/liboracle.so, test infrastructure) */
void or_dgemm_nt(int ps, int m, int n, int k, double alpha, struct blasfeo_dmat *sA, int ai, int aj, struct blasfeo_dmat *sB, int bi, int bj, double beta, struct blasfeo_dmat *sC, int ci, int cj, struct blasfeo_dmat *sD, int di, int dj);
void or_dtrsm_rltn(int ps, int m, int n, double alpha, struct blasfeo_dmat *sA, int ai, int aj, struct blasfeo_dmat *sB, int bi, int bj, struct blasfeo_dmat *sD, int di, int dj);
void or_dpotrf_l(int ps, int m, struct blasfeo_dmat *sC, int ci, int cj, struct blasfeo_dmat *sD, int di, int dj);
void or_dsyrk_potrf_ln(int ps, int m, int k, struct blasfeo_dmat *sA, int ai, int aj, struct blasfeo_dmat *sB, int bi, int bj, struct blasfeo_dmat *sC, int ci, int cj, struct blasfeo_dmat *sD, int di, int dj);
void or_dgetrf_rp(int ps, int m, int n, struct blasfeo_dmat *sC, int ci, int cj, struct blasfeo_dmat *sD, int di, int dj, int *ipiv);

static int g_fail = 0;
static unsigned long long g_seed = 12345;

static double urand(void)
{
	g_seed = g_seed * 6364136223846793005ULL + 1442695040888963407ULL;
	return (double)(g_seed >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

/* one matrix on the GPU and its host copy for the oracle */
struct pair
{
	struct blasfeo_dmat g, h;
	void *hmem;
	int m, n;
};

static void pair_new(struct pair *p, int m, int n, const double *A)
{
	p->m = m;
	p->n = n;
	blasfeo_allocate_dmat(m, n, &p->g);
	size_t sz = blasfeo_memsize_dmat(m, n);
	p->hmem = calloc(1, sz + 64);
	blasfeo_create_dmat(m, n, &p->h, p->hmem);
	if (A)
	{
		blasfeo_pack_dmat(m, n, (double *)A, m, &p->g, 0, 0);
		blasfeo_pack_dmat(m, n, (double *)A, m, &p->h, 0, 0);
	}
}

static void pair_free(struct pair *p)
{
	blasfeo_free_dmat(&p->g);
	free(p->hmem);
}

/* column-major m x n with entries U[-1,1] (+ shift on the diagonal) */
static double *dense(int m, int n, double diag)
{
	double *A = malloc(sizeof(double) * (size_t)m * n);
	for (int j = 0; j < n; j++)
		for (int i = 0; i < m; i++)
			A[i + (size_t)j * m] = urand() + (i == j ? diag : 0.0);
	return A;
}

/* SPD n x n: B B^T + n I */
static double *spd(int n)
{
	double *B = dense(n, n, 0.0), *A = malloc(sizeof(double) * (size_t)n * n);
	for (int j = 0; j < n; j++)
		for (int i = 0; i < n; i++)
		{
			double s = (i == j) ? n : 0.0;
			for (int l = 0; l < n; l++)
				s += B[i + (size_t)l * n] * B[j + (size_t)l * n];
			A[i + (size_t)j * n] = s;
		}
	free(B);
	return A;
}

static void check(struct pair *p, const char *what)
{
	size_t sz = p->g.memsize;
	double *gb = malloc(sz);
	if (hipMemcpy(gb, p->g.mem, sz, hipMemcpyDeviceToHost) != hipSuccess)
	{
		printf("FAIL %s: device copy\n", what);
		g_fail = 1;
		free(gb);
		return;
	}
	const double *hb = (const double *)p->h.mem;
	size_t bad = 0, first = 0;
	for (size_t e = 0; e < sz / sizeof(double); e++)
	{
		double s = gb[e], r = hb[e], d = fabs(s - r);
		int nan = isnan(s) != isnan(r);
		if (nan || (d > 1e-11 * (fabs(s) + fabs(r)) && d > 1e-11))
		{
			if (!bad)
				first = e;
			bad++;
		}
	}
	if (bad || p->g.use_dA != p->h.use_dA)
	{
		printf("FAIL %s: %zu elements differ (first at %zu: gpu %.17g oracle %.17g), use_dA gpu %d oracle %d\n",
			what, bad, first, bad ? gb[first] : 0.0, bad ? hb[first] : 0.0, p->g.use_dA, p->h.use_dA);
		g_fail = 1;
	}
	else
		printf("ok %s (use_dA %d)\n", what, p->g.use_dA);
	free(gb);
}

static void run(int ps)
{
	char w[128];
	blasfeo_amd_set_ps(ps);

	/* dgemm_nt at offsets: D(2:39, 1:30) = 1.5 A B^T - 0.5 C */
	{
		double *A = dense(43, 34, 0.0), *B = dense(35, 33, 0.0), *C = dense(45, 31, 0.0);
		struct pair a, b, c, d;
		pair_new(&a, 43, 34, A);
		pair_new(&b, 35, 33, B);
		pair_new(&c, 45, 31, C);
		pair_new(&d, 41, 33, NULL);
		blasfeo_dgemm_nt(37, 29, 30, 1.5, &a.g, 3, 1, &b.g, 5, 2, -0.5, &c.g, 6, 0, &d.g, 2, 1);
		or_dgemm_nt(ps, 37, 29, 30, 1.5, &a.h, 3, 1, &b.h, 5, 2, -0.5, &c.h, 6, 0, &d.h, 2, 1);
		snprintf(w, sizeof w, "ps=%d blasfeo_dgemm_nt", ps);
		check(&d, w);
		pair_free(&a); pair_free(&b); pair_free(&c); pair_free(&d);
		free(A); free(B); free(C);
	}

	/* dpotrf_l (n = 70 into a separate D, n = 200 in place), then
	 * dtrsm_rltn with A = the n = 70 factor: its cache is valid (use_dA = 70
	 * after potrf) and is used; then with A = a fresh lower matrix whose
	 * cache is formed and stored by the call */
	{
		double *S1 = spd(70), *S2 = spd(200), *Bm = dense(33, 70, 0.0), *T = dense(70, 70, 3.0);
		struct pair c1, d1, c2, b, x, t, x2;
		pair_new(&c1, 70, 70, S1);
		pair_new(&d1, 70, 70, NULL);
		pair_new(&c2, 200, 200, S2);
		blasfeo_dpotrf_l(70, &c1.g, 0, 0, &d1.g, 0, 0);
		or_dpotrf_l(ps, 70, &c1.h, 0, 0, &d1.h, 0, 0);
		snprintf(w, sizeof w, "ps=%d blasfeo_dpotrf_l n=70", ps);
		check(&d1, w);
		blasfeo_dpotrf_l(200, &c2.g, 0, 0, &c2.g, 0, 0);
		or_dpotrf_l(ps, 200, &c2.h, 0, 0, &c2.h, 0, 0);
		snprintf(w, sizeof w, "ps=%d blasfeo_dpotrf_l n=200 in place", ps);
		check(&c2, w);
		pair_new(&b, 33, 70, Bm);
		pair_new(&x, 33, 70, NULL);
		blasfeo_dtrsm_rltn(33, 70, 2.0, &d1.g, 0, 0, &b.g, 0, 0, &x.g, 0, 0);
		or_dtrsm_rltn(ps, 33, 70, 2.0, &d1.h, 0, 0, &b.h, 0, 0, &x.h, 0, 0);
		snprintf(w, sizeof w, "ps=%d blasfeo_dtrsm_rltn (cached dA)", ps);
		check(&x, w);
		check(&d1, "  ... its A");
		pair_new(&t, 70, 70, T);
		pair_new(&x2, 33, 70, NULL);
		blasfeo_dtrsm_rltn(33, 70, 1.0, &t.g, 0, 0, &b.g, 0, 0, &x2.g, 0, 0);
		or_dtrsm_rltn(ps, 33, 70, 1.0, &t.h, 0, 0, &b.h, 0, 0, &x2.h, 0, 0);
		snprintf(w, sizeof w, "ps=%d blasfeo_dtrsm_rltn (dA formed)", ps);
		check(&x2, w);
		check(&t, "  ... its A (dA and use_dA written)");
		pair_free(&c1); pair_free(&d1); pair_free(&c2); pair_free(&b); pair_free(&x); pair_free(&t); pair_free(&x2);
		free(S1); free(S2); free(Bm); free(T);
	}

	/* dsyrk_dpotrf_ln: D = chol(C + A A^T), m = 40, k = 10 */
	{
		double *A = dense(40, 10, 0.0), *C = spd(40);
		struct pair a, c, d;
		pair_new(&a, 40, 10, A);
		pair_new(&c, 40, 40, C);
		pair_new(&d, 40, 40, NULL);
		blasfeo_dsyrk_dpotrf_ln(40, 10, &a.g, 0, 0, &a.g, 0, 0, &c.g, 0, 0, &d.g, 0, 0);
		or_dsyrk_potrf_ln(ps, 40, 10, &a.h, 0, 0, &a.h, 0, 0, &c.h, 0, 0, &d.h, 0, 0);
		snprintf(w, sizeof w, "ps=%d blasfeo_dsyrk_dpotrf_ln", ps);
		check(&d, w);
		pair_free(&a); pair_free(&c); pair_free(&d);
		free(A); free(C);
	}

	/* dgetrf_rp 50 x 50 with the pivots returned to the host */
	{
		double *A = dense(50, 50, 0.0);
		struct pair c, d;
		int gp[50], hp[50];
		memset(gp, 0xff, sizeof gp);
		memset(hp, 0xff, sizeof hp);
		pair_new(&c, 50, 50, A);
		pair_new(&d, 50, 50, NULL);
		blasfeo_dgetrf_rp(50, 50, &c.g, 0, 0, &d.g, 0, 0, gp);
		or_dgetrf_rp(ps, 50, 50, &c.h, 0, 0, &d.h, 0, 0, hp);
		snprintf(w, sizeof w, "ps=%d blasfeo_dgetrf_rp", ps);
		check(&d, w);
		if (memcmp(gp, hp, sizeof gp))
		{
			printf("FAIL ps=%d blasfeo_dgetrf_rp pivots differ\n", ps);
			g_fail = 1;
		}
		else
			printf("ok ps=%d blasfeo_dgetrf_rp pivots\n", ps);
		pair_free(&c); pair_free(&d);
		free(A);
	}

	/* one batched entry: three dpotrf_l calls from device-resident
	 * descriptors and records, then the descriptors copied back */
	{
		const int ns[3] = {20, 90, 150};
		struct pair c[3], d[3];
		struct blasfeo_dmat hd[6];
		struct blasfeo_amd_dfact_arg rec[3];
		struct blasfeo_dmat *dd = NULL;
		struct blasfeo_amd_dfact_arg *dr = NULL;
		for (int i = 0; i < 3; i++)
		{
			double *S = spd(ns[i]);
			pair_new(&c[i], ns[i], ns[i], S);
			pair_new(&d[i], ns[i] + 3, ns[i] + 1, NULL);
			free(S);
			hd[2 * i] = c[i].g;
			hd[2 * i + 1] = d[i].g;
		}
		if (hipMalloc((void **)&dd, sizeof hd) != hipSuccess || hipMalloc((void **)&dr, sizeof rec) != hipSuccess)
		{
			printf("FAIL ps=%d batched: device allocation\n", ps);
			g_fail = 1;
			return;
		}
		for (int i = 0; i < 3; i++)
		{
			rec[i].m = rec[i].n = ns[i];
			rec[i].ci = rec[i].cj = 0;
			rec[i].di = 3;
			rec[i].dj = 1;
			rec[i].sC = dd + 2 * i;
			rec[i].sD = dd + 2 * i + 1;
			rec[i].ipiv = NULL;
		}
		hipMemcpy(dd, hd, sizeof hd, hipMemcpyHostToDevice);
		hipMemcpy(dr, rec, sizeof rec, hipMemcpyHostToDevice);
		struct blasfeo_amd_hint h = {150, 150, 0};
		int rc = blasfeo_amd_dpotrf_l_batched(ps, 3, dr, &h, NULL);
		if (rc != BLASFEO_AMD_OK || hipDeviceSynchronize() != hipSuccess)
		{
			printf("FAIL ps=%d blasfeo_amd_dpotrf_l_batched: %d %s\n", ps, rc, blasfeo_amd_last_error());
			g_fail = 1;
		}
		hipMemcpy(hd, dd, sizeof hd, hipMemcpyDeviceToHost);
		for (int i = 0; i < 3; i++)
		{
			d[i].g.use_dA = hd[2 * i + 1].use_dA;
			or_dpotrf_l(ps, ns[i], &c[i].h, 0, 0, &d[i].h, 3, 1);
			snprintf(w, sizeof w, "ps=%d blasfeo_amd_dpotrf_l_batched instance %d (n=%d at (3,1))", ps, i, ns[i]);
			check(&d[i], w);
			pair_free(&c[i]);
			pair_free(&d[i]);
		}
		hipFree(dd);
		hipFree(dr);
	}
}

int main(void)
{
	printf("%s\n", blasfeo_amd_version());
	run(4);
	run(8);
	blasfeo_amd_set_ps(4);
	printf(g_fail ? "FAILED\n" : "ALL OK\n");
	return g_fail;
}
