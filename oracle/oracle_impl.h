/*
 * oracle_impl.h -- body of the CPU oracle, instantiated once for double and
 * once for float by oracle.c (REAL, MAT, PFX are defined there).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the GPU library is compared
 * against; nothing in the shipped library links or calls it.
 *
 * Each routine restates the arithmetic of the reference's REFERENCE backend
 * (blasfeo_ref/x_blas3_ref.c, blasfeo_ref/x_lapack_ref.c) on panel-major
 * storage with a runtime panel size: every output element is accumulated
 * over k in ascending order exactly as the reference's 2x2-unrolled loops do,
 * reciprocals are formed as 1/x and applied by multiplication, and the
 * zero-pivot conventions are the reference's.  The dA / use_dA side effects
 * follow the HIGH_PERFORMANCE drivers (blasfeo_hp_pm/d_lapack_lib4.c:947-950,
 * :1185, :1882, :2421; blasfeo_hp_pm/d_blas3_lib4.c:4162-4178), which are the
 * ones the drop-in replaces.
 */

#define EL(sM, i, j) ((sM)->pA[((i) - ((i) % ps)) * (sM)->cn + (j) * ps + ((i) % ps)])

/* D <= beta*C + alpha*A*B^T ; blasfeo_ref/x_blas3_ref.c:142-235 */
void PFX(gemm_nt)(int ps, int m, int n, int k, REAL alpha, MAT *sA, int ai, int aj, MAT *sB, int bi, int bj, REAL beta, MAT *sC, int ci, int cj, MAT *sD, int di, int dj)
	{
	if(m <= 0 || n <= 0)
		return;
	sD->use_dA = 0;
	for(int j = 0; j < n; j++)
		for(int i = 0; i < m; i++)
			{
			REAL c = 0;
			for(int l = 0; l < k; l++)
				c += EL(sA, ai+i, aj+l) * EL(sB, bi+j, bj+l);
			EL(sD, di+i, dj+j) = alpha * c + beta * EL(sC, ci+i, cj+j);
			}
	}

/* D <= beta*C + alpha*A*B ; blasfeo_ref/x_blas3_ref.c:41-135 */
void PFX(gemm_nn)(int ps, int m, int n, int k, REAL alpha, MAT *sA, int ai, int aj, MAT *sB, int bi, int bj, REAL beta, MAT *sC, int ci, int cj, MAT *sD, int di, int dj)
	{
	if(m <= 0 || n <= 0)
		return;
	sD->use_dA = 0;
	for(int j = 0; j < n; j++)
		for(int i = 0; i < m; i++)
			{
			REAL c = 0;
			for(int l = 0; l < k; l++)
				c += EL(sA, ai+i, aj+l) * EL(sB, bi+l, bj+j);
			EL(sD, di+i, dj+j) = alpha * c + beta * EL(sC, ci+i, cj+j);
			}
	}

/* lower trapezoid of D <= beta*C + alpha*A*B^T, n columns of an m x n block
 * (n == m is blasfeo_ref/x_blas3_ref.c:4130-4222 syrk_ln, otherwise syrk_ln_mn :4229) */
void PFX(syrk_ln_mn)(int ps, int m, int n, int k, REAL alpha, MAT *sA, int ai, int aj, MAT *sB, int bi, int bj, REAL beta, MAT *sC, int ci, int cj, MAT *sD, int di, int dj)
	{
	if(m <= 0 || n <= 0)
		return;
	sD->use_dA = 0;
	for(int j = 0; j < n; j++)
		for(int i = j; i < m; i++)
			{
			REAL c = 0;
			for(int l = 0; l < k; l++)
				c += EL(sA, ai+i, aj+l) * EL(sB, bi+j, bj+l);
			EL(sD, di+i, dj+j) = beta * EL(sC, ci+i, cj+j) + alpha * c;
			}
	}

void PFX(syrk_ln)(int ps, int m, int k, REAL alpha, MAT *sA, int ai, int aj, MAT *sB, int bi, int bj, REAL beta, MAT *sC, int ci, int cj, MAT *sD, int di, int dj)
	{
	PFX(syrk_ln_mn)(ps, m, m, k, alpha, sA, ai, aj, sB, bi, bj, beta, sC, ci, cj, sD, di, dj);
	}

/* D <= alpha * B * A^{-T}, A lower n x n with reciprocal diagonal cached in
 * sA->dA (blasfeo_ref/x_blas3_ref.c:1642-1752; cache rule of
 * blasfeo_hp_pm/d_blas3_lib4.c:4162-4178) */
void PFX(trsm_rltn)(int ps, int m, int n, REAL alpha, MAT *sA, int ai, int aj, MAT *sB, int bi, int bj, MAT *sD, int di, int dj)
	{
	if(m <= 0 || n <= 0)
		return;
	sD->use_dA = 0;
	REAL *dA = sA->dA;
	if(ai == 0 && aj == 0)
		{
		if(sA->use_dA < n)
			{
			for(int i = 0; i < n; i++)
				dA[i] = (REAL)1.0 / EL(sA, i, i);
			sA->use_dA = n;
			}
		}
	else
		{
		for(int i = 0; i < n; i++)
			dA[i] = (REAL)1.0 / EL(sA, ai+i, aj+i);
		sA->use_dA = 0;
		}
	for(int j = 0; j < n; j++)
		{
		REAL inv = dA[j];
		for(int i = 0; i < m; i++)
			{
			REAL c = alpha * EL(sB, bi+i, bj+j);
			for(int l = 0; l < j; l++)
				c -= EL(sD, di+i, dj+l) * EL(sA, ai+j, aj+l);
			EL(sD, di+i, dj+j) = c * inv;
			}
		}
	}

/* dA / use_dA handling shared by the left-side solves (blasfeo_ref/x_blas3_ref.c:472-486, 701-715) */
static REAL *PFX(trsm_left_diag)(int ps, int m, MAT *sA, int ai, int aj, MAT *sD)
	{
	sD->use_dA = 0;
	REAL *dA = sA->dA;
	if(ai == 0 && aj == 0)
		{
		if(sA->use_dA < m)
			{
			for(int i = 0; i < m; i++)
				dA[i] = (REAL)1.0 / EL(sA, i, i);
			sA->use_dA = m;
			}
		}
	else
		{
		for(int i = 0; i < m; i++)
			dA[i] = (REAL)1.0 / EL(sA, ai+i, aj+i);
		sA->use_dA = 0;
		}
	return dA;
	}

/* D <= alpha * A^{-1} * B, A lower m x m ; blasfeo_ref/x_blas3_ref.c:445-673.
 * Every element accumulates its k terms in ascending order (the 2x2 blocks of
 * the reference subtract the in-block term last, which is also ascending). */
void PFX(trsm_llnn)(int ps, int m, int n, REAL alpha, MAT *sA, int ai, int aj, MAT *sB, int bi, int bj, MAT *sD, int di, int dj)
	{
	if(m <= 0 || n <= 0)
		return;
	REAL *dA = PFX(trsm_left_diag)(ps, m, sA, ai, aj, sD);
	for(int j = 0; j < n; j++)
		for(int i = 0; i < m; i++)
			{
			REAL c = alpha * EL(sB, bi+i, bj+j);
			for(int l = 0; l < i; l++)
				c -= EL(sA, ai+i, aj+l) * EL(sD, di+l, dj+j);
			EL(sD, di+i, dj+j) = c * dA[i];
			}
	}

/* D <= alpha * A^{-T} * B, A lower m x m ; blasfeo_ref/x_blas3_ref.c:674-900.
 * Rows are solved bottom-up in pairs (id, id+1), id = m-2, m-4, ...: the
 * upper row of a pair accumulates k = id+2 .. m-1 ascending and subtracts the
 * in-pair term k = id+1 last; the lower row of a pair, and the single row 0
 * left over when m is odd, accumulate k ascending from the row below. */
void PFX(trsm_lltn)(int ps, int m, int n, REAL alpha, MAT *sA, int ai, int aj, MAT *sB, int bi, int bj, MAT *sD, int di, int dj)
	{
	if(m <= 0 || n <= 0)
		return;
	REAL *dA = PFX(trsm_left_diag)(ps, m, sA, ai, aj, sD);
	for(int j = 0; j < n; j++)
		for(int i = m - 1; i >= 0; i--)
			{
			REAL c = alpha * EL(sB, bi+i, bj+j);
			const int pair_top = ((m - i) % 2 == 0);
			for(int l = pair_top ? i + 2 : i + 1; l < m; l++)
				c -= EL(sA, ai+l, aj+i) * EL(sD, di+l, dj+j);
			if(pair_top)
				c -= EL(sA, ai+i+1, aj+i) * EL(sD, di+i+1, dj+j);
			EL(sD, di+i, dj+j) = c * dA[i];
			}
	}

/* D <= alpha * A^{-1} * B, A unit lower m x m (diagonal not read), sD->use_dA = 0 and
 * sA untouched; blasfeo_ref/x_blas3_ref.c:572-672 (k ascending; the in-pair term of a
 * row pair is the last, which is also ascending) */
void PFX(trsm_llnu)(int ps, int m, int n, REAL alpha, MAT *sA, int ai, int aj, MAT *sB, int bi, int bj, MAT *sD, int di, int dj)
	{
	if(m <= 0 || n <= 0)
		return;
	sD->use_dA = 0;
	for(int j = 0; j < n; j++)
		for(int i = 0; i < m; i++)
			{
			REAL c = alpha * EL(sB, bi+i, bj+j);
			for(int l = 0; l < i; l++)
				c -= EL(sA, ai+i, aj+l) * EL(sD, di+l, dj+j);
			EL(sD, di+i, dj+j) = c;
			}
	}

/* D <= alpha * U^{-1} * B, U the upper triangle of A (m x m, non-unit);
 * blasfeo_ref/x_blas3_ref.c:920-1100: the dA rule and the pair order of
 * trsm_lltn with A(k, i) read as A(i, k) */
void PFX(trsm_lunn)(int ps, int m, int n, REAL alpha, MAT *sA, int ai, int aj, MAT *sB, int bi, int bj, MAT *sD, int di, int dj)
	{
	if(m <= 0 || n <= 0)
		return;
	REAL *dA = PFX(trsm_left_diag)(ps, m, sA, ai, aj, sD);
	for(int j = 0; j < n; j++)
		for(int i = m - 1; i >= 0; i--)
			{
			REAL c = alpha * EL(sB, bi+i, bj+j);
			const int pair_top = ((m - i) % 2 == 0);
			for(int l = pair_top ? i + 2 : i + 1; l < m; l++)
				c -= EL(sA, ai+i, aj+l) * EL(sD, di+l, dj+j);
			if(pair_top)
				c -= EL(sA, ai+i, aj+i+1) * EL(sD, di+i+1, dj+j);
			EL(sD, di+i, dj+j) = c * dA[i];
			}
	}

/* blasfeo_drowpe: rows ii and ipiv[ii] swapped over all columns, ii ascending;
 * use_dA = 0; auxiliary/d_aux_lib4.c:2767-2780 */
void PFX(rowpe)(int ps, int kmax, int *ipiv, MAT *sA)
	{
	sA->use_dA = 0;
	for(int ii = 0; ii < kmax; ii++)
		if(ipiv[ii] != ii)
			for(int j = 0; j < sA->n; j++)
				{
				REAL t = EL(sA, ii, j);
				EL(sA, ii, j) = EL(sA, ipiv[ii], j);
				EL(sA, ipiv[ii], j) = t;
				}
	}

/* D <= alpha * B * A, A lower n x n ; blasfeo_ref/x_blas3_ref.c:3218-3330 */
void PFX(trmm_rlnn)(int ps, int m, int n, REAL alpha, MAT *sA, int ai, int aj, MAT *sB, int bi, int bj, MAT *sD, int di, int dj)
	{
	if(m <= 0 || n <= 0)
		return;
	sD->use_dA = 0;
	for(int j = 0; j < n; j++)
		for(int i = 0; i < m; i++)
			{
			REAL c = 0;
			for(int l = j; l < n; l++)
				c += EL(sB, bi+i, bj+l) * EL(sA, ai+l, aj+j);
			EL(sD, di+i, dj+j) = alpha * c;
			}
	}

/* one left-looking Cholesky column sweep shared by potrf_l, potrf_l_mn and
 * syrk_dpotrf_ln(_mn): column j of D <= (C + A*B^T)(.,j) - sum_{l<j} D(.,l)D(j,l),
 * then scaled by 1/sqrt(pivot) (zero if pivot <= 0 or NaN).
 * blasfeo_ref/x_lapack_ref.c:41-176, :185-326, :399-566 */
static void PFX(chol_core)(int ps, int m, int n, int k, MAT *sA, int ai, int aj, MAT *sB, int bi, int bj, MAT *sC, int ci, int cj, MAT *sD, int di, int dj)
	{
	REAL *dD = sD->dA;
	for(int j = 0; j < n; j++)
		{
		REAL c = EL(sC, ci+j, cj+j);
		for(int l = 0; l < k; l++)
			c += EL(sA, ai+j, aj+l) * EL(sB, bi+j, bj+l);
		for(int l = 0; l < j; l++)
			c -= EL(sD, di+j, dj+l) * EL(sD, di+j, dj+l);
		REAL inv = c > 0 ? (REAL)1.0 / SQRT(c) : (REAL)0.0;
		dD[j] = inv;
		EL(sD, di+j, dj+j) = c * inv;
		for(int i = j+1; i < m; i++)
			{
			REAL d = EL(sC, ci+i, cj+j);
			for(int l = 0; l < k; l++)
				d += EL(sA, ai+i, aj+l) * EL(sB, bi+j, bj+l);
			for(int l = 0; l < j; l++)
				d -= EL(sD, di+i, dj+l) * EL(sD, di+j, dj+l);
			EL(sD, di+i, dj+j) = d * inv;
			}
		}
	}

/* blasfeo_ref/x_lapack_ref.c:41-176; use_dA rule blasfeo_hp_pm/d_lapack_lib4.c:947-950 */
void PFX(potrf_l)(int ps, int m, MAT *sC, int ci, int cj, MAT *sD, int di, int dj)
	{
	if(m <= 0)
		return;
	sD->use_dA = (di == 0 && dj == 0) ? m : 0;
	PFX(chol_core)(ps, m, m, 0, sD, 0, 0, sD, 0, 0, sC, ci, cj, sD, di, dj);
	}

/* blasfeo_ref/x_lapack_ref.c:185-326; use_dA rule blasfeo_hp_pm/d_lapack_lib4.c:1185 */
void PFX(potrf_l_mn)(int ps, int m, int n, MAT *sC, int ci, int cj, MAT *sD, int di, int dj)
	{
	if(m <= 0 || n <= 0)
		return;
	if(n > m)
		n = m;
	sD->use_dA = (di == 0 && dj == 0) ? 1 : 0;
	PFX(chol_core)(ps, m, n, 0, sD, 0, 0, sD, 0, 0, sC, ci, cj, sD, di, dj);
	}

/* D <= chol(C + A*B^T) on the m x n lower trapezoid;
 * blasfeo_ref/x_lapack_ref.c:399-566 (n == m) and :570-738;
 * use_dA rule blasfeo_hp_pm/d_lapack_lib4.c:1545,1882 */
void PFX(syrk_potrf_ln_mn)(int ps, int m, int n, int k, MAT *sA, int ai, int aj, MAT *sB, int bi, int bj, MAT *sC, int ci, int cj, MAT *sD, int di, int dj)
	{
	if(m <= 0 || n <= 0)
		return;
	if(n > m)
		n = m;
	sD->use_dA = (di == 0 && dj == 0) ? 1 : 0;
	PFX(chol_core)(ps, m, n, k, sA, ai, aj, sB, bi, bj, sC, ci, cj, sD, di, dj);
	}

void PFX(syrk_potrf_ln)(int ps, int m, int k, MAT *sA, int ai, int aj, MAT *sB, int bi, int bj, MAT *sC, int ci, int cj, MAT *sD, int di, int dj)
	{
	PFX(syrk_potrf_ln_mn)(ps, m, m, k, sA, ai, aj, sB, bi, bj, sC, ci, cj, sD, di, dj);
	}

/* LU with partial row pivoting, left-looking (Crout) column sweep of
 * blasfeo_ref/x_lapack_ref.c:1156-1536: for each column j the upper part is
 * corrected row by row, the lower part is corrected, the pivot is the first
 * row of largest |.| (strict '>' from dmax = 0, so an all-zero or NaN column
 * keeps row j), the whole rows j and ip are swapped, and the sub-diagonal is
 * scaled by the reciprocal pivot (0 if the pivot is 0), which is stored in dA.
 * ipiv[j] is the 0-based absolute row swapped with row j.  Columns j >= m (m < n)
 * receive only the upper correction. */
void PFX(getrf_rp)(int ps, int m, int n, MAT *sC, int ci, int cj, MAT *sD, int di, int dj, int *ipiv)
	{
	sD->use_dA = (di == 0 && dj == 0) ? 1 : 0;
	if(m <= 0 || n <= 0)
		return;
	REAL *dD = sD->dA;
	if(&EL(sC, ci, cj) != &EL(sD, di, dj))
		for(int j = 0; j < n; j++)
			for(int i = 0; i < m; i++)
				EL(sD, di+i, dj+j) = EL(sC, ci+i, cj+j);
	int p = m < n ? m : n;
	for(int j = 0; j < n; j++)
		{
		int iu = j < m ? j : m;
		for(int i = 0; i < iu; i++)
			{
			REAL d = EL(sD, di+i, dj+j);
			for(int l = 0; l < i; l++)
				d -= EL(sD, di+i, dj+l) * EL(sD, di+l, dj+j);
			EL(sD, di+i, dj+j) = d;
			}
		if(j >= p)
			continue;
		REAL dmax = 0;
		int ip = j;
		for(int i = j; i < m; i++)
			{
			REAL d = EL(sD, di+i, dj+j);
			for(int l = 0; l < j; l++)
				d -= EL(sD, di+i, dj+l) * EL(sD, di+l, dj+j);
			EL(sD, di+i, dj+j) = d;
			REAL a = d > 0 ? d : -d;
			if(a > dmax)
				{
				dmax = a;
				ip = i;
				}
			}
		ipiv[j] = ip;
		if(ip != j)
			for(int l = 0; l < n; l++)
				{
				REAL t = EL(sD, di+j, dj+l);
				EL(sD, di+j, dj+l) = EL(sD, di+ip, dj+l);
				EL(sD, di+ip, dj+l) = t;
				}
		REAL piv = EL(sD, di+j, dj+j);
		REAL inv = piv != 0 ? (REAL)1.0 / piv : (REAL)0.0;
		dD[j] = inv;
		for(int i = j+1; i < m; i++)
			EL(sD, di+i, dj+j) *= inv;
		}
	}

/* LU without pivoting, blasfeo_ref/x_lapack_ref.c:880-1150: column by column,
 * U(i,j) = C(i,j) - sum_{k<i} L(i,k) U(k,j) for i <= j, then
 * L(i,j) = (C(i,j) - sum_{k<j} L(i,k) U(k,j)) * (1/U(j,j)) for i > j, every sum in
 * ascending k (the reference's 2x2 blocks subtract the in-block term last, which
 * is also ascending); dA = 1/U(j,j) with no zero guard; use_dA = 1 at offset (0,0) */
void PFX(getrf_np)(int ps, int m, int n, MAT *sC, int ci, int cj, MAT *sD, int di, int dj)
	{
	sD->use_dA = (di == 0 && dj == 0) ? 1 : 0;
	if(m <= 0 || n <= 0)
		return;
	REAL *dD = sD->dA;
	for(int j = 0; j < n; j++)
		{
		for(int i = 0; i <= j && i < m; i++)
			{
			REAL c = EL(sC, ci+i, cj+j);
			for(int k = 0; k < i; k++)
				c -= EL(sD, di+i, dj+k) * EL(sD, di+k, dj+j);
			EL(sD, di+i, dj+j) = c;
			}
		if(j >= m)
			continue;
		REAL inv = (REAL)1.0 / EL(sD, di+j, dj+j);
		dD[j] = inv;
		for(int i = j+1; i < m; i++)
			{
			REAL c = EL(sC, ci+i, cj+j);
			for(int k = 0; k < j; k++)
				c -= EL(sD, di+i, dj+k) * EL(sD, di+k, dj+j);
			EL(sD, di+i, dj+j) = c * inv;
			}
		}
	}

/* backward Riccati factorisation d_back_ric_trf_libstr,
 * examples/example_d_riccati_recursion.c:115-140 (fused syrk_dpotrf branch) */
void PFX(ric_trf)(int ps, int N, const int *nx, const int *nu, MAT *BAbt, MAT *RSQrq, MAT *L, MAT *work)
	{
	PFX(potrf_l)(ps, nx[N], &RSQrq[N], 0, 0, &L[N], 0, 0);
	for(int s = 0; s < N; s++)
		{
		int k = N - s - 1;
		int nk = nu[k] + nx[k];
		PFX(trmm_rlnn)(ps, nk, nx[k+1], (REAL)1.0, &L[k+1], nu[k+1], nu[k+1], &BAbt[k], 0, 0, work, 0, 0);
		PFX(syrk_potrf_ln)(ps, nk, nx[k+1], work, 0, 0, work, 0, 0, &RSQrq[k], 0, 0, &L[k], 0, 0);
		}
	}

/* ---- column-major <-> panel-major (auxiliary/d_aux_lib4.c) ---- */

/* sA(ai+i, aj+j) = A[i + j*lda]; use_dA = 0; auxiliary/d_aux_lib4.c:1607-1640 */
void PFX(pack)(int ps, int m, int n, REAL *A, int lda, MAT *sA, int ai, int aj)
	{
	if(m <= 0 || n <= 0)
		return;
	sA->use_dA = 0;
	for(int j = 0; j < n; j++)
		for(int i = 0; i < m; i++)
			EL(sA, ai+i, aj+j) = A[i + (size_t)j*lda];
	}

/* sA(ai+i, aj+j) = A[j + i*lda], i < n, j < m (the transpose of the m x n A);
 * use_dA = 0; auxiliary/d_aux_lib4.c:2086-2125 */
void PFX(pack_tran)(int ps, int m, int n, REAL *A, int lda, MAT *sA, int ai, int aj)
	{
	sA->use_dA = 0;
	for(int i = 0; i < n; i++)
		for(int j = 0; j < m; j++)
			EL(sA, ai+i, aj+j) = A[j + (size_t)i*lda];
	}

/* A[i + j*lda] = sA(ai+i, aj+j); auxiliary/d_aux_lib4.c:2238-2270 */
void PFX(unpack)(int ps, int m, int n, MAT *sA, int ai, int aj, REAL *A, int lda)
	{
	for(int j = 0; j < n; j++)
		for(int i = 0; i < m; i++)
			A[i + (size_t)j*lda] = EL(sA, ai+i, aj+j);
	}

/* ---- level 2 (blasfeo_ref/x_blas2_ref.c), used by the Riccati solve ----
 * Rows (gemv_n, trmv, trsv) or columns (gemv_t) are taken in pairs; inside a
 * pair each partial sum adds two products at a time (a*x0 + b*x1, rounded,
 * then accumulated), and an odd last row or column accumulates one product
 * at a time, as the reference's loops do. */

/* z <= beta*y + alpha*A*x, A m x n; blasfeo_ref/x_blas2_ref.c:39-98 */
void PFX(gemv_n)(int ps, int m, int n, REAL alpha, MAT *sA, int ai, int aj, VEC *sx, int xi, REAL beta, VEC *sy, int yi, VEC *sz, int zi)
	{
	REAL *x = sx->pa + xi, *y = sy->pa + yi, *z = sz->pa + zi;
	for(int i = 0; i < m; i++)
		z[i] = beta == 0 ? (REAL)0.0 : beta * y[i];
	int i = 0;
	for(; i < m-1; i += 2)
		{
		REAL y0 = 0, y1 = 0;
		int j = 0;
		for(; j < n-1; j += 2)
			{
			y0 += EL(sA, ai+i, aj+j) * x[j] + EL(sA, ai+i, aj+j+1) * x[j+1];
			y1 += EL(sA, ai+i+1, aj+j) * x[j] + EL(sA, ai+i+1, aj+j+1) * x[j+1];
			}
		if(j < n)
			{
			y0 += EL(sA, ai+i, aj+j) * x[j];
			y1 += EL(sA, ai+i+1, aj+j) * x[j];
			}
		z[i] += alpha * y0;
		z[i+1] += alpha * y1;
		}
	for(; i < m; i++)
		{
		REAL y0 = 0;
		for(int j = 0; j < n; j++)
			y0 += EL(sA, ai+i, aj+j) * x[j];
		z[i] += alpha * y0;
		}
	}

/* z <= beta*y + alpha*A^T*x, A m x n; blasfeo_ref/x_blas2_ref.c:146-205 */
void PFX(gemv_t)(int ps, int m, int n, REAL alpha, MAT *sA, int ai, int aj, VEC *sx, int xi, REAL beta, VEC *sy, int yi, VEC *sz, int zi)
	{
	REAL *x = sx->pa + xi, *y = sy->pa + yi, *z = sz->pa + zi;
	for(int j = 0; j < n; j++)
		z[j] = beta == 0 ? (REAL)0.0 : beta * y[j];
	int j = 0;
	for(; j < n-1; j += 2)
		{
		REAL y0 = 0, y1 = 0;
		int i = 0;
		for(; i < m-1; i += 2)
			{
			y0 += EL(sA, ai+i, aj+j) * x[i] + EL(sA, ai+i+1, aj+j) * x[i+1];
			y1 += EL(sA, ai+i, aj+j+1) * x[i] + EL(sA, ai+i+1, aj+j+1) * x[i+1];
			}
		if(i < m)
			{
			y0 += EL(sA, ai+i, aj+j) * x[i];
			y1 += EL(sA, ai+i, aj+j+1) * x[i];
			}
		z[j] += alpha * y0;
		z[j+1] += alpha * y1;
		}
	for(; j < n; j++)
		{
		REAL y0 = 0;
		for(int i = 0; i < m; i++)
			y0 += EL(sA, ai+i, aj+j) * x[i];
		z[j] += alpha * y0;
		}
	}

/* z <= A*x, A lower m x m, non-unit; bottom-up so z may alias x;
 * blasfeo_ref/x_blas2_ref.c:470-530 (the _mn form with n = m) */
void PFX(trmv_lnn)(int ps, int m, MAT *sA, int ai, int aj, VEC *sx, int xi, VEC *sz, int zi)
	{
	REAL *x = sx->pa + xi, *z = sz->pa + zi;
	int n = m;
	if(n % 2 != 0)
		{
		int i = n-1;
		REAL y0 = x[i];
		y0 *= EL(sA, ai+i, aj+i);
		for(int j = 0; j < i; j++)
			y0 += EL(sA, ai+i, aj+j) * x[j];
		z[i] = y0;
		n -= 1;
		}
	for(int i = n-2; i >= 0; i -= 2)
		{
		REAL y0 = x[i], y1 = x[i+1];
		y1 *= EL(sA, ai+i+1, aj+i+1);
		y1 += EL(sA, ai+i+1, aj+i) * y0;
		y0 *= EL(sA, ai+i, aj+i);
		for(int j = 0; j < i-1; j += 2)
			{
			y0 += EL(sA, ai+i, aj+j) * x[j] + EL(sA, ai+i, aj+j+1) * x[j+1];
			y1 += EL(sA, ai+i+1, aj+j) * x[j] + EL(sA, ai+i+1, aj+j+1) * x[j+1];
			}
		z[i] = y0;
		z[i+1] = y1;
		}
	}

/* z <= A^T*x, A lower m x m, non-unit; top-down so z may alias x;
 * blasfeo_ref/x_blas2_ref.c:535-590 (the _mn form with n = m) */
void PFX(trmv_ltn)(int ps, int m, MAT *sA, int ai, int aj, VEC *sx, int xi, VEC *sz, int zi)
	{
	REAL *x = sx->pa + xi, *z = sz->pa + zi;
	int j = 0;
	for(; j < m-1; j += 2)
		{
		REAL y0 = x[j], y1 = x[j+1];
		y0 *= EL(sA, ai+j, aj+j);
		y0 += EL(sA, ai+j+1, aj+j) * y1;
		y1 *= EL(sA, ai+j+1, aj+j+1);
		int i = j+2;
		for(; i < m-1; i += 2)
			{
			y0 += EL(sA, ai+i, aj+j) * x[i] + EL(sA, ai+i+1, aj+j) * x[i+1];
			y1 += EL(sA, ai+i, aj+j+1) * x[i] + EL(sA, ai+i+1, aj+j+1) * x[i+1];
			}
		for(; i < m; i++)
			{
			y0 += EL(sA, ai+i, aj+j) * x[i];
			y1 += EL(sA, ai+i, aj+j+1) * x[i];
			}
		z[j] = y0;
		z[j+1] = y1;
		}
	for(; j < m; j++)
		{
		REAL y0 = x[j];
		y0 *= EL(sA, ai+j, aj+j);
		for(int i = j+1; i < m; i++)
			y0 += EL(sA, ai+i, aj+j) * x[i];
		z[j] = y0;
		}
	}

/* Reciprocal diagonal used by trsv: the cache when ai = aj = 0 and use_dA = 1,
 * else formed as 1/a_jj (stored; use_dA = 1 at offset 0, 0 elsewhere);
 * blasfeo_ref/x_blas2_ref.c:773-787, blasfeo_hp_pm/d_blas2_lib4.c:1300-1316 */
static REAL *PFX(trsv_diag)(int ps, int n, MAT *sA, int ai, int aj)
	{
	REAL *dA = sA->dA;
	if(ai == 0 && aj == 0)
		{
		if(sA->use_dA != 1)
			{
			for(int i = 0; i < n; i++)
				dA[i] = (REAL)1.0 / EL(sA, ai+i, aj+i);
			sA->use_dA = 1;
			}
		}
	else
		{
		for(int i = 0; i < n; i++)
			dA[i] = (REAL)1.0 / EL(sA, ai+i, aj+i);
		sA->use_dA = 0;
		}
	return dA;
	}

/* z <= solution of A(0:n,0:n) z = x(0:n), then z(n:m) = x(n:m) - A(n:m,0:n) z(0:n);
 * A lower m x n, non-unit; blasfeo_ref/x_blas2_ref.c:734-840 (trsv_lnn is n = m,
 * :990-1050, the same loops) */
void PFX(trsv_lnn_mn)(int ps, int m, int n, MAT *sA, int ai, int aj, VEC *sx, int xi, VEC *sz, int zi)
	{
	if(m == 0 || n == 0)
		return;
	REAL *dA = PFX(trsv_diag)(ps, n, sA, ai, aj);
	REAL *x = sx->pa + xi, *z = sz->pa + zi;
	int i = 0;
	for(; i < n-1; i += 2)
		{
		REAL y0 = x[i], y1 = x[i+1];
		int j = 0;
		for(; j < i-1; j += 2)
			{
			y0 -= EL(sA, ai+i, aj+j) * z[j] + EL(sA, ai+i, aj+j+1) * z[j+1];
			y1 -= EL(sA, ai+i+1, aj+j) * z[j] + EL(sA, ai+i+1, aj+j+1) * z[j+1];
			}
		y0 *= dA[i];
		y1 -= EL(sA, ai+i+1, aj+j) * y0;
		y1 *= dA[i+1];
		z[i] = y0;
		z[i+1] = y1;
		}
	for(; i < n; i++)
		{
		REAL y0 = x[i];
		for(int j = 0; j < i; j++)
			y0 -= EL(sA, ai+i, aj+j) * z[j];
		y0 *= dA[i];
		z[i] = y0;
		}
	for(; i < m-1; i += 2)
		{
		REAL y0 = x[i], y1 = x[i+1];
		int j = 0;
		for(; j < n-1; j += 2)
			{
			y0 -= EL(sA, ai+i, aj+j) * z[j] + EL(sA, ai+i, aj+j+1) * z[j+1];
			y1 -= EL(sA, ai+i+1, aj+j) * z[j] + EL(sA, ai+i+1, aj+j+1) * z[j+1];
			}
		if(j < n)
			{
			y0 -= EL(sA, ai+i, aj+j) * z[j];
			y1 -= EL(sA, ai+i+1, aj+j) * z[j];
			}
		z[i] = y0;
		z[i+1] = y1;
		}
	for(; i < m; i++)
		{
		REAL y0 = x[i];
		for(int j = 0; j < n; j++)
			y0 -= EL(sA, ai+i, aj+j) * z[j];
		z[i] = y0;
		}
	}

/* z(0:n) <= solution of A(0:n,0:n)^T z(0:n) = x(0:n) - A(n:m,0:n)^T x(n:m),
 * z(n:m) = x(n:m); A lower m x n, non-unit; blasfeo_ref/x_blas2_ref.c:888-985
 * (trsv_ltn is n = m, :1115-1175, the same arithmetic) */
void PFX(trsv_ltn_mn)(int ps, int m, int n, MAT *sA, int ai, int aj, VEC *sx, int xi, VEC *sz, int zi)
	{
	if(m == 0)
		return;
	REAL *dA = PFX(trsv_diag)(ps, n, sA, ai, aj);
	REAL *x = sx->pa + xi, *z = sz->pa + zi;
	if(x != z)
		for(int i = 0; i < m; i++)
			z[i] = x[i];
	int j;
	if(n % 2 != 0)
		{
		j = n-1;
		REAL y0 = z[j];
		for(int i = j+1; i < m; i++)
			y0 -= EL(sA, ai+i, aj+j) * z[i];
		y0 *= dA[j];
		z[j] = y0;
		j -= 2;
		}
	else
		j = n-2;
	for(; j >= 0; j -= 2)
		{
		REAL y0 = z[j], y1 = z[j+1];
		int i = j+2;
		for(; i < m-1; i += 2)
			{
			y0 -= EL(sA, ai+i, aj+j) * z[i] + EL(sA, ai+i+1, aj+j) * z[i+1];
			y1 -= EL(sA, ai+i, aj+j+1) * z[i] + EL(sA, ai+i+1, aj+j+1) * z[i+1];
			}
		if(i < m)
			{
			y0 -= EL(sA, ai+i, aj+j) * z[i];
			y1 -= EL(sA, ai+i, aj+j+1) * z[i];
			}
		y1 *= dA[j+1];
		y0 -= EL(sA, ai+j+1, aj+j) * y1;
		y0 *= dA[j];
		z[j] = y0;
		z[j+1] = y1;
		}
	}

/* blasfeo_dveccp, blasfeo_daxpy (z = alpha*x + y), blasfeo_dvecsc;
 * blasfeo_ref/x_blas1_ref.c */
static void PFX(veccp)(int m, VEC *sx, int xi, VEC *sz, int zi)
	{
	for(int i = 0; i < m; i++)
		sz->pa[zi+i] = sx->pa[xi+i];
	}
static void PFX(axpy)(int m, REAL alpha, VEC *sx, int xi, VEC *sy, int yi, VEC *sz, int zi)
	{
	for(int i = 0; i < m; i++)
		sz->pa[zi+i] = sy->pa[yi+i] + alpha * sx->pa[xi+i];
	}
static void PFX(vecsc)(int m, REAL alpha, VEC *sx, int xi)
	{
	for(int i = 0; i < m; i++)
		sx->pa[xi+i] *= alpha;
	}

/* Riccati solve d_back_ric_trs_libstr, examples/example_d_riccati_recursion.c:203-267:
 * backward substitution (Pb, ux), then forward substitution (ux, pi).  The
 * stage-0 steps of the reference's loops (full trsv on L[0], all of ux[0]
 * negated) are the k = 0 cases here. */
void PFX(ric_trs)(int ps, int N, const int *nx, const int *nu, MAT *BAbt, VEC *b, VEC *rq, MAT *L, VEC *Pb, VEC *ux, VEC *pi, VEC *work)
	{
	if(N < 0)
		return;
	PFX(veccp)(nu[N]+nx[N], &rq[N], 0, &ux[N], 0);
	for(int k = N-1; k >= 0; k--)
		{
		int nk = nu[k] + nx[k];
		PFX(trmv_ltn)(ps, nx[k+1], &L[k+1], nu[k+1], nu[k+1], &b[k], 0, &Pb[k], 0);
		PFX(trmv_lnn)(ps, nx[k+1], &L[k+1], nu[k+1], nu[k+1], &Pb[k], 0, &Pb[k], 0);
		PFX(veccp)(nk, &rq[k], 0, &ux[k], 0);
		PFX(veccp)(nx[k+1], &Pb[k], 0, work, 0);
		PFX(axpy)(nx[k+1], (REAL)1.0, &ux[k+1], nu[k+1], work, 0, work, 0);
		PFX(gemv_n)(ps, nk, nx[k+1], (REAL)1.0, &BAbt[k], 0, 0, work, 0, (REAL)1.0, &ux[k], 0, &ux[k], 0);
		PFX(trsv_lnn_mn)(ps, nk, k == 0 ? nk : nu[k], &L[k], 0, 0, &ux[k], 0, &ux[k], 0);
		}
	for(int k = 0; k < N; k++)
		{
		int nk = nu[k] + nx[k];
		PFX(veccp)(nx[k+1], &ux[k+1], nu[k+1], &pi[k], 0);
		PFX(vecsc)(k == 0 ? nk : nu[k], (REAL)-1.0, &ux[k], 0);
		PFX(trsv_ltn_mn)(ps, nk, k == 0 ? nk : nu[k], &L[k], 0, 0, &ux[k], 0, &ux[k], 0);
		PFX(gemv_t)(ps, nk, nx[k+1], (REAL)1.0, &BAbt[k], 0, 0, &ux[k], 0, (REAL)1.0, &b[k], 0, &ux[k+1], nu[k+1]);
		PFX(veccp)(nx[k+1], &ux[k+1], nu[k+1], work, 0);
		PFX(trmv_ltn)(ps, nx[k+1], &L[k+1], nu[k+1], nu[k+1], work, 0, work, 0);
		PFX(trmv_lnn)(ps, nx[k+1], &L[k+1], nu[k+1], nu[k+1], work, 0, work, 0);
		PFX(axpy)(nx[k+1], (REAL)1.0, work, 0, &pi[k], 0, &pi[k], 0);
		}
	}

/* d_back_ric_sv_libstr (examples/example_d_riccati_recursion.c:58-111): the
 * backward factorisation on matrices augmented by one row, which carries the
 * linear terms through the recursion (BAbt[k]: row nk = b[k]^T; RSQrq[k]: row
 * nk = [r; q]^T; L[k] the (nk+1) x nk trapezoid, nk = nu[k] + nx[k]), then the
 * forward substitution for ux[0..N] and pi[0..N-1].  work: (nk+1) x nx[k+1]
 * for every stage; wvec: nx[k+1]. */
void PFX(ric_sv)(int ps, int N, const int *nx, const int *nu, MAT *BAbt, MAT *RSQrq, MAT *L, VEC *ux, VEC *pi, MAT *work, VEC *wvec)
	{
	if(N < 0)
		return;
	PFX(potrf_l_mn)(ps, nx[N]+1, nx[N], &RSQrq[N], 0, 0, &L[N], 0, 0);
	for(int k = N-1; k >= 0; k--)
		{
		int nk = nu[k] + nx[k], nx1 = nx[k+1], nu1 = nu[k+1];
		PFX(trmm_rlnn)(ps, nk+1, nx1, (REAL)1.0, &L[k+1], nu1, nu1, &BAbt[k], 0, 0, work, 0, 0);
		/* blasfeo_dgead(1, nx1, 1.0, L[k+1], nu1+nx1, nu1, work, nk, 0) */
		for(int j = 0; j < nx1; j++)
			EL(work, nk, j) = EL(work, nk, j) + EL(&L[k+1], nu1+nx1, nu1+j);
		PFX(syrk_potrf_ln_mn)(ps, nk+1, nk, nx1, work, 0, 0, work, 0, 0, &RSQrq[k], 0, 0, &L[k], 0, 0);
		}
	for(int k = 0; k < N; k++)
		{
		int nk = nu[k] + nx[k], nx1 = nx[k+1], nu1 = nu[k+1];
		int ns = k == 0 ? nk : nu[k];
		/* blasfeo_drowex(ns, -1.0, L[k], nk, 0, ux[k], 0) */
		for(int j = 0; j < ns; j++)
			ux[k].pa[j] = -EL(&L[k], nk, j);
		PFX(trsv_ltn_mn)(ps, nk, ns, &L[k], 0, 0, &ux[k], 0, &ux[k], 0);
		/* blasfeo_drowex(nx1, 1.0, BAbt[k], nk, 0, ux[k+1], nu1) */
		for(int j = 0; j < nx1; j++)
			ux[k+1].pa[nu1+j] = EL(&BAbt[k], nk, j);
		PFX(gemv_t)(ps, nk, nx1, (REAL)1.0, &BAbt[k], 0, 0, &ux[k], 0, (REAL)1.0, &ux[k+1], nu1, &ux[k+1], nu1);
		PFX(veccp)(nx1, &ux[k+1], nu1, &pi[k], 0);
		/* blasfeo_drowex(nx1, 1.0, L[k+1], nu1+nx1, nu1, wvec, 0) */
		for(int j = 0; j < nx1; j++)
			wvec->pa[j] = EL(&L[k+1], nu1+nx1, nu1+j);
		PFX(trmv_ltn)(ps, nx1, &L[k+1], nu1, nu1, &pi[k], 0, &pi[k], 0);
		PFX(axpy)(nx1, (REAL)1.0, wvec, 0, &pi[k], 0, &pi[k], 0);
		PFX(trmv_lnn)(ps, nx1, &L[k+1], nu1, nu1, &pi[k], 0, &pi[k], 0);
		}
	}

#undef EL
