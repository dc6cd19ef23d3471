"""Writes tests/golden/test_x_data_n12.npz: the reference's integer-valued
test matrices (tests/test_x.c:144-151: A[i] = i+1 with diag(A) += 1,
B[i] = 2(i+1), C[i] = 0.5(i+1), column-major) at n = 12, and the exact
results of gemm_nt, gemm_nn, syrk_ln (lower part, upper keeps the -1
sentinel of tests/test_x.c:354) with alpha = 1, beta = 1.5, and trmm_rlnn
with alpha = 1, computed in exact integer/rational arithmetic (Python ints
and fractions), independent of the oracle."""
import os
from fractions import Fraction

import numpy as np

n = 12
A = [[Fraction(0)] * n for _ in range(n)]
B = [[Fraction(0)] * n for _ in range(n)]
C = [[Fraction(0)] * n for _ in range(n)]
for idx in range(n * n):
    i, j = idx % n, idx // n
    A[i][j] = Fraction(idx + 1)
    B[i][j] = Fraction(2 * (idx + 1))
    C[i][j] = Fraction(idx + 1, 2)
for i in range(n):
    A[i][i] += 1
alpha, beta = Fraction(1), Fraction(3, 2)


def mat(f):
    return np.array([[float(f(i, j)) for j in range(n)] for i in range(n)])


gemm_nt = mat(lambda i, j: alpha * sum(A[i][l] * B[j][l] for l in range(n)) + beta * C[i][j])
gemm_nn = mat(lambda i, j: alpha * sum(A[i][l] * B[l][j] for l in range(n)) + beta * C[i][j])
syrk_ln = mat(lambda i, j: (alpha * sum(A[i][l] * B[j][l] for l in range(n)) + beta * C[i][j]) if i >= j else Fraction(-1))
# trmm_rlnn: D = alpha * B * A with A lower
trmm_rlnn = mat(lambda i, j: alpha * sum(B[i][l] * A[l][j] for l in range(j, n)))
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "test_x_data_n12.npz")
np.savez(out, A=mat(lambda i, j: A[i][j]), B=mat(lambda i, j: B[i][j]), C=mat(lambda i, j: C[i][j]),
         gemm_nt=gemm_nt, gemm_nn=gemm_nn, syrk_ln=syrk_ln, trmm_rlnn=trmm_rlnn)
print(out)
