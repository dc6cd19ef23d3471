// Emulator build of bfa_lapack.hip (TEST INFRASTRUCTURE ONLY): the library's
// batched factorisation entries, plus a direct launch of the global-memory
// Cholesky kernel k_potrf for any size bound.
#include "bfa_lapack.hip"

extern "C" int emu_k_potrf(int ps, int batch, const blasfeo_amd_dfact_arg *args, int mmax)
{
	using namespace bfa;
	const size_t shm = sizeof(double) * PW * ((mmax + 15) / 16 * 16 + 1);
	if (ps == 4)
		hipLaunchKernelGGL((k_potrf<double, 4>), dim3(batch), dim3(NTL), shm, 0, args, mmax);
	else if (ps == 8)
		hipLaunchKernelGGL((k_potrf<double, 8>), dim3(batch), dim3(NTL), shm, 0, args, mmax);
	else
		return -1;
	return 0;
}
