// Emulator build of bfa_blas3.hip (TEST INFRASTRUCTURE ONLY).
#include "bfa_blas3.hip"
