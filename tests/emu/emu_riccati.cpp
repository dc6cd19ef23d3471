// Emulator build of bfa_riccati.hip (TEST INFRASTRUCTURE ONLY).
#include "bfa_riccati.hip"
