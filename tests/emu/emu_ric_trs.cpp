// Emulator build of bfa_ric_trs.hip (TEST INFRASTRUCTURE ONLY).
#include "bfa_ric_trs.hip"
