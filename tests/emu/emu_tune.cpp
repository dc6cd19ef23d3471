// Emulator build of bfa_tune.cpp (TEST INFRASTRUCTURE ONLY): the tuning knobs.
#include "bfa_tune.cpp"
