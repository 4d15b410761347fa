// Stockham pass kernels for Goldilocks (see engine.h).
#define FIELD_T Goldilocks
#define FIELD_FN launch_pass_goldilocks
#define FIELD_MAXLR 12
#include "pass_inst.inc"
