// Stockham pass kernels for Goldilocks, any root (register-DFT constants from the root table).
#define FIELD_T Goldilocks
#define FIELD_U 0
#define FIELD_FN launch_pass_goldilocks_u0
#define FIELD_MAXLR 12
#include "pass_inst.inc"
