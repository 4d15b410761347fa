// Stockham pass kernels for Goldilocks, roots whose 64th root is 2^153 (its inverse).
#define FIELD_T Goldilocks
#define FIELD_U 153
#define FIELD_FN launch_pass_goldilocks_u153
#define FIELD_MAXLR 12
#include "pass_inst.inc"
