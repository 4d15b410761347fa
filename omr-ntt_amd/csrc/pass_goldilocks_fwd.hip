// Stockham pass kernels for Goldilocks, roots whose 64th root is 2^39 (the ark-convention forward root).
#define FIELD_T Goldilocks
#define FIELD_U 39
#define FIELD_FN launch_pass_goldilocks_u39
#define FIELD_MAXLR 12
#include "pass_inst.inc"
