// Stockham pass kernels for BabyBear (see engine.h).
#define FIELD_T BabyBear
#define FIELD_U 0
#define FIELD_FN launch_pass_babybear
#define FIELD_MAXLR 13
#include "pass_inst.inc"
