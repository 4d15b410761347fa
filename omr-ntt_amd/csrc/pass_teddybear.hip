// Stockham pass kernels for TeddyBear (see engine.h).
#define FIELD_T TeddyBear
#define FIELD_U 0
#define FIELD_FN launch_pass_teddybear
#define FIELD_MAXLR 13
#include "pass_inst.inc"
