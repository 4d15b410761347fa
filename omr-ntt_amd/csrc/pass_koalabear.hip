// Stockham pass kernels for KoalaBear (see engine.h).
#define FIELD_T KoalaBear
#define FIELD_U 0
#define FIELD_FN launch_pass_koalabear
#define FIELD_MAXLR 13
#include "pass_inst.inc"
