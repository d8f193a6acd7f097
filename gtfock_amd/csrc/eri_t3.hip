#define GTF_BRA_TYPE 3
#include "eri_inst.inc"
