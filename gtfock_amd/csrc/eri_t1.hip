#define GTF_BRA_TYPE 1
#include "eri_inst.inc"
