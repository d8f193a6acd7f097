#define GTF_BRA_TYPE 0
#include "eri_inst.inc"
