#define GTF_BRA_TYPE 2
#include "eri_inst.inc"
