#define GTF_BRA_TYPE 4
#include "eri_inst.inc"
