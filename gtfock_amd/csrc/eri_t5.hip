#define GTF_BRA_TYPE 5
#include "eri_inst.inc"
