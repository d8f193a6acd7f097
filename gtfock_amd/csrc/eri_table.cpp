// eri_table.cpp -- the class table of ERI kernels (filled once).
#include "eri_kernels.h"

void gtf_eri_register_t0(EriClassInfo*);
void gtf_eri_register_t1(EriClassInfo*);
void gtf_eri_register_t2(EriClassInfo*);
void gtf_eri_register_t3(EriClassInfo*);
void gtf_eri_register_t4(EriClassInfo*);
void gtf_eri_register_t5(EriClassInfo*);

namespace {
struct EriTable {
    EriClassInfo tab[GTF_NCLASSES];
    EriTable()
    {
        gtf_eri_register_t0(tab);
        gtf_eri_register_t1(tab);
        gtf_eri_register_t2(tab);
        gtf_eri_register_t3(tab);
        gtf_eri_register_t4(tab);
        gtf_eri_register_t5(tab);
    }
};
}  // namespace

const EriClassInfo* gtf_eri_class(int cls)
{
    // filled once, thread-safely (several batches may start at the same time)
    static const EriTable t;
    return (cls >= 0 && cls < GTF_NCLASSES) ? &t.tab[cls] : nullptr;
}
