// oneel.h -- launchers of the one-electron kernels.
#pragma once
#include "gtf_internal.h"

int gtf_oneel(hipStream_t st, const BasisView& b, const double* boys, int ld, double* S, double* T, double* V);

// one molecule of a batched one-electron launch: its first 64-thread block,
// shell and atom offsets into the shared basis arrays, sizes, the leading
// dimension of its matrices and their offset in S / T / V
struct OneelMol {
    int blk0, shell_off, nshells, atom_off, natoms, ld;
    long moff;
};

// S, T, V of every listed molecule in one launch; d_mol holds nmol + 1
// entries, entry nmol's blk0 = nblocks
int gtf_oneel_batch(hipStream_t st, const BasisView& b, const double* boys, const OneelMol* d_mol, int nmol,
                    int nblocks, double* S, double* T, double* V);
