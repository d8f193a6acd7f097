// oneel.h -- launcher of the one-electron kernel.
#pragma once
#include "gtf_internal.h"

int gtf_oneel(hipStream_t st, const BasisView& b, const double* boys, int ld, double* S, double* T, double* V);
