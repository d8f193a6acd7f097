// basis_internal.h -- host representation behind the opaque BasisSet_t.
#pragma once
#include <vector>

struct BasisSet {
    int pure = 1;
    int natoms = 0;
    int charge = 0;
    int nelec_neutral = 0;
    std::vector<int> Z;
    std::vector<double> xyz;  // bohr, 3 per atom
    int nshells = 0;
    std::vector<int> L, nprim, poff, atom, fstart, s_startind;
    std::vector<double> alpha, coef;  // coef includes primitive normalisation
    int nbf = 0;
    int maxdim = 0;
    std::vector<std::vector<double>> guess;  // per element Z, nbf_atom^2 values
};

void gtf_boys_table(double* tab, int ngrid, int ncol, double delta);
// BOYS8_NM slabs of BOYS8_NGRID x BOYS8_NTERM Taylor coefficients (gtf_internal.h)
void gtf_boys_taylor_tables(double* t8);
