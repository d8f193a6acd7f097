/*
 * gtf_oracle.h -- CPU FP64 restatement of GTFock's Fock build, Schwarz
 * screening, McWeeny/Palser-Manolopoulos purification and Pulay DIIS.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker for the HIP library
 * (gtfock_amd/csrc) and the timed CPU baseline of bench.py.  Nothing in the
 * product library links, loads or calls it.
 *
 * Reference files restated (semantics, not code):
 *   pfock/screening.c:15-180      Schwarz screening and the pair CSR
 *   pfock/fock_task.c:29-219      quartet enumeration and update_F digestion
 *   pfock/pfock.c:1274-1729       PFock_computeFock task loop, symmetrisation
 *   pfock/pfock.c:1134-1188       PFock_getMat (F = J + K)
 *   pfock/pfock.c:1782-1867       X = U diag(lambda^-1/2) U^T
 *   pscf/scf.c:80-127,348-459     energy, fock_build, SCF iteration
 *   pscf/purif.c:257-488          compute_purification
 *   pscf/purif.c:491-641          compute_diis
 *   pscf/pdgemm.c:164-238         C = A * B^T convention of pdgemm3D_2
 * The integral library GTFock calls (OptErd behind libcint/include/CInt.h)
 * ships only as closed archives; the integrals here are written from the
 * published methods: Obara-Saika (J. Chem. Phys. 84, 3963 (1986)) vertical
 * recursion with the Head-Gordon-Pople (J. Chem. Phys. 89, 5777 (1988))
 * horizontal recursion, and the Obara-Saika one-electron recursions.
 */
#ifndef GTF_ORACLE_H
#define GTF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ob_basis ob_basis;

/* Basis set: Gaussian94 .gbs (spherical) + .xyz in Angstrom. */
ob_basis *ob_load(const char *gbs_path, const char *xyz_path);
void ob_free(ob_basis *b);
int ob_nshells(const ob_basis *b);
int ob_nbf(const ob_basis *b);
int ob_natoms(const ob_basis *b);
int ob_nocc(const ob_basis *b);
int ob_shell_L(const ob_basis *b, int s);
int ob_shell_nprim(const ob_basis *b, int s);
int ob_shell_fstart(const ob_basis *b, int s);
int ob_shell_atom(const ob_basis *b, int s);
int ob_shell_dim(const ob_basis *b, int s);
int ob_max_shell_dim(const ob_basis *b);
double ob_nuc_energy(const ob_basis *b);
void ob_atom_xyz(const ob_basis *b, int a, double *xyz_bohr);

/* Boys function F_m(T), m = 0..mmax. */
void ob_boys(int mmax, double T, double *F);

/* Contracted, normalised, spherical ERIs of one shell quartet in the
 * reference layout integrals[iM + dM*(iN + dN*(iP + dP*iQ))]
 * (pfock/fock_task.c:73-74).  Returns the number of integrals. */
int ob_eri_quartet(const ob_basis *b, int M, int N, int P, int Q, double *out);

/* One-electron matrices, nbf x nbf row-major. */
void ob_overlap(const ob_basis *b, double *S);
void ob_kinetic(const ob_basis *b, double *T);
void ob_potential(const ob_basis *b, double *V);
void ob_core_hamiltonian(const ob_basis *b, double *H);

/* Schwarz screening.  vpair (nshells^2, may be NULL) receives
 * max |(mn|mn)| per ordered shell pair.  The CSR arrays are allocated by the
 * caller with room for nshells^2 entries.  Returns nnz. */
int ob_screening(const ob_basis *b, double tol, double *vpair, double *vmax,
                 int *shellptr, int *shellid, int *shellrid, double *shellvalue);

/* Fock build statistics. */
typedef struct {
    double usq;     /* shell quartets computed */
    double uitl;    /* integrals computed */
    double nnz;     /* significant shell pairs */
    double screen_s; /* wall seconds of the screening (CSR) pass */
    double build_s;  /* wall seconds of the rest: quartet loop, reduction, symmetrisation */
} ob_fock_stats;

/* J and K accumulators after symmetrisation: Jacc = 2J(D), Kacc = -K(D),
 * so F = Jacc + Kacc = G(D).  D, Jacc, Kacc are nbf x nbf row-major. */
void ob_fock_build(const ob_basis *b, double tol, int nblocks, const double *D,
                   double *Jacc, double *Kacc, ob_fock_stats *st);

/* List the unique screened quartets (M,N,P,Q) as the reference task loop
 * enumerates them.  quartets (4 ints each) may be NULL to count only. */
long ob_enumerate_quartets(const ob_basis *b, double tol, int *quartets, long cap);

/* Dense helpers with the pdgemm3D_2 convention: C = A * B^T (n x n). */
void ob_gemm_abt(int n, const double *A, const double *B, double *C);

/* X = S^{-1/2} by a cyclic Jacobi eigensolver. */
void ob_inv_sqrt(int n, const double *S, double *X);

/* Palser-Manolopoulos canonical purification of F (orthogonal basis);
 * D_out = X D X.  Returns the iteration count. */
int ob_purification(int n, int nocc, const double *F, const double *X,
                    double *D_out);

/* Pulay DIIS state (pscf/purif.c:491-641). */
typedef struct ob_diis ob_diis;
ob_diis *ob_diis_create(int n);
void ob_diis_free(ob_diis *d);
/* F is overwritten with the orthogonalised / extrapolated Fock matrix. */
void ob_diis_step(ob_diis *d, const double *X, const double *S,
                  const double *D, double *F, int iter);

/* Whole SCF as pscf/scf.c runs it with a zero initial density
 * (core-Hamiltonian guess).  energies[niters] receives the total energy of
 * each iteration; returns the number of iterations run.  Dfinal (nbf^2, may
 * be NULL) receives the last density; Ffinal the last F = H + G(D). */
int ob_scf(const ob_basis *b, double tol, int niters, double *energies,
           double *Dfinal, double *Ffinal);

#ifdef __cplusplus
}
#endif
#endif
