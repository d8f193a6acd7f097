/*
 * CInt.h -- basis-set and integral API of the MI355X Fock engine.
 *
 * Same names, argument meaning and status codes as GTFock's
 * libcint/include/CInt.h:45-181 (whose implementation ships only as closed
 * archives); this library implements them from scratch on top of its own
 * basis parser and HIP integral kernels.  Each declaration cites the
 * reference declaration it replaces.
 */
#ifndef __CINT_H__
#define __CINT_H__

#include <stdint.h>

struct ERD;
struct OED;
struct BasisSet;

typedef struct OED *OED_t;
typedef struct ERD *ERD_t;
typedef struct BasisSet *BasisSet_t;

/* libcint/include/CInt.h:18-28 */
typedef enum {
    CINT_STATUS_SUCCESS = 0,
    CINT_STATUS_NOT_INITIALIZED = 1,
    CINT_STATUS_ALLOC_FAILED = 2,
    CINT_STATUS_INVALID_VALUE = 3,
    CINT_STATUS_EXECUTION_FAILED = 4,
    CINT_STATUS_INTERNAL_ERROR = 5,
    CINT_STATUS_FILEIO_FAILED = 6,
    CINT_STATUS_OFFLOAD_ERROR = 7
} CIntStatus_t;

#ifdef __cplusplus
extern "C" {
#endif

/* ---- basis set: libcint/include/CInt.h:45-145 ---- */
CIntStatus_t CInt_createBasisSet(BasisSet_t *basis);                      /* CInt.h:45 */
CIntStatus_t CInt_destroyBasisSet(BasisSet_t basis);                      /* CInt.h:47 */
CIntStatus_t CInt_unpackBasisSet(BasisSet_t basis, void *buf);            /* CInt.h:49 */
CIntStatus_t CInt_importBasisSet(BasisSet_t basis, int natoms, int *Zs,   /* CInt.h:56 */
                                 double *X, double *Y, double *Z, int nprims, int nshells,
                                 int pure, int *shells_p_atom, int *prims_p_shell, int *L,
                                 double *cc, double *alpha);
CIntStatus_t CInt_loadBasisSet(BasisSet_t basis, char *bsfile, char *xyzfile); /* CInt.h:64 */
CIntStatus_t CInt_packBasisSet(BasisSet_t basis, void **buf, int *bufsize);    /* CInt.h:68 */
int CInt_getNumShells(BasisSet_t basis);                                  /* CInt.h:72 */
int CInt_getNumFuncs(BasisSet_t basis);                                   /* CInt.h:74 */
int CInt_getNumAtoms(BasisSet_t basis);                                   /* CInt.h:76 */
int CInt_getMaxShellDim(BasisSet_t basis);                                /* CInt.h:78 */
int CInt_getNumOccOrb(BasisSet_t basis);                                  /* CInt.h:80 */
int CInt_getFuncEndInd(BasisSet_t basis, int shellid);                    /* CInt.h:82 */
int CInt_getAtomStartInd(BasisSet_t basis, int atomid);                   /* CInt.h:85 */
void CInt_getShellxyz(BasisSet_t basis, int shellid, double *x, double *y, double *z); /* CInt.h:123 */
int CInt_getShellDim(BasisSet_t basis, int shellid);                      /* CInt.h:129 */
int CInt_getFuncStartInd(BasisSet_t basis, int shellid);                  /* CInt.h:132 */
void CInt_getInitialGuess(BasisSet_t basis, int atomid, double **guess,   /* CInt.h:135 */
                          int *spos, int *epos);
int CInt_getTotalCharge(BasisSet_t basis);                                /* CInt.h:141 */
int CInt_getNneutral(BasisSet_t basis);                                   /* CInt.h:143 */
double CInt_getNucEnergy(BasisSet_t basis);                               /* CInt.h:145 */

/* ---- one-electron integrals: libcint/include/CInt.h:90-121 ----
 * Blocks are column-major within the shell pair (row index fastest), as
 * pfock/one_electron.c:14-25 reads them.  The pointer is owned by the OED
 * handle and valid until its next call. */
CIntStatus_t CInt_createOED(BasisSet_t basis, OED_t *oed);                /* CInt.h:90 */
CIntStatus_t CInt_destroyOED(OED_t oed);                                  /* CInt.h:93 */
CIntStatus_t CInt_computePairKin(BasisSet_t basis, OED_t oed, int A, int B,   /* CInt.h:95 */
                                 double **integrals, int *nints);
CIntStatus_t CInt_computePairOvl(BasisSet_t basis, OED_t oed, int A, int B,   /* CInt.h:102 */
                                 double **integrals, int *nints);
CIntStatus_t CInt_computePairPot(BasisSet_t basis, OED_t oed, int A, int B,   /* CInt.h:109 */
                                 double **integrals, int *nints);
CIntStatus_t CInt_computePairCoreH(BasisSet_t basis, OED_t oed, int A, int B, /* CInt.h:116 */
                                   double **integrals, int *nints);

/* ---- two-electron integrals: libcint/include/CInt.h:152-181 ----
 * integrals[iM + dM*(iN + dN*(iP + dP*iQ))] (pfock/fock_task.c:73-74). */
CIntStatus_t CInt_createERD(BasisSet_t basis, ERD_t *erd, int nthreads);  /* CInt.h:152 */
CIntStatus_t CInt_destroyERD(ERD_t erd);                                  /* CInt.h:156 */
CIntStatus_t CInt_computeShellQuartet(BasisSet_t basis, ERD_t erd, int tid,   /* CInt.h:159 */
                                      int A, int B, int C, int D, double **integrals, int *nints);
CIntStatus_t CInt_computeShellQuartets(BasisSet_t basis, ERD_t erd,       /* CInt.h:169 */
                                       uint32_t threadId, uint32_t shellIndixA,
                                       const uint32_t *shellIndicesB, uint32_t shellIndixC,
                                       const uint32_t *shellIndicesD, uint32_t shellIndicesCount,
                                       double **integrals, int *integralsCount);
void CInt_getMaxMemory(ERD_t erd, double *memsize);                       /* CInt.h:180 */

#ifdef __cplusplus
}
#endif

#endif /* __CINT_H__ */
