/*
 * pfock.h -- the PFock Fock-matrix engine API, on one MI355X GPU.
 *
 * Same function names, argument meaning, status codes and error behaviour as
 * GTFock's pfock/pfock.h:188-565, so that pscf/scf.c's call sequence
 * (PFock_create, PFock_putDenMat, PFock_commitDenMats, PFock_computeFock,
 * PFock_getMat, ...) runs unchanged.  The engine struct is opaque: D, J, K,
 * the Schwarz data and the screened quartet lists stay resident in HBM and
 * the hot path (screening, ERIs, J/K digestion) runs as HIP kernels.
 *
 * Single process, single GPU: nprow = npcol = 1.  Calls are host-synchronous
 * at PFock_getMat and on return from PFock_computeFock.
 */
#ifndef __PFOCK_H__
#define __PFOCK_H__

#include "CInt.h"

struct PFock;
typedef struct PFock *PFock_t;   /* pfock/pfock.h:181 */

/* pfock/pfock.h:188-202 */
typedef enum {
    PFOCK_STATUS_SUCCESS = 0,
    PFOCK_STATUS_INIT_FAILED = 1,
    PFOCK_STATUS_ALLOC_FAILED = 2,
    PFOCK_STATUS_INVALID_VALUE = 3,
    PFOCK_STATUS_EXECUTION_FAILED = 4,
    PFOCK_STATUS_INTERNAL_ERROR = 5
} PFockStatus_t;

/* pfock/pfock.h:209-219 */
typedef enum {
    PFOCK_MAT_TYPE_D = 0,   /* density matrix */
    PFOCK_MAT_TYPE_F = 1,   /* Fock two-electron part G(D) = 2J - K */
    PFOCK_MAT_TYPE_J = 2,   /* Coulomb accumulator, 2J(D) */
    PFOCK_MAT_TYPE_K = 3    /* exchange accumulator, -K(D) */
} PFockMatType_t;

#ifdef __cplusplus
extern "C" {
#endif

/* pfock/pfock.h:241-243 -- screening, quartet lists, device state */
PFockStatus_t PFock_create(BasisSet_t basis, int nprow, int npcol, int ntasks,
                           double tolscr, int max_numdmat, int symm, PFock_t *pfock);
/* pfock/pfock.h:252 */
PFockStatus_t PFock_destroy(PFock_t pfock);
/* pfock/pfock.h:262 */
PFockStatus_t PFock_setNumDenMat(int numdmat, PFock_t pfock);
/* pfock/pfock.h:282-285 -- inclusive row/column ranges, row-major, leading dimension stride */
PFockStatus_t PFock_putDenMat(int rowstart, int rowend, int colstart, int colend,
                              int stride, double *dmat, int index, PFock_t pfock);
/* pfock/pfock.h:296 -- Global Arrays handles do not exist here */
PFockStatus_t PFock_putDenMatGA(int ga, int index, PFock_t pfock);
/* pfock/pfock.h:307 */
PFockStatus_t PFock_fillDenMat(double value, int index, PFock_t pfock);
/* pfock/pfock.h:320 */
PFockStatus_t PFock_commitDenMats(PFock_t pfock);
/* pfock/pfock.h:329 */
PFockStatus_t PFock_sync(PFock_t pfock);
/* pfock/pfock.h:350-358 */
PFockStatus_t PFock_getMat(PFock_t pfock, PFockMatType_t type, int index,
                           int rowstart, int rowend, int colstart, int colend,
                           int stride, double *mat);
/* pfock/pfock.h:370 */
PFockStatus_t PFock_getMatGA(PFock_t pfock, PFockMatType_t type, int index, int ga);
/* pfock/pfock.h:386 */
PFockStatus_t PFock_getLocalMatInds(PFock_t pfock, int *rowstart, int *rowend,
                                    int *colstart, int *colend);
/* pfock/pfock.h:409 -- returns a borrowed host mirror of the matrix */
PFockStatus_t PFock_getLocalMatPtr(PFock_t pfock, PFockMatType_t type, int index,
                                   int *rowstart, int *rowend, int *colstart, int *colend,
                                   int *stride, double **mat);
/* pfock/pfock.h:426 */
PFockStatus_t PFock_getMatGAHandle(PFock_t pfock, PFockMatType_t type, int index, int *ga);
/* pfock/pfock.h:441 */
PFockStatus_t PFock_computeFock(BasisSet_t basis, PFock_t pfock);
/* pfock/pfock.h:455-485 */
PFockStatus_t PFock_createCoreHMat(PFock_t pfock, BasisSet_t basis);
PFockStatus_t PFock_destroyCoreHMat(PFock_t pfock);
PFockStatus_t PFock_getCoreHMat(PFock_t pfock, int rowstart, int rowend,
                                int colstart, int colend, int stride, double *mat);
/* pfock/pfock.h:497-546 */
PFockStatus_t PFock_createOvlMat(PFock_t pfock, BasisSet_t basis);
PFockStatus_t PFock_destroyOvlMat(PFock_t pfock);
PFockStatus_t PFock_getOvlMat(PFock_t pfock, int rowstart, int rowend,
                              int colstart, int colend, int stride, double *mat);
PFockStatus_t PFock_getOvlMat2(PFock_t pfock, int rowstart, int rowend,
                               int colstart, int colend, int stride, double *mat);
/* pfock/pfock.h:556 -- bytes of host memory; device bytes via gtf_pfock_device_bytes */
PFockStatus_t PFock_getMemorySize(PFock_t pfock, double *mem_cpu);
/* pfock/pfock.h:565 */
PFockStatus_t PFock_getStatistics(PFock_t pfock);

#ifdef __cplusplus
}
#endif

#endif /* __PFOCK_H__ */
