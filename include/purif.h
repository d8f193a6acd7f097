/*
 * purif.h -- density purification and DIIS driver on one MI355X GPU.
 *
 * Replaces GTFock's pscf/purif.h:16-86.  The fields pscf/scf.c reads keep
 * their names and types (runpurif, srow_purif, scol_purif, nrows_purif,
 * ncols_purif, ldx, D_block, F_block, H_block, S_block, X_block, np_purif,
 * timedgemm, timepdgemm, timepass, timetr); the MPI communicators and the
 * 3-D process cube of the reference are gone: one GPU holds the whole
 * matrices and the DGEMMs run as FP64 MFMA kernels.  Host blocks are the
 * full nbf x nbf matrices, row-major with leading dimension ldx.
 *
 * X_block and S_block are uploaded to the GPU on the first compute_diis /
 * compute_purification call after create_purif.
 */
#ifndef __PURIF_H__
#define __PURIF_H__

#include "CInt.h"
#include "pfock.h"

#define MAX_DIIS 10   /* pscf/purif.h:12 */
#define LDBMAT 128    /* pscf/purif.h:13 */

struct purif_dev;

typedef struct _purif_t {
    int runpurif;
    int rundgemm;
    int nbf;
    int nobtls;
    int np_purif;
    int nprow_purif;
    int npcol_purif;
    int npgrd_purif;
    int srow_purif;
    int scol_purif;
    int nrows_purif;
    int ncols_purif;
    int ldx;
    int meshsize;
    double *H_block;
    double *X_block;
    double *S_block;
    double *F_block;
    double *D_block;
    /* DIIS bookkeeping (pscf/purif.h:62-70) */
    int len_diis;
    int bmax_id;
    double bmax;
    double b_mat[(MAX_DIIS + 1) * (MAX_DIIS + 1)];
    /* statistics (pscf/purif.h:72-76); timedgemm is GEMM kernel time */
    double timedgemm;
    double timepdgemm;
    double timetr;
    double timepass;
    /* device state */
    struct purif_dev *dev;
} purif_t;

#ifdef __cplusplus
extern "C" {
#endif

/* pscf/purif.h:78-79; the process-grid sizes must all be 1 */
purif_t *create_purif(BasisSet_t basis, int nprow_purif, int npcol_purif, int npgrd_purif);
/* pscf/purif.h:81 */
void destroy_purif(purif_t *purif);
/* pscf/purif.h:83 -- F_block: orthogonal-basis Fock matrix (X F X); on return
 * D_block = X D X; returns the purification iteration count */
int compute_purification(purif_t *purif, double *F_block, double *D_block);
/* pscf/purif.h:85-86 -- on return F_block = X F X, DIIS-extrapolated for iter > 1 */
void compute_diis(PFock_t pfock, purif_t *purif, double *D_block, double *F_block, int iter);

#ifdef __cplusplus
}
#endif

#endif /* __PURIF_H__ */
