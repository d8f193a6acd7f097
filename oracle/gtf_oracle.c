/*
 * gtf_oracle.c -- CPU FP64 restatement of GTFock's Fock build and density
 * purification.  TEST INFRASTRUCTURE ONLY: the checker for the HIP library and
 * the timed CPU baseline of bench.py.  The product library never links it.
 *
 * See gtf_oracle.h for the list of reference files each part restates.
 */
#include "gtf_oracle.h"

#include <ctype.h>
#include <math.h>
#include <time.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define LMAXS 2                   /* highest shell angular momentum (d) */
#define LMAXP (2 * LMAXS)         /* highest pair angular momentum      */
#define LMAXQ (4 * LMAXS)         /* highest quartet angular momentum   */
#define NCU(l) (((l) + 1) * ((l) + 2) * ((l) + 3) / 6) /* cart comps, order <= l */
#define NC(l) (((l) + 1) * ((l) + 2) / 2)               /* cart comps, order == l */
#define BOHR_PER_ANGSTROM (1.0 / 0.529177210903)         /* CODATA 2018 */

/* ------------------------------------------------------------------------ */
/* basis set                                                                  */

struct ob_basis {
    int natoms;
    int *Z;
    double *xyz;      /* bohr, 3 per atom */
    int charge;
    int nshells;
    int *L, *nprim, *poff, *atom, *fstart;
    double *alpha, *coef;  /* coef includes primitive normalisation */
    int nbf;
    int maxdim;
};

static const char *ELEMS[] = {
    "X",  "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg",
    "Al", "Si", "P",  "S",  "Cl", "Ar", "K",  "Ca", "Sc", "Ti", "V",  "Cr", "Mn",
    "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr"};
#define NELEMS 37

static int elem_z(const char *s)
{
    for (int z = 1; z < NELEMS; z++) {
        if (strlen(s) == strlen(ELEMS[z])) {
            int ok = 1;
            for (size_t i = 0; i < strlen(s); i++)
                if (tolower((unsigned char)s[i]) != tolower((unsigned char)ELEMS[z][i])) ok = 0;
            if (ok) return z;
        }
    }
    return -1;
}

typedef struct {
    int L, nprim;
    double alpha[32], coef[32];
} gshell;
typedef struct {
    int nsh;
    gshell sh[32];
} gelem;

static double dfact(int n) /* (n)!! with (-1)!! = 1 */
{
    double r = 1.0;
    for (int k = n; k > 1; k -= 2) r *= k;
    return r;
}

static void fix_d_exponent(char *s)
{
    for (; *s; s++)
        if (*s == 'D' || *s == 'd') *s = 'E';
}

/* Parse a Gaussian94 basis file into per-element shell lists. */
static int parse_gbs(const char *path, gelem *elems)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return -1;
    char line[512];
    int cur = -1;
    memset(elems, 0, sizeof(gelem) * NELEMS);
    while (fgets(line, sizeof line, fp)) {
        char t0[64] = {0}, t1[64] = {0}, t2[64] = {0};
        int nt = sscanf(line, "%63s %63s %63s", t0, t1, t2);
        if (nt <= 0 || t0[0] == '!') continue;
        if (strcmp(t0, "****") == 0) { cur = -1; continue; }
        if (cur < 0) {
            if (nt == 2 && strcmp(t1, "0") == 0 && elem_z(t0) > 0) cur = elem_z(t0);
            continue;
        }
        /* shell header: TYPE nprim scale */
        if (nt >= 2 && isalpha((unsigned char)t0[0])) {
            int np = atoi(t1);
            double scale = (nt >= 3) ? atof(t2) : 1.0;
            int isSP = (strcasecmp(t0, "SP") == 0);
            int L = -1;
            switch (toupper((unsigned char)t0[0])) {
            case 'S': L = 0; break;
            case 'P': L = 1; break;
            case 'D': L = 2; break;
            case 'F': L = 3; break;
            }
            if (L < 0 || np <= 0 || np > 32) { fclose(fp); return -2; }
            gelem *e = &elems[cur];
            gshell *s1 = &e->sh[e->nsh++];
            gshell *s2 = isSP ? &e->sh[e->nsh++] : NULL;
            s1->L = isSP ? 0 : L;
            s1->nprim = np;
            if (s2) { s2->L = 1; s2->nprim = np; }
            for (int k = 0; k < np; k++) {
                if (!fgets(line, sizeof line, fp)) { fclose(fp); return -3; }
                fix_d_exponent(line);
                double a, c1, c2 = 0.0;
                int n = sscanf(line, "%lf %lf %lf", &a, &c1, &c2);
                if (n < 2) { fclose(fp); return -3; }
                s1->alpha[k] = a * scale * scale;
                s1->coef[k] = c1;
                if (s2) { s2->alpha[k] = a * scale * scale; s2->coef[k] = c2; }
            }
        }
    }
    fclose(fp);
    return 0;
}

ob_basis *ob_load(const char *gbs_path, const char *xyz_path)
{
    gelem *elems = (gelem *)calloc(NELEMS, sizeof(gelem));
    if (parse_gbs(gbs_path, elems) != 0) { free(elems); return NULL; }
    FILE *fp = fopen(xyz_path, "r");
    if (!fp) { free(elems); return NULL; }
    char line[512];
    int natoms = 0;
    if (!fgets(line, sizeof line, fp) || sscanf(line, "%d", &natoms) != 1 || natoms <= 0) {
        fclose(fp); free(elems); return NULL;
    }
    ob_basis *b = (ob_basis *)calloc(1, sizeof(ob_basis));
    b->natoms = natoms;
    b->Z = (int *)calloc(natoms, sizeof(int));
    b->xyz = (double *)calloc(3 * natoms, sizeof(double));
    /* second line: optional integer total charge */
    b->charge = 0;
    if (fgets(line, sizeof line, fp)) {
        char tok[64], extra[64];
        int n = sscanf(line, "%63s %63s", tok, extra);
        if (n == 1) {
            char *end;
            long q = strtol(tok, &end, 10);
            if (*end == '\0') b->charge = (int)q;
        }
    }
    for (int a = 0; a < natoms; a++) {
        char sym[16];
        double x, y, z;
        if (!fgets(line, sizeof line, fp) || sscanf(line, "%15s %lf %lf %lf", sym, &x, &y, &z) != 4) {
            fclose(fp); ob_free(b); free(elems); return NULL;
        }
        b->Z[a] = elem_z(sym);
        if (b->Z[a] < 0 || elems[b->Z[a]].nsh == 0) { fclose(fp); ob_free(b); free(elems); return NULL; }
        b->xyz[3 * a + 0] = x * BOHR_PER_ANGSTROM;
        b->xyz[3 * a + 1] = y * BOHR_PER_ANGSTROM;
        b->xyz[3 * a + 2] = z * BOHR_PER_ANGSTROM;
    }
    fclose(fp);
    int ns = 0, np = 0;
    for (int a = 0; a < natoms; a++) {
        gelem *e = &elems[b->Z[a]];
        ns += e->nsh;
        for (int s = 0; s < e->nsh; s++) np += e->sh[s].nprim;
    }
    b->nshells = ns;
    b->L = (int *)calloc(ns, sizeof(int));
    b->nprim = (int *)calloc(ns, sizeof(int));
    b->poff = (int *)calloc(ns, sizeof(int));
    b->atom = (int *)calloc(ns, sizeof(int));
    b->fstart = (int *)calloc(ns + 1, sizeof(int));
    b->alpha = (double *)calloc(np, sizeof(double));
    b->coef = (double *)calloc(np, sizeof(double));
    int si = 0, pi = 0, fi = 0;
    b->maxdim = 0;
    for (int a = 0; a < natoms; a++) {
        gelem *e = &elems[b->Z[a]];
        for (int s = 0; s < e->nsh; s++) {
            gshell *g = &e->sh[s];
            if (g->L > LMAXS) { ob_free(b); free(elems); return NULL; }
            b->L[si] = g->L;
            b->nprim[si] = g->nprim;
            b->poff[si] = pi;
            b->atom[si] = a;
            b->fstart[si] = fi;
            int l = g->L;
            /* primitive normalisation of x^l exp(-a r^2), then contraction
             * renormalisation to unit self-overlap */
            double norm2 = 0.0;
            double cn[32];
            for (int k = 0; k < g->nprim; k++) {
                double al = g->alpha[k];
                cn[k] = g->coef[k] * sqrt(pow(4.0 * al, l) * pow(2.0 * al / M_PI, 1.5) / dfact(2 * l - 1));
            }
            for (int i = 0; i < g->nprim; i++)
                for (int j = 0; j < g->nprim; j++) {
                    double p = g->alpha[i] + g->alpha[j];
                    norm2 += cn[i] * cn[j] * dfact(2 * l - 1) / pow(2.0 * p, l) * pow(M_PI / p, 1.5);
                }
            for (int k = 0; k < g->nprim; k++) {
                b->alpha[pi + k] = g->alpha[k];
                b->coef[pi + k] = cn[k] / sqrt(norm2);
            }
            pi += g->nprim;
            fi += 2 * l + 1;
            if (2 * l + 1 > b->maxdim) b->maxdim = 2 * l + 1;
            si++;
        }
    }
    b->fstart[ns] = fi;
    b->nbf = fi;
    free(elems);
    return b;
}

void ob_free(ob_basis *b)
{
    if (!b) return;
    free(b->Z); free(b->xyz); free(b->L); free(b->nprim); free(b->poff);
    free(b->atom); free(b->fstart); free(b->alpha); free(b->coef);
    free(b);
}

int ob_nshells(const ob_basis *b) { return b->nshells; }
int ob_nbf(const ob_basis *b) { return b->nbf; }
int ob_natoms(const ob_basis *b) { return b->natoms; }
int ob_shell_L(const ob_basis *b, int s) { return b->L[s]; }
int ob_shell_nprim(const ob_basis *b, int s) { return b->nprim[s]; }
int ob_shell_fstart(const ob_basis *b, int s) { return b->fstart[s]; }
int ob_shell_atom(const ob_basis *b, int s) { return b->atom[s]; }
int ob_shell_dim(const ob_basis *b, int s) { return 2 * b->L[s] + 1; }
int ob_max_shell_dim(const ob_basis *b) { return b->maxdim; }
void ob_atom_xyz(const ob_basis *b, int a, double *xyz)
{
    xyz[0] = b->xyz[3 * a]; xyz[1] = b->xyz[3 * a + 1]; xyz[2] = b->xyz[3 * a + 2];
}

int ob_nocc(const ob_basis *b)
{
    int ne = -b->charge;
    for (int a = 0; a < b->natoms; a++) ne += b->Z[a];
    return ne / 2;
}

double ob_nuc_energy(const ob_basis *b)
{
    double e = 0.0;
    for (int i = 0; i < b->natoms; i++)
        for (int j = 0; j < i; j++) {
            double dx = b->xyz[3 * i] - b->xyz[3 * j];
            double dy = b->xyz[3 * i + 1] - b->xyz[3 * j + 1];
            double dz = b->xyz[3 * i + 2] - b->xyz[3 * j + 2];
            e += b->Z[i] * b->Z[j] / sqrt(dx * dx + dy * dy + dz * dz);
        }
    return e;
}

/* ------------------------------------------------------------------------ */
/* Boys function: series F_m(T) = exp(-T) sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1))
 * for the highest order, downward recursion F_m = (2T F_{m+1} + exp(-T))/(2m+1);
 * for T > 35 the asymptotic F_0 = sqrt(pi/T)/2 with upward recursion. */

void ob_boys(int mmax, double T, double *F)
{
    if (T > 35.0) {
        double e = exp(-T);
        F[0] = 0.5 * sqrt(M_PI / T);
        for (int m = 1; m <= mmax; m++) F[m] = ((2 * m - 1) * F[m - 1] - e) / (2.0 * T);
        return;
    }
    double e = exp(-T);
    double term = 1.0 / (2 * mmax + 1), sum = term;
    for (int k = 1; k < 400; k++) {
        term *= 2.0 * T / (2 * mmax + 2 * k + 1);
        sum += term;
        if (term < 1e-18 * sum) break;
    }
    F[mmax] = e * sum;
    for (int m = mmax - 1; m >= 0; m--) F[m] = (2.0 * T * F[m + 1] + e) / (2 * m + 1);
}

/* ------------------------------------------------------------------------ */
/* cartesian component bookkeeping: components of order l are ordered
 * lx = l..0, then lz = 0..l-lx; global index = NCU(l-1) + within-order index */

static int cart_lx[NCU(LMAXQ)], cart_ly[NCU(LMAXQ)], cart_lz[NCU(LMAXQ)], cart_l[NCU(LMAXQ)];
static int cart_tab_ready = 0;

static int cidx(int lx, int ly, int lz)
{
    int l = lx + ly + lz;
    int off = l * (l + 1) * (l + 2) / 6;
    return off + ((l - lx) * (l - lx + 1)) / 2 + lz;
}

static void cart_init(void)
{
    if (cart_tab_ready) return;
    for (int l = 0; l <= LMAXQ; l++)
        for (int lx = l; lx >= 0; lx--)
            for (int lz = 0; lz <= l - lx; lz++) {
                int ly = l - lx - lz;
                int k = cidx(lx, ly, lz);
                cart_lx[k] = lx; cart_ly[k] = ly; cart_lz[k] = lz; cart_l[k] = l;
            }
    cart_tab_ready = 1;
}

static int comp(int k, int d) { return d == 0 ? cart_lx[k] : (d == 1 ? cart_ly[k] : cart_lz[k]); }
static int minus1(int k, int d)
{
    int lx = cart_lx[k] - (d == 0), ly = cart_ly[k] - (d == 1), lz = cart_lz[k] - (d == 2);
    return (lx < 0 || ly < 0 || lz < 0) ? -1 : cidx(lx, ly, lz);
}
static int plus1(int k, int d)
{
    return cidx(cart_lx[k] + (d == 0), cart_ly[k] + (d == 1), cart_lz[k] + (d == 2));
}
static int first_dir(int k) { return cart_lx[k] ? 0 : (cart_ly[k] ? 1 : 2); }

/* real solid harmonics, rows m = -l..l over cartesian components of order l */
static void sph_matrix(int l, double *T /* (2l+1) x NC(l) */)
{
    int nc = NC(l);
    memset(T, 0, sizeof(double) * (2 * l + 1) * nc);
    if (l == 0) { T[0] = 1.0; return; }
    if (l == 1) { T[0] = T[4] = T[8] = 1.0; return; }  /* x, y, z */
    /* l == 2, cartesian order xx, xy, xz, yy, yz, zz */
    const double r3 = sqrt(3.0);
    T[0 * 6 + 1] = r3;                              /* xy          */
    T[1 * 6 + 4] = r3;                              /* yz          */
    T[2 * 6 + 0] = -0.5; T[2 * 6 + 3] = -0.5; T[2 * 6 + 5] = 1.0; /* z^2 - r^2/2 */
    T[3 * 6 + 2] = r3;                              /* xz          */
    T[4 * 6 + 0] = 0.5 * r3; T[4 * 6 + 3] = -0.5 * r3;            /* (x^2-y^2)   */
}

/* ------------------------------------------------------------------------ */
/* two-electron integrals: Obara-Saika VRR on (e0|f0), contraction, then
 * Head-Gordon-Pople HRR on both sides, then the spherical transform */

typedef struct {
    double X[NCU(LMAXP) * NCU(LMAXP) * (LMAXQ + 1)];
    double C[NCU(LMAXP) * NCU(LMAXP)];
    double H1[NCU(LMAXP) * NCU(LMAXS) * NCU(LMAXP)];
    double H2[NC(LMAXS) * NC(LMAXS) * NCU(LMAXP) * NCU(LMAXS)];
    double cart[NC(LMAXS) * NC(LMAXS) * NC(LMAXS) * NC(LMAXS)];
    double tmp[NC(LMAXS) * NC(LMAXS) * NC(LMAXS) * NC(LMAXS)];
    double W[NCU(LMAXP) * NCU(LMAXS) * NC(LMAXS) * NC(LMAXS)];
} eri_ws;

/* HRR: in[e][nv] for e with order in [L1, L1+L2] (global cart index) ->
 * out[a][b][nv] for order(a) == L1, order(b) == L2 (within-order indices).
 * AB = A - B.  work holds (L1+L2 range) x NCU(L2) x nv. */
static void hrr(int L1, int L2, const double *AB, const double *in, int nv, double *out, double *work)
{
    int ne = NCU(L1 + L2), nb = NCU(L2);
    /* work[e][b][v]; b = 0 copies the input */
    for (int e = 0; e < ne; e++)
        if (cart_l[e] >= L1)
            for (int v = 0; v < nv; v++) work[((size_t)e * nb + 0) * nv + v] = in[(size_t)e * nv + v];
    for (int lb = 1; lb <= L2; lb++)
        for (int b = NCU(lb - 1); b < NCU(lb); b++) {
            int d = first_dir(b), b1 = minus1(b, d);
            for (int a = NCU(L1 - 1 < 0 ? -1 : L1 - 1); a < NCU(L1 + L2 - lb); a++) {
                if (cart_l[a] < L1) continue;
                int a1 = plus1(a, d);
                for (int v = 0; v < nv; v++)
                    work[((size_t)a * nb + b) * nv + v] =
                        work[((size_t)a1 * nb + b1) * nv + v] + AB[d] * work[((size_t)a * nb + b1) * nv + v];
            }
        }
    int a0 = NCU(L1) - NC(L1), b0 = NCU(L2) - NC(L2);
    for (int a = 0; a < NC(L1); a++)
        for (int b = 0; b < NC(L2); b++)
            for (int v = 0; v < nv; v++)
                out[((size_t)a * NC(L2) + b) * nv + v] = work[((size_t)(a0 + a) * nb + (b0 + b)) * nv + v];
}

static int eri_quartet_ws(const ob_basis *b, int M, int N, int P, int Q, double *out, eri_ws *ws)
{
    cart_init();
    int la = b->L[M], lb = b->L[N], lc = b->L[P], ld = b->L[Q];
    int lab = la + lb, lcd = lc + ld, L = lab + lcd;
    const double *A = &b->xyz[3 * b->atom[M]], *B = &b->xyz[3 * b->atom[N]];
    const double *Cc = &b->xyz[3 * b->atom[P]], *Dd = &b->xyz[3 * b->atom[Q]];
    int ne = NCU(lab), nf = NCU(lcd);
    double AB[3], CD[3], AB2 = 0, CD2 = 0;
    for (int d = 0; d < 3; d++) {
        AB[d] = A[d] - B[d]; CD[d] = Cc[d] - Dd[d];
        AB2 += AB[d] * AB[d]; CD2 += CD[d] * CD[d];
    }
    memset(ws->C, 0, sizeof(double) * ne * nf);
    double F[LMAXQ + 1];
    for (int i = 0; i < b->nprim[M]; i++)
        for (int j = 0; j < b->nprim[N]; j++) {
            double a = b->alpha[b->poff[M] + i], bb = b->alpha[b->poff[N] + j];
            double z = a + bb;
            double Kab = exp(-a * bb / z * AB2) * b->coef[b->poff[M] + i] * b->coef[b->poff[N] + j];
            double Pp[3], PA[3];
            for (int d = 0; d < 3; d++) { Pp[d] = (a * A[d] + bb * B[d]) / z; PA[d] = Pp[d] - A[d]; }
            for (int k = 0; k < b->nprim[P]; k++)
                for (int l = 0; l < b->nprim[Q]; l++) {
                    double c = b->alpha[b->poff[P] + k], dd = b->alpha[b->poff[Q] + l];
                    double eta = c + dd;
                    double Kcd = exp(-c * dd / eta * CD2) * b->coef[b->poff[P] + k] * b->coef[b->poff[Q] + l];
                    double Qq[3], QC[3], W[3], WP[3], WQ[3], PQ2 = 0;
                    for (int d = 0; d < 3; d++) {
                        Qq[d] = (c * Cc[d] + dd * Dd[d]) / eta;
                        QC[d] = Qq[d] - Cc[d];
                        W[d] = (z * Pp[d] + eta * Qq[d]) / (z + eta);
                        WP[d] = W[d] - Pp[d];
                        WQ[d] = W[d] - Qq[d];
                        PQ2 += (Pp[d] - Qq[d]) * (Pp[d] - Qq[d]);
                    }
                    double rho = z * eta / (z + eta);
                    double pref = 2.0 * pow(M_PI, 2.5) / (z * eta * sqrt(z + eta)) * Kab * Kcd;
                    ob_boys(L, rho * PQ2, F);
                    double *X = ws->X;
                    const int sm = L + 1;
#define XI(e, f, m) X[((size_t)(e) * nf + (f)) * sm + (m)]
                    for (int m = 0; m <= L; m++) XI(0, 0, m) = pref * F[m];
                    double oo2z = 0.5 / z, oo2e = 0.5 / eta, oo2ze = 0.5 / (z + eta);
                    double rz = rho / z, re = rho / eta;
                    for (int e = 1; e < ne; e++) {
                        int d = first_dir(e), e1 = minus1(e, d), e2 = minus1(e1, d);
                        int ei = comp(e, d) - 1;
                        for (int m = 0; m <= L - cart_l[e]; m++) {
                            double v = PA[d] * XI(e1, 0, m) + WP[d] * XI(e1, 0, m + 1);
                            if (ei > 0) v += ei * oo2z * (XI(e2, 0, m) - rz * XI(e2, 0, m + 1));
                            XI(e, 0, m) = v;
                        }
                    }
                    for (int f = 1; f < nf; f++) {
                        int d = first_dir(f), f1 = minus1(f, d), f2 = minus1(f1, d);
                        int fi = comp(f, d) - 1;
                        for (int e = 0; e < ne; e++) {
                            int em = minus1(e, d);
                            int ei = comp(e, d);
                            for (int m = 0; m <= L - cart_l[e] - cart_l[f]; m++) {
                                double v = QC[d] * XI(e, f1, m) + WQ[d] * XI(e, f1, m + 1);
                                if (fi > 0) v += fi * oo2e * (XI(e, f2, m) - re * XI(e, f2, m + 1));
                                if (ei > 0) v += ei * oo2ze * XI(em, f1, m + 1);
                                XI(e, f, m) = v;
                            }
                        }
                    }
                    for (int e = 0; e < ne; e++) {
                        if (cart_l[e] < la) continue;
                        for (int f = 0; f < nf; f++)
                            if (cart_l[f] >= lc) ws->C[e * nf + f] += XI(e, f, 0);
                    }
#undef XI
                }
        }
    /* bra HRR: C[e][f] -> H1[a][b][f] */
    int nca = NC(la), ncb = NC(lb), ncc = NC(lc), ncd = NC(ld);
    double *work = ws->W;
    hrr(la, lb, AB, ws->C, nf, ws->H1, work);
    /* ket HRR per (a,b): need in[f][v] layout -> transpose */
    int nab = nca * ncb;
    /* H2[f][ab] */
    for (int ab = 0; ab < nab; ab++)
        for (int f = 0; f < nf; f++) ws->H2[(size_t)f * nab + ab] = ws->H1[(size_t)ab * nf + f];
    hrr(lc, ld, CD, ws->H2, nab, ws->tmp, work); /* tmp[c][d][ab] */
    /* cart[a][b][c][d] */
    int ncd_ = ncc * ncd;
    for (int cd = 0; cd < ncd_; cd++)
        for (int ab = 0; ab < nab; ab++) ws->cart[(size_t)ab * ncd_ + cd] = ws->tmp[(size_t)cd * nab + ab];
    /* spherical transform on each index */
    double Ta[5 * 6], Tb[5 * 6], Tc[5 * 6], Td[5 * 6];
    sph_matrix(la, Ta); sph_matrix(lb, Tb); sph_matrix(lc, Tc); sph_matrix(ld, Td);
    int sa = 2 * la + 1, sb = 2 * lb + 1, sc = 2 * lc + 1, sd = 2 * ld + 1;
    for (int ia = 0; ia < sa; ia++)
        for (int ib = 0; ib < sb; ib++)
            for (int ic = 0; ic < sc; ic++)
                for (int id = 0; id < sd; id++) {
                    double v = 0.0;
                    for (int a = 0; a < nca; a++) {
                        double ta = Ta[ia * nca + a];
                        if (ta == 0.0) continue;
                        for (int bq = 0; bq < ncb; bq++) {
                            double tb = ta * Tb[ib * ncb + bq];
                            if (tb == 0.0) continue;
                            for (int c = 0; c < ncc; c++) {
                                double tc = tb * Tc[ic * ncc + c];
                                if (tc == 0.0) continue;
                                for (int d = 0; d < ncd; d++) {
                                    double td = Td[id * ncd + d];
                                    if (td == 0.0) continue;
                                    v += tc * td * ws->cart[((size_t)(a * ncb + bq) * ncc + c) * ncd + d];
                                }
                            }
                        }
                    }
                    out[ia + sa * (ib + sb * (ic + sc * id))] = v;
                }
    return sa * sb * sc * sd;
}

int ob_eri_quartet(const ob_basis *b, int M, int N, int P, int Q, double *out)
{
    eri_ws *ws = (eri_ws *)malloc(sizeof(eri_ws));
    int n = eri_quartet_ws(b, M, N, P, Q, out, ws);
    free(ws);
    return n;
}

/* ------------------------------------------------------------------------ */
/* one-electron integrals (Obara-Saika) -- layout like pfock/one_electron.c:14-25
 * is handled by the callers; these return full nbf x nbf row-major matrices */

static void overlap1d(int la, int lb, double PA, double PB, double p, double s00, double S[][LMAXS + 3])
{
    /* S[i][j], i <= la, j <= lb + 2 */
    double oo2p = 0.5 / p;
    for (int i = 0; i <= la; i++)
        for (int j = 0; j <= lb + 2; j++) {
            double v;
            if (i == 0 && j == 0) v = s00;
            else if (i > 0) {
                v = PA * S[i - 1][j];
                if (i > 1) v += (i - 1) * oo2p * S[i - 2][j];
                if (j > 0) v += j * oo2p * S[i - 1][j - 1];
            } else {
                v = PB * S[i][j - 1];
                if (j > 1) v += (j - 1) * oo2p * S[i][j - 2];
            }
            S[i][j] = v;
        }
}

static void sph_block(int la, int lb, const double *cartab, double *out /* sa x sb */)
{
    double Ta[30], Tb[30];
    sph_matrix(la, Ta); sph_matrix(lb, Tb);
    int nca = NC(la), ncb = NC(lb), sa = 2 * la + 1, sb = 2 * lb + 1;
    for (int i = 0; i < sa; i++)
        for (int j = 0; j < sb; j++) {
            double v = 0;
            for (int a = 0; a < nca; a++)
                for (int c = 0; c < ncb; c++) v += Ta[i * nca + a] * Tb[j * ncb + c] * cartab[a * ncb + c];
            out[i * sb + j] = v;
        }
}

/* which: 0 overlap, 1 kinetic */
static void st_matrix(const ob_basis *b, double *Mout, int which)
{
    cart_init();
    int n = b->nbf;
    #pragma omp parallel for schedule(dynamic)
    for (int s1 = 0; s1 < b->nshells; s1++)
        for (int s2 = 0; s2 < b->nshells; s2++) {
            int la = b->L[s1], lb = b->L[s2];
            const double *A = &b->xyz[3 * b->atom[s1]], *B = &b->xyz[3 * b->atom[s2]];
            double cartab[36] = {0};
            for (int i = 0; i < b->nprim[s1]; i++)
                for (int j = 0; j < b->nprim[s2]; j++) {
                    double a = b->alpha[b->poff[s1] + i], bb = b->alpha[b->poff[s2] + j];
                    double p = a + bb, mu = a * bb / p;
                    double cc = b->coef[b->poff[s1] + i] * b->coef[b->poff[s2] + j];
                    double S[3][LMAXS + 1][LMAXS + 3], T1[3][LMAXS + 1][LMAXS + 1];
                    for (int d = 0; d < 3; d++) {
                        double P = (a * A[d] + bb * B[d]) / p;
                        double X = A[d] - B[d];
                        overlap1d(la, lb, P - A[d], P - B[d], p, sqrt(M_PI / p) * exp(-mu * X * X), S[d]);
                        for (int ii = 0; ii <= la; ii++)
                            for (int jj = 0; jj <= lb; jj++) {
                                double t = -2.0 * bb * bb * S[d][ii][jj + 2] + bb * (2 * jj + 1) * S[d][ii][jj];
                                if (jj >= 2) t -= 0.5 * jj * (jj - 1) * S[d][ii][jj - 2];
                                T1[d][ii][jj] = t;
                            }
                    }
                    for (int ca = 0; ca < NC(la); ca++)
                        for (int cb = 0; cb < NC(lb); cb++) {
                            int ka = NCU(la) - NC(la) + ca, kb = NCU(lb) - NC(lb) + cb;
                            int ix = cart_lx[ka], iy = cart_ly[ka], iz = cart_lz[ka];
                            int jx = cart_lx[kb], jy = cart_ly[kb], jz = cart_lz[kb];
                            double v;
                            if (which == 0) v = S[0][ix][jx] * S[1][iy][jy] * S[2][iz][jz];
                            else
                                v = T1[0][ix][jx] * S[1][iy][jy] * S[2][iz][jz] +
                                    S[0][ix][jx] * T1[1][iy][jy] * S[2][iz][jz] +
                                    S[0][ix][jx] * S[1][iy][jy] * T1[2][iz][jz];
                            cartab[ca * NC(lb) + cb] += cc * v;
                        }
                }
            double blk[25];
            sph_block(la, lb, cartab, blk);
            int sb = 2 * lb + 1;
            for (int i = 0; i < 2 * la + 1; i++)
                for (int j = 0; j < sb; j++) Mout[(size_t)(b->fstart[s1] + i) * n + b->fstart[s2] + j] = blk[i * sb + j];
        }
}

void ob_overlap(const ob_basis *b, double *S) { st_matrix(b, S, 0); }
void ob_kinetic(const ob_basis *b, double *T) { st_matrix(b, T, 1); }

void ob_potential(const ob_basis *b, double *V)
{
    cart_init();
    int n = b->nbf;
    #pragma omp parallel for schedule(dynamic)
    for (int s1 = 0; s1 < b->nshells; s1++)
        for (int s2 = 0; s2 < b->nshells; s2++) {
            int la = b->L[s1], lb = b->L[s2], lab = la + lb;
            const double *A = &b->xyz[3 * b->atom[s1]], *B = &b->xyz[3 * b->atom[s2]];
            double AB[3], AB2 = 0;
            for (int d = 0; d < 3; d++) { AB[d] = A[d] - B[d]; AB2 += AB[d] * AB[d]; }
            double acc[NCU(LMAXP)];
            memset(acc, 0, sizeof acc);
            for (int i = 0; i < b->nprim[s1]; i++)
                for (int j = 0; j < b->nprim[s2]; j++) {
                    double a = b->alpha[b->poff[s1] + i], bb = b->alpha[b->poff[s2] + j];
                    double p = a + bb;
                    double cc = b->coef[b->poff[s1] + i] * b->coef[b->poff[s2] + j];
                    double Kab = exp(-a * bb / p * AB2);
                    double Pp[3], PA[3];
                    for (int d = 0; d < 3; d++) { Pp[d] = (a * A[d] + bb * B[d]) / p; PA[d] = Pp[d] - A[d]; }
                    for (int c = 0; c < b->natoms; c++) {
                        const double *Cn = &b->xyz[3 * c];
                        double PC[3], PC2 = 0;
                        for (int d = 0; d < 3; d++) { PC[d] = Pp[d] - Cn[d]; PC2 += PC[d] * PC[d]; }
                        double F[LMAXP + 1];
                        ob_boys(lab, p * PC2, F);
                        double th[NCU(LMAXP)][LMAXP + 1];
                        for (int m = 0; m <= lab; m++) th[0][m] = 2.0 * M_PI / p * Kab * F[m];
                        for (int e = 1; e < NCU(lab); e++) {
                            int d = first_dir(e), e1 = minus1(e, d), e2 = minus1(e1, d);
                            int ei = comp(e, d) - 1;
                            for (int m = 0; m <= lab - cart_l[e]; m++) {
                                double v = PA[d] * th[e1][m] - PC[d] * th[e1][m + 1];
                                if (ei > 0) v += ei * 0.5 / p * (th[e2][m] - th[e2][m + 1]);
                                th[e][m] = v;
                            }
                        }
                        for (int e = 0; e < NCU(lab); e++) acc[e] -= b->Z[c] * cc * th[e][0];
                    }
                }
            double out[NC(LMAXS) * NC(LMAXS)];
            double work[NCU(LMAXP) * NCU(LMAXS)];
            hrr(la, lb, AB, acc, 1, out, work);
            double blk[25];
            sph_block(la, lb, out, blk);
            int sb = 2 * lb + 1;
            for (int ii = 0; ii < 2 * la + 1; ii++)
                for (int jj = 0; jj < sb; jj++) V[(size_t)(b->fstart[s1] + ii) * n + b->fstart[s2] + jj] = blk[ii * sb + jj];
        }
}

void ob_core_hamiltonian(const ob_basis *b, double *H)
{
    int n = b->nbf;
    double *V = (double *)malloc(sizeof(double) * n * n);
    ob_kinetic(b, H);
    ob_potential(b, V);
    for (long i = 0; i < (long)n * n; i++) H[i] += V[i];
    free(V);
}

/* ------------------------------------------------------------------------ */
/* Schwarz screening: pfock/screening.c:15-180 */

int ob_screening(const ob_basis *b, double tol, double *vpair, double *vmax_out,
                 int *shellptr, int *shellid, int *shellrid, double *shellvalue)
{
    int ns = b->nshells;
    double *v = (double *)calloc((size_t)ns * ns, sizeof(double));
    double vmax = 0.0;
    #pragma omp parallel
    {
        eri_ws *ws = (eri_ws *)malloc(sizeof(eri_ws));
        double *buf = (double *)malloc(sizeof(double) * 625);
        #pragma omp for schedule(dynamic) reduction(max : vmax)
        for (int M = 0; M < ns; M++)
            for (int N = 0; N < ns; N++) {
                int dM = 2 * b->L[M] + 1, dN = 2 * b->L[N] + 1;
                eri_quartet_ws(b, M, N, M, N, buf, ws);
                double mx = 0.0;
                /* the (m,n,m,n) diagonal: screening.c:75-76 */
                for (int iM = 0; iM < dM; iM++)
                    for (int iN = 0; iN < dN; iN++) {
                        double x = fabs(buf[iM + dM * (iN + dN * (iM + dM * iN))]);
                        if (x > mx) mx = x;
                    }
                v[(size_t)M * ns + N] = mx;
                if (mx > vmax) vmax = mx;
            }
        free(ws); free(buf);
    }
    double eta = tol * tol / vmax;
    int nnz = 0;
    for (int A = 0; A < ns; A++) {
        shellptr[A] = nnz;
        for (int B = 0; B < ns; B++) {
            double x = v[(size_t)A * ns + B];
            if (!(x > eta)) continue;
            if ((A > B && (A + B) % 2 == 1) || (A < B && (A + B) % 2 == 0)) continue;
            shellvalue[nnz] = (A == B) ? x : -x;
            shellid[nnz] = B;
            shellrid[nnz] = A;
            nnz++;
        }
    }
    shellptr[ns] = nnz;
    if (vpair) memcpy(vpair, v, sizeof(double) * ns * ns);
    if (vmax_out) *vmax_out = vmax;
    free(v);
    return nnz;
}

/* ------------------------------------------------------------------------ */
/* Fock build: pfock/pfock.c:1372-1400 task loop, pfock/fock_task.c:121-219
 * quartet filter, pfock/fock_task.c:29-117 update_F multiplicities */

typedef struct {
    int ns, nnz;
    int *ptr, *id, *rid;
    double *val;
} csr_t;

static void make_csr(const ob_basis *b, double tol, csr_t *c)
{
    int ns = b->nshells;
    c->ns = ns;
    c->ptr = (int *)malloc(sizeof(int) * (ns + 1));
    c->id = (int *)malloc(sizeof(int) * ns * ns);
    c->rid = (int *)malloc(sizeof(int) * ns * ns);
    c->val = (double *)malloc(sizeof(double) * ns * ns);
    c->nnz = ob_screening(b, tol, NULL, NULL, c->ptr, c->id, c->rid, c->val);
}

static void free_csr(csr_t *c) { free(c->ptr); free(c->id); free(c->rid); free(c->val); }

static int quartet_kept(int M, int N, int P, int Q)
{
    if ((M > P && (M + P) % 2 == 1) || (M < P && (M + P) % 2 == 0)) return 0;
    if (M == P && ((N > Q && (N + Q) % 2 == 1) || (N < Q && (N + Q) % 2 == 0))) return 0;
    return 1;
}

/* six J/K updates of one shell quartet; J[] = 2J accumulator, K[] = -K */
static void update_jk(const ob_basis *b, int M, int N, int P, int Q, int f1, int f2, int f3,
                      const double *I, const double *D, double *J, double *K)
{
    int n = b->nbf;
    int dM = 2 * b->L[M] + 1, dN = 2 * b->L[N] + 1, dP = 2 * b->L[P] + 1, dQ = 2 * b->L[Q] + 1;
    int oM = b->fstart[M], oN = b->fstart[N], oP = b->fstart[P], oQ = b->fstart[Q];
    double cJ1 = 2.0 * (1 + f1) * (1 + f2);            /* J_MN */
    double cJ2 = 2.0 * f3 * (1 + f1) * (1 + f2);       /* J_PQ */
    double cMP = (1 + f3);                             /* K_MP */
    double cNP = f1 * (1 + f3);                        /* K_NP */
    double cMQ = f2 * (1 + f3);                        /* K_MQ */
    double cNQ = f1 * f2 * (1 + f3);                   /* K_NQ */
    for (int iQ = 0; iQ < dQ; iQ++)
        for (int iP = 0; iP < dP; iP++)
            for (int iN = 0; iN < dN; iN++)
                for (int iM = 0; iM < dM; iM++) {
                    double x = I[iM + dM * (iN + dN * (iP + dP * iQ))];
                    int m = oM + iM, nn = oN + iN, p = oP + iP, q = oQ + iQ;
                    J[(size_t)m * n + nn] += cJ1 * D[(size_t)p * n + q] * x;
                    J[(size_t)p * n + q] += cJ2 * D[(size_t)m * n + nn] * x;
                    K[(size_t)m * n + p] -= cMP * D[(size_t)nn * n + q] * x;
                    K[(size_t)nn * n + p] -= cNP * D[(size_t)m * n + q] * x;
                    K[(size_t)m * n + q] -= cMQ * D[(size_t)nn * n + p] * x;
                    K[(size_t)nn * n + q] -= cNQ * D[(size_t)m * n + p] * x;
                }
}

static double wall_seconds(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void symmetrise(int n, double *A)
{
    for (int i = 0; i < n; i++)
        for (int j = 0; j < i; j++) {
            double s = 0.5 * (A[(size_t)i * n + j] + A[(size_t)j * n + i]);
            A[(size_t)i * n + j] = A[(size_t)j * n + i] = s;
        }
}

void ob_fock_build(const ob_basis *b, double tol, int nblocks, const double *D,
                   double *Jacc, double *Kacc, ob_fock_stats *st)
{
    int n = b->nbf, ns = b->nshells;
    csr_t c;
    double t0 = wall_seconds();
    make_csr(b, tol, &c);
    double t1 = wall_seconds();
    double tol2 = tol * tol;
    if (nblocks <= 0) nblocks = 4;
    if (nblocks > ns) nblocks = ns;
    int *blk = (int *)malloc(sizeof(int) * (nblocks + 1));
    for (int i = 0; i <= nblocks; i++) blk[i] = (int)((long)ns * i / nblocks);
    memset(Jacc, 0, sizeof(double) * n * n);
    memset(Kacc, 0, sizeof(double) * n * n);
    double usq = 0, uitl = 0;
    int nthr = 1;
#ifdef _OPENMP
    nthr = omp_get_max_threads();
#endif
    double *Jt = (double *)calloc((size_t)nthr * n * n, sizeof(double));
    double *Kt = (double *)calloc((size_t)nthr * n * n, sizeof(double));
    for (int task = 0; task < nblocks * nblocks; task++) {
        int rowid = task / nblocks, colid = task % nblocks;
        int startM = blk[rowid], endM = blk[rowid + 1] - 1;
        int startP = blk[colid], endP = blk[colid + 1] - 1;
        int startMN = c.ptr[startM], endMN = c.ptr[endM + 1];
        int startPQ = c.ptr[startP], endPQ = c.ptr[endP + 1];
        #pragma omp parallel reduction(+ : usq, uitl)
        {
            int t = 0;
#ifdef _OPENMP
            t = omp_get_thread_num();
#endif
            double *J = Jt + (size_t)t * n * n, *K = Kt + (size_t)t * n * n;
            eri_ws *ws = (eri_ws *)malloc(sizeof(eri_ws));
            double *buf = (double *)malloc(sizeof(double) * 625);
            #pragma omp for schedule(dynamic)
            for (int i = startMN; i < endMN; i++) {
                int M = c.rid[i], N = c.id[i];
                double v1 = c.val[i];
                int f1 = v1 < 0.0;
                for (int j = startPQ; j < endPQ; j++) {
                    int P = c.rid[j], Q = c.id[j];
                    if (!quartet_kept(M, N, P, Q)) continue;
                    double v2 = c.val[j];
                    if (!(fabs(v1 * v2) >= tol2)) continue;
                    int f2 = v2 < 0.0;
                    int f3 = !(M == P && N == Q);
                    int nints = eri_quartet_ws(b, M, N, P, Q, buf, ws);
                    usq += 1.0;
                    uitl += nints;
                    update_jk(b, M, N, P, Q, f1, f2, f3, buf, D, J, K);
                }
            }
            free(ws); free(buf);
        }
    }
    for (int t = 0; t < nthr; t++)
        for (long k = 0; k < (long)n * n; k++) {
            Jacc[k] += Jt[(size_t)t * n * n + k];
            Kacc[k] += Kt[(size_t)t * n * n + k];
        }
    free(Jt); free(Kt);
    symmetrise(n, Jacc);
    symmetrise(n, Kacc);
    if (st) {
        st->usq = usq; st->uitl = uitl; st->nnz = c.nnz;
        st->screen_s = t1 - t0;
        st->build_s = wall_seconds() - t1;
    }
    free(blk);
    free_csr(&c);
}

long ob_enumerate_quartets(const ob_basis *b, double tol, int *quartets, long cap)
{
    csr_t c;
    make_csr(b, tol, &c);
    double tol2 = tol * tol;
    long cnt = 0;
    for (int i = 0; i < c.nnz; i++)
        for (int j = 0; j < c.nnz; j++) {
            int M = c.rid[i], N = c.id[i], P = c.rid[j], Q = c.id[j];
            if (!quartet_kept(M, N, P, Q)) continue;
            if (!(fabs(c.val[i] * c.val[j]) >= tol2)) continue;
            if (quartets && cnt < cap) {
                quartets[4 * cnt] = M; quartets[4 * cnt + 1] = N;
                quartets[4 * cnt + 2] = P; quartets[4 * cnt + 3] = Q;
            }
            cnt++;
        }
    free_csr(&c);
    return cnt;
}

/* ------------------------------------------------------------------------ */
/* dense linear algebra for purification and DIIS */

void ob_gemm_abt(int n, const double *A, const double *B, double *C)
{
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double s = 0.0;
            const double *a = A + (size_t)i * n, *bb = B + (size_t)j * n;
            for (int k = 0; k < n; k++) s += a[k] * bb[k];
            C[(size_t)i * n + j] = s;
        }
}

/* cyclic Jacobi: S = U diag(w) U^T, then X = U diag(w^-1/2) U^T */
void ob_inv_sqrt(int n, const double *S, double *X)
{
    double *A = (double *)malloc(sizeof(double) * n * n);
    double *U = (double *)calloc((size_t)n * n, sizeof(double));
    memcpy(A, S, sizeof(double) * n * n);
    for (int i = 0; i < n; i++) U[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                if (i != j) off += A[(size_t)i * n + j] * A[(size_t)i * n + j];
                else diag += A[(size_t)i * n + j] * A[(size_t)i * n + j];
            }
        if (off <= 1e-30 * diag) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                double apq = A[(size_t)p * n + q];
                if (fabs(apq) < 1e-300) continue;
                double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
                for (int k = 0; k < n; k++) { /* columns p,q */
                    double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = cs * akp - sn * akq;
                    A[(size_t)k * n + q] = sn * akp + cs * akq;
                }
                for (int k = 0; k < n; k++) { /* rows p,q */
                    double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = cs * apk - sn * aqk;
                    A[(size_t)q * n + k] = sn * apk + cs * aqk;
                }
                for (int k = 0; k < n; k++) {
                    double ukp = U[(size_t)k * n + p], ukq = U[(size_t)k * n + q];
                    U[(size_t)k * n + p] = cs * ukp - sn * ukq;
                    U[(size_t)k * n + q] = sn * ukp + cs * ukq;
                }
            }
    }
    double *Us = (double *)malloc(sizeof(double) * n * n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) Us[(size_t)i * n + j] = U[(size_t)i * n + j] / sqrt(A[(size_t)j * n + j]);
    /* X = Us * U^T  (pfock.c:1860 GA_Dgemm('N','T')) */
    ob_gemm_abt(n, Us, U, X);
    free(A); free(U); free(Us);
}

int ob_purification(int n, int nocc, const double *F, const double *X, double *Dout)
{
    size_t nn = (size_t)n * n;
    double *D = (double *)malloc(sizeof(double) * nn);
    double *D2 = (double *)malloc(sizeof(double) * nn);
    double *D3 = (double *)malloc(sizeof(double) * nn);
    /* Gershgorin bounds, purif.c:302-340 */
    double hmin = 1e300, hmax = -1e300, trF = 0.0;
    for (int i = 0; i < n; i++) {
        double off = 0.0, dg = F[(size_t)i * n + i];
        for (int j = 0; j < n; j++) off += fabs(F[(size_t)j * n + i]);
        off -= fabs(dg);
        if (dg - off < hmin) hmin = dg - off;
        if (dg + off > hmax) hmax = dg + off;
        trF += dg;
    }
    double mu = trF / n;
    double l1 = (double)nocc / (hmax - mu), l2 = (double)(n - nocc) / (mu - hmin);
    double lambda = l1 < l2 ? l1 : l2;
    for (size_t k = 0; k < nn; k++) D[k] = F[k] * (-lambda / n);
    for (int i = 0; i < n; i++) D[(size_t)i * n + i] += lambda * mu / n + (double)nocc / n;
    int it;
    for (it = 0; it < 200; it++) {
        ob_gemm_abt(n, D, D, D2);   /* D symmetric: D*D */
        ob_gemm_abt(n, D, D2, D3);  /* D*D2 (D2 symmetric) */
        double err = 0.0, tr = 0.0, tr2 = 0.0;
        for (size_t k = 0; k < nn; k++) err += (D[k] - D2[k]) * (D[k] - D2[k]);
        err = sqrt(err);
        for (int i = 0; i < n; i++) {
            size_t k = (size_t)i * n + i;
            tr += D2[k] - D3[k];
            tr2 += D[k] - D2[k];
        }
        double c = tr / tr2;
        if (c < 0.5)
            for (size_t k = 0; k < nn; k++) D[k] = ((1.0 - 2.0 * c) * D[k] + (1.0 + c) * D2[k] - D3[k]) / (1.0 - c);
        else
            for (size_t k = 0; k < nn; k++) D[k] = ((1.0 + c) * D2[k] - D3[k]) / c;
        if (err < 1e-11) break;
    }
    /* D = X D X via two A*B^T products, purif.c:469-476 */
    ob_gemm_abt(n, X, D, D2);
    ob_gemm_abt(n, D2, X, Dout);
    free(D); free(D2); free(D3);
    return it;
}

#define MAX_DIIS 10
struct ob_diis {
    int n, len, bmax_id;
    double *vecs, *Fvecs;
    double b[(MAX_DIIS + 1) * (MAX_DIIS + 1)];
};

ob_diis *ob_diis_create(int n)
{
    ob_diis *d = (ob_diis *)calloc(1, sizeof(ob_diis));
    d->n = n;
    d->bmax_id = -1;
    d->vecs = (double *)calloc((size_t)MAX_DIIS * n * n, sizeof(double));
    d->Fvecs = (double *)calloc((size_t)MAX_DIIS * n * n, sizeof(double));
    for (int i = 0; i <= MAX_DIIS; i++)
        for (int j = 0; j <= MAX_DIIS; j++) d->b[i * (MAX_DIIS + 1) + j] = (i != j) ? -1.0 : 0.0;
    return d;
}

void ob_diis_free(ob_diis *d)
{
    if (!d) return;
    free(d->vecs); free(d->Fvecs); free(d);
}

/* Gauss-Jordan inverse with partial pivoting, in place (m <= 11) */
static void small_inverse(int m, double *A, int lda)
{
    double inv[(MAX_DIIS + 1) * (MAX_DIIS + 1)];
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) inv[i * m + j] = (i == j);
    for (int col = 0; col < m; col++) {
        int piv = col;
        for (int r = col + 1; r < m; r++)
            if (fabs(A[r * lda + col]) > fabs(A[piv * lda + col])) piv = r;
        if (piv != col)
            for (int k = 0; k < m; k++) {
                double t = A[col * lda + k]; A[col * lda + k] = A[piv * lda + k]; A[piv * lda + k] = t;
                t = inv[col * m + k]; inv[col * m + k] = inv[piv * m + k]; inv[piv * m + k] = t;
            }
        double d = A[col * lda + col];
        for (int k = 0; k < m; k++) { A[col * lda + k] /= d; inv[col * m + k] /= d; }
        for (int r = 0; r < m; r++) {
            if (r == col) continue;
            double f = A[r * lda + col];
            if (f == 0.0) continue;
            for (int k = 0; k < m; k++) { A[r * lda + k] -= f * A[col * lda + k]; inv[r * m + k] -= f * inv[col * m + k]; }
        }
    }
    for (int i = 0; i < m; i++)
        for (int j = 0; j < m; j++) A[i * lda + j] = inv[i * m + j];
}

void ob_diis_step(ob_diis *d, const double *X, const double *S, const double *D, double *F, int iter)
{
    int n = d->n;
    size_t nn = (size_t)n * n;
    const int ld = MAX_DIIS + 1;
    double *work = (double *)malloc(sizeof(double) * nn);
    double *curF = F;
    int cur = -1;
    if (iter > 1) {
        if (d->len < MAX_DIIS) cur = d->len++;
        else cur = d->bmax_id;
        double *e = d->vecs + (size_t)cur * nn;
        curF = d->Fvecs + (size_t)cur * nn;
        ob_gemm_abt(n, F, D, work);   /* F D^T */
        ob_gemm_abt(n, work, S, e);   /* (F D) S^T */
        memcpy(work, e, sizeof(double) * nn);
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) e[(size_t)i * n + j] -= work[(size_t)j * n + i];
        ob_gemm_abt(n, X, e, work);
        ob_gemm_abt(n, work, X, e);
        for (int i = 0; i < d->len; i++) {
            const double *e2 = d->vecs + (size_t)i * nn;
            double s = 0.0;
            for (size_t k = 0; k < nn; k++) s += e[k] * e2[k];
            d->b[cur * ld + i] = s;
            d->b[i * ld + cur] = s;
        }
        double bmax = -1e308;
        for (int i = 0; i < d->len; i++)
            if (bmax < d->b[i * ld + i]) { bmax = d->b[i * ld + i]; d->bmax_id = i; }
    }
    /* X F X^T into curF */
    ob_gemm_abt(n, X, F, work);
    ob_gemm_abt(n, work, X, curF);
    if (iter > 1) {
        int m = d->len + 1;
        double binv[(MAX_DIIS + 1) * (MAX_DIIS + 1)];
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++) binv[i * m + j] = d->b[i * ld + j];
        small_inverse(m, binv, m);
        double coeffs[MAX_DIIS + 1];
        for (int i = 0; i < m; i++) coeffs[i] = -binv[i * m + m - 1];
        memset(F, 0, sizeof(double) * nn);
        for (int i = 0; i < d->len; i++) {
            const double *Fv = d->Fvecs + (size_t)i * nn;
            for (size_t k = 0; k < nn; k++) F[k] += coeffs[i] * Fv[k];
        }
    }
    free(work);
}

/* ------------------------------------------------------------------------ */
/* SCF: pscf/scf.c:189-474 with a zero initial density */

int ob_scf(const ob_basis *b, double tol, int niters, double *energies, double *Dfinal, double *Ffinal)
{
    int n = b->nbf, nocc = ob_nocc(b);
    size_t nn = (size_t)n * n;
    double *S = (double *)malloc(sizeof(double) * nn), *H = (double *)malloc(sizeof(double) * nn);
    double *X = (double *)malloc(sizeof(double) * nn), *D = (double *)calloc(nn, sizeof(double));
    double *F = (double *)malloc(sizeof(double) * nn), *J = (double *)malloc(sizeof(double) * nn);
    double *K = (double *)malloc(sizeof(double) * nn);
    ob_overlap(b, S);
    ob_core_hamiltonian(b, H);
    ob_inv_sqrt(n, S, X);
    ob_diis *dd = ob_diis_create(n);
    double enuc = ob_nuc_energy(b), e0 = -1.0;
    int done = 0;
    for (int iter = 0; iter < niters; iter++) {
        ob_fock_build(b, tol, 4, D, J, K, NULL);
        double e = 0.0;
        for (size_t k = 0; k < nn; k++) {
            F[k] = J[k] + K[k] + H[k];
            e += D[k] * (H[k] + F[k]);
        }
        energies[iter] = e + enuc;
        done = iter + 1;
        if (Ffinal) memcpy(Ffinal, F, sizeof(double) * nn);
        if (Dfinal) memcpy(Dfinal, D, sizeof(double) * nn);
        if (iter > 0 && fabs(e - e0) < 1e-11) break;
        e0 = e;
        ob_diis_step(dd, X, S, D, F, iter);
        ob_purification(n, nocc, F, X, D);
    }
    ob_diis_free(dd);
    free(S); free(H); free(X); free(D); free(F); free(J); free(K);
    return done;
}
