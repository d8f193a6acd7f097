// basis.cpp -- the CInt basis-set API (libcint/include/CInt.h:45-145) written
// from scratch: Gaussian94 basis files, .xyz geometries, normalisation,
// per-atom initial-guess blocks, and a byte wire format for pack/unpack.
//
// Conventions (GTFock's closed CInt library does not document them; these
// are the choices of this implementation):
//  * geometry in Angstrom, converted with 1 bohr = 0.529177210903 A;
//  * shells ordered by atom (file order), then by the order in the basis file;
//  * primitives normalised for x^l exp(-a r^2), contractions renormalised to
//    unit self-overlap;
//  * spherical functions ordered m = -l..l (p as x, y, z);
//  * the .xyz second line may hold the integer total charge;
//  * initial-guess blocks are read from "<Element>.dat" files next to the
//    basis file when present (one value per line, nbf_atom^2 values),
//    otherwise the guess is zero (core-Hamiltonian start).
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/CInt.h"
#include "basis_internal.h"
#include "gtf_internal.h"

static const char* kElems[] = {"X",  "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",
                               "Ne", "Na", "Mg", "Al", "Si", "P",  "S",  "Cl", "Ar", "K",
                               "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni", "Cu",
                               "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr"};
static const int kNElems = 37;
static const double kBohrPerAngstrom = 1.0 / 0.529177210903;

static int element_z(const std::string& s)
{
    for (int z = 1; z < kNElems; z++) {
        const char* e = kElems[z];
        if (s.size() != strlen(e)) continue;
        bool ok = true;
        for (size_t i = 0; i < s.size(); i++)
            if (tolower((unsigned char)s[i]) != tolower((unsigned char)e[i])) ok = false;
        if (ok) return z;
    }
    return -1;
}

static double dfact(int n)
{
    double r = 1.0;
    for (int k = n; k > 1; k -= 2) r *= k;
    return r;
}

namespace {
struct RawShell {
    int L;
    std::vector<double> a, c;
};
}  // namespace

static int parse_gbs(const char* path, std::vector<std::vector<RawShell>>& elems, int& pure)
{
    FILE* fp = fopen(path, "r");
    if (!fp) return -1;
    elems.assign(kNElems, {});
    pure = 1;
    char line[512];
    int cur = -1;
    bool first = true;
    while (fgets(line, sizeof line, fp)) {
        char t0[64] = {0}, t1[64] = {0}, t2[64] = {0};
        int nt = sscanf(line, "%63s %63s %63s", t0, t1, t2);
        if (nt <= 0 || t0[0] == '!') continue;
        if (first) {
            first = false;
            if (!strcasecmp(t0, "cartesian")) { pure = 0; continue; }
            if (!strcasecmp(t0, "spherical")) { pure = 1; continue; }
        }
        if (!strcmp(t0, "****")) { cur = -1; continue; }
        if (cur < 0) {
            if (nt == 2 && !strcmp(t1, "0") && element_z(t0) > 0) cur = element_z(t0);
            continue;
        }
        if (nt < 2 || !isalpha((unsigned char)t0[0])) continue;
        int np = atoi(t1);
        double scale = nt >= 3 ? atof(t2) : 1.0;
        bool sp = !strcasecmp(t0, "SP");
        const char* kinds = "SPDFG";
        const char* k = strchr(kinds, toupper((unsigned char)t0[0]));
        if (!k || np <= 0) { fclose(fp); return -2; }
        RawShell s1, s2;
        s1.L = sp ? 0 : (int)(k - kinds);
        s2.L = 1;
        for (int i = 0; i < np; i++) {
            if (!fgets(line, sizeof line, fp)) { fclose(fp); return -3; }
            for (char* c = line; *c; c++)
                if (*c == 'D' || *c == 'd') *c = 'E';
            double a, c1, c2 = 0.0;
            if (sscanf(line, "%lf %lf %lf", &a, &c1, &c2) < 2) { fclose(fp); return -3; }
            s1.a.push_back(a * scale * scale);
            s1.c.push_back(c1);
            if (sp) { s2.a.push_back(a * scale * scale); s2.c.push_back(c2); }
        }
        elems[cur].push_back(s1);
        if (sp) elems[cur].push_back(s2);
    }
    fclose(fp);
    return 0;
}

static void read_guess(const std::string& dir, int z, std::vector<double>& g, int nbf_atom)
{
    g.clear();
    std::string path = dir + "/" + kElems[z] + ".dat";
    FILE* fp = fopen(path.c_str(), "r");
    if (!fp) return;
    double v;
    while (fscanf(fp, "%lf", &v) == 1) g.push_back(v);
    fclose(fp);
    if ((int)g.size() != nbf_atom * nbf_atom) g.clear();
}

static void finalize(BasisSet* b)
{
    b->nshells = (int)b->L.size();
    b->fstart.assign(b->nshells + 1, 0);
    b->maxdim = 0;
    int f = 0;
    for (int s = 0; s < b->nshells; s++) {
        b->fstart[s] = f;
        int d = b->pure ? 2 * b->L[s] + 1 : (b->L[s] + 1) * (b->L[s] + 2) / 2;
        f += d;
        if (d > b->maxdim) b->maxdim = d;
    }
    b->fstart[b->nshells] = f;
    b->nbf = f;
    b->s_startind.assign(b->natoms + 1, b->nshells);
    for (int s = b->nshells - 1; s >= 0; s--) b->s_startind[b->atom[s]] = s;
    for (int a = b->natoms - 1; a >= 0; a--)
        if (b->s_startind[a] > b->s_startind[a + 1]) b->s_startind[a] = b->s_startind[a + 1];
    b->nelec_neutral = 0;
    for (int a = 0; a < b->natoms; a++) b->nelec_neutral += b->Z[a];
}

extern "C" {

CIntStatus_t CInt_createBasisSet(BasisSet_t* basis)
{
    if (!basis) return CINT_STATUS_INVALID_VALUE;
    *basis = new (std::nothrow) BasisSet();
    return *basis ? CINT_STATUS_SUCCESS : CINT_STATUS_ALLOC_FAILED;
}

CIntStatus_t CInt_destroyBasisSet(BasisSet_t basis)
{
    delete basis;
    return CINT_STATUS_SUCCESS;
}

CIntStatus_t CInt_loadBasisSet(BasisSet_t b, char* bsfile, char* xyzfile)
{
    if (!b || !bsfile || !xyzfile) return CINT_STATUS_INVALID_VALUE;
    std::vector<std::vector<RawShell>> elems;
    int pure;
    if (parse_gbs(bsfile, elems, pure) != 0) return CINT_STATUS_FILEIO_FAILED;
    FILE* fp = fopen(xyzfile, "r");
    if (!fp) return CINT_STATUS_FILEIO_FAILED;
    char line[512];
    int natoms = 0;
    if (!fgets(line, sizeof line, fp) || sscanf(line, "%d", &natoms) != 1 || natoms <= 0) {
        fclose(fp);
        return CINT_STATUS_FILEIO_FAILED;
    }
    *b = BasisSet();
    b->pure = pure;
    b->natoms = natoms;
    b->charge = 0;
    if (fgets(line, sizeof line, fp)) {
        char tok[64], extra[64];
        if (sscanf(line, "%63s %63s", tok, extra) == 1) {
            char* end;
            long q = strtol(tok, &end, 10);
            if (*end == '\0') b->charge = (int)q;
        }
    }
    for (int a = 0; a < natoms; a++) {
        char sym[16];
        double x, y, z;
        if (!fgets(line, sizeof line, fp) || sscanf(line, "%15s %lf %lf %lf", sym, &x, &y, &z) != 4) {
            fclose(fp);
            return CINT_STATUS_FILEIO_FAILED;
        }
        int zz = element_z(sym);
        if (zz < 0 || elems[zz].empty()) {
            fclose(fp);
            fprintf(stderr, "CInt_loadBasisSet: no basis functions for element %s\n", sym);
            return CINT_STATUS_INVALID_VALUE;
        }
        b->Z.push_back(zz);
        b->xyz.push_back(x * kBohrPerAngstrom);
        b->xyz.push_back(y * kBohrPerAngstrom);
        b->xyz.push_back(z * kBohrPerAngstrom);
    }
    fclose(fp);
    for (int a = 0; a < natoms; a++) {
        for (const RawShell& rs : elems[b->Z[a]]) {
            int l = rs.L;
            if (l > 2) {
                fprintf(stderr, "CInt_loadBasisSet: shells above d are not supported\n");
                return CINT_STATUS_INVALID_VALUE;
            }
            std::vector<double> cn(rs.a.size());
            for (size_t k = 0; k < rs.a.size(); k++)
                cn[k] = rs.c[k] * std::sqrt(std::pow(4.0 * rs.a[k], l) * std::pow(2.0 * rs.a[k] / M_PI, 1.5) /
                                            dfact(2 * l - 1));
            double n2 = 0.0;
            for (size_t i = 0; i < rs.a.size(); i++)
                for (size_t j = 0; j < rs.a.size(); j++) {
                    double p = rs.a[i] + rs.a[j];
                    n2 += cn[i] * cn[j] * dfact(2 * l - 1) / std::pow(2.0 * p, l) * std::pow(M_PI / p, 1.5);
                }
            b->L.push_back(l);
            b->nprim.push_back((int)rs.a.size());
            b->poff.push_back((int)b->alpha.size());
            b->atom.push_back(a);
            for (size_t k = 0; k < rs.a.size(); k++) {
                b->alpha.push_back(rs.a[k]);
                b->coef.push_back(cn[k] / std::sqrt(n2));
            }
        }
    }
    if (!b->pure) {
        fprintf(stderr, "CInt_loadBasisSet: cartesian basis sets are not supported\n");
        return CINT_STATUS_INVALID_VALUE;
    }
    finalize(b);
    // per-atom guess blocks
    std::string dir(bsfile);
    size_t slash = dir.find_last_of('/');
    dir = slash == std::string::npos ? "." : dir.substr(0, slash);
    b->guess.assign(kNElems, {});
    for (int a = 0; a < natoms; a++) {
        int z = b->Z[a];
        if (!b->guess[z].empty()) continue;
        int nfa = 0;
        for (int s = b->s_startind[a]; s < b->s_startind[a + 1]; s++) nfa += 2 * b->L[s] + 1;
        read_guess(dir, z, b->guess[z], nfa);
        // no (usable) file: a zero block, sized here so that
        // CInt_getInitialGuess only reads (batches on several threads may
        // share a basis)
        if (b->guess[z].empty()) b->guess[z].assign((size_t)nfa * nfa, 0.0);
    }
    return CINT_STATUS_SUCCESS;
}

CIntStatus_t CInt_importBasisSet(BasisSet_t b, int natoms, int* Zs, double* X, double* Y, double* Z,
                                 int nprims, int nshells, int pure, int* shells_p_atom, int* prims_p_shell,
                                 int* L, double* cc, double* alpha)
{
    if (!b || natoms <= 0 || nshells <= 0 || !pure) return CINT_STATUS_INVALID_VALUE;
    *b = BasisSet();
    b->pure = 1;
    b->natoms = natoms;
    for (int a = 0; a < natoms; a++) {
        if (Zs[a] <= 0 || Zs[a] >= kNElems) return CINT_STATUS_INVALID_VALUE;
        b->Z.push_back(Zs[a]);
        b->xyz.push_back(X[a]);
        b->xyz.push_back(Y[a]);
        b->xyz.push_back(Z[a]);
    }
    int s = 0, p = 0;
    for (int a = 0; a < natoms; a++)
        for (int k = 0; k < shells_p_atom[a]; k++, s++) {
            if (s >= nshells || L[s] > 2) return CINT_STATUS_INVALID_VALUE;
            b->L.push_back(L[s]);
            b->nprim.push_back(prims_p_shell[s]);
            b->poff.push_back(p);
            b->atom.push_back(a);
            for (int i = 0; i < prims_p_shell[s]; i++, p++) {
                if (p >= nprims) return CINT_STATUS_INVALID_VALUE;
                b->alpha.push_back(alpha[p]);
                b->coef.push_back(cc[p]);  // taken as already normalised
            }
        }
    finalize(b);
    b->guess.assign(kNElems, {});
    return CINT_STATUS_SUCCESS;
}

// wire format: "GTFB" magic, version, then the scalar and vector fields
CIntStatus_t CInt_packBasisSet(BasisSet_t b, void** buf, int* bufsize)
{
    std::vector<char> out;
    auto put = [&](const void* p, size_t n) { out.insert(out.end(), (const char*)p, (const char*)p + n); };
    auto puti = [&](int v) { put(&v, sizeof v); };
    auto putvi = [&](const std::vector<int>& v) { puti((int)v.size()); put(v.data(), v.size() * sizeof(int)); };
    auto putvd = [&](const std::vector<double>& v) { puti((int)v.size()); put(v.data(), v.size() * sizeof(double)); };
    put("GTFB", 4);
    puti(1);
    puti(b->pure); puti(b->natoms); puti(b->charge);
    putvi(b->Z); putvd(b->xyz); putvi(b->L); putvi(b->nprim); putvi(b->poff); putvi(b->atom);
    putvd(b->alpha); putvd(b->coef);
    puti((int)b->guess.size());
    for (auto& g : b->guess) putvd(g);
    *buf = malloc(out.size());
    if (!*buf) return CINT_STATUS_ALLOC_FAILED;
    memcpy(*buf, out.data(), out.size());
    *bufsize = (int)out.size();
    return CINT_STATUS_SUCCESS;
}

CIntStatus_t CInt_unpackBasisSet(BasisSet_t b, void* buf)
{
    const char* p = (const char*)buf;
    if (!b || !p || memcmp(p, "GTFB", 4)) return CINT_STATUS_INVALID_VALUE;
    p += 4;
    auto geti = [&]() { int v; memcpy(&v, p, sizeof v); p += sizeof v; return v; };
    auto getvi = [&](std::vector<int>& v) { v.resize(geti()); memcpy(v.data(), p, v.size() * sizeof(int)); p += v.size() * sizeof(int); };
    auto getvd = [&](std::vector<double>& v) { v.resize(geti()); memcpy(v.data(), p, v.size() * sizeof(double)); p += v.size() * sizeof(double); };
    if (geti() != 1) return CINT_STATUS_INVALID_VALUE;
    *b = BasisSet();
    b->pure = geti(); b->natoms = geti(); b->charge = geti();
    getvi(b->Z); getvd(b->xyz); getvi(b->L); getvi(b->nprim); getvi(b->poff); getvi(b->atom);
    getvd(b->alpha); getvd(b->coef);
    b->guess.resize(geti());
    for (auto& g : b->guess) getvd(g);
    if ((int)b->guess.size() < kNElems) b->guess.resize(kNElems);
    for (int z : b->Z)
        if (z <= 0 || z >= kNElems) return CINT_STATUS_INVALID_VALUE;
    finalize(b);
    return CINT_STATUS_SUCCESS;
}

int CInt_getNumShells(BasisSet_t b) { return b->nshells; }
int CInt_getNumFuncs(BasisSet_t b) { return b->nbf; }
int CInt_getNumAtoms(BasisSet_t b) { return b->natoms; }
int CInt_getMaxShellDim(BasisSet_t b) { return b->maxdim; }
int CInt_getNumOccOrb(BasisSet_t b) { return (b->nelec_neutral - b->charge) / 2; }
int CInt_getFuncStartInd(BasisSet_t b, int s) { return b->fstart[s]; }
int CInt_getFuncEndInd(BasisSet_t b, int s) { return b->fstart[s + 1] - 1; }
int CInt_getShellDim(BasisSet_t b, int s) { return b->fstart[s + 1] - b->fstart[s]; }
int CInt_getAtomStartInd(BasisSet_t b, int a) { return b->s_startind[a]; }
int CInt_getTotalCharge(BasisSet_t b) { return b->charge; }
int CInt_getNneutral(BasisSet_t b) { return b->nelec_neutral; }

void CInt_getShellxyz(BasisSet_t b, int s, double* x, double* y, double* z)
{
    const double* c = &b->xyz[3 * b->atom[s]];
    *x = c[0];
    *y = c[1];
    *z = c[2];
}

double CInt_getNucEnergy(BasisSet_t b)
{
    double e = 0.0;
    for (int i = 0; i < b->natoms; i++)
        for (int j = 0; j < i; j++) {
            double dx = b->xyz[3 * i] - b->xyz[3 * j], dy = b->xyz[3 * i + 1] - b->xyz[3 * j + 1];
            double dz = b->xyz[3 * i + 2] - b->xyz[3 * j + 2];
            e += b->Z[i] * b->Z[j] / std::sqrt(dx * dx + dy * dy + dz * dz);
        }
    return e;
}

// all-zero block handed out for atoms without a guess (read only)
static const int kMaxGuessDim = 256;
static const double kZeroGuess[kMaxGuessDim * kMaxGuessDim] = {0.0};

void CInt_getInitialGuess(BasisSet_t b, int atomid, double** guess, int* spos, int* epos)
{
    int s0 = b->s_startind[atomid], s1 = b->s_startind[atomid + 1];
    *spos = b->fstart[s0];
    *epos = b->fstart[s1] - 1;
    const int nfa = *epos - *spos + 1;
    const int z = b->Z[atomid];
    // read only: several threads may hold one basis
    if (z > 0 && z < (int)b->guess.size() && (int)b->guess[z].size() == nfa * nfa)
        *guess = b->guess[z].data();
    else
        *guess = nfa <= kMaxGuessDim ? const_cast<double*>(kZeroGuess) : nullptr;
}

}  // extern "C"

// Boys table: F_m(i * delta) for m = 0..ncol-1 from the series
//   F_m(T) = exp(-T) sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1))
// for the top order, then downward recursion, in long double.
// the compact Taylor tables of the ERI kernels (boys_eval8): for order M and
// grid point i, coefficient k is F_{M+k}(i * BOYS8_DELTA) / k!
void gtf_boys_taylor_tables(double* t8)
{
    const int ncol = BOYS8_NM + BOYS8_NTERM - 1;
    std::vector<double> F((size_t)BOYS8_NGRID * ncol);
    gtf_boys_table(F.data(), BOYS8_NGRID, ncol, BOYS8_DELTA);
    for (int M = 0; M < BOYS8_NM; M++)
        for (int i = 0; i < BOYS8_NGRID; i++) {
            double fact = 1.0;
            for (int k = 0; k < BOYS8_NTERM; k++) {
                if (k > 0) fact *= k;
                const double c = F[(size_t)i * ncol + M + k] / fact;
                const size_t at = (size_t)M * BOYS8_SLAB + (size_t)i * BOYS8_NTERM + k;
                t8[at] = c;
            }
        }
}

void gtf_boys_table(double* tab, int ngrid, int ncol, double delta)
{
    for (int i = 0; i < ngrid; i++) {
        long double T = (long double)i * delta;
        int mmax = ncol - 1;
        long double term = 1.0L / (2 * mmax + 1), sum = term;
        for (int k = 1; k < 2000; k++) {
            term *= 2.0L * T / (2 * mmax + 2 * k + 1);
            sum += term;
            if (term < 1e-22L * sum) break;
        }
        long double e = expl(-T);
        long double F[64];
        F[mmax] = e * sum;
        for (int m = mmax - 1; m >= 0; m--) F[m] = (2.0L * T * F[m + 1] + e) / (2 * m + 1);
        for (int m = 0; m < ncol; m++) tab[(size_t)i * ncol + m] = (double)F[m];
    }
}
