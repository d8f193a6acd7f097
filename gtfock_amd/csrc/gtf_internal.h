// gtf_internal.h -- shared definitions of the HIP Fock engine (host + device).
//
// Angular momentum up to d (L <= 2) per shell, spherical functions.
// Shell pairs are stored canonically with L(s1) >= L(s2); a pair type is
//   0 ss, 1 ps, 2 pp, 3 ds, 4 dp, 5 dd
// and a quartet class (bra type A >= ket type B) is A*(A+1)/2 + B (21 classes).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <type_traits>
#include <utility>

#define GTF_LMAX 2
#define GTF_NPAIRTYPES 6
#define GTF_NCLASSES 21
// ket-list block re-sorted by primitive count (one wavefront)
#define GTF_KET_BLOCK 64

// Boys-function grid: F_m(T0) for T0 = i * BOYS_DELTA, i = 0..BOYS_NGRID-1,
// m = 0..BOYS_NCOL-1; T >= BOYS_TMAX uses the asymptotic form.
#define BOYS_DELTA 0.05
#define BOYS_INV_DELTA 20.0
#define BOYS_NGRID 721
#define BOYS_NCOL 16
#define BOYS_TMAX 36.0
#define BOYS_NTAYLOR 7
// compact Taylor table of the ERI kernels: orders M = 0 .. BOYS8_NM - 1,
// T0 = i * BOYS8_DELTA for i < BOYS8_NGRID, BOYS8_NTERM coefficients per row
#define BOYS8_DELTA 0.1
#define BOYS8_INV_DELTA 10.0
#define BOYS8_NGRID 361
#define BOYS8_NTERM 8
#define BOYS8_NM 9
#define BOYS8_SLAB (BOYS8_NGRID * BOYS8_NTERM)

#define GTF_CHECK(call)                                                                    \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            fprintf(stderr, "gtfock_amd: HIP error %s at %s:%d\n", hipGetErrorString(e_), \
                    __FILE__, __LINE__);                                                   \
            return -1;                                                                     \
        }                                                                                  \
    } while (0)

__host__ __device__ constexpr int nherm(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }
__host__ __device__ constexpr int nsph(int l) { return 2 * l + 1; }
__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
__host__ __device__ constexpr int pair_type(int la, int lb) { return la * (la + 1) / 2 + lb; }
__host__ __device__ constexpr int class_index(int A, int B) { return A * (A + 1) / 2 + B; }

// Hermite / cartesian component index: order k = t+u+v, then t descending,
// then v ascending.  The same ordering serves cartesian components
// (lx, ly, lz) of one shell when restricted to order k.
__host__ __device__ constexpr int hidx(int t, int u, int v)
{
    return (t + u + v) * (t + u + v + 1) * (t + u + v + 2) / 6 +
           ((u + v) * (u + v + 1)) / 2 + v;
}
__host__ __device__ constexpr int horder(int h)
{
    int k = 0;
    while (nherm(k) <= h) ++k;
    return k;
}
__host__ __device__ constexpr int herm_t(int h)
{
    int k = horder(h);
    int r = h - (k == 0 ? 0 : nherm(k - 1));
    int s = 0;  // s = u + v
    while ((s + 1) * (s + 2) / 2 <= r) ++s;
    return k - s;
}
__host__ __device__ constexpr int herm_v(int h)
{
    int k = horder(h);
    int r = h - (k == 0 ? 0 : nherm(k - 1));
    int s = 0;
    while ((s + 1) * (s + 2) / 2 <= r) ++s;
    return r - s * (s + 1) / 2;
}
__host__ __device__ constexpr int herm_u(int h) { return horder(h) - herm_t(h) - herm_v(h); }

// 1/sqrt(s) from the hardware estimate refined by two Newton steps
// y <- y (3 - s y^2) / 2 (relative error ~1e-16 in FP64)
__device__ __forceinline__ double rsqrt_nr(double s)
{
    double y = __builtin_amdgcn_rsq(s);
    double h = s * y;
    y = fma(y, fma(-0.5 * h, y, 0.5), y);
    h = s * y;
    y = fma(y, fma(-0.5 * h, y, 0.5), y);
    return y;
}
__device__ __forceinline__ float rsqrt_nr(float s)
{
    float y = __builtin_amdgcn_rsqf(s);
    y = fmaf(y, fmaf(-0.5f * s * y, y, 0.5f), y);
    return y;
}

// compile-time loop: f(std::integral_constant<int, i>) for i in [B, E)
template <int B, int E, typename F>
__host__ __device__ __forceinline__ void static_for(F&& f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// Device views of the engine's resident data.
struct PairView {
    const int* s1;       // canonical first shell
    const int* s2;       // canonical second shell
    const int* nprim;    // primitive pairs
    const int* poff;     // offset into prim[]
    const long* eoff;    // offset into E[]
    const double4* prim; // (p, Px, Py, Pz) per primitive pair
    const double* E;     // Hermite->spherical coefficients, scaled by c_i c_j K_ab / p
    const long* zoff;    // offset into Ez[]
    const double* Ez;    // the same, structural nonzeros only (herm_nz_pos order)
    const double* q;     // Schwarz bound per primitive pair (a pair's primitives in descending order)
    double peps;         // primitive quartets with q_bra q_ket < peps are skipped (0: none)
};

struct BasisView {
    int nshells, nbf, natoms;
    const int* L;
    const int* nprim;
    const int* poff;
    const int* fstart;
    const int* atom;
    const double* alpha;
    const double* coef;
    const double* center;    // 3 per shell (bohr)
    const double* atom_xyz;  // 3 per atom
    const double* atom_Z;
};

// Spherical transform rows (m = -l..l) over cartesian components of order l
// in the hidx order (xx, xy, xz, yy, yz, zz for l = 2).
__host__ __device__ constexpr double sph_coef(int l, int m, int c)
{
    if (l == 0) return 1.0;
    if (l == 1) return m == c ? 1.0 : 0.0;
    const double r3 = 1.7320508075688772;
    switch (m) {
    case 0: return c == 1 ? r3 : 0.0;                              // xy
    case 1: return c == 4 ? r3 : 0.0;                              // yz
    case 2: return c == 5 ? 1.0 : ((c == 0 || c == 3) ? -0.5 : 0.0); // z^2 - r^2/2
    case 3: return c == 2 ? r3 : 0.0;                              // xz
    case 4: return c == 0 ? 0.5 * r3 : (c == 3 ? -0.5 * r3 : 0.0); // (x^2 - y^2)
    }
    return 0.0;
}

// Structural nonzeros of the Hermite coefficients of a spherical component
// pair (as, bs) of shells (la, lb): E^{ab}_{tuv} can be nonzero only if some
// cartesian pair (a, b) in the two spherical functions has t <= ax + bx,
// u <= ay + by, v <= az + bz.
__host__ __device__ constexpr bool herm_nz(int la, int lb, int as, int bs, int h)
{
    const int t = herm_t(h), u = herm_u(h), v = herm_v(h);
    const int ca0 = la == 0 ? 0 : nherm(la - 1), cb0 = lb == 0 ? 0 : nherm(lb - 1);
    for (int ca = 0; ca < ncart(la); ca++) {
        if (sph_coef(la, as, ca) == 0.0) continue;
        for (int cb = 0; cb < ncart(lb); cb++) {
            if (sph_coef(lb, bs, cb) == 0.0) continue;
            if (t <= herm_t(ca0 + ca) + herm_t(cb0 + cb) && u <= herm_u(ca0 + ca) + herm_u(cb0 + cb) &&
                v <= herm_v(ca0 + ca) + herm_v(cb0 + cb))
                return true;
        }
    }
    return false;
}
// The structural-nonzero table of one pair type, evaluated once at compile
// time: nz[ab][h], pos[ab][h] (position in the compact layout), nnz.
template <int LA, int LB>
struct NzTab {
    static constexpr int NAB = nsph(LA) * nsph(LB), H = nherm(LA + LB);
    struct Data {
        bool nz[NAB][H];
        int pos[NAB][H];
        int nnz;
    };
    static constexpr Data make()
    {
        Data d{};
        int n = 0;
        for (int ab = 0; ab < NAB; ab++)
            for (int h = 0; h < H; h++) {
                d.nz[ab][h] = herm_nz(LA, LB, ab / nsph(LB), ab % nsph(LB), h);
                d.pos[ab][h] = n;
                if (d.nz[ab][h]) n++;
            }
        d.nnz = n;
        return d;
    }
    static constexpr Data data = make();
};

// number of structural nonzeros of a (la, lb) pair block (runtime form), and
// the position of (component ab, hermite h) in the compact layout
__host__ __device__ constexpr int herm_nnz(int la, int lb)
{
    int n = 0;
    for (int ab = 0; ab < nsph(la) * nsph(lb); ab++)
        for (int h = 0; h < nherm(la + lb); h++)
            if (herm_nz(la, lb, ab / nsph(lb), ab % nsph(lb), h)) n++;
    return n;
}
// herm_nnz read from the compile-time tables (the runtime form takes
// thousands of steps per call; set-up code calls it once per shell pair)
inline int herm_nnz_lookup(int la, int lb)
{
    static_assert(GTF_LMAX == 2, "table below lists l = 0..2");
    static constexpr int t[3][3] = {{NzTab<0, 0>::data.nnz, NzTab<0, 1>::data.nnz, NzTab<0, 2>::data.nnz},
                                    {NzTab<1, 0>::data.nnz, NzTab<1, 1>::data.nnz, NzTab<1, 2>::data.nnz},
                                    {NzTab<2, 0>::data.nnz, NzTab<2, 1>::data.nnz, NzTab<2, 2>::data.nnz}};
    return t[la][lb];
}
__host__ __device__ constexpr int herm_nz_pos(int la, int lb, int ab, int h)
{
    int n = 0;
    for (int x = 0; x < nsph(la) * nsph(lb); x++)
        for (int y = 0; y < nherm(la + lb); y++) {
            if (x == ab && y == h) return n;
            if (herm_nz(la, lb, x / nsph(lb), x % nsph(lb), y)) n++;
        }
    return n;
}
