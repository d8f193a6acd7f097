// eri_md.h -- McMurchie-Davidson electron-repulsion integrals for one shell
// quartet class, fully unrolled at compile time.
//
//   (ab|cd) = 2 pi^(5/2) / (p q sqrt(p+q))
//             sum_{tuv} E^{ab}_{tuv} sum_{t'u'v'} (-1)^{t'+u'+v'} E^{cd}_{t'u'v'}
//             R_{t+t',u+u',v+v'}(alpha, P - Q),   alpha = p q / (p + q)
//
// (McMurchie and Davidson, J. Comput. Phys. 26, 218 (1978)).  The Hermite
// coefficients E are stored per primitive pair already contracted with the
// spherical transform, the contraction coefficients, exp(-mu |AB|^2) and 1/p,
// so a primitive quartet costs the Boys function, the R recursion and two
// small matrix products.  A thread owns NC consecutive bra components of one
// contracted quartet and all ket components.
#pragma once
#include "boys.h"
#include "gtf_internal.h"

#define GTF_TWO_PI_2_5 34.986836655249725693  // 2 pi^(5/2)

// R^{0}_{tuv} for t+u+v <= L, in place (stage n = L .. 0):
//   R^n_{000} = pref (-2 alpha)^n F_n(T)
//   R^n_{t+1,u,v} = t R^{n+1}_{t-1,u,v} + X R^{n+1}_{t,u,v}   (same for u/Y, v/Z)
template <int L, typename Real>
__device__ __forceinline__ void hermite_R(Real* R, const Real* F, Real pref, Real m2a, Real X, Real Y,
                                          Real Z)
{
    Real pw[L + 1];
    pw[0] = pref;
    static_for<1, L + 1>([&](auto n_) {
        constexpr int n = decltype(n_)::value;
        pw[n] = pw[n - 1] * m2a;
    });
    R[0] = pw[L] * F[L];
    static_for<0, L>([&](auto s_) {
        constexpr int n = L - 1 - decltype(s_)::value;
        constexpr int top = nherm(L - n);
        static_for<0, top - 1>([&](auto i_) {
            constexpr int h = top - 1 - decltype(i_)::value;
            constexpr int t = herm_t(h), u = herm_u(h), v = herm_v(h);
            if constexpr (t > 0) {
                Real r = X * R[hidx(t - 1, u, v)];
                if constexpr (t > 1) r += Real(t - 1) * R[hidx(t - 2, u, v)];
                R[h] = r;
            } else if constexpr (u > 0) {
                Real r = Y * R[hidx(t, u - 1, v)];
                if constexpr (u > 1) r += Real(u - 1) * R[hidx(t, u - 2, v)];
                R[h] = r;
            } else {
                Real r = Z * R[hidx(t, u, v - 1)];
                if constexpr (v > 1) r += Real(v - 1) * R[hidx(t, u, v - 2)];
                R[h] = r;
            }
        });
        R[0] = pw[n] * F[n];
    });
}

template <int h1, int h2>
__host__ __device__ constexpr int hsum()
{
    return hidx(herm_t(h1) + herm_t(h2), herm_u(h1) + herm_u(h2), herm_v(h1) + herm_v(h2));
}

template <int LA, int LB, int LC, int LD>
struct EriShape {
    static constexpr int LAB = LA + LB, LCD = LC + LD, L = LAB + LCD;
    static constexpr int HAB = nherm(LAB), HCD = nherm(LCD), HL = nherm(L);
    static constexpr int NA = nsph(LA), NB = nsph(LB), NCc = nsph(LC), ND = nsph(LD);
    static constexpr int NAB = NA * NB, NCD = NCc * ND;
    static constexpr int ESZ_AB = NAB * HAB, ESZ_CD = NCD * HCD;
};

// Register budget heuristic: the largest divisor NC of NAB whose live set
// R + NC*(Ebra + V + acc) stays within ~128 doubles.
template <int LA, int LB, int LC, int LD>
__host__ __device__ constexpr int default_nc()
{
    using S = EriShape<LA, LB, LC, LD>;
    int best = 1;
    for (int nc = 1; nc <= S::NAB; ++nc) {
        if (S::NAB % nc) continue;
        int reg = S::HL + nc * (S::HAB + S::HCD + S::NCD) + S::L + 8;
        if (reg <= 128) best = nc;
    }
    return best;
}

// classes whose kernels apply the primitive screening of eri_block (by
// total angular momentum L)
__host__ __device__ constexpr bool prim_screened_l(int L) { return L <= 3; }
template <int LA, int LB, int LC, int LD>
__host__ __device__ constexpr bool prim_screened()
{
    return prim_screened_l(LA + LB + LC + LD);
}

// Contracted integrals (bra components c0..c0+NC-1, all ket components) of
// the quartet (bra pair | ket pair), summed over the bra primitive pairs
// [i0, i1) (i1 < 0: all) and, in the classes of prim_screened, the
// primitive quartets whose Schwarz bound q_bra q_ket is at least pv.peps.  acc[c][cd], cd = c_sph * nsph(LD) + d_sph.
// The ket coefficients are read in the compact nonzero-only layout and the
// structurally zero terms are skipped at compile time; when a thread owns the
// whole bra pair (NC == NAB, c0 == 0) the bra side is treated the same way.
template <int LA, int LB, int LC, int LD, int NC, typename Real>
__device__ __forceinline__ void eri_block(const PairView& pv, int bra, int ket, int c0,
                                          const double* __restrict__ boys_slab,
                                          Real (&acc)[NC][EriShape<LA, LB, LC, LD>::NCD], int i0 = 0,
                                          int i1 = -1)
{
    using S = EriShape<LA, LB, LC, LD>;
    constexpr int L = S::L, HAB = S::HAB, HCD = S::HCD, HL = S::HL, NCD = S::NCD;
    constexpr bool BRA_FULL = (NC == S::NAB);
    using TA = NzTab<LA, LB>;
    using TC = NzTab<LC, LD>;
    constexpr int NZAB = TA::data.nnz, NZCD = TC::data.nnz;
    static_for<0, NC>([&](auto c_) {
        static_for<0, NCD>([&](auto d_) { acc[decltype(c_)::value][decltype(d_)::value] = Real(0); });
    });
    const int nb = pv.nprim[bra], nk = pv.nprim[ket];
    const double4* __restrict__ pb = pv.prim + pv.poff[bra];
    const double4* __restrict__ pk = pv.prim + pv.poff[ket];
    const double* __restrict__ Eb = BRA_FULL ? pv.Ez + pv.zoff[bra] : pv.E + pv.eoff[bra] + (long)c0 * HAB;
    const double* __restrict__ Ek0 = pv.Ez + pv.zoff[ket];
    // primitive screening (classes of total angular momentum <= 3, whose
    // s and p shells carry the long contractions; the higher classes would
    // pay for it in registers): a pair's primitives are stored by
    // descending Schwarz bound, so the kets of bra primitive i that pass
    // q_i q_j >= peps are a prefix, found by bisection once per i, and once
    // one bra primitive has none the later ones have none either
    constexpr bool PSCREEN = prim_screened<LA, LB, LC, LD>();
    const int ie = i1 < 0 ? nb : i1;
    for (int i = i0; i < ie; ++i) {
        int nke = nk;
        if constexpr (PSCREEN) {
            if (pv.peps > 0.0) {
                const double qi = pv.q[pv.poff[bra] + i];
                const double* __restrict__ qk = pv.q + pv.poff[ket];
                int lo = 0, hi = nk;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (qi * qk[mid] >= pv.peps) lo = mid + 1;
                    else hi = mid;
                }
                nke = lo;
                if (nke == 0) break;
            }
        }
        const double4 B = pb[i];
        Real eb[NC][HAB];
        static_for<0, NC>([&](auto c_) {
            constexpr int c = decltype(c_)::value;
            static_for<0, HAB>([&](auto h_) {
                constexpr int h = decltype(h_)::value;
                if constexpr (BRA_FULL) {
                    if constexpr (TA::data.nz[c][h]) eb[c][h] = Real(Eb[(long)i * NZAB + TA::data.pos[c][h]]);
                } else {
                    eb[c][h] = Real(Eb[(long)i * S::ESZ_AB + c * HAB + h]);
                }
            });
        });
        for (int j = 0; j < nke; ++j) {
            const double4 K = pk[j];
            // T = alpha |PQ|^2 and F_0..F_L(T) in FP64 for both precisions;
            // an FP32 tile rounds them for the R recursion and contractions
            const double Xd = B.y - K.y, Yd = B.z - K.z, Zd = B.w - K.w;
            const double rd = rsqrt_nr(B.x + K.x);
            const double alphad = B.x * K.x * (rd * rd);
            const double Td = alphad * (Xd * Xd + Yd * Yd + Zd * Zd);
            double Fd[L + 1];
            boys_eval8<L>(Td, Fd, boys_slab);
            Real F[L + 1];
            static_for<0, L + 1>([&](auto m_) {
                constexpr int m = decltype(m_)::value;
                F[m] = Real(Fd[m]);
            });
            const Real pref = Real(GTF_TWO_PI_2_5 * rd);
            Real R[HL];
            hermite_R<L, Real>(R, F, pref, Real(-2.0 * alphad), Real(Xd), Real(Yd), Real(Zd));
            Real V[NC][HCD];
            static_for<0, HCD>([&](auto h2_) {
                constexpr int h2 = decltype(h2_)::value;
                constexpr bool neg = (horder(h2) & 1) != 0;
                static_for<0, NC>([&](auto c_) {
                    constexpr int c = decltype(c_)::value;
                    Real v = Real(0);
                    static_for<0, HAB>([&](auto h1_) {
                        constexpr int h1 = decltype(h1_)::value;
                        if constexpr (!BRA_FULL || TA::data.nz[c][h1])
                            v += eb[c][h1] * R[hsum<h1, h2>()];
                    });
                    V[c][h2] = neg ? -v : v;
                });
            });
            const double* __restrict__ ek = Ek0 + (long)j * NZCD;
            static_for<0, NCD>([&](auto cd_) {
                constexpr int cd = decltype(cd_)::value;
                static_for<0, HCD>([&](auto h2_) {
                    constexpr int h2 = decltype(h2_)::value;
                    if constexpr (TC::data.nz[cd][h2]) {
                        const Real e = Real(ek[TC::data.pos[cd][h2]]);
                        static_for<0, NC>([&](auto c_) {
                            constexpr int c = decltype(c_)::value;
                            acc[c][cd] += e * V[c][h2];
                        });
                    }
                });
            });
        }
    }
}
