// oneel.hip -- overlap, kinetic and nuclear-attraction matrices by the
// McMurchie-Davidson scheme (one thread per ordered shell pair; setup work,
// replaces pfock/one_electron.c:28-119 and the closed OED library):
//   S_ab = E^x_0 E^y_0 E^z_0 (pi/p)^(3/2)
//   T_ab = sum over directions of  -2b^2 S_{i,j+2} + b(2j+1) S_{ij} - j(j-1)/2 S_{i,j-2}
//   V_ab = -sum_C Z_C (2 pi / p) sum_tuv E^{ab}_{tuv} R_{tuv}(p, P - C)
#include "boys.h"
#include "gtf_internal.h"
#include "oneel.h"

#define LKIN (GTF_LMAX + 2)

// S, T, V of ordered shell pair pid = A * nshells + B of one molecule
__device__ __forceinline__ void oneel_pair(const BasisView& b, const double* __restrict__ boys, int ld,
                                           double* __restrict__ S, double* __restrict__ Tk, double* __restrict__ V,
                                           int pid)
{
    if (pid >= b.nshells * b.nshells) return;
    const int A = pid / b.nshells, B = pid % b.nshells;
    const int la = b.L[A], lb = b.L[B], lab = la + lb;
    const double* CA = b.center + 3 * A;
    const double* CB = b.center + 3 * B;
    double AB2 = 0.0;
    for (int d = 0; d < 3; d++) AB2 += (CA[d] - CB[d]) * (CA[d] - CB[d]);
    const int nca = ncart(la), ncb = ncart(lb);
    double sc[36], tc[36], vc[36];
    for (int k = 0; k < 36; k++) sc[k] = tc[k] = vc[k] = 0.0;
    const int ca0 = la == 0 ? 0 : nherm(la - 1), cb0 = lb == 0 ? 0 : nherm(lb - 1);
    for (int i = 0; i < b.nprim[A]; i++)
        for (int j = 0; j < b.nprim[B]; j++) {
            const double a = b.alpha[b.poff[A] + i], bb = b.alpha[b.poff[B] + j];
            const double p = a + bb, mu = a * bb / p;
            const double cc = b.coef[b.poff[A] + i] * b.coef[b.poff[B] + j];
            double P[3];
            // E1[d][x][y][t] for x <= la, y <= lb + 2
            double E1[3][GTF_LMAX + 1][LKIN + 1][GTF_LMAX + LKIN + 2];
            for (int d = 0; d < 3; d++) {
                P[d] = (a * CA[d] + bb * CB[d]) / p;
                const double XPA = P[d] - CA[d], XPB = P[d] - CB[d];
                for (int x = 0; x <= GTF_LMAX; x++)
                    for (int y = 0; y <= LKIN; y++)
                        for (int t = 0; t < GTF_LMAX + LKIN + 2; t++) E1[d][x][y][t] = 0.0;
                E1[d][0][0][0] = exp(-mu * (CA[d] - CB[d]) * (CA[d] - CB[d]));
                for (int x = 0; x < la; x++)
                    for (int t = 0; t <= x + 1; t++) {
                        double v = XPA * E1[d][x][0][t] + (t + 1) * E1[d][x][0][t + 1];
                        if (t > 0) v += E1[d][x][0][t - 1] / (2.0 * p);
                        E1[d][x + 1][0][t] = v;
                    }
                for (int y = 0; y < lb + 2; y++)
                    for (int x = 0; x <= la; x++)
                        for (int t = 0; t <= x + y + 1; t++) {
                            double v = XPB * E1[d][x][y][t] + (t + 1) * E1[d][x][y][t + 1];
                            if (t > 0) v += E1[d][x][y][t - 1] / (2.0 * p);
                            E1[d][x][y + 1][t] = v;
                        }
            }
            const double sq = sqrt(M_PI / p);
            // Hermite Coulomb R_{tuv}(p, P - C) for every nucleus, summed with -Z_C
            double Rsum[nherm(2 * GTF_LMAX)];
            for (int h = 0; h < nherm(lab); h++) Rsum[h] = 0.0;
            for (int c = 0; c < b.natoms; c++) {
                const double X = P[0] - b.atom_xyz[3 * c], Y = P[1] - b.atom_xyz[3 * c + 1],
                             Z = P[2] - b.atom_xyz[3 * c + 2];
                double F[2 * GTF_LMAX + 1];
                boys_eval_rt(lab, p * (X * X + Y * Y + Z * Z), F, boys);
                // R^n_{tuv} table, n = 0..lab
                double R[2 * GTF_LMAX + 1][nherm(2 * GTF_LMAX)];
                double pw = 1.0;
                for (int n = 0; n <= lab; n++) {
                    R[n][0] = pw * F[n];
                    pw *= -2.0 * p;
                }
                for (int k = 1; k <= lab; k++)
                    for (int n = 0; n <= lab - k; n++)
                        for (int h = nherm(k - 1); h < nherm(k); h++) {
                            const int t = herm_t(h), u = herm_u(h), v = herm_v(h);
                            double r;
                            if (t > 0) {
                                r = X * R[n + 1][hidx(t - 1, u, v)];
                                if (t > 1) r += (t - 1) * R[n + 1][hidx(t - 2, u, v)];
                            } else if (u > 0) {
                                r = Y * R[n + 1][hidx(t, u - 1, v)];
                                if (u > 1) r += (u - 1) * R[n + 1][hidx(t, u - 2, v)];
                            } else {
                                r = Z * R[n + 1][hidx(t, u, v - 1)];
                                if (v > 1) r += (v - 1) * R[n + 1][hidx(t, u, v - 2)];
                            }
                            R[n][h] = r;
                        }
                for (int h = 0; h < nherm(lab); h++) Rsum[h] -= b.atom_Z[c] * R[0][h];
            }
            for (int ia = 0; ia < nca; ia++) {
                const int ax = herm_t(ca0 + ia), ay = herm_u(ca0 + ia), az = herm_v(ca0 + ia);
                for (int ib = 0; ib < ncb; ib++) {
                    const int bx = herm_t(cb0 + ib), by = herm_u(cb0 + ib), bz = herm_v(cb0 + ib);
                    const int ax_[3] = {ax, ay, az}, bx_[3] = {bx, by, bz};
                    double s1[3], t1[3];
                    for (int d = 0; d < 3; d++) {
                        const int ii = ax_[d], jj = bx_[d];
                        s1[d] = E1[d][ii][jj][0] * sq;
                        double t = -2.0 * bb * bb * E1[d][ii][jj + 2][0] * sq + bb * (2 * jj + 1) * s1[d];
                        if (jj >= 2) t -= 0.5 * jj * (jj - 1) * E1[d][ii][jj - 2][0] * sq;
                        t1[d] = t;
                    }
                    sc[ia * ncb + ib] += cc * s1[0] * s1[1] * s1[2];
                    tc[ia * ncb + ib] += cc * (t1[0] * s1[1] * s1[2] + s1[0] * t1[1] * s1[2] + s1[0] * s1[1] * t1[2]);
                    double v = 0.0;
                    for (int t = 0; t <= ax + bx; t++)
                        for (int u = 0; u <= ay + by; u++)
                            for (int w = 0; w <= az + bz; w++)
                                v += E1[0][ax][bx][t] * E1[1][ay][by][u] * E1[2][az][bz][w] * Rsum[hidx(t, u, w)];
                    vc[ia * ncb + ib] += cc * 2.0 * M_PI / p * v;
                }
            }
        }
    const int na = nsph(la), nb = nsph(lb);
    const int r0 = b.fstart[A], c0 = b.fstart[B];
    for (int as = 0; as < na; as++)
        for (int bs = 0; bs < nb; bs++) {
            double s = 0.0, t = 0.0, v = 0.0;
            for (int ia = 0; ia < nca; ia++) {
                const double ta = sph_coef(la, as, ia);
                if (ta == 0.0) continue;
                for (int ib = 0; ib < ncb; ib++) {
                    const double w = ta * sph_coef(lb, bs, ib);
                    s += w * sc[ia * ncb + ib];
                    t += w * tc[ia * ncb + ib];
                    v += w * vc[ia * ncb + ib];
                }
            }
            const long k = (long)(r0 + as) * ld + c0 + bs;
            S[k] = s;
            Tk[k] = t;
            V[k] = v;
        }
}

__global__ void oneel_kernel(BasisView b, const double* __restrict__ boys, int ld, double* __restrict__ S,
                             double* __restrict__ Tk, double* __restrict__ V)
{
    oneel_pair(b, boys, ld, S, Tk, V, blockIdx.x * blockDim.x + threadIdx.x);
}

// Many molecules in one launch: molecule k owns the 64-thread blocks
// [mol[k].blk0, mol[k + 1].blk0) (mol[nmol].blk0 = the grid), one thread per
// ordered shell pair, and views the shared basis arrays from its own shell
// and atom offsets.  Each thread computes exactly what oneel_kernel does
// for that molecule alone.
__global__ __launch_bounds__(64) void oneel_batch_kernel(BasisView b, const double* __restrict__ boys,
                                                         const OneelMol* __restrict__ mol, int nmol,
                                                         double* __restrict__ S, double* __restrict__ Tk,
                                                         double* __restrict__ V)
{
    const int blk = blockIdx.x;
    int lo = 0, hi = nmol - 1;  // last k with mol[k].blk0 <= blk
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (mol[mid].blk0 <= blk) lo = mid;
        else hi = mid - 1;
    }
    const OneelMol m = mol[lo];
    const BasisView v{m.nshells,           b.nbf,          m.natoms,
                      b.L + m.shell_off,   b.nprim + m.shell_off, b.poff + m.shell_off,
                      b.fstart + m.shell_off, b.atom + m.shell_off, b.alpha,
                      b.coef,              b.center + 3 * m.shell_off, b.atom_xyz + 3 * m.atom_off,
                      b.atom_Z + m.atom_off};
    oneel_pair(v, boys, m.ld, S + m.moff, Tk + m.moff, V + m.moff, (blk - m.blk0) * 64 + threadIdx.x);
}

int gtf_oneel_batch(hipStream_t st, const BasisView& b, const double* boys, const OneelMol* d_mol, int nmol,
                    int nblocks, double* S, double* T, double* V)
{
    if (nmol <= 0 || nblocks <= 0) return 0;
    oneel_batch_kernel<<<nblocks, 64, 0, st>>>(b, boys, d_mol, nmol, S, T, V);
    GTF_CHECK(hipGetLastError());
    return 0;
}

int gtf_oneel(hipStream_t st, const BasisView& b, const double* boys, int ld, double* S, double* T, double* V)
{
    const int n = b.nshells * b.nshells;
    oneel_kernel<<<(n + 63) / 64, 64, 0, st>>>(b, boys, ld, S, T, V);
    GTF_CHECK(hipGetLastError());
    return 0;
}
