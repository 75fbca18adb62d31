// Three- and two-centre Coulomb integrals of density fitting on gfx950.
//
// (ij|P) of orbital shells i, j and an auxiliary shell P, and (P|Q) (the same
// kernel with a unit s function, exponent 0 and factor 1, in the j slot).
// Rys quadrature with the recurrences of the four-centre kernels (jk_device.h)
// and a ket made of one function on centre C:
//   u = t^2 / (p + c), B00 = u/2, B10 = (1 - u c)/(2p), B01 = (1 - u p)/(2c)
//   C00 = (P - A) - u c (P - C), C'00 = u p (P - C)
// then the transfer I(a, b+1) = I(a+1, b) + (A - B) I(a, b) on the bra.
//
// Work split: one lane per (shell pair, auxiliary shell, Cartesian component
// of the auxiliary shell).  The component is uniform over a block (grid y),
// so its exponents (kx, ky, kz) are compile-time constants in the branch the
// block takes: the lane runs the vertical recurrence only to kx, ky, kz and
// keeps ncart(li) ncart(lj) accumulators.  The lane transforms i and j to
// spherical functions itself and writes one value per (i, j) function pair
// into a Cartesian-auxiliary staging matrix; a second kernel transforms the
// auxiliary index.  Within a block consecutive lanes take consecutive
// auxiliary shells of one angular momentum, whose staging columns are
// consecutive: the stores coalesce.
#pragma once
#include <hip/hip_runtime.h>

#include "c2s.h"
#include "jk_device.h"
#include "rys_eval.h"

namespace mjk {

constexpr int DF_LMAX_AO = 3;    // orbital shells up to f
constexpr int DF_LMAX_AUX = 4;   // auxiliary shells up to g

// Cartesian -> spherical coefficient (m, c) of angular momentum L: the
// compile-time C2S tables for L <= 3 (their zeros fold away), the l = 4 rows
// (9 x 15, from the closed formula on the host) from memory.  The s and p
// factors are folded into the primitive coefficients, as in the four-centre
// path.
template <int L>
__device__ __forceinline__ double df_c2s(const double* __restrict__ t4, int m, int c)
{
    if constexpr (L <= 3) return C2S[L][m][c];
    else return t4[m * 15 + c];
}

struct Int3cArgs {
    // orbital pairs of one (li, lj) class
    const double4* ppP;     // per primitive pair {Px, Py, Pz, p}
    const double* ppK;      // per primitive pair exp(-ab/p |AB|^2) ca cb
    const int* pprim0;      // first primitive pair of each pair
    const int* pnprim;
    const double4* pA;      // centre of the first shell {Ax, Ay, Az, 0}
    const double4* pAB;     // A - B
    const int* pi0;         // first spherical AO of the first shell
    const int* pj0;         // first spherical AO of the second shell (-1: unit function)
    int npair;
    // auxiliary shells of one angular momentum
    const double4* kC;      // centre
    const int* kprim0;      // first primitive
    const int* knprim;
    const double2* kprim;   // {exponent, coefficient}
    int naux;               // shells in this angular-momentum group
    long long col0;         // staging column of (component 0, shell 0) of the group
    // output
    double* out;            // staging rows: (i, j) function pair -> row i * ld_row + j
    long long ld_row;       // rows per i (nao; 0 for (P|Q): row = i)
    long long ld;           // row length (staging columns)
    const double* rys;
    const double* c2s4;     // l = 4 rows (9 x 15)
};

template <int LI, int LJ, int LK, int KX, int KY, int KZ, bool UNITJ>
__device__ __forceinline__ void int3c_lane(const Int3cArgs& a, int pair, int aux)
{
    constexpr int NI = LI + LJ + 1;
    constexpr int NR = (LI + LJ + LK) / 2 + 1;
    constexpr int NCI = ncart(LI), NCJ = ncart(LJ);
    constexpr double PI52 = 34.98683665524972;   // 2 pi^(5/2)
    double acc[NCI][NCJ];
#pragma unroll
    for (int i = 0; i < NCI; ++i)
#pragma unroll
        for (int j = 0; j < NCJ; ++j) acc[i][j] = 0.0;
    const double4 A = a.pA[pair], AB = a.pAB[pair], C = a.kC[aux];
    const int q0 = a.pprim0[pair], nq = a.pnprim[pair];
    const int k0 = a.kprim0[aux], nk = a.knprim[aux];
    for (int qi = 0; qi < nq; ++qi) {
        const double4 P = a.ppP[q0 + qi];
        const double Kab = a.ppK[q0 + qi];
        const double p = P.w;
        const double PCx = P.x - C.x, PCy = P.y - C.y, PCz = P.z - C.z;
        const double PC2 = PCx * PCx + PCy * PCy + PCz * PCz;
        const double PAx = P.x - A.x, PAy = P.y - A.y, PAz = P.z - A.z;
        for (int ki = 0; ki < nk; ++ki) {
            const double2 e = a.kprim[k0 + ki];
            const double c = e.x;
            const double pc = p + c;
            const double ipc = 1.0 / pc;
            const double T = p * c * ipc * PC2;
            const double fac = PI52 / (p * c * sqrt(pc)) * Kab * e.y;
            RysCursor<NR> rc(a.rys, T);
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                double x, w;
                rc.root(r, x, w);
                const double u = x * ipc;
                const double B00 = 0.5 * u, B10 = (1.0 - u * c) / (2.0 * p), B01 = (1.0 - u * p) / (2.0 * c);
                double gx[NI][KX + 1], gy[NI][KY + 1], gz[NI][KZ + 1];
                vrr<NI, KX + 1>(gx, 1.0, PAx - u * c * PCx, u * p * PCx, B10, B01, B00);
                vrr<NI, KY + 1>(gy, 1.0, PAy - u * c * PCy, u * p * PCy, B10, B01, B00);
                vrr<NI, KZ + 1>(gz, w * fac, PAz - u * c * PCz, u * p * PCz, B10, B01, B00);
                // transfer to (a, b) at the lane's k
                double hx[NI][LJ + 1], hy[NI][LJ + 1], hz[NI][LJ + 1];
#pragma unroll
                for (int i = 0; i < NI; ++i) {
                    hx[i][0] = gx[i][KX];
                    hy[i][0] = gy[i][KY];
                    hz[i][0] = gz[i][KZ];
                }
#pragma unroll
                for (int b = 1; b <= LJ; ++b)
#pragma unroll
                    for (int i = 0; i < NI - b; ++i) {
                        hx[i][b] = hx[i + 1][b - 1] + AB.x * hx[i][b - 1];
                        hy[i][b] = hy[i + 1][b - 1] + AB.y * hy[i][b - 1];
                        hz[i][b] = hz[i + 1][b - 1] + AB.z * hz[i][b - 1];
                    }
#pragma unroll
                for (int ci = 0; ci < NCI; ++ci)
#pragma unroll
                    for (int cj = 0; cj < NCJ; ++cj)
                        acc[ci][cj] += hx[cart_x(LI, ci)][cart_x(LJ, cj)] * hy[cart_y(LI, ci)][cart_y(LJ, cj)] *
                                       hz[cart_z(LI, ci)][cart_z(LJ, cj)];
            }
        }
    }
    // spherical i and j, store
    constexpr int NSI = 2 * LI + 1, NSJ = 2 * LJ + 1;
    const int i0 = a.pi0[pair], j0 = a.pj0[pair];
    constexpr int CK = (LK - KX) * (LK - KX + 1) / 2 + KZ;   // component index of (KX, KY, KZ)
    const long long column = a.col0 + (long long)CK * a.naux + aux;
#pragma unroll
    for (int mi = 0; mi < NSI; ++mi) {
#pragma unroll
        for (int mj = 0; mj < NSJ; ++mj) {
            double v = 0.0;
#pragma unroll
            for (int ci = 0; ci < NCI; ++ci) {
                const double ti = df_c2s<LI>(a.c2s4, mi, ci);
                double s = 0.0;
#pragma unroll
                for (int cj = 0; cj < NCJ; ++cj) {
                    const double tj = UNITJ ? 1.0 : df_c2s<LJ>(a.c2s4, mj, cj);
                    s += tj * acc[ci][cj];
                }
                v += ti * s;
            }
            if (UNITJ) {
                a.out[(long long)(i0 + mi) * a.ld + column] = v;
            } else {
                a.out[((long long)(i0 + mi) * a.ld_row + (j0 + mj)) * a.ld + column] = v;
                // a pair of two different shells also fills the transposed rows
                if (i0 != j0) a.out[((long long)(j0 + mj) * a.ld_row + (i0 + mi)) * a.ld + column] = v;
            }
        }
    }
}

// component CK of the auxiliary shell, chosen by the block (uniform branch)
template <int LI, int LJ, int LK, bool UNITJ, int CK>
__device__ __forceinline__ void int3c_dispatch(const Int3cArgs& a, int pair, int aux, int ck)
{
    if constexpr (CK < ncart(LK)) {
        if (ck == CK) {
            int3c_lane<LI, LJ, LK, cart_x(LK, CK), cart_y(LK, CK), cart_z(LK, CK), UNITJ>(a, pair, aux);
            return;
        }
        int3c_dispatch<LI, LJ, LK, UNITJ, CK + 1>(a, pair, aux, ck);
    }
}

// grid: x over (pair, auxiliary shell) tasks, auxiliary shell fastest;
// y = Cartesian component of the auxiliary shell
template <int LI, int LJ, int LK, bool UNITJ>
__global__ void __launch_bounds__(128) k_int3c(Int3cArgs a)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)a.npair * a.naux) return;
    const int pair = (int)(t / a.naux), aux = (int)(t % a.naux);
    int3c_dispatch<LI, LJ, LK, UNITJ, 0>(a, pair, aux, (int)blockIdx.y);
}

template <int LI, int LJ, int LK, bool UNITJ>
inline int launch_int3c_t(const Int3cArgs& a, hipStream_t st)
{
    const long long n = (long long)a.npair * a.naux;
    if (n == 0) return 0;
    const dim3 grid((unsigned)((n + 127) / 128), (unsigned)ncart(LK));
    hipLaunchKernelGGL((k_int3c<LI, LJ, LK, UNITJ>), grid, dim3(128), 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// defined in the generated df_int3c_l<LK>.hip units
int launch_int3c(int li, int lj, int lk, bool unitj, const Int3cArgs& a, hipStream_t st);

}  // namespace mjk

