// Rys-quadrature ERI + J/K digestion kernels for gfx950 (MI355X).
//
// One templated kernel per (la lb | lc ld) quartet class.  A wavefront of 64
// lanes owns one 8 x 8 tile of (bra pair, ket pair) tasks; each lane
// evaluates one contracted shell quartet in registers:
//   primitive loop -> Rys roots (table) -> 2D integrals (VRR, then HRR to
//   move angular momentum onto B and D) -> 6D assembly over Cartesian
//   components -> contracted block in registers -> J/K digestion
//   (register-level partial sums, FP64 atomics into the Cartesian J'/K').
//
// Recurrences (Rys, Dupuis, King, J. Comput. Chem. 4, 154 (1983); Lindh,
// Ryu, Liu, J. Chem. Phys. 95, 5889 (1991) for the transfer step), with
// u = t^2/(p+q):
//   B00 = u/2, B10 = (1 - u q)/(2p), B01 = (1 - u p)/(2q)
//   C00 = (P - A) - u q (P - Q), C'00 = (Q - C) + u p (P - Q)
//   g(i+1,k) = C00 g(i,k) + i B10 g(i-1,k) + k B00 g(i,k-1)
//   g(i,k+1) = C'00 g(i,k) + k B01 g(i,k-1) + i B00 g(i-1,k)
//   I(a,b+1) = I(a+1,b) + (A-B) I(a,b)   (and the same for the ket)
//
// Digestion: for a unique quartet with degeneracy factor f (1/2 per
// diagonal bra pair, ket pair and bra==ket), with Ds = D + D^T,
//   J'[c,d] += f sum_ab v Ds[a,b],  J'[a,b] += f sum_cd v Ds[c,d]
//   K'[a,d] += f sum_bc v D[b,c],   K'[b,d] += f sum_ac v D[a,c]
//   K'[a,c] += f sum_bd v D[b,d],   K'[b,c] += f sum_ad v D[a,d]
// and J = J' + J'^T, K = K'(D) + K'(D^T)^T (K'(D) + K'(D)^T when D is
// symmetric).  This covers all 8 index permutations of (ab|cd)
// (reference semantics: pyscf/lib/vhf/nr_direct_dot.c:1293-1647).
#include <hip/hip_runtime.h>
#include "jk_common.h"
#include "rys_eval.h"

namespace mjk {

// Cartesian exponents in the order lx = l..0, ly = l-lx..0.
__host__ __device__ constexpr int cart_x(int l, int i)
{
    int lx = l, k = i;
    while (k > l - lx) { k -= l - lx + 1; --lx; }
    return lx;
}
__host__ __device__ constexpr int cart_y(int l, int i)
{
    int lx = l, k = i;
    while (k > l - lx) { k -= l - lx + 1; --lx; }
    return l - lx - k;
}
__host__ __device__ constexpr int cart_z(int l, int i) { return l - cart_x(l, i) - cart_y(l, i); }

// Vertical recurrence for one Cartesian direction and one root.
template <int NI, int NK>
__device__ __forceinline__ void vrr(double (&g)[NI][NK], double g00, double C, double Cp,
                                    double B10, double B01, double B00)
{
    g[0][0] = g00;
    if constexpr (NI > 1) {
        g[1][0] = C * g00;
#pragma unroll
        for (int i = 1; i < NI - 1; ++i) g[i + 1][0] = C * g[i][0] + (i * B10) * g[i - 1][0];
    }
#pragma unroll
    for (int k = 0; k < NK - 1; ++k) {
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            double v = Cp * g[i][k];
            if (k > 0) v += (k * B01) * g[i][k - 1];
            if (i > 0) v += (i * B00) * g[i - 1][k];
            g[i][k + 1] = v;
        }
    }
}

// Horizontal (transfer) recurrence: g(a+b, c+d) -> I(a, b, c, d).
template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void hrr(const double (&g)[LA + LB + 1][LC + LD + 1], double AB, double CD,
                                    double (&I)[LA + 1][LB + 1][LC + 1][LD + 1])
{
    constexpr int NI = LA + LB + 1, NK = LC + LD + 1;
    double h[NI][LB + 1][NK];
#pragma unroll
    for (int a = 0; a < NI; ++a)
#pragma unroll
        for (int k = 0; k < NK; ++k) h[a][0][k] = g[a][k];
#pragma unroll
    for (int b = 1; b <= LB; ++b)
#pragma unroll
        for (int a = 0; a < NI - b; ++a)
#pragma unroll
            for (int k = 0; k < NK; ++k) h[a][b][k] = h[a + 1][b - 1][k] + AB * h[a][b - 1][k];
#pragma unroll
    for (int a = 0; a <= LA; ++a)
#pragma unroll
        for (int b = 0; b <= LB; ++b) {
            double e[NK][LD + 1];
#pragma unroll
            for (int k = 0; k < NK; ++k) e[k][0] = h[a][b][k];
#pragma unroll
            for (int d = 1; d <= LD; ++d)
#pragma unroll
                for (int k = 0; k < NK - d; ++k) e[k][d] = e[k + 1][d - 1] + CD * e[k][d - 1];
#pragma unroll
            for (int c = 0; c <= LC; ++c)
#pragma unroll
                for (int d = 0; d <= LD; ++d) I[a][b][c][d] = e[c][d];
        }
}

__device__ __forceinline__ void atomic_add(double* p, double v)
{
    unsafeAtomicAdd(p, v);
}

// Locate the bra tile of this block: largest i with bt_off[i] <= b.
__device__ __forceinline__ int find_tile(const int* __restrict__ off, int n, int b)
{
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= b) lo = mid; else hi = mid;
    }
    return lo;
}

// Reference prescreen predicate (pyscf/lib/vhf/optimizer.c:90-117) on the
// shells of this quartet; symmetric under i<->j, k<->l and bra<->ket.
__device__ __forceinline__ bool keep_quartet(const JKArgs& a, int mol, int p, int q, double qpq)
{
    if (qpq <= a.cutoff) return false;
    const double* __restrict__ dc = a.dm_cond + a.M.cbase[mol];
    const int ns = a.M.nsh[mol];
    const int i = a.P.shA[p], j = a.P.shB[p], k = a.P.shA[q], l = a.P.shB[q];
    double d = 4.0 * fmax(dc[j * ns + i], dc[l * ns + k]);
    d = fmax(d, fmax(dc[j * ns + k], dc[j * ns + l]));
    d = fmax(d, fmax(dc[i * ns + k], dc[i * ns + l]));
    return d * qpq > a.cutoff;
}

// Wave-level tally of evaluated quartets and primitive quartets (one atomic
// per wave and counter).
__device__ __forceinline__ void count_work(unsigned long long* c, bool alive, int nprimq)
{
    const unsigned long long m = __ballot(alive);
    int np = alive ? nprimq : 0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) np += __shfl_xor(np, off);
    if (threadIdx.x == 0 && c) {
        atomicAdd(c, (unsigned long long)__popcll(m));
        atomicAdd(c + 1, (unsigned long long)np);
    }
}

// ---------------------------------------------------------------------------
// One lane per shell quartet.
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(64) void k_jk_lane(const JKArgs a)
{
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    constexpr int NF = NA * NB * NC * ND;
    constexpr int NR = (LA + LB + LC + LD) / 2 + 1;
    constexpr int NI = LA + LB + 1, NK = LC + LD + 1;

    const int bt = find_tile(a.bt_off, a.nbt, blockIdx.x);
    const int kt = blockIdx.x - a.bt_off[bt];
    const int4 t = a.bt[bt];
    const int p = t.x + (threadIdx.x >> 3);
    const int q = t.z + kt * TILE + (threadIdx.x & 7);
    const int mol = a.bt_mol[bt];
    bool alive = p < t.y && q < t.w && !(a.diag && q > p);
    double4 rA = make_double4(0, 0, 0, 0), rC = rA;
    int pb0 = 0, npb = 0, pk0 = 0, npk = 0;
    if (alive) {
        rA = a.P.rA[p];
        rC = a.P.rA[q];
        if (a.screen) alive = keep_quartet(a, mol, p, q, rA.w * rC.w);
    }
    if (alive) {
        pb0 = a.P.prim0[p];
        npb = a.P.nprim[p];
        pk0 = a.P.prim0[q];
        npk = a.P.nprim[q];
    }
    count_work(a.counters, alive, npb * npk);
    if (!alive) return;
    const double4 AB = a.P.rAB[p];
    const double4 CD = a.P.rAB[q];

    double gout[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) gout[f] = 0.0;

    for (int ib = 0; ib < npb; ++ib) {
        const double4 Pp = a.P.ppP[pb0 + ib];
        const double2 Kp = a.P.ppK[pb0 + ib];
        const double PAx = Pp.x - rA.x, PAy = Pp.y - rA.y, PAz = Pp.z - rA.z;
        for (int ik = 0; ik < npk; ++ik) {
            const double4 Qp = a.P.ppP[pk0 + ik];
            const double2 Kq = a.P.ppK[pk0 + ik];
            const double pq = Pp.w + Qp.w;
            const double ipq = 1.0 / pq;
            const double PQx = Pp.x - Qp.x, PQy = Pp.y - Qp.y, PQz = Pp.z - Qp.z;
            const double T = Pp.w * Qp.w * ipq * (PQx * PQx + PQy * PQy + PQz * PQz);
            const double pref = Kp.x * Kq.x * sqrt(ipq);
            const RysCursor<NR> rys(a.rys, T);
            const double QCx = Qp.x - rC.x, QCy = Qp.y - rC.y, QCz = Qp.z - rC.z;
#pragma unroll 1
            for (int r = 0; r < NR; ++r) {
                double rx, rw;
                rys.root(r, rx, rw);
                const double u = rx * ipq;
                const double B00 = 0.5 * u;
                const double B10 = 0.5 * Kp.y * (1.0 - u * Qp.w);
                const double B01 = 0.5 * Kq.y * (1.0 - u * Pp.w);
                const double uq = u * Qp.w, up = u * Pp.w;
                double gx[NI][NK], gy[NI][NK], gz[NI][NK];
                vrr<NI, NK>(gx, 1.0, PAx - uq * PQx, QCx + up * PQx, B10, B01, B00);
                vrr<NI, NK>(gy, 1.0, PAy - uq * PQy, QCy + up * PQy, B10, B01, B00);
                vrr<NI, NK>(gz, rw * pref, PAz - uq * PQz, QCz + up * PQz, B10, B01, B00);
                double Ix[LA + 1][LB + 1][LC + 1][LD + 1];
                double Iy[LA + 1][LB + 1][LC + 1][LD + 1];
                double Iz[LA + 1][LB + 1][LC + 1][LD + 1];
                hrr<LA, LB, LC, LD>(gx, AB.x, CD.x, Ix);
                hrr<LA, LB, LC, LD>(gy, AB.y, CD.y, Iy);
                hrr<LA, LB, LC, LD>(gz, AB.z, CD.z, Iz);
#pragma unroll
                for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                for (int ib2 = 0; ib2 < NB; ++ib2)
#pragma unroll
                for (int ic = 0; ic < NC; ++ic)
#pragma unroll
                for (int id = 0; id < ND; ++id) {
                    const int f = ((ia * NB + ib2) * NC + ic) * ND + id;
                    gout[f] = fma(Ix[cart_x(LA, ia)][cart_x(LB, ib2)][cart_x(LC, ic)][cart_x(LD, id)]
                                      * Iy[cart_y(LA, ia)][cart_y(LB, ib2)][cart_y(LC, ic)][cart_y(LD, id)],
                                  Iz[cart_z(LA, ia)][cart_z(LB, ib2)][cart_z(LC, ic)][cart_z(LD, id)],
                                  gout[f]);
                }
            }
        }
    }

    // degeneracy of the unique quartet
    double fac = 1.0;
    if (a.P.shA[p] == a.P.shB[p]) fac *= 0.5;
    if (a.P.shA[q] == a.P.shB[q]) fac *= 0.5;
    if (p == q && a.diag) fac *= 0.5;
#pragma unroll
    for (int f = 0; f < NF; ++f) gout[f] *= fac;

    const long long base = a.M.dbase[mol];
    const int ld = a.M.ld[mol];
    const int oa = a.P.cartA[p], ob = a.P.cartB[p], oc = a.P.cartA[q], od = a.P.cartB[q];

    for (int s = 0; s < a.n_j; ++s) {
        const double* __restrict__ D = a.Dj + s * a.slot + base;
        double* V = a.J + s * a.slot + base;
        double dab[NA][NB], dcd[NC][ND];
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int j = 0; j < NB; ++j) dab[i][j] = D[(oa + i) * ld + ob + j];
#pragma unroll
        for (int i = 0; i < NC; ++i)
#pragma unroll
            for (int j = 0; j < ND; ++j) dcd[i][j] = D[(oc + i) * ld + od + j];
#pragma unroll
        for (int ic = 0; ic < NC; ++ic)
#pragma unroll
            for (int id = 0; id < ND; ++id) {
                double acc = 0.0;
#pragma unroll
                for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                    for (int ib2 = 0; ib2 < NB; ++ib2)
                        acc = fma(gout[((ia * NB + ib2) * NC + ic) * ND + id], dab[ia][ib2], acc);
                atomic_add(V + (oc + ic) * ld + od + id, acc);
            }
#pragma unroll
        for (int ia = 0; ia < NA; ++ia)
#pragma unroll
            for (int ib2 = 0; ib2 < NB; ++ib2) {
                double acc = 0.0;
#pragma unroll
                for (int ic = 0; ic < NC; ++ic)
#pragma unroll
                    for (int id = 0; id < ND; ++id)
                        acc = fma(gout[((ia * NB + ib2) * NC + ic) * ND + id], dcd[ic][id], acc);
                atomic_add(V + (oa + ia) * ld + ob + ib2, acc);
            }
    }

    for (int s = 0; s < a.n_k; ++s) {
        const double* __restrict__ D = a.Dk + s * a.slot + base;
        double* V = a.K + s * a.slot + base;
        // K'[a,d] += sum_bc v D[b,c]
        {
            double dbc[NB][NC];
#pragma unroll
            for (int i = 0; i < NB; ++i)
#pragma unroll
                for (int j = 0; j < NC; ++j) dbc[i][j] = D[(ob + i) * ld + oc + j];
#pragma unroll
            for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                for (int id = 0; id < ND; ++id) {
                    double acc = 0.0;
#pragma unroll
                    for (int ib2 = 0; ib2 < NB; ++ib2)
#pragma unroll
                        for (int ic = 0; ic < NC; ++ic)
                            acc = fma(gout[((ia * NB + ib2) * NC + ic) * ND + id], dbc[ib2][ic], acc);
                    atomic_add(V + (oa + ia) * ld + od + id, acc);
                }
        }
        // K'[b,d] += sum_ac v D[a,c]
        {
            double dac[NA][NC];
#pragma unroll
            for (int i = 0; i < NA; ++i)
#pragma unroll
                for (int j = 0; j < NC; ++j) dac[i][j] = D[(oa + i) * ld + oc + j];
#pragma unroll
            for (int ib2 = 0; ib2 < NB; ++ib2)
#pragma unroll
                for (int id = 0; id < ND; ++id) {
                    double acc = 0.0;
#pragma unroll
                    for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                        for (int ic = 0; ic < NC; ++ic)
                            acc = fma(gout[((ia * NB + ib2) * NC + ic) * ND + id], dac[ia][ic], acc);
                    atomic_add(V + (ob + ib2) * ld + od + id, acc);
                }
        }
        // K'[a,c] += sum_bd v D[b,d]
        {
            double dbd[NB][ND];
#pragma unroll
            for (int i = 0; i < NB; ++i)
#pragma unroll
                for (int j = 0; j < ND; ++j) dbd[i][j] = D[(ob + i) * ld + od + j];
#pragma unroll
            for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                for (int ic = 0; ic < NC; ++ic) {
                    double acc = 0.0;
#pragma unroll
                    for (int ib2 = 0; ib2 < NB; ++ib2)
#pragma unroll
                        for (int id = 0; id < ND; ++id)
                            acc = fma(gout[((ia * NB + ib2) * NC + ic) * ND + id], dbd[ib2][id], acc);
                    atomic_add(V + (oa + ia) * ld + oc + ic, acc);
                }
        }
        // K'[b,c] += sum_ad v D[a,d]
        {
            double dad[NA][ND];
#pragma unroll
            for (int i = 0; i < NA; ++i)
#pragma unroll
                for (int j = 0; j < ND; ++j) dad[i][j] = D[(oa + i) * ld + od + j];
#pragma unroll
            for (int ib2 = 0; ib2 < NB; ++ib2)
#pragma unroll
                for (int ic = 0; ic < NC; ++ic) {
                    double acc = 0.0;
#pragma unroll
                    for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                        for (int id = 0; id < ND; ++id)
                            acc = fma(gout[((ia * NB + ib2) * NC + ic) * ND + id], dad[ia][id], acc);
                    atomic_add(V + (ob + ib2) * ld + oc + ic, acc);
                }
        }
    }
}

}  // namespace mjk

// ---------------------------------------------------------------------------
// Schwarz bound of each shell pair: q = sqrt(max_{a,b} |(ab|ab)|) over the
// spherical functions (pyscf/lib/vhf/optimizer.c:408-454), floor 1e-100.
#include "c2s.h"

namespace mjk {

template <int LA, int LB>
__global__ __launch_bounds__(64) void k_qcond(const QCondArgs a)
{
    constexpr int NR = LA + LB + 1;
    constexpr int NI = LA + LB + 1;
    constexpr int SA = 2 * LA + 1, SB = 2 * LB + 1;
    constexpr int NA = ncart(LA), NB = ncart(LB);
    const int p = a.p0 + blockIdx.x * 64 + threadIdx.x;
    if (p >= a.p0 + a.np) return;
    const double4 rA = a.P.rA[p];
    const double4 AB = a.P.rAB[p];
    double acc[SA][SB];
#pragma unroll
    for (int i = 0; i < SA; ++i)
#pragma unroll
        for (int j = 0; j < SB; ++j) acc[i][j] = 0.0;
    const int p0 = a.P.prim0[p], np = a.P.nprim[p];
    for (int ib = 0; ib < np; ++ib) {
        const double4 Pp = a.P.ppP[p0 + ib];
        const double2 Kp = a.P.ppK[p0 + ib];
        for (int ik = 0; ik < np; ++ik) {
            const double4 Qp = a.P.ppP[p0 + ik];
            const double2 Kq = a.P.ppK[p0 + ik];
            const double pq = Pp.w + Qp.w;
            const double ipq = 1.0 / pq;
            const double PQx = Pp.x - Qp.x, PQy = Pp.y - Qp.y, PQz = Pp.z - Qp.z;
            const double T = Pp.w * Qp.w * ipq * (PQx * PQx + PQy * PQy + PQz * PQz);
            const double pref = Kp.x * Kq.x * sqrt(ipq);
            const RysCursor<NR> rys(a.rys, T);
#pragma unroll 1
            for (int r = 0; r < NR; ++r) {
                double rx, rw;
                rys.root(r, rx, rw);
                const double u = rx * ipq;
                const double B00 = 0.5 * u;
                const double B10 = 0.5 * Kp.y * (1.0 - u * Qp.w);
                const double B01 = 0.5 * Kq.y * (1.0 - u * Pp.w);
                const double uq = u * Qp.w, up = u * Pp.w;
                double gx[NI][NI], gy[NI][NI], gz[NI][NI];
                vrr<NI, NI>(gx, 1.0, Pp.x - rA.x - uq * PQx, Qp.x - rA.x + up * PQx, B10, B01, B00);
                vrr<NI, NI>(gy, 1.0, Pp.y - rA.y - uq * PQy, Qp.y - rA.y + up * PQy, B10, B01, B00);
                vrr<NI, NI>(gz, rw * pref, Pp.z - rA.z - uq * PQz, Qp.z - rA.z + up * PQz, B10, B01, B00);
                double Ix[LA + 1][LB + 1][LA + 1][LB + 1];
                double Iy[LA + 1][LB + 1][LA + 1][LB + 1];
                double Iz[LA + 1][LB + 1][LA + 1][LB + 1];
                hrr<LA, LB, LA, LB>(gx, AB.x, AB.x, Ix);
                hrr<LA, LB, LA, LB>(gy, AB.y, AB.y, Iy);
                hrr<LA, LB, LA, LB>(gz, AB.z, AB.z, Iz);
#pragma unroll
                for (int ma = 0; ma < SA; ++ma)
#pragma unroll
                for (int mb = 0; mb < SB; ++mb) {
                    double v = 0.0;
#pragma unroll
                    for (int c1 = 0; c1 < NA; ++c1) {
                        if (C2S[LA][ma][c1] == 0.0) continue;
#pragma unroll
                        for (int c2 = 0; c2 < NB; ++c2) {
                            if (C2S[LB][mb][c2] == 0.0) continue;
#pragma unroll
                            for (int c3 = 0; c3 < NA; ++c3) {
                                if (C2S[LA][ma][c3] == 0.0) continue;
#pragma unroll
                                for (int c4 = 0; c4 < NB; ++c4) {
                                    if (C2S[LB][mb][c4] == 0.0) continue;
                                    const double cc = C2S[LA][ma][c1] * C2S[LB][mb][c2]
                                                    * C2S[LA][ma][c3] * C2S[LB][mb][c4];
                                    v = fma(cc * Ix[cart_x(LA, c1)][cart_x(LB, c2)][cart_x(LA, c3)][cart_x(LB, c4)]
                                                * Iy[cart_y(LA, c1)][cart_y(LB, c2)][cart_y(LA, c3)][cart_y(LB, c4)],
                                            Iz[cart_z(LA, c1)][cart_z(LB, c2)][cart_z(LA, c3)][cart_z(LB, c4)], v);
                                }
                            }
                        }
                    }
                    acc[ma][mb] += v;
                }
            }
        }
    }
    double m = 0.0;
#pragma unroll
    for (int i = 0; i < SA; ++i)
#pragma unroll
        for (int j = 0; j < SB; ++j) m = fmax(m, fabs(acc[i][j]));
    a.qout[p] = m > 1e-200 ? fmax(sqrt(m), 1e-100) : 1e-100;
}

// ---------------------------------------------------------------------------
// Spherical <-> Cartesian transforms of densities and J/K.
__constant__ double c_c2s[4][7][10];

// D_cart = C D C^T for every density, plus J density C (D + D^T) C^T and,
// for a non-symmetric D, the transposed K density C D^T C^T.
// Grid: x over all Cartesian rows of the batch, y over columns.
__global__ __launch_bounds__(256) void k_sph2cart(const XformArgs x, const double* __restrict__ Dsph,
                                                  long long sslot, int n_dm, int hermi, int with_j,
                                                  int with_k, double* __restrict__ Dj,
                                                  double* __restrict__ Dk, long long cslot)
{
    const long long I = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int jj = blockIdx.y * blockDim.y + threadIdx.y;
    if (I >= x.ncart_tot) return;
    const int m = x.cart_mol[I];
    const int ld = x.ld[m];
    if (jj >= ld) return;
    const long long J = x.cart_first[m] + jj;
    const int il = (int)(I - x.cart_first[m]);
    const int s0i = x.cart_s0[I], li = x.cart_l[I], ci = x.cart_c[I];
    const int s0j = x.cart_s0[J], lj = x.cart_l[J], cj = x.cart_c[J];
    const int nao = x.nao[m];
    const long long out = x.dbase[m] + (long long)il * ld + jj;
    for (int s = 0; s < n_dm; ++s) {
        const double* __restrict__ D = Dsph + s * sslot + x.sbase[m];
        double t1 = 0.0, t2 = 0.0;
        for (int m1 = 0; m1 < 2 * li + 1; ++m1) {
            const double a1 = c_c2s[li][m1][ci];
            if (a1 == 0.0) continue;
            for (int m2 = 0; m2 < 2 * lj + 1; ++m2) {
                const double a2 = a1 * c_c2s[lj][m2][cj];
                if (a2 == 0.0) continue;
                t1 = fma(a2, D[(long long)(s0i + m1) * nao + s0j + m2], t1);
                t2 = fma(a2, D[(long long)(s0j + m2) * nao + s0i + m1], t2);
            }
        }
        if (with_j) Dj[s * cslot + out] = t1 + t2;
        if (with_k) {
            if (hermi == 1) {
                Dk[s * cslot + out] = t1;
            } else {
                Dk[(2 * s) * cslot + out] = t1;
                Dk[(2 * s + 1) * cslot + out] = t2;
            }
        }
    }
}

// vj = C^T (J' + J'^T) C; vk = C^T (K' + K'^T) C (hermi 1) or
// C^T (K'_D + K'_{D^T}^T) C (hermi 0, 2); hermi 2 then copies minus the lower
// triangle into the upper one (pyscf/lib/numpy_helper.py:480-525).
__global__ __launch_bounds__(256) void k_cart2sph(const XformArgs x, const double* __restrict__ Jc,
                                                  const double* __restrict__ Kc, long long cslot,
                                                  int n_dm, int hermi, int with_j, int with_k,
                                                  double* __restrict__ vj, double* __restrict__ vk,
                                                  long long sslot)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int jj = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= x.nsph_tot) return;
    const int m = x.sph_mol[i];
    const int nao = x.nao[m];
    if (jj >= nao) return;
    const int ldc = x.ld[m];
    const int il = (int)(i - x.sph_first[m]);
    const long long out = x.sbase[m] + (long long)il * nao + jj;
    // hermi 2: the upper triangle is minus the transposed lower one
    const bool anti = (hermi == 2 && il < jj);
    const long long r = anti ? x.sph_first[m] + jj : i;
    const long long c = anti ? i : x.sph_first[m] + jj;
    const int c0i = x.sph_c0[r], li = x.sph_l[r], mi = x.sph_m[r];
    const int c0j = x.sph_c0[c], lj = x.sph_l[c], mj = x.sph_m[c];
    const long long cb = x.dbase[m];
    for (int s = 0; s < n_dm; ++s) {
        double tj = 0.0, tk = 0.0;
        for (int a1 = 0; a1 < ncart(li); ++a1) {
            const double w1 = c_c2s[li][mi][a1];
            if (w1 == 0.0) continue;
            for (int a2 = 0; a2 < ncart(lj); ++a2) {
                const double w = w1 * c_c2s[lj][mj][a2];
                if (w == 0.0) continue;
                const long long e1 = cb + (long long)(c0i + a1) * ldc + c0j + a2;
                const long long e2 = cb + (long long)(c0j + a2) * ldc + c0i + a1;
                if (with_j) tj = fma(w, Jc[s * cslot + e1] + Jc[s * cslot + e2], tj);
                if (with_k) {
                    if (hermi == 1)
                        tk = fma(w, Kc[s * cslot + e1] + Kc[s * cslot + e2], tk);
                    else
                        tk = fma(w, Kc[(2 * s) * cslot + e1] + Kc[(2 * s + 1) * cslot + e2], tk);
                }
            }
        }
        if (with_j) vj[s * sslot + out] = anti ? tj : tj;
        if (with_k) vk[s * sslot + out] = anti ? -tk : tk;
    }
}

// dm_cond[I,J] = 0.5 max_sets max_{i in I, j in J} (|D_ij| + |D_ji|)
// (pyscf/lib/vhf/optimizer.c:494-518), per shell pair of each molecule.
__global__ __launch_bounds__(256) void k_dmcond(const DmCondArgs x, const double* __restrict__ Dsph,
                                                long long sslot, int n_dm, double* __restrict__ out)
{
    const long long I = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int jj = blockIdx.y * blockDim.y + threadIdx.y;
    if (I >= x.nsh_tot) return;
    const int m = x.sh_mol[I];
    const int ns = x.nsh[m];
    if (jj >= ns) return;
    const long long J = x.sh_first[m] + jj;
    const int il = (int)(I - x.sh_first[m]);
    const int nao = x.nao[m];
    const int i0 = x.sh_s0[I], ni = x.sh_n[I], j0 = x.sh_s0[J], nj = x.sh_n[J];
    double v = 0.0;
    for (int s = 0; s < n_dm; ++s) {
        const double* __restrict__ D = Dsph + s * sslot + x.sbase[m];
        for (int a = 0; a < ni; ++a)
            for (int b = 0; b < nj; ++b)
                v = fmax(v, fabs(D[(long long)(i0 + a) * nao + j0 + b]) + fabs(D[(long long)(j0 + b) * nao + i0 + a]));
    }
    out[x.cbase[m] + (long long)il * ns + jj] = 0.5 * v;
}

// ---------------------------------------------------------------------------
// Host-side dispatch.
#define MJK_CLASSES_L2(X) \
    X(0,0,0,0) X(1,0,0,0) X(1,0,1,0) X(1,1,0,0) X(1,1,1,0) X(1,1,1,1) \
    X(2,0,0,0) X(2,0,1,0) X(2,0,1,1) X(2,0,2,0) \
    X(2,1,0,0) X(2,1,1,0) X(2,1,1,1) X(2,1,2,0) X(2,1,2,1) \
    X(2,2,0,0) X(2,2,1,0) X(2,2,1,1) X(2,2,2,0) X(2,2,2,1) X(2,2,2,2)

int launch_jk(int la, int lb, int lc, int ld, int nblocks, const JKArgs& a, hipStream_t st)
{
    const int code = ((la * 4 + lb) * 4 + lc) * 4 + ld;
    switch (code) {
#define X(A, B, C, D) \
    case ((A * 4 + B) * 4 + C) * 4 + D: \
        hipLaunchKernelGGL((k_jk_lane<A, B, C, D>), dim3(nblocks), dim3(64), 0, st, a); \
        return (int)hipGetLastError();
        MJK_CLASSES_L2(X)
#undef X
    default:
        return -1;
    }
}

int launch_qcond(int la, int lb, const QCondArgs& a, hipStream_t st)
{
    const int nb = (a.np + 63) / 64;
    if (nb == 0) return 0;
    const int code = la * 4 + lb;
    switch (code) {
#define Q(A, B) \
    case A * 4 + B: \
        hipLaunchKernelGGL((k_qcond<A, B>), dim3(nb), dim3(64), 0, st, a); \
        return (int)hipGetLastError();
        Q(0, 0) Q(1, 0) Q(1, 1) Q(2, 0) Q(2, 1) Q(2, 2)
#undef Q
    default:
        return -1;
    }
}

int upload_c2s(hipStream_t st)
{
    return (int)hipMemcpyToSymbolAsync(HIP_SYMBOL(c_c2s), C2S, sizeof(C2S), 0, hipMemcpyHostToDevice, st);
}

int launch_sph2cart(const XformArgs& x, int maxld, const double* Dsph, long long sslot, int n_dm,
                    int hermi, int with_j, int with_k, double* Dj, double* Dk, long long cslot,
                    hipStream_t st)
{
    dim3 blk(64, 4);
    dim3 grd((unsigned)((x.ncart_tot + 63) / 64), (unsigned)((maxld + 3) / 4));
    hipLaunchKernelGGL(k_sph2cart, grd, blk, 0, st, x, Dsph, sslot, n_dm, hermi, with_j, with_k, Dj, Dk,
                       cslot);
    return (int)hipGetLastError();
}

int launch_cart2sph(const XformArgs& x, int maxnao, const double* Jc, const double* Kc, long long cslot,
                    int n_dm, int hermi, int with_j, int with_k, double* vj, double* vk, long long sslot,
                    hipStream_t st)
{
    dim3 blk(64, 4);
    dim3 grd((unsigned)((x.nsph_tot + 63) / 64), (unsigned)((maxnao + 3) / 4));
    hipLaunchKernelGGL(k_cart2sph, grd, blk, 0, st, x, Jc, Kc, cslot, n_dm, hermi, with_j, with_k, vj, vk,
                       sslot);
    return (int)hipGetLastError();
}

int launch_dmcond(const DmCondArgs& x, const double* Dsph, long long sslot, int n_dm, double* out,
                  hipStream_t st)
{
    dim3 blk(64, 4);
    dim3 grd((unsigned)((x.nsh_tot + 63) / 64), (unsigned)((x.maxnsh + 3) / 4));
    hipLaunchKernelGGL(k_dmcond, grd, blk, 0, st, x, Dsph, sslot, n_dm, out);
    return (int)hipGetLastError();
}

}  // namespace mjk
