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
#pragma once
#include <hip/hip_runtime.h>
#include "jk_common.h"
#include "rys_eval.h"
#include "c2s.h"

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

