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

// Cartesian component i of angular momentum l, in the order xx..x, then
// decreasing lx, and within one lx decreasing ly: i = r (r + 1) / 2 + lz with
// r = l - lx = ly + lz.  Branch-free closed forms (i < 21, l <= 5), so that
// indices built from unrolled loop counters fold to constants.
__host__ __device__ constexpr int cart_row(int i)
{
    return (i >= 1) + (i >= 3) + (i >= 6) + (i >= 10) + (i >= 15);
}
__host__ __device__ constexpr int cart_x(int l, int i) { return l - cart_row(i); }
__host__ __device__ constexpr int cart_y(int l, int i)
{
    return cart_row(i) - (i - cart_row(i) * (cart_row(i) + 1) / 2);
}
__host__ __device__ constexpr int cart_z(int l, int i) { return l - cart_x(l, i) - cart_y(l, i); }

// Vertical recurrence for one Cartesian direction and one root.
template <int NI, int NK, typename R>
__device__ __forceinline__ void vrr(R (&g)[NI][NK], R g00, R C, R Cp, R B10, R B01, R B00)
{
    g[0][0] = g00;
    if constexpr (NI > 1) {
        g[1][0] = C * g00;
#pragma unroll
        for (int i = 1; i < NI - 1; ++i) g[i + 1][0] = C * g[i][0] + (R(i) * B10) * g[i - 1][0];
    }
#pragma unroll
    for (int k = 0; k < NK - 1; ++k) {
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            R v = Cp * g[i][k];
            if (k > 0) v += (R(k) * B01) * g[i][k - 1];
            if (i > 0) v += (R(i) * B00) * g[i - 1][k];
            g[i][k + 1] = v;
        }
    }
}

// Horizontal (transfer) recurrence: g(a+b, c+d) -> I(a, b, c, d).
template <int LA, int LB, int LC, int LD, typename R>
__device__ __forceinline__ void hrr(const R (&g)[LA + LB + 1][LC + LD + 1], R AB, R CD,
                                    R (&I)[LA + 1][LB + 1][LC + 1][LD + 1])
{
    constexpr int NI = LA + LB + 1, NK = LC + LD + 1;
    R h[NI][LB + 1][NK];
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
            R e[NK][LD + 1];
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

// Index of the XCD (accelerator complex die) this wave runs on.  Each XCD
// accumulates into its own copy of J'/K', so the cache lines that atomics hit
// stay in that XCD's L2 instead of moving between dies; the copies are summed
// by the output transform.  Correctness does not depend on the mapping (the
// atomics are device-coherent), only the locality does.
__device__ __forceinline__ int xcc_copy(int ncopy)
{
    const int id = __builtin_amdgcn_s_getreg((20) | (0 << 6) | ((4 - 1) << 11));   // HW_REG_XCC_ID[3:0]
    return ncopy > 1 ? (id % ncopy) : 0;
}

__device__ __forceinline__ void atomic_add(double* p, double v)
{
    unsafeAtomicAdd(p, v);
}

// Load of read-only data at a wave-uniform address through the scalar cache
// (constant address space: s_load into SGPRs), which keeps block-uniform pair
// data out of the vector registers.
template <typename T>
__device__ __forceinline__ T uload(const T* p)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return *(const __attribute__((address_space(4))) T*)(p);
#else
    return *p;
#endif
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
// Largest density factor of the predicate: the quartet survives when
// q_ij q_kl > cutoff and qpq * dmax > cutoff; qpq * dmax also bounds its
// contribution to J and K (the mixed-precision criterion).
__device__ __forceinline__ double quartet_dmax(const JKArgs& a, int mol, int p, int q)
{
    const double* __restrict__ dc = a.dm_cond + a.M.cbase[mol];
    const int ns = a.M.nsh[mol];
    const int i = a.P.oshA[p], j = a.P.oshB[p], k = a.P.oshA[q], l = a.P.oshB[q];
    // J terms only for J builds, K terms only for K builds (the J-only and
    // K-only predicates, pyscf/lib/vhf/optimizer.c:119-167)
    double d = 0.0;
    if (a.n_j) d = 4.0 * fmax(dc[j * ns + i], dc[l * ns + k]);
    if (a.n_k) {
        d = fmax(d, fmax(dc[j * ns + k], dc[j * ns + l]));
        d = fmax(d, fmax(dc[i * ns + k], dc[i * ns + l]));
    }
    return d;
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
// Quartet kernel.  A block owns one 8 x 8 tile of (bra pair, ket pair)
// tasks; lane l of every wave holds quartet (p0 + l/8, q0 + l%8).  Classes
// with many Cartesian components are cut into SPLIT slices over the leading
// components (a, then b, then c); each wave evaluates whole slices, so the
// slice index is wave-uniform and dispatched at compile time: every array
// index stays a constant and the slice's integrals live in registers.  The
// roots and the vertical recurrence are recomputed per slice.

// Slice geometry: depth 0 = all components, 1 = a block of NA / SPLIT
// consecutive a components (one a when SPLIT = NA), 2 = fixed (a, b),
// 3 = fixed (a, b, c).
template <int NA, int NB, int NC, int SPLIT>
struct SliceShape {
    static constexpr int depth = SPLIT == 1                      ? 0
                                 : (SPLIT <= NA && NA % SPLIT == 0) ? 1
                                 : SPLIT == NA * NB                 ? 2
                                                                    : 3;
};

struct QuartetCtx {
    int p, q, mol;
    bool alive, pvalid, qvalid;
    double4 rA, rC, AB, CD;
    int pb0, npb, pk0, npk;
    int tq, ts;          // ket primitives in the transposed copy: tP[tq + ik * ts]
    double fac;
    const double* rtab;  // the block's LDS copy of the Rys table section (row kernels)
};

// Rys table of the row kernels: classes of at most two roots (la + lb + lc +
// ld <= 3) read the table from a copy in LDS that each block makes once
// (11 KB for two roots).  Their integrals are light and their root series
// are most of their memory reads: from the vector cache, the 16-byte gathers
// of 64 lanes kept the texture address and data units 80-90% busy (MI355X,
// C30H62 cc-pVDZ, rocprofv3 TA_TA_BUSY / TD_TD_BUSY), and from LDS the build
// runs 1.18x faster.  Classes of three or more roots gained nothing overall
// (the larger copy costs LDS occupancy and fill time) and read the table in
// memory.  (ss|ss) needs no table.
template <int LA, int LB, int LC, int LD>
struct RowRys {
    static constexpr int L = LA + LB + LC + LD;
    static constexpr int nr = L / 2 + 1;
    static constexpr bool lds = nr <= 2 && L > 0;
    static constexpr int size = lds ? RysLds<nr>::SIZE : 0;
    static_assert(size * 8 <= 12 * 1024, "static LDS of the row kernels stays within the 16 KB the host leaves");
};

// Sum over the 8 lanes sharing the bra pair (xor 1,2,4) or the ket pair
// (xor 8,16,32) of the 8 x 8 lane tile.
template <int M0>
__device__ __forceinline__ double tile_sum(double v)
{
    v += __shfl_xor(v, M0);
    v += __shfl_xor(v, 2 * M0);
    v += __shfl_xor(v, 4 * M0);
    return v;
}

// Sum of v over the 64 lanes of a full wave, left in lane 63, by DPP moves on
// the VALU (no LDS traffic): quad swaps, half-row and row mirrors, then the
// row broadcasts of lanes 15 and 31 into the odd rows and the upper half.
// Lanes that a step does not write add 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROW_MASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ double wave_sum_last(double v)
{
    v += dpp_move<0xB1, 0xf>(v);    // quad_perm [1, 0, 3, 2]
    v += dpp_move<0x4E, 0xf>(v);    // quad_perm [2, 3, 0, 1]
    v += dpp_move<0x141, 0xf>(v);   // row_half_mirror
    v += dpp_move<0x140, 0xf>(v);   // row_mirror
    v += dpp_move<0x142, 0xa>(v);   // row_bcast:15 into rows 1 and 3
    v += dpp_move<0x143, 0xc>(v);   // row_bcast:31 into rows 2 and 3
    return v;
}

// Component ranges of slice S of a class cut into SPLIT slices.
template <int LA, int LB, int LC, int LD, int SPLIT, int S>
struct SliceRange {
    static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    static constexpr int DEPTH = SliceShape<NA, NB, NC, SPLIT>::depth;
    static constexpr int IA0 = DEPTH == 0 ? 0 : DEPTH == 1 ? S * (NA / SPLIT) : DEPTH == 2 ? S / NB : S / (NB * NC);
    static constexpr int IA1 = DEPTH == 0 ? NA : DEPTH == 1 ? IA0 + NA / SPLIT : IA0 + 1;
    static constexpr int IB0 = DEPTH <= 1 ? 0 : DEPTH == 2 ? S % NB : (S / NC) % NB;
    static constexpr int IB1 = DEPTH <= 1 ? NB : IB0 + 1;
    static constexpr int IC0 = DEPTH <= 2 ? 0 : S % NC;
    static constexpr int IC1 = DEPTH <= 2 ? NC : IC0 + 1;
    static constexpr int MA = IA1 - IA0, MB = IB1 - IB0, MC = IC1 - IC0;
    static constexpr int NFL = MA * MB * MC * ND;
};

// One primitive quartet: what the root loop needs.
struct PrimQuartet {
    double ipq, pref, p, q, ip2, iq2;     // 1/(p+q), prefactor, p, q, 1/(2p), 1/(2q)
    double PAx, PAy, PAz, QCx, QCy, QCz, PQx, PQy, PQz;
};

__device__ __forceinline__ PrimQuartet prim_quartet(const double4& Pp, const double2& Kp, const double4& Qp,
                                                    const double2& Kq, const QuartetCtx& c, double& T)
{
    PrimQuartet m;
    const double pq = Pp.w + Qp.w;
    const double sipq = rsqrt_nr(pq);   // 1/sqrt(p+q)
    m.ipq = sipq * sipq;
    m.PQx = Pp.x - Qp.x;
    m.PQy = Pp.y - Qp.y;
    m.PQz = Pp.z - Qp.z;
    T = Pp.w * Qp.w * m.ipq * (m.PQx * m.PQx + m.PQy * m.PQy + m.PQz * m.PQz);
    m.pref = Kp.x * Kq.x * sipq;
    m.p = Pp.w;
    m.q = Qp.w;
    m.ip2 = 0.5 * Kp.y;
    m.iq2 = 0.5 * Kq.y;
    m.PAx = Pp.x - c.rA.x;
    m.PAy = Pp.y - c.rA.y;
    m.PAz = Pp.z - c.rA.z;
    m.QCx = Qp.x - c.rC.x;
    m.QCy = Qp.y - c.rC.y;
    m.QCz = Qp.z - c.rC.z;
    return m;
}

// Adds root (rx, rw) of primitive quartet m to the slice's Cartesian block.
template <int LA, int LB, int LC, int LD, int SPLIT, int S, typename R>
__device__ __forceinline__ void add_root(const PrimQuartet& m, double rx, double rw, const QuartetCtx& c,
                                         R (&gout)[SliceRange<LA, LB, LC, LD, SPLIT, S>::NFL])
{
    using SR = SliceRange<LA, LB, LC, LD, SPLIT, S>;
    constexpr int ND = SR::ND;
    constexpr int NI = LA + LB + 1, NK = LC + LD + 1;
    constexpr int IA0 = SR::IA0, IA1 = SR::IA1, IB0 = SR::IB0, IB1 = SR::IB1, IC0 = SR::IC0, IC1 = SR::IC1;
    constexpr int MB = SR::MB, MC = SR::MC;
    const double u = rx * m.ipq;
    const double uq = u * m.q, up = u * m.p;
    const R b10 = R(m.ip2 * (1.0 - uq)), b01 = R(m.iq2 * (1.0 - up)), b00 = R(0.5 * u);
    R gx[NI][NK], gy[NI][NK], gz[NI][NK];
    vrr<NI, NK, R>(gx, R(1), R(m.PAx - uq * m.PQx), R(m.QCx + up * m.PQx), b10, b01, b00);
    vrr<NI, NK, R>(gy, R(1), R(m.PAy - uq * m.PQy), R(m.QCy + up * m.PQy), b10, b01, b00);
    vrr<NI, NK, R>(gz, R(rw * m.pref), R(m.PAz - uq * m.PQz), R(m.QCz + up * m.PQz), b10, b01, b00);
    R Ix[LA + 1][LB + 1][LC + 1][LD + 1];
    R Iy[LA + 1][LB + 1][LC + 1][LD + 1];
    R Iz[LA + 1][LB + 1][LC + 1][LD + 1];
    hrr<LA, LB, LC, LD, R>(gx, R(c.AB.x), R(c.CD.x), Ix);
    hrr<LA, LB, LC, LD, R>(gy, R(c.AB.y), R(c.CD.y), Iy);
    hrr<LA, LB, LC, LD, R>(gz, R(c.AB.z), R(c.CD.z), Iz);
#pragma unroll
    for (int ia = IA0; ia < IA1; ++ia)
#pragma unroll
    for (int ib2 = IB0; ib2 < IB1; ++ib2)
#pragma unroll
    for (int ic = IC0; ic < IC1; ++ic)
#pragma unroll
    for (int id = 0; id < ND; ++id) {
        const int f = (((ia - IA0) * MB + ib2 - IB0) * MC + ic - IC0) * ND + id;
        gout[f] = fma(Ix[cart_x(LA, ia)][cart_x(LB, ib2)][cart_x(LC, ic)][cart_x(LD, id)]
                          * Iy[cart_y(LA, ia)][cart_y(LB, ib2)][cart_y(LC, ic)][cart_y(LD, id)],
                      Iz[cart_z(LA, ia)][cart_z(LB, ib2)][cart_z(LC, ic)][cart_z(LD, id)], gout[f]);
    }
}

// Ket primitives evaluated two at a time (independent chains for the
// scheduler to interleave) in classes whose slice is at most this many
// components; larger slices keep one at a time for their registers.
#ifndef MJK_PAIR_NFL
#define MJK_PAIR_NFL 27
#endif

// Contracted Cartesian block of one quartet restricted to slice S, times
// the quartet's degeneracy factor, in FP64 (recurrences in precision R).
// RSH: the kernel also evaluates the range-separated operators (a.omega != 0).
template <int LA, int LB, int LC, int LD, int SPLIT, int S, typename R, bool UNIFORM_BRA = false, bool RSH = false,
          bool LT = false>
__device__ __forceinline__ void eri_slice(const JKArgs& a, const QuartetCtx& c,
                                          double (&gd)[SliceRange<LA, LB, LC, LD, SPLIT, S>::NFL])
{
    using SR = SliceRange<LA, LB, LC, LD, SPLIT, S>;
    constexpr int NR = (LA + LB + LC + LD) / 2 + 1;
    constexpr int NFL = SR::NFL;
    using Rys = RysFor<NR, LA + LB + LC + LD, LT>;
    const double* const rt = LT ? c.rtab : a.rys;

    R gout[NFL];
#pragma unroll
    for (int f = 0; f < NFL; ++f) gout[f] = R(0);

    for (int ib = 0; ib < c.npb; ++ib) {
        const double4 Pp = UNIFORM_BRA ? uload(a.P.ppP + c.pb0 + ib) : a.P.ppP[c.pb0 + ib];
        const double2 Kp = UNIFORM_BRA ? uload(a.P.ppK + c.pb0 + ib) : a.P.ppK[c.pb0 + ib];
        if (RSH && a.omega != 0.0) {   // range-separated operator, one ket primitive at a time
            for (int ik = 0; ik < c.npk; ++ik) {
                double T;
                const int t0 = c.tq + ik * c.ts;
                const PrimQuartet m = prim_quartet(Pp, Kp, a.P.tP[t0], a.P.tK[t0], c, T);
                const double rho = m.p * m.q * m.ipq;
                for (int pass = 0; pass < rsh_passes(a.omega); ++pass) {
                    double Tp = T, sx, sw;
                    rsh_scale(a.omega, rho, pass, Tp, sx, sw);
                    const typename Rys::type rys(rt, Tp);
#pragma unroll 1
                    for (int r = 0; r < NR; ++r) {
                        double rx, rw;
                        rys.root(r, rx, rw);
                        add_root<LA, LB, LC, LD, SPLIT, S, R>(m, sx * rx, sw * rw, c, gout);
                    }
                }
            }
            continue;
        }
        int ik = 0;
        if (NFL <= MJK_PAIR_NFL) {
            for (; ik + 1 < c.npk; ik += 2) {
                double T0, T1;
                const int t0 = c.tq + ik * c.ts, t1 = t0 + c.ts;
                const PrimQuartet m0 = prim_quartet(Pp, Kp, a.P.tP[t0], a.P.tK[t0], c, T0);
                const PrimQuartet m1 = prim_quartet(Pp, Kp, a.P.tP[t1], a.P.tK[t1], c, T1);
                const typename Rys::type rys0(rt, T0), rys1(rt, T1);
#pragma unroll 1
                for (int r = 0; r < NR; ++r) {
                    double rx0, rw0, rx1, rw1;
                    rys0.root(r, rx0, rw0);
                    rys1.root(r, rx1, rw1);
                    add_root<LA, LB, LC, LD, SPLIT, S, R>(m0, rx0, rw0, c, gout);
                    add_root<LA, LB, LC, LD, SPLIT, S, R>(m1, rx1, rw1, c, gout);
                }
            }
        }
        for (; ik < c.npk; ++ik) {
            double T;
            const int t0 = c.tq + ik * c.ts;
            const PrimQuartet m = prim_quartet(Pp, Kp, a.P.tP[t0], a.P.tK[t0], c, T);
            const typename Rys::type rys(rt, T);
#pragma unroll 1
            for (int r = 0; r < NR; ++r) {
                double rx, rw;
                rys.root(r, rx, rw);
                add_root<LA, LB, LC, LD, SPLIT, S, R>(m, rx, rw, c, gout);
            }
        }
    }
    // contracted block in FP64 from here on (J/K accumulation is always FP64)
#pragma unroll
    for (int f = 0; f < NFL; ++f) gd[f] = double(gout[f]) * c.fac;
}

#define GOUT(ia, ib2, ic, id) gd[((((ia) - IA0) * MB + (ib2) - IB0) * MC + (ic) - IC0) * ND + (id)]

// Tile digestion: J' summed across the lanes sharing a pair, K' per lane.
template <int LA, int LB, int LC, int LD, int SPLIT, int S, typename R>
__device__ __forceinline__ void quartet_slice(const JKArgs& a, const QuartetCtx& c)
{
    using SR = SliceRange<LA, LB, LC, LD, SPLIT, S>;
    constexpr int ND = SR::ND;
    constexpr int IA0 = SR::IA0, IA1 = SR::IA1, IB0 = SR::IB0, IB1 = SR::IB1, IC0 = SR::IC0, IC1 = SR::IC1;
    constexpr int MA = SR::MA, MB = SR::MB, MC = SR::MC;
    double gd[SR::NFL];
    eri_slice<LA, LB, LC, LD, SPLIT, S, R, false, true>(a, c, gd);

    const long long base = a.M.dbase[c.mol];
    const int ld = a.M.ld[c.mol];
    const int lane = threadIdx.x & 63;
    // pair data of dead lanes: only valid pairs' offsets are read
    const int oa = c.pvalid ? a.P.cartA[c.p] : 0, ob = c.pvalid ? a.P.cartB[c.p] : 0;
    const int oc = c.qvalid ? a.P.cartA[c.q] : 0, od = c.qvalid ? a.P.cartB[c.q] : 0;

    const int cp = xcc_copy(a.ncopy);
    for (int s = 0; s < a.n_j; ++s) {
        const double* __restrict__ D = a.Dj + s * a.slot + base;
        double* V = a.J + cp * a.jcopy + s * a.slot + base;
        // J'[c,d] += sum_ab v Ds[a,b]: summed over this slice's (a, b) and the 8 bra lanes
        double dab[MA][MB];
#pragma unroll
        for (int i = 0; i < MA; ++i)
#pragma unroll
            for (int j = 0; j < MB; ++j) dab[i][j] = c.alive ? D[(oa + IA0 + i) * ld + ob + IB0 + j] : 0.0;
#pragma unroll
        for (int ic = IC0; ic < IC1; ++ic)
#pragma unroll
            for (int id = 0; id < ND; ++id) {
                double acc = 0.0;
#pragma unroll
                for (int ia = IA0; ia < IA1; ++ia)
#pragma unroll
                    for (int ib2 = IB0; ib2 < IB1; ++ib2) acc = fma(GOUT(ia, ib2, ic, id), dab[ia - IA0][ib2 - IB0], acc);
                acc = tile_sum<8>(acc);
                if (lane < 8 && c.qvalid && acc != 0.0) atomic_add(V + (oc + ic) * ld + od + id, acc);
            }
        // J'[a,b] += sum_cd v Ds[c,d]: summed over this slice's (c, d) and the 8 ket lanes
        double dcd[MC][ND];
#pragma unroll
        for (int i = 0; i < MC; ++i)
#pragma unroll
            for (int j = 0; j < ND; ++j) dcd[i][j] = c.alive ? D[(oc + IC0 + i) * ld + od + j] : 0.0;
#pragma unroll
        for (int ia = IA0; ia < IA1; ++ia)
#pragma unroll
            for (int ib2 = IB0; ib2 < IB1; ++ib2) {
                double acc = 0.0;
#pragma unroll
                for (int ic = IC0; ic < IC1; ++ic)
#pragma unroll
                    for (int id = 0; id < ND; ++id) acc = fma(GOUT(ia, ib2, ic, id), dcd[ic - IC0][id], acc);
                acc = tile_sum<1>(acc);
                if ((lane & 7) == 0 && c.pvalid && acc != 0.0) atomic_add(V + (oa + ia) * ld + ob + ib2, acc);
            }
    }

    if (!c.alive) return;
    for (int s = 0; s < a.n_k; ++s) {
        const double* __restrict__ D = a.Dk + s * a.slot + base;
        double* V = a.K + cp * a.kcopy + s * a.slot + base;
        // K'[a,d] += sum_bc v D[b,c]
        {
            double dbc[MB][MC];
#pragma unroll
            for (int i = 0; i < MB; ++i)
#pragma unroll
                for (int j = 0; j < MC; ++j) dbc[i][j] = D[(ob + IB0 + i) * ld + oc + IC0 + j];
#pragma unroll
            for (int ia = IA0; ia < IA1; ++ia)
#pragma unroll
                for (int id = 0; id < ND; ++id) {
                    double acc = 0.0;
#pragma unroll
                    for (int ib2 = IB0; ib2 < IB1; ++ib2)
#pragma unroll
                        for (int ic = IC0; ic < IC1; ++ic)
                            acc = fma(GOUT(ia, ib2, ic, id), dbc[ib2 - IB0][ic - IC0], acc);
                    atomic_add(V + (oa + ia) * ld + od + id, acc);
                }
        }
        // K'[b,d] += sum_ac v D[a,c]
        {
            double dac[MA][MC];
#pragma unroll
            for (int i = 0; i < MA; ++i)
#pragma unroll
                for (int j = 0; j < MC; ++j) dac[i][j] = D[(oa + IA0 + i) * ld + oc + IC0 + j];
#pragma unroll
            for (int ib2 = IB0; ib2 < IB1; ++ib2)
#pragma unroll
                for (int id = 0; id < ND; ++id) {
                    double acc = 0.0;
#pragma unroll
                    for (int ia = IA0; ia < IA1; ++ia)
#pragma unroll
                        for (int ic = IC0; ic < IC1; ++ic)
                            acc = fma(GOUT(ia, ib2, ic, id), dac[ia - IA0][ic - IC0], acc);
                    atomic_add(V + (ob + ib2) * ld + od + id, acc);
                }
        }
        // K'[a,c] += sum_bd v D[b,d]
        {
            double dbd[MB][ND];
#pragma unroll
            for (int i = 0; i < MB; ++i)
#pragma unroll
                for (int j = 0; j < ND; ++j) dbd[i][j] = D[(ob + IB0 + i) * ld + od + j];
#pragma unroll
            for (int ia = IA0; ia < IA1; ++ia)
#pragma unroll
                for (int ic = IC0; ic < IC1; ++ic) {
                    double acc = 0.0;
#pragma unroll
                    for (int ib2 = IB0; ib2 < IB1; ++ib2)
#pragma unroll
                        for (int id = 0; id < ND; ++id)
                            acc = fma(GOUT(ia, ib2, ic, id), dbd[ib2 - IB0][id], acc);
                    atomic_add(V + (oa + ia) * ld + oc + ic, acc);
                }
        }
        // K'[b,c] += sum_ad v D[a,d]
        {
            double dad[MA][ND];
#pragma unroll
            for (int i = 0; i < MA; ++i)
#pragma unroll
                for (int j = 0; j < ND; ++j) dad[i][j] = D[(oa + IA0 + i) * ld + od + j];
#pragma unroll
            for (int ib2 = IB0; ib2 < IB1; ++ib2)
#pragma unroll
                for (int ic = IC0; ic < IC1; ++ic) {
                    double acc = 0.0;
#pragma unroll
                    for (int ia = IA0; ia < IA1; ++ia)
#pragma unroll
                        for (int id = 0; id < ND; ++id)
                            acc = fma(GOUT(ia, ib2, ic, id), dad[ia - IA0][id], acc);
                    atomic_add(V + (ob + ib2) * ld + oc + ic, acc);
                }
        }
    }
}

template <int LA, int LB, int LC, int LD, int SPLIT, int S>
__global__ __launch_bounds__(64) void k_jk(const JKArgs a)
{
    const int lane = threadIdx.x & 63;
    const int bt = find_tile(a.bt_off, a.nbt, blockIdx.x);
    const int kt = blockIdx.x - a.bt_off[bt];
    const int4 t = a.bt[bt];
    QuartetCtx c;
    c.p = t.x + (lane >> 3);
    c.q = t.z + kt * TILE + (lane & 7);
    c.mol = a.bt_mol[bt];
    if (a.mol_active && !a.mol_active[c.mol]) return;   // block-uniform
    c.pvalid = c.p < t.y;
    c.qvalid = c.q < t.w;
    c.alive = c.pvalid && c.qvalid && !(a.diag && c.q > c.p);
    c.rA = c.rC = c.AB = c.CD = make_double4(0, 0, 0, 0);
    c.pb0 = c.npb = c.pk0 = c.npk = 0;
    c.tq = c.ts = 0;
    bool small = true;
    if (c.alive) {
        c.rA = a.P.rA[c.p];
        c.rC = a.P.rA[c.q];
        const double qpq = c.rA.w * c.rC.w;
        if (a.screen || a.fp32_tol > 0) {
            const double dmax = quartet_dmax(a, c.mol, c.p, c.q);
            if (a.screen) c.alive = qpq > a.cutoff && dmax * qpq > a.cutoff;
            small = dmax * qpq < a.fp32_tol;
        } else {
            small = false;
        }
    }
    if (c.alive) {
        c.pb0 = a.P.prim0[c.p];
        c.npb = a.P.nprim[c.p];
        c.pk0 = a.P.prim0[c.q];
        c.npk = a.P.nprim[c.q];
        c.tq = a.P.tpos[c.q];
        c.ts = a.P.tstride[c.q];
        c.AB = a.P.rAB[c.p];
        c.CD = a.P.rAB[c.q];
        double f = 1.0;
        if (a.P.shA[c.p] == a.P.shB[c.p]) f *= 0.5;
        if (a.P.shA[c.q] == a.P.shB[c.q]) f *= 0.5;
        if (c.p == c.q && a.diag) f *= 0.5;
        c.fac = f;
    } else {
        c.fac = 0.0;
    }
    if (!__ballot(c.alive)) return;   // whole tile screened out (wave-uniform exit)
    // FP32 only when every live quartet of the wave is below the bound
    const bool use32 = a.fp32_tol > 0 && __ballot(c.alive && !small) == 0;
    if (S == 0 && a.counters) count_work(a.counters + (use32 ? 2 : 0), c.alive, c.npb * c.npk);
    if (use32) quartet_slice<LA, LB, LC, LD, SPLIT, S, float>(a, c);
    else quartet_slice<LA, LB, LC, LD, SPLIT, S, double>(a, c);
}

// ---------------------------------------------------------------------------
// Cooperative kernel for classes with many Cartesian components.  A block of
// 64 NW threads evaluates one quartet at a time (every barrier is
// block-uniform); a block owns one bra pair of a tile and walks its 8 kets.
// Per primitive quartet:
//   phase A: item (root r, direction d) recomputes T and root r, runs the
//            vertical recurrence and the transfer recurrence, streaming
//            I_d(a, b, c, d') to LDS row by row;
//   phase B: thread t accumulates components f = t, t + 64 NW, ... as
//            sum_r Ix Iy Iz read from LDS.
// Digestion: the contracted block goes to LDS (reusing the I region), the
// D blocks are staged, and each J'/K' element is computed by one owner
// thread and added with one FP64 atomic.
// Cooperative classes run at three waves per SIMD (measured on MI355X, ms per
// launch at two -> three waves): (dp|dp) C30H62 cc-pVDZ 2.37 -> 2.16, (dd|dp)
// 1.11 -> 0.87, (dd|pp) 1.45 -> 1.39; (fp|dp) and (fd|pp) of a def2-TZVP
// chain 1.23x and 1.22x.  Classes already at three waves by registers lost up
// to 7% to the smaller primitive batches that three waves need in LDS.
__host__ __device__ constexpr int coop_key(int la, int lb, int lc, int ld) { return ((la * 4 + lb) * 4 + lc) * 4 + ld; }
__host__ __device__ constexpr bool coop_three_waves(int la, int lb, int lc, int ld)
{
    const int k = coop_key(la, lb, lc, ld);
    return k == coop_key(2, 1, 2, 1) || k == coop_key(2, 2, 1, 1) || k == coop_key(2, 2, 2, 1) ||
           k == coop_key(3, 0, 2, 2) || k == coop_key(3, 1, 1, 1) || k == coop_key(3, 1, 2, 1) ||
           k == coop_key(3, 2, 1, 1) || k == coop_key(3, 2, 3, 0) || k == coop_key(3, 3, 2, 0) ||
           k == coop_key(3, 3, 3, 0);
}

// Cooperative classes at two waves per SIMD by registers (169-256 per lane)
// whose 48 KB primitive batches left LDS room for only 3-5 blocks of 64 lanes
// per CU: batches sized to 1/8 of the CU's LDS per 64 lanes.  Measured on
// MI355X (ms per launch): (dd|dd) C30H62 cc-pVDZ 0.257 -> 0.147; def2-TZVP
// chain (1442 AOs) (fp|dd) 19.4 -> 10.2, (fp|fp) 20.1 -> 13.9, (ff|pp) 11.2
// -> 7.9, (fd|dp) 38.8 -> 23.1 (with a two-wave request: its smaller batches
// took 258 registers; the same request made the others 10-50% slower).
__host__ __device__ constexpr bool coop_two_wave_lds(int la, int lb, int lc, int ld)
{
    const int k = coop_key(la, lb, lc, ld);
    return k == coop_key(2, 2, 2, 2) || k == coop_key(3, 1, 2, 2) || k == coop_key(3, 1, 3, 1) ||
           k == coop_key(3, 2, 2, 1) || k == coop_key(3, 2, 2, 2) || k == coop_key(3, 2, 3, 1) ||
           k == coop_key(3, 3, 1, 1) || k == coop_key(3, 3, 2, 1);
}
// Cooperative classes at three waves by registers whose batches held them to
// 2.25 (LDS): batches within 1/12 of the CU's LDS per 64 lanes ((fp|fs) 1.07x,
// (ff|ps) 1.05x on the def2-TZVP chain).
__host__ __device__ constexpr bool coop_three_wave_lds(int la, int lb, int lc, int ld)
{
    const int k = coop_key(la, lb, lc, ld);
    return k == coop_key(3, 1, 3, 0) || k == coop_key(3, 3, 1, 0);
}

template <int LA, int LB, int LC, int LD>
struct CoopShape {
    static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    static constexpr int NF = NA * NB * NC * ND;
    static constexpr int NR = (LA + LB + LC + LD) / 2 + 1;
    static constexpr int ISZ = (LA + 1) * (LB + 1) * (LC + 1) * (LD + 1);
    static constexpr int SI = NR * 3 * ISZ;
    // J'[a,b] of the block's bra pair summed over its kets in LDS (one flush per
    // block instead of one per quartet): (fd|ps) 1.22x, (ff|ps) 1.30x, (dd|ds)
    // 1.11x on a def2-TZVP chain; not for (ff|dd), whose registers it takes
    // from 255 to 260 (one wave per SIMD, 0.57x)
    static constexpr bool JAB = !(LA == 3 && LB == 3 && LC == 2 && LD == 2);
    // staged D blocks, then the J'[a,b] sums (two densities)
    static constexpr int NDB = NA * NB + NC * ND + NB * NC + NA * NC + NB * ND + NA * ND + (JAB ? 2 * NA * NB : 0);
    static constexpr int NW = NF <= 64 * 24 ? 1 : NF <= 128 * 24 ? 2 : 4;
    static constexpr int NT = 64 * NW;
    static constexpr int CH = (NF + NT - 1) / NT;
    // waves per SIMD asked of the compiler: 3 for the classes whose registers
    // otherwise land between 169 and 256 per lane (two waves) and fit 168 with
    // at most 22 spilled, 2 for (fd|dp) (its LDS-sized batches otherwise take
    // 258), 1 (no request) for the rest
    static constexpr int WPE = coop_three_waves(LA, LB, LC, LD)               ? 3
                               : coop_key(LA, LB, LC, LD) == coop_key(3, 2, 2, 1) ? 2
                                                                                 : 1;
    // primitive quartets per phase: enough (root, direction) items to occupy the
    // block, within 6144 doubles (48 KB) of I arrays, and at two or three waves
    // per SIMD within the block's share of LDS (160 KB over 8 / NW or 12 / NW
    // blocks)
    static constexpr int LDSI = WPE == 3 || coop_three_wave_lds(LA, LB, LC, LD) ? 1700 * NW - NDB
                                : coop_two_wave_lds(LA, LB, LC, LD)           ? 2560 * NW - NDB
                                                                              : 6144;
    static constexpr int PB0 = NT / (NR * 3) < LDSI / SI ? NT / (NR * 3) : LDSI / SI;
    static constexpr int PB = PB0 < 1 ? 1 : PB0;
    static constexpr int SU = PB * SI > NF ? PB * SI : NF;
};

// Transfer recurrence g(a+b, c+d) -> I(a, b, c, d) written to LDS as
// dst[((a (LB+1) + b) (LC+1) + c) (LD+1) + d]; only two bra slabs live.
template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void hrr_to_lds(const double (&g)[LA + LB + 1][LC + LD + 1], double AB, double CD,
                                           double* __restrict__ dst)
{
    constexpr int NI = LA + LB + 1, NK = LC + LD + 1;
    double h[NI][NK];
#pragma unroll
    for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int k = 0; k < NK; ++k) h[i][k] = g[i][k];
#pragma unroll
    for (int b = 0; b <= LB; ++b) {
        if (b > 0) {
            // h_b(a) = h_{b-1}(a+1) + AB h_{b-1}(a), valid for a <= NI-1-b (ascending a: in place)
#pragma unroll
            for (int i = 0; i < NI - b; ++i)
#pragma unroll
                for (int k = 0; k < NK; ++k) h[i][k] = h[i + 1][k] + AB * h[i][k];
        }
#pragma unroll
        for (int ia = 0; ia <= LA; ++ia) {
            double e[NK][LD + 1];
#pragma unroll
            for (int k = 0; k < NK; ++k) e[k][0] = h[ia][k];
#pragma unroll
            for (int d = 1; d <= LD; ++d)
#pragma unroll
                for (int k = 0; k < NK - d; ++k) e[k][d] = e[k + 1][d - 1] + CD * e[k][d - 1];
#pragma unroll
            for (int c = 0; c <= LC; ++c)
#pragma unroll
                for (int d = 0; d <= LD; ++d)
                    dst[((ia * (LB + 1) + b) * (LC + 1) + c) * (LD + 1) + d] = e[c][d];
        }
    }
}

__device__ __forceinline__ int cart_x_rt(int l, int i)
{
    int lx = l, k = i;
    while (k > l - lx) { k -= l - lx + 1; --lx; }
    return lx;
}
__device__ __forceinline__ int cart_y_rt(int l, int i)
{
    int lx = l, k = i;
    while (k > l - lx) { k -= l - lx + 1; --lx; }
    return l - lx - k;
}

template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__((CoopShape<LA, LB, LC, LD>::NT))
__attribute__((amdgpu_waves_per_eu(CoopShape<LA, LB, LC, LD>::WPE))) void k_jk_coop(const JKArgs a)
{
    using Sh = CoopShape<LA, LB, LC, LD>;
    constexpr int NA = Sh::NA, NB = Sh::NB, NC = Sh::NC, ND = Sh::ND, NF = Sh::NF, NR = Sh::NR;
    constexpr int NI = LA + LB + 1, NK = LC + LD + 1;
    constexpr int ISZ = Sh::ISZ, NT = Sh::NT, CH = Sh::CH;
    constexpr int NITEM = NR * 3;
    __shared__ double lds[Sh::SU + Sh::NDB];
    double* const sblk = lds;                 // I arrays, later the contracted block
    double* const sD = lds + Sh::SU;          // staged D blocks
    double* const sJab = sD + Sh::NDB - (Sh::JAB ? 2 * NA * NB : 0);   // J'[a,b] over the block's kets
    const int tid = threadIdx.x;

    const int tb = blockIdx.x >> 3;
    const int row = blockIdx.x & 7;
    const int bt = find_tile(a.bt_off, a.nbt, tb);
    const int kt = tb - a.bt_off[bt];
    const int4 t = a.bt[bt];
    const int p = t.x + row;
    if (p >= t.y) return;
    const int mol = a.bt_mol[bt];
    if (a.mol_active && !a.mol_active[mol]) return;   // block-uniform
    const double4 rA = a.P.rA[p];
    const double4 AB = a.P.rAB[p];
    const int pb0 = a.P.prim0[p], npb = a.P.nprim[p];
    const bool diagA = a.P.shA[p] == a.P.shB[p];
    const int oa = a.P.cartA[p], ob = a.P.cartB[p];
    const long long base = a.M.dbase[mol];
    const int ld = a.M.ld[mol];

    // LDS offsets of this thread's components in the three I arrays of a root
    int ox[CH], oy[CH], oz[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        int f = tid + NT * c;
        if (f >= NF) f = NF - 1;
        const int id = f % ND, ic = (f / ND) % NC, ib = (f / (ND * NC)) % NB, ia = f / (ND * NC * NB);
        const int xa = cart_x_rt(LA, ia), xb = cart_x_rt(LB, ib), xc = cart_x_rt(LC, ic), xd = cart_x_rt(LD, id);
        const int ya = cart_y_rt(LA, ia), yb = cart_y_rt(LB, ib), yc = cart_y_rt(LC, ic), yd = cart_y_rt(LD, id);
        const int za = LA - xa - ya, zb = LB - xb - yb, zc = LC - xc - yc, zd = LD - xd - yd;
        ox[c] = ((xa * (LB + 1) + xb) * (LC + 1) + xc) * (LD + 1) + xd;
        oy[c] = ISZ + ((ya * (LB + 1) + yb) * (LC + 1) + yc) * (LD + 1) + yd;
        oz[c] = 2 * ISZ + ((za * (LB + 1) + zb) * (LC + 1) + zc) * (LD + 1) + zd;
    }

    // J'[a,b] of up to two densities summed over the block's kets in LDS (each
    // element by the thread that owns it in the digestion), one flush at the end
    const bool jab_lds = Sh::JAB && a.n_j <= 2;
    if (jab_lds)
        for (int e = tid; e < 2 * NA * NB; e += NT) sJab[e] = 0.0;
    unsigned long long nq_blk = 0, np_blk = 0;
    for (int col = 0; col < TILE; ++col) {
        const int q = t.z + kt * TILE + col;
        if (q >= t.w || (a.diag && q > p)) break;
        const double4 rC = a.P.rA[q];
        if (a.screen) {
            const double qpq = rA.w * rC.w;
            if (!(qpq > a.cutoff && quartet_dmax(a, mol, p, q) * qpq > a.cutoff)) continue;   // block-uniform
        }
        const int pk0 = a.P.prim0[q], npk = a.P.nprim[q];
        nq_blk += 1;
        np_blk += (unsigned long long)(npb * npk);
        const double4 CD = a.P.rAB[q];
        double gout[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) gout[c] = 0.0;

        constexpr int PB = Sh::PB;
        const int npq = npb * npk;
        // erfc operator: every primitive quartet twice (two root sets)
        const int nvq = rsh_passes(a.omega) * npq;
        for (int j0 = 0; j0 < nvq; j0 += PB) {      // PB primitive quartets per phase
            const int nb = min(PB, nvq - j0);
            {
                for (int it = tid; it < nb * NITEM; it += NT) {
                    const int j = it / NITEM;
                    const int r = (it - j * NITEM) / 3;
                    const int d = it - j * NITEM - 3 * r;
                    const int pass = j0 + j >= npq;
                    const int jq = j0 + j - pass * npq;
                    const int ib = jq / npk, ik = jq - ib * npk;
                    const double4 Pp = a.P.ppP[pb0 + ib];
                    const double2 Kp = a.P.ppK[pb0 + ib];
                    const double4 Qp = a.P.ppP[pk0 + ik];
                    const double2 Kq = a.P.ppK[pk0 + ik];
                    const double pq = Pp.w + Qp.w;
                    const double sipq = rsqrt_nr(pq);
                    const double ipq = sipq * sipq;
                    const double PQx = Pp.x - Qp.x, PQy = Pp.y - Qp.y, PQz = Pp.z - Qp.z;
                    double T = Pp.w * Qp.w * ipq * (PQx * PQx + PQy * PQy + PQz * PQz);
                    double sx, sw;
                    rsh_scale(a.omega, Pp.w * Qp.w * ipq, pass, T, sx, sw);
                    const RysCursor<NR> rys(a.rys, T);
                    double rx, rw;
                    rys.root(r, rx, rw);
                    rx *= sx;
                    rw *= sw;
                    const double u = rx * ipq;
                    const double PQd = d == 0 ? PQx : d == 1 ? PQy : PQz;
                    const double PAd = d == 0 ? Pp.x - rA.x : d == 1 ? Pp.y - rA.y : Pp.z - rA.z;
                    const double QCd = d == 0 ? Qp.x - rC.x : d == 1 ? Qp.y - rC.y : Qp.z - rC.z;
                    const double ABd = d == 0 ? AB.x : d == 1 ? AB.y : AB.z;
                    const double CDd = d == 0 ? CD.x : d == 1 ? CD.y : CD.z;
                    const double g00 = d == 2 ? rw * Kp.x * Kq.x * sipq : 1.0;
                    double g[NI][NK];
                    vrr<NI, NK, double>(g, g00, PAd - u * Qp.w * PQd, QCd + u * Pp.w * PQd,
                                        0.5 * Kp.y * (1.0 - u * Qp.w), 0.5 * Kq.y * (1.0 - u * Pp.w), 0.5 * u);
                    hrr_to_lds<LA, LB, LC, LD>(g, ABd, CDd, sblk + it * ISZ);
                }
                __syncthreads();
#pragma unroll
                for (int c = 0; c < CH; ++c) {
                    double s2 = 0.0;
                    for (int j = 0; j < nb; ++j) {
#pragma unroll
                        for (int r = 0; r < NR; ++r) {
                            const double* Ir = sblk + (j * NR + r) * 3 * ISZ;
                            s2 = fma(Ir[ox[c]] * Ir[oy[c]], Ir[oz[c]], s2);
                        }
                    }
                    gout[c] += s2;
                }
                __syncthreads();
            }
        }

        // contracted block (with its degeneracy factor) to LDS
        double fac = diagA ? 0.5 : 1.0;
        if (a.P.shA[q] == a.P.shB[q]) fac *= 0.5;
        if (p == q && a.diag) fac *= 0.5;
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int f = tid + NT * c;
            if (f < NF) sblk[f] = gout[c] * fac;
        }
        const int oc = a.P.cartA[q], od = a.P.cartB[q];
        const int cp = xcc_copy(a.ncopy);
#define BLK(ia, ib2, ic, id) sblk[(((ia) * NB + (ib2)) * NC + (ic)) * ND + (id)]
        for (int s = 0; s < a.n_j; ++s) {
            const double* __restrict__ D = a.Dj + s * a.slot + base;
            double* V = a.J + cp * a.jcopy + s * a.slot + base;
            double* dab = sD;
            double* dcd = sD + NA * NB;
            for (int e = tid; e < NA * NB; e += NT) dab[e] = D[(oa + e / NB) * ld + ob + e % NB];
            for (int e = tid; e < NC * ND; e += NT) dcd[e] = D[(oc + e / ND) * ld + od + e % ND];
            __syncthreads();
            for (int e = tid; e < NC * ND; e += NT) {
                const int ic = e / ND, id = e % ND;
                double acc = 0.0;
                for (int ia = 0; ia < NA; ++ia)
                    for (int ib2 = 0; ib2 < NB; ++ib2) acc = fma(BLK(ia, ib2, ic, id), dab[ia * NB + ib2], acc);
                atomic_add(V + (oc + ic) * ld + od + id, acc);
            }
            for (int e = tid; e < NA * NB; e += NT) {
                const int ia = e / NB, ib2 = e % NB;
                double acc = 0.0;
                for (int ic = 0; ic < NC; ++ic)
                    for (int id = 0; id < ND; ++id) acc = fma(BLK(ia, ib2, ic, id), dcd[ic * ND + id], acc);
                if (jab_lds)
                    sJab[s * NA * NB + e] += acc;
                else
                    atomic_add(V + (oa + ia) * ld + ob + ib2, acc);
            }
            __syncthreads();
        }
        for (int s = 0; s < a.n_k; ++s) {
            const double* __restrict__ D = a.Dk + s * a.slot + base;
            double* V = a.K + cp * a.kcopy + s * a.slot + base;
            double* dbc = sD;
            double* dac = dbc + NB * NC;
            double* dbd = dac + NA * NC;
            double* dad = dbd + NB * ND;
            for (int e = tid; e < NB * NC; e += NT) dbc[e] = D[(ob + e / NC) * ld + oc + e % NC];
            for (int e = tid; e < NA * NC; e += NT) dac[e] = D[(oa + e / NC) * ld + oc + e % NC];
            for (int e = tid; e < NB * ND; e += NT) dbd[e] = D[(ob + e / ND) * ld + od + e % ND];
            for (int e = tid; e < NA * ND; e += NT) dad[e] = D[(oa + e / ND) * ld + od + e % ND];
            __syncthreads();
            for (int e = tid; e < NA * ND; e += NT) {          // K'[a,d] += sum_bc v D[b,c]
                const int ia = e / ND, id = e % ND;
                double acc = 0.0;
                for (int ib2 = 0; ib2 < NB; ++ib2)
                    for (int ic = 0; ic < NC; ++ic) acc = fma(BLK(ia, ib2, ic, id), dbc[ib2 * NC + ic], acc);
                atomic_add(V + (oa + ia) * ld + od + id, acc);
            }
            for (int e = tid; e < NB * ND; e += NT) {          // K'[b,d] += sum_ac v D[a,c]
                const int ib2 = e / ND, id = e % ND;
                double acc = 0.0;
                for (int ia = 0; ia < NA; ++ia)
                    for (int ic = 0; ic < NC; ++ic) acc = fma(BLK(ia, ib2, ic, id), dac[ia * NC + ic], acc);
                atomic_add(V + (ob + ib2) * ld + od + id, acc);
            }
            for (int e = tid; e < NA * NC; e += NT) {          // K'[a,c] += sum_bd v D[b,d]
                const int ia = e / NC, ic = e % NC;
                double acc = 0.0;
                for (int ib2 = 0; ib2 < NB; ++ib2)
                    for (int id = 0; id < ND; ++id) acc = fma(BLK(ia, ib2, ic, id), dbd[ib2 * ND + id], acc);
                atomic_add(V + (oa + ia) * ld + oc + ic, acc);
            }
            for (int e = tid; e < NB * NC; e += NT) {          // K'[b,c] += sum_ad v D[a,d]
                const int ib2 = e / NC, ic = e % NC;
                double acc = 0.0;
                for (int ia = 0; ia < NA; ++ia)
                    for (int id = 0; id < ND; ++id) acc = fma(BLK(ia, ib2, ic, id), dad[ia * ND + id], acc);
                atomic_add(V + (ob + ib2) * ld + oc + ic, acc);
            }
            __syncthreads();
        }
#undef BLK
        __syncthreads();   // the block region is rewritten by the next quartet
    }
    if (jab_lds) {
        const int cp = xcc_copy(a.ncopy);
        for (int s = 0; s < a.n_j; ++s) {
            double* V = a.J + cp * a.jcopy + s * a.slot + base;
            for (int e = tid; e < NA * NB; e += NT) {
                const double v = sJab[s * NA * NB + e];
                if (v != 0.0) atomic_add(V + (oa + e / NB) * ld + ob + e % NB, v);
            }
        }
    }
    if (tid == 0 && a.counters && nq_blk) {
        atomicAdd(a.counters, nq_blk);
        atomicAdd(a.counters + 1, np_blk);
    }
}

template <int LA, int LB, int LC, int LD>
int launch_jk_coop(int nblocks, const JKArgs& a, hipStream_t st)
{
    constexpr int NT = CoopShape<LA, LB, LC, LD>::NT;
    hipLaunchKernelGGL((k_jk_coop<LA, LB, LC, LD>), dim3(nblocks * TILE), dim3(NT), 0, st, a);
    return (int)hipGetLastError();
}

// Number of slices for a class: keep <= ~54 components per lane.
__host__ __device__ constexpr int jk_split(int la, int lb, int lc, int ld)
{
    const int na = ncart(la), nb = ncart(lb), nc = ncart(lc), nd = ncart(ld);
    const int nf = na * nb * nc * nd;
    if (nf <= 81) return 1;
    // (dp|pp): 81 per lane; at 54 its registers already exceeded 256 (one wave
    // per SIMD), so the third slice only repeated the recurrences: C30H62
    // cc-pVDZ class 3.40 -> 2.40 ms on MI355X (384 registers, no spills)
    if (la == 2 && lb == 1 && lc == 1 && ld == 1) return 2;
    for (int s = 2; s <= na; ++s)   // fewest blocks of a components with at most 54 components each
        if (na % s == 0 && nf / s <= 54) return s;
    if (nf / (na * nb) <= 54) return na * nb;
    return na * nb * nc;
}

template <int LA, int LB, int LC, int LD, int S>
int launch_jk_slice(int nblocks, const JKArgs& a, hipStream_t st)
{
    constexpr int SPLIT = jk_split(LA, LB, LC, LD);
    hipLaunchKernelGGL((k_jk<LA, LB, LC, LD, SPLIT, S>), dim3(nblocks), dim3(64), 0, st, a);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// Row kernel.  Every K' contribution of a quartet (ab|cd) lands in rows a
// and b of K' (K'[a,d], K'[a,c], K'[b,d], K'[b,c]).  A block of 256 threads
// owns one bra pair and one chunk of the surviving kets of the ket class:
// the Schwarz prefix of every primitive bucket is found by binary search,
// the prefixes are concatenated, and the block's waves claim 64 consecutive
// kets of its chunk at a time from a shared counter.  Chunks of equal size keep the blocks'
// run times even whatever the bra pair's ket count.  The block accumulates its K' rows and its J'[a,b] block
// in LDS with LDS atomics and flushes the non-zero entries to HBM once; only
// J'[c,d] is added per quartet.  LDS holds n_k (NA + NB) ld + n_j NA NB
// doubles (dynamic shared memory).
__device__ __forceinline__ void lds_add(double* p, double v)
{
    atomicAdd(p, v);
}

// K' updates of the bra's B rows: in LDS, or (KBM, the rows of large
// molecules that do not fit beside the A rows) memory-side atomics into the
// accumulator copy.
template <bool KBM>
__device__ __forceinline__ void kb_add(double* p, double v)
{
    if constexpr (KBM) atomic_add(p, v);
    else lds_add(p, v);
}

template <int LA, int LB, int LC, int LD, int SPLIT, int S, typename R, bool RSH, bool KBM = false>
__device__ __forceinline__ void row_digest(const JKArgs& a, const QuartetCtx& c, double* kr, double* jr,
                                           int krow, long long base, int ld, int oa, int ob, int cp)
{
    using SR = SliceRange<LA, LB, LC, LD, SPLIT, S>;
    constexpr int NA = SR::NA, NB = SR::NB, ND = SR::ND;
    constexpr int IA0 = SR::IA0, IA1 = SR::IA1, IB0 = SR::IB0, IB1 = SR::IB1, IC0 = SR::IC0, IC1 = SR::IC1;
    constexpr int MA = SR::MA, MB = SR::MB, MC = SR::MC;
    double gd[SR::NFL];
    eri_slice<LA, LB, LC, LD, SPLIT, S, R, true, RSH, RowRys<LA, LB, LC, LD>::lds>(a, c, gd);
    const int oc = c.qvalid ? a.P.cartA[c.q] : 0, od = c.qvalid ? a.P.cartB[c.q] : 0;
    const int lane = threadIdx.x & 63;
    for (int s = 0; s < a.n_j; ++s) {
        const double* __restrict__ D = a.Dj + s * a.slot + base;
        double* V = a.J + cp * a.jcopy + s * a.slot + base;
        // J'[c,d] += sum_ab v Ds[a,b]  (per quartet, global)
        double dab[MA][MB];
#pragma unroll
        for (int i = 0; i < MA; ++i)
#pragma unroll
            for (int j = 0; j < MB; ++j) dab[i][j] = D[(oa + IA0 + i) * ld + ob + IB0 + j];
        if (c.alive) {
#pragma unroll
            for (int ic = IC0; ic < IC1; ++ic)
#pragma unroll
                for (int id = 0; id < ND; ++id) {
                    double acc = 0.0;
#pragma unroll
                    for (int ia = IA0; ia < IA1; ++ia)
#pragma unroll
                        for (int ib2 = IB0; ib2 < IB1; ++ib2)
                            acc = fma(GOUT(ia, ib2, ic, id), dab[ia - IA0][ib2 - IB0], acc);
                    atomic_add(V + (oc + ic) * ld + od + id, acc);
                }
        }
        // J'[a,b] += sum_cd v Ds[c,d]  (summed over the wave, then into LDS)
        double dcd[MC][ND];
#pragma unroll
        for (int i = 0; i < MC; ++i)
#pragma unroll
            for (int j = 0; j < ND; ++j) dcd[i][j] = c.alive ? D[(oc + IC0 + i) * ld + od + j] : 0.0;
#pragma unroll
        for (int ia = IA0; ia < IA1; ++ia)
#pragma unroll
            for (int ib2 = IB0; ib2 < IB1; ++ib2) {
                double acc = 0.0;
#pragma unroll
                for (int ic = IC0; ic < IC1; ++ic)
#pragma unroll
                    for (int id = 0; id < ND; ++id) acc = fma(GOUT(ia, ib2, ic, id), dcd[ic - IC0][id], acc);
                acc = wave_sum_last(acc);
                if (lane == 63 && acc != 0.0) lds_add(jr + s * NA * NB + ia * NB + ib2, acc);
            }
    }
    if (!c.alive) return;
    for (int s = 0; s < a.n_k; ++s) {
        const double* __restrict__ D = a.Dk + s * a.slot + base;
        double* KA = kr + s * krow;              // rows a: K'[a, :]
        double* KB = KBM ? a.K + cp * a.kcopy + s * a.slot + base + (long long)ob * ld   // rows b: K'[b, :]
                         : KA + NA * ld;
        // K'[a,d] += sum_bc v D[b,c];  K'[b,d] += sum_ac v D[a,c]
        // K'[a,c] += sum_bd v D[b,d];  K'[b,c] += sum_ad v D[a,d]
        double dbc[MB][MC], dac[MA][MC], dbd[MB][ND], dad[MA][ND];
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
            for (int j = 0; j < MC; ++j) dbc[i][j] = D[(ob + IB0 + i) * ld + oc + IC0 + j];
#pragma unroll
        for (int i = 0; i < MA; ++i)
#pragma unroll
            for (int j = 0; j < MC; ++j) dac[i][j] = D[(oa + IA0 + i) * ld + oc + IC0 + j];
#pragma unroll
        for (int i = 0; i < MB; ++i)
#pragma unroll
            for (int j = 0; j < ND; ++j) dbd[i][j] = D[(ob + IB0 + i) * ld + od + j];
#pragma unroll
        for (int i = 0; i < MA; ++i)
#pragma unroll
            for (int j = 0; j < ND; ++j) dad[i][j] = D[(oa + IA0 + i) * ld + od + j];
#pragma unroll
        for (int ia = IA0; ia < IA1; ++ia)
#pragma unroll
            for (int id = 0; id < ND; ++id) {
                double acc = 0.0;
#pragma unroll
                for (int ib2 = IB0; ib2 < IB1; ++ib2)
#pragma unroll
                    for (int ic = IC0; ic < IC1; ++ic) acc = fma(GOUT(ia, ib2, ic, id), dbc[ib2 - IB0][ic - IC0], acc);
                lds_add(KA + ia * ld + od + id, acc);
            }
#pragma unroll
        for (int ib2 = IB0; ib2 < IB1; ++ib2)
#pragma unroll
            for (int id = 0; id < ND; ++id) {
                double acc = 0.0;
#pragma unroll
                for (int ia = IA0; ia < IA1; ++ia)
#pragma unroll
                    for (int ic = IC0; ic < IC1; ++ic) acc = fma(GOUT(ia, ib2, ic, id), dac[ia - IA0][ic - IC0], acc);
                kb_add<KBM>(KB + ib2 * ld + od + id, acc);
            }
#pragma unroll
        for (int ia = IA0; ia < IA1; ++ia)
#pragma unroll
            for (int ic = IC0; ic < IC1; ++ic) {
                double acc = 0.0;
#pragma unroll
                for (int ib2 = IB0; ib2 < IB1; ++ib2)
#pragma unroll
                    for (int id = 0; id < ND; ++id) acc = fma(GOUT(ia, ib2, ic, id), dbd[ib2 - IB0][id], acc);
                lds_add(KA + ia * ld + oc + ic, acc);
            }
#pragma unroll
        for (int ib2 = IB0; ib2 < IB1; ++ib2)
#pragma unroll
            for (int ic = IC0; ic < IC1; ++ic) {
                double acc = 0.0;
#pragma unroll
                for (int ia = IA0; ia < IA1; ++ia)
#pragma unroll
                    for (int id = 0; id < ND; ++id) acc = fma(GOUT(ia, ib2, ic, id), dad[ia - IA0][id], acc);
                kb_add<KBM>(KB + ib2 * ld + oc + ic, acc);
            }
    }
}

// Minimum waves per SIMD the compiler must allow for the row kernel of a
// class (caps its VGPR budget at 512 / value).  Two for the (dp|p s) and
// (ds|pp) classes, whose unconstrained allocation leaves one wave per SIMD
// (measured on MI355X: 10-15% faster); the others run best unconstrained.
#ifndef MJK_ROW_WPE
#define MJK_ROW_WPE 1
#endif
template <int LA, int LB, int LC, int LD>
struct RowOccupancy {
    static constexpr int value =
        (LA == 2 && LB == 1 && LC == 1 && LD == 0) || (LA == 2 && LB == 0 && LC == 1 && LD == 1) ? 2 : MJK_ROW_WPE;
};

// Row kernels.  RSH: the variant for the range-separated operators
// (a.omega != 0); the Coulomb variant keeps its registers free of that path.
template <int LA, int LB, int LC, int LD, int SPLIT, int S, bool RSH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RSH ? 1 : RowOccupancy<LA, LB, LC, LD>::value)))
void k_jk_row(const JKArgs a)
{
    constexpr bool KBM = false;
#include "jk_row_body.inc"
}

// Row kernel of classes whose full K' rows exceed a workgroup's LDS in large
// molecules (f-shell bras): the A rows in LDS, the B rows' updates as
// memory-side atomics.
template <int LA, int LB, int LC, int LD, int SPLIT, int S>
__global__ __launch_bounds__(256) void k_jk_row_m(const JKArgs a)
{
    constexpr bool KBM = true, RSH = false;
#include "jk_row_body.inc"
}

template <int LA, int LB, int LC, int LD, int S, bool RSH>
int launch_jk_row(int nblocks, const JKArgs& a, size_t shmem, hipStream_t st)
{
    constexpr int SPLIT = jk_split(LA, LB, LC, LD);
    hipLaunchKernelGGL((k_jk_row<LA, LB, LC, LD, SPLIT, S, RSH>), dim3(nblocks), dim3(256), shmem, st, a);
    return (int)hipGetLastError();
}

template <int LA, int LB, int LC, int LD, int S>
int launch_jk_row_m(int nblocks, const JKArgs& a, size_t shmem, hipStream_t st)
{
    constexpr int SPLIT = jk_split(LA, LB, LC, LD);
    hipLaunchKernelGGL((k_jk_row_m<LA, LB, LC, LD, SPLIT, S>), dim3(nblocks), dim3(256), shmem, st, a);
    return (int)hipGetLastError();
}



// ---------------------------------------------------------------------------
// Row kernel over merged pair groups.  A merged pair holds a generally
// contracted s shell as one unit of two contractions sharing their
// primitives; a pair group has CA x CB contraction combinations (CA, CB in
// {1, 2}).  One primitive-quartet loop computes the primitive integral block
// once and adds it, weighted by the coefficient products of each combination,
// into the NIJ NKL contracted blocks of the quartet.  Contraction ci of a unit
// sits ci ncart(l) Cartesian functions after its first one, so the bra's K'
// rows are CA NA + CB NB consecutive rows of the arena.  No slices, no
// range-separated variant: the classes that use it are small.
template <int LA, int LB, int LC, int LD, int CA, int CB, int CC, int CD>
struct GcShape {
    static constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    static constexpr int NF = NA * NB * NC * ND;
    static constexpr int NIJ = CA * CB, NKL = CC * CD, NCOMB = NIJ * NKL;
    static constexpr int NRA = CA * NA, NRB = CB * NB;   // bra K' rows of A and of B
};

// Coefficient products of a pair's combinations: a plain pair (N = 1) has
// them folded into its K; a merged pair's are in its W record.
template <int N>
__device__ __forceinline__ void combo_weights(const double4& w, double (&out)[N])
{
    if constexpr (N == 1) {
        out[0] = 1.0;
    } else {
        out[0] = w.x;
        out[1] = w.y;
        if constexpr (N > 2) {
            out[2] = w.z;
            out[3] = w.w;
        }
    }
}

template <int LA, int LB, int LC, int LD, int CA, int CB, int CC, int CD, typename R>
__device__ __forceinline__ void eri_gc(const JKArgs& a, const QuartetCtx& c,
                                       double (&gd)[GcShape<LA, LB, LC, LD, CA, CB, CC, CD>::NCOMB]
                                                   [GcShape<LA, LB, LC, LD, CA, CB, CC, CD>::NF])
{
    using G = GcShape<LA, LB, LC, LD, CA, CB, CC, CD>;
    constexpr int NR = (LA + LB + LC + LD) / 2 + 1;
    constexpr int NF = G::NF, NIJ = G::NIJ, NKL = G::NKL;
    R gout[NIJ * NKL][NF];
#pragma unroll
    for (int k = 0; k < NIJ * NKL; ++k)
#pragma unroll
        for (int f = 0; f < NF; ++f) gout[k][f] = R(0);
    for (int ib = 0; ib < c.npb; ++ib) {
        const double4 Pp = uload(a.P.ppP + c.pb0 + ib);
        const double2 Kp = uload(a.P.ppK + c.pb0 + ib);
        double wp[NIJ];
        combo_weights<NIJ>(NIJ > 1 ? uload(a.P.ppW + c.pb0 + ib) : make_double4(1, 0, 0, 0), wp);
        for (int ik = 0; ik < c.npk; ++ik) {
            double T;
            const int t0 = c.tq + ik * c.ts;
            const PrimQuartet m = prim_quartet(Pp, Kp, a.P.tP[t0], a.P.tK[t0], c, T);
            double wq[NKL];
            combo_weights<NKL>(NKL > 1 ? a.P.tW[t0] : make_double4(1, 0, 0, 0), wq);
            R tmp[NF];
#pragma unroll
            for (int f = 0; f < NF; ++f) tmp[f] = R(0);
            using Rys = RysFor<NR, LA + LB + LC + LD, RowRys<LA, LB, LC, LD>::lds>;
            const typename Rys::type rys(Rys::lds > 0 ? c.rtab : a.rys, T);
#pragma unroll 1
            for (int r = 0; r < NR; ++r) {
                double rx, rw;
                rys.root(r, rx, rw);
                add_root<LA, LB, LC, LD, 1, 0, R>(m, rx, rw, c, tmp);
            }
#pragma unroll
            for (int i = 0; i < NIJ; ++i)
#pragma unroll
                for (int k = 0; k < NKL; ++k) {
                    const R w = R(wp[i] * wq[k]);
#pragma unroll
                    for (int f = 0; f < NF; ++f) gout[i * NKL + k][f] = fma(w, tmp[f], gout[i * NKL + k][f]);
                }
        }
    }
#pragma unroll
    for (int k = 0; k < NIJ * NKL; ++k)
#pragma unroll
        for (int f = 0; f < NF; ++f) gd[k][f] = double(gout[k][f]) * c.fac;
}

// Merged patterns whose K digestion loads its four D blocks one by one: the
// registers of the blocks not yet needed then stay free, and the kernel
// drops below 168 VGPRs (three waves per SIMD).  Measured on MI355X
// (C30H62 cc-pVDZ, each class alone): (ps|ps) with a merged bra s 1.14x and
// (pp|ss) with a merged ket s 1.20x faster; (ds|ss) and (pp|ps) patterns
// lost 3-19% to the serialised loads and keep them together.
template <int LA, int LB, int LC, int LD, int CA, int CB, int CC, int CD>
struct GcStagedLoads {
    static constexpr bool value =
        (LA == 1 && LB == 0 && LC == 1 && LD == 0 && CA == 1 && CB == 2 && CC == 1 && CD == 1) ||
        (LA == 1 && LB == 1 && LC == 0 && LD == 0 && CA == 1 && CB == 1 && CC == 2 && CD == 1);
};

template <int LA, int LB, int LC, int LD, int CA, int CB, int CC, int CD, typename R>
__device__ __forceinline__ void row_digest_gc(const JKArgs& a, const QuartetCtx& c, double* kr, double* jr,
                                              int krow, long long base, int ld, int oa, int ob, int cp)
{
    using G = GcShape<LA, LB, LC, LD, CA, CB, CC, CD>;
    constexpr int NA = G::NA, NB = G::NB, NC = G::NC, ND = G::ND, NF = G::NF;
    constexpr int NKL = G::NKL, NRA = G::NRA, NRB = G::NRB;
    double gd[G::NCOMB][NF];
    eri_gc<LA, LB, LC, LD, CA, CB, CC, CD, R>(a, c, gd);
    const int oc = c.qvalid ? a.P.cartA[c.q] : 0, od = c.qvalid ? a.P.cartB[c.q] : 0;
    const int lane = threadIdx.x & 63;
#define GV(ci, cj, ck, cl, ia, ib2, ic, id)                                                         \
    gd[((ci) * CB + (cj)) * NKL + (ck) * CD + (cl)][(((ia) * NB + (ib2)) * NC + (ic)) * ND + (id)]
    for (int s = 0; s < a.n_j; ++s) {
        const double* __restrict__ D = a.Dj + s * a.slot + base;
        double* V = a.J + cp * a.jcopy + s * a.slot + base;
        // J'[c,d] += sum_ab v Ds[a,b]  (per quartet and ket combination, global)
        double dab[NRA][NRB];
#pragma unroll
        for (int i = 0; i < NRA; ++i)
#pragma unroll
            for (int j = 0; j < NRB; ++j) dab[i][j] = D[(oa + i) * ld + ob + j];
        if (c.alive) {
#pragma unroll
            for (int ck = 0; ck < CC; ++ck)
#pragma unroll
            for (int cl = 0; cl < CD; ++cl)
#pragma unroll
            for (int ic = 0; ic < NC; ++ic)
#pragma unroll
            for (int id = 0; id < ND; ++id) {
                double acc = 0.0;
#pragma unroll
                for (int ci = 0; ci < CA; ++ci)
#pragma unroll
                for (int cj = 0; cj < CB; ++cj)
#pragma unroll
                for (int ia = 0; ia < NA; ++ia)
#pragma unroll
                for (int ib2 = 0; ib2 < NB; ++ib2)
                    acc = fma(GV(ci, cj, ck, cl, ia, ib2, ic, id), dab[ci * NA + ia][cj * NB + ib2], acc);
                atomic_add(V + (oc + ck * NC + ic) * ld + od + cl * ND + id, acc);
            }
        }
        // J'[a,b] += sum_cd v Ds[c,d]  (summed over the wave, then into LDS)
        double dcd[CC * NC][CD * ND];
#pragma unroll
        for (int i = 0; i < CC * NC; ++i)
#pragma unroll
            for (int j = 0; j < CD * ND; ++j) dcd[i][j] = c.alive ? D[(oc + i) * ld + od + j] : 0.0;
#pragma unroll
        for (int ci = 0; ci < CA; ++ci)
#pragma unroll
        for (int cj = 0; cj < CB; ++cj)
#pragma unroll
        for (int ia = 0; ia < NA; ++ia)
#pragma unroll
        for (int ib2 = 0; ib2 < NB; ++ib2) {
            double acc = 0.0;
#pragma unroll
            for (int ck = 0; ck < CC; ++ck)
#pragma unroll
            for (int cl = 0; cl < CD; ++cl)
#pragma unroll
            for (int ic = 0; ic < NC; ++ic)
#pragma unroll
            for (int id = 0; id < ND; ++id)
                acc = fma(GV(ci, cj, ck, cl, ia, ib2, ic, id), dcd[ck * NC + ic][cl * ND + id], acc);
            acc = wave_sum_last(acc);
            if (lane == 63 && acc != 0.0) lds_add(jr + s * NRA * NRB + (ci * NA + ia) * NRB + cj * NB + ib2, acc);
        }
    }
    if (!c.alive) return;
    for (int s = 0; s < a.n_k; ++s) {
        const double* __restrict__ D = a.Dk + s * a.slot + base;
        double* KA = kr + s * krow;              // rows a (all contractions of A): K'[a, :]
        double* KB = KA + NRA * ld;              // rows b: K'[b, :]
        // the four D blocks: loaded together, or (GcStagedLoads) each just before
        // its products, the scheduling barriers keeping the loads apart
        constexpr bool STAGED = GcStagedLoads<LA, LB, LC, LD, CA, CB, CC, CD>::value;
        double dbc[NRB][CC * NC], dac[NRA][CC * NC], dbd[NRB][CD * ND], dad[NRA][CD * ND];
#pragma unroll
        for (int i = 0; i < NRB; ++i)
#pragma unroll
            for (int j = 0; j < CC * NC; ++j) dbc[i][j] = D[(ob + i) * ld + oc + j];
        if constexpr (!STAGED) {
#pragma unroll
            for (int i = 0; i < NRA; ++i)
#pragma unroll
                for (int j = 0; j < CC * NC; ++j) dac[i][j] = D[(oa + i) * ld + oc + j];
#pragma unroll
            for (int i = 0; i < NRB; ++i)
#pragma unroll
                for (int j = 0; j < CD * ND; ++j) dbd[i][j] = D[(ob + i) * ld + od + j];
#pragma unroll
            for (int i = 0; i < NRA; ++i)
#pragma unroll
                for (int j = 0; j < CD * ND; ++j) dad[i][j] = D[(oa + i) * ld + od + j];
        }
        // K'[a,d] += sum_bc v D[b,c]
#pragma unroll
        for (int ci = 0; ci < CA; ++ci)
#pragma unroll
        for (int cl = 0; cl < CD; ++cl)
#pragma unroll
        for (int ia = 0; ia < NA; ++ia)
#pragma unroll
        for (int id = 0; id < ND; ++id) {
            double acc = 0.0;
#pragma unroll
            for (int cj = 0; cj < CB; ++cj)
#pragma unroll
            for (int ck = 0; ck < CC; ++ck)
#pragma unroll
            for (int ib2 = 0; ib2 < NB; ++ib2)
#pragma unroll
            for (int ic = 0; ic < NC; ++ic)
                acc = fma(GV(ci, cj, ck, cl, ia, ib2, ic, id), dbc[cj * NB + ib2][ck * NC + ic], acc);
            lds_add(KA + (ci * NA + ia) * ld + od + cl * ND + id, acc);
        }
        if constexpr (STAGED) {
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int i = 0; i < NRA; ++i)
#pragma unroll
                for (int j = 0; j < CC * NC; ++j) dac[i][j] = D[(oa + i) * ld + oc + j];
        }
        // K'[b,d] += sum_ac v D[a,c]
#pragma unroll
        for (int cj = 0; cj < CB; ++cj)
#pragma unroll
        for (int cl = 0; cl < CD; ++cl)
#pragma unroll
        for (int ib2 = 0; ib2 < NB; ++ib2)
#pragma unroll
        for (int id = 0; id < ND; ++id) {
            double acc = 0.0;
#pragma unroll
            for (int ci = 0; ci < CA; ++ci)
#pragma unroll
            for (int ck = 0; ck < CC; ++ck)
#pragma unroll
            for (int ia = 0; ia < NA; ++ia)
#pragma unroll
            for (int ic = 0; ic < NC; ++ic)
                acc = fma(GV(ci, cj, ck, cl, ia, ib2, ic, id), dac[ci * NA + ia][ck * NC + ic], acc);
            lds_add(KB + (cj * NB + ib2) * ld + od + cl * ND + id, acc);
        }
        if constexpr (STAGED) {
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int i = 0; i < NRB; ++i)
#pragma unroll
                for (int j = 0; j < CD * ND; ++j) dbd[i][j] = D[(ob + i) * ld + od + j];
        }
        // K'[a,c] += sum_bd v D[b,d]
#pragma unroll
        for (int ci = 0; ci < CA; ++ci)
#pragma unroll
        for (int ck = 0; ck < CC; ++ck)
#pragma unroll
        for (int ia = 0; ia < NA; ++ia)
#pragma unroll
        for (int ic = 0; ic < NC; ++ic) {
            double acc = 0.0;
#pragma unroll
            for (int cj = 0; cj < CB; ++cj)
#pragma unroll
            for (int cl = 0; cl < CD; ++cl)
#pragma unroll
            for (int ib2 = 0; ib2 < NB; ++ib2)
#pragma unroll
            for (int id = 0; id < ND; ++id)
                acc = fma(GV(ci, cj, ck, cl, ia, ib2, ic, id), dbd[cj * NB + ib2][cl * ND + id], acc);
            lds_add(KA + (ci * NA + ia) * ld + oc + ck * NC + ic, acc);
        }
        if constexpr (STAGED) {
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int i = 0; i < NRA; ++i)
#pragma unroll
                for (int j = 0; j < CD * ND; ++j) dad[i][j] = D[(oa + i) * ld + od + j];
        }
        // K'[b,c] += sum_ad v D[a,d]
#pragma unroll
        for (int cj = 0; cj < CB; ++cj)
#pragma unroll
        for (int ck = 0; ck < CC; ++ck)
#pragma unroll
        for (int ib2 = 0; ib2 < NB; ++ib2)
#pragma unroll
        for (int ic = 0; ic < NC; ++ic) {
            double acc = 0.0;
#pragma unroll
            for (int ci = 0; ci < CA; ++ci)
#pragma unroll
            for (int cl = 0; cl < CD; ++cl)
#pragma unroll
            for (int ia = 0; ia < NA; ++ia)
#pragma unroll
            for (int id = 0; id < ND; ++id)
                acc = fma(GV(ci, cj, ck, cl, ia, ib2, ic, id), dad[ci * NA + ia][cl * ND + id], acc);
            lds_add(KB + (cj * NB + ib2) * ld + oc + ck * NC + ic, acc);
        }
    }
#undef GV
}

template <int LA, int LB, int LC, int LD, int CA, int CB, int CC, int CD>
__global__ __launch_bounds__(256) void k_jk_row_gc(const JKArgs a)
{
    using G = GcShape<LA, LB, LC, LD, CA, CB, CC, CD>;
    constexpr int NRA = G::NRA, NRB = G::NRB, NROW = NRA + NRB;
    extern __shared__ double lds_rows[];
    __shared__ int s_start[ROW_MAXBUCKET], s_pre[ROW_MAXBUCKET + 1];
    __shared__ int s_next;
    const int tid = threadIdx.x, lane = tid & 63;
    using RT = RowRys<LA, LB, LC, LD>;
    __shared__ double s_rys[RT::size > 0 ? RT::size : 1];
    if constexpr (RT::size > 0) {
        rys_lds_fill<RT::nr>(s_rys, a.rys, tid, 256);
        __syncthreads();
    }
    __shared__ int s_task;
    unsigned int cnt[4] = {0, 0, 0, 0};   // per-lane work tallies: a lane's share of a launch stays far below 2^32
    const bool persist = a.thead != nullptr;
    const int ntask = persist ? (a.tcount ? *a.tcount : a.ntask) : 0;
    for (int it = 0;; ++it) {
    int tb;
    if (persist) {
        if (tid == 0) s_task = atomicAdd(a.thead, 1);
        __syncthreads();
        const int ti = __builtin_amdgcn_readfirstlane(s_task);
        __syncthreads();
        if (ti >= ntask) break;   // block-uniform
        tb = a.tlist ? uload(a.tlist + ti) : ti;
    } else {
        if (it > 0) break;
        tb = blockIdx.x;
    }
    const int4 task = uload(a.rt + tb);
    const int p = task.x;
    const int mol = uload(a.pmol + p);
    if (a.mol_active && !uload(a.mol_active + mol)) continue;   // block-uniform
    const int ld = uload(a.M.ld + mol);
    const long long base = uload(a.M.dbase + mol);
    const int krow = NROW * ld;
    double* kr = lds_rows;
    double* jr = lds_rows + a.n_k * krow;
    const int nz = a.n_k * krow + a.n_j * NRA * NRB;
    for (int e = tid; e < nz; e += 256) lds_rows[e] = 0.0;
    const double qp = uload(a.P.rA + p).w;
    if (tid < a.nbucket) {
        const int2 r = a.rr[(long long)task.w * a.nbucket + tid];
        s_start[tid] = r.x;
        s_pre[tid + 1] = r.y - r.x;
    }
    __syncthreads();
    if (tid == 0) {
        s_pre[0] = 0;
        for (int k = 1; k <= a.nbucket; ++k) s_pre[k] += s_pre[k - 1];
        s_next = task.y;
    }
    __syncthreads();
    const int iend = min(s_pre[a.nbucket], task.z);

    QuartetCtx c;
    c.p = p;
    c.mol = mol;
    c.pvalid = true;
    c.rtab = s_rys;
    c.rA = uload(a.P.rA + p);
    c.AB = uload(a.P.rAB + p);
    const int pbo = uload(a.rtp + tb);
    c.pb0 = uload(a.P.prim0 + p) + pbo;
    c.npb = min(a.rprim, uload(a.P.nprim + p) - pbo);
    // a merged quartet counts as its NCOMB quartets of segmented shells, once
    const unsigned long long quartet_tally = pbo == 0 ? G::NCOMB : 0;
    const bool diagA = uload(a.P.shA + p) == uload(a.P.shB + p);
    const int oa = uload(a.P.cartA + p), ob = uload(a.P.cartB + p);
    const int cp = xcc_copy(a.ncopy);
    for (;;) {
        int i0 = 0;
        if (lane == 0) i0 = atomicAdd(&s_next, 64);
        i0 = __shfl(i0, 0);
        if (i0 >= iend) break;   // wave-uniform
        const int idx = i0 + lane;
        c.qvalid = idx < iend;
        c.q = 0;
        if (c.qvalid) {
            int lo = 0, hi = a.nbucket;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_pre[mid] <= idx) lo = mid; else hi = mid;
            }
            c.q = s_start[lo] + idx - s_pre[lo];
        }
        c.alive = c.qvalid;
        c.rC = c.CD = make_double4(0, 0, 0, 0);
        c.pk0 = c.npk = 0;
        c.tq = c.ts = 0;
        bool small = true;
        if (c.alive) {
            c.rC = a.P.rA[c.q];
            const double qpq = qp * c.rC.w;
            if (a.screen || a.fp32_tol > 0) {
                const double dmax = quartet_dmax(a, mol, p, c.q);
                if (a.screen) c.alive = qpq > a.cutoff && dmax * qpq > a.cutoff;
                small = dmax * qpq < a.fp32_tol;
            } else {
                small = false;
            }
        }
        c.fac = 0.0;
        if (c.alive) {
            c.pk0 = a.P.prim0[c.q];
            c.npk = a.P.nprim[c.q];
            c.tq = a.P.tpos[c.q];
            c.ts = a.P.tstride[c.q];
            c.CD = a.P.rAB[c.q];
            double f = diagA ? 0.5 : 1.0;
            if (a.P.shA[c.q] == a.P.shB[c.q]) f *= 0.5;
            if (c.q == p && a.diag) f *= 0.5;
            c.fac = f;
        }
        if (!__ballot(c.alive)) continue;   // wave-uniform
        const bool use32 = a.fp32_tol > 0 && __ballot(c.alive && !small) == 0;
        if (c.alive) {
            cnt[use32 ? 2 : 0] += (unsigned)quartet_tally;
            cnt[use32 ? 3 : 1] += (unsigned)(c.npb * c.npk);
        }
        if (use32) row_digest_gc<LA, LB, LC, LD, CA, CB, CC, CD, float>(a, c, kr, jr, krow, base, ld, oa, ob, cp);
        else row_digest_gc<LA, LB, LC, LD, CA, CB, CC, CD, double>(a, c, kr, jr, krow, base, ld, oa, ob, cp);
    }
    __syncthreads();
    // flush: rows r < NRA are A's rows oa + r, the others B's rows ob + r - NRA
    for (int e = tid; e < a.n_k * krow; e += 256) {
        const double v = kr[e];
        if (v == 0.0) continue;
        const int s = e / krow;
        const int r = (e - s * krow) / ld;
        const int col = e - s * krow - r * ld;
        const int row = r < NRA ? oa + r : ob + (r - NRA);
        atomic_add(a.K + cp * a.kcopy + s * a.slot + base + (long long)row * ld + col, v);
    }
    for (int e = tid; e < a.n_j * NRA * NRB; e += 256) {
        const double v = jr[e];
        if (v == 0.0) continue;
        const int s = e / (NRA * NRB);
        const int ra = (e / NRB) % NRA, rb = e % NRB;
        atomic_add(a.J + cp * a.jcopy + s * a.slot + base + (long long)(oa + ra) * ld + ob + rb, v);
    }
    __syncthreads();   // the next task rewrites the LDS rows
    }
    if (a.counters) {   // one atomic per counter and block
        __shared__ unsigned long long s_cnt[4];
        if (tid < 4) s_cnt[tid] = 0;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            unsigned long long v = cnt[k];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
            if (lane == 0 && v) atomicAdd(s_cnt + k, v);
        }
        __syncthreads();
        if (tid < 4 && s_cnt[tid]) atomicAdd(a.counters + tid, s_cnt[tid]);
    }
}

template <int LA, int LB, int LC, int LD, int CA, int CB, int CC, int CD>
int launch_jk_row_gc(int nblocks, const JKArgs& a, size_t shmem, hipStream_t st)
{
    hipLaunchKernelGGL((k_jk_row_gc<LA, LB, LC, LD, CA, CB, CC, CD>), dim3(nblocks), dim3(256), shmem, st, a);
    return (int)hipGetLastError();
}

}  // namespace mjk

