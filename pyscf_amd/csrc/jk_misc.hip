// Schwarz-bound, density-screening and basis-transform kernels.
#include <algorithm>
#include "jk_device.h"

// Schwarz bound of each shell pair: q = sqrt(max_{a,b} |(ab|ab)|) over the
// spherical functions (pyscf/lib/vhf/optimizer.c:408-454), floor 1e-100.

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
            for (int pass = 0; pass < rsh_passes(a.omega); ++pass) {   // range-separated operators
            double Tp = T, sx, sw;
            rsh_scale(a.omega, Pp.w * Qp.w * ipq, pass, Tp, sx, sw);
            const typename RysFor<NR, 2 * (LA + LB)>::type rys(a.rys, Tp);
#pragma unroll 1
            for (int r = 0; r < NR; ++r) {
                double rx, rw;
                rys.root(r, rx, rw);
                rx *= sx;
                rw *= sw;
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
                                                  long long sslot, int ncopy, long long jcopy,
                                                  long long kcopy)
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
                for (int cp = 0; cp < ncopy; ++cp) {
                    if (with_j) tj = fma(w, Jc[cp * jcopy + s * cslot + e1] + Jc[cp * jcopy + s * cslot + e2], tj);
                    if (with_k) {
                        const double* Kp = Kc + cp * kcopy;
                        if (hermi == 1)
                            tk = fma(w, Kp[s * cslot + e1] + Kp[s * cslot + e2], tk);
                        else
                            tk = fma(w, Kp[(2 * s) * cslot + e1] + Kp[(2 * s + 1) * cslot + e2], tk);
                    }
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


// Plain device-to-device copy (also the reference kernel of known traffic
// for calibrating the HBM byte counters).
__global__ __launch_bounds__(256) void k_copy(const double* __restrict__ src, double* __restrict__ dst, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Zero fill (a kernel rather than a memset node, so that a captured graph
// replays it like any other kernel).
__global__ __launch_bounds__(256) void k_zero(double* __restrict__ dst, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dst[i] = 0.0;
}

// ---------------------------------------------------------------------------
// Host-side launchers.
// Row tasks of active molecules: list[count++] = t for every task t whose bra
// pair belongs to an active molecule (one wave-aggregated atomic per wave;
// the order within a wave is kept).  count must be zero before the launch.
__global__ __launch_bounds__(256) void k_compact_rows(const int4* __restrict__ rt, int nrt,
                                                       const int* __restrict__ pmol, const int* __restrict__ active,
                                                       int* __restrict__ list, int* __restrict__ count)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool on = t < nrt && active[pmol[rt[t].x]] != 0;
    const unsigned long long mask = __ballot(on);
    if (!mask) return;
    const int lane = threadIdx.x & 63;
    int base = 0;
    const int leader = __ffsll((long long)mask) - 1;
    if (lane == leader) base = atomicAdd(count, __popcll(mask));
    base = __shfl(base, leader);
    if (on) list[base + __popcll(mask & ((1ull << lane) - 1))] = t;
}

int launch_compact_rows(const int4* rt, int nrt, const int* pmol, const int* active, int* list, int* count,
                        hipStream_t st)
{
    if (nrt <= 0) return 0;
    hipLaunchKernelGGL(k_compact_rows, dim3((unsigned)((nrt + 255) / 256)), dim3(256), 0, st, rt, nrt, pmol, active,
                       list, count);
    return (int)hipGetLastError();
}

int launch_zero(double* dst, long long n, hipStream_t st)
{
    if (n <= 0) return 0;
    const long long nb = std::min<long long>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_zero, dim3((unsigned)nb), dim3(256), 0, st, dst, n);
    return (int)hipGetLastError();
}

int launch_copy(const double* src, double* dst, long long n, hipStream_t st)
{
    if (n <= 0) return 0;
    const long long nb = std::min<long long>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_copy, dim3((unsigned)nb), dim3(256), 0, st, src, dst, n);
    return (int)hipGetLastError();
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
        Q(0, 0) Q(1, 0) Q(1, 1) Q(2, 0) Q(2, 1) Q(2, 2) Q(3, 0) Q(3, 1) Q(3, 2) Q(3, 3)
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
                    int ncopy, long long jcopy, long long kcopy, hipStream_t st)
{
    dim3 blk(64, 4);
    dim3 grd((unsigned)((x.nsph_tot + 63) / 64), (unsigned)((maxnao + 3) / 4));
    hipLaunchKernelGGL(k_cart2sph, grd, blk, 0, st, x, Jc, Kc, cslot, n_dm, hermi, with_j, with_k, vj, vk,
                       sslot, ncopy, jcopy, kcopy);
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
