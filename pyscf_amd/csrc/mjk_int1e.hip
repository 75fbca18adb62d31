// One-electron integrals on the device: overlap, kinetic energy and nuclear
// attraction (mol.intor('int1e_ovlp' / 'int1e_kin' / 'int1e_nuc'),
// pyscf/gto/moleintor.py; used by the core Hamiltonian of
// pyscf/scf/hf.py:315-345), for SCF drivers that keep everything else on the
// device and for batches of many molecules.
//
// One lane per (shell pair, integral): the pair's primitive pairs in a loop,
// 1D overlap integrals by the Obara-Saika recursion
//   S(i+1, j) = PA S(i, j) + (i S(i-1, j) + j S(i, j-1)) / 2p
//   S(i, j+1) = PB S(i, j) + (i S(i-1, j) + j S(i, j-1)) / 2p
// with S(0, 0) = sqrt(pi/p) (times exp(-ab/p AB^2) in x); kinetic energy per
// direction -1/2 (4 b^2 S(i, j+2) - 2 b (2j+1) S(i, j) + j (j-1) S(i, j-2));
// nuclear attraction -Z_C <a|1/r_C|b> by Rys quadrature over T = p |PC|^2,
// g(i+1) = (PA - t^2 PC) g(i) + i (1 - t^2) / 2p g(i-1), then the transfer
// I(i, j+1) = I(i+1, j) + (A - B) I(i, j), prefactor 2 pi / p (point nuclei).
// The lane transforms both indices to spherical functions and writes the
// (i, j) and (j, i) blocks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "c2s.h"
#include "jk_device.h"
#include "mjk.h"
#include "mjk_internal.h"
#include "rys_eval.h"
#include "units.h"

using mjk_internal::ctx_fail;
using namespace mjk;

namespace {

struct Int1eArgs {
    const double4* ppP;     // per primitive pair {Px, Py, Pz, p}
    const double2* ppK;     // per primitive pair {exp(-ab/p AB^2) ca cb, b}
    const int* pprim0;
    const int* pnprim;
    const double4* pA;
    const double4* pAB;     // A - B
    const int* pi0;
    const int* pj0;
    int npair;
    const double4* nuc;     // {x, y, z, Z}
    int natm;
    double* out;
    int nao;
    const double* rys;
};

template <int LA, int LB, int WHICH>
__global__ void __launch_bounds__(128) k_int1e(Int1eArgs a)
{
    const int pair = blockIdx.x * blockDim.x + threadIdx.x;
    if (pair >= a.npair) return;
    constexpr int NCA = ncart(LA), NCB = ncart(LB);
    double acc[NCA][NCB];
#pragma unroll
    for (int i = 0; i < NCA; ++i)
#pragma unroll
        for (int j = 0; j < NCB; ++j) acc[i][j] = 0.0;
    const double4 A = a.pA[pair], AB = a.pAB[pair];
    const int q0 = a.pprim0[pair], nq = a.pnprim[pair];
    for (int qi = 0; qi < nq; ++qi) {
        const double4 P = a.ppP[q0 + qi];
        const double2 K = a.ppK[q0 + qi];
        const double p = P.w, ip2 = 0.5 / p;
        const double PA[3] = {P.x - A.x, P.y - A.y, P.z - A.z};
        const double PB[3] = {PA[0] + AB.x, PA[1] + AB.y, PA[2] + AB.z};   // P - B = P - A + (A - B)
        if constexpr (WHICH <= 1) {
            constexpr int NI = LA + 1, NJ = LB + 1 + (WHICH == 1 ? 2 : 0);
            double S[3][NI][NJ];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                S[d][0][0] = sqrt(M_PI / p) * (d == 0 ? K.x : 1.0);
#pragma unroll
                for (int i = 0; i < NI; ++i)
#pragma unroll
                    for (int j = 0; j < NJ; ++j) {
                        if (i == 0 && j == 0) continue;
                        double v;
                        if (j == 0) {
                            v = PA[d] * S[d][i - 1][0];
                            if (i >= 2) v += (i - 1) * ip2 * S[d][i - 2][0];
                        } else {
                            v = PB[d] * S[d][i][j - 1];
                            if (i >= 1) v += i * ip2 * S[d][i - 1][j - 1];
                            if (j >= 2) v += (j - 1) * ip2 * S[d][i][j - 2];
                        }
                        S[d][i][j] = v;
                    }
            }
#pragma unroll
            for (int ca = 0; ca < NCA; ++ca)
#pragma unroll
                for (int cb = 0; cb < NCB; ++cb) {
                    const int ex[3] = {cart_x(LA, ca), cart_y(LA, ca), cart_z(LA, ca)};
                    const int ey[3] = {cart_x(LB, cb), cart_y(LB, cb), cart_z(LB, cb)};
                    if constexpr (WHICH == 0) {
                        acc[ca][cb] += S[0][ex[0]][ey[0]] * S[1][ex[1]][ey[1]] * S[2][ex[2]][ey[2]];
                    } else {
                        const double b = K.y;
                        double v = 0.0;
#pragma unroll
                        for (int d = 0; d < 3; ++d) {
                            const int u = ex[d], t = ey[d];
                            double kd = -2.0 * b * (2 * t + 1) * S[d][u][t] + 4.0 * b * b * S[d][u][t + 2];
                            if (t >= 2) kd += t * (t - 1) * S[d][u][t - 2];
                            double o = -0.5 * kd;
#pragma unroll
                            for (int d2 = 0; d2 < 3; ++d2)
                                if (d2 != d) o *= S[d2][ex[d2]][ey[d2]];
                            v += o;
                        }
                        acc[ca][cb] += v;
                    }
                }
        } else {
            constexpr int NR = (LA + LB) / 2 + 1, NG = LA + LB + 1;
            for (int c = 0; c < a.natm; ++c) {
                const double4 C = a.nuc[c];
                const double PC[3] = {P.x - C.x, P.y - C.y, P.z - C.z};
                const double T = p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
                const double fac = -C.w * 2.0 * M_PI / p * K.x;
                RysCursor<NR> rc(a.rys, T);
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    double x, w;
                    rc.root(r, x, w);
                    const double B10 = ip2 * (1.0 - x);
                    double I[3][LA + 1][LB + 1];
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        double g[NG];
                        const double C00 = PA[d] - x * PC[d];
                        g[0] = d == 2 ? w * fac : 1.0;
                        if (NG > 1) g[1] = C00 * g[0];
#pragma unroll
                        for (int k = 1; k < NG - 1; ++k) g[k + 1] = C00 * g[k] + k * B10 * g[k - 1];
                        const double abd = d == 0 ? AB.x : d == 1 ? AB.y : AB.z;
                        double h[NG][LB + 1];
#pragma unroll
                        for (int k = 0; k < NG; ++k) h[k][0] = g[k];
#pragma unroll
                        for (int t = 1; t <= LB; ++t)
#pragma unroll
                            for (int k = 0; k < NG - t; ++k) h[k][t] = h[k + 1][t - 1] + abd * h[k][t - 1];
#pragma unroll
                        for (int u = 0; u <= LA; ++u)
#pragma unroll
                            for (int t = 0; t <= LB; ++t) I[d][u][t] = h[u][t];
                    }
#pragma unroll
                    for (int ca = 0; ca < NCA; ++ca)
#pragma unroll
                        for (int cb = 0; cb < NCB; ++cb)
                            acc[ca][cb] += I[0][cart_x(LA, ca)][cart_x(LB, cb)] * I[1][cart_y(LA, ca)][cart_y(LB, cb)] *
                                           I[2][cart_z(LA, ca)][cart_z(LB, cb)];
                }
            }
        }
    }
    const int i0 = a.pi0[pair], j0 = a.pj0[pair];
#pragma unroll
    for (int mi = 0; mi < 2 * LA + 1; ++mi)
#pragma unroll
        for (int mj = 0; mj < 2 * LB + 1; ++mj) {
            double v = 0.0;
#pragma unroll
            for (int ca = 0; ca < NCA; ++ca) {
                double s = 0.0;
#pragma unroll
                for (int cb = 0; cb < NCB; ++cb) s += C2S[LB][mj][cb] * acc[ca][cb];
                v += C2S[LA][mi][ca] * s;
            }
            a.out[(long long)(i0 + mi) * a.nao + j0 + mj] = v;
            a.out[(long long)(j0 + mj) * a.nao + i0 + mi] = v;
        }
}

template <int LA, int LB>
int launch_t(int which, const Int1eArgs& a, hipStream_t st)
{
    const dim3 grid((unsigned)((a.npair + 127) / 128));
    if (which == 0) hipLaunchKernelGGL((k_int1e<LA, LB, 0>), grid, dim3(128), 0, st, a);
    else if (which == 1) hipLaunchKernelGGL((k_int1e<LA, LB, 1>), grid, dim3(128), 0, st, a);
    else hipLaunchKernelGGL((k_int1e<LA, LB, 2>), grid, dim3(128), 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_int1e(int la, int lb, int which, const Int1eArgs& a, hipStream_t st)
{
    switch (la * 4 + lb) {
    case 0: return launch_t<0, 0>(which, a, st);
    case 4: return launch_t<1, 0>(which, a, st);
    case 5: return launch_t<1, 1>(which, a, st);
    case 8: return launch_t<2, 0>(which, a, st);
    case 9: return launch_t<2, 1>(which, a, st);
    case 10: return launch_t<2, 2>(which, a, st);
    case 12: return launch_t<3, 0>(which, a, st);
    case 13: return launch_t<3, 1>(which, a, st);
    case 14: return launch_t<3, 2>(which, a, st);
    case 15: return launch_t<3, 3>(which, a, st);
    default: return -1;
    }
}

template <typename T>
int up(mjk_ctx* ctx, const std::vector<T>& h, T** d, std::vector<void*>& owned)
{
    *d = nullptr;
    if (h.empty()) return MJK_OK;
    int rc = mjk_device_alloc(ctx, h.size() * sizeof(T), (void**)d);
    if (rc) return rc;
    owned.push_back(*d);
    if (hipMemcpyAsync(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, mjk_internal::ctx_stream(ctx)) !=
        hipSuccess)
        return ctx_fail(ctx, MJK_EDEVICE, "upload failed");
    return MJK_OK;
}

int run(mjk_ctx* ctx, const int* atm, int natm, const int* bas, int nbas, const double* env, double* const out[3],
        bool device_out, std::vector<void*>& owned)
{
    hipStream_t st = mjk_internal::ctx_stream(ctx);
    std::vector<Unit> u;
    std::vector<double> ex, co;
    int nao = 0, rc;
    if ((rc = fold_units(atm, natm, bas, nbas, env, 3, u, ex, co, nao)))
        return ctx_fail(ctx, rc, "mjk_int1e_device: shells must be l <= 3 with valid atoms");
    constexpr double EXPCUT = 60.0;   // primitive pairs below e^-60 contribute nothing at double precision
    const int ncls = 10;
    std::vector<std::vector<int>> cls(ncls);
    for (int i = 0; i < (int)u.size(); ++i)
        for (int j = 0; j <= i; ++j) {
            int a = i, b = j;
            if (u[b].l > u[a].l) std::swap(a, b);
            const int c = u[a].l * (u[a].l + 1) / 2 + u[b].l;
            cls[c].push_back(a);
            cls[c].push_back(b);
        }
    std::vector<double4> ppP, pA, pAB, nuc;
    std::vector<double2> ppK;
    std::vector<int> pp0, pnp, pi0, pj0, cstart(ncls + 1, 0);
    for (int c = 0; c < ncls; ++c) {
        cstart[c] = (int)pA.size();
        for (size_t k = 0; k < cls[c].size(); k += 2) {
            const Unit& A = u[cls[c][k]];
            const Unit& B = u[cls[c][k + 1]];
            const double ab[3] = {A.r[0] - B.r[0], A.r[1] - B.r[1], A.r[2] - B.r[2]};
            const double r2 = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
            const int first = (int)ppP.size();
            for (int x = 0; x < A.nprim; ++x)
                for (int y = 0; y < B.nprim; ++y) {
                    const double ea = ex[A.p0 + x], eb = ex[B.p0 + y], p = ea + eb;
                    const double t = ea * eb / p * r2;
                    if (t > EXPCUT) continue;
                    ppP.push_back(make_double4((ea * A.r[0] + eb * B.r[0]) / p, (ea * A.r[1] + eb * B.r[1]) / p,
                                               (ea * A.r[2] + eb * B.r[2]) / p, p));
                    ppK.push_back(make_double2(std::exp(-t) * co[A.p0 + x] * co[B.p0 + y], eb));
                }
            pp0.push_back(first);
            pnp.push_back((int)ppP.size() - first);
            pA.push_back(make_double4(A.r[0], A.r[1], A.r[2], 0.0));
            pAB.push_back(make_double4(ab[0], ab[1], ab[2], 0.0));
            pi0.push_back(A.sph0);
            pj0.push_back(B.sph0);
        }
    }
    cstart[ncls] = (int)pA.size();
    for (int c = 0; c < natm; ++c) {
        const double* r = env + atm[6 * c + 1];
        nuc.push_back(make_double4(r[0], r[1], r[2], (double)atm[6 * c]));
    }
    double4 *d_ppP, *d_pA, *d_pAB, *d_nuc;
    double2* d_ppK;
    int *d_pp0, *d_pnp, *d_pi0, *d_pj0;
    if ((rc = up(ctx, ppP, &d_ppP, owned)) || (rc = up(ctx, pA, &d_pA, owned)) || (rc = up(ctx, pAB, &d_pAB, owned)) ||
        (rc = up(ctx, nuc, &d_nuc, owned)) || (rc = up(ctx, ppK, &d_ppK, owned)) ||
        (rc = up(ctx, pp0, &d_pp0, owned)) || (rc = up(ctx, pnp, &d_pnp, owned)) ||
        (rc = up(ctx, pi0, &d_pi0, owned)) || (rc = up(ctx, pj0, &d_pj0, owned)))
        return rc;
    const size_t nn = (size_t)nao * nao;
    double* dout[3] = {nullptr, nullptr, nullptr};
    for (int w = 0; w < 3; ++w) {
        if (!out[w]) continue;
        if (device_out) {
            dout[w] = out[w];
        } else {
            if ((rc = mjk_device_alloc(ctx, sizeof(double) * nn, (void**)&dout[w]))) return rc;
            owned.push_back(dout[w]);
        }
        for (int la = 0; la <= 3; ++la)
            for (int lb = 0; lb <= la; ++lb) {
                const int c = la * (la + 1) / 2 + lb;
                const int np = cstart[c + 1] - cstart[c];
                if (!np) continue;
                Int1eArgs a{};
                a.ppP = d_ppP;
                a.ppK = d_ppK;
                a.pprim0 = d_pp0 + cstart[c];
                a.pnprim = d_pnp + cstart[c];
                a.pA = d_pA + cstart[c];
                a.pAB = d_pAB + cstart[c];
                a.pi0 = d_pi0 + cstart[c];
                a.pj0 = d_pj0 + cstart[c];
                a.npair = np;
                a.nuc = d_nuc;
                a.natm = natm;
                a.out = dout[w];
                a.nao = nao;
                a.rys = mjk_internal::ctx_rys(ctx);
                if (launch_int1e(la, lb, w, a, st)) return ctx_fail(ctx, MJK_EDEVICE, "int1e launch failed");
            }
    }
    if (!device_out)
        for (int w = 0; w < 3; ++w)
            if (out[w] && hipMemcpyAsync(out[w], dout[w], sizeof(double) * nn, hipMemcpyDeviceToHost, st) != hipSuccess)
                return ctx_fail(ctx, MJK_EDEVICE, "download failed");
    if (hipStreamSynchronize(st) != hipSuccess) return ctx_fail(ctx, MJK_EDEVICE, "int1e synchronisation failed");
    return MJK_OK;
}

}  // namespace

extern "C" {

int mjk_int1e_device(mjk_ctx* ctx, const int* atm, int natm, const int* bas, int nbas, const double* env, double* s,
                     double* t, double* v, int device_out)
{
    if (!ctx || !atm || !bas || !env || natm <= 0 || nbas <= 0 || (!s && !t && !v))
        return ctx_fail(ctx, MJK_EINVAL, "bad arguments to mjk_int1e_device");
    std::lock_guard<std::recursive_mutex> lock(mjk_internal::ctx_mutex(ctx));
    if (hipSetDevice(mjk_internal::ctx_device(ctx)) != hipSuccess)
        return ctx_fail(ctx, MJK_EDEVICE, "hipSetDevice failed");
    std::vector<void*> owned;
    double* const out[3] = {s, t, v};
    const int rc = run(ctx, atm, natm, bas, nbas, env, out, device_out != 0, owned);
    (void)hipStreamSynchronize(mjk_internal::ctx_stream(ctx));
    for (void* p : owned) (void)mjk_device_free(ctx, p);
    return rc;
}

}  // extern "C"
