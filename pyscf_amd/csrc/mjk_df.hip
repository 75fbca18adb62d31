// Density-fitted J/K (pyscf/df/incore.py:129-214 cholesky_eri, pyscf/df/df_jk.py:278-418
// get_jk) on one device.
//
// Build (once per molecule and auxiliary basis):
//   * (P|Q) and (ij|P) by the Rys kernels of df_device.h into staging
//     matrices with Cartesian auxiliary columns, then the auxiliary index to
//     spherical functions (k_df_auxsph);
//   * (P|Q) = L L^T (rocsolver_dpotrf), B = L^-1 (P|ij) (rocblas_dtrsm); when
//     the factorisation fails, B = (V w^-1/2)^T (P|ij) over the eigenpairs of
//     (P|Q) with w above lindep, as the reference falls back.
//   B stays in HBM as a column-major (naux x nao^2) matrix: for every orbital
//   pair (i, j), both triangles, the naux fitted values are contiguous.
// Builds (per call):
//   * J: rho = B vec(D) and vec(J) = B^T rho, two bandwidth-bound passes over
//     the lower-triangle rows of B (k_df_rho, k_df_j);
//   * K of a density: Y_i = B_i D^T for every i (one strided-batched GEMM,
//     B_i = the naux x nao block of row i), then K = B^T Y as one GEMM over
//     (P, l): K_ik = sum_Pl B_k[P,l] Y_i[P,l] = sum_P (B_P D B_P)_ik;
//   * K of occupied orbitals X (D = X X^T): Y_i = B_i X, K = Y^T Y.
// All on rocBLAS/rocSOLVER on the context's stream; FP64 throughout.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "df_device.h"
#include "mjk.h"
#include "mjk_internal.h"
#include "units.h"

using mjk_internal::ctx_fail;
using namespace mjk;

struct mjk_df {
    mjk_ctx* ctx = nullptr;
    int nao = 0, naux_raw = 0, naux = 0;
    int fitted = 1;            // 0: B holds the raw (ij|P) (lindep < 0)
    double* B = nullptr;       // naux x nao^2, column-major
    double* j2c = nullptr;     // raw (P|Q), naux_raw^2
    // call buffers (grown on demand)
    double* Y = nullptr;
    size_t capY = 0;
    double* D = nullptr;       // densities / orbitals, outputs
    size_t capD = 0;
    double* rho = nullptr;
    size_t caprho = 0;
    void* ws = nullptr;
    size_t capws = 0;
    // blocks of the exchange product from the occupied orbitals (0: by size;
    // MJK_DF_KBLOCKS overrides)
    int kblocks = 0;
};

namespace {

constexpr double EXPCUT = 40.0;   // primitive pairs with exp(-ab/p |AB|^2) < e^-40 are dropped

// rows of the Cartesian -> spherical transform for l <= 4, [l][m][c] (9 x 15
// per l); l <= 1 are identities (their factors are in the coefficients)
void c2s_rows(std::vector<double>& tab)
{
    tab.assign(5 * 9 * 15, 0.0);
    for (int l = 0; l <= 4; ++l) {
        const int nc = (l + 1) * (l + 2) / 2;
        std::vector<double> t((2 * l + 1) * nc);
        if (l <= 1) {
            for (int m = 0; m < 2 * l + 1; ++m) t[m * nc + m] = 1.0;
        } else {
            mjk_c2s(l, t.data());
        }
        for (int m = 0; m < 2 * l + 1; ++m)
            for (int c = 0; c < nc; ++c) tab[(l * 9 + m) * 15 + c] = t[m * nc + c];
    }
}

// B[row][P] = sum_c tab[l][m][c] S[row][col0[l] + c n[l] + u], per final auxiliary function
// P = {l, m, u}; rows x naux outputs
__global__ void k_df_auxsph(const double* __restrict__ S, long long lds, double* __restrict__ B, long long rows,
                            int naux, const int4* __restrict__ map, const long long* __restrict__ col0,
                            const int* __restrict__ ngrp, const double* __restrict__ tab)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= rows * naux) return;
    const long long row = t / naux;
    const int P = (int)(t % naux);
    const int4 m = map[P];   // {l, m, unit in group, 0}
    const int l = m.x, nc = (l + 1) * (l + 2) / 2;
    const double* s = S + row * lds + col0[l] + m.z;
    const double* tr = tab + (l * 9 + m.y) * 15;
    double v = 0.0;
    for (int c = 0; c < nc; ++c) v += tr[c] * s[(long long)c * ngrp[l]];
    B[row * naux + P] = v;
}

// J of one density from the fitted tensor, two passes over the lower-triangle
// rows (i >= j) of B, which is symmetric in (i, j): half the tensor's bytes
// per pass.  These are bandwidth-bound matrix-vector products; as GEMMs with
// one column they ran at a few percent of HBM bandwidth.
//   rho[P] = sum_{i>=j} w_ij B[ij][P],  w_ij = D_ij + D_ji (i > j), D_ii (i = j)
// Grid: x = chunks of DF_ROWS lower-triangle rows, y = column tiles of
// DF_COLS columns; a lane keeps DF_COLS / 256 column sums over the chunk's
// rows, then adds them into rho (FP64 atomics; rho is zeroed first).
constexpr int DF_ROWS = 256;
constexpr int DF_CPT = 8;                 // columns per lane
constexpr int DF_COLS = 256 * DF_CPT;     // columns per block

__device__ __forceinline__ void tri_index(long long t, int& i, int& j)
{
    int r = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((long long)(r + 1) * (r + 2) / 2 <= t) ++r;
    while ((long long)r * (r + 1) / 2 > t) --r;
    i = r;
    j = (int)(t - (long long)r * (r + 1) / 2);
}

__global__ void __launch_bounds__(256) k_df_rho(const double* __restrict__ B, int nao, int naux,
                                                const double* __restrict__ D, double* __restrict__ rho)
{
    const long long ntri = (long long)nao * (nao + 1) / 2;
    const long long t0 = (long long)blockIdx.x * DF_ROWS;
    const long long t1 = t0 + DF_ROWS < ntri ? t0 + DF_ROWS : ntri;
    const int c0 = blockIdx.y * DF_COLS + threadIdx.x;
    double acc[DF_CPT];
#pragma unroll
    for (int k = 0; k < DF_CPT; ++k) acc[k] = 0.0;
    int i, j;
    tri_index(t0, i, j);
    for (long long t = t0; t < t1; ++t) {
        const double w = i == j ? D[(long long)i * nao + i] : D[(long long)i * nao + j] + D[(long long)j * nao + i];
        const double* row = B + ((long long)i * nao + j) * naux;
#pragma unroll
        for (int k = 0; k < DF_CPT; ++k) {
            const int c = c0 + k * 256;
            if (c < naux) acc[k] += w * row[c];
        }
        if (++j > i) {
            ++i;
            j = 0;
        }
    }
#pragma unroll
    for (int k = 0; k < DF_CPT; ++k) {
        const int c = c0 + k * 256;
        if (c < naux) atomicAdd(rho + c, acc[k]);
    }
}

// J_ij = J_ji = sum_P B[ij][P] rho[P] for the lower-triangle rows; one wave per
// row, rho staged in LDS (naux doubles)
__global__ void __launch_bounds__(256) k_df_j(const double* __restrict__ B, int nao, int naux,
                                              const double* __restrict__ rho, double* __restrict__ J)
{
    extern __shared__ double srho[];
    for (int c = threadIdx.x; c < naux; c += blockDim.x) srho[c] = rho[c];
    __syncthreads();
    const long long ntri = (long long)nao * (nao + 1) / 2;
    const int lane = threadIdx.x & 63;
    const long long nw = (long long)gridDim.x * (blockDim.x >> 6);
    for (long long t = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); t < ntri; t += nw) {
        int i, j;
        tri_index(t, i, j);
        const double* row = B + ((long long)i * nao + j) * naux;
        double v = 0.0;
        for (int c = lane; c < naux; c += 64) v += row[c] * srho[c];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) {
            J[(long long)i * nao + j] = v;
            J[(long long)j * nao + i] = v;
        }
    }
}

template <typename T>
int up(mjk_ctx* ctx, const std::vector<T>& h, T** d)
{
    *d = nullptr;
    if (h.empty()) return MJK_OK;
    int rc = mjk_device_alloc(ctx, h.size() * sizeof(T), (void**)d);
    if (rc) return rc;
    if (hipMemcpyAsync(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, mjk_internal::ctx_stream(ctx)) !=
        hipSuccess)
        return ctx_fail(ctx, MJK_EDEVICE, "upload failed");
    return MJK_OK;
}

struct DevFree {
    mjk_ctx* ctx;
    std::vector<void*> ptrs;
    ~DevFree()
    {
        for (void* p : ptrs)
            if (p) (void)mjk_device_free(ctx, p);
    }
};

#define RB(ctx, x)                                                                                    \
    do {                                                                                              \
        rocblas_status s_ = (x);                                                                      \
        if (s_ != rocblas_status_success)                                                             \
            return ctx_fail(ctx, MJK_EDEVICE, "%s failed: %s", #x, rocblas_status_to_string(s_));     \
    } while (0)

#define HC(ctx, x)                                                                                    \
    do {                                                                                              \
        hipError_t e_ = (x);                                                                          \
        if (e_ != hipSuccess) return ctx_fail(ctx, MJK_EDEVICE, "%s failed: %s", #x, hipGetErrorString(e_)); \
    } while (0)

int grow(mjk_ctx* ctx, double** p, size_t* cap, size_t elems)
{
    if (*cap >= elems) return MJK_OK;
    if (*p) (void)mjk_device_free(ctx, *p);
    *p = nullptr;
    *cap = 0;
    int rc = mjk_device_alloc(ctx, elems * sizeof(double), (void**)p);
    if (rc == MJK_OK) *cap = elems;
    return rc;
}

// rocBLAS workspace of at least `bytes` for the handle
int workspace(mjk_df* df, rocblas_handle h, size_t bytes)
{
    bytes = std::max(bytes, (size_t)32 << 20);
    if (df->capws < bytes) {
        if (df->ws) (void)mjk_device_free(df->ctx, df->ws);
        df->ws = nullptr;
        df->capws = 0;
        int rc = mjk_device_alloc(df->ctx, bytes, &df->ws);
        if (rc) return rc;
        df->capws = bytes;
    }
    RB(df->ctx, rocblas_set_workspace(h, df->ws, df->capws));
    return MJK_OK;
}

int build(mjk_ctx* ctx, mjk_df* df, const int* atm, int natm, const int* bas, int nbas, const double* env,
          const int* aatm, int anatm, const int* abas, int anbas, const double* aenv, double lindep)
{
    hipStream_t st = mjk_internal::ctx_stream(ctx);
    std::vector<Unit> ao, ax;
    std::vector<double> aoe, aoc, axe, axc;
    int nao = 0, naux = 0, rc;
    if ((rc = fold_units(atm, natm, bas, nbas, env, DF_LMAX_AO, ao, aoe, aoc, nao)))
        return ctx_fail(ctx, rc, "mjk_df_build: orbital shells must be l <= %d with valid atoms", DF_LMAX_AO);
    if ((rc = fold_units(aatm, anatm, abas, anbas, aenv, DF_LMAX_AUX, ax, axe, axc, naux)))
        return ctx_fail(ctx, rc, "mjk_df_build: auxiliary shells must be l <= %d with valid atoms", DF_LMAX_AUX);
    df->nao = nao;
    df->naux_raw = naux;
    // orbital pairs by class (li >= lj)
    const int ncls = (DF_LMAX_AO + 1) * (DF_LMAX_AO + 2) / 2;
    std::vector<std::vector<int>> cls(ncls);   // flat records: i unit, j unit
    for (int u = 0; u < (int)ao.size(); ++u)
        for (int v = 0; v <= u; ++v) {
            int i = u, j = v;
            if (ao[j].l > ao[i].l) std::swap(i, j);
            cls[ao[i].l * (ao[i].l + 1) / 2 + ao[j].l].push_back(i);
            cls[ao[i].l * (ao[i].l + 1) / 2 + ao[j].l].push_back(j);
        }
    std::vector<double4> ppP, pA, pAB;
    std::vector<double> ppK;
    std::vector<int> pp0, pnp, pi0, pj0;
    std::vector<int> cstart(ncls + 1, 0);
    for (int c = 0; c < ncls; ++c) {
        cstart[c] = (int)pA.size();
        for (size_t k = 0; k < cls[c].size(); k += 2) {
            const Unit& I = ao[cls[c][k]];
            const Unit& Jn = ao[cls[c][k + 1]];
            const double ab[3] = {I.r[0] - Jn.r[0], I.r[1] - Jn.r[1], I.r[2] - Jn.r[2]};
            const double r2 = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
            const int first = (int)ppP.size();
            for (int a = 0; a < I.nprim; ++a)
                for (int b = 0; b < Jn.nprim; ++b) {
                    const double ea = aoe[I.p0 + a], eb = aoe[Jn.p0 + b], p = ea + eb;
                    const double x = ea * eb / p * r2;
                    if (x > EXPCUT) continue;
                    ppP.push_back(make_double4((ea * I.r[0] + eb * Jn.r[0]) / p, (ea * I.r[1] + eb * Jn.r[1]) / p,
                                               (ea * I.r[2] + eb * Jn.r[2]) / p, p));
                    ppK.push_back(std::exp(-x) * aoc[I.p0 + a] * aoc[Jn.p0 + b]);
                }
            if ((int)ppP.size() == first) continue;   // no primitive pair survives: the pair is zero
            pp0.push_back(first);
            pnp.push_back((int)ppP.size() - first);
            pA.push_back(make_double4(I.r[0], I.r[1], I.r[2], 0.0));
            pAB.push_back(make_double4(ab[0], ab[1], ab[2], 0.0));
            pi0.push_back(I.sph0);
            pj0.push_back(Jn.sph0);
        }
    }
    cstart[ncls] = (int)pA.size();
    // auxiliary shells grouped by l; (P|Q) "pairs" with the unit function
    std::vector<int> gstart(DF_LMAX_AUX + 2, 0), gn(DF_LMAX_AUX + 1, 0);
    std::vector<std::vector<int>> grp(DF_LMAX_AUX + 1);
    for (int u = 0; u < (int)ax.size(); ++u) grp[ax[u].l].push_back(u);
    std::vector<double4> kC, qP, qA, qAB;
    std::vector<int> kp0, knp, q0, qnp, qi0, qj0;
    std::vector<double2> kprim;
    std::vector<double> qK;
    std::vector<long long> col0(DF_LMAX_AUX + 1, 0);
    std::vector<int4> map(naux);
    long long ncol = 0;
    for (int l = 0; l <= DF_LMAX_AUX; ++l) {
        gstart[l] = (int)kC.size();
        gn[l] = (int)grp[l].size();
        col0[l] = ncol;
        ncol += (long long)(l + 1) * (l + 2) / 2 * gn[l];
        for (int k = 0; k < gn[l]; ++k) {
            const Unit& U = ax[grp[l][k]];
            kC.push_back(make_double4(U.r[0], U.r[1], U.r[2], 0.0));
            kp0.push_back((int)kprim.size());
            knp.push_back(U.nprim);
            q0.push_back((int)qP.size());
            qnp.push_back(U.nprim);
            for (int p = 0; p < U.nprim; ++p) {
                kprim.push_back(make_double2(axe[U.p0 + p], axc[U.p0 + p]));
                qP.push_back(make_double4(U.r[0], U.r[1], U.r[2], axe[U.p0 + p]));
                qK.push_back(axc[U.p0 + p]);
            }
            qA.push_back(make_double4(U.r[0], U.r[1], U.r[2], 0.0));
            qAB.push_back(make_double4(0.0, 0.0, 0.0, 0.0));
            qi0.push_back(U.sph0);
            qj0.push_back(-1);
            for (int m = 0; m < 2 * l + 1; ++m) map[U.sph0 + m] = make_int4(l, m, k, 0);
        }
    }
    gstart[DF_LMAX_AUX + 1] = (int)kC.size();
    std::vector<double> tab;
    c2s_rows(tab);
    DevFree tmp{ctx, {}};
    double4 *d_ppP, *d_pA, *d_pAB, *d_kC, *d_qP, *d_qA, *d_qAB;
    double *d_ppK, *d_qK, *d_tab;
    double2* d_kprim;
    int *d_pp0, *d_pnp, *d_pi0, *d_pj0, *d_kp0, *d_knp, *d_q0, *d_qnp, *d_qi0, *d_qj0, *d_gn;
    int4* d_map;
    long long* d_col0;
#define UP(v, d)                                   \
    if ((rc = up(ctx, v, &d))) return rc;          \
    tmp.ptrs.push_back((void*)d);
    UP(ppP, d_ppP) UP(pA, d_pA) UP(pAB, d_pAB) UP(kC, d_kC) UP(qP, d_qP) UP(qA, d_qA) UP(qAB, d_qAB)
    UP(ppK, d_ppK) UP(qK, d_qK) UP(tab, d_tab) UP(kprim, d_kprim) UP(pp0, d_pp0) UP(pnp, d_pnp)
    UP(pi0, d_pi0) UP(pj0, d_pj0) UP(kp0, d_kp0) UP(knp, d_knp) UP(q0, d_q0) UP(qnp, d_qnp) UP(qi0, d_qi0)
    UP(qj0, d_qj0) UP(gn, d_gn) UP(map, d_map) UP(col0, d_col0)
#undef UP
    const double* t4 = d_tab + 4 * 9 * 15;
    const long long nn = (long long)nao * nao;
    // (P|Q)
    double *S2 = nullptr, *S3 = nullptr;
    if ((rc = mjk_device_alloc(ctx, sizeof(double) * (size_t)naux * ncol, (void**)&S2))) return rc;
    tmp.ptrs.push_back(S2);
    HC(ctx, hipMemsetAsync(S2, 0, sizeof(double) * (size_t)naux * ncol, st));
    if ((rc = mjk_device_alloc(ctx, sizeof(double) * (size_t)naux * naux, (void**)&df->j2c))) return rc;
    for (int li = 0; li <= DF_LMAX_AUX; ++li) {
        if (!gn[li]) continue;
        for (int lk = 0; lk <= DF_LMAX_AUX; ++lk) {
            if (!gn[lk]) continue;
            Int3cArgs a{};
            a.ppP = d_qP;
            a.ppK = d_qK;
            a.pprim0 = d_q0 + gstart[li];
            a.pnprim = d_qnp + gstart[li];
            a.pA = d_qA + gstart[li];
            a.pAB = d_qAB + gstart[li];
            a.pi0 = d_qi0 + gstart[li];
            a.pj0 = d_qj0 + gstart[li];
            a.npair = gn[li];
            a.kC = d_kC + gstart[lk];
            a.kprim0 = d_kp0 + gstart[lk];
            a.knprim = d_knp + gstart[lk];
            a.kprim = d_kprim;
            a.naux = gn[lk];
            a.col0 = col0[lk];
            a.out = S2;
            a.ld_row = 0;
            a.ld = ncol;
            a.rys = mjk_internal::ctx_rys(ctx);
            a.c2s4 = t4;
            if (launch_int3c(li, 0, lk, true, a, st)) return ctx_fail(ctx, MJK_EDEVICE, "(P|Q) launch failed");
        }
    }
    {
        const long long n = (long long)naux * naux;
        k_df_auxsph<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(S2, ncol, df->j2c, naux, naux, d_map, d_col0, d_gn,
                                                                 d_tab);
        HC(ctx, hipGetLastError());
    }
    // (ij|P)
    if ((rc = mjk_device_alloc(ctx, sizeof(double) * (size_t)nn * ncol, (void**)&S3))) return rc;
    tmp.ptrs.push_back(S3);
    HC(ctx, hipMemsetAsync(S3, 0, sizeof(double) * (size_t)nn * ncol, st));
    for (int li = 0; li <= DF_LMAX_AO; ++li)
        for (int lj = 0; lj <= li; ++lj) {
            const int c = li * (li + 1) / 2 + lj;
            const int np = cstart[c + 1] - cstart[c];
            if (!np) continue;
            for (int lk = 0; lk <= DF_LMAX_AUX; ++lk) {
                if (!gn[lk]) continue;
                Int3cArgs a{};
                a.ppP = d_ppP;
                a.ppK = d_ppK;
                a.pprim0 = d_pp0 + cstart[c];
                a.pnprim = d_pnp + cstart[c];
                a.pA = d_pA + cstart[c];
                a.pAB = d_pAB + cstart[c];
                a.pi0 = d_pi0 + cstart[c];
                a.pj0 = d_pj0 + cstart[c];
                a.npair = np;
                a.kC = d_kC + gstart[lk];
                a.kprim0 = d_kp0 + gstart[lk];
                a.knprim = d_knp + gstart[lk];
                a.kprim = d_kprim;
                a.naux = gn[lk];
                a.col0 = col0[lk];
                a.out = S3;
                a.ld_row = nao;
                a.ld = ncol;
                a.rys = mjk_internal::ctx_rys(ctx);
                a.c2s4 = t4;
                if (launch_int3c(li, lj, lk, false, a, st))
                    return ctx_fail(ctx, MJK_EDEVICE, "(ij|P) launch failed");
            }
        }
    if ((rc = mjk_device_alloc(ctx, sizeof(double) * (size_t)nn * naux, (void**)&df->B))) return rc;
    {
        const long long n = nn * naux;
        k_df_auxsph<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(S3, ncol, df->B, nn, naux, d_map, d_col0, d_gn,
                                                                 d_tab);
        HC(ctx, hipGetLastError());
    }
    HC(ctx, hipStreamSynchronize(st));
    (void)mjk_device_free(ctx, S3);
    tmp.ptrs.back() = nullptr;
    df->naux = naux;
    if (lindep < 0) {   // raw (ij|P), no fit
        df->fitted = 0;
        return MJK_OK;
    }
    // B <- L^-1 B with (P|Q) = L L^T, or the eigenvector fit
    rocblas_handle h = (rocblas_handle)mjk_internal::ctx_rocblas(ctx);
    if (!h) return ctx_fail(ctx, MJK_EDEVICE, "rocblas_create_handle failed");
    RB(ctx, rocblas_set_stream(h, st));
    double* L = nullptr;
    if ((rc = mjk_device_alloc(ctx, sizeof(double) * (size_t)naux * naux, (void**)&L))) return rc;
    tmp.ptrs.push_back(L);
    rocblas_int* info = nullptr;
    if ((rc = mjk_device_alloc(ctx, sizeof(rocblas_int) * 4, (void**)&info))) return rc;
    tmp.ptrs.push_back(info);
    HC(ctx, hipMemcpyAsync(L, df->j2c, sizeof(double) * (size_t)naux * naux, hipMemcpyDeviceToDevice, st));
    size_t need = 0;
    double* wv = nullptr;
    if ((rc = mjk_device_alloc(ctx, sizeof(double) * 2 * (size_t)naux, (void**)&wv))) return rc;
    tmp.ptrs.push_back(wv);
    RB(ctx, rocblas_start_device_memory_size_query(h));
    {
        const double one = 1.0;
        (void)rocsolver_dpotrf(h, rocblas_fill_lower, naux, L, naux, info);
        (void)rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, naux, L, naux, wv, wv + naux, info);
        (void)rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, rocblas_diagonal_non_unit,
                            naux, (rocblas_int)nn, &one, L, naux, df->B, naux);
    }
    RB(ctx, rocblas_stop_device_memory_size_query(h, &need));
    if ((rc = workspace(df, h, need))) return rc;
    RB(ctx, rocblas_set_pointer_mode(h, rocblas_pointer_mode_host));
    RB(ctx, rocsolver_dpotrf(h, rocblas_fill_lower, naux, L, naux, info));
    rocblas_int hinfo = 0;
    HC(ctx, hipMemcpyAsync(&hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost, st));
    HC(ctx, hipStreamSynchronize(st));
    if (hinfo == 0) {
        const double one = 1.0;
        RB(ctx, rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none,
                              rocblas_diagonal_non_unit, naux, (rocblas_int)nn, &one, L, naux, df->B, naux));
    } else {
        // eigenvector fit over w > lindep
        HC(ctx, hipMemcpyAsync(L, df->j2c, sizeof(double) * (size_t)naux * naux, hipMemcpyDeviceToDevice, st));
        RB(ctx, rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, naux, L, naux, wv, wv + naux, info));
        std::vector<double> w(naux);
        HC(ctx, hipMemcpyAsync(w.data(), wv, sizeof(double) * naux, hipMemcpyDeviceToHost, st));
        HC(ctx, hipMemcpyAsync(&hinfo, info, sizeof(hinfo), hipMemcpyDeviceToHost, st));
        HC(ctx, hipStreamSynchronize(st));
        if (hinfo != 0) return ctx_fail(ctx, MJK_EDEVICE, "mjk_df_build: eigensolver info %d", (int)hinfo);
        int first = 0;
        while (first < naux && !(w[first] > lindep)) ++first;   // eigenvalues ascend
        const int keep = naux - first;
        if (keep == 0) return ctx_fail(ctx, MJK_EINVAL, "mjk_df_build: no auxiliary eigenvalue above lindep");
        std::vector<double> sc(keep);
        for (int k = 0; k < keep; ++k) sc[k] = 1.0 / std::sqrt(w[first + k]);
        double* dsc = nullptr;
        if ((rc = up(ctx, sc, &dsc))) return rc;
        tmp.ptrs.push_back(dsc);
        // W = V[:, first:] diag(w^-1/2) (in place on L's columns)
        RB(ctx, rocblas_ddgmm(h, rocblas_side_right, naux, keep, L + (size_t)first * naux, naux, dsc, 1,
                              L + (size_t)first * naux, naux));
        double* B2 = nullptr;
        if ((rc = mjk_device_alloc(ctx, sizeof(double) * (size_t)nn * keep, (void**)&B2))) return rc;
        const double one = 1.0, zero = 0.0;
        RB(ctx, rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, keep, (rocblas_int)nn, naux, &one,
                              L + (size_t)first * naux, naux, df->B, naux, &zero, B2, keep));
        HC(ctx, hipStreamSynchronize(st));
        (void)mjk_device_free(ctx, df->B);
        df->B = B2;
        df->naux = keep;
    }
    HC(ctx, hipStreamSynchronize(st));
    return MJK_OK;
}

// rocBLAS's DGEMM for the long-inner-dimension shapes of the exchange
// product (m, n = nao, inner dimension naux nocc or naux nao) runs at about
// 50 TFLOP/s for even m and n and at a fraction of that for odd ones
// (measured on MI355X, inner dimension 426646 and a 512 MiB workspace: m = n
// = 730 in 9.1 ms, 729 and 731 in 68.7 ms; every odd m in 183..731 tried took
// 68.4-68.8 ms, every even one 0.9-9.1 ms).  So the GEMMs below see even
// dimensions only; an odd last AO is a matrix-vector product.

// Block of an index in the lower-block layout of df_syrk: blocks of bs over
// [0, ne), the odd last index (ne = nao - 1) a block of its own after them.
__device__ __forceinline__ int df_blk(int x, int ne, int bs) { return x >= ne ? 1 << 30 : x / bs; }

// Upper part of a symmetric nao x nao matrix (column-major view) whose lower
// blocks (block of the row >= block of the column) are computed.
__global__ void k_df_mirror(double* K, int nao, int ne, int bs)
{
    const long long n = (long long)nao * nao;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(e / nao), r = (int)(e - (long long)c * nao);
        if (df_blk(r, ne, bs) < df_blk(c, ne, bs)) K[e] = K[(long long)r * nao + c];
    }
}

// Inside a workspace-size query rocBLAS calls only report sizes (statuses
// other than success): failures count outside a query only.
#define RBQ(ctx, query, x)                                                                            \
    do {                                                                                              \
        rocblas_status s_ = (x);                                                                      \
        if (!(query) && s_ != rocblas_status_success)                                                 \
            return ctx_fail(ctx, MJK_EDEVICE, "%s failed: %s", #x, rocblas_status_to_string(s_));     \
    } while (0)

// Blocks of the symmetric exchange product for nao AOs: one block (a single
// GEMM) below 640 AOs, else about 250 AOs per block, a width at which the
// GEMMs keep the full product's rate (49-57 TFLOP/s for 242-256 and 486-512
// rows, 30 for 184, 15 for 122) while the lower triangle needs (1 + 1/nb) / 2
// of its FLOPs.  C30H62 cc-pVDZ / cc-pVDZ-JKFIT (730 AOs, blocks 244, 244,
// 242): the build from the occupied orbitals 21.1 -> 17.8 ms.  kblocks > 0
// overrides.
int df_kblocks(int nao, int kblocks)
{
    if (kblocks > 0) return kblocks;
    return nao < 640 ? 1 : (nao + 125) / 250;
}

// K = Y^T Y (column-major nao x nao; Y: kin x nao, column-major) as the
// lower blocks of nb x nb even-sized blocks over the even part of the AOs, the
// odd last AO's row by DGEMV, then the mirror.
int df_syrk(mjk_ctx* ctx, rocblas_handle h, int nb, int nao, int kin, const double* Y, double* K, hipStream_t st)
{
    const double one = 1.0, zero = 0.0;
    const bool query = rocblas_is_device_memory_size_query(h);
    const int ne = nao & ~1;
    nb = std::max(1, std::min(nb, ne / 2));
    int bs = (ne + nb - 1) / nb;
    bs += bs & 1;   // even blocks (the last one too: ne is even)
    for (int i0 = 0; i0 < ne; i0 += bs) {
        const int ni = std::min(bs, ne - i0);
        for (int j0 = 0; j0 <= i0; j0 += bs) {
            const int nj = std::min(bs, ne - j0);
            RBQ(ctx, query, rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, ni, nj, kin, &one,
                                          Y + (size_t)i0 * kin, kin, Y + (size_t)j0 * kin, kin, &zero,
                                          K + (size_t)j0 * nao + i0, nao));
        }
    }
    if (ne < nao)   // row nao - 1: K[nao-1, c] = y_{nao-1} . y_c, all c
        RBQ(ctx, query, rocblas_dgemv(h, rocblas_operation_transpose, kin, nao, &one, Y, kin, Y + (size_t)ne * kin, 1,
                                      &zero, K + ne, nao));
    if (query || (bs >= ne && ne == nao)) return MJK_OK;
    k_df_mirror<<<1024, 256, 0, st>>>(K, nao, ne, bs);
    HC(ctx, hipGetLastError());
    return MJK_OK;
}

// K = B^T Y (column-major nao x nao; B, Y: kin x nao, column-major), even
// GEMM dimensions: the ne x ne block by DGEMM, an odd last AO's column and
// row by DGEMV.
int df_gemm_tn(mjk_ctx* ctx, rocblas_handle h, int nao, int kin, const double* B, const double* Y, double* K)
{
    const double one = 1.0, zero = 0.0;
    const bool query = rocblas_is_device_memory_size_query(h);
    const int ne = nao & ~1;
    if (ne > 0)
        RBQ(ctx, query, rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, ne, ne, kin, &one, B,
                                      kin, Y, kin, &zero, K, nao));
    if (ne < nao) {
        // column nao - 1: K[r, nao-1] = b_r . y_{nao-1}, all r
        RBQ(ctx, query, rocblas_dgemv(h, rocblas_operation_transpose, kin, nao, &one, B, kin, Y + (size_t)ne * kin, 1,
                                      &zero, K + (size_t)ne * nao, 1));
        // row nao - 1: K[nao-1, c] = b_{nao-1} . y_c, c < nao - 1
        RBQ(ctx, query, rocblas_dgemv(h, rocblas_operation_transpose, kin, ne, &one, Y, kin, B + (size_t)ne * kin, 1,
                                      &zero, K + ne, nao));
    }
    return MJK_OK;
}

// J and K of n_dm densities (device, nao x nao each, row-major), optional
// occupied orbitals X (n_dm x nao x nocc, row-major) for K
int get_jk_dev(mjk_df* df, const double* d_dms, int n_dm, const double* d_orbo, int nocc, int with_j, int with_k,
               double* d_vj, double* d_vk, bool orbo_colmajor = false)
{
    mjk_ctx* ctx = df->ctx;
    hipStream_t st = mjk_internal::ctx_stream(ctx);
    rocblas_handle h = (rocblas_handle)mjk_internal::ctx_rocblas(ctx);
    if (!h) return ctx_fail(ctx, MJK_EDEVICE, "rocblas_create_handle failed");
    RB(ctx, rocblas_set_stream(h, st));
    RB(ctx, rocblas_set_pointer_mode(h, rocblas_pointer_mode_host));
    int rc;
    const int nao = df->nao, naux = df->naux;
    const long long nn = (long long)nao * nao;
    const double one = 1.0, zero = 0.0;
    const int ncol = d_orbo ? nocc : nao;
    // workspace of these calls (sized by a query, once per shape)
    {
        size_t need = 0;
        RB(ctx, rocblas_start_device_memory_size_query(h));
        if (ncol > 0) {
            (void)rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_transpose, naux, ncol, nao,
                                                &one, df->B, naux, (rocblas_stride)naux * nao, d_dms, ncol, 0, &zero,
                                                df->Y, naux, (rocblas_stride)naux * ncol, nao);
            if (d_orbo)
                (void)df_syrk(ctx, h, df_kblocks(nao, df->kblocks), nao, naux * ncol, df->Y, d_vk, st);
            else
                (void)df_gemm_tn(ctx, h, nao, naux * ncol, df->B, df->Y, d_vk);
        }
        RB(ctx, rocblas_stop_device_memory_size_query(h, &need));
        if ((rc = workspace(df, h, need))) return rc;
    }
    if (with_j) {
        if ((rc = grow(ctx, &df->rho, &df->caprho, (size_t)naux * n_dm))) return rc;
        HC(ctx, hipMemsetAsync(df->rho, 0, sizeof(double) * (size_t)naux * n_dm, st));
        const long long ntri = (long long)nao * (nao + 1) / 2;
        const dim3 grid_rho((unsigned)((ntri + DF_ROWS - 1) / DF_ROWS), (unsigned)((naux + DF_COLS - 1) / DF_COLS));
        const size_t lds = sizeof(double) * naux;
        if (lds > 160 * 1024) return ctx_fail(ctx, MJK_EUNSUPPORTED, "mjk_df: %d fitting functions exceed LDS", naux);
        if (lds > 64 * 1024)
            HC(ctx, hipFuncSetAttribute((const void*)k_df_j, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        for (int k = 0; k < n_dm; ++k) {
            k_df_rho<<<grid_rho, 256, 0, st>>>(df->B, nao, naux, d_dms + (size_t)k * nn, df->rho + (size_t)k * naux);
            k_df_j<<<2048, 256, lds, st>>>(df->B, nao, naux, df->rho + (size_t)k * naux, d_vj + (size_t)k * nn);
        }
        HC(ctx, hipGetLastError());
    }
    if (with_k) {
        if ((rc = grow(ctx, &df->Y, &df->capY, (size_t)nao * naux * std::max(ncol, 1)))) return rc;
        for (int k = 0; k < n_dm; ++k) {
            double* K = d_vk + (size_t)k * nn;
            if (ncol == 0) {
                HC(ctx, hipMemsetAsync(K, 0, sizeof(double) * nn, st));
                continue;
            }
            // Y_i (naux x ncol) = B_i (naux x nao) op(M), M = D (row-major: op T) or X (row-major nao x nocc:
            // op T; column-major, leading dimension nao: op N)
            const double* M = d_orbo ? d_orbo + (size_t)k * nao * nocc : d_dms + (size_t)k * nn;
            const bool cm = d_orbo && orbo_colmajor;
            RB(ctx, rocblas_dgemm_strided_batched(h, rocblas_operation_none,
                                                  cm ? rocblas_operation_none : rocblas_operation_transpose, naux,
                                                  ncol, nao, &one, df->B, naux, (rocblas_stride)naux * nao, M,
                                                  cm ? nao : ncol, 0, &zero, df->Y, naux, (rocblas_stride)naux * ncol,
                                                  nao));
            if (d_orbo) {
                // K = Y^T Y over (P, o), symmetric
                if ((rc = df_syrk(ctx, h, df_kblocks(nao, df->kblocks), nao, naux * ncol, df->Y, K, st))) return rc;
            } else {
                // K (row-major) = B^T Y over (P, l): column-major C(k, i) = K[i][k]
                if ((rc = df_gemm_tn(ctx, h, nao, naux * nao, df->B, df->Y, K))) return rc;
            }
        }
    }
    return MJK_OK;
}

void release(mjk_df* df)
{
    if (!df) return;
    if (df->ctx) {
        (void)hipStreamSynchronize(mjk_internal::ctx_stream(df->ctx));
        if (rocblas_handle h = (rocblas_handle)mjk_internal::ctx_blas(df->ctx))
            if (df->ws) (void)rocblas_set_workspace(h, nullptr, 0);
        void* ps[] = {df->B, df->j2c, df->Y, df->D, df->rho, df->ws};
        for (void* p : ps)
            if (p) (void)mjk_device_free(df->ctx, p);
    }
    delete df;
}

}  // namespace

namespace mjk_internal {

int df_get_jk_dev(mjk_df* df, const double* d_dms, const double* d_orbo_colmajor, int nocc, double* d_vj, double* d_vk)
{
    return get_jk_dev(df, d_dms, 1, d_orbo_colmajor, nocc, 1, 1, d_vj, d_vk, true);
}

int df_nao(const mjk_df* df)
{
    return df->nao;
}

mjk_ctx* df_ctx(const mjk_df* df)
{
    return df->ctx;
}

}  // namespace mjk_internal

extern "C" {

int mjk_df_build(mjk_ctx* ctx, const int* atm, int natm, const int* bas, int nbas, const double* env,
                 const int* aux_atm, int aux_natm, const int* aux_bas, int aux_nbas, const double* aux_env,
                 double lindep, mjk_df** out)
{
    if (!ctx || !atm || !bas || !env || !aux_atm || !aux_bas || !aux_env || !out || natm <= 0 || nbas <= 0 ||
        aux_natm <= 0 || aux_nbas <= 0)
        return ctx_fail(ctx, MJK_EINVAL, "bad arguments to mjk_df_build");
    *out = nullptr;
    std::lock_guard<std::recursive_mutex> lock(mjk_internal::ctx_mutex(ctx));
    if (hipSetDevice(mjk_internal::ctx_device(ctx)) != hipSuccess)
        return ctx_fail(ctx, MJK_EDEVICE, "hipSetDevice failed");
    if (env[8] != 0.0) return ctx_fail(ctx, MJK_EUNSUPPORTED, "mjk_df_build: Coulomb operator only (omega = 0)");
    mjk_df* df = new mjk_df();
    df->ctx = ctx;
    if (const char* kb = getenv("MJK_DF_KBLOCKS")) df->kblocks = std::max(1, std::min(16, atoi(kb)));
    const int rc = build(ctx, df, atm, natm, bas, nbas, env, aux_atm, aux_natm, aux_bas, aux_nbas, aux_env, lindep);
    if (rc) {
        release(df);
        return rc;
    }
    *out = df;
    return MJK_OK;
}

int mjk_df_naux(const mjk_df* df)
{
    return df ? df->naux : MJK_EINVAL;
}

int mjk_df_get_jk_device(mjk_ctx* ctx, mjk_df* df, const double* d_dms, int n_dm, const double* d_orbo, int nocc,
                         int with_j, int with_k, double* d_vj, double* d_vk)
{
    if (!ctx || !df || df->ctx != ctx || !d_dms || n_dm < 1 || (with_j && !d_vj) || (with_k && !d_vk) ||
        (d_orbo && nocc < 0) || !df->fitted)
        return ctx_fail(ctx, MJK_EINVAL, "bad arguments to mjk_df_get_jk_device");
    std::lock_guard<std::recursive_mutex> lock(mjk_internal::ctx_mutex(ctx));
    if (hipSetDevice(mjk_internal::ctx_device(ctx)) != hipSuccess)
        return ctx_fail(ctx, MJK_EDEVICE, "hipSetDevice failed");
    return get_jk_dev(df, d_dms, n_dm, d_orbo, nocc, with_j, with_k, d_vj, d_vk);
}

int mjk_df_get_jk(mjk_ctx* ctx, mjk_df* df, const double* dms, int n_dm, const double* orbo, int nocc, int with_j,
                  int with_k, double* vj, double* vk)
{
    if (!ctx || !df || df->ctx != ctx || !dms || n_dm < 1 || (with_j && !vj) || (with_k && !vk) ||
        (orbo && nocc < 0) || !df->fitted)
        return ctx_fail(ctx, MJK_EINVAL, "bad arguments to mjk_df_get_jk");
    std::lock_guard<std::recursive_mutex> lock(mjk_internal::ctx_mutex(ctx));
    if (hipSetDevice(mjk_internal::ctx_device(ctx)) != hipSuccess)
        return ctx_fail(ctx, MJK_EDEVICE, "hipSetDevice failed");
    hipStream_t st = mjk_internal::ctx_stream(ctx);
    const size_t nn = (size_t)df->nao * df->nao, no = orbo ? (size_t)df->nao * nocc : 0;
    // layout of df->D: densities, orbitals, J, K
    const size_t need = n_dm * (nn + no + (with_j ? nn : 0) + (with_k ? nn : 0));
    int rc = grow(ctx, &df->D, &df->capD, need);
    if (rc) return rc;
    double* dD = df->D;
    double* dX = dD + n_dm * nn;
    double* dJ = dX + n_dm * no;
    double* dK = dJ + (with_j ? n_dm * nn : 0);
    HC(ctx, hipMemcpyAsync(dD, dms, sizeof(double) * n_dm * nn, hipMemcpyHostToDevice, st));
    if (orbo) HC(ctx, hipMemcpyAsync(dX, orbo, sizeof(double) * n_dm * no, hipMemcpyHostToDevice, st));
    if ((rc = get_jk_dev(df, dD, n_dm, orbo ? dX : nullptr, nocc, with_j, with_k, dJ, dK))) return rc;
    if (with_j) HC(ctx, hipMemcpyAsync(vj, dJ, sizeof(double) * n_dm * nn, hipMemcpyDeviceToHost, st));
    if (with_k) HC(ctx, hipMemcpyAsync(vk, dK, sizeof(double) * n_dm * nn, hipMemcpyDeviceToHost, st));
    HC(ctx, hipStreamSynchronize(st));
    return MJK_OK;
}

int mjk_df_fetch(mjk_ctx* ctx, mjk_df* df, int which, double* out)
{
    if (!ctx || !df || df->ctx != ctx || !out || which < 0 || which > 1)
        return ctx_fail(ctx, MJK_EINVAL, "bad arguments to mjk_df_fetch");
    std::lock_guard<std::recursive_mutex> lock(mjk_internal::ctx_mutex(ctx));
    hipStream_t st = mjk_internal::ctx_stream(ctx);
    if (which == 0)
        HC(ctx, hipMemcpyAsync(out, df->j2c, sizeof(double) * (size_t)df->naux_raw * df->naux_raw,
                               hipMemcpyDeviceToHost, st));
    else
        HC(ctx, hipMemcpyAsync(out, df->B, sizeof(double) * (size_t)df->naux * df->nao * df->nao,
                               hipMemcpyDeviceToHost, st));
    HC(ctx, hipStreamSynchronize(st));
    return MJK_OK;
}

void mjk_df_destroy(mjk_df* df)
{
    if (!df) return;
    mjk_ctx* ctx = df->ctx;
    std::lock_guard<std::recursive_mutex> lock(mjk_internal::ctx_mutex(ctx));
    release(df);
}

}  // extern "C"
