// Device-resident closed-shell SCF: the RHF loop of pyscf/scf/hf.py:47-242
// with every n x n matrix in HBM.  Per cycle the host sees a handful of
// scalars (the DIIS overlaps, the energy terms, the gradient norm) and sends
// back the DIIS coefficients; the density, Fock matrix, DIIS history and MO
// coefficients never leave the device.
//
//   * F = H + V, CDIIS on the commutator error S D F - F D S (the error
//     vectors' overlaps are all the extrapolation needs; pyscf/scf/diis.py:40-66,
//     pyscf/lib/diis.py:202-275), from cycle 1, space 8 by default;
//   * optional level shift F + shift (S - S D S / 2) after DIIS
//     (pyscf/scf/hf.py:766-786);
//   * F C = S C e through X = V w^-1/2 from the overlap's eigenpairs (one
//     standard eigenproblem of X^T F X per cycle), or the generalized solver
//     when the smallest overlap eigenvalue is below 1e-10 (pyscf/scf/hf.py:1340-1347);
//   * aufbau occupation D = 2 C_occ C_occ^T (pyscf/scf/hf.py:840);
//   * incremental V = V_last + J(dD) - K(dD)/2 on the J/K engine, screened at
//     direct_scf_tol (pyscf/scf/hf.py:2416-2427);
//   * converged when |dE| < conv_tol and 2 |C_vir^T F C_occ| < conv_tol_grad
//     (pyscf/scf/hf.py:194-197).
// Matrices are column-major n x n; the symmetric ones are layout-free, the MO
// coefficients are fetched as row-major (nao, nmo).  rocBLAS and rocSOLVER
// run on the context's stream with a workspace allocated through the
// context (tracked against its budget), so the loop allocates nothing after
// set-up.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <shared_mutex>
#include <vector>

#include "mjk.h"
#include "mjk_internal.h"

using mjk_internal::ctx_fail;

struct mjk_scf {
    mjk_ctx* ctx = nullptr;
    mjk_basis* basis = nullptr;
    int n = 0, nocc = 0;
    double enuc = 0;
    rocblas_handle h = nullptr;
    void* ws = nullptr;
    size_t ws_bytes = 0;
    // n x n matrices
    double *S = nullptr, *H = nullptr, *X = nullptr, *D = nullptr, *Dlast = nullptr, *dD = nullptr;
    double *V = nullptr, *F = nullptr, *Fx = nullptr, *T1 = nullptr, *T2 = nullptr, *J = nullptr, *K = nullptr;
    double *C = nullptr, *Sc = nullptr;
    double* w = nullptr;        // n eigenvalues
    double* e = nullptr;        // n off-diagonal work of the eigensolvers
    double* ring_f = nullptr;   // space x n^2 Fock history
    double* ring_e = nullptr;   // space x n^2 error history
    double* scal = nullptr;     // device scalars: [0, space) overlaps, space.. energy terms, gradient
    double* coef = nullptr;     // space DIIS coefficients
    rocblas_int* info = nullptr;
    bool orth = true;           // X valid (else the generalized solver)
    int space = 8;
    int nvec = 0, next = 0;     // stored vectors, next slot
    std::vector<double> B;      // space x space overlaps of the stored errors
    // results
    double e_tot = 0, e_last = 0, gnorm = 0;
    int cycles = 0, converged = 0;
    // density fitting (mjk_rhf_set_df): full J/K of D every cycle, K from the
    // occupied orbitals X = sqrt(2) C_occ
    mjk_df* df = nullptr;
    double* Xo = nullptr;
};

namespace {

__global__ void k_lincomb(long long n, double a, const double* x, double b, const double* y, double* out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a * x[i] + (y ? b * y[i] : 0.0);
}

// out = out + a x + b y
__global__ void k_accum2(long long n, double* out, double a, const double* x, double b, const double* y)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += a * x[i] + b * y[i];
}

// err = M^T - M
__global__ void k_commutator(int n, const double* M, double* err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i < n) err[(size_t)j * n + i] = M[(size_t)i * n + j] - M[(size_t)j * n + i];
}

// D = (D + D^T) / 2
__global__ void k_symmetrize(int n, double* D)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i < j) {
        const double v = 0.5 * (D[(size_t)j * n + i] + D[(size_t)i * n + j]);
        D[(size_t)j * n + i] = v;
        D[(size_t)i * n + j] = v;
    }
}

// X[:, j] *= w[j]^-1/2
__global__ void k_scale_cols(int n, double* X, const double* w)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i < n) X[(size_t)j * n + i] *= 1.0 / sqrt(w[j]);
}

constexpr int NT = 256;

inline dim3 grid1(long long n) { return dim3((unsigned)((n + NT - 1) / NT)); }
inline dim3 grid2(int n) { return dim3((unsigned)((n + NT - 1) / NT), (unsigned)n); }

#define RBCHK(s, x)                                                                                  \
    do {                                                                                             \
        rocblas_status st_ = (x);                                                                    \
        if (st_ != rocblas_status_success)                                                           \
            return ctx_fail((s)->ctx, MJK_EDEVICE, "%s failed: %s", #x, rocblas_status_to_string(st_)); \
    } while (0)

#define HCHK(s, x)                                                                                   \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) return ctx_fail((s)->ctx, MJK_EDEVICE, "%s failed: %s", #x, hipGetErrorString(e_)); \
    } while (0)

// c = alpha op(a) op(b) + beta c, all n x n unless given
int gemm(mjk_scf* s, rocblas_operation ta, rocblas_operation tb, int m, int nn, int k, double alpha,
         const double* a, int lda, const double* b, int ldb, double beta, double* c, int ldc)
{
    RBCHK(s, rocblas_dgemm(s->h, ta, tb, m, nn, k, &alpha, a, lda, b, ldb, &beta, c, ldc));
    return MJK_OK;
}

// Symmetric solve of the (m+1) x (m+1) DIIS system by Gaussian elimination
// with partial pivoting; a singular system falls back to the least-squares
// solution through the eigenpairs of the matrix (Jacobi rotations), as the
// reference falls back to a least-squares solver.
void diis_solve(int m, const std::vector<double>& Bm, std::vector<double>& c)
{
    const int N = m + 1;
    std::vector<double> A(N * N), rhs(N, 0.0);
    A[0] = 0;
    for (int i = 1; i < N; ++i) A[i] = A[i * N] = 1.0;
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) A[(i + 1) * N + j + 1] = Bm[i * m + j];
    rhs[0] = 1.0;
    std::vector<double> M = A, r = rhs;
    bool ok = true;
    double amax = 0;
    for (double v : A) amax = std::max(amax, std::fabs(v));
    for (int col = 0; col < N && ok; ++col) {
        int piv = col;
        for (int i = col + 1; i < N; ++i)
            if (std::fabs(M[i * N + col]) > std::fabs(M[piv * N + col])) piv = i;
        if (M[piv * N + col] == 0.0) {   // exactly singular, as a direct solver reports it
            ok = false;
            break;
        }
        if (piv != col) {
            for (int j = 0; j < N; ++j) std::swap(M[col * N + j], M[piv * N + j]);
            std::swap(r[col], r[piv]);
        }
        for (int i = col + 1; i < N; ++i) {
            const double f = M[i * N + col] / M[col * N + col];
            for (int j = col; j < N; ++j) M[i * N + j] -= f * M[col * N + j];
            r[i] -= f * r[col];
        }
    }
    std::vector<double> x(N, 0.0);
    if (ok) {
        for (int i = N - 1; i >= 0; --i) {
            double v = r[i];
            for (int j = i + 1; j < N; ++j) v -= M[i * N + j] * x[j];
            x[i] = v / M[i * N + i];
        }
        for (double v : x) ok = ok && std::isfinite(v);
    }
    if (!ok) {
        // pseudo-inverse: A = Q diag(l) Q^T by cyclic Jacobi
        std::vector<double> a = A, q(N * N, 0.0);
        for (int i = 0; i < N; ++i) q[i * N + i] = 1.0;
        for (int sweep = 0; sweep < 100; ++sweep) {
            double off = 0;
            for (int i = 0; i < N; ++i)
                for (int j = i + 1; j < N; ++j) off += a[i * N + j] * a[i * N + j];
            if (off < 1e-30 * (amax * amax + 1e-300)) break;
            for (int p = 0; p < N; ++p)
                for (int qq = p + 1; qq < N; ++qq) {
                    const double apq = a[p * N + qq];
                    if (apq == 0.0) continue;
                    const double th = 0.5 * (a[qq * N + qq] - a[p * N + p]) / apq;
                    const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                    const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
                    for (int k = 0; k < N; ++k) {
                        const double akp = a[k * N + p], akq = a[k * N + qq];
                        a[k * N + p] = cs * akp - sn * akq;
                        a[k * N + qq] = sn * akp + cs * akq;
                    }
                    for (int k = 0; k < N; ++k) {
                        const double apk = a[p * N + k], aqk = a[qq * N + k];
                        a[p * N + k] = cs * apk - sn * aqk;
                        a[qq * N + k] = sn * apk + cs * aqk;
                    }
                    for (int k = 0; k < N; ++k) {
                        const double qkp = q[k * N + p], qkq = q[k * N + qq];
                        q[k * N + p] = cs * qkp - sn * qkq;
                        q[k * N + qq] = sn * qkp + cs * qkq;
                    }
                }
        }
        double lmax = 0;
        for (int i = 0; i < N; ++i) lmax = std::max(lmax, std::fabs(a[i * N + i]));
        const double cut = 1e-15 * N * lmax;
        std::fill(x.begin(), x.end(), 0.0);
        for (int k = 0; k < N; ++k) {
            const double l = a[k * N + k];
            if (std::fabs(l) <= cut) continue;
            double qr = 0;
            for (int i = 0; i < N; ++i) qr += q[i * N + k] * rhs[i];
            for (int i = 0; i < N; ++i) x[i] += q[i * N + k] * qr / l;
        }
    }
    c.assign(x.begin() + 1, x.end());
}

// The context's rocBLAS handle, made on first use.  Creating or destroying a
// handle may synchronise the device, which would break another thread's
// stream capture: both run with no capture in flight.
rocblas_handle blas_of(mjk_ctx* ctx)
{
    return (rocblas_handle)mjk_internal::ctx_rocblas(ctx);
}

int alloc(mjk_scf* s, double** p, size_t elems)
{
    return mjk_device_alloc(s->ctx, elems * sizeof(double), (void**)p);
}

// F C = S C e for the Fock matrix in Fx: eigenvalues to w, vectors to C
int solve_fock(mjk_scf* s)
{
    const int n = s->n;
    const size_t nn = (size_t)n * n;
    hipStream_t st = mjk_internal::ctx_stream(s->ctx);
    if (s->orth) {
        int rc;
        if ((rc = gemm(s, rocblas_operation_none, rocblas_operation_none, n, n, n, 1.0, s->Fx, n, s->X, n, 0.0,
                       s->T1, n)))
            return rc;
        if ((rc = gemm(s, rocblas_operation_transpose, rocblas_operation_none, n, n, n, 1.0, s->X, n, s->T1, n,
                       0.0, s->T2, n)))
            return rc;
        RBCHK(s, rocsolver_dsyevd(s->h, rocblas_evect_original, rocblas_fill_upper, n, s->T2, n, s->w, s->e,
                                  s->info));
        return gemm(s, rocblas_operation_none, rocblas_operation_none, n, n, n, 1.0, s->X, n, s->T2, n, 0.0, s->C,
                    n);
    }
    HCHK(s, hipMemcpyAsync(s->C, s->Fx, nn * sizeof(double), hipMemcpyDeviceToDevice, st));
    HCHK(s, hipMemcpyAsync(s->Sc, s->S, nn * sizeof(double), hipMemcpyDeviceToDevice, st));
    RBCHK(s, rocsolver_dsygvd(s->h, rocblas_eform_ax, rocblas_evect_original, rocblas_fill_upper, n, s->C, n, s->Sc,
                              n, s->w, s->e, s->info));
    return MJK_OK;
}

// D = 2 C_occ C_occ^T (eigenvalues ascending: the first nocc columns)
int make_density(mjk_scf* s)
{
    const int n = s->n;
    int rc;
    if ((rc = gemm(s, rocblas_operation_none, rocblas_operation_transpose, n, n, s->nocc, 2.0, s->C, n, s->C, n,
                   0.0, s->D, n)))
        return rc;
    k_symmetrize<<<grid2(n), NT, 0, mjk_internal::ctx_stream(s->ctx)>>>(n, s->D);
    return MJK_OK;
}

int check_info(mjk_scf* s, const char* what)
{
    rocblas_int info = 0;
    hipStream_t st = mjk_internal::ctx_stream(s->ctx);
    HCHK(s, hipMemcpyAsync(&info, s->info, sizeof(info), hipMemcpyDeviceToHost, st));
    HCHK(s, hipStreamSynchronize(st));
    if (info != 0) return ctx_fail(s->ctx, MJK_EDEVICE, "%s: eigensolver info %d", what, (int)info);
    return MJK_OK;
}

__global__ void k_scale(long long n, double a, const double* x, double* out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a * x[i];
}

// J/K of src into V (V = J - K/2, or V += J - K/2 when accumulate)
int veff(mjk_scf* s, const double* src, bool accumulate, double direct_scf_tol, double fp32_tol)
{
    const long long nn = (long long)s->n * s->n;
    int rc;
    if (s->df) {
        // density fitting: J and K of the full density from the orbitals (src is D)
        const long long no = (long long)s->n * s->nocc;
        k_scale<<<grid1(no), NT, 0, mjk_internal::ctx_stream(s->ctx)>>>(no, std::sqrt(2.0), s->C, s->Xo);
        rc = mjk_internal::df_get_jk_dev(s->df, src, s->Xo, s->nocc, s->J, s->K);
    } else {
        rc = mjk_get_jk_device(s->ctx, s->basis, src, 1, 1, 1, 1, direct_scf_tol, fp32_tol, s->J, s->K, nullptr);
    }
    if (rc) return rc;
    hipStream_t st = mjk_internal::ctx_stream(s->ctx);
    if (accumulate)
        k_accum2<<<grid1(nn), NT, 0, st>>>(nn, s->V, 1.0, s->J, -0.5, s->K);
    else
        k_lincomb<<<grid1(nn), NT, 0, st>>>(nn, 1.0, s->J, -0.5, s->K, s->V);
    return MJK_OK;
}

// energy terms tr(H D), tr(V D) and F = H + V; scalars into scal[space], scal[space + 1]
int energy_and_fock(mjk_scf* s)
{
    const long long nn = (long long)s->n * s->n;
    hipStream_t st = mjk_internal::ctx_stream(s->ctx);
    RBCHK(s, rocblas_set_pointer_mode(s->h, rocblas_pointer_mode_device));
    rocblas_status a = rocblas_ddot(s->h, (rocblas_int)nn, s->H, 1, s->D, 1, s->scal + s->space);
    rocblas_status b = rocblas_ddot(s->h, (rocblas_int)nn, s->V, 1, s->D, 1, s->scal + s->space + 1);
    RBCHK(s, rocblas_set_pointer_mode(s->h, rocblas_pointer_mode_host));
    RBCHK(s, a);
    RBCHK(s, b);
    k_lincomb<<<grid1(nn), NT, 0, st>>>(nn, 1.0, s->H, 1.0, s->V, s->F);
    return MJK_OK;
}

// |g| = 2 |C_vir^T F C_occ| into scal[space + 2]
int gradient(mjk_scf* s)
{
    const int n = s->n, nocc = s->nocc, nvir = n - nocc;
    if (nvir <= 0 || nocc <= 0) {
        HCHK(s, hipMemsetAsync(s->scal + s->space + 2, 0, sizeof(double), mjk_internal::ctx_stream(s->ctx)));
        return MJK_OK;
    }
    int rc;
    if ((rc = gemm(s, rocblas_operation_none, rocblas_operation_none, n, nocc, n, 1.0, s->F, n, s->C, n, 0.0, s->T1,
                   n)))
        return rc;
    if ((rc = gemm(s, rocblas_operation_transpose, rocblas_operation_none, nvir, nocc, n, 2.0, s->C + (size_t)nocc * n,
                   n, s->T1, n, 0.0, s->T2, nvir)))
        return rc;
    RBCHK(s, rocblas_set_pointer_mode(s->h, rocblas_pointer_mode_device));
    rocblas_status g = rocblas_dnrm2(s->h, (rocblas_int)((long long)nvir * nocc), s->T2, 1, s->scal + s->space + 2);
    RBCHK(s, rocblas_set_pointer_mode(s->h, rocblas_pointer_mode_host));
    RBCHK(s, g);
    return MJK_OK;
}

// CDIIS: stores F and its error, extrapolates into Fx
int diis(mjk_scf* s)
{
    const int n = s->n;
    const long long nn = (long long)n * n;
    hipStream_t st = mjk_internal::ctx_stream(s->ctx);
    int rc;
    // S D F: T1 = S D, T2 = T1 F
    if ((rc = gemm(s, rocblas_operation_none, rocblas_operation_none, n, n, n, 1.0, s->S, n, s->D, n, 0.0, s->T1, n)))
        return rc;
    if ((rc = gemm(s, rocblas_operation_none, rocblas_operation_none, n, n, n, 1.0, s->T1, n, s->F, n, 0.0, s->T2, n)))
        return rc;
    const int slot = s->next;
    double* ef = s->ring_e + (size_t)slot * nn;
    HCHK(s, hipMemcpyAsync(s->ring_f + (size_t)slot * nn, s->F, nn * sizeof(double), hipMemcpyDeviceToDevice, st));
    k_commutator<<<grid2(n), NT, 0, st>>>(n, s->T2, ef);
    s->next = (s->next + 1) % s->space;
    s->nvec = std::min(s->nvec + 1, s->space);
    RBCHK(s, rocblas_set_pointer_mode(s->h, rocblas_pointer_mode_device));
    rocblas_status ds = rocblas_status_success;
    for (int j = 0; j < s->nvec && ds == rocblas_status_success; ++j)
        ds = rocblas_ddot(s->h, (rocblas_int)nn, ef, 1, s->ring_e + (size_t)j * nn, 1, s->scal + j);
    RBCHK(s, rocblas_set_pointer_mode(s->h, rocblas_pointer_mode_host));
    RBCHK(s, ds);
    std::vector<double> row(s->nvec);
    HCHK(s, hipMemcpyAsync(row.data(), s->scal, s->nvec * sizeof(double), hipMemcpyDeviceToHost, st));
    HCHK(s, hipStreamSynchronize(st));
    for (int j = 0; j < s->nvec; ++j) s->B[slot * s->space + j] = s->B[j * s->space + slot] = row[j];
    if (s->nvec < 2) {
        HCHK(s, hipMemcpyAsync(s->Fx, s->F, nn * sizeof(double), hipMemcpyDeviceToDevice, st));
        return MJK_OK;
    }
    const int m = s->nvec;
    std::vector<double> Bm(m * m), c;
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) Bm[i * m + j] = s->B[i * s->space + j];
    diis_solve(m, Bm, c);
    HCHK(s, hipMemcpyAsync(s->coef, c.data(), m * sizeof(double), hipMemcpyHostToDevice, st));
    const double one = 1.0, zero = 0.0;
    // Fx = sum_j c_j F_j: the history as an (n^2 x m) matrix times c
    RBCHK(s, rocblas_dgemv(s->h, rocblas_operation_none, (rocblas_int)nn, m, &one, s->ring_f, (rocblas_int)nn, s->coef,
                           1, &zero, s->Fx, 1));
    HCHK(s, hipStreamSynchronize(st));   // c lives on this stack frame
    return MJK_OK;
}

int level_shift(mjk_scf* s, double shift)
{
    // Fx += shift S - shift/2 (S D) S; T1 = S D from the DIIS step
    const int n = s->n;
    const long long nn = (long long)n * n;
    int rc;
    if ((rc = gemm(s, rocblas_operation_none, rocblas_operation_none, n, n, n, 1.0, s->S, n, s->D, n, 0.0, s->T1, n)))
        return rc;
    if ((rc = gemm(s, rocblas_operation_none, rocblas_operation_none, n, n, n, 1.0, s->T1, n, s->S, n, 0.0, s->T2, n)))
        return rc;
    k_accum2<<<grid1(nn), NT, 0, mjk_internal::ctx_stream(s->ctx)>>>(nn, s->Fx, shift, s->S, -0.5 * shift, s->T2);
    return MJK_OK;
}

int read_scalars(mjk_scf* s, double* e1, double* e2, double* g)
{
    double v[3];
    hipStream_t st = mjk_internal::ctx_stream(s->ctx);
    HCHK(s, hipMemcpyAsync(v, s->scal + s->space, 3 * sizeof(double), hipMemcpyDeviceToHost, st));
    HCHK(s, hipStreamSynchronize(st));
    *e1 = v[0];
    *e2 = v[1];
    *g = v[2];
    return MJK_OK;
}

void release(mjk_scf* s)
{
    if (!s) return;
    if (s->ctx) {
        (void)hipStreamSynchronize(mjk_internal::ctx_stream(s->ctx));
        double** bufs[] = {&s->S, &s->H, &s->X, &s->D, &s->Dlast, &s->dD, &s->V, &s->F, &s->Fx, &s->T1,
                           &s->T2, &s->J, &s->K, &s->C, &s->Sc, &s->w, &s->e, &s->ring_f, &s->ring_e,
                           &s->scal, &s->coef, &s->Xo};
        for (double** b : bufs)
            if (*b) (void)mjk_device_free(s->ctx, *b);
        if (s->info) (void)mjk_device_free(s->ctx, s->info);
        if (s->h) (void)rocblas_set_workspace(s->h, nullptr, 0);   // the handle outlives this SCF
        if (s->ws) (void)mjk_device_free(s->ctx, s->ws);
    }
    delete s;
}

int create(mjk_ctx* ctx, mjk_basis* basis, const double* s1e, const double* h1e, int nocc, double enuc,
           int diis_space, mjk_scf** out)
{
    const int n = mjk_basis_nao(basis, 0);
    if (n <= 0 || mjk_basis_slot_size(basis) != (long long)n * n)
        return ctx_fail(ctx, MJK_EINVAL, "mjk_rhf_create: the basis must hold one molecule");
    if (nocc < 0 || nocc > n) return ctx_fail(ctx, MJK_EINVAL, "mjk_rhf_create: nocc %d outside [0, %d]", nocc, n);
    if (diis_space < 1 || diis_space > 64)
        return ctx_fail(ctx, MJK_EINVAL, "mjk_rhf_create: diis_space %d outside [1, 64]", diis_space);
    if ((long long)n * n > INT32_MAX) return ctx_fail(ctx, MJK_EINVAL, "mjk_rhf_create: %d AOs is too many", n);
    mjk_scf* s = new mjk_scf();
    s->ctx = ctx;
    s->basis = basis;
    s->n = n;
    s->nocc = nocc;
    s->enuc = enuc;
    s->space = diis_space;
    s->B.assign((size_t)diis_space * diis_space, 0.0);
    hipStream_t st = mjk_internal::ctx_stream(ctx);
    const size_t nn = (size_t)n * n;
    int rc = MJK_OK;
    if (!(s->h = blas_of(ctx))) {
        release(s);
        return ctx_fail(ctx, MJK_EDEVICE, "rocblas_create_handle failed");
    }
    (void)rocblas_set_stream(s->h, st);
    double** mats[] = {&s->S, &s->H, &s->X, &s->D, &s->Dlast, &s->dD, &s->V, &s->F, &s->Fx, &s->T1, &s->T2,
                       &s->J, &s->K, &s->C, &s->Sc};
    for (double** m : mats)
        if (!rc) rc = alloc(s, m, nn);
    if (!rc) rc = alloc(s, &s->w, n);
    if (!rc) rc = alloc(s, &s->e, n);
    if (!rc) rc = alloc(s, &s->ring_f, nn * diis_space);
    if (!rc) rc = alloc(s, &s->ring_e, nn * diis_space);
    if (!rc) rc = alloc(s, &s->scal, diis_space + 8);
    if (!rc) rc = alloc(s, &s->coef, diis_space);
    if (!rc) rc = mjk_device_alloc(ctx, 16 * sizeof(rocblas_int), (void**)&s->info);
    if (rc) {
        release(s);
        return rc;
    }
    // workspace of every rocBLAS / rocSOLVER call of the loop, sized once
    size_t ws = 0;
    if (rocblas_start_device_memory_size_query(s->h) == rocblas_status_success) {
        const double one = 1.0, zero = 0.0;
        (void)rocsolver_dsyevd(s->h, rocblas_evect_original, rocblas_fill_upper, n, s->T2, n, s->w, s->e, s->info);
        (void)rocsolver_dsygvd(s->h, rocblas_eform_ax, rocblas_evect_original, rocblas_fill_upper, n, s->C, n, s->Sc,
                               n, s->w, s->e, s->info);
        (void)rocblas_dgemm(s->h, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, s->S, n, s->D, n,
                            &zero, s->T1, n);
        (void)rocblas_dgemm(s->h, rocblas_operation_transpose, rocblas_operation_none, n, n, n, &one, s->S, n, s->D, n,
                            &zero, s->T1, n);
        (void)rocblas_dgemm(s->h, rocblas_operation_none, rocblas_operation_transpose, n, n, std::max(nocc, 1), &one,
                            s->C, n, s->C, n, &zero, s->D, n);
        (void)rocblas_dgemv(s->h, rocblas_operation_none, (rocblas_int)nn, diis_space, &one, s->ring_f,
                            (rocblas_int)nn, s->coef, 1, &zero, s->Fx, 1);
        (void)rocblas_set_pointer_mode(s->h, rocblas_pointer_mode_device);
        (void)rocblas_ddot(s->h, (rocblas_int)nn, s->H, 1, s->D, 1, s->scal);
        (void)rocblas_dnrm2(s->h, (rocblas_int)nn, s->T2, 1, s->scal);
        (void)rocblas_set_pointer_mode(s->h, rocblas_pointer_mode_host);
        if (rocblas_stop_device_memory_size_query(s->h, &ws) != rocblas_status_success) ws = 0;
    }
    // at least 4 n^2 doubles and 32 MiB, should the query come back short
    ws = std::max(ws, nn * 4 * sizeof(double) + ((size_t)32 << 20));
    if ((rc = mjk_device_alloc(ctx, ws, &s->ws))) {
        release(s);
        return rc;
    }
    s->ws_bytes = ws;
    if (rocblas_set_workspace(s->h, s->ws, ws) != rocblas_status_success) {
        release(s);
        return ctx_fail(ctx, MJK_EDEVICE, "rocblas_set_workspace failed");
    }
    (void)rocblas_set_pointer_mode(s->h, rocblas_pointer_mode_host);
    if (hipMemcpyAsync(s->S, s1e, nn * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(s->H, h1e, nn * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(s->X, s1e, nn * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess) {
        release(s);
        return ctx_fail(ctx, MJK_EDEVICE, "mjk_rhf_create: upload failed");
    }
    // X = V w^-1/2 from the overlap's eigenpairs, unless S is nearly singular
    if (rocsolver_dsyevd(s->h, rocblas_evect_original, rocblas_fill_upper, n, s->X, n, s->w, s->e, s->info) !=
            rocblas_status_success ||
        (rc = check_info(s, "overlap eigenpairs"))) {
        release(s);
        return rc ? rc : ctx_fail(ctx, MJK_EDEVICE, "rocsolver_dsyevd failed");
    }
    double wmin = 0;
    if (hipMemcpyAsync(&wmin, s->w, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        release(s);
        return ctx_fail(ctx, MJK_EDEVICE, "mjk_rhf_create: download failed");
    }
    s->orth = wmin > 1e-10;
    if (s->orth) k_scale_cols<<<grid2(n), NT, 0, st>>>(n, s->X, s->w);
    *out = s;
    return MJK_OK;
}

int run(mjk_scf* s, const double* dm0, int max_cycle, double conv_tol, double conv_tol_grad, double direct_scf_tol,
        double fp32_tol, double shift, int* cycles_out)
{
    const int n = s->n;
    const long long nn = (long long)n * n;
    hipStream_t st = mjk_internal::ctx_stream(s->ctx);
    int rc;
    s->nvec = s->next = 0;
    std::fill(s->B.begin(), s->B.end(), 0.0);
    s->converged = 0;
    s->cycles = 0;
    // initial density: the caller's, or the core Hamiltonian's ('1e' guess)
    if (dm0) {
        if (s->df)
            return ctx_fail(s->ctx, MJK_EINVAL, "mjk_rhf_run: a density-fitted run starts from the core guess");
        HCHK(s, hipMemcpyAsync(s->D, dm0, nn * sizeof(double), hipMemcpyHostToDevice, st));
    } else {
        HCHK(s, hipMemcpyAsync(s->Fx, s->H, nn * sizeof(double), hipMemcpyDeviceToDevice, st));
        if ((rc = solve_fock(s)) || (rc = check_info(s, "core guess")) || (rc = make_density(s))) return rc;
    }
    if ((rc = veff(s, s->D, false, direct_scf_tol, fp32_tol)) || (rc = energy_and_fock(s))) return rc;
    double e1, e2, g;
    if ((rc = read_scalars(s, &e1, &e2, &g))) return rc;
    s->e_tot = e1 + 0.5 * e2 + s->enuc;
    for (int cycle = 0; cycle < max_cycle; ++cycle) {
        const double last_e = s->e_tot;
        if ((rc = diis(s))) return rc;
        if (std::fabs(shift) > 1e-4 && (rc = level_shift(s, shift))) return rc;
        if ((rc = solve_fock(s)) || (rc = check_info(s, "Fock eigenpairs"))) return rc;
        std::swap(s->D, s->Dlast);
        if ((rc = make_density(s))) return rc;
        if (s->df) {   // density fitting: full builds (the reference turns direct_scf off for DF)
            if ((rc = veff(s, s->D, false, direct_scf_tol, fp32_tol))) return rc;
        } else {
            k_lincomb<<<grid1(nn), NT, 0, st>>>(nn, 1.0, s->D, -1.0, s->Dlast, s->dD);
            if ((rc = veff(s, s->dD, true, direct_scf_tol, fp32_tol))) return rc;
        }
        if ((rc = energy_and_fock(s)) || (rc = gradient(s)) || (rc = read_scalars(s, &e1, &e2, &g))) return rc;
        s->e_tot = e1 + 0.5 * e2 + s->enuc;
        s->gnorm = g;
        s->cycles = cycle + 1;
        if (std::fabs(s->e_tot - last_e) < conv_tol && g < conv_tol_grad) {
            s->converged = 1;
            break;
        }
    }
    if (cycles_out) *cycles_out = s->cycles;
    return MJK_OK;
}

}  // namespace

namespace mjk_internal {

void* ctx_rocblas(mjk_ctx* ctx)
{
    if (void* h = ctx_blas(ctx)) return h;
    std::unique_lock<std::shared_mutex> lk(capture_mutex());
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) return nullptr;
    ctx_set_blas(ctx, h, [](void* p) { (void)rocblas_destroy_handle((rocblas_handle)p); });
    return h;
}

}  // namespace mjk_internal

extern "C" {

int mjk_rhf_create(mjk_ctx* ctx, mjk_basis* basis, const double* s1e, const double* h1e, int nocc, double enuc,
                   int diis_space, mjk_scf** out)
{
    if (!ctx || !basis || !s1e || !h1e || !out) return ctx_fail(ctx, MJK_EINVAL, "bad arguments to mjk_rhf_create");
    *out = nullptr;
    std::lock_guard<std::recursive_mutex> lock(mjk_internal::ctx_mutex(ctx));
    if (hipSetDevice(mjk_internal::ctx_device(ctx)) != hipSuccess)
        return ctx_fail(ctx, MJK_EDEVICE, "hipSetDevice failed");
    return create(ctx, basis, s1e, h1e, nocc, enuc, diis_space, out);
}

int mjk_rhf_run(mjk_ctx* ctx, mjk_scf* scf, const double* dm0, int max_cycle, double conv_tol, double conv_tol_grad,
                double direct_scf_tol, double fp32_tol, double level_shift, mjk_rhf_result* res)
{
    if (!ctx || !scf || scf->ctx != ctx || max_cycle < 0 || !(conv_tol > 0) || !(conv_tol_grad > 0))
        return ctx_fail(ctx, MJK_EINVAL, "bad arguments to mjk_rhf_run");
    std::lock_guard<std::recursive_mutex> lock(mjk_internal::ctx_mutex(ctx));
    if (hipSetDevice(mjk_internal::ctx_device(ctx)) != hipSuccess)
        return ctx_fail(ctx, MJK_EDEVICE, "hipSetDevice failed");
    int cycles = 0;
    const int rc = run(scf, dm0, max_cycle, conv_tol, conv_tol_grad, direct_scf_tol, fp32_tol, level_shift, &cycles);
    if (res) {
        res->e_tot = scf->e_tot;
        res->gnorm = scf->gnorm;
        res->cycles = scf->cycles;
        res->converged = scf->converged;
    }
    return rc;
}

int mjk_rhf_fetch(mjk_ctx* ctx, mjk_scf* scf, double* dm, double* mo_energy, double* mo_coeff, double* fock)
{
    if (!ctx || !scf || scf->ctx != ctx) return ctx_fail(ctx, MJK_EINVAL, "bad arguments to mjk_rhf_fetch");
    std::lock_guard<std::recursive_mutex> lock(mjk_internal::ctx_mutex(ctx));
    hipStream_t st = mjk_internal::ctx_stream(ctx);
    const int n = scf->n;
    const size_t nn = (size_t)n * n;
    if (dm) HCHK(scf, hipMemcpyAsync(dm, scf->D, nn * sizeof(double), hipMemcpyDeviceToHost, st));
    if (mo_energy) HCHK(scf, hipMemcpyAsync(mo_energy, scf->w, n * sizeof(double), hipMemcpyDeviceToHost, st));
    if (fock) HCHK(scf, hipMemcpyAsync(fock, scf->F, nn * sizeof(double), hipMemcpyDeviceToHost, st));
    std::vector<double> c;
    if (mo_coeff) {
        c.resize(nn);
        HCHK(scf, hipMemcpyAsync(c.data(), scf->C, nn * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    HCHK(scf, hipStreamSynchronize(st));
    if (mo_coeff)   // column-major C to row-major (nao, nmo)
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) mo_coeff[(size_t)i * n + j] = c[(size_t)j * n + i];
    return MJK_OK;
}

int mjk_rhf_set_df(mjk_ctx* ctx, mjk_scf* scf, mjk_df* df)
{
    if (!ctx || !scf || scf->ctx != ctx || (df && (mjk_internal::df_ctx(df) != ctx || mjk_internal::df_nao(df) != scf->n)))
        return ctx_fail(ctx, MJK_EINVAL, "bad arguments to mjk_rhf_set_df (the fit must be of this context and basis)");
    std::lock_guard<std::recursive_mutex> lock(mjk_internal::ctx_mutex(ctx));
    if (df && !scf->Xo) {
        const int rc = alloc(scf, &scf->Xo, (size_t)scf->n * std::max(scf->nocc, 1));
        if (rc) return rc;
    }
    scf->df = df;
    return MJK_OK;
}

const double* mjk_rhf_device_density(const mjk_scf* scf)
{
    return scf ? scf->D : nullptr;
}

void mjk_rhf_destroy(mjk_scf* scf)
{
    if (!scf) return;
    std::lock_guard<std::recursive_mutex> lock(mjk_internal::ctx_mutex(scf->ctx));
    release(scf);
}

}  // extern "C"
