// Rys roots and weights from the interpolation table of gen_rys_table.py.
// Usable on host (unit tests through the C-ABI) and device (kernels).
//
// For n roots and T < 36 + 6n: interval iv = floor(T/2), local variable
// s = T - 2 iv - 1 in [-1, 1], each root x_i (= t_i^2) and weight w_i is a
// 14-term Chebyshev series; the buffer the kernels read holds it in the power
// basis of s (converted once on the host, rys_fill_buffer), evaluated by
// Horner's rule: 13 FMAs against Clenshaw's 26 operations.  The conversion is
// well conditioned here: over every series of the table, sum_j |a_j| is at
// most 6.4 times the smallest |f(s)| on [-1, 1].  For larger T the half-range
// Gauss-Hermite limit x_i = s_i^2/T, w_i = W_i/sqrt(T).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include <type_traits>

#define RYS_STORAGE static constexpr
#include "rys_table.inc"
#undef RYS_STORAGE

namespace mjk {

// 1/sqrt(x) for normal positive x: on the device the hardware estimate
// (v_rsq_f64) refined by two Newton-Raphson steps y <- y (3 - x y^2) / 2, each
// of which squares the relative error (written as y + y (1/2 - (x y / 2) y));
// on the host the library square root.
__host__ __device__ __forceinline__ double rsqrt_nr(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double y = __builtin_amdgcn_rsq(x);
    double h = 0.5 * x * y;
    y = fma(y, fma(-h, y, 0.5), y);
    h = 0.5 * x * y;
    y = fma(y, fma(-h, y, 0.5), y);
    return y;
#else
    return 1.0 / sqrt(x);
#endif
}

// Offset of the Hermite constants behind the Chebyshev coefficients in the
// single device buffer holding both.
constexpr int RYS_HERM_BASE = RYS_TABLE_SIZE;
constexpr int RYS_BUF_SIZE = RYS_TABLE_SIZE + RYS_HOFF[RYS_NMAX + 1];

// sum_j a_j s^j, j < RYS_NCOEF (Horner)
__host__ __device__ __forceinline__ double poly14(const double* __restrict__ a, double s)
{
    double v = a[RYS_NCOEF - 1];
#pragma unroll
    for (int k = RYS_NCOEF - 2; k >= 0; --k) v = fma(v, s, a[k]);
    return v;
}

// tab: RYS_COEF followed by RYS_HERM (RYS_BUF_SIZE doubles).
template <int N>
__host__ __device__ __forceinline__ void rys_roots(const double* __restrict__ tab, double T,
                                                   double (&x)[N], double (&w)[N])
{
    if (!(T < 36.0 + 6.0 * N)) {   // also catches NaN: constant indices only
        const double it = 1.0 / T;
        const double st = sqrt(it);
        const double* h = tab + RYS_HERM_BASE + RYS_HOFF[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            x[i] = h[2 * i] * it;
            w[i] = h[2 * i + 1] * st;
        }
    } else {
        int iv = (int)(T * 0.5);
        iv = iv < 0 ? 0 : (iv > 17 + 3 * N ? 17 + 3 * N : iv);
        const double s = T - 2.0 * iv - 1.0;
        const double* c = tab + RYS_OFF[N] + iv * (N * 2 * RYS_NCOEF);
#pragma unroll
        for (int i = 0; i < N; ++i) {
            x[i] = poly14(c + (2 * i) * RYS_NCOEF, s);
            w[i] = poly14(c + (2 * i + 1) * RYS_NCOEF, s);
        }
    }
}

// Root-by-root evaluation: prepare once per T, then root(i) gives (x_i, w_i).
// Lets kernels keep a rolled root loop (one root's working set live at a time).
template <int N>
struct RysCursor {
    const double* c;   // interval coefficients, or Hermite constants
    double s;          // local variable (series) or 1/T (asymptotic)
    double st;         // 1/sqrt(T) (asymptotic)
    bool asym;
    __host__ __device__ __forceinline__ RysCursor(const double* __restrict__ tab, double T)
    {
        asym = !(T < 36.0 + 6.0 * N);   // also catches NaN: constant indices only
        if (asym) {
            st = rsqrt_nr(T);
            s = st * st;
            c = tab + RYS_HERM_BASE + RYS_HOFF[N];
        } else {
            int iv = (int)(T * 0.5);
            iv = iv < 0 ? 0 : (iv > 17 + 3 * N ? 17 + 3 * N : iv);
            s = T - 2.0 * iv - 1.0;
            st = 0.0;
            c = tab + RYS_OFF[N] + iv * (N * 2 * RYS_NCOEF);
        }
    }
    __host__ __device__ __forceinline__ void root(int i, double& x, double& w) const
    {
        if (asym) {
            x = c[2 * i] * s;
            w = c[2 * i + 1] * st;
        } else {
            x = poly14(c + (2 * i) * RYS_NCOEF, s);
            w = poly14(c + (2 * i + 1) * RYS_NCOEF, s);
        }
    }
};

// The n-root section of the table copied into LDS by a block (kernels whose
// roots come from one section): every interval's 2n series followed by the
// Hermite constants.  The interval stride is padded from 28n doubles (a
// multiple of 4) to 28n + 2: intervals then start 4 LDS banks apart modulo
// 32, so lanes of one wave reading the same coefficient of different
// intervals spread over 8 bank groups of their 16-byte reads.
template <int N>
struct RysLds {
    static constexpr int NINT = 18 + 3 * N;          // intervals of T < 36 + 6n
    static constexpr int SER = N * 2 * RYS_NCOEF;    // doubles of one interval
    static constexpr int STRIDE = SER + 2;
    static constexpr int SIZE = NINT * STRIDE + 2 * N;
};

// Copy of the n-root section into lt (RysLds<N>::SIZE doubles) by nthr
// threads; the caller synchronises the block before the first cursor reads it.
template <int N>
__host__ __device__ __forceinline__ void rys_lds_fill(double* lt, const double* __restrict__ tab, int tid, int nthr)
{
    using L = RysLds<N>;
    const double* src = tab + RYS_OFF[N];
    for (int e = tid; e < L::NINT * L::SER; e += nthr) {
        const int iv = e / L::SER;
        lt[iv * L::STRIDE + (e - iv * L::SER)] = src[e];
    }
    for (int e = tid; e < 2 * N; e += nthr) lt[L::NINT * L::STRIDE + e] = tab[RYS_HERM_BASE + RYS_HOFF[N] + e];
}

// RysCursor over the LDS copy (or a host buffer of the same layout).
template <int N>
struct RysCursorLds {
    const double* c;
    double s, st;
    bool asym;
    __host__ __device__ __forceinline__ RysCursorLds(const double* lt, double T)
    {
        using L = RysLds<N>;
        asym = !(T < 36.0 + 6.0 * N);
        if (asym) {
            st = rsqrt_nr(T);
            s = st * st;
            c = lt + L::NINT * L::STRIDE;
        } else {
            int iv = (int)(T * 0.5);
            iv = iv < 0 ? 0 : (iv > L::NINT - 1 ? L::NINT - 1 : iv);
            s = T - 2.0 * iv - 1.0;
            st = 0.0;
            c = lt + iv * L::STRIDE;
        }
    }
    __host__ __device__ __forceinline__ void root(int i, double& x, double& w) const
    {
        if (asym) {
            x = c[2 * i] * s;
            w = c[2 * i + 1] * st;
        } else {
            x = poly14(c + (2 * i) * RYS_NCOEF, s);
            w = poly14(c + (2 * i + 1) * RYS_NCOEF, s);
        }
    }
};

// Range-separated operators, selected by omega (env[PTR_RANGE_OMEGA] of the
// basis): 0 the Coulomb 1/r, > 0 erf(omega r)/r, < 0 erfc(|omega| r)/r.  The
// erf kernel restricts the Rys variable to t^2 <= theta = omega^2 / (omega^2 +
// rho), rho = pq / (p + q); substituting t = sqrt(theta) s gives the
// full-range roots at theta T, scaled by theta, with weights scaled by
// sqrt(theta).  erfc = 1 - erf takes two root sets: the full-range set (pass
// 0) and the erf set with negated weights (pass 1).
__host__ __device__ __forceinline__ int rsh_passes(double omega) { return omega < 0.0 ? 2 : 1; }

// T of the root set of this pass, and the factors for its roots and weights.
__host__ __device__ __forceinline__ void rsh_scale(double omega, double rho, int pass, double& T, double& sx,
                                                   double& sw)
{
    sx = 1.0;
    sw = 1.0;
    if (omega == 0.0 || (omega < 0.0 && pass == 0)) return;
    const double w2 = omega * omega;
    const double th = w2 / (w2 + rho);
    T *= th;
    sx = th;
    sw = omega > 0.0 ? sqrt(th) : -sqrt(th);
}

// The weight of the single root, F_0(T) = sqrt(pi/T) erf(sqrt T) / 2, in
// closed form (below T = 0.2 its Taylor series sum_k (-T)^k / (k! (2k + 1)),
// 13 terms, the next below 1e-19), for (ss|ss), whose integrals never use the
// root itself.  It replaces 14 table loads and the series by erf: 11% faster
// (ss|ss) kernels on MI355X.  (With the root needed, F_1 / F_0 in closed form
// cost (ps|ss) 13%: there the table stays.)
struct RysWeightOnly {
    double w0;
    __host__ __device__ __forceinline__ RysWeightOnly(const double* __restrict__, double T)
    {
        if (T < 0.2) {
            double term = 1.0;   // (-T)^k / k!
            w0 = 0.0;
#pragma unroll
            for (int k = 0; k < 13; ++k) {
                w0 = fma(term, 1.0 / (2 * k + 1), w0);
                term *= -T * (1.0 / (k + 1));
            }
        } else {
            const double st = sqrt(T);
            w0 = 0.886226925452758014 * erf(st) / st;   // sqrt(pi) / 2
        }
    }
    __host__ __device__ __forceinline__ void root(int, double& x, double& w) const
    {
        x = 0.0;
        w = w0;
    }
};

// Root evaluator of a class: weight only when L = la + lb + lc + ld = 0;
// LDS: the cursor reads the block's LDS copy of the section (lds doubles of
// it, 0 when the class needs no table).
template <int NR, int L, bool LDS = false>
struct RysFor {
    static constexpr bool weight_only = NR == 1 && L == 0;
    using type = typename std::conditional<weight_only, RysWeightOnly,
                                           typename std::conditional<LDS, RysCursorLds<NR>, RysCursor<NR>>::type>::type;
    static constexpr int lds = (LDS && !weight_only) ? RysLds<NR>::SIZE : 0;
};

// Host-side helper: the table as one flat array (series then Hermite
// constants), every 14-term Chebyshev series sum_k c_k T_k(s) rewritten as
// sum_j a_j s^j.  The integer coefficients of T_k (T_0 = 1, T_1 = s,
// T_k+1 = 2 s T_k - T_k-1) are exact in long double, and each a_j is summed in
// long double and rounded once.
inline void rys_fill_buffer(double* buf)
{
    static_assert(RYS_TABLE_SIZE % RYS_NCOEF == 0, "whole series");
    long double t[RYS_NCOEF][RYS_NCOEF] = {};   // t[k][j]: s^j coefficient of T_k
    t[0][0] = 1;
    t[1][1] = 1;
    for (int k = 1; k + 1 < RYS_NCOEF; ++k)
        for (int j = 0; j < RYS_NCOEF; ++j)
            t[k + 1][j] = (j > 0 ? 2 * t[k][j - 1] : 0) - t[k - 1][j];
    for (int b = 0; b < RYS_TABLE_SIZE; b += RYS_NCOEF)
        for (int j = 0; j < RYS_NCOEF; ++j) {
            long double a = 0;
            for (int k = j; k < RYS_NCOEF; ++k) a += (long double)RYS_COEF[b + k] * t[k][j];
            buf[b + j] = (double)a;
        }
    for (int i = 0; i < RYS_HOFF[RYS_NMAX + 1]; ++i) buf[RYS_HERM_BASE + i] = RYS_HERM[i];
}

}  // namespace mjk
