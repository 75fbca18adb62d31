// Cartesian -> real solid harmonic coefficients, C2S[l][m][c].
//
// Cartesian order inside a shell: lx = l..0, ly = l-lx..0 (d: xx xy xz yy yz
// zz).  Spherical order m = -l..l (d: xy, yz, 3z^2-r^2, xz, x^2-y^2), the
// functions being r^l Y_lm with Y_lm normalised on the unit sphere
// (closed forms in pyscf/symm/test/test_sph.py:43-100).  For l = 0, 1 the
// factors 1/(2 sqrt(pi)) and sqrt(3/(4 pi)) (pyscf/gto/mole.py:174-178) are
// folded into the contraction coefficients and p is ordered x, y, z, so the
// l <= 1 blocks are identities.
#pragma once

namespace mjk {

constexpr double C2S_S0 = 0.28209479177387814;   // 1/(2 sqrt(pi))
constexpr double C2S_P1 = 0.4886025119029199;    // sqrt(3/(4 pi))
constexpr double D_A = 1.0925484305920792;       // sqrt(15/pi)/2
constexpr double D_B = 0.31539156525252005;      // sqrt(5/pi)/4
constexpr double D_C = 0.5462742152960396;       // sqrt(15/pi)/4
constexpr double F_3 = 0.5900435899266435;       // sqrt(70/pi)/8
constexpr double F_2 = 2.890611442640554;        // sqrt(105/pi)/2
constexpr double F_1 = 0.4570457994644658;       // sqrt(42/pi)/8
constexpr double F_0 = 0.3731763325901154;       // sqrt(7/pi)/4
constexpr double F_G = 1.445305721320277;        // sqrt(105/pi)/4

constexpr int C2S_MAXL = 3;

constexpr double C2S[4][7][10] = {
    {{1.0}},
    {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}},
    {
        {0.0, D_A, 0.0, 0.0, 0.0, 0.0},
        {0.0, 0.0, 0.0, 0.0, D_A, 0.0},
        {-D_B, 0.0, 0.0, -D_B, 0.0, 2.0 * D_B},
        {0.0, 0.0, D_A, 0.0, 0.0, 0.0},
        {D_C, 0.0, 0.0, -D_C, 0.0, 0.0},
    },
    {
        {0.0, 3.0 * F_3, 0.0, 0.0, 0.0, 0.0, -F_3, 0.0, 0.0, 0.0},
        {0.0, 0.0, 0.0, 0.0, F_2, 0.0, 0.0, 0.0, 0.0, 0.0},
        {0.0, -F_1, 0.0, 0.0, 0.0, 0.0, -F_1, 0.0, 4.0 * F_1, 0.0},
        {0.0, 0.0, -3.0 * F_0, 0.0, 0.0, 0.0, 0.0, -3.0 * F_0, 0.0, 2.0 * F_0},
        {-F_1, 0.0, 0.0, -F_1, 0.0, 4.0 * F_1, 0.0, 0.0, 0.0, 0.0},
        {0.0, 0.0, F_G, 0.0, 0.0, 0.0, 0.0, -F_G, 0.0, 0.0},
        {F_3, 0.0, 0.0, -3.0 * F_3, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    },
};

}  // namespace mjk
