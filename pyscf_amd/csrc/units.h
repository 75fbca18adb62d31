// Basis shells unrolled into contraction units for the density-fitting and
// one-electron kernels: one unit per (shell, contraction), primitives with
// their coefficients, the s and p angular factors folded in (the l <= 1
// spherical transforms are then identities, as in the four-centre path).
#pragma once
#include <vector>

#include "c2s.h"
#include "mjk.h"

namespace mjk {

struct Unit {
    double r[3];
    int l, nprim, p0, sph0;
};

inline int fold_units(const int* atm, int natm, const int* bas, int nbas, const double* env, int lmax,
               std::vector<Unit>& units, std::vector<double>& ex, std::vector<double>& co, int& nsph)
{
    nsph = 0;
    for (int s = 0; s < nbas; ++s) {
        const int* b = bas + 8 * s;
        const int ia = b[0], l = b[1], np = b[2], nc = b[3];
        if (ia < 0 || ia >= natm || l < 0 || l > lmax || np <= 0 || nc <= 0) return MJK_EUNSUPPORTED;
        const double* c = env + atm[6 * ia + 1];
        const double f = l == 0 ? C2S_S0 : l == 1 ? C2S_P1 : 1.0;
        for (int k = 0; k < nc; ++k) {
            Unit u;
            u.r[0] = c[0];
            u.r[1] = c[1];
            u.r[2] = c[2];
            u.l = l;
            u.nprim = np;
            u.p0 = (int)ex.size();
            u.sph0 = nsph;
            for (int p = 0; p < np; ++p) {
                ex.push_back(env[b[5] + p]);
                co.push_back(env[b[6] + k * np + p] * f);
            }
            units.push_back(u);
            nsph += 2 * l + 1;
        }
    }
    return MJK_OK;
}

}  // namespace mjk
