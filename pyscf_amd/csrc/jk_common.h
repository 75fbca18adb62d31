// Data layout shared by the host driver and the HIP kernels.
//
// Shell pairs live in HBM as structure-of-arrays, grouped by pair class
// (la >= lb) and pair group, and, inside a group, sorted by descending
// Schwarz bound.  Every primitive pair (exponent sum p, centre P, folded
// prefactor K, 1/p) is stored once; a quartet task only reads two pair
// records.
//
// Generally contracted s shells with two contractions sharing their
// primitives (cc-pVXZ cores) appear twice: as two segmented shells in
// "split" pairs, and as one two-contraction unit in "merged" pairs, whose
// primitive pairs carry the products of the contraction coefficients of
// every contraction combination (ppW) so that one primitive loop serves all
// combinations.  A quartet class runs either its split or its merged pair
// groups, never both.
//
// Densities and outputs are kept in a Cartesian AO arena: for molecule m of
// a batch, the block [dbase[m], dbase[m] + ld[m]^2) of every slot holds its
// (ncart x ncart) matrix.  Spherical <-> Cartesian transforms run before and
// after the quartet kernels, so the kernels never apply the cart->sph
// transform per quartet.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mjk {

constexpr int LMAX = 3;                       // s, p, d, f
constexpr int NPAIRCLASS = (LMAX + 1) * (LMAX + 2) / 2;
constexpr int TILE = 8;                       // quartet tile: 8 bra x 8 ket pairs
constexpr int ROW_MAXBUCKET = 32;             // primitive-count buckets a row-kernel block can hold
constexpr int ROW_TCTL = 8;                   // ints per class of the persistent launches: count, 7 slice heads

__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
__host__ __device__ constexpr int nsph(int l) { return 2 * l + 1; }

// pair class index for la >= lb: (0,0)=0 (1,0)=1 (1,1)=2 (2,0)=3 ...
__host__ __device__ constexpr int pair_class(int la, int lb) { return la * (la + 1) / 2 + lb; }

struct PairSet {
    const double4* ppP;    // per primitive pair {Px, Py, Pz, p}
    const double2* ppK;    // per primitive pair {K, 1/p}; K = ca cb exp(-ab/p AB^2) sqrt(2) pi^(5/4) / p
                           // (merged pairs: without ca cb, which are in ppW)
    const double4* ppW;    // merged pairs: ca[ci] cb[cj] of contraction combination ci * CB + cj
    const int* prim0;      // first primitive pair of each shell pair
    const int* nprim;      // primitive pairs of each shell pair
    const double4* rA;     // {Ax, Ay, Az, Schwarz q}
    const double4* rAB;    // {ABx, ABy, ABz, 0}  (A - B)
    const int* shA;        // segmented shell index of A within its molecule (first contraction)
    const int* shB;
    const int* oshA;       // shell index of A in the molecule's input basis (dm_cond rows)
    const int* oshB;
    const int* cartA;      // Cartesian AO offset of A within its molecule
    const int* cartB;
    // Ket-side copy of the primitive pairs, transposed per (class, molecule,
    // bucket) segment: primitive ik of pair q sits at tP[tpos[q] + ik * tstride[q]]
    // (tstride = the segment's pair count), so lanes holding consecutive kets
    // read consecutive addresses.
    const double4* tP;
    const double2* tK;
    const double4* tW;     // transposed ppW
    const int* tpos;
    const int* tstride;
};

struct MolSet {
    const long long* dbase;   // arena offset of the molecule's Cartesian block
    const int* ld;            // ncart of the molecule
    const long long* cbase;   // offset of the molecule's dm_cond block
    const int* nsh;           // shells of the molecule's input basis (dm_cond dimension)
};

// One launch = one quartet class over a list of bra tiles.  A bra tile is
// TILE consecutive bra pairs; each owns bt_off[i+1]-bt_off[i] ket tiles of
// TILE pairs starting at q0 (Schwarz-pruned ket prefix [q0, qend)).
struct JKArgs {
    PairSet P;
    MolSet M;
    const int4* bt;         // {p0, p1, q0, qend}
    const int* bt_mol;
    const int* bt_off;      // prefix of ket tiles, nbt + 1 entries
    int nbt;
    int diag;               // bra and ket from the same pair class: keep q <= p
    const double* Dj;       // Cartesian J densities (D + D^T), n_j slots
    const double* Dk;       // Cartesian K densities, n_k slots
    double* J;              // accumulators J' (n_j slots); J = J' + J'^T afterwards
    double* K;              // accumulators K' (n_k slots)
    long long slot;         // slot stride (elements)
    int n_j;
    int n_k;
    int ncopy;              // private accumulator copies (one per XCD); the kernel picks XCC_ID % ncopy
    long long jcopy;        // element stride between copies of J'
    long long kcopy;        // element stride between copies of K'
    const double* dm_cond;  // per molecule (nsh x nsh) over the input basis's shells
    double cutoff;
    int screen;
    double fp32_tol;        // > 0: waves whose quartets all have q_ij q_kl dmax < fp32_tol run in FP32
    const double* rys;      // Rys table buffer
    double omega;           // operator: 0 Coulomb, > 0 erf(omega r)/r, < 0 erfc(|omega| r)/r
    unsigned long long* counters;  // {FP64 quartets, FP64 primitive quartets, FP32 quartets, FP32 primitive quartets}
    // row kernel: block b takes bra pair rt[b].x and the kets rt[b].y .. min(rt[b].z, n) - 1 of
    // the n surviving kets of the ket class, i.e. of the concatenated ket ranges
    // rr[rt[b].w * nbucket + k] = {first, end} (one per primitive bucket k), for the bra's
    // primitive pairs rtp[b] .. rtp[b] + rprim - 1 (the integrals are linear in the primitive
    // sum, so a bra pair's primitives may be shared among several tasks)
    const int4* rt;
    const int2* rr;
    const int* rtp;
    int rprim;
    const int* pmol;        // molecule of every pair
    int nbucket;
    const int* mol_active;  // per molecule: 0 skips its quartets (batched SCF; null: all active)
    // persistent row launch (thead != null): the blocks of slice kernel S claim
    // task entries ti = atomicAdd(thead + S, 1) until ti >= the task count
    // (*tcount, or ntask when tcount is null); entry ti is row task tlist[ti]
    // (ti when tlist is null).  Otherwise block b runs row task b.
    const int* tlist;
    const int* tcount;
    int* thead;
    int ntask;
};

struct QCondArgs {
    PairSet P;
    int p0, np;             // pair range of one class
    double* qout;           // Schwarz bound per pair
    const double* rys;
    double omega;           // operator, as in JKArgs
};

struct XformArgs {
    // per Cartesian AO of the whole batch
    const int* cart_mol;    // molecule
    const int* cart_s0;     // first spherical AO of its shell (molecule local)
    const int* cart_l;      // l of its shell
    const int* cart_c;      // Cartesian component index in its shell
    // per spherical AO of the whole batch
    const int* sph_mol;
    const int* sph_c0;      // first Cartesian AO of its shell (molecule local)
    const int* sph_l;
    const int* sph_m;       // m index (0..2l) in its shell
    // per molecule
    const long long* dbase; // Cartesian arena offsets
    const int* ld;          // ncart
    const long long* sbase; // spherical arena offsets
    const int* nao;         // spherical AOs
    const long long* cart_first;  // first global Cartesian AO of the molecule
    const long long* sph_first;   // first global spherical AO of the molecule
    int nmol;
    long long ncart_tot;
    long long nsph_tot;
};

struct DmCondArgs {
    const int* sh_mol;      // per shell of the batch: molecule
    const int* sh_s0;       // first spherical AO (molecule local)
    const int* sh_n;        // spherical AOs of the shell
    const long long* sh_first;  // per molecule: first global shell
    const long long* cbase; // per molecule: dm_cond block offset
    const int* nsh;         // per molecule: shells
    const long long* sbase; // per molecule: spherical arena offset
    const int* nao;
    long long nsh_tot;
    int maxnsh;
};

}  // namespace mjk
