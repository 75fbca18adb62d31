// Host driver of the direct J/K engine: C ABI declared in include/mjk.h.
//
// Basis upload (once per molecule or batch):
//   * every contracted shell with nctr > 1 is split into nctr segmented
//     shells sharing the primitives, keeping PySCF's AO order (ctr, m);
//   * shell pairs (i >= j) are built with the higher l first, primitive pairs
//     with exp(-ab/(a+b)|AB|^2) below exp(-expcutoff) are dropped, and the
//     prefactor sqrt(2) pi^(5/4) ca cb exp(-ab/p AB^2)/p is folded per pair;
//   * Schwarz bounds are computed on the GPU (k_qcond), pairs are sorted by
//     class and descending bound, and per quartet class a list of bra tiles
//     with their Schwarz-surviving ket prefix is built.
// J/K build: dm_cond -> sph->cart densities -> one kernel per quartet class
// -> cart->sph J and K with the reference's symmetrisation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <array>
#include <atomic>
#include <fstream>
#include <sstream>
#include <dlfcn.h>
#include <map>
#include <tuple>
#include <mutex>
#include <shared_mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "mjk.h"
#include "mjk_internal.h"
#include "c2s.h"
#include "jk_common.h"
#include "rys_eval.h"

namespace mjk {
int launch_jk(int la, int lb, int lc, int ld, int nblocks, const JKArgs& a, hipStream_t st);
int launch_jk_rows(int la, int lb, int lc, int ld, int nblocks, const JKArgs& a, size_t shmem, hipStream_t st);
int row_kernel_available(int la, int lb, int lc, int ld);
int launch_jk_rows_gc(int la, int lb, int lc, int ld, int ca, int cb, int cc, int cd, int nblocks, const JKArgs& a,
                      size_t shmem, hipStream_t st);
int row_gc_available(int la, int lb, int lc, int ld, int ca, int cb, int cc, int cd);
int row_m_available(int la, int lb, int lc, int ld);
int launch_jk_rows_m(int la, int lb, int lc, int ld, int nblocks, const JKArgs& a, size_t shmem, hipStream_t st);
int coop_kernel(int la, int lb, int lc, int ld);
int launch_qcond(int la, int lb, const QCondArgs& a, hipStream_t st);
int upload_c2s(hipStream_t st);
int launch_sph2cart(const XformArgs& x, int maxld, const double* Dsph, long long sslot, int n_dm,
                    int hermi, int with_j, int with_k, double* Dj, double* Dk, long long cslot,
                    hipStream_t st);
int launch_cart2sph(const XformArgs& x, int maxnao, const double* Jc, const double* Kc, long long cslot,
                    int n_dm, int hermi, int with_j, int with_k, double* vj, double* vk, long long sslot,
                    int ncopy, long long jcopy, long long kcopy, hipStream_t st);
int launch_compact_rows(const int4* rt, int nrt, const int* pmol, const int* active, int* list, int* count,
                        hipStream_t st);
int launch_dmcond(const DmCondArgs& x, const double* Dsph, long long sslot, int n_dm, double* out,
                  hipStream_t st);
int lmax_kernels();
int launch_copy(const double* src, double* dst, long long n, hipStream_t st);
int launch_zero(double* dst, long long n, hipStream_t st);
}  // namespace mjk

using namespace mjk;

namespace {

// Primitive pairs with exp(-ab/p AB^2) below e^-40 = 4e-18 are dropped (libcint
// drops them below e^-60).  Measured on (H2O)16 def2-SVP: 7% fewer primitive
// quartets, 111 -> 115 builds/s, J and K unchanged beyond 5e-14 (the
// summation-order noise of two builds); e^-30 moved J by 8e-10.  The largest
// cutoff an override may set (MJK_EXPCUTOFF) is libcint's 60.
constexpr double EXPCUTOFF_DEFAULT = 40.0;
// LDS of a gfx950 workgroup, and the most a row-kernel block's dynamic rows
// may take (less 1 KB for its task bookkeeping); a class's rows must also
// leave room for its static copy of the Rys table (row_static_lds).
constexpr int LDS_WORKGROUP = 160 * 1024;
constexpr int ROW_LDS_MAX = LDS_WORKGROUP - 1024;
constexpr double EXPCUTOFF_MAX = 60.0;
constexpr int ATM_SLOTS = 6, BAS_SLOTS = 8, PTR_COORD = 1, ATOM_OF = 0, ANG_OF = 1, NPRIM_OF = 2,
              NCTR_OF = 3, PTR_EXP = 5, PTR_COEFF = 6, PTR_RANGE_OMEGA = 8;
// Quartet classes of a basis: one per pair of pair classes and pair groups
// (at most 5 x 5 group combinations for a pair of classes with every group).
constexpr int NQCLASS_MAX = 1024;

// Pair groups within a pair class.  Shells of the input basis with one
// contraction, and generally contracted shells that are not merged (l > 0 or
// more than two contractions), are split into segmented shells (one per
// contraction).  A generally contracted s shell with two contractions is a
// "merged unit"; its pairs exist twice: split into segmented pairs (group
// SPLIT) and as unit pairs carrying both contractions (groups M21, M12, M22:
// the contraction counts of A and B).  Pairs of segmented shells not from a
// merged unit form group PLAIN, shared by both representations.
enum PairGroup { G_PLAIN = 0, G_SPLIT = 1, G_M21 = 2, G_M12 = 3, G_M22 = 4, NGT = 5 };
constexpr int NGROUP = NPAIRCLASS * NGT;
__host__ __device__ constexpr int group_ca(int t) { return t == G_M21 || t == G_M22 ? 2 : 1; }
__host__ __device__ constexpr int group_cb(int t) { return t == G_M12 || t == G_M22 ? 2 : 1; }

// Shell pairs of a class are grouped by their number of primitive pairs so
// that every tile runs a uniform primitive loop (no lane divergence).
constexpr int NBUCKET = 16;
static_assert(NBUCKET <= ROW_MAXBUCKET, "row kernel holds at most ROW_MAXBUCKET ket buckets");
int prim_bucket(int n)
{
    static const int edge[NBUCKET - 1] = {1, 2, 3, 4, 6, 9, 12, 16, 20, 25, 32, 40, 50, 64, 100};
    for (int i = 0; i < NBUCKET - 1; ++i)
        if (n <= edge[i]) return i;
    return NBUCKET - 1;
}

struct GShell {
    int l, nprim, orig;
    double r[3];
    std::vector<double> e, c;
    int s0, c0;   // molecule-local spherical / Cartesian offsets
    int unit;     // merged unit (or lone segmented shell) this shell belongs to
    int ci;       // contraction index within its shell of the input basis
    bool merged;  // part of a merged two-contraction s shell
};

// A unit is a merged two-contraction s shell or a lone segmented shell.
struct UnitHost {
    int first;    // first segmented shell
    int nctr;     // 2 for a merged unit, 1 otherwise
};

struct MolHost {
    int nao = 0, ncart = 0, nsh = 0, nbas_orig = 0;
    std::vector<GShell> sh;
    std::vector<UnitHost> units;
    std::vector<int> orig_s0, orig_n;   // per input shell: first spherical AO, spherical AOs (all contractions)
    long long sbase = 0, dbase = 0, cbase = 0;
    long long sph_first = 0, cart_first = 0, sh_first = 0;
};

struct QClass {
    int X, Y, la, lb, lc, ld;
    int GX = 0, GY = 0;       // pair groups (X * NGT + type) of bra and ket
    int ca = 1, cb = 1, cc = 1, cd = 1;   // contraction counts of the groups
    int mode = 0;             // 0 both representations, 1 split only, 2 merged only
    int base = 0;             // index of the (X, Y) class this one belongs to
    int nbt = 0;
    long long nblocks = 0;
    double work = 0;          // rough cost (ket tiles x components x roots)
    double est = 0;           // algorithmic FLOPs of the Schwarz-prescreened quartets (an upper bound)
    int4* bt = nullptr;
    int* bt_mol = nullptr;
    int* bt_off = nullptr;
    // row-kernel tasks {bra pair, first ket, end of chunk or INT_MAX for the last chunk, range set}
    // and per-bucket ket ranges {first, end} of each bra pair (range set r = rr[r * NBUCKET ...])
    int nrt = 0;
    int4* rt = nullptr;
    int2* rr = nullptr;
    int* rtp = nullptr;      // bra primitive offset of every task
    int rprim = 0;           // bra primitive pairs per task
    int bucket = 0;          // floor(log2) of the Schwarz-prescreened primitive quartets (tuning key)
    int kernel_pref = 0;     // tuning: 0 the default choice, 1 row kernel, 2 fallback (tile / cooperative)
    long long tl0 = 0;       // offset of this class's segment of the compacted task lists
};


int lmax_built()
{
    return lmax_kernels();
}

// Algorithmic FLOPs of nq evaluated quartets with nprimq primitive quartets;
// ncomb > 1: a merged class, whose primitive quartets each feed ncomb
// contraction combinations (nq counts every combination).
double flops_class(int a, int b, int c, int d, double nq, double nprimq, int ncomb = 1)
{
    const int L = a + b + c + d;
    const double nr = L / 2 + 1;
    const double Nc = (double)ncart(a) * ncart(b) * ncart(c) * ncart(d);
    const double Ns = (double)nsph(a) * nsph(b) * nsph(c) * nsph(d);
    const double fprim = 30 + 20 * nr + 12 * nr * (a + b + 1) * (c + d + 1) + 3 * nr * Nc + 2 * ncomb * Nc;
    double fc2s = 0;
    if (a >= 2) fc2s += (double)nsph(a) * ncart(a) * ncart(b) * ncart(c) * ncart(d);
    if (b >= 2) fc2s += (double)nsph(a) * nsph(b) * ncart(b) * ncart(c) * ncart(d);
    if (c >= 2) fc2s += (double)nsph(a) * nsph(b) * nsph(c) * ncart(c) * ncart(d);
    if (d >= 2) fc2s += (double)nsph(a) * nsph(b) * nsph(c) * nsph(d) * ncart(d);
    fc2s *= 2;
    const double fdig = 12 * Ns;
    return nprimq * fprim + nq * (fc2s + fdig);
}

}  // namespace

struct mjk_ctx {
    int device = 0;
    int ncopy = 8;        // private J'/K' accumulator copies (MI355X has 8 XCDs)
    // LDS bytes a row-kernel block may take; 0 = fallback kernels only.  gfx950
    // gives a workgroup up to 160 KB; 96 KB keeps room for a second block per CU
    // and let (dd|ss) on C30H62 cc-pVDZ use rows: 73.0 -> 68.9 ms per build
    size_t row_lds = ROW_LDS_MAX;
    int row_chunk = 16384;    // least primitive quartets per row-kernel block
    int row_blocks = 16384;   // row-kernel blocks per class launch to aim for ...
    int row_min_blocks = 512;    // ... and the fewest, whatever row_chunk asks
    int row_prim = 0;            // bra primitive pairs per row task (0: by class size)
    // per-class tuning table (mjk_ctx_set_tuning): key (la lb lc ld ca cb cc cd bucket) ->
    // {kernel preference, bra primitive pairs per task, task-size scale}; tune_gen counts
    // its changes, and a basis rebuilds its task lists when it sees a new generation
    std::map<std::array<int, 9>, std::tuple<int, int, double>> tune;
    int tune_gen = 0;
    hipStream_t stream = nullptr;
    void* blas = nullptr;                 // dense-algebra handle of the device-resident SCF
    void (*blas_destroy)(void*) = nullptr;
    size_t budget = 0;
    std::atomic<size_t> used{0};  // tracked device bytes (read by request admission on other threads)
    std::string err;
    double* rys = nullptr;
    std::map<void*, size_t> allocs;
    std::recursive_mutex mu;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    // extra streams (default priority) on which the quartet classes of a build
    // run concurrently; they fork from and join back into `stream`
    static constexpr int NAUX = 3;
    int nstreams = 4;                     // 1 + aux streams used, MJK_STREAMS
    int sched = 1;                        // class cost for the stream assignment: 0 work, 1 est; MJK_SCHED
    int graphs = 1;                       // replay plain builds as HIP graphs (mjk_ctx_set_graphs, MJK_GRAPHS)
    int order = 1;                        // issue order within a stream: 0 by cost, 1 heavy classes first; MJK_ORDER
    // persistent row launches (blocks claim tasks from a device counter): 0 never,
    // 1 batched bases (the tasks of inactive molecules are compacted away on the
    // device first), 2 every basis; MJK_PERSIST
    int persist = 1;
    int persist_blocks = 2048;            // blocks of a persistent row launch; MJK_PERSIST_BLOCKS
    double expcutoff = EXPCUTOFF_DEFAULT; // primitive pairs with exp(-ab/p AB^2) < e^-expcutoff dropped; MJK_EXPCUTOFF
    hipStream_t aux[NAUX] = {nullptr, nullptr, nullptr};
    hipEvent_t ev_fork = nullptr;
    hipEvent_t ev_join[NAUX] = {nullptr, nullptr, nullptr};
};

namespace {

int fail(mjk_ctx* ctx, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    return code;
}

#define HIPCHK(ctx, x)                                                                            \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) return fail(ctx, MJK_EDEVICE, "%s failed: %s", #x, hipGetErrorString(e_)); \
    } while (0)

int dalloc(mjk_ctx* ctx, size_t bytes, void** p)
{
    *p = nullptr;
    if (bytes == 0) return MJK_OK;
    if (ctx->used + bytes > ctx->budget)
        return fail(ctx, MJK_ENOMEM, "HBM budget exceeded: need %zu more bytes, %zu of %zu in use", bytes,
                    ctx->used.load(), ctx->budget);
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(ctx, MJK_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    ctx->used += bytes;
    ctx->allocs[*p] = bytes;
    return MJK_OK;
}

// hipFree synchronises the whole device, and a synchronisation from any thread
// while another thread captures a stream into a graph invalidates that
// capture.  Captures therefore hold g_capture shared, and a free runs only
// when it can take g_capture exclusively; otherwise the pointer waits in
// g_deferred and is released by the next free or build that finds no capture
// in flight.  (Budget accounting releases the bytes at once.)
std::shared_mutex g_capture;
std::mutex g_deferred_mu;
std::vector<std::pair<int, void*>> g_deferred;   // {device, pointer}

void free_now(int device, void* p)
{
    int cur = 0;
    (void)hipGetDevice(&cur);
    if (cur != device) (void)hipSetDevice(device);
    (void)hipFree(p);
    if (cur != device) (void)hipSetDevice(cur);
}

// Frees what earlier calls deferred; caller holds g_capture exclusively.
void drain_deferred_locked()
{
    std::vector<std::pair<int, void*>> todo;
    {
        std::lock_guard<std::mutex> lk(g_deferred_mu);
        todo.swap(g_deferred);
    }
    for (auto& d : todo) free_now(d.first, d.second);
}

void drain_deferred()
{
    std::unique_lock<std::shared_mutex> lk(g_capture, std::try_to_lock);
    if (lk.owns_lock()) drain_deferred_locked();
}

void dfree(mjk_ctx* ctx, void* p)
{
    if (!p) return;
    auto it = ctx->allocs.find(p);
    if (it != ctx->allocs.end()) {
        ctx->used -= it->second;
        ctx->allocs.erase(it);
    }
    std::unique_lock<std::shared_mutex> lk(g_capture, std::try_to_lock);
    if (!lk.owns_lock()) {
        std::lock_guard<std::mutex> dl(g_deferred_mu);
        g_deferred.push_back({ctx->device, p});
        return;
    }
    drain_deferred_locked();
    free_now(ctx->device, p);
}

// Work buffers of the no-allocation build (mjk_get_jk_device_noalloc): one
// array per device, resident from the library's load to the process's end.
// In host code the name `scratch` is a host-side stand-in and never a device
// address: each device's address is taken once with hipGetSymbolAddress while
// that device is current (mjk_ctx_create) and kept in g_scratch_addr, and only
// that address is passed to kernels.
static __device__ __align__(256) double scratch[MJK_NOALLOC_SCRATCH_BYTES / sizeof(double)];
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_addr[MAX_DEVICES] = {};
std::string g_scratch_err[MAX_DEVICES];

// Called with `device` current.
void scratch_init(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_addr[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        g_scratch_err[device] = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
        return;
    }
    g_scratch_addr[device] = static_cast<double*>(p);
}

double* scratch_address(int device, std::string* why)
{
    if (device < 0 || device >= MAX_DEVICES) {
        *why = "device index out of range";
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_addr[device])
        *why = g_scratch_err[device].empty() ? "scratch not initialised" : g_scratch_err[device];
    return g_scratch_addr[device];
}

template <typename T>
int upload(mjk_ctx* ctx, const std::vector<T>& v, T** d)
{
    int rc = dalloc(ctx, v.size() * sizeof(T), (void**)d);
    if (rc) return rc;
    if (!v.empty()) {
        // on the context's (non-blocking) stream: no null-stream synchronisation
        // that could meet another thread's graph capture
        HIPCHK(ctx, hipMemcpyAsync(*d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    }
    return MJK_OK;
}

}  // namespace

// Everything a captured build depends on: a replay is valid only for the same key.
struct GraphKey {
    const void* dms = nullptr;
    const void* vj = nullptr;
    const void* vk = nullptr;
    const void* bufs[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    int n_dm = 0, hermi = 0, with_j = 0, with_k = 0, nstreams = 0, class_filter = -1;
    double cut = 0, fp32_tol = 0;
    size_t row_lds = 0;
    bool operator==(const GraphKey& o) const
    {
        for (int i = 0; i < 5; ++i)
            if (bufs[i] != o.bufs[i]) return false;
        return dms == o.dms && vj == o.vj && vk == o.vk && n_dm == o.n_dm && hermi == o.hermi &&
               class_filter == o.class_filter &&
               with_j == o.with_j && with_k == o.with_k && nstreams == o.nstreams && cut == o.cut &&
               fp32_tol == o.fp32_tol && row_lds == o.row_lds;
    }
};

struct mjk_basis {
    mjk_ctx* ctx = nullptr;
    int nmol = 0;
    double omega = 0;   // two-electron operator, env[PTR_RANGE_OMEGA]: 0 1/r, > 0 erf, < 0 erfc
    // the last plain build (no statistics, no profiling) as an instantiated HIP graph
    hipGraphExec_t gexec = nullptr;
    GraphKey gkey;
    GraphKey last_key;          // key of the last plain build: a build is captured when its key repeats
    bool has_last_key = false;
    std::vector<MolHost> mols;
    int maxld = 0, maxnao = 0, maxnsh = 0;
    long long slot_s = 0, slot_c = 0, cond_total = 0;
    long long ncart_tot = 0, nsph_tot = 0, nsh_tot = 0;
    // pairs (device)
    int npairs = 0;
    double4* ppP = nullptr;
    double2* ppK = nullptr;
    double4* ppW = nullptr;          // merged pairs' coefficient products (PairSet::ppW)
    int *prim0 = nullptr, *nprim = nullptr, *shA = nullptr, *shB = nullptr, *cartA = nullptr, *cartB = nullptr;
    int *oshA = nullptr, *oshB = nullptr;
    double4* tP = nullptr;           // ket-side transposed primitive pairs (PairSet::tP)
    double2* tK = nullptr;
    double4* tW = nullptr;
    int *tpos = nullptr, *tstride = nullptr;
    double4 *rA = nullptr, *rAB = nullptr;
    std::vector<int> seg;            // [(group * nmol + mol) * NBUCKET + bucket] -> first pair, group = class * NGT + type
    int* d_pmol = nullptr;           // molecule of every pair
    int* d_active = nullptr;         // per molecule: 0 skips its quartets (mjk_basis_set_active)
    std::vector<int> h_active;       // host copy of d_active
    std::vector<long long> h_sbase;  // per molecule: spherical arena offset (block of nao^2)
    std::vector<int> h_nao;
    int* d_tlist = nullptr;          // compacted row-task lists, one segment per class (QClass::tl0)
    int* d_tctl = nullptr;           // per class ROW_TCTL ints: task count, one claim counter per slice kernel
    int class_filter = -1;           // >= 0: builds run only this quartet class (mjk_basis_set_class_filter)
    int tune_gen = -1;               // ctx->tune_gen the task lists were built with
    std::vector<double> q;           // Schwarz bound per pair (sorted order)
    std::vector<int> h_shA, h_shB, h_nprim;
    std::vector<long long> pcost;    // prefix sums of h_nprim (row-kernel task sizes)
    // per molecule / per AO device arrays
    long long *d_dbase = nullptr, *d_cbase = nullptr, *d_sbase = nullptr, *d_cart_first = nullptr,
              *d_sph_first = nullptr, *d_sh_first = nullptr;
    int *d_ld = nullptr, *d_nsh = nullptr, *d_nao = nullptr;
    int *d_cart_mol = nullptr, *d_cart_s0 = nullptr, *d_cart_l = nullptr, *d_cart_c = nullptr;
    int *d_sph_mol = nullptr, *d_sph_c0 = nullptr, *d_sph_l = nullptr, *d_sph_m = nullptr;
    int *d_sh_mol = nullptr, *d_sh_s0 = nullptr, *d_sh_n = nullptr;
    // tiles
    double tile_cutoff = -1.0;
    std::vector<QClass> qcls;
    // work buffers
    double *Dsph = nullptr, *Vj = nullptr, *Vk = nullptr;
    double *Dj = nullptr, *Dk = nullptr, *J = nullptr, *K = nullptr, *dmcond = nullptr;
    long long cap_ds = 0, cap_vj = 0, cap_vk = 0, cap_dj = 0, cap_j = 0, cap_dk = 0, cap_k = 0, cap_cond = 0;
    unsigned long long* counters = nullptr;
    std::vector<void*> owned;
    // per-class profiling (HIP events around every class launch)
    bool profiling = false;
    bool prof_counts = false;             // tallies (quartets, primitive quartets) while profiling
    std::vector<hipEvent_t> pev;          // 2 per quartet class
    std::vector<double> prof_ms;
    std::vector<long long> prof_nq, prof_np, prof_launches;
    std::vector<int> prof_kind;           // kernel of the class's last launch: 0 tile, 1 row, 2 cooperative,
                                          // 3 merged row, 4 row with the B rows in memory
    double prof_span_ms = 0;              // quartet phase of the profiled builds (all classes, all streams)
    long long prof_builds = 0;
};

namespace {

void drop_graph(mjk_basis* b)
{
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
}

void free_tiles(mjk_basis* b)
{
    drop_graph(b);   // its kernels read the task lists
    for (auto& c : b->qcls) {
        dfree(b->ctx, c.bt);
        dfree(b->ctx, c.bt_mol);
        dfree(b->ctx, c.bt_off);
        dfree(b->ctx, c.rt);
        dfree(b->ctx, c.rr);
        dfree(b->ctx, c.rtp);
    }
    dfree(b->ctx, b->d_tlist);
    dfree(b->ctx, b->d_tctl);
    b->d_tlist = nullptr;
    b->d_tctl = nullptr;
    b->qcls.clear();
    b->tile_cutoff = -1.0;
}

// Tuning entry of a class: the entry of its (angular momenta, contraction
// pattern) with the nearest work bucket, at most 2 buckets away; null if none.
const std::tuple<int, int, double>* find_tune(const mjk_ctx* ctx, const QClass& q)
{
    const std::tuple<int, int, double>* best = nullptr;
    int dist = 3;
    for (int d = 0; d <= 2 && !best; ++d)
        for (int sgn : {1, -1}) {
            const std::array<int, 9> key = {q.la, q.lb, q.lc, q.ld, q.ca, q.cb, q.cc, q.cd, q.bucket + sgn * d};
            auto it = ctx->tune.find(key);
            if (it != ctx->tune.end() && d < dist) {
                best = &it->second;
                dist = d;
            }
        }
    return best;
}

// Bra tiles of TILE pairs; each keeps the ket prefix whose Schwarz product
// with the tile's largest bound survives the cutoff.  One quartet class per
// pair of pair groups: for the (X, Y) pair classes, the plain x plain class
// always runs; where merged units exist, either the classes over the split
// groups (mode 1) or those over the merged groups (mode 2) run, never both.
// Within one group (bra and ket of the same group) only q <= p is kept; across
// groups every combination, each in one class.
int build_tiles(mjk_basis* b, double cutoff)
{
    mjk_ctx* ctx = b->ctx;
    free_tiles(b);
    const int nmol = b->nmol;
    auto has = [&](int G) {
        return b->seg[(size_t)(G + 1) * nmol * NBUCKET] > b->seg[(size_t)G * nmol * NBUCKET];
    };
    // decode (la, lb) of a pair class
    auto dec = [](int c, int& la, int& lb) {
        la = 0;
        while ((la + 1) * (la + 2) / 2 <= c) ++la;
        lb = c - la * (la + 1) / 2;
    };
    auto make_class = [&](QClass& qc, int GX, int GY) -> int {
        std::vector<int4> bt;
        std::vector<int> btm, off(1, 0);
        long long total = 0;
        double est_q = 0, est_p = 0;   // Schwarz-prescreened quartets and primitive quartets
        for (int m = 0; m < nmol; ++m)
        for (int bx = 0; bx < NBUCKET; ++bx) {
            const size_t sx = ((size_t)GX * nmol + m) * NBUCKET + bx;
            const int x0 = b->seg[sx], x1 = b->seg[sx + 1];
            for (int p0 = x0; p0 < x1; p0 += TILE)
            for (int by = 0; by < NBUCKET; ++by) {
                const size_t sy = ((size_t)GY * nmol + m) * NBUCKET + by;
                const int y0 = b->seg[sy], y1 = b->seg[sy + 1];
                if (y1 <= y0) continue;
                const int p1 = std::min(p0 + TILE, x1);
                int qend = y1;
                if (cutoff > 0) {
                    const double qb = b->q[p0];
                    // first ket index with q[y] * qb <= cutoff (q sorted descending)
                    int lo = y0, hi = y1;
                    while (lo < hi) {
                        const int mid = (lo + hi) / 2;
                        if (b->q[mid] * qb > cutoff) lo = mid + 1; else hi = mid;
                    }
                    qend = lo;
                }
                if (GX == GY) qend = std::min(qend, p1);
                const int nkt = (qend - y0 + TILE - 1) / TILE;
                if (nkt <= 0) continue;
                bt.push_back(make_int4(p0, p1, y0, qend));
                btm.push_back(m);
                for (int p = p0; p < p1; ++p) {
                    const int qe = GX == GY ? std::min(qend, p + 1) : qend;
                    if (qe > y0) {
                        est_q += qe - y0;
                        est_p += (double)b->h_nprim[p] * (b->pcost[qe] - b->pcost[y0]);
                    }
                }
                total += nkt;
                if (total > 2000000000LL)
                    return fail(ctx, MJK_EUNSUPPORTED, "quartet class has more than 2^31 tiles");
                off.push_back((int)total);
            }
        }
        qc.nbt = (int)bt.size();
        qc.nblocks = total;
        qc.work = (double)total * ncart(qc.la) * ncart(qc.lb) * ncart(qc.lc) * ncart(qc.ld) *
                  ((qc.la + qc.lb + qc.lc + qc.ld) / 2 + 1);
        qc.est = flops_class(qc.la, qc.lb, qc.lc, qc.ld, est_q * qc.ca * qc.cb * qc.cc * qc.cd, est_p,
                             qc.ca * qc.cb * qc.cc * qc.cd);
        qc.bucket = est_p >= 1.0 ? (int)std::floor(std::log2(est_p)) : 0;
        const std::tuple<int, int, double>* tuned = find_tune(ctx, qc);
        qc.kernel_pref = tuned ? std::get<0>(*tuned) : 0;
        if (qc.nbt > 0) {
            int rc;
            if ((rc = upload(ctx, bt, &qc.bt))) return rc;
            if ((rc = upload(ctx, btm, &qc.bt_mol))) return rc;
            if ((rc = upload(ctx, off, &qc.bt_off))) return rc;
        }
        // Row-kernel tasks: the same Schwarz prefixes per bra pair, cut into
        // chunks of about equal primitive-quartet count (the kets' primitive
        // counts differ up to a hundredfold).
        if (qc.nbt > 0 && (qc.ca * qc.cb * qc.cc * qc.cd > 1
                              ? row_gc_available(qc.la, qc.lb, qc.lc, qc.ld, qc.ca, qc.cb, qc.cc, qc.cd)
                              : row_kernel_available(qc.la, qc.lb, qc.lc, qc.ld))) {
            std::vector<int2> rr;
            std::vector<int> bra;
            long long cost_all = 0;
            for (int m = 0; m < nmol; ++m) {
                const size_t sx = ((size_t)GX * nmol + m) * NBUCKET, sy = ((size_t)GY * nmol + m) * NBUCKET;
                for (int p = b->seg[sx]; p < b->seg[sx + NBUCKET]; ++p) {
                    long long np_p = 0;
                    for (int by = 0; by < NBUCKET; ++by) {
                        const int y0 = b->seg[sy + by], y1 = b->seg[sy + by + 1];
                        int yend = y1;
                        if (cutoff > 0) {
                            int lo = y0, hi = y1;
                            while (lo < hi) {
                                const int mid = (lo + hi) / 2;
                                if (b->q[mid] * b->q[p] > cutoff) lo = mid + 1; else hi = mid;
                            }
                            yend = lo;
                        }
                        if (GX == GY) yend = std::min(yend, p + 1);
                        yend = std::max(yend, y0);
                        rr.push_back(make_int2(y0, yend));
                        np_p += b->pcost[yend] - b->pcost[y0];
                    }
                    if (np_p == 0) {
                        rr.resize(rr.size() - NBUCKET);
                        continue;
                    }
                    bra.push_back(p);
                    cost_all += (long long)b->h_nprim[p] * np_p;
                }
            }
            // enough blocks to fill the chip, each long enough to amortise its
            // fixed cost (LDS clear, range set-up, row flush)
            long long budget = std::min(std::max(cost_all / ctx->row_blocks, (long long)ctx->row_chunk),
                                        std::max(cost_all / ctx->row_min_blocks, 1LL));
            if (tuned && std::get<2>(*tuned) > 0)
                budget = std::max(1LL, (long long)((double)budget * std::get<2>(*tuned)));
            std::vector<std::pair<long long, int4>> tasks;
            std::vector<int> tprim;   // bra primitive offset of every task (parallel to tasks)
            // bra primitive pairs per task: small classes give every bra primitive
            // pair its own task (a small molecule's long contractions are its
            // critical path: H2O cc-pVDZ 0.69 -> 0.52 ms per build from 2 to 1);
            // large ones share them in eights (measured on MI355X, (H2O)16 def2-SVP
            // with the classes on four streams: 2 -> 92, 4 -> 107, 8 -> 110 builds/s)
            const int P = ctx->row_prim > 0                     ? ctx->row_prim
                          : tuned && std::get<1>(*tuned) > 0 ? std::get<1>(*tuned)
                          : cost_all < (1LL << 18)            ? 1
                          : cost_all < (1LL << 20)            ? 2
                                                              : 8;
            qc.rprim = P;
            for (size_t ib = 0; ib < bra.size(); ++ib)
            for (int pbo = 0; pbo < b->h_nprim[bra[ib]]; pbo += P) {
                // bra pair ib, its primitive pairs [pbo, pbo + P)
                const int p = bra[ib];
                const long long npb = std::min(P, b->h_nprim[p] - pbo);
                int flat = 0, start = 0;
                long long cur = 0;
                for (int k = 0; k < NBUCKET; ++k) {
                    int y = rr[ib * NBUCKET + k].x;
                    const int y1 = rr[ib * NBUCKET + k].y;
                    while (y < y1) {
                        const long long rest = npb * (b->pcost[y1] - b->pcost[y]);
                        if (cur + rest < budget) {
                            cur += rest;
                            flat += y1 - y;
                            break;
                        }
                        // smallest y' with npb * (pcost[y'] - pcost[y]) >= budget - cur
                        int lo = y + 1, hi = y1;
                        while (lo < hi) {
                            const int mid = (lo + hi) / 2;
                            if (npb * (b->pcost[mid] - b->pcost[y]) >= budget - cur) hi = mid; else lo = mid + 1;
                        }
                        cur += npb * (b->pcost[lo] - b->pcost[y]);
                        flat += lo - y;
                        y = lo;
                        tasks.push_back({cur, make_int4(p, start, flat, (int)ib)});
                        tprim.push_back(pbo);
                        start = flat;
                        cur = 0;
                    }
                }
                if (flat > start) {
                    tasks.push_back({cur, make_int4(p, start, flat, (int)ib)});
                    tprim.push_back(pbo);
                }
                tasks.back().second.z = INT_MAX;   // the last chunk runs to the end of the kets
            }
            // heaviest tasks first (the last chunk of a bra pair is the short one)
            std::vector<size_t> order(tasks.size());
            std::iota(order.begin(), order.end(), 0);
            std::stable_sort(order.begin(), order.end(),
                             [&](size_t u, size_t v) { return tasks[u].first > tasks[v].first; });
            if (tasks.size() > 2000000000ULL) return fail(ctx, MJK_EUNSUPPORTED, "too many row-kernel tasks");
            std::vector<int4> rt(tasks.size());
            std::vector<int> rtp(tasks.size());
            for (size_t t = 0; t < tasks.size(); ++t) {
                rt[t] = tasks[order[t]].second;
                rtp[t] = tprim[order[t]];
            }
            qc.nrt = (int)rt.size();
            int rc;
            if (qc.nrt > 0 && ((rc = upload(ctx, rt, &qc.rt)) || (rc = upload(ctx, rr, &qc.rr)) ||
                               (rc = upload(ctx, rtp, &qc.rtp))))
                return rc;
        }
        return MJK_OK;
    };
    int nbase = 0;
    for (int X = 0; X < NPAIRCLASS; ++X) {
        for (int Y = 0; Y <= X; ++Y, ++nbase) {
            int la, lb, lc, ld;
            dec(X, la, lb);
            dec(Y, lc, ld);
            const bool gc = has(X * NGT + G_SPLIT) || has(Y * NGT + G_SPLIT);
            std::vector<std::pair<int, int>> split, merged;   // (bra type, ket type)
            for (int tx = 0; tx < NGT; ++tx)
                for (int ty = 0; ty < NGT; ++ty) {
                    if (X == Y && ty > tx) continue;
                    if (!has(X * NGT + tx) || !has(Y * NGT + ty)) continue;
                    if (tx == G_PLAIN && ty == G_PLAIN) continue;
                    const bool sx = tx <= G_SPLIT, sy = ty <= G_SPLIT;
                    if (sx && sy) split.push_back({tx, ty});
                    else if ((tx == G_PLAIN || !sx) && (ty == G_PLAIN || !sy)) merged.push_back({tx, ty});
                }
            // merged mode needs a merged row kernel for every one of its classes
            // (the Coulomb operator only)
            bool merged_ok = gc && b->omega == 0.0 && !merged.empty();
            for (auto& t : merged)
                merged_ok = merged_ok && row_gc_available(la, lb, lc, ld, group_ca(t.first), group_cb(t.first),
                                                          group_ca(t.second), group_cb(t.second));
            std::vector<std::tuple<int, int, int>> subs;     // (bra type, ket type, mode)
            subs.push_back(std::make_tuple((int)G_PLAIN, (int)G_PLAIN, 0));
            for (auto& t : split) subs.push_back(std::make_tuple(t.first, t.second, merged_ok ? 1 : 0));
            if (merged_ok)
                for (auto& t : merged) subs.push_back(std::make_tuple(t.first, t.second, 2));
            for (auto& sub : subs) {
                QClass qc;
                qc.X = X;
                qc.Y = Y;
                qc.la = la, qc.lb = lb, qc.lc = lc, qc.ld = ld;
                const int tx = std::get<0>(sub), ty = std::get<1>(sub);
                qc.GX = X * NGT + tx;
                qc.GY = Y * NGT + ty;
                qc.ca = group_ca(tx), qc.cb = group_cb(tx), qc.cc = group_ca(ty), qc.cd = group_cb(ty);
                qc.mode = std::get<2>(sub);
                qc.base = nbase;
                int rc;
                if ((rc = make_class(qc, qc.GX, qc.GY))) return rc;
                b->qcls.push_back(qc);
                if ((int)b->qcls.size() > NQCLASS_MAX)
                    return fail(ctx, MJK_EUNSUPPORTED, "more than %d quartet classes", NQCLASS_MAX);
            }
        }
    }
    // compacted task lists and the persistent launches' counters
    long long ntl = 0;
    for (auto& qc : b->qcls) {
        qc.tl0 = ntl;
        ntl += qc.nrt;
    }
    int rc2;
    if (ntl > 0 && (rc2 = dalloc(ctx, sizeof(int) * (size_t)ntl, (void**)&b->d_tlist))) return rc2;
    if ((rc2 = dalloc(ctx, sizeof(int) * ROW_TCTL * NQCLASS_MAX, (void**)&b->d_tctl))) return rc2;
    b->tile_cutoff = cutoff;
    b->tune_gen = ctx->tune_gen;
    return MJK_OK;
}

PairSet pairset(const mjk_basis* b)
{
    PairSet P;
    P.ppP = b->ppP;
    P.ppK = b->ppK;
    P.ppW = b->ppW;
    P.prim0 = b->prim0;
    P.nprim = b->nprim;
    P.rA = b->rA;
    P.rAB = b->rAB;
    P.shA = b->shA;
    P.shB = b->shB;
    P.oshA = b->oshA;
    P.oshB = b->oshB;
    P.cartA = b->cartA;
    P.cartB = b->cartB;
    P.tP = b->tP;
    P.tK = b->tK;
    P.tW = b->tW;
    P.tpos = b->tpos;
    P.tstride = b->tstride;
    return P;
}

XformArgs xform(const mjk_basis* b)
{
    XformArgs x;
    x.cart_mol = b->d_cart_mol;
    x.cart_s0 = b->d_cart_s0;
    x.cart_l = b->d_cart_l;
    x.cart_c = b->d_cart_c;
    x.sph_mol = b->d_sph_mol;
    x.sph_c0 = b->d_sph_c0;
    x.sph_l = b->d_sph_l;
    x.sph_m = b->d_sph_m;
    x.dbase = b->d_dbase;
    x.ld = b->d_ld;
    x.sbase = b->d_sbase;
    x.nao = b->d_nao;
    x.cart_first = b->d_cart_first;
    x.sph_first = b->d_sph_first;
    x.nmol = b->nmol;
    x.ncart_tot = b->ncart_tot;
    x.nsph_tot = b->nsph_tot;
    return x;
}

DmCondArgs dmcond_args(const mjk_basis* b)
{
    DmCondArgs d;
    d.sh_mol = b->d_sh_mol;
    d.sh_s0 = b->d_sh_s0;
    d.sh_n = b->d_sh_n;
    d.sh_first = b->d_sh_first;
    d.cbase = b->d_cbase;
    d.nsh = b->d_nsh;
    d.sbase = b->d_sbase;
    d.nao = b->d_nao;
    d.nsh_tot = b->nsh_tot;
    d.maxnsh = b->maxnsh;
    return d;
}

template <typename T>
int track_upload(mjk_basis* b, const std::vector<T>& v, T** d)
{
    int rc = upload(b->ctx, v, d);
    if (rc == MJK_OK && *d) b->owned.push_back((void*)*d);
    return rc;
}

// Dynamic LDS of a row-kernel block of class q for n_j J and n_k K densities:
// the bra's K' rows and its J'[a,b] block (every contraction of a merged unit).
size_t row_shmem(const QClass& q, int n_j, int n_k, int maxld)
{
    const size_t ra = (size_t)q.ca * ncart(q.la), rb = (size_t)q.cb * ncart(q.lb);
    return sizeof(double) * ((size_t)n_k * (ra + rb) * maxld + (size_t)n_j * ra * rb);
}

// Static LDS of a row-kernel block of class q beside its dynamic rows: the
// copy of the Rys table that the one- and two-root classes read (the rule of
// RowRys in jk_device.h) and 1 KB for the task bookkeeping.
size_t row_static_lds(const QClass& q)
{
    const int L = q.la + q.lb + q.lc + q.ld, nr = L / 2 + 1;
    size_t t = 0;
    if (L > 0 && nr == 1) t = sizeof(double) * RysLds<1>::SIZE;
    if (L > 0 && nr == 2) t = sizeof(double) * RysLds<2>::SIZE;
    return t + 1024;
}

// Whether class q runs its row kernel as far as LDS goes: its rows within the
// context's limit, rows and static LDS within the workgroup's.
bool row_fits(const mjk_ctx* ctx, const QClass& q, int n_j, int n_k, int maxld)
{
    const size_t sh = row_shmem(q, n_j, n_k, maxld);
    return sh <= ctx->row_lds && sh + row_static_lds(q) <= (size_t)LDS_WORKGROUP;
}

// Dynamic LDS of the row kernel that keeps only the A rows of K' (the B rows'
// updates as memory-side atomics): plain classes of f-shell bras.
size_t row_shmem_m(const QClass& q, int n_j, int n_k, int maxld)
{
    const size_t ra = ncart(q.la), rb = ncart(q.lb);
    return sizeof(double) * ((size_t)n_k * ra * maxld + (size_t)n_j * ra * rb);
}

// Row kernel a plain class runs in a build: 1 with its full K' rows in LDS,
// 2 with only the A rows there (the Coulomb operator; when the full rows do
// not fit), 0 none (its fallback kernel).
int row_choice(const mjk_ctx* ctx, const mjk_basis* b, const QClass& q, int n_j, int n_k)
{
    if (q.nrt == 0 || q.kernel_pref == 2) return 0;
    if (row_fits(ctx, q, n_j, n_k, b->maxld)) return 1;
    if (q.mode != 0 || b->omega != 0.0 || !row_m_available(q.la, q.lb, q.lc, q.ld)) return 0;
    const size_t sh = row_shmem_m(q, n_j, n_k, b->maxld);
    return sh <= ctx->row_lds && sh + row_static_lds(q) <= (size_t)LDS_WORKGROUP ? 2 : 0;
}

int parse_molecule(mjk_ctx* ctx, const int* atm, int natm, const int* bas, int nbas, const double* env,
                   MolHost& m)
{
    if (natm <= 0 || nbas <= 0 || !atm || !bas || !env) return fail(ctx, MJK_EINVAL, "empty molecule");
    m.nbas_orig = nbas;
    int s0 = 0, c0 = 0;
    for (int i = 0; i < nbas; ++i) {
        const int* r = bas + i * BAS_SLOTS;
        const int l = r[ANG_OF], np = r[NPRIM_OF], nc = r[NCTR_OF], ia = r[ATOM_OF];
        if (l < 0 || l > LMAX) return fail(ctx, MJK_EUNSUPPORTED, "shell %d has l=%d (max %d)", i, l, LMAX);
        if (np <= 0 || nc <= 0 || ia < 0 || ia >= natm) return fail(ctx, MJK_EINVAL, "bad shell %d", i);
        const double* xyz = env + atm[ia * ATM_SLOTS + PTR_COORD];
        const double ang = (l == 0) ? C2S_S0 : (l == 1) ? C2S_P1 : 1.0;
        // a two-contraction s shell is kept whole as a merged unit
        const bool merged = l == 0 && nc == 2;
        m.orig_s0.push_back(s0);
        m.orig_n.push_back(nc * nsph(l));
        if (merged) m.units.push_back({(int)m.sh.size(), 2});
        for (int c = 0; c < nc; ++c) {
            GShell g;
            g.l = l;
            g.nprim = np;
            g.orig = i;
            g.ci = c;
            g.merged = merged;
            if (!merged) m.units.push_back({(int)m.sh.size(), 1});
            g.unit = (int)m.units.size() - 1;
            g.r[0] = xyz[0];
            g.r[1] = xyz[1];
            g.r[2] = xyz[2];
            g.e.assign(env + r[PTR_EXP], env + r[PTR_EXP] + np);
            g.c.resize(np);
            for (int k = 0; k < np; ++k) g.c[k] = env[r[PTR_COEFF] + c * np + k] * ang;
            g.s0 = s0;
            g.c0 = c0;
            s0 += nsph(l);
            c0 += ncart(l);
            m.sh.push_back(g);
        }
    }
    m.nao = s0;
    m.ncart = c0;
    m.nsh = (int)m.sh.size();
    return MJK_OK;
}

struct PairHost {
    int A, B;          // molecule-local (first) shell indices, l(A) >= l(B)
    int mol;
    int type = G_PLAIN;   // PairGroup
    int oA = 0, oB = 0;   // shells of the input basis (dm_cond)
    int uA = 0, uB = 0;   // units
    std::vector<double4> P;
    std::vector<double2> K;
    std::vector<double4> W;   // merged pairs: coefficient products of the combinations
};

int build_basis(mjk_ctx* ctx, mjk_basis* b, double cutoff)
{
    const int nmol = b->nmol;
    // offsets
    long long sb = 0, cb = 0, kb = 0, sf = 0, cf = 0, hf = 0;
    for (auto& m : b->mols) {
        m.sbase = sb;
        m.dbase = cb;
        m.cbase = kb;
        m.sph_first = sf;
        m.cart_first = cf;
        m.sh_first = hf;
        sb += (long long)m.nao * m.nao;
        cb += (long long)m.ncart * m.ncart;
        kb += (long long)m.nbas_orig * m.nbas_orig;   // dm_cond over the input basis's shells
        sf += m.nao;
        cf += m.ncart;
        hf += m.nbas_orig;
        b->maxld = std::max(b->maxld, m.ncart);
        b->maxnao = std::max(b->maxnao, m.nao);
        b->maxnsh = std::max(b->maxnsh, m.nbas_orig);
    }
    b->slot_s = sb;
    b->slot_c = cb;
    b->cond_total = kb;
    b->nsph_tot = sf;
    b->ncart_tot = cf;
    b->nsh_tot = hf;

    // shell pairs, grouped by class and pair group, then molecule
    std::vector<std::vector<PairHost>> byseg((size_t)NGROUP * nmol * NBUCKET);
    const double sq2pi54 = std::sqrt(2.0) * std::pow(M_PI, 1.25);
    for (int mi = 0; mi < nmol; ++mi) {
        const MolHost& m = b->mols[mi];
        for (int i = 0; i < m.nsh; ++i) {
            for (int j = 0; j <= i; ++j) {
                int A = i, B = j;
                if (m.sh[A].l < m.sh[B].l) std::swap(A, B);
                const GShell& sa = m.sh[A];
                const GShell& sbb = m.sh[B];
                const double ABx = sa.r[0] - sbb.r[0], ABy = sa.r[1] - sbb.r[1], ABz = sa.r[2] - sbb.r[2];
                const double rr = ABx * ABx + ABy * ABy + ABz * ABz;
                PairHost ph;
                ph.A = A;
                ph.B = B;
                ph.mol = mi;
                ph.type = sa.merged || sbb.merged ? G_SPLIT : G_PLAIN;
                ph.oA = sa.orig;
                ph.oB = sbb.orig;
                ph.uA = sa.unit;
                ph.uB = sbb.unit;
                for (int ia = 0; ia < sa.nprim; ++ia)
                    for (int ib = 0; ib < sbb.nprim; ++ib) {
                        const double ea = sa.e[ia], eb = sbb.e[ib];
                        const double p = ea + eb;
                        const double E = ea * eb / p * rr;
                        if (E > ctx->expcutoff) continue;
                        const double cc = sa.c[ia] * sbb.c[ib];
                        if (cc == 0.0) continue;
                        double4 P;
                        P.x = (ea * sa.r[0] + eb * sbb.r[0]) / p;
                        P.y = (ea * sa.r[1] + eb * sbb.r[1]) / p;
                        P.z = (ea * sa.r[2] + eb * sbb.r[2]) / p;
                        P.w = p;
                        ph.P.push_back(P);
                        ph.K.push_back(make_double2(cc * std::exp(-E) * sq2pi54 / p, 1.0 / p));
                    }
                if (ph.P.empty()) continue;
                const int bk = prim_bucket((int)ph.P.size());
                const int G = pair_class(sa.l, sbb.l) * NGT + ph.type;
                byseg[((size_t)G * nmol + mi) * NBUCKET + bk].push_back(std::move(ph));
            }
        }
        // merged pairs: pairs of units at least one of which is a merged s shell;
        // l(A) >= l(B), and for two s units the merged one (or the later one) first
        const int nu = (int)m.units.size();
        for (int u = 0; u < nu; ++u) {
            for (int v = 0; v <= u; ++v) {
                if (m.units[u].nctr == 1 && m.units[v].nctr == 1) continue;
                int uA = u, uB = v;
                const int lu = m.sh[m.units[u].first].l, lv = m.sh[m.units[v].first].l;
                if (lu < lv || (lu == lv && m.units[u].nctr < m.units[v].nctr)) std::swap(uA, uB);
                const UnitHost& UA = m.units[uA];
                const UnitHost& UB = m.units[uB];
                const GShell& sa = m.sh[UA.first];
                const GShell& sbb = m.sh[UB.first];
                const double ABx = sa.r[0] - sbb.r[0], ABy = sa.r[1] - sbb.r[1], ABz = sa.r[2] - sbb.r[2];
                const double rr = ABx * ABx + ABy * ABy + ABz * ABz;
                PairHost ph;
                ph.A = UA.first;
                ph.B = UB.first;
                ph.mol = mi;
                ph.type = UA.nctr == 2 && UB.nctr == 2 ? G_M22 : UA.nctr == 2 ? G_M21 : G_M12;
                ph.oA = sa.orig;
                ph.oB = sbb.orig;
                ph.uA = uA;
                ph.uB = uB;
                for (int ia = 0; ia < sa.nprim; ++ia)
                    for (int ib = 0; ib < sbb.nprim; ++ib) {
                        const double ea = sa.e[ia], eb = sbb.e[ib];
                        const double p = ea + eb;
                        const double E = ea * eb / p * rr;
                        if (E > ctx->expcutoff) continue;
                        double w[4] = {0, 0, 0, 0};
                        bool any = false;
                        for (int ci = 0; ci < UA.nctr; ++ci)
                            for (int cj = 0; cj < UB.nctr; ++cj) {
                                w[ci * UB.nctr + cj] = m.sh[UA.first + ci].c[ia] * m.sh[UB.first + cj].c[ib];
                                any = any || w[ci * UB.nctr + cj] != 0.0;
                            }
                        if (!any) continue;
                        double4 P;
                        P.x = (ea * sa.r[0] + eb * sbb.r[0]) / p;
                        P.y = (ea * sa.r[1] + eb * sbb.r[1]) / p;
                        P.z = (ea * sa.r[2] + eb * sbb.r[2]) / p;
                        P.w = p;
                        ph.P.push_back(P);
                        ph.K.push_back(make_double2(std::exp(-E) * sq2pi54 / p, 1.0 / p));
                        ph.W.push_back(make_double4(w[0], w[1], w[2], w[3]));
                    }
                if (ph.P.empty()) continue;
                const int bk = prim_bucket((int)ph.P.size());
                const int G = pair_class(sa.l, sbb.l) * NGT + ph.type;
                byseg[((size_t)G * nmol + mi) * NBUCKET + bk].push_back(std::move(ph));
            }
        }
    }
    // flatten (unsorted) for the Schwarz kernel
    auto flatten = [&](std::vector<int>& order_seg_start) {
        order_seg_start.assign((size_t)NGROUP * nmol * NBUCKET + 1, 0);
        int n = 0;
        for (size_t s = 0; s < byseg.size(); ++s) {
            order_seg_start[s] = n;
            n += (int)byseg[s].size();
        }
        order_seg_start.back() = n;
        return n;
    };
    auto upload_pairs = [&](std::vector<double>* qv) -> int {
        std::vector<double4> hP, hA, hAB, hW;
        std::vector<double2> hK;
        std::vector<int> hp0, hnp, hsa, hsb, hca, hcb, hoa, hob;
        for (auto& segv : byseg)
            for (auto& ph : segv) {
                const MolHost& m = b->mols[ph.mol];
                const GShell& sa = m.sh[ph.A];
                const GShell& sbb = m.sh[ph.B];
                hp0.push_back((int)hP.size());
                hnp.push_back((int)ph.P.size());
                hP.insert(hP.end(), ph.P.begin(), ph.P.end());
                hK.insert(hK.end(), ph.K.begin(), ph.K.end());
                if (ph.W.empty()) hW.resize(hW.size() + ph.P.size(), make_double4(0, 0, 0, 0));
                else hW.insert(hW.end(), ph.W.begin(), ph.W.end());
                hoa.push_back(ph.oA);
                hob.push_back(ph.oB);
                const double qq = qv ? (*qv)[hsa.size()] : 0.0;
                hA.push_back(make_double4(sa.r[0], sa.r[1], sa.r[2], qq));
                hAB.push_back(make_double4(sa.r[0] - sbb.r[0], sa.r[1] - sbb.r[1], sa.r[2] - sbb.r[2], 0.0));
                hsa.push_back(ph.A);
                hsb.push_back(ph.B);
                hca.push_back(sa.c0);
                hcb.push_back(sbb.c0);
            }
        if (hP.size() > (size_t)2000000000) return fail(ctx, MJK_EUNSUPPORTED, "too many primitive pairs");
        int rc;
        if ((rc = track_upload(b, hP, &b->ppP))) return rc;
        if ((rc = track_upload(b, hK, &b->ppK))) return rc;
        if ((rc = track_upload(b, hW, &b->ppW))) return rc;
        if ((rc = track_upload(b, hoa, &b->oshA))) return rc;
        if ((rc = track_upload(b, hob, &b->oshB))) return rc;
        if ((rc = track_upload(b, hp0, &b->prim0))) return rc;
        if ((rc = track_upload(b, hnp, &b->nprim))) return rc;
        if ((rc = track_upload(b, hA, &b->rA))) return rc;
        if ((rc = track_upload(b, hAB, &b->rAB))) return rc;
        if ((rc = track_upload(b, hsa, &b->shA))) return rc;
        if ((rc = track_upload(b, hsb, &b->shB))) return rc;
        if ((rc = track_upload(b, hca, &b->cartA))) return rc;
        if ((rc = track_upload(b, hcb, &b->cartB))) return rc;
        if (qv) {
            // ket-side transposed copy: per segment, primitive ik of its pair j at
            // base + ik * n + j (n pairs, padded to the segment's largest count)
            std::vector<double4> tP, tW;
            std::vector<double2> tK;
            std::vector<int> tpos, tstr;
            tpos.reserve(hp0.size());
            tstr.reserve(hp0.size());
            for (auto& segv : byseg) {
                const int n = (int)segv.size();
                size_t mx = 0;
                for (auto& ph : segv) mx = std::max(mx, ph.P.size());
                const size_t base = tP.size();
                tP.resize(base + mx * n, make_double4(0, 0, 0, 1));
                tK.resize(base + mx * n, make_double2(0, 1));
                tW.resize(base + mx * n, make_double4(0, 0, 0, 0));
                for (int j = 0; j < n; ++j) {
                    const PairHost& ph = segv[j];
                    for (size_t ik = 0; ik < ph.P.size(); ++ik) {
                        tP[base + ik * n + j] = ph.P[ik];
                        tK[base + ik * n + j] = ph.K[ik];
                        if (!ph.W.empty()) tW[base + ik * n + j] = ph.W[ik];
                    }
                    tpos.push_back((int)(base + j));
                    tstr.push_back(n);
                }
            }
            if (tP.size() > (size_t)2000000000) return fail(ctx, MJK_EUNSUPPORTED, "too many primitive pairs");
            if ((rc = track_upload(b, tP, &b->tP))) return rc;
            if ((rc = track_upload(b, tK, &b->tK))) return rc;
            if ((rc = track_upload(b, tW, &b->tW))) return rc;
            if ((rc = track_upload(b, tpos, &b->tpos))) return rc;
            if ((rc = track_upload(b, tstr, &b->tstride))) return rc;
        }
        b->h_shA = hsa;
        b->h_shB = hsb;
        b->h_nprim = hnp;
        return MJK_OK;
    };
    auto release_pairs = [&]() {
        void* ps[] = {b->ppP, b->ppK, b->ppW, b->prim0, b->nprim, b->rA, b->rAB, b->shA, b->shB, b->oshA, b->oshB,
                      b->cartA, b->cartB};
        for (void* p : ps) {
            dfree(ctx, p);
            b->owned.erase(std::remove(b->owned.begin(), b->owned.end(), p), b->owned.end());
        }
    };

    std::vector<int> seg0;
    const int np = flatten(seg0);
    b->npairs = np;
    int rc = upload_pairs(nullptr);
    if (rc) return rc;
    // Schwarz bounds on the device, one launch per pair class
    double* dq = nullptr;
    if ((rc = dalloc(ctx, sizeof(double) * std::max(np, 1), (void**)&dq))) return rc;
    for (int X = 0; X < NPAIRCLASS; ++X) {
        // every group of the class (the bounds of merged pairs are replaced below)
        const int p0 = seg0[(size_t)X * NGT * nmol * NBUCKET], p1 = seg0[(size_t)(X + 1) * NGT * nmol * NBUCKET];
        if (p1 <= p0) continue;
        int la = 0;
        while ((la + 1) * (la + 2) / 2 <= X) ++la;
        const int lb = X - la * (la + 1) / 2;
        if (la > lmax_built()) {
            dfree(ctx, dq);
            return fail(ctx, MJK_EUNSUPPORTED, "basis has l=%d shells; kernels are built for l<=%d", la,
                        lmax_built());
        }
        QCondArgs qa;
        qa.P = pairset(b);
        qa.p0 = p0;
        qa.np = p1 - p0;
        qa.qout = dq;
        qa.rys = ctx->rys;
        qa.omega = b->omega;
        if (launch_qcond(la, lb, qa, ctx->stream)) {
            dfree(ctx, dq);
            return fail(ctx, MJK_EDEVICE, "q_cond launch failed for class (%d,%d)", la, lb);
        }
    }
    std::vector<double> qflat(np);
    HIPCHK(ctx, hipMemcpyAsync(qflat.data(), dq, sizeof(double) * np, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    dfree(ctx, dq);
    release_pairs();
    // the bound of a merged pair: the largest bound of its split pairs (the
    // reference's q_cond over the whole shell, pyscf/lib/vhf/optimizer.c:408-454)
    {
        std::map<std::tuple<int, int, int>, double> qmax;
        int k = 0;
        for (auto& segv : byseg)
            for (auto& ph : segv) {
                if (ph.type == G_SPLIT) {
                    const auto key = std::make_tuple(ph.mol, std::max(ph.uA, ph.uB), std::min(ph.uA, ph.uB));
                    auto it = qmax.find(key);
                    if (it == qmax.end()) qmax[key] = qflat[k];
                    else it->second = std::max(it->second, qflat[k]);
                }
                ++k;
            }
        k = 0;
        for (auto& segv : byseg)
            for (auto& ph : segv) {
                if (ph.type >= G_M21) {
                    const auto it = qmax.find(std::make_tuple(ph.mol, std::max(ph.uA, ph.uB), std::min(ph.uA, ph.uB)));
                    qflat[k] = it == qmax.end() ? 1e-100 : it->second;
                }
                ++k;
            }
    }
    // sort each (class, molecule) segment by descending bound
    {
        int k = 0;
        for (auto& segv : byseg) {
            std::vector<int> idx(segv.size());
            std::iota(idx.begin(), idx.end(), 0);
            std::vector<double> qs(qflat.begin() + k, qflat.begin() + k + segv.size());
            std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return qs[x] > qs[y]; });
            std::vector<PairHost> sorted;
            sorted.reserve(segv.size());
            for (int i : idx) sorted.push_back(std::move(segv[i]));
            segv = std::move(sorted);
            for (size_t i = 0; i < idx.size(); ++i) qflat[k + i] = qs[idx[i]];
            k += (int)segv.size();
        }
    }
    b->q = qflat;
    flatten(b->seg);
    if ((rc = upload_pairs(&b->q))) return rc;
    b->pcost.assign(np + 1, 0);
    for (int p = 0; p < np; ++p) b->pcost[p + 1] = b->pcost[p] + b->h_nprim[p];
    {
        std::vector<int> pmol(std::max(np, 1), 0);
        for (int G = 0; G < NGROUP; ++G)
            for (int m = 0; m < nmol; ++m)
                for (int p = b->seg[((size_t)G * nmol + m) * NBUCKET]; p < b->seg[((size_t)G * nmol + m + 1) * NBUCKET]; ++p)
                    pmol[p] = m;
        if ((rc = track_upload(b, pmol, &b->d_pmol))) return rc;
        std::vector<int> act(nmol, 1);
        if ((rc = track_upload(b, act, &b->d_active))) return rc;
    }

    // per molecule / AO maps
    std::vector<long long> dbase, cbase, sbase, cfirst, sfirst, hfirst;
    std::vector<int> ld, nsh, nao, cmol, cs0, cl, cc, smol, sc0, sl, sm, hmol, hs0, hn;
    for (int mi = 0; mi < nmol; ++mi) {
        const MolHost& m = b->mols[mi];
        dbase.push_back(m.dbase);
        cbase.push_back(m.cbase);
        sbase.push_back(m.sbase);
        cfirst.push_back(m.cart_first);
        sfirst.push_back(m.sph_first);
        hfirst.push_back(m.sh_first);
        ld.push_back(m.ncart);
        nsh.push_back(m.nbas_orig);
        nao.push_back(m.nao);
        for (const GShell& g : m.sh) {
            for (int c = 0; c < ncart(g.l); ++c) {
                cmol.push_back(mi);
                cs0.push_back(g.s0);
                cl.push_back(g.l);
                cc.push_back(c);
            }
            for (int s = 0; s < nsph(g.l); ++s) {
                smol.push_back(mi);
                sc0.push_back(g.c0);
                sl.push_back(g.l);
                sm.push_back(s);
            }
        }
        // dm_cond shells: the input basis's shells, all contractions
        for (int I = 0; I < m.nbas_orig; ++I) {
            hmol.push_back(mi);
            hs0.push_back(m.orig_s0[I]);
            hn.push_back(m.orig_n[I]);
        }
    }
    if ((rc = track_upload(b, dbase, &b->d_dbase))) return rc;
    if ((rc = track_upload(b, cbase, &b->d_cbase))) return rc;
    if ((rc = track_upload(b, sbase, &b->d_sbase))) return rc;
    b->h_sbase = sbase;
    b->h_nao = nao;
    b->h_active.assign(b->nmol, 1);
    if ((rc = track_upload(b, cfirst, &b->d_cart_first))) return rc;
    if ((rc = track_upload(b, sfirst, &b->d_sph_first))) return rc;
    if ((rc = track_upload(b, hfirst, &b->d_sh_first))) return rc;
    if ((rc = track_upload(b, ld, &b->d_ld))) return rc;
    if ((rc = track_upload(b, nsh, &b->d_nsh))) return rc;
    if ((rc = track_upload(b, nao, &b->d_nao))) return rc;
    if ((rc = track_upload(b, cmol, &b->d_cart_mol))) return rc;
    if ((rc = track_upload(b, cs0, &b->d_cart_s0))) return rc;
    if ((rc = track_upload(b, cl, &b->d_cart_l))) return rc;
    if ((rc = track_upload(b, cc, &b->d_cart_c))) return rc;
    if ((rc = track_upload(b, smol, &b->d_sph_mol))) return rc;
    if ((rc = track_upload(b, sc0, &b->d_sph_c0))) return rc;
    if ((rc = track_upload(b, sl, &b->d_sph_l))) return rc;
    if ((rc = track_upload(b, sm, &b->d_sph_m))) return rc;
    if ((rc = track_upload(b, hmol, &b->d_sh_mol))) return rc;
    if ((rc = track_upload(b, hs0, &b->d_sh_s0))) return rc;
    if ((rc = track_upload(b, hn, &b->d_sh_n))) return rc;
    if ((rc = dalloc(ctx, sizeof(unsigned long long) * 4 * NQCLASS_MAX, (void**)&b->counters))) return rc;
    b->owned.push_back(b->counters);
    return build_tiles(b, cutoff);
}

// Classes a build of n_j J and n_k K densities runs: the plain classes, and
// per (X, Y) pair of pair classes either its merged classes -- when every one
// of them has row tasks whose LDS fits -- or its split classes.
std::vector<char> classes_to_run(const mjk_ctx* ctx, const mjk_basis* b, int n_j, int n_k)
{
    const int nb = b->qcls.empty() ? 0 : b->qcls.back().base + 1;
    std::vector<char> merged(nb, 0);
    for (const QClass& q : b->qcls)
        if (q.mode == 2) merged[q.base] = 1;
    for (const QClass& q : b->qcls)
        if (q.mode == 2 && q.nblocks > 0 &&
            (q.nrt == 0 || q.kernel_pref == 2 || !row_fits(ctx, q, n_j, n_k, b->maxld)))
            merged[q.base] = 0;
    std::vector<char> run(b->qcls.size(), 0);
    for (size_t c = 0; c < b->qcls.size(); ++c) {
        const QClass& q = b->qcls[c];
        run[c] = q.nblocks > 0 && (q.mode == 0 || (q.mode == 1) != (merged[q.base] != 0));
    }
    return run;
}

int ensure(mjk_ctx* ctx, double** p, long long& cap, long long need)
{
    if (need <= cap && *p) return MJK_OK;
    dfree(ctx, *p);
    *p = nullptr;
    cap = 0;
    int rc = dalloc(ctx, sizeof(double) * std::max(need, 1LL), (void**)p);
    if (rc) return rc;
    cap = need;
    return MJK_OK;
}

}  // namespace

namespace {

// Real solid harmonics r^l Y_lm (unit-sphere normalised) as rows m = -l..l over
// the Cartesian monomials of angular momentum l (order lx = l..0, ly =
// l-lx..0), from the closed formula (Helgaker, Jorgensen, Olsen, Molecular
// Electronic-Structure Theory, eqs. 6.4.47-6.4.50):
//   S_lm = N_lm sum_{t, u, v} (-1)^(t+v-v_m) 4^-t C(l,t) C(l-t,|m|+t) C(t,u) C(|m|,2v)
//          x^(2t+|m|-2(u+v)) y^(2(u+v)) z^(l-2t-|m|),
//   v = v_m, v_m + 1, ... <= |m|/2, v_m = 0 (m >= 0) or 1/2 (m < 0),
//   N_lm = sqrt(2 (l+|m|)! (l-|m|)! / 2^delta(m,0)) / (2^|m| l!),
// times sqrt((2l+1)/(4 pi)).  Used for l = 4 (auxiliary g functions); the
// l <= 3 tables of c2s.h agree with it.
void solid_harmonic_rows(int l, double* out)
{
    auto binom = [](int n, int k) {
        if (k < 0 || k > n) return 0.0;
        double r = 1.0;
        for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
        return r;
    };
    auto fact = [](int n) {
        double r = 1.0;
        for (int i = 2; i <= n; ++i) r *= i;
        return r;
    };
    const int nc = (l + 1) * (l + 2) / 2;
    std::fill(out, out + (2 * l + 1) * nc, 0.0);
    for (int m = -l; m <= l; ++m) {
        const int am = std::abs(m), v0 = m < 0 ? 1 : 0;
        const double N = std::sqrt(2.0 * fact(l + am) * fact(l - am) / (m == 0 ? 2.0 : 1.0)) /
                         (std::pow(2.0, am) * fact(l)) * std::sqrt((2.0 * l + 1.0) / (4.0 * M_PI));
        for (int t = 0; t <= (l - am) / 2; ++t)
            for (int u = 0; u <= t; ++u)
                for (int v2 = v0; v2 <= am; v2 += 2) {
                    const double sgn = ((t + (v2 - v0) / 2) % 2) ? -1.0 : 1.0;
                    const double C = sgn * std::pow(0.25, t) * binom(l, t) * binom(l - t, am + t) * binom(t, u) *
                                     binom(am, v2);
                    const int px = 2 * t + am - 2 * u - v2, py = 2 * u + v2;
                    const int r = l - px;   // = py + pz
                    const int pz = l - px - py;
                    out[(m + l) * nc + r * (r + 1) / 2 + pz] += N * C;
                }
    }
}

// The tuning table the context starts with: the file MJK_TUNING names, else
// tuning/mi355x.txt beside the library's lib/ folder; none when MJK_TUNE=0.
// Lines: la lb lc ld ca cb cc cd bucket kernel rprim scale [anything], '#'
// starts a comment.  An unreadable or malformed file leaves the table empty.
void load_tuning_file(mjk_ctx* ctx)
{
    if (const char* t = getenv("MJK_TUNE"))
        if (atoi(t) == 0) return;
    std::string path;
    if (const char* f = getenv("MJK_TUNING")) {
        path = f;
    } else {
        Dl_info info;
        if (!dladdr((void*)&mjk_ctx_create, &info) || !info.dli_fname) return;
        std::string lib = info.dli_fname;
        const size_t slash = lib.find_last_of('/');
        const std::string dir = slash == std::string::npos ? std::string(".") : lib.substr(0, slash);
        path = dir + "/../tuning/mi355x.txt";
    }
    std::ifstream in(path);
    if (!in) return;
    std::vector<int> keys;
    std::vector<double> scale;
    std::string line;
    while (std::getline(in, line)) {
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
        std::istringstream ls(line);
        int k[11];
        double sc;
        for (int i = 0; i < 11; ++i)
            if (!(ls >> k[i])) return;
        if (!(ls >> sc)) return;
        keys.insert(keys.end(), k, k + 11);
        scale.push_back(sc);
    }
    if (mjk_ctx_set_tuning(ctx, (int)scale.size(), keys.data(), scale.data()) != MJK_OK) {
        ctx->err.clear();
        ctx->tune.clear();
    }
}

}  // namespace

namespace mjk_internal {

hipStream_t ctx_stream(mjk_ctx* ctx) { return ctx->stream; }
int ctx_device(const mjk_ctx* ctx) { return ctx->device; }
size_t ctx_budget(const mjk_ctx* ctx) { return ctx->budget; }
std::recursive_mutex& ctx_mutex(mjk_ctx* ctx) { return ctx->mu; }
std::shared_mutex& capture_mutex() { return g_capture; }
void* ctx_blas(mjk_ctx* ctx) { return ctx->blas; }
const double* ctx_rys(const mjk_ctx* ctx) { return ctx->rys; }
void ctx_set_blas(mjk_ctx* ctx, void* handle, void (*destroy)(void*))
{
    ctx->blas = handle;
    ctx->blas_destroy = destroy;
}

int ctx_fail(mjk_ctx* ctx, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    return code;
}

}  // namespace mjk_internal

extern "C" {

int mjk_ctx_create(mjk_ctx** out, int device, size_t hbm_budget)
{
    if (!out) return MJK_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return MJK_EDEVICE;
    if (device < 0 || device >= ndev) return MJK_EINVAL;
    mjk_ctx* ctx = new mjk_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return MJK_EDEVICE;
    }
    size_t fr = 0, tot = 0;
    (void)hipMemGetInfo(&fr, &tot);
    ctx->budget = hbm_budget ? hbm_budget : (size_t)(0.8 * (double)fr);
    if (const char* nc = getenv("MJK_NCOPY")) ctx->ncopy = std::max(1, std::min(16, atoi(nc)));
    if (const char* rc = getenv("MJK_ROW_CHUNK")) ctx->row_chunk = std::max(256, std::min(1 << 24, atoi(rc)));
    if (const char* rm = getenv("MJK_ROW_MIN_BLOCKS")) ctx->row_min_blocks = std::max(1, std::min(1 << 20, atoi(rm)));
    if (const char* rq = getenv("MJK_ROW_PRIM")) ctx->row_prim = std::max(0, std::min(4096, atoi(rq)));
    if (const char* rb = getenv("MJK_ROW_BLOCKS")) ctx->row_blocks = std::max(256, std::min(1 << 20, atoi(rb)));
    if (const char* rl = getenv("MJK_ROW_LDS")) ctx->row_lds = (size_t)std::max(0, std::min(ROW_LDS_MAX, atoi(rl)));
    if (const char* ns = getenv("MJK_STREAMS")) ctx->nstreams = std::max(1, std::min(1 + mjk_ctx::NAUX, atoi(ns)));
    if (const char* sc = getenv("MJK_SCHED")) ctx->sched = atoi(sc);
    if (const char* gr = getenv("MJK_GRAPHS")) ctx->graphs = atoi(gr) != 0;
    if (const char* od = getenv("MJK_ORDER")) ctx->order = atoi(od);
    if (const char* pe = getenv("MJK_PERSIST")) ctx->persist = std::max(0, std::min(2, atoi(pe)));
    if (const char* pb = getenv("MJK_PERSIST_BLOCKS")) ctx->persist_blocks = std::max(64, std::min(1 << 20, atoi(pb)));
    if (const char* ec = getenv("MJK_EXPCUTOFF")) ctx->expcutoff = std::max(20.0, std::min(EXPCUTOFF_MAX, atof(ec)));
    for (auto& e : ctx->ev) (void)hipEventCreate(&e);
    bool streams_ok = hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; i < mjk_ctx::NAUX; ++i)
        streams_ok = streams_ok && hipStreamCreateWithFlags(&ctx->aux[i], hipStreamNonBlocking) == hipSuccess &&
                     hipEventCreateWithFlags(&ctx->ev_join[i], hipEventDisableTiming) == hipSuccess;
    if (!streams_ok) {
        mjk_ctx_destroy(ctx);
        return MJK_EDEVICE;
    }
    scratch_init(device);   // once per device; a failure is reported by the no-allocation build
    load_tuning_file(ctx);
    std::vector<double> tab(RYS_BUF_SIZE);
    rys_fill_buffer(tab.data());
    if (upload(ctx, tab, &ctx->rys) || upload_c2s(ctx->stream) || hipStreamSynchronize(ctx->stream) != hipSuccess) {
        mjk_ctx_destroy(ctx);
        return MJK_EDEVICE;
    }
    *out = ctx;
    return MJK_OK;
}

void mjk_ctx_destroy(mjk_ctx* ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    {
        std::vector<void*> ps;
        for (auto& kv : ctx->allocs) ps.push_back(kv.first);
        for (void* p : ps) dfree(ctx, p);
    }
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (int i = 0; i < mjk_ctx::NAUX; ++i) {
        if (ctx->aux[i]) {
            (void)hipStreamSynchronize(ctx->aux[i]);
            (void)hipStreamDestroy(ctx->aux[i]);
        }
        if (ctx->ev_join[i]) (void)hipEventDestroy(ctx->ev_join[i]);
    }
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->blas && ctx->blas_destroy) {
        std::unique_lock<std::shared_mutex> lk(g_capture);
        ctx->blas_destroy(ctx->blas);
    }
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* mjk_last_error(const mjk_ctx* ctx)
{
    return ctx ? ctx->err.c_str() : "no context";
}

size_t mjk_ctx_hbm_used(const mjk_ctx* ctx)
{
    return ctx ? ctx->used.load() : 0;
}

int mjk_ctx_set_row_lds(mjk_ctx* ctx, size_t bytes)
{
    if (!ctx) return MJK_EINVAL;
    if (bytes > (size_t)ROW_LDS_MAX) return fail(ctx, MJK_EINVAL, "row kernel LDS limit is %d bytes", ROW_LDS_MAX);
    std::lock_guard<std::recursive_mutex> lock(ctx->mu);
    ctx->row_lds = bytes;
    return MJK_OK;
}

int mjk_ctx_set_graphs(mjk_ctx* ctx, int enable)
{
    if (!ctx) return MJK_EINVAL;
    std::lock_guard<std::recursive_mutex> lock(ctx->mu);
    ctx->graphs = enable != 0;
    return MJK_OK;
}

int mjk_basis_upload_batch(mjk_ctx* ctx, int nmol, const int* natm, const int* nbas, const int* const* atm,
                           const int* const* bas, const double* const* env, double cutoff, mjk_basis** out)
{
    if (!ctx || !out || nmol <= 0) return fail(ctx, MJK_EINVAL, "bad arguments to mjk_basis_upload_batch");
    std::lock_guard<std::recursive_mutex> lock(ctx->mu);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    *out = nullptr;
    mjk_basis* b = new mjk_basis();
    b->ctx = ctx;
    b->nmol = nmol;
    b->mols.resize(nmol);
    int rc = MJK_OK;
    for (int m = 0; m < nmol && rc == MJK_OK; ++m)
        rc = parse_molecule(ctx, atm[m], natm[m], bas[m], nbas[m], env[m], b->mols[m]);
    // the operator of every two-electron integral, as libcint reads it
    if (rc == MJK_OK) {
        b->omega = env[0][PTR_RANGE_OMEGA];
        for (int m = 1; m < nmol; ++m)
            if (env[m][PTR_RANGE_OMEGA] != b->omega) {
                rc = fail(ctx, MJK_EINVAL, "molecules of one batch must share env[PTR_RANGE_OMEGA]");
                break;
            }
        if (rc == MJK_OK && !std::isfinite(b->omega))
            rc = fail(ctx, MJK_EINVAL, "env[PTR_RANGE_OMEGA] is not finite");
    }
    if (rc == MJK_OK) rc = build_basis(ctx, b, cutoff);
    if (rc != MJK_OK) {
        std::string keep = ctx->err;
        mjk_basis_destroy(b);
        ctx->err = keep;
        return rc;
    }
    *out = b;
    return MJK_OK;
}

int mjk_basis_upload(mjk_ctx* ctx, const int* atm, int natm, const int* bas, int nbas, const double* env,
                     double cutoff, mjk_basis** out)
{
    return mjk_basis_upload_batch(ctx, 1, &natm, &nbas, &atm, &bas, &env, cutoff, out);
}

void mjk_basis_destroy(mjk_basis* b)
{
    if (!b) return;
    mjk_ctx* ctx = b->ctx;
    std::lock_guard<std::recursive_mutex> lock(ctx->mu);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    free_tiles(b);
    for (void* p : b->owned) dfree(ctx, p);
    double* work[] = {b->Dsph, b->Vj, b->Vk, b->Dj, b->Dk, b->J, b->K, b->dmcond};
    for (double* p : work) dfree(ctx, p);
    for (hipEvent_t e : b->pev) (void)hipEventDestroy(e);
    delete b;
}

int mjk_basis_set_active(mjk_ctx* ctx, mjk_basis* b, const int* active)
{
    if (!ctx || !b || b->ctx != ctx || !active) return fail(ctx, MJK_EINVAL, "bad arguments to mjk_basis_set_active");
    std::lock_guard<std::recursive_mutex> lock(ctx->mu);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    std::vector<int> act(b->nmol);
    for (int m = 0; m < b->nmol; ++m) act[m] = active[m] != 0;
    // stream-ordered before the builds that follow (a replayed graph reads the same buffer)
    HIPCHK(ctx, hipMemcpyAsync(b->d_active, act.data(), sizeof(int) * b->nmol, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    b->h_active = act;
    return MJK_OK;
}

int mjk_basis_nao(const mjk_basis* b, int mol)
{
    if (!b || mol < 0 || mol >= b->nmol) return MJK_EINVAL;
    return b->mols[mol].nao;
}

long long mjk_basis_slot_size(const mjk_basis* b)
{
    return b ? b->slot_s : MJK_EINVAL;
}

int mjk_basis_q_cond(mjk_ctx* ctx, const mjk_basis* b, int mol, double* q_cond)
{
    if (!ctx || !b || mol < 0 || mol >= b->nmol || !q_cond) return fail(ctx, MJK_EINVAL, "bad arguments");
    const MolHost& m = b->mols[mol];
    const int n = m.nbas_orig;
    for (int i = 0; i < n * n; ++i) q_cond[i] = 1e-100;
    for (int G = 0; G < NGROUP; ++G) {
        const int p0 = b->seg[((size_t)G * b->nmol + mol) * NBUCKET];
        const int p1 = b->seg[((size_t)G * b->nmol + mol + 1) * NBUCKET];
        for (int p = p0; p < p1; ++p) {
            const int i = m.sh[b->h_shA[p]].orig, j = m.sh[b->h_shB[p]].orig;
            q_cond[i * n + j] = std::max(q_cond[i * n + j], b->q[p]);
            q_cond[j * n + i] = q_cond[i * n + j];
        }
    }
    return MJK_OK;
}

// One J/K build.  scratch == nullptr: the basis's own work buffers (grown on
// demand), task lists rebuilt when the cutoff changes, graph replay, profiling
// and statistics.  scratch != nullptr (this device's address of the
// no-allocation scratch array): the work buffers are carved out of it and no
// device memory is allocated or freed -- the task lists must already exist
// for this cutoff, and the build runs eagerly without statistics or profiling.
static int get_jk_impl(mjk_ctx* ctx, mjk_basis* b, const double* d_dms, int n_dm, int hermi, int with_j,
                       int with_k, double direct_scf_tol, double fp32_tol, double* d_vj, double* d_vk,
                       mjk_stats* stats, double* scratch)
{
    if (!ctx || !b || b->ctx != ctx) return fail(ctx, MJK_EINVAL, "context/basis mismatch");
    if (n_dm <= 0 || !d_dms) return fail(ctx, MJK_EINVAL, "n_dm must be positive");
    if (hermi < 0 || hermi > 2) return fail(ctx, MJK_EINVAL, "hermi must be 0, 1 or 2");
    if ((with_j && !d_vj) || (with_k && !d_vk)) return fail(ctx, MJK_EINVAL, "missing output buffer");
    if (!with_j && !with_k) return MJK_OK;
    std::lock_guard<std::recursive_mutex> lock(ctx->mu);
    auto t0 = std::chrono::steady_clock::now();
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const bool noalloc = scratch != nullptr;
    if (!noalloc) drain_deferred();
    const bool screen = direct_scf_tol > 0;
    const double cut = screen ? direct_scf_tol : 0.0;
    int rc;
    if (b->tile_cutoff != cut || b->tune_gen != ctx->tune_gen) {
        if (noalloc)
            return fail(ctx, MJK_EINVAL,
                        "no-allocation build: the basis's quartet task lists are built for direct_scf_tol %g "
                        "and tuning table %d, not %g and %d (upload the basis with this cutoff, or run one "
                        "mjk_get_jk_device build)",
                        b->tile_cutoff, b->tune_gen, cut, ctx->tune_gen);
        if ((rc = build_tiles(b, cut))) return rc;
    }
    const int n_j = with_j ? n_dm : 0;
    const int n_k = with_k ? (hermi == 1 ? n_dm : 2 * n_dm) : 0;
    const long long jcopy = (long long)std::max(n_j, 1) * b->slot_c;
    const long long kcopy = (long long)std::max(n_k, 1) * b->slot_c;
    const bool need_cond = screen || fp32_tol > 0;
    // work buffers: Cartesian densities, private J'/K' copies, dm_cond
    double *Dj, *Dk, *Jacc, *Kacc, *dmcond;
    int ncopy = ctx->ncopy;
    if (!noalloc) {
        if ((rc = ensure(ctx, &b->Dj, b->cap_dj, jcopy))) return rc;
        if ((rc = ensure(ctx, &b->J, b->cap_j, ncopy * jcopy))) return rc;
        if ((rc = ensure(ctx, &b->Dk, b->cap_dk, kcopy))) return rc;
        if ((rc = ensure(ctx, &b->K, b->cap_k, ncopy * kcopy))) return rc;
        if (need_cond && (rc = ensure(ctx, &b->dmcond, b->cap_cond, b->cond_total))) return rc;
        Dj = b->Dj, Dk = b->Dk, Jacc = b->J, Kacc = b->K, dmcond = b->dmcond;
    } else {
        // each buffer starts on a 256-byte boundary; as many accumulator copies
        // as fit (at most the context's), at least one
        auto up = [](long long n) { return (n + 31) / 32 * 32; };
        const long long cap = (long long)(MJK_NOALLOC_SCRATCH_BYTES / sizeof(double));
        auto need = [&](int nc) {
            return up(jcopy) + up(kcopy) + up(nc * jcopy) + up(nc * kcopy) + (need_cond ? up(b->cond_total) : 0);
        };
        while (ncopy > 1 && need(ncopy) > cap) --ncopy;
        if (need(ncopy) > cap)
            return fail(ctx, MJK_ENOMEM,
                        "no-allocation build needs %lld bytes of scratch; its static scratch holds %lld bytes",
                        need(1) * (long long)sizeof(double), cap * (long long)sizeof(double));
        double* q = scratch;
        Dj = q, q += up(jcopy);
        Dk = q, q += up(kcopy);
        Jacc = q, q += up(ncopy * jcopy);
        Kacc = q, q += up(ncopy * kcopy);
        dmcond = need_cond ? q : nullptr;
    }
    hipStream_t st = ctx->stream;
    std::vector<char> run = classes_to_run(ctx, b, n_j, n_k);
    if (b->class_filter >= 0)
        for (size_t c = 0; c < run.size(); ++c) run[c] = run[c] && (int)c == b->class_filter;
    const bool prof = !noalloc && b->profiling && b->pev.size() >= 2 * b->qcls.size();
    const bool tally = !noalloc && (stats || (prof && b->prof_counts));
    // The whole build on the context's stream and the streams it forks.
    // capture: recorded into a graph, without the timing events.  The
    // accumulators are cleared by a kernel, not a memset node.
    auto enqueue = [&](bool capture) -> int {
    if (!capture) HIPCHK(ctx, hipEventRecord(ctx->ev[0], st));
    if (n_j && launch_zero(Jacc, ncopy * jcopy, st)) return fail(ctx, MJK_EDEVICE, "zero fill launch failed");
    if (n_k && launch_zero(Kacc, ncopy * kcopy, st)) return fail(ctx, MJK_EDEVICE, "zero fill launch failed");
    // (the 64-bit counters are cleared as doubles: both are all-zero bits)
    if (tally && launch_zero(reinterpret_cast<double*>(b->counters), 4 * NQCLASS_MAX, st))
        return fail(ctx, MJK_EDEVICE, "zero fill launch failed");
    if (need_cond && launch_dmcond(dmcond_args(b), d_dms, b->slot_s, n_dm, dmcond, st))
        return fail(ctx, MJK_EDEVICE, "dm_cond launch failed");
    if (launch_sph2cart(xform(b), b->maxld, d_dms, b->slot_s, n_dm, hermi, with_j, with_k, Dj, Dk,
                        b->slot_c, st))
        return fail(ctx, MJK_EDEVICE, "density transform launch failed");
    if (!capture) HIPCHK(ctx, hipEventRecord(ctx->ev[1], st));
    JKArgs a;
    a.P = pairset(b);
    a.M.dbase = b->d_dbase;
    a.M.ld = b->d_ld;
    a.M.cbase = b->d_cbase;
    a.M.nsh = b->d_nsh;
    a.Dj = Dj;
    a.Dk = Dk;
    a.J = Jacc;
    a.K = Kacc;
    a.slot = b->slot_c;
    a.n_j = n_j;
    a.n_k = n_k;
    a.ncopy = ncopy;
    a.jcopy = jcopy;
    a.kcopy = kcopy;
    a.dm_cond = dmcond;
    a.cutoff = cut;
    a.screen = screen ? 1 : 0;
    a.fp32_tol = fp32_tol > 0 ? fp32_tol : 0.0;
    a.rys = ctx->rys;
    a.omega = b->omega;
    a.pmol = b->d_pmol;
    a.nbucket = NBUCKET;
    a.mol_active = b->nmol > 1 ? b->d_active : nullptr;
    a.tlist = nullptr;
    a.tcount = nullptr;
    a.thead = nullptr;
    a.ntask = 0;
    // persistent row launches: the blocks claim tasks from a counter; a batched
    // basis first compacts the tasks of its active molecules into a list
    const bool persist = ctx->persist == 2 || (ctx->persist == 1 && b->nmol > 1);
    const bool compact = persist && b->nmol > 1;
    if (persist && launch_zero(reinterpret_cast<double*>(b->d_tctl), ROW_TCTL / 2 * NQCLASS_MAX, st))
        return fail(ctx, MJK_EDEVICE, "zero fill launch failed");
    // classes, heaviest first, each on the least loaded of the streams (they all
    // add into the same J'/K' with atomics, so their order does not matter)
    const int nst = ctx->nstreams;
    hipStream_t streams[1 + mjk_ctx::NAUX] = {st, ctx->aux[0], ctx->aux[1], ctx->aux[2]};
    if (nst > 1) {
        HIPCHK(ctx, hipEventRecord(ctx->ev_fork, st));
        for (int i = 1; i < nst; ++i) HIPCHK(ctx, hipStreamWaitEvent(streams[i], ctx->ev_fork, 0));
    }
    std::vector<size_t> corder(b->qcls.size());
    std::iota(corder.begin(), corder.end(), 0);
    auto cost = [&](const QClass& q) { return ctx->sched == 0 ? q.work : q.est; };
    std::stable_sort(corder.begin(), corder.end(),
                     [&](size_t u, size_t v) { return cost(b->qcls[u]) > cost(b->qcls[v]); });
    std::vector<int> cstream(b->qcls.size(), -1);
    {
        double load[1 + mjk_ctx::NAUX] = {0, 0, 0, 0};
        for (size_t c : corder) {
            if (!run[c]) continue;
            const int si = (int)(std::min_element(load, load + nst) - load);
            load[si] += cost(b->qcls[c]);
            cstream[c] = si;
        }
    }
    // issue order (the order within each stream): the register-heavy classes (most
    // Cartesian components x roots) first, so that they hold CUs before the light
    // classes fill them (MJK_ORDER=0: by cost)
    std::vector<size_t> lorder = corder;
    if (ctx->order == 1) {
        auto heavy = [&](const QClass& q) {
            return ncart(q.la) * ncart(q.lb) * ncart(q.lc) * ncart(q.ld) * ((q.la + q.lb + q.lc + q.ld) / 2 + 1);
        };
        std::stable_sort(lorder.begin(), lorder.end(),
                         [&](size_t u, size_t v) { return heavy(b->qcls[u]) > heavy(b->qcls[v]); });
    }
    for (size_t c : lorder) {
        const QClass& qc = b->qcls[c];
        if (cstream[c] < 0) continue;
        hipStream_t st = streams[cstream[c]];
        if (prof) HIPCHK(ctx, hipEventRecord(b->pev[2 * c], st));
        a.bt = qc.bt;
        a.bt_mol = qc.bt_mol;
        a.bt_off = qc.bt_off;
        a.nbt = qc.nbt;
        a.diag = qc.GX == qc.GY;
        a.counters = tally ? b->counters + 4 * c : nullptr;   // work tallies only on request
        // the row launches' size: one block per task, or the persistent grid
        int nrow_blocks = qc.nrt;
        if (persist && qc.nrt > 0) {
            a.thead = b->d_tctl + ROW_TCTL * c + 1;
            a.ntask = qc.nrt;
            a.tcount = compact ? b->d_tctl + ROW_TCTL * c : nullptr;
            a.tlist = compact ? b->d_tlist + qc.tl0 : nullptr;
            nrow_blocks = std::min(qc.nrt, ctx->persist_blocks);
        } else {
            a.thead = nullptr;
            a.tcount = nullptr;
            a.tlist = nullptr;
            a.ntask = 0;
        }
        const int rowk = qc.mode == 2 ? 0 : row_choice(ctx, b, qc, n_j, n_k);
        const bool row_class = qc.mode == 2 || rowk > 0;
        if (compact && row_class && qc.nrt > 0 &&
            launch_compact_rows(qc.rt, qc.nrt, b->d_pmol, b->d_active, b->d_tlist + qc.tl0, b->d_tctl + ROW_TCTL * c, st))
            return fail(ctx, MJK_EDEVICE, "task compaction launch failed");
        if (qc.mode == 2) {
            // merged classes: the merged row kernel (its LDS fit was checked for this build)
            a.rt = qc.rt;
            a.rr = qc.rr;
            a.rtp = qc.rtp;
            a.rprim = qc.rprim;
            if (launch_jk_rows_gc(qc.la, qc.lb, qc.lc, qc.ld, qc.ca, qc.cb, qc.cc, qc.cd, nrow_blocks, a,
                                  row_shmem(qc, n_j, n_k, b->maxld), st))
                return fail(ctx, MJK_EDEVICE, "merged row kernel launch failed for class (%d%d|%d%d)", qc.la, qc.lb,
                            qc.lc, qc.ld);
            if (prof) b->prof_kind[c] = 3;
            if (prof) HIPCHK(ctx, hipEventRecord(b->pev[2 * c + 1], st));
            continue;
        }
        // row kernel (one block per bra pair, K' rows in LDS) when the rows fit
        int rr = -1;   // -1: no row kernel for this class (cooperative classes)
        if (rowk > 0) {
            a.rt = qc.rt;
            a.rr = qc.rr;
            a.rtp = qc.rtp;
            a.rprim = qc.rprim;
            rr = rowk == 1 ? launch_jk_rows(qc.la, qc.lb, qc.lc, qc.ld, nrow_blocks, a, row_shmem(qc, n_j, n_k, b->maxld), st)
                           : launch_jk_rows_m(qc.la, qc.lb, qc.lc, qc.ld, nrow_blocks, a,
                                              row_shmem_m(qc, n_j, n_k, b->maxld), st);
            if (rr > 0) return fail(ctx, MJK_EDEVICE, "row kernel launch failed for class (%d%d|%d%d)", qc.la, qc.lb, qc.lc, qc.ld);
        }
        if (rr != 0 && launch_jk(qc.la, qc.lb, qc.lc, qc.ld, (int)qc.nblocks, a, st))
            return fail(ctx, MJK_EUNSUPPORTED, "no kernel for class (%d%d|%d%d)", qc.la, qc.lb, qc.lc, qc.ld);
        if (prof) b->prof_kind[c] = rr == 0 ? (rowk == 2 ? 4 : 1) : coop_kernel(qc.la, qc.lb, qc.lc, qc.ld) ? 2 : 0;
        if (prof) HIPCHK(ctx, hipEventRecord(b->pev[2 * c + 1], st));
    }
    for (int i = 1; i < nst; ++i) {
        HIPCHK(ctx, hipEventRecord(ctx->ev_join[i - 1], streams[i]));
        HIPCHK(ctx, hipStreamWaitEvent(st, ctx->ev_join[i - 1], 0));
    }
    if (!capture) HIPCHK(ctx, hipEventRecord(ctx->ev[2], st));
    if (launch_cart2sph(xform(b), b->maxnao, Jacc, Kacc, b->slot_c, n_dm, hermi, with_j, with_k, d_vj, d_vk,
                        b->slot_s, ncopy, jcopy, kcopy, st))
        return fail(ctx, MJK_EDEVICE, "output transform launch failed");
    if (!capture) HIPCHK(ctx, hipEventRecord(ctx->ev[3], st));
    return MJK_OK;
    };
    // Plain builds (no statistics, no profiling) replay one HIP graph of the
    // whole build while nothing it depends on changes: one launch instead of
    // one per kernel (the launch-bound small molecules gain most).
    if (ctx->graphs && !prof && !tally && !noalloc) {
        GraphKey key;
        key.dms = d_dms;
        key.vj = with_j ? d_vj : nullptr;
        key.vk = with_k ? d_vk : nullptr;
        const void* bufs[5] = {Jacc, Kacc, Dj, Dk, dmcond};
        for (int i = 0; i < 5; ++i) key.bufs[i] = bufs[i];
        key.n_dm = n_dm;
        key.hermi = hermi;
        key.with_j = with_j;
        key.with_k = with_k;
        key.nstreams = ctx->nstreams;
        key.cut = cut;
        key.fp32_tol = fp32_tol > 0 ? fp32_tol : 0.0;
        key.row_lds = ctx->row_lds;
        key.class_filter = b->class_filter;
        const bool repeat = b->has_last_key && b->last_key == key;
        b->last_key = key;
        b->has_last_key = true;
        if (!(b->gexec && b->gkey == key) && !repeat) {
            // a build seen once runs eagerly: one-shot requests never pay for a capture
            drop_graph(b);
            if ((rc = enqueue(false))) return rc;
            return MJK_OK;
        }
        if (!(b->gexec && b->gkey == key)) {
            drop_graph(b);
            std::shared_lock<std::shared_mutex> cap(g_capture);   // no hipFree meanwhile
            HIPCHK(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            rc = enqueue(true);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(st, &g);
            if (rc == MJK_OK && e != hipSuccess)
                rc = fail(ctx, MJK_EDEVICE, "graph capture failed: %s", hipGetErrorString(e));
            if (rc == MJK_OK) {
                const hipError_t ei = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
                if (ei != hipSuccess) {
                    b->gexec = nullptr;
                    rc = fail(ctx, MJK_EDEVICE, "graph instantiation failed: %s", hipGetErrorString(ei));
                }
            }
            if (g) (void)hipGraphDestroy(g);
            if (rc != MJK_OK) return rc;
            b->gkey = key;
        }
        HIPCHK(ctx, hipGraphLaunch(b->gexec, st));
        return MJK_OK;
    }
    if ((rc = enqueue(false))) return rc;
    if (prof) {
        HIPCHK(ctx, hipStreamSynchronize(st));
        float span = 0;
        HIPCHK(ctx, hipEventElapsedTime(&span, ctx->ev[1], ctx->ev[2]));
        b->prof_span_ms += span;
        b->prof_builds += 1;
        std::vector<unsigned long long> cnt(4 * NQCLASS_MAX, 0);
        if (tally)
            HIPCHK(ctx, hipMemcpy(cnt.data(), b->counters, sizeof(unsigned long long) * cnt.size(),
                                  hipMemcpyDeviceToHost));
        for (size_t c = 0; c < b->qcls.size(); ++c) {
            if (!run[c]) continue;
            float ms = 0;
            HIPCHK(ctx, hipEventElapsedTime(&ms, b->pev[2 * c], b->pev[2 * c + 1]));
            b->prof_ms[c] += ms;
            b->prof_nq[c] += (long long)(cnt[4 * c] + cnt[4 * c + 2]);
            b->prof_np[c] += (long long)(cnt[4 * c + 1] + cnt[4 * c + 3]);
            b->prof_launches[c] += 1;
        }
    }
    if (stats) {
        HIPCHK(ctx, hipStreamSynchronize(st));
        float t01 = 0, t12 = 0, t23 = 0;
        HIPCHK(ctx, hipEventElapsedTime(&t01, ctx->ev[0], ctx->ev[1]));
        HIPCHK(ctx, hipEventElapsedTime(&t12, ctx->ev[1], ctx->ev[2]));
        HIPCHK(ctx, hipEventElapsedTime(&t23, ctx->ev[2], ctx->ev[3]));
        std::vector<unsigned long long> cnt(4 * NQCLASS_MAX);
        HIPCHK(ctx, hipMemcpy(cnt.data(), b->counters, sizeof(unsigned long long) * cnt.size(),
                              hipMemcpyDeviceToHost));
        stats->ms_quartets = t12;
        stats->ms_transform = t01 + t23;
        stats->quartets = 0;
        stats->prim_quartets = 0;
        stats->flops = 0;
        stats->quartets_fp32 = 0;
        stats->flops_fp32 = 0;
        for (size_t c = 0; c < b->qcls.size(); ++c) {
            const QClass& qc = b->qcls[c];
            const int nco = qc.ca * qc.cb * qc.cc * qc.cd;
            const double f64 =
                flops_class(qc.la, qc.lb, qc.lc, qc.ld, (double)cnt[4 * c], (double)cnt[4 * c + 1], nco);
            const double f32 =
                flops_class(qc.la, qc.lb, qc.lc, qc.ld, (double)cnt[4 * c + 2], (double)cnt[4 * c + 3], nco);
            stats->quartets += (long long)(cnt[4 * c] + cnt[4 * c + 2]);
            stats->prim_quartets += (long long)(cnt[4 * c + 1] + cnt[4 * c + 3]);
            stats->quartets_fp32 += (long long)cnt[4 * c + 2];
            stats->flops += f64 + f32;
            stats->flops_fp32 += f32;
        }
        stats->ms_total =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return MJK_OK;
}

int mjk_get_jk_device(mjk_ctx* ctx, mjk_basis* b, const double* d_dms, int n_dm, int hermi, int with_j,
                      int with_k, double direct_scf_tol, double fp32_tol, double* d_vj, double* d_vk,
                      mjk_stats* stats)
{
    return get_jk_impl(ctx, b, d_dms, n_dm, hermi, with_j, with_k, direct_scf_tol, fp32_tol, d_vj, d_vk, stats,
                       nullptr);
}

int mjk_get_jk_device_noalloc(mjk_ctx* ctx, mjk_basis* b, const double* d_dms, int n_dm, int hermi, int with_j,
                              int with_k, double direct_scf_tol, double fp32_tol, double* d_vj, double* d_vk)
{
    if (!ctx) return MJK_EINVAL;
    std::string why;
    double* ptr = scratch_address(ctx->device, &why);
    if (!ptr) return fail(ctx, MJK_EDEVICE, "no scratch on device %d: %s", ctx->device, why.c_str());
    return get_jk_impl(ctx, b, d_dms, n_dm, hermi, with_j, with_k, direct_scf_tol, fp32_tol, d_vj, d_vk, nullptr,
                       ptr);
}

int mjk_get_jk(mjk_ctx* ctx, mjk_basis* b, const double* dms, int n_dm, int hermi, int with_j, int with_k,
               double direct_scf_tol, double fp32_tol, double* vj, double* vk, mjk_stats* stats)
{
    if (!ctx || !b || !dms || n_dm <= 0) return fail(ctx, MJK_EINVAL, "bad arguments to mjk_get_jk");
    auto t0 = std::chrono::steady_clock::now();
    int rc;
    std::lock_guard<std::recursive_mutex> outer(ctx->mu);
    // A batch with inactive molecules moves only the active molecules' blocks
    // over PCIe (runs of consecutive active molecules, one copy per run and
    // density): the inactive molecules' blocks of dms are not read, and their
    // blocks of vj/vk are not written.
    std::vector<std::pair<long long, long long>> runs;   // {offset, length} within a slot
    {
        const int nm = b->nmol;
        bool partial = false;
        for (int m = 0; m < nm && nm > 1; ++m) partial = partial || !b->h_active[m];
        if (partial) {
            for (int m = 0; m < nm;) {
                if (!b->h_active[m]) {
                    ++m;
                    continue;
                }
                int e = m;
                while (e < nm && b->h_active[e]) ++e;
                const long long lo = b->h_sbase[m];
                const long long hi = b->h_sbase[e - 1] + (long long)b->h_nao[e - 1] * b->h_nao[e - 1];
                runs.push_back({lo, hi - lo});
                m = e;
            }
        } else {
            runs.push_back({0, b->slot_s});
        }
    }
    const bool partial = !(runs.size() == 1 && runs[0].first == 0 && runs[0].second == b->slot_s);
    {
        HIPCHK(ctx, hipSetDevice(ctx->device));
        const long long n = (long long)n_dm * b->slot_s;
        if ((rc = ensure(ctx, &b->Dsph, b->cap_ds, n))) return rc;
        if (with_j && (rc = ensure(ctx, &b->Vj, b->cap_vj, n))) return rc;
        if (with_k && (rc = ensure(ctx, &b->Vk, b->cap_vk, n))) return rc;
        if (!partial) {
            HIPCHK(ctx, hipMemcpyAsync(b->Dsph, dms, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
        } else {
            for (int s = 0; s < n_dm; ++s)
                for (const auto& r : runs)
                    HIPCHK(ctx, hipMemcpyAsync(b->Dsph + s * b->slot_s + r.first, dms + s * b->slot_s + r.first,
                                               sizeof(double) * r.second, hipMemcpyHostToDevice, ctx->stream));
        }
    }
    rc = mjk_get_jk_device(ctx, b, b->Dsph, n_dm, hermi, with_j, with_k, direct_scf_tol, fp32_tol,
                           with_j ? b->Vj : nullptr, with_k ? b->Vk : nullptr, stats);
    if (rc) return rc;
    const long long n = (long long)n_dm * b->slot_s;
    if (!partial) {
        if (with_j) HIPCHK(ctx, hipMemcpyAsync(vj, b->Vj, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
        if (with_k) HIPCHK(ctx, hipMemcpyAsync(vk, b->Vk, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    } else {
        for (int s = 0; s < n_dm; ++s)
            for (const auto& r : runs) {
                const long long o = s * b->slot_s + r.first;
                if (with_j)
                    HIPCHK(ctx, hipMemcpyAsync(vj + o, b->Vj + o, sizeof(double) * r.second, hipMemcpyDeviceToHost,
                                               ctx->stream));
                if (with_k)
                    HIPCHK(ctx, hipMemcpyAsync(vk + o, b->Vk + o, sizeof(double) * r.second, hipMemcpyDeviceToHost,
                                               ctx->stream));
            }
    }
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (stats)
        stats->ms_total =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return MJK_OK;
}

int mjk_get_jk_resident(mjk_ctx* ctx, mjk_basis* b, int n_dm, int hermi, int with_j, int with_k,
                        double direct_scf_tol, double fp32_tol, mjk_stats* stats)
{
    if (!ctx || !b || b->ctx != ctx || n_dm <= 0) return fail(ctx, MJK_EINVAL, "bad arguments to mjk_get_jk_resident");
    std::lock_guard<std::recursive_mutex> outer(ctx->mu);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const long long n = (long long)n_dm * b->slot_s;
    if (!b->Dsph || b->cap_ds < n)
        return fail(ctx, MJK_EINVAL, "no resident density of %d slots: mjk_get_jk uploads one", n_dm);
    int rc;
    if (with_j && (rc = ensure(ctx, &b->Vj, b->cap_vj, n))) return rc;
    if (with_k && (rc = ensure(ctx, &b->Vk, b->cap_vk, n))) return rc;
    rc = mjk_get_jk_device(ctx, b, b->Dsph, n_dm, hermi, with_j, with_k, direct_scf_tol, fp32_tol,
                           with_j ? b->Vj : nullptr, with_k ? b->Vk : nullptr, stats);
    if (rc) return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return MJK_OK;
}

int mjk_basis_fetch_jk(mjk_ctx* ctx, mjk_basis* b, int n_dm, double* vj, double* vk)
{
    if (!ctx || !b || b->ctx != ctx || n_dm <= 0) return fail(ctx, MJK_EINVAL, "bad arguments to mjk_basis_fetch_jk");
    std::lock_guard<std::recursive_mutex> outer(ctx->mu);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const long long n = (long long)n_dm * b->slot_s;
    if ((vj && (!b->Vj || b->cap_vj < n)) || (vk && (!b->Vk || b->cap_vk < n)))
        return fail(ctx, MJK_EINVAL, "no resident J/K of %d slots", n_dm);
    if (vj) HIPCHK(ctx, hipMemcpyAsync(vj, b->Vj, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    if (vk) HIPCHK(ctx, hipMemcpyAsync(vk, b->Vk, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return MJK_OK;
}

int mjk_basis_set_class_filter(mjk_ctx* ctx, mjk_basis* b, int cls)
{
    if (!ctx || !b || b->ctx != ctx) return fail(ctx, MJK_EINVAL, "context/basis mismatch");
    std::lock_guard<std::recursive_mutex> lock(ctx->mu);
    if (cls < -1 || cls >= (int)b->qcls.size())
        return fail(ctx, MJK_EINVAL, "class %d out of range (-1 .. %d)", cls, (int)b->qcls.size() - 1);
    b->class_filter = cls;
    return MJK_OK;
}

int mjk_ctx_set_tuning(mjk_ctx* ctx, int n, const int* keys, const double* scale)
{
    if (!ctx || n < 0 || (n > 0 && (!keys || !scale))) return fail(ctx, MJK_EINVAL, "bad arguments to mjk_ctx_set_tuning");
    std::map<std::array<int, 9>, std::tuple<int, int, double>> t;
    for (int i = 0; i < n; ++i) {
        const int* r = keys + 11 * i;
        for (int j = 0; j < 4; ++j)
            if (r[j] < 0 || r[j] > LMAX || r[4 + j] < 1 || r[4 + j] > 2)
                return fail(ctx, MJK_EINVAL, "tuning entry %d: bad class key", i);
        if (r[9] < 0 || r[9] > 2 || r[10] < 0 || r[10] > 4096 || !(scale[i] >= 0 && scale[i] <= 1e6))
            return fail(ctx, MJK_EINVAL, "tuning entry %d: bad setting", i);
        t[{r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8]}] = std::make_tuple(r[9], r[10], scale[i]);
    }
    std::lock_guard<std::recursive_mutex> lock(ctx->mu);
    ctx->tune.swap(t);
    ++ctx->tune_gen;   // every basis rebuilds its task lists at its next build
    return MJK_OK;
}

int mjk_ctx_tuning_entries(mjk_ctx* ctx)
{
    if (!ctx) return MJK_EINVAL;
    std::lock_guard<std::recursive_mutex> lock(ctx->mu);
    return (int)ctx->tune.size();
}

int mjk_basis_class_key(mjk_ctx* ctx, const mjk_basis* b, int k, int* key9)
{
    if (!ctx || !b || b->ctx != ctx || k < 0 || k >= (int)b->qcls.size() || !key9)
        return fail(ctx, MJK_EINVAL, "bad arguments to mjk_basis_class_key");
    const QClass& q = b->qcls[k];
    const int v[9] = {q.la, q.lb, q.lc, q.ld, q.ca, q.cb, q.cc, q.cd, q.bucket};
    for (int i = 0; i < 9; ++i) key9[i] = v[i];
    return MJK_OK;
}

int mjk_basis_nclasses(const mjk_basis* b)
{
    return b ? (int)b->qcls.size() : MJK_EINVAL;
}

int mjk_basis_class(mjk_ctx* ctx, const mjk_basis* b, int k, int n_dm, int hermi, int* l4, int* ctr4,
                    long long* nblocks, int* kernel)
{
    if (!ctx || !b || b->ctx != ctx || k < 0 || k >= (int)b->qcls.size() || !l4 || !ctr4 || !nblocks || !kernel ||
        n_dm <= 0)
        return fail(ctx, MJK_EINVAL, "bad arguments to mjk_basis_class");
    const QClass& qc = b->qcls[k];
    l4[0] = qc.la, l4[1] = qc.lb, l4[2] = qc.lc, l4[3] = qc.ld;
    ctr4[0] = qc.ca, ctr4[1] = qc.cb, ctr4[2] = qc.cc, ctr4[3] = qc.cd;
    *nblocks = qc.nblocks;
    // the choice a build of J and K of n_dm densities makes
    const int n_j = n_dm, n_k = hermi == 1 ? n_dm : 2 * n_dm;
    const std::vector<char> run = classes_to_run(ctx, b, n_j, n_k);
    if (!run[k]) {
        *kernel = -1;
    } else if (qc.mode == 2) {
        *kernel = 3;
    } else {
        const int rowk = row_choice(ctx, b, qc, n_j, n_k);
        *kernel = rowk == 1 ? 1 : rowk == 2 ? 4 : coop_kernel(qc.la, qc.lb, qc.lc, qc.ld) ? 2 : 0;
    }
    return MJK_OK;
}

int mjk_profile(mjk_ctx* ctx, mjk_basis* b, int enable)
{
    if (!ctx || !b) return MJK_EINVAL;
    std::lock_guard<std::recursive_mutex> lock(ctx->mu);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const size_t n = NQCLASS_MAX;
    if (b->pev.empty()) {
        b->pev.resize(2 * n);
        for (auto& e : b->pev) HIPCHK(ctx, hipEventCreate(&e));
    }
    b->prof_ms.assign(n, 0.0);
    b->prof_nq.assign(n, 0);
    b->prof_np.assign(n, 0);
    b->prof_launches.assign(n, 0);
    b->prof_kind.assign(n, -1);
    b->prof_span_ms = 0;
    b->prof_builds = 0;
    if (enable < 0 || enable > 2) return fail(ctx, MJK_EINVAL, "profile mode must be 0, 1 or 2");
    b->profiling = enable != 0;
    b->prof_counts = enable == 1;
    return MJK_OK;
}

int mjk_profile_read(mjk_ctx* ctx, const mjk_basis* b, int max_classes, int* cls, double* ms, long long* nq,
                     long long* nprimq, double* flops, long long* launches)
{
    if (!ctx || !b) return MJK_EINVAL;
    int k = 0;
    for (size_t c = 0; c < b->qcls.size() && c < b->prof_ms.size(); ++c) {
        if (b->prof_launches[c] == 0) continue;
        if (k >= max_classes) break;
        const QClass& qc = b->qcls[c];
        cls[4 * k] = qc.la;
        cls[4 * k + 1] = qc.lb;
        cls[4 * k + 2] = qc.lc;
        cls[4 * k + 3] = qc.ld;
        ms[k] = b->prof_ms[c];
        nq[k] = b->prof_nq[c];
        nprimq[k] = b->prof_np[c];
        flops[k] = flops_class(qc.la, qc.lb, qc.lc, qc.ld, (double)b->prof_nq[c], (double)b->prof_np[c],
                               qc.ca * qc.cb * qc.cc * qc.cd);
        launches[k] = b->prof_launches[c];
        ++k;
    }
    return k;
}

int mjk_profile_span(mjk_ctx* ctx, const mjk_basis* b, double* ms, long long* builds)
{
    if (!ctx || !b || !ms || !builds) return MJK_EINVAL;
    *ms = b->prof_span_ms;
    *builds = b->prof_builds;
    return MJK_OK;
}

int mjk_profile_kinds(mjk_ctx* ctx, const mjk_basis* b, int max_classes, int* kind)
{
    if (!ctx || !b || !kind) return MJK_EINVAL;
    int k = 0;
    for (size_t c = 0; c < b->qcls.size() && c < b->prof_ms.size(); ++c) {
        if (b->prof_launches[c] == 0) continue;
        if (k >= max_classes) break;
        kind[k++] = b->prof_kind[c];
    }
    return k;
}

int mjk_device_alloc(mjk_ctx* ctx, size_t bytes, void** ptr)
{
    if (!ctx || !ptr) return MJK_EINVAL;
    std::lock_guard<std::recursive_mutex> lock(ctx->mu);
    HIPCHK(ctx, hipSetDevice(ctx->device));
    return dalloc(ctx, bytes, ptr);
}

int mjk_device_free(mjk_ctx* ctx, void* ptr)
{
    if (!ctx) return MJK_EINVAL;
    std::lock_guard<std::recursive_mutex> lock(ctx->mu);
    dfree(ctx, ptr);
    return MJK_OK;
}

int mjk_memcpy_h2d(mjk_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    if (!ctx) return MJK_EINVAL;
    HIPCHK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return MJK_OK;
}

int mjk_memcpy_d2h(mjk_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    if (!ctx) return MJK_EINVAL;
    HIPCHK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return MJK_OK;
}

int mjk_device_copy(mjk_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    if (!ctx || (bytes % sizeof(double)) != 0) return MJK_EINVAL;
    if (launch_copy((const double*)src, (double*)dst, (long long)(bytes / sizeof(double)), ctx->stream))
        return fail(ctx, MJK_EDEVICE, "copy launch failed");
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return MJK_OK;
}

int mjk_mem_info(int device, size_t* free_bytes, size_t* total_bytes)
{
    if (!free_bytes || !total_bytes) return MJK_EINVAL;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return MJK_EDEVICE;
    if (hipSetDevice(device) != hipSuccess) return MJK_EDEVICE;
    const hipError_t e = hipMemGetInfo(free_bytes, total_bytes);
    (void)hipSetDevice(cur);
    return e == hipSuccess ? MJK_OK : MJK_EDEVICE;
}

int mjk_synchronize(mjk_ctx* ctx)
{
    if (!ctx) return MJK_EINVAL;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return MJK_OK;
}

int mjk_c2s(int l, double* out)
{
    if (l < 0 || l > 4 || !out) return MJK_EINVAL;
    if (l > C2S_MAXL) {
        solid_harmonic_rows(l, out);
        return MJK_OK;
    }
    // l <= 1: identity blocks here, the angular factors sit in the contraction coefficients
    const double f = l == 0 ? C2S_S0 : l == 1 ? C2S_P1 : 1.0;
    for (int m = 0; m < nsph(l); ++m)
        for (int c = 0; c < ncart(l); ++c) out[m * ncart(l) + c] = f * C2S[l][m][c];
    return MJK_OK;
}

int mjk_rys_roots(int n, double T, double* x, double* w)
{
    if (n < 1 || n > RYS_NMAX || !(T >= 0) || !x || !w) return MJK_EINVAL;
    static std::vector<double> tab;
    if (tab.empty()) {
        tab.resize(RYS_BUF_SIZE);
        rys_fill_buffer(tab.data());
    }
    // the kernels' cursor over the table in memory and over the padded layout
    // the row kernels copy to LDS must agree bit for bit (same coefficients,
    // same evaluation order)
    switch (n) {
#define R(N)                                                          \
    case N: {                                                         \
        double xx[N], ww[N];                                          \
        rys_roots<N>(tab.data(), T, xx, ww);                          \
        std::vector<double> lt(RysLds<N>::SIZE);                      \
        rys_lds_fill<N>(lt.data(), tab.data(), 0, 1);                 \
        const RysCursorLds<N> cur(lt.data(), T);                      \
        const RysCursor<N> ref(tab.data(), T);                        \
        for (int i = 0; i < N; ++i) {                                 \
            double xl, wl, xg, wg;                                    \
            cur.root(i, xl, wl);                                      \
            ref.root(i, xg, wg);                                      \
            if (xl != xg || wl != wg) return MJK_EINTERNAL;           \
            x[i] = xx[i];                                             \
            w[i] = ww[i];                                             \
        }                                                             \
        return MJK_OK;                                                \
    }
        R(1) R(2) R(3) R(4) R(5) R(6) R(7) R(8) R(9)
#undef R
    }
    return MJK_EINVAL;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// One-electron integrals on the host for the SCF caller (overlap, kinetic,
// nuclear attraction), spherical AOs in PySCF order.  Obara-Saika 1D overlap
// recurrences; nuclear attraction by Rys quadrature with the same root
// table as the ERI kernels: V = -Z 2pi/p K_ab sum_i w_i Ix Iy Iz with
// C00 = (P-A) - t^2 (P-C), B10 = (1-t^2)/(2p).
namespace {

void overlap_1d(double S[8][8], int la, int lb, double p, double PA, double PB, double s00)
{
    const double h = 0.5 / p;
    S[0][0] = s00;
    for (int a = 0; a <= la; ++a)
        for (int b = 0; b <= lb; ++b) {
            if (a == 0 && b == 0) continue;
            double v;
            if (a > 0) {
                v = PA * S[a - 1][b];
                if (a > 1) v += h * (a - 1) * S[a - 2][b];
                if (b > 0) v += h * b * S[a - 1][b - 1];
            } else {
                v = PB * S[a][b - 1];
                if (b > 1) v += h * (b - 1) * S[a][b - 2];
            }
            S[a][b] = v;
        }
}

template <int N>
void nuc_roots(const double* tab, double T, double* x, double* w)
{
    double xx[N], ww[N];
    rys_roots<N>(tab, T, xx, ww);
    for (int i = 0; i < N; ++i) {
        x[i] = xx[i];
        w[i] = ww[i];
    }
}

void roots_n(const double* tab, int n, double T, double* x, double* w)
{
    switch (n) {
    case 1: nuc_roots<1>(tab, T, x, w); break;
    case 2: nuc_roots<2>(tab, T, x, w); break;
    case 3: nuc_roots<3>(tab, T, x, w); break;
    case 4: nuc_roots<4>(tab, T, x, w); break;
    default: nuc_roots<5>(tab, T, x, w); break;
    }
}

int cart_xyz(int l, int i, int& lx, int& ly)
{
    lx = l;
    int k = i;
    while (k > l - lx) { k -= l - lx + 1; --lx; }
    ly = l - lx - k;
    return l - lx - ly;
}

}  // namespace

extern "C" int mjk_int1e(int which, const int* atm, int natm, const int* bas, int nbas, const double* env,
                         double* out)
{
    if (which < 0 || which > 2 || !out) return MJK_EINVAL;
    MolHost m;
    if (parse_molecule(nullptr, atm, natm, bas, nbas, env, m) != MJK_OK) return MJK_EINVAL;
    static std::vector<double> tab;
    static std::mutex tab_mu;
    {
        std::lock_guard<std::mutex> lk(tab_mu);
        if (tab.empty()) {
            tab.resize(RYS_BUF_SIZE);
            rys_fill_buffer(tab.data());
        }
    }
    const int nao = m.nao;
    // the three matrices are symmetric: blocks j <= i, mirrored; shell rows
    // interleaved over host threads (each block written by one thread)
    auto shell_row = [&](int i) {
        for (int j = 0; j <= i; ++j) {
            const GShell& A = m.sh[i];
            const GShell& B = m.sh[j];
            const int la = A.l, lb = B.l, na = ncart(la), nb = ncart(lb);
            double cart[10][10];
            for (auto& r : cart) for (double& v : r) v = 0.0;
            double AB2 = 0;
            for (int d = 0; d < 3; ++d) AB2 += (A.r[d] - B.r[d]) * (A.r[d] - B.r[d]);
            for (int ia = 0; ia < A.nprim; ++ia)
                for (int ib = 0; ib < B.nprim; ++ib) {
                    const double a = A.e[ia], b = B.e[ib], p = a + b;
                    const double cc = A.c[ia] * B.c[ib];
                    const double K = std::exp(-a * b / p * AB2);
                    double P[3];
                    for (int d = 0; d < 3; ++d) P[d] = (a * A.r[d] + b * B.r[d]) / p;
                    if (which <= 1) {
                        double S[3][8][8];
                        for (int d = 0; d < 3; ++d)
                            overlap_1d(S[d], la + 1, lb + 2, p, P[d] - A.r[d], P[d] - B.r[d],
                                       std::sqrt(M_PI / p) * (d == 0 ? K : 1.0));
                        for (int x = 0; x < na; ++x)
                            for (int y = 0; y < nb; ++y) {
                                int e1[3], e2[3];
                                e1[2] = cart_xyz(la, x, e1[0], e1[1]);
                                e2[2] = cart_xyz(lb, y, e2[0], e2[1]);
                                double v;
                                if (which == 0) {
                                    v = S[0][e1[0]][e2[0]] * S[1][e1[1]][e2[1]] * S[2][e1[2]][e2[2]];
                                } else {
                                    v = 0.0;
                                    for (int d = 0; d < 3; ++d) {
                                        const int u = e1[d], t = e2[d];
                                        double kd = -2.0 * b * (2 * t + 1) * S[d][u][t] + 4.0 * b * b * S[d][u][t + 2];
                                        if (t >= 2) kd += t * (t - 1) * S[d][u][t - 2];
                                        kd *= -0.5;
                                        double o = 1.0;
                                        for (int d2 = 0; d2 < 3; ++d2)
                                            if (d2 != d) o *= S[d2][e1[d2]][e2[d2]];
                                        v += kd * o;
                                    }
                                }
                                cart[x][y] += cc * v;
                            }
                    } else {
                        const int nr = (la + lb) / 2 + 1;
                        for (int c = 0; c < natm; ++c) {
                            const double* C = env + atm[c * ATM_SLOTS + PTR_COORD];
                            const double Z = atm[c * ATM_SLOTS];
                            double PC[3], PC2 = 0;
                            for (int d = 0; d < 3; ++d) {
                                PC[d] = P[d] - C[d];
                                PC2 += PC[d] * PC[d];
                            }
                            double x[9], w[9];
                            roots_n(tab.data(), nr, p * PC2, x, w);
                            const double fac = -Z * 2.0 * M_PI / p * K * cc;
                            for (int r = 0; r < nr; ++r) {
                                double I[3][4][4];
                                for (int d = 0; d < 3; ++d) {
                                    double g[8];
                                    const double C00 = P[d] - A.r[d] - x[r] * PC[d];
                                    const double B10 = 0.5 / p * (1.0 - x[r]);
                                    g[0] = (d == 2) ? w[r] * fac : 1.0;
                                    if (la + lb > 0) g[1] = C00 * g[0];
                                    for (int k = 1; k < la + lb; ++k) g[k + 1] = C00 * g[k] + k * B10 * g[k - 1];
                                    double h[8][4];
                                    for (int k = 0; k <= la + lb; ++k) h[k][0] = g[k];
                                    for (int t = 1; t <= lb; ++t)
                                        for (int k = 0; k <= la + lb - t; ++k)
                                            h[k][t] = h[k + 1][t - 1] + (A.r[d] - B.r[d]) * h[k][t - 1];
                                    for (int u = 0; u <= la; ++u)
                                        for (int t = 0; t <= lb; ++t) I[d][u][t] = h[u][t];
                                }
                                for (int xx = 0; xx < na; ++xx)
                                    for (int y = 0; y < nb; ++y) {
                                        int e1[3], e2[3];
                                        e1[2] = cart_xyz(la, xx, e1[0], e1[1]);
                                        e2[2] = cart_xyz(lb, y, e2[0], e2[1]);
                                        cart[xx][y] += I[0][e1[0]][e2[0]] * I[1][e1[1]][e2[1]] * I[2][e1[2]][e2[2]];
                                    }
                            }
                        }
                    }
                }
            for (int mi = 0; mi < nsph(la); ++mi)
                for (int mj = 0; mj < nsph(lb); ++mj) {
                    double v = 0.0;
                    for (int x = 0; x < na; ++x) {
                        const double c1 = C2S[la][mi][x];
                        if (c1 == 0.0) continue;
                        for (int y = 0; y < nb; ++y) v += c1 * C2S[lb][mj][y] * cart[x][y];
                    }
                    out[(size_t)(A.s0 + mi) * nao + B.s0 + mj] = v;
                    out[(size_t)(B.s0 + mj) * nao + A.s0 + mi] = v;
                }
        }
    };
    const int nt = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t)
        pool.emplace_back([&, t] {
            for (int i = t; i < m.nsh; i += nt) shell_row(i);
        });
    for (auto& th : pool) th.join();
    return MJK_OK;
}
