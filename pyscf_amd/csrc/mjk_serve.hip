// Request server: continuous batching of J/K builds and whole RHF solves on
// one device, behind mjk_submit / mjk_poll.
//
// * nslots worker threads, each owning a context (its own stream of default
//   priority); requests that run at the same time overlap on the device.
// * A request is copied at submission (molecule arrays, densities) and runs
//   to completion on one slot: shell pairs, Schwarz bounds, densities and, for
//   an RHF, the whole SCF state stay resident in that slot's context for the
//   request's life (mjk_rhf_*: D, F, the DIIS history and C never leave HBM).
// * Admission in arrival order: the head of the queue starts when a slot is
//   free and the device bytes the contexts have tracked for their running
//   requests (each request counted at least at its estimate until its
//   buffers exist), plus the head's estimate, stay within the server's budget.
//   A request whose estimate alone exceeds the budget is refused at
//   submission.
// * Results stay in the server until mjk_result collects them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "mjk.h"

namespace {

constexpr int ATM_SLOTS = 6, BAS_SLOTS = 8;     // pyscf/gto/mole.py:80-101
constexpr int CHARGE_OF = 0, PTR_COORD = 1;
constexpr int ANG_OF = 1, NPRIM_OF = 2, NCTR_OF = 3;

struct Req {
    long long ticket = 0;
    int kind = 0;
    std::vector<int> atm, bas;
    std::vector<double> env, dm;
    int natm = 0, nbas = 0, n_dm = 0, hermi = 1, with_j = 1, with_k = 1, nelectron = 0, max_cycle = 0;
    double direct_scf_tol = 0, fp32_tol = 0, conv_tol = 0;
    int nao = 0;
    size_t estimate = 0;
    // results
    int state = MJK_REQ_QUEUED;
    std::vector<double> vj, vk, mo_energy;
    double e_tot = 0;
    int cycles = 0, converged = 0;
    double t_submit = 0, t_start = 0, t_done = 0;
    size_t peak = 0;
    int slot = -1;
    std::string error;
};

struct Slot {
    mjk_ctx* ctx = nullptr;
    size_t idle_bytes = 0;      // what the context holds with no request (Rys table)
    std::shared_ptr<Req> running;
    std::thread th;
};

}  // namespace

struct mjk_server {
    int device = 0;
    size_t budget = 0;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::shared_ptr<Req>> queue;
    std::map<long long, std::shared_ptr<Req>> reqs;   // submitted, not yet collected
    std::deque<long long> finished;                   // finished, not yet returned by mjk_wait_any
    long long next_ticket = 1;
    bool closing = false;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
};

namespace {

double now(const mjk_server* s)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - s->t0).count();
}

int nao_of(const std::vector<int>& bas, int nbas)
{
    int n = 0;
    for (int i = 0; i < nbas; ++i) n += (2 * bas[i * BAS_SLOTS + ANG_OF] + 1) * bas[i * BAS_SLOTS + NCTR_OF];
    return n;
}

// Upper estimate of the device bytes a request holds: primitive pairs (split
// and merged representations), shell pairs, quartet task lists, the
// Cartesian arenas (densities, the accumulator copies), dm_cond and, for an
// RHF, the SCF state.
size_t estimate_bytes(const Req& r)
{
    double sp = 0, sp2 = 0;
    long long nsh = 0, ncart = 0;
    for (int i = 0; i < r.nbas; ++i) {
        const int l = r.bas[i * BAS_SLOTS + ANG_OF], np = r.bas[i * BAS_SLOTS + NPRIM_OF],
                  nc = r.bas[i * BAS_SLOTS + NCTR_OF];
        sp += (double)np * nc;
        sp2 += (double)np * np * nc;
        nsh += nc;
        ncart += (long long)nc * (l + 1) * (l + 2) / 2;
    }
    const double nprim_pairs = 2.0 * (sp * sp + sp2) / 2;     // split + merged copies
    const double npairs = 2.0 * (double)nsh * (nsh + 1) / 2;
    const double n = r.nao, nc = (double)ncart;
    const int ndm = std::max(r.n_dm, 1);
    double bytes = 48 * nprim_pairs + 160 * npairs + 16.0 * 55 * 16 * (npairs / 8 + 1) * 2;
    bytes += 8.0 * (ndm * (3 + 2 * 8) * nc * nc + 3.0 * ndm * n * n + (double)nsh * nsh);
    // the SCF state: 15 n^2 matrices, two histories of 8, the solver workspace
    // (at least 4 n^2 doubles and 32 MiB), small vectors
    if (r.kind == MJK_REQ_RHF) bytes += 8.0 * (15 + 2 * 8 + 4 + 1) * n * n + (double)(32 << 20);
    return (size_t)bytes + ((size_t)16 << 20);
}

double nuclear_repulsion(const Req& r)
{
    double e = 0;
    for (int i = 0; i < r.natm; ++i)
        for (int j = 0; j < i; ++j) {
            const double* a = &r.env[r.atm[i * ATM_SLOTS + PTR_COORD]];
            const double* b = &r.env[r.atm[j * ATM_SLOTS + PTR_COORD]];
            const double d = std::sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) +
                                       (a[2] - b[2]) * (a[2] - b[2]));
            e += (double)r.atm[i * ATM_SLOTS + CHARGE_OF] * r.atm[j * ATM_SLOTS + CHARGE_OF] / d;
        }
    return e;
}

// bytes the running requests hold: tracked, but at least their estimates
size_t held_locked(const mjk_server* s)
{
    size_t tot = 0;
    for (const Slot& sl : s->slots)
        if (sl.running) {
            const size_t used = mjk_ctx_hbm_used(sl.ctx);
            const size_t tracked = used > sl.idle_bytes ? used - sl.idle_bytes : 0;
            tot += std::max(tracked, sl.running->estimate);
        }
    return tot;
}

std::string ctx_error(mjk_ctx* ctx, int rc, const char* what)
{
    char buf[600];
    snprintf(buf, sizeof(buf), "%s failed (code %d): %s", what, rc, mjk_last_error(ctx));
    return buf;
}

// runs one request on a slot's context; returns an error text or ""
std::string execute(mjk_ctx* ctx, Req& r, size_t idle)
{
    mjk_basis* b = nullptr;
    int rc = mjk_basis_upload(ctx, r.atm.data(), r.natm, r.bas.data(), r.nbas, r.env.data(), r.direct_scf_tol, &b);
    if (rc) return ctx_error(ctx, rc, "mjk_basis_upload");
    std::string err;
    const size_t nn = (size_t)r.nao * r.nao;
    if (r.kind == MJK_REQ_JK) {
        if (r.with_j) r.vj.assign(nn * r.n_dm, 0.0);
        if (r.with_k) r.vk.assign(nn * r.n_dm, 0.0);
        rc = mjk_get_jk(ctx, b, r.dm.data(), r.n_dm, r.hermi, r.with_j, r.with_k, r.direct_scf_tol, r.fp32_tol,
                        r.with_j ? r.vj.data() : nullptr, r.with_k ? r.vk.data() : nullptr, nullptr);
        if (rc) err = ctx_error(ctx, rc, "mjk_get_jk");
    } else {
        std::vector<double> s1e(nn), t(nn), v(nn);
        rc = mjk_int1e_device(ctx, r.atm.data(), r.natm, r.bas.data(), r.nbas, r.env.data(), s1e.data(), t.data(),
                              v.data(), 0);
        if (rc) {
            err = ctx_error(ctx, rc, "mjk_int1e_device");
        } else {
            for (size_t i = 0; i < nn; ++i) t[i] += v[i];
            mjk_scf* scf = nullptr;
            rc = mjk_rhf_create(ctx, b, s1e.data(), t.data(), r.nelectron / 2, nuclear_repulsion(r), 8, &scf);
            if (rc) {
                err = ctx_error(ctx, rc, "mjk_rhf_create");
            } else {
                mjk_rhf_result res{};
                const double grad = std::sqrt(r.conv_tol);
                rc = mjk_rhf_run(ctx, scf, nullptr, r.max_cycle, r.conv_tol, grad, r.direct_scf_tol, r.fp32_tol, 0.0,
                                 &res);
                if (rc) {
                    err = ctx_error(ctx, rc, "mjk_rhf_run");
                } else {
                    r.e_tot = res.e_tot;
                    r.cycles = res.cycles;
                    r.converged = res.converged;
                    r.dm.assign(nn, 0.0);
                    r.mo_energy.assign(r.nao, 0.0);
                    rc = mjk_rhf_fetch(ctx, scf, r.dm.data(), r.mo_energy.data(), nullptr, nullptr);
                    if (rc) err = ctx_error(ctx, rc, "mjk_rhf_fetch");
                }
                const size_t used = mjk_ctx_hbm_used(ctx);
                r.peak = std::max(r.peak, used > idle ? used - idle : 0);
                mjk_rhf_destroy(scf);
            }
        }
    }
    const size_t used = mjk_ctx_hbm_used(ctx);
    r.peak = std::max(r.peak, used > idle ? used - idle : 0);
    mjk_basis_destroy(b);
    return err;
}

void worker(mjk_server* s, int k)
{
    Slot& sl = s->slots[k];
    for (;;) {
        std::shared_ptr<Req> r;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] {
                return (s->closing && s->queue.empty()) ||
                       (!s->queue.empty() && held_locked(s) + s->queue.front()->estimate <= s->budget);
            });
            if (s->queue.empty()) return;
            r = s->queue.front();
            s->queue.pop_front();
            sl.running = r;
            r->state = MJK_REQ_RUNNING;
            r->slot = k;
            r->t_start = now(s);
        }
        std::string err = execute(sl.ctx, *r, sl.idle_bytes);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            r->t_done = now(s);
            r->error = err;
            r->state = err.empty() ? MJK_REQ_DONE : MJK_REQ_FAILED;
            if (!err.empty()) {
                r->vj.clear();
                r->vk.clear();
            }
            if (r->kind == MJK_REQ_JK) std::vector<double>().swap(r->dm);   // the input densities
            sl.running.reset();
            s->finished.push_back(r->ticket);
        }
        s->cv.notify_all();
    }
}

void fill_info(const Req& r, mjk_request_info* info)
{
    info->state = r.state;
    info->nao = r.nao;
    info->n_dm = r.n_dm;
    info->e_tot = r.e_tot;
    info->cycles = r.cycles;
    info->converged = r.converged;
    info->t_submit = r.t_submit;
    info->t_start = r.t_start;
    info->t_done = r.t_done;
    info->hbm_estimate = r.estimate;
    info->hbm_peak = r.peak;
    info->slot = r.slot;
    snprintf(info->error, sizeof(info->error), "%s", r.error.c_str());
}

thread_local std::string t_err;   // the calling thread's last server error

int sfail(mjk_server*, int code, const std::string& msg)
{
    t_err = msg;
    return code;
}

}  // namespace

extern "C" {

int mjk_server_create(int device, int nslots, size_t hbm_budget, mjk_server** out)
{
    if (!out || nslots < 1 || nslots > 64) return MJK_EINVAL;
    *out = nullptr;
    size_t fr = 0, tot = 0;
    if (mjk_mem_info(device, &fr, &tot) != MJK_OK) return MJK_EDEVICE;
    mjk_server* s = new mjk_server();
    s->device = device;
    s->budget = hbm_budget ? hbm_budget : (size_t)(0.8 * (double)fr);
    s->slots.resize(nslots);
    for (int k = 0; k < nslots; ++k) {
        // each context may take the whole budget; admission keeps the sum within it
        const int rc = mjk_ctx_create(&s->slots[k].ctx, device, 0);
        if (rc) {
            for (int j = 0; j < k; ++j) mjk_ctx_destroy(s->slots[j].ctx);
            delete s;
            return rc;
        }
        s->slots[k].idle_bytes = mjk_ctx_hbm_used(s->slots[k].ctx);
    }
    for (int k = 0; k < nslots; ++k) s->slots[k].th = std::thread(worker, s, k);
    *out = s;
    return MJK_OK;
}

int mjk_submit(mjk_server* s, const mjk_request* q, long long* ticket)
{
    if (!s || !q || !ticket) return sfail(s, MJK_EINVAL, "bad arguments to mjk_submit");
    *ticket = 0;
    if (q->kind != MJK_REQ_JK && q->kind != MJK_REQ_RHF) return sfail(s, MJK_EINVAL, "unknown request kind");
    if (!q->atm || !q->bas || !q->env || q->natm <= 0 || q->nbas <= 0 || q->env_size <= 0)
        return sfail(s, MJK_EINVAL, "request without a molecule");
    auto r = std::make_shared<Req>();
    r->kind = q->kind;
    r->natm = q->natm;
    r->nbas = q->nbas;
    r->atm.assign(q->atm, q->atm + (size_t)q->natm * ATM_SLOTS);
    r->bas.assign(q->bas, q->bas + (size_t)q->nbas * BAS_SLOTS);
    r->env.assign(q->env, q->env + q->env_size);
    for (int i = 0; i < r->natm; ++i)
        if (r->atm[i * ATM_SLOTS + PTR_COORD] < 0 || r->atm[i * ATM_SLOTS + PTR_COORD] + 3 > q->env_size)
            return sfail(s, MJK_EINVAL, "atm points outside env");
    r->nao = nao_of(r->bas, r->nbas);
    r->direct_scf_tol = q->direct_scf_tol;
    r->fp32_tol = q->fp32_tol;
    if (q->kind == MJK_REQ_JK) {
        if (!q->dm || q->n_dm < 1 || (!q->with_j && !q->with_k) || (q->hermi != 0 && q->hermi != 1))
            return sfail(s, MJK_EINVAL, "J/K request needs densities, n_dm >= 1, hermi 0/1 and J or K");
        r->n_dm = q->n_dm;
        r->hermi = q->hermi;
        r->with_j = q->with_j != 0;
        r->with_k = q->with_k != 0;
        r->dm.assign(q->dm, q->dm + (size_t)q->n_dm * r->nao * r->nao);
    } else {
        if (q->nelectron <= 0 || q->nelectron % 2 || q->nelectron / 2 > r->nao || q->max_cycle < 0 || !(q->conv_tol > 0))
            return sfail(s, MJK_EINVAL, "RHF request needs an even electron count within the basis and conv_tol > 0");
        r->nelectron = q->nelectron;
        r->max_cycle = q->max_cycle;
        r->conv_tol = q->conv_tol;
    }
    r->estimate = estimate_bytes(*r);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->closing) return sfail(s, MJK_EINVAL, "server is closing");
    if (r->estimate > s->budget) {
        char buf[200];
        snprintf(buf, sizeof(buf), "request needs about %zu bytes, the server's budget is %zu", r->estimate, s->budget);
        return sfail(s, MJK_ENOMEM, buf);
    }
    r->ticket = s->next_ticket++;
    r->t_submit = now(s);
    s->reqs[r->ticket] = r;
    s->queue.push_back(r);
    *ticket = r->ticket;
    s->cv.notify_all();
    return MJK_OK;
}

int mjk_poll(mjk_server* s, long long ticket, mjk_request_info* info)
{
    if (!s || !info) return MJK_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return sfail(s, MJK_EINVAL, "unknown ticket");
    fill_info(*it->second, info);
    return MJK_OK;
}

int mjk_wait_any(mjk_server* s, double timeout_s, long long* ticket)
{
    if (!s || !ticket) return MJK_EINVAL;
    *ticket = 0;
    std::unique_lock<std::mutex> lk(s->mu);
    const auto until = std::chrono::steady_clock::now() +
                       std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                           std::chrono::duration<double>(std::max(0.0, timeout_s)));
    if (!s->cv.wait_until(lk, until, [&] { return !s->finished.empty(); })) return MJK_OK;
    *ticket = s->finished.front();
    s->finished.pop_front();
    return MJK_OK;
}

int mjk_result(mjk_server* s, long long ticket, double* vj, double* vk, double* dm, double* mo_energy)
{
    if (!s) return MJK_EINVAL;
    std::shared_ptr<Req> r;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) return sfail(s, MJK_EINVAL, "unknown ticket");
        r = it->second;
        if (r->state != MJK_REQ_DONE && r->state != MJK_REQ_FAILED)
            return sfail(s, MJK_EINVAL, "request has not finished");
        s->reqs.erase(it);
    }
    if (r->state == MJK_REQ_FAILED) return sfail(s, MJK_EDEVICE, r->error);
    if (vj && !r->vj.empty()) std::memcpy(vj, r->vj.data(), r->vj.size() * sizeof(double));
    if (vk && !r->vk.empty()) std::memcpy(vk, r->vk.data(), r->vk.size() * sizeof(double));
    if (dm && r->kind == MJK_REQ_RHF) std::memcpy(dm, r->dm.data(), r->dm.size() * sizeof(double));
    if (mo_energy && !r->mo_energy.empty()) std::memcpy(mo_energy, r->mo_energy.data(), r->mo_energy.size() * sizeof(double));
    return MJK_OK;
}

size_t mjk_server_hbm_held(mjk_server* s)
{
    if (!s) return 0;
    std::lock_guard<std::mutex> lk(s->mu);
    return held_locked(s);
}

const char* mjk_server_error(const mjk_server*)
{
    return t_err.c_str();
}

void mjk_server_destroy(mjk_server* s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->closing = true;
    }
    s->cv.notify_all();
    for (Slot& sl : s->slots)
        if (sl.th.joinable()) sl.th.join();
    for (Slot& sl : s->slots) mjk_ctx_destroy(sl.ctx);
    delete s;
}

}  // extern "C"
