// lbm_server.cpp -- continuous batching of independent lattice-Boltzmann cases
// on one device (new; the reference runs one simulation per process,
// src/main.cpp:96-434).
//
// Resident state lives in one device arena sized by the memory budget, so
// admitting and retiring cases never allocates or frees device memory (a
// free would wait for the whole device).  Cases of one storage type are the
// members of one group: a chunk of K steps is one graph-replayed sequence of
// fused launches over all of them, and every member carries its own stop, so
// it freezes at exactly its step count inside the chunk.  Between chunks the
// frozen members leave the group and newly admitted ones join; while a chunk
// runs, retired cases are read back and queued cases are initialised on an
// auxiliary stream.  Built only on the C-ABI of lbm_capi.h plus HIP streams
// and events.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/lbm_capi.h"
#include "launch_shape.hpp"

namespace {

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// First-fit allocator over [0, size) with coalescing of freed ranges.
class Arena {
  public:
    void reset(size_t size) {
        free_.clear();
        if (size) free_[0] = size;
    }
    bool alloc(size_t n, size_t &off) {
        n = align256(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            off = it->first;
            const size_t rest = it->second - n, next = it->first + n;
            free_.erase(it);
            if (rest) free_[next] = rest;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t n) {
        n = align256(n);
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

  private:
    std::map<size_t, size_t> free_;
};

struct Case {
    long long id = 0;
    lbm_case c{};
    int state = LBM_CASE_QUEUED;
    size_t bytes = 0, offset = 0;
    lbm_solver *solver = nullptr;
    long long counted = 0;  // steps already added to the statistics
    long long steps_done = 0;
    long long cells = 0;
    double submit_s = 0, admit_s = 0, done_s = 0;
    std::vector<float> rho, j;
    std::string error;  // why a FAILED case failed
};

struct Pool {
    lbm_group *g = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<Case *> members;  // in group member order
    bool tuned = false;
    bool launched = false;
    double ns_per_cell_step = 0.04;  // running estimate of the fused launch cost
    long long chunk_cells = 0;
    int chunk_steps = 0;
};

int storage_slot(int storage) {
    switch (storage) {
        case LBM_STORAGE_F32: return 0;
        case LBM_STORAGE_F16_SHIFTED: return 1;
        case LBM_STORAGE_BF16_SHIFTED: return 2;
        case LBM_STORAGE_F64: return 3;
        default: return -1;
    }
}
const int kStorages[4] = {LBM_STORAGE_F32, LBM_STORAGE_F16_SHIFTED, LBM_STORAGE_BF16_SHIFTED, LBM_STORAGE_F64};

void case_solver_config(const lbm_case &c, int device, int wg, int variant, lbm_solver_config &cfg) {
    lbm_solver_config_default(&cfg);
    for (int a = 0; a < 3; ++a) cfg.dims[a] = c.dims[a];
    cfg.storage = c.storage;
    cfg.tau = c.tau;
    cfg.inv_tau = 1.0 / c.tau;
    cfg.smagorinsky_cs = c.smagorinsky_cs;
    cfg.device = device;
    cfg.work_group_size = wg;
    cfg.kernel_variant = variant;
    cfg.store_velocity = 0;
    cfg.store_density = 0;
    if (c.kind == LBM_CASE_CHANNEL) {
        cfg.bc[0] = cfg.bc[1] = LBM_FLAG_GHOST_LAYER;  // periodic x: ghost faces fed by the exchange
        cfg.bc[2] = cfg.bc[3] = cfg.bc[4] = cfg.bc[5] = LBM_FLAG_OBSTACLE;
        cfg.forcing = LBM_FORCING_BODY;
        cfg.gravitation[0] = c.body_force;
        cfg.lid_velocity = 0.0;
    } else {
        for (int i = 0; i < 6; ++i) cfg.bc[i] = LBM_FLAG_OBSTACLE;
        cfg.forcing = LBM_FORCING_REFERENCE;
        cfg.lid_velocity = c.lid_velocity;
    }
}

}  // namespace

struct lbm_server {
    lbm_server_config cfg{};
    std::string err;
    void *arena_base = nullptr;
    Arena arena;
    hipStream_t aux = nullptr;
    Pool pools[4];
    std::map<long long, std::unique_ptr<Case>> cases;
    std::deque<Case *> queue;
    std::vector<Case *> ready;   // initialised, joining at the next chunk boundary
    std::vector<Case *> retiring;
    std::deque<long long> done_order;
    bool inflight = false;
    long long next_id = 1;
    lbm_server_stats stats{};
    std::chrono::steady_clock::time_point t0;

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
    int fail(int code, const std::string &m) {
        err = m;
        return code;
    }
    int fail_lib(int code, const char *what, const char *detail) {
        err = std::string(what) + ": " + (detail ? detail : "");
        return code;
    }
};

namespace {

int active_count(const lbm_server *s) {
    int n = 0;
    for (const Pool &p : s->pools) n += (int)p.members.size();
    return n;
}

// cells of the cases resident in the pools or initialised for the next chunk
long long resident_cells(const lbm_server *s) {
    long long n = 0;
    for (const Pool &p : s->pools)
        for (const Case *c : p.members) n += c->cells;
    for (const Case *c : s->ready) n += c->cells;
    return n;
}

int pool_open(lbm_server *s, int slot) {
    Pool &p = s->pools[slot];
    if (p.g) return LBM_OK;
    int rc = lbm_group_create_empty(s->cfg.device, kStorages[slot], s->cfg.work_group_size, s->cfg.kernel_variant, &p.g);
    if (rc) return s->fail_lib(rc, "group creation", lbm_last_error());
    rc = lbm_group_set_exchange_mode(p.g, 1);
    // periodic channel seams inside the step kernel for FP32/FP64, copy launch for FP16/BF16
    if (!rc) rc = lbm_group_set_fused_x(p.g, LBM_FUSED_X_AUTO);
    // membership changes at most chunk boundaries, and a changed membership
    // needs new graphs: one captured pair, replayed K/2 times, keeps the
    // capture cost per chunk small
    if (!rc) rc = lbm_group_set_graph_pairs(p.g, 1);
    if (rc) return s->fail_lib(rc, "exchange mode", lbm_group_error(p.g));
    void *st = nullptr;
    lbm_group_get_stream(p.g, &st);
    p.stream = (hipStream_t)st;
    if (hipEventCreate(&p.ev0) != hipSuccess || hipEventCreate(&p.ev1) != hipSuccess)
        return s->fail(LBM_ERR_DEVICE, "hipEventCreate failed");
    return LBM_OK;
}

// Initial state of a case in its arena block, on the auxiliary stream.
int case_setup(lbm_server *s, Case *c) {
    lbm_solver_config cfg;
    case_solver_config(c->c, s->cfg.device, s->cfg.work_group_size, s->cfg.kernel_variant, cfg);
    int rc = lbm_solver_create_in(&cfg, (char *)s->arena_base + c->offset, c->bytes, (void *)s->aux, &c->solver);
    if (rc) return s->fail_lib(rc, "case solver", lbm_last_error());
    lbm_solver *sv = c->solver;
    const int X = c->c.dims[0], Y = c->c.dims[1], Z = c->c.dims[2];
    if (c->c.perturbation > 0.0) {
        rc = lbm_solver_init_perturbed(sv, c->c.perturbation, c->c.seed, 0, nullptr, nullptr);
        if (rc) return s->fail_lib(rc, "perturbation", lbm_solver_error(sv));
    }
    if (c->c.kind == LBM_CASE_CAVITY) {
        // VELOCITY_INJECTION layer at y = Y-2, x and z interior (src/CController.hpp:531-546)
        const int o[3] = {1, Y - 2, 1}, sz[3] = {X - 2, 1, Z - 2};
        std::vector<int> fl((size_t)sz[0] * sz[2], LBM_FLAG_VELOCITY_INJECTION);
        rc = lbm_solver_set(sv, LBM_FIELD_FLAGS, o, sz, fl.data(), nullptr);
        if (rc) return s->fail_lib(rc, "lid layer", lbm_solver_error(sv));
    } else {
        if (c->c.obstacle_fill > 0.0) {
            double achieved = 0.0;
            rc = lbm_solver_place_spheres(sv, c->c.seed, c->c.obstacle_fill, c->c.r_min, c->c.r_max, nullptr, nullptr,
                                          &achieved);
            if (rc) return s->fail_lib(rc, "spheres", lbm_solver_error(sv));
        }
        if (c->c.perturbation > 0.0) {
            // ghost planes start as copies of their periodic images (all 19 populations)
            std::vector<double> plane((size_t)19 * Y * Z);  // room for the compute type of any storage
            const int sz[3] = {1, Y, Z};
            const int pairs[2][2] = {{X - 2, 0}, {1, X - 1}};
            for (const auto &pr : pairs) {
                const int so[3] = {pr[0], 0, 0}, ro[3] = {pr[1], 0, 0};
                rc = lbm_solver_store(sv, LBM_FIELD_DD, so, sz, plane.data());
                if (!rc) rc = lbm_solver_set(sv, LBM_FIELD_DD, ro, sz, plane.data(), nullptr);
                if (rc) return s->fail_lib(rc, "periodic ghost planes", lbm_solver_error(sv));
            }
        }
    }
    return LBM_OK;
}

// Destroys the case's solver and returns its arena block.  A solver that
// cannot be destroyed (still a group member) keeps its block allocated, so
// admission can never hand memory that a group still steps to another case.
bool case_release(lbm_server *s, Case *c) {
    if (c->solver) {
        if (lbm_solver_destroy(c->solver) != LBM_OK) return false;
        c->solver = nullptr;
    }
    s->arena.release(c->offset, c->bytes);
    s->stats.bytes_in_use -= align256(c->bytes);
    return true;
}

// The case leaves the server as failed: its solver and arena block go (see
// case_release), its error is the server's last one.
void case_fail(lbm_server *s, Case *c) {
    (void)case_release(s, c);
    c->state = LBM_CASE_FAILED;
    c->error = s->err;
    c->done_s = s->now();
    s->done_order.push_back(c->id);
    s->stats.failed++;
}

// Joins the case's storage pool; on failure the group is left as it was (the
// member and the comms naming it are removed again).  If that removal fails
// too, the case stays a member (in step with the group's member indices, its
// block allocated), *stuck is set and the server reports the error.
int pool_join(lbm_server *s, Case *c, bool *stuck) {
    *stuck = false;
    const int slot = storage_slot(c->c.storage);
    int rc = pool_open(s, slot);
    if (rc) return rc;
    Pool &p = s->pools[slot];
    int idx = -1;
    rc = lbm_group_add_member(p.g, c->solver, &idx);
    if (rc) return s->fail_lib(rc, "group join", lbm_group_error(p.g));
    const char *step = nullptr;
    if (c->c.kind == LBM_CASE_CHANNEL) {
        // periodic x: the case is its own neighbour across both x faces
        const int X = c->c.dims[0], Y = c->c.dims[1], Z = c->c.dims[2];
        const int size[3] = {1, Y, Z};
        const int so0[3] = {1, 0, 0}, ro0[3] = {0, 0, 0}, n0[3] = {1, 0, 0};
        const int so1[3] = {X - 2, 0, 0}, ro1[3] = {X - 1, 0, 0}, n1[3] = {-1, 0, 0};
        rc = lbm_group_add_comm(p.g, idx, idx, so0, ro0, size, n0);
        if (!rc) rc = lbm_group_add_comm(p.g, idx, idx, so1, ro1, size, n1);
        if (rc) step = "periodic comms";
    }
    if (!rc) {
        rc = lbm_group_set_member_stop(p.g, idx, c->c.steps);
        if (rc) step = "member stop";
    }
    if (rc) {
        const int code = s->fail_lib(rc, step, lbm_group_error(p.g));
        const std::string why = s->err;
        if (lbm_group_remove_member(p.g, idx) != LBM_OK) {
            p.members.push_back(c);
            c->state = LBM_CASE_RUNNING;
            s->err = why + "; leaving the group again failed: " + lbm_group_error(p.g);
            *stuck = true;
        }
        return code;
    }
    p.members.push_back(c);
    c->state = LBM_CASE_RUNNING;
    return LBM_OK;
}

// rho and j of the whole case, z-slab by z-slab through the solver's staging
// area (one face of 19 values holds the 4 moment values of >= 4 planes).
int case_readback(lbm_server *s, Case *c) {
    const int X = c->c.dims[0], Y = c->c.dims[1], Z = c->c.dims[2];
    const size_t plane = (size_t)X * Y;
    const bool f64 = c->c.storage == LBM_STORAGE_F64;
    c->rho.resize((size_t)c->cells);
    c->j.resize((size_t)c->cells * 3);
    const int slab = 4;
    std::vector<double> rb, jb;
    std::vector<float> rf, jf;
    for (int z0 = 0; z0 < Z; z0 += slab) {
        const int nz = std::min(slab, Z - z0);
        const int o[3] = {0, 0, z0}, sz[3] = {X, Y, nz};
        const size_t n = plane * nz;
        int rc;
        if (f64) {
            rb.resize(n);
            jb.resize(3 * n);
            rc = lbm_solver_moments(c->solver, o, sz, rb.data(), jb.data());
        } else {
            rf.resize(n);
            jf.resize(3 * n);
            rc = lbm_solver_moments(c->solver, o, sz, rf.data(), jf.data());
        }
        if (rc) return s->fail_lib(rc, "readback", lbm_solver_error(c->solver));
        const size_t at = plane * z0;
        for (size_t i = 0; i < n; ++i) c->rho[at + i] = f64 ? (float)rb[i] : rf[i];
        for (int a = 0; a < 3; ++a)
            for (size_t i = 0; i < n; ++i)
                c->j[(size_t)a * c->cells + at + i] = f64 ? (float)jb[a * n + i] : jf[a * n + i];
    }
    if (!case_release(s, c)) return s->fail(LBM_ERR_STATE, "readback: the case's solver cannot be released");
    c->state = LBM_CASE_DONE;
    c->done_s = s->now();
    s->done_order.push_back(c->id);
    s->stats.completed++;
    return LBM_OK;
}

int admit(lbm_server *s) {
    long long cells = s->cfg.max_active_cells > 0 ? resident_cells(s) : 0;
    for (auto it = s->queue.begin(); it != s->queue.end();) {
        if (s->cfg.max_active > 0 && active_count(s) + (int)s->ready.size() >= s->cfg.max_active) break;
        Case *c = *it;
        // the cell limit admits in order: a case that would exceed it waits,
        // and so does everything behind it, unless nothing is resident
        if (s->cfg.max_active_cells > 0 && cells > 0 && cells + c->cells > s->cfg.max_active_cells) break;
        size_t off = 0;
        if (!s->arena.alloc(c->bytes, off)) {
            if (!s->cfg.backfill) break;
            ++it;
            continue;
        }
        c->offset = off;
        s->stats.bytes_in_use += align256(c->bytes);
        s->stats.bytes_peak = std::max(s->stats.bytes_peak, s->stats.bytes_in_use);
        int rc = case_setup(s, c);
        it = s->queue.erase(it);
        if (rc) {
            // the case fails on its own: its block returns to the arena and
            // the server goes on with the others
            case_fail(s, c);
            continue;
        }
        c->admit_s = s->now();
        s->ready.push_back(c);
        cells += c->cells;
    }
    s->stats.cells_peak = std::max(s->stats.cells_peak, resident_cells(s));
    return LBM_OK;
}

}  // namespace

extern "C" void lbm_server_config_default(lbm_server_config *c) {
    if (!c) return;
    *c = lbm_server_config{};
    c->device = 0;
    c->memory_budget = (size_t)8 << 30;
    c->max_active = 0;
    c->chunk_ms = 25.0;
    c->max_chunk_steps = 1000;
    c->work_group_size = 256;
    c->kernel_variant = 0;
    c->autotune = 0;
    c->backfill = 1;
    c->max_active_cells = 0;
}

extern "C" void lbm_case_default(lbm_case *c) {
    if (!c) return;
    *c = lbm_case{};
    c->kind = LBM_CASE_CAVITY;
    c->dims[0] = c->dims[1] = c->dims[2] = 32;
    c->steps = 100;
    c->storage = LBM_STORAGE_F32;
    c->tau = 0.52;
    c->smagorinsky_cs = 0.17;
    c->lid_velocity = 0.05;
    c->body_force = 1e-6;
    c->obstacle_fill = 0.05;
    c->r_min = 3.0;
    c->r_max = 8.0;
    c->perturbation = 0.0;
    c->seed = 1;
}

extern "C" int lbm_server_case_bytes(const lbm_case *c, size_t *bytes) {
    if (!c || !bytes) return LBM_ERR_INVALID_ARGUMENT;
    if (c->kind != LBM_CASE_CAVITY && c->kind != LBM_CASE_CHANNEL) return LBM_ERR_INVALID_ARGUMENT;
    if (storage_slot(c->storage) < 0) return LBM_ERR_INVALID_ARGUMENT;
    // the solver's size does not depend on the launch shape the server uses
    lbm_solver_config cfg;
    case_solver_config(*c, 0, 256, 0, cfg);
    return lbm_solver_required_bytes(&cfg, bytes);
}

extern "C" int lbm_server_create(const lbm_server_config *cfg, lbm_server **out) {
    if (!cfg || !out) return LBM_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (cfg->memory_budget < ((size_t)1 << 20) || cfg->max_active < 0 || cfg->max_active_cells < 0 ||
        !(cfg->chunk_ms > 0.0) ||
        cfg->max_chunk_steps < 2 || !lbm_work_group_supported(cfg->work_group_size) ||
        !lbm_variant_supported(cfg->kernel_variant))
        return LBM_ERR_INVALID_ARGUMENT;
    std::unique_ptr<lbm_server> s(new lbm_server());
    s->cfg = *cfg;
    s->cfg.max_chunk_steps &= ~1;
    s->t0 = std::chrono::steady_clock::now();
    if (hipSetDevice(cfg->device) != hipSuccess) return LBM_ERR_DEVICE;
    if (hipMalloc(&s->arena_base, cfg->memory_budget) != hipSuccess) {
        s->arena_base = nullptr;
        return LBM_ERR_OUT_OF_MEMORY;
    }
    s->arena.reset(cfg->memory_budget);
    if (hipStreamCreateWithFlags(&s->aux, hipStreamNonBlocking) != hipSuccess) {
        (void)hipFree(s->arena_base);
        return LBM_ERR_DEVICE;
    }
    *out = s.release();
    return LBM_OK;
}

extern "C" int lbm_server_destroy(lbm_server *s) {
    if (!s) return LBM_OK;
    (void)hipSetDevice(s->cfg.device);
    for (Pool &p : s->pools) {
        if (p.g) {
            (void)lbm_group_wait(p.g);
            lbm_group_destroy(p.g);
        }
        if (p.ev0) (void)hipEventDestroy(p.ev0);
        if (p.ev1) (void)hipEventDestroy(p.ev1);
    }
    for (auto &kv : s->cases)
        if (kv.second->solver) lbm_solver_destroy(kv.second->solver);
    if (s->aux) (void)hipStreamDestroy(s->aux);
    if (s->arena_base) (void)hipFree(s->arena_base);
    delete s;
    return LBM_OK;
}

extern "C" const char *lbm_server_error(const lbm_server *s) { return s ? s->err.c_str() : ""; }

extern "C" double lbm_server_time(const lbm_server *s) { return s ? s->now() : 0.0; }

extern "C" int lbm_server_submit(lbm_server *s, const lbm_case *c, long long *id) {
    if (!s || !c) return LBM_ERR_INVALID_ARGUMENT;
    size_t bytes = 0;
    int rc = lbm_server_case_bytes(c, &bytes);
    if (rc) return s->fail(rc, std::string("case rejected: ") + lbm_last_error());
    if (c->steps < 0 || !(c->tau > 0.5) || c->smagorinsky_cs < 0.0)
        return s->fail(LBM_ERR_INVALID_ARGUMENT, "case needs steps >= 0 and tau > 0.5");
    if (c->kind == LBM_CASE_CHANNEL && c->obstacle_fill > 0.0 &&
        !(c->obstacle_fill < 0.9 && c->r_min > 0.0 && c->r_max >= c->r_min))
        return s->fail(LBM_ERR_INVALID_ARGUMENT, "channel spheres need 0 <= fill < 0.9 and 0 < r_min <= r_max");
    if (align256(bytes) > s->cfg.memory_budget)
        return s->fail(LBM_ERR_CAPACITY, "case needs " + std::to_string(bytes) + " device bytes, more than the budget");
    std::unique_ptr<Case> cs(new Case());
    cs->id = s->next_id++;
    cs->c = *c;
    cs->bytes = bytes;
    cs->cells = (long long)c->dims[0] * c->dims[1] * c->dims[2];
    cs->submit_s = s->now();
    s->queue.push_back(cs.get());
    if (id) *id = cs->id;
    s->cases[cs->id] = std::move(cs);
    s->stats.submitted++;
    return LBM_OK;
}

extern "C" int lbm_server_progress(lbm_server *s, int *completed) {
    if (!s) return LBM_ERR_INVALID_ARGUMENT;
    if (completed) *completed = 0;
    if (hipSetDevice(s->cfg.device) != hipSuccess) return s->fail(LBM_ERR_DEVICE, "hipSetDevice failed");
    const long long done_before = s->stats.completed;
    // 1. the chunk in flight
    if (s->inflight) {
        for (Pool &p : s->pools) {
            if (!p.launched) continue;
            if (hipEventSynchronize(p.ev1) != hipSuccess) return s->fail(LBM_ERR_DEVICE, "chunk wait failed");
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, p.ev0, p.ev1) == hipSuccess && p.chunk_cells > 0 && p.chunk_steps > 0) {
                s->stats.device_ms += ms;
                const double est = (double)ms * 1e6 / ((double)p.chunk_cells * p.chunk_steps);
                p.ns_per_cell_step = 0.5 * p.ns_per_cell_step + 0.5 * est;
            }
            p.launched = false;
        }
        s->inflight = false;
    }
    // 2. retire members that reached their step count
    for (Pool &p : s->pools) {
        for (int i = (int)p.members.size() - 1; i >= 0; --i) {
            Case *c = p.members[i];
            long long done = 0;
            lbm_solver_get_step_counter(c->solver, &done);
            s->stats.steps_cells += (done - c->counted) * c->cells;
            c->counted = done;
            c->steps_done = done;
            if (done < c->c.steps) continue;
            int rc = lbm_group_remove_member(p.g, i);
            if (rc) return s->fail_lib(rc, "group leave", lbm_group_error(p.g));
            p.members.erase(p.members.begin() + i);
            s->retiring.push_back(c);
        }
    }
    // 3. cases initialised during the last chunk join; one that cannot join
    // fails on its own, as in admission
    std::vector<Case *> ready;
    ready.swap(s->ready);
    for (size_t i = 0; i < ready.size(); ++i) {
        Case *c = ready[i];
        if (c->c.steps == 0) {
            s->retiring.push_back(c);
            continue;
        }
        bool stuck = false;
        const int jr = pool_join(s, c, &stuck);
        if (jr == LBM_OK) continue;
        if (stuck) {  // the group no longer matches the pool: stop here, keep the rest ready
            s->ready.insert(s->ready.end(), ready.begin() + (long)i + 1, ready.end());
            return jr;
        }
        case_fail(s, c);
    }
    // 4. the next chunk, one graph-replayed step sequence per storage
    for (Pool &p : s->pools) {
        if (p.members.empty()) continue;
        if (s->cfg.autotune && !p.tuned) {
            int wg = 0, var = 0;
            double ms = 0;
            int rc = lbm_group_autotune(p.g, 2, &wg, &var, &ms);
            if (rc) return s->fail_lib(rc, "autotune", lbm_group_error(p.g));
            p.tuned = true;
        }
        long long cells = 0, remaining = 0;
        for (Case *c : p.members) {
            long long done = 0;
            lbm_solver_get_step_counter(c->solver, &done);
            if (done < c->c.steps) cells += c->cells;
            remaining = std::max(remaining, (long long)c->c.steps - done);
        }
        if (remaining <= 0) continue;
        long long k = (long long)(s->cfg.chunk_ms * 1e6 / (p.ns_per_cell_step * (double)std::max(cells, 1ll)));
        k = std::max(2ll, std::min<long long>(k, s->cfg.max_chunk_steps));
        k = std::min(k, remaining + (remaining & 1));
        k += k & 1;
        if (hipEventRecord(p.ev0, p.stream) != hipSuccess) return s->fail(LBM_ERR_DEVICE, "event record failed");
        int rc = lbm_group_step(p.g, (int)k);
        if (rc) return s->fail_lib(rc, "chunk", lbm_group_error(p.g));
        if (hipEventRecord(p.ev1, p.stream) != hipSuccess) return s->fail(LBM_ERR_DEVICE, "event record failed");
        p.launched = true;
        p.chunk_cells = cells;
        p.chunk_steps = (int)k;
        s->inflight = true;
        s->stats.chunks++;
    }
    // 5. while it runs: results of the retired cases (a case whose readback
    // fails fails on its own), then admissions
    std::vector<Case *> retiring;
    retiring.swap(s->retiring);
    for (Case *c : retiring)
        if (case_readback(s, c) != LBM_OK) case_fail(s, c);
    int rc = admit(s);
    if (rc) return rc;
    if (completed) *completed = (int)(s->stats.completed - done_before);
    return LBM_OK;
}

extern "C" int lbm_server_next_done(lbm_server *s, lbm_case_status *st) {
    if (!s || !st) return LBM_ERR_INVALID_ARGUMENT;
    while (!s->done_order.empty()) {
        const long long id = s->done_order.front();
        s->done_order.pop_front();
        if (s->cases.count(id)) return lbm_server_status(s, id, st);
    }
    return LBM_ERR_STATE;
}

extern "C" int lbm_server_status(const lbm_server *s, long long id, lbm_case_status *st) {
    if (!s || !st) return LBM_ERR_INVALID_ARGUMENT;
    auto it = s->cases.find(id);
    if (it == s->cases.end()) return LBM_ERR_OUT_OF_RANGE;
    const Case *c = it->second.get();
    st->id = c->id;
    st->state = c->state;
    st->steps_done = c->steps_done;
    st->cells = c->cells;
    st->submit_s = c->submit_s;
    st->admit_s = c->admit_s;
    st->done_s = c->done_s;
    return LBM_OK;
}

extern "C" int lbm_server_fetch(lbm_server *s, long long id, float *rho, float *j) {
    if (!s) return LBM_ERR_INVALID_ARGUMENT;
    auto it = s->cases.find(id);
    if (it == s->cases.end()) return s->fail(LBM_ERR_OUT_OF_RANGE, "unknown case id");
    Case *c = it->second.get();
    if (c->state == LBM_CASE_FAILED) {
        const std::string why = "case " + std::to_string(id) + " failed: " + c->error;
        s->cases.erase(it);
        return s->fail(LBM_ERR_STATE, why);
    }
    if (c->state != LBM_CASE_DONE) return s->fail(LBM_ERR_STATE, "case not completed");
    if (rho) std::copy(c->rho.begin(), c->rho.end(), rho);
    if (j) std::copy(c->j.begin(), c->j.end(), j);
    s->cases.erase(it);
    return LBM_OK;
}

extern "C" int lbm_server_stats_get(const lbm_server *s, lbm_server_stats *st) {
    if (!s || !st) return LBM_ERR_INVALID_ARGUMENT;
    *st = s->stats;
    st->queued = (int)s->queue.size();
    st->active = active_count(s);
    st->ready = (int)s->ready.size();
    return LBM_OK;
}
