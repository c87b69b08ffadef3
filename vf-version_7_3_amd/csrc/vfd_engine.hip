// vfd_engine.hip -- resident runs, batched ticks (hipGraph), and the serving
// loop (continuous batching under an HBM budget).  See include/vfd.h.
//
// The reference diffuses the 8 cytokine fields of one world per tick, one
// field at a time per GPU (src/Diffusion/Diffuser.cpp:300-421).  Here every
// field of every resident run of one grid shape goes into the same launch of
// each pass; shapes run on parallel stream branches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/vfd.h"
#include "vfd_internal.h"

using namespace vfd;
using Clock = std::chrono::steady_clock;

namespace {

constexpr int kMaxRunFields = 32;
constexpr int kAuxStreams = 8;
constexpr int kMaxGraphs = 8;
// work buffers one shape group of a batch may hold at once: fields beyond it
// go in further chunks, each through all five passes before the next, so on
// mid-size grids a chunk's work arrays can stay in the 256 MB last-level
// cache between passes (measured: ensemble ticks 4 % faster than with an
// 8 GiB cap).
// VFD_GROUP_WORK_BYTES overrides it.
constexpr size_t kGroupWorkBytes = size_t(512) << 20;
static size_t group_work_bytes() {
    static const size_t v = [] {
        const char* s = getenv("VFD_GROUP_WORK_BYTES");
        const long long b = s ? atoll(s) : 0;
        return b > 0 ? (size_t)b : kGroupWorkBytes;
    }();
    return v;
}

struct ShapeKey {
    int nz, ny, nx, prec;
    bool operator<(const ShapeKey& o) const {
        if (nz != o.nz) return nz < o.nz;
        if (ny != o.ny) return ny < o.ny;
        if (nx != o.nx) return nx < o.nx;
        return prec < o.prec;
    }
};

struct Shape {
    Plan plan{};
    std::vector<void*> pool;  // work buffers of plan.work_bytes
    std::vector<char> held;   // 1 while a batch or a cached graph owns the buffer
    int nruns = 0;
};

struct Run {
    ShapeKey key{};
    int nfields = 0;
    std::vector<void*> field, kspec;
    size_t bytes = 0;
    bool busy = false;  // in an in-flight serving batch
};

struct Held {  // work buffers a batch or graph owns
    Shape* shape;
    std::vector<int> idx;
};

struct GraphRec {
    hipGraphExec_t exec = nullptr;
    std::vector<Held> held;
    uint64_t last_use = 0;
};

struct Request {
    uint64_t ticket;
    int run;
    Clock::time_point t0;
};

// streams + events for the per-shape parallel branches of one batch
struct Branches {
    hipStream_t aux[kAuxStreams] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t join[kAuxStreams] = {};
    int create() {
        for (int i = 0; i < kAuxStreams; ++i) {
            if (int rc = hip_check(hipStreamCreateWithFlags(&aux[i], hipStreamNonBlocking), "branch stream")) return rc;
            if (int rc = hip_check(hipEventCreateWithFlags(&join[i], hipEventDisableTiming), "join event")) return rc;
        }
        return hip_check(hipEventCreateWithFlags(&fork, hipEventDisableTiming), "fork event");
    }
    void destroy() {
        for (int i = 0; i < kAuxStreams; ++i) {
            if (aux[i]) (void)hipStreamDestroy(aux[i]);
            if (join[i]) (void)hipEventDestroy(join[i]);
            aux[i] = nullptr;
            join[i] = nullptr;
        }
        if (fork) (void)hipEventDestroy(fork);
        fork = nullptr;
    }
};

// the fields of one batch grouped by shape, with the work buffers they own
struct Batch {
    std::vector<std::pair<Shape*, std::vector<vfd_field>>> groups;
    std::vector<Held> held;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    Branches br;
    bool busy = false;
    std::vector<Request> reqs;
    Batch batch;
};

struct ReqState {
    bool done = false;
    double latency_ms = 0.0;
};

}  // namespace

struct vfd_engine {
    std::mutex mu;
    std::condition_variable cv;
    int dev = 0;
    size_t budget = 0, used = 0, peak = 0;
    uint64_t admit_waits = 0;  // runs admitted after waiting for memory
    std::map<ShapeKey, Shape> shapes;
    std::map<int, Run> runs;
    int next_run = 1;
    Branches br;  // branches of vfd_step
    std::map<std::vector<int>, GraphRec> graphs;
    uint64_t graph_clock = 0;
    // serving
    std::thread server;
    bool serving = false, stop = false;
    int max_batch_fields = 256;
    std::vector<Slot> slots;
    std::deque<Request> pending;
    std::unordered_map<uint64_t, ReqState> reqs;
    uint64_t next_ticket = 1;
    uint64_t batches = 0, ticks = 0, done_count = 0;
};

namespace {

size_t run_bytes_for(const Plan& p, int nfields) { return (size_t)nfields * (p.field_bytes + p.kspec_bytes); }

int make_plan_for(const ShapeKey& k, Plan* p) {
    vfd_plan_desc d{k.nz, k.ny, k.nx, 0, 0, 0, k.prec, VFD_PAD_ZERO, VFD_KSPEC_SEPARABLE};
    return plan_init(&d, p);
}

// Reserve `want` work buffers of a shape (growing the pool within the budget);
// returns how many were reserved (>= 1 unless the budget is exhausted).
int acquire(vfd_engine* e, Shape& s, int want, Held* h) {
    h->shape = &s;
    h->idx.clear();
    for (int i = 0; i < (int)s.pool.size() && (int)h->idx.size() < want; ++i)
        if (!s.held[i]) {
            s.held[i] = 1;
            h->idx.push_back(i);
        }
    while ((int)h->idx.size() < want && e->used + s.plan.work_bytes <= e->budget) {
        void* p = nullptr;
        if (hipMalloc(&p, s.plan.work_bytes) != hipSuccess) {
            (void)hipGetLastError();
            break;
        }
        e->used += s.plan.work_bytes;
        e->peak = std::max(e->peak, e->used);
        s.pool.push_back(p);
        s.held.push_back(1);
        h->idx.push_back((int)s.pool.size() - 1);
    }
    return (int)h->idx.size();
}

void release(Held& h) {
    for (int i : h.idx) h.shape->held[i] = 0;
    h.idx.clear();
}

// Give the idle work buffers at the end of each shape's pool back to the
// budget (each shape keeps its first one): pools that earlier batches grew
// would otherwise hold memory an admission is waiting for.  Only trailing
// entries go, so the indices that in-flight batches and graphs hold stay valid.
bool trim_pools(vfd_engine* e) {
    bool any = false;
    for (auto& kv : e->shapes) {
        Shape& s = kv.second;
        while (s.pool.size() > 1 && !s.held.back()) {
            if (!any) (void)hipDeviceSynchronize();  // released buffers may still be read by queued ticks
            any = true;
            (void)hipFree(s.pool.back());
            s.pool.pop_back();
            s.held.pop_back();
            e->used -= s.plan.work_bytes;
        }
    }
    return any;
}

// Group the listed runs' fields by shape and reserve their work buffers
// (every allocation happens here, never while a stream is being captured).
int build_batch(vfd_engine* e, const std::vector<int>& run_ids, Batch* bt) {
    std::map<ShapeKey, std::vector<vfd_field>> groups;
    for (int id : run_ids) {
        auto it = e->runs.find(id);
        if (it == e->runs.end()) return set_error(VFD_ERR_NOTFOUND, "step: unknown run %d", id);
        Run& r = it->second;
        auto& g = groups[r.key];
        for (int f = 0; f < r.nfields; ++f) g.push_back(vfd_field{r.field[f], r.field[f], nullptr, r.kspec[f], 0.0, 0, 0});
    }
    for (auto& kv : groups) {
        Shape& s = e->shapes[kv.first];
        Held h;
        const int by_bytes = (int)std::max<size_t>(1, group_work_bytes() / s.plan.work_bytes);
        const int want = std::min({(int)kv.second.size(), kMaxFieldsPerLaunch, by_bytes});
        if (acquire(e, s, want, &h) < 1) {
            release(h);
            for (auto& hh : bt->held) release(hh);
            bt->held.clear();
            bt->groups.clear();
            return set_error(VFD_ERR_BUDGET, "step: no work buffer fits the HBM budget for %dx%dx%d", kv.first.nx,
                             kv.first.ny, kv.first.nz);
        }
        bt->held.push_back(h);
        bt->groups.emplace_back(&s, std::move(kv.second));
    }
    return VFD_OK;
}

void release_batch(Batch* bt) {
    for (auto& h : bt->held) release(h);
    bt->held.clear();
    bt->groups.clear();
}

// Launch `ticks` ticks of a built batch: shapes run on branch streams forked
// from / joined to `st` with events (also valid while `st` is captured);
// within a shape, fields go in chunks of the reserved work-buffer count.
int launch_batch(Batch& bt, int ticks, hipStream_t st, Branches& br) {
    const int ng = (int)bt.groups.size();
    if (ng == 0) return VFD_OK;
    const bool fork = ng > 1;
    if (fork) {
        if (int rc = hip_check(hipEventRecord(br.fork, st), "fork")) return rc;
    }
    for (int gi = 0; gi < ng; ++gi) {
        Shape& s = *bt.groups[gi].first;
        std::vector<vfd_field>& fl = bt.groups[gi].second;
        const Held& h = bt.held[gi];
        const size_t k = h.idx.size();
        hipStream_t gs = fork ? br.aux[gi % kAuxStreams] : st;
        if (fork) {
            if (int rc = hip_check(hipStreamWaitEvent(gs, br.fork, 0), "fork wait")) return rc;
        }
        for (int t = 0; t < ticks; ++t)
            for (size_t c = 0; c < fl.size(); c += k) {
                const int n = (int)std::min(k, fl.size() - c);
                for (int i = 0; i < n; ++i) fl[c + i].work = s.pool[h.idx[i]];
                if (int rc = plan_diffuse(s.plan, fl.data() + c, n, gs)) return rc;
            }
        if (fork) {
            if (int rc = hip_check(hipEventRecord(br.join[gi % kAuxStreams], gs), "join")) return rc;
            if (int rc = hip_check(hipStreamWaitEvent(st, br.join[gi % kAuxStreams], 0), "join wait")) return rc;
        }
    }
    return VFD_OK;
}

void drop_graph(GraphRec& g) {
    for (auto& h : g.held) release(h);
    g.held.clear();
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g.exec = nullptr;
}

void drop_graphs_with(vfd_engine* e, int run_id) {
    for (auto it = e->graphs.begin(); it != e->graphs.end();) {
        if (std::find(it->first.begin(), it->first.end(), run_id) != it->first.end()) {
            drop_graph(it->second);
            it = e->graphs.erase(it);
        } else {
            ++it;
        }
    }
}

// ------------------------------------------------------------------ serving
void server_loop(vfd_engine* e) {
    std::unique_lock<std::mutex> lk(e->mu);
    (void)hipSetDevice(e->dev);
    while (!e->stop) {
        bool progressed = false;
        // completions
        for (Slot& s : e->slots) {
            if (!s.busy) continue;
            if (hipEventQuery(s.done) != hipSuccess) continue;
            const auto now = Clock::now();
            for (const Request& r : s.reqs) {
                ReqState& st = e->reqs[r.ticket];
                st.done = true;
                st.latency_ms = std::chrono::duration<double, std::milli>(now - r.t0).count();
                e->runs[r.run].busy = false;
                ++e->done_count;
            }
            release_batch(&s.batch);
            s.reqs.clear();
            s.busy = false;
            progressed = true;
            e->cv.notify_all();
        }
        // admissions: a free slot takes every pending request whose run is idle
        for (Slot& s : e->slots) {
            if (s.busy || e->pending.empty()) continue;
            std::vector<int> ids;
            int fields = 0;
            for (auto it = e->pending.begin(); it != e->pending.end();) {
                Run& r = e->runs[it->run];
                const bool dup = std::find(ids.begin(), ids.end(), it->run) != ids.end();
                if (!r.busy && !dup && fields + r.nfields <= e->max_batch_fields) {
                    ids.push_back(it->run);
                    fields += r.nfields;
                    r.busy = true;
                    s.reqs.push_back(*it);
                    it = e->pending.erase(it);
                } else {
                    ++it;
                }
            }
            if (ids.empty()) continue;
            int rc = build_batch(e, ids, &s.batch);
            if (!rc) rc = launch_batch(s.batch, 1, s.stream, s.br);
            if (!rc) rc = hip_check(hipEventRecord(s.done, s.stream), "batch event");
            if (rc) {
                // give the requests back, keep the loop alive
                (void)hipStreamSynchronize(s.stream);
                release_batch(&s.batch);
                for (auto it = s.reqs.rbegin(); it != s.reqs.rend(); ++it) {
                    e->runs[it->run].busy = false;
                    e->pending.push_front(*it);
                }
                s.reqs.clear();
                continue;
            }
            s.busy = true;
            ++e->batches;
            ++e->ticks;
            progressed = true;
        }
        if (!progressed) {
            bool any_busy = false;
            for (Slot& s : e->slots) any_busy |= s.busy;
            // poll in-flight batches every 20 us; sleep until a submission otherwise
            if (any_busy) e->cv.wait_for(lk, std::chrono::microseconds(20));
            else e->cv.wait_for(lk, std::chrono::milliseconds(5));
        }
    }
}

// One run of a shape leaves (or a run being built gives up): after the last
// one the shape's work buffers go to `frees` (their bytes added to `bytes`) and
// the shape is dropped.  Buffers a batch or graph still holds keep the shape.
void leave_shape(vfd_engine* e, const ShapeKey& key, std::vector<void*>* frees, size_t* bytes) {
    auto sit = e->shapes.find(key);
    if (sit == e->shapes.end() || --sit->second.nruns > 0) return;
    for (char c : sit->second.held)
        if (c) return;
    for (void* p : sit->second.pool) {
        frees->push_back(p);
        *bytes += sit->second.plan.work_bytes;
    }
    e->shapes.erase(sit);
}

// Free device memory without the engine lock (hipFree waits for the device's
// queued work; the server keeps collecting and launching batches meanwhile),
// then return the bytes to the budget and wake admissions waiting for them.
void free_and_return(vfd_engine* e, const std::vector<void*>& frees, size_t bytes) {
    if (!frees.empty()) (void)hipDeviceSynchronize();
    for (void* p : frees) (void)hipFree(p);
    std::lock_guard<std::mutex> lk(e->mu);
    e->used -= bytes;
    e->cv.notify_all();
}

}  // namespace

// ====================================================================== C-ABI
extern "C" {

int vfd_engine_create(size_t hbm_budget_bytes, vfd_engine** engine) {
    if (!engine) return set_error(VFD_ERR_INVALID, "engine_create: null output");
    *engine = nullptr;
    auto* e = new vfd_engine();
    e->budget = hbm_budget_bytes;
    int rc = hip_check(hipGetDevice(&e->dev), "hipGetDevice");
    if (!rc) rc = e->br.create();
    if (rc) {
        vfd_engine_destroy(e);
        return rc;
    }
    *engine = e;
    return VFD_OK;
}

void vfd_engine_destroy(vfd_engine* e) {
    if (!e) return;
    vfd_serve_stop(e);
    (void)hipDeviceSynchronize();
    for (auto& g : e->graphs) drop_graph(g.second);
    for (auto& r : e->runs) {
        for (void* p : r.second.field) (void)hipFree(p);
        for (void* p : r.second.kspec) (void)hipFree(p);
    }
    for (auto& s : e->shapes)
        for (void* p : s.second.pool) (void)hipFree(p);
    e->br.destroy();
    delete e;
}

size_t vfd_run_bytes(const vfd_run_desc* d) {
    if (!d) return 0;
    Plan p;
    vfd_plan_desc pd{d->nz, d->ny, d->nx, 0, 0, 0, d->precision, VFD_PAD_ZERO, VFD_KSPEC_SEPARABLE};
    if (plan_init(&pd, &p)) return 0;
    return run_bytes_for(p, d->nfields);
}

// Admit a run: wait_s <= 0 refuses at once when the run does not fit the
// budget; otherwise the call waits up to wait_s seconds for destroyed runs to
// free the memory (a run larger than the whole budget is refused at once).
// The run's bytes are reserved under the engine lock; its fields and kernel
// spectra are then allocated and written without it.
static int run_create(vfd_engine* e, const vfd_run_desc* d, int* run_id, double wait_s) {
    if (!e || !d || !run_id || !d->sigma2 || !d->decay)
        return set_error(VFD_ERR_INVALID, "run_create: null argument");
    if (d->nfields < 1 || d->nfields > kMaxRunFields)
        return set_error(VFD_ERR_INVALID, "run_create: nfields %d outside 1..%d", d->nfields, kMaxRunFields);
    const ShapeKey key{d->nz, d->ny, d->nx, d->precision};
    Plan plan;
    size_t reserved = 0, reserved_work = 0;
    {
        std::unique_lock<std::mutex> lk(e->mu);
        const auto deadline = Clock::now() + std::chrono::duration_cast<Clock::duration>(
                                                 std::chrono::duration<double>(wait_s > 0 ? wait_s : 0.0));
        bool waited = false;
        for (;;) {  // the shape may appear or go while this call waits
            auto sit = e->shapes.find(key);
            const bool new_shape = sit == e->shapes.end();
            if (new_shape) {
                if (int rc = make_plan_for(key, &plan)) return rc;
            } else {
                plan = sit->second.plan;
            }
            // a new shape also needs its first work buffer
            const size_t need = run_bytes_for(plan, d->nfields) + (new_shape ? plan.work_bytes : 0);
            if (e->used + need > e->budget && need <= e->budget) trim_pools(e);
            if (e->used + need <= e->budget) {
                if (new_shape) {
                    Shape s;
                    s.plan = plan;
                    sit = e->shapes.emplace(key, s).first;
                    reserved_work = plan.work_bytes;
                }
                sit->second.nruns++;  // the shape stays while this run is built
                reserved = need;
                e->used += need;
                e->peak = std::max(e->peak, e->used);
                e->admit_waits += waited;
                break;
            }
            bool refuse = need > e->budget || wait_s <= 0;
            if (!refuse) {
                waited = true;
                refuse = e->cv.wait_until(lk, deadline) == std::cv_status::timeout && e->used + need > e->budget;
            }
            if (refuse)
                return set_error(VFD_ERR_BUDGET, "run_create: %zu bytes needed, %zu of %zu in use", need, e->used,
                                 e->budget);
        }
    }
    Run r;
    r.key = key;
    r.nfields = d->nfields;
    r.bytes = run_bytes_for(plan, d->nfields);
    int rc = VFD_OK;
    for (int f = 0; f < d->nfields && !rc; ++f) {
        void *fp = nullptr, *kp = nullptr;
        rc = hip_check(hipMalloc(&fp, plan.field_bytes), "run field");
        if (!rc) {
            r.field.push_back(fp);
            rc = hip_check(hipMalloc(&kp, plan.kspec_bytes), "run spectrum");
        }
        if (!rc) {
            r.kspec.push_back(kp);
            rc = hip_check(hipMemset(fp, 0, plan.field_bytes), "run field clear");
        }
        if (!rc) rc = plan_gauss_spectrum(plan, d->sigma2[f], d->decay[f], kp, nullptr);
    }
    std::vector<void*> frees;
    size_t unreserve = 0;
    {
        std::lock_guard<std::mutex> lk(e->mu);
        Shape& s = e->shapes.find(key)->second;  // kept by the nruns count taken above
        if (reserved_work) {
            // schedule for a new shape, tuned once per process on scratch buffers
            // when they fit (plan_init took a cached schedule: a shape whose runs
            // all left and that comes back is not tuned again)
            const int nf = std::min(8, d->nfields);
            if (!rc && !getenv("VFD_NO_AUTOTUNE") && !tuned_known(s.plan) &&
                e->used + (size_t)nf * (plan.work_bytes + plan.field_bytes + plan.kspec_bytes) <= e->budget)
                (void)plan_autotune(s.plan, nf, nullptr, nullptr);
            e->used -= reserved_work;  // the shape's first work buffer is charged by acquire
            reserved -= reserved_work;
            Held h;
            if (!rc && acquire(e, s, 1, &h) < 1)  // every shape keeps one work buffer
                rc = set_error(VFD_ERR_BUDGET, "run_create: no room for a work buffer");
            release(h);
        }
        if (!rc) {
            const int id = e->next_run++;
            e->runs[id] = r;
            *run_id = id;
            return VFD_OK;
        }
        frees = r.field;
        frees.insert(frees.end(), r.kspec.begin(), r.kspec.end());
        unreserve = reserved;
        leave_shape(e, key, &frees, &unreserve);
    }
    free_and_return(e, frees, unreserve);
    return rc;
}

int vfd_run_create(vfd_engine* e, const vfd_run_desc* d, int* run_id) { return run_create(e, d, run_id, 0.0); }

int vfd_run_create_wait(vfd_engine* e, const vfd_run_desc* d, double timeout_s, int* run_id) {
    return run_create(e, d, run_id, timeout_s);
}

int vfd_run_destroy(vfd_engine* e, int run_id) {
    if (!e) return set_error(VFD_ERR_INVALID, "run_destroy: null engine");
    std::vector<void*> frees;
    size_t bytes = 0;
    {
        std::lock_guard<std::mutex> lk(e->mu);
        auto it = e->runs.find(run_id);
        if (it == e->runs.end()) return set_error(VFD_ERR_NOTFOUND, "run_destroy: unknown run %d", run_id);
        if (it->second.busy) return set_error(VFD_ERR_STATE, "run_destroy: run %d has a tick in flight", run_id);
        for (const Request& r : e->pending)
            if (r.run == run_id) return set_error(VFD_ERR_STATE, "run_destroy: run %d has pending requests", run_id);
        drop_graphs_with(e, run_id);
        const Run& r = it->second;
        frees = r.field;
        frees.insert(frees.end(), r.kspec.begin(), r.kspec.end());
        bytes = r.bytes;
        leave_shape(e, r.key, &frees, &bytes);
        e->runs.erase(it);
    }
    free_and_return(e, frees, bytes);  // admissions waiting for memory recheck
    return VFD_OK;
}

int vfd_run_field_ptr(vfd_engine* e, int run_id, int field, void** d_field) {
    if (!e || !d_field) return set_error(VFD_ERR_INVALID, "run_field_ptr: null argument");
    std::lock_guard<std::mutex> lk(e->mu);
    auto it = e->runs.find(run_id);
    if (it == e->runs.end()) return set_error(VFD_ERR_NOTFOUND, "run_field_ptr: unknown run %d", run_id);
    if (field < 0 || field >= it->second.nfields) return set_error(VFD_ERR_INVALID, "run_field_ptr: bad field");
    *d_field = it->second.field[field];
    return VFD_OK;
}

static int run_copy(vfd_engine* e, int run_id, int field, void* dst, const void* src, bool in, void* stream) {
    if (!e || (in ? !src : !dst)) return set_error(VFD_ERR_INVALID, "run field copy: null argument");
    std::lock_guard<std::mutex> lk(e->mu);
    auto it = e->runs.find(run_id);
    if (it == e->runs.end()) return set_error(VFD_ERR_NOTFOUND, "run field copy: unknown run %d", run_id);
    if (field < 0 || field >= it->second.nfields) return set_error(VFD_ERR_INVALID, "run field copy: bad field");
    const size_t bytes = e->shapes[it->second.key].plan.field_bytes;
    void* f = it->second.field[field];
    return hip_check(hipMemcpyAsync(in ? f : dst, in ? src : f, bytes, hipMemcpyDefault, (hipStream_t)stream),
                     "run field copy");
}

int vfd_run_write_field(vfd_engine* e, int run_id, int field, const void* src, void* stream) {
    return run_copy(e, run_id, field, nullptr, src, true, stream);
}

int vfd_run_read_field(vfd_engine* e, int run_id, int field, void* dst, void* stream) {
    return run_copy(e, run_id, field, dst, nullptr, false, stream);
}

int vfd_step(vfd_engine* e, const int* run_ids, int nruns, int ticks, int use_graph, void* stream) {
    if (!e || (nruns > 0 && !run_ids) || nruns < 0 || ticks < 1) return set_error(VFD_ERR_INVALID, "step: bad arguments");
    std::lock_guard<std::mutex> lk(e->mu);
    if (e->serving) return set_error(VFD_ERR_STATE, "step: engine is serving; submit requests instead");
    std::vector<int> ids(run_ids, run_ids + nruns);
    std::sort(ids.begin(), ids.end());
    ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
    hipStream_t st = (hipStream_t)stream;
    if (!use_graph) {
        Batch bt;
        int rc = build_batch(e, ids, &bt);
        if (!rc) rc = launch_batch(bt, ticks, st, e->br);
        // the work buffers are reused by later work of this engine only after
        // this tick's launches: calls are ordered on `st`
        release_batch(&bt);
        return rc;
    }
    std::vector<int> key = ids;
    key.push_back(-ticks);
    auto it = e->graphs.find(key);
    if (it == e->graphs.end()) {
        if ((int)e->graphs.size() >= kMaxGraphs) {  // evict the least recently used graph
            auto lru = e->graphs.begin();
            for (auto g = e->graphs.begin(); g != e->graphs.end(); ++g)
                if (g->second.last_use < lru->second.last_use) lru = g;
            drop_graph(lru->second);
            e->graphs.erase(lru);
        }
        // capture on a private stream, then launch the instantiated graph on `st`
        hipStream_t cs = nullptr;
        int rc = hip_check(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "capture stream");
        GraphRec g;
        Batch bt;
        hipGraph_t graph = nullptr;
        if (!rc) rc = build_batch(e, ids, &bt);  // work buffers stay owned by the graph
        g.held = bt.held;
        if (!rc) rc = hip_check(hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed), "begin capture");
        if (!rc) {
            int lrc = launch_batch(bt, ticks, cs, e->br);
            hipError_t ce = hipStreamEndCapture(cs, &graph);
            rc = lrc ? lrc : hip_check(ce, "end capture");
        }
        if (!rc) rc = hip_check(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0), "graph instantiate");
        if (graph) (void)hipGraphDestroy(graph);
        if (cs) (void)hipStreamDestroy(cs);
        if (rc) {
            drop_graph(g);
            return rc;
        }
        it = e->graphs.emplace(key, g).first;
    }
    it->second.last_use = ++e->graph_clock;
    return hip_check(hipGraphLaunch(it->second.exec, st), "graph launch");
}

int vfd_step_plain(vfd_engine* e, const int* run_ids, int nruns, int ticks, void* stream) {
    if (!e || (nruns > 0 && !run_ids) || nruns < 0 || ticks < 1) return set_error(VFD_ERR_INVALID, "step: bad arguments");
    std::lock_guard<std::mutex> lk(e->mu);
    hipStream_t st = (hipStream_t)stream;
    for (int t = 0; t < ticks; ++t)
        for (int i = 0; i < nruns; ++i) {
            auto it = e->runs.find(run_ids[i]);
            if (it == e->runs.end()) return set_error(VFD_ERR_NOTFOUND, "step_plain: unknown run %d", run_ids[i]);
            Run& r = it->second;
            Shape& s = e->shapes[r.key];
            Held h;
            if (acquire(e, s, 1, &h) < 1) return set_error(VFD_ERR_BUDGET, "step_plain: no work buffer");
            for (int f = 0; f < r.nfields; ++f) {
                vfd_field fl{r.field[f], r.field[f], s.pool[h.idx[0]], r.kspec[f], 0.0, 0, 0};
                for (int pass = 0; pass < 5; ++pass)
                    if (int rc = plan_diffuse_passes(s.plan, &fl, 1, pass, pass + 1, st)) {
                        release(h);
                        return rc;
                    }
            }
            release(h);
        }
    return VFD_OK;
}

int vfd_serve_start(vfd_engine* e, int slots, int max_batch_fields) {
    if (!e || slots < 1 || slots > 16) return set_error(VFD_ERR_INVALID, "serve_start: slots must be 1..16");
    std::lock_guard<std::mutex> lk(e->mu);
    if (e->serving) return set_error(VFD_ERR_STATE, "serve_start: already serving");
    e->slots.assign(slots, Slot());
    for (Slot& s : e->slots) {
        if (int rc = hip_check(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), "slot stream")) return rc;
        if (int rc = hip_check(hipEventCreateWithFlags(&s.done, hipEventDisableTiming), "slot event")) return rc;
        if (int rc = s.br.create()) return rc;
    }
    e->max_batch_fields = max_batch_fields > 0 ? max_batch_fields : 256;
    e->stop = false;
    e->serving = true;
    e->server = std::thread(server_loop, e);
    return VFD_OK;
}

int vfd_serve_stop(vfd_engine* e) {
    if (!e) return set_error(VFD_ERR_INVALID, "serve_stop: null engine");
    {
        std::unique_lock<std::mutex> lk(e->mu);
        if (!e->serving) return VFD_OK;
        // drain: wait until every submitted request has completed
        e->cv.wait(lk, [&] {
            if (!e->pending.empty()) return false;
            for (Slot& s : e->slots)
                if (s.busy) return false;
            return true;
        });
        e->stop = true;
    }
    e->cv.notify_all();
    if (e->server.joinable()) e->server.join();
    std::lock_guard<std::mutex> lk(e->mu);
    for (Slot& s : e->slots) {
        (void)hipStreamSynchronize(s.stream);
        (void)hipStreamDestroy(s.stream);
        (void)hipEventDestroy(s.done);
        s.br.destroy();
    }
    e->slots.clear();
    e->serving = false;
    return VFD_OK;
}

int vfd_submit(vfd_engine* e, int run_id, uint64_t* ticket) {
    if (!e || !ticket) return set_error(VFD_ERR_INVALID, "submit: null argument");
    {
        std::lock_guard<std::mutex> lk(e->mu);
        if (!e->serving) return set_error(VFD_ERR_STATE, "submit: engine is not serving");
        if (!e->runs.count(run_id)) return set_error(VFD_ERR_NOTFOUND, "submit: unknown run %d", run_id);
        const uint64_t t = e->next_ticket++;
        e->pending.push_back(Request{t, run_id, Clock::now()});
        e->reqs[t] = ReqState{};
        *ticket = t;
    }
    e->cv.notify_all();
    return VFD_OK;
}

int vfd_poll(vfd_engine* e, uint64_t ticket, int* done, double* latency_ms) {
    if (!e || !done) return set_error(VFD_ERR_INVALID, "poll: null argument");
    std::lock_guard<std::mutex> lk(e->mu);
    auto it = e->reqs.find(ticket);
    if (it == e->reqs.end()) return set_error(VFD_ERR_NOTFOUND, "poll: unknown ticket");
    *done = it->second.done ? 1 : 0;
    if (latency_ms) *latency_ms = it->second.latency_ms;
    return VFD_OK;
}

int vfd_wait(vfd_engine* e, uint64_t ticket, double* latency_ms) {
    if (!e) return set_error(VFD_ERR_INVALID, "wait: null engine");
    std::unique_lock<std::mutex> lk(e->mu);
    auto it = e->reqs.find(ticket);
    if (it == e->reqs.end()) return set_error(VFD_ERR_NOTFOUND, "wait: unknown ticket");
    e->cv.wait(lk, [&] { return e->reqs[ticket].done || !e->serving; });
    if (!e->reqs[ticket].done) return set_error(VFD_ERR_STATE, "wait: serving stopped");
    if (latency_ms) *latency_ms = e->reqs[ticket].latency_ms;
    return VFD_OK;
}

int vfd_engine_get_stats(vfd_engine* e, vfd_engine_stats* s) {
    if (!e || !s) return set_error(VFD_ERR_INVALID, "engine_get_stats: null argument");
    std::lock_guard<std::mutex> lk(e->mu);
    s->budget_bytes = e->budget;
    s->used_bytes = e->used;
    s->runs = (int)e->runs.size();
    s->shapes = (int)e->shapes.size();
    s->batches = e->batches;
    s->ticks = e->ticks;
    s->requests_done = e->done_count;
    s->requests_pending = e->pending.size();
    s->peak_used_bytes = e->peak;
    s->runs_admitted_after_wait = e->admit_waits;
    return VFD_OK;
}

}  // extern "C"
