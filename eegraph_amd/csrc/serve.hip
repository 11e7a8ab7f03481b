// Serving loop over one device-resident graph with continuous batching.
//
// Traversal queries (BFS, SSSP, SSWP) go to one engine per algorithm.  An engine owns a
// vertex-major label plane of B columns (and the bit masks of batch.hip); each query occupies
// one column from admission until its answer has been read.  The engine thread loops:
//   1. reset the columns given back since the last round (labels to their initial value,
//      bits cleared), then admit waiting queries into free columns: seed the source label and
//      bits, and append the source to the frontier of the next iteration;
//   2. launch one iteration (BFS level or plane iteration) over every column in flight, with
//      a per-column activity mask written by the kernels;
//   3. with two iterations in flight, read the older one's activity: a column whose bit is
//      clear has converged (no vertex holds pending work for it), so its query is done: its
//      labels are gathered into a result buffer (original vertex order) and the column is
//      given back at once.
// Queries that arrive while others run therefore join the next iteration and share every
// launch with them.  A BFS column admitted before level L reports levels relative to L-1.
// The label planes are the admission limit: a query waits until its engine has a free column
// (label_budget_bytes sizes the engines).  Result buffers come from a pool of one per column;
// when every pooled buffer holds an unread answer, further answers go to host memory, so a
// client that reads answers in any order never blocks admission.
//
// Queries without an engine go to single-query lanes: each lane is a host thread with its own
// traversal state and stream, so several such queries run at once and overlap on the device
// with each other and with the engines.  PPR, PR and CC always take a lane; SSSP and SSWP
// take the label-plane engines only with EEG_SERVER_PLANE_ENGINES, since planes pay off when
// many queries sweep the same vertices together and queries admitted at different times
// rarely do; BFS takes its engine unless EEG_SERVER_NO_BFS_ENGINE.
//
// Every answer equals eeg_run's for the same query: the BFS/SSSP/SSWP fixpoints are unique
// and do not depend on which other columns share the iterations.
//
// No counterpart exists in the reference, whose drivers run one query per process
// (src/main.cu:101-121, src/eegraph.cu:2040-2247).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch.hpp"
#include "ppr_plane.hpp"

namespace eeg {
namespace {

using Clock = std::chrono::steady_clock;

enum TState { T_PENDING = 0, T_RUNNING, T_GATHER, T_DONE, T_CANCELLED };
enum EngKind { E_BFS = 0, E_SSSP = 1, E_SSWP = 2, E_PPR = 3, E_COUNT = 4 };

struct EvHold {  // one event shared by the answers gathered in one round
    hipEvent_t e = nullptr;
    ~EvHold() {
        // the last holder may be a query that was never waited for: its record may still be
        // pending on the engine stream, and an event is not destroyed while one is
        if (e) {
            (void)hipEventSynchronize(e);
            (void)hipEventDestroy(e);
        }
    }
};

struct Ticket {
    uint64_t id = 0;
    uint32_t algo = 0, source = 0;
    TState state = T_PENDING;
    eeg_status status = EEG_OK;
    int engine = -1;  // -1: single-query lane
    int col = -1;
    uint32_t offset = 0;    // BFS: level before the first level of this column
    uint32_t first_it = 0;  // first iteration that includes this column
    uint32_t batch = 0;     // most columns in flight in an iteration this query took part in
    bool cancel = false;
    double t_submit = 0, t_start = 0, t_done = 0;
    std::vector<uint32_t> host;  // answer in host memory (lane, or result pool exhausted)
    std::string err;             // why the query failed
    DevMem* res = nullptr;       // answer in a pooled device buffer
    std::shared_ptr<EvHold> ev;  // completes when `res` holds the answer
};
using TicketP = std::shared_ptr<Ticket>;

// Device control block of an engine: Ctl followed by the 4 x BATCH_MAXW activity words.
struct EngCtl {
    Ctl ctl;
    uint64_t act[4 * BATCH_MAXW];
};

}  // namespace
}  // namespace eeg

struct eeg_server;

namespace eeg {
namespace {

// What every batching engine shares: its thread, its queue of waiting queries, the occupant of
// each column and the pool of result buffers answers are gathered into.
struct EngineBase {
    eeg_server* srv = nullptr;
    int kind = E_BFS;
    uint32_t n = 0, K = 8;
    uint32_t ncols = 0;
    hipStream_t st = nullptr;
    std::thread th;
    std::condition_variable cv;
    std::deque<TicketP> pending;      // guarded by the server mutex
    std::vector<TicketP> col;         // occupant of each column
    uint32_t running = 0;             // columns in T_RUNNING
    bool broken = false;
    std::vector<std::unique_ptr<DevMem>> res_all;  // result pool (at most ncols buffers)
    std::vector<DevMem*> res_free;                 // guarded by the server mutex
    virtual void loop() = 0;
    virtual ~EngineBase() {
        if (st) (void)hipStreamDestroy(st);
    }
};

struct Engine : EngineBase {
    uint32_t W = 0, B = 0, qcap = 0;
    uint32_t R = 0;  // mask record stride: BFS 4W (seen, next, visit, stamp per vertex), planes W
    std::vector<uint8_t> released;    // column given back, reset pending
    // device state
    DevMem planes, m1, m3, q[2], s[2], cand, stamp, ctlbuf, stage, spill;
    EngCtl* hctl = nullptr;           // pinned, 4 slots
    uint32_t* hstage = nullptr;       // pinned, 4 slots of stage_words
    size_t stage_words = 0;
    hipEvent_t ev[4] = {};
    uint32_t it = 1;                  // next iteration (BFS: level)
    uint32_t base = 1;                // stamp base
    unsigned grid = 0;                // resident workgroups of the iteration kernels
    // The iteration counter only grows while queries run: BFS levels are stored as
    // it - 1 + depth and stamps as base + it.  From kItLimit on, no query is admitted until
    // the engine drains; then every column is reset and the counter restarts at 1, long before
    // a level or stamp could reach 2^32.
    static constexpr uint32_t kItLimit = 0xf0000000u;

    uint32_t fill() const { return kind == E_SSWP ? 0u : INF; }
    uint32_t seed_val() const { return kind == E_SSWP ? INF : 0u; }
    eeg_status init(eeg_graph* g);
    eeg_status reset_device();
    void loop() override;
    eeg_status launch_iteration(eeg_graph* g, uint32_t j);
    bool gather(std::vector<std::pair<TicketP, uint32_t>>& fin);
    ~Engine() override;
};

// Personalised PageRank engine: B columns of rank planes (ppr_plane.hip).  A query occupies a
// column from admission until its own L1 change drops below the server's epsilon; columns join
// and leave between iterations, every iteration is one gather pass for all columns in flight.
struct PprEngine : EngineBase {
    PprPlane pl;
    PprCols cur{};                    // restart vertex of every column (PPR_IDLE when free)
    double* hres = nullptr;           // pinned, 4 slots of 32 residuals
    hipEvent_t ev[4] = {};
    uint32_t it = 1;
    eeg_status init(eeg_graph* g, uint32_t B, float damping);
    void loop() override;
    bool gather(std::vector<std::pair<TicketP, uint32_t>>& fin, uint32_t j);
    ~PprEngine() override;
};

// A single-query lane: a host thread with its own traversal state and stream, so that several
// queries run at once on one graph and their launches overlap on the device.
struct Lane {
    Work w;
    hipStream_t st = nullptr;
    DevMem out;
    std::shared_ptr<BatchWork> bw;  // multi-source passes of this lane (made on first use)
    std::thread th;
    ~Lane() {
        if (st) (void)hipStreamDestroy(st);
    }
};

}  // namespace
}  // namespace eeg

struct eeg_server {
    eeg_graph* g = nullptr;
    eeg_server_cfg cfg{};
    uint32_t K = 8;
    bool discard = false;
    std::mutex mu;
    std::condition_variable cv_done;
    std::unordered_map<uint64_t, eeg::TicketP> tickets;
    uint64_t next_id = 1;
    eeg::Clock::time_point t0;
    std::unique_ptr<eeg::EngineBase> eng[eeg::E_COUNT];
    bool use_eng[eeg::E_COUNT] = {false, false, false, false};  // algorithms served by a batching engine
    bool stop = false;
    // single-query lanes
    std::deque<eeg::TicketP> lane_q;
    std::condition_variable lane_cv;
    uint32_t lanes_busy = 0;  // lanes running a query (guarded by mu)
    bool lane_batch_off = false;  // a lane's multi-source pass found no memory (guarded by mu)
    std::vector<std::unique_ptr<eeg::Lane>> lanes;
    std::mutex prep_mu;  // builds of the derived tables a lane needs
    // answer extraction (labels of a finished column to the caller)
    std::mutex io_mu;
    hipStream_t io_st = nullptr;
    eeg::DevMem io_buf;

    double now_ms() const {
        return std::chrono::duration<double, std::milli>(eeg::Clock::now() - t0).count();
    }
};

namespace eeg {
namespace {

Engine::~Engine() {
    if (st) (void)hipStreamSynchronize(st);
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    if (hctl) (void)hipHostFree(hctl);
    if (hstage) (void)hipHostFree(hstage);
}

__global__ void k_eng_fill32(uint32_t* a, uint64_t n, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        a[i] = v;
}

unsigned grid_sv(uint64_t work) {
    uint64_t g = (work + WG - 1) / WG;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(g, 65536));
}

eeg_status Engine::init(eeg_graph* g) {
    n = g->n;
    grid = (unsigned)batch_cu_count(g->device) * 16;
    W = (ncols + 63) / 64;
    B = ncols;
    qcap = n + ncols + 16;
    EEG_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    for (auto& e : ev) EEG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    const uint64_t nb = (uint64_t)n * B, nw = (uint64_t)n * W;
    EEG_TRY(planes.alloc(nb * 4 + 4));
    R = (kind == E_BFS) ? 4 * W : W;
    EEG_TRY(m1.alloc((uint64_t)n * R * 8 + 8));
    if (kind != E_BFS) {
        EEG_TRY(m3.alloc(nw * 8 + 8));
        EEG_TRY(stamp.alloc((size_t)n * 4 + 4));
    }
    for (int i = 0; i < 2; ++i) {
        EEG_TRY(q[i].alloc((size_t)qcap * 4));
        EEG_TRY(s[i].alloc((size_t)qcap * 4));
    }
    if (kind == E_BFS) EEG_TRY(cand.alloc((size_t)qcap * 4));
    EEG_TRY(spill.alloc((size_t)n * 4 + 4));
    EEG_TRY(ctlbuf.alloc(sizeof(EngCtl)));
    stage_words = 3 * (size_t)ncols + 4;
    EEG_TRY(stage.alloc(4 * stage_words * 4));
    EEG_HIP(hipHostMalloc((void**)&hctl, 4 * sizeof(EngCtl), hipHostMallocDefault));
    EEG_HIP(hipHostMalloc((void**)&hstage, 4 * stage_words * 4, hipHostMallocDefault));
    EEG_TRY(reset_device());
    EEG_HIP(hipStreamSynchronize(st));
    col.assign(ncols, nullptr);
    released.assign(ncols, 0);
    return EEG_OK;
}

// Every column to its initial labels, bits and stamps cleared, control block drained (on the
// engine stream, after everything queued before).
eeg_status Engine::reset_device() {
    const uint64_t nb = (uint64_t)n * B;
    k_eng_fill32<<<grid_sv(nb), WG, 0, st>>>(planes.as<uint32_t>(), nb, fill());
    EEG_HIP(hipGetLastError());
    EEG_HIP(hipMemsetAsync(m1.p, 0, m1.bytes, st));
    if (m3.p) EEG_HIP(hipMemsetAsync(m3.p, 0, m3.bytes, st));
    if (stamp.p) EEG_HIP(hipMemsetAsync(stamp.p, 0, stamp.bytes, st));
    EEG_HIP(hipMemsetAsync(ctlbuf.p, 0, sizeof(EngCtl), st));
    // a drained engine: the first admission starts from empty frontier slots
    k_eng_fill32<<<1, 64, 0, st>>>(&ctlbuf.as<EngCtl>()->ctl.done, 1, 1u);
    EEG_HIP(hipGetLastError());
    return EEG_OK;
}

eeg_status Engine::launch_iteration(eeg_graph* g, uint32_t j) {
    const Tigr& t = g->tout[k_index(K)];
    EngCtl* ec = ctlbuf.as<EngCtl>();
    const int cur = (int)((j - 1) & 1);
    if (kind == E_BFS) {
        MsArgs a{};
        a.row = g->out.rowp();
        a.col = g->out.colp();
        a.vnc = t.vnc.as<uint32_t>();
        uint64_t* rec = m1.as<uint64_t>();
        a.seen = rec;
        a.next = rec + W;
        a.visit = rec + 2 * W;
        a.stamp = reinterpret_cast<uint32_t*>(rec + 3 * W);
        a.cand = cand.as<uint32_t>();
        a.dist = planes.as<uint32_t>();
        a.ctl = &ec->ctl;
        a.act = ec->act;
        a.W = W;
        a.B = B;
        a.n = n;
        a.R = R;
        a.sst = 2 * R;
        a.qcap = qcap;
        a.level = j;
        a.stampv = base + j;
        a.qin = q[cur].as<uint32_t>();
        a.sin = s[cur].as<uint32_t>();
        a.qout = q[cur ^ 1].as<uint32_t>();
        a.sout = s[cur ^ 1].as<uint32_t>();
        return launch_ms_level(a, K, grid, st);
    }
    PlaneArgs a{};
    a.row = g->out.rowp();
    a.col = g->out.colp();
    a.wgt = g->out.wp();
    a.vnc = t.vnc.as<uint32_t>();
    a.lab = planes.as<uint32_t>();
    a.chg = m1.as<uint64_t>();
    a.nchg = m3.as<uint64_t>();
    a.stamp = stamp.as<uint32_t>();
    a.ctl = &ec->ctl;
    a.act = ec->act;
    a.W = W;
    a.B = B;
    a.n = n;
    a.qcap = qcap;
    a.it = j;
    a.stampv = base + j;
    a.widest = kind == E_SSWP ? 1 : 0;
    a.qin = q[cur].as<uint32_t>();
    a.sin = s[cur].as<uint32_t>();
    a.qout = q[cur ^ 1].as<uint32_t>();
    a.sout = s[cur ^ 1].as<uint32_t>();
    return launch_plane_iteration(a, K, grid, st, running);
}

void fail_running(eeg_server* srv, EngineBase* e, eeg_status why) {
    // caller holds srv->mu
    for (uint32_t c = 0; c < e->ncols; ++c) {
        TicketP& t = e->col[c];
        if (t && t->state == T_RUNNING) {
            t->state = T_DONE;
            t->status = why;
            t->t_done = srv->now_ms();
        }
    }
    for (auto& t : e->pending) {
        t->state = T_DONE;
        t->status = why;
    }
    e->pending.clear();
    e->broken = true;
    srv->cv_done.notify_all();
}

void Engine::loop() {
    eeg_graph* g = srv->g;
    (void)hipSetDevice(g->device);
    const Tigr& t = g->tout[k_index(K)];
    std::deque<uint32_t> inflight;
    EngCtl* ec = ctlbuf.as<EngCtl>();
    while (true) {
        std::vector<uint32_t> pv, pc;
        uint64_t clr[BATCH_MAXW] = {0, 0, 0, 0};
        bool any_clr = false, restart = false;
        uint32_t active = 0;
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            auto count_free = [&] { return (size_t)std::count(col.begin(), col.end(), nullptr); };
            auto count_released = [&] { return (size_t)std::count(released.begin(), released.end(), (uint8_t)1); };
            while (!srv->stop && inflight.empty() && running == 0 && count_released() == 0 &&
                   (pending.empty() || count_free() == 0))
                cv.wait(lk);
            if (srv->stop) break;
            if (it >= kItLimit && running == 0 && inflight.empty()) {
                // drained past the counter limit: every column is free or given back (its
                // answer already gathered), so all of them are reset and the counter restarts
                for (uint32_t c = 0; c < ncols; ++c) {
                    released[c] = 0;
                    col[c] = nullptr;
                }
                it = 1;
                restart = true;
            }
            // Columns given back are reset (one pass over the plane for all of them, in stream
            // order before any reuse) only when waiting queries need them or the engine is idle.
            const size_t nfree = count_free(), nrel = count_released();
            if (nrel && (pending.size() > nfree || (running == 0 && pending.empty()))) {
                for (uint32_t c = 0; c < ncols; ++c)
                    if (released[c]) {
                        released[c] = 0;
                        col[c] = nullptr;
                        clr[c / 64] |= 1ull << (c % 64);
                        any_clr = true;
                    }
            }
            uint32_t c = 0;
            const double tnow = srv->now_ms();
            while (!pending.empty() && !broken && it < kItLimit) {
                while (c < ncols && col[c]) ++c;
                if (c >= ncols) break;
                TicketP tk = pending.front();
                pending.pop_front();
                if (tk->cancel) {
                    tk->state = T_CANCELLED;
                    tk->status = EEG_ERR_CANCELLED;
                    srv->cv_done.notify_all();
                    continue;
                }
                col[c] = tk;
                tk->col = (int)c;
                tk->state = T_RUNNING;
                tk->t_start = tnow;
                tk->first_it = it;
                tk->offset = (kind == E_BFS) ? it - 1 : 0u;
                pv.push_back(g->h_perm[tk->source]);
                pc.push_back(c);
                ++running;
            }
            active = running;
            for (uint32_t k = 0; k < ncols; ++k)
                if (col[k] && col[k]->state == T_RUNNING) col[k]->batch = std::max(col[k]->batch, active);
        }
        eeg_status rc = EEG_OK;
        if (restart) rc = reset_device();
        if (rc == EEG_OK && any_clr) {
            // a converged plane column has no change bits left; a BFS column keeps its seen
            // bits and visit bits on vertices without out-edges
            rc = launch_release(n, W, R, B, clr, planes.as<uint32_t>(), fill(),
                                kind == E_BFS ? m1.as<uint64_t>() : nullptr,                  // seen
                                kind == E_BFS ? m1.as<uint64_t>() + 2 * W : nullptr, nullptr,  // visit
                                st);
        }
        if (rc == EEG_OK && !pv.empty()) {
            uint32_t* hs = hstage + (size_t)(it % 4) * stage_words;
            uint32_t* ds = stage.as<uint32_t>() + (size_t)(it % 4) * stage_words;
            const uint32_t np = (uint32_t)pv.size();
            std::vector<uint32_t> uniq(pv);
            std::sort(uniq.begin(), uniq.end());
            uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
            memcpy(hs, pv.data(), np * 4);
            memcpy(hs + np, pc.data(), np * 4);
            memcpy(hs + 2 * np, uniq.data(), uniq.size() * 4);
            const size_t words = 2 * (size_t)np + uniq.size();
            rc = hipMemcpyAsync(ds, hs, words * 4, hipMemcpyHostToDevice, st) == hipSuccess ? EEG_OK : EEG_ERR_DEVICE;
            const int in_slot = (kind == E_BFS) ? (int)((2 * it - 2) % 3) : (int)((it - 1) % 3);
            const int cur = (int)((it - 1) & 1);
            if (rc == EEG_OK)
                rc = launch_admit(&ec->ctl, in_slot, ds, ds + np, np, ds + 2 * np, (uint32_t)uniq.size(),
                                  t.vnc.as<uint32_t>(), R, B, kind == E_BFS ? m1.as<uint64_t>() : nullptr,
                                  kind == E_BFS ? m1.as<uint64_t>() + 2 * W : m1.as<uint64_t>(), planes.as<uint32_t>(),
                                  kind == E_BFS ? it - 1 : seed_val(),  // BFS: levels are global
                                  q[cur].as<uint32_t>(), s[cur].as<uint32_t>(), qcap, st);
        }
        if (rc == EEG_OK && active > 0) {
            rc = launch_iteration(g, it);
            if (rc == EEG_OK &&
                (hipMemcpyAsync(&hctl[it % 4], ec, sizeof(EngCtl), hipMemcpyDeviceToHost, st) != hipSuccess ||
                 hipEventRecord(ev[it % 4], st) != hipSuccess))
                rc = EEG_ERR_DEVICE;
            if (rc == EEG_OK) inflight.push_back(it++);
        }
        if (rc != EEG_OK) {
            std::lock_guard<std::mutex> lk(srv->mu);
            fail_running(srv, this, rc);
            break;
        }
        // inspect the older iteration while the newer one runs
        while (!inflight.empty() && (inflight.size() >= 2 || active == 0)) {
            const uint32_t j = inflight.front();
            inflight.pop_front();
            if (hipEventSynchronize(ev[j % 4]) != hipSuccess) {
                std::lock_guard<std::mutex> lk(srv->mu);
                fail_running(srv, this, fail(EEG_ERR_DEVICE, "serving engine: device error"));
                return;
            }
            const EngCtl& h = hctl[j % 4];
            std::vector<std::pair<TicketP, uint32_t>> fin;
            {
                std::lock_guard<std::mutex> lk(srv->mu);
                if (h.ctl.err) {
                    fail_running(srv, this, fail(EEG_ERR_DEVICE, "serving engine: frontier overflow (err=%u)", h.ctl.err));
                    return;
                }
                const uint64_t* act = h.act + (j % 4) * BATCH_MAXW;
                const double tnow = srv->now_ms();
                bool any = false;
                for (uint32_t c = 0; c < ncols; ++c) {
                    TicketP& tk = col[c];
                    if (!tk || tk->state != T_RUNNING || tk->first_it > j) continue;
                    if ((act[c / 64] >> (c % 64)) & 1ull) continue;
                    --running;
                    if (srv->discard || tk->cancel) {
                        tk->state = T_DONE;
                        tk->status = EEG_OK;
                        tk->t_done = tnow;
                        released[c] = 1;
                        any = true;
                    } else {
                        tk->state = T_GATHER;
                        fin.emplace_back(tk, c);
                    }
                }
                if (any) srv->cv_done.notify_all();
            }
            if (!fin.empty() && !gather(fin)) return;
        }
    }
}

// Answers of converged columns to result buffers (or host memory), then the columns are given
// back.  Runs on the engine stream after the iterations already queued, none of which writes
// a converged column.
bool Engine::gather(std::vector<std::pair<TicketP, uint32_t>>& fin) {
    eeg_graph* g = srv->g;
    auto ev = std::make_shared<EvHold>();
    eeg_status rc = hipEventCreateWithFlags(&ev->e, hipEventDisableTiming) == hipSuccess ? EEG_OK : EEG_ERR_DEVICE;
    std::vector<DevMem*> bufs(fin.size(), nullptr);
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        for (size_t i = 0; i < fin.size(); ++i) {
            if (!res_free.empty()) {
                bufs[i] = res_free.back();
                res_free.pop_back();
            } else if (res_all.size() < ncols) {
                res_all.emplace_back(new DevMem());
                bufs[i] = res_all.back().get();
                bufs[i]->alloc((size_t)n * 4 + 4);
                if (!bufs[i]->p) bufs[i] = nullptr;  // out of memory: host copy below
            }
        }
    }
    for (size_t i = 0; i < fin.size() && rc == EEG_OK; ++i) {
        TicketP& tk = fin[i].first;
        const uint32_t c = fin[i].second;
        if (bufs[i]) {
            rc = launch_extract(planes.as<uint32_t>(), B, c, g->perm.as<uint32_t>(), n, tk->offset,
                                bufs[i]->as<uint32_t>(), st);
        } else {
            tk->host.resize(n);
            rc = launch_extract(planes.as<uint32_t>(), B, c, g->perm.as<uint32_t>(), n, tk->offset,
                                spill.as<uint32_t>(), st);
            if (rc == EEG_OK &&
                (hipMemcpyAsync(tk->host.data(), spill.p, (size_t)n * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
                 hipStreamSynchronize(st) != hipSuccess))
                rc = EEG_ERR_DEVICE;
        }
    }
    if (rc == EEG_OK && hipEventRecord(ev->e, st) != hipSuccess) rc = EEG_ERR_DEVICE;
    std::lock_guard<std::mutex> lk(srv->mu);
    const double tnow = srv->now_ms();
    for (size_t i = 0; i < fin.size(); ++i) {
        TicketP& tk = fin[i].first;
        tk->state = T_DONE;
        tk->status = rc == EEG_OK ? EEG_OK : fail(EEG_ERR_DEVICE, "gathering an answer failed");
        tk->t_done = tnow;
        tk->res = bufs[i];
        tk->ev = ev;
        if (tk->cancel && tk->res) {  // cancelled while gathering: nobody will read it
            res_free.push_back(tk->res);
            tk->res = nullptr;
        }
        released[fin[i].second] = 1;
    }
    srv->cv_done.notify_all();
    if (rc != EEG_OK) {
        fail_running(srv, this, EEG_ERR_DEVICE);
        return false;
    }
    return true;
}

PprEngine::~PprEngine() {
    if (st) (void)hipStreamSynchronize(st);
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    if (hres) (void)hipHostFree(hres);
}

eeg_status PprEngine::init(eeg_graph* g, uint32_t B, float damping) {
    n = g->n;
    ncols = B;
    EEG_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    for (auto& e : ev) EEG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    EEG_HIP(hipHostMalloc((void**)&hres, 4 * 32 * sizeof(double), hipHostMallocDefault));
    EEG_TRY(pl.init(g, B, damping, st));
    for (uint32_t q = 0; q < 32; ++q) cur.src[q] = PPR_IDLE;
    cur.fresh = 0;
    col.assign(ncols, nullptr);
    return EEG_OK;
}

void PprEngine::loop() {
    eeg_graph* g = srv->g;
    (void)hipSetDevice(g->device);
    const double eps = srv->cfg.ppr_epsilon;
    constexpr uint32_t kMaxIters = 10000;
    std::deque<uint32_t> inflight;
    while (true) {
        uint32_t fresh = 0, active = 0;
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            while (!srv->stop && inflight.empty() && running == 0 && pending.empty()) cv.wait(lk);
            if (srv->stop) break;
            const double tnow = srv->now_ms();
            for (uint32_t c = 0; c < ncols && !pending.empty() && !broken; ++c) {
                if (col[c]) continue;
                TicketP tk;
                while (!pending.empty() && !tk) {
                    tk = pending.front();
                    pending.pop_front();
                    if (tk->cancel) {
                        tk->state = T_CANCELLED;
                        tk->status = EEG_ERR_CANCELLED;
                        srv->cv_done.notify_all();
                        tk = nullptr;
                    }
                }
                if (!tk) break;
                col[c] = tk;
                tk->col = (int)c;
                tk->state = T_RUNNING;
                tk->t_start = tnow;
                tk->first_it = it;
                cur.src[c] = g->h_perm[tk->source];
                fresh |= 1u << c;
                ++running;
            }
            active = running;
            for (uint32_t k = 0; k < ncols; ++k)
                if (col[k] && col[k]->state == T_RUNNING) col[k]->batch = std::max(col[k]->batch, active);
        }
        eeg_status rc = EEG_OK;
        if (active > 0) {
            PprCols cols = cur;
            cols.fresh = fresh;
            // columns are taken lowest first: the pass covers only up to the highest busy one
            rc = pl.iterate(g, it, cols, st, pl.width_for(cols));
            if (rc == EEG_OK &&
                (hipMemcpyAsync(hres + (it % 4) * 32, pl.resid.p, (size_t)ncols * 8, hipMemcpyDeviceToHost, st) !=
                     hipSuccess ||
                 hipEventRecord(ev[it % 4], st) != hipSuccess))
                rc = EEG_ERR_DEVICE;
            if (rc == EEG_OK) inflight.push_back(it++);
        }
        if (rc != EEG_OK) {
            std::lock_guard<std::mutex> lk(srv->mu);
            fail_running(srv, this, rc);
            break;
        }
        // inspect the older iteration while the newer one runs: a column whose L1 change is
        // below epsilon has converged; its ranks (the buffer that iteration wrote, intact until
        // the iteration after the next) are gathered and the column is free for the next query
        while (!inflight.empty() && (inflight.size() >= 2 || active == 0)) {
            const uint32_t j = inflight.front();
            inflight.pop_front();
            if (hipEventSynchronize(ev[j % 4]) != hipSuccess) {
                std::lock_guard<std::mutex> lk(srv->mu);
                fail_running(srv, this, fail(EEG_ERR_DEVICE, "PPR engine: device error"));
                return;
            }
            const double* r = hres + (j % 4) * 32;
            std::vector<std::pair<TicketP, uint32_t>> fin;
            {
                std::lock_guard<std::mutex> lk(srv->mu);
                const double tnow = srv->now_ms();
                bool any = false;
                for (uint32_t c = 0; c < ncols; ++c) {
                    TicketP& tk = col[c];
                    if (!tk || tk->state != T_RUNNING || tk->first_it > j) continue;
                    const bool conv = r[c] < eps;
                    if (!conv && j - tk->first_it + 1 < kMaxIters) continue;
                    --running;
                    if (!conv) tk->err = "PPR residual above epsilon after the iteration limit";
                    if (srv->discard || tk->cancel || !conv) {
                        tk->state = T_DONE;
                        tk->status = conv ? EEG_OK : EEG_ERR_NOT_CONVERGED;
                        tk->t_done = tnow;
                        cur.src[c] = PPR_IDLE;
                        tk = nullptr;
                        any = true;
                    } else {
                        tk->state = T_GATHER;
                        fin.emplace_back(tk, c);
                    }
                }
                if (any) srv->cv_done.notify_all();
            }
            if (!fin.empty() && !gather(fin, j)) return;
        }
    }
}

bool PprEngine::gather(std::vector<std::pair<TicketP, uint32_t>>& fin, uint32_t j) {
    eeg_graph* g = srv->g;
    auto evh = std::make_shared<EvHold>();
    eeg_status rc = hipEventCreateWithFlags(&evh->e, hipEventDisableTiming) == hipSuccess ? EEG_OK : EEG_ERR_DEVICE;
    std::vector<DevMem*> bufs(fin.size(), nullptr);
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        for (size_t i = 0; i < fin.size(); ++i) {
            if (!res_free.empty()) {
                bufs[i] = res_free.back();
                res_free.pop_back();
            } else {
                res_all.emplace_back(new DevMem());
                bufs[i] = res_all.back().get();
                bufs[i]->alloc((size_t)n * 4 + 4);
                if (!bufs[i]->p) bufs[i] = nullptr;
            }
        }
    }
    for (size_t i = 0; i < fin.size() && rc == EEG_OK; ++i) {
        TicketP& tk = fin[i].first;
        const uint32_t c = fin[i].second;
        if (bufs[i]) {
            rc = pl.extract(g, j, c, cur.src[c], bufs[i]->as<float>(), st);
        } else {  // no device memory left for a pooled answer: host copy through the engine stream
            DevMem tmp;
            rc = tmp.alloc((size_t)n * 4 + 4);
            tk->host.resize(n);
            if (rc == EEG_OK) rc = pl.extract(g, j, c, cur.src[c], tmp.as<float>(), st);
            if (rc == EEG_OK &&
                (hipMemcpyAsync(tk->host.data(), tmp.p, (size_t)n * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
                 hipStreamSynchronize(st) != hipSuccess))
                rc = EEG_ERR_DEVICE;
        }
    }
    if (rc == EEG_OK && hipEventRecord(evh->e, st) != hipSuccess) rc = EEG_ERR_DEVICE;
    std::lock_guard<std::mutex> lk(srv->mu);
    const double tnow = srv->now_ms();
    for (size_t i = 0; i < fin.size(); ++i) {
        TicketP& tk = fin[i].first;
        const uint32_t c = fin[i].second;
        tk->state = T_DONE;
        tk->status = rc == EEG_OK ? EEG_OK : fail(EEG_ERR_DEVICE, "gathering an answer failed");
        tk->t_done = tnow;
        tk->res = bufs[i];
        tk->ev = evh;
        if (tk->cancel && tk->res) {
            res_free.push_back(tk->res);
            tk->res = nullptr;
        }
        cur.src[c] = PPR_IDLE;  // the column takes a new query from the next iteration on
        col[c] = nullptr;
    }
    srv->cv_done.notify_all();
    if (rc != EEG_OK) {
        fail_running(srv, this, EEG_ERR_DEVICE);
        return false;
    }
    return true;
}

// Tables a query reads, built once (under the graph's lock) before any lane uses them.
eeg_status prepare_for(eeg_server* srv, uint32_t algo) {
    eeg_graph* g = srv->g;
    std::lock_guard<std::mutex> lk(g->mu);
    if (algo == EEG_CC) {
        EEG_TRY(ensure_sym(g));
        return ensure_tigr(g, g->sym, g->tsym, srv->K);
    }
    if (algo == EEG_PR || algo == EEG_PPR) {
        EEG_TRY(ensure_in(g));
        EEG_TRY(ensure_tigr(g, g->in, g->tin, srv->K));
        return ensure_pr_span(g, srv->K);
    }
    return ensure_tigr(g, g->out, g->tout, srv->K);
}

// Up to 256 waiting SSSP/SSWP queries answered in one multi-source pass on the lane's own
// state; false when the pass found no device memory (the queries are handed back to the queue).
bool lane_batch(eeg_server* srv, Lane* ln, std::vector<TicketP>& batch) {
    eeg_graph* g = srv->g;
    const uint32_t a = batch[0]->algo, k = (uint32_t)batch.size(), n = g->n;
    eeg_opts o;
    memset(&o, 0, sizeof(o));
    o.variant = EEG_SYNC_PUSH_DD;
    o.K = srv->K;
    std::vector<uint32_t> srcs(k);
    for (uint32_t i = 0; i < k; ++i) srcs[i] = batch[i]->source;
    std::vector<std::vector<uint32_t>> outs(k);
    std::vector<uint32_t*> rows(k, nullptr);
    eeg_status rc = EEG_OK;
    if (!srv->discard) {
        try {  // k answers of n words each wait in host memory for their clients
            for (uint32_t i = 0; i < k; ++i) {
                outs[i].resize(n);
                rows[i] = outs[i].data();
            }
        } catch (const std::bad_alloc&) {
            outs.clear();
            outs.resize(k);
            rc = fail(EEG_ERR_OOM, "no host memory for %u answers", k);
        }
    }
    if (rc == EEG_OK) rc = prepare_for(srv, a);
    if (rc == EEG_OK && !ln->bw) ln->bw = make_batch_work();
    if (rc == EEG_OK) rc = run_batch_rows(g, *ln->bw, ln->w, ln->st, a, srcs.data(), k, o, rows.data(), nullptr);
    std::lock_guard<std::mutex> lk(srv->mu);
    --srv->lanes_busy;
    if (rc == EEG_ERR_OOM) {  // device or host memory ran short: one at a time from now on
        ln->bw.reset();
        srv->lane_batch_off = true;
        for (auto it = batch.rbegin(); it != batch.rend(); ++it) {
            (*it)->state = T_PENDING;
            srv->lane_q.push_front(*it);
        }
        srv->lane_cv.notify_all();
        return false;
    }
    const double tnow = srv->now_ms();
    const std::string err = rc != EEG_OK ? eeg_last_error() : "";
    for (uint32_t i = 0; i < k; ++i) {
        TicketP& tk = batch[i];
        tk->status = rc;
        if (rc != EEG_OK) tk->err = err;
        tk->state = T_DONE;
        tk->t_done = tnow;
        tk->host.swap(outs[i]);
    }
    srv->cv_done.notify_all();
    return true;
}

void lane_loop(eeg_server* srv, Lane* ln) {
    eeg_graph* g = srv->g;
    (void)hipSetDevice(g->device);
    // a multi-source pass costs several single queries, even with few sources, so a lane
    // batches only a queue of 32 or more (RMAT-26 mix: 17 raised the SSSP p50 at 200 offered
    // q/s, 32 served 1.90-2.16x at 200-1000 q/s against 1.80-2.02x for 64)
    constexpr size_t kLaneBatchMin = 32, kLaneBatchMax = 64 * BATCH_MAXW;
    while (true) {
        TicketP tk;
        std::vector<TicketP> batch;
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            srv->lane_cv.wait(lk, [&] { return srv->stop || !srv->lane_q.empty(); });
            if (srv->stop) return;
            tk = srv->lane_q.front();
            srv->lane_q.pop_front();
            if (tk->cancel) {
                tk->state = T_CANCELLED;
                tk->status = EEG_ERR_CANCELLED;
                srv->cv_done.notify_all();
                continue;
            }
            tk->state = T_RUNNING;
            tk->t_start = srv->now_ms();
            tk->batch = 1;
            ++srv->lanes_busy;
            if ((tk->algo == EEG_SSSP || tk->algo == EEG_SSWP) && !srv->lane_batch_off &&
                (srv->cfg.flags & EEG_SERVER_NO_LANE_BATCH) == 0) {
                size_t same = 1;
                for (const TicketP& t : srv->lane_q)
                    if (t->algo == tk->algo && !t->cancel) ++same;
                if (same >= kLaneBatchMin) {
                    batch.push_back(tk);
                    bool cancelled = false;
                    for (auto it = srv->lane_q.begin(); it != srv->lane_q.end() && batch.size() < kLaneBatchMax;) {
                        if ((*it)->algo != tk->algo) {
                            ++it;
                            continue;
                        }
                        TicketP t = *it;
                        it = srv->lane_q.erase(it);
                        if (t->cancel) {
                            t->state = T_CANCELLED;
                            t->status = EEG_ERR_CANCELLED;
                            cancelled = true;
                            continue;
                        }
                        t->state = T_RUNNING;
                        t->t_start = tk->t_start;
                        batch.push_back(t);
                    }
                    for (TicketP& t : batch) t->batch = (uint32_t)batch.size();
                    if (cancelled) srv->cv_done.notify_all();
                }
            }
        }
        if (!batch.empty()) {
            lane_batch(srv, ln, batch);
            continue;
        }
        eeg_opts o;
        memset(&o, 0, sizeof(o));
        o.variant = EEG_SYNC_PUSH_DD;
        o.K = srv->K;
        // a partial resident grid leaves room on every CU for the other lanes and the engines
        o.wgs_per_cu = srv->cfg.lane_wgs_per_cu ? srv->cfg.lane_wgs_per_cu : 16;
        o.damping = srv->cfg.damping;
        o.epsilon = srv->cfg.ppr_epsilon;
        const uint32_t a = tk->algo;
        const bool sourced = a == EEG_BFS || a == EEG_SSSP || a == EEG_SSWP || a == EEG_PPR;
        const uint32_t isrc = sourced ? g->h_perm[tk->source] : 0u;
        std::vector<uint32_t> out;
        eeg_status rc = prepare_for(srv, a);
        if (rc == EEG_OK)
            rc = (a == EEG_PR || a == EEG_PPR) ? run_pagerank(g, a, isrc, o, nullptr, &ln->w, ln->st)
                                               : run_traversal(g, a, isrc, o, nullptr, nullptr, 0, &ln->w, ln->st);
        if (rc == EEG_OK && !srv->discard) {  // answer in original vertex order
            out.resize(g->n);
            rc = ln->out.alloc((size_t)g->n * 4 + 4);
            if (rc == EEG_OK) rc = unpermute(g, ln->w.lab.p, ln->out.p, ln->st);
            if (rc == EEG_OK &&
                (hipMemcpyAsync(out.data(), ln->out.p, (size_t)g->n * 4, hipMemcpyDeviceToHost, ln->st) != hipSuccess ||
                 hipStreamSynchronize(ln->st) != hipSuccess))
                rc = fail(EEG_ERR_DEVICE, "copying the answer failed");
        }
        std::lock_guard<std::mutex> lk(srv->mu);
        --srv->lanes_busy;
        tk->status = rc;
        if (rc != EEG_OK) tk->err = eeg_last_error();
        tk->state = T_DONE;
        tk->t_done = srv->now_ms();
        tk->host.swap(out);
        srv->cv_done.notify_all();
    }
}

}  // namespace
}  // namespace eeg

using namespace eeg;

extern "C" eeg_status eeg_server_create(eeg_graph* g, const eeg_server_cfg* cfg, eeg_server** out) {
    clear_error();
    if (!g || !out) return fail(EEG_ERR_INVALID_ARG, "null graph or result pointer");
    *out = nullptr;
    eeg_server_cfg c;
    memset(&c, 0, sizeof(c));
    if (cfg) c = *cfg;
    if (c.max_batch == 0) c.max_batch = 256;
    if (c.max_batch > 64 * BATCH_MAXW) return fail(EEG_ERR_INVALID_ARG, "max_batch must be <= %u", 64 * BATCH_MAXW);
    const uint32_t K = c.K ? c.K : 8;
    if (k_index(K) < 0) return fail(EEG_ERR_INVALID_ARG, "K must be 4, 8, 16 or 32");
    if (c.damping == 0.0f) c.damping = 0.85f;
    if (c.ppr_epsilon == 0.0) c.ppr_epsilon = 1e-7;
    const uint32_t known = EEG_SERVER_DISCARD_LABELS | EEG_SERVER_PLANE_ENGINES | EEG_SERVER_NO_BFS_ENGINE |
                           EEG_SERVER_NO_PPR_ENGINE | EEG_SERVER_NO_LANE_BATCH;
    if (c.flags & ~known) return fail(EEG_ERR_INVALID_ARG, "unknown flags 0x%x", c.flags);
    if (c.lanes > 64) return fail(EEG_ERR_INVALID_ARG, "lanes must be <= 64");
    DeviceGuard dg;
    EEG_TRY(select_device(g->device));
    {
        std::lock_guard<std::mutex> lk(g->mu);
        EEG_TRY(ensure_tigr(g, g->out, g->tout, K));
    }
    std::unique_ptr<eeg_server> srv(new (std::nothrow) eeg_server());
    if (!srv) return fail(EEG_ERR_OOM, "host allocation failed");
    srv->g = g;
    srv->cfg = c;
    srv->K = K;
    srv->discard = (c.flags & EEG_SERVER_DISCARD_LABELS) != 0;
    srv->t0 = Clock::now();
    const uint32_t n = g->n;
    if (c.label_budget_bytes == 0) {
        size_t fr = 0, tot = 0;
        EEG_HIP(hipMemGetInfo(&fr, &tot));
        c.label_budget_bytes = std::min<uint64_t>(fr / 4, 48ull << 30);
        srv->cfg.label_budget_bytes = c.label_budget_bytes;
    }
    // engines in use: BFS unless disabled, SSSP/SSWP label planes only when asked for
    bool use[3] = {(c.flags & EEG_SERVER_NO_BFS_ENGINE) == 0,
                   g->weighted && (c.flags & EEG_SERVER_PLANE_ENGINES) != 0,
                   g->weighted && (c.flags & EEG_SERVER_PLANE_ENGINES) != 0};
    const int neng = (int)use[0] + (int)use[1] + (int)use[2];
    // label budget split evenly over the engines; a column costs its labels plus its share of
    // the bit masks
    const double per_col = (double)n * 4.0 + (double)n * 8.0 * 3.0 / 64.0;
    const uint64_t cols =
        neng ? (uint64_t)((double)(c.label_budget_bytes / neng) / std::max(per_col, 1.0)) : 0;
    uint64_t fit = cols;
    if (fit >= 16 && fit < c.max_batch) fit -= fit % 16;  // 64-byte aligned label rows
    const uint32_t ncols = (uint32_t)std::min<uint64_t>(fit, c.max_batch);
    for (int k = 0; k < 3; ++k) srv->use_eng[k] = use[k];
    for (int k = 0; k < 3 && ncols > 0; ++k) {
        if (!use[k]) continue;
        Engine* e = new (std::nothrow) Engine();
        if (!e) return fail(EEG_ERR_OOM, "host allocation failed");
        srv->eng[k].reset(e);
        e->srv = srv.get();
        e->kind = k;
        e->K = K;
        e->ncols = ncols;
        if (c.first_iteration) e->it = c.first_iteration;
        EEG_TRY(e->init(g));
    }
    // PPR engine: rank planes of cfg.ppr_columns (8, 16 or 32, default 16) columns, fewer when
    // those do not fit a third of the free memory.  Default from the RMAT-26 load sweep (served
    // over one-at-a-time queries/s at 100 / 400 / 5000 offered q/s): 8 columns 1.49 / 1.79 /
    // 2.10, 16 columns 1.30 / 1.55 / 2.33, 32 columns 0.94 / 1.07 / 2.45 - a query's latency is
    // its ~67 passes, and a pass of 32 columns costs ~3x one of 8.
    if (c.ppr_columns != 0 && c.ppr_columns != 8 && c.ppr_columns != 16 && c.ppr_columns != 32)
        return fail(EEG_ERR_INVALID_ARG, "ppr_columns must be 8, 16 or 32");
    if ((c.flags & EEG_SERVER_NO_PPR_ENGINE) == 0 && n > 0) {
        {
            std::lock_guard<std::mutex> lk(g->mu);
            EEG_TRY(ensure_in(g));
            EEG_TRY(ensure_ppr_seg(g));
        }
        size_t fr = 0, tot = 0;
        EEG_HIP(hipMemGetInfo(&fr, &tot));
        const uint32_t nsp = g->pprseg.nspill;
        const uint32_t want = c.ppr_columns ? c.ppr_columns : 16u;
        uint32_t B = 0;
        for (uint32_t b = want; b >= 8 && !B; b /= 2)
            if (PprPlane::bytes_for(n, b, nsp) <= fr / 3) B = b;
        if (B) {
            PprEngine* e = new (std::nothrow) PprEngine();
            if (!e) return fail(EEG_ERR_OOM, "host allocation failed");
            srv->eng[E_PPR].reset(e);
            e->srv = srv.get();
            e->kind = E_PPR;
            e->K = K;
            EEG_TRY(e->init(g, B, c.damping));
            srv->use_eng[E_PPR] = true;
        }
    }
    EEG_HIP(hipStreamCreateWithFlags(&srv->io_st, hipStreamNonBlocking));
    EEG_TRY(srv->io_buf.alloc((size_t)n * 4 + 4));
    const uint32_t nl = c.lanes ? c.lanes : 2;
    for (uint32_t i = 0; i < nl; ++i) {
        srv->lanes.emplace_back(new (std::nothrow) Lane());
        if (!srv->lanes.back()) return fail(EEG_ERR_OOM, "host allocation failed");
        EEG_HIP(hipStreamCreateWithFlags(&srv->lanes.back()->st, hipStreamNonBlocking));
    }
    for (int k = 0; k < E_COUNT; ++k)
        if (srv->eng[k]) srv->eng[k]->th = std::thread([e = srv->eng[k].get()] { e->loop(); });
    for (auto& ln : srv->lanes) ln->th = std::thread([s = srv.get(), l = ln.get()] { lane_loop(s, l); });
    *out = srv.release();
    return EEG_OK;
}

extern "C" eeg_status eeg_server_submit(eeg_server* s, uint32_t algo, uint32_t source, uint64_t* ticket) {
    clear_error();
    if (!s || !ticket) return fail(EEG_ERR_INVALID_ARG, "null server or ticket");
    if (algo > EEG_PPR) return fail(EEG_ERR_INVALID_ARG, "unknown algorithm %u", algo);
    const bool sourced = algo == EEG_BFS || algo == EEG_SSSP || algo == EEG_SSWP || algo == EEG_PPR;
    if (sourced && source >= s->g->n) return fail(EEG_ERR_INVALID_ARG, "source %u out of range (n = %u)", source, s->g->n);
    if ((algo == EEG_SSSP || algo == EEG_SSWP) && !s->g->weighted)
        return fail(EEG_ERR_INVALID_ARG, "SSSP/SSWP need a weighted graph");
    auto tk = std::make_shared<Ticket>();
    tk->algo = algo;
    tk->source = sourced ? source : 0;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return fail(EEG_ERR_INVALID_ARG, "server is shutting down");
    tk->id = s->next_id++;
    tk->t_submit = s->now_ms();
    const int e = algo == EEG_BFS    ? E_BFS
                  : algo == EEG_SSSP ? E_SSSP
                  : algo == EEG_SSWP ? E_SSWP
                  : algo == EEG_PPR  ? E_PPR
                                     : -1;
    bool to_engine = e >= 0 && s->use_eng[e];
    if (to_engine && e == E_PPR) {
        // A pass over all PPR columns costs several single-query passes, so a lone PPR query is
        // answered sooner by a free lane; once lanes are busy or the engine already runs PPR
        // queries, new ones share its passes.
        EngineBase* pe = s->eng[E_PPR].get();
        if (pe && pe->running == 0 && pe->pending.empty() && s->lane_q.empty() && s->lanes_busy < s->lanes.size())
            to_engine = false;
    }
    if (to_engine) {
        EngineBase* eng = s->eng[e].get();
        if (!eng) return fail(EEG_ERR_REJECTED, "label budget holds no column for this algorithm");
        if (eng->broken) return fail(EEG_ERR_DEVICE, "serving engine stopped after a device error");
        tk->engine = e;
        eng->pending.push_back(tk);
        eng->cv.notify_one();
    } else {
        s->lane_q.push_back(tk);
        s->lane_cv.notify_one();
    }
    s->tickets[tk->id] = tk;
    *ticket = tk->id;
    return EEG_OK;
}

extern "C" eeg_status eeg_server_wait(eeg_server* s, uint64_t ticket, void* labels_out, double timeout_ms,
                                      eeg_query_result* res) {
    clear_error();
    if (!s) return fail(EEG_ERR_INVALID_ARG, "null server");
    if (labels_out && s->discard) return fail(EEG_ERR_INVALID_ARG, "server discards labels (EEG_SERVER_DISCARD_LABELS)");
    TicketP tk;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return fail(EEG_ERR_INVALID_ARG, "unknown ticket %llu", (unsigned long long)ticket);
    tk = it->second;
    auto finished = [&] { return tk->state == T_DONE || tk->state == T_CANCELLED; };
    if (timeout_ms < 0) {
        s->cv_done.wait(lk, finished);
    } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), finished)) {
        return fail(EEG_ERR_TIMEOUT, "ticket %llu not done after %.1f ms", (unsigned long long)ticket, timeout_ms);
    }
    s->tickets.erase(ticket);
    if (res) {
        res->algo = tk->algo;
        res->source = tk->source;
        res->batch_size = tk->batch;
        res->status = (uint32_t)tk->status;
        res->t_submit_ms = tk->t_submit;
        res->t_start_ms = tk->t_start;
        res->t_done_ms = tk->t_done;
    }
    if (tk->state == T_CANCELLED) return fail(EEG_ERR_CANCELLED, "ticket %llu was cancelled", (unsigned long long)ticket);
    EngineBase* e = tk->engine >= 0 ? s->eng[tk->engine].get() : nullptr;
    DevMem* buf = tk->res;
    tk->res = nullptr;
    if (tk->status != EEG_OK) {
        if (buf && e) e->res_free.push_back(buf);
        return fail((eeg_status)tk->status, "query failed: %s", tk->err.empty() ? "device error" : tk->err.c_str());
    }
    if (!buf) {  // answer in host memory (or labels discarded)
        if (labels_out && !tk->host.empty()) memcpy(labels_out, tk->host.data(), tk->host.size() * 4);
        return EEG_OK;
    }
    lk.unlock();
    eeg_status rc = EEG_OK;
    if (labels_out) {
        std::lock_guard<std::mutex> io(s->io_mu);
        DeviceGuard dg;
        if (hipSetDevice(s->g->device) != hipSuccess || hipStreamWaitEvent(s->io_st, tk->ev->e, 0) != hipSuccess ||
            hipMemcpyAsync(labels_out, buf->p, (size_t)s->g->n * 4, hipMemcpyDeviceToHost, s->io_st) != hipSuccess ||
            hipStreamSynchronize(s->io_st) != hipSuccess)
            rc = fail(EEG_ERR_DEVICE, "copying the answer failed");
    }
    lk.lock();
    e->res_free.push_back(buf);  // later gathers into it follow on the engine stream
    return rc;
}

extern "C" eeg_status eeg_server_cancel(eeg_server* s, uint64_t ticket) {
    clear_error();
    if (!s) return fail(EEG_ERR_INVALID_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return fail(EEG_ERR_INVALID_ARG, "unknown ticket %llu", (unsigned long long)ticket);
    TicketP tk = it->second;
    if (tk->state == T_PENDING) {
        tk->cancel = true;  // dropped at admission
        if (tk->engine >= 0) {
            auto& pq = s->eng[tk->engine]->pending;
            auto pos = std::find(pq.begin(), pq.end(), tk);
            if (pos != pq.end()) {
                pq.erase(pos);
                tk->state = T_CANCELLED;
                tk->status = EEG_ERR_CANCELLED;
            }
        } else {
            auto pos = std::find(s->lane_q.begin(), s->lane_q.end(), tk);
            if (pos != s->lane_q.end()) {
                s->lane_q.erase(pos);
                tk->state = T_CANCELLED;
                tk->status = EEG_ERR_CANCELLED;
            }
        }
        s->cv_done.notify_all();
    } else if ((tk->state == T_RUNNING || tk->state == T_GATHER) && tk->engine >= 0) {
        tk->cancel = true;  // column given back at convergence, a gathered answer dropped
        s->tickets.erase(it);
    } else if (tk->state == T_DONE) {
        if (tk->res) s->eng[tk->engine]->res_free.push_back(tk->res);
        tk->res = nullptr;
        s->tickets.erase(it);
    }
    return EEG_OK;
}

extern "C" eeg_status eeg_server_destroy(eeg_server* s) {
    clear_error();
    if (!s) return fail(EEG_ERR_INVALID_ARG, "null server");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        for (auto& e : s->eng)
            if (e) e->cv.notify_all();
        s->lane_cv.notify_all();
        for (auto& kv : s->tickets) {
            TicketP& t = kv.second;
            if (t->state == T_PENDING || t->state == T_RUNNING) {
                t->state = T_CANCELLED;
                t->status = EEG_ERR_CANCELLED;
            }
        }
        s->cv_done.notify_all();
    }
    for (auto& e : s->eng)
        if (e && e->th.joinable()) e->th.join();
    for (auto& ln : s->lanes)
        if (ln && ln->th.joinable()) ln->th.join();
    for (auto& e : s->eng)
        if (e && e->st) (void)hipStreamSynchronize(e->st);
    if (s->io_st) {
        (void)hipStreamSynchronize(s->io_st);
        (void)hipStreamDestroy(s->io_st);
    }
    delete s;
    return EEG_OK;
}
