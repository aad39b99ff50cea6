// server.cpp -- continuous-batching request loop (see server.hpp).
#include "rvs/server.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <iterator>
#include <map>
#include <random>
#include <stdexcept>
#include <string>

namespace rvs {

namespace {

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t plane_bytes(const rvs_target& t) { return t.bitdepth <= 8 ? 1 : 2; }

size_t yuv_bytes(const rvs_target& t) { return (size_t)t.cam.w * t.cam.h * 3 / 2 * plane_bytes(t); }

// Entry-list capacity of a batch of n views (0 = the library's default).
uint32_t batch_cap(uint32_t per_view, size_t n) {
    if (!per_view) return 0u;
    const uint64_t c = (uint64_t)per_view * n;
    return c > 0xffffffffull ? 0xffffffffu : (uint32_t)c;
}

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

// First-fit allocator over the server's one device arena (offsets, 256-byte
// granules); neighbouring free ranges merge on release.
class Server::Pool {
public:
    explicit Pool(size_t bytes) : m_total(bytes & ~(size_t)255) {
        if (m_total) m_free[0] = m_total;
    }
    size_t total() const { return m_total; }
    size_t largest() const {
        size_t m = 0;
        for (const auto& f : m_free) m = std::max(m, f.second);
        return m;
    }
    // Returns the offset, or SIZE_MAX when no free range holds `bytes`.
    size_t take(size_t bytes) {
        bytes = align256(std::max<size_t>(bytes, 256));
        for (auto it = m_free.begin(); it != m_free.end(); ++it) {
            if (it->second < bytes) continue;
            const size_t off = it->first, len = it->second;
            m_free.erase(it);
            if (len > bytes) m_free[off + bytes] = len - bytes;
            return off;
        }
        return SIZE_MAX;
    }
    void give(size_t off, size_t bytes) {
        bytes = align256(std::max<size_t>(bytes, 256));
        auto it = m_free.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != m_free.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            m_free.erase(nx);
        }
        if (it != m_free.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                m_free.erase(it);
            }
        }
    }

private:
    size_t m_total;
    std::map<size_t, size_t> m_free;   // offset -> length
};

// A slot holds one batch from launch until its outputs are on the host: its
// launch set runs on a compute stream shared with other slots (with one such
// stream the batches of all slots run in order, so the device is never split
// between them; several let long tails overlap), and its YUV planes
// (contiguous at the start of the batch's allocation, laid out as in the host
// staging buffer) are copied back on the slot's own stream, overlapping later
// batches.  The allocation comes from the server's shared pool at launch and
// returns to it when the batch completes.
struct Server::Slot {
    hipStream_t stream = nullptr;      // host copies (and, split, inpainting + encoding)
    hipEvent_t computed = nullptr;     // the batch's launch set has finished
    hipEvent_t done = nullptr;         // ... and its outputs are on the host
    char* arena = nullptr;             // the batch's allocation (pool base + alloc_off)
    size_t alloc_off = 0, alloc_bytes = 0;
    uint8_t* pinned = nullptr;
    size_t pinned_bytes = 0;
    bool busy = false;
    int index = 0;           // slot number (picks the compute stream)
    int scene = 0;
    std::vector<int> reqs;
    std::vector<rvs_target> tgts;
    std::vector<rvs_view_out> outs;
    std::vector<size_t> host_off;
    size_t yuv_bytes = 0;    // the batch's encoded planes at the start of the allocation
    size_t ws_off = 0;
    uint32_t cap = 0;        // entry capacity the current batch was launched with
};

Server::Server(std::vector<std::vector<rvs_source>> scenes, const rvs_serve_opts& opts)
    : m_scenes(std::move(scenes)), m_opt(opts) {
    if (m_scenes.empty()) throw std::invalid_argument("server: no scene");
    if (m_opt.slots < 1) m_opt.slots = 1;
    if (m_opt.max_batch < 1) m_opt.max_batch = 1;
    if (m_opt.memory_budget < 4096) throw std::invalid_argument("server: memory budget too small");
    m_per_view = m_opt.entry_cap;
    m_pool = new Pool(m_opt.memory_budget);
    hip_check(hipMalloc((void**)&m_base, m_pool->total()), "server arena");
    const int ncs = std::max(1, std::min(m_opt.compute_streams, m_opt.slots));
    for (int k = 0; k < ncs; ++k) {
        hipStream_t cs = nullptr;
        hip_check(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "stream create");
        m_compute.push_back(cs);
    }
    for (int k = 0; k < m_opt.slots; ++k) {
        Slot* s = new Slot;
        s->index = k;
        m_slots.push_back(s);
        hip_check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "stream create");
        hip_check(hipEventCreateWithFlags(&s->computed, hipEventDisableTiming), "event create");
        hip_check(hipEventCreateWithFlags(&s->done, hipEventDisableTiming), "event create");
        if (m_opt.copy_to_host) {
            s->pinned_bytes = m_opt.host_staging_bytes
                                  ? m_opt.host_staging_bytes
                                  : std::min(m_pool->total() / (4 * (size_t)m_opt.slots), (size_t)1 << 30);
            hip_check(hipHostMalloc((void**)&s->pinned, s->pinned_bytes, 0), "pinned staging");
        }
    }
    m_thread = std::thread([this] { loop(); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(m_mu);
        m_stop = true;
    }
    m_cv.notify_all();
    if (m_thread.joinable()) m_thread.join();
    for (void* cs : m_compute) (void)hipStreamSynchronize((hipStream_t)cs);
    for (Slot* s : m_slots) {
        if (s->stream) (void)hipStreamSynchronize(s->stream);
        if (s->pinned) (void)hipHostFree(s->pinned);
        if (s->done) (void)hipEventDestroy(s->done);
        if (s->computed) (void)hipEventDestroy(s->computed);
        if (s->stream) (void)hipStreamDestroy(s->stream);
        delete s;
    }
    for (void* cs : m_compute) (void)hipStreamDestroy((hipStream_t)cs);
    if (m_base) (void)hipFree(m_base);
    delete m_pool;
}

double Server::now_ms() const {
    const int64_t t = std::chrono::steady_clock::now().time_since_epoch().count();
    return (double)(t - m_t0) * 1e-6;
}

std::string Server::error() const {
    std::lock_guard<std::mutex> lk(const_cast<std::mutex&>(m_mu));
    return m_errmsg;
}

void Server::reset(const rvs_request* reqs, int n) {
    m_reqs = reqs;
    m_n = n;
    m_done = 0;
    m_batches = 0;
    m_error = 0;
    m_errmsg.clear();
    m_pending.clear();
    m_stats.assign(n, rvs_request_stats{});
    m_out.assign(m_opt.keep_outputs ? n : 0, std::vector<uint8_t>());
    for (int k = 0; k < n; ++k) {
        m_stats[k].rc = 1;   // not done
        if (reqs[k].scene < 0 || reqs[k].scene >= (int)m_scenes.size())
            throw std::invalid_argument("server: request with an unknown scene");
    }
    m_t0 = std::chrono::steady_clock::now().time_since_epoch().count();
}

void Server::submit_locked(int k, double now) {
    m_stats[k].arrival_ms = now;
    m_pending.push_back(k);
}

// Every request ends here, served or failed: its outcome is recorded and, in a
// closed loop, the next request is submitted in its place, so the number of
// requests outstanding never shrinks.
void Server::complete_locked(int k, int rc, double now, const std::string& msg) {
    m_stats[k].rc = rc;
    m_stats[k].done_ms = now;
    if (rc != RVS_OK) {
        m_error = rc;
        m_errmsg = msg;
    }
    ++m_done;
    if (m_closed && m_next_closed < m_n) submit_locked(m_next_closed++, now);
}

int Server::run_open(const rvs_request* reqs, int n, double rate, uint64_t seed) {
    if (n <= 0) return 0;
    if (!(rate > 0.0)) throw std::invalid_argument("server: arrival rate must be > 0");
    {
        std::lock_guard<std::mutex> lk(m_mu);
        reset(reqs, n);
        m_closed = false;
    }
    std::mt19937_64 rng(seed);
    std::exponential_distribution<double> gap(rate);
    const auto t0 = std::chrono::steady_clock::time_point(std::chrono::steady_clock::duration(m_t0));
    double t = 0.0;
    for (int k = 0; k < n; ++k) {
        t += gap(rng);
        std::this_thread::sleep_until(t0 + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                               std::chrono::duration<double>(t)));
        {
            std::lock_guard<std::mutex> lk(m_mu);
            submit_locked(k, now_ms());
        }
        m_cv.notify_all();
    }
    std::unique_lock<std::mutex> lk(m_mu);
    m_cv.wait(lk, [this] { return m_done == m_n; });
    return m_error;
}

int Server::run_closed(const rvs_request* reqs, int n, int concurrency) {
    if (n <= 0) return 0;
    {
        std::lock_guard<std::mutex> lk(m_mu);
        reset(reqs, n);
        m_closed = true;
        const int c = std::max(1, std::min(concurrency, n));
        const double t = now_ms();
        for (int k = 0; k < c; ++k) submit_locked(k, t);
        m_next_closed = c;
    }
    m_cv.notify_all();
    std::unique_lock<std::mutex> lk(m_mu);
    m_cv.wait(lk, [this] { return m_done == m_n; });
    return m_error;
}

// Device bytes of a batch: its YUV planes and workspace (requests return only
// the planes, so their colour maps live in the workspace); 0 with rc set when
// the plan itself fails.
size_t Server::batch_bytes(const std::vector<rvs_source>& S, const std::vector<rvs_target>& tgts, uint32_t cap,
                           size_t* ws_out, int* rc_out) const {
    size_t ws = 0;
    const int rc = rvs_synth_plan(S.data(), (int)S.size(), tgts.data(), (int)tgts.size(), cap, &ws);
    if (rc_out) *rc_out = rc;
    if (rc != RVS_OK) return 0;
    size_t h = 0;
    for (const rvs_target& t : tgts) h += yuv_bytes(t);
    if (ws_out) *ws_out = ws;
    return align256(h) + ws;
}

// Forms and enqueues the next batch on a free slot.  Admission: the oldest
// pending request fixes the scene; further pending requests of that scene
// join in arrival order while the batch fits the largest free range of the
// pool (and the slot's host staging).  A request that could not fit even the
// whole budget, or whose plan is invalid, fails at once; one that only waits
// for memory stays queued until a batch completes.
bool Server::launch(Slot& s) {
    std::vector<int> batch;
    std::vector<rvs_target> tgts;
    size_t ws = 0, need = 0, hostb = 0;
    {
        std::lock_guard<std::mutex> lk(m_mu);
        while (!m_pending.empty()) {
            // requests that can never be served fail here, whatever else is in flight
            const int k = m_pending.front();
            const rvs_request& r = m_reqs[k];
            const std::vector<rvs_target> one(1, r.target);
            int prc = RVS_OK;
            const size_t b1 = batch_bytes(m_scenes[r.scene], one, batch_cap(m_per_view, 1), nullptr, &prc);
            std::string why;
            int code = RVS_OK;
            if (prc != RVS_OK) {
                code = prc;
                why = "request " + std::to_string(k) + ": " + rvs_last_error();
            } else if (align256(b1) > m_pool->total()) {
                code = RVS_E_CAPACITY;
                why = "request " + std::to_string(k) + " needs " + std::to_string(b1) +
                      " device bytes, more than the memory budget";
            } else if ((r.target.cam.w & 1) || (r.target.cam.h & 1)) {
                code = RVS_E_UNSUPPORTED;
                why = "request " + std::to_string(k) + ": an odd target size cannot be encoded as YUV420";
            } else if (r.target.bitdepth < 1 || r.target.bitdepth > 16) {
                code = RVS_E_INVALID;
                why = "request " + std::to_string(k) + ": output bit depth " + std::to_string(r.target.bitdepth);
            } else if (s.pinned && yuv_bytes(r.target) > s.pinned_bytes) {
                code = RVS_E_CAPACITY;
                why = "request " + std::to_string(k) + " does not fit the host staging buffer";
            }
            if (code == RVS_OK) break;
            m_pending.pop_front();
            complete_locked(k, code, now_ms(), why);
            m_cv.notify_all();
        }
        if (m_pending.empty()) return false;
        const size_t room = m_pool->largest();
        const int scene = m_reqs[m_pending.front()].scene;
        const std::vector<rvs_source>& S = m_scenes[scene];
        for (auto it = m_pending.begin(); it != m_pending.end() && (int)batch.size() < m_opt.max_batch;) {
            const rvs_request& r = m_reqs[*it];
            if (r.scene != scene) { ++it; continue; }
            tgts.push_back(r.target);
            size_t w2 = 0;
            int prc = RVS_OK;
            const size_t n2 = batch_bytes(S, tgts, batch_cap(m_per_view, tgts.size()), &w2, &prc);
            const size_t h2 = hostb + yuv_bytes(r.target);
            if (prc != RVS_OK || n2 > room || (s.pinned && h2 > s.pinned_bytes)) {
                // a batch too large (in memory, views or pixels) ends here; the
                // request waits for the next batch
                tgts.pop_back();
                break;
            }
            ws = w2; need = n2; hostb = h2;
            batch.push_back(*it);
            it = m_pending.erase(it);
        }
        if (batch.empty()) return false;   // the oldest request waits for memory
        s.alloc_off = m_pool->take(need);
        if (s.alloc_off == SIZE_MAX) {     // cannot happen: need <= largest free range
            for (auto it = batch.rbegin(); it != batch.rend(); ++it) m_pending.push_front(*it);
            return false;
        }
        s.alloc_bytes = need;
        s.arena = m_base + s.alloc_off;
        s.scene = scene;
    }
    const std::vector<rvs_source>& S = m_scenes[s.scene];
    s.reqs = batch;
    s.tgts = tgts;
    s.outs.assign(batch.size(), rvs_view_out{});
    s.host_off.assign(batch.size(), 0);
    // the batch's YUV planes first, contiguous as in the host staging buffer;
    // then the colour maps; the workspace after them
    size_t h = 0;
    for (size_t k = 0; k < batch.size(); ++k) {
        const rvs_target& t = tgts[k];
        const size_t px = (size_t)t.cam.w * t.cam.h, b = plane_bytes(t);
        rvs_view_out& vo = s.outs[k];
        vo.y = s.arena + h;
        vo.u = s.arena + h + px * b;
        vo.v = s.arena + h + px * b + px / 4 * b;
        s.host_off[k] = h;
        h += yuv_bytes(t);
    }
    s.yuv_bytes = h;
    for (size_t k = 0; k < batch.size(); ++k) s.outs[k].color = nullptr;   // planes only
    s.ws_off = align256(h);
    const double t_launch = now_ms();
    s.cap = batch_cap(m_per_view, tgts.size());
    hipStream_t cs = (hipStream_t)m_compute[s.index % m_compute.size()];
    int rc;
    if (m_opt.split_finish) {
        // binning, raster and replay on the compute stream; inpainting and
        // encoding on the slot's stream, overlapping the next batch
        rc = rvs_synthesize_batch_async_split(S.data(), (int)S.size(), tgts.data(), (int)tgts.size(),
                                              s.arena + s.ws_off, ws, s.cap, s.outs.data(), cs, s.stream);
    } else {
        rc = rvs_synthesize_batch_async(S.data(), (int)S.size(), tgts.data(), (int)tgts.size(), s.arena + s.ws_off,
                                        ws, s.cap, s.outs.data(), cs);
        if (rc == RVS_OK && hipEventRecord(s.computed, cs) != hipSuccess) rc = RVS_E_HIP;
        if (rc == RVS_OK && hipStreamWaitEvent(s.stream, s.computed, 0) != hipSuccess) rc = RVS_E_HIP;
    }
    if (rc == RVS_OK && s.pinned &&
        hipMemcpyAsync(s.pinned, s.arena, s.yuv_bytes, hipMemcpyDeviceToHost, s.stream) != hipSuccess)
        rc = RVS_E_HIP;
    if (rc == RVS_OK && hipEventRecord(s.done, s.stream) != hipSuccess) rc = RVS_E_HIP;
    const std::string msg = rc == RVS_OK ? std::string() : std::string(rvs_last_error());
    if (rc != RVS_OK) {
        (void)hipStreamSynchronize(cs);
        (void)hipStreamSynchronize(s.stream);
    }
    std::lock_guard<std::mutex> lk(m_mu);
    const int id = m_batches++;
    for (int k : batch) {
        m_stats[k].launch_ms = t_launch;
        m_stats[k].batch = id;
        m_stats[k].batch_views = (int)batch.size();
    }
    if (rc != RVS_OK) {
        m_pool->give(s.alloc_off, s.alloc_bytes);
        const double t = now_ms();
        for (int k : batch) complete_locked(k, rc, t, msg.empty() ? "HIP error while enqueueing a batch" : msg);
        m_cv.notify_all();
        return true;
    }
    s.busy = true;
    return true;
}

// A completed batch (query = what hipEventQuery returned for its done event).
void Server::finish(Slot& s, int query_rc) {
    const hipError_t query = (hipError_t)query_rc;
    const std::vector<rvs_source>& S = m_scenes[s.scene];
    uint32_t needed = 0;
    int rc;
    std::string msg;
    if (query != hipSuccess) {
        rc = RVS_E_HIP;
        msg = std::string("batch failed on the device: ") + hipGetErrorString(query);
    } else {
        rc = rvs_synth_check(S.data(), (int)S.size(), s.tgts.data(), (int)s.tgts.size(), s.arena + s.ws_off, s.cap,
                             &needed, s.stream);
        if (rc != RVS_OK) msg = rvs_last_error();
    }
    const double t = now_ms();
    {
        std::lock_guard<std::mutex> lk(m_mu);
        m_pool->give(s.alloc_off, s.alloc_bytes);
        s.busy = false;
        if (rc == RVS_E_CAPACITY && needed > 0 && needed != 0xffffffffu) {
            // The entry list overflowed: raise the per-view capacity and queue the
            // batch's requests again at the front (arrival times unchanged); their
            // next launch plans with the larger list, and a request that then no
            // longer fits the budget fails at admission.
            const uint64_t nv = s.tgts.size();
            const uint64_t want = ((uint64_t)needed + needed / 8 + 1024 + nv - 1) / nv;
            const uint32_t pv = (uint32_t)(want > 0xffffffffull ? 0xffffffffull : want);
            if (pv > m_per_view) m_per_view = pv;
            for (auto it = s.reqs.rbegin(); it != s.reqs.rend(); ++it) m_pending.push_front(*it);
            ++m_requeued;
        } else {
            for (size_t k = 0; k < s.reqs.size(); ++k) {
                const int r = s.reqs[k];
                if (rc == RVS_OK && m_opt.keep_outputs && s.pinned) {
                    const uint8_t* src = s.pinned + s.host_off[k];
                    m_out[r].assign(src, src + yuv_bytes(s.tgts[k]));
                }
                complete_locked(r, rc, t, msg.empty() ? "error in a batch" : msg);
            }
        }
    }
    m_cv.notify_all();
}

void Server::loop() {
    (void)hipSetDevice(0);
    while (true) {
        bool progressed = false;
        for (Slot* s : m_slots) {
            if (!s->busy) continue;
            const hipError_t q = hipEventQuery(s->done);
            if (q == hipErrorNotReady) continue;
            finish(*s, q);
            progressed = true;
        }
        for (Slot* s : m_slots) {
            if (!s->busy && launch(*s)) progressed = true;
        }
        if (!progressed) {
            std::unique_lock<std::mutex> lk(m_mu);
            if (m_stop) break;
            bool any_busy = false;
            for (Slot* s : m_slots) any_busy = any_busy || s->busy;
            if (any_busy)
                m_cv.wait_for(lk, std::chrono::microseconds(20));
            else
                m_cv.wait_for(lk, std::chrono::milliseconds(2), [this] { return m_stop || !m_pending.empty(); });
        }
    }
}

}  // namespace rvs
