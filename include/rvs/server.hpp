// server.hpp -- serving loop: resident source sets, memory-budget admission,
// continuous batching of target-view requests onto the batched device path.
//
// Requests (scene, target pose and camera) are queued on arrival.  A scheduler
// thread keeps up to `slots` batches in flight.  Device memory is one arena of
// `memory_budget` bytes shared by all batches: whenever a slot is free it takes
// the oldest pending request's scene and admits further pending requests of
// that scene in arrival order while the batch's workspace plus outputs fit the
// largest free range of the arena (and at most max_batch views); the batch's
// memory returns to the arena when it completes, so requests that arrive while
// others are on the device are admitted as slots and memory free up.  A
// request that could not fit the whole budget, or whose plan is invalid, fails
// with the plan's own error code.  A batch is one rvs_synthesize_batch_async
// launch set followed by device-to-host copies of the YUV planes, so batches
// on different slots overlap on the device.  A batch whose binned-entry list
// overflowed is queued again, with a larger list, without blocking the other
// slots.  Each request's latency runs from its arrival to the completion of
// its host copy.  Results equal those of a lone rvs_synthesize call for the
// same target.
#ifndef RVS_SERVER_HPP
#define RVS_SERVER_HPP

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "rvs_hip.h"
#include "rvs_host.h"

namespace rvs {

class Server {
public:
    Server(std::vector<std::vector<rvs_source>> scenes, const rvs_serve_opts& opts);
    ~Server();
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    /** Open loop: request k arrives at the k-th event of a Poisson process of
     * `rate` per second (seeded); blocks until every request is done. */
    int run_open(const rvs_request* reqs, int n, double rate, uint64_t seed);
    /** Closed loop: `concurrency` requests outstanding at all times. */
    int run_closed(const rvs_request* reqs, int n, int concurrency);
    const std::vector<rvs_request_stats>& stats() const { return m_stats; }
    /** Host copy of request k's YUV planes (keep_outputs). */
    const std::vector<uint8_t>& output(int k) const { return m_out[k]; }
    /** Message of the last failure in the current run (empty when none). */
    std::string error() const;
    /** Batches queued again after their binned-entry list overflowed (all runs). */
    int requeued() const { return m_requeued; }

private:
    struct Slot;
    class Pool;
    void reset(const rvs_request* reqs, int n);
    void submit_locked(int k, double now);
    void complete_locked(int k, int rc, double now, const std::string& msg);
    size_t batch_bytes(const std::vector<rvs_source>& S, const std::vector<rvs_target>& tgts, uint32_t cap,
                       size_t* ws_out, int* rc_out) const;
    void loop();
    bool launch(Slot& s);
    void finish(Slot& s, int query);
    double now_ms() const;

    std::vector<std::vector<rvs_source>> m_scenes;
    rvs_serve_opts m_opt;
    std::vector<Slot*> m_slots;
    Pool* m_pool = nullptr;         // free ranges of the device arena
    char* m_base = nullptr;         // the arena (memory_budget bytes)
    uint32_t m_per_view = 0;        // binned entries per view of a batch (0: library default); grows on overflow
    int m_requeued = 0;             // batches queued again after a list overflow
    std::vector<void*> m_compute;   // hipStream_t: slot k's launch sets run on m_compute[k % size]
    std::mutex m_mu;
    std::condition_variable m_cv;
    std::thread m_thread;
    bool m_stop = false;
    // current run
    const rvs_request* m_reqs = nullptr;
    int m_n = 0;
    int m_done = 0;
    int m_next_closed = 0;    // closed loop: next request to submit on completion
    bool m_closed = false;
    int m_batches = 0;
    std::deque<int> m_pending;
    std::vector<rvs_request_stats> m_stats;
    std::vector<std::vector<uint8_t>> m_out;
    int64_t m_t0 = 0;
    int m_error = 0;
    std::string m_errmsg;
};

}  // namespace rvs

#endif
