// Commit server: continuous batching of LDE + Merkle commit requests on one
// device under an HBM budget.
//
// Requests arrive at any time (submit() from any thread).  The server owns
// `nstreams` commit engines (one stream and one completion thread each).
// Admission runs whenever a request arrives, a batch completes or a client
// releases a result, and hands batches to free engines:
//
//   * small requests (LDE of at most kLargeWork elements) are taken in
//     arrival order and packed into batches of at most kLargeWork elements
//     in total and max_batch_jobs jobs, so a small request never waits for a
//     large one in its own batch;
//   * large requests run one per batch, in arrival order among themselves,
//     on engines 1..nstreams-1: engine 0 only ever takes small batches, so
//     a small request is never queued behind large ones for an engine; and
//     a large batch's leaf launch (its longest kernel by far) is issued in
//     pieces of kLargeMerkleChunk work-groups and its NTT passes in ranges
//     of kLargeNttChunk tiles: the device takes another stream's kernels
//     between two of a stream's kernels, not while one still has
//     work-groups to dispatch;
//   * the budget is one device allocation made when the server starts; a
//     request is admitted only when its working set -- output buffers (leaf
//     rows, digests, caps) followed by its LDE scratch -- can be placed in it
//     next to everything admitted and not yet released.  Large requests are
//     placed from the top of the region, small ones from the bottom.  While
//     the oldest large request waits for room, small requests are placed
//     only below the top `bytes` of the region that request needs, so it
//     starts as soon as the large requests ahead of it finish: no request
//     starves.
//
// Serving never allocates device memory: placement is host bookkeeping (a
// request's scratch returns to the region when its batch completes, its
// outputs when the client releases it).  Batches on different engines
// overlap on the device.  A request's result is the same alone or in any mix
// (each job's kernels touch only its own buffers).  Twiddle, coset and LDE
// factor tables stay resident on the device between requests.
#pragma once
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "batch.h"
#include "placement.h"

namespace zkx {

struct ServeRequest {
    uint64_t id = 0;
    // job shape; coeffs is a device pointer owned by the client and must stay
    // valid until the request completes
    CommitJob job;
    bool want_leaves = true;
    // filled by the server
    bool large = false;
    bool collected = false;  // returned by wait() or wait_any()
    size_t bytes = 0;        // working set: outputs + scratch
    size_t out_bytes = 0;    // outputs, at the start of the working set
    size_t off = 0;          // working set's offset in the region
    // seconds, steady clock: submitted, admitted, kernels enqueued, finished
    double t_submit = 0, t_start = 0, t_enqueued = 0, t_done = 0;
    int state = 0;                                  // 0 queued, 1 running, 2 done, 3 failed
    std::string error;
};

struct ServeStats {
    uint64_t submitted = 0, completed = 0, failed = 0, batches = 0;
    size_t bytes_in_use = 0, peak_bytes = 0;
    uint64_t max_batch = 0;
};

class Server {
public:
    // LDE elements (cols x 2^(lg_n + rate)) above which a request is large;
    // also the work cap of a batch of small requests (about 20 ms of device
    // time at the batched rate).
    static constexpr uint64_t kLargeWork = 1ull << 28;
    // A large batch's leaf launch is issued in pieces of this many
    // work-groups (about four times what the device holds at once), so small
    // batches' kernels are dispatched between them.  Served at 160 requests/s
    // (20 s Poisson streams), against one launch: small-request p99 230 ->
    // 134 ms, mean latency 34 -> 24.5 ms, large requests' median service
    // 47 -> 50 ms; 2048 work-groups gave 114 ms small p99 but 56 ms large.
    static constexpr uint64_t kLargeMerkleChunk = 4096;
    // Its NTT passes likewise, in ranges of this many 4096-element
    // tiles (2 GB of LDE per piece, 1-2 ms).
    static constexpr uint64_t kLargeNttChunk = 65536;

    Server(Device &d, size_t budget_bytes, int nstreams, size_t max_batch_jobs);
    ~Server();

    // Returns the request id; throws Error(1) when the request alone can never
    // fit the budget or is not a valid commit job.
    uint64_t submit(const CommitJob &job, bool want_leaves);
    // Blocks until the request is done (or failed); returns its record.
    ServeRequest wait(uint64_t id);
    // Blocks until some finished request that no wait()/wait_any() call has
    // returned yet exists; returns it.  Throws Error(1) when none is pending.
    ServeRequest wait_any();
    bool poll(uint64_t id, ServeRequest *out);
    // Frees the request's output buffers and its budget charge.
    void release(uint64_t id);
    ServeStats stats();
    static size_t footprint(const CommitJob &job, bool want_leaves);
    static uint64_t work(const CommitJob &job);
    static double now();

private:
    struct Slot {
        std::vector<ServeRequest *> batch;
        hipEvent_t done = nullptr;
        bool busy = false;      // batch assigned and not yet completed
        bool assigned = false;  // batch handed over, completion thread not yet started on it
        bool large = false;
        std::condition_variable cv;
        std::thread th;
    };

    void admit_locked();
    void assign_locked(size_t e, std::vector<ServeRequest *> batch, bool large);
    void run(size_t e);
    bool launch(size_t e, std::vector<ServeRequest *> &batch, std::string &err);

    Device &dev_;
    size_t budget_;
    size_t max_batch_;
    std::vector<std::unique_ptr<Engine>> engines_;
    std::vector<std::unique_ptr<Slot>> slots_;
    DevBuf region_;     // the budget, allocated once
    Placement place_;
    std::mutex mu_;
    std::condition_variable done_cv_;  // request completion
    std::deque<ServeRequest *> small_q_, large_q_;
    std::deque<uint64_t> finished_;    // completion order, for wait_any
    std::map<uint64_t, std::unique_ptr<ServeRequest>> reqs_;
    uint64_t next_id_ = 1;
    size_t used_ = 0;
    int large_running_ = 0;
    ServeStats stats_;
    bool stop_ = false;
};

}  // namespace zkx
