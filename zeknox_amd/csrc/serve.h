// Commit server: continuous batching of LDE + Merkle commit requests on one
// device under an HBM budget.
//
// Requests arrive at any time (submit() from any thread).  A scheduler thread
// owns `nstreams` commit engines (one stream each).  Whenever an engine is
// free it admits the oldest queued requests whose working set -- LDE scratch
// plus output buffers (leaf rows, digests, caps) -- still fits in the budget
// next to everything admitted and not yet released, and launches them as one
// batch; batches on different engines overlap on the device.  Outputs live in
// device memory from a stream-ordered pool until the client releases the
// request, so a request's result is the same alone or in any mix (each job's
// kernels touch only its own buffers).  Twiddle, coset and LDE factor tables
// stay resident on the device between requests.
#pragma once
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "batch.h"

namespace zkx {

struct ServeRequest {
    uint64_t id = 0;
    // job shape; coeffs is a device pointer owned by the client and must stay
    // valid until the request completes
    CommitJob job;
    bool want_leaves = true;
    // filled by the server
    size_t bytes = 0;  // budget charge while admitted
    void *out = nullptr;
    double t_submit = 0, t_start = 0, t_done = 0;  // seconds, steady clock
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
    Server(Device &d, size_t budget_bytes, int nstreams, size_t max_batch_jobs);
    ~Server();

    // Returns the request id; throws Error(1) when the request alone can never
    // fit the budget or is not a valid commit job.
    uint64_t submit(const CommitJob &job, bool want_leaves);
    // Blocks until the request is done (or failed); returns its record.
    ServeRequest wait(uint64_t id);
    bool poll(uint64_t id, ServeRequest *out);
    // Frees the request's output buffers and its budget charge.
    void release(uint64_t id);
    ServeStats stats();
    static size_t footprint(const CommitJob &job, bool want_leaves);
    static double now();

private:
    void loop();
    void launch(size_t engine, std::vector<ServeRequest *> batch);

    Device &dev_;
    size_t budget_;
    size_t max_batch_;
    std::vector<std::unique_ptr<Engine>> engines_;
    struct Slot {
        std::vector<ServeRequest *> batch;
        void *scratch = nullptr;
        hipEvent_t done = nullptr;
        bool busy = false;
    };
    std::vector<Slot> slots_;
    std::mutex mu_;
    std::condition_variable cv_;       // scheduler wake-up
    std::condition_variable done_cv_;  // request completion
    std::deque<ServeRequest *> queue_;
    std::map<uint64_t, std::unique_ptr<ServeRequest>> reqs_;
    uint64_t next_id_ = 1;
    size_t used_ = 0;
    ServeStats stats_;
    bool stop_ = false;
    std::thread worker_;
};

}  // namespace zkx
