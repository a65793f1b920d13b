// Commit server (see serve.h).
#include "serve.h"

#include <chrono>
#include <string>

namespace zkx {

namespace {

constexpr size_t kAlign = 256;
size_t align_up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

struct OutLayout {
    size_t leaves = 0, digests = 0, caps = 0, total = 0;
};

OutLayout out_layout(const CommitJob &j, bool want_leaves)
{
    OutLayout o;
    const uint64_t N = 1ull << (j.lg_n + j.rate), nc = 1ull << j.cap_height;
    const uint64_t nd = nc == N ? 0 : 2 * (N - nc);
    o.leaves = want_leaves ? align_up((size_t)j.cols * N * 8) : 0;
    o.digests = align_up((size_t)(nd ? nd : 1) * 32);
    o.caps = align_up((size_t)nc * 32);
    o.total = o.leaves + o.digests + o.caps;
    return o;
}

}  // namespace

double Server::now()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

size_t Server::footprint(const CommitJob &job, bool want_leaves)
{
    return commit_scratch_bytes(&job, 1) + out_layout(job, want_leaves).total;
}

uint64_t Server::work(const CommitJob &job) { return (uint64_t)job.cols << (job.lg_n + job.rate); }

// ---- server ---------------------------------------------------------------------

Server::Server(Device &d, size_t budget_bytes, int nstreams, size_t max_batch_jobs)
    : dev_(d), budget_(budget_bytes & ~(kAlign - 1)), max_batch_(max_batch_jobs ? max_batch_jobs : 64)
{
    if (nstreams < 1 || nstreams > 8) invalid("server: 1..8 streams");
    if (budget_ == 0) invalid("server: budget below 256 bytes");
    ZKX_HIP(hipSetDevice(d.id));
    region_ = DevBuf(budget_);
    place_.reset(budget_);
    for (int i = 0; i < nstreams; i++) {
        engines_.push_back(std::make_unique<Engine>(d));
        engines_.back()->set_graphs(false);  // batch compositions rarely repeat
        slots_.push_back(std::make_unique<Slot>());
        ZKX_HIP(hipEventCreateWithFlags(&slots_.back()->done, hipEventDisableTiming));
    }
    for (size_t e = 0; e < slots_.size(); e++) slots_[e]->th = std::thread([this, e] { run(e); });
}

Server::~Server()
{
    {
        std::lock_guard<std::mutex> g(mu_);
        stop_ = true;
    }
    // each completion thread finishes the batch it holds, then exits
    for (auto &s : slots_) s->cv.notify_all();
    for (auto &s : slots_)
        if (s->th.joinable()) s->th.join();
    (void)hipSetDevice(dev_.id);
    for (auto &e : engines_) (void)hipStreamSynchronize(e->stream());
    for (auto &s : slots_) (void)hipEventDestroy(s->done);
}

uint64_t Server::submit(const CommitJob &job, bool want_leaves)
{
    check_commit_job(job);
    const size_t bytes = footprint(job, want_leaves);
    if (bytes > budget_) invalid("server: request needs " + std::to_string(bytes) + " bytes, budget is " +
                                 std::to_string(budget_));
    auto r = std::make_unique<ServeRequest>();
    r->job = job;
    r->want_leaves = want_leaves;
    r->bytes = bytes;
    r->out_bytes = out_layout(job, want_leaves).total;
    r->large = work(job) > kLargeWork;
    r->t_submit = now();
    std::lock_guard<std::mutex> g(mu_);
    const uint64_t id = r->id = next_id_++;
    (r->large ? large_q_ : small_q_).push_back(r.get());
    reqs_[id] = std::move(r);
    stats_.submitted++;
    admit_locked();
    return id;
}

bool Server::poll(uint64_t id, ServeRequest *out)
{
    std::lock_guard<std::mutex> g(mu_);
    auto it = reqs_.find(id);
    if (it == reqs_.end()) invalid("server: unknown request");
    if (it->second->state < 2) return false;
    if (out) *out = *it->second;
    return true;
}

ServeRequest Server::wait(uint64_t id)
{
    std::unique_lock<std::mutex> g(mu_);
    auto it = reqs_.find(id);
    if (it == reqs_.end()) invalid("server: unknown request");
    ServeRequest *r = it->second.get();
    done_cv_.wait(g, [&] { return r->state >= 2; });
    r->collected = true;
    return *r;
}

ServeRequest Server::wait_any()
{
    std::unique_lock<std::mutex> g(mu_);
    while (true) {
        while (!finished_.empty()) {
            const uint64_t id = finished_.front();
            finished_.pop_front();
            auto it = reqs_.find(id);
            if (it == reqs_.end() || it->second->collected) continue;
            it->second->collected = true;
            return *it->second;
        }
        if (stats_.submitted == stats_.completed + stats_.failed) invalid("server: no request pending");
        done_cv_.wait(g);
    }
}

void Server::release(uint64_t id)
{
    std::unique_ptr<ServeRequest> r;
    {
        std::lock_guard<std::mutex> g(mu_);
        auto it = reqs_.find(id);
        if (it == reqs_.end()) invalid("server: unknown request");
        if (it->second->state < 2) invalid("server: request still running");
        r = std::move(it->second);
        reqs_.erase(it);
        // the request's scratch went back when its batch completed (or
        // failed); its outputs go back now
        place_.give(r->off, r->out_bytes);
        used_ -= r->out_bytes;
        stats_.bytes_in_use = used_;
        admit_locked();  // room freed: queued requests may fit now
    }
}

ServeStats Server::stats()
{
    std::lock_guard<std::mutex> g(mu_);
    return stats_;
}

void Server::assign_locked(size_t e, std::vector<ServeRequest *> batch, bool large)
{
    Slot &s = *slots_[e];
    const double t = now();
    for (ServeRequest *r : batch) {
        r->state = 1;
        r->t_start = t;
    }
    stats_.batches++;
    stats_.max_batch = std::max<uint64_t>(stats_.max_batch, batch.size());
    stats_.bytes_in_use = used_;
    stats_.peak_bytes = std::max(stats_.peak_bytes, used_);
    s.batch = std::move(batch);
    s.large = large;
    s.busy = true;
    s.assigned = true;
    s.cv.notify_one();
}

void Server::admit_locked()
{
    if (stop_) return;
    // the first free engine from `first` on (small batches try engine 0 first)
    auto free_slot = [&](size_t first) -> long {
        for (size_t e = first; e < slots_.size(); e++)
            if (!slots_[e]->busy) return (long)e;
        return -1;
    };
    // large requests: arrival order, one per batch, placed from the top of
    // the region, on engines 1.. (engine 0 stays with small batches when
    // there is more than one engine)
    const size_t large_first = slots_.size() > 1 ? 1 : 0;
    while (!large_q_.empty()) {
        ServeRequest *r = large_q_.front();
        const long e = free_slot(large_first);
        if (e < 0 || !place_.take_high(r->bytes, r->off)) break;
        large_q_.pop_front();
        used_ += r->bytes;
        large_running_++;
        assign_locked((size_t)e, {r}, true);
    }
    // small requests: arrival order, packed up to kLargeWork elements, placed
    // from the bottom; while a large request waits, only below the room it
    // needs at the top
    const ServeRequest *blocked = large_q_.empty() ? nullptr : large_q_.front();
    const size_t limit = blocked ? budget_ - blocked->bytes : budget_;
    while (!small_q_.empty()) {
        const long e = free_slot(0);
        if (e < 0) break;
        std::vector<ServeRequest *> batch;
        uint64_t w = 0;
        while (!small_q_.empty() && batch.size() < max_batch_) {
            ServeRequest *r = small_q_.front();
            const uint64_t rw = work(r->job);
            if (!batch.empty() && w + rw > kLargeWork) break;
            if (!place_.take_low(r->bytes, limit, r->off)) break;
            small_q_.pop_front();
            used_ += r->bytes;
            w += rw;
            batch.push_back(r);
        }
        if (batch.empty()) break;
        assign_locked((size_t)e, std::move(batch), false);
    }
}

bool Server::launch(size_t e, std::vector<ServeRequest *> &batch, std::string &err)
{
    Engine &eng = *engines_[e];
    Slot &slot = *slots_[e];
    hipStream_t st = eng.stream();
    try {
        ZKX_HIP(hipSetDevice(dev_.id));
        std::vector<CommitJob> jobs;
        std::vector<uint64_t *> lde;
        for (ServeRequest *r : batch) {
            const OutLayout o = out_layout(r->job, r->want_leaves);
            char *p = region_.as<char>() + r->off;
            r->job.leaves = r->want_leaves ? reinterpret_cast<uint64_t *>(p) : nullptr;
            r->job.digests = reinterpret_cast<uint64_t *>(p + o.leaves);
            r->job.caps = reinterpret_cast<uint64_t *>(p + o.leaves + o.digests);
            jobs.push_back(r->job);
            lde.push_back(reinterpret_cast<uint64_t *>(p + r->out_bytes));
        }
        {
            std::lock_guard<std::mutex> g(eng.mu);
            try {
                LaunchPieces pieces;
                if (slot.large) pieces = LaunchPieces{kLargeNttChunk, kLargeMerkleChunk};
                eng.commit(jobs.data(), jobs.size(), nullptr, nullptr, lde.data(), pieces);
            } catch (...) {
                eng.mark_failed();
                throw;
            }
        }
        ZKX_HIP(hipEventRecord(slot.done, st));
        const double t = now();
        for (ServeRequest *r : batch) r->t_enqueued = t;
        return true;
    } catch (const std::exception &ex) {
        (void)hipStreamSynchronize(st);
        err = ex.what();
        return false;
    }
}

void Server::run(size_t e)
{
    Slot &s = *slots_[e];
    (void)hipSetDevice(dev_.id);
    std::unique_lock<std::mutex> g(mu_);
    while (true) {
        s.cv.wait(g, [&] { return stop_ || s.assigned; });
        if (!s.assigned) return;
        s.assigned = false;
        std::vector<ServeRequest *> batch = s.batch;
        g.unlock();
        // enqueue, then wait for the batch on this thread: completion is seen
        // as soon as the device signals it
        std::string err;
        bool ok = launch(e, batch, err);
        if (ok) {
            const hipError_t rc = hipEventSynchronize(s.done);
            if (rc != hipSuccess) {
                ok = false;
                err = hipGetErrorString(rc);
                std::lock_guard<std::mutex> ge(engines_[e]->mu);
                engines_[e]->mark_failed();
            }
        }
        g.lock();
        const double t = now();
        for (ServeRequest *r : batch) {
            const size_t sb = r->bytes - r->out_bytes;
            place_.give(r->off + r->out_bytes, sb);
            used_ -= sb;
            r->t_done = t;
            if (ok) {
                r->state = 2;
                stats_.completed++;
            } else {
                r->state = 3;
                r->error = err;
                stats_.failed++;
            }
            finished_.push_back(r->id);
        }
        if (s.large) large_running_--;
        s.batch.clear();
        s.busy = false;
        stats_.bytes_in_use = used_;
        admit_locked();
        done_cv_.notify_all();
    }
}

}  // namespace zkx
