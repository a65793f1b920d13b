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

Server::Server(Device &d, size_t budget_bytes, int nstreams, size_t max_batch_jobs)
    : dev_(d), budget_(budget_bytes), max_batch_(max_batch_jobs ? max_batch_jobs : 64)
{
    if (nstreams < 1 || nstreams > 8) invalid("server: 1..8 streams");
    ZKX_HIP(hipSetDevice(d.id));
    // keep freed pool memory cached on the device between batches
    hipMemPool_t pool;
    ZKX_HIP(hipDeviceGetDefaultMemPool(&pool, d.id));
    uint64_t threshold = UINT64_MAX;
    ZKX_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &threshold));
    for (int i = 0; i < nstreams; i++) {
        engines_.push_back(std::make_unique<Engine>(d));
        Slot s;
        ZKX_HIP(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
        slots_.push_back(s);
    }
    worker_ = std::thread([this] { loop(); });
}

Server::~Server()
{
    {
        std::lock_guard<std::mutex> g(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    if (worker_.joinable()) worker_.join();
    (void)hipSetDevice(dev_.id);
    for (auto &e : engines_) (void)hipStreamSynchronize(e->stream());
    for (auto &kv : reqs_)
        if (kv.second->out) (void)hipFreeAsync(kv.second->out, engines_[0]->stream());
    for (auto &e : engines_) (void)hipStreamSynchronize(e->stream());
    for (auto &s : slots_) (void)hipEventDestroy(s.done);
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
    r->t_submit = now();
    uint64_t id;
    {
        std::lock_guard<std::mutex> g(mu_);
        id = r->id = next_id_++;
        queue_.push_back(r.get());
        reqs_[id] = std::move(r);
        stats_.submitted++;
    }
    cv_.notify_all();
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
    return *r;
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
        used_ -= out_layout(r->job, r->want_leaves).total;
        stats_.bytes_in_use = used_;
    }
    if (r->out) {
        ZKX_HIP(hipSetDevice(dev_.id));
        ZKX_HIP(hipFreeAsync(r->out, engines_[0]->stream()));
    }
    cv_.notify_all();  // budget freed: queued requests may fit now
}

ServeStats Server::stats()
{
    std::lock_guard<std::mutex> g(mu_);
    return stats_;
}

void Server::launch(size_t e, std::vector<ServeRequest *> batch)
{
    Engine &eng = *engines_[e];
    Slot &slot = slots_[e];
    hipStream_t st = eng.stream();
    std::vector<CommitJob> jobs;
    try {
        ZKX_HIP(hipSetDevice(dev_.id));
        for (ServeRequest *r : batch) {
            const OutLayout o = out_layout(r->job, r->want_leaves);
            ZKX_HIP(hipMallocAsync(&r->out, o.total, st));
            char *p = static_cast<char *>(r->out);
            r->job.leaves = r->want_leaves ? reinterpret_cast<uint64_t *>(p) : nullptr;
            r->job.digests = reinterpret_cast<uint64_t *>(p + o.leaves);
            r->job.caps = reinterpret_cast<uint64_t *>(p + o.leaves + o.digests);
            jobs.push_back(r->job);
        }
        const size_t sbytes = commit_scratch_bytes(jobs.data(), jobs.size());
        ZKX_HIP(hipMallocAsync(&slot.scratch, sbytes, st));
        {
            std::lock_guard<std::mutex> g(eng.mu);
            eng.commit(jobs.data(), jobs.size(), slot.scratch);
        }
        ZKX_HIP(hipFreeAsync(slot.scratch, st));
        slot.scratch = nullptr;
        ZKX_HIP(hipEventRecord(slot.done, st));
    } catch (const std::exception &ex) {
        (void)hipStreamSynchronize(st);
        if (slot.scratch) (void)hipFreeAsync(slot.scratch, st);
        slot.scratch = nullptr;
        std::lock_guard<std::mutex> g(mu_);
        for (ServeRequest *r : batch) {
            r->state = 3;
            r->error = ex.what();
            r->t_done = now();
            used_ -= commit_scratch_bytes(&r->job, 1);
            stats_.failed++;
        }
        slot.busy = false;
        done_cv_.notify_all();
        return;
    }
    const double t = now();
    std::lock_guard<std::mutex> g(mu_);
    for (ServeRequest *r : batch) r->t_start = t;
    slot.batch = std::move(batch);
}

void Server::loop()
{
    (void)hipSetDevice(dev_.id);
    while (true) {
        std::vector<std::pair<size_t, std::vector<ServeRequest *>>> to_launch;
        {
            std::unique_lock<std::mutex> g(mu_);
            bool any_busy = false;
            for (auto &s : slots_) any_busy |= s.busy;
            const bool idle = queue_.empty() && !any_busy;
            if (stop_ && idle) return;
            // nothing running: sleep until a submit; else poll completions often
            if (idle) cv_.wait(g, [&] { return stop_ || !queue_.empty(); });
            else cv_.wait_for(g, std::chrono::microseconds(100));

            // completions
            for (auto &s : slots_) {
                if (!s.busy || s.batch.empty()) continue;
                if (hipEventQuery(s.done) != hipSuccess) continue;
                const double t = now();
                for (ServeRequest *r : s.batch) {
                    r->state = 2;
                    r->t_done = t;
                    used_ -= commit_scratch_bytes(&r->job, 1);
                    stats_.completed++;
                }
                s.batch.clear();
                s.busy = false;
                done_cv_.notify_all();
            }
            // admission: FIFO, as many queued requests as fit the budget
            for (size_t e = 0; e < slots_.size(); e++) {
                if (slots_[e].busy || queue_.empty()) continue;
                std::vector<ServeRequest *> batch;
                while (!queue_.empty() && batch.size() < max_batch_) {
                    ServeRequest *r = queue_.front();
                    if (used_ + r->bytes > budget_) break;
                    queue_.pop_front();
                    used_ += r->bytes;
                    r->state = 1;
                    batch.push_back(r);
                }
                if (batch.empty()) break;
                slots_[e].busy = true;
                stats_.batches++;
                stats_.max_batch = std::max<uint64_t>(stats_.max_batch, batch.size());
                stats_.bytes_in_use = used_;
                stats_.peak_bytes = std::max(stats_.peak_bytes, used_);
                to_launch.emplace_back(e, std::move(batch));
            }
        }
        for (auto &l : to_launch) launch(l.first, std::move(l.second));
    }
}

}  // namespace zkx
