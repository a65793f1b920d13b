// The exported C ABI (include/zeknox.h): argument checks, host/device
// buffer handling and error mapping around the drivers in ntt.hip and
// merkle.hip.  Mirrors native/lib.cu:36-175 and the entry points of
// native/ntt/ntt.cu and native/merkle/merkle.cu.
#include <stdio.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/zeknox.h"
#include "merkle.h"
#include "merkle_cpu.h"
#include "ntt.h"

using namespace zkx;

namespace {

RustError ok() { return RustError{0, nullptr}; }

RustError err(int code, const char *msg) { return RustError{code, msg ? strdup(msg) : nullptr}; }

template <class F>
RustError guarded(F &&f)
{
    try {
        f();
        return ok();
    } catch (const Error &e) {
        return err(e.code, e.what());
    } catch (const std::exception &e) {
        return err(-1, e.what());
    }
}

// Device view of a caller buffer: the pointer itself when it is on the
// device, else a library allocation (copied in when `copy_in`).
struct View {
    DevBuf own;
    uint64_t *p = nullptr;
    View(void *user, bool on_device, size_t elems, bool copy_in, Device &d)
    {
        if (on_device) {
            p = static_cast<uint64_t *>(user);
        } else {
            own = DevBuf(elems * sizeof(uint64_t));
            p = own.as<uint64_t>();
            if (copy_in && elems)
                ZKX_HIP(hipMemcpyAsync(p, user, elems * sizeof(uint64_t), hipMemcpyHostToDevice, d.stream));
        }
    }
};

void copy_out(void *user, bool on_device, const uint64_t *dev, size_t elems, Device &d)
{
    if (!on_device && elems)
        ZKX_HIP(hipMemcpyAsync(user, dev, elems * sizeof(uint64_t), hipMemcpyDeviceToHost, d.stream));
}

// salt_seed: null for the reference's clock-seeded salt (ntt.cu:541-546),
// else the seed of a reproducible salt.
void lde_entry(size_t device_id, void *output, void *input, uint32_t lg_n, NTT_Direction dir, NTT_Config cfg,
               const uint64_t *salt_seed = nullptr)
{
    if (lg_n == 0 || cfg.extension_rate_bits < 1) invalid("invalid args");
    if (!output || !input) invalid("null buffer");
    const uint32_t lg = lg_n + cfg.extension_rate_bits;
    if (lg > 32) invalid("extended log size exceeds 32");
    Device &d = device(device_id);
    std::lock_guard<std::mutex> g(d.mu);
    const size_t n_in = (size_t)cfg.batches << lg_n;
    const size_t n_out = ((size_t)cfg.batches + cfg.salt_size) << lg;
    View in(input, cfg.are_inputs_on_device, n_in, true, d);
    View out(output, cfg.are_outputs_on_device, n_out, false, d);
    if (cfg.batches) lde(d, out.p, in.p, lg_n, cfg.extension_rate_bits, cfg.batches, dir == inverse, cfg.with_coset);
    if (cfg.salt_size) {
        const uint64_t count = (uint64_t)cfg.salt_size << lg;
        const uint64_t seed =
            salt_seed ? *salt_seed : (uint64_t)time(nullptr) * 0x9E3779B97F4A7C15ull ^ (uint64_t)clock();
        fill_salt(out.p + ((size_t)cfg.batches << lg), count, seed, d.stream);
    }
    copy_out(output, cfg.are_outputs_on_device, out.p, n_out, d);
    d.sync();
}

void merkle_entry(void *dig, void *cap, const void *leaves, uint64_t nd, uint64_t nc, uint64_t nl, uint64_t ls,
                  uint64_t h, uint64_t ht, uint64_t gpu_id)
{
    MerkleSpec s;
    s.num_digests = nd;
    s.num_caps = nc;
    s.num_leaves = nl;
    s.leaf_size = ls;
    s.cap_height = h;
    s.hash_type = ht;
    s.leaf_stride = ls;
    s.col_stride = 1;
    s.digests = static_cast<uint64_t *>(dig);
    s.caps = static_cast<uint64_t *>(cap);
    s.leaves = static_cast<const uint64_t *>(leaves);
    check_merkle(s);
    Device &d = device(gpu_id);
    std::lock_guard<std::mutex> g(d.mu);
    try {
        merkle_build(d, s);
        d.sync();
    } catch (...) {
        d.merkle_ctr_dirty = true;  // a failed launch may leave arrival counts behind
        throw;
    }
}

void merkle_void(void *dig, void *cap, const void *leaves, uint64_t nd, uint64_t nc, uint64_t nl, uint64_t ls,
                 uint64_t h, uint64_t ht, uint64_t gpu_id)
{
    try {
        merkle_entry(dig, cap, leaves, nd, nc, nl, ls, h, ht, gpu_id);
        set_last_error("");
    } catch (const std::exception &e) {
        set_last_error(e.what());
        fprintf(stderr, "zeknox: fill_digests_buf_linear: %s\n", e.what());
    }
}

}  // namespace

extern "C" {

RustError get_number_of_gpus(size_t *ngpus)
{
    return guarded([&] {
        if (!ngpus) invalid("null pointer");
        *ngpus = (size_t)device_count();
    });
}

RustError list_devices_info(void)
{
    return guarded([&] {
        for (int i = 0; i < device_count(); i++) {
            Device &d = device((size_t)i);
            const hipDeviceProp_t &p = d.prop;
            printf("Device %d: %s (%s)\n", i, p.name, p.gcnArchName);
            printf("  compute units        : %d\n", p.multiProcessorCount);
            printf("  clock (kHz)          : %d\n", p.clockRate);
            printf("  global memory (MiB)  : %zu\n", (size_t)(p.totalGlobalMem >> 20));
            printf("  LDS per block (KiB)  : %zu\n", (size_t)(p.sharedMemPerBlock >> 10));
            printf("  L2 cache (KiB)       : %d\n", p.l2CacheSize >> 10);
            printf("  wavefront size       : %d\n", p.warpSize);
        }
    });
}

bool cuda_available(void) { return device_count() > 0; }

RustError init_twiddle_factors(size_t device_id, size_t lg_n)
{
    return guarded([&] {
        if (lg_n > 32) invalid("log size exceeds 32");
        Device &d = device(device_id);
        std::lock_guard<std::mutex> g(d.mu);
        d.twiddle_table((uint32_t)lg_n);
        d.sync();
    });
}

RustError init_coset(size_t device_id, size_t lg_domain_size, const uint64_t coset_gen)
{
    return guarded([&] {
        if (lg_domain_size > 32) invalid("log size exceeds 32");
        Device &d = device(device_id);
        std::lock_guard<std::mutex> g(d.mu);
        d.build_coset((uint32_t)lg_domain_size, coset_gen % 0xFFFFFFFF00000001ull);
        d.sync();
    });
}

void init_cuda_degree_and_generator(const uint32_t max_degree, const uint64_t group_generator)
{
    // lib.cu:137-155: coset table of 2^max_degree powers and the twiddle
    // tables of every log size 2..max_degree on every device.  The function
    // returns nothing, so a failure is recorded for zkx_last_error and
    // reported on stderr; the first one stops the initialisation.
    auto check = [](RustError e) {
        if (e.code == 0) return true;
        const std::string m = e.message ? e.message : "init_cuda: unknown error";
        free(const_cast<char *>(e.message));
        set_last_error(m);
        fprintf(stderr, "zeknox: init_cuda: %s\n", m.c_str());
        return false;
    };
    set_last_error("");
    if (max_degree > 32) {
        check(err(1, "init_cuda: max_degree exceeds 32"));
        return;
    }
    for (int i = 0; i < device_count(); i++) {
        if (!check(init_coset((size_t)i, max_degree, group_generator))) return;
        for (uint32_t k = 2; k <= max_degree; k++)
            if (!check(init_twiddle_factors((size_t)i, k))) return;
    }
}

void init_cuda_degree(const uint32_t max_degree) { init_cuda_degree_and_generator(max_degree, 7); }

void init_cuda(void) { init_cuda_degree(24); }

RustError compute_batched_ntt(size_t device_id, void *inout, uint32_t lg_domain_size, NTT_Direction ntt_direction,
                              NTT_Config cfg)
{
    return guarded([&] {
        if (lg_domain_size == 0 || lg_domain_size > 32) invalid("invalid args");
        if (!inout) invalid("null buffer");
        Device &d = device(device_id);
        std::lock_guard<std::mutex> g(d.mu);
        const size_t elems = (size_t)cfg.batches << lg_domain_size;
        View v(inout, cfg.are_inputs_on_device, elems, true, d);
        if (cfg.batches) ntt(d, v.p, lg_domain_size, cfg.batches, ntt_direction == inverse, cfg.with_coset);
        // results go back to the caller's buffer unless it asked to keep
        // them on the device (ntt.cu:290-297)
        if (!cfg.are_inputs_on_device) copy_out(inout, false, v.p, elems, d);
        d.sync();
    });
}

RustError compute_batched_lde(size_t device_id, void *output, void *input, uint32_t lg_domain_size,
                              NTT_Direction ntt_direction, NTT_Config cfg)
{
    return guarded([&] { lde_entry(device_id, output, input, lg_domain_size, ntt_direction, cfg); });
}

RustError zkx_compute_batched_lde_seeded(size_t device_id, void *output, void *input, uint32_t lg_domain_size,
                                         NTT_Direction ntt_direction, NTT_Config cfg, uint64_t salt_seed)
{
    return guarded([&] { lde_entry(device_id, output, input, lg_domain_size, ntt_direction, cfg, &salt_seed); });
}

RustError compute_batched_lde_multi_gpu(void *output, void *input, uint32_t num_gpu, NTT_Direction ntt_direction,
                                        NTT_Config cfg, uint32_t lg_domain_size, size_t total_num_input_elements,
                                        size_t total_num_output_elements)
{
    return guarded([&] {
        if (total_num_input_elements == 0 || lg_domain_size == 0 || cfg.extension_rate_bits < 1)
            invalid("invalid args");
        if (num_gpu != 1) invalid("this build drives one device: num_gpu must be 1");
        if (total_num_input_elements != ((size_t)cfg.batches << lg_domain_size) ||
            total_num_output_elements < (((size_t)cfg.batches + cfg.salt_size) << (lg_domain_size + cfg.extension_rate_bits)))
            invalid("element counts do not match the configuration");
        lde_entry(0, output, input, lg_domain_size, ntt_direction, cfg);
    });
}

RustError compute_transpose_rev(size_t device_id, void *output, void *input, uint32_t lg_n, NTT_TransposeConfig cfg)
{
    return guarded([&] {
        if (lg_n > 32) invalid("log size exceeds 32");
        if (!output || !input) invalid("null buffer");
        Device &d = device(device_id);
        std::lock_guard<std::mutex> g(d.mu);
        const size_t elems = (size_t)cfg.batches << lg_n;
        View in(input, cfg.are_inputs_on_device, elems, true, d);
        View out(output, cfg.are_outputs_on_device, elems, false, d);
        if (elems) transpose_rev(d, out.p, in.p, lg_n, cfg.batches, true);
        copy_out(output, cfg.are_outputs_on_device, out.p, elems, d);
        d.sync();
    });
}

void fill_digests_buf_linear_gpu_with_gpu_ptr(void *digests_buf_gpu_ptr, void *cap_buf_gpu_ptr,
                                              void *leaves_buf_gpu_ptr, uint64_t digests_buf_size,
                                              uint64_t cap_buf_size, uint64_t leaves_buf_size, uint64_t leaf_size,
                                              uint64_t cap_height, uint64_t hash_type, uint64_t gpu_id)
{
    merkle_void(digests_buf_gpu_ptr, cap_buf_gpu_ptr, leaves_buf_gpu_ptr, digests_buf_size, cap_buf_size,
                leaves_buf_size, leaf_size, cap_height, hash_type, gpu_id);
}

void fill_digests_buf_linear_multigpu_with_gpu_ptr(void *digests_buf_gpu_ptr, void *cap_buf_gpu_ptr,
                                                   void *leaves_buf_gpu_ptr, uint64_t digests_buf_size,
                                                   uint64_t cap_buf_size, uint64_t leaves_buf_size,
                                                   uint64_t leaf_size, uint64_t cap_height, uint64_t hash_type)
{
    merkle_void(digests_buf_gpu_ptr, cap_buf_gpu_ptr, leaves_buf_gpu_ptr, digests_buf_size, cap_buf_size,
                leaves_buf_size, leaf_size, cap_height, hash_type, 0);
}

void fill_digests_buf_linear_cpu(void *digests_buf_ptr, void *cap_buf_ptr, const void *leaves_buf_ptr,
                                 uint64_t digests_buf_size, uint64_t cap_buf_size, uint64_t leaves_buf_size,
                                 uint64_t leaf_size, uint64_t cap_height, uint64_t hash_type)
{
    // host buffers, host threads: no HIP device is used or needed
    try {
        cpu::fill_digests_linear(static_cast<uint64_t *>(digests_buf_ptr), static_cast<uint64_t *>(cap_buf_ptr),
                                 static_cast<const uint64_t *>(leaves_buf_ptr), digests_buf_size, cap_buf_size,
                                 leaves_buf_size, leaf_size, cap_height, hash_type);
        set_last_error("");
    } catch (const std::exception &e) {
        set_last_error(e.what());
        fprintf(stderr, "zeknox: fill_digests_buf_linear_cpu: %s\n", e.what());
    }
}

}  // extern "C"

// ---- batched commits (zeknox_ext.h) -----------------------------------------

#include "../../include/zeknox_ext.h"
#include "batch.h"

namespace {

std::vector<CommitJob> to_jobs(const ZkxCommitJob *jobs, size_t n)
{
    if (n && !jobs) invalid("null job array");
    std::vector<CommitJob> v(n);
    for (size_t i = 0; i < n; i++) {
        v[i].coeffs = static_cast<const uint64_t *>(jobs[i].coeffs);
        v[i].leaves = static_cast<uint64_t *>(jobs[i].leaves);
        v[i].digests = static_cast<uint64_t *>(jobs[i].digests);
        v[i].caps = static_cast<uint64_t *>(jobs[i].caps);
        v[i].cols = jobs[i].cols;
        v[i].lg_n = jobs[i].lg_n;
        v[i].rate = jobs[i].rate;
        v[i].cap_height = jobs[i].cap_height;
        v[i].hash_type = jobs[i].hash_type;
    }
    return v;
}

}  // namespace

extern "C" {

RustError zkx_commit_batch(size_t device_id, const ZkxCommitJob *jobs, size_t njobs)
{
    return zkx_commit_batch_pieces(device_id, jobs, njobs, 0, 0);
}

RustError zkx_commit_batch_pieces(size_t device_id, const ZkxCommitJob *jobs, size_t njobs, uint64_t ntt_tiles,
                                  uint64_t merkle_groups)
{
    return guarded([&] {
        std::vector<CommitJob> v = to_jobs(jobs, njobs);
        Device &d = device(device_id);
        Engine &e = default_engine(d);
        std::lock_guard<std::mutex> g(e.mu);
        try {
            e.commit(v.data(), v.size(), nullptr, nullptr, nullptr, LaunchPieces{ntt_tiles, merkle_groups});
            e.sync();
        } catch (...) {
            e.mark_failed();
            throw;
        }
    });
}

RustError zkx_commit_trim(size_t device_id)
{
    return guarded([&] {
        Engine &e = default_engine(device(device_id));
        std::lock_guard<std::mutex> g(e.mu);
        e.trim();
    });
}

RustError zkx_autotune(size_t device_id, const ZkxCommitJob *jobs, size_t njobs, int iters)
{
    return guarded([&] {
        std::vector<CommitJob> v = to_jobs(jobs, njobs);
        Device &d = device(device_id);
        Engine &e = default_engine(d);
        std::lock_guard<std::mutex> g(e.mu);
        std::vector<TuneKey> done;
        for (const CommitJob &j : v) {
            const TuneKey k = tune_key(j);
            if (std::find(done.begin(), done.end(), k) != done.end()) continue;
            autotune(e, j, iters);
            done.push_back(k);
        }
        e.sync();
    });
}

RustError zkx_get_tuning(size_t device_id, const ZkxCommitJob *job, ZkxTuning *out)
{
    return guarded([&] {
        if (!out) invalid("null output");
        Device &d = device(device_id);
        Tuning t = d.tuning_default();
        if (job) t = d.tuning_for(tune_key(to_jobs(job, 1)[0]));
        *out = ZkxTuning{t.ntt_strided_max, t.ntt_even, t.merkle_tile_log, t.merkle_valu};
    });
}

RustError zkx_set_tuning(size_t device_id, const ZkxCommitJob *job, const ZkxTuning *t)
{
    return guarded([&] {
        if (!t) invalid("null tuning");
        Tuning v;
        v.ntt_strided_max = t->ntt_strided_max;
        v.ntt_even = t->ntt_even;
        v.merkle_tile_log = t->merkle_tile_log;
        v.merkle_valu = t->merkle_valu;
        if (!valid_tuning(v)) invalid("schedule parameters out of range");
        Device &d = device(device_id);
        std::lock_guard<std::mutex> g(d.tune_mu);
        if (job) {
            d.tuned[tune_key(to_jobs(job, 1)[0])] = v;
        } else {
            d.tune_default = v;
            d.tune_default_set = true;
        }
    });
}

int zkx_ntt_pass_sizes(uint32_t lg, const ZkxTuning *t, uint32_t *sizes, int cap)
{
    try {
        Tuning v;
        if (t) {
            v.ntt_strided_max = t->ntt_strided_max;
            v.ntt_even = t->ntt_even;
            v.merkle_tile_log = t->merkle_tile_log;
            v.merkle_valu = t->merkle_valu;
        }
        if (!valid_tuning(v) || lg > 32) return -1;
        const std::vector<uint32_t> p = ntt_pass_sizes(lg, v);
        for (size_t i = 0; i < p.size() && (int)i < cap; i++) sizes[i] = p[i];
        return (int)p.size();
    } catch (...) {
        return -1;
    }
}

int zkx_merkle_plan(uint64_t nl, uint64_t nc, uint32_t tile_log, uint64_t *count, uint32_t *lg_w, uint32_t *tile_logs,
                    uint32_t *first_ctr, int cap, uint64_t *counters)
{
    try {
        MerkleSpec s;
        s.num_leaves = nl;
        s.num_caps = nc;
        s.cap_height = nc ? (uint64_t)(63 - __builtin_clzll(nc)) : 0;
        s.num_digests = nc == nl ? 0 : 2 * (nl - nc);
        s.leaf_size = 8;
        // the plan needs sizes only: placeholder buffers pass the checks
        s.leaves = s.caps = s.digests = reinterpret_cast<uint64_t *>(256);
        check_merkle(s);
        std::vector<MerkleArgs> passes;
        plan_merkle(s, passes, tile_log);
        MerkleArgs f;
        const uint64_t n = fuse_merkle_passes(passes, f);
        for (size_t i = 0; i < passes.size() && (int)i < cap; i++) {
            count[i] = passes[i].count;
            lg_w[i] = passes[i].lg_w;
            tile_logs[i] = passes[i].tile_log;
            first_ctr[i] = i ? f.next_ctr[i - 1] : 0;
        }
        if (counters) *counters = n;
        return (int)passes.size();
    } catch (...) {
        return -1;
    }
}

RustError zkx_set_commit_graphs(size_t device_id, int on)
{
    return guarded([&] { device(device_id).commit_graphs = on != 0; });
}

RustError zkx_set_level_launches(size_t device_id, uint64_t min_leaves)
{
    return guarded([&] { device(device_id).level_launch_leaves = min_leaves; });
}

RustError zkx_set_lone_level_launches(size_t device_id, uint64_t min_leaves)
{
    return guarded([&] { device(device_id).lone_level_leaves = min_leaves; });
}

size_t zkx_commit_graphs_captured(size_t device_id)
{
    try {
        Engine &e = default_engine(device(device_id));
        std::lock_guard<std::mutex> g(e.mu);
        return e.graphs_captured();
    } catch (...) {
        return 0;
    }
}

RustError zkx_set_autotune(size_t device_id, int on)
{
    return guarded([&] {
        Device &d = device(device_id);
        std::lock_guard<std::mutex> g(d.tune_mu);
        d.autotune_new_shapes = on != 0;
    });
}

RustError zkx_tuning_clear(size_t device_id)
{
    return guarded([&] {
        Device &d = device(device_id);
        std::lock_guard<std::mutex> g(d.tune_mu);
        d.tuned.clear();
        d.tune_default = Tuning{};
        d.tune_default_set = false;
    });
}

RustError zkx_tuning_save(size_t device_id, const char *path)
{
    return guarded([&] {
        if (!path) invalid("null path");
        Device &d = device(device_id);
        std::lock_guard<std::mutex> g(d.tune_mu);
        FILE *f = fopen(path, "w");
        if (!f) invalid(std::string("cannot open ") + path);
        fprintf(f, "# cols lg_n rate hash leaves : ntt_strided_max ntt_even merkle_tile_log merkle_valu\n");
        for (const auto &kv : d.tuned) {
            const TuneKey &k = kv.first;
            const Tuning &t = kv.second;
            fprintf(f, "%u %u %u %u %d : %u %u %u %u\n", std::get<0>(k), std::get<1>(k), std::get<2>(k),
                    std::get<3>(k), (int)std::get<4>(k), t.ntt_strided_max, t.ntt_even, t.merkle_tile_log,
                    t.merkle_valu);
        }
        const bool ok = fclose(f) == 0;
        if (!ok) invalid(std::string("cannot write ") + path);
    });
}

RustError zkx_tuning_load(size_t device_id, const char *path)
{
    return guarded([&] {
        if (!path) invalid("null path");
        Device &d = device(device_id);
        FILE *f = fopen(path, "r");
        if (!f) invalid(std::string("cannot open ") + path);
        std::vector<std::pair<TuneKey, Tuning>> rows;
        char line[256];
        int lineno = 0;
        bool bad = false;
        while (fgets(line, sizeof line, f)) {
            lineno++;
            if (line[0] == '#' || line[0] == '\n') continue;
            unsigned c, n, r, h, s, e, t, mv = 0;
            int lv;
            // the merkle_valu column is optional (tables written before it existed)
            const int got = sscanf(line, "%u %u %u %u %d : %u %u %u %u", &c, &n, &r, &h, &lv, &s, &e, &t, &mv);
            if (got != 8 && got != 9) {
                bad = true;
                break;
            }
            Tuning v;
            v.ntt_strided_max = s;
            v.ntt_even = e;
            v.merkle_tile_log = t;
            v.merkle_valu = mv;
            if (!valid_tuning(v)) {
                bad = true;
                break;
            }
            rows.push_back({TuneKey{c, n, r, h, lv != 0}, v});
        }
        fclose(f);
        if (bad) invalid(std::string(path) + ": bad line " + std::to_string(lineno));
        std::lock_guard<std::mutex> g(d.tune_mu);
        for (auto &kv : rows) d.tuned[kv.first] = kv.second;
    });
}

size_t zkx_commit_scratch_bytes(const ZkxCommitJob *jobs, size_t njobs)
{
    try {
        std::vector<CommitJob> v = to_jobs(jobs, njobs);
        return commit_scratch_bytes(v.data(), v.size());
    } catch (...) {
        return 0;
    }
}

}  // extern "C"

// ---- commit server (zeknox_ext.h) -------------------------------------------

#include "serve.h"

extern "C" {

RustError zkx_server_create(size_t device_id, size_t budget_bytes, int nstreams, size_t max_batch_jobs,
                            void **server)
{
    return guarded([&] {
        if (!server) invalid("null server handle");
        Device &d = device(device_id);
        *server = new Server(d, budget_bytes, nstreams, max_batch_jobs);
    });
}

RustError zkx_server_destroy(void *server)
{
    return guarded([&] { delete static_cast<Server *>(server); });
}

size_t zkx_server_footprint(const ZkxCommitJob *job, int want_leaves)
{
    try {
        std::vector<CommitJob> v = to_jobs(job, 1);
        return Server::footprint(v[0], want_leaves != 0);
    } catch (...) {
        return 0;
    }
}

RustError zkx_server_submit(void *server, const ZkxCommitJob *job, int want_leaves, uint64_t *id)
{
    return guarded([&] {
        if (!server || !id) invalid("null argument");
        std::vector<CommitJob> v = to_jobs(job, 1);
        v[0].leaves = v[0].digests = v[0].caps = nullptr;
        // output pointers are the server's; a placeholder passes the job check
        v[0].caps = reinterpret_cast<uint64_t *>(1);
        v[0].digests = reinterpret_cast<uint64_t *>(1);
        *id = static_cast<Server *>(server)->submit(v[0], want_leaves != 0);
    });
}

static void fill_result(const ServeRequest &r, ZkxServeResult *result)
{
    if (result) {
        result->leaves = r.job.leaves;
        result->digests = r.job.digests;
        result->caps = r.job.caps;
        result->t_submit = r.t_submit;
        result->t_start = r.t_start;
        result->t_done = r.t_done;
        result->state = r.state;
        result->reserved = 0;
        result->t_enqueued = r.t_enqueued;
    }
    if (r.state == 3) throw Error(-1, "request failed: " + r.error);
}

RustError zkx_server_wait(void *server, uint64_t id, ZkxServeResult *result)
{
    return guarded([&] {
        if (!server) invalid("null server");
        fill_result(static_cast<Server *>(server)->wait(id), result);
    });
}

RustError zkx_server_wait_any(void *server, uint64_t *id, ZkxServeResult *result)
{
    return guarded([&] {
        if (!server || !id) invalid("null argument");
        ServeRequest r = static_cast<Server *>(server)->wait_any();
        *id = r.id;
        fill_result(r, result);
    });
}

RustError zkx_server_release(void *server, uint64_t id)
{
    return guarded([&] {
        if (!server) invalid("null server");
        static_cast<Server *>(server)->release(id);
    });
}

RustError zkx_server_stats(void *server, ZkxServeStats *stats)
{
    return guarded([&] {
        if (!server || !stats) invalid("null argument");
        ServeStats s = static_cast<Server *>(server)->stats();
        stats->submitted = s.submitted;
        stats->completed = s.completed;
        stats->failed = s.failed;
        stats->batches = s.batches;
        stats->max_batch = s.max_batch;
        stats->bytes_in_use = s.bytes_in_use;
        stats->peak_bytes = s.peak_bytes;
    });
}

}  // extern "C"
