// Served clouds: a request loop with continuous batching over resident octrees.
//
// Requests are new full builds of a cloud (its particles come from the host) or incremental updates of a resident
// cloud (a normal displacement of N(0, (sigma h)^2) per particle, then one warm step: keys, sort through the previous
// leaves, one updateOctree step, link, centers, layout, neighbors). Each cloud keeps its state between requests in a
// plain pipeline object on the device (the warm-start state the reference keeps inside Domain / GlobalAssignment,
// domain/domain.hpp:169-218, assignment.hpp:54-144), so an update warm-starts from the cloud's previous tree.
//
// Scheduling: requests are released at their arrival times (or all at once) into one FIFO. The head of the FIFO is
// admitted as soon as (a) one of the `slots` workers is free, (b) its cloud is not busy with an earlier request (the
// requests of one cloud run in order), and (c) for a build, its modelled device memory fits the HBM budget next to the
// resident clouds, after evicting least-recently-used idle clouds that no queued request needs. Every worker runs its
// requests on its own HIP stream, so requests admitted while others are on the GPU overlap with them on the device.
// An update of an evicted cloud rebuilds it from the request's copy of the cloud's build particles first.
//
// A request's result does not depend on what else runs: its cloud's state is touched only by that cloud's requests, in
// order, and every kernel of a pipeline step runs on the worker's stream. The optional digest (an order-independent
// 64-bit sum over every particle of its key, count and stored neighbor rows) lets a test compare each request's
// result alone and in the mix.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.hpp"
#include "runtime.hpp"
#include "host_util.hpp"

namespace hipcstone
{
namespace serve
{

// order-independent digest: the wrapping sum over particles of a hash of (i, key, count, stored neighbors)
__global__ __launch_bounds__(256) void digestKernel(const uint64_t* __restrict__ keys,
                                                     const uint32_t* __restrict__ counts,
                                                     const uint32_t* __restrict__ nb, size_t n, uint32_t ngmax,
                                                     unsigned long long* __restrict__ out)
{
    unsigned long long acc = 0;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    {
        uint32_t c = counts[i];
        uint64_t d = mix64(uint64_t(i) * 0x9E3779B97F4A7C15ull ^ keys[i]) + mix64((uint64_t(i) << 32) | c);
        uint32_t m = c < ngmax ? c : ngmax;
        for (uint32_t k = 0; k < m; ++k)
            d += mix64((uint64_t(nb[i * size_t(ngmax) + k]) << 20) ^ (uint64_t(k) << 52) ^ uint64_t(i));
        acc += d;
    }
    // wave sum, then one atomic per wave
    for (int off = 32; off > 0; off >>= 1)
        acc += __shfl_down(acc, off, kWave);
    if (laneId() == 0) atomicAdd(out, acc);
}

__global__ void zeroU64Kernel(unsigned long long* p) { *p = 0; }

struct Cloud
{
    hipcstone_pipeline* pipe = nullptr;
    DeviceBuf nb, nc;
    uint64_t n       = 0;
    size_t bytes     = 0;  // device memory accounted to the cloud (reserved at admission, actual once it ran)
    double lastUse   = 0;
    bool busy        = false;
    bool resident    = false;
    // an evicted cloud's particles as they stood (x, y, z, h in the cloud's last order): its next request rebuilds
    // the octree from them, so eviction drops the tree, never the cloud's evolution
    std::vector<float> spill[4];
};

struct Job
{
    size_t req;
    bool build;   // a full build first (a build request, or an update of an evicted cloud)
    bool update;  // then one warm step
    size_t reserved;
};

// device bytes of a cloud of n particles: the pipeline's particle buffers, its sort workspace, its leaf and node
// arrays (about one leaf per 10 particles at bucket 64) and the neighbor rows, grown by the buffers' 1/8 slack
inline size_t modelBytes(uint64_t n, uint32_t ngmax)
{
    double perParticle = 16 /* input copy */ + 2 * 16 /* x,y,z,h sets */ + 8 + 8 + 4 + 4 /* keys, perm, buffers */ +
                         32 + 8 /* sort workspace incl. packed records */ + 40 /* bucket-sort scratch */ +
                         12 /* leaf array, nodes, layout */ + 4.0 * ngmax + 4 /* neighbor rows, counts */;
    return size_t(double(n) * perParticle * 1.125) + (size_t(4) << 20);
}

class Server
{
public:
    Server(const hipcstone_serve_config& cfg, const hipcstone_serve_request* req, size_t num,
           hipcstone_serve_result* res)
        : cfg_(cfg)
        , req_(req)
        , num_(num)
        , res_(res)
    {
    }

    hipcstone_status run(double* makespan)
    {
        for (size_t i = 0; i < num_; ++i)
            order_.push_back(i);
        std::stable_sort(order_.begin(), order_.end(),
                         [&](size_t a, size_t b) { return req_[a].arrival_s < req_[b].arrival_s; });
        for (size_t i = 0; i < num_; ++i)
        {
            res_[i]          = hipcstone_serve_result{};
            res_[i].id       = req_[i].id;
            // without realtime release every request is released at the start: latencies count from there
            res_[i].arrival_s = cfg_.realtime ? req_[i].arrival_s : 0.0;
        }
        int dev = 0;
        HC_CHECK(hipGetDevice(&dev));
        HC_CHECK(hipStreamCreate(&dispatchStream_));
        start_ = Clock::now();
        std::vector<std::thread> workers;
        for (int w = 0; w < cfg_.slots; ++w)
            workers.emplace_back([this, w, dev] { worker(w, dev); });
        dispatch();
        {
            std::lock_guard<std::mutex> lock(m_);
            stop_ = true;
        }
        cvWork_.notify_all();
        for (auto& t : workers)
            t.join();
        if (makespan) *makespan = now();
        for (auto& kv : clouds_)
            if (kv.second.pipe) hipcstone_pipeline_destroy(kv.second.pipe);
        clouds_.clear();
        (void)hipStreamDestroy(dispatchStream_);
        if (firstError_ != HIPCSTONE_OK) setError(firstErrorText_);
        return firstError_;
    }

private:
    using Clock = std::chrono::steady_clock;
    double now() const { return std::chrono::duration<double>(Clock::now() - start_).count(); }

    // the number of queued (released, not admitted) requests of each cloud: such clouds are never evicted
    bool queuedFor(uint32_t cloud) const
    {
        for (size_t q : pending_)
            if (req_[q].cloud == cloud) return true;
        return false;
    }

    // evict least-recently-used idle clouds no queued request needs until `need` fits; false if it cannot
    bool makeRoom(size_t need, uint32_t keep)
    {
        while (used_ + need > cfg_.hbm_budget)
        {
            Cloud* victim  = nullptr;
            uint32_t vid   = 0;
            for (auto& kv : clouds_)
            {
                Cloud& c = kv.second;
                if (!c.resident || c.busy || kv.first == keep || queuedFor(kv.first)) continue;
                if (!victim || c.lastUse < victim->lastUse)
                {
                    victim = &c;
                    vid    = kv.first;
                }
            }
            if (!victim) return false;
            // spill the idle cloud's particles to the host (on the dispatcher's own stream), then free its device
            // memory (safe while other streams run: nothing of this cloud is in flight)
            hipcstone_pipeline_view v;
            if (hipcstone_pipeline_get_view(victim->pipe, &v) == HIPCSTONE_OK && v.num_particles == victim->n)
            {
                void* src[4] = {v.x, v.y, v.z, v.h};
                for (int a = 0; a < 4; ++a)
                {
                    victim->spill[a].resize(victim->n);
                    if (hipMemcpyAsync(victim->spill[a].data(), src[a], victim->n * sizeof(float),
                                       hipMemcpyDeviceToHost, dispatchStream_) != hipSuccess)
                        failLocked(HIPCSTONE_HIP_ERROR, "serve: spill copy failed");
                }
                if (hipStreamSynchronize(dispatchStream_) != hipSuccess)
                    failLocked(HIPCSTONE_HIP_ERROR, "serve: spill synchronization failed");
            }
            hipcstone_pipeline_destroy(victim->pipe);
            victim->pipe = nullptr;
            {
                DeviceBuf a, b; // freed at the end of this scope
                swap(a, victim->nb);
                swap(b, victim->nc);
            }
            used_ -= victim->bytes;
            victim->bytes    = 0;
            victim->resident = false;
            ++evictions_;
            (void)vid;
        }
        return true;
    }

    void dispatch()
    {
        std::unique_lock<std::mutex> lock(m_);
        size_t next = 0;
        while (true)
        {
            double t = now();
            while (next < num_ && (!cfg_.realtime || req_[order_[next]].arrival_s <= t))
                pending_.push_back(order_[next++]);
            bool admitted = false;
            // scan the queue in arrival order: a request is admissible when a slot is free and its cloud has neither
            // a running request nor an earlier queued one (so each cloud's requests run in order); a build that does
            // not fit the budget stops the scan (later builds never overtake it)
            std::vector<uint32_t> blocked;
            for (auto it = pending_.begin(); it != pending_.end() && inflight_ < cfg_.slots;)
            {
                size_t r                         = *it;
                const hipcstone_serve_request& q = req_[r];
                Cloud& c                         = clouds_[q.cloud];
                if (c.busy || std::find(blocked.begin(), blocked.end(), q.cloud) != blocked.end())
                {
                    blocked.push_back(q.cloud);
                    ++it;
                    continue;
                }
                Job job{r, q.kind == HIPCSTONE_SERVE_BUILD || !c.resident, q.kind == HIPCSTONE_SERVE_UPDATE, 0};
                if (q.kind == HIPCSTONE_SERVE_UPDATE && !c.resident && c.spill[0].size() != q.n && !q.x)
                {
                    // an update of a cloud that never became resident (its build was rejected) and brings no
                    // particles: nothing to update
                    it             = pending_.erase(it);
                    res_[r].status = HIPCSTONE_SERVE_REJECTED;
                    res_[r].admit_s = res_[r].start_s = res_[r].end_s = t;
                    continue;
                }
                if (job.build)
                {
                    size_t need = modelBytes(q.n, cfg_.ngmax);
                    // a build replaces the cloud's resident state
                    size_t freed = c.resident ? c.bytes : 0;
                    if (freed) used_ -= freed;
                    if (!makeRoom(need, q.cloud))
                    {
                        if (freed) used_ += freed;
                        if (inflight_ == 0)
                        {
                            // larger than the budget even with nothing else resident
                            it             = pending_.erase(it);
                            res_[r].status = HIPCSTONE_SERVE_REJECTED;
                            res_[r].admit_s = res_[r].start_s = res_[r].end_s = t;
                            continue;
                        }
                        break; // wait for running requests to finish and become evictable
                    }
                    if (freed) c.bytes = 0;
                    used_ += need;
                    job.reserved = need;
                    c.bytes      = need;
                }
                c.busy = true;
                it     = pending_.erase(it);
                res_[r].admit_s = t;
                jobs_.push_back(job);
                ++inflight_;
                admitted = true;
            }
            if (admitted) cvWork_.notify_all();
            if (next >= num_ && pending_.empty() && inflight_ == 0) break;
            // sleep until the next arrival or a completion
            if (next < num_ && cfg_.realtime)
            {
                double wait = req_[order_[next]].arrival_s - now();
                if (wait > 0) cvDone_.wait_for(lock, std::chrono::duration<double>(wait));
            }
            else if (!admitted)
                cvDone_.wait(lock);
        }
    }

    void worker(int slot, int dev)
    {
        (void)hipSetDevice(dev);
        hipStream_t s = nullptr;
        if (hipStreamCreate(&s) != hipSuccess)
        {
            fail(HIPCSTONE_HIP_ERROR, "serve: hipStreamCreate failed");
            return;
        }
        DeviceBuf digest;
        (void)digest.reserve(64);
        while (true)
        {
            Job job;
            {
                std::unique_lock<std::mutex> lock(m_);
                cvWork_.wait(lock, [&] { return stop_ || !jobs_.empty(); });
                if (jobs_.empty()) break;
                job = jobs_.front();
                jobs_.pop_front();
            }
            const hipcstone_serve_request& q = req_[job.req];
            hipcstone_serve_result& out      = res_[job.req];
            out.slot                         = uint32_t(slot);
            out.start_s                      = now();
            Cloud* c;
            {
                std::lock_guard<std::mutex> lock(m_);
                c = &clouds_[q.cloud];
            }
            hipcstone_status st = execute(job, q, *c, s, digest, out);
            double end          = now();
            size_t actual       = c->pipe ? hipcstone_pipeline_device_bytes(c->pipe) + c->nb.bytes + c->nc.bytes : 0;
            {
                std::lock_guard<std::mutex> lock(m_);
                out.end_s = end;
                if (st != HIPCSTONE_OK) failLocked(st, hipcstone_last_error());
                if (job.build || c->resident)
                {
                    // the reservation becomes the actual size
                    used_ -= c->bytes;
                    c->bytes = actual;
                    used_ += actual;
                }
                c->busy          = false;
                c->lastUse       = end;
                out.resident_bytes = used_;
                --inflight_;
            }
            cvDone_.notify_all();
        }
        (void)hipStreamDestroy(s);
    }

    hipcstone_status execute(const Job& job, const hipcstone_serve_request& q, Cloud& c, hipStream_t s,
                             DeviceBuf& digest, hipcstone_serve_result& out)
    {
        hipcstone_box_f32 box;
        hipcstone_make_box_f32(cfg_.box_lo, cfg_.box_hi, cfg_.box_lo, cfg_.box_hi, cfg_.box_lo, cfg_.box_hi, 0, 0, 0,
                               &box);
        const size_t n = size_t(q.n);
        if (job.build)
        {
            // a build request brings its particles; an evicted cloud comes back from its spilled particles
            const bool fromSpill = q.kind == HIPCSTONE_SERVE_UPDATE && c.spill[0].size() == n;
            if (!fromSpill && (!q.x || !q.y || !q.z || !q.h || n == 0))
                return invalid("serve: a build needs its particles");
            if (c.pipe) hipcstone_pipeline_destroy(c.pipe);
            c.pipe = nullptr;
            if (auto e = hipcstone_pipeline_create(&c.pipe, cfg_.curve, 64, 32, cfg_.bucket_size, cfg_.ngmax)) return e;
            DeviceBuf in[4];
            const float* src[4] = {q.x, q.y, q.z, q.h};
            if (fromSpill)
                for (int a = 0; a < 4; ++a)
                    src[a] = c.spill[a].data();
            for (int a = 0; a < 4; ++a)
            {
                if (auto e = in[a].reserve(n * sizeof(float))) return e;
                HC_CHECK(hipMemcpyAsync(in[a].ptr, src[a], n * sizeof(float), hipMemcpyHostToDevice, s));
            }
            if (auto e = c.nb.reserve(std::max<size_t>(1, n * cfg_.ngmax) * sizeof(uint32_t))) return e;
            if (auto e = c.nc.reserve(n * sizeof(uint32_t))) return e;
            if (auto e = hipcstone_pipeline_sync(c.pipe, s, in[0].ptr, in[1].ptr, in[2].ptr, in[3].ptr, n, &box, 1,
                                                 c.nb.as<uint32_t>(), c.nc.as<uint32_t>()))
                return e;
            HC_CHECK(hipStreamSynchronize(s)); // the input copies are freed here
            for (auto& v : c.spill)
                std::vector<float>().swap(v);
            c.n        = q.n;
            c.resident = true;
            out.status = q.kind == HIPCSTONE_SERVE_BUILD ? HIPCSTONE_SERVE_BUILT : HIPCSTONE_SERVE_REBUILT;
        }
        if (job.update)
        {
            if (q.n != c.n) return invalid("serve: an update must keep the cloud's particle count");
            hipcstone_pipeline_view v;
            if (auto e = hipcstone_pipeline_get_view(c.pipe, &v)) return e;
            if (auto e = hipcstone_synth_displace_normal_f32(s, static_cast<float*>(v.x), static_cast<float*>(v.y),
                                                             static_cast<float*>(v.z), static_cast<const float*>(v.h),
                                                             n, cfg_.sigma, q.seed, q.id, &box))
                return e;
            if (auto e = hipcstone_pipeline_sync(c.pipe, s, v.x, v.y, v.z, v.h, n, &box, 0, c.nb.as<uint32_t>(),
                                                 c.nc.as<uint32_t>()))
                return e;
            if (!job.build) out.status = HIPCSTONE_SERVE_UPDATED;
        }
        hipcstone_pipeline_view v;
        if (auto e = hipcstone_pipeline_get_view(c.pipe, &v)) return e;
        out.num_leaves = v.num_leaves;
        if (cfg_.digests)
        {
            auto* d = digest.as<unsigned long long>();
            zeroU64Kernel<<<1, 1, 0, s>>>(d);
            digestKernel<<<std::max<unsigned>(1, std::min<unsigned>(1024, ceilDiv(n, 256))), 256, 0, s>>>(
                static_cast<const uint64_t*>(v.keys), c.nc.as<uint32_t>(), c.nb.as<uint32_t>(), n, cfg_.ngmax, d);
            HC_CHECK_LAUNCH("digestKernel");
            unsigned long long h = 0;
            HC_CHECK(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, s));
            HC_CHECK(hipStreamSynchronize(s));
            out.digest = h;
        }
        else
            HC_CHECK(hipStreamSynchronize(s));
        return HIPCSTONE_OK;
    }

    void fail(hipcstone_status st, const std::string& text)
    {
        std::lock_guard<std::mutex> lock(m_);
        failLocked(st, text);
    }
    void failLocked(hipcstone_status st, const std::string& text)
    {
        if (firstError_ == HIPCSTONE_OK)
        {
            firstError_     = st;
            firstErrorText_ = text;
        }
    }

    hipcstone_serve_config cfg_;
    const hipcstone_serve_request* req_;
    size_t num_;
    hipcstone_serve_result* res_;
    std::vector<size_t> order_;
    std::deque<size_t> pending_;
    std::deque<Job> jobs_;
    std::unordered_map<uint32_t, Cloud> clouds_;
    std::mutex m_;
    std::condition_variable cvWork_, cvDone_;
    Clock::time_point start_;
    hipStream_t dispatchStream_ = nullptr;
    size_t used_        = 0;
    int inflight_       = 0;
    bool stop_          = false;
    uint64_t evictions_ = 0;
    hipcstone_status firstError_ = HIPCSTONE_OK;
    std::string firstErrorText_;

public:
    uint64_t evictions() const { return evictions_; }
};

} // namespace serve
} // namespace hipcstone

using namespace hipcstone;

extern "C" hipcstone_status hipcstone_serve_run(const hipcstone_serve_config* cfg,
                                                const hipcstone_serve_request* requests, size_t num_requests,
                                                hipcstone_serve_result* results, hipcstone_serve_stats* stats)
{
    if (!cfg || (num_requests && (!requests || !results))) return invalid("serve_run: null argument");
    if (cfg->slots <= 0 || cfg->slots > 64) return invalid("serve_run: 1 to 64 slots");
    if (cfg->bucket_size == 0) return invalid("serve_run: bucket size 0");
    if (!(cfg->box_hi > cfg->box_lo)) return invalid("serve_run: empty box");
    serve::Server server(*cfg, requests, num_requests, results);
    double makespan = 0;
    hipcstone_status st = server.run(&makespan);
    if (stats)
    {
        stats->makespan_s = makespan;
        stats->evictions  = server.evictions();
    }
    return st;
}
