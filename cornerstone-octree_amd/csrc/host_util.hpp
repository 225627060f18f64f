// Host-side helpers of the C++ orchestration objects: growable device buffers and per-stage HIP-event timing.
#pragma once

#include <string>
#include <utility>
#include <vector>

#include "runtime.hpp"

namespace hipcstone
{

// owns one device allocation; not copyable (a copy would free the allocation twice), swappable
struct DeviceBuf
{
    void* ptr    = nullptr;
    size_t bytes = 0;
    DeviceBuf()  = default;
    DeviceBuf(const DeviceBuf&)            = delete;
    DeviceBuf& operator=(const DeviceBuf&) = delete;
    ~DeviceBuf()
    {
        if (ptr) (void)hipFree(ptr);
    }
    friend void swap(DeviceBuf& a, DeviceBuf& b) noexcept
    {
        std::swap(a.ptr, b.ptr);
        std::swap(a.bytes, b.bytes);
    }
    hipcstone_status reserve(size_t b)
    {
        if (b <= bytes) return HIPCSTONE_OK;
        if (ptr) HC_CHECK(hipFree(ptr));
        ptr   = nullptr;
        bytes = 0;
        size_t grown = b + b / 8 + 256;
        hipError_t e = hipMalloc(&ptr, grown);
        if (e != hipSuccess)
        {
            ptr = nullptr;
            setError(std::string("pipeline: hipMalloc failed: ") + hipGetErrorString(e));
            return HIPCSTONE_OUT_OF_MEMORY;
        }
        bytes = grown;
        return HIPCSTONE_OK;
    }
    template<class V>
    V* as() const
    {
        return static_cast<V*>(ptr);
    }
};

// HIP events around each stage of a sync step, read back lazily (no synchronization while recording)
template<int kStages>
class StageTimer
{
public:
    ~StageTimer()
    {
        for (auto& r : pending_)
        {
            (void)hipEventDestroy(r.start);
            (void)hipEventDestroy(r.stop);
        }
        for (auto e : pool_)
            (void)hipEventDestroy(e);
    }
    void enable(bool on)
    {
        on_ = on;
        for (int i = 0; i < kStages; ++i)
        {
            ms_[i]  = 0;
            cnt_[i] = 0;
        }
    }
    hipcstone_status begin(int stage, hipcstone_stream s)
    {
        if (!on_) return HIPCSTONE_OK;
        Rec r{stage, get(), get()};
        if (!r.start || !r.stop) return invalid("stage timer: event creation failed");
        HC_CHECK(hipEventRecord(r.start, stream(s)));
        pending_.push_back(r);
        return HIPCSTONE_OK;
    }
    hipcstone_status end(hipcstone_stream s)
    {
        if (!on_ || pending_.empty()) return HIPCSTONE_OK;
        HC_CHECK(hipEventRecord(pending_.back().stop, stream(s)));
        return HIPCSTONE_OK;
    }
    hipcstone_status collect(double* ms, int32_t* cnt)
    {
        for (auto& r : pending_)
        {
            HC_CHECK(hipEventSynchronize(r.stop));
            float t = 0;
            HC_CHECK(hipEventElapsedTime(&t, r.start, r.stop));
            ms_[r.stage] += t;
            cnt_[r.stage] += 1;
            pool_.push_back(r.start);
            pool_.push_back(r.stop);
        }
        pending_.clear();
        for (int i = 0; i < kStages; ++i)
        {
            if (ms) ms[i] = ms_[i];
            if (cnt) cnt[i] = cnt_[i];
        }
        return HIPCSTONE_OK;
    }

private:
    struct Rec
    {
        int stage;
        hipEvent_t start;
        hipEvent_t stop;
    };
    hipEvent_t get()
    {
        if (!pool_.empty())
        {
            hipEvent_t e = pool_.back();
            pool_.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    bool on_ = false;
    double ms_[kStages]{};
    int32_t cnt_[kStages]{};
    std::vector<Rec> pending_;
    std::vector<hipEvent_t> pool_;
};

} // namespace hipcstone
