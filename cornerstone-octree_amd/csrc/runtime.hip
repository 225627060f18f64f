// Runtime part of the C ABI: errors, device memory, streams, events, box construction, byte fill.
#include "runtime.hpp"

#include "common.hpp"

#include <atomic>
#include <mutex>

namespace hipcstone
{

static thread_local std::string g_lastError;

void setError(const std::string& msg) { g_lastError = msg; }

// the per-device fault word; host code uses the address hipGetSymbolAddress returns for the current device, never
// the name
static __device__ __align__(256) uint32_t g_faultWord[64];

constexpr int kMaxDevices = 64;
static std::mutex g_faultMutex;
static uint32_t* g_faultPtr[kMaxDevices];

uint32_t* faultWord()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    {
        setError("fault word: no current device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_faultMutex);
    if (!g_faultPtr[dev])
    {
        void* p      = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(g_faultWord));
        if (e != hipSuccess)
        {
            setError(std::string("fault word: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
            return nullptr;
        }
        g_faultPtr[dev] = static_cast<uint32_t*>(p);
    }
    return g_faultPtr[dev];
}

__global__ void takeFaultsKernel(uint32_t* word, uint32_t* out)
{
    *out  = *word;
    *word = 0;
}

hipcstone_status enqueueFaultRead(hipStream_t s, uint32_t* host)
{
    uint32_t* w = faultWord();
    if (!w) return HIPCSTONE_HIP_ERROR;
    HC_CHECK(hipMemcpyAsync(host, w, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    return HIPCSTONE_OK;
}

uint32_t* pinnedStaging()
{
    static thread_local uint32_t* p = nullptr;
    if (!p)
    {
        void* q      = nullptr;
        hipError_t e = hipHostMalloc(&q, kPinnedWords * sizeof(uint32_t), hipHostMallocDefault);
        if (e != hipSuccess)
        {
            setError(std::string("pinned staging: hipHostMalloc failed: ") + hipGetErrorString(e));
            return nullptr;
        }
        p = static_cast<uint32_t*>(q);
    }
    return p;
}

hipcstone_status checkFaults(uint32_t faults)
{
    if (faults == 0) return HIPCSTONE_OK;
    std::string msg = "device fault:";
    if (faults & HIPCSTONE_FAULT_SORT_LOOKBACK) msg += " radix sort look-back gave up;";
    if (faults & HIPCSTONE_FAULT_WALK_BUDGET) msg += " neighbor walk exceeded its step budget;";
    if (faults & HIPCSTONE_FAULT_MALFORMED_TREE) msg += " tree walk met an out-of-range node;";
    if (faults & HIPCSTONE_FAULT_CAPACITY) msg += " batched leaf array exceeded its capacity;";
    setError(msg);
    return HIPCSTONE_DEVICE_FAULT;
}

static std::atomic<int> g_neighborWaves{4};

int neighborWaves() { return g_neighborWaves.load(); }

static std::atomic<uint32_t> g_lookbackLimit{0};

uint32_t sortLookbackLimit()
{
    uint32_t v = g_lookbackLimit.load();
    return v ? v : (1u << 22);
}

__global__ void fillBytesKernel(uint8_t* dst, uint8_t value, size_t bytes)
{
    size_t i      = (size_t(blockIdx.x) * blockDim.x + threadIdx.x) * 16;
    size_t stride = size_t(gridDim.x) * blockDim.x * 16;
    uint32_t v4   = 0x01010101u * value;
    for (; i < bytes; i += stride)
    {
        if (i + 16 <= bytes && (reinterpret_cast<uintptr_t>(dst + i) & 15) == 0)
        {
            *reinterpret_cast<uint4*>(dst + i) = make_uint4(v4, v4, v4, v4);
        }
        else
        {
            for (size_t k = i; k < i + 16 && k < bytes; ++k)
                dst[k] = value;
        }
    }
}

template<class T, class B>
static void makeBox(T xmin, T xmax, T ymin, T ymax, T zmin, T zmax, int bx, int by, int bz, B* b)
{
    b->lim[0]  = xmin;
    b->lim[1]  = xmax;
    b->lim[2]  = ymin;
    b->lim[3]  = ymax;
    b->lim[4]  = zmin;
    b->lim[5]  = zmax;
    b->len[0]  = xmax - xmin;
    b->len[1]  = ymax - ymin;
    b->len[2]  = zmax - zmin;
    b->ilen[0] = T(1.) / (xmax - xmin);
    b->ilen[1] = T(1.) / (ymax - ymin);
    b->ilen[2] = T(1.) / (zmax - zmin);
    b->bnd[0]  = uint8_t(bx);
    b->bnd[1]  = uint8_t(by);
    b->bnd[2]  = uint8_t(bz);
}

} // namespace hipcstone

using namespace hipcstone;

extern "C"
{

const char* hipcstone_last_error(void) { return g_lastError.c_str(); }

hipcstone_status hipcstone_device_faults(hipcstone_stream s, uint32_t* faults)
{
    uint32_t* w = faultWord();
    if (!w) return HIPCSTONE_HIP_ERROR;
    // a one-word staging slot next to the word itself: read and clear it in one kernel, then copy the slot out
    takeFaultsKernel<<<1, 1, 0, stream(s)>>>(w, w + 32);
    HC_CHECK_LAUNCH("takeFaultsKernel");
    uint32_t v = 0;
    HC_CHECK(hipMemcpyAsync(&v, w + 32, sizeof(v), hipMemcpyDeviceToHost, stream(s)));
    HC_CHECK(hipStreamSynchronize(stream(s)));
    if (faults) *faults = v;
    return checkFaults(v);
}

hipcstone_status hipcstone_fault_word(uint32_t** word)
{
    if (!word) return invalid("fault_word: null pointer");
    *word = faultWord();
    return *word ? HIPCSTONE_OK : HIPCSTONE_HIP_ERROR;
}

hipcstone_status hipcstone_set_neighbor_waves(int32_t waves)
{
    if (waves == 0) waves = 4;
    if (waves != 2 && waves != 4 && waves != 8) return invalid("set_neighbor_waves: 2, 4 or 8 waves per workgroup");
    g_neighborWaves.store(waves);
    return HIPCSTONE_OK;
}

int32_t hipcstone_get_neighbor_waves(void) { return g_neighborWaves.load(); }

hipcstone_status hipcstone_set_sort_lookback_limit(uint32_t polls)
{
    g_lookbackLimit.store(polls);
    return HIPCSTONE_OK;
}

const char* hipcstone_build_info(void) { return "hipcstone 0.1 (gfx950, wave64, -ffp-contract=off)"; }

hipcstone_status hipcstone_device_alloc(void** ptr, size_t bytes)
{
    if (!ptr) return invalid("ptr");
    *ptr = nullptr;
    if (bytes == 0) return HIPCSTONE_OK;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess)
    {
        setError(std::string("hipMalloc failed: ") + hipGetErrorString(e));
        return HIPCSTONE_OUT_OF_MEMORY;
    }
    return HIPCSTONE_OK;
}

hipcstone_status hipcstone_device_free(void* ptr)
{
    if (ptr) HC_CHECK(hipFree(ptr));
    return HIPCSTONE_OK;
}

hipcstone_status hipcstone_memcpy_h2d(void* dst, const void* src, size_t bytes, hipcstone_stream s)
{
    if (bytes == 0) return HIPCSTONE_OK;
    HC_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream(s)));
    HC_CHECK(hipStreamSynchronize(stream(s)));
    return HIPCSTONE_OK;
}

hipcstone_status hipcstone_memcpy_d2h(void* dst, const void* src, size_t bytes, hipcstone_stream s)
{
    if (bytes == 0) return HIPCSTONE_OK;
    HC_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream(s)));
    HC_CHECK(hipStreamSynchronize(stream(s)));
    return HIPCSTONE_OK;
}

hipcstone_status hipcstone_memcpy_d2d(void* dst, const void* src, size_t bytes, hipcstone_stream s)
{
    if (bytes == 0) return HIPCSTONE_OK;
    HC_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream(s)));
    return HIPCSTONE_OK;
}

hipcstone_status hipcstone_fill_bytes(void* dst, uint8_t value, size_t bytes, hipcstone_stream s)
{
    if (bytes == 0) return HIPCSTONE_OK;
    if (!dst) return invalid("dst");
    unsigned blocks = ceilDiv(bytes, 256 * 16);
    if (blocks > 4096) blocks = 4096;
    fillBytesKernel<<<blocks, 256, 0, stream(s)>>>(static_cast<uint8_t*>(dst), value, bytes);
    HC_CHECK_LAUNCH("fillBytesKernel");
    return HIPCSTONE_OK;
}

hipcstone_status hipcstone_stream_create(hipcstone_stream* s)
{
    hipStream_t h;
    HC_CHECK(hipStreamCreateWithFlags(&h, hipStreamNonBlocking));
    *s = h;
    return HIPCSTONE_OK;
}

hipcstone_status hipcstone_stream_destroy(hipcstone_stream s)
{
    if (s) HC_CHECK(hipStreamDestroy(stream(s)));
    return HIPCSTONE_OK;
}

hipcstone_status hipcstone_stream_sync(hipcstone_stream s)
{
    HC_CHECK(hipStreamSynchronize(stream(s)));
    return HIPCSTONE_OK;
}

hipcstone_status hipcstone_device_sync(void)
{
    HC_CHECK(hipDeviceSynchronize());
    return HIPCSTONE_OK;
}

hipcstone_status hipcstone_event_create(void** ev)
{
    hipEvent_t e;
    HC_CHECK(hipEventCreate(&e));
    *ev = e;
    return HIPCSTONE_OK;
}

hipcstone_status hipcstone_event_destroy(void* ev)
{
    if (ev) HC_CHECK(hipEventDestroy(static_cast<hipEvent_t>(ev)));
    return HIPCSTONE_OK;
}

hipcstone_status hipcstone_event_record(void* ev, hipcstone_stream s)
{
    HC_CHECK(hipEventRecord(static_cast<hipEvent_t>(ev), stream(s)));
    return HIPCSTONE_OK;
}

hipcstone_status hipcstone_event_elapsed_ms(void* start, void* stop, float* ms)
{
    HC_CHECK(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
    HC_CHECK(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
    return HIPCSTONE_OK;
}

void hipcstone_make_box_f32(float xmin, float xmax, float ymin, float ymax, float zmin, float zmax, int bx, int by,
                            int bz, hipcstone_box_f32* box)
{
    makeBox<float>(xmin, xmax, ymin, ymax, zmin, zmax, bx, by, bz, box);
}

void hipcstone_make_box_f64(double xmin, double xmax, double ymin, double ymax, double zmin, double zmax, int bx,
                            int by, int bz, hipcstone_box_f64* box)
{
    makeBox<double>(xmin, xmax, ymin, ymax, zmin, zmax, bx, by, bz, box);
}

} // extern "C"
