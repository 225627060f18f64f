// Runtime part of the C ABI: errors, device memory, streams, events, box construction, byte fill.
#include "runtime.hpp"

#include "common.hpp"

namespace hipcstone
{

static thread_local std::string g_lastError;

void setError(const std::string& msg) { g_lastError = msg; }

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
