// common.hip -- errors, contexts, the grow-only workspace and fill kernels.
#include "nbh_internal.hpp"

#include <cstring>
#include <mutex>

namespace nbh {

[[noreturn]] void fail(int status, const std::string& message, int code)
{
    throw Error{status, code, message};
}

void hip_check(hipError_t e, const char* what)
{
    if (e == hipSuccess)
        return;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        fail(NBH_OUT_OF_MEMORY, std::string(what) + ": " + hipGetErrorString(e));
    fail(NBH_HIP_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

void* Workspace::get(int slot, size_t need)
{
    if (need == 0)
        need = 256;
    if (bytes[slot] >= need)
        return ptr[slot];
    if (ptr[slot]) {
        NBH_HIP(hipFree(ptr[slot]));
        ptr[slot] = nullptr;
        bytes[slot] = 0;
    }
    // 12.5% head-room so that slowly growing inputs do not reallocate on
    // every call; 256-byte granularity keeps every slot aligned.
    size_t want = need + need / 8;
    want = (want + 255) & ~size_t(255);
    NBH_HIP(hipMalloc(&ptr[slot], want));
    bytes[slot] = want;
    return ptr[slot];
}

void Workspace::release()
{
    for (int s = 0; s < SLOT_COUNT; ++s) {
        if (ptr[s])
            (void)hipFree(ptr[s]);
        ptr[s] = nullptr;
        bytes[s] = 0;
    }
}

int* error_words(nbh_context* ctx)
{
    return static_cast<int*>(ctx->ws.get(SLOT_ERRORS, 64 * sizeof(int)));
}

void sync_read(nbh_context* ctx, const int* d_src, int* h_dst, int count)
{
    NBH_HIP(hipMemcpyAsync(ctx->host_pinned, d_src, count * sizeof(int),
                           hipMemcpyDeviceToHost, ctx->stream));
    NBH_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(h_dst, ctx->host_pinned, count * sizeof(int));
}

// ------------------------------------------------------------ fills
// Plain fill kernels (used instead of memsets so that every clear is a
// kernel node when a caller captures the stream into a graph).
template <typename T>
__global__ void fill_kernel(T* __restrict__ p, T v, size_t n)
{
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
         i < n; i += stride)
        p[i] = v;
}

template <typename T>
static void fill(hipStream_t s, T* p, T v, size_t n)
{
    if (n == 0)
        return;
    const unsigned threads = 256;
    unsigned blocks = blocks_for(n, threads);
    if (blocks > 4096)
        blocks = 4096;
    fill_kernel<T><<<blocks, threads, 0, s>>>(p, v, n);
    NBH_HIP(hipGetLastError());
}

void launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t v, size_t n)
{
    fill<uint32_t>(s, p, v, n);
}
void launch_fill_i32(hipStream_t s, int* p, int v, size_t n)
{
    fill<int>(s, p, v, n);
}
void launch_fill_f64(hipStream_t s, double* p, double v, size_t n)
{
    fill<double>(s, p, v, n);
}
void launch_fill_u64(hipStream_t s, unsigned long long* p,
                     unsigned long long v, size_t n)
{
    fill<unsigned long long>(s, p, v, n);
}

} // namespace nbh

extern "C" {


int nbh_context_create(nbh_context** out, void* hip_stream)
{
    return nbh::guard([&] {
        if (!out)
            nbh::fail(NBH_INVALID_ARGUMENT, "context output pointer is null");
        auto* ctx = new nbh_context();
        ctx->stream = static_cast<hipStream_t>(hip_stream);
        hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&ctx->host_pinned),
                                     64 * sizeof(int), hipHostMallocDefault);
        if (e != hipSuccess) {
            delete ctx;
            nbh::hip_check(e, "hipHostMalloc");
        }
        *out = ctx;
    });
}

int nbh_copy_to_host(void* h_dst, const void* d_src, size_t bytes)
{
    return nbh::guard([&] {
        if (bytes == 0)
            return;
        if (!h_dst || !d_src)
            nbh::fail(NBH_INVALID_ARGUMENT, "copy pointers must not be null");
        NBH_HIP(hipDeviceSynchronize());
        NBH_HIP(hipMemcpy(h_dst, d_src, bytes, hipMemcpyDeviceToHost));
    });
}

int nbh_context_destroy(nbh_context* ctx)
{
    return nbh::guard([&] {
        if (!ctx)
            return;
        ctx->ws.release();
        if (ctx->host_pinned)
            (void)hipHostFree(ctx->host_pinned);
        delete ctx;
    });
}

} // extern "C"
