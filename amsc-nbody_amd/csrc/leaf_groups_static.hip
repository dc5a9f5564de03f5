// leaf_groups_static.hip -- the leaf-group build without allocation: the
// run-length encode + exclusive scan of nbh_leaf_groups_build with its
// scratch words (the scan's look-back status words, its tile ticket and the
// group count) in a file-scope device array instead of a workspace.
//
// The array is sized for kStaticLeafGroupParticles particles and lives on
// every device the library runs on until the process ends.  Its device
// address is looked up once per device with hipGetSymbolAddress (the array's
// name in host code is not a device address) and kept per device.  Each call
// clears the status words it uses with a fill kernel, so a call never sees
// an earlier call's words; calls that overlap on one device share the array
// and race.
#include "device_utils.hpp"
#include "nbh_internal.hpp"

#include <mutex>

namespace nbh {

namespace {

constexpr int kStaticLeafGroupParticles = NBH_LEAF_GROUPS_NOALLOC_MAX;
constexpr size_t kStaticScanTile = 2048; // the scan's smallest tile (checked per call)
constexpr size_t kStaticStatusWords =
    (static_cast<size_t>(kStaticLeafGroupParticles) + kStaticScanTile - 1) / kStaticScanTile + 64;
// status words, then one word whose low half holds the group count
constexpr size_t kStaticWords = kStaticStatusWords + 1;

static __device__ __align__(256) unsigned long long scratch[kStaticWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned long long* g_scratch[kMaxDevices] = {};

void need(bool ok, int status, const char* message)
{
    if (!ok)
        fail(status, message);
}

// Device address of `scratch` on the current device (looked up on the
// first call made on that device).
unsigned long long* scratch_on_current_device()
{
    int dev = 0;
    NBH_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        fail(NBH_HIP_ERROR, "device ordinal outside the static scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch[dev] == nullptr) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            fail(NBH_HIP_ERROR, std::string("hipGetSymbolAddress of the static scratch failed: ") +
                                    hipGetErrorString(e));
        g_scratch[dev] = static_cast<unsigned long long*>(p);
    }
    return g_scratch[dev];
}

} // namespace

} // namespace nbh

using namespace nbh;

extern "C" int nbh_leaf_groups_build_noalloc(void* hip_stream, const void* d_sorted_keys,
                                             int key_bytes, int n_particles, int capacity,
                                             void* d_unique_keys, int* d_first_particle,
                                             int* d_particle_count, int* d_n_groups,
                                             int* h_n_groups)
{
    return guard([&] {
        need(n_particles > 0, NBH_INVALID_ARGUMENT,
             "buildMortonLeafGroups requires nParticles >= 1");
        need(n_particles <= capacity, NBH_INVALID_ARGUMENT,
             "nParticles exceeds MortonLeafGroups capacity");
        need(n_particles <= kStaticLeafGroupParticles, NBH_INVALID_ARGUMENT,
             "nParticles exceeds the allocation-free leaf-group limit of 2^28 particles");
        need(d_sorted_keys != nullptr, NBH_INVALID_ARGUMENT,
             "d_sortedKeys must not be null");
        need(d_unique_keys && d_first_particle && d_particle_count,
             NBH_LOGIC_ERROR, "MortonLeafGroups memory has not been allocated");
        need(key_bytes == 4 || key_bytes == 8, NBH_INVALID_ARGUMENT,
             "key_bytes must be 4 or 8");
        const size_t words = scan_status_words(static_cast<size_t>(n_particles));
        need(words <= kStaticStatusWords, NBH_INVALID_ARGUMENT,
             "leaf-group scan needs more status words than the static scratch holds");
        const hipStream_t st = static_cast<hipStream_t>(hip_stream);
        unsigned long long* status = scratch_on_current_device();
        int* d_ng = d_n_groups ? d_n_groups
                               : reinterpret_cast<int*>(status + kStaticStatusWords);
        launch_fill_u64(st, status, 0ull, words);
        if (key_bytes == 4)
            launch_leaf_groups<uint32_t>(st, static_cast<const uint32_t*>(d_sorted_keys),
                                         n_particles, static_cast<uint32_t*>(d_unique_keys),
                                         d_first_particle, d_particle_count, d_ng, status);
        else
            launch_leaf_groups<uint64_t>(st, static_cast<const uint64_t*>(d_sorted_keys),
                                         n_particles, static_cast<uint64_t*>(d_unique_keys),
                                         d_first_particle, d_particle_count, d_ng, status);
        if (h_n_groups) {
            int g = 0;
            NBH_HIP(hipMemcpyAsync(&g, d_ng, sizeof(int), hipMemcpyDeviceToHost, st));
            NBH_HIP(hipStreamSynchronize(st));
            if (g <= 0 || g > n_particles)
                fail(NBH_RUNTIME_ERROR, "DeviceRunLengthEncode produced an invalid group count");
            *h_n_groups = g;
        }
    });
}
