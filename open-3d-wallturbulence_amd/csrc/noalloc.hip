// Scratch memory of the allocation-free entry points: one file-scope device
// array, which every GPU the library runs on holds until the process ends. Its
// device address differs per GPU and is looked up once per device, with that
// device current, when a channel is created there (hipGetSymbolAddress); host
// code uses only that address, never the array's name.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "device.hpp"
#include "owt_internal.hpp"

namespace owt {

// 4 Mi complex doubles = 64 MiB
constexpr size_t kScratchElems = (size_t)4 << 20;
static __device__ __align__(256) double2 scratch[kScratchElems];

namespace {
std::mutex g_scratch_mu;
std::vector<double2 *> g_scratch_ptr;  // per device ordinal
}  // namespace

size_t noalloc_scratch_bytes() { return kScratchElems * sizeof(double2); }

double2 *noalloc_scratch(int device) {
    std::lock_guard<std::mutex> l(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error(3, std::string("HIP error: scratch address on device ") + std::to_string(device) + ": " +
                               hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<double2 *>(p);
    }
    return g_scratch_ptr[device];
}

}  // namespace owt
