// Device scratch of the no-allocation entry point (paddi_range_discords_noalloc).
//
// A file-scope device array: each GPU the library runs on holds its own copy, from the
// library's load until the process ends.  Host code never passes the array's name to a
// kernel or a copy (on the host that name is a stand-in, not a device address); it uses the
// address hipGetSymbolAddress returns for the current device, taken once per device by
// scratch_setup (called from paddi_create) and kept in `addr`.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "engine.h"

namespace paddi {

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex mtx;
void *addr[MAX_DEVICES] = {};
}  // namespace

void *scratch_setup(int device)
{
    if (device < 0 || device >= MAX_DEVICES)
        throw HipError(PADDI_ERR_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(mtx);
    if (!addr[device]) {
        int cur = -1;
        HIPC(hipGetDevice(&cur));
        if (cur != device) throw HipError(PADDI_ERR_DEVICE, "scratch_setup: device not current");
        void *p = nullptr;
        HIPC(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw HipError(PADDI_ERR_DEVICE, "hipGetSymbolAddress returned no address");
        addr[device] = p;
    }
    return addr[device];
}

}  // namespace paddi
