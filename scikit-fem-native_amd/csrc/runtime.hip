// Error reporting and device/stream helpers of the C-ABI.
#include <hip/hip_runtime.h>

#include <string>

#include "skfa_internal.h"
#include "../../include/skfem_amd.h"

namespace skfa {
static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
}  // namespace skfa

extern "C" {

const char* skfa_last_error(void) { return skfa::g_last_error.c_str(); }

int skfa_abi_version(void) { return SKFA_ABI_VERSION; }

int skfa_device_name(char* buf, int32_t len) {
    SKFA_REQUIRE(buf != nullptr && len > 0, "null buffer");
    int dev = 0;
    SKFA_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    SKFA_HIP(hipGetDeviceProperties(&prop, dev));
    std::string name = prop.gcnArchName;
    name = name.substr(0, size_t(len - 1));
    for (size_t i = 0; i < name.size(); ++i) buf[i] = name[i];
    buf[name.size()] = 0;
    return skfa::kOk;
}

// Timing of a stream region with HIP events (used by the benchmark so that
// kernel time is measured on the stream the kernels are launched on).
int skfa_event_create(void** ev) {
    SKFA_REQUIRE(ev != nullptr, "null argument");
    hipEvent_t e;
    SKFA_HIP(hipEventCreate(&e));
    *ev = e;
    return skfa::kOk;
}

int skfa_event_destroy(void* ev) {
    SKFA_HIP(hipEventDestroy(static_cast<hipEvent_t>(ev)));
    return skfa::kOk;
}

int skfa_event_record(void* ev, void* stream) {
    SKFA_HIP(hipEventRecord(static_cast<hipEvent_t>(ev), static_cast<hipStream_t>(stream)));
    return skfa::kOk;
}

int skfa_event_elapsed_ms(void* start, void* stop, float* ms) {
    SKFA_REQUIRE(ms != nullptr, "null argument");
    SKFA_HIP(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
    SKFA_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
    return skfa::kOk;
}

int skfa_stream_synchronize(void* stream) {
    SKFA_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    return skfa::kOk;
}

}  // extern "C"
