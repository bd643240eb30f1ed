// Internal helpers shared by the MI355X assembly library translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace skfa {

void set_error(const std::string& msg);

// Status codes returned by every extern "C" entry point.
enum Status : int {
    kOk = 0,
    kInvalid = -1,    // invalid argument (shape, size, pointer)
    kHip = -2,        // HIP runtime error
    kUnsupported = -3,
    kCapacity = -4,   // a documented size limit was exceeded
};

#define SKFA_HIP(call)                                                                  \
    do {                                                                                \
        hipError_t err_ = (call);                                                       \
        if (err_ != hipSuccess) {                                                       \
            ::skfa::set_error(std::string(#call) + ": " + hipGetErrorString(err_));     \
            return ::skfa::kHip;                                                        \
        }                                                                               \
    } while (0)

#define SKFA_REQUIRE(cond, msg)                 \
    do {                                        \
        if (!(cond)) {                          \
            ::skfa::set_error(msg);             \
            return ::skfa::kInvalid;            \
        }                                       \
    } while (0)

#define SKFA_LAUNCH_CHECK()                                                              \
    do {                                                                                 \
        hipError_t err_ = hipGetLastError();                                             \
        if (err_ != hipSuccess) {                                                        \
            ::skfa::set_error(std::string("kernel launch: ") + hipGetErrorString(err_)); \
            return ::skfa::kHip;                                                         \
        }                                                                                \
    } while (0)

// Exclusive prefix sum over int64 (n items -> n+1 outputs, out[n] = total).
int exclusive_scan_i64(const int64_t* in, int64_t* out, int64_t n, hipStream_t s);
// Same over int32 inputs with int64 outputs.
int exclusive_scan_i32_to_i64(const int32_t* in, int64_t* out, int64_t n, hipStream_t s);

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    int alloc(size_t count) {
        release();
        n = count;
        if (count == 0) return kOk;
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
            return kHip;
        }
        return kOk;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
};

}  // namespace skfa
