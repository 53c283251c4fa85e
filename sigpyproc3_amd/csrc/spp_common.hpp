// Shared definitions for the libspp_hip C-ABI: status codes, the thread-local
// last-error string, and HIP call checking. Every extern "C" entry point returns
// an int status (SPP_OK on success) and never throws across the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/spp_hip.h"

namespace spp {

void set_error(const std::string& msg);
void clear_error();

struct Error {
    int code;
    std::string msg;
};

}  // namespace spp

// Validate an argument: on failure record the message and return SPP_EINVAL.
#define SPP_REQUIRE(cond, ...)                                   \
    do {                                                          \
        if (!(cond)) {                                            \
            char _b[512];                                         \
            std::snprintf(_b, sizeof(_b), __VA_ARGS__);           \
            spp::set_error(_b);                                   \
            return SPP_EINVAL;                                    \
        }                                                         \
    } while (0)

// Check a HIP runtime call: on failure record it and return SPP_EHIP.
#define SPP_HIP(call)                                                              \
    do {                                                                           \
        hipError_t _e = (call);                                                    \
        if (_e != hipSuccess) {                                                    \
            char _b[512];                                                          \
            std::snprintf(_b, sizeof(_b), "%s failed: %s (%s:%d)", #call,          \
                          hipGetErrorString(_e), __FILE__, __LINE__);              \
            spp::set_error(_b);                                                    \
            return SPP_EHIP;                                                       \
        }                                                                          \
    } while (0)

// Check the launch status of the kernel just enqueued.
#define SPP_LAUNCHED(name)                                                         \
    do {                                                                           \
        hipError_t _e = hipGetLastError();                                         \
        if (_e != hipSuccess) {                                                    \
            char _b[512];                                                          \
            std::snprintf(_b, sizeof(_b), "launch of %s failed: %s", name,         \
                          hipGetErrorString(_e));                                  \
            spp::set_error(_b);                                                    \
            return SPP_EHIP;                                                       \
        }                                                                          \
    } while (0)

namespace spp {

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Number of compute units of the current device (cached per device id).
int device_cu_count();

}  // namespace spp
