// Shared plumbing of the C ABI: per-thread error text, status codes, and HIP call checking.
// Entry points wrap their body in guard(), which turns any C++ exception into a status code and a
// message for owe_hip_last_error(); nothing throws across the extern "C" boundary.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "owe_hip.h"

namespace owehip {

// A failure with the status code the C ABI reports for it.
struct Failure : std::runtime_error {
    int code;
    Failure(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};

void setLastError(const std::string& what);
void clearLastError();

[[noreturn]] inline void fail(int code, const std::string& what) { throw Failure(code, what); }
inline void require(bool ok, const std::string& what) {
    if (!ok) fail(OWE_HIP_ERR_INVALID, what);
}

inline void hipCheck(hipError_t e, const char* call) {
    if (e != hipSuccess) fail(OWE_HIP_ERR_DEVICE, std::string(call) + ": " + hipGetErrorString(e));
}
#define OWE_HIP_CHECK(call) ::owehip::hipCheck((call), #call)

template <class F>
int guard(F&& body) {
    try {
        clearLastError();
        body();
        return OWE_HIP_OK;
    } catch (const Failure& f) {
        setLastError(f.what());
        return f.code;
    } catch (const std::exception& e) {
        setLastError(e.what());
        return OWE_HIP_ERR_INVALID;
    } catch (...) {
        setLastError("unknown error");
        return OWE_HIP_ERR_INVALID;
    }
}

// Makes `device` current after checking that it exists and is a gfx950.
void selectDevice(int device);

// Fails unless p points into device memory (of `device` when it is not negative): kernels must not
// be handed host pointers.
inline void requireDevicePointer(const void* p, const char* what, int device = -1) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess || a.type != hipMemoryTypeDevice ||
        (device >= 0 && a.device != device)) {
        (void)hipGetLastError();
        fail(OWE_HIP_ERR_INVALID, std::string(what) + " must point to device memory");
    }
}

}  // namespace owehip
