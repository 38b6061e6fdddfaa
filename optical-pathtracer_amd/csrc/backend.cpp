// Backend identity, availability and device enumeration: the C-ABI side of owe::Backend's name(),
// summary(), available() and devices() (src/owe/render/backend.hpp:34-43), and the per-thread error
// text of every entry point.
#include <cstdio>
#include <cstring>
#include <string>

#include "common.hpp"

namespace owehip {

namespace {
thread_local std::string g_lastError;

bool isGfx950(const hipDeviceProp_t& p) { return std::strncmp(p.gcnArchName, "gfx950", 6) == 0; }

void copyText(char* dst, size_t cap, const std::string& s) {
    if (!dst || cap == 0) return;
    std::snprintf(dst, cap, "%s", s.c_str());
}
}  // namespace

void setLastError(const std::string& what) { g_lastError = what; }
void clearLastError() { g_lastError.clear(); }

void selectDevice(int device) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) fail(OWE_HIP_ERR_UNAVAILABLE, "no HIP device is visible");
    if (device < 0 || device >= n) fail(OWE_HIP_ERR_INVALID, "device index " + std::to_string(device) + " out of range");
    hipDeviceProp_t p;
    OWE_HIP_CHECK(hipGetDeviceProperties(&p, device));
    if (!isGfx950(p))
        fail(OWE_HIP_ERR_UNAVAILABLE, std::string("device ") + std::to_string(device) + " is " + p.gcnArchName +
                                          "; this backend is built for gfx950 only");
    OWE_HIP_CHECK(hipSetDevice(device));
}

}  // namespace owehip

using namespace owehip;

extern "C" {

const char* owe_hip_last_error(void) { return g_lastError.c_str(); }

const char* owe_hip_backend_name(void) { return "hip"; }

const char* owe_hip_backend_summary(void) {
    return "MI355X (gfx950) HIP kernels: float32 camera-relative wavefront transport; validated against the "
           "IEEE-754 double reference";
}

const char* owe_hip_backend_integrators(void) { return "path,light,sppm,vcm,bdpt,hybrid"; }

int owe_hip_available(char* why, size_t cap) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        copyText(why, cap, "no HIP device is visible (hipGetDeviceCount found none)");
        return 0;
    }
    for (int i = 0; i < n; ++i) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, i) == hipSuccess && isGfx950(p)) {
            copyText(why, cap, "");
            return 1;
        }
    }
    copyText(why, cap, "no gfx950 (MI355X) device among the visible HIP devices");
    return 0;
}

int owe_hip_device_count(int* count) {
    return guard([&] {
        require(count != nullptr, "count is null");
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
        *count = n;
    });
}

int owe_hip_device_info_get(int index, owe_hip_device_info* out) {
    return guard([&] {
        require(out != nullptr, "out is null");
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || index < 0 || index >= n)
            fail(OWE_HIP_ERR_INVALID, "no device " + std::to_string(index));
        hipDeviceProp_t p;
        OWE_HIP_CHECK(hipGetDeviceProperties(&p, index));
        std::memset(out, 0, sizeof *out);
        out->index = index;
        copyText(out->name, sizeof out->name, p.name);
        copyText(out->arch, sizeof out->arch, p.gcnArchName);
        copyText(out->type, sizeof out->type, p.integrated ? "integrated" : "discrete");
        out->hbm_bytes = uint64_t(p.totalGlobalMem);
        out->compute_units = p.multiProcessorCount;
        out->xcds = isGfx950(p) ? 8 : 1;
    });
}

}  // extern "C"
