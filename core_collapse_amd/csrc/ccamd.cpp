// ccamd.cpp -- library-wide C ABI: error reporting, version, device information.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

#include "common.hpp"

namespace cc {
static thread_local std::string g_last_error;
void set_error(cc_status code, const std::string& msg) {
    g_last_error = "[" + std::to_string((int)code) + "] " + msg;
}
}  // namespace cc

extern "C" {

const char* cc_last_error(void) { return cc::g_last_error.c_str(); }

int cc_version(void) { return 10000; }

cc_status cc_device_info(char* name, int name_len, int* cu_count) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        cc::set_error(CC_ERR_NO_DEVICE, "no HIP device");
        return CC_ERR_NO_DEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        cc::set_error(CC_ERR_HIP, "hipGetDeviceProperties failed");
        return CC_ERR_HIP;
    }
    if (name && name_len > 0) {
        std::string s = std::string(prop.name) + " (" + prop.gcnArchName + ")";
        std::strncpy(name, s.c_str(), (size_t)name_len - 1);
        name[name_len - 1] = 0;
    }
    if (cu_count) *cu_count = prop.multiProcessorCount;
    return CC_OK;
}

}  // extern "C"
