// Error reporting shared by every entry point of libafqmc_hip.so.
#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <unordered_set>

#include "afqmc_hip.h"
#include "common.h"

namespace {
thread_local char g_last_error[512] = "";
}

namespace afq {

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

int check_launch(const char* what) {
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        set_error("%s: launch failed: %s", what, hipGetErrorString(err));
        return -2;
    }
    return 0;
}

void allow_full_lds(const void* fn) {
    static std::mutex mu;
    static std::unordered_set<const void*> done;
    std::lock_guard<std::mutex> lock(mu);
    if (done.count(fn)) return;
    (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    (void)hipGetLastError();
    done.insert(fn);
}

int cu_count() {
    static std::mutex mu;
    static int counts[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        (void)hipGetLastError();
        return 256;
    }
    std::lock_guard<std::mutex> lock(mu);
    if (counts[dev] == 0) {
        int c = 0;
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0) {
            (void)hipGetLastError();
            c = 256;
        }
        counts[dev] = c;
    }
    return counts[dev];
}

}  // namespace afq

extern "C" const char* afq_last_error(void) { return g_last_error; }

extern "C" int afq_abi_version(void) { return 23; }
