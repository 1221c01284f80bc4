// rs_runtime.cpp -- library runtime: device selection, the library stream, error text and
// plain device-memory helpers.
#include <cstdio>
#include <mutex>

#include "rs_internal.hpp"

namespace rs {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg)
{
    set_error(msg);
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    std::string msg = std::string(what) + ": " + hipGetErrorString(e);
    set_error(msg);
    if (e == hipErrorOutOfMemory) return RS_ERR_OUT_OF_MEMORY;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return RS_ERR_NO_DEVICE;
    return RS_ERR_DEVICE;
}

Runtime &runtime()
{
    static Runtime rt;
    return rt;
}

int require_runtime()
{
    Runtime &rt = runtime();
    if (!rt.ready) {
        int rc = rs_init(0);
        if (rc != RS_OK) return rc;
    }
    RS_HIP(hipSetDevice(rt.device));
    return RS_OK;
}

}  // namespace rs

using namespace rs;

extern "C" const char *rs_last_error(void) { return g_last_error.c_str(); }

extern "C" int rs_version(int *major, int *minor)
{
    if (major) *major = 0;
    if (minor) *minor = 1;
    return RS_OK;
}

extern "C" int rs_device_count(int *count)
{
    if (!count) return fail(RS_ERR_INVALID_ARG, "rs_device_count: null");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return RS_OK;
}

extern "C" int rs_init(int device)
{
    static std::mutex mu;
    std::lock_guard<std::mutex> lock(mu);
    Runtime &rt = runtime();
    if (rt.ready && rt.device == device) return RS_OK;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return fail(RS_ERR_NO_DEVICE, "rs_init: no HIP device visible");
    if (device < 0 || device >= n) return fail(RS_ERR_INVALID_ARG, "rs_init: device index out of range");
    RS_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    RS_HIP(hipGetDeviceProperties(&prop, device));
    if (rt.stream) (void)hipStreamDestroy(rt.stream);
    RS_HIP(hipStreamCreateWithFlags(&rt.stream, hipStreamNonBlocking));
    rt.device = device;
    rt.n_cu = prop.multiProcessorCount;
    rt.ready = true;
    // the allocation-free solve's workspace address on this device; a failure leaves it unset
    // and rs_solver_solve_device_noalloc tries again and reports the error
    if (scratch_setup(device) != RS_OK) set_error("");
    return RS_OK;
}

extern "C" int rs_finalize(void)
{
    Runtime &rt = runtime();
    if (!rt.ready) return RS_OK;
    (void)hipStreamSynchronize(rt.stream);
    (void)hipStreamDestroy(rt.stream);
    rt.stream = nullptr;
    rt.ready = false;
    return RS_OK;
}

extern "C" int rs_dmalloc(void **ptr, size_t bytes)
{
    if (!ptr) return fail(RS_ERR_INVALID_ARG, "rs_dmalloc: null");
    int rc = require_runtime();
    if (rc) return rc;
    RS_HIP(hipMalloc(ptr, bytes ? bytes : 1));
    return RS_OK;
}

extern "C" int rs_dfree(void *ptr)
{
    if (!ptr) return RS_OK;
    RS_HIP(hipFree(ptr));
    return RS_OK;
}

static int copy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind)
{
    if (!bytes) return RS_OK;
    if (!dst || !src) return fail(RS_ERR_INVALID_ARG, "rs_memcpy: null pointer");
    int rc = require_runtime();
    if (rc) return rc;
    RS_HIP(hipMemcpyAsync(dst, src, bytes, kind, runtime().stream));
    RS_HIP(hipStreamSynchronize(runtime().stream));
    return RS_OK;
}

extern "C" int rs_memcpy_htod(void *dst, const void *src, size_t bytes) { return copy(dst, src, bytes, hipMemcpyHostToDevice); }
extern "C" int rs_memcpy_dtoh(void *dst, const void *src, size_t bytes) { return copy(dst, src, bytes, hipMemcpyDeviceToHost); }
extern "C" int rs_memcpy_dtod(void *dst, const void *src, size_t bytes) { return copy(dst, src, bytes, hipMemcpyDeviceToDevice); }

extern "C" int rs_synchronize(void)
{
    int rc = require_runtime();
    if (rc) return rc;
    RS_HIP(hipStreamSynchronize(runtime().stream));
    return RS_OK;
}
