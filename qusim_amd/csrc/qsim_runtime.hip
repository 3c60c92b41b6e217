// qsim_runtime.hip -- error state, device selection, memory, streams, events
// and the per-device reduction scratch of libqsim_hip.so.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <set>
#include <utility>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "qsim_common.h"

namespace qsim {

static thread_local std::string g_last_error;

int set_error(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

static bool trace_enabled()
{
    static const bool on = [] {
        const char *v = std::getenv("QSIM_ROCTX");
        return v && v[0] == '1';
    }();
    return on;
}

TraceRange::TraceRange(const char *name) : on_(trace_enabled())
{
    if (on_) roctxRangePushA(name);
}

TraceRange::~TraceRange()
{
    if (on_) roctxRangePop();
}

int hip_error(hipError_t e, const char *what)
{
    std::string m = std::string("HIP error ") + hipGetErrorName(e) + " (" +
                    hipGetErrorString(e) + ") in " + what;
    // the runtime also keeps the error as the thread's last error; clear it,
    // or the launch check (hipGetLastError) after the next kernel would
    // report it again as that launch's failure
    (void)hipGetLastError();
    return set_error(e == hipErrorOutOfMemory ? QSIM_ERR_NOMEM : QSIM_ERR_HIP, m);
}

hipStream_t as_stream(qsim_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// One scratch block per device, grown on demand; reductions serialise on it.
struct DevScratch {
    void *dev = nullptr;
    size_t bytes = 0;
    void *host = nullptr;
};
static std::mutex g_scratch_mu;
static DevScratch g_scratch[64];

std::recursive_mutex &device_wait_lock()
{
    static std::recursive_mutex mu;
    return mu;
}

int scratch_acquire(size_t bytes, Scratch *out)
{
    QSIM_DEVICE_WAIT_GUARD;  // before g_scratch_mu: one lock order everywhere
    int d = 0;
    hipError_t e = hipGetDevice(&d);
    if (e != hipSuccess) return hip_error(e, "hipGetDevice");
    if (d < 0 || d >= 64) return set_error(QSIM_ERR_INVALID, "device index out of range");
    g_scratch_mu.lock();
    DevScratch &s = g_scratch[d];
    if (s.bytes < bytes) {
        if (s.dev) (void)hipFree(s.dev);
        s.dev = nullptr;
        s.bytes = 0;
        size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
        e = hipMalloc(&s.dev, want);
        if (e != hipSuccess) {
            g_scratch_mu.unlock();
            return hip_error(e, "hipMalloc(reduction scratch)");
        }
        s.bytes = want;
    }
    if (!s.host) {
        e = hipHostMalloc(&s.host, 256, hipHostMallocDefault);
        if (e != hipSuccess) {
            g_scratch_mu.unlock();
            return hip_error(e, "hipHostMalloc(reduction result)");
        }
    }
    out->dev = s.dev;
    out->host = s.host;
    return QSIM_OK;
}

void scratch_release() { g_scratch_mu.unlock(); }

int lds_limit(const void *kernel, size_t bytes)
{
    if (bytes <= 65536) return QSIM_OK;
    static std::mutex mu;
    static std::set<std::pair<const void *, int>> done;
    int d = 0;
    hipError_t e = hipGetDevice(&d);
    if (e != hipSuccess) return hip_error(e, "hipGetDevice");
    std::lock_guard<std::mutex> lk(mu);
    if (done.count({kernel, d})) return QSIM_OK;
    e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return hip_error(e, "hipFuncSetAttribute");
    done.insert({kernel, d});
    return QSIM_OK;
}

int device_cu_count()
{
    static int cached[64] = {0};
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess || d < 0 || d >= 64) return 256;
    if (!cached[d]) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, d) != hipSuccess) return 256;
        cached[d] = p.multiProcessorCount;
    }
    return cached[d];
}

}  // namespace qsim

using namespace qsim;

extern "C" {

const char *qsim_hip_last_error(void) { return g_last_error.c_str(); }

int qsim_hip_init(int device)
{
    int n = 0;
    QSIM_HIP(hipGetDeviceCount(&n));
    if (n <= 0) return set_error(QSIM_ERR_HIP, "no HIP device visible");
    QSIM_CHECK(device >= 0 && device < n, "qsim_hip_init: device index out of range");
    QSIM_DEVICE_WAIT_GUARD;
    QSIM_HIP(hipSetDevice(device));
    QSIM_HIP(hipFree(nullptr));  // create the context now, not at the first kernel
    return QSIM_OK;
}

int qsim_hip_device_info(char *name, size_t name_len, int *cu_count, size_t *mem_free,
                         size_t *mem_total)
{
    int d = 0;
    QSIM_HIP(hipGetDevice(&d));
    hipDeviceProp_t p;
    QSIM_HIP(hipGetDeviceProperties(&p, d));
    if (name && name_len) {
        std::snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
    }
    if (cu_count) *cu_count = p.multiProcessorCount;
    size_t f = 0, t = 0;
    QSIM_HIP(hipMemGetInfo(&f, &t));
    if (mem_free) *mem_free = f;
    if (mem_total) *mem_total = t;
    return QSIM_OK;
}

int qsim_hip_synchronize(qsim_stream_t stream)
{
    if (!stream) {
        QSIM_DEVICE_WAIT_GUARD;
        QSIM_HIP(hipStreamSynchronize(nullptr));
        return QSIM_OK;
    }
    QSIM_HIP(hipStreamSynchronize(as_stream(stream)));
    return QSIM_OK;
}

int qsim_hip_stream_create(qsim_stream_t *stream)
{
    QSIM_CHECK(stream, "null stream out-pointer");
    QSIM_DEVICE_WAIT_GUARD;
    hipStream_t s;
    QSIM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = reinterpret_cast<qsim_stream_t>(s);
    return QSIM_OK;
}

int qsim_hip_stream_destroy(qsim_stream_t stream)
{
    QSIM_DEVICE_WAIT_GUARD;
    if (stream) QSIM_HIP(hipStreamDestroy(as_stream(stream)));
    return QSIM_OK;
}

int qsim_hip_malloc(void **ptr, size_t bytes)
{
    QSIM_CHECK(ptr, "null out-pointer");
    *ptr = nullptr;
    if (bytes == 0) return QSIM_OK;
    QSIM_DEVICE_WAIT_GUARD;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        *ptr = nullptr;
        (void)hipGetLastError();  // not the next launch's error (see hip_error)
        return set_error(QSIM_ERR_NOMEM, "can't alloc " + std::to_string(bytes) +
                                             " bytes on the device: " + hipGetErrorString(e));
    }
    return QSIM_OK;
}

int qsim_hip_free(void *ptr)
{
    QSIM_DEVICE_WAIT_GUARD;
    if (ptr) QSIM_HIP(hipFree(ptr));
    return QSIM_OK;
}

int qsim_hip_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
    if (!bytes) return QSIM_OK;
    QSIM_CHECK(dst && src, "null pointer in memcpy_h2d");
    QSIM_DEVICE_WAIT_GUARD;
    QSIM_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return QSIM_OK;
}

int qsim_hip_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    if (!bytes) return QSIM_OK;
    QSIM_CHECK(dst && src, "null pointer in memcpy_d2h");
    QSIM_DEVICE_WAIT_GUARD;
    QSIM_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return QSIM_OK;
}

int qsim_hip_memcpy_d2d(void *dst, const void *src, size_t bytes, qsim_stream_t s)
{
    if (!bytes) return QSIM_OK;
    QSIM_CHECK(dst && src, "null pointer in memcpy_d2d");
    QSIM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(s)));
    return QSIM_OK;
}

int qsim_hip_event_create(void **ev)
{
    QSIM_CHECK(ev, "null out-pointer");
    hipEvent_t e;
    QSIM_HIP(hipEventCreate(&e));
    *ev = reinterpret_cast<void *>(e);
    return QSIM_OK;
}

int qsim_hip_event_destroy(void *ev)
{
    if (ev) QSIM_HIP(hipEventDestroy(reinterpret_cast<hipEvent_t>(ev)));
    return QSIM_OK;
}

int qsim_hip_event_record(void *ev, qsim_stream_t s)
{
    QSIM_CHECK(ev, "null event");
    QSIM_HIP(hipEventRecord(reinterpret_cast<hipEvent_t>(ev), as_stream(s)));
    return QSIM_OK;
}

int qsim_hip_event_elapsed_ms(void *start, void *stop, float *ms)
{
    QSIM_CHECK(start && stop && ms, "null event");
    QSIM_HIP(hipEventSynchronize(reinterpret_cast<hipEvent_t>(stop)));
    QSIM_HIP(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start),
                                 reinterpret_cast<hipEvent_t>(stop)));
    return QSIM_OK;
}

int qsim_hip_event_query(void *ev, int *done)
{
    QSIM_CHECK(ev && done, "null event");
    const hipError_t e = hipEventQuery(reinterpret_cast<hipEvent_t>(ev));
    if (e == hipErrorNotReady) {
        *done = 0;
        return QSIM_OK;
    }
    QSIM_HIP(e);
    *done = 1;
    return QSIM_OK;
}

int qsim_hip_event_synchronize(void *ev)
{
    QSIM_CHECK(ev, "null event");
    QSIM_HIP(hipEventSynchronize(reinterpret_cast<hipEvent_t>(ev)));
    return QSIM_OK;
}

}  // extern "C"
