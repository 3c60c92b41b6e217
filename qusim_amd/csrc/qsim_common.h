// qsim_common.h -- shared device/host helpers of libqsim_hip.so (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>
#include <cstdint>
#include <string>

#include "qsim_hip.h"

// ------------------------------------------------------------------ complex
// Interleaved complex of T (double -> 16 B, float -> 8 B), aligned so one
// element is one vector load.
template <typename T>
struct alignas(2 * sizeof(T)) cx {
    T x, y;
};
using c128 = cx<double>;
using c64 = cx<float>;

template <typename T>
__host__ __device__ __forceinline__ cx<T> mk(T a, T b)
{
    cx<T> r;
    r.x = a;
    r.y = b;
    return r;
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b)
{
    return mk(a.x + b.x, a.y + b.y);
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b)
{
    return mk(a.x - b.x, a.y - b.y);
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator*(cx<T> a, cx<T> b)
{
    return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator*(cx<T> a, T s)
{
    return mk(a.x * s, a.y * s);
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator*(T s, cx<T> a)
{
    return mk(s * a.x, s * a.y);
}
// a * conj(b)
template <typename T>
__host__ __device__ __forceinline__ cx<T> mulconj(cx<T> a, cx<T> b)
{
    return mk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename T>
__host__ __device__ __forceinline__ T abs2(cx<T> a)
{
    return a.y * a.y + a.x * a.x;
}
// i * c  (reference ExecTrait::IMul, src/DeviceTemplate.h:454-457)
template <typename T>
__host__ __device__ __forceinline__ cx<T> imul(cx<T> c)
{
    return mk(-c.y, c.x);
}
// exp(z) = e^re (cos im, sin im)
__device__ __forceinline__ c128 cexp(c128 z)
{
    double s, c;
    sincos(z.y, &s, &c);
    double e = exp(z.x);
    return mk(e * c, e * s);
}
// a / b, Smith's algorithm
__device__ __forceinline__ c128 cdiv(c128 a, c128 b)
{
    if (fabs(b.x) >= fabs(b.y)) {
        double r = b.y / b.x, d = b.x + b.y * r;
        return mk((a.x + a.y * r) / d, (a.y - a.x * r) / d);
    }
    double r = b.x / b.y, d = b.x * r + b.y;
    return mk((a.x * r + a.y) / d, (a.y * r - a.x) / d);
}
template <typename D, typename S>
__host__ __device__ __forceinline__ cx<D> cvt(cx<S> a)
{
    return mk<D>((D)a.x, (D)a.y);
}

// fold_half(i, N) = i > N/2 ? N - i : i  (src/DeviceTemplate.h:544-547)
// Non-temporal store for sweep outputs: the next sweep reads them from
// other XCDs, so keeping the lines in this XCD's L2 gains nothing (measured:
// alternating row / column passes run faster with them).
template <typename T>
__device__ __forceinline__ void store_nt(cx<T> *p, cx<T> v)
{
    // global address space: a global (not flat) store, no LDS counter wait
    using GT = __attribute__((address_space(1))) T;
    __builtin_nontemporal_store(v.x, (GT *)&p->x);
    __builtin_nontemporal_store(v.y, (GT *)&p->y);
}

// Pointers that kernels read from descriptor structs are generic (flat) to
// the compiler, and flat accesses count against the LDS counter too: an LDS
// wait inside an FFT exchange would also wait for every prefetched global
// load.  ld_g / st_g access through global-address-space pointers instead.
template <typename T>
__device__ __forceinline__ cx<T> ld_g(const cx<T> *p)
{
    using GT = __attribute__((address_space(1))) const T;
    const GT *q = (const GT *)&p->x;
    return mk(q[0], q[1]);
}
// Non-temporal load of data read once per sweep (psi lines).
template <typename T>
__device__ __forceinline__ cx<T> ld_nt(const cx<T> *p)
{
    using GT = __attribute__((address_space(1))) const T;
    const GT *q = (const GT *)&p->x;
    return mk(__builtin_nontemporal_load(q), __builtin_nontemporal_load(q + 1));
}
__device__ __forceinline__ double ld_g(const double *p)
{
    using GT = __attribute__((address_space(1))) const double;
    return *(const GT *)p;
}
template <typename T>
__device__ __forceinline__ void st_g(cx<T> *p, cx<T> v)
{
    using GT = __attribute__((address_space(1))) T;
    GT *q = (GT *)&p->x;
    q[0] = v.x;
    q[1] = v.y;
}

__host__ __device__ __forceinline__ size_t fold_half(size_t i, size_t n)
{
    return (i > n / 2) ? (n - i) : i;
}

// Preconditioner gamma of op_Vellekoop (KIND 1), op_Hao1 (2), op_Hao2 (3),
// src/DeviceTemplate.h:1020-1077
template <int KIND>
__device__ __forceinline__ c128 born_gamma(c128 v, double slow, double eps)
{
    c128 g;
    if (KIND == 1) {
        c128 iv = imul(v);
        g = mk(1. - iv.x / eps, -iv.y / eps);
    } else if (KIND == 2) {
        c128 iv = imul(mk(v.x, -v.y));
        g = mk(1. - iv.x / eps, -iv.y / eps);
    } else {
        c128 iv = imul(v);
        g = cdiv(mk(2., 0.), mk(1. + iv.x / eps, iv.y / eps));
    }
    return g * slow;
}

static inline c128 to_c(qsim_c128 a) { return mk(a.re, a.im); }

// ------------------------------------------------------------------ errors
namespace qsim {
int set_error(int code, const std::string &msg);
int hip_error(hipError_t e, const char *what);
hipStream_t as_stream(qsim_stream_t s);
// Per-device scratch for reductions: at least `bytes` of device memory and a
// pinned host slot; the caller holds the returned lock while using it.
struct Scratch {
    void *dev;
    void *host;
};
int scratch_acquire(size_t bytes, Scratch *out);
void scratch_release();
int device_cu_count();
// Raise a kernel's dynamic-LDS limit to `bytes` (no-op up to 64 KiB), once per
// (kernel, device) for the whole process and under a lock: the attribute is
// process state, and request loops launch from several threads at once.
int lds_limit(const void *kernel, size_t bytes);
// Graph capture and the calls that wait on every stream of the device
// (hipFree, hipDeviceSynchronize, synchronous copies, null-stream waits,
// stream creation and destruction) never overlap: a device-wide wait issued
// by one thread reaches the streams another thread is capturing and
// invalidates that capture (seen in a request loop whose Born workers
// created and freed jobs while the main thread captured a group advance).
// Both sides hold this process-wide lock; it is recursive because entry
// points nest.
std::recursive_mutex &device_wait_lock();
#define QSIM_DEVICE_WAIT_GUARD \
    std::lock_guard<std::recursive_mutex> qsim_dwg_(qsim::device_wait_lock())
// A roctx range (rocprofv3 --marker-trace) around an entry point, recorded
// only when the environment sets QSIM_ROCTX=1; otherwise it costs a branch.
// Asynchronous entry points mark the host-side enqueue, synchronous ones the
// whole call.
class TraceRange {
public:
    explicit TraceRange(const char *name);
    ~TraceRange();
    TraceRange(const TraceRange &) = delete;
    TraceRange &operator=(const TraceRange &) = delete;

private:
    bool on_;
};
}  // namespace qsim

#define QSIM_HIP(call)                                                \
    do {                                                              \
        hipError_t e_ = (call);                                       \
        if (e_ != hipSuccess) return qsim::hip_error(e_, #call);      \
    } while (0)

#define QSIM_CHECK(cond, msg)                                         \
    do {                                                              \
        if (!(cond)) return qsim::set_error(QSIM_ERR_INVALID, (msg)); \
    } while (0)

#define QSIM_TRY(expr)                    \
    do {                                  \
        int rc_ = (expr);                 \
        if (rc_ != QSIM_OK) return rc_;   \
    } while (0)
