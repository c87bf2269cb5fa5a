// gpe_get_state_noalloc: gpe_get_state without any allocation.  gpe_get_state stages the level
// through a device buffer it allocates per call (hipMallocAsync); this variant streams the level
// through `scratch`, a file-scope device array of SCRATCH_BYTES that every GPU the library runs
// on holds until the process ends, in chunks of at most that size: unpack a chunk of the ABI
// index range into the scratch, copy it to the caller's host array, wait, next chunk.
//
// In host code `scratch` names a host-side stand-in, not device memory: its device address is
// taken once per device with hipGetSymbolAddress (with that device current) and only that
// address is passed to kernels and copies.  Calls on one device share its scratch, so calls
// that overlap on one device race; every chunk is written in full before it is copied, so
// nothing needs clearing between calls.
#include "gpe_device.h"

#include <mutex>

namespace gpe {

constexpr long long SCRATCH_BYTES = 32ll << 20;  // 32 MiB: 2 Mi complex128 values per chunk
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) double2 scratch[SCRATCH_BYTES / sizeof(double2)];

// Device address of `scratch` on each device, looked up on the first call on that device.
static void *g_scratch_addr[MAX_DEVICES];
static std::mutex g_scratch_mutex;

static hipError_t scratch_address(int device, void **out)
{
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if (!g_scratch_addr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) {
            const hipError_t back = hipSetDevice(cur);
            if (e == hipSuccess) e = back;
        }
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_addr[device] = p;
    }
    *out = g_scratch_addr[device];
    return hipSuccess;
}

// ABI elements [first, first + count) of a level (ABI order c, z, y, x, v) into out[0, count).
template <typename R>
__global__ __launch_bounds__(256) void k_unpack_range(Geometry g, const vec2_t<R> *dev, double2 *out, long long first,
                                                      long long count)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    long long r = first + i;
    const int v = (int)(r % g.V);
    r /= g.V;
    const int x = (int)(r % g.nx);
    r /= g.nx;
    const int y = (int)(r % g.ny);
    r /= g.ny;
    const int z = (int)(r % g.nz);
    const int c = (int)(r / g.nz);
    const vec2_t<R> s = dev[line_index(g, c, z, y, v) + x];
    out[i] = make_double2((double)s.x, (double)s.y);
}

}  // namespace gpe

using namespace gpe;

extern "C" int gpe_get_state_noalloc(gpe_ctx *ctx, int level, double *psi)
{
    CtxView cv;
    if (ctx_view(ctx, &cv) != GPE_OK) return GPE_EINVAL;
    if (!psi) return ctx_fail(ctx, GPE_EINVAL, "null state pointer");
    if (level != 0 && level != 1) return ctx_fail(ctx, GPE_EINVAL, "level must be 0 or 1");
    void *sp = nullptr;
    hipError_t e = scratch_address(cv.device, &sp);
    if (e != hipSuccess) return ctx_fail(ctx, GPE_EDEVICE, "no device address for the scratch array");
    const Geometry &g = cv.g;
    const long long total = (long long)cv.ncomp * g.nz * g.ny * g.nx * g.V;
    const long long chunk = SCRATCH_BYTES / (long long)sizeof(double2);
    double2 *stage = (double2 *)sp;
    for (long long first = 0; first < total && e == hipSuccess; first += chunk) {
        const long long n = total - first < chunk ? total - first : chunk;  // n <= chunk: within the scratch
        const unsigned grid = (unsigned)((n + 255) / 256);
        if (cv.fp64)
            hipLaunchKernelGGL(k_unpack_range<double>, dim3(grid), dim3(256), 0, cv.stream, g,
                               (const double2 *)cv.level[level], stage, first, n);
        else
            hipLaunchKernelGGL(k_unpack_range<float>, dim3(grid), dim3(256), 0, cv.stream, g,
                               (const float2 *)cv.level[level], stage, first, n);
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(psi + 2 * first, stage, (size_t)n * sizeof(double2), hipMemcpyDeviceToHost, cv.stream);
        // the next chunk overwrites the scratch: wait for this one's copy
        if (e == hipSuccess) e = hipStreamSynchronize(cv.stream);
    }
    if (e != hipSuccess) return ctx_fail(ctx, GPE_EDEVICE, hipGetErrorString(e));
    return GPE_OK;
}
