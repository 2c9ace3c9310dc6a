// Divergence error of device-resident fields without allocating
// (cmcx_div_error_fields).
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops): the kernels' partial sums live in a
// file-scope device array, `scratch`, of fixed size, which each GPU holds until
// the process ends.  Host code never uses the array's name as an address (on
// the host it is only a stand-in symbol): the address on each device comes from
// hipGetSymbolAddress on the entry point's first call on that device, with that
// device current, and is kept per device.
//
// The partials are fully rewritten by every call (one (sum, count) pair per
// block of the reduction, then the two totals), so nothing needs clearing
// between calls.  Every call on a GPU shares that GPU's array: calls that
// overlap on one GPU (other host threads, or other streams still running a
// previous call) race on it.
#include "cmcx_internal.h"

#include <mutex>

namespace cmcx {

namespace {

// 2 totals + div_partials() partial doubles of the divergence reduction
constexpr int kScratchDoubles = 2 + 2 * 2048;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::once_flag g_once[kMaxDevices];
double* g_ptr[kMaxDevices] = {nullptr};
hipError_t g_err[kMaxDevices] = {hipSuccess};

// Device address of `scratch` on the current device (looked up once per device).
int scratch_on_device(double** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        set_error(std::string("cmcx_div_error_fields: no usable current device (") +
                  hipGetErrorString(e) + ")");
        return CMCX_ERR_HIP;
    }
    std::call_once(g_once[dev], [dev]() {
        void* p = nullptr;
        g_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_ptr[dev] = g_err[dev] == hipSuccess ? static_cast<double*>(p) : nullptr;
    });
    if (g_err[dev] != hipSuccess || !g_ptr[dev]) {
        set_error(std::string("cmcx_div_error_fields: hipGetSymbolAddress failed: ") +
                  hipGetErrorString(g_err[dev]));
        return CMCX_ERR_HIP;
    }
    *out = g_ptr[dev];
    return CMCX_OK;
}

template <class T>
int div_fields(const void* u, const void* v, const void* w, const uint8_t* type, const Geom& g,
               const double* h, hipStream_t st, double* err) {
    double* sp = nullptr;
    int rc = scratch_on_device(&sp);
    if (rc != CMCX_OK) return rc;
    static_assert(kScratchDoubles >= 2, "scratch holds the totals");
    if (2 + div_partials() > kScratchDoubles) {
        set_error("cmcx_div_error_fields: scratch smaller than the reduction's partials");
        return CMCX_ERR_TOO_LARGE;
    }
    Layer4<T> L{};
    L.f[0] = const_cast<T*>(static_cast<const T*>(u));
    L.f[1] = const_cast<T*>(static_cast<const T*>(v));
    L.f[2] = const_cast<T*>(static_cast<const T*>(w));
    launch_div_error<T>(g, type, L, (T)h[0], (T)h[1], (T)h[2], sp, sp + 2, st);
    double tot[2] = {0.0, 0.0};
    CMCX_HIP(hipMemcpyAsync(tot, sp, sizeof(tot), hipMemcpyDeviceToHost, st));
    CMCX_HIP(hipStreamSynchronize(st));
    *err = tot[0] / tot[1];
    return CMCX_OK;
}

}  // namespace

}  // namespace cmcx

using namespace cmcx;

extern "C" int cmcx_div_error_fields(const void* u, const void* v, const void* w,
                                     const uint8_t* type, const int32_t* dims, const double* h,
                                     int dtype, void* stream, double* err) {
    if (!u || !v || !w || !type || !dims || !h || !err) return CMCX_ERR_ARG;
    if (dtype != CMCX_F32 && dtype != CMCX_F64) return CMCX_ERR_ARG;
    for (int a = 0; a < 3; a++)
        if (dims[a] < 3) {
            set_error("cmcx_div_error_fields: every dimension must be >= 3");
            return CMCX_ERR_ARG;
        }
    const int64_t N = (int64_t)dims[0] * dims[1] * dims[2];
    if (N > ((int64_t)1 << 31) - 1) {
        set_error("cmcx_div_error_fields: grid too large (node count must fit in int32)");
        return CMCX_ERR_TOO_LARGE;
    }
    Geom g{};
    g.nx = dims[0];
    g.ny = dims[1];
    g.nz = dims[2];
    g.sy = g.nz;
    g.sx = (int64_t)g.ny * g.nz;
    g.N = N;
    try {
        const hipStream_t st = static_cast<hipStream_t>(stream);
        return dtype == CMCX_F32 ? div_fields<float>(u, v, w, type, g, h, st, err)
                                 : div_fields<double>(u, v, w, type, g, h, st, err);
    } catch (const HipError& e) {
        set_error(std::string("HIP error ") + hipGetErrorString(e.code) + " in " + e.where);
        return CMCX_ERR_HIP;
    } catch (const std::string& e) {
        set_error(e);
        return CMCX_ERR_ARG;
    }
}
