// dcp_polar_adjoint_noalloc: the 2x2 polar rotation and its adjoint of
// dcp_polar_adjoint without any device allocation.  Its device staging
// (inputs Ft and gR, outputs gF and R) lives in the file-scope array
// `scratch` below, DCP_NOALLOC_MAX_MATRICES matrices (1 MiB per GPU, held by
// every GPU the library runs on until the process ends).  Every call on one
// GPU shares that GPU's array: calls that overlap on one GPU (threads,
// processes sharing the device) race and must be serialised by the caller.
#include "kernels.hpp"

#include <map>
#include <mutex>
#include <string>

#define DCP_NOALLOC_MAX_MATRICES 16384   // 4 arrays x 16 B per matrix = 1 MiB

// Ft | gR | gF | R, row-major 2x2 per matrix
static __device__ __align__(256) float scratch[4 * 4 * DCP_NOALLOC_MAX_MATRICES];

// the library's thread-local error message (dcp_last_error), in engine.hip
int dcp_record_error(const std::string& m);

namespace {
std::mutex g_scratch_mu;
std::map<int, float*> g_scratch_ptr;     // device -> address of its copy of `scratch`

int noalloc_err(const std::string& m) { return dcp_record_error(m); }

// Device address of the current device's `scratch`, looked up once per device
// with that device current.  The host-side name is only a stand-in: every use
// below goes through this address.
int scratch_address(float** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return noalloc_err(std::string("hipGetDevice: ") + hipGetErrorString(e));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it == g_scratch_ptr.end()) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return noalloc_err(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        it = g_scratch_ptr.emplace(dev, (float*)p).first;
    }
    *out = it->second;
    return 0;
}
}  // namespace

extern "C" int dcp_polar_adjoint_noalloc(const float* Ft, const float* gR, int32_t n, int32_t precision,
                                         int32_t svd_grad, float* gF, float* R) {
    if (!Ft || !gR || !gF || !R || n < 0 || (precision != 0 && precision != 1) || (svd_grad != 0 && svd_grad != 1))
        return noalloc_err("dcp_polar_adjoint_noalloc: bad arguments");
    if (n > DCP_NOALLOC_MAX_MATRICES)
        return noalloc_err("dcp_polar_adjoint_noalloc: " + std::to_string(n) + " matrices exceed its static scratch of " +
                           std::to_string(DCP_NOALLOC_MAX_MATRICES));
    if (n == 0) return 0;
    float* base = nullptr;
    if (scratch_address(&base)) return -1;
    float *dF = base, *dg = base + 4 * (size_t)DCP_NOALLOC_MAX_MATRICES;
    float *dgF = base + 8 * (size_t)DCP_NOALLOC_MAX_MATRICES, *dR = base + 12 * (size_t)DCP_NOALLOC_MAX_MATRICES;
    const size_t bytes = 16 * (size_t)n;
    hipError_t e = hipMemcpy(dF, Ft, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dg, gR, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const int blocks = (n + 255) / 256;
        if (precision == 1)
            hipLaunchKernelGGL(k_polar_adj<float>, dim3(blocks), dim3(256), 0, nullptr, dF, dg, n, svd_grad, dgF, dR);
        else
            hipLaunchKernelGGL(k_polar_adj<double>, dim3(blocks), dim3(256), 0, nullptr, dF, dg, n, svd_grad, dgF, dR);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(gF, dgF, bytes, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(R, dR, bytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return noalloc_err(std::string("dcp_polar_adjoint_noalloc: ") + hipGetErrorString(e));
    return 0;
}
