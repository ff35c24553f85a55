// Device memory of the entry points that allocate nothing (cles_batch_run_noalloc):
// a file-scope array per GPU, held from the first use of the library on that
// GPU until the process ends.  Its device address is looked up once per device
// with that device current (at context creation, else at the first call on
// that device) and host code only ever uses that looked-up pointer.
#include <mutex>

#include "cles_internal.h"

static __device__ __align__(256) double scratch[CLES_NOALLOC_DTS_MAX];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double *g_ptr[MAX_DEVICES] = {};

__global__ void __launch_bounds__(256) k_fill_f64(double *p, int64_t n, double v)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

}  // namespace

namespace cles {

void noalloc_setup(int device)
{
    if (device < 0 || device >= MAX_DEVICES) throw Error{CLES_ERR_INVALID, "device index beyond the scratch table"};
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return;
    int cur = 0;
    CLES_HIP(hipGetDevice(&cur));
    CLES_HIP(hipSetDevice(device));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    (void)hipSetDevice(cur);
    if (e != hipSuccess || !p)
        throw Error{CLES_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
    g_ptr[device] = (double *)p;
}

double *noalloc_dts(int device, int64_t doubles, hipStream_t stream)
{
    if (doubles < 0 || doubles > (int64_t)CLES_NOALLOC_DTS_MAX)
        throw Error{CLES_ERR_INVALID, "count * nrec = " + std::to_string(doubles) +
                                          " exceeds the allocation-free dt history (CLES_NOALLOC_DTS_MAX = " +
                                          std::to_string((int64_t)CLES_NOALLOC_DTS_MAX) + " doubles)"};
    noalloc_setup(device);
    double *p = g_ptr[device];
    // the array keeps the previous call's records: clear this call's part
    if (doubles > 0) {
        const int blocks = (int)std::min<int64_t>((doubles + 255) / 256, 1024);
        hipLaunchKernelGGL(k_fill_f64, dim3(blocks), dim3(256), 0, stream, p, doubles, 0.0);
        CLES_HIP(hipGetLastError());
    }
    return p;
}

}  // namespace cles
