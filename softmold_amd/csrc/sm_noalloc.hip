// Fixed device scratch for the entry points that allocate nothing.
//
// `scratch` is a file-scope device array that each GPU holds from the library's load until
// the process ends.  Host code never uses its name as an address: at context creation the
// device's copy is looked up once with hipGetSymbolAddress (with that device current) and
// the pointer is kept per device; the kernels take that pointer as an argument.
//
// Every call on one GPU shares that GPU's array and it keeps the last call's contents, so
// two such calls that overlap on one GPU (two host threads, or two streams) race.
#include "sm_internal.hpp"

#include <mutex>

namespace smdev {

namespace {

static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_ptr[kMaxDevices] = {};

// radius[r] = dist[r] / sums[4r + 3]: the division sm_vesicle_radius makes on the host
__global__ void k_head_radius(int nrep, const double *__restrict__ sums, const double *__restrict__ dist,
                              double *__restrict__ radius)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrep) return;
    radius[r] = dist[r] / sums[4 * r + 3];
}

}  // namespace

hipError_t noalloc_scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return hipSuccess;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_ptr[device] = static_cast<double *>(p);
    return hipSuccess;
}

double *noalloc_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_ptr[device];
}

void launch_head_radius(hipStream_t st, int nrep, const double *sums, const double *dist, double *radius)
{
    hipLaunchKernelGGL(k_head_radius, dim3((nrep + 63) / 64), dim3(64), 0, st, nrep, sums, dist, radius);
}

}  // namespace smdev
