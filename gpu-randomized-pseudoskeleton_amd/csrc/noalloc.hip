// Allocation-free pseudo-inverse: the core pinv of the CUR round (the same
// kernels as farcur_pinv) with every scratch buffer in a file-scope device
// array, for callers that cannot allocate device memory (HIP graph capture,
// serving loops that forbid hipMalloc).
//
// Scratch layout (bytes, 256-aligned sections), sized for r <= kNoAllocMaxR:
//   W copy    r^2 c128            the pinv overwrites its input with QR reflectors
//   V         pinv_v_elems(r) c128 Jacobi scratch of the global-memory variant
//   sigma     r doubles            singular values in Jacobi column order
//   job       one BlockJob record
//   list      the one-entry job list
//   out       one JobOut (kept, sweeps, ...)
// The array lives for the process on every GPU the library runs on. Its
// device address is taken once per device at context creation
// (noalloc_init) and used from then on; the array's name is only valid in
// device code.
#include <algorithm>
#include <mutex>
#include <vector>

#include "kernels.hpp"
#include "noalloc.hpp"

namespace farcur {

namespace {

constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t kOffW = 0;
constexpr size_t kOffV = align256(kOffW + (size_t)kNoAllocMaxR * kNoAllocMaxR * sizeof(c128));
constexpr size_t kOffSigma = align256(kOffV + ((size_t)kNoAllocMaxR * kNoAllocMaxR + (kNoAllocMaxR + 1) / 2) * sizeof(c128));
constexpr size_t kOffJob = align256(kOffSigma + (size_t)kNoAllocMaxR * sizeof(double));
constexpr size_t kOffList = align256(kOffJob + sizeof(BlockJob));
constexpr size_t kOffOut = align256(kOffList + 4 * sizeof(int32_t));
constexpr size_t kScratchBytes = align256(kOffOut + sizeof(JobOut));

}  // namespace

// The scratch array (about 2 MiB), one per GPU for the process lifetime.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_mu;
std::vector<unsigned char*> g_ptr;  // device address of `scratch`, per device

// Copies W into the scratch copy and writes the job record, its list and a
// zeroed JobOut (a kernel rather than host copies or memsets, so the call
// stays valid under stream capture).
__global__ void k_noalloc_setup(unsigned char* base, const c128* __restrict__ W, int r, BlockJob job) {
    c128* Wc = (c128*)(base + kOffW);
    const size_t rr = (size_t)r * r;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < rr; i += (size_t)gridDim.x * blockDim.x)
        Wc[i] = W[i];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *(BlockJob*)(base + kOffJob) = job;
        int32_t* list = (int32_t*)(base + kOffList);
        list[0] = 0;
        *(JobOut*)(base + kOffOut) = JobOut{};
    }
}

// sigma_out = the singular values sorted descending (rank by counting:
// larger values first, equal values in column order), *kept_out = kept.
__global__ void k_noalloc_finish(const unsigned char* base, int r, double* __restrict__ sigma_out,
                                 int* __restrict__ kept_out) {
    const double* s = (const double*)(base + kOffSigma);
    if (sigma_out)
        for (int i = threadIdx.x; i < r; i += blockDim.x) {
            const double si = s[i];
            int rank = 0;
            for (int j = 0; j < r; ++j) {
                const double sj = s[j];
                rank += (sj > si) || (sj == si && j < i);
            }
            sigma_out[rank] = si;
        }
    if (kept_out && threadIdx.x == 0) *kept_out = ((const JobOut*)(base + kOffOut))->kept;
}

}  // namespace

size_t noalloc_scratch_bytes() { return kScratchBytes; }

int noalloc_init(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    if ((int)g_ptr.size() <= device) g_ptr.resize(device + 1, nullptr);
    if (g_ptr[device]) return 0;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return (int)(e != hipSuccess ? e : hipErrorInvalidSymbol);
    g_ptr[device] = (unsigned char*)p;
    return 0;
}

unsigned char* noalloc_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    return device >= 0 && device < (int)g_ptr.size() ? g_ptr[device] : nullptr;
}

void launch_pinv_noalloc(unsigned char* base, const c128* W, int r, double atol, double rtol, int rmax, int threads,
                         c128* U, double* sigma, int* kept, hipStream_t s) {
    BlockJob job{};
    job.r = r;
    job.W = (c128*)(base + kOffW);
    job.V = (c128*)(base + kOffV);
    job.U = U;
    job.sigma = (double*)(base + kOffSigma);
    const size_t rr = (size_t)r * r;
    const int grid = (int)std::min<size_t>((rr + 255) / 256, 256);
    hipLaunchKernelGGL(k_noalloc_setup, dim3(grid), dim3(256), 0, s, base, W, r, job);
    launch_pinv((const BlockJob*)(base + kOffJob), (const int32_t*)(base + kOffList), 1, atol, rtol, 0.0, rmax,
                threads, 0, 0, 0, 0, (JobOut*)(base + kOffOut), s);
    hipLaunchKernelGGL(k_noalloc_finish, dim3(1), dim3(256), 0, s, (const unsigned char*)base, r, sigma, kept);
}

}  // namespace farcur
