// Density filters of filtered-design scripts: TO_convolve and its transpose
// (app/ffip.cpp:222-306), as gathers on the GPU.
//
//   out(i, j) = sum_{(m, n) in W(i, j)} in(m, n) f(m - i + r, n - j + r) / w(i, j)
//   w(i, j)   = sum_{(m, n) in W(i, j)} f(m - i + r, n - j + r)
// with W(i, j) the (2r+1)^2 window clipped to the array.  The transpose scatters
// in(i, j) f(.) / w(i, j) back onto the window; here each output element gathers
// those terms in the reference's (i, j) order, so no atomics are needed.
// Contraction is off and every sum runs in the reference's loop order, so the
// results equal the reference's operation for operation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>

namespace ffip_amd {

#define FL_OK(expr)                                                                                   \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr); \
    } while (0)

__global__ void k_window_weight(int n0, int n1, const double* __restrict__ f, int r, double* __restrict__ w)
{
#pragma clang fp contract(off)
    const int i = int(blockIdx.y) * int(blockDim.y) + int(threadIdx.y);
    const int j = int(blockIdx.x) * int(blockDim.x) + int(threadIdx.x);
    if (i >= n0 || j >= n1)
        return;
    const int fw = 2 * r + 1;
    const int l0 = max(i - r, 0), l1 = max(j - r, 0), u0 = min(i + r + 1, n0), u1 = min(j + r + 1, n1);
    double s = 0;
    for (int m = l0; m < u0; ++m)
        for (int n = l1; n < u1; ++n)
            s += f[(m - i + r) * fw + (n - j + r)];
    w[int64_t(i) * n1 + j] = s;
}

__global__ void k_to_convolve(int n0, int n1, const double* __restrict__ in, const double* __restrict__ f, int r,
                              const double* __restrict__ w, double* __restrict__ out)
{
#pragma clang fp contract(off)
    const int i = int(blockIdx.y) * int(blockDim.y) + int(threadIdx.y);
    const int j = int(blockIdx.x) * int(blockDim.x) + int(threadIdx.x);
    if (i >= n0 || j >= n1)
        return;
    const int fw = 2 * r + 1;
    const int l0 = max(i - r, 0), l1 = max(j - r, 0), u0 = min(i + r + 1, n0), u1 = min(j + r + 1, n1);
    double s = 0;
    for (int m = l0; m < u0; ++m)
        for (int n = l1; n < u1; ++n)
            s += in[int64_t(m) * n1 + n] * f[(m - i + r) * fw + (n - j + r)];
    out[int64_t(i) * n1 + j] = s / w[int64_t(i) * n1 + j];
}

__global__ void k_to_convolve_t(int n0, int n1, const double* __restrict__ in, const double* __restrict__ f, int r,
                                const double* __restrict__ w, double* __restrict__ out)
{
#pragma clang fp contract(off)
    const int m = int(blockIdx.y) * int(blockDim.y) + int(threadIdx.y);
    const int n = int(blockIdx.x) * int(blockDim.x) + int(threadIdx.x);
    if (m >= n0 || n >= n1)
        return;
    const int fw = 2 * r + 1;
    // the windows W(i, j) that contain (m, n), visited in increasing (i, j)
    const int l0 = max(m - r, 0), l1 = max(n - r, 0), u0 = min(m + r + 1, n0), u1 = min(n + r + 1, n1);
    double s = 0;
    for (int i = l0; i < u0; ++i)
        for (int j = l1; j < u1; ++j)
            s += in[int64_t(i) * n1 + j] * f[(m - i + r) * fw + (n - j + r)] / w[int64_t(i) * n1 + j];
    out[int64_t(m) * n1 + n] = s;
}

// Host driver: copies in, filters, copies out (the arrays are design-sized).
void to_convolve(const double* in, int64_t n0, int64_t n1, const double* filt, int64_t f0, int64_t f1, bool transpose,
                 double* out)
{
    if (n0 < 1 || n1 < 1 || n0 * n1 >= (int64_t(1) << 31))
        throw std::runtime_error("array size out of range");
    if (f0 % 2 == 0 || f1 % 2 == 0)
        throw std::runtime_error("filter shape is not odd");
    if (f0 != f1)
        throw std::runtime_error("filter must be square (the reference indexes both axes with the first radius)");
    const int r = int(f0 / 2);
    const size_t na = size_t(n0 * n1), nf = size_t(f0 * f1);
    double *din = nullptr, *df = nullptr, *dw = nullptr, *dout = nullptr;
    auto cleanup = [&] {
        (void)hipFree(din);
        (void)hipFree(df);
        (void)hipFree(dw);
        (void)hipFree(dout);
    };
    try {
        FL_OK(hipMalloc(&din, na * sizeof(double)));
        FL_OK(hipMalloc(&df, nf * sizeof(double)));
        FL_OK(hipMalloc(&dw, na * sizeof(double)));
        FL_OK(hipMalloc(&dout, na * sizeof(double)));
        FL_OK(hipMemcpy(din, in, na * sizeof(double), hipMemcpyHostToDevice));
        FL_OK(hipMemcpy(df, filt, nf * sizeof(double), hipMemcpyHostToDevice));
        dim3 block(32, 8);
        dim3 grid(unsigned((n1 + 31) / 32), unsigned((n0 + 7) / 8));
        k_window_weight<<<grid, block>>>(int(n0), int(n1), df, r, dw);
        FL_OK(hipGetLastError());
        if (transpose)
            k_to_convolve_t<<<grid, block>>>(int(n0), int(n1), din, df, r, dw, dout);
        else
            k_to_convolve<<<grid, block>>>(int(n0), int(n1), din, df, r, dw, dout);
        FL_OK(hipGetLastError());
        FL_OK(hipMemcpy(out, dout, na * sizeof(double), hipMemcpyDeviceToHost));
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
}

// ---------------------------------------------------------------- allocation-free form
// Window weights of the allocation-free entry point live in a file-scope device
// array (one per GPU, held until the process ends).  Host code must not pass
// the array's name to a kernel: it is a host-side stand-in.  Its device address
// is looked up once per device and kept.
constexpr int64_t NOALLOC_MAX = int64_t(1) << 21;  // elements (16 MiB of FP64 weights)
static __device__ __align__(256) double scratch[NOALLOC_MAX];

static std::mutex g_scratch_mu;
static double* g_scratch_ptr[64] = {};

static double* scratch_ptr()
{
    int dev = 0;
    FL_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64)
        throw std::runtime_error("device index out of range for the filter scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        FL_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    return g_scratch_ptr[dev];
}

// TO_convolve / TO_convolve_transpose on device arrays, asynchronously on
// `stream`, with no allocation: the window weights use the device's scratch
// array (calls that overlap on one device race on it).
void to_convolve_noalloc(const double* in, int64_t n0, int64_t n1, const double* filt, int64_t f0, int64_t f1,
                         bool transpose, double* out, void* stream_)
{
    const hipStream_t stream = static_cast<hipStream_t>(stream_);
    if (n0 < 1 || n1 < 1 || n0 * n1 >= (int64_t(1) << 31))
        throw std::runtime_error("array size out of range");
    if (n0 * n1 > NOALLOC_MAX)
        throw std::runtime_error("array of " + std::to_string(n0 * n1) + " elements exceeds the allocation-free limit of " +
                                 std::to_string(NOALLOC_MAX) + " elements");
    if (f0 % 2 == 0 || f1 % 2 == 0)
        throw std::runtime_error("filter shape is not odd");
    if (f0 != f1)
        throw std::runtime_error("filter must be square (the reference indexes both axes with the first radius)");
    const int r = int(f0 / 2);
    double* w = scratch_ptr();
    dim3 block(32, 8);
    dim3 grid(unsigned((n1 + 31) / 32), unsigned((n0 + 7) / 8));
    k_window_weight<<<grid, block, 0, stream>>>(int(n0), int(n1), filt, r, w);
    FL_OK(hipGetLastError());
    if (transpose)
        k_to_convolve_t<<<grid, block, 0, stream>>>(int(n0), int(n1), in, filt, r, w, out);
    else
        k_to_convolve<<<grid, block, 0, stream>>>(int(n0), int(n1), in, filt, r, w, out);
    FL_OK(hipGetLastError());
}

}  // namespace ffip_amd
