// Sustained HBM bandwidth of plain streaming kernels on the current device:
// the ceiling against which the half-step kernels' byte rates are read.
//
// k_stream reads R arrays and writes W arrays of the same length, element i of
// each output being the sum of element i of the inputs (R = 0: a constant fill;
// W = 0: a reduction kept alive by a store that never happens).  Every lane moves
// 16-byte words (dwordx4), issues the loads of V words per input before using
// any of them, and strides over the arrays with a grid of a few workgroups per
// CU, so that each CU keeps many kilobytes in flight.  The half-step kernels
// read about two bytes for every byte they write, so R = 2, W = 1 is their
// natural ceiling; R = 1, W = 1 is a copy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace ffip_amd {

#define HIP_CHECK(expr)                                                                               \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr); \
    } while (0)

struct StreamArgs {
    const float4* in[3];
    float4* out[2];
    int64_t n;     // float4 words per array
    float4* sink;  // written only if a reduction hits an impossible value
};

template <int R, int W, int V>
__global__ void __launch_bounds__(256) k_stream(const StreamArgs A)
{
    const int64_t stride = int64_t(gridDim.x) * blockDim.x * V;
    for (int64_t base = (int64_t(blockIdx.x) * blockDim.x) * V + threadIdx.x; base < A.n; base += stride) {
        float4 acc[V];
#pragma unroll
        for (int v = 0; v < V; ++v)
            acc[v] = make_float4(1.f, 2.f, 3.f, 4.f);
        if constexpr (R > 0) {
            float4 x[R][V];
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const int64_t q = base + int64_t(v) * blockDim.x;
                    x[r][v] = q < A.n ? A.in[r][q] : make_float4(0.f, 0.f, 0.f, 0.f);
                }
#pragma unroll
            for (int v = 0; v < V; ++v) {
                acc[v] = x[0][v];
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    acc[v].x += x[r][v].x;
                    acc[v].y += x[r][v].y;
                    acc[v].z += x[r][v].z;
                    acc[v].w += x[r][v].w;
                }
            }
        }
        if constexpr (W > 0) {
#pragma unroll
            for (int w = 0; w < W; ++w)
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const int64_t q = base + int64_t(v) * blockDim.x;
                    if (q < A.n)
                        A.out[w][q] = acc[v];
                }
        } else {
            float s = 0.f;
#pragma unroll
            for (int v = 0; v < V; ++v)
                s += acc[v].x + acc[v].y + acc[v].z + acc[v].w;
            if (s == -1234.5f)  // inputs are all 1: never true
                A.sink[threadIdx.x] = acc[0];
        }
    }
}

// Mean and best rate (GB/s of bytes read + written) of `reps` launches moving
// `bytes` per array, reads R in {0..3}, writes W in {0..2}, R + W >= 1.
void stream_bandwidth(int R, int W, size_t bytes, int reps, double out[2])
{
    if (R < 0 || R > 3 || W < 0 || W > 2 || R + W == 0)
        throw std::runtime_error("stream kernel takes 0-3 input and 0-2 output arrays, at least one in all");
    if (reps < 1)
        throw std::runtime_error("reps must be positive");
    const int64_t n = int64_t(bytes / sizeof(float4));
    if (n < 1024)
        throw std::runtime_error("arrays of at least 16 KiB");
    int dev = 0, ncu = 0;
    HIP_CHECK(hipGetDevice(&dev));
    HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    std::vector<void*> bufs;
    auto release = [&] {
        for (void* p : bufs)
            (void)hipFree(p);
    };
    try {
        StreamArgs A{};
        A.n = n;
        for (int r = 0; r < R; ++r) {
            void* p = nullptr;
            HIP_CHECK(hipMalloc(&p, size_t(n) * sizeof(float4)));
            bufs.push_back(p);
            // all ones: the reduction's sink test can never pass
            HIP_CHECK(hipMemsetD32(static_cast<hipDeviceptr_t>(p), 0x3f800000, size_t(n) * 4));
            A.in[r] = static_cast<const float4*>(p);
        }
        for (int w = 0; w < W; ++w) {
            void* p = nullptr;
            HIP_CHECK(hipMalloc(&p, size_t(n) * sizeof(float4)));
            bufs.push_back(p);
            A.out[w] = static_cast<float4*>(p);
        }
        {
            void* p = nullptr;
            HIP_CHECK(hipMalloc(&p, 256 * sizeof(float4)));
            bufs.push_back(p);
            A.sink = static_cast<float4*>(p);
        }
        constexpr int V = 4;
        // a few workgroups per CU, each lane V words per input per pass
        const int64_t want = (n + 256 * V - 1) / (256 * V);
        const unsigned grid = unsigned(std::max<int64_t>(1, std::min<int64_t>(want, int64_t(ncu) * 8)));
        auto launch = [&] {
            switch (R * 3 + W) {
            case 1: k_stream<0, 1, V><<<grid, 256>>>(A); break;
            case 2: k_stream<0, 2, V><<<grid, 256>>>(A); break;
            case 3: k_stream<1, 0, V><<<grid, 256>>>(A); break;
            case 4: k_stream<1, 1, V><<<grid, 256>>>(A); break;
            case 5: k_stream<1, 2, V><<<grid, 256>>>(A); break;
            case 6: k_stream<2, 0, V><<<grid, 256>>>(A); break;
            case 7: k_stream<2, 1, V><<<grid, 256>>>(A); break;
            case 8: k_stream<2, 2, V><<<grid, 256>>>(A); break;
            case 9: k_stream<3, 0, V><<<grid, 256>>>(A); break;
            case 10: k_stream<3, 1, V><<<grid, 256>>>(A); break;
            default: k_stream<3, 2, V><<<grid, 256>>>(A); break;
            }
            HIP_CHECK(hipGetLastError());
        };
        launch();  // warm-up
        HIP_CHECK(hipDeviceSynchronize());
        std::vector<hipEvent_t> ev(size_t(reps) + 1);
        for (auto& e : ev)
            HIP_CHECK(hipEventCreate(&e));
        HIP_CHECK(hipEventRecord(ev[0], nullptr));
        for (int q = 0; q < reps; ++q) {
            launch();
            HIP_CHECK(hipEventRecord(ev[size_t(q) + 1], nullptr));
        }
        HIP_CHECK(hipEventSynchronize(ev[size_t(reps)]));
        const double moved = double(R + W) * double(n) * sizeof(float4);
        double total = 0, best = 0;
        for (int q = 0; q < reps; ++q) {
            float ms = 0;
            HIP_CHECK(hipEventElapsedTime(&ms, ev[size_t(q)], ev[size_t(q) + 1]));
            total += ms;
            best = std::max(best, moved / (double(ms) * 1e-3) / 1e9);
        }
        for (auto& e : ev)
            (void)hipEventDestroy(e);
        out[0] = moved * reps / (total * 1e-3) / 1e9;
        out[1] = best;
    } catch (...) {
        release();
        throw;
    }
    release();
}

}  // namespace ffip_amd
