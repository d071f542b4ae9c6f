// Kernels of the design surrogate: a small 3-D convolutional network over the
// design density, trained and evaluated with BF16 MFMA (FP32 accumulation).
//
// A convolution is an im2col gather (channels-last activations, K = 27 * Cin
// taps ordered (kz, ky, kx, cin)) followed by one GEMM on the matrix cores; its
// backward pass is two more GEMMs on the same kernel (weight gradient col^T dZ,
// data gradient dZ W^T) and a col2im gather.  Everything here works on raw
// device pointers and runs on the caller's stream; no allocation.
//
// The surrogate has no counterpart in the reference: it proposes objective and
// gradient values for inner line-search steps of an adjoint optimisation, and
// every design it proposes is re-judged by the full FDTD + adjoint solve.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "surrogate.hpp"

namespace ffip_amd {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

#define SG_OK(expr)                                                                                   \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr); \
    } while (0)

// ---------------------------------------------------------------- GEMM
// C (+)= op(A) op(B): op(A) is M x K, op(B) is K x N, both bf16; C is FP32 row-major
// with leading dimension ldc.  TA: A is stored K x M (A[k*lda + m]); TB: B is
// stored N x K (B[n*ldb + k]).  A workgroup of 4 wavefronts owns a 128 x 32 tile
// of C, each wavefront a 32 x 32 MFMA tile; K advances 32 at a time through LDS.
// With ksplit > 1 the K range is split over gridDim.z.  Partial tiles are added
// atomically onto C (C must hold the initial value), or, when `part` is given,
// stored to part[z][M][N] for k_splitk_sum to add up in z order (the same bits
// on every run).
static constexpr int GBM = 128, GBN = 32, GBK = 32, GPAD = 8;

template <bool TA, bool TB>
__global__ void __launch_bounds__(256) k_gemm(int M, int N, int K, const __bf16* __restrict__ A, int lda,
                                              const __bf16* __restrict__ B, int ldb, float* __restrict__ C, int ldc,
                                              int accumulate, int kchunk, float* __restrict__ part)
{
    __shared__ __bf16 As[GBM][GBK + GPAD];
    __shared__ __bf16 Bs[GBN][GBK + GPAD];  // stored [n][k]
    const int tid = int(threadIdx.x);
    const int wave = tid >> 6, lane = tid & 63;
    const int m0 = int(blockIdx.x) * GBM, n0 = int(blockIdx.y) * GBN;
    const int kb = int(blockIdx.z) * kchunk;
    const int ke = min(K, kb + kchunk);
    f32x16 acc;
#pragma unroll
    for (int j = 0; j < 16; ++j)
        acc[j] = 0.f;
    for (int k0 = kb; k0 < ke; k0 += GBK) {
        // A tile: 128 x 32, 16 elements per thread
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int e = tid + 256 * r;
            int mm, kk;
            if (TA) {  // consecutive threads walk m (contiguous in memory)
                mm = e % GBM;
                kk = e / GBM;
            } else {  // consecutive threads walk k
                kk = e % GBK;
                mm = e / GBK;
            }
            const int gm = m0 + mm, gk = k0 + kk;
            __bf16 v = __bf16(0.f);
            if (gm < M && gk < ke)
                v = TA ? A[int64_t(gk) * lda + gm] : A[int64_t(gm) * lda + gk];
            As[mm][kk] = v;
        }
        // B tile: 32 x 32, 4 elements per thread
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int e = tid + 256 * r;
            int nn, kk;
            if (TB) {
                kk = e % GBK;
                nn = e / GBK;
            } else {
                nn = e % GBN;
                kk = e / GBN;
            }
            const int gn = n0 + nn, gk = k0 + kk;
            __bf16 v = __bf16(0.f);
            if (gn < N && gk < ke)
                v = TB ? B[int64_t(gn) * ldb + gk] : B[int64_t(gk) * ldb + gn];
            Bs[nn][kk] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GBK; kk += 16) {
            // lane: row/column (lane & 31), eight consecutive k from 8 * (lane >> 5)
            bf16x8 a, b;
            const int ko = kk + 8 * (lane >> 5);
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                a[i] = As[wave * 32 + (lane & 31)][ko + i];
                b[i] = Bs[lane & 31][ko + i];
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // result element j of this lane: row 8*(j/4) + 4*(lane/32) + j%4, column lane%32
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int m = m0 + wave * 32 + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
        const int n = n0 + (lane & 31);
        if (m < M && n < N) {
            float* c = C + int64_t(m) * ldc + n;
            if (gridDim.z > 1 && part)
                part[(int64_t(blockIdx.z) * M + m) * N + n] = acc[j];
            else if (gridDim.z > 1)
                atomicAdd(c, acc[j]);
            else
                *c = accumulate ? *c + acc[j] : acc[j];
        }
    }
}

// ---------------------------------------------------------------- im2col / col2im
// Activations are channels-last [batch][z][y][x][c].  Per axis a: stride s[a],
// dilation d[a], padding d[a] (output extent (in - 1) / s + 1).  Column row p =
// (b, oz, oy, ox), column (kz, ky, kx, c).
struct ConvGeom {
    int batch, C;
    int in[3];   // z, y, x
    int out[3];
    int s[3], d[3];
};

// zero the M x N block of C (rows ldc apart) before split-K partial tiles are
// added onto it: a kernel, not hipMemset2DAsync, so that a caller may capture the
// GEMM in a HIP graph
__global__ void k_zero_block(float* C, int64_t M, int64_t N, int64_t ldc)
{
    const int64_t total = M * N;
    for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += int64_t(gridDim.x) * blockDim.x)
        C[(t / N) * ldc + t % N] = 0.f;
}

// C (+)= the nz partial tiles part[z][M][N], added in z order
__global__ void k_splitk_sum(const float* __restrict__ part, int64_t nz, int64_t M, int64_t N, float* C, int64_t ldc,
                             int accumulate)
{
    const int64_t total = M * N;
    for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += int64_t(gridDim.x) * blockDim.x) {
        float* c = C + (t / N) * ldc + t % N;
        float v = accumulate ? *c : 0.f;
        for (int64_t z = 0; z < nz; ++z)
            v += part[z * total + t];
        *c = v;
    }
}

__global__ void k_im2col(ConvGeom g, const __bf16* __restrict__ X, __bf16* __restrict__ col)
{
    const int64_t K = 27 * g.C;
    const int64_t total = int64_t(g.batch) * g.out[0] * g.out[1] * g.out[2] * K;
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total; e += int64_t(gridDim.x) * blockDim.x) {
        const int64_t p = e / K;
        const int r = int(e % K);
        const int c = r % g.C, t = r / g.C;
        const int kx = t % 3, ky = (t / 3) % 3, kz = t / 9;
        int q = int(p);
        const int ox = q % g.out[2];
        q /= g.out[2];
        const int oy = q % g.out[1];
        q /= g.out[1];
        const int oz = q % g.out[0];
        const int b = q / g.out[0];
        const int iz = oz * g.s[0] + (kz - 1) * g.d[0];
        const int iy = oy * g.s[1] + (ky - 1) * g.d[1];
        const int ix = ox * g.s[2] + (kx - 1) * g.d[2];
        __bf16 v = __bf16(0.f);
        if (iz >= 0 && iz < g.in[0] && iy >= 0 && iy < g.in[1] && ix >= 0 && ix < g.in[2])
            v = X[(((int64_t(b) * g.in[0] + iz) * g.in[1] + iy) * g.in[2] + ix) * g.C + c];
        col[e] = v;
    }
}

// dX[b, iz, iy, ix, c] = sum over taps of dcol at the output position that read it
// (a gather, so it is deterministic); dX is FP32.
__global__ void k_col2im(ConvGeom g, const float* __restrict__ dcol, float* __restrict__ dX)
{
    const int64_t K = 27 * g.C;
    const int64_t total = int64_t(g.batch) * g.in[0] * g.in[1] * g.in[2] * g.C;
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total; e += int64_t(gridDim.x) * blockDim.x) {
        const int c = int(e % g.C);
        int64_t q = e / g.C;
        const int ix = int(q % g.in[2]);
        q /= g.in[2];
        const int iy = int(q % g.in[1]);
        q /= g.in[1];
        const int iz = int(q % g.in[0]);
        const int b = int(q / g.in[0]);
        float acc = 0.f;
        for (int kz = 0; kz < 3; ++kz) {
            const int nz = iz - (kz - 1) * g.d[0];
            if (nz < 0 || nz % g.s[0])
                continue;
            const int oz = nz / g.s[0];
            if (oz >= g.out[0])
                continue;
            for (int ky = 0; ky < 3; ++ky) {
                const int ny = iy - (ky - 1) * g.d[1];
                if (ny < 0 || ny % g.s[1])
                    continue;
                const int oy = ny / g.s[1];
                if (oy >= g.out[1])
                    continue;
                for (int kx = 0; kx < 3; ++kx) {
                    const int nx = ix - (kx - 1) * g.d[2];
                    if (nx < 0 || nx % g.s[2])
                        continue;
                    const int ox = nx / g.s[2];
                    if (ox >= g.out[2])
                        continue;
                    const int64_t p = ((int64_t(b) * g.out[0] + oz) * g.out[1] + oy) * g.out[2] + ox;
                    acc += dcol[p * K + ((kz * 3 + ky) * 3 + kx) * g.C + c];
                }
            }
        }
        dX[e] = acc;
    }
}

// ---------------------------------------------------------------- elementwise
// Z = Y + bias (per column); A = relu ? max(Z, 0) : Z stored as bf16 (next layer's
// input) and, when zkeep is given, Z kept in FP32 for the backward pass.
__global__ void k_bias_act(int64_t rows, int n, const float* __restrict__ Y, const float* __restrict__ bias, int relu,
                           __bf16* __restrict__ out, float* __restrict__ zkeep)
{
    const int64_t total = rows * n;
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total; e += int64_t(gridDim.x) * blockDim.x) {
        const float z = Y[e] + bias[e % n];
        if (zkeep)
            zkeep[e] = z;
        out[e] = __bf16(relu && z < 0.f ? 0.f : z);
    }
}

// dZ = relu ? dA * (Z > 0) : dA, as bf16 for the GEMMs; FP32 copy for the bias sum.
__global__ void k_act_back(int64_t total, const float* __restrict__ dA, const float* __restrict__ Z, int relu,
                           __bf16* __restrict__ dZb, float* __restrict__ dZf)
{
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total; e += int64_t(gridDim.x) * blockDim.x) {
        const float g = relu && Z[e] <= 0.f ? 0.f : dA[e];
        dZb[e] = __bf16(g);
        dZf[e] = g;
    }
}

// db[c] (+)= sum over rows of dZ[:, c]: one workgroup per column, shuffle reduction.
__global__ void __launch_bounds__(256) k_colsum(int64_t rows, int n, const float* __restrict__ dZ,
                                                float* __restrict__ db, int accumulate)
{
    const int c = int(blockIdx.x);
    float s = 0.f;
    for (int64_t r = threadIdx.x; r < rows; r += blockDim.x)
        s += dZ[r * n + c];
    for (int o = 32; o > 0; o >>= 1)
        s += __shfl_down(s, o);
    __shared__ float part[4];
    if ((threadIdx.x & 63) == 0)
        part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float t = part[0] + part[1] + part[2] + part[3];
        db[c] = accumulate ? db[c] + t : t;
    }
}

__global__ void k_f32_to_bf16(int64_t n, const float* __restrict__ x, __bf16* __restrict__ y)
{
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n; e += int64_t(gridDim.x) * blockDim.x)
        y[e] = __bf16(x[e]);
}

// Head: out[b, y, x, c] = mean over z of A[b, z, y0 + y, x0 + x, c] for the
// crop (ny x nx) at (y0, x0); backward spreads dout / nz back over z.
__global__ void k_head(int batch, int nz, int ny_in, int nx_in, int C, int y0, int x0, int ny, int nx,
                       const __bf16* __restrict__ A, float* __restrict__ out)
{
    const int64_t total = int64_t(batch) * ny * nx * C;
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total; e += int64_t(gridDim.x) * blockDim.x) {
        const int c = int(e % C);
        int64_t q = e / C;
        const int x = int(q % nx);
        q /= nx;
        const int y = int(q % ny);
        const int b = int(q / ny);
        float s = 0.f;
        for (int z = 0; z < nz; ++z)
            s += float(A[(((int64_t(b) * nz + z) * ny_in + (y0 + y)) * nx_in + (x0 + x)) * C + c]);
        out[e] = s / float(nz);
    }
}

__global__ void k_head_back(int batch, int nz, int ny_in, int nx_in, int C, int y0, int x0, int ny, int nx,
                            const float* __restrict__ dout, float* __restrict__ dA)
{
    const int64_t total = int64_t(batch) * nz * ny_in * nx_in * C;
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total; e += int64_t(gridDim.x) * blockDim.x) {
        const int c = int(e % C);
        int64_t q = e / C;
        const int x = int(q % nx_in);
        q /= nx_in;
        const int y = int(q % ny_in);
        q /= ny_in;
        const int b = int(q / nz);
        const int yy = y - y0, xx = x - x0;
        float g = 0.f;
        if (yy >= 0 && yy < ny && xx >= 0 && xx < nx)
            g = dout[((int64_t(b) * ny + yy) * nx + xx) * C + c] / float(nz);
        dA[e] = g;
    }
}

// Adam (Kingma & Ba 2015) on FP32 master weights; the bf16 copy used by the
// GEMMs is refreshed in the same pass.
__global__ void k_adam(int64_t n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2,
                       __bf16* __restrict__ wb)
{
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n; e += int64_t(gridDim.x) * blockDim.x) {
        const float gi = g[e];
        const float mi = b1 * m[e] + (1.f - b1) * gi;
        const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
        m[e] = mi;
        v[e] = vi;
        const float wi = w[e] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[e] = wi;
        if (wb)
            wb[e] = __bf16(wi);
    }
}

// ---------------------------------------------------------------- host wrappers
static unsigned grid_for(int64_t n)
{
    int64_t b = (n + 255) / 256;
    return unsigned(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}

static ConvGeom geom_of(const int64_t g[11])
{
    ConvGeom c;
    c.batch = int(g[0]);
    c.C = int(g[1]);
    for (int a = 0; a < 3; ++a) {
        c.in[a] = int(g[2 + a]);
        c.s[a] = int(g[5 + a]);
        c.d[a] = int(g[8 + a]);
        if (c.in[a] < 1 || c.s[a] < 1 || c.d[a] < 1)
            throw std::runtime_error("bad convolution geometry");
        c.out[a] = (c.in[a] - 1) / c.s[a] + 1;
    }
    if (c.batch < 1 || c.C < 1)
        throw std::runtime_error("bad convolution geometry");
    return c;
}

void sg_gemm(int ta, int tb, int64_t M, int64_t N, int64_t K, const void* A, int64_t lda, const void* B, int64_t ldb,
             float* C, int64_t ldc, int accumulate, int ksplit, void* stream, float* work, int64_t work_elems)
{
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (M < 0 || N < 0 || K < 0 || M >= (int64_t(1) << 31) || N >= (int64_t(1) << 31) || K >= (int64_t(1) << 31))
        throw std::runtime_error("GEMM size out of range");
    if (M == 0 || N == 0)
        return;
    if (ksplit < 1)
        ksplit = 1;
    int64_t kchunk = (K + ksplit - 1) / ksplit;
    kchunk = (kchunk + GBK - 1) / GBK * GBK;
    if (kchunk < GBK)
        kchunk = GBK;
    const int64_t nz = (K + kchunk - 1) / kchunk;
    dim3 grid(unsigned((M + GBM - 1) / GBM), unsigned((N + GBN - 1) / GBN), unsigned(nz < 1 ? 1 : nz));
    const __bf16* a = static_cast<const __bf16*>(A);
    const __bf16* b = static_cast<const __bf16*>(B);
    const int64_t blocks = std::min<int64_t>((M * N + 255) / 256, 4096);
    float* part = nullptr;
    if (nz > 1 && work) {  // partial tiles to the workspace, added up in order below
        if (work_elems < nz * M * N)
            throw std::runtime_error("GEMM workspace too small: " + std::to_string(nz * M * N) + " floats needed");
        part = work;
    } else if (nz > 1 && !accumulate) {  // partial tiles are added atomically onto C
        k_zero_block<<<unsigned(blocks), 256, 0, st>>>(C, M, N, ldc);
        SG_OK(hipGetLastError());
    }
    const int m = int(M), n = int(N), k = int(K), la = int(lda), lb = int(ldb), lc = int(ldc), kc = int(kchunk);
    if (!ta && !tb)
        k_gemm<false, false><<<grid, 256, 0, st>>>(m, n, k, a, la, b, lb, C, lc, accumulate, kc, part);
    else if (ta && !tb)
        k_gemm<true, false><<<grid, 256, 0, st>>>(m, n, k, a, la, b, lb, C, lc, accumulate, kc, part);
    else if (!ta && tb)
        k_gemm<false, true><<<grid, 256, 0, st>>>(m, n, k, a, la, b, lb, C, lc, accumulate, kc, part);
    else
        k_gemm<true, true><<<grid, 256, 0, st>>>(m, n, k, a, la, b, lb, C, lc, accumulate, kc, part);
    SG_OK(hipGetLastError());
    if (part) {
        k_splitk_sum<<<unsigned(blocks), 256, 0, st>>>(part, nz, M, N, C, ldc, accumulate);
        SG_OK(hipGetLastError());
    }
}

void sg_im2col(const int64_t geom[11], const void* X, void* col, void* stream)
{
    hipStream_t st = static_cast<hipStream_t>(stream);
    ConvGeom g = geom_of(geom);
    const int64_t total = int64_t(g.batch) * g.out[0] * g.out[1] * g.out[2] * 27 * g.C;
    k_im2col<<<grid_for(total), 256, 0, st>>>(g, static_cast<const __bf16*>(X), static_cast<__bf16*>(col));
    SG_OK(hipGetLastError());
}

void sg_col2im(const int64_t geom[11], const float* dcol, float* dX, void* stream)
{
    hipStream_t st = static_cast<hipStream_t>(stream);
    ConvGeom g = geom_of(geom);
    const int64_t total = int64_t(g.batch) * g.in[0] * g.in[1] * g.in[2] * g.C;
    k_col2im<<<grid_for(total), 256, 0, st>>>(g, dcol, dX);
    SG_OK(hipGetLastError());
}

void sg_bias_act(int64_t rows, int64_t n, const float* Y, const float* bias, int relu, void* out, float* zkeep,
                 void* stream)
{
    hipStream_t st = static_cast<hipStream_t>(stream);
    k_bias_act<<<grid_for(rows * n), 256, 0, st>>>(rows, int(n), Y, bias, relu, static_cast<__bf16*>(out), zkeep);
    SG_OK(hipGetLastError());
}

void sg_act_back(int64_t total, const float* dA, const float* Z, int relu, void* dZb, float* dZf, void* stream)
{
    hipStream_t st = static_cast<hipStream_t>(stream);
    k_act_back<<<grid_for(total), 256, 0, st>>>(total, dA, Z, relu, static_cast<__bf16*>(dZb), dZf);
    SG_OK(hipGetLastError());
}

void sg_colsum(int64_t rows, int64_t n, const float* dZ, float* db, int accumulate, void* stream)
{
    hipStream_t st = static_cast<hipStream_t>(stream);
    k_colsum<<<unsigned(n), 256, 0, st>>>(rows, int(n), dZ, db, accumulate);
    SG_OK(hipGetLastError());
}

void sg_to_bf16(int64_t n, const float* x, void* y, void* stream)
{
    hipStream_t st = static_cast<hipStream_t>(stream);
    k_f32_to_bf16<<<grid_for(n), 256, 0, st>>>(n, x, static_cast<__bf16*>(y));
    SG_OK(hipGetLastError());
}

void sg_head(const int64_t h[9], const void* A, float* out, void* stream)
{
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t total = h[0] * h[7] * h[8] * h[4];
    k_head<<<grid_for(total), 256, 0, st>>>(int(h[0]), int(h[1]), int(h[2]), int(h[3]), int(h[4]), int(h[5]),
                                            int(h[6]), int(h[7]), int(h[8]), static_cast<const __bf16*>(A), out);
    SG_OK(hipGetLastError());
}

void sg_head_back(const int64_t h[9], const float* dout, float* dA, void* stream)
{
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int64_t total = h[0] * h[1] * h[2] * h[3] * h[4];
    k_head_back<<<grid_for(total), 256, 0, st>>>(int(h[0]), int(h[1]), int(h[2]), int(h[3]), int(h[4]), int(h[5]),
                                                 int(h[6]), int(h[7]), int(h[8]), dout, dA);
    SG_OK(hipGetLastError());
}

void sg_adam(int64_t n, float* w, const float* g, float* m, float* v, const float hp[4], int64_t t, void* wb,
             void* stream)
{
    hipStream_t st = static_cast<hipStream_t>(stream);
    const float b1 = hp[1], b2 = hp[2];
    const float c1 = 1.f - powf(b1, float(t)), c2 = 1.f - powf(b2, float(t));
    k_adam<<<grid_for(n), 256, 0, st>>>(n, w, g, m, v, hp[0], b1, b2, hp[3], c1, c2, static_cast<__bf16*>(wb));
    SG_OK(hipGetLastError());
}

}  // namespace ffip_amd
