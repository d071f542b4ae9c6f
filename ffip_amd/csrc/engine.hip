// Fused FDTD half-step kernels for gfx950 and the device engine around them.
//
// One launch per half step (k_half<IS_E>) does, for every real point of the three
// components it updates, what the reference does in three passes:
//   curl + CPML psi      (PML_Stepping::step / Curl_Stepping::step, src/fields.cpp:89-148)
//   ADE poles + update   (Medium_Stepping::step, src/structure.cpp:56-85)
//   D/B DFT accumulation (DFT_Unit::update, src/dft.cpp:29-62; these need the
//                         polarization sums formed in the update)
// Point sources (Gaussian_Dipoles_Stepping::step, src/fields.cpp:17-33) touch a
// few points only and run in a small follow-up kernel that also patches the DFT
// sums of the points it changes; the ghost fills of Period/PEC/PMC boundaries
// (Fields::sync_boundary, src/fields.cpp:327-482) follow.  E/H DFT monitors
// gather each step's samples into a ring, added into their FP64 sums in
// batches (k_dft_gather, k_dft_ring).
//
// The update is written in "E form": E^{n+1} = E^n + (dt*curl - (sum P^{n+1} -
// sum P^n)) / eps_inf.  Algebraically this is the reference's D form
// (E = (D - sum P)/eps_inf with D += dt*curl, src/structure.cpp:83) since
// D^n = eps_inf E^n + sum P^n; it needs no D array.  The polarization difference is
// formed from the stored (rounded) P values so that D stays consistent when
// fields are stored in FP32.  All arithmetic is FP64; T is the storage type.
//
// Tiling: a BX x BY (32 x 8) workgroup owns an x-y tile and marches a z segment,
// carrying the two z-neighbour components in registers; x/y neighbours come from
// L1/L2.  Each batch of planes issues all its loads before computing (CPML
// auxiliaries and polarizations first, then the loads every thread makes), so a
// batch costs one memory round trip; material words are fetched a batch ahead.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <limits>
#include <map>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

#include "engine.hpp"

namespace ffip_amd {

#define HIP_OK(expr)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr); \
    } while (0)

// workgroup tile: BX consecutive x points (one wavefront row when BX = 64) by BY rows
#ifndef FFIP_BX
#define FFIP_BX 32
#endif
#ifndef FFIP_BY
#define FFIP_BY 8
#endif
static constexpr int BX = FFIP_BX;
static constexpr int BY = FFIP_BY;
static constexpr int MAXU = 16;
static constexpr int MAXS = 8;  // polarization slots per point
static constexpr int PH_CHUNK = 1024;

struct DftDev {
    int comp;  // 0..2 within the half
    int box[3][2];
    int dims[3];
    int nf, mf0, is_db;
    int fused;  // accumulated inside the half-step kernel (else through the sample ring)
    int64_t npts;
    double* re;
    double* im;
    // E/H units: the stored samples of the last steps, [DFT_BATCH][npts] in the
    // field's storage type, added into re/im by k_dft_ring every DFT_BATCH steps
    // and before any read
    void* ring;
};

// steps whose E/H monitor samples are gathered before they are added into the
// FP64 sums (one read-modify-write of the sums per batch instead of per step)
constexpr int DFT_BATCH = 16;

struct PmlDev {
    const int* smap;
    const double* b;
    const double* c;
    const double* ik;
    void* psi1;  // term-1 auxiliary of component (a+2)%3
    void* psi2;  // term-2 auxiliary of component (a+1)%3
    int W;
};

// One medium of the table.  The update reads the first 16 bytes (1/eps, pole
// count, first pole) in one LDS read; eps is read only for D monitors.
struct alignas(16) MatEntry {
    double inv;
    int np, p0;
    double eps, pad;
};
static_assert(sizeof(MatEntry) == 32 && alignof(MatEntry) == 16 && offsetof(MatEntry, p0) == 12,
              "the LDS table layout assumes 32-byte entries with inv, np, p0 in the first 16 bytes");

template <typename T>
struct HalfArgs {
    const T* u[3];  // updated components, values before the half step
    T* w[3];        // where their new values go (= u, or the other buffer of a fused step)
    const T* s[3];
    int pitch, plane;
    int tail0;  // first packed x index (Layout::tail0)
    int n[3];
    int lo[3][3], hi[3][3];
    int dlo[3], dhi[3];
    int gx, gy, gz, nblocks, lz;
    // tiles of this launch: tile t = tx + gx * (ty + gy * tz), in the order of the list
    const uint32_t* tiles;
    int ntiles;
    double inv_dx, dt;
    // materials
    double u_inv, u_eps;
    const uint32_t* ids;  // per cell: id of component 0 | 1 << 8 | 2 << 16
    int mbox[3][2];       // cells outside this box hold bg_word (not read from ids)
    uint32_t bg_word;
    const MatEntry* mats;
    int nmats;
    const double* coef;  // [npoles][4] = (c1, c2, sigma*c3, 0) of the table's poles
    int npoles;
    int d_box[3][3][2];
    int d_np[3];
    const double* d_inv[3];
    const double* d_eps[3];
    const double* d_amp[3];
    const double* d_c[3];  // [np][3] = (c1, c2, c3)
    // Polarization state.  pcur is read, pprev read and then overwritten with
    // P^{n+1}; the buffers swap roles every step.  Two layouts:
    //  full grid: pole s of component c at cell o is element (s*3 + c)*pslice + o;
    //  compact:   only points with poles hold entries; component c's start at
    //             pcomp[c], the row segment (k, j, x tile) holds its points' poles
    //             from segbase[c][((k - dlo_z) * seg_ny + (j - dlo_y)) * gx + tx],
    //             pole-major: pole 0 of the segment's points (in x order) that
    //             have it, then pole 1, and so on.  Segments are stored tile by
    //             tile (plan_polarization), which the kernels do not depend on.
    T* pcur;
    T* pprev;
    int64_t pslice;
    int64_t pcomp[3];
    const uint32_t* segbase[3];
    int seg_ny;
    int origin;  // cell offset of lattice index (0, 0, 0) in every per-cell array
    PmlDev pml[3];
    int nunits;
    DftDev units[MAXU];
    // the units each component accumulates in this kernel, and the union of their boxes
    int nfu[3];
    int fu[3][MAXU];
    int fbox[3][3][2];
    const double* ph;  // this half step's phasor row [mf][2] = (cos wt, -sin wt)
};

// Gaussian1 pulse (src/utility.cpp:260-264)
__device__ __forceinline__ double gaussian1(double t, double width)
{
    double arg = t / width;
    return arg * exp(-arg * arg / 2);
}

// Column constants of one thread: position, CPML slab positions and
// coefficients along x and y, and which of the three components live here.
struct ColCtx {
    int i, j, base;
    int dx, dy;  // element offsets of the x and y neighbours the half reads (E: +1, H: -1)
    int sx, sy;  // slab positions (-1 outside the CPML)
    double bx, cx, ikx, by, cy, iky;
    bool in[3];
};

// Plane constants (uniform over the workgroup): CPML along z.
struct ZCtx {
    int k, o, sz;
    double bz, cz, ikz;
};

// Loads and stores at an unsigned 32-bit byte offset from a workgroup-uniform
// base: the base stays in scalar registers and loads that share an offset share
// one vector register (global_load v, v_off, s[base]).
template <typename T>
__device__ __forceinline__ T ldb(const T* base, uint32_t off)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
}
template <typename T>
__device__ __forceinline__ void stb(T* base, uint32_t off, T v)
{
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + off) = v;
}

// Element of lattice index (i, j, k) relative to index (0, 0, 0) in every per-cell
// array (Layout::index less its origin): rows of `pitch`, planes of `plane`, and
// from x index tail0 on the packed last line shared by TAIL_GROUP rows.
__device__ __forceinline__ int64_t lat_off(int64_t pitch, int64_t plane, int tail0, int i, int j, int k)
{
    if (i < tail0)
        return i + pitch * j + plane * k;
    const int m = (j + 1) % TAIL_GROUP;
    return tail0 + 2 * m + (i - tail0) + pitch * (j - m) + plane * k;
}

// Everything one thread reads for one plane.  All of it is requested before any
// of it is used, so a plane costs one memory round trip whatever it holds (CPML
// auxiliaries, several poles, dense-region coefficients).
template <typename T, int NPM, bool CP>
struct PlaneIn {
    T n0, n1, v2, v2a, v2b, v1a, v0b;  // stencil operands (see k_half)
    T e[3];                            // the three updated components
    T psi[3][2];                       // CPML auxiliaries of the two curl terms of each component
    T pc[3][NPM > 0 ? NPM : 1];        // P^n of each pole
    T pp[3][NPM > 0 ? NPM : 1];        // P^{n-1}
    uint32_t poff[3][CP && NPM > 0 ? NPM : 1];  // compact P storage: entry of each pole
    double dinv[3];                    // 1/eps_inf in the dense region
};

// Rank of this lane among the lanes of its x-row segment whose predicate holds,
// and their count (a row segment is BX consecutive lanes).
__device__ __forceinline__ void row_rank(bool pred, uint32_t& rank, uint32_t& count)
{
    const uint64_t m = __ballot(pred);
    const uint32_t lane = __lane_id();
    static_assert(BX == 32 || BX == 64, "row segments are half or whole wavefronts");
    if constexpr (BX == 32) {
        const uint32_t row = (lane & 32u) ? uint32_t(m >> 32) : uint32_t(m);
        rank = __popc(row & ((1u << (lane & 31u)) - 1u));
        count = __popc(row);
    } else {
        rank = __popcll(m & ((uint64_t(1) << lane) - 1));
        count = __popcll(m);
    }
}

// CPML auxiliary offsets (slab-only arrays, src/fields.cpp:89-109): the slab along
// axis a holds the points of its W slab positions, the slab position replacing
// that axis.  The y and z slabs keep x-rows of the field arrays' padded pitch, so
// a tile's row starts on a 128-byte line as it does in the field arrays.
template <typename T>
__device__ __forceinline__ int psi_off_x(const HalfArgs<T>& A, const ColCtx& X, int k)
{
    return X.sx + A.pml[0].W * (X.j + A.n[1] * k);
}
template <typename T>
__device__ __forceinline__ int psi_off_y(const HalfArgs<T>& A, const ColCtx& X, int k)
{
    return X.i + A.pitch * (X.sy + A.pml[1].W * k);
}
template <typename T>
__device__ __forceinline__ int psi_off_z(const HalfArgs<T>& A, const ColCtx& X, int sz)
{
    return X.i + A.pitch * (X.j + A.n[1] * sz);
}

template <typename T>
__device__ __forceinline__ int64_t dense_off(const HalfArgs<T>& A, int C, int i, int j, int k)
{
    const int b00 = A.d_box[C][0][0], b10 = A.d_box[C][1][0], b20 = A.d_box[C][2][0];
    const int64_t d0 = A.d_box[C][0][1] - b00 + 1, d1n = A.d_box[C][1][1] - b10 + 1;
    return (int64_t(k - b20) * d1n + (j - b10)) * d0 + (i - b00);
}

// Stencil operands and updated components of one plane (loads every thread
// issues; o is clamped into the segment so they are always in bounds).
// dx, dy: the column's x and y neighbour offsets (ColCtx)
template <typename T, bool IS_E, int NPM, bool CP>
__device__ __forceinline__ void load_fields(const HalfArgs<T>& A, int o, int dx, int dy, PlaneIn<T, NPM, CP>& L)
{
    constexpr uint32_t E = sizeof(T);
    const uint32_t b = uint32_t(o) * E, bz = uint32_t(A.plane) * E;
    const uint32_t bx = b + uint32_t(dx) * E, by = b + uint32_t(dy) * E;
    if (IS_E) {
        L.n0 = ldb(A.s[0], b + bz);
        L.n1 = ldb(A.s[1], b + bz);
    } else {
        L.n0 = ldb(A.s[0], b);
        L.n1 = ldb(A.s[1], b);
    }
    L.v2a = ldb(A.s[2], bx);
    L.v2b = ldb(A.s[2], by);
    L.v1a = ldb(A.s[1], bx);
    L.v0b = ldb(A.s[0], by);
    L.v2 = ldb(A.s[2], b);
#pragma unroll
    for (int C = 0; C < 3; ++C)
        L.e[C] = ldb<T>(A.u[C], b);
}

// Loads that only some points need: CPML auxiliaries, polarizations, dense
// 1/eps_inf.  A batch issues all of these before its unconditional loads, so
// that waiting for the (newer) field data also covers them.  With ALL_LANES
// every lane issues every polarization load of its kernel's shape (NPM poles
// per component) and a point without the pole reads element 0 of the array and
// ignores it: under a per-point branch the compiler converts the loaded value
// to FP64 inside the branch, right after its load, and so waits for each load
// as soon as it is issued, which the fused kernel (loads and use in one plane)
// cannot hide.  The half-step kernels, whose batches wait for everything at
// once, measured faster with the branch (512^3 voxel mix, electric half step
// 2956 vs 3066 us).  Kernels with more pole slots than the half has (NPM =
// MAXS) always keep the branch.
// `id` is the packed material word of the cell and (compact layout) sb[c] the
// first P entry of its row segment, both fetched one batch earlier.  With the
// compact layout it must be called by every live lane of the wave (offsets
// come from ballots over the row).
template <typename T, bool MAT, int NPM, bool CP, bool DENSE, bool PML, bool ALL_LANES>
__device__ __forceinline__ void load_sparse(const HalfArgs<T>& A, const MatEntry* smat, const ColCtx& X,
                                            const ZCtx& Z, uint32_t id, const uint32_t sb[3], const bool act[3],
                                            PlaneIn<T, NPM, CP>& L)
{
    constexpr uint32_t E = sizeof(T);
    if constexpr (PML) {
        // component 0 differentiates along (y, z), 1 along (z, x), 2 along (x, y).
        // These stay per-point branches: on the CPML tiles (few points in a slab,
        // register-bound kernels) they measured faster than issuing all six.
        if (act[0]) {
            if (X.sy >= 0)
                L.psi[0][0] = ldb(static_cast<const T*>(A.pml[1].psi1), uint32_t(psi_off_y(A, X, Z.k)) * E);
            if (Z.sz >= 0)
                L.psi[0][1] = ldb(static_cast<const T*>(A.pml[2].psi2), uint32_t(psi_off_z(A, X, Z.sz)) * E);
        }
        if (act[1]) {
            if (Z.sz >= 0)
                L.psi[1][0] = ldb(static_cast<const T*>(A.pml[2].psi1), uint32_t(psi_off_z(A, X, Z.sz)) * E);
            if (X.sx >= 0)
                L.psi[1][1] = ldb(static_cast<const T*>(A.pml[0].psi2), uint32_t(psi_off_x(A, X, Z.k)) * E);
        }
        if (act[2]) {
            if (X.sx >= 0)
                L.psi[2][0] = ldb(static_cast<const T*>(A.pml[0].psi1), uint32_t(psi_off_x(A, X, Z.k)) * E);
            if (X.sy >= 0)
                L.psi[2][1] = ldb(static_cast<const T*>(A.pml[1].psi2), uint32_t(psi_off_y(A, X, Z.k)) * E);
        }
    }
    if (MAT) {
#pragma unroll
        for (int C = 0; C < 3; ++C) {
            const int mid = int((id >> (8 * C)) & 255u);
            int np = 0;
            bool dense_pt = false;
            if (act[C]) {
                if (!DENSE || mid != 255)
                    np = smat[mid].np;
                else {
                    np = A.d_np[C];
                    dense_pt = true;
                }
            }
            if (DENSE && (ALL_LANES || dense_pt))
                L.dinv[C] = A.d_inv[C][dense_pt ? dense_off(A, C, X.i, X.j, Z.k) : 0];
            uint32_t e = CP ? sb[C] : 0u;
#pragma unroll
            for (int s = 0; s < NPM; ++s) {
                const bool has = s < np;
                if constexpr (CP) {
                    // pole s of the row's points in x order: this lane's rank among
                    // those having it, after the entries of the lower poles
                    uint32_t rank, count;
                    row_rank(has, rank, count);
                    L.poff[C][s] = e + rank;
                    e += count;
                    if (ALL_LANES || has) {
                        const uint32_t bo = has ? L.poff[C][s] * E : 0u;
                        L.pc[C][s] = ldb<T>(A.pcur + A.pcomp[C], bo);
                        L.pp[C][s] = ldb<T>(A.pprev + A.pcomp[C], bo);
                    }
                } else if constexpr (ALL_LANES && NPM < MAXS) {
                    const uint32_t bo = has ? uint32_t(Z.o) * E : 0u;
                    L.pc[C][s] = ldb<T>(A.pcur + (s * 3 + C) * A.pslice, bo);
                    L.pp[C][s] = ldb<T>(A.pprev + (s * 3 + C) * A.pslice, bo);
                } else if (has) {
                    L.pc[C][s] = ldb<T>(A.pcur + (s * 3 + C) * A.pslice, uint32_t(Z.o) * E);
                    L.pp[C][s] = ldb<T>(A.pprev + (s * 3 + C) * A.pslice, uint32_t(Z.o) * E);
                }
            }
        }
    }
}

// element (of the polarization buffers) holding pole s of component C at this point
template <typename T, int NPM, bool CP>
__device__ __forceinline__ T* pole_ptr(const HalfArgs<T>& A, const PlaneIn<T, NPM, CP>& L, const ZCtx& Z, int C, int s)
{
    if constexpr (CP)
        return reinterpret_cast<T*>(reinterpret_cast<char*>(A.pprev + A.pcomp[C]) +
                                    L.poff[C][s] * uint32_t(sizeof(T)));
    else
        return reinterpret_cast<T*>(reinterpret_cast<char*>(A.pprev + (s * 3 + C) * A.pslice) +
                                    uint32_t(Z.o) * uint32_t(sizeof(T)));
}

// one CPML term (src/fields.cpp:89-109): psi <- b psi + c d; returns psi + d / kappa
template <typename T>
__device__ __forceinline__ double cpml_apply(void* arr, int q, T psi, double b, double c, double ik, double d)
{
    const double ps = b * double(psi) + c * d;
    stb(static_cast<T*>(arr), uint32_t(q) * sizeof(T), T(ps));
    return ps + d * ik;
}

// Update component C of one point from its preloaded plane data: CPML, ADE
// poles (src/structure.cpp:56-85, in E form), DFT accumulation, store.
template <typename T, bool MAT, int NPM, bool CP, bool DENSE, int C, bool DFT = true>
__device__ __forceinline__ void update_comp(const HalfArgs<T>& A, const MatEntry* smat, const double* spc,
                                            const ColCtx& X, const ZCtx& Z, double d1, double d2, uint32_t id,
                                            const PlaneIn<T, NPM, CP>& L)
{
    // term 1 differentiates along a1 = (C+1)%3, term 2 along a2 = (C+2)%3
    double t1 = d1, t2 = d2;
    if (C == 0) {
        if (X.sy >= 0)
            t1 = cpml_apply<T>(A.pml[1].psi1, psi_off_y(A, X, Z.k), L.psi[0][0], X.by, X.cy, X.iky, d1);
        if (Z.sz >= 0)
            t2 = cpml_apply<T>(A.pml[2].psi2, psi_off_z(A, X, Z.sz), L.psi[0][1], Z.bz, Z.cz, Z.ikz, d2);
    } else if (C == 1) {
        if (Z.sz >= 0)
            t1 = cpml_apply<T>(A.pml[2].psi1, psi_off_z(A, X, Z.sz), L.psi[1][0], Z.bz, Z.cz, Z.ikz, d1);
        if (X.sx >= 0)
            t2 = cpml_apply<T>(A.pml[0].psi2, psi_off_x(A, X, Z.k), L.psi[1][1], X.bx, X.cx, X.ikx, d2);
    } else {
        if (X.sx >= 0)
            t1 = cpml_apply<T>(A.pml[0].psi1, psi_off_x(A, X, Z.k), L.psi[2][0], X.bx, X.cx, X.ikx, d1);
        if (X.sy >= 0)
            t2 = cpml_apply<T>(A.pml[1].psi2, psi_off_y(A, X, Z.k), L.psi[2][1], X.by, X.cy, X.iky, d2);
    }
    const double acc = t1 - t2;  // increment of D/dt (B/dt)
    const double e_old = double(L.e[C]);
    double e_new, sumP = 0;
    int mid = 0;
    if (!MAT) {
        e_new = e_old + A.dt * acc * A.u_inv;
    } else {
        mid = int((id >> (8 * C)) & 255u);
        double inv, dP = 0;
        if (!DENSE || mid != 255) {
            const MatEntry m = smat[mid];
            inv = m.inv;
#pragma unroll
            for (int s = 0; s < NPM; ++s) {
                if (s < m.np) {
                    const double* cf = spc + 4 * (m.p0 + s);
                    const double2 c01 = *reinterpret_cast<const double2*>(cf);
                    const double pc = double(L.pc[C][s]);
                    const double pn = c01.x * pc + c01.y * double(L.pp[C][s]) + cf[2] * e_old;
                    const T pt = T(pn);
                    *pole_ptr<T, NPM, CP>(A, L, Z, C, s) = pt;
                    dP += double(pt) - pc;
                    sumP += double(pt);
                }
            }
        } else {
            inv = L.dinv[C];
            const int np = A.d_np[C];
            if (NPM > 0 && np > 0) {
                const int64_t q = dense_off(A, C, X.i, X.j, Z.k);
#pragma unroll
                for (int s = 0; s < NPM; ++s) {
                    if (s < np) {
                        const double* cc = A.d_c[C] + 3 * s;
                        const double amp = A.d_amp[C][q * np + s];
                        const double pc = double(L.pc[C][s]);
                        const double pn = cc[0] * pc + cc[1] * double(L.pp[C][s]) + amp * cc[2] * e_old;
                        const T pt = T(pn);
                        *pole_ptr<T, NPM, CP>(A, L, Z, C, s) = pt;
                        dP += double(pt) - pc;
                        sumP += double(pt);
                    }
                }
            }
        }
        e_new = e_old + (A.dt * acc - dP) * inv;
    }
    const T stored = T(e_new);
    // DFT monitors (k is uniform across the workgroup)
    const bool in_dft = DFT && A.nfu[C] > 0 && Z.k >= A.fbox[C][2][0] && Z.k <= A.fbox[C][2][1] &&
                        X.i >= A.fbox[C][0][0] && X.i <= A.fbox[C][0][1] && X.j >= A.fbox[C][1][0] &&
                        X.j <= A.fbox[C][1][1];
    for (int m = 0; in_dft && m < A.nfu[C]; ++m) {
        const DftDev& U = A.units[A.fu[C][m]];
        if (Z.k < U.box[2][0] || Z.k > U.box[2][1])
            continue;
        if (X.i < U.box[0][0] || X.i > U.box[0][1] || X.j < U.box[1][0] || X.j > U.box[1][1])
            continue;
        double v = double(stored);
        if (U.is_db) {
            const double eps = !MAT ? A.u_eps
                                    : ((!DENSE || mid != 255) ? smat[mid].eps
                                                              : A.d_eps[C][dense_off(A, C, X.i, X.j, Z.k)]);
            v = eps * double(stored) + sumP;
        }
        const int64_t off = (X.i - U.box[0][0]) +
                            int64_t(U.dims[0]) * ((X.j - U.box[1][0]) + int64_t(U.dims[1]) * (Z.k - U.box[2][0]));
        const double* ph = A.ph + 2 * U.mf0;
        for (int f = 0; f < U.nf; ++f) {
            const int64_t o = int64_t(f) * U.npts + off;
            U.im[o] += ph[2 * f + 1] * v;
            U.re[o] += ph[2 * f] * v;
        }
    }
    stb<T>(A.w[C], uint32_t(Z.o) * sizeof(T), stored);
}

// One half step over a (BX x BY) column tile and a z segment of A.lz planes.
// Loads for U consecutive planes are issued before any of them is computed so
// that each thread keeps several planes of requests in flight; the two
// components differentiated along z are carried between planes in registers.
// PML = false: the launch's tiles lie clear of every CPML slab (no auxiliary
// code or registers); DENSE = false: no point of this half has a per-point medium.
template <typename T, bool IS_E, bool MAT, int NPM, bool CP, int U, bool PML, bool DENSE, bool DFT>
__global__ void __launch_bounds__(BX * BY) k_half(const HalfArgs<T> A)
{
    // material table and pole coefficients, staged in LDS (sized at launch:
    // 32 bytes per table entry and per pole, read with 16-byte loads)
    extern __shared__ __align__(16) double lds_tab[];
    MatEntry* smat = reinterpret_cast<MatEntry*>(lds_tab);
    double* spc = lds_tab + 4 * A.nmats;
    if (MAT) {
        const int tid = int(threadIdx.x + BX * threadIdx.y);
        for (int q = tid; q < A.nmats; q += BX * BY)
            smat[q] = A.mats[q];
        if (NPM > 0)
            for (int q = tid; q < 4 * A.npoles; q += BX * BY)
                spc[q] = A.coef[q];
        __syncthreads();
    }
    // XCD-aware order: consecutive dispatches land on the 8 XCDs in turn, so the
    // tiles one XCD runs are made contiguous in space (shared halos stay in its L2)
    const int b = int(blockIdx.x);
    const int slot = (b & 7) * (A.nblocks >> 3) + (b >> 3);
    if (slot >= A.ntiles)
        return;
    const int t = int(A.tiles[slot]);
    const int tx = t % A.gx;
    const int tr = t / A.gx;
    const int ty = tr % A.gy;
    const int tz = tr / A.gy;
    const int i = A.dlo[0] + tx * BX + int(threadIdx.x);
    const int j = A.dlo[1] + ty * BY + int(threadIdx.y);
    const int kb = A.dlo[2] + tz * A.lz;
    const int ke = min(kb + A.lz - 1, A.dhi[2]);
    if (i > A.dhi[0] || j > A.dhi[1])
        return;
    const int sz = A.plane;
    const double inv_dx = A.inv_dx;
    ColCtx X;
    X.i = i;
    X.j = j;
    X.base = A.origin + int(lat_off(A.pitch, A.plane, A.tail0, i, j, 0));
    {
        constexpr int u = IS_E ? 1 : -1;
        X.dx = A.origin + int(lat_off(A.pitch, A.plane, A.tail0, i + u, j, 0)) - X.base;
        X.dy = A.origin + int(lat_off(A.pitch, A.plane, A.tail0, i, j + u, 0)) - X.base;
    }
#pragma unroll
    for (int C = 0; C < 3; ++C)
        X.in[C] = i >= A.lo[C][0] && i <= A.hi[C][0] && j >= A.lo[C][1] && j <= A.hi[C][1];
    X.sx = PML && A.pml[0].W ? A.pml[0].smap[i] : -1;
    X.sy = PML && A.pml[1].W ? A.pml[1].smap[j] : -1;
    X.bx = X.cx = X.by = X.cy = 0;
    X.ikx = X.iky = 1;
    if (X.sx >= 0) {
        X.bx = A.pml[0].b[i];
        X.cx = A.pml[0].c[i];
        X.ikx = A.pml[0].ik[i];
    }
    if (X.sy >= 0) {
        X.by = A.pml[1].b[j];
        X.cy = A.pml[1].c[j];
        X.iky = A.pml[1].ik[j];
    }
    T c0, c1;  // E: s0, s1 at plane k; H: at plane k-1
    {
        const int o = X.base + sz * kb;
        const uint32_t b = uint32_t(IS_E ? o : o - sz) * sizeof(T);
        c0 = ldb(A.s[0], b);
        c1 = ldb(A.s[1], b);
    }
    // material words are fetched one batch ahead of the planes they describe
    // and so are the first compact polarization entries of the row segments
    constexpr bool POL = MAT && NPM > 0 && CP;
    const int seg0 = (j - A.dlo[1]) * A.gx + tx;
    const int segplane = A.seg_ny * A.gx;
    // columns and planes outside the box of non-default material use the
    // default medium's word without loading it
    const bool mcol = i >= A.mbox[0][0] && i <= A.mbox[0][1] && j >= A.mbox[1][0] && j <= A.mbox[1][1];
    // FP64: every lane loads (the id array spans the whole lattice); cells
    // outside the box read element 0, one cached word, and take the default word
    // held in a register.  FP32: the choice between the argument word and the
    // id array compiles to one flat load through the chosen pointer.  Measured
    // on MI355X, same box, E half step: FP64 256^3 dispersive 460 us (global
    // load) vs 486 us (flat), 200x200x100 design 94.7 vs 99.0 us; FP32 the other
    // way round (256^3 241 vs 255 us, 512^3 voxel mix 2956 vs 3002 us).
    const uint32_t bg_word = __builtin_amdgcn_readfirstlane(A.bg_word);
    auto mat_word = [&](int kk) -> uint32_t {
        if (!MAT)
            return 0u;
        const bool in = mcol && kk >= A.mbox[2][0] && kk <= A.mbox[2][1];
        if constexpr (sizeof(T) == 4) {
            if (!in)
                return A.bg_word;
            return ldb(A.ids, uint32_t(X.base + sz * kk) * 4u);
        } else {
            const uint32_t w = ldb(A.ids, in ? uint32_t(X.base + sz * kk) * 4u : 0u);
            return in ? w : bg_word;
        }
    };
    uint32_t idn[U], sbn[U][3];
#pragma unroll
    for (int q = 0; q < U; ++q) {
        const int kk = min(kb + q, ke);
        idn[q] = mat_word(kk);
#pragma unroll
        for (int C = 0; C < 3; ++C)
            sbn[q][C] = POL ? A.segbase[C][(kk - A.dlo[2]) * segplane + seg0] : 0u;
    }
    for (int kc = kb; kc <= ke; kc += U) {
        uint32_t idc[U], sbc[U][3];
#pragma unroll
        for (int q = 0; q < U; ++q) {
            idc[q] = idn[q];
#pragma unroll
            for (int C = 0; C < 3; ++C)
                sbc[q][C] = sbn[q][C];
        }
        PlaneIn<T, NPM, CP> L[U];
        ZCtx Z[U];
        bool act[U][3];
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const int k = kc + q;
            Z[q].k = k;
            Z[q].o = X.base + sz * k;
            Z[q].sz = -1;
            Z[q].bz = Z[q].cz = 0;
            Z[q].ikz = 1;
            if (PML && k <= ke && A.pml[2].W) {
                Z[q].sz = A.pml[2].smap[k];
                if (Z[q].sz >= 0) {
                    Z[q].bz = A.pml[2].b[k];
                    Z[q].cz = A.pml[2].c[k];
                    Z[q].ikz = A.pml[2].ik[k];
                }
            }
#pragma unroll
            for (int C = 0; C < 3; ++C)
                act[q][C] = k <= ke && X.in[C] && k >= A.lo[C][2] && k <= A.hi[C][2];
            if (k <= ke)
                load_sparse<T, MAT, NPM, CP, DENSE, PML, false>(A, smat, X, Z[q], idc[q], sbc[q], act[q], L[q]);
        }
        // then the loads every thread makes: next batch's material words and
        // segment bases, fields.  With full-grid polarizations the next batch's
        // words go after the fields, so that waiting for the fields does not wait
        // for them (256^3 spheres: electric half step 225 -> 215 us); with the
        // compact layout that order measured no better (512^3 2625 -> 2629 us).
        auto prefetch = [&]() {
#pragma unroll
            for (int q = 0; q < U; ++q) {
                const int kk = min(kc + U + q, ke);
                idn[q] = mat_word(kk);
#pragma unroll
                for (int C = 0; C < 3; ++C)
                    sbn[q][C] = POL ? A.segbase[C][(kk - A.dlo[2]) * segplane + seg0] : 0u;
            }
        };
        if constexpr (CP)
            prefetch();
#pragma unroll
        for (int q = 0; q < U; ++q)
            load_fields<T, IS_E, NPM, CP>(A, X.base + sz * min(kc + q, ke), X.dx, X.dy, L[q]);
        if constexpr (!CP)
            prefetch();
#pragma unroll
        for (int q = 0; q < U; ++q) {
            if (kc + q > ke)
                break;
            const PlaneIn<T, NPM, CP>& P = L[q];
            // E: s0(k)=c0, s0(k+1)=n0; H: s0(k-1)=c0, s0(k)=n0 (s1 likewise)
            const double a0 = double(c0), a1 = double(c1), g0 = double(P.n0), g1 = double(P.n1);
            const double w2 = double(P.v2);
            if (act[q][0]) {
                const double d1 = (double(P.v2b) - w2) * inv_dx;
                const double d2 = (IS_E ? (g1 - a1) : (a1 - g1)) * inv_dx;
                update_comp<T, MAT, NPM, CP, DENSE, 0, DFT>(A, smat, spc, X, Z[q], d1, d2, idc[q], P);
            }
            if (act[q][1]) {
                const double d1 = (IS_E ? (g0 - a0) : (a0 - g0)) * inv_dx;
                const double d2 = (double(P.v2a) - w2) * inv_dx;
                update_comp<T, MAT, NPM, CP, DENSE, 1, DFT>(A, smat, spc, X, Z[q], d1, d2, idc[q], P);
            }
            if (act[q][2]) {
                const double d1 = (double(P.v1a) - (IS_E ? a1 : g1)) * inv_dx;
                const double d2 = (double(P.v0b) - (IS_E ? a0 : g0)) * inv_dx;
                update_comp<T, MAT, NPM, CP, DENSE, 2, DFT>(A, smat, spc, X, Z[q], d1, d2, idc[q], P);
            }
            c0 = P.n0;
            c1 = P.n1;
        }
    }
}

// ---------------------------------------------------------------- fused full step
// One launch advances the interior of the grid by a whole time step (step_m then
// step_e, src/simulation.cpp:748-766) when the magnetic half is a uniform medium
// with no magnetic sources, B monitors or ghost faces.  A workgroup owns a BX x BY
// column tile and a z segment [kb, ke] clear of every CPML slab.  Marching up in
// z, it forms H^{n+1/2} on plane k+1 (from E^n on planes k and k+1) just before
// E^{n+1} on plane k (from H^{n+1/2} on planes k and k+1), so each field element
// is read once and written once per step: 12 accesses per cell instead of the 18
// of two half-step kernels.  E^{n+1} on the tile's last column and row needs
// H^{n+1/2} one point beyond the tile; the workgroup forms those values itself
// from E^n, as the neighbouring tile (or the magnetic half-step kernel, on the
// CPML shell) does when it forms and stores them, and the segment's first H plane
// is formed the same way by the segment below.  Fields are therefore
// double-buffered: E^n and H^{n-1/2} are read from one set, E^{n+1} and
// H^{n+1/2} written to the other, so no workgroup reads what another writes in
// the same launch.  The fused kernel covers a box of whole tiles clear of every
// CPML slab; the half-step kernels cover the rest (the CPML shell), the magnetic
// one before and the electric one after it, so rim values that fall in the shell
// are read, already new, from the H^{n+1/2} buffer.  The arithmetic and its order
// are k_half's.
template <typename T>
struct StepArgs {
    HalfArgs<T> e;     // electric half: u = E^n, w = E^{n+1}
    const T* hin[3];   // H^{n-1/2}
    T* hout[3];        // H^{n+1/2}
    int hlo[3][3], hhi[3][3];  // real points of Hx, Hy, Hz
    double h_inv;      // 1/mu of the uniform magnetic medium
    // the box of lattice indices the fused kernel covers (whole tiles, clear of
    // every CPML slab): first and last index per axis, tiles per axis (z tiles of
    // e.lz planes); the half-step kernels cover the rest of the grid
    int fo[3], fe[3], fn[3];
};

// H^{n+1/2} from its curl differences d1, d2 (k_half's magnetic update in a
// uniform medium outside the CPML)
template <typename T>
__device__ __forceinline__ T h_value(const StepArgs<T>& S, double d1, double d2, T hold)
{
    const double acc = d1 - d2;
    return T(double(hold) + S.e.dt * acc * S.h_inv);
}

// k_fstep stages the field planes of its tile with direct-to-LDS loads
// (global_load_lds_dword) issued by a fifth wavefront of the workgroup, D planes
// ahead of the plane being computed.  The four computing wavefronts hold no
// field loads in registers, so the bytes a CU keeps in flight are set by the LDS
// ring (D planes of E^n with a one-point rim and of H^{n-1/2} with the far rim,
// per workgroup) rather than by the register budget.  The loading wavefront
// issues nothing but these loads, so waiting for a plane is a static count of
// outstanding loads (vmcnt), followed by the workgroup barrier that also orders
// the computing wavefronts' LDS traffic.  It addresses rows as origin + i +
// pitch * j: its box and rims stay below the packed last line (Layout::tail0).
//
// Ring slot layout, in dwords: for c = 0..2 the E^n region [BY+2][BX+2] (x from
// i0-1, y from j0-1), then for c = 0..2 the H region [BY+1][BX+1] (x from i0, y
// from j0); each region is padded to whole loads of 64 dwords, and lane l of the
// loading wavefront's q-th load of a region writes its dword 64q + l.  An H
// region is loaded with H^{n-1/2}, except the rims that lie outside the fused box
// (their H^{n+1/2} was formed by the shell's magnetic kernel) and, for the plane
// above a segment that ends at the box's top, all of it, which come from
// H^{n+1/2}; the computing wavefronts then write H^{n+1/2} over it in place (each
// element is read and written by one thread), and the next plane's electric
// update reads its neighbours there after the plane's barrier.
template <typename T, int D>
struct FRing {
    static constexpr int W = int(sizeof(T) / 4);  // dwords per element
    static constexpr int EX = BX + 2, EY = BY + 2, HX = BX + 1, HY = BY + 1;
    static constexpr int ESP = (EX * EY * W + 63) / 64 * 64;  // dwords per E region
    static constexpr int HSP = (HX * HY * W + 63) / 64 * 64;  // dwords per H region
    static constexpr int NQE = ESP / 64, NQH = HSP / 64;
    static constexpr int NQ = 3 * (NQE + NQH);    // loads per plane
    static constexpr int SLOT = 3 * (ESP + HSP);  // dwords per plane
    static constexpr int R = D + 2;               // planes in use (2) and in flight (D)
    static_assert(D >= 1 && NQ * (D - 1) <= 63, "outstanding loads exceed the vmcnt range");
};

typedef __attribute__((address_space(3))) void* lds_void_ptr;

// s_waitcnt vmcnt(n) with the other counters left alone (gfx9 encoding)
template <int N>
__device__ __forceinline__ void wait_vmcnt()
{
    static_assert(N >= 0 && N <= 63, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

template <typename T, bool MAT, int NPM, bool CP, bool DENSE, int D>
__global__ void __launch_bounds__(BX * BY + 64) k_fstep(const StepArgs<T> S)
{
    using RG = FRing<T, D>;
    const HalfArgs<T>& A = S.e;
    extern __shared__ __align__(16) double lds_tab[];
    MatEntry* smat = reinterpret_cast<MatEntry*>(lds_tab);
    double* spc = lds_tab + 4 * A.nmats;
    __shared__ uint32_t ring[RG::R * RG::SLOT];
    const int tid = int(threadIdx.x);
    const bool loader = tid >= BX * BY;
    if (MAT && !loader) {
        for (int q = tid; q < A.nmats; q += BX * BY)
            smat[q] = A.mats[q];
        if (NPM > 0)
            for (int q = tid; q < 4 * A.npoles; q += BX * BY)
                spc[q] = A.coef[q];
        // visible after the first barrier
    }
    const int b = int(blockIdx.x);
    const int slot = (b & 7) * (A.nblocks >> 3) + (b >> 3);
    if (slot >= A.ntiles)
        return;  // the whole workgroup
    const int t = int(A.tiles[slot]);
    const int ftx = t % S.fn[0];
    const int tr = t / S.fn[0];
    const int fty = tr % S.fn[1];
    const int ftz = tr / S.fn[1];
    const int i0 = S.fo[0] + ftx * BX, j0 = S.fo[1] + fty * BY;
    const int kb = S.fo[2] + ftz * A.lz;
    const int ke = min(kb + A.lz - 1, S.fe[2]);
    const int jt = j0 + BY, ir = i0 + BX;
    // rims outside the box: their H^{n+1/2} is in the new buffer already
    const bool top_sh = jt > S.fe[1], right_sh = ir > S.fe[0], above_sh = ke + 1 > S.fe[2];
    // load coordinates stay in [dlo - 1, dhi + 1], inside every array's padding
    auto clx = [&](int x) { return min(max(x, A.dlo[0] - 1), A.dhi[0] + 1); };
    auto cly = [&](int y) { return min(max(y, A.dlo[1] - 1), A.dhi[1] + 1); };
    auto clz = [&](int z) { return min(max(z, A.dlo[2] - 1), A.dhi[2] + 1); };
    const int sz = A.plane;
    // ring slot of plane p (planes kb-1 .. ke+1)
    auto slot_base = [&](int p) { return ring + ((p - kb + 1) % RG::R) * RG::SLOT; };

    if (loader) {
        // ---- the loading wavefront
        const int lane = tid - BX * BY;
        uint32_t qe[RG::NQE], qh[RG::NQH];
        uint32_t hsel = 0;  // H loads whose element is a rim taken from H^{n+1/2}
#pragma unroll
        for (int q = 0; q < RG::NQE; ++q) {
            const int d = q * 64 + lane, e = d / RG::W, dw = d % RG::W;
            int x = i0, y = j0;
            if (e < RG::EX * RG::EY) {
                x = i0 - 1 + e % RG::EX;
                y = j0 - 1 + e / RG::EX;
            }
            qe[q] = (uint32_t(A.origin + clx(x) + A.pitch * cly(y)) * RG::W + uint32_t(dw)) * 4u;
        }
#pragma unroll
        for (int q = 0; q < RG::NQH; ++q) {
            const int d = q * 64 + lane, e = d / RG::W, dw = d % RG::W;
            int x = i0, y = j0;
            if (e < RG::HX * RG::HY) {
                const int xr = e % RG::HX, yr = e / RG::HX;
                x = i0 + xr;
                y = j0 + yr;
                if ((yr == BY && top_sh) || (xr == BX && right_sh))
                    hsel |= 1u << q;
            }
            qh[q] = (uint32_t(A.origin + clx(x) + A.pitch * cly(y)) * RG::W + uint32_t(dw)) * 4u;
        }
        auto issue = [&](int p) {
            const uint32_t zoff = uint32_t(sz) * uint32_t(clz(p)) * uint32_t(sizeof(T));
            uint32_t* base = slot_base(p);
            const bool last_sh = above_sh && p == ke + 1;
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int q = 0; q < RG::NQE; ++q)
                    __builtin_amdgcn_global_load_lds(reinterpret_cast<const char*>(A.u[c]) + (qe[q] + zoff),
                                                     (lds_void_ptr)(base + c * RG::ESP + q * 64), 4, 0, 0);
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int q = 0; q < RG::NQH; ++q) {
                    const bool ho = last_sh || ((hsel >> q) & 1u);
                    const char* src = ho ? reinterpret_cast<const char*>(S.hout[c])
                                         : reinterpret_cast<const char*>(S.hin[c]);
                    __builtin_amdgcn_global_load_lds(src + (qh[q] + zoff),
                                                     (lds_void_ptr)(base + 3 * RG::ESP + c * RG::HSP + q * 64), 4, 0, 0);
                }
        };
        // wait until at most `after` planes issued after the one needed are in flight
        auto wait_planes = [&](int after) {
            if (after <= 0)
                wait_vmcnt<0>();
            else if (D >= 2 && after == 1)
                wait_vmcnt<(D >= 2 ? RG::NQ : 0)>();
            else
                wait_vmcnt<RG::NQ*(D - 1)>();
        };
        const int plast = ke + 1;
        wait_vmcnt<0>();
        for (int p = kb - 1; p <= min(kb + D - 1, plast); ++p)
            issue(p);
        wait_planes(min(D - 1, plast - kb));  // planes kb-1 and kb
        __syncthreads();
        if (kb + D <= plast)
            issue(kb + D);
        for (int k = kb; k <= ke; ++k) {
            wait_planes(min(D - 1, ke - k));  // plane k+1
            __syncthreads();
            if (k + 1 + D <= plast)
                issue(k + 1 + D);
        }
        return;
    }

    // ---- the computing wavefronts: thread (tx, ty) of the BX x BY tile
    const int tx = tid % BX, ty = tid / BX;
    const int gtx = (i0 - A.dlo[0]) / BX;  // tile column in the half step's grid
    const int i = i0 + tx, j = j0 + ty;
    const bool top = ty == BY - 1, right = tx == BX - 1;
    const int il = clx(i), jl = cly(j);
    const double inv_dx = A.inv_dx;
    ColCtx X;
    X.i = i;
    X.j = j;
    X.base = A.origin + i + A.pitch * j;
#pragma unroll
    for (int C = 0; C < 3; ++C)
        X.in[C] = i >= A.lo[C][0] && i <= A.hi[C][0] && j >= A.lo[C][1] && j <= A.hi[C][1];
    X.sx = X.sy = -1;
    X.bx = X.cx = X.by = X.cy = 0;
    X.ikx = X.iky = 1;
    auto hreal_xy = [&](int C, int x, int y) {
        return x >= S.hlo[C][0] && x <= S.hhi[C][0] && y >= S.hlo[C][1] && y <= S.hhi[C][1];
    };
    auto hreal_z = [&](int C, int z) { return z >= S.hlo[C][2] && z <= S.hhi[C][2]; };
    bool hown[3];
#pragma unroll
    for (int C = 0; C < 3; ++C)
        hown[C] = hreal_xy(C, i, j);
    const bool t_hz = top && hreal_xy(2, i, jt), t_hx = top && hreal_xy(0, i, jt);
    const bool r_hz = right && hreal_xy(2, ir, j), r_hy = right && hreal_xy(1, ir, j);

    // material words: k_half's rule (outside the material box, the default word)
    const bool mcol = i >= A.mbox[0][0] && i <= A.mbox[0][1] && j >= A.mbox[1][0] && j <= A.mbox[1][1];
    const uint32_t bg_word = __builtin_amdgcn_readfirstlane(A.bg_word);
    constexpr bool POL = MAT && NPM > 0 && CP;
    // rows past the grid (the last tile row's spare threads) read the last row's entry
    const int seg0 = (min(j, A.dhi[1]) - A.dlo[1]) * A.gx + gtx;
    const int segplane = A.seg_ny * A.gx;
    auto load_words = [&](int kk, uint32_t& id, uint32_t sb[3]) {
        const int kc = clz(kk);
        if (MAT) {
            const bool in = mcol && kc >= A.mbox[2][0] && kc <= A.mbox[2][1];
            const uint32_t w = ldb(A.ids, in ? uint32_t(A.origin + il + A.pitch * jl + sz * kc) * 4u : 0u);
            id = in ? w : bg_word;
        } else {
            id = 0u;
        }
        const int kq = min(max(kc, A.dlo[2]), A.dhi[2]);
#pragma unroll
        for (int C = 0; C < 3; ++C)
            sb[C] = POL ? A.segbase[C][(kq - A.dlo[2]) * segplane + seg0] : 0u;
    };
    // staged values: E^n component c at tile offset (x, y), x in [-1, BX], y in
    // [-1, BY]; H region component c at (x, y), x in [0, BX], y in [0, BY]
    constexpr int EC = RG::ESP / RG::W, HC = RG::HSP / RG::W;
    auto ebase = [&](int p) { return reinterpret_cast<const T*>(slot_base(p)); };
    auto hbase = [&](int p) { return reinterpret_cast<const T*>(slot_base(p) + 3 * RG::ESP); };
    auto ev = [&](const T* E, int c, int y, int x) { return double(E[c * EC + (y + 1) * RG::EX + (x + 1)]); };
    auto hix = [&](int c, int y, int x) { return c * HC + y * RG::HX + x; };
    auto hv = [&](const T* G, int c, int y, int x) { return G[hix(c, y, x)]; };

    // H^{n+1/2} of plane kk (own point and rim points) from the staged E^n planes
    // kk and kk-1 and the H region of plane kk; fills hs, stores the own point when
    // kk <= ke and returns its values in hn
    auto h_plane = [&](int kk, T hn[3]) {
        const T* E = ebase(kk);
        const T* Ep = ebase(kk - 1);
        const T* G = hbase(kk);
        // H^{n+1/2} replaces H^{n-1/2} in the same region: every element is read
        // and then written by one thread only
        T* H = const_cast<T*>(G);
        const bool st = kk <= ke;
        const double ex = ev(E, 0, ty, tx), ey = ev(E, 1, ty, tx), ez = ev(E, 2, ty, tx);
        const uint32_t o = uint32_t(X.base + sz * kk) * uint32_t(sizeof(T));
        {
            const bool r = hown[0] && hreal_z(0, kk);
            const double d1 = (ev(E, 2, ty - 1, tx) - ez) * inv_dx;
            const double d2 = (ev(Ep, 1, ty, tx) - ey) * inv_dx;
            const T h0 = hv(G, 0, ty, tx);
            const T v = r ? h_value(S, d1, d2, h0) : h0;
            hn[0] = v;
            H[hix(0, ty, tx)] = v;
            if (r && st)
                stb(S.hout[0], o, v);
        }
        {
            const bool r = hown[1] && hreal_z(1, kk);
            const double d1 = (ev(Ep, 0, ty, tx) - ex) * inv_dx;
            const double d2 = (ev(E, 2, ty, tx - 1) - ez) * inv_dx;
            const T h0 = hv(G, 1, ty, tx);
            const T v = r ? h_value(S, d1, d2, h0) : h0;
            hn[1] = v;
            H[hix(1, ty, tx)] = v;
            if (r && st)
                stb(S.hout[1], o, v);
        }
        {
            const bool r = hown[2] && hreal_z(2, kk);
            const double d1 = (ev(E, 1, ty, tx - 1) - ey) * inv_dx;
            const double d2 = (ev(E, 0, ty - 1, tx) - ex) * inv_dx;
            const T h0 = hv(G, 2, ty, tx);
            const T v = r ? h_value(S, d1, d2, h0) : h0;
            hn[2] = v;
            H[hix(2, ty, tx)] = v;
            if (r && st)
                stb(S.hout[2], o, v);
        }
        if (top && !top_sh) {
            const double txv = ev(E, 0, BY, tx), tyv = ev(E, 1, BY, tx), tzv = ev(E, 2, BY, tx);
            {  // Hz at (i, jt)
                const double d1 = (ev(E, 1, BY, tx - 1) - tyv) * inv_dx;
                const double d2 = (ev(E, 0, BY - 1, tx) - txv) * inv_dx;
                const T h0 = hv(G, 2, BY, tx);
                H[hix(2, BY, tx)] = t_hz && hreal_z(2, kk) ? h_value(S, d1, d2, h0) : h0;
            }
            {  // Hx at (i, jt)
                const double d1 = (ev(E, 2, BY - 1, tx) - tzv) * inv_dx;
                const double d2 = (ev(Ep, 1, BY, tx) - tyv) * inv_dx;
                const T h0 = hv(G, 0, BY, tx);
                H[hix(0, BY, tx)] = t_hx && hreal_z(0, kk) ? h_value(S, d1, d2, h0) : h0;
            }
        }
        if (right && !right_sh) {
            const double rxv = ev(E, 0, ty, BX), ryv = ev(E, 1, ty, BX), rzv = ev(E, 2, ty, BX);
            {  // Hz at (ir, j)
                const double d1 = (ev(E, 1, ty, BX - 1) - ryv) * inv_dx;
                const double d2 = (ev(E, 0, ty - 1, BX) - rxv) * inv_dx;
                const T h0 = hv(G, 2, ty, BX);
                H[hix(2, ty, BX)] = r_hz && hreal_z(2, kk) ? h_value(S, d1, d2, h0) : h0;
            }
            {  // Hy at (ir, j)
                const double d1 = (ev(Ep, 0, ty, BX) - rxv) * inv_dx;
                const double d2 = (ev(E, 2, ty, BX - 1) - rzv) * inv_dx;
                const T h0 = hv(G, 1, ty, BX);
                H[hix(1, ty, BX)] = r_hy && hreal_z(1, kk) ? h_value(S, d1, d2, h0) : h0;
            }
        }
    };

    uint32_t idc, sbc[3];
    load_words(kb, idc, sbc);
    __syncthreads();  // planes kb-1 and kb staged; material table visible
    {
        T h0[3];
        h_plane(kb, h0);
    }
    for (int k = kb; k <= ke; ++k) {
        // this plane's polarization loads and the next plane's material words go
        // out before the barrier
        PlaneIn<T, NPM, CP> L;
        {
            ZCtx Z;
            Z.k = k;
            Z.o = X.base + sz * k;
            Z.sz = -1;
            Z.bz = Z.cz = 0;
            Z.ikz = 1;
            bool act[3];
#pragma unroll
            for (int C = 0; C < 3; ++C)
                act[C] = X.in[C] && k >= A.lo[C][2] && k <= A.hi[C][2];
            load_sparse<T, MAT, NPM, CP, DENSE, false, true>(A, smat, X, Z, idc, sbc, act, L);
        }
        uint32_t idn, sbn[3];
        load_words(k + 1, idn, sbn);
        __syncthreads();  // plane k+1 staged; H^{n+1/2} of plane k complete in hs
        T hn[3];
        if (!(k == ke && above_sh)) {
            h_plane(k + 1, hn);
        } else {
            // the plane above is the shell's: its new H was staged into the H region
            const T* G = hbase(k + 1);
#pragma unroll
            for (int C = 0; C < 3; ++C)
                hn[C] = hv(G, C, ty, tx);
        }
        // E^{n+1} of plane k
        const T* E = ebase(k);
        const T* H = hbase(k);
        ZCtx Z;
        Z.k = k;
        Z.o = X.base + sz * k;
        Z.sz = -1;
        Z.bz = Z.cz = 0;
        Z.ikz = 1;
#pragma unroll
        for (int C = 0; C < 3; ++C)
            L.e[C] = E[C * EC + (ty + 1) * RG::EX + (tx + 1)];
        const double hx = double(hv(H, 0, ty, tx)), hy = double(hv(H, 1, ty, tx));
        const double hz = double(hv(H, 2, ty, tx));
        if (X.in[0] && k >= A.lo[0][2] && k <= A.hi[0][2]) {
            const double d1 = (double(hv(H, 2, ty + 1, tx)) - hz) * inv_dx;
            const double d2 = (double(hn[1]) - hy) * inv_dx;
            update_comp<T, MAT, NPM, CP, DENSE, 0, false>(A, smat, spc, X, Z, d1, d2, idc, L);
        }
        if (X.in[1] && k >= A.lo[1][2] && k <= A.hi[1][2]) {
            const double d1 = (double(hn[0]) - hx) * inv_dx;
            const double d2 = (double(hv(H, 2, ty, tx + 1)) - hz) * inv_dx;
            update_comp<T, MAT, NPM, CP, DENSE, 1, false>(A, smat, spc, X, Z, d1, d2, idc, L);
        }
        if (X.in[2] && k >= A.lo[2][2] && k <= A.hi[2][2]) {
            const double d1 = (double(hv(H, 1, ty, tx + 1)) - hy) * inv_dx;
            const double d2 = (double(hv(H, 0, ty + 1, tx)) - hx) * inv_dx;
            update_comp<T, MAT, NPM, CP, DENSE, 2, false>(A, smat, spc, X, Z, d1, d2, idc, L);
        }
        idc = idn;
#pragma unroll
        for (int C = 0; C < 3; ++C)
            sbc[C] = sbn[C];
    }
}

// E/H DFT monitors (DFT_Unit::update, src/dft.cpp:29-62).  Every step, after the
// point sources, k_dft_gather copies each monitored point's stored value into
// its unit's sample ring; every DFT_BATCH steps (and before any read of the
// sums) k_dft_ring adds the batch into the FP64 running sums:
//   re[f][p] += cos(w_f t_s) x_s[p],  im[f][p] += -sin(w_f t_s) x_s[p]
// for s in step order, one FP64 multiply-add per term exactly as a per-step
// accumulation would do it.  The ring holds the stored (storage-precision)
// values, which are the values the per-step form would read, so FP32 storage
// keeps FP64 sums.  Per point and frequency the sums are read and written once
// per batch: 32 B per DFT_BATCH steps instead of 32 B per step.
template <typename T>
struct RingArgs {
    const T* u[6];  // E then H components (offset to lattice index 0)
    int pitch, plane, tail0;
    int nunits;
    int64_t first[MAXU * 2 + 1];  // unit q's points are gather threads first[q] .. first[q+1]-1
    DftDev units[MAXU * 2];
    int half[MAXU * 2];           // 0: electric unit, 1: magnetic
    int slot;                     // ring slot of this step
    int cnt;                      // k_dft_ring: samples in the ring
    const double* ph[2];          // k_dft_ring: phasor rows of the batch's first step, per half
    int mf[2];                    // phasor row length (frequencies) per half
};

template <typename T>
__global__ void __launch_bounds__(256) k_dft_gather(const RingArgs<T> A)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= A.first[A.nunits])
        return;
    int q = 0;
    while (t >= A.first[q + 1])
        ++q;
    const DftDev& U = A.units[q];
    const int p = int(t - A.first[q]);  // unit sizes are checked < 2^31
    const int ii = p % U.dims[0];
    const int r = p / U.dims[0];
    const int jj = r % U.dims[1];
    const int kk = r / U.dims[1];
    const int o = int(lat_off(A.pitch, A.plane, A.tail0, U.box[0][0] + ii, U.box[1][0] + jj, U.box[2][0] + kk));
    const T* u = A.u[A.half[q] * 3 + U.comp];
    static_cast<T*>(U.ring)[int64_t(A.slot) * U.npts + p] = u[o];
}

// one thread per point of unit blockIdx.y: the batch's samples in registers, then
// every frequency's sums; the unit's phasors of the batch are staged in LDS
template <typename T>
__global__ void __launch_bounds__(256) k_dft_ring(const RingArgs<T> A)
{
    extern __shared__ double rph[];  // [cnt][nf][2]
    const DftDev& U = A.units[blockIdx.y];
    const int h = A.half[blockIdx.y];
    const int nf = U.nf, cnt = A.cnt;
    for (int q = int(threadIdx.x); q < cnt * nf; q += int(blockDim.x)) {
        const int st = q / nf, f = q - st * nf;
        const double* row = A.ph[h] + int64_t(st) * 2 * A.mf[h] + 2 * (U.mf0 + f);
        rph[2 * q] = row[0];
        rph[2 * q + 1] = row[1];
    }
    __syncthreads();
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= U.npts)
        return;
    const T* ring = static_cast<const T*>(U.ring);
    double v[DFT_BATCH];
#pragma unroll
    for (int st = 0; st < DFT_BATCH; ++st)
        v[st] = st < cnt ? double(ring[int64_t(st) * U.npts + p]) : 0.0;
    for (int f = 0; f < nf; ++f) {
        double re = U.re[int64_t(f) * U.npts + p], im = U.im[int64_t(f) * U.npts + p];
#pragma unroll
        for (int st = 0; st < DFT_BATCH; ++st)
            if (st < cnt) {
                im += rph[2 * (st * nf + f) + 1] * v[st];
                re += rph[2 * (st * nf + f)] * v[st];
            }
        U.re[int64_t(f) * U.npts + p] = re;
        U.im[int64_t(f) * U.npts + p] = im;
    }
}

// ---------------------------------------------------------------- sources
template <typename T>
struct SrcArgs {
    int ntarget;
    T* u[3];
    const int64_t* idx;
    const int* comp;
    const double* scale;  // dt / g_inf of the target point
    const int* eoff;      // [ntarget+1] into entries
    const double* amp;
    const double* start;
    const double* end;
    const double* width;
    const int* moff;      // [ntarget+1] into monitor hits
    const int* munit;
    const int64_t* mofs;
    double dt;
    DftDev units[MAXU];
    const double* ph;
};

template <typename T>
__global__ void k_sources(const SrcArgs<T> A, double t)
{
    const int q = int(blockIdx.x) * int(blockDim.x) + int(threadIdx.x);
    if (q >= A.ntarget)
        return;
    double g = 0;
    for (int e = A.eoff[q]; e < A.eoff[q + 1]; ++e)
        if (t > A.start[e] && t < A.end[e])
            g += A.amp[e] * gaussian1(t - (A.start[e] + A.end[e]) / 2, A.width[e]);
    if (g == 0)
        return;
    T* f = A.u[A.comp[q]];
    const int64_t i = A.idx[q];
    const T old = f[i];
    const T nv = T(double(old) + g * A.scale[q]);
    f[i] = nv;
    const double dv = double(nv) - double(old);
    const double dD = g * A.dt;
    for (int m = A.moff[q]; m < A.moff[q + 1]; ++m) {
        const DftDev& U = A.units[A.munit[m]];
        const double v = U.is_db ? dD : dv;
        const double* ph = A.ph + 2 * U.mf0;
        for (int fq = 0; fq < U.nf; ++fq) {
            const int64_t o = int64_t(fq) * U.npts + A.mofs[m];
            U.im[o] += ph[2 * fq + 1] * v;
            U.re[o] += ph[2 * fq] * v;
        }
    }
}

// ---------------------------------------------------------------- ghost faces
struct FaceDev {
    int comp, axis, dst, src;
    int t0lo, t0hi, t1lo, t1hi;  // transverse ranges (axes (axis+1)%3, (axis+2)%3)
    double mult;
};
static constexpr int MAXF = 16;
template <typename T>
struct FaceArgs {
    T* u[3];
    int64_t pitch, plane;
    int tail0;
    int nf;
    FaceDev f[MAXF];
};

template <typename T>
__global__ void k_faces(const FaceArgs<T> A)
{
    const FaceDev& F = A.f[blockIdx.z];
    const int t0 = F.t0lo + int(blockIdx.x) * int(blockDim.x) + int(threadIdx.x);
    const int t1 = F.t1lo + int(blockIdx.y);
    if (t0 > F.t0hi || t1 > F.t1hi)
        return;
    int c[3];
    const int a0 = (F.axis + 1) % 3, a1 = (F.axis + 2) % 3;
    c[a0] = t0;
    c[a1] = t1;
    c[F.axis] = F.src;
    const int64_t s = lat_off(A.pitch, A.plane, A.tail0, c[0], c[1], c[2]);
    c[F.axis] = F.dst;
    const int64_t d = lat_off(A.pitch, A.plane, A.tail0, c[0], c[1], c[2]);
    T* u = A.u[F.comp % 3];
    u[d] = T(F.mult * double(u[s]));
}

// ---------------------------------------------------------------- stop-condition samples
// sum over a DFT unit's points of re^2 + im^2 at one frequency
// (DFT_Unit::get_local_norm, src/dft.cpp:93-107); one workgroup, result to out[slot]
__global__ void __launch_bounds__(1024) k_dft_norm(const double* re, const double* im, int64_t n, double* out,
                                                   int64_t slot)
{
    __shared__ double part[16];
    double s = 0;
    for (int64_t q = threadIdx.x; q < n; q += blockDim.x)
        s += re[q] * re[q] + im[q] * im[q];
    for (int off = 32; off > 0; off >>= 1)
        s += __shfl_xor(s, off);
    const int w = int(threadIdx.x) / 64;
    if ((threadIdx.x & 63) == 0)
        part[w] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int q = 0; q < int(blockDim.x) / 64; ++q)
            t += part[q];
        out[slot] = t;
    }
}

// interpolated field value at a point (Fields::get_eh_helper, src/fields.cpp:245-248)
template <typename T>
__global__ void k_point_value(const T* u, const int64_t* idx, const double* w, int nterm, double* out, int64_t slot)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double v = 0;
        for (int q = 0; q < nterm; ++q)
            v += double(u[idx[q]]) * w[q];
        out[slot] = v;
    }
}

// ---------------------------------------------------------------- adjoint gradient
// Sensitivity Re(2 pi i f E_fwd E_adj eps' / norm) at the E points strictly inside
// the design box (ffip/simulation.py:841-869), carried onto the density nodes by
// the adjoint of zero-padded trilinear interpolation (app/ffip.cpp:308-352).
// Gather form: one thread per density node sums the points in its support, in a
// fixed order, so the result is deterministic.
struct GradComp {
    const double *fre, *fim, *are, *aim;  // forward / adjoint DFT at the design frequency
    int lo[3];                            // unit box origin (cell index)
    int dims[3];                          // unit dims
    int par[3];                           // parity of the component per axis
};

struct GradArgs {
    GradComp c[3];
    int M[3];
    double dx2;
    double blo[3], bhi[3];  // design box, doubled coordinates, strict interior
    double gx0[3], dxn[3];  // density nodes: origin and spacing (physical)
    int nd[3];              // density nodes per axis
    double two_pi_f, epr, epi, nr, ni;
    double* grad;  // [nz][ny][nx]
};

__global__ void k_adjoint_gradient(const GradArgs A)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t nn = int64_t(A.nd[0]) * A.nd[1] * A.nd[2];
    if (t >= nn)
        return;
    const int node[3] = {int(t % A.nd[0]), int((t / A.nd[0]) % A.nd[1]), int(t / (int64_t(A.nd[0]) * A.nd[1]))};
    // eps' / norm
    const double den = A.nr * A.nr + A.ni * A.ni;
    const double qr = (A.epr * A.nr + A.epi * A.ni) / den, qi = (A.epi * A.nr - A.epr * A.ni) / den;
    double acc = 0;
    for (int c = 0; c < 3; ++c) {
        const GradComp& G = A.c[c];
        int l0[3], l1[3];
        for (int a = 0; a < 3; ++a) {
            if (A.nd[a] == 1) {
                l0[a] = 0;
                l1[a] = G.dims[a] - 1;
                continue;
            }
            // lattice points whose normalised coordinate lies within one spacing of the node
            const double xlo = (A.gx0[a] + (node[a] - 1) * A.dxn[a]) / A.dx2;
            const double xhi = (A.gx0[a] + (node[a] + 1) * A.dxn[a]) / A.dx2;
            l0[a] = max(0, int(floor((xlo + A.M[a] - 1 + G.par[a]) / 2)) - G.lo[a] - 1);
            l1[a] = min(G.dims[a] - 1, int(ceil((xhi + A.M[a] - 1 + G.par[a]) / 2)) - G.lo[a] + 1);
        }
        for (int kk = l0[2]; kk <= l1[2]; ++kk)
            for (int jj = l0[1]; jj <= l1[1]; ++jj)
                for (int ii = l0[0]; ii <= l1[0]; ++ii) {
                    const int loc[3] = {ii, jj, kk};
                    double w = 1;
                    bool ok = true;
                    for (int a = 0; a < 3 && ok; ++a) {
                        const double xd = double(2 * (G.lo[a] + loc[a]) - A.M[a] + (1 - G.par[a]));
                        if (!(xd > A.blo[a] && xd < A.bhi[a])) {
                            ok = false;
                            break;
                        }
                        if (A.nd[a] == 1)
                            continue;
                        const double s = (xd * A.dx2 - A.gx0[a]) / A.dxn[a];
                        const double d = fabs(s - double(node[a]));
                        if (!(d < 1)) {
                            ok = false;
                            break;
                        }
                        w *= 1 - d;
                    }
                    if (!ok)
                        continue;
                    const int64_t o = ii + int64_t(G.dims[0]) * (jj + int64_t(G.dims[1]) * kk);
                    const double fr = G.fre[o], fi = G.fim[o], ar = G.are[o], ai = G.aim[o];
                    const double pr = fr * ar - fi * ai, pi = fr * ai + fi * ar;  // E_fwd E_adj
                    const double xi = pr * qi + pi * qr;                          // Im(E_fwd E_adj eps'/norm)
                    acc += w * (-A.two_pi_f * xi);                                // Re(2 pi i f (...))
                }
    }
    A.grad[t] = acc;
}

// ---------------------------------------------------------------- probes / phasors
template <typename T>
struct ProbeArgs {
    const T* u[6];
    int np;
    const int* off;
    const int* comp;
    const int64_t* idx;
    const double* w;
    double* trace;
};

template <typename T>
__global__ void k_probes(const ProbeArgs<T> A, int64_t row)
{
    const int p = int(blockIdx.x) * int(blockDim.x) + int(threadIdx.x);
    if (p >= A.np)
        return;
    double v = 0;
    for (int q = A.off[p]; q < A.off[p + 1]; ++q)
        v += double(A.u[A.comp[q]][A.idx[q]]) * A.w[q];
    A.trace[row * A.np + p] = v;
}

// table[s][mf][2] for steps step0 .. step0+count-1: time = (step + toff) * dt
__global__ void k_phasors(double* table, const double* freqs, int mf, int64_t step0, int count, double toff, double dt)
{
    const int q = int(blockIdx.x) * int(blockDim.x) + int(threadIdx.x);
    if (q >= mf * count)
        return;
    const int s = q / mf, f = q % mf;
    const double pi = 3.141592653589793e+00;
    const double time = (double(step0 + s) + toff) * dt;
    const double arg = 2 * pi * freqs[f] * time;
    table[2 * int64_t(q)] = cos(arg);
    table[2 * int64_t(q) + 1] = -sin(arg);
}

// ---------------------------------------------------------------- engine
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

// Device blocks of destroyed engines, kept for the next engine of the process:
// repeated solves of one problem (the adjoint loop's forward and adjoint runs)
// allocate the same sizes, and hipMalloc / hipFree both cost milliseconds and
// synchronise the device.  Blocks are reused at their exact (rounded) size on the
// same device; at most FFIP_DEVICE_CACHE_MB (default 4096) MiB are held, beyond
// which released blocks are freed.  The cache lives until the process ends.
class BlockCache {
public:
    static BlockCache& get()
    {
        static BlockCache* c = new BlockCache;  // never destroyed: no HIP calls at exit
        return *c;
    }
    static size_t round(size_t bytes)
    {
        const size_t g = bytes >= (size_t(1) << 20) ? (size_t(1) << 20) : 256;
        return (bytes + g - 1) / g * g;
    }
    void* take(size_t bytes, int dev)
    {
        {
            std::lock_guard<std::mutex> lk(m_);
            auto it = free_.find({dev, bytes});
            if (it != free_.end()) {
                void* p = it->second;
                free_.erase(it);
                held_ -= bytes;
                return p;
            }
        }
        void* p = nullptr;
        HIP_OK(hipMalloc(&p, bytes));
        return p;
    }
    // p was allocated on device dev, which must be current
    void give(void* p, size_t bytes, int dev)
    {
        if (bytes > cap_) {
            (void)hipFree(p);
            return;
        }
        std::lock_guard<std::mutex> lk(m_);
        while (held_ + bytes > cap_ && !free_.empty()) {  // drop the largest blocks first
            auto it = std::prev(free_.end());
            (void)hipFree(it->second);
            held_ -= it->first.second;
            free_.erase(it);
        }
        free_.emplace(std::make_pair(dev, bytes), p);
        held_ += bytes;
    }

private:
    BlockCache()
    {
        const char* e = std::getenv("FFIP_DEVICE_CACHE_MB");
        cap_ = size_t(e ? std::max(0.0, std::atof(e)) : 4096.0) << 20;
    }
    std::mutex m_;
    std::multimap<std::pair<int, size_t>, void*> free_;
    size_t held_ = 0, cap_ = 0;
};

struct EngineImpl {
    std::unique_ptr<Setup> S;
    bool fp32 = false;
    size_t tsz = 8;
    hipStream_t stream = nullptr;
    std::vector<std::pair<void*, size_t>> allocs;  // device blocks and their rounded sizes
    int device = 0;                                 // current device when the engine was built
    double bytes = 0;
    int64_t n = 0;  // time_step

    void* field[6] = {};
    // fused full-step mode: field holds E^n / H^{n-1/2}, field_b receives the
    // step's new values and the two swap after every step (the CPML auxiliaries
    // are read and written only by the shell's half-step kernels, in place)
    bool fused = false;
    // planes the fused kernel's loads run ahead of the plane being computed
    // (FFIP_FDEPTH, 1 or 2)
    int fdepth = 1;
    void* field_b[6] = {};
    int lzf = 32;
    uint32_t* ftiles = nullptr;  // fused-kernel tiles: 0 .. nft-1 over its box
    int nft = 0;
    int fbox[3][2] = {{0, -1}, {0, -1}, {0, -1}};  // its box of lattice indices
    int fcount[3] = {0, 0, 0};                       // its tiles per axis
    // the CPML shell's half-step tiles, per half: clear of that half's slabs, not
    uint32_t* stiles[2][2] = {};
    int nst[2][2] = {};
    double traffic_f[7] = {};
    void* Pa[2] = {nullptr, nullptr};
    void* Pb[2] = {nullptr, nullptr};
    int S_slots[2] = {0, 0};
    int64_t pcomp[2][3] = {};           // element offset of each component's compact P entries
    uint32_t* segbase[2][3] = {};       // first entry of every BX-point row segment
    int pseg_ny[2] = {0, 0};
    uint32_t* ids[2] = {};  // packed ids of the three components of a half
    MatEntry* mats[2] = {};
    int nmats[2] = {0, 0};
    double* p_coef[2] = {};
    int npoles[2] = {0, 0};
    int lz = 32;
    double *d_inv[2][3] = {}, *d_eps[2][3] = {}, *d_amp[2][3] = {}, *d_c[2][3] = {};
    PmlDev pml[2][3] = {};
    std::vector<DftDev> units[2];
    // E/H monitor sample rings: steps gathered since the last flush, and the
    // first of them; the units they serve (half, unit)
    int ring_cnt = 0;
    int64_t ring_n0 = 0;
    std::vector<std::pair<int, int>> ring_units;
    template <typename T>
    RingArgs<T> ring_args();
    template <typename T>
    void gather_dft();
    template <typename T>
    void flush_ring();
    void flush_dft();
    std::vector<std::vector<double>> host_freqs[2];
    double* freqs_dev[2] = {};
    int MF[2] = {0, 0};
    double* ph[2] = {};
    int64_t ph_start = -1;
    // sources
    struct Src {
        int ntarget = 0;
        int64_t* idx = nullptr;
        int *comp = nullptr, *eoff = nullptr, *moff = nullptr, *munit = nullptr;
        double *scale = nullptr, *amp = nullptr, *start = nullptr, *end = nullptr, *width = nullptr;
        int64_t* mofs = nullptr;
        double max_end = -1;
    } src[2];
    // faces
    std::vector<FaceDev> faces[2];
    // probes
    int nprobes = 0;
    int *pr_off = nullptr, *pr_comp = nullptr;
    int64_t* pr_idx = nullptr;
    double* pr_w = nullptr;
    double* trace = nullptr;
    int64_t trace_cap = 0;
    double traffic[2][7] = {};

    template <typename X>
    X* dalloc(size_t count, bool zero = true)
    {
        if (count == 0)
            count = 1;
        const size_t nb = BlockCache::round(count * sizeof(X));
        void* p = BlockCache::get().take(nb, device);
        allocs.emplace_back(p, nb);
        bytes += double(count * sizeof(X));
        if (zero)
            HIP_OK(hipMemsetAsync(p, 0, count * sizeof(X), stream));
        return static_cast<X*>(p);
    }
    // reusable device buffer for per-step stop-condition samples
    double* samples = nullptr;
    size_t samples_cap = 0;
    double* sample_buf(size_t n)
    {
        n = std::max<size_t>(n, 1);
        if (n > samples_cap) {
            HIP_OK(hipStreamSynchronize(stream));
            if (samples) {
                auto it = std::find_if(allocs.begin(), allocs.end(),
                                       [&](const std::pair<void*, size_t>& a) { return a.first == samples; });
                BlockCache::get().give(it->first, it->second, device);
                allocs.erase(it);
            }
            samples = dalloc<double>(n);
            samples_cap = n;
        }
        return samples;
    }
    template <typename X>
    X* upload(const X* v, size_t n)
    {
        // synchronous copy: the host array may be a temporary
        X* p = dalloc<X>(n, false);
        if (n)
            HIP_OK(hipMemcpy(p, v, n * sizeof(X), hipMemcpyHostToDevice));
        else {
            X zero{};
            HIP_OK(hipMemcpy(p, &zero, sizeof(X), hipMemcpyHostToDevice));
        }
        return p;
    }
    template <typename X>
    X* upload(const std::vector<X>& v)
    {
        return upload(v.data(), v.size());
    }

    void build();
    template <typename T>
    HalfArgs<T> half_args(int h);
    template <typename T>
    void launch_half(int h, hipEvent_t e0, hipEvent_t e1);
    // sources, E/H DFT monitors and ghost faces of half h after its kernels
    template <typename T>
    void post_half(int h);
    bool fused_eligible(std::string* why) const;
    template <typename T>
    StepArgs<T> step_args();
    template <typename T>
    void build_fused_tiles();
    template <typename T>
    void step_fused(hipEvent_t* ev);
    template <typename T>
    void launch_step(size_t lds, const StepArgs<T>& S0);
    template <typename T, int D>
    void launch_fstep(bool mat, int npm, size_t lds, const StepArgs<T>& X)
    {
        const dim3 grid(X.e.nblocks, 1, 1), block(BX * BY + 64, 1, 1);
        if (!mat) {
            k_fstep<T, false, 0, false, false, D><<<grid, block, 0, stream>>>(X);
        } else if (dense[0]) {
            k_fstep<T, true, 0, false, true, D><<<grid, block, lds, stream>>>(X);
        } else if (npm == 0) {
            k_fstep<T, true, 0, false, false, D><<<grid, block, lds, stream>>>(X);
        } else if (pcompact[0]) {
            switch (npm) {
            case 1: k_fstep<T, true, 1, true, false, D><<<grid, block, lds, stream>>>(X); break;
            case 2: k_fstep<T, true, 2, true, false, D><<<grid, block, lds, stream>>>(X); break;
            default: k_fstep<T, true, 3, true, false, D><<<grid, block, lds, stream>>>(X); break;
            }
        } else {
            switch (npm) {
            case 1: k_fstep<T, true, 1, false, false, D><<<grid, block, lds, stream>>>(X); break;
            case 2: k_fstep<T, true, 2, false, false, D><<<grid, block, lds, stream>>>(X); break;
            default: k_fstep<T, true, 3, false, false, D><<<grid, block, lds, stream>>>(X); break;
            }
        }
    }
    // half-step kernel of half h over a tile list (with or without the CPML paths)
    template <typename T>
    void launch_tiles(int h, const HalfArgs<T>& A, const uint32_t* list, int count, bool pml);
    // npm: most poles at any point of the half (register arrays are sized by it);
    // cp: compact polarization layout; U: planes whose loads are issued together
    // Kernel choice.  Electric material kernels with up to three poles are
    // compiled with and without the CPML and per-point-medium paths; the other
    // material kernels (magnetic media, more poles) keep both paths, which are
    // correct on any tile.
    template <typename T, bool IS_E, bool CP, bool PML, bool DENSE>
    void launch_mat(int npm, int U, size_t lds, dim3 grid, dim3 block, const HalfArgs<T>& A)
    {
#define FFIP_K(NP, UU, P, D) k_half<T, IS_E, true, NP, CP, UU, P, D, false><<<grid, block, lds, stream>>>(A)
        // two planes per batch only while the pole registers allow (three poles
        // fit without the CPML paths)
        if (U >= 2 && npm <= (PML ? 2 : 3)) {
            if (npm == 1)
                FFIP_K(1, 2, PML, DENSE);
            else if (npm == 2)
                FFIP_K(2, 2, PML, DENSE);
            else if constexpr (!PML)
                FFIP_K(3, 2, PML, DENSE);
            return;
        }
        switch (npm) {
        case 1: FFIP_K(1, 1, PML, DENSE); break;
        case 2: FFIP_K(2, 1, PML, DENSE); break;
        case 3: FFIP_K(3, 1, PML, DENSE); break;
        case 4: FFIP_K(4, 1, true, true); break;
        default: FFIP_K(MAXS, 1, true, true); break;
        }
#undef FFIP_K
    }
    template <typename T, bool IS_E, bool PML, bool DENSE>
    void launch_flags(bool mat, int npm, bool cp, int U, size_t lds, dim3 grid, dim3 block, const HalfArgs<T>& A)
    {
        if (!mat) {
            switch (U) {
            case 1: k_half<T, IS_E, false, 0, false, 1, PML, false, false><<<grid, block, 0, stream>>>(A); break;
            case 2: k_half<T, IS_E, false, 0, false, 2, PML, false, false><<<grid, block, 0, stream>>>(A); break;
            default: k_half<T, IS_E, false, 0, false, 4, PML, false, false><<<grid, block, 0, stream>>>(A); break;
            }
            return;
        }
        if constexpr (IS_E || (PML && DENSE)) {
            if (npm == 0) {
                if (U >= 2)
                    k_half<T, IS_E, true, 0, false, 2, PML, DENSE, false><<<grid, block, lds, stream>>>(A);
                else
                    k_half<T, IS_E, true, 0, false, 1, PML, DENSE, false><<<grid, block, lds, stream>>>(A);
            } else if (cp) {
                launch_mat<T, IS_E, true, PML, DENSE>(npm, U, lds, grid, block, A);
            } else {
                launch_mat<T, IS_E, false, PML, DENSE>(npm, U, lds, grid, block, A);
            }
        }
    }
    // the half-step kernel with the in-kernel (D/B) monitor path: CPML and
    // per-point-medium paths on, up to MAXS poles, one plane per load batch
    bool monitors[2] = {false, false};  // some D/B monitor of the half has points
    template <typename T, bool IS_E>
    void launch_general(bool mat, size_t lds, dim3 grid, dim3 block, const HalfArgs<T>& A)
    {
        const int h = IS_E ? 0 : 1;
        if (!mat)
            k_half<T, IS_E, false, 0, false, 1, true, false, true><<<grid, block, 0, stream>>>(A);
        else if (pcompact[h])
            k_half<T, IS_E, true, MAXS, true, 1, true, true, true><<<grid, block, lds, stream>>>(A);
        else
            k_half<T, IS_E, true, MAXS, false, 1, true, true, true><<<grid, block, lds, stream>>>(A);
    }
    template <typename T, bool IS_E>
    void launch_k(bool mat, int npm, bool cp, int U, bool pml, bool dense, size_t lds, dim3 grid, dim3 block,
                  const HalfArgs<T>& A)
    {
        if constexpr (!IS_E) {
            if (mat) {  // magnetic material kernels: general form only
                launch_flags<T, IS_E, true, true>(mat, npm, cp, U, lds, grid, block, A);
                return;
            }
        }
        if (pml) {
            if (dense)
                launch_flags<T, IS_E, true, true>(mat, npm, cp, U, lds, grid, block, A);
            else
                launch_flags<T, IS_E, true, false>(mat, npm, cp, U, lds, grid, block, A);
        } else {
            if (dense)
                launch_flags<T, IS_E, false, true>(mat, npm, cp, U, lds, grid, block, A);
            else
                launch_flags<T, IS_E, false, false>(mat, npm, cp, U, lds, grid, block, A);
        }
    }
    bool pcompact[2] = {false, false};  // polarization layout of each half
    // tiles of the half step: list a (launched without the CPML paths when
    // pml_a is false) and list b (with them)
    uint32_t* tiles_a[2] = {};
    uint32_t* tiles_b[2] = {};
    int ntiles_a[2] = {0, 0}, ntiles_b[2] = {0, 0};
    bool pml_a[2] = {true, true};
    bool dense[2] = {false, false};  // some point of the half has a per-point medium
    template <typename T>
    void build_tiles();
    int unroll = 2, unroll_mat = 2;  // z planes whose loads are issued together (pole-free / pole kernels)
    int unroll_shell = 0;            // pole kernels on CPML tiles of a split half step (0: unroll_mat)
    int unroll_plain = 1;            // material-table kernels without poles (per-point or piecewise media)
    void ensure_phasors();
    template <typename T>
    void step_once(hipEvent_t* ev, bool timed_e, bool timed_h);
    void step_n(int64_t nsteps);
    ~EngineImpl();
};

EngineImpl::~EngineImpl()
{
    if (stream)
        (void)hipStreamSynchronize(stream);
    // the stream is idle: the blocks can serve the next engine on this device
    int cur = 0;
    const bool switched = hipGetDevice(&cur) == hipSuccess && cur != device;
    if (switched)
        (void)hipSetDevice(device);
    for (auto& a : allocs)
        BlockCache::get().give(a.first, a.second, device);
    if (switched)
        (void)hipSetDevice(cur);
    if (stream)
        (void)hipStreamDestroy(stream);
}

PolLayoutPlan plan_polarization(const Setup& s, int h, int lz)
{
    const Layout& L = s.L;
    const HalfMaterial& M = s.mat[h];
    const size_t tsz = s.precision == "fp32" ? 4 : 8;
    PolLayoutPlan P;
    if (M.uniform || M.max_poles == 0)
        return P;
    for (int a = 0; a < 3; ++a) {
        P.dlo[a] = 1 << 30;
        P.dhi[a] = -1;
        for (int c = 0; c < 3; ++c)
            if (L.hi[h * 3 + c][a] >= L.lo[h * 3 + c][a]) {
                P.dlo[a] = std::min(P.dlo[a], L.lo[h * 3 + c][a]);
                P.dhi[a] = std::max(P.dhi[a], L.hi[h * 3 + c][a]);
            }
    }
    P.gx = std::max(0, (P.dhi[0] - P.dlo[0] + BX) / BX);
    P.ny = std::max(0, P.dhi[1] - P.dlo[1] + 1);
    P.nz = std::max(0, P.dhi[2] - P.dlo[2] + 1);
    const size_t nseg = size_t(P.gx) * P.ny * P.nz;
    lz = std::max(1, lz);
    for (int c = 0; c < 3; ++c) {
        const int comp = h * 3 + c;
        std::vector<uint32_t>& sb = P.segbase[c];
        sb.assign(std::max<size_t>(nseg, 1), 0u);
        int64_t run = 0;
        // entries in tile order (z segment, y tile, x tile), then plane and row
        // within the tile: the segments one workgroup reads for one plane are
        // contiguous, and so are its successive planes
        for (int z0 = P.dlo[2]; z0 <= P.dhi[2]; z0 += lz)
            for (int y0 = P.dlo[1]; y0 <= P.dhi[1]; y0 += BY)
                for (int tx = 0; tx < P.gx; ++tx)
                    for (int k = z0; k <= std::min(z0 + lz - 1, P.dhi[2]); ++k)
                        for (int j = y0; j <= std::min(y0 + BY - 1, P.dhi[1]); ++j) {
                            const size_t seg = (size_t(k - P.dlo[2]) * P.ny + size_t(j - P.dlo[1])) * P.gx + tx;
                            sb[seg] = uint32_t(run);
                            if (j < L.lo[comp][1] || j > L.hi[comp][1] || k < L.lo[comp][2] || k > L.hi[comp][2])
                                continue;
                            const int i0 = std::max(P.dlo[0] + tx * BX, L.lo[comp][0]);
                            const int i1 = std::min(P.dlo[0] + tx * BX + BX - 1, L.hi[comp][0]);
                            int most = 0, n = 0;
                            for (int i = i0; i <= i1; ++i) {
                                const int id = M.ids[c][L.index(i, j, k)];
                                const int np = id == 255 ? M.dense[c].npoles : M.table.npoles[id];
                                n += np;
                                most = std::max(most, np);
                            }
                            run += n;
                            P.full_bytes += double(most) * BX * double(tsz);
                            // a segment's entries are contiguous: whole lines but a partial one
                            P.compact_bytes += n ? double(n) * double(tsz) + 64.0 : 0.0;
                        }
        if (run * int64_t(tsz) >= (int64_t(1) << 32))
            throw std::runtime_error("polarization state of one component exceeds 4 GiB");
        P.comp_start[c] = P.total;
        P.total += run;
    }
    P.compact = P.full_bytes > 1.5 * P.compact_bytes;
    return P;
}

void EngineImpl::build()
{
    // FFIP_TIMING=1 reports where the device set-up spends its time on stderr
    static const bool report = std::getenv("FFIP_TIMING") != nullptr;
    auto last = std::chrono::steady_clock::now();
    auto mark = [&](const char* what) {
        if (!report)
            return;
        auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[ffip]   %s %.2f ms\n", what, std::chrono::duration<double, std::milli>(now - last).count());
        last = now;
    };
    Setup& s = *S;
    const Layout& L = s.L;
    fp32 = s.precision == "fp32";
    tsz = fp32 ? 4 : 8;
    // measured on MI355X (256^3 dispersive, 512^3 voxel-mix and 200x200x100
    // design workloads, FP32 and FP64): 8-plane z segments; FP32 kernels issue
    // two planes per load batch (one when a point carries more than two poles,
    // except in the CPML-free kernel), FP64 kernels one (512^3 FP64 E half step
    // 6151 -> 5630 us against two for the pole kernels).  Material-table kernels
    // without poles take one plane per batch in both precisions: at two, the FP32
    // design-box electric kernel needs 102 VGPRs (4 waves per SIMD) instead of 75
    // (6), and the 200x200x100 design gradient's electric half step took
    // 49.2 us per step instead of 44.2.
    // Segments of 12 planes on grids that still give at least 24576 tiles (96 per
    // CU): fewer planes are re-read at segment ends.  Measured: 512^3 voxel mix
    // 3.837 -> 3.775 ms per step (10 planes 3.813, 14 planes 3.775), 640^3 with
    // 8 x 64-frequency planes 7.467 -> 7.394 ms; on the 256^3 spheres (6534
    // tiles at 12 planes) 0.391 -> 0.406 ms, and the 200x200x100 design grid is
    // best at 8.
    lz = 8;
    {
        const int64_t tiles12 = int64_t((L.n[0] + BX - 1) / BX) * ((L.n[1] + BY - 1) / BY) * ((L.n[2] + 11) / 12);
        if (tiles12 >= 24576)
            lz = 12;
    }
    unroll = fp32 ? 2 : 1;
    unroll_mat = fp32 ? 2 : 1;
    unroll_shell = fp32 ? 0 : 1;
    unroll_plain = 1;
    // tuning overrides: z segment per workgroup, planes per load batch
    if (const char* e = std::getenv("FFIP_LZ"))
        lz = std::max(1, std::atoi(e));
    if (const char* e = std::getenv("FFIP_UNROLL"))
        unroll = std::atoi(e);
    if (const char* e = std::getenv("FFIP_UNROLL_MAT"))
        unroll_mat = std::atoi(e);
    if (const char* e = std::getenv("FFIP_UNROLL_SHELL"))
        unroll_shell = std::atoi(e);
    if (const char* e = std::getenv("FFIP_UNROLL_PLAIN"))
        unroll_plain = std::atoi(e);
    // the half-step kernel addresses every per-cell array with a 32-bit byte offset
    if (L.alloc * int64_t(tsz) >= (int64_t(1) << 32))
        throw std::runtime_error("grid too large: one field array must stay below 4 GiB (" +
                                 std::to_string(L.alloc) + " cells)");
    HIP_OK(hipGetDevice(&device));
    HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    // zeroed stand-in for absent dense-medium arrays, which the kernels read at
    // element 0 and ignore (see load_sparse); absent CPML auxiliaries point there too
    void* stub = dalloc<double>(32);
    for (int c = 0; c < 6; ++c) {
        field[c] = fp32 ? (void*)dalloc<float>(L.alloc) : (void*)dalloc<double>(L.alloc);
    }
    for (int h = 0; h < 2; ++h) {
        HalfMaterial& M = s.mat[h];
        S_slots[h] = M.uniform ? 0 : M.max_poles;
        if (S_slots[h] > MAXS)
            throw std::runtime_error("more than 8 polarization poles at one point");
        if (S_slots[h] > 0) {
            PolLayoutPlan P = plan_polarization(s, h, lz);
            pcompact[h] = P.compact;
            size_t cnt;
            if (P.compact) {
                for (int c = 0; c < 3; ++c) {
                    pcomp[h][c] = P.comp_start[c];
                    segbase[h][c] = upload(P.segbase[c]);
                }
                pseg_ny[h] = P.ny;
                cnt = size_t(P.total);
            } else {
                cnt = size_t(S_slots[h]) * 3 * size_t(L.alloc);
            }
            Pa[h] = fp32 ? (void*)dalloc<float>(cnt) : (void*)dalloc<double>(cnt);
            Pb[h] = fp32 ? (void*)dalloc<float>(cnt) : (void*)dalloc<double>(cnt);
        }
        mark("fields and polarization");
        if (!M.uniform) {
            const size_t na = size_t(L.alloc);
            std::unique_ptr<uint32_t[]> packed(new uint32_t[na]);
            const uint8_t *i0 = M.ids[0].data(), *i1 = M.ids[1].data(), *i2 = M.ids[2].data();
            parallel_planes(0, int(L.nplanes) - 1, [&](int p) {
                for (size_t q = size_t(p) * L.plane, e = q + size_t(L.plane); q < e; ++q)
                    packed[q] = uint32_t(i0[q]) | uint32_t(i1[q]) << 8 | uint32_t(i2[q]) << 16;
            });
            ids[h] = upload(packed.get(), na);
            std::vector<MatEntry> me(M.table.eps.size());
            for (size_t e = 0; e < me.size(); ++e)
                me[e] = {M.table.inv_eps[e], M.table.npoles[e], M.table.pole0[e], M.table.eps[e], 0.0};
            nmats[h] = int(me.size());
            mats[h] = upload(me);
            std::vector<double> cf;
            for (size_t p = 0; p < M.table.c1.size(); ++p) {
                cf.push_back(M.table.c1[p]);
                cf.push_back(M.table.c2[p]);
                cf.push_back(M.table.sc3[p]);
                cf.push_back(0.0);
            }
            npoles[h] = int(M.table.c1.size());
            p_coef[h] = upload(cf);
            if (size_t(nmats[h] + npoles[h]) * 4 * sizeof(double) > 65536)
                throw std::runtime_error("material table too large for LDS");
            for (int c = 0; c < 3; ++c) {
                DenseRegion& D = M.dense[c];
                d_inv[h][c] = static_cast<double*>(stub);
                if (!D.present)
                    continue;
                d_inv[h][c] = upload(D.inv_eps);
                d_eps[h][c] = upload(D.eps);
                d_amp[h][c] = upload(D.amp);
                std::vector<double> cc;
                for (int q = 0; q < D.npoles; ++q) {
                    cc.push_back(D.c1[q]);
                    cc.push_back(D.c2[q]);
                    cc.push_back(D.c3[q]);
                }
                d_c[h][c] = upload(cc);
            }
        }
        mark("materials");
        // CPML: E points use even-parity tables, H points odd-parity ones
        const int par = h == 0 ? 0 : 1;
        for (int a = 0; a < 3; ++a) {
            PmlAxis& P = s.pml[a];
            PmlDev& D = pml[h][a];
            D.W = P.W[par];
            D.psi1 = D.psi2 = stub;
            if (D.W == 0)
                continue;
            D.smap = upload(P.smap[par]);
            D.b = upload(P.b[par]);
            D.c = upload(P.c[par]);
            D.ik = upload(P.ik[par]);
            // [n_z][W][pitch] (y slab), [W][n_y][pitch] (z slab), [n_z][n_y][W] (x slab)
            size_t cnt = size_t(D.W);
            for (int b = 0; b < 3; ++b)
                if (b != a)
                    cnt *= b == 0 ? size_t(L.pitch) : size_t(L.n[b]);
            D.psi1 = fp32 ? (void*)dalloc<float>(cnt) : (void*)dalloc<double>(cnt);
            D.psi2 = fp32 ? (void*)dalloc<float>(cnt) : (void*)dalloc<double>(cnt);
        }
        mark("cpml");
        // DFT units
        std::vector<double> fl;
        for (auto& U : s.units[h]) {
            DftDev d{};
            d.comp = U.comp % 3;
            d.is_db = U.is_db ? 1 : 0;
            d.nf = int(U.freqs.size());
            d.mf0 = int(fl.size());
            for (double f : U.freqs)
                fl.push_back(f);
            if (U.has_local) {
                for (int a = 0; a < 3; ++a) {
                    d.box[a][0] = U.box[a][0];
                    d.box[a][1] = U.box[a][1];
                    d.dims[a] = U.dims[a];
                }
                d.npts = U.npts();
            } else {
                for (int a = 0; a < 3; ++a) {
                    d.box[a][0] = 1;
                    d.box[a][1] = 0;
                    d.dims[a] = 0;
                }
                d.npts = 0;
            }
            // E/H samples are gathered after the step into the unit's ring and
            // added into the sums in batches: inside the z-marching kernel each
            // monitored plane would add a dependent read-modify-write round trip
            // per component.  D/B samples need the polarization sums, so those
            // stay in the half step.
            d.fused = d.is_db ? 1 : 0;
            if (d.npts >= (int64_t(1) << 31))
                throw std::runtime_error("DFT monitor with 2^31 or more points");
            d.re = dalloc<double>(size_t(d.npts) * d.nf);
            d.im = dalloc<double>(size_t(d.npts) * d.nf);
            if (!d.fused && d.npts > 0) {
                d.ring = fp32 ? (void*)dalloc<float>(size_t(d.npts) * DFT_BATCH, false)
                              : (void*)dalloc<double>(size_t(d.npts) * DFT_BATCH, false);
                ring_units.emplace_back(h, int(units[h].size()));
            }
            units[h].push_back(d);
        }
        if (units[h].size() > size_t(MAXU))
            throw std::runtime_error("more than 16 DFT units on one half step");
        for (auto& U : units[h])
            monitors[h] = monitors[h] || (U.fused && U.npts > 0);
        MF[h] = int(fl.size());
        freqs_dev[h] = upload(fl);
        ph[h] = dalloc<double>(size_t(std::max(1, MF[h])) * 2 * PH_CHUNK);
        mark("dft units");
        // sources: group dipoles per target point (insertion order kept)
        std::vector<std::pair<int64_t, int>> key;  // (index, comp)
        std::vector<std::vector<int>> members;
        std::unordered_map<int64_t, size_t> where;
        for (size_t q = 0; q < s.dip[h].size(); ++q) {
            const Dipole& D = s.dip[h][q];
            int64_t idx = L.index(D.i, D.j, D.k) - L.origin;
            int64_t hk = idx * 4 + D.comp % 3;
            auto it = where.find(hk);
            size_t t;
            if (it == where.end()) {
                t = key.size();
                where.emplace(hk, t);
                key.emplace_back(idx, D.comp % 3);
                members.emplace_back();
            } else
                t = it->second;
            members[t].push_back(int(q));
        }
        Src& SR = src[h];
        SR.ntarget = int(key.size());
        if (SR.ntarget > 0) {
            std::vector<int64_t> vidx;
            std::vector<int> vcomp, eoff{0}, moff{0}, munit;
            std::vector<double> scale, amp, st, en, wd;
            std::vector<int64_t> mofs;
            for (size_t t = 0; t < key.size(); ++t) {
                vidx.push_back(key[t].first);
                vcomp.push_back(key[t].second);
                const Dipole& D0 = s.dip[h][members[t][0]];
                // point's g_inf
                double g;
                if (M.uniform)
                    g = M.u_eps;
                else {
                    int id = M.ids[key[t].second][key[t].first + L.origin];
                    if (id != 255)
                        g = M.table.eps[id];
                    else {
                        DenseRegion& DR = M.dense[key[t].second];
                        int64_t d0 = DR.box[0][1] - DR.box[0][0] + 1, d1 = DR.box[1][1] - DR.box[1][0] + 1;
                        g = DR.eps[(int64_t(D0.k - DR.box[2][0]) * d1 + (D0.j - DR.box[1][0])) * d0 +
                                   (D0.i - DR.box[0][0])];
                    }
                }
                scale.push_back(s.dt / g);
                for (int q : members[t]) {
                    const Dipole& D = s.dip[h][q];
                    amp.push_back(D.amp);
                    st.push_back(D.start);
                    en.push_back(D.end);
                    wd.push_back(D.width);
                    SR.max_end = std::max(SR.max_end, D.end);
                }
                eoff.push_back(int(amp.size()));
                for (size_t u = 0; u < s.units[h].size(); ++u) {
                    const DftUnit& U = s.units[h][u];
                    if (!U.has_local || U.comp % 3 != key[t].second || !units[h][u].fused)
                        continue;
                    if (D0.i < U.box[0][0] || D0.i > U.box[0][1] || D0.j < U.box[1][0] || D0.j > U.box[1][1] ||
                        D0.k < U.box[2][0] || D0.k > U.box[2][1])
                        continue;
                    munit.push_back(int(u));
                    mofs.push_back((D0.i - U.box[0][0]) +
                                   int64_t(U.dims[0]) * ((D0.j - U.box[1][0]) + int64_t(U.dims[1]) * (D0.k - U.box[2][0])));
                }
                moff.push_back(int(munit.size()));
            }
            SR.idx = upload(vidx);
            SR.comp = upload(vcomp);
            SR.scale = upload(scale);
            SR.eoff = upload(eoff);
            SR.amp = upload(amp);
            SR.start = upload(st);
            SR.end = upload(en);
            SR.width = upload(wd);
            SR.moff = upload(moff);
            SR.munit = upload(munit);
            SR.mofs = upload(mofs);
        }
        mark("sources");
        // faces
        for (auto& F : s.faces[h]) {
            FaceDev d{};
            d.comp = F.comp;
            d.axis = F.axis;
            d.dst = F.dst;
            d.src = F.src;
            d.mult = F.mult;
            int a0 = (F.axis + 1) % 3, a1 = (F.axis + 2) % 3;
            d.t0lo = L.lo[F.comp][a0];
            d.t0hi = L.hi[F.comp][a0];
            d.t1lo = L.lo[F.comp][a1];
            d.t1hi = L.hi[F.comp][a1];
            if (d.t0hi >= d.t0lo && d.t1hi >= d.t1lo)
                faces[h].push_back(d);
        }
        if (faces[h].size() > size_t(MAXF))
            throw std::runtime_error("too many boundary faces");
    }
    mark("faces");
    // algorithmic traffic per half-step launch
    for (int h = 0; h < 2; ++h) {
        const HalfMaterial& M = s.mat[h];
        double nreal = 0, nsrc = 0, ade = 0, psi = 0, dftb = 0;
        for (int c = 0; c < 3; ++c) {
            nreal += double(L.count(h * 3 + c));
            nsrc += double(L.count((1 - h) * 3 + c));
        }
        double mcells = 0;  // cells whose material word is read (inside M.mbox)
        if (!M.uniform) {
            int lo[3], hi[3];
            mcells = 1;
            for (int a = 0; a < 3; ++a) {
                lo[a] = M.mbox[a][0];
                hi[a] = M.mbox[a][1];
                int dlo = 1 << 30, dhi = -(1 << 30);
                for (int c = 0; c < 3; ++c) {
                    dlo = std::min(dlo, L.lo[h * 3 + c][a]);
                    dhi = std::max(dhi, L.hi[h * 3 + c][a]);
                }
                mcells *= std::max(0, std::min(hi[a], dhi) - std::max(lo[a], dlo) + 1);
            }
            const int bg_np = M.table.npoles[M.bg_id];
            for (int c = 0; c < 3; ++c) {
                const int comp = h * 3 + c;
                // points outside the box carry the default medium's poles
                double inside = 0;
                for (int k = std::max(lo[2], L.lo[comp][2]); k <= std::min(hi[2], L.hi[comp][2]); ++k)
                    for (int j = std::max(lo[1], L.lo[comp][1]); j <= std::min(hi[1], L.hi[comp][1]); ++j)
                        for (int i = std::max(lo[0], L.lo[comp][0]); i <= std::min(hi[0], L.hi[comp][0]); ++i) {
                            int id = M.ids[c][L.index(i, j, k)];
                            int np = id == 255 ? M.dense[c].npoles : M.table.npoles[id];
                            ade += 3.0 * np;  // read P, read P-, write P+
                            inside += 1;
                        }
                ade += 3.0 * bg_np * (double(L.count(comp)) - inside);
            }
        }
        const int par = h == 0 ? 0 : 1;
        for (int c = 0; c < 3; ++c) {
            const int comp = h * 3 + c;
            for (int a : {(c + 1) % 3, (c + 2) % 3}) {
                const std::vector<int>& sm = s.pml[a].smap[par];
                double inslab = 0;
                for (int q = L.lo[comp][a]; q <= L.hi[comp][a]; ++q)
                    inslab += sm[q] >= 0 ? 1 : 0;
                double other = 1;
                for (int b = 0; b < 3; ++b)
                    if (b != a)
                        other *= std::max(0, L.hi[comp][b] - L.lo[comp][b] + 1);
                psi += 2.0 * inslab * other;  // read + write psi
            }
        }
        for (auto& U : units[h])
            if (U.fused)  // the others go through the sample ring
                dftb += double(U.npts) * U.nf * 32.0;  // re, im: read + write, FP64
        double* t = traffic[h];
        t[0] = (2 * nreal + nsrc) * double(tsz);
        t[1] = 4.0 * mcells;  // one packed 32-bit id word per cell of the material box
        t[2] = ade * double(tsz);
        t[3] = psi * double(tsz);
        t[4] = dftb;
        t[5] = t[0] + t[1] + t[2] + t[3] + t[4];
        t[6] = nreal;
    }
    mark("traffic model");
    // probes
    nprobes = int(s.probes.size());
    if (nprobes) {
        std::vector<int> off{0}, comp;
        std::vector<int64_t> idx;
        std::vector<double> w;
        for (auto& P : s.probes) {
            for (auto& t : P) {
                comp.push_back(t.comp);
                idx.push_back(t.index - L.origin);
                w.push_back(t.w);
            }
            off.push_back(int(idx.size()));
        }
        pr_off = upload(off);
        pr_comp = upload(comp);
        pr_idx = upload(idx);
        pr_w = upload(w);
    }
    mark("probes");
    if (fp32)
        build_tiles<float>();
    else
        build_tiles<double>();
    mark("tile lists");
    // fused full-step kernel: only on request (FFIP_FUSED=1) and where the
    // simulation allows it.  Measured on MI355X it is slower than the two
    // half-step kernels: 512^3 voxel mix 4.26 vs 4.09 ms per step (fused
    // interior 2.76 ms, CPML shell on the half-step kernels 1.50 ms), 256^3
    // dispersive 0.425 vs 0.379 ms, design gradient 0.735 vs 0.547 s (FP32).  It
    // moves a third fewer field bytes, but every plane waits for its
    // polarization loads and a workgroup barrier, and the shell's four small
    // launches run at under 3 TB/s.
    {
        const char* fe = std::getenv("FFIP_FUSED");
        fused = fe && std::atoi(fe) != 0 && fused_eligible(nullptr);
        if (const char* d = std::getenv("FFIP_FDEPTH"))
            fdepth = std::min(2, std::max(1, std::atoi(d)));
    }
    if (fused) {
        for (int c = 0; c < 6; ++c)
            field_b[c] = fp32 ? (void*)dalloc<float>(L.alloc) : (void*)dalloc<double>(L.alloc);
        // z segment per workgroup: the longest of 32, 16, 8 planes that still gives
        // several rounds of workgroups over the CUs (FFIP_LZF overrides)
        {
            int cols = 1, nz = 0;
            for (int a = 0; a < 2; ++a) {
                int lo = 1 << 30, hi = -1;
                for (int c = 0; c < 6; ++c)
                    if (L.hi[c][a] >= L.lo[c][a]) {
                        lo = std::min(lo, L.lo[c][a]);
                        hi = std::max(hi, L.hi[c][a]);
                    }
                cols *= (hi - lo + (a == 0 ? BX : BY)) / (a == 0 ? BX : BY);
            }
            {
                int lo = 1 << 30, hi = -1;
                for (int c = 0; c < 6; ++c)
                    if (L.hi[c][2] >= L.lo[c][2]) {
                        lo = std::min(lo, L.lo[c][2]);
                        hi = std::max(hi, L.hi[c][2]);
                    }
                nz = hi - lo + 1;
            }
            // a multiple of the half-step kernels' segment, so that their shell
            // tiles never straddle a fused segment
            lzf = lz;
            for (int m : {4, 2})
                if (int64_t(cols) * ((nz + m * lz - 1) / (m * lz)) >= 4096) {
                    lzf = m * lz;
                    break;
                }
            if (const char* e = std::getenv("FFIP_LZF"))
                lzf = std::max(lz, std::atoi(e) / lz * lz);
        }
        if (fp32)
            build_fused_tiles<float>();
        else
            build_fused_tiles<double>();
        mark("fused step");
    }
    HIP_OK(hipStreamSynchronize(stream));
    mark("copies and fills");
}

template <typename T>
HalfArgs<T> EngineImpl::half_args(int h)
{
    const Setup& s = *S;
    const Layout& L = s.L;
    HalfArgs<T> A;
    std::memset(&A, 0, sizeof(A));
    for (int c = 0; c < 3; ++c) {
        A.u[c] = static_cast<const T*>(field[h * 3 + c]);
        A.w[c] = static_cast<T*>(field[h * 3 + c]);
        A.s[c] = static_cast<const T*>(field[(1 - h) * 3 + c]);
    }
    A.pitch = int(L.pitch);
    A.tail0 = int(std::min<int64_t>(L.tail0, Layout::NO_TAIL));
    A.plane = int(L.plane);
    for (int a = 0; a < 3; ++a) {
        A.n[a] = L.n[a];
        A.dlo[a] = 1 << 30;
        A.dhi[a] = -1;
    }
    for (int c = 0; c < 3; ++c)
        for (int a = 0; a < 3; ++a) {
            A.lo[c][a] = L.lo[h * 3 + c][a];
            A.hi[c][a] = L.hi[h * 3 + c][a];
            if (A.hi[c][a] >= A.lo[c][a]) {
                A.dlo[a] = std::min(A.dlo[a], A.lo[c][a]);
                A.dhi[a] = std::max(A.dhi[a], A.hi[c][a]);
            }
        }
    A.inv_dx = 1 / s.dx;
    A.dt = s.dt;
    const HalfMaterial& M = s.mat[h];
    A.u_inv = M.u_inv;
    A.u_eps = M.u_eps;
    if (!M.uniform) {
        A.ids = ids[h];
        for (int a = 0; a < 3; ++a) {
            A.mbox[a][0] = M.mbox[a][0];
            A.mbox[a][1] = M.mbox[a][1];
        }
        A.bg_word = uint32_t(M.bg_id) * 0x010101u;
        A.mats = mats[h];
        A.nmats = nmats[h];
        A.coef = p_coef[h];
        A.npoles = npoles[h];
        for (int c = 0; c < 3; ++c) {
            const DenseRegion& D = M.dense[c];
            A.d_inv[c] = d_inv[h][c];  // the zeroed stand-in when absent
            if (!D.present)
                continue;
            for (int a = 0; a < 3; ++a) {
                A.d_box[c][a][0] = D.box[a][0];
                A.d_box[c][a][1] = D.box[a][1];
            }
            A.d_np[c] = D.npoles;
            A.d_inv[c] = d_inv[h][c];
            A.d_eps[c] = d_eps[h][c];
            A.d_amp[c] = d_amp[h][c];
            A.d_c[c] = d_c[h][c];
        }
    }
    A.origin = int(L.origin);
    A.pcur = static_cast<T*>(Pa[h]);
    A.pprev = static_cast<T*>(Pb[h]);
    A.pslice = L.alloc;
    for (int c = 0; c < 3; ++c) {
        A.pcomp[c] = pcomp[h][c];
        A.segbase[c] = segbase[h][c];
    }
    A.seg_ny = pseg_ny[h];
    for (int a = 0; a < 3; ++a)
        A.pml[a] = pml[h][a];
    A.nunits = int(units[h].size());
    for (int c = 0; c < 3; ++c)
        for (int a = 0; a < 3; ++a) {
            A.fbox[c][a][0] = 1 << 30;
            A.fbox[c][a][1] = -(1 << 30);
        }
    for (int u = 0; u < A.nunits; ++u) {
        const DftDev& d = units[h][u];
        A.units[u] = d;
        if (!d.fused || d.npts == 0)
            continue;
        const int c = d.comp;
        A.fu[c][A.nfu[c]++] = u;
        for (int a = 0; a < 3; ++a) {
            A.fbox[c][a][0] = std::min(A.fbox[c][a][0], d.box[a][0]);
            A.fbox[c][a][1] = std::max(A.fbox[c][a][1], d.box[a][1]);
        }
    }
    A.ph = ph[h] + size_t(n - ph_start) * 2 * std::max(1, MF[h]);
    // tile grid, padded to a multiple of the 8 XCDs
    A.lz = lz;
    if (A.dhi[0] >= A.dlo[0] && A.dhi[1] >= A.dlo[1] && A.dhi[2] >= A.dlo[2]) {
        A.gx = (A.dhi[0] - A.dlo[0] + BX) / BX;
        A.gy = (A.dhi[1] - A.dlo[1] + BY) / BY;
        A.gz = (A.dhi[2] - A.dlo[2] + lz) / lz;
        A.nblocks = ((A.gx * A.gy * A.gz + 7) / 8) * 8;
    }
    return A;
}

void EngineImpl::ensure_phasors()
{
    if (ph_start >= 0 && n < ph_start + PH_CHUNK)
        return;
    flush_dft();  // the pending samples' phasors are rows of the current table
    ph_start = n;
    for (int h = 0; h < 2; ++h) {
        if (MF[h] == 0)
            continue;
        int total = MF[h] * PH_CHUNK;
        // H units sample at (n + 1/2) dt, E units at (n + 1) dt (src/simulation.cpp:748-766)
        k_phasors<<<(total + 255) / 256, 256, 0, stream>>>(ph[h], freqs_dev[h], MF[h], ph_start, PH_CHUNK,
                                                         h == 0 ? 1.0 : 0.5, S->dt);
        HIP_OK(hipGetLastError());
    }
}

template <typename T>
void EngineImpl::build_tiles()
{
    const Setup& s = *S;
    for (int h = 0; h < 2; ++h) {
        for (int c = 0; c < 3; ++c)
            dense[h] = dense[h] || s.mat[h].dense[c].present;
        HalfArgs<T> A = half_args<T>(h);
        if (A.dhi[0] < A.dlo[0] || A.dhi[1] < A.dlo[1] || A.dhi[2] < A.dlo[2])
            continue;
        // per axis: does tile q reach a CPML slab (of this half's parity)?
        const int par = h == 0 ? 0 : 1, step[3] = {BX, BY, A.lz}, n[3] = {A.gx, A.gy, A.gz};
        std::vector<char> touch[3];
        for (int a = 0; a < 3; ++a) {
            touch[a].assign(size_t(n[a]), 0);
            const PmlAxis& P = s.pml[a];
            if (P.W[par] == 0)
                continue;
            for (int q = 0; q < n[a]; ++q)
                for (int i = A.dlo[a] + q * step[a]; i <= std::min(A.dlo[a] + (q + 1) * step[a] - 1, A.dhi[a]); ++i)
                    if (P.smap[par][i] >= 0)
                        touch[a][q] = 1;
        }
        std::vector<uint32_t> in, sh;
        for (int tz = 0; tz < A.gz; ++tz)
            for (int ty = 0; ty < A.gy; ++ty)
                for (int tx = 0; tx < A.gx; ++tx) {
                    const uint32_t t = uint32_t(tx + A.gx * (ty + A.gy * tz));
                    (touch[0][tx] || touch[1][ty] || touch[2][tz] ? sh : in).push_back(t);
                }
        // Two launches pay only where the CPML-free kernel gains occupancy and
        // each launch still fills the GPU several times over (a 200x200x100 grid
        // of 2275 tiles lost 10 %).  Measured on MI355X (electric half step, one
        // launch -> two): FP32 with full-grid polarizations (256^3 dispersive,
        // 74 instead of 109 VGPRs) 240 -> 234 us; FP32 compact (512^3 voxel mix,
        // 112 and 109 VGPRs: no gain) 2824 -> 2911 us and 640^3 5362 -> 5472 us;
        // FP64 full grid 414 -> 431 us; FP64 compact 5633 -> 5583 us.  Measured
        // again after the compact kernels went to one plane per batch (75 VGPRs
        // without the CPML paths, 109 with them): 512^3 2747 -> 2757 us, 640^3
        // 5230 -> 5206 us.  The split is kept where it gains more than 2 %: FP32
        // with full-grid polarizations.
        // FFIP_CPML_SPLIT=1 / 0 forces / forbids the split wherever both lists are non-empty
        const char* force = std::getenv("FFIP_CPML_SPLIT");
        bool split = !s.mat[h].uniform && in.size() + sh.size() >= 4096 && fp32 && !pcompact[h];
        if (force)
            split = std::atoi(force) != 0;
        split = split && !sh.empty() && !in.empty();
        if (split) {
            pml_a[h] = false;
        } else if (sh.empty()) {
            pml_a[h] = false;  // no CPML reaches this half's tiles
        } else {
            pml_a[h] = true;
            in.insert(in.end(), sh.begin(), sh.end());
            std::sort(in.begin(), in.end());
            sh.clear();
        }
        ntiles_a[h] = int(in.size());
        ntiles_b[h] = int(sh.size());
        tiles_a[h] = upload(in);
        tiles_b[h] = upload(sh);
    }
}

template <typename T>
void EngineImpl::launch_tiles(int h, const HalfArgs<T>& A, const uint32_t* list, int count, bool pml)
{
    if (count == 0)
        return;
    const HalfMaterial& M = S->mat[h];
    const bool mat = !M.uniform;
    const size_t lds = mat ? size_t(A.nmats + A.npoles) * 4 * sizeof(double) : 0;
    // pole kernels on the CPML tiles of a split half step take unroll_shell planes per batch
    const bool split = fused || !pml_a[h];
    const int u = pml && split && unroll_shell > 0 && mat && S_slots[h] > 0
                      ? unroll_shell
                      : (mat ? (S_slots[h] > 0 ? unroll_mat : unroll_plain) : unroll);
    HalfArgs<T> X = A;
    X.tiles = list;
    X.ntiles = count;
    X.nblocks = ((count + 7) / 8) * 8;
    const dim3 grid(X.nblocks, 1, 1), block(BX, BY, 1);
    if (monitors[h]) {
        // flux-density monitors accumulate inside the kernel: the general form
        if (h == 0)
            launch_general<T, true>(mat, lds, grid, block, X);
        else
            launch_general<T, false>(mat, lds, grid, block, X);
    } else if (h == 0) {
        launch_k<T, true>(mat, S_slots[h], pcompact[h], u, pml, dense[h], lds, grid, block, X);
    } else {
        launch_k<T, false>(mat, S_slots[h], pcompact[h], u, pml, dense[h], lds, grid, block, X);
    }
    HIP_OK(hipGetLastError());
}

template <typename T>
void EngineImpl::launch_half(int h, hipEvent_t e0, hipEvent_t e1)
{
    HalfArgs<T> A = half_args<T>(h);
    if (A.dhi[0] < A.dlo[0] || A.dhi[1] < A.dlo[1] || A.dhi[2] < A.dlo[2])
        return;
    // tiles clear of the CPML without its paths (when split), then the others
    if (e0)
        HIP_OK(hipEventRecord(e0, stream));
    launch_tiles<T>(h, A, tiles_a[h], ntiles_a[h], pml_a[h]);
    launch_tiles<T>(h, A, tiles_b[h], ntiles_b[h], true);
    if (e1)
        HIP_OK(hipEventRecord(e1, stream));
    if (S_slots[h] > 0)
        std::swap(Pa[h], Pb[h]);
    post_half<T>(h);
}

// Point sources, E/H DFT monitors and ghost faces of half h, after its kernels
// (they read and write the current field buffers)
template <typename T>
void EngineImpl::post_half(int h)
{
    const double* phr = ph[h] + size_t(n - ph_start) * 2 * std::max(1, MF[h]);
    // sources of this half at its source time (src/simulation.cpp:748-766)
    const double t = h == 1 ? double(n) * S->dt : (double(n) + 0.5) * S->dt;
    Src& SR = src[h];
    if (SR.ntarget > 0 && !(t > SR.max_end)) {
        SrcArgs<T> B;
        std::memset(&B, 0, sizeof(B));
        B.ntarget = SR.ntarget;
        for (int c = 0; c < 3; ++c)
            B.u[c] = static_cast<T*>(field[h * 3 + c]) + S->L.origin;
        B.idx = SR.idx;
        B.comp = SR.comp;
        B.scale = SR.scale;
        B.eoff = SR.eoff;
        B.amp = SR.amp;
        B.start = SR.start;
        B.end = SR.end;
        B.width = SR.width;
        B.moff = SR.moff;
        B.munit = SR.munit;
        B.mofs = SR.mofs;
        B.dt = S->dt;
        for (size_t u = 0; u < units[h].size(); ++u)
            B.units[u] = units[h][u];
        B.ph = phr;
        k_sources<T><<<(SR.ntarget + 255) / 256, 256, 0, stream>>>(B, t);
        HIP_OK(hipGetLastError());
    }
    if (!faces[h].empty()) {
        FaceArgs<T> F;
        std::memset(&F, 0, sizeof(F));
        for (int c = 0; c < 3; ++c)
            F.u[c] = static_cast<T*>(field[h * 3 + c]) + S->L.origin;
        F.pitch = S->L.pitch;
        F.tail0 = S->L.tail0;
        F.plane = S->L.plane;
        F.nf = int(faces[h].size());
        int m0 = 1, m1 = 1;
        for (int q = 0; q < F.nf; ++q) {
            F.f[q] = faces[h][q];
            m0 = std::max(m0, F.f[q].t0hi - F.f[q].t0lo + 1);
            m1 = std::max(m1, F.f[q].t1hi - F.f[q].t1lo + 1);
        }
        k_faces<T><<<dim3((m0 + 255) / 256, m1, F.nf), 256, 0, stream>>>(F);
        HIP_OK(hipGetLastError());
    }
}

// Whether the fused full-step kernel can run this simulation: the magnetic half
// must be one pole-free medium with no point sources, flux-density monitors or
// ghost faces (the kernel forms magnetic values of neighbouring tiles itself and
// runs no magnetic source or face pass between the halves).
bool EngineImpl::fused_eligible(std::string* why) const
{
    auto no = [&](const char* w) {
        if (why)
            *why = w;
        return false;
    };
    const Setup& s = *S;
    if (!s.mat[1].uniform)
        return no("magnetic media");
    if (src[1].ntarget > 0)
        return no("magnetic sources");
    if (!faces[0].empty() || !faces[1].empty())
        return no("periodic or symmetry faces");
    if (monitors[0] || monitors[1])
        return no("flux-density monitors");
    if (S_slots[0] > 3)
        return no("more than three poles at a point");
    if (dense[0] && S_slots[0] > 0)
        return no("per-point media with poles");
    const Layout& L = s.L;
    for (int a = 0; a < 3; ++a) {
        int lo[2] = {1 << 30, 1 << 30}, hi[2] = {-1, -1};
        for (int h = 0; h < 2; ++h)
            for (int c = 0; c < 3; ++c)
                if (L.hi[h * 3 + c][a] >= L.lo[h * 3 + c][a]) {
                    lo[h] = std::min(lo[h], L.lo[h * 3 + c][a]);
                    hi[h] = std::max(hi[h], L.hi[h * 3 + c][a]);
                }
        if (lo[0] != lo[1] || hi[0] != hi[1] || hi[0] < lo[0])
            return no("electric and magnetic lattices cover different ranges");
    }
    return true;
}

template <typename T>
StepArgs<T> EngineImpl::step_args()
{
    StepArgs<T> S0;
    std::memset(&S0, 0, sizeof(S0));
    S0.e = half_args<T>(0);
    HalfArgs<T>& A = S0.e;
    for (int c = 0; c < 3; ++c) {
        A.w[c] = static_cast<T*>(field_b[c]);
        A.s[c] = nullptr;
        S0.hin[c] = static_cast<const T*>(field[3 + c]);
        S0.hout[c] = static_cast<T*>(field_b[3 + c]);
        for (int a = 0; a < 3; ++a) {
            S0.hlo[c][a] = S->L.lo[3 + c][a];
            S0.hhi[c][a] = S->L.hi[3 + c][a];
        }
    }
    S0.h_inv = S->mat[1].u_inv;
    A.lz = lzf;
    for (int a = 0; a < 3; ++a) {
        S0.fo[a] = fbox[a][0];
        S0.fe[a] = fbox[a][1];
        S0.fn[a] = fcount[a];
    }
    return S0;
}

// The fused kernel's box: per axis, the longest run of whole tiles (BX, BY, lz
// points, counted from the half step's first index) whose electric and magnetic
// points all lie clear of the CPML slabs.  The fused kernel tiles it with
// segments of lzf planes along z; the half-step kernels cover the rest of the
// grid (the CPML shell) with their own tiles, listed per half as clear of that
// half's slabs (stiles[h][0]) or not (stiles[h][1]).
template <typename T>
void EngineImpl::build_fused_tiles()
{
    const Setup& s = *S;
    const Layout& L = s.L;
    const HalfArgs<T> A = half_args<T>(0);
    const int step[3] = {BX, BY, A.lz};
    for (int a = 0; a < 3; ++a) {
        const PmlAxis& P = s.pml[a];
        auto clear = [&](int x) {
            return (P.W[0] == 0 || P.smap[0][x] < 0) && (P.W[1] == 0 || P.smap[1][x] < 0);
        };
        // longest run of clear indices
        int best0 = 0, best1 = -1;
        for (int x = A.dlo[a]; x <= A.dhi[a];) {
            if (!clear(x)) {
                ++x;
                continue;
            }
            int y = x;
            while (y + 1 <= A.dhi[a] && clear(y + 1))
                ++y;
            if (y - x > best1 - best0) {
                best0 = x;
                best1 = y;
            }
            x = y + 1;
        }
        // the fused kernel addresses rows without the packed last line: its box and
        // the one-point rims around it stay below the packed x indices
        if (a == 0 && L.tail0 != Layout::NO_TAIL)
            best1 = std::min(best1, L.tail0 - 2);
        // round inwards to whole tiles (a run that reaches the grid's end keeps it)
        const int first = A.dlo[a] + (best0 - A.dlo[a] + step[a] - 1) / step[a] * step[a];
        const int last = best1 == A.dhi[a] ? A.dhi[a] : A.dlo[a] + (best1 + 1 - A.dlo[a]) / step[a] * step[a] - 1;
        fbox[a][0] = first;
        fbox[a][1] = last;
    }
    const bool empty = fbox[0][1] < fbox[0][0] || fbox[1][1] < fbox[1][0] || fbox[2][1] < fbox[2][0];
    for (int a = 0; a < 3; ++a)
        fcount[a] = empty ? 0 : (fbox[a][1] - fbox[a][0] + (a == 0 ? BX : a == 1 ? BY : lzf)) / (a == 0 ? BX : a == 1 ? BY : lzf);
    std::vector<uint32_t> in;
    for (int q = 0; q < fcount[0] * fcount[1] * fcount[2]; ++q)
        in.push_back(uint32_t(q));
    nft = int(in.size());
    ftiles = upload(in);
    auto inbox = [&](int a, int lo, int hi) { return !empty && lo >= fbox[a][0] && hi <= fbox[a][1]; };
    for (int h = 0; h < 2; ++h) {
        HalfArgs<T> B = half_args<T>(h);
        const int par = h == 0 ? 0 : 1, bstep[3] = {BX, BY, B.lz}, bn[3] = {B.gx, B.gy, B.gz};
        std::vector<char> touch[3], inside[3];
        for (int a = 0; a < 3; ++a) {
            touch[a].assign(size_t(bn[a]), 0);
            inside[a].assign(size_t(bn[a]), 0);
            const PmlAxis& P = s.pml[a];
            for (int q = 0; q < bn[a]; ++q) {
                const int lo = B.dlo[a] + q * bstep[a], hi = std::min(lo + bstep[a] - 1, B.dhi[a]);
                inside[a][q] = inbox(a, lo, hi);
                for (int x = lo; x <= hi && P.W[par]; ++x)
                    if (P.smap[par][x] >= 0)
                        touch[a][q] = 1;
            }
        }
        std::vector<uint32_t> la, lb;
        for (int tz = 0; tz < B.gz; ++tz)
            for (int ty = 0; ty < B.gy; ++ty)
                for (int tx = 0; tx < B.gx; ++tx) {
                    if (inside[0][tx] && inside[1][ty] && inside[2][tz])
                        continue;  // the fused kernel's
                    const uint32_t t = uint32_t(tx + B.gx * (ty + B.gy * tz));
                    (touch[0][tx] || touch[1][ty] || touch[2][tz] ? lb : la).push_back(t);
                }
        nst[h][0] = int(la.size());
        nst[h][1] = int(lb.size());
        stiles[h][0] = upload(la);
        stiles[h][1] = upload(lb);
    }
    // algorithmic bytes of one fused-kernel launch: every field element of its
    // points read and written once, and the electric half's material words,
    // polarization state and flux-density monitor sums there
    auto cnt = [&](int comp, int a, int lo, int hi) {
        if (empty)
            return int64_t(0);
        const int l = std::max({lo, L.lo[comp][a], fbox[a][0]}), u = std::min({hi, L.hi[comp][a], fbox[a][1]});
        return int64_t(std::max(0, u - l + 1));
    };
    const int big = 1 << 30;
    double pts = 0, mcells = 0, ade = 0, dftb = 0;
    for (int c = 0; c < 6; ++c)
        pts += double(cnt(c, 0, -big, big)) * double(cnt(c, 1, -big, big)) * double(cnt(c, 2, -big, big));
    const HalfMaterial& M = s.mat[0];
    if (!M.uniform && !empty) {
        mcells = 1;
        for (int a = 0; a < 3; ++a)
            mcells *= double(std::max(0, std::min(M.mbox[a][1], fbox[a][1]) - std::max(M.mbox[a][0], fbox[a][0]) + 1));
        std::vector<double> per(size_t(L.n[2]), 0.0);
        parallel_planes(fbox[2][0], fbox[2][1], [&](int k) {
            double acc = 0;
            for (int c = 0; c < 3; ++c) {
                if (k < L.lo[c][2] || k > L.hi[c][2])
                    continue;
                for (int j = std::max(L.lo[c][1], fbox[1][0]); j <= std::min(L.hi[c][1], fbox[1][1]); ++j)
                    for (int i = std::max(L.lo[c][0], fbox[0][0]); i <= std::min(L.hi[c][0], fbox[0][1]); ++i) {
                        const int id = M.ids[c][L.index(i, j, k)];
                        acc += double(id == 255 ? M.dense[c].npoles : M.table.npoles[id]);
                    }
            }
            per[size_t(k)] = acc;
        });
        for (double v : per)
            ade += 3.0 * v;  // read P, read P-, write P+
    }
    for (auto& U : units[0])
        if (U.fused && U.npts > 0)
            dftb += double(cnt(U.comp, 0, U.box[0][0], U.box[0][1])) * double(cnt(U.comp, 1, U.box[1][0], U.box[1][1])) *
                    double(cnt(U.comp, 2, U.box[2][0], U.box[2][1])) * U.nf * 32.0;
    traffic_f[0] = 2 * pts * double(tsz);
    traffic_f[1] = 4.0 * mcells;
    traffic_f[2] = ade * double(tsz);
    traffic_f[3] = 0;
    traffic_f[4] = dftb;
    traffic_f[5] = traffic_f[0] + traffic_f[1] + traffic_f[2] + traffic_f[3] + traffic_f[4];
    traffic_f[6] = pts;
}

template <typename T>
void EngineImpl::launch_step(size_t lds, const StepArgs<T>& S0)
{
    if (nft == 0)
        return;
    StepArgs<T> X = S0;
    X.e.tiles = ftiles;
    X.e.ntiles = nft;
    X.e.nblocks = ((nft + 7) / 8) * 8;
    const bool mat = !S->mat[0].uniform;
    const int npm = S_slots[0];
    if (fdepth == 1)
        launch_fstep<T, 1>(mat, npm, lds, X);
    else
        launch_fstep<T, 2>(mat, npm, lds, X);
    HIP_OK(hipGetLastError());
}

// One time step in fused mode, from the current buffers into the other ones:
// the magnetic half-step kernel on the CPML shell, the fused kernel on the
// interior, the electric half-step kernel on the shell (it reads the new H of
// both); then the buffers swap, and the electric sources and the monitors
// follow as after the two half-step kernels.
// ev (may be null): ev[0], ev[1] bracket the fused kernel, ev[2], ev[3] the whole
// step's kernels
template <typename T>
void EngineImpl::step_fused(hipEvent_t* ev)
{
    StepArgs<T> S0 = step_args<T>();
    const size_t lds = S->mat[0].uniform ? 0 : size_t(S0.e.nmats + S0.e.npoles) * 4 * sizeof(double);
    if (ev)
        HIP_OK(hipEventRecord(ev[2], stream));
    {
        HalfArgs<T> A = half_args<T>(1);
        for (int c = 0; c < 3; ++c)
            A.w[c] = static_cast<T*>(field_b[3 + c]);
        launch_tiles<T>(1, A, stiles[1][0], nst[1][0], false);
        launch_tiles<T>(1, A, stiles[1][1], nst[1][1], true);
    }
    if (ev)
        HIP_OK(hipEventRecord(ev[0], stream));
    launch_step<T>(lds, S0);
    if (ev)
        HIP_OK(hipEventRecord(ev[1], stream));
    {
        HalfArgs<T> A = half_args<T>(0);
        for (int c = 0; c < 3; ++c) {
            A.w[c] = static_cast<T*>(field_b[c]);
            A.s[c] = static_cast<const T*>(field_b[3 + c]);
        }
        launch_tiles<T>(0, A, stiles[0][0], nst[0][0], false);
        launch_tiles<T>(0, A, stiles[0][1], nst[0][1], true);
    }
    if (ev)
        HIP_OK(hipEventRecord(ev[3], stream));
    for (int c = 0; c < 6; ++c)
        std::swap(field[c], field_b[c]);
    if (S_slots[0] > 0)
        std::swap(Pa[0], Pb[0]);
    post_half<T>(1);
    post_half<T>(0);
}

template <typename T>
void EngineImpl::step_once(hipEvent_t* ev, bool timed_e, bool timed_h)
{
    ensure_phasors();
    if (fused) {
        step_fused<T>(timed_e || timed_h ? ev : nullptr);
    } else {
        launch_half<T>(1, timed_h ? ev[2] : nullptr, timed_h ? ev[3] : nullptr);
        launch_half<T>(0, timed_e ? ev[0] : nullptr, timed_e ? ev[1] : nullptr);
    }
    if (nprobes && n < trace_cap) {
        ProbeArgs<T> P;
        std::memset(&P, 0, sizeof(P));
        for (int c = 0; c < 6; ++c)
            P.u[c] = static_cast<const T*>(field[c]) + S->L.origin;
        P.np = nprobes;
        P.off = pr_off;
        P.comp = pr_comp;
        P.idx = pr_idx;
        P.w = pr_w;
        P.trace = trace;
        k_probes<T><<<(nprobes + 63) / 64, 64, 0, stream>>>(P, n);
        HIP_OK(hipGetLastError());
    }
    if (!ring_units.empty())
        gather_dft<T>();
    ++n;
}

template <typename T>
RingArgs<T> EngineImpl::ring_args()
{
    RingArgs<T> R;
    std::memset(&R, 0, sizeof(R));
    for (int c = 0; c < 6; ++c)
        R.u[c] = static_cast<const T*>(field[c]) + S->L.origin;
    R.pitch = int(S->L.pitch);
    R.tail0 = S->L.tail0;
    R.plane = int(S->L.plane);
    R.nunits = int(ring_units.size());
    R.first[0] = 0;
    for (int q = 0; q < R.nunits; ++q) {
        const auto& hu = ring_units[size_t(q)];
        R.units[q] = units[hu.first][size_t(hu.second)];
        R.half[q] = hu.first;
        R.first[q + 1] = R.first[q] + R.units[q].npts;
    }
    return R;
}

// sample of this step (both halves' monitors read the final values of the step)
template <typename T>
void EngineImpl::gather_dft()
{
    if (ring_cnt == 0)
        ring_n0 = n;
    RingArgs<T> R = ring_args<T>();
    R.slot = ring_cnt;
    const int64_t total = R.first[R.nunits];
    k_dft_gather<T><<<unsigned((total + 255) / 256), 256, 0, stream>>>(R);
    HIP_OK(hipGetLastError());
    if (++ring_cnt == DFT_BATCH)
        flush_ring<T>();
}

template <typename T>
void EngineImpl::flush_ring()
{
    if (ring_cnt == 0)
        return;
    RingArgs<T> R = ring_args<T>();
    R.cnt = ring_cnt;
    int64_t most = 0;
    int nf = 1;
    for (int h = 0; h < 2; ++h) {
        R.ph[h] = ph[h] + size_t(ring_n0 - ph_start) * 2 * std::max(1, MF[h]);
        R.mf[h] = std::max(1, MF[h]);
    }
    for (int q = 0; q < R.nunits; ++q) {
        most = std::max(most, R.units[q].npts);
        nf = std::max(nf, R.units[q].nf);
    }
    const size_t lds = size_t(ring_cnt) * nf * 2 * sizeof(double);
    if (lds > 65536)
        throw std::runtime_error("DFT monitor with too many frequencies for one batch");
    k_dft_ring<T><<<dim3(unsigned((most + 255) / 256), unsigned(R.nunits)), 256, lds, stream>>>(R);
    HIP_OK(hipGetLastError());
    ring_cnt = 0;
}

// every gathered sample added into the FP64 sums (before any read of them)
void EngineImpl::flush_dft()
{
    if (fp32)
        flush_ring<float>();
    else
        flush_ring<double>();
}

void EngineImpl::step_n(int64_t nsteps)
{
    for (int64_t q = 0; q < nsteps; ++q) {
        if (fp32)
            step_once<float>(nullptr, false, false);
        else
            step_once<double>(nullptr, false, false);
    }
}

// ---------------------------------------------------------------- Engine API
Engine::Engine(std::unique_ptr<Setup> setup) : p_(new EngineImpl)
{
    p_->S = std::move(setup);
    p_->build();
}

Engine::~Engine() = default;

const Setup& Engine::setup() const { return *p_->S; }
int64_t Engine::time_step() const { return p_->n; }

void Engine::sync() { HIP_OK(hipStreamSynchronize(p_->stream)); }

void Engine::traffic(int h, double out[7]) const
{
    if (h < 0 || h > 2)
        throw std::runtime_error("half must be 0 (electric), 1 (magnetic) or 2 (fused step)");
    if (h == 2 && !p_->fused)
        throw std::runtime_error("this simulation does not run the fused step kernel");
    for (int q = 0; q < 7; ++q)
        out[q] = h == 2 ? p_->traffic_f[q] : p_->traffic[h][q];
}

bool Engine::fused() const { return p_->fused; }
double Engine::device_bytes() const { return p_->bytes; }

void Engine::step(int64_t nsteps)
{
    p_->step_n(nsteps);
    sync();
}

int64_t Engine::run_until_time(double t)
{
    // while (time_step * dt < time), src/simulation.cpp:768-780
    int64_t taken = 0;
    while (double(p_->n) * p_->S->dt < t) {
        p_->step_n(1);
        ++taken;
    }
    sync();
    return taken;
}

KernelTiming Engine::step_timed(int64_t nsteps, bool per_kernel)
{
    KernelTiming r;
    hipEvent_t t0, t1;
    HIP_OK(hipEventCreate(&t0));
    HIP_OK(hipEventCreate(&t1));
    std::vector<hipEvent_t> ev;
    if (per_kernel) {
        ev.resize(size_t(nsteps) * 4);
        for (auto& e : ev)
            HIP_OK(hipEventCreate(&e));
    }
    sync();
    HIP_OK(hipEventRecord(t0, p_->stream));
    for (int64_t q = 0; q < nsteps; ++q) {
        hipEvent_t* e = per_kernel ? &ev[size_t(q) * 4] : nullptr;
        if (p_->fp32)
            p_->step_once<float>(e, per_kernel, per_kernel);
        else
            p_->step_once<double>(e, per_kernel, per_kernel);
    }
    HIP_OK(hipEventRecord(t1, p_->stream));
    HIP_OK(hipEventSynchronize(t1));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, t0, t1));
    r.total_ms = ms;
    r.steps = nsteps;
    if (per_kernel) {
        for (int64_t q = 0; q < nsteps; ++q) {
            float a = 0, b = 0;
            HIP_OK(hipEventElapsedTime(&a, ev[size_t(q) * 4 + 0], ev[size_t(q) * 4 + 1]));
            HIP_OK(hipEventElapsedTime(&b, ev[size_t(q) * 4 + 2], ev[size_t(q) * 4 + 3]));
            // fused mode: a = the fused kernel, b = all kernels of the step
            r.e_ms += a;
            r.h_ms += p_->fused ? b - a : b;
        }
        for (auto& e : ev)
            (void)hipEventDestroy(e);
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    return r;
}

void Engine::dft_dims(int k, int64_t dims[4]) const
{
    const Setup& s = *p_->S;
    if (k < 0 || k >= int(s.outputs.size()))
        throw std::runtime_error("DFT output index out of range");
    auto hu = s.output_units[k];
    const DftUnit& U = s.units[hu.first][hu.second];
    dims[0] = int64_t(s.outputs[k].freqs.size());
    dims[1] = (U.q2[2] - U.q1[2]) / 2 + 1;
    dims[2] = (U.q2[1] - U.q1[1]) / 2 + 1;
    dims[3] = (U.q2[0] - U.q1[0]) / 2 + 1;
}

// output_fields_collective (src/dft.cpp:273-320) for one rank
void Engine::get_dft(int k, double* re, double* im, double* x, double* y, double* z, double* f)
{
    p_->flush_dft();
    sync();
    const Setup& s = *p_->S;
    int64_t dims[4];
    dft_dims(k, dims);
    auto hu = s.output_units[k];
    const DftUnit& U = s.units[hu.first][hu.second];
    const DftOutput& O = s.outputs[k];
    const DftDev& D = p_->units[hu.first][hu.second];
    const int64_t full = dims[1] * dims[2] * dims[3];
    if (re)
        std::fill(re, re + dims[0] * full, 0.0);
    if (im)
        std::fill(im, im + dims[0] * full, 0.0);
    if (U.has_local && D.npts > 0) {
        std::vector<double> hre(size_t(D.npts) * D.nf), him(size_t(D.npts) * D.nf);
        HIP_OK(hipMemcpy(hre.data(), D.re, hre.size() * 8, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(him.data(), D.im, him.size() * 8, hipMemcpyDeviceToHost));
        int off[3];
        for (int a = 0; a < 3; ++a)
            off[a] = (U.l1[a] - U.q1[a]) / 2;
        for (size_t fi = 0; fi < O.freqs.size(); ++fi) {
            auto it = std::lower_bound(U.freqs.begin(), U.freqs.end(), O.freqs[fi]);
            if (it == U.freqs.end() || *it != O.freqs[fi])
                continue;
            int64_t fu = it - U.freqs.begin();
            for (int kk = 0; kk < U.dims[2]; ++kk)
                for (int jj = 0; jj < U.dims[1]; ++jj)
                    for (int ii = 0; ii < U.dims[0]; ++ii) {
                        int64_t src = fu * D.npts + ii + int64_t(U.dims[0]) * (jj + int64_t(U.dims[1]) * kk);
                        int64_t dst = int64_t(fi) * full + (ii + off[0]) +
                                      dims[3] * ((jj + off[1]) + dims[2] * (kk + off[2]));
                        if (re)
                            re[dst] = hre[src];
                        if (im)
                            im[dst] = him[src];
                    }
        }
    }
    // linspace of the closure (src/dft.cpp:312-318, src/utility.cpp:281-291)
    auto lin = [](double a, double b, int64_t n, double* out) {
        if (!out)
            return;
        if (n <= 1) {
            out[0] = a;
            return;
        }
        double st = (b - a) / double(n - 1);
        for (int64_t q = 0; q < n; ++q)
            out[q] = a + double(q) * st;
    };
    lin(U.q1[0] * s.dx2, U.q2[0] * s.dx2, dims[3], x);
    lin(U.q1[1] * s.dx2, U.q2[1] * s.dx2, dims[2], y);
    lin(U.q1[2] * s.dx2, U.q2[2] * s.dx2, dims[1], z);
    if (f)
        for (size_t q = 0; q < O.freqs.size(); ++q)
            f[q] = O.freqs[q];
}

void Engine::set_trace_capacity(int64_t steps)
{
    if (p_->nprobes == 0 || steps <= p_->trace_cap)
        return;
    sync();
    double* t = p_->dalloc<double>(size_t(steps) * p_->nprobes);
    if (p_->trace && p_->trace_cap)
        HIP_OK(hipMemcpyAsync(t, p_->trace, size_t(p_->trace_cap) * p_->nprobes * 8, hipMemcpyDeviceToDevice,
                              p_->stream));
    p_->trace = t;
    p_->trace_cap = steps;
    sync();
}

int64_t Engine::trace_len() const { return std::min(p_->n, p_->trace_cap); }

void Engine::get_trace(double* out, int64_t cap)
{
    sync();
    int64_t n = std::min(trace_len(), cap);
    if (n > 0 && p_->nprobes)
        HIP_OK(hipMemcpy(out, p_->trace, size_t(n) * p_->nprobes * 8, hipMemcpyDeviceToHost));
}

void Engine::component_dims(int ct, int64_t dims[3]) const
{
    const Layout& L = p_->S->L;
    int c = comp_index(ct);
    dims[0] = std::max(0, L.hi[c][2] - L.lo[c][2] + 1);
    dims[1] = std::max(0, L.hi[c][1] - L.lo[c][1] + 1);
    dims[2] = std::max(0, L.hi[c][0] - L.lo[c][0] + 1);
}

void Engine::get_component(int ct, double* out)
{
    sync();
    const Layout& L = p_->S->L;
    int c = comp_index(ct);
    int64_t dims[3];
    component_dims(ct, dims);
    if (dims[0] * dims[1] * dims[2] == 0)
        return;
    size_t ts = p_->tsz;
    std::vector<char> host(size_t(L.alloc) * ts);
    HIP_OK(hipMemcpy(host.data(), p_->field[c], host.size(), hipMemcpyDeviceToHost));
    int64_t q = 0;
    for (int k = L.lo[c][2]; k <= L.hi[c][2]; ++k)
        for (int j = L.lo[c][1]; j <= L.hi[c][1]; ++j)
            for (int i = L.lo[c][0]; i <= L.hi[c][0]; ++i) {
                int64_t id = L.index(i, j, k);
                out[q++] = p_->fp32 ? double(reinterpret_cast<float*>(host.data())[id])
                                    : reinterpret_cast<double*>(host.data())[id];
            }
}

double Engine::probe(const double pos[3], int ct)
{
    sync();
    auto terms = probe_terms(*p_->S, pos, ct);
    double v = 0;
    for (auto& t : terms) {
        double x;
        if (p_->fp32) {
            float h;
            HIP_OK(hipMemcpy(&h, static_cast<float*>(p_->field[t.comp]) + t.index, 4, hipMemcpyDeviceToHost));
            x = h;
        } else {
            HIP_OK(hipMemcpy(&x, static_cast<double*>(p_->field[t.comp]) + t.index, 8, hipMemcpyDeviceToHost));
        }
        v += x * t.w;
    }
    return v;
}

void Engine::adjoint_gradient(const Engine& adj, const int fwd_out[3], const int adj_out[3], double freq,
                              const double eps_prime[2], const double norm[2], const double center[3],
                              const double size[3], const int64_t dim[3], double* grad)
{
    EngineImpl& F = *p_;
    const EngineImpl& Ad = *adj.p_;
    const Setup& S = *F.S;
    if (Ad.S->L.M[0] != S.L.M[0] || Ad.S->L.n[0] != S.L.n[0] || Ad.S->L.n[1] != S.L.n[1] ||
        Ad.S->L.n[2] != S.L.n[2])
        throw std::runtime_error("forward and adjoint simulations have different grids");
    if (dim[0] < 1 || dim[1] < 1 || dim[2] < 1)
        throw std::runtime_error("density grid dimension must be positive");
    GradArgs A;
    std::memset(&A, 0, sizeof(A));
    for (int c = 0; c < 3; ++c) {
        const EngineImpl* E[2] = {&F, &Ad};
        const int ks[2] = {fwd_out[c], adj_out[c]};
        const DftUnit* U[2];
        const DftDev* D[2];
        int64_t fi[2];
        for (int s = 0; s < 2; ++s) {
            const Setup& SS = *E[s]->S;
            if (ks[s] < 0 || ks[s] >= int(SS.outputs.size()))
                throw std::runtime_error("DFT output index out of range");
            auto hu = SS.output_units[ks[s]];
            U[s] = &SS.units[hu.first][hu.second];
            D[s] = &E[s]->units[hu.first][hu.second];
            if (U[s]->comp != c || !U[s]->has_local)
                throw std::runtime_error("design DFT outputs must be Ex, Ey, Ez over the design box");
            auto it = std::lower_bound(U[s]->freqs.begin(), U[s]->freqs.end(), freq);
            if (it == U[s]->freqs.end() || *it != freq)
                throw std::runtime_error("frequency is not registered");
            fi[s] = it - U[s]->freqs.begin();
        }
        for (int a = 0; a < 3; ++a)
            if (U[0]->box[a][0] != U[1]->box[a][0] || U[0]->box[a][1] != U[1]->box[a][1])
                throw std::runtime_error("forward and adjoint design DFTs cover different points");
        GradComp& G = A.c[c];
        G.fre = D[0]->re + fi[0] * D[0]->npts;
        G.fim = D[0]->im + fi[0] * D[0]->npts;
        G.are = D[1]->re + fi[1] * D[1]->npts;
        G.aim = D[1]->im + fi[1] * D[1]->npts;
        for (int a = 0; a < 3; ++a) {
            G.lo[a] = U[0]->box[a][0];
            G.dims[a] = U[0]->dims[a];
            G.par[a] = comp_parity(c, a);
        }
    }
    for (int a = 0; a < 3; ++a) {
        A.M[a] = S.L.M[a];
        const double cd = center[a] / S.dx2, sd = size[a] / S.dx2;  // as Two_Medium_Box (src/geometry.cpp:49-56)
        A.blo[a] = std::min(cd - sd / 2, cd + sd / 2);
        A.bhi[a] = std::max(cd - sd / 2, cd + sd / 2);
        // numpy.linspace(c - s/2, c + s/2, dim) and the spacing transpose() takes from its first two nodes
        const double g0 = center[a] - size[a] / 2, g1 = center[a] + size[a] / 2;
        A.gx0[a] = g0;
        A.dxn[a] = dim[a] > 1 ? (g0 + (g1 - g0) / double(dim[a] - 1)) - g0 : 1.0;
        A.nd[a] = int(dim[a]);
    }
    A.dx2 = S.dx2;
    A.two_pi_f = 2 * 3.141592653589793e+00 * freq;
    A.epr = eps_prime[0];
    A.epi = eps_prime[1];
    A.nr = norm[0];
    A.ni = norm[1];
    const int64_t nn = dim[0] * dim[1] * dim[2];
    F.flush_dft();
    adj.p_->flush_dft();
    sync();
    HIP_OK(hipStreamSynchronize(Ad.stream));  // its DFT sums are read on this engine's stream
    double* g = F.sample_buf(size_t(nn));
    A.grad = g;
    k_adjoint_gradient<<<unsigned((nn + 255) / 256), 256, 0, F.stream>>>(A);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(F.stream));
    HIP_OK(hipMemcpy(grad, g, size_t(nn) * 8, hipMemcpyDeviceToHost));
}

// Steps in examination windows with one per-step sample kept on the device and
// read back once per window; `sample` enqueues the sample of the step just taken
// into buf[slot].
template <typename F>
static std::vector<double> window_samples(EngineImpl& p, int64_t nsteps, double* buf, F&& sample)
{
    for (int64_t q = 0; q < nsteps; ++q) {
        p.step_n(1);
        sample(q);
    }
    std::vector<double> h(size_t(std::max<int64_t>(nsteps, 0)));
    if (nsteps > 0) {
        HIP_OK(hipStreamSynchronize(p.stream));
        HIP_OK(hipMemcpy(h.data(), buf, size_t(nsteps) * 8, hipMemcpyDeviceToHost));
    }
    return h;
}

static int64_t steps_until(const EngineImpl& p, double t)
{
    // number of steps the loop `while (time_step * dt < t) step` takes from now
    int64_t n = p.n, k = 0;
    while (double(n + k) * p.S->dt < t)
        ++k;
    return k;
}

// Simulation::run_until_dft, src/simulation.cpp:782-842
int64_t Engine::run_until_dft(const double c[3], const double s[3], int ct, double freq, double interval, double var,
                              std::string* log)
{
    EngineImpl& P = *p_;
    const Setup& S = *P.S;
    double p1[3], p2[3];
    for (int a = 0; a < 3; ++a) {
        p1[a] = (c[a] - s[a] / 2) / S.dx2;
        p2[a] = (c[a] + s[a] / 2) / S.dx2;
    }
    // closure of the requested box, then the registered unit with that closure
    auto lo = [](double x, int T) {
        int xi = int(std::floor(x));
        return ((xi & 1) ^ T) ? xi - 1 : xi;
    };
    auto hi = [](double x, int T) {
        int xi = int(std::ceil(x));
        return ((xi & 1) ^ T) ? xi + 1 : xi;
    };
    int q1[3], q2[3];
    for (int a = 0; a < 3; ++a) {
        int T = (ct >> a) & 1;
        q1[a] = lo(p1[a], T);
        q2[a] = hi(p2[a], T);
    }
    const int h = ctype_is_e(ct) ? 0 : 1;
    int u = -1;
    for (size_t k = 0; k < S.units[h].size(); ++k) {
        const DftUnit& U = S.units[h][k];
        if (U.ct == ct && U.q1[0] == q1[0] && U.q1[1] == q1[1] && U.q1[2] == q1[2] && U.q2[0] == q2[0] &&
            U.q2[1] == q2[1] && U.q2[2] == q2[2])
            u = int(k);
    }
    if (u < 0)
        throw std::runtime_error("dft is not registered");
    const DftUnit& U = S.units[h][u];
    auto it = std::lower_bound(U.freqs.begin(), U.freqs.end(), freq);
    if (it == U.freqs.end() || *it != freq)
        throw std::runtime_error("frequency is not registered");
    const int64_t fi = it - U.freqs.begin();
    const DftDev& D = P.units[h][u];
    const int64_t start = P.n;
    const double turnoff = S.turnoff_time;
    P.step_n(steps_until(P, turnoff));
    int extension = 1;
    double cur_var;
    do {
        double max_norm = std::numeric_limits<double>::min();
        double min_norm = std::numeric_limits<double>::max();
        const int64_t k = steps_until(P, turnoff + extension * interval);
        double* buf = P.sample_buf(size_t(k));
        auto vals = window_samples(P, k, buf, [&](int64_t q) {
            P.flush_dft();
            k_dft_norm<<<1, 1024, 0, P.stream>>>(D.re + fi * D.npts, D.im + fi * D.npts, D.npts, buf, q);
            HIP_OK(hipGetLastError());
        });
        for (double v : vals) {
            if (v > max_norm)
                max_norm = v;
            if (v < min_norm)
                min_norm = v;
        }
        ++extension;
        cur_var = (max_norm - min_norm) / max_norm;
        if (log) {
            char b[160];
            std::snprintf(b, sizeof b, "max norm=%.10g,min norm=%.10g,var=%.10g\n", max_norm, min_norm, cur_var);
            *log += b;
        }
    } while (cur_var > var);
    sync();
    return P.n - start;
}

// Simulation::run_until_fields_decayed, src/simulation.cpp:844-901
int64_t Engine::run_until_decay(const double pos[3], int ct, double interval, double decay_by, std::string* log)
{
    EngineImpl& P = *p_;
    const Setup& S = *P.S;
    if (!ctype_is_eh(ct))
        throw std::runtime_error("decay probe takes an E or H component");
    auto terms = probe_terms(S, pos, ct);
    std::vector<int64_t> ti;
    std::vector<double> tw;
    int comp = comp_index(ct);
    for (auto& t : terms) {
        ti.push_back(t.index);
        tw.push_back(t.w);
    }
    int64_t* d_idx = P.upload(ti);
    double* d_w = P.upload(tw);
    const int nterm = int(ti.size());
    auto sample = [&](double* buf) {
        return [&P, buf, d_idx, d_w, nterm, comp](int64_t q) {
            if (P.fp32)
                k_point_value<float><<<1, 64, 0, P.stream>>>(static_cast<const float*>(P.field[comp]), d_idx, d_w,
                                                             nterm, buf, q);
            else
                k_point_value<double><<<1, 64, 0, P.stream>>>(static_cast<const double*>(P.field[comp]), d_idx,
                                                              d_w, nterm, buf, q);
            HIP_OK(hipGetLastError());
        };
    };
    const int64_t start = P.n;
    const double turnoff = S.turnoff_time;
    double max_sqr = 0;
    {
        const int64_t k = steps_until(P, turnoff);
        double* buf = P.sample_buf(size_t(k));
        for (double v : window_samples(P, k, buf, sample(buf)))
            if (v * v > max_sqr)
                max_sqr = v * v;
    }
    int extension = 1;
    double decay;
    do {
        double cur_max = 0;
        const int64_t k = steps_until(P, turnoff + extension * interval);
        double* buf = P.sample_buf(size_t(k));
        for (double v : window_samples(P, k, buf, sample(buf)))
            if (v * v > cur_max)
                cur_max = v * v;
        ++extension;
        if (cur_max > max_sqr)
            max_sqr = cur_max;
        decay = cur_max / max_sqr;
        if (log) {
            char b[160];
            std::snprintf(b, sizeof b, "Field Decay = %.10g/%.10g=%.10g\n", cur_max, max_sqr, decay);
            *log += b;
        }
    } while (decay > decay_by);
    sync();
    return P.n - start;
}

}  // namespace ffip_amd
