// Host-side set-up of one simulation: turns the reference's JSON configuration
// (the document ffip/simulation.py:362-402 writes, consumed by
// Simulation::init, src/simulation.cpp:127-188) into the compact device layout
// the HIP kernels use.
//
// Device layout ("padded Yee-cell lattice").  The reference stores every field
// component interleaved on one doubled-coordinate grid (src/utility.cpp:130-140).
// Here each of the six components has its own array, all sharing one index space:
// doubled coordinate x on axis a maps to cell index ix = (x + M[a]) >> 1 with M[a]
// the smallest odd integer >= 1 - P1[a] (P1 = lower doubled bound of the domain).
// With M odd, an E point reads its H neighbours at ix+1 and ix along its two
// derivative axes and an H point reads E at ix and ix-1, for every component and
// any grid parity.  Rows carry >= 1 padding element, planes >= 1 padding row and
// the array one padding plane on each side, so every +-1 neighbour read of a real
// point stays inside the allocation and reads zero where the reference reads its
// zero ghost shell.
#pragma once

#include <algorithm>
#include <array>
#include <atomic>
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "json.hpp"

namespace ffip_amd {

// body(k) for every k in [k0, k1], spread over up to 16 host threads (the calling
// thread takes the rest if no more can start); callers write disjoint ranges.
template <class F>
void parallel_planes(int k0, int k1, F&& body)
{
    const int n = k1 - k0 + 1;
    const unsigned hw = std::thread::hardware_concurrency();
    const int nt = std::max(1, std::min({int(hw ? hw : 1u), 16, n / 2}));
    if (nt == 1) {
        for (int k = k0; k <= k1; ++k)
            body(k);
        return;
    }
    std::atomic<int> next{k0};
    auto work = [&] {
        for (int k = next.fetch_add(1); k <= k1; k = next.fetch_add(1))
            body(k);
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) {
        try {
            pool.emplace_back(work);
        } catch (const std::system_error&) {
            break;
        }
    }
    work();
    for (auto& th : pool)
        th.join();
}

// Allocator whose value-less construct leaves elements uninitialised, for
// per-point arrays that threaded plane loops fill afterwards.
template <class X>
struct UninitAlloc : std::allocator<X> {
    template <class U>
    struct rebind {
        using other = UninitAlloc<U>;
    };
    UninitAlloc() = default;
    template <class U>
    UninitAlloc(const UninitAlloc<U>&) noexcept
    {
    }
    template <class U>
    void construct(U*) noexcept
    {
    }
    template <class U, class... Args>
    void construct(U* p, Args&&... args)
    {
        ::new (static_cast<void*>(p)) U(std::forward<Args>(args)...);
    }
};
template <class X>
using RawVec = std::vector<X, UninitAlloc<X>>;

// Coord_Type bit patterns, include/utility.hpp:48-64 (reference)
enum CType : int {
    CT_EX = 0b001, CT_EY = 0b010, CT_EZ = 0b100,
    CT_HX = 0b110, CT_HY = 0b101, CT_HZ = 0b011,
    CT_DX = 0b1001, CT_DY = 0b1010, CT_DZ = 0b1100,
    CT_BX = 0b1110, CT_BY = 0b1101, CT_BZ = 0b1011,
};

int ctype_from_name(const std::string& s);  // throws on unknown names
bool ctype_is_e(int ct);
bool ctype_is_eh(int ct);
int ctype_dir(int ct);  // axis of the component (0,1,2)
// half (0 = electric, 1 = magnetic) and axis of a component index 0..5
inline int comp_index(int ct) { return (ctype_is_e(ct) ? 0 : 3) + ctype_dir(ct); }
inline int comp_parity(int comp, int axis)
{
    // E_a: odd along a, even elsewhere; H_a: even along a, odd elsewhere.
    bool along = (comp % 3) == axis;
    return comp < 3 ? (along ? 1 : 0) : (along ? 0 : 1);
}

struct Sus {
    double a0, a1, b0, b1, b2;
    bool operator==(const Sus& o) const
    {
        return a0 == o.a0 && b0 == o.b0 && b1 == o.b1 && b2 == o.b2 && a1 == o.a1;
    }
};

struct Medium {
    double eps = 1, mu = 1;
    std::vector<Sus> e_sus, m_sus;
    std::vector<double> e_amp, m_amp;
};

struct AbsMedium {
    double eps = 0, mu = 0;
    std::vector<double> e_amp, m_amp;  // indexed by susceptibility pool
};

enum Bc : int { BC_NONE = 0, BC_PERIOD = 1, BC_PEC = 2, BC_PMC = 3 };

// Rows whose last two x indices share one packed line (see Layout::tail0).
constexpr int TAIL_GROUP = 15;

struct Layout {
    int P1[3], P2[3];  // doubled-coordinate domain (sim_p1, sim_p2)
    int M[3];          // ix = (x + M) >> 1
    int n[3];          // index extent per axis, ghosts included
    int64_t pitch, rows, plane, nplanes, alloc, origin;
    int lo[6][3], hi[6][3];  // real points of each component (inclusive)
    // Rows are `pitch` elements (a whole number of 128-byte lines of FP32).  When
    // the last two x indices of a row (its last real point and the ghost after it)
    // would sit alone in the row's last line, tail0 is the first of them and
    // they are packed instead: those two indices of the TAIL_GROUP rows
    // j = 15g - 1 .. 15g + 13 of a plane are stored side by side in the last line
    // of row 15g - 1, so one line serves 15 rows.  Otherwise tail0 = NO_TAIL.
    static constexpr int NO_TAIL = 1 << 30;
    int tail0 = NO_TAIL;

    int ix(int a, int x) const { return (x + M[a]) >> 1; }
    int coord(int a, int i, int parity) const { return 2 * i - M[a] + (1 - parity); }
    // element of lattice index (i, j, k) in every per-cell array; j >= -1, k >= -1
    // (the kernels' lat_off is the same map)
    int64_t index(int i, int j, int k) const
    {
        if (i < tail0)
            return origin + i + pitch * j + plane * k;
        const int m = (j + 1) % TAIL_GROUP;
        return origin + tail0 + 2 * m + (i - tail0) + pitch * (j - m) + plane * k;
    }
    int64_t count(int comp) const;
};

struct Geometry;

struct MatTable {
    std::vector<double> inv_eps, eps;
    std::vector<int> npoles, pole0;
    std::vector<double> c1, c2, sc3;  // pole table
};

struct DenseRegion {  // per-point media of two-medium boxes (one per component)
    bool present = false;
    int box[3][2] = {{0, -1}, {0, -1}, {0, -1}};
    int npoles = 0;                   // = susceptibility pool size of this half
    std::vector<double> inv_eps, eps;  // [k][j][i] over box
    std::vector<double> amp;           // [pt][npoles], masked (|amp| <= 1e-4 -> 0)
    std::vector<double> c1, c2, c3;    // [npoles]
    int64_t npts() const
    {
        return int64_t(box[0][1] - box[0][0] + 1) * (box[1][1] - box[1][0] + 1) * (box[2][1] - box[2][0] + 1);
    }
};

struct HalfMaterial {
    bool uniform = true;
    double u_inv = 1, u_eps = 1;
    RawVec<uint8_t> ids[3];  // per component, Layout::alloc entries
    MatTable table;
    DenseRegion dense[3];
    int max_poles = 0;
    // every point whose id is not the default medium's (bg_id) lies in this box
    // of lattice indices (inclusive; empty when lo > hi)
    int bg_id = 0;
    int mbox[3][2] = {{0, -1}, {0, -1}, {0, -1}};
};

struct PmlAxis {  // per axis; [parity]
    std::vector<double> b[2], c[2], ik[2];
    std::vector<int> smap[2];  // -1 outside, else slab position
    int W[2] = {0, 0};
};

struct Dipole {
    int comp;
    int i, j, k;
    double amp, start, end, width;
};

struct DftUnit {
    int ct, comp;
    bool is_db;
    int q1[3], q2[3];  // closure (doubled coordinates)
    bool has_local = false;
    int l1[3], l2[3];  // local region inside the grid (doubled)
    int box[3][2];     // local region in index space
    int dims[3];       // local dims
    std::vector<double> freqs;  // sorted, unique after finalisation
    int64_t npts() const { return has_local ? int64_t(dims[0]) * dims[1] * dims[2] : 0; }
};

struct DftOutput {
    std::string group;
    int unit;
    std::vector<double> freqs;
};

struct ProbeTerm {
    int comp;
    int64_t index;
    double w;
};

struct FaceCopy {  // ghost fill for Period / PEC / PMC (src/fields.cpp:327-482)
    int comp, axis;
    int dst, src;  // index along axis
    double mult;
};

struct GeomInfoRequest {
    std::string group;
    int ct;
    int geom;
};

struct Setup {
    double resolution, dx, dt, dx2, courant;
    double size[3];
    int bc[3][2];
    Layout L;
    std::vector<Sus> pool[2];
    Medium bg;
    HalfMaterial mat[2];
    PmlAxis pml[3];
    std::vector<Dipole> dip[2];
    double turnoff_time = 0;
    std::vector<DftUnit> units[2];
    std::vector<DftOutput> outputs;
    std::vector<std::pair<int, int>> output_units;  // (half, unit) for each output
    std::vector<std::vector<ProbeTerm>> probes;
    std::vector<FaceCopy> faces[2];
    std::vector<GeomInfoRequest> geom_info;
    // per component: geometry owning each point of the box geom_box (lattice
    // indices, inclusive; -1 = default medium, as is every point outside the box)
    std::vector<std::vector<int>> geom_index;
    int geom_box[6][3][2] = {};
    std::string precision = "fp64";
    bool keep_geom_index = false;
};

using DatasetMap = std::map<std::string, std::vector<double>>;

struct ProbeSpec {
    double pos[3];
    int ct;
};

// Full set-up.  Throws std::runtime_error with a message on invalid input.
void build_setup(const Json& cfg, const DatasetMap& ds, const std::vector<ProbeSpec>& probes, Setup& S);

// Geometry-info request output (src/simulation.cpp:190-233): the physical
// coordinates of the points of one component owned by one geometry.
void geom_info_coords(const Setup& S, int ct, int geom, std::vector<double>& x, std::vector<double>& y,
                      std::vector<double>& z);

// Probe terms for Simulation::at (include/utility.hpp:983-1022) at a physical position.
std::vector<ProbeTerm> probe_terms(const Setup& S, const double pos[3], int ct);

}  // namespace ffip_amd
