// Host-side set-up; each function cites the reference code whose semantics it
// reproduces.  Compiled with -ffp-contract=off so that every set-up quantity
// (PML grading, source weights, material interpolation) is rounded exactly as the
// reference computes it.
#include "setup.hpp"

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <system_error>
#include <thread>

namespace ffip_amd {

static const double PI = 3.141592653589793e+00;  // src/utility.cpp:5

int ctype_from_name(const std::string& s)
{
    static const std::map<std::string, int> m = {
        {"Ex", CT_EX}, {"Ey", CT_EY}, {"Ez", CT_EZ}, {"Hx", CT_HX}, {"Hy", CT_HY}, {"Hz", CT_HZ},
        {"Dx", CT_DX}, {"Dy", CT_DY}, {"Dz", CT_DZ}, {"Bx", CT_BX}, {"By", CT_BY}, {"Bz", CT_BZ}};
    auto it = m.find(s);
    if (it == m.end())
        throw std::runtime_error("unknown field component \"" + s + "\"");
    return it->second;
}

static int bits3(int x) { return (x & 1) + ((x >> 1) & 1) + ((x >> 2) & 1); }
bool ctype_is_e(int ct) { return bits3(ct & 7) == 1; }
bool ctype_is_eh(int ct) { return ct > 0 && ct < 7; }
int ctype_dir(int ct)
{
    int c = ct & 7;
    if (bits3(c) == 2)
        c ^= 7;
    return c == 1 ? 0 : (c == 2 ? 1 : (c == 4 ? 2 : 3));
}

static int parity_of(int x) { return x & 1; }  // two's complement: (-3 & 1) == 1

// ---------------------------------------------------------------- utilities
// std::round is what the reference uses (include/utility.hpp:387-390)
static double cround(double x) { return std::round(x); }

// get_nearest_int<S,T>(double), include/utility.hpp:487-495
static int nearest_int(double x, int S, int T)
{
    int xi = S == -1 ? int(std::floor(x)) : int(std::ceil(x));
    return ((xi & 1) ^ T) ? xi + S : xi;
}

// get_nearest_point<S>, include/utility.hpp:498-534
static std::array<int, 3> nearest_point(const double p[3], int ct, int S)
{
    return {nearest_int(p[0], S, (ct & 1) ? 1 : 0), nearest_int(p[1], S, (ct & 2) ? 1 : 0),
            nearest_int(p[2], S, (ct & 4) ? 1 : 0)};
}

// Yee3::get_interp_weights, src/utility.cpp:231-241
static std::array<double, 8> interp_weights(double sx, double sy, double sz)
{
    return {(1 - sx) * (1 - sy) * (1 - sz), sx * (1 - sy) * (1 - sz), (1 - sx) * sy * (1 - sz),
            sx * sy * (1 - sz),             (1 - sx) * (1 - sy) * sz, sx * (1 - sy) * sz,
            (1 - sx) * sy * sz,             sx * sy * sz};
}

static double gaussian_width(double f) { return 1 / (2 * PI * f); }

int64_t Layout::count(int comp) const
{
    int64_t c = 1;
    for (int a = 0; a < 3; ++a)
        c *= std::max(0, hi[comp][a] - lo[comp][a] + 1);
    return c;
}

// ---------------------------------------------------------------- media
// make_*_susceptibility, src/medium.cpp:98-118
static Sus sus_lorentz(double f, double g)
{
    double omega = 2 * PI * f;
    return {omega * omega, 0, omega * omega, 2 * PI * g, 1};
}
static Sus sus_drude(double f, double g)
{
    double omega = 2 * PI * f;
    return {omega * omega, 0, 0, 2 * PI * g, 1};
}
static Sus sus_debye(double tau) { return {1, 0, 1, tau, 0}; }
static const Sus SUS_COND = {1, 0, 0, 1, 0};

// json2susceptibility, src/simulation.cpp:90-101
static Sus json2sus(const Json& j)
{
    const std::string& t = j.at("type").str();
    if (t == "Lorentz")
        return sus_lorentz(j.at("frequency").num(), j.at("gamma").num());
    if (t == "Drude")
        return sus_drude(j.at("frequency").num(), j.at("gamma").num());
    if (t == "Deybe")
        return sus_debye(j.at("tau").num());
    throw std::runtime_error("Unknonw susceptibility");
}

// json2medium, src/simulation.cpp:58-83
static Medium json2medium(const Json& j)
{
    Medium m;
    m.eps = j.at("epsilon").num();
    m.mu = j.at("mu").num();
    double ec = j.at("electric conductivity").num();
    if (ec != 0) {
        m.e_sus.push_back(SUS_COND);
        m.e_amp.push_back(ec);
    }
    double mc = j.at("magnetic conductivity").num();
    if (mc != 0) {
        m.m_sus.push_back(SUS_COND);
        m.m_amp.push_back(mc);
    }
    if (const Json* es = j.find("electric susceptibility"))
        for (auto& s : es->elems()) {
            m.e_sus.push_back(json2sus(s));
            m.e_amp.push_back(s.at("amplitude").num());
        }
    if (const Json* ms = j.find("magnetic susceptibility"))
        for (auto& s : ms->elems()) {
            m.m_sus.push_back(json2sus(s));
            m.m_amp.push_back(s.at("amplitude").num());
        }
    return m;
}

// Abstract_Susceptibility, src/medium.cpp:24-31
static void abstract_sus(const Sus& s, double dt, double& c1, double& c2, double& c3)
{
    double c0 = 2 * s.b2 + s.b1 * dt;
    c1 = (4 * s.b2 - 2 * s.b0 * dt * dt) / c0;
    c2 = (-2 * s.b2 + s.b1 * dt) / c0;
    c3 = (2 * s.a0 * dt * dt) / c0;
}

// Structure::add_to_material_pool, src/structure.cpp:95-124
static void add_to_pool(std::vector<Sus>* pool, const Medium& m)
{
    for (auto& s : m.e_sus)
        if (std::find(pool[0].begin(), pool[0].end(), s) == pool[0].end())
            pool[0].push_back(s);
    for (auto& s : m.m_sus)
        if (std::find(pool[1].begin(), pool[1].end(), s) == pool[1].end())
            pool[1].push_back(s);
}

// Structure::get_abstract_medium, src/structure.cpp:126-153
static AbsMedium abstract_medium(const std::vector<Sus>* pool, const Medium& m)
{
    AbsMedium a;
    a.eps = m.eps;
    a.mu = m.mu;
    a.e_amp.assign(pool[0].size(), 0.0);
    a.m_amp.assign(pool[1].size(), 0.0);
    for (size_t i = 0; i < m.e_sus.size(); ++i)
        for (size_t j = 0; j < pool[0].size(); ++j)
            if (pool[0][j] == m.e_sus[i]) {
                a.e_amp[j] = m.e_amp[i];
                break;
            }
    for (size_t i = 0; i < m.m_sus.size(); ++i)
        for (size_t j = 0; j < pool[1].size(); ++j)
            if (pool[1][j] == m.m_sus[i]) {
                a.m_amp[j] = m.m_amp[i];
                break;
            }
    return a;
}

// ---------------------------------------------------------------- geometry
struct Geometry {
    virtual ~Geometry() = default;
    virtual bool inside(double x, double y, double z) const = 0;
    virtual void bbox(double lo[3], double hi[3]) const = 0;
    virtual bool per_point() const { return false; }
    virtual AbsMedium medium_at(double, double, double) const { return med; }
    // epsilon (half 0) or mu (half 1) at a point, with the np susceptibility
    // amplitudes of that half written to amp: the values medium_at gives
    // true when half_at is the same at every point, with no susceptibility,
    // whatever the point: value then holds that epsilon (half 0) or mu (half 1)
    virtual bool half_constant(int, double&) const { return false; }
    virtual double half_at(double x, double y, double z, int half, double* amp, size_t np) const
    {
        AbsMedium m = medium_at(x, y, z);
        const std::vector<double>& a = half == 0 ? m.e_amp : m.m_amp;
        for (size_t p = 0; p < np; ++p)
            amp[p] = p < a.size() ? a[p] : 0.0;
        return half == 0 ? m.eps : m.mu;
    }
    AbsMedium med;
};

// Box, src/geometry.cpp:29-46 (CGAL Iso_cuboid_3 strict interior)
struct BoxGeom : Geometry {
    double p1[3], p2[3], lo[3], hi[3];
    BoxGeom(const double c[3], const double s[3])
    {
        for (int a = 0; a < 3; ++a) {
            p1[a] = c[a] - s[a] / 2;
            p2[a] = c[a] + s[a] / 2;
            lo[a] = std::min(p1[a], p2[a]);
            hi[a] = std::max(p1[a], p2[a]);
        }
    }
    bool inside(double x, double y, double z) const override
    {
        return x > lo[0] && x < hi[0] && y > lo[1] && y < hi[1] && z > lo[2] && z < hi[2];
    }
    void bbox(double l[3], double h[3]) const override
    {
        for (int a = 0; a < 3; ++a) {
            l[a] = lo[a];
            h[a] = hi[a];
        }
    }
};

// Sphere, src/geometry.cpp:13-27: squared distance < radius*radius.  CGAL decides
// this exactly; long double makes a wrong decision require a point within ~1e-19
// relative of the surface.
struct SphereGeom : Geometry {
    double c[3], r, r2;
    SphereGeom(const double cc[3], double rad) : r(rad), r2(rad * rad)
    {
        for (int a = 0; a < 3; ++a)
            c[a] = cc[a];
    }
    bool inside(double x, double y, double z) const override
    {
        long double dx = (long double)x - c[0], dy = (long double)y - c[1], dz = (long double)z - c[2];
        return dx * dx + dy * dy + dz * dz < (long double)r2;
    }
    void bbox(double l[3], double h[3]) const override
    {
        double rr = std::fabs(r);
        for (int a = 0; a < 3; ++a) {
            l[a] = c[a] - rr - 1;
            h[a] = c[a] + rr + 1;
        }
    }
};

// interpn<N>::operator() with zero-padding extrapolation, include/utility.hpp:676-700
// and 770-794; dims/q ordered from the slowest axis.
static double interpn(const double* data, const int* dims, const double* q, int nd)
{
    size_t dimn = size_t(dims[0]);
    size_t stride = 1;
    for (int d = 1; d < nd; ++d)
        stride *= size_t(dims[d]);
    auto base = [&](const double* ptr) -> double {
        return nd == 1 ? ptr[0] : interpn(ptr, dims + 1, q + 1, nd - 1);
    };
    double xq = q[0];
    if (dimn == 1)
        return base(data);
    if (xq <= 0) {
        if (xq <= -1)
            return 0.0;
        return (xq + 1) * base(data);
    }
    if (xq >= double(dimn - 1)) {
        if (xq >= double(dimn))
            return 0.0;
        return (double(dimn) - xq) * base(data + stride * (dimn - 1));
    }
    size_t index = size_t(xq);
    double tx = xq - double(index);
    return tx * base(data + stride * (index + 1)) + (1 - tx) * base(data + stride * index);
}

// Two_Medium_Box, src/geometry.cpp:49-76
struct TwoMediumGeom : BoxGeom {
    int dim[3];
    AbsMedium m1, m2;
    std::vector<double> rho;
    TwoMediumGeom(const double c[3], const double s[3], const int d[3], AbsMedium a, AbsMedium b,
                  std::vector<double> r)
        : BoxGeom(c, s), m1(std::move(a)), m2(std::move(b)), rho(std::move(r))
    {
        for (int k = 0; k < 3; ++k)
            dim[k] = d[k];
        if (int64_t(rho.size()) != int64_t(dim[0]) * dim[1] * dim[2])
            throw std::runtime_error("two medium box: density dataset size does not match dimension");
        if (dim[0] < 1 || dim[1] < 1 || dim[2] < 1)
            throw std::runtime_error("Invalid Dimension");
    }
    bool per_point() const override { return true; }
    double rho_at(double x, double y, double z) const
    {
        double q[3] = {(z - p1[2]) / (p2[2] - p1[2]) * (dim[2] - 1), (y - p1[1]) / (p2[1] - p1[1]) * (dim[1] - 1),
                       (x - p1[0]) / (p2[0] - p1[0]) * (dim[0] - 1)};
        int dz[3] = {dim[2], dim[1], dim[0]};
        return interpn(rho.data(), dz, q, 3);
    }
    bool half_constant(int half, double& value) const override
    {
        // m1 v + m2 (1 - v) is exactly 1 for m1 = m2 = 1 (1 - v rounds by at most
        // 2^-54 when v < 1/2 and is exact otherwise), and exactly 0 for zero amplitudes
        const double g1 = half == 0 ? m1.eps : m1.mu, g2 = half == 0 ? m2.eps : m2.mu;
        if (g1 != 1 || g2 != 1)
            return false;
        for (const std::vector<double>* a : {half == 0 ? &m1.e_amp : &m1.m_amp, half == 0 ? &m2.e_amp : &m2.m_amp})
            for (double x : *a)
                if (x != 0)
                    return false;
        value = 1;
        return true;
    }
    double half_at(double x, double y, double z, int half, double* amp, size_t np) const override
    {
        // the arithmetic of medium_at, without building the medium
        double val = rho_at(x, y, z);
        double om = 1 - val;
        const std::vector<double>& a1 = half == 0 ? m1.e_amp : m1.m_amp;
        const std::vector<double>& a2 = half == 0 ? m2.e_amp : m2.m_amp;
        for (size_t p = 0; p < np; ++p)
            amp[p] = p < a1.size() ? a1[p] * val + a2[p] * om : 0.0;
        return half == 0 ? m1.eps * val + m2.eps * om : m1.mu * val + m2.mu * om;
    }
    AbsMedium medium_at(double x, double y, double z) const override
    {
        double val = rho_at(x, y, z);
        AbsMedium r;
        double om = 1 - val;
        r.eps = m1.eps * val + m2.eps * om;
        r.mu = m1.mu * val + m2.mu * om;
        r.e_amp.resize(m1.e_amp.size());
        r.m_amp.resize(m1.m_amp.size());
        for (size_t i = 0; i < r.e_amp.size(); ++i)
            r.e_amp[i] = m1.e_amp[i] * val + m2.e_amp[i] * om;
        for (size_t i = 0; i < r.m_amp.size(); ++i)
            r.m_amp[i] = m1.m_amp[i] * val + m2.m_amp[i] * om;
        return r;
    }
};

static void vec3(const Json& j, double out[3])
{
    for (int a = 0; a < 3; ++a)
        out[a] = j[a].num();
}

// General_Medium_Box, src/geometry.cpp:79-107, built as src/simulation.cpp:439-514
// reads it: the density's zero-padded trilinear interpolant (as Two_Medium_Box)
// is placed on the sampled density function, taken as uniformly spaced from its
// first to its last sample, and selects a zero-padded linear interpolation
// between the sampled media (interpn<1> over Abstract_Medium, field by field).
struct GeneralMediumGeom : BoxGeom {
    int dim[3];
    std::vector<double> rho, rho_fun;
    // per sample: eps, mu and the pool amplitudes of each half, as field tables
    std::vector<double> eps, mu;
    std::vector<std::vector<double>> e_amp, m_amp;  // [pool entry][sample]
    GeneralMediumGeom(const double c[3], const double s[3], const int d[3], std::vector<double> r,
                      std::vector<double> rf, const std::vector<AbsMedium>& media)
        : BoxGeom(c, s), rho(std::move(r)), rho_fun(std::move(rf))
    {
        for (int k = 0; k < 3; ++k)
            dim[k] = d[k];
        if (dim[0] < 1 || dim[1] < 1 || dim[2] < 1)
            throw std::runtime_error("Invalid Dimension");
        if (int64_t(rho.size()) != int64_t(dim[0]) * dim[1] * dim[2])
            throw std::runtime_error("general medium box: density dataset size does not match dimension");
        if (rho_fun.empty() || media.size() != rho_fun.size())
            throw std::runtime_error("general medium box: the density function and the media tables differ in size");
        const size_t n = media.size();
        eps.resize(n);
        mu.resize(n);
        e_amp.assign(media[0].e_amp.size(), std::vector<double>(n));
        m_amp.assign(media[0].m_amp.size(), std::vector<double>(n));
        for (size_t i = 0; i < n; ++i) {
            eps[i] = media[i].eps;
            mu[i] = media[i].mu;
            for (size_t p = 0; p < e_amp.size(); ++p)
                e_amp[p][i] = media[i].e_amp[p];
            for (size_t p = 0; p < m_amp.size(); ++p)
                m_amp[p][i] = media[i].m_amp[p];
        }
    }
    bool per_point() const override { return true; }
    // position in the sampled density function of the density at (x, y, z)
    double table_pos(double x, double y, double z) const
    {
        double q[3] = {(z - p1[2]) / (p2[2] - p1[2]) * (dim[2] - 1), (y - p1[1]) / (p2[1] - p1[1]) * (dim[1] - 1),
                       (x - p1[0]) / (p2[0] - p1[0]) * (dim[0] - 1)};
        int dz[3] = {dim[2], dim[1], dim[0]};
        const double val = interpn(rho.data(), dz, q, 3);
        return (val - rho_fun.front()) / (rho_fun.back() - rho_fun.front()) * double(rho_fun.size() - 1);
    }
    double field(const std::vector<double>& tab, double xq) const
    {
        const int n = int(tab.size());
        return interpn(tab.data(), &n, &xq, 1);
    }
    double half_at(double x, double y, double z, int half, double* amp, size_t np) const override
    {
        const double xq = table_pos(x, y, z);
        const std::vector<std::vector<double>>& a = half == 0 ? e_amp : m_amp;
        for (size_t p = 0; p < np; ++p)
            amp[p] = p < a.size() ? field(a[p], xq) : 0.0;
        return field(half == 0 ? eps : mu, xq);
    }
    AbsMedium medium_at(double x, double y, double z) const override
    {
        const double xq = table_pos(x, y, z);
        AbsMedium r;
        r.eps = field(eps, xq);
        r.mu = field(mu, xq);
        for (auto& t : e_amp)
            r.e_amp.push_back(field(t, xq));
        for (auto& t : m_amp)
            r.m_amp.push_back(field(t, xq));
        return r;
    }
};

static std::unique_ptr<Geometry> build_geometry(const Json& g, const std::vector<Sus>* pool, const DatasetMap& ds,
                                                double dx2)
{
    const std::string& t = g.at("type").str();
    double c[3], s[3];
    vec3(g.at("center"), c);
    for (auto& v : c)
        v /= dx2;
    if (t == "box" || t == "two medium box" || t == "general medium box") {
        vec3(g.at("size"), s);
        for (auto& v : s)
            v /= dx2;
    }
    if (t == "box") {
        auto b = std::make_unique<BoxGeom>(c, s);
        b->med = abstract_medium(pool, json2medium(g.at("medium")));
        return b;
    }
    if (t == "sphere") {
        auto b = std::make_unique<SphereGeom>(c, g.at("radius").num() / dx2);
        b->med = abstract_medium(pool, json2medium(g.at("medium")));
        return b;
    }
    if (t == "two medium box") {
        int d[3] = {g.at("dimension")[0].integer(), g.at("dimension")[1].integer(), g.at("dimension")[2].integer()};
        const std::string& name = g.at("density dataset").str();
        auto it = ds.find(name);
        if (it == ds.end())
            throw std::runtime_error("dataset \"" + name + "\" was not provided");
        return std::make_unique<TwoMediumGeom>(c, s, d, abstract_medium(pool, json2medium(g.at("medium1"))),
                                               abstract_medium(pool, json2medium(g.at("medium2"))), it->second);
    }
    if (t == "general medium box") {
        int d[3] = {g.at("dimension")[0].integer(), g.at("dimension")[1].integer(), g.at("dimension")[2].integer()};
        auto data = [&](const char* key) -> const std::vector<double>& {
            const std::string& name = g.at(key).str();
            auto it = ds.find(name);
            if (it == ds.end())
                throw std::runtime_error("dataset \"" + name + "\" was not provided");
            return it->second;
        };
        const std::vector<double>& rho = data("density dataset");
        const std::vector<double>& rho_fun = data("density function dataset");
        const std::vector<double>& epsilon = data("epsilon dataset");
        const std::vector<double>& mu = data("mu dataset");
        const std::vector<double>& e_amp = data("electric susceptibility amplitudes dataset");
        const std::vector<double>& m_amp = data("magnetic susceptibility amplitudes dataset");
        // the medium with its permittivity, permeability and amplitudes replaced
        // by each sample's (src/simulation.cpp:498-511)
        Medium m = json2medium(g.at("medium"));
        const size_t ne = m.e_amp.size(), nm = m.m_amp.size();
        if (mu.size() < epsilon.size() || e_amp.size() < epsilon.size() * ne || m_amp.size() < epsilon.size() * nm)
            throw std::runtime_error("general medium box: a medium dataset is shorter than the epsilon dataset");
        std::vector<AbsMedium> media;
        for (size_t i = 0; i < epsilon.size(); ++i) {
            m.eps = epsilon[i];
            m.mu = mu[i];
            for (size_t j = 0; j < ne; ++j)
                m.e_amp[j] = e_amp[i * ne + j];
            for (size_t j = 0; j < nm; ++j)
                m.m_amp[j] = m_amp[i * nm + j];
            media.push_back(abstract_medium(pool, m));
        }
        return std::make_unique<GeneralMediumGeom>(c, s, d, rho, rho_fun, media);
    }
    throw std::runtime_error("geometry type \"" + t + "\" is not supported");
}

// ---------------------------------------------------------------- layout
static void make_layout(Setup& S)
{
    Layout& L = S.L;
    for (int a = 0; a < 3; ++a) {
        int m = 1 - L.P1[a];
        L.M[a] = (m & 1) ? m : m + 1;
        L.n[a] = ((L.P2[a] + 1 + L.M[a]) >> 1) + 1;
    }
    L.pitch = ((int64_t(L.n[0]) + 1 + 31) / 32) * 32;
    L.rows = L.n[1] + 1;
    L.plane = L.pitch * L.rows;
    L.nplanes = L.n[2] + 2;
    L.alloc = L.plane * L.nplanes;
    // the first row of the allocation is row j = -1 of plane k = -1, so every
    // index with i, j, k >= -1 (and the packed last line of that plane's first
    // rows) lies inside it
    L.origin = L.plane + L.pitch;
    // pack the last line of every row when it would hold only its last two indices
    // (grids of 32k cells along x); FFIP_PACK_TAIL=0 keeps one line per row
    const char* pk = std::getenv("FFIP_PACK_TAIL");
    const bool pack = !(pk && std::atoi(pk) == 0);
    L.tail0 = Layout::NO_TAIL;
    if (pack && L.n[0] > 2 && (L.n[0] - 2) % 32 == 0 && L.pitch == L.n[0] - 2 + 32)
        L.tail0 = L.n[0] - 2;
    for (int c = 0; c < 6; ++c)
        for (int a = 0; a < 3; ++a) {
            int p = comp_parity(c, a);
            int first = parity_of(L.P1[a]) == p ? L.P1[a] : L.P1[a] + 1;
            int last = parity_of(L.P2[a]) == p ? L.P2[a] : L.P2[a] - 1;
            if (first > last) {
                L.lo[c][a] = 0;
                L.hi[c][a] = -1;
            } else {
                L.lo[c][a] = L.ix(a, first);
                L.hi[c][a] = L.ix(a, last);
            }
        }
}

// ---------------------------------------------------------------- PML
// Simulation::set_pmls (src/simulation.cpp:655-726), PML getters
// (src/fields.cpp:485-506), classification (include/fields.hpp:233-262).
static void set_pmls(const Json* pmls, Setup& S)
{
    const Layout& L = S.L;
    int sim_dim[3];
    for (int a = 0; a < 3; ++a)
        sim_dim[a] = L.P2[a] - L.P1[a] + 1;
    std::vector<double> k[3], b[3], c[3];
    for (int a = 0; a < 3; ++a) {
        k[a].assign(sim_dim[a], 1.0);
        b[a].assign(sim_dim[a], 0.0);
        c[a].assign(sim_dim[a], 0.0);
    }
    const double dt = S.dt, dx2 = S.dx2;
    if (pmls)
        for (auto& p : pmls->elems()) {
            double d = p.at("thickness").num();
            double kmax = p.at("k max").num();
            double smax = p.at("sigma max").num();
            int m = p.at("order").integer();
            int dir0 = 0, dir1 = 2, side0 = -1, side1 = 1;
            if (const Json* dj = p.find("direction")) {
                const std::string& s = dj->str();
                dir0 = dir1 = s == "x" ? 0 : (s == "y" ? 1 : (s == "z" ? 2 : -1));
                if (dir0 < 0)
                    throw std::runtime_error("PML direction must be x, y or z");
            }
            if (const Json* sj = p.find("side")) {
                const std::string& s = sj->str();
                if (s == "positive")
                    side0 = side1 = 1;
                else if (s == "negative")
                    side0 = side1 = -1;
                else
                    throw std::runtime_error("PML side must be positive or negative");
            }
            auto sigma = [&](double x) { return std::pow(x / d, m) * smax; };
            auto kk = [&](double x) { return 1 + (kmax - 1) * std::pow(x / d, m); };
            auto bb = [&](double x) { return std::exp(-sigma(x) / kk(x) * dt); };
            auto cc = [&](double x) { return 1 / kk(x) * (bb(x) - 1); };
            for (int dir = dir0; dir <= dir1; ++dir)
                for (int side = side0; side <= side1; side += 2) {
                    double Lh = S.size[dir];
                    for (int i = L.P1[dir]; i <= L.P2[dir]; ++i) {
                        double x;
                        if (side == 1 && (i * dx2 >= Lh / 2 - d))
                            x = i * dx2 - (Lh / 2 - d);
                        else if (side == -1 && (i * dx2 <= -Lh / 2 + d))
                            x = -Lh / 2 + d - i * dx2;
                        else
                            continue;
                        int index = i - L.P1[dir];
                        k[dir][index] = kk(x);
                        b[dir][index] = bb(x);
                        c[dir][index] = cc(x);
                    }
                }
        }
    for (int a = 0; a < 3; ++a)
        for (int p = 0; p < 2; ++p) {
            PmlAxis& A = S.pml[a];
            A.b[p].assign(L.n[a], 0.0);
            A.c[p].assign(L.n[a], 0.0);
            A.ik[p].assign(L.n[a], 1.0);
            A.smap[p].assign(L.n[a], -1);
            int w = 0;
            for (int ix = 0; ix < L.n[a]; ++ix) {
                int x = L.coord(a, ix, p);
                if (x < L.P1[a] || x > L.P2[a])
                    continue;
                int t = x - L.P1[a];
                A.b[p][ix] = b[a][t];
                A.c[p][ix] = c[a][t];
                A.ik[p][ix] = 1 / k[a][t];
                if (c[a][t] != 0)
                    A.smap[p][ix] = w++;
            }
            A.W[p] = w;
        }
}


// ---------------------------------------------------------------- materials
// Structure::set_materials_from_geometry (src/structure.cpp:155-249) with the pole
// masks of get_non_zeros_from_array (src/structure.cpp:252-260).  Constant media
// become entries of a <= 255-entry table (uint8 id per point); points inside a
// two-medium box carry their own interpolated medium (id 255, dense region).
namespace {

struct TableBuilder {
    int half;
    size_t np;
    const std::array<double, 3>* pc;
    HalfMaterial* HM;
    std::vector<std::pair<double, std::vector<double>>> keys;

    int entry(const AbsMedium& m)
    {
        double g = half == 0 ? m.eps : m.mu;
        const std::vector<double>& amp = half == 0 ? m.e_amp : m.m_amp;
        std::vector<double> masked(np, 0.0);
        for (size_t i = 0; i < np; ++i)
            if (std::abs(amp[i]) > 1e-4)
                masked[i] = amp[i];
        for (size_t e = 0; e < keys.size(); ++e)
            if (keys[e].first == g && keys[e].second == masked)
                return int(e);
        if (keys.size() >= 255)
            throw std::runtime_error("more than 255 distinct constant media in one simulation");
        MatTable& T = HM->table;
        keys.emplace_back(g, masked);
        T.inv_eps.push_back(1 / g);
        T.eps.push_back(g);
        T.pole0.push_back(int(T.c1.size()));
        int npl = 0;
        for (size_t i = 0; i < np; ++i)
            if (masked[i] != 0) {
                T.c1.push_back(pc[i][0]);
                T.c2.push_back(pc[i][1]);
                T.sc3.push_back(masked[i] * pc[i][2]);  // amp[n] * sus.c3, src/structure.cpp:79
                ++npl;
            }
        T.npoles.push_back(npl);
        HM->max_poles = std::max(HM->max_poles, npl);
        return int(keys.size() - 1);
    }
};

}  // namespace

static void set_materials(const std::vector<std::unique_ptr<Geometry>>& geoms, Setup& S)
{
    const Layout& L = S.L;
    const AbsMedium bg = abstract_medium(S.pool, S.bg);
    std::vector<std::array<double, 3>> pc[2];
    for (int h = 0; h < 2; ++h) {
        pc[h].resize(S.pool[h].size());
        for (size_t i = 0; i < S.pool[h].size(); ++i)
            abstract_sus(S.pool[h][i], S.dt, pc[h][i][0], pc[h][i][1], pc[h][i][2]);
    }
    if (S.keep_geom_index)
        S.geom_index.assign(6, {});

    for (int h = 0; h < 2; ++h) {
        HalfMaterial& HM = S.mat[h];
        TableBuilder TB{h, S.pool[h].size(), pc[h].data(), &HM, {}};
        const int bg_id = TB.entry(bg);
        std::vector<int> geom_id(geoms.size(), -1);
        for (size_t g = 0; g < geoms.size(); ++g)
            if (!geoms[g]->per_point())
                geom_id[g] = TB.entry(geoms[g]->med);
        bool any_dense = false;
        int dense_const_id[3] = {-1, -1, -1};
        // owner[cc]: first geometry (list order) strictly containing each point of
        // the box ob[cc] that covers every geometry's bounding box; points outside
        // it belong to no geometry, so only the geometries' extent is scanned
        std::vector<RawVec<int>> owner(3);
        int ob[3][3][2];
        bool partial[3] = {false, false, false};  // some real points lie outside ob
        for (int cc = 0; cc < 3; ++cc) {
            const int comp = h * 3 + cc;
            const int lo0 = L.lo[comp][0], lo1 = L.lo[comp][1], lo2 = L.lo[comp][2];
            const int e0 = L.hi[comp][0] - lo0 + 1, e1 = L.hi[comp][1] - lo1 + 1, e2 = L.hi[comp][2] - lo2 + 1;
            for (int a = 0; a < 3; ++a) {
                ob[cc][a][0] = L.lo[comp][a];
                ob[cc][a][1] = L.lo[comp][a] - 1;
            }
            if (e0 <= 0 || e1 <= 0 || e2 <= 0)
                continue;
            int par[3] = {comp_parity(comp, 0), comp_parity(comp, 1), comp_parity(comp, 2)};
            std::vector<std::array<int, 6>> rng(geoms.size());
            std::vector<bool> hit(geoms.size(), false);
            int ub[3][2] = {{1 << 30, -(1 << 30)}, {1 << 30, -(1 << 30)}, {1 << 30, -(1 << 30)}};
            for (size_t g = 0; g < geoms.size(); ++g) {
                double bl[3], bh[3];
                geoms[g]->bbox(bl, bh);
                bool empty = false;
                for (int a = 0; a < 3; ++a) {
                    double lo_c = std::max(bl[a], double(L.P1[a]) - 2), hi_c = std::min(bh[a], double(L.P2[a]) + 2);
                    int first = int(std::floor(lo_c)), last = int(std::ceil(hi_c));
                    int lo_i = L.ix(a, std::max(first, L.P1[a] - 1)), hi_i = L.ix(a, std::min(last, L.P2[a] + 1));
                    rng[g][a] = std::max(lo_i, L.lo[comp][a]);
                    rng[g][3 + a] = std::min(hi_i, L.hi[comp][a]);
                    if (rng[g][a] > rng[g][3 + a])
                        empty = true;
                }
                if (empty)
                    continue;
                hit[g] = true;
                for (int a = 0; a < 3; ++a) {
                    ub[a][0] = std::min(ub[a][0], rng[g][a]);
                    ub[a][1] = std::max(ub[a][1], rng[g][3 + a]);
                }
            }
            if (ub[0][0] > ub[0][1]) {  // no geometry reaches this component
                partial[cc] = true;
                continue;
            }
            for (int a = 0; a < 3; ++a) {
                ob[cc][a][0] = ub[a][0];
                ob[cc][a][1] = ub[a][1];
            }
            const int n0 = ub[0][1] - ub[0][0] + 1, n1 = ub[1][1] - ub[1][0] + 1, n2 = ub[2][1] - ub[2][0] + 1;
            partial[cc] = int64_t(n0) * n1 * n2 < int64_t(e0) * e1 * e2;
            RawVec<int>& own = owner[cc];
            own.resize(size_t(n0) * n1 * n2);  // each plane is set to -1 by its thread below
            // geometries touching each plane, in list order, so that planes can be
            // filled independently with the first-containing-geometry rule
            std::vector<std::vector<int>> on_plane(static_cast<size_t>(n2));
            for (size_t g = 0; g < geoms.size(); ++g)
                if (hit[g])
                    for (int k = rng[g][2]; k <= rng[g][5]; ++k)
                        on_plane[size_t(k - ub[2][0])].push_back(int(g));
            parallel_planes(ub[2][0], ub[2][1], [&](int k) {
                const double z = L.coord(2, k, par[2]);
                std::fill_n(&own[size_t(k - ub[2][0]) * n1 * n0], size_t(n1) * n0, -1);
                for (int g : on_plane[size_t(k - ub[2][0])]) {
                    const std::array<int, 6>& r = rng[g];
                    for (int j = r[1]; j <= r[4]; ++j) {
                        double y = L.coord(1, j, par[1]);
                        int* row = &own[(size_t(k - ub[2][0]) * n1 + (j - ub[1][0])) * n0 - ub[0][0]];
                        for (int i = r[0]; i <= r[3]; ++i) {
                            if (row[i] >= 0)
                                continue;
                            double x = L.coord(0, i, par[0]);
                            if (geoms[g]->inside(x, y, z))
                                row[i] = g;
                        }
                    }
                }
            });
            // index into own of lattice point (i, j, k) inside ob
            auto oix = [&](int i, int j, int k) {
                return (size_t(k - ub[2][0]) * n1 + (j - ub[1][0])) * n0 + (i - ub[0][0]);
            };
            // dense region bounding box: points owned by two-medium boxes, searched
            // within the union of those geometries' ranges
            DenseRegion& D = HM.dense[cc];
            int bl[3] = {1 << 30, 1 << 30, 1 << 30}, bh[3] = {-1, -1, -1};
            int pl[3] = {1 << 30, 1 << 30, 1 << 30}, ph[3] = {-1, -1, -1};
            for (size_t g = 0; g < geoms.size(); ++g)
                if (hit[g] && geoms[g]->per_point())
                    for (int a = 0; a < 3; ++a) {
                        pl[a] = std::min(pl[a], rng[g][a]);
                        ph[a] = std::max(ph[a], rng[g][3 + a]);
                    }
            if (ph[2] >= pl[2]) {
                std::vector<char> pp(geoms.size());
                for (size_t g = 0; g < geoms.size(); ++g)
                    pp[g] = geoms[g]->per_point() ? 1 : 0;
                // per plane: x and y extent of the per-point-medium points (empty: lo > hi)
                const int np = ph[2] - pl[2] + 1;
                std::vector<std::array<int, 4>> ext(size_t(np), {1 << 30, -1, 1 << 30, -1});
                parallel_planes(pl[2], ph[2], [&](int k) {
                    std::array<int, 4>& e = ext[size_t(k - pl[2])];
                    for (int j = pl[1]; j <= ph[1]; ++j)
                        for (int i = pl[0]; i <= ph[0]; ++i) {
                            const int g = own[oix(i, j, k)];
                            if (g >= 0 && pp[size_t(g)]) {
                                e[0] = std::min(e[0], i);
                                e[1] = std::max(e[1], i);
                                e[2] = std::min(e[2], j);
                                e[3] = std::max(e[3], j);
                            }
                        }
                });
                for (int k = pl[2]; k <= ph[2]; ++k) {
                    const std::array<int, 4>& e = ext[size_t(k - pl[2])];
                    if (e[1] < e[0])
                        continue;
                    bl[0] = std::min(bl[0], e[0]);
                    bh[0] = std::max(bh[0], e[1]);
                    bl[1] = std::min(bl[1], e[2]);
                    bh[1] = std::max(bh[1], e[3]);
                    bl[2] = std::min(bl[2], k);
                    bh[2] = std::max(bh[2], k);
                }
            }
            // per-point media that are provably the same everywhere (mu = 1 on the
            // magnetic side, typically) collapse to a table entry without evaluation
            bool fast_const = bh[0] >= 0, have_val = false;
            double fast_val = 0;
            for (size_t g = 0; g < geoms.size() && fast_const; ++g) {
                double v;
                if (!geoms[g]->per_point())
                    continue;
                if (!geoms[g]->half_constant(h, v) || (have_val && v != fast_val))
                    fast_const = false;
                fast_val = v;
                have_val = true;
            }
            if (fast_const) {
                AbsMedium m;
                m.eps = m.mu = fast_val;
                m.e_amp.assign(S.pool[0].size(), 0.0);
                m.m_amp.assign(S.pool[1].size(), 0.0);
                dense_const_id[cc] = TB.entry(m);
            } else if (bh[0] >= 0) {
                D.present = true;
                for (int a = 0; a < 3; ++a) {
                    D.box[a][0] = bl[a];
                    D.box[a][1] = bh[a];
                }
                const size_t np = S.pool[h].size();
                D.npoles = int(np);
                for (size_t i = 0; i < np; ++i) {
                    D.c1.push_back(pc[h][i][0]);
                    D.c2.push_back(pc[h][i][1]);
                    D.c3.push_back(pc[h][i][2]);
                }
                int64_t n = D.npts();
                D.inv_eps.assign(n, 1.0);
                D.eps.assign(n, 1.0);
                D.amp.assign(n * np, 0.0);
                int d0 = bh[0] - bl[0] + 1, d1 = bh[1] - bl[1] + 1;
                parallel_planes(bl[2], bh[2], [&](int k) {
                    std::vector<double> amp(np);
                    for (int j = bl[1]; j <= bh[1]; ++j)
                        for (int i = bl[0]; i <= bh[0]; ++i) {
                            int g = own[oix(i, j, k)];
                            if (g < 0 || !geoms[g]->per_point())
                                continue;
                            double gi = geoms[g]->half_at(L.coord(0, i, par[0]), L.coord(1, j, par[1]),
                                                          L.coord(2, k, par[2]), h, amp.data(), np);
                            int64_t q = (int64_t(k - bl[2]) * d1 + (j - bl[1])) * d0 + (i - bl[0]);
                            D.inv_eps[q] = 1 / gi;
                            D.eps[q] = gi;
                            for (size_t p = 0; p < np; ++p)
                                D.amp[q * np + p] = std::abs(amp[p]) > 1e-4 ? amp[p] : 0.0;
                        }
                });
                // a region whose points all carry the same pole-free medium (for
                // instance mu = 1 on the magnetic side) collapses to a table entry
                bool constant = true;
                double g0 = 0;
                bool have = false;
                for (int k = bl[2]; k <= bh[2] && constant; ++k)
                    for (int j = bl[1]; j <= bh[1] && constant; ++j)
                        for (int i = bl[0]; i <= bh[0]; ++i) {
                            int g = own[oix(i, j, k)];
                            if (g < 0 || !geoms[g]->per_point())
                                continue;
                            int64_t q = (int64_t(k - bl[2]) * d1 + (j - bl[1])) * d0 + (i - bl[0]);
                            for (size_t p = 0; p < np; ++p)
                                if (D.amp[q * np + p] != 0)
                                    constant = false;
                            if (!have) {
                                g0 = D.eps[q];
                                have = true;
                            } else if (D.eps[q] != g0)
                                constant = false;
                            if (!constant)
                                break;
                        }
                if (constant && have) {
                    AbsMedium m;
                    m.eps = m.mu = g0;
                    m.e_amp.assign(S.pool[0].size(), 0.0);
                    m.m_amp.assign(S.pool[1].size(), 0.0);
                    dense_const_id[cc] = TB.entry(m);
                    D = DenseRegion();
                } else {
                    any_dense = true;
                }
            }
        }
        for (int cc = 0; cc < 3; ++cc)
            if (HM.dense[cc].present)
                HM.max_poles = std::max(HM.max_poles, HM.dense[cc].npoles);
        // points outside every owner box take the default medium
        HM.bg_id = bg_id;
        for (int a = 0; a < 3; ++a) {
            HM.mbox[a][0] = 1 << 30;
            HM.mbox[a][1] = -(1 << 30);
        }
        for (int cc = 0; cc < 3; ++cc)
            if (!owner[cc].empty())
                for (int a = 0; a < 3; ++a) {
                    HM.mbox[a][0] = std::min(HM.mbox[a][0], ob[cc][a][0]);
                    HM.mbox[a][1] = std::max(HM.mbox[a][1], ob[cc][a][1]);
                }
        // ids per point; uniform when every point has the same pole-free medium
        bool uniform = !any_dense;
        int seen = -1;
        auto id_of = [&](int cc, int g) {
            if (g < 0)
                return bg_id;
            if (!geoms[g]->per_point())
                return geom_id[g];
            return dense_const_id[cc] >= 0 ? dense_const_id[cc] : 255;
        };
        // the id every point would share: the default's when some real point lies
        // outside the owner box, else that of the first owner entry
        for (int cc = 0; cc < 3 && seen < 0; ++cc) {
            if (partial[cc])
                seen = bg_id;
            else if (!owner[cc].empty())
                seen = id_of(cc, owner[cc][0]);
        }
        for (int cc = 0; cc < 3 && uniform; ++cc) {
            if (partial[cc] && seen != bg_id) {  // real points outside the box take the default medium
                uniform = false;
                break;
            }
            if (owner[cc].empty())
                continue;
            std::vector<int> gid(geoms.size());
            for (size_t g = 0; g < geoms.size(); ++g)
                gid[g] = id_of(cc, int(g));
            const RawVec<int>& own = owner[cc];
            const size_t pn = size_t(ob[cc][0][1] - ob[cc][0][0] + 1) * (ob[cc][1][1] - ob[cc][1][0] + 1);
            std::atomic<bool> differ{false};
            parallel_planes(0, ob[cc][2][1] - ob[cc][2][0], [&](int k) {
                if (differ.load(std::memory_order_relaxed))
                    return;
                const int* q = own.data() + size_t(k) * pn;
                for (size_t e = 0; e < pn; ++e)
                    if ((q[e] < 0 ? bg_id : gid[size_t(q[e])]) != seen) {
                        differ.store(true, std::memory_order_relaxed);
                        return;
                    }
            });
            uniform = !differ.load();
        }
        if (uniform && seen >= 0 && HM.table.npoles[seen] != 0)
            uniform = false;
        if (uniform && seen < 0)
            seen = bg_id;
        HM.uniform = uniform;
        if (uniform) {
            HM.u_inv = HM.table.inv_eps[seen];
            HM.u_eps = HM.table.eps[seen];
            HM.max_poles = 0;
        }
        for (int cc = 0; cc < 3; ++cc) {
            const int comp = h * 3 + cc;
            const int lo0 = L.lo[comp][0], lo1 = L.lo[comp][1], lo2 = L.lo[comp][2];
            const int e0 = L.hi[comp][0] - lo0 + 1, e1 = L.hi[comp][1] - lo1 + 1, e2 = L.hi[comp][2] - lo2 + 1;
            if (S.keep_geom_index) {
                S.geom_index[comp].assign(owner[cc].begin(), owner[cc].end());
                for (int a = 0; a < 3; ++a) {
                    S.geom_box[comp][a][0] = ob[cc][a][0];
                    S.geom_box[comp][a][1] = ob[cc][a][1];
                }
            }
            if (uniform)
                continue;
            // padding cells keep id 0; real points outside ob take the default medium
            HM.ids[cc].resize(size_t(L.alloc));
            uint8_t* ids = HM.ids[cc].data();
            parallel_planes(0, int(L.nplanes) - 1,
                            [&](int p) { std::memset(ids + size_t(p) * L.plane, 0, size_t(L.plane)); });
            if (e0 <= 0 || e1 <= 0 || e2 <= 0)
                continue;
            parallel_planes(lo2, lo2 + e2 - 1, [&](int k) {
                // contiguous up to the packed indices, which are set one by one
                const int e0c = std::max(0, std::min(e0, L.tail0 - lo0));
                for (int j = lo1; j < lo1 + e1; ++j) {
                    if (e0c > 0)
                        std::memset(ids + L.index(lo0, j, k), bg_id, size_t(e0c));
                    for (int i = lo0 + e0c; i < lo0 + e0; ++i)
                        ids[L.index(i, j, k)] = uint8_t(bg_id);
                }
            });
            const int b0 = ob[cc][0][0], b1 = ob[cc][1][0], b2 = ob[cc][2][0];
            const int n0 = ob[cc][0][1] - b0 + 1, n1 = ob[cc][1][1] - b1 + 1;
            parallel_planes(b2, ob[cc][2][1], [&](int k) {
                size_t q = size_t(k - b2) * n1 * n0;
                for (int j = b1; j < b1 + n1; ++j)
                    for (int i = b0; i < b0 + n0; ++i, ++q)
                        ids[L.index(i, j, k)] = uint8_t(id_of(cc, owner[cc][q]));
            });
        }
    }
}

// ---------------------------------------------------------------- sources
// Fields::add_dipole_source_gaussian1, src/fields.cpp:183-207
static void add_dipole_g1(Setup& S, const double pos[3], int ct, double amp, double start, double freq, double cutoff)
{
    const Layout& L = S.L;
    amp = -amp;
    auto base = nearest_point(pos, ct, -1);
    auto w = interp_weights((pos[0] - base[0]) / 2, (pos[1] - base[1]) / 2, (pos[2] - base[2]) / 2);
    int half = ctype_is_e(ct) ? 0 : 1;
    int comp = half * 3 + ctype_dir(ct);
    double width = gaussian_width(freq);
    double end = start + 2 * width * cutoff;
    if (end > S.turnoff_time)
        S.turnoff_time = end;
    int index = 0;
    for (int k = 0; k < 4; k += 2)
        for (int j = 0; j < 4; j += 2)
            for (int i = 0; i < 4; i += 2) {
                int p[3] = {base[0] + i, base[1] + j, base[2] + k};
                bool inside = true;
                for (int a = 0; a < 3; ++a)
                    inside = inside && p[a] >= L.P1[a] && p[a] <= L.P2[a];
                if (inside && w[index] > 1e-4)
                    S.dip[half].push_back({comp, L.ix(0, p[0]), L.ix(1, p[1]), L.ix(2, p[2]), amp * w[index], start,
                                           end, width});
                ++index;
            }
}

// Simulation::set_volume_source / set_inhom_source, src/simulation.cpp:235-336
static void set_source(Setup& S, const Json& src, const DatasetMap& ds)
{
    const std::string& t = src.at("type").str();
    double center[3], size[3];
    vec3(src.at("center"), center);
    vec3(src.at("size"), size);
    for (int a = 0; a < 3; ++a) {
        center[a] /= S.dx2;
        size[a] /= S.dx2;
    }
    int ct = ctype_from_name(src.at("field component").str());
    const Json& fn = src.at("function");
    const std::string& ft = fn.at("type").str();
    if (ft != "Gaussian1")
        throw std::runtime_error("only Gaussian1 source functions are supported (Gaussian2 is defective in the "
                                 "reference, src/fields.cpp:229-230)");
    double freq = fn.at("frequency").num();
    double cutoff = fn.at("cutoff").num();
    double stime = fn.at("start time").num();
    double p1[3];
    for (int a = 0; a < 3; ++a)
        p1[a] = center[a] - size[a] / 2;
    if (t == "volume source") {
        double amp = src.at("amplitude").num();
        int count[3];
        double dlen[3];
        for (int a = 0; a < 3; ++a) {
            count[a] = int(cround(size[a] / 2));
            if (count[a] != 0) {
                dlen[a] = size[a] / count[a];
                amp *= dlen[a] / 2;
            } else
                dlen[a] = 0;
        }
        if (count[0] < 0 || count[1] < 0 || count[2] < 0)
            return;
        for (int k = 0; k <= count[2]; ++k)
            for (int j = 0; j <= count[1]; ++j)
                for (int i = 0; i <= count[0]; ++i) {
                    double p[3] = {p1[0] + i * dlen[0], p1[1] + j * dlen[1], p1[2] + k * dlen[2]};
                    add_dipole_g1(S, p, ct, amp, stime, freq, cutoff);
                }
    } else if (t == "inhomogeneous source") {
        int dim[3] = {src.at("dimension")[0].integer(), src.at("dimension")[1].integer(),
                      src.at("dimension")[2].integer()};
        double fcen = src.at("frequency").num();
        auto ia = ds.find(src.at("amplitude dataset").str());
        auto ip = ds.find(src.at("phase dataset").str());
        if (ia == ds.end() || ip == ds.end())
            throw std::runtime_error("inhomogeneous source datasets were not provided");
        int64_t n = int64_t(dim[0]) * dim[1] * dim[2];
        if (int64_t(ia->second.size()) < n || int64_t(ip->second.size()) < n)
            throw std::runtime_error("inhomogeneous source dataset is smaller than its dimension");
        double dlen[3] = {0, 0, 0};
        for (int a = 0; a < 3; ++a)
            if (dim[a] > 1)
                dlen[a] = size[a] / (dim[a] - 1);
        int64_t index = 0;
        for (int k = 0; k <= dim[2] - 1; ++k)
            for (int j = 0; j <= dim[1] - 1; ++j)
                for (int i = 0; i <= dim[0] - 1; ++i) {
                    double p[3] = {p1[0] + i * dlen[0], p1[1] + j * dlen[1], p1[2] + k * dlen[2]};
                    add_dipole_g1(S, p, ct, ia->second[index], stime - ip->second[index] / (2 * PI * fcen), freq,
                                  cutoff);
                    ++index;
                }
    } else {
        throw std::runtime_error("Unknown source type \"" + t + "\"");
    }
}

// ---------------------------------------------------------------- DFT
// Volume_Fields_DFT (src/simulation.cpp:103-118), DFT_Hub::register_volume_dft
// (src/dft.cpp:253-271), DFT_Unit::set_local_region (src/dft.cpp:163-180)
static void register_dft(Setup& S, const Json& fo)
{
    if (fo.at("type").str() != "volume fields dft")
        throw std::runtime_error("Unable to read fields_output configuration");
    double c[3], s[3], p1[3], p2[3];
    vec3(fo.at("center"), c);
    vec3(fo.at("size"), s);
    for (int a = 0; a < 3; ++a) {
        c[a] /= S.dx2;
        s[a] /= S.dx2;
        p1[a] = c[a] - s[a] / 2;
        p2[a] = c[a] + s[a] / 2;
    }
    int ct = ctype_from_name(fo.at("field component").str());
    auto q1 = nearest_point(p1, ct, -1), q2 = nearest_point(p2, ct, +1);
    for (int a = 0; a < 3; ++a)
        if (q1[a] > q2[a])
            throw std::runtime_error("Invali volume in DFT_Unit constructor");
    int half = ctype_is_e(ct) ? 0 : 1;
    int u = -1;
    auto& units = S.units[half];
    for (size_t k = 0; k < units.size(); ++k) {
        auto& U = units[k];
        if (U.ct == ct && U.q1[0] == q1[0] && U.q1[1] == q1[1] && U.q1[2] == q1[2] && U.q2[0] == q2[0] &&
            U.q2[1] == q2[1] && U.q2[2] == q2[2])
            u = int(k);
    }
    if (u < 0) {
        DftUnit U;
        U.ct = ct;
        U.comp = half * 3 + ctype_dir(ct);
        U.is_db = !ctype_is_eh(ct);
        for (int a = 0; a < 3; ++a) {
            U.q1[a] = q1[a];
            U.q2[a] = q2[a];
        }
        bool inter = true;
        for (int a = 0; a < 3; ++a)
            inter = inter && q1[a] <= S.L.P2[a] && S.L.P1[a] <= q2[a];
        if (inter) {
            double i1[3], i2[3];
            for (int a = 0; a < 3; ++a) {
                i1[a] = std::max(q1[a], S.L.P1[a]);
                i2[a] = std::min(q2[a], S.L.P2[a]);
            }
            auto l1 = nearest_point(i1, ct, +1), l2 = nearest_point(i2, ct, -1);
            U.has_local = true;
            for (int a = 0; a < 3; ++a) {
                U.l1[a] = l1[a];
                U.l2[a] = l2[a];
                U.dims[a] = (l2[a] - l1[a]) / 2 + 1;
                if (U.dims[a] <= 0)
                    U.has_local = false;
                U.box[a][0] = S.L.ix(a, l1[a]);
                U.box[a][1] = S.L.ix(a, l2[a]);
            }
        }
        units.push_back(U);
        u = int(units.size() - 1);
    }
    DftOutput o;
    o.group = fo.at("output group").str();
    o.unit = u;
    for (auto& f : fo.at("frequency").elems())
        o.freqs.push_back(f.num());
    for (double f : o.freqs)
        units[u].freqs.push_back(f);
    S.outputs.push_back(o);
    S.output_units.emplace_back(half, u);
}

// ---------------------------------------------------------------- probes
// Yee3::interp, include/utility.hpp:983-1022 (weights and the partially-inside case)
std::vector<ProbeTerm> probe_terms(const Setup& S, const double pos_phys[3], int ct)
{
    if (!ctype_is_eh(ct))
        throw std::runtime_error("probes take E or H components");
    const Layout& L = S.L;
    double pos[3] = {pos_phys[0] / S.dx2, pos_phys[1] / S.dx2, pos_phys[2] / S.dx2};
    auto base = nearest_point(pos, ct, -1);
    auto w = interp_weights((pos[0] - base[0]) / 2, (pos[1] - base[1]) / 2, (pos[2] - base[2]) / 2);
    bool all_in = true, inter = true;
    for (int a = 0; a < 3; ++a) {
        all_in = all_in && L.P1[a] <= base[a] && base[a] + 2 <= L.P2[a];
        inter = inter && base[a] <= L.P2[a] && L.P1[a] <= base[a] + 2;
    }
    int comp = comp_index(ct);
    std::vector<ProbeTerm> out;
    int wi = 0;
    for (int k = 0; k < 4; k += 2)
        for (int j = 0; j < 4; j += 2)
            for (int i = 0; i < 4; i += 2) {
                int p[3] = {base[0] + i, base[1] + j, base[2] + k};
                bool in = true;
                for (int a = 0; a < 3; ++a)
                    in = in && p[a] >= L.P1[a] && p[a] <= L.P2[a];
                if (all_in || (inter && in))
                    out.push_back({comp, L.index(L.ix(0, p[0]), L.ix(1, p[1]), L.ix(2, p[2])), w[wi]});
                ++wi;
            }
    return out;
}

// ---------------------------------------------------------------- boundaries
// Fields::sync_boundary on one rank (src/fields.cpp:327-482): ghost(P1-1) takes
// the value at P2-1 (Period) or mult*value at P1+1 (PEC/PMC); ghost(P2+1) takes
// the value at P1+1 (Period) or mult*value at P2-1.
static void set_faces(Setup& S)
{
    const Layout& L = S.L;
    for (int a = 0; a < 3; ++a)
        for (int side = 0; side < 2; ++side) {
            int bc = S.bc[a][side];
            if (bc == BC_NONE)
                continue;
            int ghost = side == 0 ? L.P1[a] - 1 : L.P2[a] + 1;
            int src;
            double mult = 1;
            if (bc == BC_PERIOD)
                src = side == 0 ? L.P2[a] - 1 : L.P1[a] + 1;
            else {
                src = side == 0 ? L.P1[a] + 1 : L.P2[a] - 1;
                int edge = side == 0 ? L.P1[a] : L.P2[a];
                int phase = bc == BC_PEC ? -1 : 1;
                mult = (edge & 1) ? phase : -phase;
            }
            for (int comp = 0; comp < 6; ++comp) {
                if (comp_parity(comp, a) != parity_of(ghost))
                    continue;
                S.faces[comp / 3].push_back({comp, a, L.ix(a, ghost), L.ix(a, src), mult});
            }
        }
}

// ---------------------------------------------------------------- geometry info
void geom_info_coords(const Setup& S, int ct, int geom, std::vector<double>& x, std::vector<double>& y,
                      std::vector<double>& z)
{
    int comp = comp_index(ct);
    if (S.geom_index.size() != 6)
        throw std::runtime_error("geometry info was not requested at set-up");
    const std::vector<int>& own = S.geom_index[comp];
    const Layout& L = S.L;
    const int(*box)[2] = S.geom_box[comp];
    int e0 = box[0][1] - box[0][0] + 1, e1 = box[1][1] - box[1][0] + 1;
    int par[3] = {comp_parity(comp, 0), comp_parity(comp, 1), comp_parity(comp, 2)};
    // the box is scanned x fastest, the order of the reference's Yee_Iterator
    for (size_t q = 0; q < own.size(); ++q) {
        if (own[q] != geom)
            continue;
        int i = int(q % e0) + box[0][0];
        int j = int((q / e0) % e1) + box[1][0];
        int k = int(q / (size_t(e0) * e1)) + box[2][0];
        // the reference stores these in std::vector<int> (src/simulation.cpp:201-211)
        x.push_back(double(int(L.coord(0, i, par[0]) * S.dx2)));
        y.push_back(double(int(L.coord(1, j, par[1]) * S.dx2)));
        z.push_back(double(int(L.coord(2, k, par[2]) * S.dx2)));
    }
}

// ---------------------------------------------------------------- entry
void build_setup(const Json& cfg, const DatasetMap& ds, const std::vector<ProbeSpec>& probes, Setup& S)
{
    // Simulation::init, src/simulation.cpp:127-188
    // FFIP_TIMING=1 reports where the host set-up spends its time on stderr
    static const bool report = std::getenv("FFIP_TIMING") != nullptr;
    auto tick = std::chrono::steady_clock::now();
    auto mark = [&](const char* what) {
        if (!report)
            return;
        auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[ffip]   host %s %.2f ms\n", what, std::chrono::duration<double, std::milli>(now - tick).count());
        tick = now;
    };
    S.courant = cfg.num_or("courant number", 0.5);
    S.resolution = cfg.at("resolution").num();
    if (!(S.resolution > 0))
        throw std::runtime_error("resolution must be positive");
    S.dx = 1 / S.resolution;
    S.dt = S.dx * S.courant;
    S.dx2 = S.dx / 2;
    if (const Json* p = cfg.find("precision"))
        S.precision = p->str();
    if (S.precision != "fp64" && S.precision != "fp32")
        throw std::runtime_error("precision must be \"fp64\" or \"fp32\"");
    // set_size, src/simulation.cpp:570-581
    const Json& sz = cfg.at("size");
    for (int a = 0; a < 3; ++a) {
        S.size[a] = sz[a].num();
        int span = int(cround(S.size[a] / S.dx) * 2);
        S.L.P1[a] = -span / 2;
        S.L.P2[a] = span / 2;
    }
    // read_boundary_conditions, src/simulation.cpp:583-653
    for (int a = 0; a < 3; ++a)
        S.bc[a][0] = S.bc[a][1] = BC_NONE;
    if (const Json* sym = cfg.find("symmetry"))
        for (auto& s : sym->elems()) {
            const std::string& d = s.at("direction").str();
            int a = d == "x" ? 0 : (d == "y" ? 1 : (d == "z" ? 2 : -1));
            if (a < 0)
                continue;
            S.L.P1[a] = 0;
            S.bc[a][0] = s.at("phase factor").integer() == 1 ? BC_PMC : BC_PEC;
        }
    if (const Json* pd = cfg.find("periodic"))
        if (pd->is_array())
            for (int a = 0; a < 3 && a < int(pd->size()); ++a)
                if ((*pd)[a].integer()) {
                    if (S.bc[a][0] == BC_PMC || S.bc[a][0] == BC_PEC)
                        S.bc[a][1] = S.bc[a][0];
                    else
                        S.bc[a][0] = S.bc[a][1] = BC_PERIOD;
                }
    for (int a = 0; a < 3; ++a)
        if (S.L.P2[a] < S.L.P1[a])
            throw std::runtime_error("simulation size must not be negative");
    make_layout(S);

    // default material and susceptibility pools, src/simulation.cpp:159-165
    if (const Json* dm = cfg.find("default material"))
        S.bg = json2medium(*dm);
    add_to_pool(S.pool, S.bg);
    set_pmls(cfg.find("PML"), S);
    mark("layout and cpml");

    // geometry, src/simulation.cpp:359-400
    std::vector<std::unique_ptr<Geometry>> geoms;
    if (const Json* gj = cfg.find("geometry")) {
        for (auto& g : gj->elems()) {
            if (g.at("type").str() != "two medium box")
                add_to_pool(S.pool, json2medium(g.at("medium")));
            else {
                add_to_pool(S.pool, json2medium(g.at("medium1")));
                add_to_pool(S.pool, json2medium(g.at("medium2")));
            }
        }
        // The reference builds 31-bit pole masks (src/structure.cpp:261) and is
        // undefined beyond 31 distinct susceptibilities; per-point amplitude
        // vectors have no such limit here.
        if (S.pool[0].size() > 255 || S.pool[1].size() > 255)
            throw std::runtime_error("more than 255 distinct susceptibilities");
        for (auto& g : gj->elems())
            geoms.push_back(build_geometry(g, S.pool, ds, S.dx2));
    }
    mark("geometry");
    if (const Json* gi = cfg.find("geometry info request"))
        for (auto& r : gi->elems()) {
            S.geom_info.push_back({r.at("output group").str(), ctype_from_name(r.at("field component").str()),
                                   r.at("geometry index").integer()});
            S.keep_geom_index = true;
        }
    set_materials(geoms, S);
    mark("materials");

    if (const Json* sj = cfg.find("sources"))
        for (auto& s : sj->elems())
            set_source(S, s, ds);
    mark("sources");
    if (const Json* fo = cfg.find("fields output"))
        for (auto& f : fo->elems())
            register_dft(S, f);
    for (int h = 0; h < 2; ++h)
        for (auto& U : S.units[h]) {  // DFT_Unit::unique_frequencies, src/dft.cpp:146-151
            std::sort(U.freqs.begin(), U.freqs.end());
            U.freqs.erase(std::unique(U.freqs.begin(), U.freqs.end()), U.freqs.end());
        }
    for (auto& p : probes)
        S.probes.push_back(probe_terms(S, p.pos, p.ct));
    set_faces(S);
    mark("dft units, probes and faces");
}

}  // namespace ffip_amd
