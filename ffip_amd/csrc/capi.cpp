// extern "C" boundary of libffip_hip.so (declared in include/ffip_hip.h).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>
#include <stdexcept>
#include <string>

#include "../../include/ffip_hip.h"
#include "engine.hpp"
#include "json.hpp"
#include "setup.hpp"
#include "surrogate.hpp"

namespace ffip_amd {
void stream_bandwidth(int R, int W, size_t bytes, int reps, double out[2]);
void to_convolve_noalloc(const double* in, int64_t n0, int64_t n1, const double* filt, int64_t f0, int64_t f1,
                         bool transpose, double* out, void* stream);
void to_convolve(const double* in, int64_t n0, int64_t n1, const double* filt, int64_t f0, int64_t f1, bool transpose,
                 double* out);
}

using namespace ffip_amd;

struct ffip_sim {
    Json config;
    DatasetMap datasets;
    std::vector<ProbeSpec> probes;
    std::unique_ptr<Engine> engine;
    int64_t trace_cap = 0;
    // geometry info coordinates of a finalized handle, computed at the first query
    std::vector<std::vector<double>> geom_cache[3];
};

static thread_local std::string g_err;

// a call made in the wrong state of the handle (FFIP_ESTATE)
struct StateError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

static int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

template <typename F>
static int guard(F&& f)
{
    try {
        f();
        return FFIP_OK;
    } catch (const StateError& e) {
        return fail(FFIP_ESTATE, e.what());
    } catch (const std::runtime_error& e) {
        std::string m = e.what();
        return fail(m.rfind("HIP error", 0) == 0 ? FFIP_EDEVICE : FFIP_EINVAL, m);
    } catch (const std::exception& e) {
        return fail(FFIP_EINVAL, e.what());
    } catch (...) {
        return fail(FFIP_EINVAL, "unknown error");
    }
}

static Engine& need_engine(ffip_sim* s)
{
    if (!s)
        throw std::runtime_error("null handle");
    if (!s->engine)
        throw StateError("ffip_sim_finalize_setup has not been called");
    return *s->engine;
}

extern "C" {

const char* ffip_last_error(void) { return g_err.c_str(); }

int ffip_sim_create(const char* config_json, ffip_sim** out)
{
    if (!config_json || !out)
        return fail(FFIP_EINVAL, "null argument");
    *out = nullptr;
    return guard([&] {
        auto s = std::make_unique<ffip_sim>();
        s->config = Json::parse(config_json);
        if (!s->config.is_object())
            throw std::runtime_error("configuration must be a JSON object");
        *out = s.release();
    });
}

int ffip_sim_set_dataset(ffip_sim* sim, const char* name, const double* data, size_t n)
{
    if (!sim || !name || (!data && n))
        return fail(FFIP_EINVAL, "null argument");
    if (sim->engine)
        return fail(FFIP_ESTATE, "datasets must be set before ffip_sim_finalize_setup");
    return guard([&] { sim->datasets[name] = std::vector<double>(data, data + n); });
}

int ffip_sim_add_probe(ffip_sim* sim, double x, double y, double z, const char* component)
{
    if (!sim || !component)
        return fail(FFIP_EINVAL, "null argument");
    if (sim->engine)
        return fail(FFIP_ESTATE, "probes must be added before ffip_sim_finalize_setup");
    return guard([&] {
        int ct = ctype_from_name(component);
        if (!ctype_is_eh(ct))
            throw std::runtime_error("probes take E or H components");
        sim->probes.push_back({{x, y, z}, ct});
    });
}

int ffip_sim_finalize_setup(ffip_sim* sim)
{
    if (!sim)
        return fail(FFIP_EINVAL, "null handle");
    if (sim->engine)
        return fail(FFIP_ESTATE, "already finalized");
    return guard([&] {
        // FFIP_TIMING=1 reports the host and device halves of the set-up on stderr
        static const bool report = std::getenv("FFIP_TIMING") != nullptr;
        auto t0 = std::chrono::steady_clock::now();
        auto S = std::make_unique<Setup>();
        build_setup(sim->config, sim->datasets, sim->probes, *S);
        auto t1 = std::chrono::steady_clock::now();
        sim->engine = std::make_unique<Engine>(std::move(S));
        if (sim->trace_cap)
            sim->engine->set_trace_capacity(sim->trace_cap);
        if (report) {
            auto t2 = std::chrono::steady_clock::now();
            std::fprintf(stderr, "[ffip] set-up: host %.1f ms, device %.1f ms\n",
                         std::chrono::duration<double, std::milli>(t1 - t0).count(),
                         std::chrono::duration<double, std::milli>(t2 - t1).count());
        }
    });
}

int ffip_sim_run(ffip_sim* sim, const char* stop_json, int64_t* steps_taken)
{
    if (!sim || !stop_json)
        return fail(FFIP_EINVAL, "null argument");
    if (!sim->engine)
        return fail(FFIP_ESTATE, "ffip_sim_finalize_setup has not been called");
    return guard([&] {
        Engine& E = *sim->engine;
        Json j = Json::parse(stop_json);
        const std::string& t = j.at("type").str();
        const Setup& S = E.setup();
        int64_t n;
        if (t == "time") {
            n = E.run_until_time(j.at("time").num());
        } else if (t == "decay") {
            double pos[3] = {j.at("position")[0].num(), j.at("position")[1].num(), j.at("position")[2].num()};
            n = E.run_until_decay(pos, ctype_from_name(j.at("field component").str()),
                                  j.at("time interval examined").num(), j.at("decayed by").num(), nullptr);
        } else {
            double c[3] = {j.at("center")[0].num(), j.at("center")[1].num(), j.at("center")[2].num()};
            double s[3] = {j.at("size")[0].num(), j.at("size")[1].num(), j.at("size")[2].num()};
            n = E.run_until_dft(c, s, ctype_from_name(j.at("field component").str()), j.at("frequency").num(),
                                j.at("time interval examined").num(), j.at("variation").num(), nullptr);
        }
        (void)S;
        if (steps_taken)
            *steps_taken = n;
    });
}

int ffip_sim_step(ffip_sim* sim, int64_t nsteps)
{
    if (nsteps < 0)
        return fail(FFIP_EINVAL, "negative step count");
    return guard([&] { need_engine(sim).step(nsteps); });
}

int ffip_sim_time_step(const ffip_sim* sim, int64_t* out)
{
    if (!out)
        return fail(FFIP_EINVAL, "null argument");
    return guard([&] { *out = need_engine(const_cast<ffip_sim*>(sim)).time_step(); });
}

int ffip_sim_dft_count(const ffip_sim* sim, int* out)
{
    if (!out)
        return fail(FFIP_EINVAL, "null argument");
    return guard([&] { *out = int(need_engine(const_cast<ffip_sim*>(sim)).setup().outputs.size()); });
}

int ffip_sim_dft_dims(const ffip_sim* sim, int k, int64_t dims[4])
{
    if (!dims)
        return fail(FFIP_EINVAL, "null argument");
    return guard([&] { need_engine(const_cast<ffip_sim*>(sim)).dft_dims(k, dims); });
}

int ffip_sim_dft_group(const ffip_sim* sim, int k, char* buf, size_t cap)
{
    if (!buf || cap == 0)
        return fail(FFIP_EINVAL, "null argument");
    return guard([&] {
        const Setup& S = need_engine(const_cast<ffip_sim*>(sim)).setup();
        if (k < 0 || k >= int(S.outputs.size()))
            throw std::runtime_error("DFT output index out of range");
        std::strncpy(buf, S.outputs[k].group.c_str(), cap - 1);
        buf[cap - 1] = 0;
    });
}

int ffip_sim_get_dft(ffip_sim* sim, int k, double* re, double* im, double* x, double* y, double* z, double* f)
{
    return guard([&] { need_engine(sim).get_dft(k, re, im, x, y, z, f); });
}

int ffip_sim_set_trace_capacity(ffip_sim* sim, int64_t steps)
{
    if (!sim || steps < 0)
        return fail(FFIP_EINVAL, "bad argument");
    return guard([&] {
        sim->trace_cap = steps;
        if (sim->engine)
            sim->engine->set_trace_capacity(steps);
    });
}

int ffip_sim_get_trace(ffip_sim* sim, double* out, int64_t cap_steps, int64_t* nsteps, int* nprobes)
{
    return guard([&] {
        Engine& E = need_engine(sim);
        if (nsteps)
            *nsteps = E.trace_len();
        if (nprobes)
            *nprobes = int(E.setup().probes.size());
        if (out)
            E.get_trace(out, cap_steps);
    });
}

int ffip_sim_probe(ffip_sim* sim, double x, double y, double z, const char* component, double* out)
{
    if (!component || !out)
        return fail(FFIP_EINVAL, "null argument");
    return guard([&] {
        double p[3] = {x, y, z};
        *out = need_engine(sim).probe(p, ctype_from_name(component));
    });
}

int ffip_sim_component_dims(const ffip_sim* sim, const char* component, int64_t dims[3])
{
    if (!component || !dims)
        return fail(FFIP_EINVAL, "null argument");
    return guard([&] {
        int ct = ctype_from_name(component);
        if (!ctype_is_eh(ct))
            throw std::runtime_error("component must be E or H");
        need_engine(const_cast<ffip_sim*>(sim)).component_dims(ct, dims);
    });
}

int ffip_sim_get_component(ffip_sim* sim, const char* component, double* out)
{
    if (!component || !out)
        return fail(FFIP_EINVAL, "null argument");
    return guard([&] {
        int ct = ctype_from_name(component);
        if (!ctype_is_eh(ct))
            throw std::runtime_error("component must be E or H");
        need_engine(sim).get_component(ct, out);
    });
}

// The set-up of a finalized handle, or else one built on the host for this call
// (geometry info needs no device state).
static const Setup& setup_for_query(const ffip_sim* sim, std::unique_ptr<Setup>& tmp)
{
    if (!sim)
        throw std::invalid_argument("null handle");
    if (sim->engine)
        return sim->engine->setup();
    tmp = std::make_unique<Setup>();
    build_setup(sim->config, sim->datasets, sim->probes, *tmp);
    return *tmp;
}

// Coordinates of geometry info request k: cached on a finalized handle (its
// set-up no longer changes), recomputed on the host before that.
static void geom_coords(const ffip_sim* sim, int k, std::vector<double>& x, std::vector<double>& y,
                        std::vector<double>& z)
{
    std::unique_ptr<Setup> tmp;
    const Setup& S = setup_for_query(sim, tmp);
    if (k < 0 || k >= int(S.geom_info.size()))
        throw std::runtime_error("geometry info request index out of range");
    if (!sim->engine) {
        geom_info_coords(S, S.geom_info[k].ct, S.geom_info[k].geom, x, y, z);
        return;
    }
    auto& c = const_cast<ffip_sim*>(sim)->geom_cache;
    if (c[0].size() != S.geom_info.size())
        for (auto& v : c)
            v.assign(S.geom_info.size(), {});
    if (c[0][k].empty() && c[1][k].empty())
        geom_info_coords(S, S.geom_info[k].ct, S.geom_info[k].geom, c[0][k], c[1][k], c[2][k]);
    x = c[0][k];
    y = c[1][k];
    z = c[2][k];
}

int ffip_sim_geom_info_count(const ffip_sim* sim, int k, int64_t* npts)
{
    if (!npts)
        return fail(FFIP_EINVAL, "null argument");
    return guard([&] {
        std::vector<double> x, y, z;
        geom_coords(sim, k, x, y, z);
        *npts = int64_t(x.size());
    });
}

int ffip_sim_get_geom_info(const ffip_sim* sim, int k, double* x, double* y, double* z)
{
    return guard([&] {
        std::vector<double> vx, vy, vz;
        geom_coords(sim, k, vx, vy, vz);
        if (x)
            std::memcpy(x, vx.data(), vx.size() * 8);
        if (y)
            std::memcpy(y, vy.data(), vy.size() * 8);
        if (z)
            std::memcpy(z, vz.data(), vz.size() * 8);
    });
}

int ffip_sim_step_timed(ffip_sim* sim, int64_t nsteps, int per_kernel, double out[3])
{
    if (!out || nsteps < 0)
        return fail(FFIP_EINVAL, "bad argument");
    return guard([&] {
        KernelTiming t = need_engine(sim).step_timed(nsteps, per_kernel != 0);
        out[0] = t.total_ms;
        out[1] = t.e_ms;
        out[2] = t.h_ms;
    });
}

namespace {

struct JsonOut {
    std::string s;
    void num(double v)
    {
        char b[40];
        std::snprintf(b, sizeof b, "%.17g", v);
        s += b;
    }
    void inum(int64_t v) { s += std::to_string(v); }
    void key(const char* k)
    {
        s += '"';
        s += k;
        s += "\":";
    }
};

std::string plan_json(const Setup& S)
{
    JsonOut o;
    const Layout& L = S.L;
    o.s += "{";
    // polarization storage the engine would use for each half
    o.key("polarization");
    o.s += "[";
    for (int h = 0; h < 2; ++h) {
        PolLayoutPlan P = plan_polarization(S, h);
        char b[160];
        std::snprintf(b, sizeof b, "%s{\"compact\":%d,\"full_bytes\":%.17g,\"compact_bytes\":%.17g}", h ? "," : "",
                      P.compact ? 1 : 0, P.full_bytes, P.compact_bytes);
        o.s += b;
    }
    o.s += "],";
    o.key("P1");
    o.s += "[" + std::to_string(L.P1[0]) + "," + std::to_string(L.P1[1]) + "," + std::to_string(L.P1[2]) + "],";
    o.key("P2");
    o.s += "[" + std::to_string(L.P2[0]) + "," + std::to_string(L.P2[1]) + "," + std::to_string(L.P2[2]) + "],";
    o.key("M");
    o.s += "[" + std::to_string(L.M[0]) + "," + std::to_string(L.M[1]) + "," + std::to_string(L.M[2]) + "],";
    o.key("n");
    o.s += "[" + std::to_string(L.n[0]) + "," + std::to_string(L.n[1]) + "," + std::to_string(L.n[2]) + "],";
    o.key("lo");
    o.s += "[";
    for (int c = 0; c < 6; ++c)
        o.s += std::string(c ? "," : "") + "[" + std::to_string(L.lo[c][0]) + "," + std::to_string(L.lo[c][1]) + "," +
               std::to_string(L.lo[c][2]) + "]";
    o.s += "],";
    o.key("hi");
    o.s += "[";
    for (int c = 0; c < 6; ++c)
        o.s += std::string(c ? "," : "") + "[" + std::to_string(L.hi[c][0]) + "," + std::to_string(L.hi[c][1]) + "," +
               std::to_string(L.hi[c][2]) + "]";
    o.s += "],";
    o.key("dt");
    o.num(S.dt);
    o.s += ",";
    o.key("slab");  // [axis][parity] -> list of indices with c != 0, and their (b, c, ik)
    o.s += "[";
    for (int a = 0; a < 3; ++a) {
        o.s += a ? ",[" : "[";
        for (int p = 0; p < 2; ++p) {
            o.s += p ? ",[" : "[";
            bool first = true;
            for (int q = 0; q < L.n[a]; ++q)
                if (S.pml[a].smap[p][q] >= 0) {
                    o.s += first ? "[" : ",[";
                    first = false;
                    o.inum(q);
                    o.s += ",";
                    o.num(S.pml[a].b[p][q]);
                    o.s += ",";
                    o.num(S.pml[a].c[p][q]);
                    o.s += ",";
                    o.num(S.pml[a].ik[p][q]);
                    o.s += "]";
                }
            o.s += "]";
        }
        o.s += "]";
    }
    o.s += "],";
    o.key("dipoles");
    o.s += "[";
    for (int h = 0; h < 2; ++h)
        for (size_t q = 0; q < S.dip[h].size(); ++q) {
            const Dipole& d = S.dip[h][q];
            o.s += (h == 0 && q == 0) || (h == 1 && q == 0 && S.dip[0].empty()) ? "[" : ",[";
            o.inum(d.comp);
            o.s += ",";
            o.inum(d.i);
            o.s += ",";
            o.inum(d.j);
            o.s += ",";
            o.inum(d.k);
            o.s += ",";
            o.num(d.amp);
            o.s += ",";
            o.num(d.start);
            o.s += ",";
            o.num(d.end);
            o.s += ",";
            o.num(d.width);
            o.s += "]";
        }
    o.s += "],";
    o.key("units");
    o.s += "[";
    bool first = true;
    for (int h = 0; h < 2; ++h)
        for (auto& U : S.units[h]) {
            o.s += first ? "{" : ",{";
            first = false;
            o.key("comp");
            o.inum(U.comp);
            o.s += ",";
            o.key("has_local");
            o.inum(U.has_local ? 1 : 0);
            o.s += ",";
            o.key("box");
            o.s += "[" + std::to_string(U.box[0][0]) + "," + std::to_string(U.box[0][1]) + "," +
                   std::to_string(U.box[1][0]) + "," + std::to_string(U.box[1][1]) + "," +
                   std::to_string(U.box[2][0]) + "," + std::to_string(U.box[2][1]) + "],";
            o.key("nf");
            o.inum(int64_t(U.freqs.size()));
            o.s += "}";
        }
    o.s += "],";
    o.key("materials");  // per half: uniform flag, table (g_inf, npoles), per-component id histogram
    o.s += "[";
    for (int h = 0; h < 2; ++h) {
        const HalfMaterial& M = S.mat[h];
        o.s += h ? ",{" : "{";
        o.key("uniform");
        o.inum(M.uniform ? 1 : 0);
        o.s += ",";
        o.key("default_id");
        o.inum(M.bg_id);
        o.s += ",";
        o.key("mbox");  // lattice box holding every non-default point
        o.s += "[";
        for (int a = 0; a < 3; ++a) {
            o.s += a ? ",[" : "[";
            o.inum(M.mbox[a][0]);
            o.s += ",";
            o.inum(M.mbox[a][1]);
            o.s += "]";
        }
        o.s += "],";
        o.key("table");
        o.s += "[";
        for (size_t e = 0; e < M.table.eps.size(); ++e) {
            o.s += e ? ",[" : "[";
            o.num(M.table.eps[e]);
            o.s += ",";
            o.inum(M.table.npoles[e]);
            o.s += "]";
        }
        o.s += "],";
        o.key("hist");
        o.s += "[";
        for (int c = 0; c < 3; ++c) {
            std::vector<int64_t> hist(256, 0);
            const int comp = h * 3 + c;
            if (!M.uniform)
                for (int k = L.lo[comp][2]; k <= L.hi[comp][2]; ++k)
                    for (int j = L.lo[comp][1]; j <= L.hi[comp][1]; ++j)
                        for (int i = L.lo[comp][0]; i <= L.hi[comp][0]; ++i)
                            ++hist[M.ids[c][L.index(i, j, k)]];
            o.s += c ? ",[" : "[";
            for (int e = 0; e < 256; ++e) {
                if (e)
                    o.s += ",";
                o.inum(hist[e]);
            }
            o.s += "]";
        }
        o.s += "],";
        o.key("dense_points");
        o.s += "[";
        for (int c = 0; c < 3; ++c) {
            if (c)
                o.s += ",";
            o.inum(M.dense[c].present ? M.dense[c].npts() : 0);
        }
        o.s += "]}";
    }
    o.s += "],";
    o.key("faces");
    o.inum(int64_t(S.faces[0].size() + S.faces[1].size()));
    o.s += ",";
    o.key("probe_terms");
    o.s += "[";
    for (size_t p = 0; p < S.probes.size(); ++p) {
        o.s += p ? ",[" : "[";
        for (size_t t = 0; t < S.probes[p].size(); ++t) {
            o.s += t ? ",[" : "[";
            o.inum(S.probes[p][t].comp);
            o.s += ",";
            o.inum(S.probes[p][t].index);
            o.s += ",";
            o.num(S.probes[p][t].w);
            o.s += "]";
        }
        o.s += "]";
    }
    o.s += "]}";
    return o.s;
}

}  // namespace

int ffip_sim_plan(const ffip_sim* sim, char* buf, size_t cap, size_t* needed)
{
    if (!sim)
        return fail(FFIP_EINVAL, "null handle");
    return guard([&] {
        Setup S;
        build_setup(sim->config, sim->datasets, sim->probes, S);
        std::string j = plan_json(S);
        if (needed)
            *needed = j.size() + 1;
        if (buf && cap) {
            size_t n = std::min(cap - 1, j.size());
            std::memcpy(buf, j.data(), n);
            buf[n] = 0;
        }
    });
}

int ffip_sim_traffic(const ffip_sim* sim, int half, double out[7])
{
    if (!out)
        return fail(FFIP_EINVAL, "null argument");
    return guard([&] { need_engine(const_cast<ffip_sim*>(sim)).traffic(half, out); });
}

int ffip_sim_info(const ffip_sim* sim, double out[6])
{
    if (!out)
        return fail(FFIP_EINVAL, "null argument");
    return guard([&] {
        Engine& E = need_engine(const_cast<ffip_sim*>(sim));
        const Setup& S = E.setup();
        for (int a = 0; a < 3; ++a)
            out[a] = double(S.L.P2[a] - S.L.P1[a]) / 2;
        out[3] = E.device_bytes();
        out[4] = S.dt;
        out[5] = S.dx;
    });
}

void ffip_sim_destroy(ffip_sim* sim) { delete sim; }

int ffip_adjoint_gradient(const ffip_sim* fwd, const ffip_sim* adj, const int fwd_out[3], const int adj_out[3],
                          double freq, const double eps_prime[2], const double norm[2], const double center[3],
                          const double size[3], const int64_t dim[3], double* grad)
{
    if (!fwd_out || !adj_out || !eps_prime || !norm || !center || !size || !dim || !grad)
        return fail(FFIP_EINVAL, "null argument");
    return guard([&] {
        Engine& F = need_engine(const_cast<ffip_sim*>(fwd));
        Engine& A = need_engine(const_cast<ffip_sim*>(adj));
        F.adjoint_gradient(A, fwd_out, adj_out, freq, eps_prime, norm, center, size, dim, grad);
    });
}

// interpn<N>::transpose with zero padding (include/utility.hpp:702-742, 797-828),
// driven as in app/ffip.cpp:308-352
static void transpose_axis(double* data, const int64_t* dims, const size_t* strides, const double* q, int nd,
                           double val)
{
    const int64_t n = dims[0];
    auto rest = [&](double* ptr, double v) {
        if (nd == 1)
            *ptr += v;
        else
            transpose_axis(ptr, dims + 1, strides + 1, q + 1, nd - 1, v);
    };
    const double xq = q[0];
    if (n == 1) {
        rest(data, val);
        return;
    }
    if (xq <= 0) {
        if (xq <= -1)
            return;
        rest(data, val * (1 + xq));
        return;
    }
    if (xq >= double(n - 1)) {
        if (xq >= double(n))
            return;
        rest(data + strides[0] * size_t(n - 1), val * (double(n) - xq));
        return;
    }
    const size_t index = size_t(xq);
    const double tx = xq - double(index);
    rest(data + strides[0] * (index + 1), val * tx);
    rest(data + strides[0] * index, val * (1 - tx));
}

int ffip_transpose(const double* gx, int64_t nx, const double* gy, int64_t ny, const double* gz, int64_t nz,
                   const double* pts, int64_t npts, double* out)
{
    if (!gx || !gy || !gz || !out || (npts > 0 && !pts))
        return fail(FFIP_EINVAL, "null argument");
    if (nx < 1 || ny < 1 || nz < 1)
        return fail(FFIP_EINVAL, "some grid size is 0");
    return guard([&] {
        double dx = nx > 1 ? gx[1] - gx[0] : 1, dy = ny > 1 ? gy[1] - gy[0] : 1, dz = nz > 1 ? gz[1] - gz[0] : 1;
        std::fill(out, out + nx * ny * nz, 0.0);
        const int64_t dims[3] = {nz, ny, nx};
        const size_t strides[3] = {size_t(ny * nx), size_t(nx), 1};
        for (int64_t i = 0; i < npts; ++i) {
            const double* p = pts + 4 * i;
            const double q[3] = {(p[2] - gz[0]) / dz, (p[1] - gy[0]) / dy, (p[0] - gx[0]) / dx};
            transpose_axis(out, dims, strides, q, 3, p[3]);
        }
    });
}

int ffip_to_convolve(const double* in, int64_t n0, int64_t n1, const double* filt, int64_t f0, int64_t f1,
                     int transpose, double* out)
{
    if (!in || !filt || !out)
        return fail(FFIP_EINVAL, "null argument");
    return guard([&] { to_convolve(in, n0, n1, filt, f0, f1, transpose != 0, out); });
}

int ffip_to_convolve_noalloc(const double* in, int64_t n0, int64_t n1, const double* filt, int64_t f0, int64_t f1,
                             int transpose, double* out, void* stream)
{
    if (!in || !filt || !out)
        return fail(FFIP_EINVAL, "null argument");
    return guard([&] {
        to_convolve_noalloc(in, n0, n1, filt, f0, f1, transpose != 0, out, stream);
    });
}

int ffip_stream_bandwidth(int reads, int writes, size_t bytes_per_array, int reps, double out[2])
{
    if (!out)
        return fail(FFIP_EINVAL, "null argument");
    return guard([&] { stream_bandwidth(reads, writes, bytes_per_array, reps, out); });
}

// ------------------------------------------------------------ surrogate kernels
int ffip_sg_gemm(int ta, int tb, int64_t M, int64_t N, int64_t K, const void* A, int64_t lda, const void* B,
                 int64_t ldb, float* C, int64_t ldc, int accumulate, int ksplit, void* stream)
{
    if ((M > 0 && N > 0 && (!C || (K > 0 && (!A || !B)))) || lda < 0 || ldb < 0 || ldc < N)
        return fail(FFIP_EINVAL, "bad GEMM arguments");
    return guard([&] { sg_gemm(ta, tb, M, N, K, A, lda, B, ldb, C, ldc, accumulate, ksplit, stream); });
}

int ffip_sg_gemm_ws(int ta, int tb, int64_t M, int64_t N, int64_t K, const void* A, int64_t lda, const void* B,
                    int64_t ldb, float* C, int64_t ldc, int accumulate, int ksplit, float* work, int64_t work_elems,
                    void* stream)
{
    if ((M > 0 && N > 0 && (!C || (K > 0 && (!A || !B)))) || lda < 0 || ldb < 0 || ldc < N || work_elems < 0 ||
        (work_elems > 0 && !work))
        return fail(FFIP_EINVAL, "bad GEMM arguments");
    return guard([&] {
        sg_gemm(ta, tb, M, N, K, A, lda, B, ldb, C, ldc, accumulate, ksplit, stream, work, work_elems);
    });
}

int ffip_sg_im2col(const int64_t geom[11], const void* X, void* col, void* stream)
{
    if (!geom || !X || !col)
        return fail(FFIP_EINVAL, "null argument");
    return guard([&] { sg_im2col(geom, X, col, stream); });
}

int ffip_sg_col2im(const int64_t geom[11], const float* dcol, float* dX, void* stream)
{
    if (!geom || !dcol || !dX)
        return fail(FFIP_EINVAL, "null argument");
    return guard([&] { sg_col2im(geom, dcol, dX, stream); });
}

int ffip_sg_bias_act(int64_t rows, int64_t n, const float* Y, const float* bias, int relu, void* out, float* zkeep,
                     void* stream)
{
    if (!Y || !bias || !out || rows < 0 || n < 1)
        return fail(FFIP_EINVAL, "bad arguments");
    return guard([&] { sg_bias_act(rows, n, Y, bias, relu, out, zkeep, stream); });
}

int ffip_sg_act_back(int64_t total, const float* dA, const float* Z, int relu, void* dZb, float* dZf, void* stream)
{
    if (!dA || !Z || !dZb || !dZf || total < 0)
        return fail(FFIP_EINVAL, "bad arguments");
    return guard([&] { sg_act_back(total, dA, Z, relu, dZb, dZf, stream); });
}

int ffip_sg_colsum(int64_t rows, int64_t n, const float* dZ, float* db, int accumulate, void* stream)
{
    if (!dZ || !db || rows < 0 || n < 1)
        return fail(FFIP_EINVAL, "bad arguments");
    return guard([&] { sg_colsum(rows, n, dZ, db, accumulate, stream); });
}

int ffip_sg_to_bf16(int64_t n, const float* x, void* y, void* stream)
{
    if (!x || !y || n < 0)
        return fail(FFIP_EINVAL, "bad arguments");
    return guard([&] { sg_to_bf16(n, x, y, stream); });
}

int ffip_sg_head(const int64_t h[9], const void* A, float* out, void* stream)
{
    if (!h || !A || !out)
        return fail(FFIP_EINVAL, "null argument");
    if (h[5] < 0 || h[6] < 0 || h[5] + h[7] > h[2] || h[6] + h[8] > h[3])
        return fail(FFIP_EINVAL, "head crop outside the activation");
    return guard([&] { sg_head(h, A, out, stream); });
}

int ffip_sg_head_back(const int64_t h[9], const float* dout, float* dA, void* stream)
{
    if (!h || !dout || !dA)
        return fail(FFIP_EINVAL, "null argument");
    if (h[5] < 0 || h[6] < 0 || h[5] + h[7] > h[2] || h[6] + h[8] > h[3])
        return fail(FFIP_EINVAL, "head crop outside the activation");
    return guard([&] { sg_head_back(h, dout, dA, stream); });
}

int ffip_sg_adam(int64_t n, float* w, const float* g, float* m, float* v, const float hp[4], int64_t t, void* wb,
                 void* stream)
{
    if (!w || !g || !m || !v || !hp || n < 0 || t < 1)
        return fail(FFIP_EINVAL, "bad arguments");
    return guard([&] { sg_adam(n, w, g, m, v, hp, t, wb, stream); });
}

}  // extern "C"
