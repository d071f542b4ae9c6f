// extern "C" boundary of libffip_hip.so (declared in include/ffip_hip.h).
#include <cstring>
#include <exception>
#include <memory>
#include <stdexcept>
#include <string>

#include "../../include/ffip_hip.h"
#include "engine.hpp"
#include "json.hpp"
#include "setup.hpp"

using namespace ffip_amd;

struct ffip_sim {
    Json config;
    DatasetMap datasets;
    std::vector<ProbeSpec> probes;
    std::unique_ptr<Engine> engine;
    int64_t trace_cap = 0;
};

static thread_local std::string g_err;

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
        throw std::logic_error("ffip_sim_finalize_setup has not been called");
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
        auto S = std::make_unique<Setup>();
        build_setup(sim->config, sim->datasets, sim->probes, *S);
        sim->engine = std::make_unique<Engine>(std::move(S));
        if (sim->trace_cap)
            sim->engine->set_trace_capacity(sim->trace_cap);
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

int ffip_sim_geom_info_count(const ffip_sim* sim, int k, int64_t* npts)
{
    if (!npts)
        return fail(FFIP_EINVAL, "null argument");
    return guard([&] {
        const Setup& S = need_engine(const_cast<ffip_sim*>(sim)).setup();
        if (k < 0 || k >= int(S.geom_info.size()))
            throw std::runtime_error("geometry info request index out of range");
        std::vector<double> x, y, z;
        geom_info_coords(S, S.geom_info[k].ct, S.geom_info[k].geom, x, y, z);
        *npts = int64_t(x.size());
    });
}

int ffip_sim_get_geom_info(const ffip_sim* sim, int k, double* x, double* y, double* z)
{
    return guard([&] {
        const Setup& S = need_engine(const_cast<ffip_sim*>(sim)).setup();
        if (k < 0 || k >= int(S.geom_info.size()))
            throw std::runtime_error("geometry info request index out of range");
        std::vector<double> vx, vy, vz;
        geom_info_coords(S, S.geom_info[k].ct, S.geom_info[k].geom, vx, vy, vz);
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

}  // extern "C"
