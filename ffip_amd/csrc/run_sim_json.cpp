// run_sim_json: the file-based entry point of app/run_sim_json.cpp:6-25 over
// libffip_hip.so.  ffip's Python front end (ffip/simulation.py:404-431) writes
// ./config.json and the input HDF5 file, runs "mpirun -np N run_sim_json" in
// that directory and reads the output HDF5 file; this program keeps all three.
//
//   run_sim_json [CONFIG]        CONFIG defaults to ./config.json
//   run_sim_json --plan [CONFIG] set-up only (no GPU): prints the plan JSON
//
// Files, as Simulation::init / geometry_info_request / output write them:
//   "input file"          opened read-only, created empty when missing
//                         (src/simulation.cpp:151-153); every float64 dataset
//                         under its root is handed to ffip_sim_set_dataset.
//   "fields output file"  created or truncated (:155-157); one group per
//                         "geometry info request" with x, y, z (:190-233), then
//                         one group per "fields output" entry with real and imag
//                         [nf][nz][ny][nx] and the axes x, y, z, f
//                         (src/dft.cpp:273-320).
//
// One GPU holds the whole grid, so under mpirun only rank 0 works; the other
// ranks exit at once with status 0.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ffip_hip.h"
#include "hdf5_io.hpp"
#include "json.hpp"

using ffip_amd::Json;

namespace {

int launcher_rank()
{
    for (const char* v : {"OMPI_COMM_WORLD_RANK", "PMIX_RANK", "PMI_RANK", "MV2_COMM_WORLD_RANK", "SLURM_PROCID"})
        if (const char* s = std::getenv(v))
            return std::atoi(s);
    return 0;
}

void check(int rc, const char* what)
{
    if (rc != 0)
        throw std::runtime_error(std::string(what) + ": " + ffip_last_error());
}

std::string read_text(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f)
        throw std::runtime_error("cannot read " + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

struct Sim {
    ffip_sim* p = nullptr;
    ~Sim()
    {
        if (p)
            ffip_sim_destroy(p);
    }
};

void load_inputs(ffip_sim* sim, const Json& cfg)
{
    ffip_amd::h5::File in(cfg.at("input file").str(), "r+c");
    for (const std::string& name : in.root_datasets()) {
        std::vector<double> v = in.read(name);
        check(ffip_sim_set_dataset(sim, name.c_str(), v.data(), v.size()), "ffip_sim_set_dataset");
    }
}

void write_geometry_info(ffip_sim* sim, const Json& cfg, ffip_amd::h5::File& out)
{
    const Json* req = cfg.find("geometry info request");
    if (!req)
        return;
    for (size_t k = 0; k < req->size(); ++k) {
        int64_t n = 0;
        check(ffip_sim_geom_info_count(sim, int(k), &n), "ffip_sim_geom_info_count");
        std::vector<double> x(n), y(n), z(n);
        check(ffip_sim_get_geom_info(sim, int(k), x.data(), y.data(), z.data()), "ffip_sim_get_geom_info");
        std::string g = (*req)[k].at("output group").str();
        out.create_group(g);
        std::vector<uint64_t> d{uint64_t(n)};
        out.write(g + "/x", d, x.data());
        out.write(g + "/y", d, y.data());
        out.write(g + "/z", d, z.data());
    }
}

void write_dft(ffip_sim* sim, ffip_amd::h5::File& out)
{
    int count = 0;
    check(ffip_sim_dft_count(sim, &count), "ffip_sim_dft_count");
    for (int k = 0; k < count; ++k) {
        int64_t d[4];
        char group[512];
        check(ffip_sim_dft_dims(sim, k, d), "ffip_sim_dft_dims");
        check(ffip_sim_dft_group(sim, k, group, sizeof group), "ffip_sim_dft_group");
        size_t n = size_t(d[0] * d[1] * d[2] * d[3]);
        std::vector<double> re(n), im(n), x(d[3]), y(d[2]), z(d[1]), f(d[0]);
        check(ffip_sim_get_dft(sim, k, re.data(), im.data(), x.data(), y.data(), z.data(), f.data()),
              "ffip_sim_get_dft");
        std::string g = group;
        out.create_group(g);
        out.write(g + "/real", {uint64_t(d[0]), uint64_t(d[1]), uint64_t(d[2]), uint64_t(d[3])}, re.data());
        out.write(g + "/imag", {uint64_t(d[0]), uint64_t(d[1]), uint64_t(d[2]), uint64_t(d[3])}, im.data());
        out.write(g + "/x", {uint64_t(d[3])}, x.data());
        out.write(g + "/y", {uint64_t(d[2])}, y.data());
        out.write(g + "/z", {uint64_t(d[1])}, z.data());
        out.write(g + "/f", {uint64_t(d[0])}, f.data());
    }
}

void describe_stop(const Json& st)
{
    // the messages of Simulation::run (src/simulation.cpp:903-950)
    const std::string& t = st.at("type").str();
    if (t == "time")
        std::printf("Simulation stops at time=%g\n", st.at("time").num());
    else if (t == "decay")
        std::printf("Simulation stops when the %s field at (%g,%g,%g) decays by %g\n",
                    st.at("field component").str().c_str(), st.at("position")[0].num(), st.at("position")[1].num(),
                    st.at("position")[2].num(), st.at("decayed by").num());
    else
        std::printf("Simulation stops when the %s DFT over the region (center=(%g,%g,%g), size=(%g,%g,%g)) varies "
                    "by less than %g\n",
                    st.at("field component").str().c_str(), st.at("center")[0].num(), st.at("center")[1].num(),
                    st.at("center")[2].num(), st.at("size")[0].num(), st.at("size")[1].num(),
                    st.at("size")[2].num(), st.at("variation").num());
    std::fflush(stdout);
}

int run(const std::string& config_path, bool plan_only)
{
    std::string text = read_text(config_path);
    Json cfg = Json::parse(text);
    Sim sim;
    check(ffip_sim_create(text.c_str(), &sim.p), "ffip_sim_create");
    load_inputs(sim.p, cfg);
    if (plan_only) {
        size_t need = 0;
        check(ffip_sim_plan(sim.p, nullptr, 0, &need), "ffip_sim_plan");
        std::string buf(need, '\0');
        check(ffip_sim_plan(sim.p, &buf[0], buf.size(), nullptr), "ffip_sim_plan");
        std::printf("%s\n", buf.c_str());
        return 0;
    }
    auto t0 = std::chrono::steady_clock::now();
    ffip_amd::h5::File out(cfg.at("fields output file").str(), "w");
    check(ffip_sim_finalize_setup(sim.p), "ffip_sim_finalize_setup");
    write_geometry_info(sim.p, cfg, out);
    const Json& stop = cfg.at("stop condition");
    describe_stop(stop);
    int64_t steps = 0;
    check(ffip_sim_run(sim.p, stop.dump().c_str(), &steps), "ffip_sim_run");
    write_dft(sim.p, out);
    double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("Simulation complete, Total time spent = %g s\nLast time step = %lld\n", sec, (long long)steps);
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    bool plan_only = false;
    std::string config = "config.json";
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--plan")
            plan_only = true;
        else if (a == "-h" || a == "--help") {
            std::printf("usage: run_sim_json [--plan] [CONFIG]   (CONFIG defaults to ./config.json)\n");
            return 0;
        } else
            config = a;
    }
    int rank = launcher_rank();
    if (rank > 0) {
        std::printf("rank %d: the whole grid runs on rank 0\n", rank);
        return 0;
    }
    try {
        return run(config, plan_only);
    } catch (std::exception& e) {
        std::fprintf(stderr, "run_sim_json: %s\n", e.what());
        return 1;
    }
}
