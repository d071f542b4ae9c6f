// Device engine: owns the device state of one simulation and advances it.
// Pure C++ interface (no HIP types) so the C-ABI layer stays plain.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "setup.hpp"

namespace ffip_amd {

struct EngineImpl;

// Polarization storage of one half.  The half-step tiles cut every x-row into
// segments of a fixed number of points.  Full grid: one array per (pole,
// component); a segment fetches a pole's whole span as soon as one of its points
// has that pole.  Compact: only points with poles hold entries, each segment's
// slot-major (pole 0 of its points in x order, then pole 1, ...).  Compact is
// chosen when the full grid would move more than 1.5x its bytes.
struct PolLayoutPlan {
    bool compact = false;
    double full_bytes = 0, compact_bytes = 0;  // estimated P bytes per sweep (one read)
    int dlo[3] = {0, 0, 0}, dhi[3] = {-1, -1, -1};
    int gx = 0, ny = 0, nz = 0;               // segments per row, rows, planes
    int64_t comp_start[3] = {0, 0, 0};        // compact: first entry of each component
    int64_t total = 0;                        // compact: entries of the half
    std::vector<uint32_t> segbase[3];         // compact: first entry of every segment
};
PolLayoutPlan plan_polarization(const Setup& S, int half, int lz = 8);

struct KernelTiming {
    double total_ms = 0;  // wall time of the timed region (HIP events)
    double e_ms = 0;      // summed duration of the electric half-step kernel
    double h_ms = 0;      // summed duration of the magnetic half-step kernel
    int64_t steps = 0;
};

class Engine {
public:
    explicit Engine(std::unique_ptr<Setup> setup);
    ~Engine();

    const Setup& setup() const;
    int64_t time_step() const;

    void step(int64_t nsteps);
    int64_t run_until_time(double t);
    // run_until_dft / run_until_fields_decayed (src/simulation.cpp:782-901)
    int64_t run_until_dft(const double c[3], const double s[3], int ct, double freq, double interval, double var,
                          std::string* log);
    int64_t run_until_decay(const double pos[3], int ct, double interval, double decay_by, std::string* log);

    // timed stepping for measurement (events on the engine's stream)
    KernelTiming step_timed(int64_t nsteps, bool per_kernel);

    // DFT output k (the k-th "fields output" entry): real/imag [nf, nz, ny, nx]
    // over the closure, coordinates and frequencies.
    void dft_dims(int k, int64_t dims[4]) const;
    void get_dft(int k, double* re, double* im, double* x, double* y, double* z, double* f);
    // probe traces [steps][nprobes]
    int64_t trace_len() const;
    void get_trace(double* out, int64_t cap);
    void set_trace_capacity(int64_t steps);
    // Simulation::at (include/utility.hpp:983-1022) for an E/H component
    double probe(const double pos[3], int ct);
    // raw component values on the component lattice, [nz][ny][nx]
    void component_dims(int ct, int64_t dims[3]) const;
    void get_component(int ct, double* out);
    void sync();
    double device_bytes() const;
    // Adjoint-method gradient on the density nodes of a two-medium design box:
    // Re(2 pi i f E_fwd E_adj eps' / norm) at the E points strictly inside the box
    // (forward = this engine, adjoint = adj; their design DFT outputs fwd_out /
    // adj_out are Ex, Ey, Ez), scattered by the adjoint of trilinear interpolation
    // onto dim = (nx, ny, nz) nodes; grad is [nz][ny][nx] on the host.
    void adjoint_gradient(const Engine& adj, const int fwd_out[3], const int adj_out[3], double freq,
                          const double eps_prime[2], const double norm[2], const double center[3],
                          const double size[3], const int64_t dim[3], double* grad);
    // Algorithmic HBM bytes of one launch of the half-step kernel (h = 0 electric,
    // 1 magnetic): out = {fields, material ids, polarization, CPML psi, DFT sums,
    // total, points updated}.  Each array element the half step must read or write
    // counts once; neighbour re-reads and table lookups count zero.
    void traffic(int h, double out[7]) const;
    // whether each step runs as one fused kernel (h = 2 in traffic())
    bool fused() const;

private:
    std::unique_ptr<EngineImpl> p_;
};

}  // namespace ffip_amd
