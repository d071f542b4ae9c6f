/*
 * ffip_hip.h -- C ABI of libffip_hip.so, the MI355X (gfx950) FDTD engine that
 * replaces the time-stepping core of JoeZeng133/ffip.
 *
 * The reference crosses into its C++ core through a process and file boundary:
 * ffip.Simulation.run (ffip/simulation.py:404-431) writes config.json plus an
 * input HDF5 file, runs `mpirun -np N run_sim_json` (app/run_sim_json.cpp:6-25,
 * which calls Simulation::init / run / output, include/simulation.hpp:104-167)
 * and reads the DFT groups back from the output HDF5 file.  This library offers
 * the same steps in-process: the JSON document is the one ffip's Python layer
 * writes, the input datasets are handed over as arrays, and results are copied
 * into caller-owned buffers.
 *
 * Conventions: every call returns 0 on success and a negative code on failure;
 * ffip_last_error() then holds a thread-local message.  The library owns all
 * device memory; the caller owns every host buffer it passes.  A handle owns one
 * HIP stream and must not be used from two threads at once.  The GPU is chosen
 * with HIP_VISIBLE_DEVICES; the reference's `np` / `decomposition` arguments have
 * no counterpart (one GPU holds the whole grid).
 */
#ifndef FFIP_HIP_H
#define FFIP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ffip_sim ffip_sim;

#define FFIP_OK 0
#define FFIP_EINVAL -1   /* bad argument, bad JSON or unsupported configuration */
#define FFIP_ESTATE -2   /* call not valid in the handle's current state */
#define FFIP_EDEVICE -3  /* HIP runtime error */

/* Message of the last failed call on this thread. */
const char *ffip_last_error(void);

/* Parse a configuration (replaces reading ./config.json, app/run_sim_json.cpp:12).
 * An extra key "precision": "fp64" | "fp32" selects the field storage type. */
int ffip_sim_create(const char *config_json, ffip_sim **out);

/* Provide an input dataset by name (replaces the input HDF5 file's datasets that
 * the JSON references: "two medium box <sfx>" densities, ffip/geom.py:432-433;
 * "inhom source amplitude/phase <sfx>", ffip/source.py:127-129). Data is copied. */
int ffip_sim_set_dataset(ffip_sim *sim, const char *name, const double *data, size_t n);

/* Record a time trace of one E/H component at a physical position, sampled after
 * every full step with Simulation::at semantics (src/simulation.cpp:994-997,
 * include/utility.hpp:983-1022).  Must precede ffip_sim_finalize_setup. */
int ffip_sim_add_probe(ffip_sim *sim, double x, double y, double z, const char *component);

/* Build the grid, materials, PML, sources and DFT monitors and upload them to
 * the GPU (the set-up half of Simulation::init, src/simulation.cpp:127-188). */
int ffip_sim_finalize_setup(ffip_sim *sim);

/* Run a stop condition (Simulation::run, src/simulation.cpp:903-958):
 * {"type":"time","time":T} | {"type":"decay",...} | {"type":"dft",...}.
 * Steps continue from the current time step. */
int ffip_sim_run(ffip_sim *sim, const char *stop_condition_json, int64_t *steps_taken);

/* Advance exactly nsteps full steps (step_m then step_e, src/simulation.cpp:748-766). */
int ffip_sim_step(ffip_sim *sim, int64_t nsteps);
int ffip_sim_time_step(const ffip_sim *sim, int64_t *out);

/* DFT outputs, one per "fields output" entry in the order given
 * (Volume_Fields_DFT::output -> DFT_Hub::output_fields_collective, src/dft.cpp:273-320).
 * dims = {nf, nz, ny, nx}; re/im are [nf][nz][ny][nx]; x/y/z/f are the 1-D axes. */
int ffip_sim_dft_count(const ffip_sim *sim, int *out);
int ffip_sim_dft_dims(const ffip_sim *sim, int k, int64_t dims[4]);
int ffip_sim_dft_group(const ffip_sim *sim, int k, char *buf, size_t cap);
int ffip_sim_get_dft(ffip_sim *sim, int k, double *re, double *im, double *x, double *y, double *z, double *f);

/* Probe traces: capacity in steps must be set before stepping; out is [steps][nprobes]. */
int ffip_sim_set_trace_capacity(ffip_sim *sim, int64_t steps);
int ffip_sim_get_trace(ffip_sim *sim, double *out, int64_t cap_steps, int64_t *nsteps, int *nprobes);

/* Simulation::at (src/simulation.cpp:994-997) for an E or H component. */
int ffip_sim_probe(ffip_sim *sim, double x, double y, double z, const char *component, double *out);

/* Raw values of one E/H component on its lattice inside the grid, [nz][ny][nx]
 * (the points Yee_Iterator(grid_p1, grid_p2, ctype) visits, src/utility.cpp:28-125). */
int ffip_sim_component_dims(const ffip_sim *sim, const char *component, int64_t dims[3]);
int ffip_sim_get_component(ffip_sim *sim, const char *component, double *out);

/* Geometry info request k (Simulation::geometry_info_request, src/simulation.cpp:190-233):
 * coordinates of the points of a component inside geometry "geometry index".  Before
 * ffip_sim_finalize_setup these run the host set-up for the call (no GPU needed). */
int ffip_sim_geom_info_count(const ffip_sim *sim, int k, int64_t *npts);
int ffip_sim_get_geom_info(const ffip_sim *sim, int k, double *x, double *y, double *z);

/* Step nsteps with HIP events on the handle's stream: out[0] = elapsed ms of the
 * whole region, out[1] / out[2] = summed ms of the electric / magnetic half-step
 * kernels (only when per_kernel != 0).  When the simulation runs the fused
 * full-step kernel (ffip_sim_traffic half 2 succeeds), out[1] is that kernel's
 * time and out[2] the time of the half-step kernels that cover the CPML shell.
 * No reference counterpart: measurement. */
int ffip_sim_step_timed(ffip_sim *sim, int64_t nsteps, int per_kernel, double out[3]);

/* Host-only dry run of the set-up (no GPU needed): writes a JSON summary of the
 * grid layout, PML slabs, dipole list, DFT units and material table into buf
 * (cap bytes); *needed receives the full length + 1.  Probes added before the
 * call are included.  No reference counterpart: diagnostics and tests. */
int ffip_sim_plan(const ffip_sim *sim, char *buf, size_t cap, size_t *needed);

/* Algorithmic HBM bytes of one launch of the half-step kernel over the whole
 * grid (half 0 = electric, 1 = magnetic) or of the fused full-step kernel over
 * the grid's interior clear of the CPML (half 2; FFIP_EINVAL when the
 * simulation does not run it): out = {fields, material ids, polarization
 * state, CPML auxiliaries, DFT sums, total, points updated}.  No reference
 * counterpart: measurement. */
int ffip_sim_traffic(const ffip_sim *sim, int half, double out[7]);

/* Grid facts: out[0..2] = Yee cells per axis, out[3] = device bytes held,
 * out[4] = dt, out[5] = dx. */
int ffip_sim_info(const ffip_sim *sim, double out[6]);

void ffip_sim_destroy(ffip_sim *sim);

/* Sustained HBM bandwidth of a plain streaming kernel on the current device:
 * `reads` arrays (0-3) are read and `writes` arrays (0-2) written, each of
 * bytes_per_array bytes, 16 bytes per lane per access, `reps` timed launches.
 * out[0] = mean and out[1] = best rate in GB/s of bytes read plus written.
 * Allocates and frees its arrays.  No reference counterpart: the measured
 * ceiling beside which the half-step kernels' byte rates are reported. */
int ffip_stream_bandwidth(int reads, int writes, size_t bytes_per_array, int reps, double out[2]);

/* Adjoint-method gradient of a two-medium design box on the GPU (the computation
 * of Adjoint_Volume.get_sensitivity, ffip/simulation.py:841-869, with the
 * transpose of app/ffip.cpp:308-352 done as a deterministic gather):
 *   grad[nz][ny][nx] = sum over E points p strictly inside the box of
 *       Re(2 pi i f E_fwd(p) E_adj(p) eps' / norm) * trilinear weight(p, node)
 * fwd / adj are finished simulations on the same grid; fwd_out / adj_out are
 * the indices of their Ex, Ey, Ez DFT outputs over the box; eps_prime and norm
 * are complex (re, im); center / size / dim describe the density grid as
 * Two_Medium_Box does (dim = nodes along x, y, z). */
int ffip_adjoint_gradient(const ffip_sim *fwd, const ffip_sim *adj, const int fwd_out[3], const int adj_out[3],
                          double freq, const double eps_prime[2], const double norm[2], const double center[3],
                          const double size[3], const int64_t dim[3], double *grad);

/* Adjoint of zero-padded trilinear interpolation (replaces the pybind helper
 * _ffip.transpose, app/ffip.cpp:308-352): scatters pts[i] = (x, y, z, value) onto
 * the regular grid gx[nx] x gy[ny] x gz[nz] (uniform spacing taken from the first
 * two nodes) and writes out[nz][ny][nx] (overwritten, zero-initialised). */
int ffip_transpose(const double *gx, int64_t nx, const double *gy, int64_t ny, const double *gz, int64_t nz,
                   const double *pts, int64_t npts, double *out);

/* Density filter of filtered-design scripts (replaces the pybind helpers
 * _ffip.TO_convolve / TO_convolve_transpose, app/ffip.cpp:222-306): in and out
 * are host arrays [n0][n1], filt [f0][f1] with odd, equal f0 and f1 (the
 * reference indexes both filter axes with the first radius, so non-square
 * filters read outside it there; rejected here).  transpose = 0: out(i, j) =
 * window-normalised correlation of in with filt; 1: its adjoint.  Computed on
 * the GPU in the reference's summation order. */
int ffip_to_convolve(const double *in, int64_t n0, int64_t n1, const double *filt, int64_t f0, int64_t f1,
                     int transpose, double *out);

/* Allocation-free form of ffip_to_convolve.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  in, filt and out are DEVICE
 * arrays ([n0][n1], [f0][f1] with odd, equal f0 and f1); the call is
 * asynchronous on `stream` (a hipStream_t, NULL = default stream) and allocates
 * nothing: the window weights live in a fixed device array of 2^21 elements
 * (16 MiB per GPU, held until the process ends), so n0 * n1 may not exceed
 * 2^21 (FFIP_EINVAL otherwise).  Calls that overlap on one GPU share that array
 * and race; results equal ffip_to_convolve's bit for bit. */
int ffip_to_convolve_noalloc(const double *in, int64_t n0, int64_t n1, const double *filt, int64_t f0, int64_t f1,
                             int transpose, double *out, void *stream);

/* ------------------------------------------------------------------------
 * Design surrogate kernels.  No counterpart in the reference: a 3-D
 * convolutional network over the design density, trained and evaluated on the
 * matrix cores (BF16 operands, FP32 accumulation), proposes objective and
 * gradient values for inner line-search steps; every design it proposes is
 * re-judged by the full FDTD + adjoint solve.  All pointers are device pointers;
 * `stream` is a hipStream_t (NULL = default stream); nothing is allocated and
 * every call is asynchronous on that stream.  bf16 arrays hold IEEE bfloat16.
 * ------------------------------------------------------------------------ */

/* C (+)= op(A) op(B): op(A) M x K, op(B) K x N (bf16), C FP32 row-major (ldc).
 * ta: A stored K x M (element (m, k) at A[k*lda + m]); tb: B stored N x K.
 * accumulate: add to C instead of overwriting; ksplit > 1 splits K over
 * workgroups (partial sums added atomically). */
int ffip_sg_gemm(int ta, int tb, int64_t M, int64_t N, int64_t K, const void *A, int64_t lda, const void *B,
                 int64_t ldb, float *C, int64_t ldc, int accumulate, int ksplit, void *stream);

/* ffip_sg_gemm whose split-K partial tiles go to a caller-owned device
 * workspace of work_elems floats (at least ksplit * M * N) and are added up in a
 * fixed order, so that repeated calls give the same bits (the atomic adds of
 * ffip_sg_gemm do not).  No reference counterpart (surrogate kernel). */
int ffip_sg_gemm_ws(int ta, int tb, int64_t M, int64_t N, int64_t K, const void *A, int64_t lda, const void *B,
                    int64_t ldb, float *C, int64_t ldc, int accumulate, int ksplit, float *work, int64_t work_elems,
                    void *stream);

/* 3x3x3 convolution taps of a channels-last bf16 activation [batch][z][y][x][C]
 * with stride s and dilation d per axis (padding d, output extent
 * (in - 1) / s + 1): col[(b, oz, oy, ox)][(kz, ky, kx, c)].
 * geom = {batch, C, in_z, in_y, in_x, s_z, s_y, s_x, d_z, d_y, d_x}. */
int ffip_sg_im2col(const int64_t geom[11], const void *X, void *col, void *stream);

/* Adjoint of ffip_sg_im2col on FP32 data: dX (overwritten) = sum of the column
 * entries that read each input element. */
int ffip_sg_col2im(const int64_t geom[11], const float *dcol, float *dX, void *stream);

/* out (bf16) = act(Y + bias) over rows x n (bias per column, act = ReLU if relu);
 * zkeep (FP32, may be NULL) receives Y + bias. */
int ffip_sg_bias_act(int64_t rows, int64_t n, const float *Y, const float *bias, int relu, void *out, float *zkeep,
                     void *stream);

/* dZ = relu ? dA * (Z > 0) : dA, written as bf16 (dZb) and FP32 (dZf). */
int ffip_sg_act_back(int64_t total, const float *dA, const float *Z, int relu, void *dZb, float *dZf, void *stream);

/* db[c] (+)= sum over rows of dZ[rows][n]. */
int ffip_sg_colsum(int64_t rows, int64_t n, const float *dZ, float *db, int accumulate, void *stream);

/* y (bf16) = x (FP32), n elements. */
int ffip_sg_to_bf16(int64_t n, const float *x, void *y, void *stream);

/* Network head: out[b][y][x][c] = mean over z of A[b][z][y0 + y][x0 + x][c]
 * (A bf16 [batch][nz][ny_in][nx_in][C]); h = {batch, nz, ny_in, nx_in, C, y0, x0,
 * ny, nx}.  ffip_sg_head_back writes dA (FP32, overwritten). */
int ffip_sg_head(const int64_t h[9], const void *A, float *out, void *stream);
int ffip_sg_head_back(const int64_t h[9], const float *dout, float *dA, void *stream);

/* Adam step on FP32 weights w with gradient g and moments m, v; hp = {lr,
 * beta1, beta2, eps}; t = step count from 1; wb (may be NULL) receives the
 * updated weights as bf16. */
int ffip_sg_adam(int64_t n, float *w, const float *g, float *m, float *v, const float hp[4], int64_t t, void *wb,
                 void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FFIP_HIP_H */
