/*
 * ffip_oracle.c -- CPU restatement (FP64, doubled Yee grid) of the time-stepping
 * hot loop of JoeZeng133/ffip.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP product
 * path and the "port" CPU baseline timed beside it.  Nothing in the product
 * (ffip_amd/) may link, import or call it.
 *
 * Layout follows the reference exactly: one interleaved doubled-coordinate array
 * `eh` (E and H) and one `accdb` (D/dt and B/dt accumulators) over the grid plus a
 * one-point ghost shell, x fastest (src/utility.cpp:130-140, 206-210), with point
 * lists built at setup time (oracle/oracle.py restates Simulation::init).
 *
 * Each stepping routine cites the reference function it restates:
 *   pml_step      <- PML_Stepping::step            src/fields.cpp:89-109
 *   curl_step     <- Curl_Stepping::step           src/fields.cpp:133-148
 *   dipole_step   <- Gaussian_Dipoles_Stepping::step (Gaussian1 only)
 *                                                   src/fields.cpp:17-33, src/utility.cpp:260-264
 *   medium_step   <- Medium_Stepping::step         src/structure.cpp:56-85
 *   sync_step     <- Fields::sync_boundary (single rank: None/Period/PEC/PMC)
 *                                                   src/fields.cpp:327-482
 *   dft_update    <- DFT_Unit::update              src/dft.cpp:29-62
 *   osim_step_m/e <- Simulation::step_m / step_e   src/simulation.cpp:748-766
 *   osim_run      <- Simulation::run_until_time    src/simulation.cpp:768-780
 *
 * Pinning (what fixes each function independently of this restatement) is listed
 * in tests/test_oracle_pins.py: discrete plane-wave dispersion in a periodic box,
 * the discrete susceptibility closed form, the CPML reflection, the DFT of a
 * known series, the Gaussian1 closed form, the reference's Yee3 known answers,
 * mirror planes against full grids, and the current-sheet Green's function.
 *
 * Threads (OpenMP, osim_set_threads): the per-point loops of each phase write
 * distinct points, as one MPI rank's loops do, so any thread count gives the
 * same bits as one.  The dipole list, which may name a point twice, stays serial.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static const double PI = 3.141592653589793e+00; /* src/utility.cpp:5 */

typedef struct {
    int64_t n;
    const int64_t *idx;
    const double *b1, *c1, *ik1, *b2, *c2, *ik2;
    double *phi1, *phi2;
} OPmlList;

typedef struct {
    int64_t n;
    const int64_t *idx;
} OCurlList;

typedef struct {
    int64_t n;
    const int64_t *idx;
    const double *amp, *start, *end, *width;
    double max_end;
} ODipoleList;

typedef struct {
    int64_t npts;
    int npoles;
    const double *c1, *c2, *c3; /* [npoles] */
    const int64_t *idx;         /* [npts] */
    const double *ginf;         /* [npts] */
    const double *amp;          /* [npts * npoles] */
    double *p, *p1;             /* [npts * npoles] */
} OMedGroup;

typedef struct {
    int is_db;
    int nf;
    int64_t n;
    const int64_t *idx; /* [n] Yee_Iterator order */
    const double *freqs;
    double *re, *im; /* [nf * n] */
    double *buf;     /* [n] */
} ODftUnit;

typedef struct {
    int64_t n;           /* points on the face */
    const int64_t *src;  /* extraction face (offset -1) */
    const int64_t *dst;  /* ghost face it is written to */
    double mult;
    double *buf;
} OFaceCopy;

typedef struct {
    double dt, inv_dx;
    int64_t stride[3];
    double *eh, *accdb;
    /* [0] = electric half, [1] = magnetic half */
    OPmlList pml[2][3];
    OCurlList curl[2][3];
    ODipoleList dip[2];
    int ngroups[2];
    OMedGroup *groups[2];
    int nunits[2];
    ODftUnit *units[2];
    /* per axis: two copies (negative ghost, positive ghost); n == 0 -> none */
    OFaceCopy sync[3][2];
    /* probes: nprobes, CSR offsets into (idx, w); trace[step * nprobes + p] */
    int nprobes;
    const int64_t *probe_off;
    const int64_t *probe_idx;
    const double *probe_w;
    double *trace;
    int64_t trace_cap;
    int64_t time_step;
} OSim;

/* src/utility.cpp:260-264 */
static double gaussian1(double t, double width)
{
    double arg = (t / width);
    return arg * exp(-arg * arg / 2);
}

/* src/fields.cpp:89-109; strides negated for magnetic points (src/fields.cpp:174-177) */
static void pml_step(OSim *s, int half)
{
    const double *eh = s->eh;
    double *acc = s->accdb;
    const long long sgn = half == 0 ? 1 : -1;
    for (int i = 0; i < 3; ++i) {
        OPmlList *L = &s->pml[half][i];
        long long st1 = sgn * s->stride[(i + 1) % 3];
        long long st2 = sgn * s->stride[(i + 2) % 3];
#pragma omp parallel for schedule(static)
        for (int64_t q = 0; q < L->n; ++q) {
            int64_t p = L->idx[q];
            double curl1 = (eh[p + st1] - eh[p - st1]) * s->inv_dx;
            double curl2 = (eh[p + st2] - eh[p - st2]) * s->inv_dx;
            L->phi1[q] = L->b1[q] * L->phi1[q] + L->c1[q] * curl1;
            L->phi2[q] = L->b2[q] * L->phi2[q] + L->c2[q] * curl2;
            acc[p] += L->phi1[q] - L->phi2[q] + curl1 * L->ik1[q] - curl2 * L->ik2[q];
        }
    }
}

/* src/fields.cpp:133-148 */
static void curl_step(OSim *s, int half)
{
    const double *eh = s->eh;
    double *acc = s->accdb;
    const long long sgn = half == 0 ? 1 : -1;
    for (int i = 0; i < 3; ++i) {
        OCurlList *L = &s->curl[half][i];
        long long st1 = sgn * s->stride[(i + 1) % 3];
        long long st2 = sgn * s->stride[(i + 2) % 3];
#pragma omp parallel for schedule(static)
        for (int64_t q = 0; q < L->n; ++q) {
            int64_t p = L->idx[q];
            double curl1 = (eh[p + st1] - eh[p - st1]) * s->inv_dx;
            double curl2 = (eh[p + st2] - eh[p - st2]) * s->inv_dx;
            acc[p] += (curl1 - curl2);
        }
    }
}

/* src/fields.cpp:17-33 (Gaussian1 list; Gaussian2 is rejected at setup) */
static void dipole_step(OSim *s, int half, double time)
{
    ODipoleList *L = &s->dip[half];
    if (time > L->max_end)
        return;
    for (int64_t q = 0; q < L->n; ++q) {
        if (time > L->start[q] && time < L->end[q])
            s->accdb[L->idx[q]] += L->amp[q] * gaussian1(time - (L->start[q] + L->end[q]) / 2, L->width[q]);
    }
}

/* src/structure.cpp:56-85 */
static void medium_step(OSim *s, int half)
{
    double *eh = s->eh;
    const double *acc = s->accdb;
    for (int g = 0; g < s->ngroups[half]; ++g) {
        OMedGroup *G = &s->groups[half][g];
        const int np = G->npoles;
#pragma omp parallel for schedule(static)
        for (int64_t q = 0; q < G->npts; ++q) {
            int64_t index = G->idx[q];
            double *p = G->p + q * np;
            double *p1 = G->p1 + q * np;
            const double *amp = G->amp + q * np;
            for (int n = 0; n < np; ++n) {
                double tmp = p[n];
                p[n] = G->c1[n] * p[n] + G->c2[n] * p1[n] + amp[n] * G->c3[n] * eh[index];
                p1[n] = tmp;
            }
            double sum = 0.0; /* std::accumulate(p, p + stride, 0.0) */
            for (int n = 0; n < np; ++n)
                sum += p[n];
            eh[index] = (acc[index] * s->dt - sum) / G->ginf[q];
        }
    }
}

/* src/fields.cpp:327-426 on one rank: both extractions happen before both
 * assignments; None contributes nothing. */
static void sync_step(OSim *s, int axis)
{
    for (int side = 0; side < 2; ++side) {
        OFaceCopy *F = &s->sync[axis][side];
        for (int64_t q = 0; q < F->n; ++q)
            F->buf[q] = s->eh[F->src[q]] * F->mult;
    }
    for (int side = 0; side < 2; ++side) {
        OFaceCopy *F = &s->sync[axis][side];
        for (int64_t q = 0; q < F->n; ++q)
            s->eh[F->dst[q]] = F->buf[q];
    }
}

/* src/dft.cpp:29-62 */
static void dft_update(OSim *s, int half, double time)
{
    for (int u = 0; u < s->nunits[half]; ++u) {
        ODftUnit *U = &s->units[half][u];
        if (U->n == 0)
            continue;
#pragma omp parallel for schedule(static)
        for (int64_t q = 0; q < U->n; ++q)
            U->buf[q] = U->is_db ? s->accdb[U->idx[q]] * s->dt : s->eh[U->idx[q]];
        for (int f = 0; f < U->nf; ++f) {
            double cos_wt = cos(2 * PI * U->freqs[f] * time);
            double sin_wt = -sin(2 * PI * U->freqs[f] * time);
            double *re = U->re + (int64_t)f * U->n;
            double *im = U->im + (int64_t)f * U->n;
#pragma omp parallel for schedule(static)
            for (int64_t q = 0; q < U->n; ++q) {
                im[q] += sin_wt * U->buf[q];
                re[q] += cos_wt * U->buf[q];
            }
        }
    }
}

static void half_step(OSim *s, int half, double src_time, double dft_time)
{
    pml_step(s, half);
    curl_step(s, half);
    dipole_step(s, half, src_time);
    medium_step(s, half);
    for (int a = 0; a < 3; ++a)
        sync_step(s, a);
    dft_update(s, half, dft_time);
}

/* Simulation::step_m, src/simulation.cpp:758-766 */
void osim_step_m(OSim *s)
{
    half_step(s, 1, (s->time_step) * s->dt, (s->time_step + 0.5) * s->dt);
}

/* Simulation::step_e, src/simulation.cpp:748-756 */
void osim_step_e(OSim *s)
{
    half_step(s, 0, (s->time_step + 0.5) * s->dt, (s->time_step + 1) * s->dt);
}

/* Trilinear probe, Yee3::interp (include/utility.hpp:983-1022), sampled after
 * every full step. Weights and indices are prepared at setup. */
static void record_probes(OSim *s)
{
    if (s->nprobes == 0 || s->time_step >= s->trace_cap)
        return;
    for (int p = 0; p < s->nprobes; ++p) {
        double v = 0;
        for (int64_t q = s->probe_off[p]; q < s->probe_off[p + 1]; ++q)
            v += s->eh[s->probe_idx[q]] * s->probe_w[q];
        s->trace[s->time_step * s->nprobes + p] = v;
    }
}

/* Advance exactly nsteps full steps (step_m, step_e, ++time_step), the body of
 * Simulation::run_until_time (src/simulation.cpp:768-780). */
void osim_run_steps(OSim *s, int64_t nsteps)
{
    for (int64_t n = 0; n < nsteps; ++n) {
        osim_step_m(s);
        osim_step_e(s);
        record_probes(s);
        ++s->time_step;
    }
}

/* Simulation::run_until_time: while (time_step * dt < time) step. Returns the
 * number of steps taken. */
int64_t osim_run_until_time(OSim *s, double time)
{
    int64_t taken = 0;
    while ((double)s->time_step * s->dt < time) {
        osim_run_steps(s, 1);
        ++taken;
    }
    return taken;
}

/* threads of the per-point loops (1 when built without OpenMP) */
void osim_set_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

int osim_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

size_t osim_sizeof(void) { return sizeof(OSim); }
size_t osim_sizeof_pml(void) { return sizeof(OPmlList); }
size_t osim_sizeof_group(void) { return sizeof(OMedGroup); }
size_t osim_sizeof_unit(void) { return sizeof(ODftUnit); }
