/* p3m_hip.h -- C ABI of the MI355X (gfx950) P3M k-space force library.
 *
 * Drop-in for the method table of fweik/p3m-standalone: the six method_t
 * objects below have the reference's names, ids, flags and call semantics
 * (types.h:306-316; p3m-ik.c:46-49, p3m-ik-i.c:40-43, p3m-ad.c:39-41,
 * p3m-ad-i.c:39-41, p3m-ik-real.c:43-46, p3m-ad-real.c:39-41), so a program
 * written against the reference links against libp3m_hip.so instead of the
 * reference's p3m-*.o files.  Every stage of Kspace_force runs as a HIP
 * kernel on the GPU; the host side only moves data and launches.
 *
 * All types are byte-identical to the reference's types.h built with
 * DOUBLE_PREC (FLOAT_TYPE = double).  If the reference's types.h has been
 * included first (TYPES_H defined), its definitions are used as they are.
 */
#ifndef P3M_HIP_H
#define P3M_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------- */
/* Reference types (types.h:107-316, FLOAT_TYPE = double)                  */
/* ---------------------------------------------------------------------- */
#ifndef TYPES_H

enum {
    METHOD_P3M_ik = 0,
    METHOD_P3M_ik_i = 1,
    METHOD_P3M_ad = 2,
    METHOD_P3M_ad_i = 3,
    METHOD_EWALD = 4,
    METHOD_P3M_ik_cuda = 5,
    METHOD_P3M_ik_r = 6,
    METHOD_P3M_ad_r = 7
};

/* types.h:126-132 */
typedef struct {
    int size;
    double *x;
    double *y;
    double *z;
    double **fields;
} vector_array_t;

/* types.h:152-159 */
typedef struct {
    vector_array_t *f;   /* total force */
    vector_array_t *f_k; /* k-space part (Kspace_force accumulates here) */
    vector_array_t *f_r; /* real-space part */
} forces_t;

/* types.h:163-183 */
typedef struct {
    double length;
    int nparticles;
    vector_array_t *p;
    double *q;
    double q2;
    vector_array_t *v;
    forces_t *reference;
    double epsilon;
    double energy;
} system_t;

/* types.h:187-197 */
typedef struct {
    int n;
    vector_array_t *p;
    double *q;
    int *id;
} neighbor_list_t;

/* types.h:202-215 */
typedef struct {
    double alpha;
    double rcut;
    double prefactor;
    int mesh;
    int ip;
    int cao;
    int cao3;
    double precision;
    int tuning;
} parameters_t;

/* types.h:217-226 */
typedef struct {
    int cao;
    double **interpol;
    double **interpol_d;
    double (*U_hat)(int, double);
} interpolation_t;

/* types.h:229-235: statistics of one stage's time over the tuning samples */
typedef struct {
    double avg;
    double sgm;
    double min;
    double max;
    int n;
} timing_t;

/* types.h:237-242: t = t_c + t_g + t_f (avg, min and max summed stage-wise) */
typedef struct {
    timing_t t;
    timing_t t_c;
    timing_t t_g;
    timing_t t_f;
} runtime_stat_t;

/* types.h:243-248 */
typedef struct {
    double t;
    double t_c;
    double t_g;
    double t_f;
} runtime_t;

/* types.h:252-286.  The FFTW plan handles are pointer-sized; this library
 * never uses them (its transforms are its own HIP kernels). */
typedef struct {
    int mesh;
    double *G_hat;
    double *Qmesh;
    vector_array_t *Fmesh;
    double *nshift;
    double *Dn;
    double *dQ[2];
    int *ca_ind[2];
    double *cf[2];
    interpolation_t *inter;
    int forward_plans;
    int backward_plans;
    void *forward_plan[3];
    void *backward_plan[3];
    neighbor_list_t *neighbor_list;
    double *self_force_corrections;
    void *method_data;
    runtime_t runtime;
} data_t;

/* types.h:290-303 */
enum {
    METHOD_FLAG_none = 0,
    METHOD_FLAG_ik = 1,
    METHOD_FLAG_ad = 2,
    METHOD_FLAG_interlaced = 4,
    METHOD_FLAG_nshift = 8,
    METHOD_FLAG_G_hat = 16,
    METHOD_FLAG_Qmesh = 32,
    METHOD_FLAG_ca = 64,
    METHOD_FLAG_self_force_correction = 128
};
#define METHOD_FLAG_P3M (METHOD_FLAG_nshift | METHOD_FLAG_G_hat | METHOD_FLAG_Qmesh | METHOD_FLAG_ca)

/* types.h:306-316 */
typedef struct {
    int method_id;
    const char *method_name;
    const char *method_name_short;
    char flags;
    data_t *(*Init)(system_t *, parameters_t *);
    void (*Influence_function)(system_t *, parameters_t *, data_t *);
    void (*Kspace_force)(system_t *, parameters_t *, data_t *, forces_t *);
    double (*Error)(system_t *, parameters_t *);
    double (*Error_k)(system_t *, parameters_t *);
} method_t;

#endif /* TYPES_H */

/* ---------------------------------------------------------------------- */
/* The method table (drop-in objects)                                      */
/* ---------------------------------------------------------------------- */

/* Each object replaces the reference object of the same name:
 *   method_p3m_ik    p3m-ik.c:46-49       ik, c2c
 *   method_p3m_ik_i  p3m-ik-i.c:40-43     ik, c2c, interlaced
 *   method_p3m_ad    p3m-ad.c:39-41       ad, c2c, self-force correction
 *   method_p3m_ad_i  p3m-ad-i.c:39-41     ad, c2c, interlaced
 *   method_p3m_ik_r  p3m-ik-real.c:43-46  ik, r2c
 *   method_p3m_ad_r  p3m-ad-real.c:39-41  ad, r2c, self-force correction
 * Semantics kept from the reference:
 *   Init               allocates the method state and computes the influence
 *                      function at p->alpha (p3m-common.c:145-232);
 *   Influence_function recomputes G (and, for ad / ad-r, the self-force
 *                      coefficients) for the current p->alpha; must follow
 *                      every change of alpha (main.c:378);
 *   Kspace_force       f->f_k += k-space force; ad / ad-r also subtract the
 *                      self force from f->f (p3m-ad-self-forces.c:114-139);
 *                      with p->tuning set, d->runtime.t_c/t_g/t_f receive the
 *                      assignment / FFT+multiply / interpolation times;
 *   Error, Error_k     rms force error estimates (total, k-space only).
 * The data_t returned by Init must be released with p3m_hip_free(), not the
 * reference's Free_data(). */
extern const method_t method_p3m_ik;
extern const method_t method_p3m_ik_i;
extern const method_t method_p3m_ad;
extern const method_t method_p3m_ad_i;
extern const method_t method_p3m_ik_r;
extern const method_t method_p3m_ad_r;

/* Look a method up by id (main.c:314-344 selects by id); NULL if unknown. */
const method_t *p3m_hip_method(int method_id);

/* ---------------------------------------------------------------------- */
/* Status-returning entry points (ABI additions)                           */
/* ---------------------------------------------------------------------- */

enum {
    P3M_HIP_OK = 0,
    P3M_HIP_EINVAL = 1,      /* bad argument: cao not in 1..7, odd or too small mesh, ... */
    P3M_HIP_ERANGE = 2,      /* a k-space position outside [-L, 2L) or not finite */
    P3M_HIP_ENOMEM = 3,      /* device or host allocation failed */
    P3M_HIP_EDEVICE = 4,     /* a HIP runtime call failed */
    P3M_HIP_ENODEVICE = 5    /* no usable GPU */
};

/* Arithmetic of the mesh stages. */
enum {
    P3M_HIP_FP64 = 0,   /* everything in FP64: reproduces the reference to ~1e-12 */
    P3M_HIP_MIXED = 1   /* mesh, FFT, interpolation in FP32 with compensated
                           accumulation; influence function, self force and
                           reductions in FP64 */
};

/* Message of the last failing call on this thread ("" if none). */
const char *p3m_hip_last_error(void);

/* Init of method `method_id` with the given precision; *out receives the
 * data_t.  Replaces method.Init (p3m-common.c:145-232 + Init_*). */
int p3m_hip_init(int method_id, system_t *s, parameters_t *p, int precision, data_t **out);
/* Influence_function with a status (p3m-ik.c:135-173, p3m-ik-i.c:126-166,
 * p3m-ad.c:111-158, p3m-ad-i.c:121-150). */
int p3m_hip_influence_function(system_t *s, parameters_t *p, data_t *d);
/* Kspace_force with a status (p3m-ik.c:179-245, p3m-ik-i.c:169-246,
 * p3m-ad.c:161-210, p3m-ad-i.c:153-203, p3m-ik-real.c:83-147,
 * p3m-ad-real.c:69-118).  Host arrays in, host arrays accumulated. */
int p3m_hip_kspace_force(system_t *s, parameters_t *p, data_t *d, forces_t *f);
/* Frees the data_t and its device state.  Replaces Free_data
 * (p3m-common.c:234-287) for data_t objects made by this library. */
int p3m_hip_free(data_t *d);
/* Change the precision of an initialised data_t. */
int p3m_hip_set_precision(data_t *d, int precision);

/* Device-resident Kspace_force: the same computation on device arrays
 * (SoA, length n, FP64) already in GPU memory; fk (3n: x, y, z blocks)
 * receives += k-space force, f (3n) -= self force for ad / ad-r (may be
 * NULL for other methods).  Runs on the library's stream; returns once the
 * work is enqueued.  `n` must equal s->nparticles given to Init. */
int p3m_hip_kspace_force_device(data_t *d, int n, double length, const double *x,
                                const double *y, const double *z, const double *q,
                                double *fk, double *f);
/* Wait for all work enqueued on the library's stream for d. */
int p3m_hip_synchronize(data_t *d);

/* Real-space part on the GPU: the pair set and arithmetic of Realteil
 * (realpart.c:124-167): every i != j with minimum-image distance <= rcut
 * (rcut <= L/2).  f->f_r += forces, s->energy += energy. */
int p3m_hip_realteil(system_t *s, parameters_t *p, forces_t *f);
/* Device-resident real-space part: fr (3n) += forces; energy (device
 * double, may be NULL) += energy.  Uses the stream of d. */
int p3m_hip_realteil_device(data_t *d, int n, double length, const double *x, const double *y,
                            const double *z, const double *q, double rcut, double alpha,
                            double *fr, double *energy);
/* Device-resident Calculate_forces (common.c:268-327): zero f, fk, fr
 * (3n each, device, FP64, x/y/z blocks); fr += real space if rcut != 0;
 * fk += k space (f -= self force for ad / ad-r); f += fk + fr.  energy
 * (device double, may be NULL) receives the real-space energy.  Enqueued on
 * the library's stream; see p3m_hip_synchronize. */
int p3m_hip_calculate_forces_device(data_t *d, int n, double length, const double *x,
                                    const double *y, const double *z, const double *q, double rcut,
                                    double alpha, double *f, double *fk, double *fr,
                                    double *energy);
/* Calculate_forces (common.c:268-327) on the GPU: zero f, f_k, f_r; real
 * space if p->rcut != 0; Kspace_force; f += f_k + f_r. */
int p3m_hip_calculate_forces(const method_t *m, system_t *s, parameters_t *p, data_t *d,
                             forces_t *f);

/* Adjoint (vector-Jacobian product) of Kspace_force; no counterpart in the
 * reference, whose path has no derivatives (the differentiable form of the
 * k-space force asked of this library).  For a cotangent dF of the k-space
 * force -- f_k plus, for ad / ad-r, the self-force term Kspace_force
 * subtracts from f (p3m-ad-self-forces.c:114-139) -- adds
 * sum_{i,d} dF_{i,d} dF_{i,d}/dx_j to dpos (3 arrays) and dF/dq_j to dq.
 * The field meshes are recomputed from the positions (nothing of a forward
 * call is stored).  Position derivatives are those of the B-spline weights at
 * the quantised offsets the forward force uses (charge-assign.c:94-97). */
int p3m_hip_kspace_force_vjp(system_t *s, parameters_t *p, data_t *d, const vector_array_t *dF,
                             vector_array_t *dpos, double *dq);
/* Device-resident adjoint: gF (3n, x/y/z blocks) is the cotangent; dpos
 * (3n) and dq (n) are accumulated (+=).  Enqueued on the library's stream. */
int p3m_hip_kspace_force_vjp_device(data_t *d, int n, double length, const double *x,
                                    const double *y, const double *z, const double *q,
                                    const double *gF, double *dpos, double *dq);

/* Adjoint of the real-space force (Realteil, realpart.c:124-167); no
 * counterpart in the reference (the differentiable form asked of this
 * library, for losses over total forces).  Over the pair set of Realteil,
 * with f_ij = q_i q_j s(r) d and s = (erfc(alpha r)/r + 2 alpha/sqrt(pi)
 * exp(-alpha^2 r^2)) / r^2: for a cotangent dF of f_r, adds
 * sum_j q_i q_j (s dg + s'(r)/r (d . dg) d) to dpos_i and
 * sum_j q_j s (d . dg) to dq_i, dg = dF_i - dF_j.  FP64.  Host arrays. */
int p3m_hip_realteil_vjp(system_t *s, parameters_t *p, const vector_array_t *dF,
                         vector_array_t *dpos, double *dq);
/* Device-resident: gF (3n) in, dpos (3n) and dq (n) accumulated; enqueued
 * on the library's stream of d. */
int p3m_hip_realteil_vjp_device(data_t *d, int n, double length, const double *x, const double *y,
                                const double *z, const double *q, double rcut, double alpha,
                                const double *gF, double *dpos, double *dq);

/* Error estimates with a status (Error / Error_k of the method table):
 * realpart.c:343-346 with p3m-ik.c:249-301, p3m-ik-i.c:295-320,409-414,
 * p3m-ad.c:436-505, p3m-ad-i.c:205-294.  For methods 0, 1 and 6 at the
 * reference's 83 tabulated meshes (q.c:1) and alpha*L <= 100 the k-space
 * estimate comes from the built-in Q_ik / Q_ik_i tables, interpolated as
 * find_error.c:84-107 does; otherwise the k-space sums run on the GPU. */
int p3m_hip_error(int method_id, system_t *s, parameters_t *p, double *total, double *kspace);
/* Replace one row of the tabulated ik / ik-i error kernel (101 values at
 * alpha*L = 0..100) for (mesh, cao); kind 0 = Q_ik (methods 0, 6), 1 =
 * Q_ik_i (method 1).  No counterpart in the reference (its tables are
 * linked in); a registered row takes precedence over the built-in one, and
 * also serves meshes the built-in tables do not hold. */
int p3m_hip_set_error_table(int kind, int mesh, int cao, const double *table101);

/* FP64 Ewald sum on sampled particles: the reference's Ewald method
 * (ewald.c:127-172 k space, realpart.c:124-167 real space) evaluated only for
 * the particles idx[0..nsample) of a system in device memory, as the accuracy
 * oracle of large systems.  f_total (device, 3 nsample, x/y/z blocks) += the
 * Ewald force (lattice vectors n^2 <= kmax^2 at alpha, minimum-image pairs
 * within rcut <= L/2); f_kspace (may be NULL) += that force minus the
 * real-space sum at alpha_split within rcut, i.e. the k-space force at
 * alpha_split when rcut covers the erfc range of alpha_split.  Synchronous. */
int p3m_hip_ewald_sample_device(int n, double length, const double *x, const double *y,
                                const double *z, const double *q, int nsample, const int *idx,
                                double alpha, double rcut, int kmax, double alpha_split,
                                double *f_total, double *f_kspace);
/* Allocation-free form of p3m_hip_ewald_sample_device.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops).  Same arguments and
 * result, enqueued on `stream` (a hipStream_t; NULL = the null stream)
 * without allocating, synchronising or creating streams; the caller waits
 * on the stream.  Its working memory is one device array of the library,
 * ~12 MiB per GPU, held until the process ends; limits: kmax <= 63
 * (<= 2^18 lattice vectors of the half sphere) and nsample <= 32768,
 * beyond which it returns P3M_HIP_EINVAL.  The first call on a device
 * resolves the array's address there (make it outside graph capture).
 * Calls that overlap on one GPU share the array and race. */
int p3m_hip_ewald_sample_device_noalloc(int n, double length, const double *x, const double *y,
                                        const double *z, const double *q, int nsample,
                                        const int *idx, double alpha, double rcut, int kmax,
                                        double alpha_split, double *f_total, double *f_kspace,
                                        void *stream);

/* ---------------------------------------------------------------------- */
/* Batches of independent systems                                          */
/* ---------------------------------------------------------------------- */

/* No counterpart in the reference, whose driver evaluates one system at a
 * time (main.c:361-386); for callers with many small systems (ensembles,
 * replicas, parameter scans) that cannot fill a GPU one by one.  All systems
 * share one method id and precision; each has its own size, box, mesh, order,
 * alpha and cutoff.  Particles are concatenated system by system: system s
 * holds slots [n_0 + ... + n_{s-1}, + n_s) of x, y, z, q (device, FP64), and
 * force arrays use the same slots in three blocks (x, y, z) of the total
 * count.  Positions must lie in [0, L_s).  Every stage is one launch over all
 * systems, on the batch's own stream. */
typedef struct {
    int nparticles;
    double length;
    int mesh; /* even, factors 2, 3, 5, 7, 11, 13 */
    int cao;  /* 1..7 */
    double alpha;
    double rcut; /* 0: no real-space part; else <= length / 2 */
} p3m_hip_batch_system;
typedef struct p3m_hip_batch p3m_hip_batch;

/* Make a batch; computes every system's influence function (and, for ad /
 * ad-r, its self-force coefficients) as Influence_function does. */
int p3m_hip_batch_create(int method_id, int precision, int nsys, const p3m_hip_batch_system *sys,
                         p3m_hip_batch **out);
int p3m_hip_batch_free(p3m_hip_batch *b);
/* New alpha for system s (recomputes its influence function). */
int p3m_hip_batch_set_alpha(p3m_hip_batch *b, int s, double alpha);
/* Kspace_force of every system: fk += k-space force, f -= self force (ad,
 * ad-r).  Enqueued on the batch's stream. */
int p3m_hip_batch_kspace_force_device(p3m_hip_batch *b, const double *x, const double *y,
                                      const double *z, const double *q, double *fk, double *f);
/* Calculate_forces (common.c:268-327) of every system: zero f, fk, fr;
 * fr += real space (systems with rcut > 0); fk += k space; f += fk + fr;
 * energy[s] (device, nsys doubles, may be NULL) receives system s's
 * real-space energy.  Enqueued on the batch's stream. */
int p3m_hip_batch_calculate_forces_device(p3m_hip_batch *b, const double *x, const double *y,
                                          const double *z, const double *q, double *f, double *fk,
                                          double *fr, double *energy);
/* Capture p3m_hip_batch_calculate_forces_device with these arrays into a HIP
 * graph; p3m_hip_batch_replay launches it (same arrays, new contents). */
int p3m_hip_batch_capture(p3m_hip_batch *b, const double *x, const double *y, const double *z,
                          const double *q, double *f, double *fk, double *fr, double *energy);
int p3m_hip_batch_replay(p3m_hip_batch *b);
int p3m_hip_batch_synchronize(p3m_hip_batch *b);
/* Choose the launch shapes (assignment tile side, real-space grid) by timing
 * force evaluations of the given input; drops a captured graph.  Results are
 * unchanged by the choice (to rounding of the summation order). */
int p3m_hip_batch_autotune(p3m_hip_batch *b, const double *x, const double *y, const double *z,
                           const double *q, double *f, double *fk, double *fr, int *tile_side,
                           int *rs_grid);
/* Mean time (ms, HIP events on the batch stream) of the batched half-shell
 * real-space kernel alone over `reps` eager force evaluations of the given
 * input (the forces are computed as by the captured step).  No counterpart
 * in the reference. */
int p3m_hip_batch_time_pairs(p3m_hip_batch *b, const double *x, const double *y, const double *z,
                             const double *q, double *f, double *fk, double *fr, int reps,
                             double *ms);
/* alpha for a real-space rms force error `precision` at cutoff rcut, the
 * formula of the reference's tuner (tuning.c:293). */
double p3m_hip_tune_alpha(int n, double length, double q2, double rcut, double precision);

/* ---------------------------------------------------------------------- */
/* Tuner                                                                    */
/* ---------------------------------------------------------------------- */

/* Tune (tuning.c:250-370) with the GPU as the timer: for a target rms force
 * error `precision` at the cutoff p->rcut, alpha follows from the real-space
 * estimate (tuning.c:293); meshes of the reference's candidate list
 * (tuning.c:38, 8 .. 300) are scanned upwards and, per mesh, orders from the
 * last accepted one downwards (from 7, or only p->cao when it is nonzero)
 * while Error(mesh, cao) <= precision; every accepted candidate is timed by
 * 50 Kspace_force calls (time_full, tuning.c:92-147: stage times of a data_t
 * made with p->tuning set) and the fastest by summed minimum stage times
 * wins.  The scan stops at the first mesh whose cheapest order (cao 2) is
 * already slower than the best.  On success *p receives mesh, cao, ip, cao3,
 * alpha and precision (= its error estimate) and the winner's statistics are
 * returned; if no candidate meets the precision, p is unchanged and the
 * returned t.avg is -1.  On a failure it prints the reason and exits (as the
 * method table entry points do).  Timed in FP64. */
runtime_stat_t p3m_hip_tune(const method_t *m, system_t *s, parameters_t *p, double precision);

/* The timer of one candidate: fill *out (seconds) for the parameters p
 * (mesh, cao, alpha, rcut set); return P3M_HIP_OK or an error code. */
typedef int (*p3m_hip_tune_timer)(void *ctx, int method_id, system_t *s, parameters_t *p,
                                  runtime_stat_t *out);
/* The scan of p3m_hip_tune with a status and a choice of timer: timer NULL
 * times on the GPU in `form` (P3M_HIP_FP64 or P3M_HIP_MIXED); otherwise
 * timer(ctx, ...) is called for every timed candidate.  *out receives the
 * winner's statistics (t.avg = -1 if none met the precision). */
int p3m_hip_tune_scan(int method_id, system_t *s, parameters_t *p, double precision, int form,
                      p3m_hip_tune_timer timer, void *ctx, runtime_stat_t *out);

/* ---------------------------------------------------------------------- */
/* Inhomogeneous error estimate and error map                               */
/* ---------------------------------------------------------------------- */

/* Generic_error_estimate_inhomo (p3m-common.c:437-719), the driver's
 * optional diagnostic (main.c:436-440): the k-space error of a clustered
 * system from its squared-charge density rho^2 (order-cao assignment of q_i^2
 * on a mesh of `mesh` points, or q2 / mesh^3 everywhere if `uniform`) and the
 * squared error kernel of the aliasing copies m in [-mc, mc]^3 (the copy-free
 * homogeneous kernel vanishes to rounding, as in the reference); *estimate =
 * sqrt(sum_i rms_i / N) with rms_i the gathered error density times q_i^2.
 * Uses p->alpha.  out_file (may be NULL) receives per mesh point
 * "nx ny nz rho^2 dens.re dens.im K^2.re K^2.im dens.re*rho^2" as
 * p3m-common.c:686-696 writes it.  The gather uses the estimate's mesh and
 * order (the reference passes the method's, p3m-common.c:681). */
int p3m_hip_error_inhomo(system_t *s, parameters_t *p, int uniform, int mesh, int cao, int mc,
                         const char *out_file, double *estimate);
/* Error_map (p3m-common.c:402-435): |f_ref->f_k - f->f_k|^2 of every particle
 * assigned with order cao to a complex mesh of `mesh` points;
 * error_mesh (host, 2 mesh^3 doubles, interleaved re/im, x-major) receives
 * it (real parts; imaginary parts 0). */
int p3m_hip_error_map(system_t *s, forces_t *f, forces_t *f_ref, int mesh, int cao,
                      double *error_mesh);

/* ---------------------------------------------------------------------- */
/* On-disk formats of the reference driver                                 */
/* ---------------------------------------------------------------------- */

/* A host system_t laid out as Init_system (common.c:192-220) makes it:
 * positions and reference forces (f, f_k, f_r) as vector_array_t with
 * fields[3] aliasing x, y, z, zeroed; release with p3m_hip_system_free. */
int p3m_hip_system_new(int n, double length, system_t **out);
int p3m_hip_system_free(system_t *s);
/* Read_system (io.c:100-157): "# Teilchenzahl: N", "# Len: L", then N lines
 * "x y z q" (any white space); q2 = sum q^2.  *out is a p3m_hip_system_new
 * system.  Errors are returned (the reference exits). */
int p3m_hip_read_system(const char *path, system_t **out);
/* Write_system (io.c:199-225): the same format, six decimals. */
int p3m_hip_write_system(const system_t *s, const char *path);
/* Read_exact_forces (io.c:70-98) into s->reference->f: per line "i fx fy fz";
 * lines with the seven columns Write_exact_forces writes also fill
 * s->reference->f_k from columns 5-7 (the reference's four-field read loses
 * step on them). */
int p3m_hip_read_exact_forces(system_t *s, const char *path);
/* Write_exact_forces (io.c:159-177): "i fx fy fz fkx fky fkz" of
 * s->reference (total and k-space), six decimals. */
int p3m_hip_write_exact_forces(const system_t *s, const char *path);
/* One row of the driver's out.dat (main.c:452-455): v = alpha, DeltaF,
 * estimate, real-space estimate, k-space estimate, k-space DeltaF, Ewald
 * k-space estimate, inhomogeneous estimate, as "% lf\t% e ... % e\n". */
int p3m_hip_format_out_row(char *buf, size_t size, const double v[8]);

/* Stage times (ms, HIP events on the library's stream) of the last calls:
 * [0] charge assignment, [1] FFTs + influence-function multiply,
 * [2] back-interpolation, [3] whole Kspace_force, [4] last real-space call
 * (0 if none).  Waits for the stream. */
int p3m_hip_stage_times(data_t *d, double out[5]);
/* Per-step stage times without synchronising between steps: after
 * p3m_hip_stage_ring_start(d, n), each of the next n calls of
 * p3m_hip_calculate_forces_device records its stage events into a set of its
 * own (n = 0 stops recording).  p3m_hip_stage_ring_read waits for the stream
 * and writes, for each recorded step s < *nsteps <= max_steps, the five
 * times of p3m_hip_stage_times to out[5 s .. 5 s + 4].  Fills the role of the
 * reference's runtime_t timers (p3m-common.h:48-57) for a sequence of steps. */
int p3m_hip_stage_ring_start(data_t *d, int nsteps);
int p3m_hip_stage_ring_read(data_t *d, double *out, int max_steps, int *nsteps);
/* For the same recorded steps: out[s] = time (ms) of the real-space pair
 * kernel alone (the cell-list build excluded), 0 for a step without real
 * space.  No counterpart in the reference (its runtime_t has no finer
 * timers). */
int p3m_hip_stage_ring_read_pairs(data_t *d, double *out, int max_steps, int *nsteps);
/* Copy the influence function (M^3 doubles, r_ind order) to the host. */
int p3m_hip_get_influence_function(data_t *d, double *G);
/* Copy the self-force coefficients (27 x 3 doubles) to the host. */
int p3m_hip_get_self_force_coefficients(data_t *d, double *b);

#ifdef __cplusplus
}
#endif

#endif /* P3M_HIP_H */
