/* Internal C ABI between the host C layer (host/) and the HIP kernels.
 *
 * The host layer owns the method logic (which stages run, prefactors,
 * argument checks, the data_t); everything below is a thin launcher around
 * one HIP kernel or runtime call.  Device pointers are plain void* / double*;
 * no HIP type crosses this header.  All launchers return P3MK_OK or an error
 * code and leave a message for p3mk_error().
 */
#ifndef P3M_KERNELS_H
#define P3M_KERNELS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    P3MK_OK = 0,
    P3MK_EINVAL = 1,
    P3MK_ENOMEM = 3,
    P3MK_EDEVICE = 4,
    P3MK_ENODEVICE = 5
};

#define P3MK_NUM_EVENTS 8
/* slots P3MK_EV_PAIRS and P3MK_EV_PAIRS + 1: around the real-space pair kernel */
#define P3MK_EV_PAIRS 6

const char *p3mk_error(void);

/* ---- runtime -------------------------------------------------------- */
typedef struct p3mk_queue p3mk_queue; /* one HIP stream + timing events */

int p3mk_device_check(void);
int p3mk_queue_create(p3mk_queue **q);
void p3mk_queue_destroy(p3mk_queue *q);
int p3mk_queue_sync(p3mk_queue *q);
int p3mk_record(p3mk_queue *q, int slot);
int p3mk_elapsed_ms(p3mk_queue *q, int a, int b, float *ms);
/* Event ring: after p3mk_ring_start(q, n), every p3mk_record also records
 * into set k (k = number of p3mk_ring_advance calls so far, while k < n);
 * n = 0 removes the ring.  Sets are read with p3mk_ring_elapsed_ms. */
int p3mk_ring_start(p3mk_queue *q, int nsets);
void p3mk_ring_advance(p3mk_queue *q);
int p3mk_ring_filled(p3mk_queue *q);
int p3mk_ring_elapsed_ms(p3mk_queue *q, int set, int a, int b, float *ms);
int p3mk_alloc(void **p, size_t bytes);
void p3mk_release(void *p);
int p3mk_alloc_host(void **p, size_t bytes);
void p3mk_release_host(void *p);
int p3mk_copy_h2d(void *dst, const void *src, size_t bytes, p3mk_queue *q);
int p3mk_copy_d2h(void *dst, const void *src, size_t bytes, p3mk_queue *q);
int p3mk_copy_d2d(void *dst, const void *src, size_t bytes, p3mk_queue *q);
/* zero fill by a kernel (safe inside captured graphs) */
int p3mk_fill_zero(void *p, size_t bytes, p3mk_queue *q);
/* y[i] += a * x[i] for i < n (FP64) */
int p3mk_axpy(size_t n, double a, const double *x, double *y, p3mk_queue *q);
/* f[i] += fk[i] + fr[i] for i < n (FP64; Calculate_forces' order) */
int p3mk_sum_parts(size_t n, const double *fk, const double *fr, double *f, p3mk_queue *q);
/* positions inside [lo, hi) (not NaN)?  *bad receives the count of others */
int p3mk_check_positions(int n, double lo, double hi, const double *x, const double *y,
                         const double *z, int *bad, int *scratch, p3mk_queue *q);

/* ---- mesh geometry --------------------------------------------------- */
typedef struct {
    int M;          /* mesh points per side (even) */
    int cao;        /* charge assignment order 1..7 */
    int r2c;        /* 1: real mesh + half spectrum; 0: complex mesh */
    int interlaced; /* second mesh, shifted by h/2, in the imaginary part */
    int fp32;       /* mesh, FFT and interpolation arithmetic in FP32 */
    double L;       /* box length */
} p3mk_geom;

/* bytes of the charge / potential mesh (complex M^3 for c2c, real M^3 for r2c) */
size_t p3mk_mesh_bytes(const p3mk_geom *g);
/* bytes of one spectrum (complex M^3 for c2c, complex M*M*(M/2+1) for r2c) */
size_t p3mk_spec_bytes(const p3mk_geom *g);

/* ---- FFT --------------------------------------------------------------- */
typedef struct p3mk_fft p3mk_fft;
int p3mk_fft_create(p3mk_fft **f, const p3mk_geom *g);
void p3mk_fft_destroy(p3mk_fft *f);
/* forward transform of the mesh along z and y (x is done by convolve):
 * c2c in place (spec == mesh); r2c mesh (real) -> spec */
int p3mk_fft_forward_zy(p3mk_fft *f, void *mesh, void *spec, p3mk_queue *q);
/* x-axis forward pass, multiplication, x-axis backward pass.
 * ad (nout = 1): out[0] <- G * spec (out[0] may equal spec)
 * ik (nout = 3): out[d] <- i (2 pi / L) Dn[n_d] G * spec, d = 0, 1, 2 */
int p3mk_convolve(p3mk_fft *f, int nout, const double *G, const double *Dn, void *spec,
                  void *const out[3], p3mk_queue *q);
/* transpose of the ik gradient: x-axis forward pass of the three spectra
 * in[d] (after p3mk_fft_forward_zy), out <- sum_d -i (2 pi / L) Dn[n_d] G in[d],
 * x-axis backward pass (out may equal in[0]) */
int p3mk_divergence(p3mk_fft *f, const double *G, const double *Dn, void *const in[3], void *out,
                    p3mk_queue *q);
/* backward transform along y and z: c2c in place (mesh == spec);
 * r2c spec -> real mesh */
int p3mk_fft_backward_yz(p3mk_fft *f, void *spec, void *mesh, p3mk_queue *q);
/* plain 3D transforms of a complex M^3 (c2c geometry) array, for tests:
 * sign -1 forward, +1 backward, unnormalised, in place */
int p3mk_fft3d(p3mk_fft *f, void *data, int sign, p3mk_queue *q);

/* ---- particle <-> mesh -------------------------------------------------- */
/* Particles sorted by the mesh tile of their stencil base: made by
 * p3mk_assign, reused by the interpolation of the same particles. */
typedef struct p3mk_psort p3mk_psort;
int p3mk_psort_create(p3mk_psort **s, const p3mk_geom *g, int n);
void p3mk_psort_destroy(p3mk_psort *s);
/* mesh = charge assignment of n particles (both meshes if interlaced);
 * every mesh point is written, no prior zeroing needed */
int p3mk_assign(const p3mk_geom *g, int n, const double *x, const double *y, const double *z,
                const double *q, p3mk_psort *s, void *mesh, p3mk_queue *qu);
/* ik back-interpolation: fk[d*n + i] += -prefac * sum W F_d (meshes F[0..2]) */
/* assignment with the sort of the last p3mk_assign (same particles, n):
 * mode 1: particle i deposits q_i gv[i] W_i; mode 2: q_i sum_d gv[d n + i]
 * W^(d)_i, W^(d) = W with axis d's weight replaced by its derivative in mesh
 * units.  Zero-fills the mesh first. */
int p3mk_assign_weighted(const p3mk_geom *g, int n, p3mk_psort *s, int mode, const double *gv,
                         void *mesh, p3mk_queue *q);
/* adjoint gather of the k-space force for the cotangent gv (3n, caller's
 * order): dq (n) and dpos (3n) += (-prefac') sums over the stencil of the
 * field meshes F (ik: 3 meshes; ad: F[0] = phi) and lambda; for ad, prefac
 * includes the 1/L of the derivative weights and selfb (may be NULL) adds
 * the self-force term. */
int p3mk_adjoint_gather(const p3mk_geom *g, int ad, int n, const p3mk_psort *s, const double *q,
                        const double *gv, void *const F[3], const void *lam, double prefac,
                        const double *selfb, double *dpos, double *dq, p3mk_queue *q_);
int p3mk_interp_ik(const p3mk_geom *g, int n, const p3mk_psort *s, void *const F[3],
                   double prefac, double *fk, p3mk_queue *qu);
/* ad back-interpolation: fk -= prefac * sum phi dQ; with selfb (81 FP64 on
 * the device) also f -= self force */
int p3mk_interp_ad(const p3mk_geom *g, int n, const p3mk_psort *s, const void *phi,
                   double prefac, const double *selfb, double *fk, double *f, p3mk_queue *qu);

/* ---- inhomogeneous error estimate and error map (p3m-common.c:402-719) ---- */
/* complex FP64 M^3 meshes, x-major; scratch of p3mk_inhomo_scratch_doubles() */
int p3mk_inhomo_scratch_doubles(void);
/* homogeneous error kernel K_d(n) = i (-2 pi n_d / L) (U^2 G - G_m), mc = 0 */
int p3mk_inhomo_homo(int M, double L, double alpha, int cao, void *const K[3], p3mk_queue *q);
/* S_d = sum_{n != 0} (n_d / L) 2 pi U(n) U(n + m M) G(n) (host result) */
int p3mk_inhomo_sums(int M, double L, double alpha, int cao, const int m[3], double S[3],
                     double *scratch, p3mk_queue *q);
/* aliasing-copy kernel K_d(n) = i (-(n_d / L) 2 pi U(n) U(-n - m M) G(n) S_d) */
int p3mk_inhomo_fill(int M, double L, double alpha, int cao, const int m[3], const double S[3],
                     void *const K[3], p3mk_queue *q);
/* K3 = (first) or += sum_d Re(K_d)^2 in the real part */
int p3mk_inhomo_square_add(int M, void *const K[3], void *K3, int first, p3mk_queue *q);
/* Km.re *= K3.re, Km.im *= K3.im */
int p3mk_inhomo_product(int M, void *Km, const void *K3, p3mk_queue *q);
/* mesh = v + 0 i at every point */
int p3mk_fill_real(int M, double v, void *mesh, p3mk_queue *q);
/* out_i = q_i^2; out_i = |b_i - a_i|^2 of two 3n force arrays */
int p3mk_square_charges(int n, const double *q, double *out, p3mk_queue *qu);
int p3mk_diff_charges(int n, const double *a, const double *b, double *out, p3mk_queue *qu);
/* rms_i = sum over the order-cao stencil of q_i^2 W Re(mesh), spacing L/M */
int p3mk_collect_rms(int M, int cao, double L, int n, const double *x, const double *y,
                     const double *z, const double *q, const void *mesh, double *rms,
                     p3mk_queue *qu);

/* ---- influence function, self force, estimates --------------------------- */
/* G (M^3 FP64, r_ind order) of method id 0..3, 6, 7 at alpha; mc aliasing
 * copies per side.  Dn and nshift are computed on the device too. */
int p3mk_influence(int method, const p3mk_geom *g, double alpha, int mc, double *G,
                   double *Dn, p3mk_queue *q);
/* self-force coefficients b (27 x 3 FP64, device) from G; scratch must
 * hold p3mk_reduce_scratch_doubles(M) doubles */
int p3mk_self_force_coeffs(const p3mk_geom *g, const double *G, double *b, double *scratch,
                           p3mk_queue *q);
size_t p3mk_reduce_scratch_doubles(int M);
/* k-space error sum of the estimate of method id (p3m_k_space_error_*),
 * summed over the mesh on the device; *he_q receives the raw sum */
int p3mk_kspace_error_sum(int method, int M, int cao, double L, double alpha, int mc,
                          double *he_q, double *scratch, p3mk_queue *q);

/* ---- real space ---------------------------------------------------------- */
typedef struct p3mk_cells p3mk_cells; /* cell-list workspace */
int p3mk_cells_create(p3mk_cells **c, int n, double L, double rcut);
void p3mk_cells_destroy(p3mk_cells *c);
/* fr += real-space force of every pair i != j with minimum-image distance
 * <= rcut; *energy (device, may be NULL) += energy.  fp32 selects the
 * mixed-precision pair arithmetic. */
int p3mk_realspace(p3mk_cells *c, int n, double L, const double *x, const double *y,
                   const double *z, const double *q, double rcut, double alpha, int fp32,
                   double *fr, double *energy, p3mk_queue *qu);
/* adjoint over the same pairs (FP64): for the cotangent g (3n) of the
 * real-space force, dpos (3n) += d/d pos and dq (n) += d/d q of sum g . f_r */
int p3mk_realspace_vjp(p3mk_cells *c, int n, double L, const double *x, const double *y,
                       const double *z, const double *q, double rcut, double alpha,
                       const double *g, double *dpos, double *dq, p3mk_queue *qu);

/* ---- batch of independent systems ------------------------------------ */
typedef struct {
    int n;         /* particles (slots [sum of earlier n, + n) of the arrays) */
    double L;      /* box length */
    int M, cao;    /* mesh (even, 13-smooth), assignment order 1..7 */
    double rcut;   /* real-space cutoff (0: no real-space part), <= L/2 */
    double alpha;  /* splitting parameter */
    double prefac; /* back-interpolation prefactor */
} p3mk_bsys_in;
typedef struct p3mk_batch p3mk_batch;
int p3mk_batch_create(p3mk_batch **b, int r2c, int interlaced, int ad, int fp32, int nsys,
                      const p3mk_bsys_in *sys);
void p3mk_batch_destroy(p3mk_batch *b);
/* device arrays of system s: influence function (M^3), Dn (M), self-force
 * coefficients (81) */
int p3mk_batch_system_arrays(p3mk_batch *b, int s, double **G, double **Dn, double **selfb);
/* precision, self-force correction, assignment tile side (4 or 8) and the
 * real-space persistent grid (0: keep) */
int p3mk_batch_configure(p3mk_batch *b, int fp32, int use_self_force, int tile_side, int rs_grid);
int p3mk_batch_set_alpha(p3mk_batch *b, int s, double alpha);
int p3mk_batch_info(const p3mk_batch *b, int *ntot, long *ntiles, long *ncell, int *tile_side,
                    int *rs_grid);
/* fk (3 blocks of ntot) += k-space force; ftot -= self force when enabled */
int p3mk_batch_kspace(p3mk_batch *b, const double *x, const double *y, const double *z,
                      const double *q, double *fk, double *ftot, p3mk_queue *q_);
/* fr (3 blocks of ntot) += real-space force; energy[s] += system energy */
int p3mk_batch_realspace(p3mk_batch *b, const double *x, const double *y, const double *z,
                         const double *q, double *fr, double *energy, p3mk_queue *q_);
/* record events P3MK_EV_PAIRS, +1 around the batched half-shell kernel in
 * the following calls (on != 0); not for calls captured into a graph */
int p3mk_batch_time_pairs(p3mk_batch *b, int on);

/* ---- Ewald sum at sampled particles --------------------------------------- */
/* f_total (3 ns) += Ewald force (k space: lattice vectors n^2 <= kmax^2 at
 * alpha; real space: pairs within rcut <= L/2) on particles idx[0..ns);
 * f_split (3 ns, may be NULL) += that force minus the real-space sum at
 * alpha2 within rcut, i.e. the k-space force at alpha2.  Synchronises q. */
int p3mk_ewald_sample(int n, double L, const double *x, const double *y, const double *z,
                      const double *q, int ns, const int *idx, double alpha, double rcut, int kmax,
                      double alpha2, double *f_total, double *f_split, p3mk_queue *q_);
/* The same on `stream` (a hipStream_t, NULL = the null stream) without any
 * allocation or synchronisation: the working memory is a file-scope device
 * array of k_ewald.hip (kmax <= 63, <= 2^18 lattice vectors, ns <= 32768;
 * larger requests return P3MK_EINVAL).  Concurrent calls on one device share
 * that array and race. */
int p3mk_ewald_sample_noalloc(int n, double L, const double *x, const double *y, const double *z,
                              const double *q, int ns, const int *idx, double alpha, double rcut,
                              int kmax, double alpha2, double *f_total, double *f_split,
                              void *stream);

/* ---- graphs ------------------------------------------------------------- */
typedef struct p3mk_graph p3mk_graph;
/* start capturing the work enqueued on q; end it into an executable graph */
int p3mk_capture_begin(p3mk_queue *q);
int p3mk_capture_end(p3mk_queue *q, p3mk_graph **g);
int p3mk_graph_launch(p3mk_graph *g, p3mk_queue *q);
void p3mk_graph_destroy(p3mk_graph *g);

#ifdef __cplusplus
}
#endif
#endif
