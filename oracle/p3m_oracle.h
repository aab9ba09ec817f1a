/* CPU oracle for the P3M k-space force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64, single-threaded
 * restatement of the reference program's algorithm (fweik/p3m-standalone),
 * used by the test-suite and by bench.py's CPU-baseline leg as the checker.
 * It is never linked into, loaded by, or called from the product library.
 *
 * Conventions: particle data and forces are SoA arrays of length N; a force
 * array `f` holds fx[0..N-1], fy[0..N-1], fz[0..N-1] back to back.
 * Meshes use the reference's layouts: complex interleaved (re,im) with
 * index 2*((i*M + j)*M + k) for c2c, and M x M x (M+2) padded real for r2c.
 * Method ids follow types.h:109-118 (0 ik, 1 ik-i, 2 ad, 3 ad-i, 6 ik-r,
 * 7 ad-r).
 */
#ifndef P3M_ORACLE_H
#define P3M_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_INTERPOL 65536

/* B-spline weights (de Boor recursion; window-functions.c:46-184) */
double orc_weight(int j, double x, int cao);
double orc_weight_d(int j, double x, int cao);

/* Mesh-coordinate quantisation (charge-assign.c:46-52,94-97,153-157):
 * returns nmp, writes the wrapped base index and the table index arg. */
int orc_locate(double r, double Hi, int cao, int ii, int M, int *base, int *arg);
double orc_quantized_offset(int arg);

/* FFTs: unnormalised, sign -1 forward (FFTW_FORWARD) / +1 backward. */
void orc_fft1d(double *c, int n, int stride, int sign);
void orc_fft3d(double *c, int M, int sign);
void orc_rfft3d(double *r, int M);   /* in-place r2c, FFTW padded layout */
void orc_irfft3d(double *c, int M);  /* in-place c2r, FFTW padded layout */

/* Differential operator and shifted wave numbers (p3m-common.c:110-143) */
void orc_dn(int M, double *Dn);
void orc_nshift(int M, double *ns);

/* Influence functions (p3m-ik.c:93-173, p3m-ik-i.c:78-166, p3m-ad.c:69-158,
 * p3m-ad-i.c:70-150).  `method` is a method id; 6 uses ik, 7 uses ad. */
int orc_influence(int method, int M, int cao, double L, double alpha, int mc, double *G);

/* Self-force coefficients b_d(m'), 27 x 3 doubles (p3m-ad-self-forces.c:30-112) */
void orc_self_force_coeffs(int M, int cao, double L, const double *G, double *b);
/* f -= sum_m' b(m') q^2 sin(2 pi m'.r/h) (p3m-ad-self-forces.c:114-139) */
void orc_subtract_self_forces(int N, double L, int M, const double *x, const double *y,
                              const double *z, const double *q, const double *b, double *f);

/* One k-space force evaluation (Kspace_force of the method table):
 * fk += k-space force; for methods 2 and 7 also f -= self force.
 * G is the influence function from orc_influence, b the self-force
 * coefficients (methods 2, 7 only; may be NULL otherwise). */
int orc_kspace_force(int method, int N, double L, const double *x, const double *y,
                     const double *z, const double *q, int M, int cao, const double *G,
                     const double *b, double *fk, double *f);

/* Stage access for tests: the assigned charge mesh (after assignment,
 * before the FFT) in the layout of `method`. */
int orc_assign_mesh(int method, int N, double L, const double *x, const double *y,
                    const double *z, const double *q, int M, int cao, double *mesh);

/* Real-space part (realpart.c:124-167): all i != j, minimum image,
 * r <= rcut.  fr += forces; *energy += energy.  O(N^2). */
void orc_realteil(int N, double L, const double *x, const double *y, const double *z,
                  const double *q, double rcut, double alpha, double *fr, double *energy);
/* Same pair set and arithmetic via a linked-cell list (needs rcut < L/2). */
void orc_realteil_cells(int N, double L, const double *x, const double *y, const double *z,
                        const double *q, double rcut, double alpha, double *fr, double *energy,
                        int i_begin, int i_end);

/* Ewald reference (ewald.c:84-229, common.c:329-363) */
void orc_ewald_influence(int kmax, double L, double alpha, double *G);
void orc_ewald_kspace(int N, double L, const double *x, const double *y, const double *z,
                      const double *q, double alpha, int kmax, const double *G, double *fk);
double orc_ewald_error_r(int N, double L, double q2, double rcut, double alpha);
double orc_ewald_error_k(int N, double L, double q2, int kmax, double alpha);
double orc_ewald_optimal_alpha(int N, double L, double q2, double rcut, int kmax);

/* Error estimates (realpart.c:343-346, p3m-ik.c:249-337, p3m-ik-i.c:249-320,
 * p3m-ad.c:436-497, p3m-ad-i.c:205-284, find_error.c:84-107).
 * `table` (101 values at alpha*L = 0,1,..,100) replaces the q_*.c lookup for
 * methods 0/1/6 when non-NULL; pass NULL for an untabulated mesh. */
double orc_realspace_error(int N, double L, double q2, double rcut, double alpha);
double orc_kspace_error(int method, int N, double L, double q2, int M, int cao,
                        double alpha, int mc, const double *table);
double orc_table_lookup(const double *table, double alphaL);

/* rms error sqrt(sum |f - fref|^2) (error.c:21-34) */
double orc_force_error(int N, const double *f, const double *fref);

#ifdef __cplusplus
}
#endif
#endif
