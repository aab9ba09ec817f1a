/* CPU oracle for the P3M k-space force path -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain C99, FP64, single-threaded restatement of the reference program
 * fweik/p3m-standalone.  Every function cites the reference file:line it
 * restates.  The oracle is the checker used by tests/ and by bench.py's CPU
 * baseline; the product library never links or calls it.
 *
 * Differences from the reference that do not change results beyond
 * rounding:
 *   - B-spline weights are evaluated by the de Boor recursion at the same
 *     quantised offsets the reference tabulates, instead of its Horner
 *     polynomials (same function, ~1 ulp apart);
 *   - the FFT is a plain mixed-radix Cooley-Tukey transform instead of FFTW
 *     (same unnormalised convention, rounding-level differences);
 *   - the per-particle caches cf/dQ/ca_ind are recomputed instead of stored.
 *
 * Build: see oracle/Makefile (gcc -O2 -std=c99 -ffp-contract=off).
 */
#include "p3m_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define SQR(a) ((a) * (a))

/* ------------------------------------------------------------------------ */
/* B-spline weights                                                          */
/* ------------------------------------------------------------------------ */

/* c[m] = N_k(y + m), m = 0..k-1, for the uniform cardinal B-spline N_k of
 * support [0, k], local coordinate y = x + 1/2 in [0, 1]. */
static void cardinal_values(int k, double y, double *c)
{
    double prev[8];
    int order, m;
    c[0] = 1.0;
    for (order = 2; order <= k; order++) {
        for (m = 0; m < order - 1; m++)
            prev[m] = c[m];
        for (m = 0; m < order; m++) {
            double left = (m < order - 1) ? prev[m] : 0.0;
            double right = (m >= 1) ? prev[m - 1] : 0.0;
            c[m] = ((y + m) * left + ((double)order - y - m) * right) / (double)(order - 1);
        }
    }
}

/* W_j(x) = B_cao(x - j + (cao-1)/2), the function window-functions.c:108-184
 * tabulates (caf_bspline). */
double orc_weight(int j, double x, int cao)
{
    double c[8];
    if (cao < 1 || cao > 7 || j < 0 || j >= cao)
        return 0.0;
    cardinal_values(cao, x + 0.5, c);
    return c[cao - 1 - j];
}

/* d/dx W_j(x) (window-functions.c:46-106, caf_bspline_d; zero for cao 1,
 * where the reference reports an unknown order and returns 0). */
double orc_weight_d(int j, double x, int cao)
{
    double c[8];
    int m;
    if (cao < 2 || cao > 7 || j < 0 || j >= cao)
        return 0.0;
    cardinal_values(cao - 1, x + 0.5, c);
    m = cao - 1 - j;
    return ((m <= cao - 2) ? c[m] : 0.0) - ((m >= 1) ? c[m - 1] : 0.0);
}

/* wrap by one period (charge-assign.c:37-39) */
static int wrap1(int ind, int M)
{
    return ind + (ind < 0) * M - (ind >= M) * M;
}

/* Position -> nearest mesh point and table index
 * (charge-assign.c:89,153-157; int_floor :46-52). */
int orc_locate(double r, double Hi, int cao, int ii, int M, int *base, int *arg)
{
    double pos = r * Hi - (double)((cao - 1) / 2) + 0.5 * ii;
    int nmp = (int)floor(pos + 0.5);
    *base = wrap1(nmp, M);
    *arg = (int)floor((pos - nmp + 0.5) * (2.0 * ORC_MAX_INTERPOL));
    return nmp;
}

/* Offset at which the reference evaluates table row `arg`
 * (interpol.c:53-57: x = i / (2 MaxInterpol), i = arg - MaxInterpol). */
double orc_quantized_offset(int arg)
{
    return (double)(arg - ORC_MAX_INTERPOL) / (2.0 * ORC_MAX_INTERPOL);
}

/* weights and derivative weights of one particle along one axis */
static void axis_weights(double r, double Hi, int cao, int ii, int M, int want_d,
                         int *base, double *w, double *dw)
{
    int arg, j;
    double xq;
    orc_locate(r, Hi, cao, ii, M, base, &arg);
    xq = orc_quantized_offset(arg);
    for (j = 0; j < cao; j++) {
        w[j] = orc_weight(j, xq, cao);
        if (want_d)
            dw[j] = orc_weight_d(j, xq, cao);
    }
}

/* ------------------------------------------------------------------------ */
/* FFT                                                                       */
/* ------------------------------------------------------------------------ */

static int smallest_factor(int n)
{
    int f;
    for (f = 2; f * f <= n; f++)
        if (n % f == 0)
            return f;
    return n;
}

/* out[k] = sum_t in[t*stride] exp(sign 2 pi i t k / n), recursive DIT */
static void fft_rec(const double *in, int stride, double *out, int n, int sign)
{
    int r, m, q, k, s;
    double *tmp;
    if (n == 1) {
        out[0] = in[0];
        out[1] = in[1];
        return;
    }
    r = smallest_factor(n);
    m = n / r;
    for (q = 0; q < r; q++)
        fft_rec(in + 2 * q * stride, stride * r, out + 2 * q * m, m, sign);
    tmp = (double *)malloc(sizeof(double) * 2 * n);
    for (s = 0; s < r; s++) {
        for (k = 0; k < m; k++) {
            int kk = k + m * s;
            double re = 0.0, im = 0.0;
            for (q = 0; q < r; q++) {
                long e = ((long)q * kk) % n;
                double ang = sign * 2.0 * M_PI * (double)e / (double)n;
                double c = cos(ang), sn = sin(ang);
                double yr = out[2 * (q * m + k)], yi = out[2 * (q * m + k) + 1];
                re += yr * c - yi * sn;
                im += yr * sn + yi * c;
            }
            tmp[2 * kk] = re;
            tmp[2 * kk + 1] = im;
        }
    }
    memcpy(out, tmp, sizeof(double) * 2 * n);
    free(tmp);
}

void orc_fft1d(double *c, int n, int stride, int sign)
{
    double *line, *res;
    int t;
    if (n <= 0)
        return;
    line = (double *)malloc(sizeof(double) * 2 * n);
    res = (double *)malloc(sizeof(double) * 2 * n);
    for (t = 0; t < n; t++) {
        line[2 * t] = c[2 * (long)t * stride];
        line[2 * t + 1] = c[2 * (long)t * stride + 1];
    }
    fft_rec(line, 1, res, n, sign);
    for (t = 0; t < n; t++) {
        c[2 * (long)t * stride] = res[2 * t];
        c[2 * (long)t * stride + 1] = res[2 * t + 1];
    }
    free(line);
    free(res);
}

/* 3D c2c on a complex n0 x n1 x n2 array (last index fastest) */
static void fft3d_dims(double *c, int n0, int n1, int n2, int sign)
{
    int a, b;
    for (a = 0; a < n0; a++)
        for (b = 0; b < n1; b++)
            orc_fft1d(c + 2 * ((long)a * n1 + b) * n2, n2, 1, sign);
    for (a = 0; a < n0; a++)
        for (b = 0; b < n2; b++)
            orc_fft1d(c + 2 * ((long)a * n1 * n2 + b), n1, n2, sign);
    for (a = 0; a < n1; a++)
        for (b = 0; b < n2; b++)
            orc_fft1d(c + 2 * ((long)a * n2 + b), n0, n1 * n2, sign);
}

void orc_fft3d(double *c, int M, int sign)
{
    fft3d_dims(c, M, M, M, sign);
}

/* in-place r2c with the padded layout FFTW uses for fftw_plan_dft_r2c_3d on
 * an M x M x (M+2) real array (p3m-ik-real.c:72, p3m-ad-real.c:62) */
void orc_rfft3d(double *r, int M)
{
    int H = M / 2 + 1, i, j, k;
    double *line = (double *)malloc(sizeof(double) * 2 * M);
    for (i = 0; i < M; i++)
        for (j = 0; j < M; j++) {
            double *row = r + ((long)i * M + j) * (M + 2);
            for (k = 0; k < M; k++) {
                line[2 * k] = row[k];
                line[2 * k + 1] = 0.0;
            }
            orc_fft1d(line, M, 1, -1);
            for (k = 0; k < H; k++) {
                row[2 * k] = line[2 * k];
                row[2 * k + 1] = line[2 * k + 1];
            }
        }
    free(line);
    /* remaining two axes on the M x M x H complex array */
    for (i = 0; i < M; i++)
        for (k = 0; k < H; k++)
            orc_fft1d(r + 2 * ((long)i * M * H + k), M, H, -1);
    for (j = 0; j < M; j++)
        for (k = 0; k < H; k++)
            orc_fft1d(r + 2 * ((long)j * H + k), M, M * H, -1);
}

/* in-place c2r, inverse of orc_rfft3d up to the factor M^3 */
void orc_irfft3d(double *c, int M)
{
    int H = M / 2 + 1, i, j, k;
    double *line;
    for (i = 0; i < M; i++)
        for (k = 0; k < H; k++)
            orc_fft1d(c + 2 * ((long)i * M * H + k), M, H, +1);
    for (j = 0; j < M; j++)
        for (k = 0; k < H; k++)
            orc_fft1d(c + 2 * ((long)j * H + k), M, M * H, +1);
    line = (double *)malloc(sizeof(double) * 2 * M);
    for (i = 0; i < M; i++)
        for (j = 0; j < M; j++) {
            double *row = c + 2 * ((long)i * M + j) * H;
            for (k = 0; k < M; k++) {
                if (k < H) {
                    line[2 * k] = row[2 * k];
                    line[2 * k + 1] = row[2 * k + 1];
                } else {
                    line[2 * k] = row[2 * (M - k)];
                    line[2 * k + 1] = -row[2 * (M - k) + 1];
                }
            }
            orc_fft1d(line, M, 1, +1);
            for (k = 0; k < M; k++)
                row[k] = line[2 * k];
        }
    free(line);
}

/* ------------------------------------------------------------------------ */
/* Operators and influence functions                                         */
/* ------------------------------------------------------------------------ */

/* p3m-common.c:110-131 */
void orc_dn(int M, double *Dn)
{
    int i;
    for (i = 0; i < M; i++)
        Dn[i] = (double)i - round((double)i / (double)M) * (double)M;
    Dn[M / 2] = 0.0;
}

/* p3m-common.c:133-143 */
void orc_nshift(int M, double *ns)
{
    int i;
    for (i = 0; i < M; i++)
        ns[i] = i - round(i / (double)M) * (double)M;
}

/* sinc (p3m-common.c:71-75) */
static double sinc(double d)
{
    double pd = M_PI * d;
    return (d == 0.0) ? 1.0 : sin(pd) / pd;
}

static double ipow(double b, int e)
{
    double r = 1.0;
    while (e-- > 0)
        r *= b;
    return r;
}

/* U_hat of the B-spline (window-functions.c:40-44) */
static double bspline_k(int cao, double d)
{
    double pd = M_PI * d;
    return (d == 0.0) ? 1.0 : pow(sin(pd) / pd, cao);
}

/* ik (p3m-ik.c:93-173): uses U_hat^2 */
static double influence_ik_point(int NX, int NY, int NZ, int M, int cao, double L,
                                 double alpha, int mc, const double *ns, const double *Dn)
{
    double Leni = 1.0 / L, fak1 = 1.0 / (double)M, fak2 = SQR(M_PI / alpha);
    double Z[3] = {0.0, 0.0, 0.0}, Nen = 0.0, zwi;
    int MX, MY, MZ;
    for (MX = -mc; MX <= mc; MX++) {
        double NMX = ns[NX] + M * MX;
        double S1 = ipow(bspline_k(cao, fak1 * NMX), 2);
        for (MY = -mc; MY <= mc; MY++) {
            double NMY = ns[NY] + M * MY;
            double S2 = S1 * ipow(bspline_k(cao, fak1 * NMY), 2);
            for (MZ = -mc; MZ <= mc; MZ++) {
                double NMZ = ns[NZ] + M * MZ;
                double S3 = S2 * ipow(bspline_k(cao, fak1 * NMZ), 2);
                double NM2 = SQR(NMX * Leni) + SQR(NMY * Leni) + SQR(NMZ * Leni);
                double t = S3 * exp(-fak2 * NM2) / NM2;
                Nen += S3;
                Z[0] += NMX * t * Leni;
                Z[1] += NMY * t * Leni;
                Z[2] += NMZ * t * Leni;
            }
        }
    }
    zwi = Dn[NX] * Z[0] * Leni + Dn[NY] * Z[1] * Leni + Dn[NZ] * Z[2] * Leni;
    zwi /= (SQR(Dn[NX] * Leni) + SQR(Dn[NY] * Leni) + SQR(Dn[NZ] * Leni)) * SQR(Nen);
    return 2.0 * zwi / M_PI;
}

/* ik-i (p3m-ik-i.c:78-166) */
static double influence_iki_point(int NX, int NY, int NZ, int M, int cao, double L,
                                  double alpha, int mc, const double *ns, const double *Dn)
{
    double Leni = 1.0 / L, fak1 = 1.0 / (double)M, fak2 = SQR(M_PI / alpha);
    double Z[3] = {0.0, 0.0, 0.0}, N1 = 0.0, N2 = 0.0, zwi;
    int MX, MY, MZ;
    for (MX = -mc; MX <= mc; MX++) {
        double NMX = ns[NX] + M * MX;
        double S1 = ipow(sinc(fak1 * NMX), 2 * cao);
        for (MY = -mc; MY <= mc; MY++) {
            double NMY = ns[NY] + M * MY;
            double S2 = S1 * ipow(sinc(fak1 * NMY), 2 * cao);
            for (MZ = -mc; MZ <= mc; MZ++) {
                double NMZ = ns[NZ] + M * MZ;
                double S3 = S2 * ipow(sinc(fak1 * NMZ), 2 * cao);
                double NM2 = SQR(NMX * Leni) + SQR(NMY * Leni) + SQR(NMZ * Leni);
                double t = S3 * exp(-fak2 * NM2) / NM2;
                N1 += S3;
                Z[0] += NMX * t * Leni;
                Z[1] += NMY * t * Leni;
                Z[2] += NMZ * t * Leni;
                if (((MX + MY + MZ) % 2) == 0)
                    N2 += S3;
                else
                    N2 -= S3;
            }
        }
    }
    zwi = Dn[NX] * Z[0] * Leni + Dn[NY] * Z[1] * Leni + Dn[NZ] * Z[2] * Leni;
    zwi /= (SQR(Dn[NX] * Leni) + SQR(Dn[NY] * Leni) + SQR(Dn[NZ] * Leni));
    zwi /= 0.5 * (SQR(N1) + SQR(N2));
    return 2.0 * zwi / M_PI;
}

/* ad (p3m-ad.c:69-158) and ad-i (p3m-ad-i.c:70-150) */
static double influence_ad_point(int interlaced, int NX, int NY, int NZ, int M, int cao,
                                 double L, double alpha, int mc, const double *ns)
{
    double Leni = 1.0 / L, fak1 = 1.0 / (double)M, fak2 = SQR(M_PI / alpha);
    double Z = 0.0, N1 = 0.0, N2 = 0.0, N3 = 0.0, N4 = 0.0;
    int MX, MY, MZ;
    for (MX = -mc; MX <= mc; MX++) {
        double NMX = ns[NX] + M * MX;
        double S1 = ipow(sinc(fak1 * NMX), 2 * cao);
        for (MY = -mc; MY <= mc; MY++) {
            double NMY = ns[NY] + M * MY;
            double S2 = S1 * ipow(sinc(fak1 * NMY), 2 * cao);
            for (MZ = -mc; MZ <= mc; MZ++) {
                double NMZ = ns[NZ] + M * MZ;
                double S3 = S2 * ipow(sinc(fak1 * NMZ), 2 * cao);
                double NM2 = SQR(NMX * Leni) + SQR(NMY * Leni) + SQR(NMZ * Leni);
                N1 += S3;
                N2 += S3 * NM2;
                if (((MX + MY + MZ) % 2) == 0) {
                    N3 += S3;
                    N4 += S3 * NM2;
                } else {
                    N3 -= S3;
                    N4 -= S3 * NM2;
                }
                Z += S3 * exp(-fak2 * NM2);
            }
        }
    }
    if (interlaced)
        return Z / (0.5 * M_PI * (N1 * N2 + N3 * N4));
    return Z / (M_PI * N1 * N2);
}

int orc_influence(int method, int M, int cao, double L, double alpha, int mc, double *G)
{
    double *ns = (double *)malloc(sizeof(double) * M);
    double *Dn = (double *)malloc(sizeof(double) * M);
    int NX, NY, NZ, h = M / 2;
    orc_nshift(M, ns);
    orc_dn(M, Dn);
    if ((method == 2 || method == 7) && alpha == 0.0) {
        /* p3m-ad.c:120-123 */
        memset(G, 0, sizeof(double) * M * M * M);
        free(ns);
        free(Dn);
        return 0;
    }
    for (NX = 0; NX < M; NX++)
        for (NY = 0; NY < M; NY++)
            for (NZ = 0; NZ < M; NZ++) {
                long ind = ((long)NX * M + NY) * M + NZ;
                int corner = (NX % h == 0) && (NY % h == 0) && (NZ % h == 0);
                double g;
                if (NX == 0 && NY == 0 && NZ == 0) {
                    g = 0.0;
                } else if (method == 0 || method == 6) {
                    g = corner ? 0.0 : influence_ik_point(NX, NY, NZ, M, cao, L, alpha, mc, ns, Dn);
                } else if (method == 1) {
                    g = corner ? 0.0 : influence_iki_point(NX, NY, NZ, M, cao, L, alpha, mc, ns, Dn);
                } else if (method == 2 || method == 7) {
                    g = influence_ad_point(0, NX, NY, NZ, M, cao, L, alpha, mc, ns);
                } else if (method == 3) {
                    g = influence_ad_point(1, NX, NY, NZ, M, cao, L, alpha, mc, ns);
                } else {
                    free(ns);
                    free(Dn);
                    return -1;
                }
                G[ind] = g;
            }
    free(ns);
    free(Dn);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Self-force correction (p3m-ad-self-forces.c:30-139)                       */
/* ------------------------------------------------------------------------ */

void orc_self_force_coeffs(int M, int cao, double L, const double *G, double *b)
{
    double *ns = (double *)malloc(sizeof(double) * M);
    double S[27];
    double mesh_i = 1.0 / M;
    int nx, ny, nz, a, c;
    orc_nshift(M, ns);
    memset(S, 0, sizeof(S));
    for (nx = 0; nx < M; nx++)
        for (ny = 0; ny < M; ny++)
            for (nz = 0; nz < M; nz++) {
                double g = G[((long)nx * M + ny) * M + nz];
                int tx = (int)ns[nx], ty = (int)ns[ny], tz = (int)ns[nz];
                int m1, m2, m3;
                if (g == 0.0)
                    continue;
                for (m1 = -1; m1 <= 1; m1++)
                    for (m2 = -1; m2 <= 1; m2++)
                        for (m3 = -1; m3 <= 1; m3++) {
                            int ind = (m1 + 1) * 9 + (m2 + 1) * 3 + (m3 + 1);
                            int mx, my, mz;
                            double sum = 0.0;
                            for (mx = -1; mx <= 1; mx++)
                                for (my = -1; my <= 1; my++)
                                    for (mz = -1; mz <= 1; mz++) {
                                        double U = pow(sinc(mesh_i * (tx + mx * M)) *
                                                           sinc(mesh_i * (ty + my * M)) *
                                                           sinc(mesh_i * (tz + mz * M)),
                                                       cao);
                                        double Us = pow(sinc(mesh_i * (tx + (mx + m1) * M)) *
                                                            sinc(mesh_i * (ty + (my + m2) * M)) *
                                                            sinc(mesh_i * (tz + (mz + m3) * M)),
                                                        cao);
                                        sum += g * U * Us;
                                    }
                            S[ind] += sum;
                        }
            }
    for (a = 0; a < 27; a++) {
        int m[3] = {a / 9 - 1, (a / 3) % 3 - 1, a % 3 - 1};
        for (c = 0; c < 3; c++)
            b[3 * a + c] = (m[c] == 0) ? 0.0 : (M_PI * m[c] * M) / SQR(SQR(L)) * S[a];
    }
    free(ns);
}

void orc_subtract_self_forces(int N, double L, int M, const double *x, const double *y,
                              const double *z, const double *q, const double *b, double *f)
{
    double h = L / M;
    int id, a, c;
    for (id = 0; id < N; id++)
        for (a = 0; a < 27; a++) {
            int m0 = a / 9 - 1, m1 = (a / 3) % 3 - 1, m2 = a % 3 - 1;
            double st = SQR(q[id]) * sin(2 * M_PI * (m0 * x[id] / h + m1 * y[id] / h + m2 * z[id] / h));
            for (c = 0; c < 3; c++)
                f[(long)c * N + id] -= b[3 * a + c] * st;
        }
}

/* ------------------------------------------------------------------------ */
/* Charge assignment and back-interpolation                                  */
/* ------------------------------------------------------------------------ */

static int is_ad(int method) { return method == 2 || method == 3 || method == 7; }
static int is_interlaced(int method) { return method == 1 || method == 3; }
static int is_r2c(int method) { return method == 6 || method == 7; }

/* mesh index of point (i,j,k) for the method's layout; `part` selects the
 * real (0) or imaginary (1) part of a complex mesh */
static long mesh_index(int method, int M, int i, int j, int k, int part)
{
    if (is_r2c(method))
        return ((long)i * M + j) * (M + 2) + k;
    return 2 * (((long)i * M + j) * M + k) + part;
}

/* assign_charge / assign_charge_real / assign_charge_and_derivatives(_real)
 * (charge-assign.c:119-176, 215-267, 1067-1133, 1243-1311) */
static void assign(int method, int N, double L, const double *x, const double *y,
                   const double *z, const double *q, int M, int cao, int ii, double *mesh)
{
    double Hi = (double)M / L;
    int id, a, bb, c;
    for (id = 0; id < N; id++) {
        int base[3];
        double w[3][8], dw[3][8];
        axis_weights(x[id], Hi, cao, ii, M, 0, &base[0], w[0], dw[0]);
        axis_weights(y[id], Hi, cao, ii, M, 0, &base[1], w[1], dw[1]);
        axis_weights(z[id], Hi, cao, ii, M, 0, &base[2], w[2], dw[2]);
        for (a = 0; a < cao; a++) {
            int i = wrap1(base[0] + a, M);
            double t0 = q[id] * w[0][a];
            for (bb = 0; bb < cao; bb++) {
                int j = wrap1(base[1] + bb, M);
                double t1 = t0 * w[1][bb];
                for (c = 0; c < cao; c++) {
                    int k = wrap1(base[2] + c, M);
                    mesh[mesh_index(method, M, i, j, k, ii)] += t1 * w[2][c];
                }
            }
        }
    }
}

int orc_assign_mesh(int method, int N, double L, const double *x, const double *y,
                    const double *z, const double *q, int M, int cao, double *mesh)
{
    memset(mesh, 0, sizeof(double) * 2 * M * M * M);
    assign(method, N, L, x, y, z, q, M, cao, 0, mesh);
    if (is_interlaced(method))
        assign(method, N, L, x, y, z, q, M, cao, 1, mesh);
    return 0;
}

/* ik gather: assign_forces / assign_forces_interlacing / assign_forces_real
 * (charge-assign.c:601-643, 646-688, 849-885) */
static void gather_ik(int method, int N, double L, const double *x, const double *y,
                      const double *z, const double *q, int M, int cao, double prefac,
                      double *const F[3], double *fk)
{
    double Hi = (double)M / L;
    int nm = is_interlaced(method) ? 2 : 1;
    int id, a, bb, c, ii;
    for (id = 0; id < N; id++) {
        double field[3] = {0.0, 0.0, 0.0};
        int base[2][3];
        double w[2][3][8], dw[8];
        for (ii = 0; ii < nm; ii++) {
            axis_weights(x[id], Hi, cao, ii, M, 0, &base[ii][0], w[ii][0], dw);
            axis_weights(y[id], Hi, cao, ii, M, 0, &base[ii][1], w[ii][1], dw);
            axis_weights(z[id], Hi, cao, ii, M, 0, &base[ii][2], w[ii][2], dw);
        }
        for (a = 0; a < cao; a++)
            for (bb = 0; bb < cao; bb++)
                for (c = 0; c < cao; c++) {
                    if (nm == 1) {
                        double cf = q[id] * w[0][0][a] * w[0][1][bb] * w[0][2][c];
                        double B = prefac * cf;
                        long l = mesh_index(method, M, wrap1(base[0][0] + a, M),
                                            wrap1(base[0][1] + bb, M), wrap1(base[0][2] + c, M), 0);
                        field[0] -= F[0][l] * B;
                        field[1] -= F[1][l] * B;
                        field[2] -= F[2][l] * B;
                    } else {
                        double cf1 = q[id] * w[0][0][a] * w[0][1][bb] * w[0][2][c];
                        double cf2 = q[id] * w[1][0][a] * w[1][1][bb] * w[1][2][c];
                        double B1 = 0.5 * prefac * cf1, B2 = 0.5 * prefac * cf2;
                        long l1 = mesh_index(method, M, wrap1(base[0][0] + a, M),
                                             wrap1(base[0][1] + bb, M), wrap1(base[0][2] + c, M), 0);
                        long l2 = mesh_index(method, M, wrap1(base[1][0] + a, M),
                                             wrap1(base[1][1] + bb, M), wrap1(base[1][2] + c, M), 1);
                        field[0] -= F[0][l1] * B1 + F[0][l2] * B2;
                        field[1] -= F[1][l1] * B1 + F[1][l2] * B2;
                        field[2] -= F[2][l1] * B1 + F[2][l2] * B2;
                    }
                }
        fk[id] += field[0];
        fk[(long)N + id] += field[1];
        fk[2L * N + id] += field[2];
    }
}

/* ad gather: assign_forces_ad / assign_forces_interlacing_ad /
 * assign_forces_ad_real (charge-assign.c:1352-1391, 727-769, 1468-1502) */
static void gather_ad(int method, int N, double L, const double *x, const double *y,
                      const double *z, const double *q, int M, int cao, double prefac,
                      const double *phi, double *fk)
{
    double Hi = (double)M / L, Leni = 1.0 / L;
    int nm = is_interlaced(method) ? 2 : 1;
    int id, a, bb, c, ii;
    for (id = 0; id < N; id++) {
        double force[3] = {0.0, 0.0, 0.0};
        double qL = q[id] * Leni;
        int base[2][3];
        double w[2][3][8], dw[2][3][8];
        for (ii = 0; ii < nm; ii++) {
            axis_weights(x[id], Hi, cao, ii, M, 1, &base[ii][0], w[ii][0], dw[ii][0]);
            axis_weights(y[id], Hi, cao, ii, M, 1, &base[ii][1], w[ii][1], dw[ii][1]);
            axis_weights(z[id], Hi, cao, ii, M, 1, &base[ii][2], w[ii][2], dw[ii][2]);
        }
        for (a = 0; a < cao; a++)
            for (bb = 0; bb < cao; bb++)
                for (c = 0; c < cao; c++) {
                    double dq[2][3];
                    double B[2];
                    for (ii = 0; ii < nm; ii++) {
                        long l = mesh_index(method, M, wrap1(base[ii][0] + a, M),
                                            wrap1(base[ii][1] + bb, M), wrap1(base[ii][2] + c, M), ii);
                        B[ii] = phi[l];
                        dq[ii][0] = dw[ii][0][a] * w[ii][1][bb] * w[ii][2][c] * qL;
                        dq[ii][1] = w[ii][0][a] * dw[ii][1][bb] * w[ii][2][c] * qL;
                        dq[ii][2] = w[ii][0][a] * w[ii][1][bb] * dw[ii][2][c] * qL;
                    }
                    if (nm == 1) {
                        double Bp = prefac * B[0];
                        force[0] -= Bp * dq[0][0];
                        force[1] -= Bp * dq[0][1];
                        force[2] -= Bp * dq[0][2];
                    } else {
                        force[0] -= 0.5 * prefac * (B[0] * dq[0][0] + B[1] * dq[1][0]);
                        force[1] -= 0.5 * prefac * (B[0] * dq[0][1] + B[1] * dq[1][1]);
                        force[2] -= 0.5 * prefac * (B[0] * dq[0][2] + B[1] * dq[1][2]);
                    }
                }
        fk[id] += force[0];
        fk[(long)N + id] += force[1];
        fk[2L * N + id] += force[2];
    }
}

/* Kspace_force of methods 0,1,2,3,6,7 (p3m-ik.c:179-245, p3m-ik-i.c:169-246,
 * p3m-ad.c:161-210, p3m-ad-i.c:153-203, p3m-ik-real.c:83-147,
 * p3m-ad-real.c:69-118) */
int orc_kspace_force(int method, int N, double L, const double *x, const double *y,
                     const double *z, const double *q, int M, int cao, const double *G,
                     const double *b, double *fk, double *f)
{
    long M3 = (long)M * M * M;
    double Leni = 1.0 / L;
    double *Q, *Dn;
    int i, j, k;
    if (cao < 1 || cao > 7 || M < 2 || (M % 2) != 0)
        return -1;
    if (method != 0 && method != 1 && method != 2 && method != 3 && method != 6 && method != 7)
        return -1;
    Q = (double *)calloc(2 * M3, sizeof(double));
    Dn = (double *)malloc(sizeof(double) * M);
    orc_dn(M, Dn);
    orc_assign_mesh(method, N, L, x, y, z, q, M, cao, Q);

    if (is_r2c(method))
        orc_rfft3d(Q, M);
    else
        orc_fft3d(Q, M, -1);

    if (!is_ad(method)) {
        double *F[3];
        int d;
        int H = is_r2c(method) ? M / 2 + 1 : M;
        for (d = 0; d < 3; d++)
            F[d] = (double *)calloc(2 * M3, sizeof(double));
        for (i = 0; i < M; i++)
            for (j = 0; j < M; j++)
                for (k = 0; k < H; k++) {
                    long ci = 2 * (((long)i * M + j) * H + k);
                    double T1 = G[((long)i * M + j) * M + k];
                    double dop[3] = {Dn[i], Dn[j], Dn[k]};
                    if (method == 0) {
                        /* p3m-ik.c:211-231 */
                        double qr = 2.0 * M_PI * Leni * Q[ci] * T1;
                        double qi = -2.0 * M_PI * Leni * Q[ci + 1] * T1;
                        for (d = 0; d < 3; d++) {
                            F[d][ci] = dop[d] * qi;
                            F[d][ci + 1] = dop[d] * qr;
                        }
                    } else if (method == 1) {
                        /* p3m-ik-i.c:202-227 */
                        Q[ci] *= T1;
                        Q[ci + 1] *= T1;
                        for (d = 0; d < 3; d++) {
                            F[d][ci] = -2.0 * M_PI * Leni * dop[d] * Q[ci + 1];
                            F[d][ci + 1] = 2.0 * M_PI * Leni * dop[d] * Q[ci];
                        }
                    } else {
                        /* p3m-ik-real.c:113-136 */
                        double qr = Q[ci] * T1, qi = Q[ci + 1] * T1;
                        for (d = 0; d < 3; d++) {
                            F[d][ci] = -2.0 * M_PI * Leni * dop[d] * qi;
                            F[d][ci + 1] = 2.0 * M_PI * Leni * dop[d] * qr;
                        }
                    }
                }
        for (d = 0; d < 3; d++) {
            if (is_r2c(method))
                orc_irfft3d(F[d], M);
            else
                orc_fft3d(F[d], M, +1);
        }
        gather_ik(method, N, L, x, y, z, q, M, cao, 1.0 / (2.0 * L * L * L), F, fk);
        for (d = 0; d < 3; d++)
            free(F[d]);
    } else {
        int H = is_r2c(method) ? M / 2 + 1 : M;
        for (i = 0; i < M; i++)
            for (j = 0; j < M; j++)
                for (k = 0; k < H; k++) {
                    long ci = 2 * (((long)i * M + j) * H + k);
                    double T1 = G[((long)i * M + j) * M + k];
                    Q[ci] *= T1;
                    Q[ci + 1] *= T1;
                }
        if (is_r2c(method))
            orc_irfft3d(Q, M);
        else
            orc_fft3d(Q, M, +1);
        gather_ad(method, N, L, x, y, z, q, M, cao, M * Leni * Leni * Leni, Q, fk);
        if ((method == 2 || method == 7) && b != NULL && f != NULL)
            orc_subtract_self_forces(N, L, M, x, y, z, q, b, f);
    }
    free(Q);
    free(Dn);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Real space                                                                */
/* ------------------------------------------------------------------------ */

static const double WUPI = 1.77245385090551602729816748334; /* realpart.c:135 */

static void pair_force(int t1, int t2, double L, double Li, const double *x, const double *y,
                       const double *z, const double *q, double rcut, double alpha,
                       double *acc, double *energy)
{
    double dx = x[t1] - x[t2], dy, dz, r;
    dx -= round(dx * Li) * L;
    dy = y[t1] - y[t2];
    dy -= round(dy * Li) * L;
    dz = z[t1] - z[t2];
    dz -= round(dz * Li) * L;
    r = sqrt(SQR(dx) + SQR(dy) + SQR(dz));
    if (r <= rcut) {
        double ar = alpha * r;
        double erfc_teil = erfc(ar);
        double r2 = SQR(r);
        double fak = q[t1] * q[t2] * (erfc_teil / r + (2.0 * alpha / WUPI) * exp(-ar * ar)) / r2;
        acc[0] += fak * dx;
        acc[1] += fak * dy;
        acc[2] += fak * dz;
        *energy += 0.5 * q[t1] * q[t2] * erfc_teil / r;
    }
}

/* realpart.c:124-167 */
void orc_realteil(int N, double L, const double *x, const double *y, const double *z,
                  const double *q, double rcut, double alpha, double *fr, double *energy)
{
    double Li = 1.0 / L;
    int t1, t2;
    for (t1 = 0; t1 < N; t1++) {
        double acc[3] = {0.0, 0.0, 0.0};
        for (t2 = 0; t2 < N; t2++) {
            if (t1 == t2)
                continue;
            pair_force(t1, t2, L, Li, x, y, z, q, rcut, alpha, acc, energy);
        }
        fr[t1] += acc[0];
        fr[(long)N + t1] += acc[1];
        fr[2L * N + t1] += acc[2];
    }
}

static int cell_coord(double r, double L, int nc)
{
    int c = (int)floor(r / L * nc);
    c %= nc;
    if (c < 0)
        c += nc;
    return c;
}

/* Same pair set as orc_realteil: every j != i whose minimum-image distance
 * is <= rcut.  Cells of side >= rcut, 27 neighbour cells (deduplicated when
 * there are fewer than 3 cells per side).  Only particles i_begin..i_end-1
 * receive forces (used for a bounded CPU-timing sample). */
void orc_realteil_cells(int N, double L, const double *x, const double *y, const double *z,
                        const double *q, double rcut, double alpha, double *fr, double *energy,
                        int i_begin, int i_end)
{
    double Li = 1.0 / L;
    int nc = (int)floor(L / rcut);
    long ncell, c;
    int *head, *next, t1;
    if (nc < 1)
        nc = 1;
    if (nc > 256)
        nc = 256;
    ncell = (long)nc * nc * nc;
    head = (int *)malloc(sizeof(int) * ncell);
    next = (int *)malloc(sizeof(int) * (N > 0 ? N : 1));
    for (c = 0; c < ncell; c++)
        head[c] = -1;
    for (t1 = N - 1; t1 >= 0; t1--) {
        long cc = ((long)cell_coord(x[t1], L, nc) * nc + cell_coord(y[t1], L, nc)) * nc +
                  cell_coord(z[t1], L, nc);
        next[t1] = head[cc];
        head[cc] = t1;
    }
    for (t1 = i_begin; t1 < i_end; t1++) {
        int cx = cell_coord(x[t1], L, nc), cy = cell_coord(y[t1], L, nc), cz = cell_coord(z[t1], L, nc);
        double acc[3] = {0.0, 0.0, 0.0};
        long seen[27];
        int nseen = 0, a, bb, cz2, s;
        for (a = -1; a <= 1; a++)
            for (bb = -1; bb <= 1; bb++)
                for (cz2 = -1; cz2 <= 1; cz2++) {
                    long cc = ((long)((cx + a + nc) % nc) * nc + (cy + bb + nc) % nc) * nc +
                              (cz + cz2 + nc) % nc;
                    int dup = 0, t2;
                    for (s = 0; s < nseen; s++)
                        if (seen[s] == cc)
                            dup = 1;
                    if (dup)
                        continue;
                    seen[nseen++] = cc;
                    for (t2 = head[cc]; t2 >= 0; t2 = next[t2]) {
                        if (t2 == t1)
                            continue;
                        pair_force(t1, t2, L, Li, x, y, z, q, rcut, alpha, acc, energy);
                    }
                }
        fr[t1] += acc[0];
        fr[(long)N + t1] += acc[1];
        fr[2L * N + t1] += acc[2];
    }
    free(head);
    free(next);
}

/* ------------------------------------------------------------------------ */
/* Ewald reference                                                           */
/* ------------------------------------------------------------------------ */

/* ewald.c:84-109.  G has (kmax+1)^3 entries; entries with n^2 > kmax^2 are 0. */
void orc_ewald_influence(int kmax, double L, double alpha, double *G)
{
    double fak1 = 2.0 / SQR(L), fak2 = SQR(M_PI / (alpha * L));
    int K = kmax + 1, nx, ny, nz;
    for (nx = 0; nx <= kmax; nx++)
        for (ny = 0; ny <= kmax; ny++)
            for (nz = 0; nz <= kmax; nz++) {
                double n2 = SQR(nx) + SQR(ny) + SQR(nz);
                long ind = ((long)nx * K + ny) * K + nz;
                if ((nx == 0 && ny == 0 && nz == 0) || n2 > (double)kmax * kmax)
                    G[ind] = 0.0;
                else
                    G[ind] = fak1 / n2 * exp(-fak2 * n2);
            }
}

/* ewald.c:122-163: fk += sum_n n q_i G(n) (Re S sin(k.r_i) + Im S cos(k.r_i)),
 * S = sum_j q_j exp(-i k.r_j), over the cube |n_d| <= kmax, n^2 <= kmax^2. */
void orc_ewald_kspace(int N, double L, const double *x, const double *y, const double *z,
                      const double *q, double alpha, int kmax, const double *G, double *fk)
{
    double Leni = 1.0 / L;
    int K = kmax + 1, nx, ny, nz, i;
    (void)alpha;
    for (nx = -kmax; nx <= kmax; nx++)
        for (ny = -kmax; ny <= kmax; ny++)
            for (nz = -kmax; nz <= kmax; nz++) {
                double re = 0.0, im = 0.0, g;
                if (nx * nx + ny * ny + nz * nz > kmax * kmax)
                    continue;
                g = G[((long)abs(nx) * K + abs(ny)) * K + abs(nz)];
                if (g == 0.0)
                    continue;
                for (i = 0; i < N; i++) {
                    double kr = 2.0 * M_PI * Leni * (nx * x[i] + ny * y[i] + nz * z[i]);
                    re += q[i] * cos(kr);
                    im += q[i] * -sin(kr);
                }
                for (i = 0; i < N; i++) {
                    double kr = 2.0 * M_PI * Leni * (nx * x[i] + ny * y[i] + nz * z[i]);
                    double ff = q[i] * g * (re * sin(kr) + im * cos(kr));
                    fk[i] += nx * ff;
                    fk[(long)N + i] += ny * ff;
                    fk[2L * N + i] += nz * ff;
                }
            }
}

/* Kolafa-Perram real-space estimate (ewald.c:175-182) */
double orc_ewald_error_r(int N, double L, double q2, double rcut, double alpha)
{
    (void)N;
    return q2 * sqrt(rcut / (2.0 * L * L * L)) * exp(-SQR(alpha) * rcut * rcut) /
           (SQR(alpha) * rcut * rcut);
}

/* Petersen k-space estimate (ewald.c:184-198) */
double orc_ewald_error_k(int N, double L, double q2, int kmax, double alpha)
{
    return 2.0 * q2 * alpha * (1.0 / L) * sqrt(1.0 / (M_PI * kmax * N)) *
           exp(-SQR(M_PI * kmax / (alpha * L)));
}

/* bisection (ewald.c:204-229) */
double orc_ewald_optimal_alpha(int N, double L, double q2, double rcut, int kmax)
{
    double lo = 0.01, hi = 10.0, flo, fhi, g, fg;
    flo = orc_ewald_error_r(N, L, q2, rcut, lo) - orc_ewald_error_k(N, L, q2, kmax, lo);
    fhi = orc_ewald_error_r(N, L, q2, rcut, hi) - orc_ewald_error_k(N, L, q2, kmax, hi);
    if (flo * fhi > 0.0)
        return -1.0;
    do {
        g = 0.5 * (lo + hi);
        fg = orc_ewald_error_r(N, L, q2, rcut, g) - orc_ewald_error_k(N, L, q2, kmax, g);
        if (flo * fg < 0.0)
            hi = g;
        else
            lo = g;
    } while (fabs(lo - hi) > 1.0e-10);
    return 0.5 * (lo + hi);
}

/* ------------------------------------------------------------------------ */
/* Error estimates                                                           */
/* ------------------------------------------------------------------------ */

/* realpart.c:343-346 */
double orc_realspace_error(int N, double L, double q2, double rcut, double alpha)
{
    return (2.0 * q2 * exp(-SQR(rcut * alpha))) / sqrt((double)N * rcut * L * L * L);
}

/* analytic_cotangent_sum (p3m-common.c:77-107) */
static double cotangent_sum(int n, double mesh_i, int cao)
{
    double c = SQR(cos(M_PI * mesh_i * (double)n));
    switch (cao) {
    case 1: return 1.0;
    case 2: return (1.0 + c * 2.0) / 3.0;
    case 3: return (2.0 + c * (11.0 + c * 2.0)) / 15.0;
    case 4: return (17.0 + c * (180.0 + c * (114.0 + c * 4.0))) / 315.0;
    case 5: return (62.0 + c * (1072.0 + c * (1452.0 + c * (247.0 + c * 2.0)))) / 2835.0;
    case 6: return (1382.0 + c * (35396.0 + c * (83021.0 + c * (34096.0 + c * (2026.0 + c * 4.0))))) / 155925.0;
    case 7: return (21844.0 + c * (776661.0 + c * (2801040.0 + c * (2123860.0 + c * (349500.0 + c * (8166.0 + c * 4.0)))))) / 6081075.0;
    }
    return 0.0;
}

/* find_error.c:39-48: linear interpolation on alpha*L = 0,1,..,100 */
double orc_table_lookup(const double *table, double alphaL)
{
    double d = alphaL / 1.0;
    int l = (int)floor(d);
    d -= l;
    if (l >= 100)
        return table[100];
    return (1. - d) * table[l] + d * table[l + 1];
}

/* tuning alias sums of p3m-ik.c:304-337, p3m-ik-i.c:249-291,
 * p3m-ad.c:436-458, p3m-ad-i.c:205-250 */
static void alias_sums(int nx, int ny, int nz, int M, int cao, double L, double alpha, int mc,
                       double *a)
{
    double mesh_i = 1.0 / M, f1 = SQR(M_PI / (alpha * L));
    int mx, my, mz, t;
    for (t = 0; t < 8; t++)
        a[t] = 0.0;
    for (mx = -mc; mx <= mc; mx++) {
        double nmx = nx + mx * M, fnmx = mesh_i * nmx;
        for (my = -mc; my <= mc; my++) {
            double nmy = ny + my * M, fnmy = mesh_i * nmy;
            for (mz = -mc; mz <= mc; mz++) {
                double nmz = nz + mz * M, fnmz = mesh_i * nmz;
                double nm2 = SQR(nmx) + SQR(nmy) + SQR(nmz);
                double ex = exp(-f1 * nm2), ex2 = SQR(ex);
                double U2 = ipow(sinc(fnmx) * sinc(fnmy) * sinc(fnmz), 2 * cao);
                int even = ((mx + my + mz) % 2) == 0;
                a[0] += ex2 / nm2;                                 /* alias1 */
                a[1] += U2 * ex * (nx * nmx + ny * nmy + nz * nmz) / nm2; /* ik alias2 */
                a[2] += U2 * ex;                                   /* ad alias2 */
                a[3] += U2 * nm2;                                  /* ad alias3 */
                a[4] += U2;                                        /* alias3/4 */
                a[5] += even ? U2 : -U2;                           /* ik-i alias4, ad-i alias6 */
                a[6] += even ? U2 * nm2 : -U2 * nm2;               /* ad-i alias5 */
            }
        }
    }
}

double orc_kspace_error(int method, int N, double L, double q2, int M, int cao,
                        double alpha, int mc, const double *table)
{
    double he_q = 0.0, mesh_i = 1.0 / M, a[8];
    int nx, ny, nz, h = M / 2;
    if (method == 0 || method == 6) {
        /* p3m-ik.c:263-301: a tabulated value >= 0 is returned as is; otherwise
         * the direct sum starts from the (negative) lookup result, -1 when the
         * mesh is not tabulated */
        he_q = (table != NULL) ? orc_table_lookup(table, alpha * L) : -1.0;
        if (he_q >= 0.0)
            return 2.0 * q2 * sqrt(he_q / (double)N) / SQR(L);
        for (nx = -h; nx < h; nx++) {
            double cx = cotangent_sum(nx, mesh_i, cao);
            for (ny = -h; ny < h; ny++) {
                double cy = cx * cotangent_sum(ny, mesh_i, cao);
                for (nz = -h; nz < h; nz++) {
                    if (nx != 0 || ny != 0 || nz != 0) {
                        double n2 = SQR(nx) + SQR(ny) + SQR(nz);
                        double cs = cotangent_sum(nz, mesh_i, cao) * cy;
                        alias_sums(nx, ny, nz, M, cao, L, alpha, mc, a);
                        he_q += (a[0] - SQR(a[1] / cs) / n2);
                    }
                }
            }
        }
        he_q = fabs(he_q);
    } else if (method == 1) {
        /* p3m-ik-i.c:295-320: a negative lookup restarts from 0 */
        he_q = (table != NULL) ? orc_table_lookup(table, alpha * L) : -1.0;
        if (he_q >= 0.0)
            return 2.0 * q2 * sqrt(he_q / (double)N) / SQR(L);
        he_q = 0.0;
        for (nx = -h; nx < h; nx++)
            for (ny = -h; ny < h; ny++)
                for (nz = -h; nz < h; nz++)
                    if (nx != 0 || ny != 0 || nz != 0) {
                        double n2 = SQR(nx) + SQR(ny) + SQR(nz);
                        alias_sums(nx, ny, nz, M, cao, L, alpha, mc, a);
                        he_q += a[0] - SQR(a[1]) / (0.5 * n2 * (SQR(a[4]) + SQR(a[5])));
                    }
        he_q = fabs(he_q);
    } else if (method == 2 || method == 7) {
        /* p3m-ad.c:475-497: no fabs, any zero component skipped */
        for (nx = -h; nx < h; nx++)
            for (ny = -h; ny < h; ny++)
                for (nz = -h; nz < h; nz++)
                    if (nx != 0 && ny != 0 && nz != 0) {
                        alias_sums(nx, ny, nz, M, cao, L, alpha, mc, a);
                        if (a[3] == 0.0 || a[4] == 0.0)
                            continue;
                        he_q += a[0] - (SQR(a[2]) / (a[3] * a[4]));
                    }
    } else if (method == 3) {
        /* p3m-ad-i.c:252-284 */
        for (nx = -h; nx < h; nx++)
            for (ny = -h; ny < h; ny++)
                for (nz = -h; nz < h; nz++)
                    if (nx != 0 && ny != 0 && nz != 0) {
                        alias_sums(nx, ny, nz, M, cao, L, alpha, mc, a);
                        he_q += (a[0] - SQR(a[2]) / (0.5 * (a[3] * a[4] + a[6] * a[5])));
                    }
        he_q = fabs(he_q);
    } else {
        return -1.0;
    }
    return 2.0 * q2 * sqrt(he_q / (double)N) / SQR(L);
}

/* error.c:21-34 */
double orc_force_error(int N, const double *f, const double *fref)
{
    double e = 0.0;
    long i;
    for (i = 0; i < 3L * N; i++)
        e += SQR(f[i] - fref[i]);
    return sqrt(e);
}
