/*
 * Inhomogeneous error estimate and error map of the reference
 * (Generic_error_estimate_inhomo, p3m-common.c:437-719; Error_map,
 * p3m-common.c:402-435), on the GPU.
 *
 * The estimate:
 *   1. rho^2: q_i^2 assigned with order cao on the estimate's mesh
 *      (assign_charge_q2, charge-assign.c:1544-1597), or q2 / M^3 at every
 *      point when `uniform` is set;
 *   2. K^2(r) = sum_d Re(K_d(r))^2 of the homogeneous error kernel (no
 *      aliasing: it vanishes to rounding, as in the reference) plus, for every
 *      aliasing copy m in [-mc, mc]^3 \ {0}, that of the copy's kernel; each
 *      K_d(k) goes through a backward transform;
 *   3. forward transforms of K^2 and rho^2, the reference's part-wise product
 *      (real with real, imaginary with imaginary), a backward transform;
 *   4. rms_i = q_i^2 sum_stencil W Re(...) (collect_rms_nocf,
 *      charge-assign.c:1655-1702); the estimate is sqrt(sum_i rms_i / N).
 * One deliberate difference: the reference gathers with the method's mesh
 * and order (p3m-common.c:681 passes the method's parameters with the
 * estimate's mesh and weight table, reading past the table's rows when the
 * method's order exceeds the estimate's by two); here the gather uses the
 * estimate's own mesh and order, as its assignment does.  Both agree when
 * the method's mesh and order equal the estimate's.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "host_common.h"
#include "p3m_hip.h"
#include "p3m_kernels.h"

typedef struct {
    p3mk_queue *qu;
    p3mk_fft *fft;
    p3mk_psort *ps;
    double *x, *y, *z, *q, *w, *rms, *scratch;
    void *Q, *Km, *K[3], *K3;
} inhomo_ws;

static void ws_release(inhomo_ws *w) {
    void *bufs[] = {w->x, w->y, w->z, w->q, w->w, w->rms, w->scratch, w->Q, w->Km,
                    w->K[0], w->K[1], w->K[2], w->K3};
    size_t k;
    for (k = 0; k < sizeof bufs / sizeof bufs[0]; k++)
        p3mk_release(bufs[k]);
    if (w->ps)
        p3mk_psort_destroy(w->ps);
    if (w->fft)
        p3mk_fft_destroy(w->fft);
    if (w->qu)
        p3mk_queue_destroy(w->qu);
    memset(w, 0, sizeof *w);
}

static const double *coord(const vector_array_t *v, int d) {
    if (v->fields)
        return v->fields[d];
    return d == 0 ? v->x : (d == 1 ? v->y : v->z);
}

/* device copies of the positions (and charges), queue, FFT and sort for a
 * complex M^3 mesh of order cao */
static int ws_setup(inhomo_ws *w, const system_t *s, const p3mk_geom *g, int nmeshes) {
    const size_t n = (size_t)s->nparticles, b = n * sizeof(double);
    const size_t mb = p3mk_mesh_bytes(g);
    int bad = 0, rc, k, *flag = NULL;
    memset(w, 0, sizeof *w);
    TRYK(p3mk_device_check(), "no usable GPU");
    rc = p3mk_queue_create(&w->qu);
    if (rc == P3MK_OK) rc = p3mk_fft_create(&w->fft, g);
    if (rc == P3MK_OK) rc = p3mk_psort_create(&w->ps, g, s->nparticles);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&w->x, b);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&w->y, b);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&w->z, b);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&w->q, b);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&w->w, b);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&w->rms, b);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&w->scratch,
                                       sizeof(double) * (size_t)p3mk_inhomo_scratch_doubles());
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&w->Q, mb);
    for (k = 0; k < nmeshes - 1 && rc == P3MK_OK; k++) {
        void **slot = k == 0 ? &w->Km : (k <= 3 ? &w->K[k - 1] : &w->K3);
        rc = p3mk_alloc(slot, mb);
    }
    if (rc == P3MK_OK) rc = p3mk_copy_h2d(w->x, coord(s->p, 0), b, w->qu);
    if (rc == P3MK_OK) rc = p3mk_copy_h2d(w->y, coord(s->p, 1), b, w->qu);
    if (rc == P3MK_OK) rc = p3mk_copy_h2d(w->z, coord(s->p, 2), b, w->qu);
    if (rc == P3MK_OK) rc = p3mk_copy_h2d(w->q, s->q, b, w->qu);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&flag, sizeof(int));
    if (rc == P3MK_OK)
        rc = p3mk_check_positions(s->nparticles, -s->length, 2.0 * s->length, w->x, w->y, w->z,
                                  &bad, flag, w->qu);
    p3mk_release(flag);
    if (rc != P3MK_OK) {
        ws_release(w);
        return p3mh_kfail(rc, "inhomogeneous estimate: device set-up");
    }
    if (bad) {
        ws_release(w);
        return p3mh_fail(P3M_HIP_ERANGE, "%d positions outside [-L, 2L) or not finite", bad);
    }
    return P3M_HIP_OK;
}

static int check_mesh_args(const system_t *s, int mesh, int cao) {
    if (s == NULL || s->p == NULL || s->q == NULL || s->nparticles <= 0 || !(s->length > 0.0))
        return p3mh_fail(P3M_HIP_EINVAL, "bad system");
    if (mesh < 2 || mesh % 2 != 0 || cao < 1 || cao > 7 || cao > mesh)
        return p3mh_fail(P3M_HIP_EINVAL, "mesh must be even (factors 2, 3, 5, 7, 11, 13) and cao 1..7");
    return P3M_HIP_OK;
}

static int write_density(const char *path, int M, const double *Q, const double *Km,
                         const double *K3) {
    FILE *f = fopen(path, "w");
    long ind;
    int nx, ny, nz;
    if (f == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "cannot open '%s' for writing", path);
    /* p3m-common.c:686-696 */
    for (nx = 0; nx < M; nx++)
        for (ny = 0; ny < M; ny++)
            for (nz = 0; nz < M; nz++) {
                ind = 2 * (((long)M * M * nx) + (long)M * ny + nz);
                fprintf(f, "%d %d %d %e %e %e %e %e %e\n", nx, ny, nz, Q[ind], Km[ind], Km[ind + 1],
                        K3[ind], K3[ind + 1], Km[ind] * Q[ind]);
            }
    if (fclose(f) != 0)
        return p3mh_fail(P3M_HIP_EINVAL, "writing '%s' failed", path);
    return P3M_HIP_OK;
}

static int estimate(inhomo_ws *w, const system_t *s, const p3mk_geom *g, double alpha, int uniform,
                    int mc, const char *out_file, double *result) {
    const int M = g->M, n = s->nparticles;
    const long M3 = (long)M * M * M;
    int m[3], k, first = 1;
    double S[3], sum = 0.0, *host = NULL;

    /* rho^2 (p3m-common.c:474-480) */
    if (uniform) {
        TRYK(p3mk_fill_real(M, s->q2 / (double)M3, w->Q, w->qu), "uniform density");
    } else {
        TRYK(p3mk_square_charges(n, w->q, w->w, w->qu), "squared charges");
        TRYK(p3mk_assign(g, n, w->x, w->y, w->z, w->w, w->ps, w->Q, w->qu), "q^2 assignment");
    }
    /* homogeneous kernel (p3m-common.c:490-560) */
    TRYK(p3mk_inhomo_homo(M, g->L, alpha, g->cao, w->K, w->qu), "homogeneous kernel");
    for (k = 0; k < 3; k++)
        TRYK(p3mk_fft3d(w->fft, w->K[k], +1, w->qu), "kernel transform");
    TRYK(p3mk_inhomo_square_add(M, w->K, w->K3, first, w->qu), "squared kernel");
    first = 0;
    /* aliasing copies (p3m-common.c:567-640) */
    for (m[0] = -mc; m[0] <= mc; m[0]++)
        for (m[1] = -mc; m[1] <= mc; m[1]++)
            for (m[2] = -mc; m[2] <= mc; m[2]++) {
                if (m[0] == 0 && m[1] == 0 && m[2] == 0)
                    continue;
                TRYK(p3mk_inhomo_sums(M, g->L, alpha, g->cao, m, S, w->scratch, w->qu), "kernel sums");
                TRYK(p3mk_inhomo_fill(M, g->L, alpha, g->cao, m, S, w->K, w->qu), "aliasing kernel");
                for (k = 0; k < 3; k++)
                    TRYK(p3mk_fft3d(w->fft, w->K[k], +1, w->qu), "kernel transform");
                TRYK(p3mk_inhomo_square_add(M, w->K, w->K3, first, w->qu), "squared kernel");
            }
    /* K^2(k), rho^2(k), product, back (p3m-common.c:644-668) */
    TRYK(p3mk_fft3d(w->fft, w->K3, -1, w->qu), "K^2 transform");
    TRYK(p3mk_copy_d2d(w->Km, w->Q, p3mk_mesh_bytes(g), w->qu), "copy");
    TRYK(p3mk_fft3d(w->fft, w->Km, -1, w->qu), "rho^2 transform");
    TRYK(p3mk_inhomo_product(M, w->Km, w->K3, w->qu), "product");
    TRYK(p3mk_fft3d(w->fft, w->Km, +1, w->qu), "error density transform");
    /* gather onto the particles (p3m-common.c:676-681, 701-706) */
    TRYK(p3mk_collect_rms(M, g->cao, g->L, n, w->x, w->y, w->z, w->q, w->Km, w->rms, w->qu),
         "gather");
    host = (double *)malloc(sizeof(double) * (size_t)(n > 2 * M3 ? n : 2 * M3) * 3);
    if (host == NULL)
        return p3mh_fail(P3M_HIP_ENOMEM, "host buffer");
    if (p3mk_copy_d2h(host, w->rms, sizeof(double) * (size_t)n, w->qu) != P3MK_OK ||
        p3mk_queue_sync(w->qu) != P3MK_OK) {
        free(host);
        return p3mh_kfail(P3MK_EDEVICE, "copy");
    }
    for (k = 0; k < n; k++)
        sum += host[k];
    *result = sqrt(sum / n);
    if (out_file != NULL) {
        double *Qh = host, *Kh = host + 2 * M3, *K3h = host + 4 * M3;
        int rc = P3M_HIP_OK;
        if (p3mk_copy_d2h(Qh, w->Q, p3mk_mesh_bytes(g), w->qu) != P3MK_OK ||
            p3mk_copy_d2h(Kh, w->Km, p3mk_mesh_bytes(g), w->qu) != P3MK_OK ||
            p3mk_copy_d2h(K3h, w->K3, p3mk_mesh_bytes(g), w->qu) != P3MK_OK ||
            p3mk_queue_sync(w->qu) != P3MK_OK)
            rc = p3mh_kfail(P3MK_EDEVICE, "copy");
        if (rc == P3M_HIP_OK)
            rc = write_density(out_file, M, Qh, Kh, K3h);
        free(host);
        return rc;
    }
    free(host);
    return P3M_HIP_OK;
}

int p3m_hip_error_inhomo(system_t *s, parameters_t *p, int uniform, int mesh, int cao, int mc,
                         const char *out_file, double *result) {
    p3mk_geom g;
    inhomo_ws w;
    int rc;
    if (p == NULL || result == NULL || mc < 0)
        return p3mh_fail(P3M_HIP_EINVAL, "error_inhomo: bad arguments");
    TRY(check_mesh_args(s, mesh, cao));
    if (!(p->alpha > 0.0))
        return p3mh_fail(P3M_HIP_EINVAL, "error_inhomo: alpha must be positive");
    memset(&g, 0, sizeof g);
    g.M = mesh;
    g.cao = cao;
    g.L = s->length;
    TRY(ws_setup(&w, s, &g, 6));
    rc = estimate(&w, s, &g, p->alpha, uniform, mc, out_file, result);
    ws_release(&w);
    return rc;
}

int p3m_hip_error_map(system_t *s, forces_t *f, forces_t *f_ref, int mesh, int cao,
                      double *error_mesh) {
    p3mk_geom g;
    inhomo_ws w;
    double *dev_f = NULL, *dev_r = NULL;
    int rc, d;
    const size_t n = s ? (size_t)s->nparticles : 0;
    if (f == NULL || f_ref == NULL || f->f_k == NULL || f_ref->f_k == NULL || error_mesh == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "error_map: bad arguments");
    TRY(check_mesh_args(s, mesh, cao));
    memset(&g, 0, sizeof g);
    g.M = mesh;
    g.cao = cao;
    g.L = s->length;
    TRY(ws_setup(&w, s, &g, 1));
    rc = p3mk_alloc((void **)&dev_f, 3 * n * sizeof(double));
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&dev_r, 3 * n * sizeof(double));
    for (d = 0; d < 3 && rc == P3MK_OK; d++) {
        rc = p3mk_copy_h2d(dev_f + d * n, coord(f->f_k, d), n * sizeof(double), w.qu);
        if (rc == P3MK_OK)
            rc = p3mk_copy_h2d(dev_r + d * n, coord(f_ref->f_k, d), n * sizeof(double), w.qu);
    }
    /* charges |f_ref,k - f_k|^2, assigned (p3m-common.c:413-427) */
    if (rc == P3MK_OK) rc = p3mk_diff_charges((int)n, dev_f, dev_r, w.w, w.qu);
    if (rc == P3MK_OK) rc = p3mk_assign(&g, (int)n, w.x, w.y, w.z, w.w, w.ps, w.Q, w.qu);
    if (rc == P3MK_OK) rc = p3mk_copy_d2h(error_mesh, w.Q, p3mk_mesh_bytes(&g), w.qu);
    if (rc == P3MK_OK) rc = p3mk_queue_sync(w.qu);
    p3mk_release(dev_f);
    p3mk_release(dev_r);
    ws_release(&w);
    if (rc != P3MK_OK)
        return p3mh_kfail(rc, "error map");
    return P3M_HIP_OK;
}
