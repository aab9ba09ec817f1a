/* Host side of the MI355X P3M library: the method table and its entry points.
 *
 * This layer is plain C.  It owns the method logic of the reference
 * (which stages run for which method id, prefactors, argument checks,
 * accumulation semantics) and reaches the GPU only through the launchers of
 * p3m_kernels.h.  There is no CPU fallback: every numeric stage of
 * Kspace_force, the influence function, the self-force coefficients, the
 * error sums and the real-space part run as HIP kernels.
 */
#include "p3m_hip.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../p3m_kernels.h"
#include "host_common.h"
#include "q_tables.h"

#define PI_D 3.14159265358979323846

/* ---------------------------------------------------------------------- */
/* errors                                                                  */
/* ---------------------------------------------------------------------- */

static _Thread_local char g_err[512];

const char *p3m_hip_last_error(void) { return g_err; }

int p3mh_fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int p3mh_kfail(int code, const char *what) {
    return p3mh_fail(code, "%s: %s", what, p3mk_error());
}

/* ---------------------------------------------------------------------- */
/* method descriptions                                                     */
/* ---------------------------------------------------------------------- */

#define is_ad p3mh_is_ad
#define is_interlaced p3mh_is_interlaced
#define is_r2c p3mh_is_r2c
#define has_self_force p3mh_has_self_force
#define known_method p3mh_known_method

/* device state hanging off data_t.method_data */
typedef struct {
    int method;
    int precision;
    p3mk_geom geom;
    int n;
    p3mk_queue *q;
    p3mk_fft *fft[2]; /* [0] FP64, [1] FP32 */
    void *mesh;
    void *spec[3];
    void *F[3];
    double *G, *Dn, *selfb, *scratch;
    double *pos;    /* x, y, z, q blocks of n */
    double *force;  /* fk (3n), f (3n), fr (3n) */
    double *hforce; /* pinned staging, 9n */
    double *energy; /* device scalar */
    int *iscratch;
    p3mk_cells *cells;
    p3mk_psort *psort;
    double cells_L, cells_rc;
    int have_G;
    int have_ks_events;
    int have_rs_events;
    float stage_ms[5];
    unsigned char *ring_rs; /* per event-ring set: real space ran */
    int ring_sets;
    void *adj[3]; /* adjoint meshes, made by the first VJP call */
} p3m_ctx;

static p3m_ctx *ctx_of(data_t *d) { return d ? (p3m_ctx *)d->method_data : NULL; }

static void ctx_free(p3m_ctx *c) {
    int k;
    if (!c)
        return;
    if (c->q)
        p3mk_queue_sync(c->q);
    p3mk_fft_destroy(c->fft[0]);
    p3mk_fft_destroy(c->fft[1]);
    p3mk_release(c->mesh);
    for (k = 0; k < 3; k++) {
        p3mk_release(c->spec[k]);
        if (c->F[k] != c->mesh)
            p3mk_release(c->F[k]);
    }
    for (k = 0; k < 3; k++)
        p3mk_release(c->adj[k]);
    p3mk_release(c->G);
    p3mk_release(c->Dn);
    p3mk_release(c->selfb);
    p3mk_release(c->scratch);
    p3mk_release(c->pos);
    p3mk_release(c->force);
    p3mk_release(c->energy);
    p3mk_release(c->iscratch);
    p3mk_release_host(c->hforce);
    p3mk_cells_destroy(c->cells);
    p3mk_psort_destroy(c->psort);
    p3mk_queue_destroy(c->q);
    free(c->ring_rs);
    free(c);
}

static int check_params(int id, const system_t *s, const parameters_t *p) {
    if (!known_method(id))
        return p3mh_fail(P3M_HIP_EINVAL, "method id %d is not a P3M method of this library", id);
    if (s == NULL || p == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "system and parameters must not be NULL");
    if (p->cao < 1 || p->cao > 7)
        return p3mh_fail(P3M_HIP_EINVAL, "charge assignment order %d not in 1..7", p->cao);
    if (p->mesh < 2 || (p->mesh % 2) != 0)
        return p3mh_fail(P3M_HIP_EINVAL, "mesh %d must be even and >= 2", p->mesh);
    if (p->mesh < p->cao)
        return p3mh_fail(P3M_HIP_EINVAL, "mesh %d smaller than cao %d", p->mesh, p->cao);
    if (s->nparticles < 0)
        return p3mh_fail(P3M_HIP_EINVAL, "negative particle count");
    if (!(s->length > 0.0) || !isfinite(s->length))
        return p3mh_fail(P3M_HIP_EINVAL, "box length must be positive");
    return P3M_HIP_OK;
}

static int ctx_alloc(p3m_ctx *c) {
    p3mk_geom g64 = c->geom;
    size_t mesh_b, spec_b, M3 = (size_t)c->geom.M * c->geom.M * c->geom.M;
    size_t n = (size_t)(c->n > 0 ? c->n : 1);
    int k;
    g64.fp32 = 0; /* FP64 sizes also hold the FP32 form */
    mesh_b = p3mk_mesh_bytes(&g64);
    spec_b = p3mk_spec_bytes(&g64);
    TRYK(p3mk_queue_create(&c->q), "stream");
    {
        p3mk_geom g32 = g64;
        g32.fp32 = 1;
        TRYK(p3mk_fft_create(&c->fft[0], &g64), "FFT plan");
        TRYK(p3mk_fft_create(&c->fft[1], &g32), "FFT plan");
    }
    TRYK(p3mk_alloc(&c->mesh, mesh_b), "mesh");
    TRYK(p3mk_psort_create(&c->psort, &g64, c->n), "particle sort");
    if (is_r2c(c->method)) {
        int nspec = is_ad(c->method) ? 1 : 3;
        for (k = 0; k < nspec; k++)
            TRYK(p3mk_alloc(&c->spec[k], spec_b), "spectrum");
        if (!is_ad(c->method)) {
            c->F[0] = c->mesh;
            TRYK(p3mk_alloc(&c->F[1], mesh_b), "field mesh");
            TRYK(p3mk_alloc(&c->F[2], mesh_b), "field mesh");
        }
    } else if (!is_ad(c->method)) {
        for (k = 0; k < 3; k++)
            TRYK(p3mk_alloc(&c->F[k], mesh_b), "field mesh");
    }
    TRYK(p3mk_alloc((void **)&c->G, sizeof(double) * M3), "influence function");
    TRYK(p3mk_alloc((void **)&c->Dn, sizeof(double) * c->geom.M), "Dn");
    TRYK(p3mk_alloc((void **)&c->selfb, sizeof(double) * 81), "self-force coefficients");
    TRYK(p3mk_alloc((void **)&c->scratch, sizeof(double) * p3mk_reduce_scratch_doubles(c->geom.M)),
         "reduction scratch");
    TRYK(p3mk_alloc((void **)&c->pos, sizeof(double) * 4 * n), "positions");
    TRYK(p3mk_alloc((void **)&c->force, sizeof(double) * 9 * n), "forces");
    TRYK(p3mk_alloc((void **)&c->energy, sizeof(double)), "energy");
    TRYK(p3mk_alloc((void **)&c->iscratch, sizeof(int) * 16), "int scratch");
    TRYK(p3mk_alloc_host((void **)&c->hforce, sizeof(double) * 9 * n), "pinned staging");
    TRYK(p3mk_fill_zero(c->selfb, sizeof(double) * 81, c->q), "zero");
    TRYK(p3mk_fill_zero(c->G, sizeof(double) * M3, c->q), "zero");
    TRYK(p3mk_queue_sync(c->q), "sync");
    return P3M_HIP_OK;
}

static p3mk_fft *ctx_fft(p3m_ctx *c) { return c->fft[c->geom.fp32 ? 1 : 0]; }

/* ---------------------------------------------------------------------- */
/* Influence function                                                      */
/* ---------------------------------------------------------------------- */

int p3m_hip_influence_function(system_t *s, parameters_t *p, data_t *d) {
    p3m_ctx *c = ctx_of(d);
    if (c == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "data_t was not made by p3m_hip_init");
    TRY(check_params(c->method, s, p));
    if (p->mesh != c->geom.M || p->cao != c->geom.cao)
        return p3mh_fail(P3M_HIP_EINVAL, "mesh/cao differ from the values given to Init");
    c->geom.L = s->length;
    TRYK(p3mk_influence(c->method, &c->geom, p->alpha, 1, c->G, c->Dn, c->q), "influence function");
    /* ad / ad-r also recompute the self-force coefficients, except at
     * alpha == 0 where the reference returns early (p3m-ad.c:120-123,138-142) */
    if (has_self_force(c->method) && p->alpha != 0.0) {
        TRYK(p3mk_self_force_coeffs(&c->geom, c->G, c->selfb, c->scratch, c->q), "self force");
        if (d->self_force_corrections != NULL)
            TRYK(p3mk_copy_d2h(d->self_force_corrections, c->selfb, sizeof(double) * 81, c->q),
                 "copy");
    }
    TRYK(p3mk_queue_sync(c->q), "sync");
    c->have_G = 1;
    return P3M_HIP_OK;
}

/* ---------------------------------------------------------------------- */
/* Init / free                                                             */
/* ---------------------------------------------------------------------- */

int p3m_hip_init(int method_id, system_t *s, parameters_t *p, int precision, data_t **out) {
    data_t *d;
    p3m_ctx *c;
    int rc;
    if (out == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "out must not be NULL");
    *out = NULL;
    TRY(check_params(method_id, s, p));
    if (precision != P3M_HIP_FP64 && precision != P3M_HIP_MIXED)
        return p3mh_fail(P3M_HIP_EINVAL, "unknown precision %d", precision);
    if (p3mk_device_check() != P3MK_OK)
        return p3mh_kfail(P3M_HIP_ENODEVICE, "device");
    d = (data_t *)calloc(1, sizeof(data_t));
    c = (p3m_ctx *)calloc(1, sizeof(p3m_ctx));
    if (d == NULL || c == NULL) {
        free(d);
        free(c);
        return p3mh_fail(P3M_HIP_ENOMEM, "host allocation");
    }
    d->mesh = p->mesh;
    d->method_data = c;
    c->method = method_id;
    c->precision = precision;
    c->n = s->nparticles;
    c->geom.M = p->mesh;
    c->geom.cao = p->cao;
    c->geom.r2c = is_r2c(method_id);
    c->geom.interlaced = is_interlaced(method_id);
    c->geom.fp32 = precision == P3M_HIP_MIXED;
    c->geom.L = s->length;
    if (has_self_force(method_id)) {
        d->self_force_corrections = (double *)calloc(81, sizeof(double));
        if (d->self_force_corrections == NULL) {
            p3m_hip_free(d);
            return p3mh_fail(P3M_HIP_ENOMEM, "host allocation");
        }
    }
    rc = ctx_alloc(c);
    if (rc == P3M_HIP_OK)
        rc = p3m_hip_influence_function(s, p, d); /* p3m-common.c:215-222 */
    if (rc != P3M_HIP_OK) {
        char keep[512];
        memcpy(keep, g_err, sizeof(keep));
        p3m_hip_free(d);
        memcpy(g_err, keep, sizeof(keep));
        return rc;
    }
    *out = d;
    return P3M_HIP_OK;
}

int p3m_hip_free(data_t *d) {
    if (d == NULL)
        return P3M_HIP_OK;
    ctx_free(ctx_of(d));
    free(d->self_force_corrections);
    free(d);
    return P3M_HIP_OK;
}

int p3m_hip_set_precision(data_t *d, int precision) {
    p3m_ctx *c = ctx_of(d);
    if (c == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "data_t was not made by p3m_hip_init");
    if (precision != P3M_HIP_FP64 && precision != P3M_HIP_MIXED)
        return p3mh_fail(P3M_HIP_EINVAL, "unknown precision %d", precision);
    c->precision = precision;
    c->geom.fp32 = precision == P3M_HIP_MIXED;
    return P3M_HIP_OK;
}

/* ---------------------------------------------------------------------- */
/* Kspace_force                                                            */
/* ---------------------------------------------------------------------- */

/* Assignment, forward FFT, influence-function multiply (ik: and gradient),
 * backward FFT: leaves the field meshes c->F[0..2] (ik) or the potential
 * mesh c->mesh (ad).  Records events 1 and 2. */
static int field_enqueue(p3m_ctx *c, int n, const double *x, const double *y, const double *z,
                         const double *q) {
    p3mk_fft *fft = ctx_fft(c);
    const p3mk_geom *g = &c->geom;
    void *spec = is_r2c(c->method) ? c->spec[0] : c->mesh;
    void *outs[3];
    int k;
    TRYK(p3mk_assign(g, n, x, y, z, q, c->psort, c->mesh, c->q), "charge assignment");
    TRYK(p3mk_record(c->q, 1), "event");
    TRYK(p3mk_fft_forward_zy(fft, c->mesh, spec, c->q), "forward FFT");
    if (!is_ad(c->method)) {
        /* ik: F_d = i (2 pi/L) Dn_d G Q^ (p3m-ik.c:211-231, p3m-ik-i.c:202-227,
         * p3m-ik-real.c:113-136), three backward transforms */
        for (k = 0; k < 3; k++)
            outs[k] = is_r2c(c->method) ? c->spec[k] : c->F[k];
        TRYK(p3mk_convolve(fft, 3, c->G, c->Dn, spec, outs, c->q), "convolution");
        for (k = 0; k < 3; k++)
            TRYK(p3mk_fft_backward_yz(fft, outs[k], c->F[k], c->q), "backward FFT");
    } else {
        /* ad: phi = FFT^-1[G Q^] (p3m-ad.c:182-196, p3m-ad-i.c:176-190,
         * p3m-ad-real.c:90-104) */
        outs[0] = spec;
        outs[1] = outs[2] = NULL;
        TRYK(p3mk_convolve(fft, 1, c->G, c->Dn, spec, outs, c->q), "convolution");
        TRYK(p3mk_fft_backward_yz(fft, spec, c->mesh, c->q), "backward FFT");
    }
    TRYK(p3mk_record(c->q, 2), "event");
    return P3M_HIP_OK;
}

/* prefactor of the back-interpolation sums: 1/(2 L^3) for ik (p3m-ik.c:242),
 * M/L^3 for ad (p3m-ad.c:202) */
static double interp_prefactor(const p3m_ctx *c) {
    const double L = c->geom.L;
    return is_ad(c->method) ? (double)c->geom.M / (L * L * L) : 1.0 / (2.0 * L * L * L);
}

static int kspace_enqueue(p3m_ctx *c, int n, const double *x, const double *y, const double *z,
                          const double *q, double *fk, double *f) {
    const p3mk_geom *g = &c->geom;
    TRYK(p3mk_record(c->q, 0), "event");
    TRY(field_enqueue(c, n, x, y, z, q));
    if (!is_ad(c->method)) {
        TRYK(p3mk_interp_ik(g, n, c->psort, c->F, interp_prefactor(c), fk, c->q),
             "back-interpolation");
    } else {
        /* assign_forces_ad* and the self-force subtraction of methods 2 and 7
         * (p3m-ad.c:204-206) */
        TRYK(p3mk_interp_ad(g, n, c->psort, c->mesh, interp_prefactor(c),
                            has_self_force(c->method) ? c->selfb : NULL, fk, f, c->q),
             "back-interpolation");
    }
    TRYK(p3mk_record(c->q, 3), "event");
    c->have_ks_events = 1;
    return P3M_HIP_OK;
}

static int collect_stage_times(p3m_ctx *c) {
    float t;
    int k;
    for (k = 0; k < 4; k++)
        c->stage_ms[k] = 0.0f;
    if (c->have_ks_events) {
        for (k = 0; k < 3; k++) {
            TRYK(p3mk_elapsed_ms(c->q, k, k + 1, &t), "event timing");
            c->stage_ms[k] = t;
        }
        TRYK(p3mk_elapsed_ms(c->q, 0, 3, &t), "event timing");
        c->stage_ms[3] = t;
    }
    c->stage_ms[4] = 0.0f;
    if (c->have_rs_events) {
        TRYK(p3mk_elapsed_ms(c->q, 4, 5, &t), "event timing");
        c->stage_ms[4] = t;
    }
    return P3M_HIP_OK;
}

int p3m_hip_kspace_force_device(data_t *d, int n, double length, const double *x,
                                const double *y, const double *z, const double *q, double *fk,
                                double *f) {
    p3m_ctx *c = ctx_of(d);
    if (c == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "data_t was not made by p3m_hip_init");
    if (n != c->n)
        return p3mh_fail(P3M_HIP_EINVAL, "particle count %d differs from Init's %d", n, c->n);
    if (length != c->geom.L)
        return p3mh_fail(P3M_HIP_EINVAL, "box length differs from the influence function's");
    if (!c->have_G)
        return p3mh_fail(P3M_HIP_EINVAL, "influence function not computed");
    if (has_self_force(c->method) && f == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "method %d needs the total-force array f", c->method);
    return kspace_enqueue(c, n, x, y, z, q, fk, f);
}

int p3m_hip_synchronize(data_t *d) {
    p3m_ctx *c = ctx_of(d);
    if (c == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "data_t was not made by p3m_hip_init");
    TRYK(p3mk_queue_sync(c->q), "sync");
    return P3M_HIP_OK;
}

static int check_host_system(const system_t *s) {
    if (s->p == NULL || s->p->x == NULL || s->p->y == NULL || s->p->z == NULL || s->q == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "system arrays must not be NULL");
    return P3M_HIP_OK;
}

static int upload_positions(p3m_ctx *c, const system_t *s) {
    const size_t n = (size_t)s->nparticles, b = sizeof(double) * n;
    int bad = 0;
    TRYK(p3mk_copy_h2d(c->pos, s->p->x, b, c->q), "upload");
    TRYK(p3mk_copy_h2d(c->pos + n, s->p->y, b, c->q), "upload");
    TRYK(p3mk_copy_h2d(c->pos + 2 * n, s->p->z, b, c->q), "upload");
    TRYK(p3mk_copy_h2d(c->pos + 3 * n, s->q, b, c->q), "upload");
    /* the reference wraps mesh indices by one period (charge-assign.c:37-39),
     * so it takes positions up to one box length outside [0, L) */
    TRYK(p3mk_check_positions((int)n, -s->length, 2.0 * s->length, c->pos, c->pos + n,
                              c->pos + 2 * n, &bad, c->iscratch, c->q),
         "position check");
    if (bad)
        return p3mh_fail(P3M_HIP_ERANGE,
                         "%d particles outside [-L, 2L) (the one-period tolerance of the "
                         "reference's mesh wrap) or not finite",
                         bad);
    return P3M_HIP_OK;
}

int p3m_hip_kspace_force(system_t *s, parameters_t *p, data_t *d, forces_t *f) {
    p3m_ctx *c = ctx_of(d);
    size_t n, i;
    int dim;
    if (c == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "data_t was not made by p3m_hip_init");
    TRY(check_params(c->method, s, p));
    TRY(check_host_system(s));
    if (f == NULL || f->f_k == NULL || (has_self_force(c->method) && f->f == NULL))
        return p3mh_fail(P3M_HIP_EINVAL, "forces must not be NULL");
    if (s->nparticles != c->n)
        return p3mh_fail(P3M_HIP_EINVAL, "particle count %d differs from Init's %d", s->nparticles, c->n);
    if (s->length != c->geom.L)
        return p3mh_fail(P3M_HIP_EINVAL, "box length differs from the influence function's");
    n = (size_t)s->nparticles;
    if (n == 0)
        return P3M_HIP_OK;
    TRY(upload_positions(c, s));
    TRYK(p3mk_fill_zero(c->force, sizeof(double) * 6 * n, c->q), "zero");
    TRY(kspace_enqueue(c, (int)n, c->pos, c->pos + n, c->pos + 2 * n, c->pos + 3 * n, c->force,
                       c->force + 3 * n));
    TRYK(p3mk_copy_d2h(c->hforce, c->force, sizeof(double) * 6 * n, c->q), "download");
    TRYK(p3mk_queue_sync(c->q), "sync");
    TRY(collect_stage_times(c));
    for (dim = 0; dim < 3; dim++) {
        double *fk = f->f_k->fields ? f->f_k->fields[dim] : (dim == 0 ? f->f_k->x : dim == 1 ? f->f_k->y : f->f_k->z);
        const double *src = c->hforce + dim * n;
        for (i = 0; i < n; i++)
            fk[i] += src[i];
    }
    if (has_self_force(c->method)) {
        for (dim = 0; dim < 3; dim++) {
            double *ft = f->f->fields ? f->f->fields[dim] : (dim == 0 ? f->f->x : dim == 1 ? f->f->y : f->f->z);
            const double *src = c->hforce + (3 + dim) * n;
            for (i = 0; i < n; i++)
                ft[i] += src[i];
        }
    }
    if (p->tuning) {
        /* p3m-common.h:48-57: seconds of assignment, FFT + multiply,
         * back-interpolation */
        d->runtime.t_c = 1e-3 * c->stage_ms[0];
        d->runtime.t_g = 1e-3 * c->stage_ms[1];
        d->runtime.t_f = 1e-3 * c->stage_ms[2];
        d->runtime.t = 1e-3 * c->stage_ms[3];
    }
    return P3M_HIP_OK;
}

int p3m_hip_stage_times(data_t *d, double out[5]) {
    p3m_ctx *c = ctx_of(d);
    int k;
    if (c == NULL || out == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "bad arguments");
    TRYK(p3mk_queue_sync(c->q), "sync");
    TRY(collect_stage_times(c));
    for (k = 0; k < 5; k++)
        out[k] = c->stage_ms[k];
    return P3M_HIP_OK;
}

int p3m_hip_get_influence_function(data_t *d, double *G) {
    p3m_ctx *c = ctx_of(d);
    size_t M3;
    if (c == NULL || G == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "bad arguments");
    M3 = (size_t)c->geom.M * c->geom.M * c->geom.M;
    TRYK(p3mk_copy_d2h(G, c->G, sizeof(double) * M3, c->q), "download");
    TRYK(p3mk_queue_sync(c->q), "sync");
    return P3M_HIP_OK;
}

int p3m_hip_get_self_force_coefficients(data_t *d, double *b) {
    p3m_ctx *c = ctx_of(d);
    if (c == NULL || b == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "bad arguments");
    TRYK(p3mk_copy_d2h(b, c->selfb, sizeof(double) * 81, c->q), "download");
    TRYK(p3mk_queue_sync(c->q), "sync");
    return P3M_HIP_OK;
}

/* ---------------------------------------------------------------------- */
/* Real space and Calculate_forces                                         */
/* ---------------------------------------------------------------------- */

static int ensure_cells(p3m_ctx *c, int n, double L, double rcut) {
    if (c->cells != NULL && c->cells_L == L && c->cells_rc == rcut)
        return P3M_HIP_OK;
    p3mk_cells_destroy(c->cells);
    c->cells = NULL;
    TRYK(p3mk_cells_create(&c->cells, n, L, rcut), "cell list");
    c->cells_L = L;
    c->cells_rc = rcut;
    return P3M_HIP_OK;
}

int p3m_hip_realteil_device(data_t *d, int n, double length, const double *x, const double *y,
                            const double *z, const double *q, double rcut, double alpha,
                            double *fr, double *energy) {
    p3m_ctx *c = ctx_of(d);
    if (c == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "data_t was not made by p3m_hip_init");
    if (n > c->n)
        return p3mh_fail(P3M_HIP_EINVAL, "more particles than Init's %d", c->n);
    if (!(rcut > 0.0) || rcut > 0.5 * length)
        return p3mh_fail(P3M_HIP_EINVAL, "real space needs 0 < rcut <= L/2");
    TRY(ensure_cells(c, c->n, length, rcut));
    TRYK(p3mk_record(c->q, 4), "event");
    TRYK(p3mk_realspace(c->cells, n, length, x, y, z, q, rcut, alpha, c->geom.fp32, fr, energy,
                        c->q),
         "real space");
    TRYK(p3mk_record(c->q, 5), "event");
    c->have_rs_events = 1;
    return P3M_HIP_OK;
}

int p3m_hip_calculate_forces_device(data_t *d, int n, double length, const double *x,
                                    const double *y, const double *z, const double *q, double rcut,
                                    double alpha, double *f, double *fk, double *fr,
                                    double *energy) {
    p3m_ctx *c = ctx_of(d);
    if (c == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "data_t was not made by p3m_hip_init");
    if (f == NULL || fk == NULL || fr == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "force arrays must not be NULL");
    if (n != c->n)
        return p3mh_fail(P3M_HIP_EINVAL, "particle count %d differs from Init's %d", n, c->n);
    /* common.c:274-278 */
    TRYK(p3mk_fill_zero(f, sizeof(double) * 3 * (size_t)n, c->q), "zero");
    TRYK(p3mk_fill_zero(fk, sizeof(double) * 3 * (size_t)n, c->q), "zero");
    TRYK(p3mk_fill_zero(fr, sizeof(double) * 3 * (size_t)n, c->q), "zero");
    if (energy != NULL)
        TRYK(p3mk_fill_zero(energy, sizeof(double), c->q), "zero");
    if (rcut != 0.0) /* common.c:290-294 */
        TRY(p3m_hip_realteil_device(d, n, length, x, y, z, q, rcut, alpha, fr, energy));
    TRY(p3m_hip_kspace_force_device(d, n, length, x, y, z, q, fk, f));
    /* common.c:315-319: f += f_k + f_r */
    TRYK(p3mk_sum_parts(3 * (size_t)n, fk, fr, f, c->q), "sum");
    if (c->ring_rs != NULL) {
        int k = p3mk_ring_filled(c->q);
        if (k < c->ring_sets)
            c->ring_rs[k] = rcut != 0.0;
        p3mk_ring_advance(c->q);
    }
    return P3M_HIP_OK;
}

int p3m_hip_stage_ring_start(data_t *d, int nsteps) {
    p3m_ctx *c = ctx_of(d);
    if (c == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "data_t was not made by p3m_hip_init");
    if (nsteps < 0)
        return p3mh_fail(P3M_HIP_EINVAL, "negative step count");
    TRYK(p3mk_ring_start(c->q, nsteps), "event ring");
    free(c->ring_rs);
    c->ring_rs = NULL;
    c->ring_sets = 0;
    if (nsteps > 0) {
        c->ring_rs = (unsigned char *)calloc((size_t)nsteps, 1);
        if (c->ring_rs == NULL) {
            p3mk_ring_start(c->q, 0);
            return p3mh_fail(P3M_HIP_ENOMEM, "host allocation failed");
        }
        c->ring_sets = nsteps;
    }
    return P3M_HIP_OK;
}

int p3m_hip_stage_ring_read(data_t *d, double *out, int max_steps, int *nsteps) {
    p3m_ctx *c = ctx_of(d);
    int k, s, filled;
    float t;
    if (c == NULL || out == NULL || nsteps == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "bad arguments");
    TRYK(p3mk_queue_sync(c->q), "sync");
    filled = p3mk_ring_filled(c->q);
    if (filled > max_steps)
        filled = max_steps;
    for (s = 0; s < filled; s++) {
        double *o = out + 5 * (size_t)s;
        for (k = 0; k < 3; k++) {
            TRYK(p3mk_ring_elapsed_ms(c->q, s, k, k + 1, &t), "event timing");
            o[k] = t;
        }
        TRYK(p3mk_ring_elapsed_ms(c->q, s, 0, 3, &t), "event timing");
        o[3] = t;
        o[4] = 0.0;
        if (c->ring_rs[s]) {
            TRYK(p3mk_ring_elapsed_ms(c->q, s, 4, 5, &t), "event timing");
            o[4] = t;
        }
    }
    *nsteps = filled;
    return P3M_HIP_OK;
}

int p3m_hip_stage_ring_read_pairs(data_t *d, double *out, int max_steps, int *nsteps) {
    p3m_ctx *c = ctx_of(d);
    int s, filled;
    float t;
    if (c == NULL || out == NULL || nsteps == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "bad arguments");
    TRYK(p3mk_queue_sync(c->q), "sync");
    filled = p3mk_ring_filled(c->q);
    if (filled > max_steps)
        filled = max_steps;
    for (s = 0; s < filled; s++) {
        out[s] = 0.0;
        if (c->ring_rs[s]) {
            TRYK(p3mk_ring_elapsed_ms(c->q, s, P3MK_EV_PAIRS, P3MK_EV_PAIRS + 1, &t), "event timing");
            out[s] = t;
        }
    }
    *nsteps = filled;
    return P3M_HIP_OK;
}

static void add_host(double *dst, const double *src, size_t n) {
    size_t i;
    for (i = 0; i < n; i++)
        dst[i] += src[i];
}

static double *field(vector_array_t *v, int dim) {
    if (v->fields)
        return v->fields[dim];
    return dim == 0 ? v->x : (dim == 1 ? v->y : v->z);
}

int p3m_hip_realteil(system_t *s, parameters_t *p, forces_t *f) {
    p3mk_queue *q = NULL;
    p3mk_cells *cells = NULL;
    double *pos = NULL, *fr = NULL, *en = NULL, *h = NULL;
    size_t n;
    int rc = P3M_HIP_OK, dim;
    double energy = 0.0;
    if (s == NULL || p == NULL || f == NULL || f->f_r == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "bad arguments");
    TRY(check_host_system(s));
    n = (size_t)s->nparticles;
    if (n == 0)
        return P3M_HIP_OK;
    if (!(p->rcut > 0.0) || p->rcut > 0.5 * s->length)
        return p3mh_fail(P3M_HIP_EINVAL, "real space needs 0 < rcut <= L/2");
    if (p3mk_queue_create(&q) != P3MK_OK)
        return p3mh_kfail(P3M_HIP_ENODEVICE, "stream");
#define CHK(expr, what)                     \
    if (rc == P3M_HIP_OK) {                 \
        int r_ = (expr);                    \
        if (r_ != P3MK_OK)                  \
            rc = p3mh_kfail(r_, what);           \
    }
    CHK(p3mk_cells_create(&cells, (int)n, s->length, p->rcut), "cell list");
    CHK(p3mk_alloc((void **)&pos, sizeof(double) * 4 * n), "positions");
    CHK(p3mk_alloc((void **)&fr, sizeof(double) * 3 * n), "forces");
    CHK(p3mk_alloc((void **)&en, sizeof(double)), "energy");
    CHK(p3mk_alloc_host((void **)&h, sizeof(double) * 3 * n), "staging");
    CHK(p3mk_copy_h2d(pos, s->p->x, sizeof(double) * n, q), "upload");
    CHK(p3mk_copy_h2d(pos + n, s->p->y, sizeof(double) * n, q), "upload");
    CHK(p3mk_copy_h2d(pos + 2 * n, s->p->z, sizeof(double) * n, q), "upload");
    CHK(p3mk_copy_h2d(pos + 3 * n, s->q, sizeof(double) * n, q), "upload");
    CHK(p3mk_fill_zero(fr, sizeof(double) * 3 * n, q), "zero");
    CHK(p3mk_fill_zero(en, sizeof(double), q), "zero");
    CHK(p3mk_realspace(cells, (int)n, s->length, pos, pos + n, pos + 2 * n, pos + 3 * n, p->rcut,
                       p->alpha, 0, fr, en, q),
        "real space");
    CHK(p3mk_copy_d2h(h, fr, sizeof(double) * 3 * n, q), "download");
    CHK(p3mk_copy_d2h(&energy, en, sizeof(double), q), "download");
    CHK(p3mk_queue_sync(q), "sync");
#undef CHK
    if (rc == P3M_HIP_OK) {
        for (dim = 0; dim < 3; dim++)
            add_host(field(f->f_r, dim), h + dim * n, n);
        s->energy += energy;
    }
    p3mk_release(pos);
    p3mk_release(fr);
    p3mk_release(en);
    p3mk_release_host(h);
    p3mk_cells_destroy(cells);
    p3mk_queue_destroy(q);
    return rc;
}

int p3m_hip_realteil_vjp_device(data_t *d, int n, double length, const double *x, const double *y,
                                const double *z, const double *q, double rcut, double alpha,
                                const double *gF, double *dpos, double *dq) {
    p3m_ctx *c = ctx_of(d);
    if (c == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "data_t was not made by p3m_hip_init");
    if (gF == NULL || dpos == NULL || dq == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "cotangent and result arrays must not be NULL");
    if (n > c->n)
        return p3mh_fail(P3M_HIP_EINVAL, "more particles than Init's %d", c->n);
    if (!(rcut > 0.0) || rcut > 0.5 * length)
        return p3mh_fail(P3M_HIP_EINVAL, "real space needs 0 < rcut <= L/2");
    TRY(ensure_cells(c, c->n, length, rcut));
    TRYK(p3mk_realspace_vjp(c->cells, n, length, x, y, z, q, rcut, alpha, gF, dpos, dq, c->q),
         "real-space adjoint");
    return P3M_HIP_OK;
}

int p3m_hip_realteil_vjp(system_t *s, parameters_t *p, const vector_array_t *dF,
                         vector_array_t *dpos, double *dq) {
    p3mk_queue *q = NULL;
    p3mk_cells *cells = NULL;
    double *pos = NULL, *g = NULL, *out = NULL, *h = NULL;
    size_t n;
    int rc = P3M_HIP_OK, dim;
    if (s == NULL || p == NULL || dF == NULL || dpos == NULL || dq == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "bad arguments");
    TRY(check_host_system(s));
    n = (size_t)s->nparticles;
    if (n == 0)
        return P3M_HIP_OK;
    if (!(p->rcut > 0.0) || p->rcut > 0.5 * s->length)
        return p3mh_fail(P3M_HIP_EINVAL, "real space needs 0 < rcut <= L/2");
    if (p3mk_queue_create(&q) != P3MK_OK)
        return p3mh_kfail(P3M_HIP_ENODEVICE, "stream");
#define CHK(expr, what)                     \
    if (rc == P3M_HIP_OK) {                 \
        int r_ = (expr);                    \
        if (r_ != P3MK_OK)                  \
            rc = p3mh_kfail(r_, what);      \
    }
    CHK(p3mk_cells_create(&cells, (int)n, s->length, p->rcut), "cell list");
    CHK(p3mk_alloc((void **)&pos, sizeof(double) * 4 * n), "positions");
    CHK(p3mk_alloc((void **)&g, sizeof(double) * 3 * n), "cotangent");
    CHK(p3mk_alloc((void **)&out, sizeof(double) * 4 * n), "adjoint");
    CHK(p3mk_alloc_host((void **)&h, sizeof(double) * 4 * n), "staging");
    CHK(p3mk_copy_h2d(pos, s->p->x, sizeof(double) * n, q), "upload");
    CHK(p3mk_copy_h2d(pos + n, s->p->y, sizeof(double) * n, q), "upload");
    CHK(p3mk_copy_h2d(pos + 2 * n, s->p->z, sizeof(double) * n, q), "upload");
    CHK(p3mk_copy_h2d(pos + 3 * n, s->q, sizeof(double) * n, q), "upload");
    for (dim = 0; dim < 3; dim++)
        CHK(p3mk_copy_h2d(g + dim * n, field((vector_array_t *)dF, dim), sizeof(double) * n, q),
            "upload");
    CHK(p3mk_fill_zero(out, sizeof(double) * 4 * n, q), "zero");
    CHK(p3mk_realspace_vjp(cells, (int)n, s->length, pos, pos + n, pos + 2 * n, pos + 3 * n,
                           p->rcut, p->alpha, g, out, out + 3 * n, q),
        "real-space adjoint");
    CHK(p3mk_copy_d2h(h, out, sizeof(double) * 4 * n, q), "download");
    CHK(p3mk_queue_sync(q), "sync");
#undef CHK
    if (rc == P3M_HIP_OK) {
        for (dim = 0; dim < 3; dim++)
            add_host(field(dpos, dim), h + dim * n, n);
        add_host(dq, h + 3 * n, n);
    }
    p3mk_release(pos);
    p3mk_release(g);
    p3mk_release(out);
    p3mk_release_host(h);
    p3mk_cells_destroy(cells);
    p3mk_queue_destroy(q);
    return rc;
}

int p3m_hip_calculate_forces(const method_t *m, system_t *s, parameters_t *p, data_t *d,
                             forces_t *f) {
    p3m_ctx *c = ctx_of(d);
    size_t n, i;
    int dim;
    double energy = 0.0;
    if (c == NULL || m == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "bad arguments");
    if (m->method_id != c->method)
        return p3mh_fail(P3M_HIP_EINVAL, "data_t belongs to method %d, not %d", c->method, m->method_id);
    TRY(check_params(c->method, s, p));
    TRY(check_host_system(s));
    if (f == NULL || f->f == NULL || f->f_k == NULL || f->f_r == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "forces must not be NULL");
    n = (size_t)s->nparticles;
    /* common.c:274-278: all three force arrays start from zero */
    for (dim = 0; dim < 3; dim++) {
        memset(field(f->f, dim), 0, sizeof(double) * n);
        memset(field(f->f_k, dim), 0, sizeof(double) * n);
        memset(field(f->f_r, dim), 0, sizeof(double) * n);
    }
    if (n == 0)
        return P3M_HIP_OK;
    if (s->nparticles != c->n)
        return p3mh_fail(P3M_HIP_EINVAL, "particle count %d differs from Init's %d", s->nparticles, c->n);
    if (s->length != c->geom.L)
        return p3mh_fail(P3M_HIP_EINVAL, "box length differs from the influence function's");
    TRY(upload_positions(c, s));
    TRYK(p3mk_fill_zero(c->force, sizeof(double) * 9 * n, c->q), "zero");
    TRYK(p3mk_fill_zero(c->energy, sizeof(double), c->q), "zero");
    if (p->rcut != 0.0) /* common.c:275-279 */
        TRY(p3m_hip_realteil_device(d, (int)n, s->length, c->pos, c->pos + n, c->pos + 2 * n,
                                    c->pos + 3 * n, p->rcut, p->alpha, c->force + 6 * n, c->energy));
    TRY(kspace_enqueue(c, (int)n, c->pos, c->pos + n, c->pos + 2 * n, c->pos + 3 * n, c->force,
                       c->force + 3 * n));
    TRYK(p3mk_copy_d2h(c->hforce, c->force, sizeof(double) * 9 * n, c->q), "download");
    TRYK(p3mk_copy_d2h(&energy, c->energy, sizeof(double), c->q), "download");
    TRYK(p3mk_queue_sync(c->q), "sync");
    for (dim = 0; dim < 3; dim++) {
        double *fk = field(f->f_k, dim), *fr = field(f->f_r, dim), *ft = field(f->f, dim);
        const double *hk = c->hforce + dim * n, *hs = c->hforce + (3 + dim) * n,
                     *hr = c->hforce + (6 + dim) * n;
        for (i = 0; i < n; i++) {
            fk[i] = hk[i];
            fr[i] = hr[i];
            ft[i] = hs[i] + (fk[i] + fr[i]); /* common.c:315-319 */
        }
    }
    s->energy += energy;
    return P3M_HIP_OK;
}

/* ---------------------------------------------------------------------- */
/* Error estimates                                                         */
/* ---------------------------------------------------------------------- */

#define MAX_TABLES 256
static struct {
    int kind, mesh, cao;
    double v[101];
} g_tables[MAX_TABLES];
static int g_ntables;

int p3m_hip_set_error_table(int kind, int mesh, int cao, const double *table101) {
    int k;
    if ((kind != 0 && kind != 1) || table101 == NULL || cao < 1 || cao > 7 || mesh < 2)
        return p3mh_fail(P3M_HIP_EINVAL, "bad error-table arguments");
    for (k = 0; k < g_ntables; k++)
        if (g_tables[k].kind == kind && g_tables[k].mesh == mesh && g_tables[k].cao == cao)
            break;
    if (k == g_ntables) {
        if (g_ntables == MAX_TABLES)
            return p3mh_fail(P3M_HIP_ENOMEM, "too many error tables");
        g_ntables++;
    }
    g_tables[k].kind = kind;
    g_tables[k].mesh = mesh;
    g_tables[k].cao = cao;
    memcpy(g_tables[k].v, table101, sizeof(double) * 101);
    return P3M_HIP_OK;
}

static int compare_ints(const void *a, const void *b) {
    const int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* find_error.c:84-107: -1 when (mesh, cao) is not tabulated or alpha L > 100;
 * otherwise the linear interpolation of the table row on the grid alpha L =
 * 0, 1, ..., 100 (find_error.c:38-58).  Rows registered with
 * p3m_hip_set_error_table take precedence over the built-in tables (the
 * reference's Q_ik / Q_ik_i for its 83 meshes, q_tables.c).  At alpha L =
 * 100 the reference reads one entry past the row; the last entry is used
 * here. */
static double table_lookup(int kind, int mesh, int cao, double alphaL) {
    const double *row = NULL;
    const int *mp;
    int k, l;
    double dd;
    if (alphaL > 100.0 || cao < 1 || cao > 7)
        return -1.0;
    for (k = 0; k < g_ntables; k++)
        if (g_tables[k].kind == kind && g_tables[k].mesh == mesh && g_tables[k].cao == cao) {
            row = g_tables[k].v;
            break;
        }
    if (row == NULL) {
        mp = bsearch(&mesh, p3m_q_meshes, P3M_Q_NMESHES, sizeof(int), compare_ints);
        if (mp == NULL)
            return -1.0;
        row = (kind == 1 ? p3m_q_ik_i : p3m_q_ik)[mp - p3m_q_meshes][cao - 1];
    }
    dd = alphaL;
    l = (int)floor(dd);
    dd -= l;
    if (l >= 100)
        return row[100];
    return (1. - dd) * row[l] + dd * row[l + 1];
}

static int kspace_error(int id, const system_t *s, const parameters_t *p, double *out) {
    p3mk_queue *q = NULL;
    double *scratch = NULL, he_q = 0.0;
    int rc = P3M_HIP_OK, kr;
    const int M = p->mesh;
    const double L = s->length, N = (double)s->nparticles;
    if (id == 0 || id == 6 || id == 1) {
        /* p3m-ik.c:263-301, p3m-ik-i.c:295-320 */
        const double t = table_lookup(id == 1 ? 1 : 0, M, p->cao, p->alpha * L);
        if (t >= 0.0) {
            *out = 2.0 * s->q2 * sqrt(t / N) / (L * L);
            return P3M_HIP_OK;
        }
        he_q = (id == 1) ? 0.0 : t; /* ik keeps the lookup's -1 (p3m-ik.c:267-297) */
    }
    if (p3mk_queue_create(&q) != P3MK_OK)
        return p3mh_kfail(P3M_HIP_ENODEVICE, "stream");
    kr = p3mk_alloc((void **)&scratch, sizeof(double) * p3mk_reduce_scratch_doubles(M));
    if (kr == P3MK_OK) {
        double sum = 0.0;
        kr = p3mk_kspace_error_sum(id, M, p->cao, L, p->alpha, 1, &sum, scratch, q);
        he_q += sum;
    }
    if (kr != P3MK_OK)
        rc = p3mh_kfail(kr, "error estimate");
    p3mk_release(scratch);
    p3mk_queue_destroy(q);
    if (rc != P3M_HIP_OK)
        return rc;
    if (id != 2 && id != 7) /* ad has no fabs (p3m-ad.c:496) */
        he_q = fabs(he_q);
    *out = 2.0 * s->q2 * sqrt(he_q / N) / (L * L);
    return P3M_HIP_OK;
}

/* realpart.c:343-346 */
static double realspace_error(const system_t *s, const parameters_t *p) {
    return (2.0 * s->q2 * exp(-(p->rcut * p->alpha) * (p->rcut * p->alpha))) /
           sqrt((double)s->nparticles * p->rcut * s->length * s->length * s->length);
}

int p3m_hip_error(int method_id, system_t *s, parameters_t *p, double *total, double *kspace) {
    double k = 0.0, r;
    TRY(check_params(method_id, s, p));
    TRY(kspace_error(method_id, s, p, &k));
    r = realspace_error(s, p);
    if (kspace)
        *kspace = k;
    if (total)
        *total = sqrt(r * r + k * k);
    return P3M_HIP_OK;
}

/* ---------------------------------------------------------------------- */
/* Adjoint of Kspace_force                                                 */
/* ---------------------------------------------------------------------- */

static int ensure_adjoint_buffers(p3m_ctx *c) {
    p3mk_geom g64 = c->geom;
    int k, nb = is_ad(c->method) ? 1 : 3;
    g64.fp32 = 0;
    for (k = 0; k < nb; k++)
        if (c->adj[k] == NULL)
            TRYK(p3mk_alloc(&c->adj[k], p3mk_mesh_bytes(&g64)), "adjoint mesh");
    return P3M_HIP_OK;
}

/* The field meshes are recomputed from the positions (no forward state is
 * kept), then the cotangent is assigned, transformed with the transposed
 * k-space operator and gathered together with the fields. */
int p3m_hip_kspace_force_vjp_device(data_t *d, int n, double length, const double *x,
                                    const double *y, const double *z, const double *q,
                                    const double *gF, double *dpos, double *dq) {
    p3m_ctx *c = ctx_of(d);
    p3mk_fft *fft;
    const p3mk_geom *g;
    const double L = length;
    void *in[3];
    int k;
    if (c == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "data_t was not made by p3m_hip_init");
    if (n != c->n)
        return p3mh_fail(P3M_HIP_EINVAL, "particle count %d differs from Init's %d", n, c->n);
    if (length != c->geom.L)
        return p3mh_fail(P3M_HIP_EINVAL, "box length differs from the influence function's");
    if (!c->have_G)
        return p3mh_fail(P3M_HIP_EINVAL, "influence function not computed");
    if (gF == NULL || dpos == NULL || dq == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "cotangent and gradient arrays must not be NULL");
    if (n == 0)
        return P3M_HIP_OK;
    TRY(ensure_adjoint_buffers(c));
    fft = ctx_fft(c);
    g = &c->geom;
    TRYK(p3mk_record(c->q, 0), "event");
    TRY(field_enqueue(c, n, x, y, z, q));
    if (!is_ad(c->method)) {
        /* psi_d = sum_i q_i g_{i,d} W_i; lambda = sum_d K_d^T psi_d */
        for (k = 0; k < 3; k++) {
            in[k] = is_r2c(c->method) ? c->spec[k] : c->adj[k];
            TRYK(p3mk_assign_weighted(g, n, c->psort, 1, gF + (size_t)k * n, c->adj[k], c->q),
                 "cotangent assignment");
            TRYK(p3mk_fft_forward_zy(fft, c->adj[k], in[k], c->q), "forward FFT");
        }
        TRYK(p3mk_divergence(fft, c->G, c->Dn, in, in[0], c->q), "adjoint multiply");
        TRYK(p3mk_fft_backward_yz(fft, in[0], c->adj[0], c->q), "backward FFT");
        TRYK(p3mk_adjoint_gather(g, 0, n, c->psort, q, gF, c->F, c->adj[0], interp_prefactor(c),
                                 NULL, dpos, dq, c->q),
             "adjoint gather");
    } else {
        /* psi = sum_i q_i sum_d g_{i,d} W^(d)_i; lambda = K psi */
        void *F[3];
        void *outs[3];
        in[0] = is_r2c(c->method) ? c->spec[0] : c->adj[0];
        outs[0] = in[0];
        outs[1] = outs[2] = NULL;
        TRYK(p3mk_assign_weighted(g, n, c->psort, 2, gF, c->adj[0], c->q), "cotangent assignment");
        TRYK(p3mk_fft_forward_zy(fft, c->adj[0], in[0], c->q), "forward FFT");
        TRYK(p3mk_convolve(fft, 1, c->G, c->Dn, in[0], outs, c->q), "adjoint multiply");
        TRYK(p3mk_fft_backward_yz(fft, in[0], c->adj[0], c->q), "backward FFT");
        F[0] = c->mesh;
        F[1] = F[2] = NULL;
        /* the ad derivative weights carry 1/L (p3m-ad.c:202 with
         * charge-assign.c:1123-1125) */
        TRYK(p3mk_adjoint_gather(g, 1, n, c->psort, q, gF, F, c->adj[0], interp_prefactor(c) / L,
                                 has_self_force(c->method) ? c->selfb : NULL, dpos, dq, c->q),
             "adjoint gather");
    }
    TRYK(p3mk_record(c->q, 3), "event");
    c->have_ks_events = 1;
    return P3M_HIP_OK;
}

int p3m_hip_kspace_force_vjp(system_t *s, parameters_t *p, data_t *d, const vector_array_t *dF,
                             vector_array_t *dpos, double *dq) {
    p3m_ctx *c = ctx_of(d);
    size_t n, b;
    int dim;
    if (c == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "data_t was not made by p3m_hip_init");
    TRY(check_params(c->method, s, p));
    TRY(check_host_system(s));
    if (dF == NULL || dpos == NULL || dq == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "cotangent and gradient arrays must not be NULL");
    if (s->nparticles != c->n)
        return p3mh_fail(P3M_HIP_EINVAL, "particle count %d differs from Init's %d", s->nparticles, c->n);
    n = (size_t)s->nparticles;
    if (n == 0)
        return P3M_HIP_OK;
    b = sizeof(double) * n;
    TRY(upload_positions(c, s));
    for (dim = 0; dim < 3; dim++)
        TRYK(p3mk_copy_h2d(c->force + dim * n, field((vector_array_t *)dF, dim), b, c->q), "upload");
    TRYK(p3mk_fill_zero(c->force + 3 * n, sizeof(double) * 4 * n, c->q), "zero");
    TRY(p3m_hip_kspace_force_vjp_device(d, (int)n, s->length, c->pos, c->pos + n, c->pos + 2 * n,
                                        c->pos + 3 * n, c->force, c->force + 3 * n,
                                        c->force + 6 * n));
    TRYK(p3mk_copy_d2h(c->hforce, c->force + 3 * n, sizeof(double) * 4 * n, c->q), "download");
    TRYK(p3mk_queue_sync(c->q), "sync");
    for (dim = 0; dim < 3; dim++)
        add_host(field(dpos, dim), c->hforce + dim * n, n);
    add_host(dq, c->hforce + 3 * n, n);
    return P3M_HIP_OK;
}

/* ---------------------------------------------------------------------- */
/* The method table                                                        */
/* ---------------------------------------------------------------------- */

/* The method-table entry points have no status channel (types.h:306-316):
 * a failure is reported and ends the program, as a failed allocation in the
 * reference's Init does, instead of returning wrong forces. */
static void report(const char *where) {
    fprintf(stderr, "p3m_hip: %s: %s\n", where, g_err);
    fflush(stderr);
    exit(EXIT_FAILURE);
}

#define DEFINE_METHOD(NAME, ID)                                                                \
    static data_t *init_##NAME(system_t *s, parameters_t *p) {                                 \
        data_t *d = NULL;                                                                      \
        if (p3m_hip_init(ID, s, p, P3M_HIP_FP64, &d) != P3M_HIP_OK)                            \
            report("Init");                                                                    \
        return d;                                                                              \
    }                                                                                          \
    static void influence_##NAME(system_t *s, parameters_t *p, data_t *d) {                    \
        if (p3m_hip_influence_function(s, p, d) != P3M_HIP_OK)                                 \
            report("Influence_function");                                                      \
    }                                                                                          \
    static void kspace_##NAME(system_t *s, parameters_t *p, data_t *d, forces_t *f) {          \
        if (p3m_hip_kspace_force(s, p, d, f) != P3M_HIP_OK)                                    \
            report("Kspace_force");                                                            \
    }                                                                                          \
    static double error_##NAME(system_t *s, parameters_t *p) {                                 \
        double t = -1.0;                                                                       \
        if (p3m_hip_error(ID, s, p, &t, NULL) != P3M_HIP_OK)                                   \
            report("Error");                                                                   \
        return t;                                                                              \
    }                                                                                          \
    static double error_k_##NAME(system_t *s, parameters_t *p) {                               \
        double k = -1.0;                                                                       \
        if (p3m_hip_error(ID, s, p, NULL, &k) != P3M_HIP_OK)                                   \
            report("Error_k");                                                                 \
        return k;                                                                              \
    }

DEFINE_METHOD(ik, 0)
DEFINE_METHOD(ik_i, 1)
DEFINE_METHOD(ad, 2)
DEFINE_METHOD(ad_i, 3)
DEFINE_METHOD(ik_r, 6)
DEFINE_METHOD(ad_r, 7)

/* names, ids and flags as the reference declares them */
const method_t method_p3m_ik = {METHOD_P3M_ik, "P3M with ik differentiation, not intelaced.",
                                "p3m-ik", METHOD_FLAG_P3M | METHOD_FLAG_ik, &init_ik,
                                &influence_ik, &kspace_ik, &error_ik, &error_k_ik};
const method_t method_p3m_ik_i = {METHOD_P3M_ik_i, "P3M with ik differentiation, interlaced.",
                                  "p3m-ik-i", METHOD_FLAG_P3M | METHOD_FLAG_ik | METHOD_FLAG_interlaced,
                                  &init_ik_i, &influence_ik_i, &kspace_ik_i, &error_ik_i,
                                  &error_k_ik_i};
const method_t method_p3m_ad = {METHOD_P3M_ad, "P3M with analytic differentiation, not intelaced.",
                                "p3m-ad",
                                (char)(METHOD_FLAG_P3M | METHOD_FLAG_ad | METHOD_FLAG_self_force_correction),
                                &init_ad, &influence_ad, &kspace_ad, &error_ad, &error_k_ad};
const method_t method_p3m_ad_i = {METHOD_P3M_ad_i, "P3M with analytic differentiation, intelaced.",
                                  "p3m-ad-i", METHOD_FLAG_P3M | METHOD_FLAG_ad | METHOD_FLAG_interlaced,
                                  &init_ad_i, &influence_ad_i, &kspace_ad_i, &error_ad_i,
                                  &error_k_ad_i};
const method_t method_p3m_ik_r = {METHOD_P3M_ik_r,
                                  "P3M with ik differentiation, not intelaced, real input.",
                                  "p3m-ik-r", METHOD_FLAG_P3M | METHOD_FLAG_ik, &init_ik_r,
                                  &influence_ik_r, &kspace_ik_r, &error_ik_r, &error_k_ik_r};
const method_t method_p3m_ad_r = {METHOD_P3M_ad_r,
                                  "P3M with analytic differentiation, not intelaced, real input.",
                                  "p3m-ad-r",
                                  (char)(METHOD_FLAG_P3M | METHOD_FLAG_ad | METHOD_FLAG_self_force_correction),
                                  &init_ad_r, &influence_ad_r, &kspace_ad_r, &error_ad_r,
                                  &error_k_ad_r};

const method_t *p3m_hip_method(int method_id) {
    switch (method_id) {
    case 0: return &method_p3m_ik;
    case 1: return &method_p3m_ik_i;
    case 2: return &method_p3m_ad;
    case 3: return &method_p3m_ad_i;
    case 6: return &method_p3m_ik_r;
    case 7: return &method_p3m_ad_r;
    }
    return NULL;
}
