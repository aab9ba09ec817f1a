/*
 * Tuner: the parameter search of the reference's Tune (tuning.c:250-370)
 * with the GPU as the timer.
 *
 * Search order and pruning (kept as the reference has them, since they decide
 * which parameters come out):
 *   - alpha from the real-space error target precision / sqrt(2) at the fixed
 *     cutoff (tuning.c:293);
 *   - meshes from the candidate list (tuning.c:38) in ascending order; the
 *     first mesh starts at the highest order, later meshes at one below the
 *     last order that met the precision, and the scan ends once that order
 *     is the lowest allowed;
 *   - per mesh, the cheapest order is timed first, and if even it is slower
 *     than the best so far, no larger mesh can win and the scan ends;
 *   - orders descend while the estimate meets the precision; a candidate no
 *     smaller in mesh and order than the best is not timed;
 *   - the winner has the smallest sum of per-stage minimum times.
 * One deliberate difference: the reference's mesh loop runs one index past
 * its list (tuning.c:295-296 reads smooth_numbers[smooth_numbers_n]); here the
 * scan ends with the list, and the "no best yet" mesh is INT_MAX.
 *
 * Timing (time_full, tuning.c:92-147): a data_t made with p->tuning set,
 * 50 Kspace_force calls on device copies of the positions, per-stage
 * statistics from the library's HIP events (assignment, FFTs + multiply,
 * back-interpolation).
 */
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "host_common.h"
#include "p3m_hip.h"
#include "p3m_kernels.h"

enum { TUNE_CAO_MIN = 2, TUNE_CAO_MAX = 7, TUNE_SAMPLES = 50 }; /* tuning.h:24-27 */

/* tuning.c:38: even meshes up to 300 whose prime factors are 2, 3, 5, 7,
 * 11 and 13, ascending */
static const int tune_meshes[] = {
    8,   10,  12,  14,  16,  18,  20,  22,  24,  26,  28,  30,  32,  36,  40,  42,  44,  48,  50,
    52,  54,  56,  60,  64,  66,  70,  72,  78,  80,  84,  88,  90,  96,  98,  100, 104, 108, 110,
    112, 120, 126, 128, 130, 132, 140, 144, 150, 154, 156, 160, 162, 168, 176, 180, 182, 192, 196,
    198, 200, 208, 210, 216, 220, 224, 234, 240, 242, 250, 252, 256, 260, 264, 270, 280, 288, 294,
    300};
#define N_TUNE_MESHES ((int)(sizeof tune_meshes / sizeof tune_meshes[0]))

static void timing_begin(timing_t *t) {
    t->avg = 0.0;
    t->sgm = 0.0;
    t->min = 1e99;
    t->max = 0.0;
    t->n = TUNE_SAMPLES;
}

static void timing_add(timing_t *t, double v) {
    t->avg += v;
    if (v < t->min)
        t->min = v;
    if (v > t->max)
        t->max = v;
}

/* the device copies one tuning run times on */
typedef struct {
    int form, n;
    double L;
    double *x, *y, *z, *q, *fk, *f;
} gpu_timer;

static int gpu_time_full(void *vctx, int method_id, system_t *s, parameters_t *p,
                         runtime_stat_t *out) {
    gpu_timer *g = (gpu_timer *)vctx;
    parameters_t mp = *p;
    data_t *d = NULL;
    double ms[5];
    int i, rc;
    mp.tuning = 1;
    timing_begin(&out->t);
    timing_begin(&out->t_c);
    timing_begin(&out->t_g);
    timing_begin(&out->t_f);
    rc = p3m_hip_init(method_id, s, &mp, g->form, &d);
    if (rc != P3M_HIP_OK)
        return rc;
    for (i = 0; i < TUNE_SAMPLES && rc == P3M_HIP_OK; i++) {
        rc = p3m_hip_kspace_force_device(d, g->n, g->L, g->x, g->y, g->z, g->q, g->fk, g->f);
        if (rc == P3M_HIP_OK)
            rc = p3m_hip_stage_times(d, ms);
        if (rc == P3M_HIP_OK) {
            timing_add(&out->t_c, 1e-3 * ms[0]);
            timing_add(&out->t_g, 1e-3 * ms[1]);
            timing_add(&out->t_f, 1e-3 * ms[2]);
        }
    }
    p3m_hip_free(d);
    if (rc != P3M_HIP_OK)
        return rc;
    out->t_c.avg /= out->t_c.n;
    out->t_g.avg /= out->t_g.n;
    out->t_f.avg /= out->t_f.n;
    out->t.avg = out->t_c.avg + out->t_g.avg + out->t_f.avg;
    out->t.min = out->t_c.min + out->t_g.min + out->t_f.min;
    out->t.max = out->t_c.max + out->t_g.max + out->t_f.max;
    return P3M_HIP_OK;
}

static void gpu_timer_release(gpu_timer *g) {
    double *ptrs[] = {g->x, g->y, g->z, g->q, g->fk, g->f};
    size_t k;
    for (k = 0; k < sizeof ptrs / sizeof ptrs[0]; k++)
        p3mk_release(ptrs[k]);
    memset(g, 0, sizeof *g);
}

static const double *coord(const vector_array_t *v, int d) {
    if (v->fields)
        return v->fields[d];
    return d == 0 ? v->x : (d == 1 ? v->y : v->z);
}

/* device copies of the system's positions and charges, and force scratch */
static int gpu_timer_setup(gpu_timer *g, system_t *s, int form) {
    const size_t n = (size_t)s->nparticles, b = n * sizeof(double);
    p3mk_queue *qu = NULL;
    int rc;
    memset(g, 0, sizeof *g);
    g->form = form;
    g->n = s->nparticles;
    g->L = s->length;
    TRYK(p3mk_device_check(), "no usable GPU");
    rc = p3mk_alloc((void **)&g->x, b);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&g->y, b);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&g->z, b);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&g->q, b);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&g->fk, 3 * b);
    if (rc == P3MK_OK) rc = p3mk_alloc((void **)&g->f, 3 * b);
    if (rc == P3MK_OK) rc = p3mk_queue_create(&qu);
    if (rc == P3MK_OK) rc = p3mk_copy_h2d(g->x, coord(s->p, 0), b, qu);
    if (rc == P3MK_OK) rc = p3mk_copy_h2d(g->y, coord(s->p, 1), b, qu);
    if (rc == P3MK_OK) rc = p3mk_copy_h2d(g->z, coord(s->p, 2), b, qu);
    if (rc == P3MK_OK) rc = p3mk_copy_h2d(g->q, s->q, b, qu);
    if (rc == P3MK_OK) rc = p3mk_fill_zero(g->fk, 3 * b, qu);
    if (rc == P3MK_OK) rc = p3mk_fill_zero(g->f, 3 * b, qu);
    if (rc == P3MK_OK) rc = p3mk_queue_sync(qu);
    if (qu)
        p3mk_queue_destroy(qu);
    if (rc != P3MK_OK) {
        gpu_timer_release(g);
        return p3mh_kfail(rc, "tuner device buffers");
    }
    return P3M_HIP_OK;
}

static void set_order(parameters_t *p, int cao) {
    p->cao = cao;
    p->ip = cao - 1;
    p->cao3 = cao * cao * cao;
}

static int scan(int method_id, system_t *s, parameters_t *p, double precision,
                p3m_hip_tune_timer timer, void *ctx, runtime_stat_t *out) {
    parameters_t it, best;
    runtime_stat_t best_time, t;
    int k, cao, lowest, highest, last_ok, first, found = 0;
    double err;

    memset(&it, 0, sizeof it);
    memset(&best, 0, sizeof best);
    memset(&best_time, 0, sizeof best_time);
    memset(out, 0, sizeof *out);
    out->t.avg = -1.0;
    best_time.t.min = DBL_MAX;
    best.mesh = INT_MAX;
    best.cao = TUNE_CAO_MAX + 1;

    it.prefactor = 1.0;
    it.rcut = p->rcut;
    it.tuning = p->tuning;
    it.alpha = p3m_hip_tune_alpha(s->nparticles, s->length, s->q2, p->rcut, precision);
    if (p->cao != 0)
        lowest = highest = last_ok = p->cao;
    else {
        lowest = TUNE_CAO_MIN;
        highest = last_ok = TUNE_CAO_MAX;
    }

    for (k = 0; k < N_TUNE_MESHES; k++) {
        it.mesh = tune_meshes[k];
        if (found) {
            if (last_ok <= lowest)
                break;
            first = last_ok - 1;
        } else {
            first = last_ok = highest;
        }
        /* this mesh's cheapest order bounds its time from below */
        set_order(&it, lowest);
        TRY(timer(ctx, method_id, s, &it, &t));
        if (t.t.min > best_time.t.min)
            break;
        for (cao = first; cao >= lowest; cao--) {
            set_order(&it, cao);
            TRY(p3m_hip_error(method_id, s, &it, &err, NULL));
            if (!(err <= precision))
                break;
            last_ok = cao;
            found = 1;
            if (it.mesh >= best.mesh && it.cao >= best.cao)
                break;
            TRY(timer(ctx, method_id, s, &it, &t));
            if (t.t.min < best_time.t.min) {
                best = it;
                best.precision = err;
                best_time = t;
            }
        }
    }
    if (!found)
        return P3M_HIP_OK;
    *p = best;
    set_order(p, best.cao);
    *out = best_time;
    return P3M_HIP_OK;
}

int p3m_hip_tune_scan(int method_id, system_t *s, parameters_t *p, double precision, int form,
                      p3m_hip_tune_timer timer, void *ctx, runtime_stat_t *out) {
    gpu_timer g;
    int rc;
    if (s == NULL || p == NULL || out == NULL || !p3mh_known_method(method_id))
        return p3mh_fail(P3M_HIP_EINVAL, "tune: bad arguments");
    if (!(precision > 0.0) || !(p->rcut > 0.0) || !(s->q2 > 0.0) || s->nparticles <= 0)
        return p3mh_fail(P3M_HIP_EINVAL, "tune: needs precision > 0, rcut > 0 and charges");
    if (p->cao < 0 || p->cao > TUNE_CAO_MAX)
        return p3mh_fail(P3M_HIP_EINVAL, "tune: cao must be 0 (scan) or 1..7");
    if (form != P3M_HIP_FP64 && form != P3M_HIP_MIXED)
        return p3mh_fail(P3M_HIP_EINVAL, "tune: unknown form");
    if (timer != NULL)
        return scan(method_id, s, p, precision, timer, ctx, out);
    TRY(gpu_timer_setup(&g, s, form));
    rc = scan(method_id, s, p, precision, gpu_time_full, &g, out);
    gpu_timer_release(&g);
    return rc;
}

runtime_stat_t p3m_hip_tune(const method_t *m, system_t *s, parameters_t *p, double precision) {
    runtime_stat_t out;
    const int rc = p3m_hip_tune_scan(m ? m->method_id : -1, s, p, precision, P3M_HIP_FP64, NULL,
                                     NULL, &out);
    if (rc != P3M_HIP_OK) {
        fprintf(stderr, "p3m_hip_tune: %s\n", p3m_hip_last_error());
        exit(1);
    }
    return out;
}
