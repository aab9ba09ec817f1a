/* Host side of batches of independent systems.
 *
 * All systems of a batch use one method id and one precision; each has its
 * own particle count, box, mesh, order, alpha and cutoff.  Particles are
 * concatenated system by system (system s holds slots [sum_{r<s} n_r, + n_s)
 * of x, y, z, q) and forces use the same layout in three blocks of the total
 * count.  Influence functions and self-force coefficients are made per
 * system with the single-system kernels; every per-call stage is one batched
 * launch over all systems (csrc/k_batch.hip).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "p3m_hip.h"

#include "../p3m_kernels.h"
#include "host_common.h"

struct p3m_hip_batch {
    int method, precision, nsys, ntot;
    p3m_hip_batch_system *sys;
    p3mk_queue *q;
    p3mk_batch *kb;
    double *scratch; /* self-force reduction scratch, largest mesh */
    p3mk_graph *graph;
};

/* alpha from the real-space error target (tuning.c:293) */
double p3m_hip_tune_alpha(int n, double length, double q2, double rcut, double precision) {
    const double V = length * length * length;
    return sqrt(-log((precision * sqrt((double)n * rcut * V)) / (2.0 * sqrt(2.0) * q2))) / rcut;
}

static int batch_influence(p3m_hip_batch *b, int s) {
    const p3m_hip_batch_system *S = &b->sys[s];
    p3mk_geom g;
    double *G, *Dn, *selfb;
    g.M = S->mesh;
    g.cao = S->cao;
    g.r2c = p3mh_is_r2c(b->method);
    g.interlaced = p3mh_is_interlaced(b->method);
    g.fp32 = b->precision == P3M_HIP_MIXED;
    g.L = S->length;
    TRYK(p3mk_batch_system_arrays(b->kb, s, &G, &Dn, &selfb), "batch arrays");
    TRYK(p3mk_influence(b->method, &g, S->alpha, 1, G, Dn, b->q), "influence function");
    if (p3mh_has_self_force(b->method)) {
        /* p3m-ad.c:120-123: no self-force coefficients at alpha == 0 */
        if (S->alpha != 0.0)
            TRYK(p3mk_self_force_coeffs(&g, G, selfb, b->scratch, b->q), "self force");
        else
            TRYK(p3mk_fill_zero(selfb, sizeof(double) * 81, b->q), "zero");
    }
    return P3M_HIP_OK;
}

int p3m_hip_batch_free(p3m_hip_batch *b) {
    if (b == NULL)
        return P3M_HIP_OK;
    if (b->q)
        p3mk_queue_sync(b->q);
    p3mk_graph_destroy(b->graph);
    p3mk_batch_destroy(b->kb);
    p3mk_release(b->scratch);
    p3mk_queue_destroy(b->q);
    free(b->sys);
    free(b);
    return P3M_HIP_OK;
}

int p3m_hip_batch_create(int method_id, int precision, int nsys, const p3m_hip_batch_system *sys,
                         p3m_hip_batch **out) {
    p3m_hip_batch *b;
    p3mk_bsys_in *in;
    int s, maxM = 2, rc;
    long ntot = 0;
    if (out == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "out must not be NULL");
    *out = NULL;
    if (!p3mh_known_method(method_id))
        return p3mh_fail(P3M_HIP_EINVAL, "method id %d is not a P3M method of this library", method_id);
    if (precision != P3M_HIP_FP64 && precision != P3M_HIP_MIXED)
        return p3mh_fail(P3M_HIP_EINVAL, "unknown precision %d", precision);
    if (nsys < 1 || sys == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "a batch needs at least one system");
    for (s = 0; s < nsys; s++) {
        const p3m_hip_batch_system *S = &sys[s];
        if (S->cao < 1 || S->cao > 7)
            return p3mh_fail(P3M_HIP_EINVAL, "system %d: cao %d not in 1..7", s, S->cao);
        if (S->mesh < 2 || S->mesh % 2 != 0 || S->mesh < S->cao)
            return p3mh_fail(P3M_HIP_EINVAL, "system %d: mesh %d must be even and >= cao", s, S->mesh);
        if (!(S->length > 0.0) || !isfinite(S->length) || S->nparticles < 0)
            return p3mh_fail(P3M_HIP_EINVAL, "system %d: bad box or particle count", s);
        if (S->rcut < 0.0 || S->rcut > 0.5 * S->length)
            return p3mh_fail(P3M_HIP_EINVAL, "system %d: real space needs 0 <= rcut <= L/2", s);
        ntot += S->nparticles;
        if (S->mesh > maxM)
            maxM = S->mesh;
    }
    if (ntot > 0x3fffffffL)
        return p3mh_fail(P3M_HIP_EINVAL, "batch too large");
    b = (p3m_hip_batch *)calloc(1, sizeof(*b));
    in = (p3mk_bsys_in *)calloc((size_t)nsys, sizeof(*in));
    if (b == NULL || in == NULL) {
        free(b);
        free(in);
        return p3mh_fail(P3M_HIP_ENOMEM, "host allocation failed");
    }
    b->method = method_id;
    b->precision = precision;
    b->nsys = nsys;
    b->ntot = (int)ntot;
    b->sys = (p3m_hip_batch_system *)malloc(sizeof(*sys) * (size_t)nsys);
    if (b->sys == NULL) {
        free(in);
        free(b);
        return p3mh_fail(P3M_HIP_ENOMEM, "host allocation failed");
    }
    memcpy(b->sys, sys, sizeof(*sys) * (size_t)nsys);
    for (s = 0; s < nsys; s++) {
        const double L = sys[s].length;
        in[s].n = sys[s].nparticles;
        in[s].L = L;
        in[s].M = sys[s].mesh;
        in[s].cao = sys[s].cao;
        in[s].rcut = sys[s].rcut;
        in[s].alpha = sys[s].alpha;
        /* back-interpolation prefactors: ik 1/(2 L^3) (p3m-ik.c:242),
         * ad M/L^3 (p3m-ad.c:202) */
        in[s].prefac = p3mh_is_ad(method_id) ? (double)sys[s].mesh / (L * L * L) : 1.0 / (2.0 * L * L * L);
    }
    rc = p3mk_queue_create(&b->q);
    if (rc == P3MK_OK)
        rc = p3mk_batch_create(&b->kb, p3mh_is_r2c(method_id), p3mh_is_interlaced(method_id),
                               p3mh_is_ad(method_id), precision == P3M_HIP_MIXED, nsys, in);
    if (rc == P3MK_OK)
        rc = p3mk_batch_configure(b->kb, precision == P3M_HIP_MIXED,
                                  p3mh_has_self_force(method_id), 8, 0);
    if (rc == P3MK_OK)
        rc = p3mk_alloc((void **)&b->scratch, sizeof(double) * p3mk_reduce_scratch_doubles(maxM));
    free(in);
    if (rc != P3MK_OK) {
        int code = p3mh_kfail(rc == P3MK_EINVAL ? P3M_HIP_EINVAL
                              : rc == P3MK_ENOMEM ? P3M_HIP_ENOMEM
                              : rc == P3MK_ENODEVICE ? P3M_HIP_ENODEVICE : P3M_HIP_EDEVICE,
                              "batch");
        p3m_hip_batch_free(b);
        return code;
    }
    for (s = 0; s < nsys; s++) {
        rc = batch_influence(b, s);
        if (rc != P3M_HIP_OK) {
            p3m_hip_batch_free(b);
            return rc;
        }
    }
    rc = p3mk_queue_sync(b->q);
    if (rc != P3MK_OK) {
        p3m_hip_batch_free(b);
        return p3mh_kfail(P3M_HIP_EDEVICE, "sync");
    }
    *out = b;
    return P3M_HIP_OK;
}

int p3m_hip_batch_set_alpha(p3m_hip_batch *b, int s, double alpha) {
    if (b == NULL || s < 0 || s >= b->nsys)
        return p3mh_fail(P3M_HIP_EINVAL, "bad batch or system index");
    b->sys[s].alpha = alpha;
    TRYK(p3mk_queue_sync(b->q), "sync");
    TRYK(p3mk_batch_set_alpha(b->kb, s, alpha), "batch alpha");
    TRY(batch_influence(b, s));
    TRYK(p3mk_queue_sync(b->q), "sync");
    return P3M_HIP_OK;
}

int p3m_hip_batch_kspace_force_device(p3m_hip_batch *b, const double *x, const double *y,
                                      const double *z, const double *q, double *fk, double *f) {
    if (b == NULL || x == NULL || y == NULL || z == NULL || q == NULL || fk == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "bad arguments");
    if (p3mh_has_self_force(b->method) && f == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "method %d needs the total-force array f", b->method);
    TRYK(p3mk_batch_kspace(b->kb, x, y, z, q, fk, f, b->q), "batched k-space force");
    return P3M_HIP_OK;
}

static int batch_forces_enqueue(p3m_hip_batch *b, const double *x, const double *y, const double *z,
                                const double *q, double *f, double *fk, double *fr,
                                double *energy) {
    const size_t n3 = 3 * (size_t)b->ntot;
    int s, any_rs = 0;
    /* common.c:274-278 */
    TRYK(p3mk_fill_zero(f, sizeof(double) * n3, b->q), "zero");
    TRYK(p3mk_fill_zero(fk, sizeof(double) * n3, b->q), "zero");
    TRYK(p3mk_fill_zero(fr, sizeof(double) * n3, b->q), "zero");
    if (energy != NULL)
        TRYK(p3mk_fill_zero(energy, sizeof(double) * (size_t)b->nsys, b->q), "zero");
    for (s = 0; s < b->nsys; s++)
        any_rs |= b->sys[s].rcut > 0.0;
    if (any_rs)
        TRYK(p3mk_batch_realspace(b->kb, x, y, z, q, fr, energy, b->q), "batched real space");
    TRYK(p3mk_batch_kspace(b->kb, x, y, z, q, fk, f, b->q), "batched k-space force");
    /* common.c:315-319 */
    TRYK(p3mk_sum_parts(n3, fk, fr, f, b->q), "sum");
    return P3M_HIP_OK;
}

int p3m_hip_batch_calculate_forces_device(p3m_hip_batch *b, const double *x, const double *y,
                                          const double *z, const double *q, double *f, double *fk,
                                          double *fr, double *energy) {
    if (b == NULL || x == NULL || y == NULL || z == NULL || q == NULL || f == NULL || fk == NULL ||
        fr == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "bad arguments");
    return batch_forces_enqueue(b, x, y, z, q, f, fk, fr, energy);
}

int p3m_hip_batch_capture(p3m_hip_batch *b, const double *x, const double *y, const double *z,
                          const double *q, double *f, double *fk, double *fr, double *energy) {
    int rc;
    if (b == NULL || x == NULL || y == NULL || z == NULL || q == NULL || f == NULL || fk == NULL ||
        fr == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "bad arguments");
    TRYK(p3mk_queue_sync(b->q), "sync");
    p3mk_graph_destroy(b->graph);
    b->graph = NULL;
    TRYK(p3mk_capture_begin(b->q), "graph capture");
    rc = batch_forces_enqueue(b, x, y, z, q, f, fk, fr, energy);
    {
        p3mk_graph *g = NULL;
        int rc2 = p3mk_capture_end(b->q, &g);
        if (rc != P3M_HIP_OK) {
            p3mk_graph_destroy(g);
            return rc;
        }
        if (rc2 != P3MK_OK)
            return p3mh_kfail(P3M_HIP_EDEVICE, "graph capture");
        b->graph = g;
    }
    return P3M_HIP_OK;
}

int p3m_hip_batch_replay(p3m_hip_batch *b) {
    if (b == NULL || b->graph == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "no captured graph");
    TRYK(p3mk_graph_launch(b->graph, b->q), "graph launch");
    return P3M_HIP_OK;
}

int p3m_hip_batch_synchronize(p3m_hip_batch *b) {
    if (b == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "bad batch");
    TRYK(p3mk_queue_sync(b->q), "sync");
    return P3M_HIP_OK;
}

/* time `reps` force evaluations (ms each, HIP events on the batch stream) */
static int time_forces(p3m_hip_batch *b, const double *x, const double *y, const double *z,
                       const double *q, double *f, double *fk, double *fr, int reps, double *ms) {
    float t = 0.0f;
    int k;
    TRY(batch_forces_enqueue(b, x, y, z, q, f, fk, fr, NULL)); /* warm-up */
    TRYK(p3mk_record(b->q, 0), "event");
    for (k = 0; k < reps; k++)
        TRY(batch_forces_enqueue(b, x, y, z, q, f, fk, fr, NULL));
    TRYK(p3mk_record(b->q, 1), "event");
    TRYK(p3mk_queue_sync(b->q), "sync");
    TRYK(p3mk_elapsed_ms(b->q, 0, 1, &t), "event timing");
    *ms = (double)t / reps;
    return P3M_HIP_OK;
}

int p3m_hip_batch_time_pairs(p3m_hip_batch *b, const double *x, const double *y, const double *z,
                             const double *q, double *f, double *fk, double *fr, int reps,
                             double *ms) {
    float t = 0.0f;
    double sum = 0.0;
    int k, rc = P3M_HIP_OK;
    if (b == NULL || x == NULL || y == NULL || z == NULL || q == NULL || f == NULL || fk == NULL ||
        fr == NULL || ms == NULL || reps < 1)
        return p3mh_fail(P3M_HIP_EINVAL, "bad arguments");
    TRYK(p3mk_queue_sync(b->q), "sync");
    TRYK(p3mk_batch_time_pairs(b->kb, 1), "events");
    for (k = 0; k < reps && rc == P3M_HIP_OK; k++) {
        int kr;
        rc = batch_forces_enqueue(b, x, y, z, q, f, fk, fr, NULL);
        if (rc == P3M_HIP_OK && (kr = p3mk_queue_sync(b->q)) != P3MK_OK)
            rc = p3mh_kfail(kr, "sync");
        if (rc == P3M_HIP_OK &&
            (kr = p3mk_elapsed_ms(b->q, P3MK_EV_PAIRS, P3MK_EV_PAIRS + 1, &t)) != P3MK_OK)
            rc = p3mh_kfail(kr, "event timing");
        sum += t;
    }
    (void)p3mk_batch_time_pairs(b->kb, 0);
    if (rc == P3M_HIP_OK)
        *ms = sum / reps;
    return rc;
}

/* Autotuner of the launch shapes: the assignment tile side (4 or 8 mesh
 * points: smaller tiles give more, lighter work items for small meshes and
 * dense clusters) and the size of the persistent real-space grid (8 to 64
 * waves per CU), every pair timed over whole force evaluations of the given
 * input (12 shapes). */
int p3m_hip_batch_autotune(p3m_hip_batch *b, const double *x, const double *y, const double *z,
                           const double *q, double *f, double *fk, double *fr, int *tile_side,
                           int *rs_grid) {
    static const int tiles[] = {8, 4};
    static const int waves_per_cu[] = {8, 16, 24, 32, 48, 64};
    const int ntiles = (int)(sizeof(tiles) / sizeof(tiles[0]));
    const int ngrids = (int)(sizeof(waves_per_cu) / sizeof(waves_per_cu[0]));
    int best_tile = 8, best_grid = 0, cus_grid = 0, k, j, fp32, sf;
    double best = 1e300, ms;
    if (b == NULL || x == NULL || y == NULL || z == NULL || q == NULL || f == NULL || fk == NULL ||
        fr == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "bad arguments");
    fp32 = b->precision == P3M_HIP_MIXED;
    sf = p3mh_has_self_force(b->method);
    TRYK(p3mk_queue_sync(b->q), "sync");
    p3mk_graph_destroy(b->graph); /* captured shapes become stale */
    b->graph = NULL;
    TRYK(p3mk_batch_info(b->kb, NULL, NULL, NULL, NULL, &cus_grid), "batch info");
    cus_grid /= 32; /* the default grid is 32 waves per CU */
    best_grid = 32 * cus_grid;
    /* every (tile side, grid) pair: the two interact through the claims,
     * whose size follows the grid */
    for (k = 0; k < ntiles; k++) {
        for (j = 0; j < ngrids; j++) {
            const int grid = waves_per_cu[j] * cus_grid;
            TRYK(p3mk_queue_sync(b->q), "sync");
            TRYK(p3mk_batch_configure(b->kb, fp32, sf, tiles[k], grid), "configure");
            TRY(time_forces(b, x, y, z, q, f, fk, fr, 3, &ms));
            if (ms < best) {
                best = ms;
                best_tile = tiles[k];
                best_grid = grid;
            }
        }
    }
    TRYK(p3mk_queue_sync(b->q), "sync");
    TRYK(p3mk_batch_configure(b->kb, fp32, sf, best_tile, best_grid), "configure");
    if (tile_side)
        *tile_side = best_tile;
    if (rs_grid)
        *rs_grid = best_grid;
    return P3M_HIP_OK;
}
