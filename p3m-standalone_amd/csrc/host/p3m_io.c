/*
 * On-disk formats of the reference's driver, for loading its systems and
 * force files and writing files it reads:
 *   system file   io.c:100-157 (read), io.c:199-225 (write):
 *                 "# Teilchenzahl: N", "# Len: L", then N lines "x y z q";
 *   force file    io.c:70-98 (read), io.c:159-177 (write): per particle
 *                 "i fx fy fz" (+ "fkx fky fkz" as written);
 *   out.dat row   main.c:452-455: alpha, DeltaF, estimate, real-space
 *                 estimate, k-space estimate, k-space DeltaF, Ewald k-space
 *                 estimate, inhomogeneous estimate.
 * The host system_t made here has the layout of Init_system (common.c:
 * 192-220): positions and the three reference force arrays as
 * vector_array_t with fields[3] aliasing x, y, z.
 *
 * Reading differs from the reference in two deliberate ways: lines are read
 * whole, so a force file with the seven columns Write_exact_forces produces
 * (or the reference's tests-wall.for) is read row by row -- the reference's
 * four-field fscanf loses step with it -- and columns 5-7, when present,
 * fill the k-space reference forces; errors are returned, not exit()ed.
 */
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "host_common.h"
#include "p3m_hip.h"

static void vector_array_free(vector_array_t *v) {
    int d;
    if (v == NULL)
        return;
    if (v->fields) {
        for (d = 0; d < 3; d++)
            free(v->fields[d]);
        free(v->fields);
    }
    free(v);
}

static vector_array_t *vector_array_new(int n) {
    vector_array_t *v = (vector_array_t *)calloc(1, sizeof *v);
    int d;
    if (v == NULL)
        return NULL;
    v->fields = (double **)calloc(3, sizeof(double *));
    if (v->fields == NULL) {
        free(v);
        return NULL;
    }
    for (d = 0; d < 3; d++) {
        v->fields[d] = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
        if (v->fields[d] == NULL) {
            vector_array_free(v);
            return NULL;
        }
    }
    v->x = v->fields[0];
    v->y = v->fields[1];
    v->z = v->fields[2];
    v->size = n;
    return v;
}

static void forces_free(forces_t *f) {
    if (f == NULL)
        return;
    vector_array_free(f->f);
    vector_array_free(f->f_k);
    vector_array_free(f->f_r);
    free(f);
}

static forces_t *forces_new(int n) {
    forces_t *f = (forces_t *)calloc(1, sizeof *f);
    if (f == NULL)
        return NULL;
    f->f = vector_array_new(n);
    f->f_k = vector_array_new(n);
    f->f_r = vector_array_new(n);
    if (!f->f || !f->f_k || !f->f_r) {
        forces_free(f);
        return NULL;
    }
    return f;
}

int p3m_hip_system_free(system_t *s) {
    if (s == NULL)
        return P3M_HIP_OK;
    vector_array_free(s->p);
    forces_free(s->reference);
    free(s->q);
    free(s);
    return P3M_HIP_OK;
}

int p3m_hip_system_new(int n, double length, system_t **out) {
    system_t *s;
    if (out == NULL || n <= 0)
        return p3mh_fail(P3M_HIP_EINVAL, "system_new: needs n > 0");
    *out = NULL;
    s = (system_t *)calloc(1, sizeof *s);
    if (s == NULL)
        return p3mh_fail(P3M_HIP_ENOMEM, "system_new");
    s->nparticles = n;
    s->length = length;
    s->p = vector_array_new(n);
    s->reference = forces_new(n);
    s->q = (double *)calloc((size_t)n, sizeof(double));
    if (!s->p || !s->reference || !s->q) {
        p3m_hip_system_free(s);
        return p3mh_fail(P3M_HIP_ENOMEM, "system_new");
    }
    *out = s;
    return P3M_HIP_OK;
}

static double *col(vector_array_t *v, int d) {
    if (v->fields)
        return v->fields[d];
    return d == 0 ? v->x : (d == 1 ? v->y : v->z);
}

/* one line of any length; 1 on success, 0 at the end of the file */
static int read_line(FILE *f, char **buf, size_t *cap) {
    size_t len = 0;
    int c;
    if (*buf == NULL) {
        *cap = 256;
        *buf = (char *)malloc(*cap);
        if (*buf == NULL)
            return -1;
    }
    while ((c = fgetc(f)) != EOF) {
        if (len + 1 >= *cap) {
            char *nb = (char *)realloc(*buf, 2 * *cap);
            if (nb == NULL)
                return -1;
            *buf = nb;
            *cap *= 2;
        }
        if (c == '\n')
            break;
        (*buf)[len++] = (char)c;
    }
    (*buf)[len] = '\0';
    return (c == EOF && len == 0) ? 0 : 1;
}

/* a line of only white space? */
static int blank(const char *s) {
    for (; *s; s++)
        if (*s != ' ' && *s != '\t' && *s != '\r')
            return 0;
    return 1;
}

/* next data line (blank lines skipped); 1, 0 at the end, -1 on failure */
static int next_data_line(FILE *f, char **buf, size_t *cap) {
    int r;
    while ((r = read_line(f, buf, cap)) == 1)
        if (!blank(*buf))
            return 1;
    return r;
}

int p3m_hip_read_system(const char *path, system_t **out) {
    FILE *f;
    char *line = NULL;
    size_t cap = 0;
    int n = 0, i, rc = P3M_HIP_OK;
    double L = 0.0, v[4];
    system_t *s = NULL;
    if (path == NULL || out == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "read_system: bad arguments");
    *out = NULL;
    f = fopen(path, "r");
    if (f == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "read_system: cannot open '%s': %s", path, strerror(errno));
    /* io.c:126-127 */
    if (fscanf(f, "# Teilchenzahl: %d\n", &n) != 1 || fscanf(f, "# Len: %lf\n", &L) != 1 || n <= 0) {
        fclose(f);
        return p3mh_fail(P3M_HIP_EINVAL, "read_system: '%s' lacks the '# Teilchenzahl:' / '# Len:' header",
                         path);
    }
    rc = p3m_hip_system_new(n, L, &s);
    for (i = 0; i < n && rc == P3M_HIP_OK; i++) {
        const int r = next_data_line(f, &line, &cap);
        if (r != 1 || sscanf(line, "%lf %lf %lf %lf", &v[0], &v[1], &v[2], &v[3]) != 4) {
            rc = p3mh_fail(P3M_HIP_EINVAL, "read_system: '%s' particle %d: expected 'x y z q'", path, i);
            break;
        }
        col(s->p, 0)[i] = v[0];
        col(s->p, 1)[i] = v[1];
        col(s->p, 2)[i] = v[2];
        s->q[i] = v[3];
        s->q2 += v[3] * v[3];
    }
    free(line);
    fclose(f);
    if (rc != P3M_HIP_OK) {
        p3m_hip_system_free(s);
        return rc;
    }
    *out = s;
    return P3M_HIP_OK;
}

int p3m_hip_write_system(const system_t *s, const char *path) {
    FILE *f;
    int i;
    if (s == NULL || path == NULL || s->p == NULL || s->q == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "write_system: bad arguments");
    f = fopen(path, "w");
    if (f == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "write_system: cannot open '%s': %s", path, strerror(errno));
    /* io.c:215-221: six decimals */
    fprintf(f, "# Teilchenzahl: %d\n", s->nparticles);
    fprintf(f, "# Len: %lf\n", s->length);
    for (i = 0; i < s->nparticles; i++)
        fprintf(f, "%lf\t%lf\t%lf\t%lf\n", col(s->p, 0)[i], col(s->p, 1)[i], col(s->p, 2)[i], s->q[i]);
    if (fclose(f) != 0)
        return p3mh_fail(P3M_HIP_EINVAL, "write_system: writing '%s' failed", path);
    return P3M_HIP_OK;
}

int p3m_hip_read_exact_forces(system_t *s, const char *path) {
    FILE *f;
    char *line = NULL;
    size_t cap = 0;
    int i, rc = P3M_HIP_OK, idx, got, d;
    double v[6];
    if (s == NULL || path == NULL || s->reference == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "read_exact_forces: bad arguments");
    f = fopen(path, "r");
    if (f == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "read_exact_forces: cannot open '%s': %s", path,
                         strerror(errno));
    for (i = 0; i < s->nparticles; i++) {
        const int r = next_data_line(f, &line, &cap);
        got = (r == 1) ? sscanf(line, "%d %lf %lf %lf %lf %lf %lf", &idx, &v[0], &v[1], &v[2], &v[3],
                                &v[4], &v[5])
                       : 0;
        if (got != 4 && got != 7) {
            rc = p3mh_fail(P3M_HIP_EINVAL, "read_exact_forces: '%s' particle %d: expected 'i fx fy fz'"
                           " (optionally 'fkx fky fkz')", path, i);
            break;
        }
        for (d = 0; d < 3; d++)
            col(s->reference->f, d)[i] = v[d];
        if (got == 7 && s->reference->f_k)
            for (d = 0; d < 3; d++)
                col(s->reference->f_k, d)[i] = v[3 + d];
    }
    free(line);
    fclose(f);
    return rc;
}

int p3m_hip_write_exact_forces(const system_t *s, const char *path) {
    FILE *f;
    int i;
    if (s == NULL || path == NULL || s->reference == NULL || !s->reference->f || !s->reference->f_k)
        return p3mh_fail(P3M_HIP_EINVAL, "write_exact_forces: bad arguments");
    f = fopen(path, "w");
    if (f == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "write_exact_forces: cannot open '%s': %s", path,
                         strerror(errno));
    /* io.c:171-174: index, total and k-space reference force, six decimals */
    for (i = 0; i < s->nparticles; i++)
        fprintf(f, "%d\t%lf\t%lf\t%lf\t%lf\t%lf\t%lf\n", i, col(s->reference->f, 0)[i],
                col(s->reference->f, 1)[i], col(s->reference->f, 2)[i], col(s->reference->f_k, 0)[i],
                col(s->reference->f_k, 1)[i], col(s->reference->f_k, 2)[i]);
    if (fclose(f) != 0)
        return p3mh_fail(P3M_HIP_EINVAL, "write_exact_forces: writing '%s' failed", path);
    return P3M_HIP_OK;
}

int p3m_hip_format_out_row(char *buf, size_t size, const double v[8]) {
    int w;
    if (buf == NULL || v == NULL || size == 0)
        return p3mh_fail(P3M_HIP_EINVAL, "format_out_row: bad arguments");
    /* main.c:452-455 */
    w = snprintf(buf, size, "% lf\t% e\t% e\t% e\t% e\t% e\t% e\t% e\n", v[0], v[1], v[2], v[3], v[4],
                 v[5], v[6], v[7]);
    if (w < 0 || (size_t)w >= size)
        return p3mh_fail(P3M_HIP_EINVAL, "format_out_row: buffer of %zu bytes too small", size);
    return P3M_HIP_OK;
}
