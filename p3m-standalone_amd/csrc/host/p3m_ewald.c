/* Host side of the sampled FP64 Ewald sum (the reference's Ewald method,
 * ewald.c:84-229, evaluated on a subset of particles as an accuracy oracle
 * for systems too large for a full Ewald evaluation). */
#include <stddef.h>

#include "p3m_hip.h"

#include "../p3m_kernels.h"
#include "host_common.h"

int p3m_hip_ewald_sample_device(int n, double length, const double *x, const double *y,
                                const double *z, const double *q, int nsample, const int *idx,
                                double alpha, double rcut, int kmax, double alpha_split,
                                double *f_total, double *f_kspace) {
    p3mk_queue *qu = NULL;
    int rc;
    if (n < 0 || nsample < 0 || x == NULL || y == NULL || z == NULL || q == NULL || idx == NULL ||
        f_total == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "bad arguments");
    if (!(alpha > 0.0) || !(length > 0.0) || rcut < 0.0 || rcut > 0.5 * length || kmax < 1 ||
        kmax > 512)
        return p3mh_fail(P3M_HIP_EINVAL, "Ewald sample needs alpha > 0, 0 <= rcut <= L/2, "
                                         "1 <= kmax <= 512");
    if (f_kspace != NULL && !(alpha_split > 0.0))
        return p3mh_fail(P3M_HIP_EINVAL, "the k-space split needs alpha_split > 0");
    TRYK(p3mk_queue_create(&qu), "stream");
    rc = p3mk_ewald_sample(n, length, x, y, z, q, nsample, idx, alpha, rcut, kmax,
                           f_kspace != NULL ? alpha_split : 0.0, f_total, f_kspace, qu);
    p3mk_queue_destroy(qu);
    if (rc != P3MK_OK)
        return p3mh_kfail(rc == P3MK_EINVAL ? P3M_HIP_EINVAL : P3M_HIP_EDEVICE, "Ewald sample");
    return P3M_HIP_OK;
}

int p3m_hip_ewald_sample_device_noalloc(int n, double length, const double *x, const double *y,
                                        const double *z, const double *q, int nsample,
                                        const int *idx, double alpha, double rcut, int kmax,
                                        double alpha_split, double *f_total, double *f_kspace,
                                        void *stream) {
    int rc;
    if (n < 0 || nsample < 0 || x == NULL || y == NULL || z == NULL || q == NULL || idx == NULL ||
        f_total == NULL)
        return p3mh_fail(P3M_HIP_EINVAL, "bad arguments");
    if (!(alpha > 0.0) || !(length > 0.0) || rcut < 0.0 || rcut > 0.5 * length || kmax < 1)
        return p3mh_fail(P3M_HIP_EINVAL, "Ewald sample needs alpha > 0, 0 <= rcut <= L/2, kmax >= 1");
    if (f_kspace != NULL && !(alpha_split > 0.0))
        return p3mh_fail(P3M_HIP_EINVAL, "the k-space split needs alpha_split > 0");
    rc = p3mk_ewald_sample_noalloc(n, length, x, y, z, q, nsample, idx, alpha, rcut, kmax,
                                   f_kspace != NULL ? alpha_split : 0.0, f_total, f_kspace, stream);
    if (rc != P3MK_OK)
        return p3mh_kfail(rc == P3MK_EINVAL ? P3M_HIP_EINVAL : P3M_HIP_EDEVICE,
                          "allocation-free Ewald sample");
    return P3M_HIP_OK;
}
