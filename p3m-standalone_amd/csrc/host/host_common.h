/* Shared helpers of the host C layer: error reporting and the method
 * properties of the reference's method ids. */
#ifndef P3M_HOST_COMMON_H
#define P3M_HOST_COMMON_H

#include "p3m_hip.h"

/* set the thread's last-error message; returns code */
int p3mh_fail(int code, const char *fmt, ...);
/* p3mh_fail with the kernel layer's message appended */
int p3mh_kfail(int code, const char *what);

#define TRYK(expr, what)                          \
    do {                                          \
        int rc_ = (expr);                         \
        if (rc_ != P3MK_OK)                       \
            return p3mh_kfail(rc_, what);         \
    } while (0)

#define TRY(expr)                                 \
    do {                                          \
        int rc_ = (expr);                         \
        if (rc_ != P3M_HIP_OK)                    \
            return rc_;                           \
    } while (0)

/* method ids of the reference (types.h:290-303, p3m-*.c method tables) */
static inline int p3mh_is_ad(int id) { return id == 2 || id == 3 || id == 7; }
static inline int p3mh_is_interlaced(int id) { return id == 1 || id == 3; }
static inline int p3mh_is_r2c(int id) { return id == 6 || id == 7; }
static inline int p3mh_has_self_force(int id) { return id == 2 || id == 7; }
static inline int p3mh_known_method(int id) {
    return id == 0 || id == 1 || id == 2 || id == 3 || id == 6 || id == 7;
}

#endif
