/* Built-in k-space error tables of the ik and interlaced-ik methods
 * (the reference's Q_ik / Q_ik_i of q_ik.c, q_ik_i.c, keyed by the mesh list
 * of q.c:1).  Data written by csrc/gen_qtables.py. */
#ifndef P3M_Q_TABLES_H
#define P3M_Q_TABLES_H

#define P3M_Q_NMESHES 83

extern const int p3m_q_meshes[P3M_Q_NMESHES]; /* ascending */
extern const double p3m_q_ik[P3M_Q_NMESHES][7][101];
extern const double p3m_q_ik_i[P3M_Q_NMESHES][7][101];

#endif
