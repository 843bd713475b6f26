/*
 * aic_oracle.h -- panel record and vector helpers shared by the CPU oracle's
 * panel (aic_oracle.c) and element (elem_oracle.c) restatements.
 *
 * TEST INFRASTRUCTURE ONLY; see aic_oracle.c.
 */
#ifndef DUST_AIC_ORACLE_H
#define DUST_AIC_ORACLE_H
#include <math.h>

#define AO_PI 3.1415926535897932384626433832795028841971693993751058209749445923

/* Panel geometry record; layout shared with the tests (74 doubles). */
enum {
  G_NVER = 0,       /* number of vertices (3 or 4), stored as double     */
  G_CEN = 1,        /* 3                                                */
  G_NOR = 4,        /* 3                                                */
  G_TANG = 7,       /* tang(3,2) column-major: t1 then t2               */
  G_AREA = 13,      /* 1                                                */
  G_VER = 14,       /* ver(3,4)                                         */
  G_EVEC = 26,      /* edge_vec(3,4)                                    */
  G_ELEN = 38,      /* edge_len(4)                                      */
  G_EUNI = 42,      /* edge_uni(3,4)                                    */
  G_COST = 54,      /* cosTi(4)                                         */
  G_SINT = 58,      /* sinTi(4)                                         */
  G_VERP = 62,      /* verp(3,4)                                        */
  G_SIZE = 74
};

static inline void ao_cross(const double *a, const double *b, double *r) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}
static inline double ao_dot(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline double ao_norm(const double *a) { return sqrt(ao_dot(a, a)); }

static const int next_tri[3] = {1, 2, 0}, next_qua[4] = {1, 2, 3, 0};
static const int prev_tri[3] = {2, 0, 1}, prev_qua[4] = {3, 0, 1, 2};

static inline double maxedge(const double *g) {
  int nv = (int)g[G_NVER];
  double m = g[G_ELEN];
  for (int i = 1; i < nv; ++i) m = g[G_ELEN + i] > m ? g[G_ELEN + i] : m;
  return m;
}

void ao_geo(int n, const int *nver, const double *ver, double *G);
double ao_pot_doublet(const double *g, const double pos[3], double ff_ratio);
double ao_pot_source(const double *g, double dou, const double pos[3], int *warn);
void ao_vel_doublet(const double *g, const double pos[3], double ff_ratio, double r_rankine,
                    double r_cutoff, double v[3]);

#endif
