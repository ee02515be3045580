/*
 * lto_internal.h -- 3x3 complex arithmetic for the CPU oracle (test infrastructure only).
 *
 * The reference does all of this with Eigen fixed-size matrices
 * (LettuceGauge/defines.hpp:122).  Here a matrix is 9 C99 complex doubles, row-major,
 * which is byte-identical to the reference's AoS link storage.
 */
#ifndef LTO_INTERNAL_H
#define LTO_INTERNAL_H

#include <complex.h>
#include <math.h>
#include <string.h>
#include "lto.h"

typedef double complex cplx;
typedef struct { cplx m[9]; } mat3;

#define LTO_PI 3.14159265358979323846

static inline mat3 m_zero(void) { mat3 r; for (int i = 0; i < 9; ++i) r.m[i] = 0; return r; }
static inline mat3 m_id(void) { mat3 r = m_zero(); r.m[0] = r.m[4] = r.m[8] = 1; return r; }

static inline mat3 m_mul(mat3 a, mat3 b)
{
    mat3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r.m[3*i+j] = a.m[3*i] * b.m[j] + a.m[3*i+1] * b.m[3+j] + a.m[3*i+2] * b.m[6+j];
    return r;
}
static inline mat3 m_dag(mat3 a)
{
    mat3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[3*i+j] = conj(a.m[3*j+i]);
    return r;
}
/* a * b^dagger, a^dagger * b */
static inline mat3 m_mul_nd(mat3 a, mat3 b) { return m_mul(a, m_dag(b)); }
static inline mat3 m_mul_dn(mat3 a, mat3 b) { return m_mul(m_dag(a), b); }
static inline mat3 m_add(mat3 a, mat3 b) { for (int i = 0; i < 9; ++i) a.m[i] += b.m[i]; return a; }
static inline mat3 m_sub(mat3 a, mat3 b) { for (int i = 0; i < 9; ++i) a.m[i] -= b.m[i]; return a; }
static inline mat3 m_scale(cplx s, mat3 a) { for (int i = 0; i < 9; ++i) a.m[i] *= s; return a; }
static inline mat3 m_rscale(double s, mat3 a) { for (int i = 0; i < 9; ++i) a.m[i] *= s; return a; }
static inline cplx m_tr(mat3 a) { return a.m[0] + a.m[4] + a.m[8]; }
static inline double m_retr_mul(mat3 a, mat3 b)
{
    double s = 0;
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) s += creal(a.m[3*i+k] * b.m[3*k+i]);
    return s;
}

static inline mat3 m_load(const double *p) { mat3 r; memcpy(&r, p, sizeof r); return r; }
static inline void m_store(double *p, mat3 a) { memcpy(p, &a, sizeof a); }

/* SU3::Projection::Algebra / TracelessAntihermitian (LettuceGauge/math/su3.hpp:422-433) */
static inline mat3 m_algebra(mat3 a)
{
    mat3 b = m_sub(a, m_dag(a));
    cplx tr = m_tr(b);
    mat3 r = m_rscale(0.5, b);
    for (int i = 0; i < 3; ++i) r.m[4*i] -= (1.0/6.0) * tr;
    return r;
}
/* SU3::Projection::Antihermitian (su3.hpp:437-441) */
static inline mat3 m_antiherm(mat3 a) { return m_rscale(0.5, m_sub(a, m_dag(a))); }
/* SU3::Projection::TracelessHermitian (su3.hpp:444-448) */
static inline mat3 m_tl_herm(mat3 a)
{
    mat3 b = m_add(a, m_dag(a));
    cplx tr = m_tr(b);
    mat3 r = m_rscale(0.5, b);
    for (int i = 0; i < 3; ++i) r.m[4*i] -= (1.0/6.0) * tr;
    return r;
}
/* SU3::Projection::Traceless (su3.hpp:458-462) */
static inline mat3 m_traceless(mat3 a)
{
    cplx tr = m_tr(a);
    for (int i = 0; i < 3; ++i) a.m[4*i] -= (1.0/3.0) * tr;
    return a;
}

/* SU3::Projection::GramSchmidt + RestoreLastRow (su3.hpp:362-394) */
static inline void m_gram_schmidt(mat3 *M)
{
    cplx *m = M->m;
    double n0 = 1.0 / sqrt(creal(m[0]*conj(m[0]) + m[1]*conj(m[1]) + m[2]*conj(m[2])));
    for (int n = 0; n < 3; ++n) m[n] = n0 * m[n];
    cplx psi = m[3]*conj(m[0]) + m[4]*conj(m[1]) + m[5]*conj(m[2]);
    for (int n = 0; n < 3; ++n) m[3+n] = m[3+n] - psi * m[n];
    double n1 = 1.0 / sqrt(creal(m[3]*conj(m[3]) + m[4]*conj(m[4]) + m[5]*conj(m[5])));
    for (int n = 0; n < 3; ++n) m[3+n] = n1 * m[3+n];
    m[6] = conj(m[1]*m[5] - m[2]*m[4]);
    m[7] = conj(m[2]*m[3] - m[0]*m[5]);
    m[8] = conj(m[0]*m[4] - m[1]*m[3]);
}

/* Cayley-Hamilton coefficients of exp(iQ) (LettuceGauge/math/su3_exp.hpp:204-283) */
typedef struct {
    mat3 Q, Q2;
    double c0, c1, c0max, theta, u, w, u2, w2;
    int signflip;
    cplx exp_miu, exp_2iu, cosw, i_xi0;
    double denom;
    cplx h0, h1, h2;
} lto_expc;

/* return_xi_0 (su3_exp.hpp:55-68) */
static inline double lto_xi0(double w)
{
    if (fabs(w) <= 0.05) {
        double w2 = w * w;
        return 1.0 - (1.0/6.0) * w2 * (1.0 - (1.0/20.0) * w2 * (1.0 - (1.0/42.0) * w2));
    }
    return sin(w) / w;
}
/* return_xi_1 (su3_exp.hpp:70-82) */
static inline double lto_xi1(double w)
{
    if (fabs(w) <= 0.115) {
        double w2 = w * w;
        return -(1.0/3.0) + (1.0/30.0) * w2 * (1.0 - (1.0/28.0) * w2 * (1.0 - (1.0/54.0) * w2));
    }
    return cos(w) / (w * w) - sin(w) / (w * w * w);
}

void lto_exp_constants(mat3 Q, lto_expc *e);
mat3 lto_exp_from(const lto_expc *e);
mat3 lto_exp(mat3 Q);

/* lattice indexing (lattice.hpp:366-382) */
typedef struct { int c[4]; } lto_site;
static inline size_t lto_site_index(lto_dims d, int t, int x, int y, int z)
{
    return (((size_t)t * d.nx + x) * d.ny + y) * d.nz + z;
}
static inline lto_site lto_site_coord(lto_dims d, size_t s)
{
    lto_site r;
    r.c[3] = (int)(s % d.nz); s /= d.nz;
    r.c[2] = (int)(s % d.ny); s /= d.ny;
    r.c[1] = (int)(s % d.nx); s /= d.nx;
    r.c[0] = (int)s;
    return r;
}
static inline int lto_len(lto_dims d, int mu) { return mu == 0 ? d.nt : mu == 1 ? d.nx : mu == 2 ? d.ny : d.nz; }
/* GaugeField4D::Move<dist> (lattice.hpp:249-296) */
static inline size_t lto_move(lto_dims d, size_t s, int mu, int dist)
{
    lto_site c = lto_site_coord(d, s);
    int L = lto_len(d, mu);
    c.c[mu] = ((c.c[mu] + dist) % L + L) % L;
    return lto_site_index(d, c.c[0], c.c[1], c.c[2], c.c[3]);
}
static inline mat3 lto_link(const double *U, size_t s, int mu) { return m_load(U + 18 * (4 * s + mu)); }
static inline void lto_set_link(double *U, size_t s, int mu, mat3 a) { m_store(U + 18 * (4 * s + mu), a); }

mat3 lto_staple_m(const double *U, lto_dims d, size_t s, int mu);

#endif
