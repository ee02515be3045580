/*
 * lto_core.c -- CPU oracle (test infrastructure only): RNG, lattice set-up, the
 * Cayley-Hamilton exponential, the Wilson staple and the local-update sweeps
 * (Cabibbo-Marinari heatbath and SU(2)-subgroup overrelaxation).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this.
 */
#include "lto_internal.h"
#include <omp.h>
#include <float.h>

/* ------------------------------------------------------------------------- */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Round constants are the
 * published ones; the reference uses pcg64 per link instead
 * (LettuceGauge/random/prng_wrapper.hpp:43-379), so streams are not comparable
 * with the reference and heatbath parity against it is statistical.          */
void lto_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 53-bit uniform in [0,1) from two 32-bit words */
double lto_u01(uint32_t a, uint32_t b)
{
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

/* counter layout shared with the device generator:
 *   key = seed, ctr = {site, stream<<24 | k<<4 | sub<<2 | mu, counter_lo, counter_hi} */
static inline void lto_draw2(uint64_t seed, uint32_t site, uint32_t stream, uint32_t k,
                             uint32_t sub, uint32_t mu, uint64_t counter, double u[2])
{
    uint32_t ctr[4] = { site, (stream << 24) | (k << 4) | (sub << 2) | mu,
                        (uint32_t)counter, (uint32_t)(counter >> 32) };
    uint32_t key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) };
    uint32_t o[4];
    lto_philox4x32_10(ctr, key, o);
    u[0] = lto_u01(o[0], o[1]);
    u[1] = lto_u01(o[2], o[3]);
}

static inline void lto_gauss2(const double u[2], double g[2])
{
    double r = sqrt(-2.0 * log(1.0 - u[0]));
    double ph = 2.0 * LTO_PI * u[1];
    g[0] = r * cos(ph);
    g[1] = r * sin(ph);
}

/* ------------------------------------------------------------------------- */
size_t lto_volume(lto_dims d) { return (size_t)d.nt * d.nx * d.ny * d.nz; }

/* OpenMP threads a parallel region of the oracle uses */
int lto_threads(void) { return omp_get_max_threads(); }

/* GaugeField4D::SetToIdentity (lattice.hpp:170-183) */
void lto_set_identity(double *U, lto_dims d)
{
    size_t n = 4 * lto_volume(d);
    #pragma omp parallel for
    for (size_t i = 0; i < n; ++i) m_store(U + 18 * i, m_id());
}

/* Haar-random SU(3) start: Gram-Schmidt of a complex Gaussian matrix (rows 0,1),
 * row 2 = conj(row0 x row1).  The reference has no Haar start; parity unpinned
 * against it, pinned by distribution tests (tests/test_oracle_invariants.py). */
void lto_haar(double *U, lto_dims d, uint64_t seed)
{
    size_t V = lto_volume(d);
    #pragma omp parallel for
    for (size_t s = 0; s < V; ++s)
        for (int mu = 0; mu < 4; ++mu) {
            mat3 M = m_zero();
            for (int k = 0; k < 6; ++k) {
                double u[2], g[2];
                lto_draw2(seed, (uint32_t)s, 2u, (uint32_t)k, 0u, (uint32_t)mu, 0, u);
                lto_gauss2(u, g);
                M.m[k] = g[0] + I * g[1];
            }
            m_gram_schmidt(&M);
            lto_set_link(U, s, mu, M);
        }
}

void lto_reunitarize(double *U, lto_dims d)
{
    size_t n = 4 * lto_volume(d);
    #pragma omp parallel for
    for (size_t i = 0; i < n; ++i) {
        mat3 M = m_load(U + 18 * i);
        m_gram_schmidt(&M);
        m_store(U + 18 * i, M);
    }
}

void lto_gram_schmidt(double *M) { mat3 a = m_load(M); m_gram_schmidt(&a); m_store(M, a); }

/* conj(x y - z w) with the rounding steps of the device (one product, then three
 * fused multiply-adds per part) */
static void conj_cross_fma(const double *x, const double *y, const double *z, const double *w, double *out)
{
    double re = z[1] * w[1];
    re = fma(-z[0], w[0], re);
    re = fma(-x[1], y[1], re);
    re = fma(x[0], y[0], re);
    double im = z[1] * w[0];
    im = fma(z[0], w[1], im);
    im = fma(-x[1], y[0], im);
    im = fma(-x[0], y[1], im);
    out[0] = re;
    out[1] = im;
}

/* Row 2 of every link rebuilt from rows 0 and 1 as conj(row0 x row1)
 * (SU3::Projection::RestoreLastRow, math/su3.hpp:362-367) with the device's rounding:
 * the links a device field holds after an upload of rows 0 and 1. */
void lto_third_row_fma(double *U, size_t nlinks)
{
    for (size_t i = 0; i < nlinks; ++i) {
        double *m = U + 18 * i;   /* entry k at m + 2k, row-major */
        conj_cross_fma(m + 2, m + 10, m + 4, m + 8, m + 12);
        conj_cross_fma(m + 4, m + 6, m + 0, m + 10, m + 14);
        conj_cross_fma(m + 0, m + 8, m + 2, m + 6, m + 16);
    }
}
void lto_algebra(const double *M, double *out) { m_store(out, m_algebra(m_load(M))); }

/* ------------------------------------------------------------------------- */
/* SU3::ExpConstants (LettuceGauge/math/su3_exp.hpp:204-283), uDerivedConstants
 * and wDerivedConstants (:84-119) */
void lto_exp_constants(mat3 Q, lto_expc *e)
{
    e->Q = Q;
    e->Q2 = m_mul(Q, Q);
    double c0s = (1.0/3.0) * creal(m_tr(m_mul(Q, e->Q2)));
    e->c0 = fabs(c0s);
    e->c1 = 0.5 * creal(m_tr(e->Q2));
    e->c0max = 2.0 * pow(e->c1 / 3.0, 1.5);
    e->signflip = c0s < 0.0;
    e->theta = acos(fmin(e->c0 / e->c0max, 1.0));
    e->u = sqrt((1.0/3.0) * e->c1) * cos((1.0/3.0) * e->theta);
    e->u2 = e->u * e->u;
    e->exp_miu = cexp(-I * e->u);
    e->exp_2iu = conj(e->exp_miu * e->exp_miu);
    e->w = sqrt(e->c1) * sin((1.0/3.0) * e->theta);
    e->w2 = e->w * e->w;
    e->cosw = cos(e->w);
    e->i_xi0 = I * lto_xi0(e->w);
    if (e->c0max < DBL_MIN) {
        e->denom = 1.0; e->h0 = 1.0; e->h1 = 0.0; e->h2 = 0.0;
        return;
    }
    double u = e->u, u2 = e->u2, w2 = e->w2;
    e->denom = 1.0 / (9.0 * u2 - w2);
    cplx h0 = (u2 - w2) * e->exp_2iu + e->exp_miu * (8.0 * u2 * e->cosw + 2.0 * u * e->i_xi0 * (3.0 * u2 + w2));
    cplx h1 = 2.0 * u * e->exp_2iu - e->exp_miu * (2.0 * u * e->cosw - e->i_xi0 * (3.0 * u2 - w2));
    cplx h2 = e->exp_2iu - e->exp_miu * (e->cosw + 3.0 * u * e->i_xi0);
    if (e->signflip) { h0 = conj(h0); h1 = -conj(h1); h2 = conj(h2); }
    e->h0 = h0; e->h1 = h1; e->h2 = h2;
}

/* SU3::exp (su3_exp.hpp:420-432) */
mat3 lto_exp_from(const lto_expc *e)
{
    mat3 r;
    for (int i = 0; i < 9; ++i) r.m[i] = e->h1 * e->Q.m[i] + e->h2 * e->Q2.m[i];
    for (int i = 0; i < 3; ++i) r.m[4*i] += e->h0;
    return m_rscale(e->denom, r);
}
mat3 lto_exp(mat3 Q) { lto_expc e; lto_exp_constants(Q, &e); return lto_exp_from(&e); }
void lto_su3_exp(const double *Q, double *out) { m_store(out, lto_exp(m_load(Q))); }

/* ------------------------------------------------------------------------- */
/* Rectangular<1>::StaplePlaq (LettuceGauge/actions/gauge/rectangular_action.hpp:103-184),
 * identical to WilsonAction::Staple (wilson_action.hpp:77-172):
 *   sum over nu != mu, nu ascending, of
 *   U_nu(x) U_mu(x+nu) U_nu(x+mu)^+  +  U_nu(x-nu)^+ U_mu(x-nu) U_nu(x-nu+mu)        */
mat3 lto_staple_m(const double *U, lto_dims d, size_t s, int mu)
{
    mat3 st = m_zero();
    size_t smu = lto_move(d, s, mu, 1);
    for (int nu = 0; nu < 4; ++nu) {
        if (nu == mu) continue;
        size_t snu = lto_move(d, s, nu, 1);
        size_t snd = lto_move(d, s, nu, -1);
        size_t snd_mu = lto_move(d, snd, mu, 1);
        mat3 a = m_mul_nd(m_mul(lto_link(U, s, nu), lto_link(U, snu, mu)), lto_link(U, smu, nu));
        mat3 b = m_mul(m_mul_dn(lto_link(U, snd, nu), lto_link(U, snd, mu)), lto_link(U, snd_mu, nu));
        st = m_add(st, m_add(a, b));
    }
    return st;
}

/* Rectangular<2>::StapleRect (rectangular_action.hpp:186-225): for nu = mu+1, mu+2, mu+3
 * (mod 4) the six 1x2 staples of the link (x, mu) -- two orthogonal to it, two
 * extending it forward, two backward -- each the path from x+mu to x.  The
 * reference leaves the accumulator uninitialised (Eigen); it starts from zero here. */
static mat3 staple_rect_m(const double *U, lto_dims d, size_t s, int mu)
{
    mat3 st = m_zero();
    for (int off = 1; off < 4; ++off) {
        const int nu = (mu + off) % 4;
        size_t nup = lto_move(d, s, nu, 1), nupp = lto_move(d, nup, nu, 1), mup_nup = lto_move(d, nup, mu, 1);
        size_t mup = lto_move(d, s, mu, 1);
        size_t nud = lto_move(d, s, nu, -1), nudd = lto_move(d, nud, nu, -1), mup_nudd = lto_move(d, nudd, mu, 1);
        size_t mup_nud = lto_move(d, nud, mu, 1);
        /* orthogonal: U_nu(x) U_nu(x+nu) U_mu(x+2nu) U_nu(x+mu+nu)^+ U_nu(x+mu)^+ and its mirror */
        mat3 a = m_mul(m_mul(lto_link(U, s, nu), lto_link(U, nup, nu)), lto_link(U, nupp, mu));
        a = m_mul_nd(m_mul_nd(a, lto_link(U, mup_nup, nu)), lto_link(U, mup, nu));
        mat3 b = m_mul(m_mul_dn(lto_link(U, nud, nu), m_dag(lto_link(U, nudd, nu))), lto_link(U, nudd, mu));
        b = m_mul(m_mul(b, lto_link(U, mup_nudd, nu)), lto_link(U, mup_nud, nu));
        st = m_add(st, m_add(a, b));
        /* forward: [U_nu(x) U_mu(x+nu) U_mu(x+mu+nu) U_nu(x+2mu)^+ + mirror] U_mu(x+mu)^+ */
        size_t mupp = lto_move(d, mup, mu, 1), mupp_nud = lto_move(d, mup_nud, mu, 1);
        mat3 c = m_mul_nd(m_mul(m_mul(lto_link(U, s, nu), lto_link(U, nup, mu)), lto_link(U, mup_nup, mu)),
                          lto_link(U, mupp, nu));
        mat3 e = m_mul(m_mul(m_mul_dn(lto_link(U, nud, nu), lto_link(U, nud, mu)), lto_link(U, mup_nud, mu)),
                       lto_link(U, mupp_nud, nu));
        st = m_add(st, m_mul_nd(m_add(c, e), lto_link(U, mup, mu)));
        /* backward: U_mu(x-mu)^+ [U_nu(x-mu) U_mu(x-mu+nu) U_mu(x+nu) U_nu(x+mu)^+ + mirror] */
        size_t mud = lto_move(d, s, mu, -1), mud_nup = lto_move(d, mud, nu, 1), mud_nud = lto_move(d, mud, nu, -1);
        mat3 f = m_mul_nd(m_mul(m_mul(lto_link(U, mud, nu), lto_link(U, mud_nup, mu)), lto_link(U, nup, mu)),
                          lto_link(U, mup, nu));
        mat3 h = m_mul(m_mul(m_mul_dn(lto_link(U, mud_nud, nu), lto_link(U, mud_nud, mu)), lto_link(U, nud, mu)),
                       lto_link(U, mup_nud, nu));
        st = m_add(st, m_mul_dn(lto_link(U, mud, mu), m_add(f, h)));
    }
    return st;
}

/* Rectangular<R>::Staple (rectangular_action.hpp:232-242): c_plaq StaplePlaq + c_rect
 * StapleRect, or StaplePlaq alone for c_rect = 0 (the Wilson action, c_plaq = 1) */
static mat3 staple_action_m(const double *U, lto_dims d, size_t s, int mu, double c_plaq, double c_rect)
{
    mat3 p = lto_staple_m(U, d, s, mu);
    if (c_rect == 0.0) return c_plaq == 1.0 ? p : m_rscale(c_plaq, p);
    return m_add(m_rscale(c_plaq, p), m_rscale(c_rect, staple_rect_m(U, d, s, mu)));
}

void lto_staple_action_field(const double *U, lto_dims d, double c_plaq, double c_rect, double *ST)
{
    size_t V = lto_volume(d);
    #pragma omp parallel for
    for (size_t s = 0; s < V; ++s)
        for (int mu = 0; mu < 4; ++mu) lto_set_link(ST, s, mu, staple_action_m(U, d, s, mu, c_plaq, c_rect));
}

void lto_staple(const double *U, lto_dims d, size_t site, int mu, double *st)
{
    m_store(st, lto_staple_m(U, d, site, mu));
}

void lto_staple_field(const double *U, lto_dims d, double *ST)
{
    size_t V = lto_volume(d);
    #pragma omp parallel for
    for (size_t s = 0; s < V; ++s)
        for (int mu = 0; mu < 4; ++mu) lto_set_link(ST, s, mu, lto_staple_m(U, d, s, mu));
}

/* ------------------------------------------------------------------------- */
/* SU(2) subgroup helpers (LettuceGauge/math/su2.hpp:21-174).  A subgroup is the
 * index pair (i,j) in {(0,1),(0,2),(1,2)}; an SU2 element is (e11, e12).       */
typedef struct { cplx e11, e12; } su2;
static const int SUB_I[3] = {0, 0, 1}, SUB_J[3] = {1, 2, 2};

/* Extract01/02/12 of U*st_adj (su2.hpp:151-174) */
static inline su2 su2_extract(mat3 W, int g)
{
    int i = SUB_I[g], j = SUB_J[g];
    su2 r = { 0.5 * (W.m[3*i+i] + conj(W.m[3*j+j])), 0.5 * (W.m[3*i+j] - conj(W.m[3*j+i])) };
    return r;
}
/* Embed0x(X) * U (su2.hpp:118-146): rows i and j of U mix */
static inline mat3 su2_embed_mul(su2 X, mat3 U, int g)
{
    int i = SUB_I[g], j = SUB_J[g];
    mat3 r = U;
    for (int n = 0; n < 3; ++n) {
        r.m[3*i+n] = X.e11 * U.m[3*i+n] + X.e12 * U.m[3*j+n];
        r.m[3*j+n] = -conj(X.e12) * U.m[3*i+n] + conj(X.e11) * U.m[3*j+n];
    }
    return r;
}
/* SU2_comp::operator* (su2.hpp:38-41) */
static inline su2 su2_mul(su2 a, su2 b)
{
    su2 r = { a.e11 * b.e11 - a.e12 * conj(b.e12), a.e11 * b.e12 + a.e12 * conj(b.e11) };
    return r;
}
static inline double su2_det(su2 a)
{
    return creal(a.e11) * creal(a.e11) + cimag(a.e11) * cimag(a.e11)
         + creal(a.e12) * creal(a.e12) + cimag(a.e12) * cimag(a.e12);
}

/* OverrelaxationSubgroupKernel::OverrelaxationSU2 (overrelaxation.hpp:76-86) */
static inline su2 su2_overrelax(su2 A)
{
    double a_norm = 1.0 / sqrt(su2_det(A));
    su2 V = { a_norm * A.e11, a_norm * A.e12 };
    su2 VV = su2_mul(V, V);
    su2 r = { conj(VV.e11), -VV.e12 };
    return r;
}

/* HeatbathKernel::HeatbathSU2 (heatbath.hpp:46-104) with Philox draws:
 * attempt a uses draw blocks k=2a (r1,r2) and k=2a+1 (r3,r0); direction uses k=0xFFFFF. */
static su2 su2_heatbath(su2 A, double prefactor, int max_iter, uint64_t seed, uint32_t site,
                        int sub, int mu, uint64_t counter, int *attempts)
{
    double a_norm = 1.0 / sqrt(su2_det(A));
    su2 V = { a_norm * A.e11, a_norm * A.e12 };
    double lambda_sq, r0;
    int it = 0;
    for (;;) {
        double ua[2], ub[2];
        lto_draw2(seed, site, 0u, (uint32_t)(2 * it), (uint32_t)sub, (uint32_t)mu, counter, ua);
        lto_draw2(seed, site, 0u, (uint32_t)(2 * it + 1), (uint32_t)sub, (uint32_t)mu, counter, ub);
        double r1 = 1.0 - ua[0];
        double x1 = log(r1);
        double r2 = 1.0 - ua[1];
        double x2 = cos(2.0 * LTO_PI * r2);
        double r3 = 1.0 - ub[0];
        double x3 = log(r3);
        lambda_sq = (-0.25 * prefactor * a_norm) * (x1 + x2 * x2 * x3);
        r0 = ub[1];
        if (it > max_iter) {
            if (attempts) *attempts = it + 1;
            su2 id = { 1.0, 0.0 };
            return id;
        }
        ++it;
        if (!(r0 * r0 + lambda_sq >= 1.0)) break;
    }
    if (attempts) *attempts = it;
    double x0 = 1.0 - 2.0 * lambda_sq;
    double abs_x = sqrt(1.0 - x0 * x0);
    double ud[2];
    lto_draw2(seed, site, 0u, 0xFFFFFu, (uint32_t)sub, (uint32_t)mu, counter, ud);
    double phi = ud[0];
    double cos_theta = 1.0 - 2.0 * ud[1];
    double vec_norm = abs_x * sqrt(1.0 - cos_theta * cos_theta);
    double x1 = vec_norm * cos(2.0 * LTO_PI * phi);
    double x2 = vec_norm * sin(2.0 * LTO_PI * phi);
    double x3 = abs_x * cos_theta;
    su2 X = { x0 + I * x1, x2 + I * x3 };
    su2 Vd = { conj(V.e11), -V.e12 };
    return su2_mul(X, Vd);
}

/* the gauge action of the sweeps: (c_plaq, c_rect) of Rectangular<R> */
static double g_c_plaq = 1.0, g_c_rect = 0.0;

/* OverrelaxationSubgroupKernel::operator() (overrelaxation.hpp:91-116) */
static void or_link(double *U, lto_dims d, size_t s, int mu)
{
    mat3 st_adj = m_dag(staple_action_m(U, d, s, mu, g_c_plaq, g_c_rect));
    mat3 L = lto_link(U, s, mu);
    for (int g = 0; g < 3; ++g) {
        su2 A = su2_extract(m_mul(L, st_adj), g);
        L = su2_embed_mul(su2_overrelax(A), L, g);
    }
    m_gram_schmidt(&L);
    lto_set_link(U, s, mu, L);
}

/* HeatbathKernel::operator() (heatbath.hpp:110-129) */
static void hb_link(double *U, lto_dims d, size_t s, int mu, double prefactor, int max_iter,
                    uint64_t seed, uint64_t counter, int *attempts)
{
    mat3 st_adj = m_dag(staple_action_m(U, d, s, mu, g_c_plaq, g_c_rect));
    mat3 L = lto_link(U, s, mu);
    for (int g = 0; g < 3; ++g) {
        su2 A = su2_extract(m_mul(L, st_adj), g);
        int a = 0;
        su2 X = su2_heatbath(A, prefactor, max_iter, seed, (uint32_t)s, g, mu, counter, &a);
        if (attempts) attempts[g] = a;
        L = su2_embed_mul(X, L, g);
    }
    m_gram_schmidt(&L);
    lto_set_link(U, s, mu, L);
}

/* colour of a link: Checkerboard (iterators.hpp:65-87) uses parity (t+x+y+z)&1,
 * Checkerboard4 (iterators.hpp:141-169) uses (sum + x_mu) % 4 */
static inline int link_colour(lto_dims d, size_t s, int mu, int schedule)
{
    lto_site c = lto_site_coord(d, s);
    int sum = c.c[0] + c.c[1] + c.c[2] + c.c[3];
    return schedule == LTO_CB4 ? (sum + c.c[mu]) % 4 : (sum & 1);
}

void lto_or_sweep(double *U, lto_dims d, int n_sweep, int schedule)
{
    size_t V = lto_volume(d);
    int ncol = schedule == LTO_CB4 ? 4 : 2;
    for (int sw = 0; sw < n_sweep; ++sw)
        for (int mu = 0; mu < 4; ++mu)
            for (int col = 0; col < ncol; ++col) {
                #pragma omp parallel for schedule(static)
                for (size_t s = 0; s < V; ++s)
                    if (link_colour(d, s, mu, schedule) == col) or_link(U, d, s, mu);
            }
}

void lto_hb_sweep(double *U, lto_dims d, double beta, int n_sweep, int schedule,
                  uint64_t seed, uint64_t *counter)
{
    size_t V = lto_volume(d);
    int ncol = schedule == LTO_CB4 ? 4 : 2;
    double prefactor = 3.0 / beta;   /* heatbath.hpp:39-40: N_col / beta */
    for (int sw = 0; sw < n_sweep; ++sw) {
        uint64_t c = (*counter)++;
        for (int mu = 0; mu < 4; ++mu)
            for (int col = 0; col < ncol; ++col) {
                #pragma omp parallel for schedule(static)
                for (size_t s = 0; s < V; ++s)
                    if (link_colour(d, s, mu, schedule) == col)
                        hb_link(U, d, s, mu, prefactor, 10, seed, c, NULL);
            }
    }
}

/* Sweeps with Rectangular<R>(beta, c_plaq, c_rect) (rectangular_action.hpp:29-60):
 * the action only enters through the staple; R = 2 needs the four-colour schedule. */
void lto_or_sweep_action(double *U, lto_dims d, double c_plaq, double c_rect, int n_sweep, int schedule)
{
    g_c_plaq = c_plaq;
    g_c_rect = c_rect;
    lto_or_sweep(U, d, n_sweep, schedule);
    g_c_plaq = 1.0;
    g_c_rect = 0.0;
}

void lto_hb_sweep_action(double *U, lto_dims d, double beta, double c_plaq, double c_rect, int n_sweep, int schedule,
                         uint64_t seed, uint64_t *counter)
{
    g_c_plaq = c_plaq;
    g_c_rect = c_rect;
    lto_hb_sweep(U, d, beta, n_sweep, schedule, seed, counter);
    g_c_plaq = 1.0;
    g_c_rect = 0.0;
}

void lto_hb_link(double *U, lto_dims d, double beta, size_t site, int mu, int max_iter,
                 uint64_t seed, uint64_t counter, int *attempts_out)
{
    hb_link(U, d, site, mu, 3.0 / beta, max_iter, seed, counter, attempts_out);
}

/* ------------------------------------------------------------------------- */
/* HMC pieces (LettuceGauge/updates/hmc_gauge.hpp) */

/* HMCKernel::RandomMomentum (hmc_gauge.hpp:400-428) with Philox Gaussians
 * (stream 1, draw blocks k=0..3 give phi1..phi8 in order) */
void lto_random_momentum(double *P, lto_dims d, uint64_t seed, uint64_t counter)
{
    size_t V = lto_volume(d);
    const double s3 = sqrt(3.0);
    #pragma omp parallel for
    for (size_t s = 0; s < V; ++s)
        for (int mu = 0; mu < 4; ++mu) {
            double phi[9];
            for (int k = 0; k < 4; ++k) {
                double u[2], g[2];
                lto_draw2(seed, (uint32_t)s, 1u, (uint32_t)k, 0u, (uint32_t)mu, counter, u);
                lto_gauss2(u, g);
                phi[1 + 2 * k] = g[0];
                phi[2 + 2 * k] = g[1];
            }
            mat3 A;
            A.m[0] = I * (phi[3] + phi[8] / s3);
            A.m[1] = phi[2] + I * phi[1];
            A.m[2] = phi[5] + I * phi[4];
            A.m[3] = -phi[2] + I * phi[1];
            A.m[4] = I * (-phi[3] + phi[8] / s3);
            A.m[5] = phi[7] + I * phi[6];
            A.m[6] = -phi[5] + I * phi[4];
            A.m[7] = -phi[7] + I * phi[6];
            A.m[8] = I * (-2.0 * phi[8] / s3);
            lto_set_link(P, s, mu, m_rscale(0.5, A));
        }
}

/* HMCKernel::UpdateMomenta (hmc_gauge.hpp:450-464): P += eps*beta/6*Algebra(st U^+) */
void lto_gauge_force_update(const double *U, double *P, lto_dims d, double eps, double beta)
{
    size_t V = lto_volume(d);
    #pragma omp parallel for
    for (size_t s = 0; s < V; ++s)
        for (int mu = 0; mu < 4; ++mu) {
            mat3 tmp = m_mul_nd(lto_staple_m(U, d, s, mu), lto_link(U, s, mu));
            mat3 p = m_add(lto_link(P, s, mu), m_rscale(eps * beta / 6.0, m_algebra(tmp)));
            lto_set_link(P, s, mu, p);
        }
}

/* HMCKernel::UpdateFields (hmc_gauge.hpp:505-523): U <- exp(-i*eps*P)... */
void lto_field_update(double *U, const double *P, lto_dims d, double eps)
{
    size_t n = 4 * lto_volume(d);
    #pragma omp parallel for
    for (size_t i = 0; i < n; ++i) {
        mat3 Q = m_scale(-I * eps, m_load(P + 18 * i));
        mat3 L = m_mul(lto_exp(Q), m_load(U + 18 * i));
        m_gram_schmidt(&L);
        m_store(U + 18 * i, L);
    }
}

/* kinetic part of HMCKernel::Hamiltonian (hmc_gauge.hpp:525-541) */
double lto_kinetic(const double *P, lto_dims d)
{
    size_t n = 4 * lto_volume(d);
    double k = 0.0;
    #pragma omp parallel for reduction(+:k)
    for (size_t i = 0; i < n; ++i) {
        mat3 p = m_load(P + 18 * i);
        k -= m_retr_mul(p, p);
    }
    return k;
}
