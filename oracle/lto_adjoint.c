/*
 * lto_adjoint.c -- CPU oracle (test infrastructure only): the analytic derivative
 * of the clover topological charge and its back-propagation through stout
 * smearing levels (the metadynamics force of LettuceGauge/updates/hmc_metadynamics.hpp).
 */
#include "lto_internal.h"
#include <omp.h>
#include <float.h>
#include <stdlib.h>

/* R_rhosigma(x) = C_rhosigma - C_sigmarho, with only rho<sigma stored in C6
 * (C_sigmarho = C_rhosigma^+), clover.hpp:246-253 */
static inline int plane_index(int a, int b)
{
    static const int P[4][4] = {{-1, 0, 1, 2}, {0, -1, 3, 4}, {1, 3, -1, 5}, {2, 4, 5, -1}};
    return P[a][b];
}
static inline mat3 R_at(const double *C6, size_t s, int rho, int sigma)
{
    int lo = rho < sigma ? rho : sigma, hi = rho < sigma ? sigma : rho;
    mat3 C = m_load(C6 + 18 * (6 * s + plane_index(lo, hi)));
    mat3 R = m_sub(C, m_dag(C));
    return rho < sigma ? R : m_rscale(-1.0, R);
}

/* CloverDerivativeComponent (clover.hpp:238-276) */
static mat3 clover_deriv_component(const double *U, const double *C6, lto_dims d, size_t s,
                                   int mu, int nu, int rho, int sigma)
{
    size_t smu = lto_move(d, s, mu, 1), snu = lto_move(d, s, nu, 1), snd = lto_move(d, s, nu, -1);
    size_t smu_nu = lto_move(d, smu, nu, 1), smu_nd = lto_move(d, smu, nu, -1);
    mat3 Rx = R_at(C6, s, rho, sigma), Rnu = R_at(C6, snu, rho, sigma), Rmunu = R_at(C6, smu_nu, rho, sigma);
    mat3 Rmu = R_at(C6, smu, rho, sigma), Rnd = R_at(C6, snd, rho, sigma), Rmund = R_at(C6, smu_nd, rho, sigma);
    mat3 Unu_x = lto_link(U, s, nu), Unu_mu = lto_link(U, smu, nu), Umu_nu = lto_link(U, snu, mu);
    mat3 Unu_mund = lto_link(U, smu_nd, nu), Umu_nd = lto_link(U, snd, mu), Unu_nd = lto_link(U, snd, nu);
    mat3 tmp1 = m_mul_nd(Unu_mu, Umu_nu);
    mat3 tmp2 = m_mul_nd(m_dag(Umu_nu), Unu_x);
    mat3 tmp3 = m_mul_nd(m_dag(Unu_mund), Umu_nd);
    mat3 tmp4 = m_mul_dn(Umu_nd, Unu_nd);
    mat3 a = m_mul(tmp1, m_add(m_mul_dn(Unu_x, Rx), m_mul_nd(Rnu, Unu_x)));
    mat3 b = m_mul(m_add(m_mul(Unu_mu, Rmunu), m_mul(Rmu, Unu_mu)), tmp2);
    mat3 c = m_mul(tmp3, m_add(m_mul(Unu_nd, Rx), m_mul(Rnd, Unu_nd)));
    mat3 e = m_mul(m_add(m_mul_dn(Unu_mund, Rmund), m_mul_nd(Rmu, Unu_mund)), tmp4);
    return m_sub(m_add(a, b), m_add(c, e));
}

/* CloverDerivative (clover.hpp:278-321): eps-tensor terms, common factor U_mu(x),
 * prefactor -2/(512 pi^2), Algebra projection */
static mat3 clover_derivative_link(const double *U, const double *C6, lto_dims d, size_t s, int mu)
{
    static const int TERMS[4][3][4] = {
        {{+1, 1, 2, 3}, {-1, 2, 1, 3}, {+1, 3, 1, 2}},
        {{-1, 0, 2, 3}, {+1, 2, 0, 3}, {-1, 3, 0, 2}},
        {{+1, 0, 1, 3}, {-1, 1, 0, 3}, {+1, 3, 0, 1}},
        {{-1, 0, 1, 2}, {+1, 1, 0, 2}, {-1, 2, 0, 1}}};
    mat3 acc = m_zero();
    for (int k = 0; k < 3; ++k) {
        const int *t = TERMS[mu][k];
        mat3 c = clover_deriv_component(U, C6, d, s, mu, t[1], t[2], t[3]);
        acc = t[0] > 0 ? m_add(acc, c) : m_sub(acc, c);
    }
    acc = m_mul(lto_link(U, s, mu), acc);
    return m_rscale(-2.0 / (512.0 * LTO_PI * LTO_PI), m_algebra(acc));
}

void lto_clover_derivative(const double *U, const double *C6, lto_dims d, double *F)
{
    size_t V = lto_volume(d);
    #pragma omp parallel for
    for (size_t s = 0; s < V; ++s)
        for (int mu = 0; mu < 4; ++mu) lto_set_link(F, s, mu, clover_derivative_link(U, C6, d, s, mu));
}

/* ExpDerivativeConstants (su3_exp.hpp:286-416), rConstants (:121-144),
 * bConstants (:146-201); only the pieces CalculateLambda needs */
typedef struct { cplx f1, f2; mat3 B1, B2; } lto_expd;

static void exp_derivative_constants(const lto_expc *e, lto_expd *o)
{
    double u = e->u, u2 = e->u2, w2 = e->w2;
    cplx xi0 = lto_xi0(e->w), xi1 = lto_xi1(e->w);
    cplx f0 = e->denom * e->h0, f1 = e->denom * e->h1, f2 = e->denom * e->h2;
    cplx em = e->exp_miu, e2 = e->exp_2iu, cw = e->cosw, ix = e->i_xi0;
    cplx r10 = 2.0 * (u + I * (u2 - w2)) * e2
             + 2.0 * em * (4.0 * u * (2.0 - I * u) * cw + ix * (9.0 * u2 + w2 - I * u * (3.0 * u2 + w2)));
    cplx r11 = 2.0 * (1.0 + 2.0 * I * u) * e2 + em * (-2.0 * (1.0 - I * u) * cw + ix * (6.0 * u + I * (w2 - 3.0 * u2)));
    cplx r12 = 2.0 * I * e2 + I * em * (cw - 3.0 * (1.0 - I * u) * xi0);
    cplx r20 = -2.0 * e2 + 2.0 * I * u * em * (cw + (1.0 + 4.0 * I * u) * xi0 + 3.0 * u2 * xi1);
    cplx r21 = -I * em * (cw + (1.0 + 2.0 * I * u) * xi0 - 3.0 * u2 * xi1);
    cplx r22 = em * (xi0 - 3.0 * I * u * xi1);
    cplx b10, b11, b12, b20, b21, b22;
    if (e->c0max < DBL_MIN) {
        b10 = b11 = b12 = b20 = b21 = b22 = 0.0;
    } else {
        double bd = 9.0 * u2 - w2;
        bd *= 2.0 * bd;
        bd = 1.0 / bd;
        double p = 3.0 * u2 - w2, q = 15.0 * u2 + w2;
        if (e->signflip) {
            b10 =  bd * conj(2.0 * u * r10 + p * r20 - 2.0 * q * conj(f0));
            b11 = -bd * conj(2.0 * u * r11 + p * r21 - 2.0 * q * -conj(f1));
            b12 =  bd * conj(2.0 * u * r12 + p * r22 - 2.0 * q * conj(f2));
            b20 = -bd * conj(r10 - 3.0 * u * r20 - 24.0 * u * conj(f0));
            b21 =  bd * conj(r11 - 3.0 * u * r21 - 24.0 * u * -conj(f1));
            b22 = -bd * conj(r12 - 3.0 * u * r22 - 24.0 * u * conj(f2));
        } else {
            b10 = bd * (2.0 * u * r10 + p * r20 - 2.0 * q * f0);
            b11 = bd * (2.0 * u * r11 + p * r21 - 2.0 * q * f1);
            b12 = bd * (2.0 * u * r12 + p * r22 - 2.0 * q * f2);
            b20 = bd * (r10 - 3.0 * u * r20 - 24.0 * u * f0);
            b21 = bd * (r11 - 3.0 * u * r21 - 24.0 * u * f1);
            b22 = bd * (r12 - 3.0 * u * r22 - 24.0 * u * f2);
        }
    }
    for (int i = 0; i < 9; ++i) {
        o->B1.m[i] = b11 * e->Q.m[i] + b12 * e->Q2.m[i];
        o->B2.m[i] = b21 * e->Q.m[i] + b22 * e->Q2.m[i];
    }
    for (int i = 0; i < 3; ++i) { o->B1.m[4*i] += b10; o->B2.m[4*i] += b20; }
    o->f1 = f1;
    o->f2 = f2;
}

/* StoutForceRecursion (stout_smearing.hpp:152-234) with CalculateLambda (:122-150).
 * U = level n-1, Uprev = level n, Sigma in/out, Qtape = exponent of level n-1 -> n.
 * Lambda is caller-provided scratch (the reference uses a static buffer, :170). */
void lto_stout_force_recursion(const double *U, const double *Uprev, double *Sigma,
                               const double *Qtape, lto_dims d, double rho, double *Lambda)
{
    size_t V = lto_volume(d);
    size_t nl = 4 * V;
    #pragma omp parallel for
    for (size_t i = 0; i < nl; ++i) {
        mat3 sp = m_mul_dn(m_load(Uprev + 18 * i), m_load(Sigma + 18 * i));
        m_store(Sigma + 18 * i, sp);
        lto_expc e;
        lto_exp_constants(m_load(Qtape + 18 * i), &e);
        lto_expd x;
        exp_derivative_constants(&e, &x);
        mat3 tmp = m_mul(m_load(U + 18 * i), sp);
        mat3 G = m_add(m_scale(m_tr(m_mul(x.B1, tmp)), e.Q), m_scale(m_tr(m_mul(x.B2, tmp)), e.Q2));
        G = m_add(G, m_scale(x.f1, tmp));
        G = m_add(G, m_scale(x.f2, m_add(m_mul(e.Q, tmp), m_mul(tmp, e.Q))));
        m_store(Lambda + 18 * i, m_tl_herm(G));
    }
    double *out = (double *)malloc(sizeof(double) * 18 * nl);
    #pragma omp parallel for
    for (size_t s = 0; s < V; ++s)
        for (int mu = 0; mu < 4; ++mu) {
            mat3 fs = m_zero();
            size_t smu = lto_move(d, s, mu, 1);
            for (int nu = 0; nu < 4; ++nu) {
                if (nu == mu) continue;
                size_t snu = lto_move(d, s, nu, 1), snd = lto_move(d, s, nu, -1);
                size_t smu_nd = lto_move(d, smu, nu, -1);
                mat3 Unu_x = lto_link(U, s, nu), Unu_mu = lto_link(U, smu, nu), Umu_nu = lto_link(U, snu, mu);
                mat3 Umu_nd = lto_link(U, snd, mu), Unu_nd = lto_link(U, snd, nu), Unu_mund = lto_link(U, smu_nd, nu);
                mat3 tmp1 = m_mul_nd(Unu_mu, Umu_nu);
                mat3 tmp2 = m_mul_dn(Umu_nd, m_sub(lto_link(Lambda, snd, mu), lto_link(Lambda, snd, nu)));
                mat3 a = m_mul(tmp1, m_add(m_mul_dn(Unu_x, lto_link(Lambda, s, nu)),
                                           m_mul_nd(lto_link(Lambda, snu, mu), Unu_x)));
                mat3 b = m_mul(m_mul_dn(Unu_mund, m_add(tmp2, m_mul_nd(lto_link(Lambda, smu_nd, nu), Umu_nd))), Unu_nd);
                mat3 c = m_mul_nd(m_mul(lto_link(Lambda, smu, nu), tmp1), Unu_x);
                fs = m_add(fs, m_sub(m_add(a, b), c));
            }
            mat3 Cadj = m_rscale(rho, m_dag(lto_staple_m(U, d, s, mu)));
            lto_expc e;
            lto_exp_constants(lto_link(Qtape, s, mu), &e);
            mat3 inner = m_mul(lto_link(Sigma, s, mu), lto_exp_from(&e));
            inner = m_add(inner, m_scale(I, m_mul(Cadj, lto_link(Lambda, s, mu))));
            inner = m_sub(inner, m_scale(I * rho, fs));
            m_store(out + 18 * (4 * s + mu), m_algebra(m_mul(lto_link(U, s, mu), inner)));
        }
    memcpy(Sigma, out, sizeof(double) * 18 * nl);
    free(out);
}

/* HMCMetaDKernel::MetaCharge (hmc_metadynamics.hpp:302-310) */
double lto_cv_charge(const double *U, lto_dims d, double rho, int n_levels)
{
    size_t n = 18 * 4 * lto_volume(d);
    double *a = (double *)malloc(sizeof(double) * n), *b = (double *)malloc(sizeof(double) * n);
    memcpy(a, U, sizeof(double) * n);
    for (int k = 0; k < n_levels; ++k) { lto_stout(a, b, d, rho); double *t = a; a = b; b = t; }
    double Q = lto_topcharge_clover(a, d);
    free(a); free(b);
    return Q;
}

/* HMCMetaDKernel::CalculateTopologicalForce (hmc_metadynamics.hpp:374-436):
 * F = dV/dQ * dQ/dU through the clover derivative and n stout recursions.
 * Returns the CV (clover Q of the most smeared field, FullTensor route). */
double lto_cv_force(const double *U, lto_dims d, double rho, int n_levels, double dvdq, double *F)
{
    size_t V = lto_volume(d);
    size_t n = 18 * 4 * V;
    double *lev = (double *)malloc(sizeof(double) * n * (n_levels + 1));
    double *tape = (double *)malloc(sizeof(double) * n * (n_levels > 0 ? n_levels : 1));
    double *C6 = (double *)malloc(sizeof(double) * 18 * 6 * V);
    double *Lam = (double *)malloc(sizeof(double) * n);
    memcpy(lev, U, sizeof(double) * n);
    for (int k = 0; k < n_levels; ++k)
        lto_stout_tape(lev + k * n, lev + (k + 1) * n, tape + k * n, d, rho);
    const double *top = lev + (size_t)n_levels * n;
    lto_clover(top, d, C6);
    double Q = lto_topcharge_clover_tensor(C6, d);
    lto_clover_derivative(top, C6, d, F);
    #pragma omp parallel for
    for (size_t i = 0; i < n; ++i) F[i] *= dvdq;
    for (int k = n_levels; k > 0; --k)
        lto_stout_force_recursion(lev + (k - 1) * n, lev + k * n, F, tape + (k - 1) * n, d, rho, Lam);
    free(lev); free(tape); free(C6); free(Lam);
    return Q;
}
