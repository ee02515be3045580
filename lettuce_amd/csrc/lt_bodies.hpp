// lt_bodies.hpp -- per-thread bodies of the HMC and collective-variable kernels.
// Single-lattice kernels and the replica-batched kernels call the same bodies;
// link bodies take the logical block index b of their lattice (8 consecutive
// blocks = the (mu, parity) classes of one tile), site bodies take the site.
#pragma once

#include "lt_device.hpp"

namespace lt {

// ---------------------------------------------------------------- HMC

// P = 1/2 * A with A the anti-hermitian combination of the eight generators used
// by the reference's RandomMomentum; phi_1..phi_8 come from Philox stream
// RNG_MOMENTUM, draw blocks k = 0..3 (Box-Muller pairs).
__device__ __forceinline__ void body_momentum(double2* __restrict__ P, const Geom& g, const LinkMap& m, int b,
                                              uint64_t seed, uint64_t counter)
{
    int mu, par, c[4], s;
    if (!map_link_b(g, m, b, mu, par, c, s)) return;
    double phi[9];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double u0, u1, g0, g1;
        draw2(seed, (uint32_t)s, RNG_MOMENTUM, (uint32_t)k, 0u, (uint32_t)mu, counter, u0, u1);
        gauss2(u0, u1, g0, g1);
        phi[1 + 2 * k] = g0;
        phi[2 + 2 * k] = g1;
    }
    const double s3 = sqrt(3.0);
    M3 A;
    A.m[0] = cd{0.0, phi[3] + phi[8] / s3};
    A.m[1] = cd{phi[2], phi[1]};
    A.m[2] = cd{phi[5], phi[4]};
    A.m[3] = cd{-phi[2], phi[1]};
    A.m[4] = cd{0.0, -phi[3] + phi[8] / s3};
    A.m[5] = cd{phi[7], phi[6]};
    A.m[6] = cd{-phi[5], phi[4]};
    A.m[7] = cd{-phi[7], phi[6]};
    A.m[8] = cd{0.0, -2.0 * phi[8] / s3};
    st_link(P, g, s, mu, par, scale(0.5, A));
}

// P += a * Algebra(st U^+) + bf * F   (F optional: metadynamics force)
__device__ __forceinline__ void body_force_update(const double2* __restrict__ U, double2* __restrict__ P,
                                                  const double2* __restrict__ F, const Geom& g, const LinkMap& m, int b,
                                                  double a, double bf)
{
    int mu, par, c[4], s;
    if (!map_link_b(g, m, b, mu, par, c, s)) return;
    M3 st = staple(U, g, s, c, mu);
    M3 A = algebra(mul_nd(st, ldu(U, g, s, mu, par)));
    M3 p = ld(P, g, s, mu, par);
    if (F) {
        M3 f = ld(F, g, s, mu, par);
#pragma unroll
        for (int k = 0; k < 9; ++k) p.m[k] += a * A.m[k] + bf * f.m[k];
    } else {
#pragma unroll
        for (int k = 0; k < 9; ++k) p.m[k] += a * A.m[k];
    }
    st_link(P, g, s, mu, par, p);
}

// U <- exp(eps P) U, Gram-Schmidt (local per link, in place)
__device__ __forceinline__ void body_field_update(double2* __restrict__ U, const double2* __restrict__ P, const Geom& g,
                                                  const LinkMap& m, int b, double eps)
{
    int mu, par, c[4], s;
    if (!map_link_b(g, m, b, mu, par, c, s)) return;
    M3 W = mul(exp_ah(ld(P, g, s, mu, par), eps), ldu(U, g, s, mu, par));
    gram_schmidt(W);
    stu(U, g, s, mu, par, W);
}

// sum over the 6 planes at site s of Re tr P_munu
__device__ __forceinline__ double body_plaq_site(const double2* __restrict__ U, const Geom& g, int s)
{
    int c[4];
    coords_of(g, s, c);
    const int par = parity_of(c);
    double acc = 0.0;
    for (int nu = 1; nu < 4; ++nu) {
        const M3 Unu = ldu(U, g, s, nu, par);
        const int snu = up(g, s, c[nu], nu);
        for (int mu = 0; mu < nu; ++mu) {
            const int smu = up(g, s, c[mu], mu);
            M3 l = mul_nd(mul(ldu(U, g, s, mu, par), ldu(U, g, smu, nu, par ^ 1)), ldu(U, g, snu, mu, par ^ 1));
            acc += retr_mul(l, dag(Unu));
        }
    }
    return acc;
}

// -Re tr(P P) of link id = 4*s + mu
__device__ __forceinline__ double body_kinetic_link(const double2* __restrict__ P, const Geom& g, long long id)
{
    const int s = (int)(id >> 2), mu = (int)(id & 3);
    int c[4];
    coords_of(g, s, c);
    M3 p = ld(P, g, s, mu, parity_of(c));
    return -retr_mul(p, p);
}

// ------------------------------------------------- collective variable

// plane index of (a, b), a < b: 01,02,03,12,13,23
__device__ __forceinline__ int plane_of(int a, int b)
{
    return a == 0 ? b - 1 : (a == 1 ? b + 1 : 5);
}

__device__ __forceinline__ M3 ld_plane(const double2* __restrict__ C, const Geom& g, int s, int plane, int par)
{
    const double2* p = C + (size_t)((plane * 2 + par) * 9) * g.Vh + (s >> 1);
    M3 r;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        double2 v = p[(size_t)k * g.Vh];
        r.m[k] = cd{v.x, v.y};
    }
    return r;
}
__device__ __forceinline__ void st_plane(double2* __restrict__ C, const Geom& g, int s, int plane, int par, const M3& a)
{
    double2* p = C + (size_t)((plane * 2 + par) * 9) * g.Vh + (s >> 1);
#pragma unroll
    for (int k = 0; k < 9; ++k) p[(size_t)k * g.Vh] = make_double2(a.m[k].re, a.m[k].im);
}

// R_rhosigma(y) = C_rhosigma - C_sigmarho with C_sigmarho = C_rhosigma^+
__device__ __forceinline__ M3 R_at(const double2* __restrict__ C, const Geom& g, int y, int ypar, int rho, int sigma)
{
    const int lo = rho < sigma ? rho : sigma, hi = rho < sigma ? sigma : rho;
    M3 c = ld_plane(C, g, y, plane_of(lo, hi), ypar);
    M3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.m[3 * i + j] = c.m[3 * i + j] - conj(c.m[3 * j + i]);
    return rho < sigma ? r : scale(-1.0, r);
}

// stout level with tape: Q = -i rho Algebra(st U^+), U' = exp(iQ) U, Gram-Schmidt
__device__ __forceinline__ void body_stout_tape(const double2* __restrict__ Uin, double2* __restrict__ Uout,
                                                double2* __restrict__ T, const Geom& g, const LinkMap& m, int b,
                                                double rho)
{
    int mu, par, c[4], s;
    if (!map_link_b(g, m, b, mu, par, c, s)) return;
    M3 st = staple(Uin, g, s, c, mu);
    M3 L = ldu(Uin, g, s, mu, par);
    M3 A = algebra(mul_nd(st, L));
    M3 Q;
#pragma unroll
    for (int k = 0; k < 9; ++k) Q.m[k] = cd{rho * A.m[k].im, -rho * A.m[k].re};
    st_link(T, g, s, mu, par, Q);
    M3 W = mul(su3_exp(Q), L);
    gram_schmidt(W);
    stu(Uout, g, s, mu, par, W);
}

// the clover term of one plane at site s
__device__ __forceinline__ void body_clover_store(const double2* __restrict__ U, double2* __restrict__ C, const Geom& g,
                                                  int s, int plane)
{
    int c[4];
    coords_of(g, s, c);
    const int par = parity_of(c);
    double pl;
    M3 cl;
    switch (plane) {
        case 0: cl = clover_plane(U, g, s, c, par, 0, 1, pl); break;
        case 1: cl = clover_plane(U, g, s, c, par, 0, 2, pl); break;
        case 2: cl = clover_plane(U, g, s, c, par, 0, 3, pl); break;
        case 3: cl = clover_plane(U, g, s, c, par, 1, 2, pl); break;
        case 4: cl = clover_plane(U, g, s, c, par, 1, 3, pl); break;
        default: cl = clover_plane(U, g, s, c, par, 2, 3, pl); break;
    }
    st_plane(C, g, s, plane, par, cl);
}

// Re tr(F01 F23 - F02 F13 + F03 F12) at site s, F = C - C^+
__device__ __forceinline__ double body_qdensity_site(const double2* __restrict__ C, const Geom& g, int s)
{
    int c[4];
    coords_of(g, s, c);
    const int par = parity_of(c);
    return retr_mul(R_at(C, g, s, par, 0, 1), R_at(C, g, s, par, 2, 3))
         - retr_mul(R_at(C, g, s, par, 0, 2), R_at(C, g, s, par, 1, 3))
         + retr_mul(R_at(C, g, s, par, 0, 3), R_at(C, g, s, par, 1, 2));
}

// One epsilon-term of the clover-charge derivative for link (x, mu):
//   tmp1 (U_nu(x)^+ R(x) + R(x+nu) U_nu(x)^+) + (U_nu(x+mu) R(x+mu+nu) + R(x+mu) U_nu(x+mu)) tmp2
//   - tmp3 (U_nu(x-nu) R(x) + R(x-nu) U_nu(x-nu)) - (U_nu(x+mu-nu)^+ R(x+mu-nu) + R(x+mu) U_nu(x+mu-nu)^+) tmp4
// with tmp1 = U_nu(x+mu) U_mu(x+nu)^+, tmp2 = U_mu(x+nu)^+ U_nu(x)^+,
//      tmp3 = U_nu(x+mu-nu)^+ U_mu(x-nu)^+, tmp4 = U_mu(x-nu)^+ U_nu(x-nu).
__device__ __forceinline__ M3 cderiv_term(const double2* __restrict__ U, const double2* __restrict__ C, const Geom& g,
                                          int s, const int c[4], int par, int mu, int nu, int rho, int sigma)
{
    const int q = par ^ 1;
    const int smu = up(g, s, c[mu], mu), snu = up(g, s, c[nu], nu), snd = dn(g, s, c[nu], nu);
    const int smu_nu = up(g, smu, c[nu], nu), smu_nd = dn(g, smu, c[nu], nu);
    const M3 Unu_x = ldu(U, g, s, nu, par), Unu_mu = ldu(U, g, smu, nu, q), Umu_nu = ldu(U, g, snu, mu, q);
    const M3 Rmu = R_at(C, g, smu, q, rho, sigma);
    M3 acc = mul(mul_nd(Unu_mu, Umu_nu), add(mul_dn(Unu_x, R_at(C, g, s, par, rho, sigma)),
                                               mul_nd(R_at(C, g, snu, q, rho, sigma), Unu_x)));
    {
        M3 a = add(mul(Unu_mu, R_at(C, g, smu_nu, par, rho, sigma)), mul(Rmu, Unu_mu));
        M3 tmp2 = mul_nd(dag(Umu_nu), Unu_x);
        acc = add(acc, mul(a, tmp2));
    }
    const M3 Unu_mund = ldu(U, g, smu_nd, nu, par), Umu_nd = ldu(U, g, snd, mu, q), Unu_nd = ldu(U, g, snd, nu, q);
    {
        M3 tmp3 = mul_nd(dag(Unu_mund), Umu_nd);
        M3 bb = add(mul(Unu_nd, R_at(C, g, s, par, rho, sigma)), mul(R_at(C, g, snd, q, rho, sigma), Unu_nd));
        acc = sub(acc, mul(tmp3, bb));
    }
    {
        M3 tmp4 = mul_dn(Umu_nd, Unu_nd);
        M3 e = add(mul_dn(Unu_mund, R_at(C, g, smu_nd, par, rho, sigma)), mul_nd(Rmu, Unu_mund));
        acc = sub(acc, mul(e, tmp4));
    }
    return acc;
}

// F(x, mu) = scale * (-2/(512 pi^2)) Algebra(U_mu(x) * sum of the three epsilon terms)
__device__ __forceinline__ void body_clover_deriv(const double2* __restrict__ U, const double2* __restrict__ C,
                                                  double2* __restrict__ F, const Geom& g, const LinkMap& m, int b,
                                                  double scale_factor)
{
    int mu, par, c[4], s;
    if (!map_link_b(g, m, b, mu, par, c, s)) return;
    M3 acc;
    switch (mu) {
        case 0:
            acc = cderiv_term(U, C, g, s, c, par, 0, 1, 2, 3);
            acc = sub(acc, cderiv_term(U, C, g, s, c, par, 0, 2, 1, 3));
            acc = add(acc, cderiv_term(U, C, g, s, c, par, 0, 3, 1, 2));
            break;
        case 1:
            acc = scale(-1.0, cderiv_term(U, C, g, s, c, par, 1, 0, 2, 3));
            acc = add(acc, cderiv_term(U, C, g, s, c, par, 1, 2, 0, 3));
            acc = sub(acc, cderiv_term(U, C, g, s, c, par, 1, 3, 0, 2));
            break;
        case 2:
            acc = cderiv_term(U, C, g, s, c, par, 2, 0, 1, 3);
            acc = sub(acc, cderiv_term(U, C, g, s, c, par, 2, 1, 0, 3));
            acc = add(acc, cderiv_term(U, C, g, s, c, par, 2, 3, 0, 1));
            break;
        default:
            acc = scale(-1.0, cderiv_term(U, C, g, s, c, par, 3, 0, 1, 2));
            acc = add(acc, cderiv_term(U, C, g, s, c, par, 3, 1, 0, 2));
            acc = sub(acc, cderiv_term(U, C, g, s, c, par, 3, 2, 0, 1));
            break;
    }
    acc = mul(ldu(U, g, s, mu, par), acc);
    const double pref = -2.0 / (512.0 * 3.14159265358979323846 * 3.14159265358979323846);
    st_link(F, g, s, mu, par, scale(scale_factor * pref, algebra(acc)));
}

// Sigma' = Uprev^+ Sigma (stored in place) and
// Lambda = TracelessHermitian(tr(B1 G) Q + tr(B2 G) Q^2 + f1 G + f2 (Q G + G Q)), G = U Sigma'
__device__ __forceinline__ void body_lambda(const double2* __restrict__ U, const double2* __restrict__ Uprev,
                                            double2* __restrict__ Sig, const double2* __restrict__ T,
                                            double2* __restrict__ Lam, const Geom& g, const LinkMap& m, int b)
{
    int mu, par, c[4], s;
    if (!map_link_b(g, m, b, mu, par, c, s)) return;
    M3 sp = mul_dn(ldu(Uprev, g, s, mu, par), ld(Sig, g, s, mu, par));
    st_link(Sig, g, s, mu, par, sp);
    M3 Q = ld(T, g, s, mu, par);
    M3 Q2 = mul(Q, Q);
    ExpC e;
    exp_constants(Q, Q2, e);
    M3 B1, B2;
    cd f1, f2;
    exp_deriv(e, Q, Q2, B1, B2, f1, f2);
    M3 G = mul(ldu(U, g, s, mu, par), sp);
    cd t1 = trace(mul(B1, G)), t2 = trace(mul(B2, G));
    M3 QG = add(mul(Q, G), mul(G, Q));
    M3 Gam;
#pragma unroll
    for (int k = 0; k < 9; ++k) Gam.m[k] = t1 * Q.m[k] + t2 * Q2.m[k] + f1 * G.m[k] + f2 * QG.m[k];
    st_link(Lam, g, s, mu, par, tl_herm(Gam));
}

// Sigma_new = Algebra(U (Sigma' exp(iQ) + i rho st^+ Lambda - i rho sum_nu [six-term staple of Lambda]))
__device__ __forceinline__ void body_sigma(const double2* __restrict__ U, double2* __restrict__ Sig,
                                           const double2* __restrict__ T, const double2* __restrict__ Lam,
                                           const Geom& g, const LinkMap& m, int b, double rho)
{
    int mu, par, c[4], s;
    if (!map_link_b(g, m, b, mu, par, c, s)) return;
    const int q = par ^ 1;
    const int smu = up(g, s, c[mu], mu);
    M3 fs = m3_zero();
#pragma unroll
    for (int nu = 0; nu < 4; ++nu) {
        if (nu == mu) continue;
        const int snu = up(g, s, c[nu], nu), snd = dn(g, s, c[nu], nu);
        const int smu_nd = dn(g, smu, c[nu], nu);
        const M3 Unu_x = ldu(U, g, s, nu, par);
        const M3 Unu_mu = ldu(U, g, smu, nu, q);
        const M3 tmp1 = mul_nd(Unu_mu, ldu(U, g, snu, mu, q));
        // rows 1 and 6: tmp1 (U_nu(x)^+ L_nu(x) + L_mu(x+nu) U_nu(x)^+)
        acc(fs, mul(tmp1, add(mul_dn(Unu_x, ld(Lam, g, s, nu, par)), mul_nd(ld(Lam, g, snu, mu, q), Unu_x))));
        // row 5: - L_nu(x+mu) tmp1 U_nu(x)^+
        fs = sub(fs, mul_nd(mul(ld(Lam, g, smu, nu, q), tmp1), Unu_x));
        // rows 2, 4, 3: U_nu(x+mu-nu)^+ (U_mu(x-nu)^+ (L_mu(x-nu) - L_nu(x-nu)) + L_nu(x+mu-nu) U_mu(x-nu)^+) U_nu(x-nu)
        const M3 Umu_nd = ldu(U, g, snd, mu, q);
        M3 tmp2 = mul_dn(Umu_nd, sub(ld(Lam, g, snd, mu, q), ld(Lam, g, snd, nu, q)));
        M3 inner = add(tmp2, mul_nd(ld(Lam, g, smu_nd, nu, par), Umu_nd));
        acc(fs, mul(mul_dn(ldu(U, g, smu_nd, nu, par), inner), ldu(U, g, snd, nu, q)));
    }
    M3 st = staple(U, g, s, c, mu);
    M3 Q = ld(T, g, s, mu, par);
    M3 inner = mul(ld(Sig, g, s, mu, par), su3_exp(Q));
    M3 cl = mul(dag(st), ld(Lam, g, s, mu, par));   // st^+ Lambda
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        // + i rho (st^+ Lambda) - i rho fs
        cd d = cl.m[k] - fs.m[k];
        inner.m[k] += cd{-rho * d.im, rho * d.re};
    }
    st_link(Sig, g, s, mu, par, algebra(mul(ldu(U, g, s, mu, par), inner)));
}

}  // namespace lt
