// lt_bodies.hpp -- per-thread bodies of the HMC and collective-variable kernels.
// Single-lattice kernels and the replica-batched kernels call the same bodies;
// link bodies take the logical block index b of their lattice (8 consecutive
// blocks = the (mu, parity) classes of one tile), site bodies take the site.
#pragma once

#include "lt_device.hpp"

// the MetaD stout tape kernel with the software-pipelined staple
#ifndef LT_MD_PIPE
#define LT_MD_PIPE 1
#endif

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

// The clover planes are kept as R_munu = C_munu - C_munu^+ (mu < nu), which is all
// the charge and its derivative use.  R is anti-hermitian: 9 reals stored exactly,
//   R[((plane*2 + parity)*5 + k)*Vh + cb],
//   k: 0 r01, 1 r02, 2 r12, 3 (Im r00, Im r11), 4 (Im r22, 0);
// Re r_ii = 0 and r_ji = -conj(r_ij) hold exactly for C - C^+ in floating point.
__device__ __forceinline__ M3 ld_plane(const double2* __restrict__ R, const Geom& g, int s, int plane, int par)
{
    const double2* p = R + slot_base(g, plane * 2 + par, R_SLOTS, cbi(g, s));
    const double2 a01 = p[0], a02 = p[slot_stride(g)], a12 = p[(size_t)2 * slot_stride(g)];
    const double2 d01 = p[(size_t)3 * slot_stride(g)], d2 = p[(size_t)4 * slot_stride(g)];
    M3 r;
    r.m[0] = cd{0.0, d01.x};
    r.m[4] = cd{0.0, d01.y};
    r.m[8] = cd{0.0, d2.x};
    r.m[1] = cd{a01.x, a01.y};
    r.m[2] = cd{a02.x, a02.y};
    r.m[5] = cd{a12.x, a12.y};
    r.m[3] = cd{-a01.x, a01.y};
    r.m[6] = cd{-a02.x, a02.y};
    r.m[7] = cd{-a12.x, a12.y};
    return r;
}
// stores R = C - C^+ of the clover term C
__device__ __forceinline__ void st_plane(double2* __restrict__ R, const Geom& g, int s, int plane, int par, const M3& c)
{
    double2* p = R + slot_base(g, plane * 2 + par, R_SLOTS, cbi(g, s));
    const cd r01 = c.m[1] - conj(c.m[3]), r02 = c.m[2] - conj(c.m[6]), r12 = c.m[5] - conj(c.m[7]);
    p[0] = make_double2(r01.re, r01.im);
    p[slot_stride(g)] = make_double2(r02.re, r02.im);
    p[(size_t)2 * slot_stride(g)] = make_double2(r12.re, r12.im);
    p[(size_t)3 * slot_stride(g)] = make_double2(c.m[0].im + c.m[0].im, c.m[4].im + c.m[4].im);
    p[(size_t)4 * slot_stride(g)] = make_double2(c.m[8].im + c.m[8].im, 0.0);
}

// R_rhosigma(y) = C_rhosigma - C_sigmarho with C_sigmarho = C_rhosigma^+
__device__ __forceinline__ M3 R_at(const double2* __restrict__ C, const Geom& g, int y, int ypar, int rho, int sigma)
{
    const int lo = rho < sigma ? rho : sigma, hi = rho < sigma ? sigma : rho;
    M3 r = ld_plane(C, g, y, plane_of(lo, hi), ypar);
    return rho < sigma ? r : scale(-1.0, r);
}

// stout level with tape: Q = -i rho Algebra(st U^+), U' = exp(iQ) U, Gram-Schmidt
__device__ __forceinline__ void body_stout_tape(const double2* __restrict__ Uin, double2* __restrict__ Uout,
                                                double2* __restrict__ T, const Geom& g, const LinkMap& m, int b,
                                                double rho)
{
    int mu, par, c[4], s;
    if (!map_link_b(g, m, b, mu, par, c, s)) return;
#if LT_MD_PIPE
    Raw6 own;
    const M3 st = staple_pipelined(Uin, g, s, c, mu, par, own);
    const M3 L = su3_of(own);
#else
    M3 st = staple(Uin, g, s, c, mu);
    M3 L = ldu(Uin, g, s, mu, par);
#endif
    M3 A = algebra(mul_nd(st, L));
    M3 Q;
#pragma unroll
    for (int k = 0; k < 9; ++k) Q.m[k] = cd{rho * A.m[k].im, -rho * A.m[k].re};
    Q = her8_round(Q);   // forward and adjoint use exactly the stored exponent
    st_her8(T, g, s, mu, par, Q);
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

// Clover-charge derivative for link (x, mu) (CloverDerivativeComponent,
// LettuceGauge/observables/clover.hpp:238-321): for each of the three nu != mu,
// with (rho, sigma) the remaining directions and s = eps(mu, nu, rho, sigma),
//   upper: tmp1 (U_nu(x)^+ R(x) + R(x+nu) U_nu(x)^+) + (U_nu(x+mu) R(x+mu+nu) + R(x+mu) U_nu(x+mu)) tmp2
//   lower: tmp3 (U_nu(x-nu) R(x) + R(x-nu) U_nu(x-nu)) + (U_nu(x+mu-nu)^+ R(x+mu-nu) + R(x+mu) U_nu(x+mu-nu)^+) tmp4
// with tmp1 = U_nu(x+mu) U_mu(x+nu)^+, tmp2 = U_mu(x+nu)^+ U_nu(x)^+,
//      tmp3 = U_nu(x+mu-nu)^+ U_mu(x-nu)^+, tmp4 = U_mu(x-nu)^+ U_nu(x-nu),
// accumulate sum_nu s (upper - lower) and end with
//   F(x, mu) = scale * (-2/(512 pi^2)) Algebra(U_mu(x) * sum).
// One half-term alone needs ~210 VGPRs, so the whole sum straight through one thread
// spills; body_clover_deriv_all runs the halves as a loop with the running sum and one
// partial product parked in LDS.

// neighbour steps in a direction known only at run time (selects instead of indexing
// the extents, which would put them in scratch)
__device__ __forceinline__ int up_rt(const Geom& g, int s, int cm, int d)
{
    const int L = pick4(g.L, d), st = pick4(g.st, d);
    return (cm + 1 < L) ? s + st : s - (L - 1) * st;
}
__device__ __forceinline__ int dn_rt(const Geom& g, int s, int cm, int d)
{
    const int L = pick4(g.L, d), st = pick4(g.st, d);
    return (cm > 0) ? s - st : s + (L - 1) * st;
}

// Re tr of the plaquette U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+ (plaquette.hpp:20-24)
// as Re tr(A B^+) with A = U_mu(x) U_nu(x+mu), B = U_nu(x) U_mu(x+nu), both in SU(3), of
// which only rows 0 and 1 are multiplied out (retr_su3_nd_r01).
__device__ __forceinline__ double plaq_plane(const double2* __restrict__ U, const Geom& g, int s, const int c[4],
                                            int par, int mu, int nu)
{
    const int smu = up_rt(g, s, pick4(c, mu), mu), snu = up_rt(g, s, pick4(c, nu), nu);
    const M3 A = mul_r01(ldu2(U, g, s, mu, par), ldu(U, g, smu, nu, par ^ 1));
    const M3 B = mul_r01(ldu2(U, g, s, nu, par), ldu(U, g, snu, mu, par ^ 1));
    return retr_su3_nd_r01(A, B);
}
// The plaquettes of link (x, mu) in a launch over the field's link map: the planes
// (mu, mu+1 mod 4) and, for mu < 2, (mu, mu+2) -- every plane of every site once -- in the
// tile order of the link kernels, so that the neighbour links a tile reads are still in
// the L2 when the neighbouring tiles read them.
__device__ __forceinline__ double body_plaq_link(const double2* __restrict__ U, const Geom& g, const LinkMap& m, int b)
{
    int mu, par, c[4], s;
    if (!map_link_b(g, m, b, mu, par, c, s)) return 0.0;
    double v = plaq_plane(U, g, s, c, par, mu, (mu + 1) & 3);
    if (mu < 2) v += plaq_plane(U, g, s, c, par, mu, mu + 2);
    return v;
}
// block sums of body_plaq_link, part[blockIdx.x]: a butterfly in each wavefront, then the
// m.nthreads / 64 wavefront sums in wave order
__device__ __forceinline__ void plaq_link_partials(const double2* __restrict__ U, const Geom& g, const LinkMap& m,
                                                   double* part)
{
    __shared__ double sh[LT_LINK_BS / 64];
    double v = body_plaq_link(U, g, m, link_block(m, blockIdx.x, gridDim.x));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (m.nthreads >> 6); ++w) v += sh[w];
        part[blockIdx.x] = v;
    }
}

// Parking slot (park_store / park_load): the first of the two products of a half waits
// in LDS while the second is formed, so it holds no registers meanwhile.  park is never
// null in the kernels: testing it anyway puts the parking in a branch of its own, and the
// compiler does not hoist the second product's link loads above a branch (19 spilled
// VGPRs in clover_deriv_all_kernel without the tests, 2 with them).

__device__ __forceinline__ M3 cderiv_upper(const double2* __restrict__ U, const double2* __restrict__ C, const Geom& g,
                                           int s, const int c[4], int par, int MU, int NU, int RHO, int SIGMA,
                                           double2* park, int n)
{
    const int q = par ^ 1;
    const int smu = up_rt(g, s, pick4(c, MU), MU), snu = up_rt(g, s, pick4(c, NU), NU);
    const M3 Unu_x = ldu(U, g, s, NU, par);
    M3 tmp1, tmp2;
    {
        const M3 Umu_nu = ldu(U, g, snu, MU, q);
        tmp2 = mul_nd(dag(Umu_nu), Unu_x);
        tmp1 = mul_nd(ldu(U, g, smu, NU, q), Umu_nu);
    }
    M3 S = mul_dn(Unu_x, R_at(C, g, s, par, RHO, SIGMA));
    S = add(S, mul_nd(R_at(C, g, snu, q, RHO, SIGMA), Unu_x));
    M3 acc = mul(tmp1, S);
    if (park) park_store(park, n, acc);
    const M3 Unu_mu = ldu(U, g, smu, NU, q);
    S = mul(Unu_mu, R_at(C, g, up_rt(g, smu, pick4(c, NU), NU), par, RHO, SIGMA));
    S = add(S, mul(R_at(C, g, smu, q, RHO, SIGMA), Unu_mu));
    const M3 second = mul(S, tmp2);
    if (park) acc = park_load(park, n);
    return add(acc, second);
}

__device__ __forceinline__ M3 cderiv_lower(const double2* __restrict__ U, const double2* __restrict__ C, const Geom& g,
                                           int s, const int c[4], int par, int MU, int NU, int RHO, int SIGMA,
                                           double2* park, int n)
{
    const int q = par ^ 1;
    const int smu = up_rt(g, s, pick4(c, MU), MU);
    const int snd = dn_rt(g, s, pick4(c, NU), NU), smu_nd = dn_rt(g, smu, pick4(c, NU), NU);
    const M3 Unu_nd = ldu(U, g, snd, NU, q);
    const M3 Unu_mund = ldu(U, g, smu_nd, NU, par);
    M3 tmp3, tmp4;
    {
        const M3 Umu_nd = ldu(U, g, snd, MU, q);
        tmp3 = mul_nd(dag(Unu_mund), Umu_nd);
        tmp4 = mul_dn(Umu_nd, Unu_nd);
    }
    M3 S = mul(Unu_nd, R_at(C, g, s, par, RHO, SIGMA));
    S = add(S, mul(R_at(C, g, snd, q, RHO, SIGMA), Unu_nd));
    M3 acc = mul(tmp3, S);
    if (park) park_store(park, n, acc);
    S = mul_dn(Unu_mund, R_at(C, g, smu_nd, par, RHO, SIGMA));
    S = add(S, mul_nd(R_at(C, g, smu, q, RHO, SIGMA), Unu_mund));
    const M3 second = mul(S, tmp4);
    if (park) acc = park_load(park, n);
    return add(acc, second);
}

// All six half terms of link (x, mu) in one pass: upper then lower half of the k-th
// direction nu != mu, k = 0, 1, 2, summed in that order.  The running sum waits in LDS
// (acc: entries e * nthreads + lane, e < 9) and the first product of each half in the
// parking slot after it (acc + 9 * nthreads), so that each half has the registers to
// itself; the empty asm statements with memory clobbers keep the compiler from carrying
// them in registers instead.  The directions are loop variables (one copy of the code
// for every mu and k).  LDS: 18 x 16 bytes per lane of the block.
__device__ __forceinline__ void lds_acc(double2* acc, int n, const M3& a, int first, int negate)
{
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        const cd v = a.m[e];
        if (first) {
            acc[e * n] = make_double2(v.re, v.im);
        } else {
            const double2 o = acc[e * n];
            acc[e * n] = negate ? make_double2(o.x - v.re, o.y - v.im) : make_double2(o.x + v.re, o.y + v.im);
        }
    }
    asm volatile("" ::: "memory");
}
__device__ __forceinline__ void lambda_link(const double2* __restrict__ U, const M3& Lprev, const M3& sig,
                                            double2* __restrict__ Sig, const double2* __restrict__ T,
                                            double2* __restrict__ Lam, const Geom& g, int s, int mu, int par);
// LAMBDA: the force is the Sigma of the top stout level; instead of storing it the kernel
// goes on with that level's Lambda (lambda_link, next: the level below, its tape and the
// Lambda buffer), as the first launch of the stout-force recursion would.
template <bool LAMBDA>
__device__ __forceinline__ void body_clover_deriv_all(const double2* __restrict__ U, const double2* __restrict__ C,
                                                      double2* __restrict__ F, const Geom& g, const LinkMap& m, int b,
                                                      double scale_factor, double2* acc, const SigmaNext& next)
{
    int mu, par, c[4], s;
    if (!map_link_b(g, m, b, mu, par, c, s)) return;
    acc += threadIdx.x;
#pragma unroll 1
    for (int k = 0; k < 3; ++k) {
        const int nu = k < mu ? k : k + 1;
        const int rho = (mu != 0 && nu != 0) ? 0 : ((mu != 1 && nu != 1) ? 1 : 2);
        const int sigma = (mu != 3 && nu != 3) ? 3 : ((mu != 2 && nu != 2) ? 2 : 1);
        const int inv = (mu > nu) + (mu > rho) + (mu > sigma) + (nu > rho) + (nu > sigma) + (rho > sigma);
        const double sign = (inv & 1) ? -1.0 : 1.0;
        double2* park = acc + 9 * m.nthreads;
        lds_acc(acc, m.nthreads, scale(sign, cderiv_upper(U, C, g, s, c, par, mu, nu, rho, sigma, park, m.nthreads)),
                k == 0, 0);
        lds_acc(acc, m.nthreads, scale(sign, cderiv_lower(U, C, g, s, c, par, mu, nu, rho, sigma, park, m.nthreads)), 0,
                1);
    }
    M3 a;
#pragma unroll
    for (int e = 0; e < 9; ++e) a.m[e] = cd{acc[e * m.nthreads].x, acc[e * m.nthreads].y};
    const double pref = -2.0 / (512.0 * 3.14159265358979323846 * 3.14159265358979323846);
    const M3 L = ldu(U, g, s, mu, par);
    a = scale(scale_factor * pref, algebra(mul(L, a)));
    if (LAMBDA)
        lambda_link(next.U, L, a, F, next.T, next.Lam, g, s, mu, par);
    else
        st_link(F, g, s, mu, par, a);
}

// Sigma' = Uprev^+ Sigma (stored in place) and
// Lambda = TracelessHermitian(tr(B1 G) Q + tr(B2 G) Q^2 + f1 G + f2 (Q G + G Q)), G = U Sigma'
// one link, with Uprev's link and Sigma given
__device__ __forceinline__ void lambda_link(const double2* __restrict__ U, const M3& Lprev, const M3& sig,
                                            double2* __restrict__ Sig, const double2* __restrict__ T,
                                            double2* __restrict__ Lam, const Geom& g, int s, int mu, int par)
{
    M3 sp = mul_dn(Lprev, sig);
    st_link(Sig, g, s, mu, par, sp);
    M3 Q = ld_her8(T, g, s, mu, par);
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
    st_her8(Lam, g, s, mu, par, tl_herm(Gam));
}
__device__ __forceinline__ void body_lambda(const double2* __restrict__ U, const double2* __restrict__ Uprev,
                                            double2* __restrict__ Sig, const double2* __restrict__ T,
                                            double2* __restrict__ Lam, const Geom& g, const LinkMap& m, int b)
{
    int mu, par, c[4], s;
    if (!map_link_b(g, m, b, mu, par, c, s)) return;
    lambda_link(U, ldu(Uprev, g, s, mu, par), ld(Sig, g, s, mu, par), Sig, T, Lam, g, s, mu, par);
}

// body_sigma_fs and body_sigma in one pass: the six-term sum and the staple st are
// built in the same loop over nu != mu from the same neighbour links (the staple with
// the products of staple(), in its order), each waiting in LDS between the directions
// (fs at lds + lane, st after it: entries e * nthreads, e < 9), then the sigma update.
// The same operations in the same order as the two launches: the same Sigma bit for
// bit, without the scratch field Fs and the second reading of the neighbours.
// LDS: 18 x 16 bytes per lane of the block.
// MODE SIGMA_STORE stores the result in Sig.  SIGMA_LAMBDA goes on with the Lambda of the
// level below for the same link (lambda_link: its Uprev is this level's U, whose link is
// at hand, and Sigma the result in registers), with that level's field, tape and Lambda
// buffer in next (not the Lambda buffer this launch reads: other lanes still read their
// neighbours' entries there); the same operations as the two launches, without the store
// and load of Sigma between them.
// SIGMA_KICK (the last level of the recursion inside a metadynamics momentum update): the
// result is the bias force at level 0, and instead of storing it the kernel applies
// UpdateMomenta to the link, P += a Algebra(st U^+) + bf * force, with the staple it
// already holds -- body_force_update's formula, without the force field's store and
// load and the staple formed again (the momenta agree with the separate update to
// rounding, not bit for bit: the products are contracted into fused multiply-adds
// differently).
enum SigmaMode { SIGMA_STORE = 0, SIGMA_KICK = 1, SIGMA_LAMBDA = 2 };
template <int MODE>
__device__ __forceinline__ void body_sigma_all(const double2* __restrict__ U, double2* __restrict__ Sig,
                                               const double2* __restrict__ T, const double2* __restrict__ Lam,
                                               const Geom& g, const LinkMap& m, int b, double rho, double2* lds,
                                               const MomentumKick& kick, const SigmaNext& next)
{
    int mu, par, c[4], s;
    if (!map_link_b(g, m, b, mu, par, c, s)) return;
    const int n = m.nthreads;
    double2* fs_slot = lds + threadIdx.x;
    double2* st_slot = fs_slot + 9 * n;
    const int q = par ^ 1;
    const int smu = up_rt(g, s, pick4(c, mu), mu);
#pragma unroll 1
    for (int k = 0; k < 3; ++k) {
        const int nu = k < mu ? k : k + 1;
        const int cnu = pick4(c, nu);
        const int snu = up_rt(g, s, cnu, nu), snd = dn_rt(g, s, cnu, nu);
        const int smu_nd = dn_rt(g, smu, cnu, nu);
        const M3 Unu_x = ldu(U, g, s, nu, par);
        const M3 Unu_mu = ldu(U, g, smu, nu, q);
        M3 tmp1;
        {
            const M3 Umu_nu = ldu(U, g, snu, mu, q);
            // staple, upper term: U_nu(x) U_mu(x+nu) U_nu(x+mu)^+
            const M3 t = mul_nd_r01(mul_r01(Unu_x, Umu_nu), Unu_mu);
            M3 st;
            if (k == 0) {
                st = t;
                third_row(st);
            } else {
                st = park_load(st_slot, n);
                acc_su3_r01(st, t);
            }
            park_store(st_slot, n, st);
            tmp1 = mul_nd(Unu_mu, Umu_nu);
        }
        {   // six-term sum: + tmp1 (U_nu(x)^+ L_nu(x) + L_mu(x+nu) U_nu(x)^+) - L_nu(x+mu) tmp1 U_nu(x)^+
            const M3 x1 = mul(tmp1, add(mul_dn(Unu_x, ld_her8(Lam, g, s, nu, par)), mul_nd(ld_her8(Lam, g, snu, mu, q), Unu_x)));
            M3 fs = k == 0 ? m3_zero() : park_load(fs_slot, n);
            acc(fs, x1);
            park_store(fs_slot, n, fs);
            const M3 x2 = mul_nd(mul(ld_her8(Lam, g, smu, nu, q), tmp1), Unu_x);
            fs = park_load(fs_slot, n);
            fs = sub(fs, x2);
            park_store(fs_slot, n, fs);
        }
        const M3 Umu_nd = ldu(U, g, snd, mu, q);
        const M3 Unu_mund = ldu(U, g, smu_nd, nu, par);
        const M3 Unu_nd = ldu(U, g, snd, nu, q);
        {   // staple, lower term: U_nu(x-nu)^+ U_mu(x-nu) U_nu(x-nu+mu)
            M3 ef = mul_r01(Umu_nd, Unu_mund);
            third_row(ef);
            M3 st = park_load(st_slot, n);
            mul_dn_acc(st, Unu_nd, ef);
            park_store(st_slot, n, st);
        }
        {   // + U_nu(x+mu-nu)^+ (U_mu(x-nu)^+ (L_mu(x-nu) - L_nu(x-nu)) + L_nu(x+mu-nu) U_mu(x-nu)^+) U_nu(x-nu)
            M3 tmp2 = mul_dn(Umu_nd, sub(ld_her8(Lam, g, snd, mu, q), ld_her8(Lam, g, snd, nu, q)));
            M3 inner = add(tmp2, mul_nd(ld_her8(Lam, g, smu_nd, nu, par), Umu_nd));
            const M3 x3 = mul(mul_dn(Unu_mund, inner), Unu_nd);
            M3 fs = park_load(fs_slot, n);
            acc(fs, x3);
            park_store(fs_slot, n, fs);
        }
    }
    M3 cl = mul_dn(park_load(st_slot, n), ld_her8(Lam, g, s, mu, par));   // st^+ Lambda
    const M3 fs = park_load(fs_slot, n);
    M3 inner = mul(ld(Sig, g, s, mu, par), su3_exp(ld_her8(T, g, s, mu, par)));
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        cd d = cl.m[e] - fs.m[e];
        inner.m[e] += cd{-rho * d.im, rho * d.re};
    }
    const M3 L = ldu(U, g, s, mu, par);
    const M3 sig = algebra(mul(L, inner));
    if (MODE == SIGMA_STORE) {
        st_link(Sig, g, s, mu, par, sig);
    } else if (MODE == SIGMA_LAMBDA) {
        lambda_link(next.U, L, sig, Sig, next.T, next.Lam, g, s, mu, par);
    } else {
        const M3 A = algebra(mul_nd(park_load(st_slot, n), L));
        M3 p = ld(kick.P, g, s, mu, par);
#pragma unroll
        for (int e = 0; e < 9; ++e) p.m[e] += kick.a * A.m[e] + kick.bf * sig.m[e];
        st_link(kick.P, g, s, mu, par, p);
    }
}

}  // namespace lt
