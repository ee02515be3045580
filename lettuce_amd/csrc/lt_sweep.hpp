// lt_sweep.hpp -- per-link bodies of the local updates, shared by the
// two-colour kernels (lt_update.hip) and the four-colour kernels (lt_cb4.hip):
// SU(2)-subgroup extraction and embedding, overrelaxation, the Kennedy-Pendleton /
// Fabricius-Haan heatbath sampler with Philox draws, the Checkerboard4 site map and
// the link update with the Wilson or the rectangular staple.
#pragma once

#include "lt_device.hpp"

namespace lt {

struct SU2 {
    cd e11, e12;
};

__device__ __forceinline__ SU2 su2_mul(SU2 a, SU2 b)
{
    return SU2{a.e11 * b.e11 - mul_c(a.e12, b.e12), a.e11 * b.e12 + mul_c(a.e12, b.e11)};
}
__device__ __forceinline__ double su2_det(SU2 a)
{
    return a.e11.re * a.e11.re + a.e11.im * a.e11.im + a.e12.re * a.e12.re + a.e12.im * a.e12.im;
}

// SU(2) part of (L * st^+) restricted to rows/cols (i,j)
__device__ __forceinline__ SU2 extract(const M3& L, const M3& st, int i, int j)
{
    cd wii = mul_c(L.m[3 * i], st.m[3 * i]) + mul_c(L.m[3 * i + 1], st.m[3 * i + 1]) + mul_c(L.m[3 * i + 2], st.m[3 * i + 2]);
    cd wij = mul_c(L.m[3 * i], st.m[3 * j]) + mul_c(L.m[3 * i + 1], st.m[3 * j + 1]) + mul_c(L.m[3 * i + 2], st.m[3 * j + 2]);
    cd wji = mul_c(L.m[3 * j], st.m[3 * i]) + mul_c(L.m[3 * j + 1], st.m[3 * i + 1]) + mul_c(L.m[3 * j + 2], st.m[3 * i + 2]);
    cd wjj = mul_c(L.m[3 * j], st.m[3 * j]) + mul_c(L.m[3 * j + 1], st.m[3 * j + 1]) + mul_c(L.m[3 * j + 2], st.m[3 * j + 2]);
    return SU2{0.5 * (wii + conj(wjj)), 0.5 * (wij - conj(wji))};
}

// rows i and j of L <- embed(X) * L
__device__ __forceinline__ void embed_mul(M3& L, SU2 X, int i, int j)
{
#pragma unroll
    for (int n = 0; n < 3; ++n) {
        cd a = L.m[3 * i + n], b = L.m[3 * j + n];
        L.m[3 * i + n] = X.e11 * a + X.e12 * b;
        L.m[3 * j + n] = conj(X.e11) * b - conj(X.e12) * a;
    }
}

__device__ __forceinline__ SU2 overrelax(SU2 A)
{
    double an = 1.0 / sqrt(su2_det(A));
    SU2 V{an * A.e11, an * A.e12};
    SU2 VV = su2_mul(V, V);
    return SU2{conj(VV.e11), -VV.e12};
}

// One Kennedy-Pendleton / Fabricius-Haan attempt k of the SU(2) sampler in the
// reference's draw order: Philox blocks 2k (r1, r2) and 2k+1 (r3, r0) of the site;
// coef = -0.25 prefactor / sqrt(det A).  Returns whether the attempt is accepted.
__device__ __forceinline__ bool kp_attempt(uint64_t seed, uint32_t site, uint32_t sub, uint32_t mu, uint64_t counter,
                                           int k, double coef, double& lambda_sq)
{
    double ua0, ua1, ub0, ub1;
    draw2(seed, site, RNG_HEATBATH, (uint32_t)(2 * k), sub, mu, counter, ua0, ua1);
    draw2(seed, site, RNG_HEATBATH, (uint32_t)(2 * k + 1), sub, mu, counter, ub0, ub1);
    const double x1 = log(1.0 - ua0);
    const double x2 = cos(2.0 * 3.14159265358979323846 * (1.0 - ua1));
    const double x3 = log(1.0 - ub0);
    lambda_sq = coef * (x1 + x2 * x2 * x3);
    const double r0 = ub1;
    return !(r0 * r0 + lambda_sq >= 1.0);
}

// The accepted SU(2) draw X (direction from Philox block 0xFFFFF) times V^+, V = A / sqrt(det A)
__device__ __forceinline__ SU2 kp_finish(SU2 V, double lambda_sq, uint64_t seed, uint32_t site, uint32_t sub,
                                         uint32_t mu, uint64_t counter)
{
    double x0 = 1.0 - 2.0 * lambda_sq;
    double abs_x = sqrt(1.0 - x0 * x0);
    double phi, ct;
    draw2(seed, site, RNG_HEATBATH, 0xFFFFFu, sub, mu, counter, phi, ct);
    double cos_theta = 1.0 - 2.0 * ct;
    double vec_norm = abs_x * sqrt(1.0 - cos_theta * cos_theta);
    double sp, cp;
    sincos(2.0 * 3.14159265358979323846 * phi, &sp, &cp);
    SU2 X{cd{x0, vec_norm * cp}, cd{vec_norm * sp, abs_x * cos_theta}};
    SU2 Vd{conj(V.e11), -V.e12};
    return su2_mul(X, Vd);
}

// Kennedy-Pendleton / Fabricius-Haan SU(2) sampler in the reference's draw order: the
// first accepted attempt k = 0, 1, ..., max_iter; after max_iter+1 rejected attempts
// the identity is returned.
constexpr int KP_MAX_ITER = 10;
__device__ __forceinline__ SU2 heatbath_su2(SU2 A, double prefactor, uint64_t seed, uint32_t site, uint32_t sub,
                                            uint32_t mu, uint64_t counter)
{
    const double an = 1.0 / sqrt(su2_det(A));
    SU2 V{an * A.e11, an * A.e12};
    const double coef = -0.25 * prefactor * an;
    double lambda_sq = 0.0;
    for (int k = 0;; ++k) {
        if (k > KP_MAX_ITER) return SU2{cd{1.0, 0.0}, cd{0.0, 0.0}};
        if (kp_attempt(seed, site, sub, mu, counter, k, coef, lambda_sq)) break;
    }
    return kp_finish(V, lambda_sq, seed, site, sub, mu, counter);
}

// Site i in [0, V/4) of colour k of direction mu under Checkerboard4, where the colour
// is (t + x + y + z + x_mu) mod 4 (iterators.hpp:141-169; extents divisible by 4):
//   mu != z: every (t, x, y) row holds Nz/4 of them, every fourth z starting at
//            (k - t - x - y - x_mu) mod 4;
//   mu == z: the colour is (t + x + y + 2z) mod 4, so only rows with t + x + y = k mod 2
//            hold any, Nz/2 each, every second z.
// One thread per site of the class; consecutive threads walk z, then y, x, t.
__device__ __forceinline__ bool cb4_site(const Geom& g, int mu, int k, int i, int c[4], int& s)
{
    if (i >= (g.V >> 2)) return false;
    if (mu != 3) {
        const int per_row = g.L[3] >> 2;
        int row = i / per_row;
        const int j = i - row * per_row;
        c[2] = row % g.L[2];
        row /= g.L[2];
        c[1] = row % g.L[1];
        c[0] = row / g.L[1];
        c[3] = ((k - c[0] - c[1] - c[2] - c[mu]) & 3) + 4 * j;
    } else {
        const int per_row = g.L[3] >> 1;
        int row = i / per_row;
        const int j = i - row * per_row;
        const int ny2 = g.L[2] >> 1;
        const int yh = row % ny2;
        row /= ny2;
        c[1] = row % g.L[1];
        c[0] = row / g.L[1];
        c[2] = 2 * yh + ((k - c[0] - c[1]) & 1);
        c[3] = (((k - c[0] - c[1] - c[2]) & 3) >> 1) + 2 * j;
    }
    s = ((c[0] * g.L[1] + c[1]) * g.L[2] + c[2]) * g.L[3] + c[3];
    return true;
}

// overrelaxation with the software-pipelined staple (the heatbath's sampler
// leaves no registers for it)
#ifndef LT_OR_PIPE
#define LT_OR_PIPE 1
#endif
template <bool HEATBATH>
__device__ __forceinline__ void update_link(double2* __restrict__ U, const Geom& g, int s, const int c[4], int mu,
                                            int par, double prefactor, uint64_t seed, uint64_t counter)
{
    M3 st, L;
    if (LT_OR_PIPE && !HEATBATH) {
        Raw6 own;
        st = staple_pipelined(U, g, s, c, mu, par, own);
        L = su3_of(own);
    } else {
        st = staple(U, g, s, c, mu);
        L = ldu(U, g, s, mu, par);
    }
#pragma unroll
    for (int sub = 0; sub < 3; ++sub) {
        const int i = sub == 2 ? 1 : 0;
        const int j = sub == 0 ? 1 : 2;
        SU2 A = extract(L, st, i, j);
        SU2 X;
        if (HEATBATH)
            X = heatbath_su2(A, prefactor, seed, (uint32_t)s, (uint32_t)sub, (uint32_t)mu, counter);
        else
            X = overrelax(A);
        embed_mul(L, X, i, j);
    }
    gram_schmidt(L);
    // the updated link is not read again in this launch: a streaming store keeps the
    // L2 for the neighbour links (-0.3 % overrelaxation, -0.8 % heatbath launch at 32^4)
    stu_stream(U, g, s, mu, par, L);
}

// Checkerboard4 sweeps with Rectangular<R>(beta, c_plaq, c_rect): the action enters
// through the staple only (heatbath.hpp:110-129, overrelaxation.hpp:91-116)
template <bool HEATBATH>
__device__ __forceinline__ void update_link_action(double2* __restrict__ U, const Geom& g, int s, const int c[4],
                                                   int mu, int par, double prefactor, uint64_t seed, uint64_t counter,
                                                   double c_plaq, double c_rect, double2* park = nullptr, int n = 0)
{
    const M3 st = staple_action(U, g, s, c, mu, par, c_plaq, c_rect, park, n);
    M3 L = ldu(U, g, s, mu, par);
#pragma unroll
    for (int sub = 0; sub < 3; ++sub) {
        const int i = sub == 2 ? 1 : 0;
        const int j = sub == 0 ? 1 : 2;
        SU2 A = extract(L, st, i, j);
        SU2 X;
        if (HEATBATH)
            X = heatbath_su2(A, prefactor, seed, (uint32_t)s, (uint32_t)sub, (uint32_t)mu, counter);
        else
            X = overrelax(A);
        embed_mul(L, X, i, j);
    }
    gram_schmidt(L);
    stu(U, g, s, mu, par, L);
}

}  // namespace lt
