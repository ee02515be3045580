// lt_device.hpp -- device-side building blocks for the gfx950 gauge-field kernels:
// FP64 complex and 3x3 matrix arithmetic, the even/odd structure-of-arrays link
// layout, lattice neighbour arithmetic, the Philox4x32-10 counter RNG, SU(2)
// subgroup helpers, Gram-Schmidt reunitarisation and the Cayley-Hamilton
// exponential.  Everything here is FP64 VALU code; MFMA is not used because a
// 3x3 complex product is not a dense contraction worth tiling.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Launch-shape knobs of the link-update kernels (sweeps, stout, flow stages).
// LT_LINK_BS is the largest thread block they are compiled for; the host never
// launches them with more threads (tiles are validated against it).
#ifndef LT_LINK_BS
#define LT_LINK_BS 256
#endif
#ifndef LT_LINK_WAVES
#define LT_LINK_WAVES 2
#endif
#ifndef LT_XCD_REMAP
#define LT_XCD_REMAP 1
#endif
// order the t tiles of each XCD's block range fastest (LinkMap::tc)
#ifndef LT_TCHUNK
#define LT_TCHUNK 1
#endif
// y tiles per chunk (LinkMap::yc; 0: all y tiles in one chunk).  Two y tiles per
// chunk at 32^4 (8 tiles in flight per x step in each XCD): OR sweep -2 %, RK3 flow
// step -1 % against one chunk; one y tile per chunk is slower.
#ifndef LT_YCHUNK
#define LT_YCHUNK 2
#endif
#define LT_LINK_KERNEL __launch_bounds__(LT_LINK_BS) __attribute__((amdgpu_waves_per_eu(LT_LINK_WAVES)))

namespace lt {

// Workgroups are dispatched round-robin over the 8 XCDs of the MI355X, each with
// its own L2.  Remapping the block index so that every XCD works on one
// contiguous range of sites keeps the z/y neighbours of a block in the same L2.
__device__ __forceinline__ int xcd_block(int b, int nb)
{
#if LT_XCD_REMAP
    if ((nb & 7) == 0) return (b & 7) * (nb >> 3) + (b >> 3);
#endif
    return b;
}
// The same with each XCD's range walked backwards when rev is set.  Consecutive
// launches over the same links alternate direction, so each starts on the links
// the previous one touched last (still in the L2 and the memory-side cache).
__device__ __forceinline__ int xcd_block_dir(int b, int nb, int rev)
{
#if LT_XCD_REMAP
    if ((nb & 7) == 0) {
        const int per = nb >> 3, loc = b >> 3;
        return (b & 7) * per + (rev ? per - 1 - loc : loc);
    }
#endif
    return rev ? nb - 1 - b : b;
}

// ---------------------------------------------------------------------------
// complex double
struct cd {
    double re, im;
};

__device__ __forceinline__ cd mk(double r, double i) { return cd{r, i}; }
__device__ __forceinline__ cd operator+(cd a, cd b) { return cd{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cd operator-(cd a, cd b) { return cd{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cd operator-(cd a) { return cd{-a.re, -a.im}; }
__device__ __forceinline__ cd operator*(cd a, cd b)
{
    return cd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cd operator*(double s, cd a) { return cd{s * a.re, s * a.im}; }
__device__ __forceinline__ cd operator*(cd a, double s) { return cd{s * a.re, s * a.im}; }
__device__ __forceinline__ cd conj(cd a) { return cd{a.re, -a.im}; }
__device__ __forceinline__ cd mul_i(cd a) { return cd{-a.im, a.re}; }
__device__ __forceinline__ cd& operator+=(cd& a, cd b) { a.re += b.re; a.im += b.im; return a; }
__device__ __forceinline__ cd& operator-=(cd& a, cd b) { a.re -= b.re; a.im -= b.im; return a; }
// a * conj(b)
__device__ __forceinline__ cd mul_c(cd a, cd b)
{
    return cd{a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}
// conj(a) * b
__device__ __forceinline__ cd cmul(cd a, cd b)
{
    return cd{a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}
// conj(x * y - z * w) as two chains of fused multiply-adds (the third row of an
// SU(3) matrix from its first two rows)
__device__ __forceinline__ cd conj_cross(cd x, cd y, cd z, cd w)
{
    double re = z.im * w.im;
    re = fma(-z.re, w.re, re);
    re = fma(-x.im, y.im, re);
    re = fma(x.re, y.re, re);
    double im = z.im * w.re;
    im = fma(z.re, w.im, im);
    im = fma(-x.im, y.re, im);
    im = fma(-x.re, y.im, im);
    return cd{re, im};
}

// ---------------------------------------------------------------------------
// 3x3 complex matrix, row-major
struct M3 {
    cd m[9];
};

__device__ __forceinline__ M3 m3_zero()
{
    M3 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.m[i] = cd{0.0, 0.0};
    return r;
}
__device__ __forceinline__ M3 m3_id()
{
    M3 r = m3_zero();
    r.m[0].re = r.m[4].re = r.m[8].re = 1.0;
    return r;
}
// sum_k op(a_k) op(b_k) over three complex pairs as two chains of one multiply
// and five fused multiply-adds (op = conj where CA / CB is set): 12 FP64
// instructions per complex entry instead of 16 for separate products and sums.
template <bool CA, bool CB>
__device__ __forceinline__ cd dot3(cd a0, cd b0, cd a1, cd b1, cd a2, cd b2)
{
    // (x.re + i sx x.im)(y.re + i sy y.im): re = x.re y.re - sx sy x.im y.im, im = sy x.re y.im + sx x.im y.re
    constexpr double sa = CA ? -1.0 : 1.0, sb = CB ? -1.0 : 1.0, sab = -sa * sb;
    double re = a0.re * b0.re;
    re = fma(sab * a0.im, b0.im, re);
    re = fma(a1.re, b1.re, re);
    re = fma(sab * a1.im, b1.im, re);
    re = fma(a2.re, b2.re, re);
    re = fma(sab * a2.im, b2.im, re);
    double im = (sb * a0.re) * b0.im;
    im = fma(sa * a0.im, b0.re, im);
    im = fma(sb * a1.re, b1.im, im);
    im = fma(sa * a1.im, b1.re, im);
    im = fma(sb * a2.re, b2.im, im);
    im = fma(sa * a2.im, b2.re, im);
    return cd{re, im};
}
// acc + sum_k op(a_k) op(b_k): the same chains started from acc (6 FMAs per part)
template <bool CA, bool CB>
__device__ __forceinline__ cd dot3_acc(cd acc, cd a0, cd b0, cd a1, cd b1, cd a2, cd b2)
{
    constexpr double sa = CA ? -1.0 : 1.0, sb = CB ? -1.0 : 1.0, sab = -sa * sb;
    double re = fma(a0.re, b0.re, acc.re);
    re = fma(sab * a0.im, b0.im, re);
    re = fma(a1.re, b1.re, re);
    re = fma(sab * a1.im, b1.im, re);
    re = fma(a2.re, b2.re, re);
    re = fma(sab * a2.im, b2.im, re);
    double im = fma(sb * a0.re, b0.im, acc.im);
    im = fma(sa * a0.im, b0.re, im);
    im = fma(sb * a1.re, b1.im, im);
    im = fma(sa * a1.im, b1.re, im);
    im = fma(sb * a2.re, b2.im, im);
    im = fma(sa * a2.im, b2.re, im);
    return cd{re, im};
}
// acc += a * b
__device__ __forceinline__ void mul_acc(M3& acc, const M3& a, const M3& b)
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            acc.m[3 * i + j] = dot3_acc<false, false>(acc.m[3 * i + j], a.m[3 * i], b.m[j], a.m[3 * i + 1], b.m[3 + j],
                                                      a.m[3 * i + 2], b.m[6 + j]);
}
// acc += a * b^+
__device__ __forceinline__ void mul_nd_acc(M3& acc, const M3& a, const M3& b)
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            acc.m[3 * i + j] = dot3_acc<false, true>(acc.m[3 * i + j], a.m[3 * i], b.m[3 * j], a.m[3 * i + 1],
                                                     b.m[3 * j + 1], a.m[3 * i + 2], b.m[3 * j + 2]);
}
// acc += a^+ * b
__device__ __forceinline__ void mul_dn_acc(M3& acc, const M3& a, const M3& b)
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            acc.m[3 * i + j] = dot3_acc<true, false>(acc.m[3 * i + j], a.m[i], b.m[j], a.m[3 + i], b.m[3 + j],
                                                     a.m[6 + i], b.m[6 + j]);
}
// a * b
__device__ __forceinline__ M3 mul(const M3& a, const M3& b)
{
    M3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r.m[3 * i + j] = dot3<false, false>(a.m[3 * i], b.m[j], a.m[3 * i + 1], b.m[3 + j], a.m[3 * i + 2], b.m[6 + j]);
    return r;
}
// a * b^+
__device__ __forceinline__ M3 mul_nd(const M3& a, const M3& b)
{
    M3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r.m[3 * i + j] = dot3<false, true>(a.m[3 * i], b.m[3 * j], a.m[3 * i + 1], b.m[3 * j + 1], a.m[3 * i + 2],
                                               b.m[3 * j + 2]);
    return r;
}
// a^+ * b
__device__ __forceinline__ M3 mul_dn(const M3& a, const M3& b)
{
    M3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r.m[3 * i + j] = dot3<true, false>(a.m[i], b.m[j], a.m[3 + i], b.m[3 + j], a.m[6 + i], b.m[6 + j]);
    return r;
}
// Rows 0 and 1 of a * b and of a * b^+ (row 2 of a is not read; row 2 of the
// result is zero), for SU(3) products whose row 2 is rebuilt by third_row.
__device__ __forceinline__ M3 mul_r01(const M3& a, const M3& b)
{
    M3 r;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r.m[3 * i + j] = dot3<false, false>(a.m[3 * i], b.m[j], a.m[3 * i + 1], b.m[3 + j], a.m[3 * i + 2], b.m[6 + j]);
    r.m[6] = r.m[7] = r.m[8] = cd{0.0, 0.0};
    return r;
}
__device__ __forceinline__ M3 mul_nd_r01(const M3& a, const M3& b)
{
    M3 r;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r.m[3 * i + j] = dot3<false, true>(a.m[3 * i], b.m[3 * j], a.m[3 * i + 1], b.m[3 * j + 1], a.m[3 * i + 2],
                                               b.m[3 * j + 2]);
    r.m[6] = r.m[7] = r.m[8] = cd{0.0, 0.0};
    return r;
}
// rows 0 and 1 of a^+ b and of a^+ b^+ (a, b complete)
__device__ __forceinline__ M3 mul_dn_r01(const M3& a, const M3& b)
{
    M3 r;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r.m[3 * i + j] = dot3<true, false>(a.m[i], b.m[j], a.m[3 + i], b.m[3 + j], a.m[6 + i], b.m[6 + j]);
    r.m[6] = r.m[7] = r.m[8] = cd{0.0, 0.0};
    return r;
}
__device__ __forceinline__ M3 mul_dd_r01(const M3& a, const M3& b)
{
    M3 r;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r.m[3 * i + j] = dot3<true, true>(a.m[i], b.m[3 * j], a.m[3 + i], b.m[3 * j + 1], a.m[6 + i], b.m[3 * j + 2]);
    r.m[6] = r.m[7] = r.m[8] = cd{0.0, 0.0};
    return r;
}
__device__ __forceinline__ M3 dag(const M3& a)
{
    M3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.m[3 * i + j] = conj(a.m[3 * j + i]);
    return r;
}
__device__ __forceinline__ M3 add(const M3& a, const M3& b)
{
    M3 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.m[i] = a.m[i] + b.m[i];
    return r;
}
__device__ __forceinline__ M3 sub(const M3& a, const M3& b)
{
    M3 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.m[i] = a.m[i] - b.m[i];
    return r;
}
__device__ __forceinline__ void acc(M3& a, const M3& b)
{
#pragma unroll
    for (int i = 0; i < 9; ++i) a.m[i] += b.m[i];
}
__device__ __forceinline__ M3 scale(double s, const M3& a)
{
    M3 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.m[i] = s * a.m[i];
    return r;
}
__device__ __forceinline__ M3 scale(cd s, const M3& a)
{
    M3 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.m[i] = s * a.m[i];
    return r;
}
__device__ __forceinline__ cd trace(const M3& a) { return a.m[0] + a.m[4] + a.m[8]; }
// Re tr(a b)
__device__ __forceinline__ double retr_mul(const M3& a, const M3& b)
{
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) s += a.m[3 * i + k].re * b.m[3 * k + i].re - a.m[3 * i + k].im * b.m[3 * k + i].im;
    return s;
}
// Re tr(a b^+) for a, b in SU(3) given by rows 0 and 1: with G_ij = row_i(a) . conj(row_j(b))
// (i, j < 2) the row-2 term conj(a_0 x a_1) . (b_0 x b_1) equals conj(G_00 G_11 - G_01 G_10)
// (Binet-Cauchy), so Re tr(a b^+) = Re(G_00 + G_11 + G_00 G_11 - G_01 G_10).
__device__ __forceinline__ double retr_su3_nd_r01(const M3& a, const M3& b)
{
    cd G[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
            G[i][j] = dot3_acc<false, true>(cd{0.0, 0.0}, a.m[3 * i], b.m[3 * j], a.m[3 * i + 1], b.m[3 * j + 1],
                                            a.m[3 * i + 2], b.m[3 * j + 2]);
    const cd d = G[0][0] * G[1][1] - G[0][1] * G[1][0];
    return G[0][0].re + G[1][1].re + d.re;
}
// traceless anti-hermitian part: 1/2 (A - A^+) - 1/6 tr(A - A^+)
__device__ __forceinline__ M3 algebra(const M3& a)
{
    M3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.m[3 * i + j] = 0.5 * (a.m[3 * i + j] - conj(a.m[3 * j + i]));
    // tr(A - A^+) = 2i * sum Im A_ii
    double t = (2.0 * (a.m[0].im + a.m[4].im + a.m[8].im)) * (1.0 / 6.0);
    r.m[0].im -= t;
    r.m[4].im -= t;
    r.m[8].im -= t;
    return r;
}
// traceless hermitian part: 1/2 (A + A^+) - 1/6 tr(A + A^+)
__device__ __forceinline__ M3 tl_herm(const M3& a)
{
    M3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.m[3 * i + j] = 0.5 * (a.m[3 * i + j] + conj(a.m[3 * j + i]));
    double t = (2.0 * (a.m[0].re + a.m[4].re + a.m[8].re)) * (1.0 / 6.0);
    r.m[0].re -= t;
    r.m[4].re -= t;
    r.m[8].re -= t;
    return r;
}

// Normalise row 0, orthogonalise and normalise row 1, row 2 = conj(row0 x row1)
__device__ __forceinline__ void gram_schmidt(M3& a)
{
    cd* m = a.m;
    double n0 = 1.0 / sqrt(m[0].re * m[0].re + m[0].im * m[0].im + m[1].re * m[1].re + m[1].im * m[1].im
                           + m[2].re * m[2].re + m[2].im * m[2].im);
    m[0] = n0 * m[0];
    m[1] = n0 * m[1];
    m[2] = n0 * m[2];
    cd psi = mul_c(m[3], m[0]) + mul_c(m[4], m[1]) + mul_c(m[5], m[2]);
    m[3] = m[3] - psi * m[0];
    m[4] = m[4] - psi * m[1];
    m[5] = m[5] - psi * m[2];
    double n1 = 1.0 / sqrt(m[3].re * m[3].re + m[3].im * m[3].im + m[4].re * m[4].re + m[4].im * m[4].im
                           + m[5].re * m[5].re + m[5].im * m[5].im);
    m[3] = n1 * m[3];
    m[4] = n1 * m[4];
    m[5] = n1 * m[5];
    m[6] = conj_cross(m[1], m[5], m[2], m[4]);
    m[7] = conj_cross(m[2], m[3], m[0], m[5]);
    m[8] = conj_cross(m[0], m[4], m[1], m[3]);
}

// ---------------------------------------------------------------------------
// Cayley-Hamilton exp(iQ) for traceless hermitian Q (Morningstar-Peardon),
// with the same branch structure and Taylor switch-over points as the
// reference's SU3::ExpConstants.
struct ExpC {
    double u, w, u2, w2, c0max;
    int signflip;
    cd exp_miu, exp_2iu, h0, h1, h2;
    double denom, cosw, xi0;
};

__device__ __forceinline__ double xi0_of(double w)
{
    if (fabs(w) <= 0.05) {
        double w2 = w * w;
        return 1.0 - (1.0 / 6.0) * w2 * (1.0 - (1.0 / 20.0) * w2 * (1.0 - (1.0 / 42.0) * w2));
    }
    return sin(w) / w;
}
__device__ __forceinline__ double xi1_of(double w)
{
    if (fabs(w) <= 0.115) {
        double w2 = w * w;
        return -(1.0 / 3.0) + (1.0 / 30.0) * w2 * (1.0 - (1.0 / 28.0) * w2 * (1.0 - (1.0 / 54.0) * w2));
    }
    return cos(w) / (w * w) - sin(w) / (w * w * w);
}

// From the invariants c0s = tr(Q^3)/3 = det Q and c1 = tr(Q^2)/2.
__device__ __forceinline__ void exp_constants_c(double c0s, double c1, ExpC& e)
{
    double c0 = fabs(c0s);
    const double c1_3 = c1 / 3.0;
    e.c0max = 2.0 * (c1_3 * sqrt(c1_3));   // 2 (c1/3)^(3/2)
    e.signflip = c0s < 0.0;
    double theta = acos(fmin(c0 / e.c0max, 1.0));
    double st, ct;
    sincos((1.0 / 3.0) * theta, &st, &ct);
    e.u = sqrt((1.0 / 3.0) * c1) * ct;
    e.w = sqrt(c1) * st;
    e.u2 = e.u * e.u;
    e.w2 = e.w * e.w;
    double su, cu;
    sincos(e.u, &su, &cu);
    e.exp_miu = cd{cu, -su};
    e.exp_2iu = conj(e.exp_miu * e.exp_miu);
    e.cosw = cos(e.w);
    e.xi0 = xi0_of(e.w);
    if (e.c0max < 2.2250738585072014e-308) {
        e.denom = 1.0;
        e.h0 = cd{1.0, 0.0};
        e.h1 = cd{0.0, 0.0};
        e.h2 = cd{0.0, 0.0};
        return;
    }
    const double u = e.u, u2 = e.u2, w2 = e.w2;
    const cd ixi0 = cd{0.0, e.xi0};
    e.denom = 1.0 / (9.0 * u2 - w2);
    cd h0 = (u2 - w2) * e.exp_2iu + e.exp_miu * (cd{8.0 * u2 * e.cosw, 0.0} + (2.0 * u) * ixi0 * (3.0 * u2 + w2));
    cd h1 = (2.0 * u) * e.exp_2iu - e.exp_miu * (cd{2.0 * u * e.cosw, 0.0} - ixi0 * (3.0 * u2 - w2));
    cd h2 = e.exp_2iu - e.exp_miu * (cd{e.cosw, 0.0} + (3.0 * u) * ixi0);
    if (e.signflip) {
        h0 = conj(h0);
        h1 = -conj(h1);
        h2 = conj(h2);
    }
    e.h0 = h0;
    e.h1 = h1;
    e.h2 = h2;
}

// Q2 = Q*Q must be supplied by the caller (it is needed again for the exponential).
__device__ __forceinline__ void exp_constants(const M3& Q, const M3& Q2, ExpC& e)
{
    exp_constants_c((1.0 / 3.0) * retr_mul(Q, Q2), 0.5 * (Q2.m[0].re + Q2.m[4].re + Q2.m[8].re), e);
}

__device__ __forceinline__ M3 exp_from(const ExpC& e, const M3& Q, const M3& Q2)
{
    M3 r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.m[i] = e.denom * (e.h1 * Q.m[i] + e.h2 * Q2.m[i]);
    r.m[0] += e.denom * e.h0;
    r.m[4] += e.denom * e.h0;
    r.m[8] += e.denom * e.h0;
    return r;
}

// exp(iQ) = f0 + f1 Q + f2 Q^2 for small Q by the power series reduced with
// Cayley-Hamilton, Q^3 = c1 Q + c0 (c1 = tr Q^2 / 2, c0 = det Q): with
// Q^k / k! = a_k + b_k Q + g_k Q^2,
//   a_{k+1} = g_k c0 / (k+1),  b_{k+1} = (a_k + g_k c1) / (k+1),  g_{k+1} = b_k / (k+1),
// f_j = sum_k i^k (a_k, b_k, g_k).  For c1 <= EXP_SERIES_C1 (|Q|_F <= 0.2) the 13
// terms k = 0..12 leave a remainder below 0.2^13 / 13! ~ 1e-19; no transcendental
// functions are needed (the Wilson-flow exponents are always this small).
constexpr double EXP_SERIES_C1 = 0.02;
__device__ __forceinline__ void exp_series_coeffs(double c0, double c1, cd& f0, cd& f1, cd& f2)
{
    double a = 1.0, b = 0.0, gg = 0.0;
    double f0r = 0.0, f0i = 0.0, f1r = 0.0, f1i = 0.0, f2r = 0.0, f2i = 0.0;
#pragma unroll
    for (int k = 0; k <= 12; ++k) {
        switch (k & 3) {   // i^k
            case 0: f0r += a; f1r += b; f2r += gg; break;
            case 1: f0i += a; f1i += b; f2i += gg; break;
            case 2: f0r -= a; f1r -= b; f2r -= gg; break;
            default: f0i -= a; f1i -= b; f2i -= gg; break;
        }
        const double inv = 1.0 / (double)(k + 1);
        const double an = gg * c0 * inv, bn = fma(gg, c1, a) * inv, gn = b * inv;
        a = an;
        b = bn;
        gg = gn;
    }
    f0 = cd{f0r, f0i};
    f1 = cd{f1r, f1i};
    f2 = cd{f2r, f2i};
}

// exp(iQ)
__device__ __forceinline__ M3 su3_exp(const M3& Q)
{
    M3 Q2 = mul(Q, Q);
    const double c1 = 0.5 * (Q2.m[0].re + Q2.m[4].re + Q2.m[8].re);
    if (c1 <= EXP_SERIES_C1) {
        cd f0, f1, f2;
        exp_series_coeffs((1.0 / 3.0) * retr_mul(Q, Q2), c1, f0, f1, f2);
        M3 r;
#pragma unroll
        for (int i = 0; i < 9; ++i) r.m[i] = f1 * Q.m[i] + f2 * Q2.m[i];
        r.m[0] += f0;
        r.m[4] += f0;
        r.m[8] += f0;
        return r;
    }
    ExpC e;
    exp_constants(Q, Q2, e);
    return exp_from(e, Q, Q2);
}

// exp(c * A) for traceless anti-hermitian A: exp(i * (-i c A))
__device__ __forceinline__ M3 exp_ah(const M3& A, double c)
{
    M3 Q;
#pragma unroll
    for (int i = 0; i < 9; ++i) Q.m[i] = cd{c * A.m[i].im, -c * A.m[i].re};   // -i c A
    return su3_exp(Q);
}

// ---------------------------------------------------------------------------
// Block reductions: a butterfly over the 64 lanes of each wavefront (cross-lane
// exchanges, no LDS round trip and no barrier per level), then the per-wave values
// in LDS combined in wave order by thread 0, which alone holds the result.  At each
// butterfly level lane i forms op(v_i, v_j) and lane j forms op(v_j, v_i); op is
// commutative, so all lanes of a wave agree and the result does not depend on the
// scheduling.  Every thread of the block must call it; the leading barrier lets a
// kernel call it more than once (thread 0 has read the previous call's LDS values).
struct OpAdd {
    template <typename T>
    __device__ __forceinline__ T operator()(T a, T b) const { return a + b; }
};
struct OpMax {
    __device__ __forceinline__ double operator()(double a, double b) const { return fmax(a, b); }
};
template <int BS, typename T, int N, typename Op>
__device__ __forceinline__ void block_reduce(T (&v)[N], Op op)
{
    static_assert(BS % 64 == 0 && BS <= 1024, "block of whole wavefronts");
    constexpr int NW = BS / 64;
    __shared__ T sh[N][NW];
    __syncthreads();
#pragma unroll
    for (int n = 0; n < N; ++n)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v[n] = op(v[n], __shfl_xor(v[n], o, 64));
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int n = 0; n < N; ++n) sh[n][threadIdx.x >> 6] = v[n];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int n = 0; n < N; ++n)
#pragma unroll
            for (int w = 1; w < NW; ++w) v[n] = op(v[n], sh[n][w]);
}
// sum of v over the block, written to *out by thread 0
template <int BS>
__device__ __forceinline__ void block_sum_to(double v, double* out)
{
    double a[1] = {v};
    block_reduce<BS>(a, OpAdd{});
    if (threadIdx.x == 0) *out = a[0];
}

// ---------------------------------------------------------------------------
// Philox4x32-10 counter RNG (Salmon et al., SC'11).  Counter layout:
//   key = seed, ctr = {site, stream<<24 | k<<4 | sub<<2 | mu, counter_lo, counter_hi}
enum : uint32_t { RNG_HEATBATH = 0, RNG_MOMENTUM = 1, RNG_HAAR = 2 };

__host__ __device__ __forceinline__ uint32_t mulhi32(uint32_t a, uint32_t b)
{
    return (uint32_t)(((uint64_t)a * b) >> 32);
}

__host__ __device__ __forceinline__ void philox10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint32_t hi0 = mulhi32(0xD2511F53u, c[0]);
        uint32_t lo0 = 0xD2511F53u * c[0];
        uint32_t hi1 = mulhi32(0xCD9E8D57u, c[2]);
        uint32_t lo1 = 0xCD9E8D57u * c[2];
        uint32_t n0 = hi1 ^ c[1] ^ k0;
        uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
    }
}

__host__ __device__ __forceinline__ double u01(uint32_t a, uint32_t b)
{
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

__host__ __device__ __forceinline__ void draw2(uint64_t seed, uint32_t site, uint32_t stream, uint32_t k,
                                               uint32_t sub, uint32_t mu, uint64_t counter, double& u0, double& u1)
{
    uint32_t c[4] = {site, (stream << 24) | (k << 4) | (sub << 2) | mu, (uint32_t)counter, (uint32_t)(counter >> 32)};
    philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    u0 = u01(c[0], c[1]);
    u1 = u01(c[2], c[3]);
}

// Box-Muller pair from two uniforms
__device__ __forceinline__ void gauss2(double u0, double u1, double& g0, double& g1)
{
    double r = sqrt(-2.0 * log(1.0 - u0));
    double s, c;
    sincos(2.0 * 3.14159265358979323846 * u1, &s, &c);
    g0 = r * c;
    g1 = r * s;
}

// ---------------------------------------------------------------------------
// Lattice geometry.  Sites are lexicographic s = ((t*Nx+x)*Ny+y)*Nz+z; the parity
// of a site is (t+x+y+z)&1 and, because Nz is even, its checkerboard index is s>>1.
// Links live in a structure-of-arrays buffer of complex doubles:
//   U[((mu*2 + parity)*9 + c)*Vh + cbi(s)],  c = 3*row + col,  Vh = V/2,
// where cbi(s) is the checkerboard index s>>1 (plain order) or, in the four-colour
// sweep kernels, that index with each group of 16 ordered even ones first (below).
struct Geom {
    int L[4];       // Nt, Nx, Ny, Nz
    int st[4];      // lexicographic strides of t, x, y, z
    int V, Vh;
};

// Storage slot of site s in its parity array.  Plain order: cb = s>>1.  Split order
// (translation units compiled with LT_SPLIT_CB = 1, the four-colour sweeps): within
// every aligned group of 16 checkerboard indices the 8 even ones come first, then the
// 8 odd ones, so a 256-byte group of one component becomes two 128-byte lines, one
// per cb parity (a last group of 8, when Vh is not a multiple of 16, keeps its
// order).  With Nz divisible by 4, cb & 1 is (z>>1) & 1, so a colour class of the
// four-colour schedule (every fourth z of a row) reads and writes whole lines instead
// of every other 16-byte slot.  The split order costs the two-colour and Jacobi
// kernels a few address instructions per link (measured at 32^4: RK3 flow stage
// +13 %, observables +22 %), so they keep the plain order.  Kernels address links
// only through this function: the order never changes a result.
#ifndef LT_SPLIT_CB
#define LT_SPLIT_CB 0
#endif
__device__ __forceinline__ int cbi(const Geom& g, int s)
{
    const int cb = s >> 1;
#if LT_SPLIT_CB
    const int p = (cb & ~15) | ((cb & 1) << 3) | ((cb >> 1) & 7);
    return cb < (g.Vh & ~15) ? p : cb;
#else
    return cb;
#endif
}

// Offset, in complex doubles, of slot 0 of storage position pos of array a (a =
// mu*2 + parity, or plane*2 + parity) in a buffer of nslot slots per entry, and the
// distance between consecutive slots of one entry: one array of Vh entries per slot.
// (Blocks of 8 positions holding their slots one after another were measured at 32^4:
// overrelaxation +3 %, heatbath -3 %, flow unchanged; not kept.)
__device__ __forceinline__ size_t slot_base(const Geom& g, int a, int nslot, int pos)
{
    return (size_t)a * (size_t)nslot * (size_t)g.Vh + (size_t)pos;
}
__device__ __forceinline__ size_t slot_stride(const Geom& g) { return (size_t)g.Vh; }

__device__ __forceinline__ void coords_of(const Geom& g, int s, int c[4])
{
    c[3] = s % g.L[3];
    int r = s / g.L[3];
    c[2] = r % g.L[2];
    r /= g.L[2];
    c[1] = r % g.L[1];
    c[0] = r / g.L[1];
}
// site index one step along +mu / -mu given the coordinate of that direction
__device__ __forceinline__ int up(const Geom& g, int s, int cmu, int mu)
{
    return (cmu + 1 < g.L[mu]) ? s + g.st[mu] : s - (g.L[mu] - 1) * g.st[mu];
}
__device__ __forceinline__ int dn(const Geom& g, int s, int cmu, int mu)
{
    return (cmu > 0) ? s - g.st[mu] : s + (g.L[mu] - 1) * g.st[mu];
}

// Element i of a buffer.  With LT_ADDR32 the byte offset is a 32-bit value added to
// the (wave-uniform) base pointer, which lets loads use the scalar-base +
// vector-offset form; buffers must then stay below 4 GiB.
#ifndef LT_ADDR32
#define LT_ADDR32 0
#endif
__device__ __forceinline__ const double2& elem(const double2* __restrict__ base, size_t i)
{
#if LT_ADDR32
    return *reinterpret_cast<const double2*>(reinterpret_cast<const char*>(base) + (uint32_t)((uint32_t)i * 16u));
#else
    return base[i];
#endif
}
__device__ __forceinline__ double2& elem(double2* __restrict__ base, size_t i)
{
#if LT_ADDR32
    return *reinterpret_cast<double2*>(reinterpret_cast<char*>(base) + (uint32_t)((uint32_t)i * 16u));
#else
    return base[i];
#endif
}

__device__ __forceinline__ M3 ld(const double2* __restrict__ U, const Geom& g, int s, int mu, int par)
{
    const size_t o = slot_base(g, mu * 2 + par, 9, cbi(g, s));
    M3 r;
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        double2 v = elem(U, o + (size_t)c * slot_stride(g));
        r.m[c] = cd{v.x, v.y};
    }
    return r;
}
__device__ __forceinline__ void st_link(double2* __restrict__ U, const Geom& g, int s, int mu, int par, const M3& a)
{
    const size_t o = slot_base(g, mu * 2 + par, 9, cbi(g, s));
#pragma unroll
    for (int c = 0; c < 9; ++c) elem(U, o + (size_t)c * slot_stride(g)) = make_double2(a.m[c].re, a.m[c].im);
}
// su(3) algebra fields (anti-hermitian, traceless: the Wilson-flow Z) in 8 reals,
//   Z[((mu*2 + parity)*4 + c)*Vh + cb],  c: 0 = (Im a00, Im a11), 1 = a01, 2 = a02, 3 = a12;
// Re a_ii = 0, a_22 = -(a_00 + a_11) and a_ji = -conj(a_ij) are implied.
struct Alg8 {
    double d0, d1;
    cd a01, a02, a12;
};
__device__ __forceinline__ Alg8 to_alg8(const M3& a)
{
    return Alg8{a.m[0].im, a.m[4].im, a.m[1], a.m[2], a.m[5]};
}
__device__ __forceinline__ M3 from_alg8(const Alg8& z)
{
    M3 r;
    r.m[0] = cd{0.0, z.d0};
    r.m[4] = cd{0.0, z.d1};
    r.m[8] = cd{0.0, -(z.d0 + z.d1)};
    r.m[1] = z.a01;
    r.m[2] = z.a02;
    r.m[5] = z.a12;
    r.m[3] = cd{-z.a01.re, z.a01.im};
    r.m[6] = cd{-z.a02.re, z.a02.im};
    r.m[7] = cd{-z.a12.re, z.a12.im};
    return r;
}
__device__ __forceinline__ Alg8 scale(double a, const Alg8& x)
{
    return Alg8{a * x.d0, a * x.d1, a * x.a01, a * x.a02, a * x.a12};
}
// a * x + b * y
__device__ __forceinline__ Alg8 axpby(double a, const Alg8& x, double b, const Alg8& y)
{
    return Alg8{a * x.d0 + b * y.d0, a * x.d1 + b * y.d1, a * x.a01 + b * y.a01, a * x.a02 + b * y.a02,
                a * x.a12 + b * y.a12};
}
__device__ __forceinline__ Alg8 ld_alg8(const double2* __restrict__ Z, const Geom& g, int s, int mu, int par)
{
    const double2* p = Z + slot_base(g, mu * 2 + par, 4, cbi(g, s));
    const double2 v0 = p[0], v1 = p[slot_stride(g)], v2 = p[(size_t)2 * slot_stride(g)], v3 = p[(size_t)3 * slot_stride(g)];
    return Alg8{v0.x, v0.y, cd{v1.x, v1.y}, cd{v2.x, v2.y}, cd{v3.x, v3.y}};
}
__device__ __forceinline__ void st_alg8(double2* __restrict__ Z, const Geom& g, int s, int mu, int par, const Alg8& z)
{
    double2* p = Z + slot_base(g, mu * 2 + par, 4, cbi(g, s));
    p[0] = make_double2(z.d0, z.d1);
    p[slot_stride(g)] = make_double2(z.a01.re, z.a01.im);
    p[(size_t)2 * slot_stride(g)] = make_double2(z.a02.re, z.a02.im);
    p[(size_t)3 * slot_stride(g)] = make_double2(z.a12.re, z.a12.im);
}

// exp(c Z) for Z anti-hermitian traceless in the 8-real form (the flow's Z):
// exp(cZ) = exp(iQ) with Q = -i c Z hermitian traceless.  Q^2 is built from its six
// independent entries and c0 = det Q, c1 = tr(Q^2)/2 directly from Q's entries.
__device__ __forceinline__ M3 exp_alg8(const Alg8& z, double c)
{
    const double q0 = c * z.d0, q1 = c * z.d1, q2 = -(q0 + q1);
    const cd q01 = cd{c * z.a01.im, -c * z.a01.re}, q02 = cd{c * z.a02.im, -c * z.a02.re},
             q12 = cd{c * z.a12.im, -c * z.a12.re};
    const double n01 = q01.re * q01.re + q01.im * q01.im, n02 = q02.re * q02.re + q02.im * q02.im,
                 n12 = q12.re * q12.re + q12.im * q12.im;
    // Q^2: diagonal and upper triangle
    const double s00 = fma(q0, q0, n01 + n02), s11 = fma(q1, q1, n01 + n12), s22 = fma(q2, q2, n02 + n12);
    const cd s01 = (q0 + q1) * q01 + mul_c(q02, q12);   // q02 conj(q12)
    const cd s02 = (q0 + q2) * q02 + q01 * q12;
    const cd s12 = (q1 + q2) * q12 + cmul(q01, q02);    // conj(q01) q02
    const double c1 = 0.5 * (s00 + s11 + s22);
    const cd t = q01 * q12;
    const double c0s = q0 * q1 * q2 - q0 * n12 - q1 * n02 - q2 * n01 + 2.0 * (t.re * q02.re + t.im * q02.im);
    cd f0, f1, f2;
    if (c1 <= EXP_SERIES_C1) {
        exp_series_coeffs(c0s, c1, f0, f1, f2);
    } else {
        ExpC e;
        exp_constants_c(c0s, c1, e);
        f0 = e.denom * e.h0;
        f1 = e.denom * e.h1;
        f2 = e.denom * e.h2;
    }
    M3 r;
    r.m[0] = f0 + q0 * f1 + s00 * f2;
    r.m[4] = f0 + q1 * f1 + s11 * f2;
    r.m[8] = f0 + q2 * f1 + s22 * f2;
    r.m[1] = f1 * q01 + f2 * s01;
    r.m[2] = f1 * q02 + f2 * s02;
    r.m[5] = f1 * q12 + f2 * s12;
    r.m[3] = f1 * conj(q01) + f2 * conj(s01);
    r.m[6] = f1 * conj(q02) + f2 * conj(s02);
    r.m[7] = f1 * conj(q12) + f2 * conj(s12);
    return r;
}

// Streaming (non-temporal) accesses for data touched once per launch, so that it
// does not displace the L2-resident neighbourhood of the link reads.
typedef double lt_d2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 ld_stream(const double2* p)
{
    const lt_d2v v = __builtin_nontemporal_load(reinterpret_cast<const lt_d2v*>(p));
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ void st_stream(double2* p, double2 v)
{
    const lt_d2v w = {v.x, v.y};
    __builtin_nontemporal_store(w, reinterpret_cast<lt_d2v*>(p));
}
__device__ __forceinline__ Alg8 ld_alg8_stream(const double2* __restrict__ Z, const Geom& g, int s, int mu, int par)
{
    const double2* p = Z + slot_base(g, mu * 2 + par, 4, cbi(g, s));
    const double2 v0 = ld_stream(p), v1 = ld_stream(p + slot_stride(g)), v2 = ld_stream(p + (size_t)2 * slot_stride(g)),
                  v3 = ld_stream(p + (size_t)3 * slot_stride(g));
    return Alg8{v0.x, v0.y, cd{v1.x, v1.y}, cd{v2.x, v2.y}, cd{v3.x, v3.y}};
}
__device__ __forceinline__ void st_alg8_stream(double2* __restrict__ Z, const Geom& g, int s, int mu, int par,
                                               const Alg8& z)
{
    double2* p = Z + slot_base(g, mu * 2 + par, 4, cbi(g, s));
    st_stream(p, make_double2(z.d0, z.d1));
    st_stream(p + slot_stride(g), make_double2(z.a01.re, z.a01.im));
    st_stream(p + (size_t)2 * slot_stride(g), make_double2(z.a02.re, z.a02.im));
    st_stream(p + (size_t)3 * slot_stride(g), make_double2(z.a12.re, z.a12.im));
}

// Traceless hermitian fields (the stout exponents Q of the CV tape, the stout-force
// Lambda) in 8 reals, slots as above:
//   c: 0 = (Re q00, Re q11), 1 = q01, 2 = q02, 3 = q12;
// Im q_ii = 0, q_22 = -(q_00 + q_11) and q_ji = conj(q_ij) are implied.
__device__ __forceinline__ M3 ld_her8(const double2* __restrict__ Z, const Geom& g, int s, int mu, int par)
{
    const double2* p = Z + slot_base(g, mu * 2 + par, 4, cbi(g, s));
    const double2 v0 = p[0], v1 = p[slot_stride(g)], v2 = p[(size_t)2 * slot_stride(g)], v3 = p[(size_t)3 * slot_stride(g)];
    M3 r;
    r.m[0] = cd{v0.x, 0.0};
    r.m[4] = cd{v0.y, 0.0};
    r.m[8] = cd{-(v0.x + v0.y), 0.0};
    r.m[1] = cd{v1.x, v1.y};
    r.m[2] = cd{v2.x, v2.y};
    r.m[5] = cd{v3.x, v3.y};
    r.m[3] = cd{v1.x, -v1.y};
    r.m[6] = cd{v2.x, -v2.y};
    r.m[7] = cd{v3.x, -v3.y};
    return r;
}
__device__ __forceinline__ void st_her8(double2* __restrict__ Z, const Geom& g, int s, int mu, int par, const M3& q)
{
    double2* p = Z + slot_base(g, mu * 2 + par, 4, cbi(g, s));
    p[0] = make_double2(q.m[0].re, q.m[4].re);
    p[slot_stride(g)] = make_double2(q.m[1].re, q.m[1].im);
    p[(size_t)2 * slot_stride(g)] = make_double2(q.m[2].re, q.m[2].im);
    p[(size_t)3 * slot_stride(g)] = make_double2(q.m[5].re, q.m[5].im);
}
// the matrix that st_her8 followed by ld_her8 reproduces
__device__ __forceinline__ M3 her8_round(const M3& q)
{
    M3 r;
    r.m[0] = cd{q.m[0].re, 0.0};
    r.m[4] = cd{q.m[4].re, 0.0};
    r.m[8] = cd{-(q.m[0].re + q.m[4].re), 0.0};
    r.m[1] = q.m[1];
    r.m[2] = q.m[2];
    r.m[5] = q.m[5];
    r.m[3] = conj(q.m[1]);
    r.m[6] = conj(q.m[2]);
    r.m[7] = conj(q.m[5]);
    return r;
}
// complex-double offset of a replica's 8-real field given its link-field offset
__host__ __device__ constexpr long long alg_offset(long long link_off) { return link_off / 72 * 32; }

// SU(3) links are stored as rows 0 and 1 only: every update ends with the
// Gram-Schmidt step whose last row is conj(row0 x row1), so row 2 is rebuilt on
// load with exactly that formula and never travels through memory.
__device__ __forceinline__ void third_row(M3& a)
{
    cd* m = a.m;
    m[6] = conj_cross(m[1], m[5], m[2], m[4]);
    m[7] = conj_cross(m[2], m[3], m[0], m[5]);
    m[8] = conj_cross(m[0], m[4], m[1], m[3]);
}
// acc + conj(x * y - z * w)
__device__ __forceinline__ cd conj_cross_acc(cd acc, cd x, cd y, cd z, cd w)
{
    double re = fma(z.im, w.im, acc.re);
    re = fma(-z.re, w.re, re);
    re = fma(-x.im, y.im, re);
    re = fma(x.re, y.re, re);
    double im = fma(z.im, w.re, acc.im);
    im = fma(z.re, w.im, im);
    im = fma(-x.im, y.re, im);
    im = fma(-x.re, y.im, im);
    return cd{re, im};
}
// acc += t for an SU(3) matrix t given by rows 0 and 1 (row 2 = conj(row0 x row1))
__device__ __forceinline__ void acc_su3_r01(M3& acc, const M3& t)
{
    const cd* m = t.m;
#pragma unroll
    for (int i = 0; i < 6; ++i) acc.m[i] += m[i];
    acc.m[6] = conj_cross_acc(acc.m[6], m[1], m[5], m[2], m[4]);
    acc.m[7] = conj_cross_acc(acc.m[7], m[2], m[3], m[0], m[5]);
    acc.m[8] = conj_cross_acc(acc.m[8], m[0], m[4], m[1], m[3]);
}
__device__ __forceinline__ M3 ldu(const double2* __restrict__ U, const Geom& g, int s, int mu, int par)
{
    const size_t o = slot_base(g, mu * 2 + par, 9, cbi(g, s));
    M3 r;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        double2 v = elem(U, o + (size_t)c * slot_stride(g));
        r.m[c] = cd{v.x, v.y};
    }
    third_row(r);
    return r;
}
// rows 0 and 1 of an SU(3) link; row 2 is left zero (callers never read it)
__device__ __forceinline__ M3 ldu2(const double2* __restrict__ U, const Geom& g, int s, int mu, int par)
{
    const size_t o = slot_base(g, mu * 2 + par, 9, cbi(g, s));
    M3 r;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        double2 v = elem(U, o + (size_t)c * slot_stride(g));
        r.m[c] = cd{v.x, v.y};
    }
    r.m[6] = r.m[7] = r.m[8] = cd{0.0, 0.0};
    return r;
}
__device__ __forceinline__ void stu(double2* __restrict__ U, const Geom& g, int s, int mu, int par, const M3& a)
{
    const size_t o = slot_base(g, mu * 2 + par, 9, cbi(g, s));
#pragma unroll
    for (int c = 0; c < 6; ++c) elem(U, o + (size_t)c * slot_stride(g)) = make_double2(a.m[c].re, a.m[c].im);
}

__device__ __forceinline__ void stu_stream(double2* __restrict__ U, const Geom& g, int s, int mu, int par, const M3& a)
{
    double2* p = U + slot_base(g, mu * 2 + par, 9, cbi(g, s));
#pragma unroll
    for (int c = 0; c < 6; ++c) st_stream(p + (size_t)c * slot_stride(g), make_double2(a.m[c].re, a.m[c].im));
}

// Clover planes R_munu = C_munu - C_munu^+ (mu < nu) are stored anti-hermitian in
// 9 reals, R_SLOTS complex slots per (plane, parity): 12 * R_SLOTS * Vh complex
// doubles per lattice; a replica at link offset off (72 * Vh units) has its
// planes at clover_offset(off).
constexpr int R_SLOTS = 5;
__host__ __device__ constexpr long long clover_offset(long long link_off) { return link_off / 72 * (12 * R_SLOTS); }

__device__ __forceinline__ int parity_of(const int c[4]) { return (c[0] + c[1] + c[2] + c[3]) & 1; }

// Thread -> site map of the per-parity link kernels.  Blocks cover 4D tiles of
// (t, x, y, zc) with zc = z/2 the checkerboard coordinate, z fastest inside a tile.
// The t tiles are split into chunks of tc consecutive tiles; tiles are ordered
// (t within the chunk) fastest, then z, y, x, chunk.  With tc = N[0]/8 each XCD's
// contiguous block range is one t-slab and the t-neighbours of a tile run a few
// blocks apart, inside the set of blocks resident in that XCD's L2.  The block
// index is first spread over XCDs.
//
// Linear maps (lin = 1) cover instead nthreads consecutive checkerboard sites of the
// parity per block (T = (1, 1, 1, nthreads), N = (blocks, 1, 1, 1)).  They serve extents
// whose z/2 has few factors of two (12^4, 20^4, 24^4: a dividing power-of-two box holds 2
// or 4 sites of a row, so a wavefront reads 32- or 64-byte pieces of 32 rows), where
// consecutive sites read whole rows and the t-slab of each XCD stays contiguous.
struct LinkMap {
    int T[4];     // tile extent in t, x, y, zc
    int N[4];     // number of tiles in t, x, y, zc
    int nthreads; // T[0]*T[1]*T[2]*T[3] == blockDim.x
    int tc;       // t tiles per chunk (divides N[0])
    int yc;       // y tiles per chunk (divides N[2]) of the Jacobi kernels (flow, stout, MetaD)
    int lin;      // linear map: block b covers checkerboard sites [b * nthreads, (b + 1) * nthreads)
    int ycs;      // y tiles per chunk of the colour sweeps
    int xr;       // spread the block index over the XCDs (xcd_block)
};

// Block index of a link kernel in its map's order: spread over the XCDs when m.xr is
// set, walked backwards when rev is set.
__device__ __forceinline__ int link_block(const LinkMap& m, int b, int nb, int rev = 0)
{
    if (m.xr) return xcd_block_dir(b, nb, rev);
    return rev ? nb - 1 - b : b;
}

// tile order, fastest first: t within its chunk, z, y within its chunk, x, y chunk, t chunk
__device__ __forceinline__ int site_of_cb(const Geom& g, int cb, int par, int c[4]);
__device__ __forceinline__ bool map_site_b(const Geom& g, const LinkMap& m, int b, int par, int c[4], int& s)
{
    int tid = threadIdx.x;
    if (m.lin) {
        const int cb = b * m.nthreads + tid;
        if (tid >= m.nthreads || cb >= g.Vh) return false;
        s = site_of_cb(g, cb, par, c);
        return true;
    }
    const int bt_lo = b % m.tc;
    b /= m.tc;
    int bz = b % m.N[3];
    b /= m.N[3];
    const int by_lo = b % m.yc;
    b /= m.yc;
    int bx = b % m.N[1];
    b /= m.N[1];
    const int ny_hi = m.N[2] / m.yc;
    int by = (b % ny_hi) * m.yc + by_lo;
    int bt = (b / ny_hi) * m.tc + bt_lo;
    if (bt >= m.N[0] || tid >= m.nthreads) return false;
    int dz = tid % m.T[3];
    tid /= m.T[3];
    int dy = tid % m.T[2];
    tid /= m.T[2];
    int dx = tid % m.T[1];
    int dt = tid / m.T[1];
    c[0] = bt * m.T[0] + dt;
    c[1] = bx * m.T[1] + dx;
    c[2] = by * m.T[2] + dy;
    int zc = bz * m.T[3] + dz;
    int flip = (c[0] + c[1] + c[2] + par) & 1;
    c[3] = 2 * zc + flip;
    s = ((c[0] * g.L[1] + c[1]) * g.L[2] + c[2]) * g.L[3] + c[3];
    return true;
}

__device__ __forceinline__ bool map_site(const Geom& g, const LinkMap& m, int par, int c[4], int& s)
{
    return map_site_b(g, m, link_block(m, blockIdx.x, gridDim.x), par, c, s);
}
__device__ __forceinline__ bool map_site_dir(const Geom& g, const LinkMap& m, int par, int rev, int c[4], int& s)
{
    return map_site_b(g, m, link_block(m, blockIdx.x, gridDim.x, rev), par, c, s);
}

// Jacobi kernels over all links: the 8 (mu, parity) classes of one tile are
// consecutive blocks of the same XCD, so the staples of all directions at a tile
// are built from one L2-resident neighbourhood.  Grid: 8 * tiles blocks.
__device__ __forceinline__ bool map_link_b(const Geom& g, const LinkMap& m, int b, int& mu, int& par, int c[4], int& s)
{
    int mp = b & 7;
    mu = mp >> 1;
    par = mp & 1;
    return map_site_b(g, m, b >> 3, par, c, s);
}
__device__ __forceinline__ bool map_link_all(const Geom& g, const LinkMap& m, int& mu, int& par, int c[4], int& s)
{
    return map_link_b(g, m, link_block(m, blockIdx.x, gridDim.x), mu, par, c, s);
}

// Site of the cb-th entry of a parity: s = 2*cb or 2*cb+1 with the right parity.
__device__ __forceinline__ int site_of_cb(const Geom& g, int cb, int par, int c[4])
{
    int s = 2 * cb;
    coords_of(g, s, c);
    int flip = (c[0] + c[1] + c[2] + par) & 1;   // z of s is even
    c[3] += flip;
    return s + flip;
}

// Staple of link (s, mu): sum over nu != mu (ascending) of
//   U_nu(x) U_mu(x+nu) U_nu(x+mu)^+ + U_nu(x-nu)^+ U_mu(x-nu) U_nu(x-nu+mu)
// Every partial product of links is in SU(3), so only its rows 0 and 1 are
// multiplied out and row 2 follows as conj(row0 x row1) (24 instead of 36 FP64
// instructions); the leading link of each product is read as rows 0 and 1 only.
// Upper term: (a b) c^+ with rows 0, 1 of both products; lower term:
// d^+ (e f) with rows 0, 1 of e f.  About 20 % fewer FP64 instructions than
// three full products per term.
__device__ __forceinline__ M3 staple(const double2* __restrict__ U, const Geom& g, int s, const int c[4], int mu)
{
    const int par = (c[0] + c[1] + c[2] + c[3]) & 1;
    const int s_mu = up(g, s, c[mu], mu);
    M3 st;
#pragma unroll
    for (int nu = 0, k = 0; nu < 4; ++nu) {
        if (nu == mu) continue;
        const int cnu = c[nu];
        const int s_nu = up(g, s, cnu, nu);
        const int s_nd = dn(g, s, cnu, nu);
        const int s_nd_mu = up(g, s_nd, c[mu], mu);
        {
            const M3 ab = mul_r01(ldu2(U, g, s, nu, par), ldu(U, g, s_nu, mu, par ^ 1));
            const M3 t = mul_nd_r01(ab, ldu(U, g, s_mu, nu, par ^ 1));
            if (k == 0) {
                st = t;
                third_row(st);
            } else {
                acc_su3_r01(st, t);
            }
        }
        {
            M3 ef = mul_r01(ldu2(U, g, s_nd, mu, par ^ 1), ldu(U, g, s_nd_mu, nu, par));
            third_row(ef);
            mul_dn_acc(st, ldu(U, g, s_nd, nu, par ^ 1), ef);
        }
        ++k;
    }
    return st;
}

// Rows 0 and 1 of one link as loaded (six complex entries).
struct Raw6 {
    double2 v[6];
};
__device__ __forceinline__ void ld_raw(const double2* __restrict__ U, const Geom& g, int s, int mu, int par, Raw6& r)
{
    const double2* p = U + slot_base(g, mu * 2 + par, 9, cbi(g, s));
#pragma unroll
    for (int c = 0; c < 6; ++c) r.v[c] = p[(size_t)c * slot_stride(g)];
}
__device__ __forceinline__ M3 rows01(const Raw6& r)
{
    M3 a;
#pragma unroll
    for (int c = 0; c < 6; ++c) a.m[c] = cd{r.v[c].x, r.v[c].y};
    a.m[6] = a.m[7] = a.m[8] = cd{0.0, 0.0};
    return a;
}
__device__ __forceinline__ M3 su3_of(const Raw6& r)
{
    M3 a = rows01(r);
    third_row(a);
    return a;
}
__device__ __forceinline__ int pick4(const int v[4], int i) { return i == 0 ? v[0] : i == 1 ? v[1] : i == 2 ? v[2] : v[3]; }

// The staple of staple() as six terms (upper and lower for each nu != mu, nu
// ascending, the same products and summation order), software-pipelined: the
// three links of term j+1 are loaded before term j is multiplied out, with
// scheduling barriers so the compiler keeps that order instead of hoisting all
// 19 link loads (which spills).  The own link is loaded with the last term.
__device__ __forceinline__ M3 staple_pipelined(const double2* __restrict__ U, const Geom& g, int s, const int c[4],
                                               int mu, int par, Raw6& own)
{
    const int q = par ^ 1;
    const int s_mu = up(g, s, c[mu], mu);
    int s_nu[3], s_nd[3], s_ndmu[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int nu = k + (k >= mu);
        const int cnu = pick4(c, nu), L = pick4(g.L, nu), st = pick4(g.st, nu);
        s_nu[k] = (cnu + 1 < L) ? s + st : s - (L - 1) * st;
        s_nd[k] = (cnu > 0) ? s - st : s + (L - 1) * st;
        s_ndmu[k] = up(g, s_nd[k], c[mu], mu);
    }
    // term j = 2k: a = U_nu(x) [rows 0, 1], b = U_mu(x+nu), cc = U_nu(x+mu):  (a b) cc^+
    // term j = 2k+1: a = U_mu(x-nu) [rows 0, 1], b = U_nu(x-nu+mu), cc = U_nu(x-nu):  cc^+ (a b)
    auto load = [&](int j, Raw6& a, Raw6& b, Raw6& cc) {
        const int k = j >> 1, nu = k + (k >= mu);
        if ((j & 1) == 0) {
            ld_raw(U, g, s, nu, par, a);
            ld_raw(U, g, s_nu[k], mu, q, b);
            ld_raw(U, g, s_mu, nu, q, cc);
        } else {
            ld_raw(U, g, s_nd[k], mu, q, a);
            ld_raw(U, g, s_ndmu[k], nu, par, b);
            ld_raw(U, g, s_nd[k], nu, q, cc);
        }
    };
    Raw6 A[2], B[2], C[2];
    load(0, A[0], B[0], C[0]);
    M3 st;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        if (j < 5) load(j + 1, A[(j + 1) & 1], B[(j + 1) & 1], C[(j + 1) & 1]);
        if (j == 4) ld_raw(U, g, s, mu, par, own);
        __builtin_amdgcn_sched_barrier(0);
        const int i = j & 1;
        if ((j & 1) == 0) {
            const M3 t = mul_nd_r01(mul_r01(rows01(A[i]), su3_of(B[i])), su3_of(C[i]));
            if (j == 0) {
                st = t;
                third_row(st);
            } else {
                acc_su3_r01(st, t);
            }
        } else {
            M3 ef = mul_r01(rows01(A[i]), su3_of(B[i]));
            third_row(ef);
            mul_dn_acc(st, su3_of(C[i]), ef);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    return st;
}

// site x + a e_mu + b e_nu (|a|, |b| <= 2, mu != nu) of site s with coordinates c
__device__ __forceinline__ int wrap_add(int v, int a, int L)
{
    v += a;
    return v < 0 ? v + L : (v >= L ? v - L : v);
}
__device__ __forceinline__ int shift2(const Geom& g, int s, const int c[4], int mu, int a, int nu, int b)
{
    const int cm = pick4(c, mu), cn = pick4(c, nu);
    return s + (wrap_add(cm, a, pick4(g.L, mu)) - cm) * pick4(g.st, mu) +
           (wrap_add(cn, b, pick4(g.L, nu)) - cn) * pick4(g.st, nu);
}

// Parking: a matrix waits in LDS (9 entries at park[e * n], one slot per lane) while
// other work runs, so it holds no registers meanwhile; the empty asm statements with
// memory clobbers keep the compiler from carrying it in registers anyway.
__device__ __forceinline__ void park_store(double2* park, int n, const M3& a)
{
#pragma unroll
    for (int e = 0; e < 9; ++e) park[e * n] = make_double2(a.m[e].re, a.m[e].im);
    asm volatile("" ::: "memory");
}
__device__ __forceinline__ M3 park_load(const double2* park, int n)
{
    asm volatile("" ::: "memory");
    M3 a;
#pragma unroll
    for (int e = 0; e < 9; ++e) a.m[e] = cd{park[e * n].x, park[e * n].y};
    return a;
}

// Rectangular<2>::StapleRect (rectangular_action.hpp:186-225): for nu = mu+1, mu+2,
// mu+3 (mod 4) the six 1x2 staples of link (x, mu) -- orthogonal to it (two), extending
// it forward (two), backward (two) -- in the reference's grouping and order.  Every
// rectangle is a product of five links, an SU(3) matrix, so only rows 0 and 1 are
// carried along each chain and row 2 is added as conj(row0 x row1) (acc_su3_r01); the
// forward pair is taken as a U^+ + b U^+ and the backward pair as U^+ a + U^+ b (the
// reference's (a + b) U^+ and U^+ (a + b)), so every accumulated term is a unitary chain.
__device__ __forceinline__ M3 staple_rect(const double2* __restrict__ U, const Geom& g, int s, const int c[4], int mu,
                                          int par, double2* park = nullptr, int n = 0)
{
    const int q = par ^ 1;
    M3 st = m3_zero();
    // park (optional, see park_store): the sum waits in LDS between the terms, so that
    // one term at a time holds registers; the same additions in the same order
    if (park) park_store(park, n, st);
    auto add_term = [&](const M3& r) {
        if (park) st = park_load(park, n);
        acc_su3_r01(st, r);
        if (park) park_store(park, n, st);
    };
#pragma unroll 1
    for (int off = 1; off < 4; ++off) {
        const int nu = (mu + off) & 3;
        auto at = [&](int a, int b) { return shift2(g, s, c, mu, a, nu, b); };
        const int nup = at(0, 1), nupp = at(0, 2), mup_nup = at(1, 1), mup = at(1, 0);
        const int nud = at(0, -1), nudd = at(0, -2), mup_nudd = at(1, -2), mup_nud = at(1, -1);
        // orthogonal: U_nu(x) U_nu(x+nu) U_mu(x+2nu) U_nu(x+mu+nu)^+ U_nu(x+mu)^+ and its mirror
        {
            M3 r = mul_r01(ldu2(U, g, s, nu, par), ldu(U, g, nup, nu, q));
            r = mul_r01(r, ldu(U, g, nupp, mu, par));
            r = mul_nd_r01(r, ldu(U, g, mup_nup, nu, par));
            r = mul_nd_r01(r, ldu(U, g, mup, nu, q));
            add_term(r);
            r = mul_dd_r01(ldu(U, g, nud, nu, q), ldu(U, g, nudd, nu, par));
            r = mul_r01(r, ldu(U, g, nudd, mu, par));
            r = mul_r01(r, ldu(U, g, mup_nudd, nu, q));
            r = mul_r01(r, ldu(U, g, mup_nud, nu, par));
            add_term(r);
        }
        // forward: [U_nu(x) U_mu(x+nu) U_mu(x+mu+nu) U_nu(x+2mu)^+ + mirror] U_mu(x+mu)^+
        {
            const int mupp = at(2, 0), mupp_nud = at(2, -1);
            const M3 ump = ldu(U, g, mup, mu, q);
            M3 r = mul_r01(ldu2(U, g, s, nu, par), ldu(U, g, nup, mu, q));
            r = mul_r01(r, ldu(U, g, mup_nup, mu, par));
            r = mul_nd_r01(r, ldu(U, g, mupp, nu, par));
            r = mul_nd_r01(r, ump);
            add_term(r);
            r = mul_dn_r01(ldu(U, g, nud, nu, q), ldu(U, g, nud, mu, q));
            r = mul_r01(r, ldu(U, g, mup_nud, mu, par));
            r = mul_r01(r, ldu(U, g, mupp_nud, nu, q));
            r = mul_nd_r01(r, ump);
            add_term(r);
        }
        // backward: U_mu(x-mu)^+ [U_nu(x-mu) U_mu(x-mu+nu) U_mu(x+nu) U_nu(x+mu)^+ + mirror]
        {
            const int mud = at(-1, 0), mud_nup = at(-1, 1), mud_nud = at(-1, -1);
            const M3 umd = ldu(U, g, mud, mu, q);
            M3 r = mul_dn_r01(umd, ldu(U, g, mud, nu, q));
            r = mul_r01(r, ldu(U, g, mud_nup, mu, par));
            r = mul_r01(r, ldu(U, g, nup, mu, q));
            r = mul_nd_r01(r, ldu(U, g, mup, nu, q));
            add_term(r);
            r = mul_dd_r01(umd, ldu(U, g, mud_nud, nu, par));
            r = mul_r01(r, ldu(U, g, mud_nud, mu, par));
            r = mul_r01(r, ldu(U, g, nud, mu, q));
            r = mul_r01(r, ldu(U, g, mup_nud, nu, par));
            add_term(r);
        }
    }
    if (park) st = park_load(park, n);
    return st;
}

// Rectangular<R>::Staple (rectangular_action.hpp:232-242)
__device__ __forceinline__ M3 staple_action(const double2* __restrict__ U, const Geom& g, int s, const int c[4], int mu,
                                            int par, double c_plaq, double c_rect, double2* park = nullptr, int n = 0)
{
    const M3 p = staple(U, g, s, c, mu);
    if (c_rect == 0.0) return c_plaq == 1.0 ? p : scale(c_plaq, p);
    return add(scale(c_plaq, p), scale(c_rect, staple_rect(U, g, s, c, mu, par, park, n)));
}

// Sum of the four clover leaves of plane (mu, nu) at site s (leaf orientation of
// the reference's CalculateClover<1,1>); leaf I (the plaquette) is returned too.

// park (optional, see park_store): the running sum of the leaves waits in LDS while the
// next leaf is formed, and the leaves read their links afresh, so that one leaf at a
// time holds registers (the tests of park are branches the compiler does not hoist the
// next leaf's loads above); the same operations in the same order either way.
__device__ __forceinline__ M3 clover_plane(const double2* __restrict__ U, const Geom& g, int s, const int c[4], int par,
                                           int mu, int nu, double& retr_plaq, double2* park = nullptr, int n = 0)
{
    // every leaf is a product of four links, an SU(3) matrix: rows 0 and 1 are
    // multiplied out along the chain, row 2 is conj(row0 x row1)
    const int q = par ^ 1;
    const int smu = up(g, s, c[mu], mu), snu = up(g, s, c[nu], nu);
    const int smd = dn(g, s, c[mu], mu), snd = dn(g, s, c[nu], nu);
    const int smd_nu = up(g, smd, c[nu], nu), smd_nd = dn(g, smd, c[nu], nu), smu_nd = up(g, snd, c[mu], mu);
    // leaf I: U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
    M3 C = mul_nd_r01(mul_nd_r01(mul_r01(ldu(U, g, s, mu, par), ldu(U, g, smu, nu, q)), ldu(U, g, snu, mu, q)),
                      ldu(U, g, s, nu, par));
    third_row(C);
    retr_plaq = C.m[0].re + C.m[4].re + C.m[8].re;
    if (park) park_store(park, n, C);
    // leaf II: U_nu(x) U_mu(x-mu+nu)^+ U_nu(x-mu)^+ U_mu(x-mu)
    M3 t = mul_r01(mul_nd_r01(mul_nd_r01(ldu(U, g, s, nu, par), ldu(U, g, smd_nu, mu, par)), ldu(U, g, smd, nu, q)),
                   ldu(U, g, smd, mu, q));
    if (park) C = park_load(park, n);
    acc_su3_r01(C, t);
    if (park) park_store(park, n, C);
    // leaf III: U_mu(x-mu)^+ U_nu(x-mu-nu)^+ U_mu(x-mu-nu) U_nu(x-nu)
    t = mul_r01(mul_r01(mul_dd_r01(ldu(U, g, smd, mu, q), ldu(U, g, smd_nd, nu, par)), ldu(U, g, smd_nd, mu, par)),
                ldu(U, g, snd, nu, q));
    if (park) C = park_load(park, n);
    acc_su3_r01(C, t);
    if (park) park_store(park, n, C);
    // leaf IV: U_nu(x-nu)^+ U_mu(x-nu) U_nu(x+mu-nu) U_mu(x)^+
    t = mul_nd_r01(mul_r01(mul_dn_r01(ldu(U, g, snd, nu, q), ldu(U, g, snd, mu, q)), ldu(U, g, smu_nd, nu, par)),
                   ldu(U, g, s, mu, par));
    if (park) C = park_load(park, n);
    acc_su3_r01(C, t);
    return C;
}


// Derivative constants of the Cayley-Hamilton exponential (Morningstar-Peardon
// r_ij, b_ij, B_1, B_2 with the sign-flip conjugation rules); returns B1, B2 and
// f1, f2 for the stout-force Lambda.
__device__ __forceinline__ void exp_deriv(const ExpC& e, const M3& Q, const M3& Q2, M3& B1, M3& B2, cd& f1, cd& f2)
{
    const double u = e.u, u2 = e.u2, w2 = e.w2;
    const double xi0 = e.xi0, xi1 = xi1_of(e.w), cw = e.cosw;
    const cd em = e.exp_miu, e2 = e.exp_2iu;
    const cd ix = cd{0.0, xi0};
    const cd I = cd{0.0, 1.0};
    const cd f0 = e.denom * e.h0;
    f1 = e.denom * e.h1;
    f2 = e.denom * e.h2;
    const cd r10 = 2.0 * (cd{u, u2 - w2} * e2)
                 + 2.0 * (em * ((4.0 * u) * cd{2.0, -u} * cw + ix * cd{9.0 * u2 + w2, -u * (3.0 * u2 + w2)}));
    const cd r11 = 2.0 * (cd{1.0, 2.0 * u} * e2) + em * (-2.0 * cd{1.0, -u} * cw + ix * cd{6.0 * u, w2 - 3.0 * u2});
    const cd r12 = 2.0 * (I * e2) + I * (em * (cd{cw, 0.0} - 3.0 * cd{1.0, -u} * xi0));
    const cd r20 = -2.0 * e2 + cd{0.0, 2.0 * u} * (em * (cd{cw, 0.0} + cd{1.0, 4.0 * u} * xi0 + cd{3.0 * u2 * xi1, 0.0}));
    const cd r21 = -(I * (em * (cd{cw, 0.0} + cd{1.0, 2.0 * u} * xi0 - cd{3.0 * u2 * xi1, 0.0})));
    const cd r22 = em * cd{xi0, -3.0 * u * xi1};
    cd b10, b11, b12, b20, b21, b22;
    if (e.c0max < 2.2250738585072014e-308) {
        b10 = b11 = b12 = b20 = b21 = b22 = cd{0.0, 0.0};
    } else {
        double bd = 9.0 * u2 - w2;
        bd *= 2.0 * bd;
        bd = 1.0 / bd;
        const double p = 3.0 * u2 - w2, q = 15.0 * u2 + w2;
        if (e.signflip) {
            b10 = bd * conj((2.0 * u) * r10 + p * r20 - (2.0 * q) * conj(f0));
            b11 = -bd * conj((2.0 * u) * r11 + p * r21 + (2.0 * q) * conj(f1));
            b12 = bd * conj((2.0 * u) * r12 + p * r22 - (2.0 * q) * conj(f2));
            b20 = -bd * conj(r10 - (3.0 * u) * r20 - (24.0 * u) * conj(f0));
            b21 = bd * conj(r11 - (3.0 * u) * r21 + (24.0 * u) * conj(f1));
            b22 = -bd * conj(r12 - (3.0 * u) * r22 - (24.0 * u) * conj(f2));
        } else {
            b10 = bd * ((2.0 * u) * r10 + p * r20 - (2.0 * q) * f0);
            b11 = bd * ((2.0 * u) * r11 + p * r21 - (2.0 * q) * f1);
            b12 = bd * ((2.0 * u) * r12 + p * r22 - (2.0 * q) * f2);
            b20 = bd * (r10 - (3.0 * u) * r20 - (24.0 * u) * f0);
            b21 = bd * (r11 - (3.0 * u) * r21 - (24.0 * u) * f1);
            b22 = bd * (r12 - (3.0 * u) * r22 - (24.0 * u) * f2);
        }
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        B1.m[k] = b11 * Q.m[k] + b12 * Q2.m[k];
        B2.m[k] = b21 * Q.m[k] + b22 * Q2.m[k];
    }
    B1.m[0] += b10;
    B1.m[4] += b10;
    B1.m[8] += b10;
    B2.m[0] += b20;
    B2.m[4] += b20;
    B2.m[8] += b20;
}

}  // namespace lt
