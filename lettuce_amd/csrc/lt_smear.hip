// lt_smear.hip -- stout smearing and the fused Wilson-flow stage.
//
// Both are Jacobi updates over all links: the new field depends only on the old
// one, so every kernel reads one link buffer and writes another.  A flow stage
// fuses the force, the Z-combination of the Runge-Kutta scheme and the
// exponential update:
//     Z <- c_old Z + c_new Algebra(st(U) U^+),   U' = exp(c_exp Z) U,  GramSchmidt(U')
// so one RK3 step is three launches and reads/writes exactly U, U' and Z.
#include "lt_internal.hpp"

namespace lt {

namespace {

__global__ void LT_LINK_KERNEL stout_kernel(const double2* __restrict__ Uin, double2* __restrict__ Uout, Geom g,
                                            LinkMap m, double rho)
{
    int mu, par, c[4], s;
    if (!map_link_all(g, m, mu, par, c, s)) return;
    M3 st = staple(Uin, g, s, c, mu);
    M3 L = ldu(Uin, g, s, mu, par);
    M3 A = algebra(mul_nd(st, L));   // C = 1/2 (A - A^+) - 1/6 tr(A - A^+)
    M3 W = mul(exp_alg8(to_alg8(A), rho), L);   // exp(rho C)
    gram_schmidt(W);
    stu(Uout, g, s, mu, par, W);
}

// The flow stage is VALU- and latency-bound.  Pipelined staple (default): two
// waves per SIMD (240 VGPRs, no spills); the plain staple() loop fits three waves
// (measured at 32^4: pipelined at two waves 3 % faster than plain at three; four
// waves spill and run 1.4x slower).
#ifndef LT_FLOW_PIPE
#define LT_FLOW_PIPE 1
#endif
#ifndef LT_FLOW_WAVES
#if LT_FLOW_PIPE
#define LT_FLOW_WAVES 2
#else
#define LT_FLOW_WAVES 3
#endif
#endif
// Z and the output links are touched once per stage: streaming accesses
#ifndef LT_FLOW_STREAM
#define LT_FLOW_STREAM 1
#endif
__global__ void __launch_bounds__(LT_LINK_BS) __attribute__((amdgpu_waves_per_eu(LT_FLOW_WAVES)))
flow_stage_kernel(const double2* __restrict__ Uin, double2* __restrict__ Uout,
                                                 double2* __restrict__ Z, Geom g, LinkMap m, double c_old, double c_new,
                                                 double c_exp, int read_z, int write_z)
{
    int mu, par, c[4], s;
    if (!map_link_all(g, m, mu, par, c, s)) return;
#if LT_FLOW_PIPE
    Raw6 own;
    const M3 st = staple_pipelined(Uin, g, s, c, mu, par, own);
    const M3 L = su3_of(own);
#else
    M3 st = staple(Uin, g, s, c, mu);
    M3 L = ldu(Uin, g, s, mu, par);
#endif
    const Alg8 F = to_alg8(algebra(mul_nd(st, L)));
#if LT_FLOW_STREAM
    const Alg8 z = read_z ? axpby(c_old, ld_alg8_stream(Z, g, s, mu, par), c_new, F) : scale(c_new, F);
    if (write_z) st_alg8_stream(Z, g, s, mu, par, z);
    M3 W = mul(exp_alg8(z, c_exp), L);
    gram_schmidt(W);
    stu_stream(Uout, g, s, mu, par, W);
#else
    const Alg8 z = read_z ? axpby(c_old, ld_alg8(Z, g, s, mu, par), c_new, F) : scale(c_new, F);
    if (write_z) st_alg8(Z, g, s, mu, par, z);
    M3 W = mul(exp_alg8(z, c_exp), L);
    gram_schmidt(W);
    stu(Uout, g, s, mu, par, W);
#endif
}

// Site-centric flow stage: one thread updates the four links of one site.  The
// force of link (x, mu) is Algebra(st U_mu(x)^+), and each upper staple term times
// U_mu(x)^+ is the plaquette P_numu(x) = P_munu(x)^+ (rectangular_action.hpp:103-184
// restated on loops), so Algebra(P_numu) = -Algebra(P_munu): the six plaquettes at x
// serve both directions of their plane.  The lower terms are summed as a staple and
// multiplied by U_mu(x)^+ once per direction.  The thread reads 52 links (4 own, 12
// forward neighbours, 36 for the lower terms) where four link threads read 76.
// Directions are finished in order 0..3.  The algebra parts of the plaquettes that
// later directions still need (at most four at a time) wait in the thread's own LDS
// slots (no sharing between threads, no barriers): plane 01 -> slot 0, 02 -> 1,
// 03 -> 2, 12 -> 0 (after 01 was read), 13 -> 3, 23 -> 1 (after 02 was read).
__device__ __forceinline__ void alg8_acc(Alg8& a, const Alg8& b, double sgn)
{
    a.d0 += sgn * b.d0;
    a.d1 += sgn * b.d1;
    a.a01 += sgn * b.a01;
    a.a02 += sgn * b.a02;
    a.a12 += sgn * b.a12;
}
struct Alg8Slots {
    double* base;   // [slot][8][n] doubles
    int n, t;
    __device__ void put(int slot, const Alg8& a) const
    {
        double* p = base + (size_t)slot * 8 * n + t;
        p[0] = a.d0;
        p[n] = a.d1;
        p[2 * n] = a.a01.re;
        p[3 * n] = a.a01.im;
        p[4 * n] = a.a02.re;
        p[5 * n] = a.a02.im;
        p[6 * n] = a.a12.re;
        p[7 * n] = a.a12.im;
    }
    __device__ Alg8 get(int slot) const
    {
        const double* p = base + (size_t)slot * 8 * n + t;
        return Alg8{p[0], p[n], cd{p[2 * n], p[3 * n]}, cd{p[4 * n], p[5 * n]}, cd{p[6 * n], p[7 * n]}};
    }
};
constexpr int FLOW_SITE_SLOTS = 4;

// Algebra(P_munu(x)), P_munu(x) = U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
__device__ __forceinline__ Alg8 plaq_alg8(const double2* __restrict__ U, const Geom& g, int s, int par, int s_mu,
                                          int s_nu, int mu, int nu)
{
    const M3 a = mul_r01(ldu2(U, g, s, mu, par), ldu(U, g, s_mu, nu, par ^ 1));
    M3 b = mul_nd_r01(a, ldu(U, g, s_nu, mu, par ^ 1));
    third_row(b);
    return to_alg8(algebra(mul_nd(b, ldu(U, g, s, nu, par))));
}

// slot of plane (a, b), a < b, in the thread's LDS slots
__device__ __forceinline__ int plane_slot(int a, int b)
{
    // 01 -> 0, 02 -> 1, 03 -> 2, 12 -> 0, 13 -> 3, 23 -> 1
    return a == 0 ? b - 1 : (a == 1 ? (b == 2 ? 0 : 3) : 1);
}

__global__ void __launch_bounds__(LT_LINK_BS) __attribute__((amdgpu_waves_per_eu(LT_FLOW_WAVES)))
flow_site_kernel(const double2* __restrict__ Uin, double2* __restrict__ Uout, double2* __restrict__ Z, Geom g,
                 LinkMap m, double c_old, double c_new, double c_exp, int read_z, int write_z)
{
    extern __shared__ double lds_slots[];
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int par = b & 1, q = par ^ 1;
    int c[4], s;
    if (!map_site_b(g, m, b >> 1, par, c, s)) return;
    const Alg8Slots slots{lds_slots, m.nthreads, (int)threadIdx.x};
    const int s0 = s;
    int sp0[4], sm0[4];
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        sp0[d] = up(g, s, c[d], d);
        sm0[d] = dn(g, s, c[d], d);
    }
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
        // opaque copies of the site indices for this direction: the same links recur in
        // other directions' terms, and merged loads would stay live across all of them
        int s = s0, sp[4], sm[4];
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            sp[d] = sp0[d];
            sm[d] = sm0[d];
            asm volatile("" : "+v"(sp[d]), "+v"(sm[d]));
        }
        asm volatile("" : "+v"(s));
        // lower staple: sum over nu != mu of U_nu(x-nu)^+ U_mu(x-nu) U_nu(x-nu+mu)
        M3 lo;
#pragma unroll
        for (int nu = 0, k = 0; nu < 4; ++nu) {
            if (nu == mu) continue;
            const int s_nd_mu = up(g, sm[nu], c[mu], mu);
            M3 ef = mul_r01(ldu2(Uin, g, sm[nu], mu, q), ldu(Uin, g, s_nd_mu, nu, par));
            third_row(ef);
            if (k == 0)
                lo = mul_dn(ldu(Uin, g, sm[nu], nu, q), ef);
            else
                mul_dn_acc(lo, ldu(Uin, g, sm[nu], nu, q), ef);
            ++k;
            __builtin_amdgcn_sched_barrier(0);
        }
        Alg8 F = to_alg8(algebra(mul_nd(lo, ldu(Uin, g, s, mu, par))));
        __builtin_amdgcn_sched_barrier(0);
        // planes (nu, mu), nu < mu, met before: Algebra(P_numu) from the slots
#pragma unroll
        for (int nu = 0; nu < mu; ++nu) alg8_acc(F, slots.get(plane_slot(nu, mu)), 1.0);
        // new planes (mu, nu), nu > mu: -Algebra(P_munu) here, kept for direction nu
#pragma unroll
        for (int nu = mu + 1; nu < 4; ++nu) {
            const Alg8 P = plaq_alg8(Uin, g, s, par, sp[mu], sp[nu], mu, nu);
            alg8_acc(F, P, -1.0);
            slots.put(plane_slot(mu, nu), P);
            __builtin_amdgcn_sched_barrier(0);
        }
        const Alg8 z = read_z ? axpby(c_old, ld_alg8_stream(Z, g, s, mu, par), c_new, F) : scale(c_new, F);
        if (write_z) st_alg8_stream(Z, g, s, mu, par, z);
        // the own link again: an opaque copy of the site index keeps the compiler from
        // holding the first load's registers across the plaquettes
        int s2 = s0;
        asm volatile("" : "+v"(s2));
        M3 W = mul(exp_alg8(z, c_exp), ldu(Uin, g, s2, mu, par));
        gram_schmidt(W);
        stu_stream(Uout, g, s, mu, par, W);
        __builtin_amdgcn_sched_barrier(0);
    }
}

}  // namespace

void k_stout(lt_ctx ctx, const Geom& g, const LinkMap& m, const double2* in, double2* out, double rho)
{
    launch(ctx, "stout", stout_kernel, dim3(8 * map_blocks(m)), dim3(m.nthreads), 0, in, out, g, m, rho);
}

#ifndef LT_FLOW_SITE
#define LT_FLOW_SITE 0
#endif
void k_flow_stage(lt_ctx ctx, const Geom& g, const LinkMap& m, const double2* in, double2* out, double2* z, double c_old,
                  double c_new, double c_exp, int read_z, int write_z)
{
    if (LT_FLOW_SITE)
        launch(ctx, "flow_stage", flow_site_kernel, dim3(2 * map_blocks(m)), dim3(m.nthreads),
               (size_t)FLOW_SITE_SLOTS * 8 * m.nthreads * sizeof(double), in, out, z, g, m, c_old, c_new, c_exp,
               read_z, write_z);
    else
        launch(ctx, "flow_stage", flow_stage_kernel, dim3(8 * map_blocks(m)), dim3(m.nthreads), 0, in, out, z, g, m,
               c_old, c_new, c_exp, read_z, write_z);
}

}  // namespace lt
