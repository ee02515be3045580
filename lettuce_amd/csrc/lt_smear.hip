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

}  // namespace

void k_stout(lt_ctx ctx, const Geom& g, const LinkMap& m, const double2* in, double2* out, double rho)
{
    launch(ctx, "stout", stout_kernel, dim3(8 * map_blocks(m)), dim3(m.nthreads), 0, in, out, g, m, rho);
}

void k_flow_stage(lt_ctx ctx, const Geom& g, const LinkMap& m, const double2* in, double2* out, double2* z, double c_old,
                  double c_new, double c_exp, int read_z, int write_z)
{
    launch(ctx, "flow_stage", flow_stage_kernel, dim3(8 * map_blocks(m)), dim3(m.nthreads), 0, in, out, z, g, m, c_old,
           c_new, c_exp, read_z, write_z);
}

}  // namespace lt
