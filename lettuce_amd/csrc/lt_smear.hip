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

constexpr int BS = 128;

__global__ void __launch_bounds__(BS) stout_kernel(const double2* __restrict__ Uin, double2* __restrict__ Uout, Geom g,
                                                   double rho)
{
    int cb = blockIdx.x * BS + threadIdx.x;
    if (cb >= g.Vh) return;
    int mu = blockIdx.y >> 1, par = blockIdx.y & 1;
    int c[4];
    int s = site_of_cb(g, cb, par, c);
    M3 st = staple(Uin, g, s, c, mu);
    M3 L = ld(Uin, g, s, mu, par);
    M3 A = algebra(mul_nd(st, L));   // C = 1/2 (A - A^+) - 1/6 tr(A - A^+)
    M3 W = mul(exp_ah(A, rho), L);   // exp(-i rho C) applied as exp(i * (-i rho C))
    gram_schmidt(W);
    st_link(Uout, g, s, mu, par, W);
}

__global__ void __launch_bounds__(BS) flow_stage_kernel(const double2* __restrict__ Uin, double2* __restrict__ Uout,
                                                        double2* __restrict__ Z, Geom g, double c_old, double c_new,
                                                        double c_exp, int read_z, int write_z)
{
    int cb = blockIdx.x * BS + threadIdx.x;
    if (cb >= g.Vh) return;
    int mu = blockIdx.y >> 1, par = blockIdx.y & 1;
    int c[4];
    int s = site_of_cb(g, cb, par, c);
    M3 st = staple(Uin, g, s, c, mu);
    M3 L = ld(Uin, g, s, mu, par);
    M3 F = algebra(mul_nd(st, L));
    M3 z;
    if (read_z) {
        M3 zo = ld(Z, g, s, mu, par);
#pragma unroll
        for (int k = 0; k < 9; ++k) z.m[k] = c_old * zo.m[k] + c_new * F.m[k];
    } else {
        z = scale(c_new, F);
    }
    if (write_z) st_link(Z, g, s, mu, par, z);
    M3 W = mul(exp_ah(z, c_exp), L);
    gram_schmidt(W);
    st_link(Uout, g, s, mu, par, W);
}

}  // namespace

void k_stout(lt_ctx ctx, const Geom& g, const double2* in, double2* out, double rho)
{
    launch(ctx, "stout", stout_kernel, dim3(nblocks(g.Vh, BS), 8), dim3(BS), 0, in, out, g, rho);
}

void k_flow_stage(lt_ctx ctx, const Geom& g, const double2* in, double2* out, double2* z, double c_old, double c_new,
                  double c_exp, int read_z, int write_z)
{
    launch(ctx, "flow_stage", flow_stage_kernel, dim3(nblocks(g.Vh, BS), 8), dim3(BS), 0, in, out, z, g, c_old, c_new,
           c_exp, read_z, write_z);
}

}  // namespace lt
