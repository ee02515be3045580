// lt_hmc.hip -- HMC building blocks: Gaussian su(3) momenta from Philox, the gauge
// (and metadynamics) momentum update, the exponential field update and the
// deterministic reductions of the Hamiltonian (plaquette sum, kinetic energy).
//
// Momenta and forces are general 3x3 fields (all 9 entries stored).
#include "lt_internal.hpp"

namespace lt {

namespace {

constexpr int RBS = 256;   // reduction block size

// P = 1/2 * A with A the anti-hermitian combination of the eight generators used
// by the reference's RandomMomentum; phi_1..phi_8 come from Philox stream
// RNG_MOMENTUM, draw blocks k = 0..3 (Box-Muller pairs).
__global__ void LT_LINK_KERNEL momentum_kernel(double2* __restrict__ P, Geom g, LinkMap m, uint64_t seed,
                                               uint64_t counter)
{
    int mu, par, c[4], s;
    if (!map_link_all(g, m, mu, par, c, s)) return;
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

// P += a * Algebra(st U^+) + b * F   (F optional: metadynamics force)
__global__ void LT_LINK_KERNEL force_update_kernel(const double2* __restrict__ U, double2* __restrict__ P,
                                                   const double2* __restrict__ F, Geom g, LinkMap m, double a, double b)
{
    int mu, par, c[4], s;
    if (!map_link_all(g, m, mu, par, c, s)) return;
    M3 st = staple(U, g, s, c, mu);
    M3 A = algebra(mul_nd(st, ldu(U, g, s, mu, par)));
    M3 p = ld(P, g, s, mu, par);
    if (F) {
        M3 f = ld(F, g, s, mu, par);
#pragma unroll
        for (int k = 0; k < 9; ++k) p.m[k] += a * A.m[k] + b * f.m[k];
    } else {
#pragma unroll
        for (int k = 0; k < 9; ++k) p.m[k] += a * A.m[k];
    }
    st_link(P, g, s, mu, par, p);
}

// U <- exp(eps P) U, Gram-Schmidt (local per link, in place)
__global__ void LT_LINK_KERNEL field_update_kernel(double2* __restrict__ U, const double2* __restrict__ P, Geom g,
                                                   LinkMap m, double eps)
{
    int mu, par, c[4], s;
    if (!map_link_all(g, m, mu, par, c, s)) return;
    M3 W = mul(exp_ah(ld(P, g, s, mu, par), eps), ldu(U, g, s, mu, par));
    gram_schmidt(W);
    stu(U, g, s, mu, par, W);
}

__device__ __forceinline__ void block_sum(double v, double* out)
{
    __shared__ double sh[RBS];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = RBS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// per site: sum over the 6 planes of Re tr P_munu
__global__ void __launch_bounds__(RBS) plaq_kernel(const double2* __restrict__ U, Geom g, double* part)
{
    const int s = blockIdx.x * RBS + threadIdx.x;
    double acc = 0.0;
    if (s < g.V) {
        int c[4];
        coords_of(g, s, c);
        const int par = parity_of(c);
#pragma unroll
        for (int nu = 1; nu < 4; ++nu) {
            const M3 Unu = ldu(U, g, s, nu, par);
            const int snu = up(g, s, c[nu], nu);
#pragma unroll
            for (int mu = 0; mu < nu; ++mu) {
                const int smu = up(g, s, c[mu], mu);
                M3 l = mul_nd(mul(ldu(U, g, s, mu, par), ldu(U, g, smu, nu, par ^ 1)), ldu(U, g, snu, mu, par ^ 1));
                acc += retr_mul(l, dag(Unu));
            }
        }
    }
    block_sum(acc, part + blockIdx.x);
}

// per link: -Re tr(P P)
__global__ void __launch_bounds__(RBS) kinetic_kernel(const double2* __restrict__ P, Geom g, double* part)
{
    const long long id = (long long)blockIdx.x * RBS + threadIdx.x;
    double acc = 0.0;
    if (id < 4LL * g.V) {
        const int s = (int)(id >> 2), mu = (int)(id & 3);
        int c[4];
        coords_of(g, s, c);
        M3 p = ld(P, g, s, mu, parity_of(c));
        acc = -retr_mul(p, p);
    }
    block_sum(acc, part + blockIdx.x);
}

// single block: ordered sum of n partials
__global__ void __launch_bounds__(RBS) finish_kernel(const double* part, int n, double* out)
{
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += RBS) acc += part[i];
    block_sum(acc, out);
}

double reduce(lt_ctx ctx, const char* name, unsigned nb, double* part, double* fin)
{
    launch(ctx, "reduce_finish", finish_kernel, dim3(1), dim3(RBS), 0, (const double*)part, (int)nb, fin);
    double* h = ctx->host_for(1);
    LT_HIP_CHECK(hipMemcpyAsync(h, fin, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    LT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    (void)name;
    return h[0];
}

}  // namespace

void k_random_momentum(lt_field P, uint64_t seed, uint64_t counter)
{
    const LinkMap& m = P->map;
    launch(P->ctx, "momentum", momentum_kernel, dim3(8 * map_blocks(m)), dim3(m.nthreads), 0, P->d, P->g, m, seed,
           counter);
}

void k_force_update(lt_field U, lt_field P, const double2* F, double a, double b)
{
    const LinkMap& m = U->map;
    launch(U->ctx, "force_update", force_update_kernel, dim3(8 * map_blocks(m)), dim3(m.nthreads), 0,
           (const double2*)U->d, P->d, F, U->g, m, a, b);
}

void k_field_update(lt_field U, lt_field P, double eps)
{
    const LinkMap& m = U->map;
    launch(U->ctx, "field_update", field_update_kernel, dim3(8 * map_blocks(m)), dim3(m.nthreads), 0, U->d,
           (const double2*)P->d, U->g, m, eps);
}

double k_plaquette_sum(lt_field U)
{
    lt_ctx ctx = U->ctx;
    unsigned nb = nblocks(U->g.V, RBS);
    double* part = ctx->scratch_for(nb + 1);
    launch(ctx, "plaquette", plaq_kernel, dim3(nb), dim3(RBS), 0, (const double2*)U->d, U->g, part);
    return reduce(ctx, "plaquette", nb, part, part + nb);
}

double k_kinetic(lt_field P)
{
    lt_ctx ctx = P->ctx;
    unsigned nb = nblocks(4LL * P->g.V, RBS);
    double* part = ctx->scratch_for(nb + 1);
    launch(ctx, "kinetic", kinetic_kernel, dim3(nb), dim3(RBS), 0, (const double2*)P->d, P->g, part);
    return reduce(ctx, "kinetic", nb, part, part + nb);
}

}  // namespace lt
