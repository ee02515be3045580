// lt_hmc.hip -- HMC building blocks: Gaussian su(3) momenta from Philox, the gauge
// (and metadynamics) momentum update, the exponential field update and the
// deterministic reductions of the Hamiltonian (plaquette sum, kinetic energy).
//
// Momenta and forces are general 3x3 fields (all 9 entries stored). The per-link
// bodies live in lt_bodies.hpp and are shared with the replica-batched kernels.
#include "lt_internal.hpp"
#include "lt_bodies.hpp"

namespace lt {

namespace {

constexpr int RBS = 256;   // reduction block size

__global__ void LT_LINK_KERNEL momentum_kernel(double2* __restrict__ P, Geom g, LinkMap m, uint64_t seed,
                                               uint64_t counter)
{
    body_momentum(P, g, m, link_block(m, blockIdx.x, gridDim.x), seed, counter);
}

__global__ void LT_LINK_KERNEL force_update_kernel(const double2* __restrict__ U, double2* __restrict__ P,
                                                   const double2* __restrict__ F, Geom g, LinkMap m, double a, double b)
{
    body_force_update(U, P, F, g, m, link_block(m, blockIdx.x, gridDim.x), a, b);
}

__global__ void LT_LINK_KERNEL field_update_kernel(double2* __restrict__ U, const double2* __restrict__ P, Geom g,
                                                   LinkMap m, double eps)
{
    body_field_update(U, P, g, m, link_block(m, blockIdx.x, gridDim.x), eps);
}

__device__ __forceinline__ void block_sum(double v, double* out)
{
    block_sum_to<RBS>(v, out);
}

__global__ void __launch_bounds__(RBS) plaq_kernel(const double2* __restrict__ U, Geom g, double* part)
{
    const int s = blockIdx.x * RBS + threadIdx.x;
    block_sum(s < g.V ? body_plaq_site(U, g, s) : 0.0, part + blockIdx.x);
}

__global__ void __launch_bounds__(RBS) kinetic_kernel(const double2* __restrict__ P, Geom g, double* part)
{
    const long long id = (long long)blockIdx.x * RBS + threadIdx.x;
    block_sum(id < 4LL * g.V ? body_kinetic_link(P, g, id) : 0.0, part + blockIdx.x);
}

// single block: ordered sum of n partials
__global__ void __launch_bounds__(RBS) finish_kernel(const double* part, int n, double* out)
{
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += RBS) acc += part[i];
    block_sum(acc, out);
}

double reduce(lt_ctx ctx, unsigned nb, double* part, double* fin)
{
    launch(ctx, "reduce_finish", finish_kernel, dim3(1), dim3(RBS), 0, (const double*)part, (int)nb, fin);
    double* h = ctx->host_for(1);
    LT_HIP_CHECK(hipMemcpyAsync(h, fin, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    LT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
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
    return reduce(ctx, nb, part, part + nb);
}

double k_kinetic(lt_field P)
{
    lt_ctx ctx = P->ctx;
    unsigned nb = nblocks(4LL * P->g.V, RBS);
    double* part = ctx->scratch_for(nb + 1);
    launch(ctx, "kinetic", kinetic_kernel, dim3(nb), dim3(RBS), 0, (const double2*)P->d, P->g, part);
    return reduce(ctx, nb, part, part + nb);
}

}  // namespace lt
