// lt_adjoint.hip -- the collective-variable forward pass (stout levels with the
// exponent tape, clover terms, clover topological charge) and its hand-written
// adjoint (clover-charge derivative, then the stout force recursion level by
// level), i.e. the metadynamics force dV/dQ * dQ/dU.
//
// Buffers (general 3x3 layout unless noted):
//   levels    n+1 link fields (two-row SU(3))
//   tape      per level, per link: the hermitian exponent Q = -i rho Algebra(st U^+), 8 reals (ld_her8)
//   clover    6 planes per site, R = C - C^+ in 9 reals: C[((plane*2 + parity)*5 + k)*Vh + cb]
//   sigma     force being propagated (in place, level by level)
//   lambda    per-link Lambda of the current level, 8 reals (ld_her8)
// The per-thread bodies live in lt_bodies.hpp and are shared with the
// replica-batched kernels.
#include "lt_internal.hpp"
#include "lt_bodies.hpp"

namespace lt {

namespace {

constexpr int CBS = 256;

__global__ void LT_LINK_KERNEL stout_tape_kernel(const double2* __restrict__ Uin, double2* __restrict__ Uout,
                                                 double2* __restrict__ T, Geom g, LinkMap m, double rho)
{
    body_stout_tape(Uin, Uout, T, g, m, link_block(m, blockIdx.x, gridDim.x), rho);
}

// one thread per (site, plane): the clover term of that plane.  Block b (XCD-remapped)
// takes plane b % 6 of site chunk b / 6, so the six planes of a chunk run in neighbouring
// blocks of one XCD and share its L2 instead of streaming the field once per plane.
__global__ void __launch_bounds__(CBS) clover_store_kernel(const double2* __restrict__ U, double2* __restrict__ C, Geom g)
{
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int s = (b / 6) * CBS + threadIdx.x;
    if (s < g.V) body_clover_store(U, C, g, s, b % 6);
}

// per site charge density, block partial sums
__global__ void __launch_bounds__(CBS) qdensity_kernel(const double2* __restrict__ C, Geom g, double* part)
{
    const int s = blockIdx.x * CBS + threadIdx.x;
    block_sum_to<CBS>(s < g.V ? body_qdensity_site(C, g, s) : 0.0, part + blockIdx.x);
}

__global__ void __launch_bounds__(CBS) sum_kernel(const double* part, int n, double* out)
{
    double a = 0.0;
    for (int i = threadIdx.x; i < n; i += CBS) a += part[i];
    block_sum_to<CBS>(a, out);
}

template <bool LAMBDA>
__global__ void LT_LINK_KERNEL clover_deriv_all_kernel(const double2* __restrict__ U, const double2* __restrict__ C,
                                                       double2* __restrict__ F, Geom g, LinkMap m, double scale_factor,
                                                       SigmaNext next)
{
    extern __shared__ double2 acc[];
    body_clover_deriv_all<LAMBDA>(U, C, F, g, m, link_block(m, blockIdx.x, gridDim.x), scale_factor, acc, next);
}

__global__ void LT_LINK_KERNEL lambda_kernel(const double2* __restrict__ U, const double2* __restrict__ Uprev,
                                             double2* __restrict__ Sig, const double2* __restrict__ T,
                                             double2* __restrict__ Lam, Geom g, LinkMap m)
{
    body_lambda(U, Uprev, Sig, T, Lam, g, m, link_block(m, blockIdx.x, gridDim.x));
}

template <int MODE>
__global__ void LT_LINK_KERNEL sigma_all_kernel(const double2* __restrict__ U, double2* __restrict__ Sig,
                                                const double2* __restrict__ T, const double2* __restrict__ Lam, Geom g,
                                                LinkMap m, double rho, MomentumKick kick, SigmaNext next)
{
    extern __shared__ double2 lds[];
    body_sigma_all<MODE>(U, Sig, T, Lam, g, m, link_block(m, blockIdx.x, gridDim.x), rho, lds, kick, next);
}

}  // namespace

void k_stout_tape(lt_ctx ctx, const Geom& g, const LinkMap& m, const double2* in, double2* out, double2* tape, double rho)
{
    launch(ctx, "stout_tape", stout_tape_kernel, dim3(8 * map_blocks(m)), dim3(m.nthreads), 0, in, out, tape, g, m, rho);
}

double k_clover_charge(lt_ctx ctx, const Geom& g, const double2* U, double2* C)
{
    launch(ctx, "clover_store", clover_store_kernel, dim3(6 * nblocks(g.V, CBS)), dim3(CBS), 0, U, C, g);
    unsigned nb = nblocks(g.V, CBS);
    double* part = ctx->scratch_for(nb + 1);
    launch(ctx, "qdensity", qdensity_kernel, dim3(nb), dim3(CBS), 0, (const double2*)C, g, part);
    launch(ctx, "qsum", sum_kernel, dim3(1), dim3(CBS), 0, (const double*)part, (int)nb, part + nb);
    double* h = ctx->host_for(1);
    LT_HIP_CHECK(hipMemcpyAsync(h, part + nb, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    LT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return -1.0 / (64.0 * 4.0 * M_PI * M_PI) * h[0];
}

void k_clover_deriv(lt_ctx ctx, const Geom& g, const LinkMap& m, const double2* U, const double2* C, double2* F,
                    double scale_factor, const SigmaNext* next)
{
    const dim3 grid(8 * map_blocks(m)), block(m.nthreads);
    const size_t lds = 18 * m.nthreads * sizeof(double2);
    if (next)
        launch(ctx, "clover_deriv", clover_deriv_all_kernel<true>, grid, block, lds, U, C, F, g, m, scale_factor, *next);
    else
        launch(ctx, "clover_deriv", clover_deriv_all_kernel<false>, grid, block, lds, U, C, F, g, m, scale_factor,
               SigmaNext{});
}

void k_stout_recursion_run(lt_ctx ctx, const Geom& g, const LinkMap& m, const LevelFn& level, const LevelFn& tape,
                           int hi, int lo, double2* Sig, double2* const lam[2], double rho, const MomentumKick* kick,
                           bool top_lambda_done)
{
    const dim3 grid(8 * map_blocks(m)), block(m.nthreads);
    const size_t lds = 18 * m.nthreads * sizeof(double2);
    int cur = 0;
    if (!top_lambda_done)
        launch(ctx, "stout_lambda", lambda_kernel, grid, block, 0, level(hi - 1), level(hi), Sig, tape(hi - 1), lam[cur],
               g, m);
    for (int k = hi; k > lo; --k) {
        const double2* U = level(k - 1);
        const double2* T = tape(k - 1);
        const double2* L = lam[cur];
        if (k - 1 > lo) {
            const SigmaNext next{level(k - 2), tape(k - 2), lam[cur ^ 1]};
            launch(ctx, "stout_sigma", sigma_all_kernel<SIGMA_LAMBDA>, grid, block, lds, U, Sig, T, L, g, m, rho,
                   MomentumKick{}, next);
            cur ^= 1;
        } else if (kick && k == 1) {
            launch(ctx, "stout_sigma", sigma_all_kernel<SIGMA_KICK>, grid, block, lds, U, Sig, T, L, g, m, rho, *kick,
                   SigmaNext{});
        } else {
            launch(ctx, "stout_sigma", sigma_all_kernel<SIGMA_STORE>, grid, block, lds, U, Sig, T, L, g, m, rho,
                   MomentumKick{}, SigmaNext{});
        }
    }
}

}  // namespace lt
