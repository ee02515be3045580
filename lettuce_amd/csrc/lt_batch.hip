// lt_batch.hip -- batched kernels: one launch advances every replica of a batch.
// Blocks are assigned to replicas by the prefix tables of BatchGrid (replicas may
// have different extents and tiles); inside a replica a block runs exactly the
// body the single-lattice kernel runs for the same logical block, so a replica's
// fields evolve bit for bit as they would alone.  Reductions keep the per-replica
// partial layout and finishing order of the single-lattice reductions.
//
// The scalar stages (dV/dQ, Hamiltonians, Metropolis, bias deposits, replica
// exchange) also run on the device, so a whole MetaD-HMC step needs no host
// round trip and can be captured in one graph.
#include "lt_batch.hpp"
#include "lt_bodies.hpp"

namespace lt {

namespace {

constexpr int RBS = 256;   // site / reduction / copy block size
constexpr int SBS = 64;    // scalar-stage block size

// the replica owning block b: the largest r with pre[r] <= b (every replica owns at least one block)
__device__ __forceinline__ int rep_of(const int* __restrict__ pre, int n, int b)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// ------------------------------------------------------------ link kernels

// Hardware blocks go round-robin over the 8 XCDs and every replica's block range
// starts at a multiple of 8, so remapping inside each replica gives every XCD a
// contiguous eighth of every replica: the L2 locality of a lattice run alone, and
// equal work per XCD whatever the mix of extents.
#define LT_BATCH_LINK_PROLOGUE                                                  \
    const int r = rep_of(bg.lb, bg.n, blockIdx.x);                              \
    const Geom g = bg.reps[r].g;                                                \
    const LinkMap m = bg.reps[r].m;                                             \
    const long long off = bg.reps[r].off;                                       \
    const int lb = link_block(m, blockIdx.x - bg.lb[r], bg.lb[r + 1] - bg.lb[r]);

__global__ void LT_LINK_KERNEL bk_momentum(double2* __restrict__ P, BatchGrid bg, const uint64_t* __restrict__ counter)
{
    LT_BATCH_LINK_PROLOGUE
    body_momentum(P + off, g, m, lb, bg.reps[r].seed, counter[r]);
}

__global__ void LT_LINK_KERNEL bk_force_update(const double2* __restrict__ U, double2* __restrict__ P,
                                               const double2* __restrict__ F, BatchGrid bg, double a, double bf)
{
    LT_BATCH_LINK_PROLOGUE
    body_force_update(U + off, P + off, F ? F + off : nullptr, g, m, lb, a, bf);
}

__global__ void LT_LINK_KERNEL bk_field_update(double2* __restrict__ U, const double2* __restrict__ P, BatchGrid bg,
                                               double eps)
{
    LT_BATCH_LINK_PROLOGUE
    body_field_update(U + off, P + off, g, m, lb, eps);
}

__global__ void LT_LINK_KERNEL bk_stout_tape(const double2* __restrict__ Uin, double2* __restrict__ Uout,
                                             double2* __restrict__ T, BatchGrid bg, double rho)
{
    LT_BATCH_LINK_PROLOGUE
    body_stout_tape(Uin + off, Uout + off, T + alg_offset(off), g, m, lb, rho);
}

template <bool LAMBDA>
__global__ void LT_LINK_KERNEL bk_clover_deriv_all(const double2* __restrict__ U, const double2* __restrict__ C,
                                                   double2* __restrict__ F, BatchGrid bg, const double* __restrict__ scale,
                                                   SigmaNext next)
{
    extern __shared__ double2 acc[];
    LT_BATCH_LINK_PROLOGUE
    if (LAMBDA) next = SigmaNext{next.U + off, next.T + alg_offset(off), next.Lam + alg_offset(off)};
    body_clover_deriv_all<LAMBDA>(U + off, C + clover_offset(off), F + off, g, m, lb, scale[r], acc, next);
}

__global__ void LT_LINK_KERNEL bk_lambda(const double2* __restrict__ U, const double2* __restrict__ Uprev,
                                         double2* __restrict__ Sig, const double2* __restrict__ T,
                                         double2* __restrict__ Lam, BatchGrid bg)
{
    LT_BATCH_LINK_PROLOGUE
    body_lambda(U + off, Uprev + off, Sig + off, T + alg_offset(off), Lam + alg_offset(off), g, m, lb);
}

template <int MODE>
__global__ void LT_LINK_KERNEL bk_sigma_all(const double2* __restrict__ U, double2* __restrict__ Sig,
                                            const double2* __restrict__ T, const double2* __restrict__ Lam, BatchGrid bg,
                                            double rho, MomentumKick kick, SigmaNext next)
{
    extern __shared__ double2 lds[];
    LT_BATCH_LINK_PROLOGUE
    if (MODE == SIGMA_KICK) kick.P += off;
    if (MODE == SIGMA_LAMBDA) next = SigmaNext{next.U + off, next.T + alg_offset(off), next.Lam + alg_offset(off)};
    body_sigma_all<MODE>(U + off, Sig + off, T + alg_offset(off), Lam + alg_offset(off), g, m, lb, rho, lds, kick, next);
}

#undef LT_BATCH_LINK_PROLOGUE

// ------------------------------------------------- site kernels, reductions

__device__ __forceinline__ void block_sum(double v, double* out)
{
    block_sum_to<RBS>(v, out);
}

__global__ void __launch_bounds__(RBS) bk_copy(const double2* __restrict__ src, double2* __restrict__ dst, BatchGrid bg,
                                               const int* __restrict__ accepted)
{
    const int r = rep_of(bg.cb, bg.n, blockIdx.x);
    if (accepted && accepted[r]) return;
    const long long n = 72LL * bg.reps[r].g.Vh;
    const long long i = (long long)(blockIdx.x - bg.cb[r]) * RBS + threadIdx.x;
    if (i < n) {
        const long long off = bg.reps[r].off;
        dst[off + i] = src[off + i];
    }
}

// block b (XCD-remapped) takes plane b % 6 of site block b / 6: the six planes of a site
// block in neighbouring blocks of one XCD
__global__ void __launch_bounds__(RBS) bk_clover_store(const double2* __restrict__ U, double2* __restrict__ C, BatchGrid bg)
{
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int sbk = b / 6;
    const int r = rep_of(bg.sb, bg.n, sbk);
    const Geom g = bg.reps[r].g;
    const long long off = bg.reps[r].off;
    const int s = (sbk - bg.sb[r]) * RBS + threadIdx.x;
    if (s < g.V) body_clover_store(U + off, C + clover_offset(off), g, s, b % 6);
}

__global__ void __launch_bounds__(RBS) bk_qdensity(const double2* __restrict__ C, BatchGrid bg, double* part)
{
    const int r = rep_of(bg.sb, bg.n, blockIdx.x);
    const Geom g = bg.reps[r].g;
    const long long off = bg.reps[r].off;
    const int s = (blockIdx.x - bg.sb[r]) * RBS + threadIdx.x;
    block_sum(s < g.V ? body_qdensity_site(C + clover_offset(off), g, s) : 0.0, part + blockIdx.x);
}

__global__ void __launch_bounds__(RBS) bk_plaq(const double2* __restrict__ U, BatchGrid bg, double* part)
{
    const int r = rep_of(bg.sb, bg.n, blockIdx.x);
    const Geom g = bg.reps[r].g;
    const long long off = bg.reps[r].off;
    const int s = (blockIdx.x - bg.sb[r]) * RBS + threadIdx.x;
    block_sum(s < g.V ? body_plaq_site(U + off, g, s) : 0.0, part + blockIdx.x);
}

__global__ void __launch_bounds__(RBS) bk_kinetic(const double2* __restrict__ P, BatchGrid bg, double* part)
{
    const int r = rep_of(bg.kb, bg.n, blockIdx.x);
    const Geom g = bg.reps[r].g;
    const long long off = bg.reps[r].off;
    const long long id = (long long)(blockIdx.x - bg.kb[r]) * RBS + threadIdx.x;
    block_sum(id < 4LL * g.V ? body_kinetic_link(P + off, g, id) : 0.0, part + blockIdx.x);
}

// block r: ordered sum of replica r's partials (same order as the single-lattice finish)
__global__ void __launch_bounds__(RBS) bk_finish(const double* __restrict__ part, const int* __restrict__ pre,
                                                 double* out)
{
    const int r = blockIdx.x;
    const double* p = part + pre[r];
    const int cnt = pre[r + 1] - pre[r];
    double acc = 0.0;
    for (int i = threadIdx.x; i < cnt; i += RBS) acc += p[i];
    block_sum(acc, out + r);
}

// ------------------------------------------------------------- scalar stages

// ReturnPotential / ReturnDerivative (metadynamics.hpp:151-187) with the host's
// operation order (no contraction), so host and device agree bit for bit.
__device__ double bias_potential(const DevBias& b, double cv)
{
#pragma clang fp contract(off)
    const int i = (int)floor((cv - b.cv_min) * b.bin_width_inverse);
    if ((unsigned)i < (unsigned)b.bins) {
        const double ic = (cv - (b.cv_min + i * b.bin_width)) * b.bin_width_inverse;
        return b.grid[i] * (1.0 - ic) + ic * b.grid[i + 1];
    }
    if (i < 0) {
        const double d = cv - b.cv_min;
        return b.grid[0] + b.threshold * d * d;
    }
    const double d = cv - b.cv_max;
    return b.grid[b.bins] + b.threshold * d * d;
}

__device__ double bias_derivative(const DevBias& b, double cv)
{
#pragma clang fp contract(off)
    const double h = 0.5 * b.bin_width;
    return b.bin_width_inverse * (bias_potential(b, cv + h) - bias_potential(b, cv - h));
}

// Q of the clover route: -1/(64 * 4 pi^2) * sum (topological_charge.hpp:199,230)
__device__ __forceinline__ double charge_of(double raw)
{
    return -1.0 / (64.0 * 4.0 * 3.14159265358979323846 * 3.14159265358979323846) * raw;
}

__device__ __forceinline__ double strength_of(const BatchState& st, int r) { return st.strength[st.label[r]]; }

__global__ void __launch_bounds__(SBS) bk_dvdq(BatchGrid bg, BatchState st, DevBias bias, int sample_index)
{
#pragma clang fp contract(off)
    for (int r = threadIdx.x; r < bg.n; r += SBS) {
        const double q = charge_of(st.qraw[r]);
        if (sample_index >= 0) st.samples[(size_t)r * st.n_sub + sample_index] = q;
        st.dvdq[r] = strength_of(st, r) * bias_derivative(bias, q);
    }
}

// H = beta (6V - plaq/3) + kinetic + strength * V(Q)   (hmc_gauge.hpp:525-541, hmc_metadynamics.hpp:557-559)
__device__ __forceinline__ double hamiltonian_of(const BatchGrid& bg, const BatchState& st, const DevBias& bias,
                                                 double beta, int r, double q)
{
#pragma clang fp contract(off)
    const double S = beta * (6.0 * (double)bg.reps[r].g.V - st.plaq[r] / 3.0);
    return S + st.kin[r] + strength_of(st, r) * bias_potential(bias, q);
}

__global__ void __launch_bounds__(SBS) bk_h0(BatchGrid bg, BatchState st, DevBias bias, double beta)
{
    for (int r = threadIdx.x; r < bg.n; r += SBS) {
        const double q = charge_of(st.qraw[r]);
        st.cv_old[r] = q;
        st.H0[r] = hamiltonian_of(bg, st, bias, beta, r, q);
    }
}

// Metropolis step of HMCMetaDKernel (hmc_metadynamics.hpp:615-637): accept if
// u <= exp(-(H1 - H0)) with u the Philox stream-3 draw of this trajectory; a
// rejected trajectory reports CV_old and replaces its path samples by CV_old.
__global__ void __launch_bounds__(SBS) bk_h1_accept(BatchGrid bg, BatchState st, DevBias bias, double beta,
                                                    int metropolis)
{
#pragma clang fp contract(off)
    for (int r = threadIdx.x; r < bg.n; r += SBS) {
        const double q = charge_of(st.qraw[r]);
        const double h1 = hamiltonian_of(bg, st, bias, beta, r, q);
        const double d = h1 - st.H0[r];
        int acc = 1;
        if (metropolis) {
            double u0, u1;
            draw2(bg.reps[r].seed, 0xFFFFFFFFu, 3u, 0u, 0u, 0u, st.counter[r], u0, u1);
            acc = u0 <= exp(-d);
        }
        st.H1[r] = h1;
        st.dH[r] = d;
        st.accepted[r] = acc;
        st.cv_cur[r] = acc ? q : st.cv_old[r];
        if (!acc)
            for (int k = 0; k < st.n_sub; ++k) st.samples[(size_t)r * st.n_sub + k] = st.cv_old[r];
    }
}

__device__ __forceinline__ int slot_of_label(const BatchState& st, int n, int k)
{
    for (int r = 0; r < n; ++r)
        if (st.label[r] == k) return r;
    return -1;
}

// UpdatePotentialSymmetric over the path samples of the replica holding the top
// label (metadynamics.hpp:103-124,126-149): per sample s, deposit at s then at -s.
// One thread per grid point; every point sees the deposits in the host's order.
__global__ void __launch_bounds__(RBS) bk_deposit(BatchGrid bg, BatchState st, DevBias bias, int n_labels)
{
#pragma clang fp contract(off)
    const int src = slot_of_label(st, bg.n, n_labels - 1);
    if (src < 0) return;
    const double wis = 1.0 / (bias.width * bias.width);
    const double* smp = st.samples + (size_t)src * st.n_sub;
    for (int b = threadIdx.x; b <= bias.bins; b += RBS) {
        const double x = bias.cv_min + b * bias.bin_width;
        double v = bias.grid[b];
        for (int k = 0; k < st.n_sub; ++k) {
            for (int sgn = 0; sgn < 2; ++sgn) {
                const double cv = sgn ? -smp[k] : smp[k];
                const double dist = cv - x;
                if (bias.well_tempered)
                    v += bias.weight * exp(-0.5 * wis * dist * dist - v / bias.wt_param);
                else
                    v += bias.weight * exp(-0.5 * wis * dist * dist);
            }
        }
        bias.grid[b] = v;
    }
    if (threadIdx.x == 0) {
        unsigned long long out = 0;
        for (int k = 0; k < st.n_sub; ++k) {
            if (smp[k] < bias.cv_min || smp[k] >= bias.cv_max) ++out;
            if (-smp[k] < bias.cv_min || -smp[k] >= bias.cv_max) ++out;
        }
        *st.exceeded += out;
    }
}

// Replica exchange by label swap, generalising MetadynamicsTemperingKernel
// (tempering.hpp:57-72) to a ladder of bias strengths s_0 < ... < s_{n-1}: the
// replicas i, j holding adjacent labels k, k+1 (and equal extents) swap labels
// with probability min(1, exp(-(s_k - s_{k+1}) (V(Q_j) - V(Q_i)))), u from Philox
// stream 4 at site k.  For s = (0, 1) this is the reference's rule
// exp(-(V(Q_unbiased) - V(Q_biased))).  Then every replica's counter advances.
//   do_swap: 0 none, 1 pairs (0,1),(2,3)..., 2 pairs (1,2),(3,4)..., 3 alternate by step
__global__ void __launch_bounds__(SBS) bk_swap_advance(BatchGrid bg, BatchState st, DevBias bias, int n_labels,
                                                       int do_swap)
{
#pragma clang fp contract(off)
    if (threadIdx.x != 0) return;
    const uint64_t c = *st.pt_counter;
    if (do_swap) {
        const int first = do_swap == 3 ? (int)(c & 1) : do_swap - 1;
        for (int k = first; k + 1 < n_labels; k += 2) {
            const int i = slot_of_label(st, bg.n, k), j = slot_of_label(st, bg.n, k + 1);
            if (i < 0 || j < 0) continue;
            bool same = true;
            for (int d = 0; d < 4; ++d) same = same && bg.reps[i].g.L[d] == bg.reps[j].g.L[d];
            if (!same) continue;
            const double delta = (st.strength[k] - st.strength[k + 1]) *
                                 (bias_potential(bias, st.cv_cur[j]) - bias_potential(bias, st.cv_cur[i]));
            double u0, u1;
            draw2(st.pt_seed, (uint32_t)k, 4u, 0u, 0u, 0u, c, u0, u1);
            const bool acc = u0 <= exp(-delta);
            st.swap_stats[2 * k] += 1;
            if (acc) {
                st.swap_stats[2 * k + 1] += 1;
                st.label[i] = k + 1;
                st.label[j] = k;
            }
        }
    }
    const int top = slot_of_label(st, bg.n, n_labels - 1);
    if (top >= 0) *st.bias_cv = st.cv_cur[top];
    for (int r = 0; r < bg.n; ++r) st.counter[r] += 1;
    *st.pt_counter = c + 1;
}

}  // namespace

void kb_copy_links(lt_ctx ctx, const BatchGrid& bg, int total_cb, const double2* src, double2* dst, const int* only_rejected)
{
    launch(ctx, "batch_copy", bk_copy, dim3(total_cb), dim3(RBS), 0, src, dst, bg, only_rejected);
}

static unsigned link_threads(const BatchGrid& bg, int)
{
    if (bg.nthreads < 1 || bg.nthreads > LT_LINK_BS) throw LtError{LT_ESTATE, "invalid batch tile"};
    return (unsigned)bg.nthreads;
}

void kb_momentum(lt_ctx ctx, const BatchGrid& bg, int total_lb, double2* P, const uint64_t* counter)
{
    launch(ctx, "batch_momentum", bk_momentum, dim3(total_lb), dim3(link_threads(bg, total_lb)), 0, P, bg, counter);
}

void kb_force_update(lt_ctx ctx, const BatchGrid& bg, int total_lb, const double2* U, double2* P, const double2* F,
                     double a, double b)
{
    launch(ctx, "batch_force_update", bk_force_update, dim3(total_lb), dim3(link_threads(bg, total_lb)), 0, U, P, F,
           bg, a, b);
}

void kb_field_update(lt_ctx ctx, const BatchGrid& bg, int total_lb, double2* U, const double2* P, double eps)
{
    launch(ctx, "batch_field_update", bk_field_update, dim3(total_lb), dim3(link_threads(bg, total_lb)), 0, U, P, bg,
           eps);
}

void kb_stout_tape(lt_ctx ctx, const BatchGrid& bg, int total_lb, const double2* in, double2* out, double2* tape,
                   double rho)
{
    launch(ctx, "batch_stout_tape", bk_stout_tape, dim3(total_lb), dim3(link_threads(bg, total_lb)), 0, in, out, tape,
           bg, rho);
}

void kb_clover_charge(lt_ctx ctx, const BatchGrid& bg, int total_sb, const double2* U, double2* C, double* part,
                      double* qraw)
{
    launch(ctx, "batch_clover_store", bk_clover_store, dim3(6 * total_sb), dim3(RBS), 0, U, C, bg);
    launch(ctx, "batch_qdensity", bk_qdensity, dim3(total_sb), dim3(RBS), 0, (const double2*)C, bg, part);
    launch(ctx, "batch_finish", bk_finish, dim3(bg.n), dim3(RBS), 0, (const double*)part, bg.sb, qraw);
}

void kb_clover_deriv(lt_ctx ctx, const BatchGrid& bg, int total_lb, const double2* U, const double2* C, double2* F,
                     const double* scale, const SigmaNext* next)
{
    const dim3 grid(total_lb), block(link_threads(bg, total_lb));
    const size_t lds = 18 * block.x * sizeof(double2);
    if (next)
        launch(ctx, "batch_clover_deriv", bk_clover_deriv_all<true>, grid, block, lds, U, C, F, bg, scale, *next);
    else
        launch(ctx, "batch_clover_deriv", bk_clover_deriv_all<false>, grid, block, lds, U, C, F, bg, scale, SigmaNext{});
}

void kb_stout_recursion_run(lt_ctx ctx, const BatchGrid& bg, int total_lb, const LevelFn& level, const LevelFn& tape,
                            int hi, int lo, double2* Sig, double2* const lam[2], double rho, const MomentumKick* kick,
                            bool top_lambda_done)
{
    const dim3 grid(total_lb), block(link_threads(bg, total_lb));
    const size_t lds = 18 * block.x * sizeof(double2);
    int cur = 0;
    if (!top_lambda_done)
        launch(ctx, "batch_stout_lambda", bk_lambda, grid, block, 0, level(hi - 1), level(hi), Sig, tape(hi - 1),
               lam[cur], bg);
    for (int k = hi; k > lo; --k) {
        const double2* U = level(k - 1);
        const double2* T = tape(k - 1);
        const double2* L = lam[cur];
        if (k - 1 > lo) {
            const SigmaNext next{level(k - 2), tape(k - 2), lam[cur ^ 1]};
            launch(ctx, "batch_stout_sigma", bk_sigma_all<SIGMA_LAMBDA>, grid, block, lds, U, Sig, T, L, bg, rho,
                   MomentumKick{}, next);
            cur ^= 1;
        } else if (kick && k == 1) {
            launch(ctx, "batch_stout_sigma", bk_sigma_all<SIGMA_KICK>, grid, block, lds, U, Sig, T, L, bg, rho, *kick,
                   SigmaNext{});
        } else {
            launch(ctx, "batch_stout_sigma", bk_sigma_all<SIGMA_STORE>, grid, block, lds, U, Sig, T, L, bg, rho,
                   MomentumKick{}, SigmaNext{});
        }
    }
}

void kb_plaquette(lt_ctx ctx, const BatchGrid& bg, int total_sb, const double2* U, double* part, double* out)
{
    launch(ctx, "batch_plaquette", bk_plaq, dim3(total_sb), dim3(RBS), 0, U, bg, part);
    launch(ctx, "batch_finish", bk_finish, dim3(bg.n), dim3(RBS), 0, (const double*)part, bg.sb, out);
}

void kb_kinetic(lt_ctx ctx, const BatchGrid& bg, int total_kb, const double2* P, double* part, double* out)
{
    launch(ctx, "batch_kinetic", bk_kinetic, dim3(total_kb), dim3(RBS), 0, P, bg, part);
    launch(ctx, "batch_finish", bk_finish, dim3(bg.n), dim3(RBS), 0, (const double*)part, bg.kb, out);
}

void kb_dvdq(lt_ctx ctx, const BatchGrid& bg, const BatchState& st, const DevBias& bias, int sample_index)
{
    launch(ctx, "batch_dvdq", bk_dvdq, dim3(1), dim3(SBS), 0, bg, st, bias, sample_index);
}

void kb_h0(lt_ctx ctx, const BatchGrid& bg, const BatchState& st, const DevBias& bias, double beta)
{
    launch(ctx, "batch_h0", bk_h0, dim3(1), dim3(SBS), 0, bg, st, bias, beta);
}

void kb_h1_accept(lt_ctx ctx, const BatchGrid& bg, const BatchState& st, const DevBias& bias, double beta,
                  int metropolis)
{
    launch(ctx, "batch_accept", bk_h1_accept, dim3(1), dim3(SBS), 0, bg, st, bias, beta, metropolis);
}

void kb_deposit(lt_ctx ctx, const BatchGrid& bg, const BatchState& st, const DevBias& bias, int n_labels)
{
    launch(ctx, "batch_deposit", bk_deposit, dim3(1), dim3(RBS), 0, bg, st, bias, n_labels);
}

void kb_swap_advance(lt_ctx ctx, const BatchGrid& bg, const BatchState& st, const DevBias& bias, int n_labels,
                     int do_swap)
{
    launch(ctx, "batch_swap", bk_swap_advance, dim3(1), dim3(SBS), 0, bg, st, bias, n_labels, do_swap);
}

}  // namespace lt
