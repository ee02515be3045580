// lt_update.hip -- checkerboarded local updates: Wilson staple, SU(2)-subgroup
// overrelaxation and Cabibbo-Marinari pseudo-heatbath with Philox draws.
//
// One launch updates one colour class of one direction mu.  With the 2-colour
// schedule a colour class is exactly one parity array, so every thread owns one
// link and writes it in place; the staple only reads links of other directions or
// of the opposite parity, so the update is race free.  The 4-colour schedule
// (colour (sum + x_mu) mod 4) is visited with one thread per site of the class.
#include "lt_internal.hpp"
#include "lt_bodies.hpp"

namespace lt {

namespace {

constexpr int BS = LT_LINK_BS;

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

// Kennedy-Pendleton / Fabricius-Haan SU(2) sampler in the reference's draw order:
// attempt a uses Philox blocks 2a (r1, r2) and 2a+1 (r3, r0); the direction uses
// block 0xFFFFF.  After max_iter+1 rejected attempts the identity is returned.
__device__ __forceinline__ SU2 heatbath_su2(SU2 A, double prefactor, uint64_t seed, uint32_t site, uint32_t sub,
                                            uint32_t mu, uint64_t counter)
{
    constexpr int max_iter = 10;
    const double an = 1.0 / sqrt(su2_det(A));
    SU2 V{an * A.e11, an * A.e12};
    double lambda_sq = 0.0;
    int it = 0;
    for (;;) {
        double ua0, ua1, ub0, ub1;
        draw2(seed, site, RNG_HEATBATH, (uint32_t)(2 * it), sub, mu, counter, ua0, ua1);
        draw2(seed, site, RNG_HEATBATH, (uint32_t)(2 * it + 1), sub, mu, counter, ub0, ub1);
        double x1 = log(1.0 - ua0);
        double x2 = cos(2.0 * 3.14159265358979323846 * (1.0 - ua1));
        double x3 = log(1.0 - ub0);
        lambda_sq = (-0.25 * prefactor * an) * (x1 + x2 * x2 * x3);
        double r0 = ub1;
        if (it > max_iter) return SU2{cd{1.0, 0.0}, cd{0.0, 0.0}};
        ++it;
        if (!(r0 * r0 + lambda_sq >= 1.0)) break;
    }
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
    stu(U, g, s, mu, par, L);
}

#ifndef LT_SWEEP_WAVES
#define LT_SWEEP_WAVES LT_LINK_WAVES
#endif
template <bool HEATBATH>
__global__ void __launch_bounds__(LT_LINK_BS) __attribute__((amdgpu_waves_per_eu(LT_SWEEP_WAVES)))
sweep_cb2_kernel(double2* __restrict__ U, Geom g, LinkMap m, int mu, int par, double prefactor, uint64_t seed,
                 uint64_t counter)
{
    int c[4], s;
    if (!map_site(g, m, par, c, s)) return;
    update_link<HEATBATH>(U, g, s, c, mu, par, prefactor, seed, counter);
}

template <bool HEATBATH>
__global__ void LT_LINK_KERNEL sweep_cb4_kernel(double2* __restrict__ U, Geom g, int mu, int colour,
                                                 double prefactor, uint64_t seed, uint64_t counter)
{
    int c[4], s;
    if (!cb4_site(g, mu, colour, xcd_block(blockIdx.x, gridDim.x) * BS + threadIdx.x, c, s)) return;
    update_link<HEATBATH>(U, g, s, c, mu, parity_of(c), prefactor, seed, counter);
}

// Heatbath colour kernels of a captured update: the Philox counter is read from
// device memory (state[0], advanced by the graph itself) instead of a launch argument,
// so one captured graph serves every replay; sweep k of an update draws with
// counter state[0] + k, as the eager sweeps draw with the field's counter.
__global__ void __launch_bounds__(LT_LINK_BS) __attribute__((amdgpu_waves_per_eu(LT_SWEEP_WAVES)))
hb_cb2_graph_kernel(double2* __restrict__ U, Geom g, LinkMap m, int mu, int par, double prefactor, uint64_t seed,
                    const uint64_t* __restrict__ state, int k)
{
    int c[4], s;
    if (!map_site(g, m, par, c, s)) return;
    update_link<true>(U, g, s, c, mu, par, prefactor, seed, state[0] + (uint64_t)k);
}

__global__ void LT_LINK_KERNEL hb_cb4_graph_kernel(double2* __restrict__ U, Geom g, int mu, int colour, double prefactor,
                                                   uint64_t seed, const uint64_t* __restrict__ state, int k)
{
    int c[4], s;
    if (!cb4_site(g, mu, colour, xcd_block(blockIdx.x, gridDim.x) * BS + threadIdx.x, c, s)) return;
    update_link<true>(U, g, s, c, mu, parity_of(c), prefactor, seed, state[0] + (uint64_t)k);
}

// PlaquetteSum of the captured update: per-block sums (the block tree of the HMC
// reductions), then one ordered sum written to plaq[state[1]]
constexpr int RBS = 256;
__device__ __forceinline__ void block_sum1(double v, double* out)
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
__global__ void __launch_bounds__(RBS) upd_plaq_kernel(const double2* __restrict__ U, Geom g, double* part)
{
    const int s = blockIdx.x * RBS + threadIdx.x;
    block_sum1(s < g.V ? body_plaq_site(U, g, s) : 0.0, part + blockIdx.x);
}
__global__ void __launch_bounds__(RBS) upd_plaq_finish_kernel(const double* __restrict__ part, int n,
                                                              const uint64_t* __restrict__ state, double* plaq)
{
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += RBS) acc += part[i];
    block_sum1(acc, plaq + state[1]);
}
// end of a captured update: the next one draws n_hb counters later and writes the next slot
__global__ void upd_advance_kernel(uint64_t* state, int n_hb)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        state[0] += (uint64_t)n_hb;
        state[1] += 1;
    }
}

// Checkerboard4 sweeps with Rectangular<R>(beta, c_plaq, c_rect): the action enters
// through the staple only (heatbath.hpp:110-129, overrelaxation.hpp:91-116)
template <bool HEATBATH>
__device__ __forceinline__ void update_link_action(double2* __restrict__ U, const Geom& g, int s, const int c[4],
                                                   int mu, int par, double prefactor, uint64_t seed, uint64_t counter,
                                                   double c_plaq, double c_rect)
{
    const M3 st = staple_action(U, g, s, c, mu, par, c_plaq, c_rect);
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

template <bool HEATBATH>
__global__ void __launch_bounds__(BS) sweep_action_kernel(double2* __restrict__ U, Geom g, int mu, int colour,
                                                          double prefactor, uint64_t seed, uint64_t counter,
                                                          double c_plaq, double c_rect)
{
    int c[4], s;
    if (!cb4_site(g, mu, colour, xcd_block(blockIdx.x, gridDim.x) * BS + threadIdx.x, c, s)) return;
    update_link_action<HEATBATH>(U, g, s, c, mu, parity_of(c), prefactor, seed, counter, c_plaq, c_rect);
}

__global__ void __launch_bounds__(BS) staple_action_kernel(const double2* __restrict__ U, double2* __restrict__ out,
                                                           Geom g, double c_plaq, double c_rect)
{
    const long long id = (long long)blockIdx.x * BS + threadIdx.x;
    if (id >= 4LL * g.V) return;
    const int s = (int)(id >> 2), mu = (int)(id & 3);
    int c[4];
    coords_of(g, s, c);
    const int par = parity_of(c);
    st_link(out, g, s, mu, par, staple_action(U, g, s, c, mu, par, c_plaq, c_rect));
}

__global__ void LT_LINK_KERNEL staple_field_kernel(const double2* __restrict__ U, double2* __restrict__ out, Geom g,
                                                   LinkMap m)
{
    int mu = blockIdx.y >> 1, par = blockIdx.y & 1;
    int c[4], s;
    if (!map_site(g, m, par, c, s)) return;
    st_link(out, g, s, mu, par, staple(U, g, s, c, mu));
}

template <bool HEATBATH>
void sweeps(lt_field u, int n, int schedule, double prefactor, const char* name)
{
    lt_ctx ctx = u->ctx;
    const Geom& g = u->g;
    for (int k = 0; k < n; ++k) {
        uint64_t counter = HEATBATH ? u->counter++ : 0;
        for (int mu = 0; mu < 4; ++mu) {
            if (schedule == LT_CB4) {
                for (int col = 0; col < 4; ++col)
                    launch(ctx, name, sweep_cb4_kernel<HEATBATH>, dim3(nblocks(g.V / 4, BS)), dim3(BS), 0, u->d, g, mu, col,
                           prefactor, u->seed, counter);
            } else {
                const LinkMap& m = u->map;
                for (int par = 0; par < 2; ++par)
                    launch(ctx, name, sweep_cb2_kernel<HEATBATH>, dim3(map_blocks(m)), dim3(m.nthreads), 0, u->d, g, m, mu,
                           par, prefactor, u->seed, counter);
            }
        }
    }
}

template <bool HEATBATH>
void sweeps_action(lt_field u, int n, double prefactor, double c_plaq, double c_rect, const char* name)
{
    lt_ctx ctx = u->ctx;
    const Geom& g = u->g;
    for (int k = 0; k < n; ++k) {
        uint64_t counter = HEATBATH ? u->counter++ : 0;
        for (int mu = 0; mu < 4; ++mu)
            for (int col = 0; col < 4; ++col)
                launch(ctx, name, sweep_action_kernel<HEATBATH>, dim3(nblocks(g.V / 4, BS)), dim3(BS), 0, u->d, g, mu, col,
                       prefactor, u->seed, counter, c_plaq, c_rect);
    }
}

}  // namespace

void k_staple_field_action(lt_field u, lt_field out, double c_plaq, double c_rect)
{
    launch(u->ctx, "staple_action", staple_action_kernel, dim3(nblocks(4LL * u->g.V, BS)), dim3(BS), 0,
           (const double2*)u->d, out->d, u->g, c_plaq, c_rect);
}

void k_or_sweep_action(lt_field u, int n, double c_plaq, double c_rect)
{
    sweeps_action<false>(u, n, 0.0, c_plaq, c_rect, "or_action");
}

void k_heatbath_sweep_action(lt_field u, double beta, int n, double c_plaq, double c_rect)
{
    sweeps_action<true>(u, n, 3.0 / beta, c_plaq, c_rect, "hb_action");
}

void k_staple_field(lt_field u, lt_field out)
{
    const LinkMap& m = u->map;
    launch(u->ctx, "staple_field", staple_field_kernel, dim3(map_blocks(m), 8), dim3(m.nthreads), 0,
           (const double2*)u->d, out->d, u->g, m);
}

// ---- the local-update loop body as one replayed graph --------------------------
struct UpdateGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // what the captured launches depend on
    const double2* d = nullptr;
    LinkMap m{};
    uint64_t seed = 0;
    double beta = 0.0;
    int n_hb = -1, n_or = -1, schedule = 0, with_plaq = 0;
    // device memory: state[0] = Philox counter, state[1] = plaquette slot; partials; plaquette sums
    uint64_t* state = nullptr;
    double* part = nullptr;
    double* plaq = nullptr;
    size_t cap = 0;
    unsigned nb = 0;

    void drop_exec()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~UpdateGraph()
    {
        drop_exec();
        if (state) (void)hipFree(state);
        if (part) (void)hipFree(part);
        if (plaq) (void)hipFree(plaq);
    }
};

void drop_update_graph(lt_field f)
{
    delete f->ug;
    f->ug = nullptr;
}

namespace {

bool same_map(const LinkMap& a, const LinkMap& b)
{
    for (int i = 0; i < 4; ++i)
        if (a.T[i] != b.T[i] || a.N[i] != b.N[i]) return false;
    return a.nthreads == b.nthreads && a.tc == b.tc && a.yc == b.yc;
}

// the launches of one update, in the order of the eager sweeps
void enqueue_update(lt_field u, UpdateGraph& G, double prefactor)
{
    lt_ctx ctx = u->ctx;
    const Geom& g = u->g;
    const LinkMap& m = u->map;
    for (int k = 0; k < G.n_hb; ++k)
        for (int mu = 0; mu < 4; ++mu) {
            if (G.schedule == LT_CB4) {
                for (int col = 0; col < 4; ++col)
                    launch(ctx, "hb_cb4", hb_cb4_graph_kernel, dim3(nblocks(g.V / 4, BS)), dim3(BS), 0, u->d, g, mu,
                           col, prefactor, u->seed, (const uint64_t*)G.state, k);
            } else {
                for (int par = 0; par < 2; ++par)
                    launch(ctx, "hb_cb2", hb_cb2_graph_kernel, dim3(map_blocks(m)), dim3(m.nthreads), 0, u->d, g, m,
                           mu, par, prefactor, u->seed, (const uint64_t*)G.state, k);
            }
        }
    sweeps<false>(u, G.n_or, G.schedule, 0.0, G.schedule == LT_CB4 ? "or_cb4" : "or_cb2");
    if (G.with_plaq) {
        launch(ctx, "plaquette", upd_plaq_kernel, dim3(G.nb), dim3(RBS), 0, (const double2*)u->d, g, G.part);
        launch(ctx, "reduce_finish", upd_plaq_finish_kernel, dim3(1), dim3(RBS), 0, (const double*)G.part, (int)G.nb,
               (const uint64_t*)G.state, G.plaq);
    }
    launch(ctx, "update_advance", upd_advance_kernel, dim3(1), dim3(64), 0, G.state, G.n_hb);
}

}  // namespace

void k_update_graph(lt_field u, double beta, int n_hb, int n_or, int schedule, int n_update, double* plaq_host)
{
    lt_ctx ctx = u->ctx;
    if (!u->ug) u->ug = new UpdateGraph();
    UpdateGraph& G = *u->ug;
    const int with_plaq = plaq_host ? 1 : 0;
    if (!G.state) {
        G.nb = nblocks(u->g.V, RBS);
        if (hipMalloc(&G.state, 2 * sizeof(uint64_t)) != hipSuccess || hipMalloc(&G.part, G.nb * sizeof(double)) != hipSuccess)
            throw LtError{LT_ENOMEM, "update graph allocation failed"};
    }
    if (with_plaq && (size_t)n_update > G.cap) {
        G.drop_exec();
        if (G.plaq) LT_HIP_CHECK(hipFree(G.plaq));
        G.plaq = nullptr;
        G.cap = 0;
        if (hipMalloc(&G.plaq, (size_t)n_update * sizeof(double)) != hipSuccess)
            throw LtError{LT_ENOMEM, "update graph allocation failed"};
        G.cap = (size_t)n_update;
    }
    if (!G.exec || G.d != u->d || !same_map(G.m, u->map) || G.seed != u->seed || G.beta != beta || G.n_hb != n_hb ||
        G.n_or != n_or || G.schedule != schedule || G.with_plaq != with_plaq) {
        G.drop_exec();
        G.d = u->d;
        G.m = u->map;
        G.seed = u->seed;
        G.beta = beta;
        G.n_hb = n_hb;
        G.n_or = n_or;
        G.schedule = schedule;
        G.with_plaq = with_plaq;
        const bool prof = ctx->profiling;
        ctx->profiling = false;
        LT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        LT_HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        try {
            enqueue_update(u, G, 3.0 / beta);
        } catch (...) {
            hipGraph_t gr = nullptr;
            (void)hipStreamEndCapture(ctx->stream, &gr);
            if (gr) (void)hipGraphDestroy(gr);
            ctx->profiling = prof;
            throw;
        }
        ctx->profiling = prof;
        LT_HIP_CHECK(hipStreamEndCapture(ctx->stream, &G.graph));
        LT_HIP_CHECK(hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0));
    }
    const uint64_t st[2] = {u->counter, 0};
    LT_HIP_CHECK(hipMemcpyAsync(G.state, st, sizeof(st), hipMemcpyHostToDevice, ctx->stream));
    for (int i = 0; i < n_update; ++i) LT_HIP_CHECK(hipGraphLaunch(G.exec, ctx->stream));
    u->counter += (uint64_t)n_hb * (uint64_t)n_update;
    if (plaq_host)
        LT_HIP_CHECK(hipMemcpyAsync(plaq_host, G.plaq, (size_t)n_update * sizeof(double), hipMemcpyDeviceToHost,
                                    ctx->stream));
    LT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

void k_or_sweep(lt_field u, int n, int schedule)
{
    sweeps<false>(u, n, schedule, 0.0, schedule == LT_CB4 ? "or_cb4" : "or_cb2");
}

void k_hb_sweep(lt_field u, double beta, int n, int schedule)
{
    sweeps<true>(u, n, schedule, 3.0 / beta, schedule == LT_CB4 ? "hb_cb4" : "hb_cb2");
}

}  // namespace lt
