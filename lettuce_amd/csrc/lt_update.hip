// lt_update.hip -- checkerboarded local updates: Wilson staple, SU(2)-subgroup
// overrelaxation and Cabibbo-Marinari pseudo-heatbath with Philox draws.
//
// One launch updates one colour class of one direction mu.  With the 2-colour
// schedule a colour class is exactly one parity array, so every thread owns one
// link and writes it in place; the staple only reads links of other directions or
// of the opposite parity, so the update is race free.  The 4-colour schedule
// (colour (sum + x_mu) mod 4) runs in lt_cb4.hip, one thread per site of the class,
// on the field reordered for the duration of the call.
#include "lt_internal.hpp"
#include "lt_bodies.hpp"
#include "lt_sweep.hpp"

namespace lt {

namespace {

constexpr int BS = LT_LINK_BS;

#ifndef LT_SWEEP_WAVES
#define LT_SWEEP_WAVES LT_LINK_WAVES
#endif
// Every launch of a sweep reads all eight half-fields, so the parity-1 launch of each
// direction walks the tiles backwards (rev, xcd_block_dir) and starts on the lines the
// parity-0 launch touched last, which are still in the L2 and the memory-side cache
// (overrelaxation colour launch -7 % at 32^4).
template <bool HEATBATH>
__global__ void __launch_bounds__(LT_LINK_BS) __attribute__((amdgpu_waves_per_eu(LT_SWEEP_WAVES)))
sweep_cb2_kernel(double2* __restrict__ U, Geom g, LinkMap m, int mu, int par, double prefactor, uint64_t seed,
                 uint64_t counter, int rev)
{
    int c[4], s;
    if (!map_site_dir(g, m, par, rev, c, s)) return;
    update_link<HEATBATH>(U, g, s, c, mu, par, prefactor, seed, counter);
}

// Heatbath colour kernels of a captured update: the Philox counter is read from
// device memory (state[0], advanced by the graph itself) instead of a launch argument,
// so one captured graph serves every replay; sweep k of an update draws with
// counter state[0] + k, as the eager sweeps draw with the field's counter.
__global__ void __launch_bounds__(LT_LINK_BS) __attribute__((amdgpu_waves_per_eu(LT_SWEEP_WAVES)))
hb_cb2_graph_kernel(double2* __restrict__ U, Geom g, LinkMap m, int mu, int par, double prefactor, uint64_t seed,
                    const uint64_t* __restrict__ state, int k, int rev)
{
    int c[4], s;
    if (!map_site_dir(g, m, par, rev, c, s)) return;
    update_link<true>(U, g, s, c, mu, par, prefactor, seed, state[0] + (uint64_t)k);
}

// PlaquetteSum of the captured update: per-block sums (the block tree of the HMC
// reductions), then one ordered sum written to plaq[state[1]]
constexpr int RBS = 256;
__device__ __forceinline__ void block_sum1(double v, double* out)
{
    block_sum_to<RBS>(v, out);
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

// The colour launches walk all y tiles of an x step in one chunk (LinkMap::yc = N[2]):
// with the alternating walk (rev) that is 2 % faster per HB + 4 OR update at 32^4 than
// the two-tile chunks the Jacobi kernels keep (the RK3 flow stage is 1.5 % slower with one
// chunk).  The tile order never changes a result.
LinkMap sweep_map(const LinkMap& m)
{
    LinkMap r = m;
    r.yc = r.N[2];
    return r;
}

// n sweeps; the four-colour ones (split) expect the field in the split order
template <bool HEATBATH>
void sweep_launches(lt_field u, int n, int schedule, double prefactor, const char* name)
{
    lt_ctx ctx = u->ctx;
    const Geom& g = u->g;
    for (int k = 0; k < n; ++k) {
        uint64_t counter = HEATBATH ? u->counter++ : 0;
        for (int mu = 0; mu < 4; ++mu) {
            if (schedule == LT_CB4) {
                for (int col = 0; col < 4; ++col) k_cb4_colour(u, HEATBATH, mu, col, prefactor, counter, name);
            } else {
                const LinkMap m = sweep_map(u->map);
                for (int par = 0; par < 2; ++par)
                    launch(ctx, name, sweep_cb2_kernel<HEATBATH>, dim3(map_blocks(m)), dim3(m.nthreads), 0, u->d, g, m, mu,
                           par, prefactor, u->seed, counter, par);
            }
        }
    }
}

template <bool HEATBATH>
void sweeps(lt_field u, int n, int schedule, double prefactor, const char* name)
{
    if (schedule == LT_CB4 && n > 0) {
        k_cb_layout(u, 1);
        sweep_launches<HEATBATH>(u, n, schedule, prefactor, name);
        k_cb_layout(u, 0);
    } else {
        sweep_launches<HEATBATH>(u, n, schedule, prefactor, name);
    }
}

template <bool HEATBATH>
void sweeps_action(lt_field u, int n, double prefactor, double c_plaq, double c_rect, const char* name)
{
    if (n <= 0) return;
    k_cb_layout(u, 1);
    for (int k = 0; k < n; ++k) {
        uint64_t counter = HEATBATH ? u->counter++ : 0;
        for (int mu = 0; mu < 4; ++mu)
            for (int col = 0; col < 4; ++col) k_cb4_colour_action(u, HEATBATH, mu, col, prefactor, counter, c_plaq, c_rect, name);
    }
    k_cb_layout(u, 0);
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
    const LinkMap m = sweep_map(u->map);
    for (int k = 0; k < G.n_hb; ++k)
        for (int mu = 0; mu < 4; ++mu) {
            if (G.schedule == LT_CB4) {
                for (int col = 0; col < 4; ++col) k_cb4_colour_graph_hb(u, mu, col, prefactor, G.state, k);
            } else {
                for (int par = 0; par < 2; ++par)
                    launch(ctx, "hb_cb2", hb_cb2_graph_kernel, dim3(map_blocks(m)), dim3(m.nthreads), 0, u->d, g, m,
                           mu, par, prefactor, u->seed, (const uint64_t*)G.state, k, par);
            }
        }
    sweep_launches<false>(u, G.n_or, G.schedule, 0.0, G.schedule == LT_CB4 ? "or_cb4" : "or_cb2");
    if (G.with_plaq) {
        if (G.schedule == LT_CB4)
            k_cb4_plaq_partials(u, G.part, G.nb);
        else
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
    // a four-colour graph runs on the split order (its kernels are lt_cb4.hip's)
    if (schedule == LT_CB4) k_cb_layout(u, 1);
    for (int i = 0; i < n_update; ++i) LT_HIP_CHECK(hipGraphLaunch(G.exec, ctx->stream));
    if (schedule == LT_CB4) k_cb_layout(u, 0);
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
