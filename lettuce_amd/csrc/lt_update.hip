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

// ---- small lattices ---------------------------------------------------------------
// A colour launch of a few thousand links leaves most of the GPU idle, and its duration
// is one wavefront's serial work: about 2400 VALU instructions for an overrelaxation
// link, twice that with the heatbath sampler (8^4: 7.8 and 17.8 us per launch).  Up to
// SMALL_VH links per colour launch, four lanes share a link: lane j < 3 builds the upper
// and the lower staple term of the j-th direction nu != mu, a two-step sum inside the
// quad (DPP quad permutations; every lane ends with the same ((t0 + t1) + (t2 + 0)),
// since a + b = b + a) gives all four lanes the staple, all four apply the SU(2) steps
// -- the heatbath's attempts base + j run side by side and the first accepted one is
// taken, exactly the attempt the sequential loop would take -- and lane 0 stores.
#ifndef LT_SMALL_VH
#define LT_SMALL_VH 16384
#endif
constexpr int SMALL_VH = LT_SMALL_VH;
constexpr int SBS = 256;

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
// a <- a + a of lane (lane ^ 1) (CTRL 0xB1) or (lane ^ 2) (CTRL 0x4E)
template <int CTRL>
__device__ __forceinline__ void quad_add(M3& a)
{
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        a.m[k].re += dpp_f64<CTRL>(a.m[k].re);
        a.m[k].im += dpp_f64<CTRL>(a.m[k].im);
    }
}

// upper + lower staple term of direction nu for link (s, mu)
__device__ __forceinline__ M3 staple_pair(const double2* __restrict__ U, const Geom& g, int s, const int c[4], int mu,
                                          int par, int nu)
{
    const int q = par ^ 1;
    const int cnu = pick4(c, nu), Ln = pick4(g.L, nu), stn = pick4(g.st, nu);
    const int s_nu = (cnu + 1 < Ln) ? s + stn : s - (Ln - 1) * stn;
    const int s_nd = (cnu > 0) ? s - stn : s + (Ln - 1) * stn;
    const int cmu = pick4(c, mu), Lm = pick4(g.L, mu), stm = pick4(g.st, mu);
    const int dmu = (cmu + 1 < Lm) ? stm : -(Lm - 1) * stm;   // site step +mu
    const int s_mu = s + dmu, s_nd_mu = s_nd + dmu;
    M3 t = mul_nd_r01(mul_r01(ldu2(U, g, s, nu, par), ldu(U, g, s_nu, mu, q)), ldu(U, g, s_mu, nu, q));
    third_row(t);
    M3 ef = mul_r01(ldu2(U, g, s_nd, mu, q), ldu(U, g, s_nd_mu, nu, par));
    third_row(ef);
    mul_dn_acc(t, ldu(U, g, s_nd, nu, q), ef);
    return t;
}

// One colour launch (mu, parity) of a small lattice, four lanes per link; the heatbath
// draws with counter state[0] + k when state is given (captured updates), else counter.
template <bool HEATBATH>
__global__ void __launch_bounds__(SBS) sweep_small_kernel(double2* __restrict__ U, Geom g, int mu, int par,
                                                          double prefactor, uint64_t seed, uint64_t counter,
                                                          const uint64_t* __restrict__ state, int k)
{
    const int tid = blockIdx.x * SBS + threadIdx.x;
    const int l = tid >> 2, j = tid & 3;
    if (l >= g.Vh) return;   // whole quads (4 Vh threads)
    if (HEATBATH && state) counter = state[0] + (uint64_t)k;
    int c[4];
    const int s = site_of_cb(g, l, par, c);
    M3 st = j < 3 ? staple_pair(U, g, s, c, mu, par, j + (j >= mu)) : m3_zero();
    quad_add<0xB1>(st);
    quad_add<0x4E>(st);
    M3 L = ldu(U, g, s, mu, par);
    const int lane = __lane_id();
#pragma unroll
    for (int sub = 0; sub < 3; ++sub) {
        const int i = sub == 2 ? 1 : 0;
        const int jj = sub == 0 ? 1 : 2;
        const SU2 A = extract(L, st, i, jj);
        SU2 X;
        if (HEATBATH) {
            const double an = 1.0 / sqrt(su2_det(A));
            const SU2 V{an * A.e11, an * A.e12};
            const double coef = -0.25 * prefactor * an;
            double lambda_sq = 0.0;
            bool found = false;
            for (int base = 0; base <= KP_MAX_ITER; base += 4) {
                double lsq = 0.0;
                const int kk = base + j;
                const bool acc = kk <= KP_MAX_ITER &&
                                 kp_attempt(seed, (uint32_t)s, (uint32_t)sub, (uint32_t)mu, counter, kk, coef, lsq);
                const int qbits = (int)((__ballot(acc) >> (lane & ~3)) & 0xFull);
                const double first = __shfl(lsq, qbits ? (lane & ~3) + __ffs(qbits) - 1 : lane);
                if (qbits) {
                    lambda_sq = first;
                    found = true;
                    break;
                }
            }
            X = found ? kp_finish(V, lambda_sq, seed, (uint32_t)s, (uint32_t)sub, (uint32_t)mu, counter)
                      : SU2{cd{1.0, 0.0}, cd{0.0, 0.0}};
        } else {
            X = overrelax(A);
        }
        embed_mul(L, X, i, jj);
    }
    gram_schmidt(L);
    if (j == 0) stu(U, g, s, mu, par, L);
}

// PlaquetteSum of the captured update: per-block sums (the block tree of the HMC
// reductions), then one ordered sum written to plaq[state[1]]
constexpr int RBS = 256;
__device__ __forceinline__ void block_sum1(double v, double* out)
{
    block_sum_to<RBS>(v, out);
}
// plaquette sum over the field's link map, one partial per block
__global__ void LT_LINK_KERNEL upd_plaq_kernel(const double2* __restrict__ U, Geom g, LinkMap m, double* part)
{
    plaq_link_partials(U, g, m, part);
}
unsigned plaq_parts(const LinkMap& m) { return 8u * map_blocks(m); }
// the sum of part[0, n) over one block, in a fixed order: lane i keeps eight running sums
// of the entries i + RBS u (u < 8) of each group of 8 RBS entries (eight loads in flight
// instead of one), the rest into the first, then a fixed pairing and the block tree
__device__ __forceinline__ double partials_sum(const double* __restrict__ part, int n)
{
    double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int i = threadIdx.x;
    for (; i + 7 * RBS < n; i += 8 * RBS)
#pragma unroll
        for (int u = 0; u < 8; ++u) a[u] += part[i + u * RBS];
    for (; i < n; i += RBS) a[0] += part[i];
    return ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
}
__global__ void __launch_bounds__(RBS) plaq_sum_kernel(const double* __restrict__ part, int n, double* out)
{
    block_sum1(partials_sum(part, n), out);
}
// the same sum into the plaquette slot of a captured update
__global__ void __launch_bounds__(RBS) upd_plaq_finish_kernel(const double* __restrict__ part, int n,
                                                              const uint64_t* __restrict__ state, double* plaq)
{
    block_sum1(partials_sum(part, n), plaq + state[1]);
}
// end of a captured update: the next one draws n_hb counters later and writes the next slot
__global__ void upd_advance_kernel(uint64_t* state, int n_hb)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        state[0] += (uint64_t)n_hb;
        state[1] += 1;
    }
}

// the rectangle staple's sum waits in LDS between its terms (9 x 16 bytes per lane)
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(2)))
staple_action_kernel(const double2* __restrict__ U, double2* __restrict__ out, Geom g, double c_plaq, double c_rect)
{
    extern __shared__ double2 lds[];
    const long long id = (long long)blockIdx.x * BS + threadIdx.x;
    if (id >= 4LL * g.V) return;
    const int s = (int)(id >> 2), mu = (int)(id & 3);
    int c[4];
    coords_of(g, s, c);
    const int par = parity_of(c);
    st_link(out, g, s, mu, par, staple_action(U, g, s, c, mu, par, c_plaq, c_rect, lds + threadIdx.x, BS));
}

__global__ void LT_LINK_KERNEL staple_field_kernel(const double2* __restrict__ U, double2* __restrict__ out, Geom g,
                                                   LinkMap m)
{
    int mu = blockIdx.y >> 1, par = blockIdx.y & 1;
    int c[4], s;
    if (!map_site(g, m, par, c, s)) return;
    st_link(out, g, s, mu, par, staple(U, g, s, c, mu));
}

// The colour launches chunk the y tiles by LinkMap::ycs (by default all y tiles of an x
// step in one chunk: with the alternating walk (rev) 2 % faster per HB + 4 OR update at
// 32^4 than the two-tile chunks the Jacobi kernels keep; the RK3 flow stage is 1.5 %
// slower with one chunk).  The tile order never changes a result.
LinkMap sweep_map(const LinkMap& m)
{
    LinkMap r = m;
    r.yc = r.ycs;
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
            } else if (g.Vh <= SMALL_VH) {
                for (int par = 0; par < 2; ++par)
                    launch(ctx, name, sweep_small_kernel<HEATBATH>, dim3(nblocks(4LL * g.Vh, SBS)), dim3(SBS), 0, u->d, g,
                           mu, par, prefactor, u->seed, counter, (const uint64_t*)nullptr, 0);
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
    launch(u->ctx, "staple_action", staple_action_kernel, dim3(nblocks(4LL * u->g.V, BS)), dim3(BS),
           9 * BS * sizeof(double2), (const double2*)u->d, out->d, u->g, c_plaq, c_rect);
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
    size_t cap = 0;        // plaquette sums held by plaq
    unsigned part_cap = 0;  // partials held by part
    unsigned nb = 0;        // partials of the present launch map

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
            } else if (g.Vh <= SMALL_VH) {
                for (int par = 0; par < 2; ++par)
                    launch(ctx, "hb_cb2", sweep_small_kernel<true>, dim3(nblocks(4LL * g.Vh, SBS)), dim3(SBS), 0, u->d, g,
                           mu, par, prefactor, u->seed, (uint64_t)0, (const uint64_t*)G.state, k);
            } else {
                for (int par = 0; par < 2; ++par)
                    launch(ctx, "hb_cb2", hb_cb2_graph_kernel, dim3(map_blocks(m)), dim3(m.nthreads), 0, u->d, g, m,
                           mu, par, prefactor, u->seed, (const uint64_t*)G.state, k, par);
            }
        }
    sweep_launches<false>(u, G.n_or, G.schedule, 0.0, G.schedule == LT_CB4 ? "or_cb4" : "or_cb2");
    if (G.with_plaq) {
        if (G.schedule == LT_CB4)
            k_cb4_plaq_partials(u, G.part);
        else
            launch(ctx, "plaquette", upd_plaq_kernel, dim3(8 * map_blocks(u->map)), dim3(u->map.nthreads), 0,
                   (const double2*)u->d, g, u->map, G.part);
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
    if (!G.state && hipMalloc(&G.state, 2 * sizeof(uint64_t)) != hipSuccess)
        throw LtError{LT_ENOMEM, "update graph allocation failed"};
    // plaquette partials: one per block of the field's present link map
    const unsigned np = plaq_parts(u->map);
    if (np > G.part_cap) {
        G.drop_exec();
        if (G.part) LT_HIP_CHECK(hipFree(G.part));
        G.part = nullptr;
        G.part_cap = 0;
        if (hipMalloc(&G.part, (size_t)np * sizeof(double)) != hipSuccess)
            throw LtError{LT_ENOMEM, "update graph allocation failed"};
        G.part_cap = np;
    }
    G.nb = np;
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

double k_plaquette_planes(lt_field u)
{
    lt_ctx ctx = u->ctx;
    const unsigned nb = plaq_parts(u->map);
    double* part = ctx->scratch_for((size_t)nb + 1);
    launch(ctx, "plaquette", upd_plaq_kernel, dim3(8 * map_blocks(u->map)), dim3(u->map.nthreads), 0,
           (const double2*)u->d, u->g, u->map, part);
    launch(ctx, "plaquette_sum", plaq_sum_kernel, dim3(1), dim3(RBS), 0, (const double*)part, (int)nb, part + nb);
    double* h = ctx->host_for(1);
    LT_HIP_CHECK(hipMemcpyAsync(h, part + nb, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    LT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return h[0];
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
