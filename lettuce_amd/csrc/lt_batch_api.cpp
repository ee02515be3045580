// lt_batch_api.cpp -- C ABI of replica batches and the batched parallel-tempering
// metadynamics (PT-MetaD) step.
//
// A batch holds n lattices (equal or mixed even extents).  Each kind of field
// (U, U_copy, P, force, stout levels, exponent tapes, clover planes, Lambda) is
// one allocation with the replicas back to back, so every stage of a MetaD-HMC
// trajectory is one launch for all replicas.  Replica r carries a label: the
// index of the bias strength it currently runs with.  One step =
//   one MetaD-HMC trajectory per replica with bias strength[label[r]] * V(Q),
//   Metropolis per replica, restore of rejected replicas,
//   optional symmetric path-sample deposits from the replica holding the top label,
//   optional replica exchange of adjacent labels,
// all enqueued without host synchronisation; with graphs enabled a step is one
// hipGraphLaunch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "lt_batch.hpp"
#include "lt_tune.hpp"

namespace {

struct StepKey {
    double beta = 0.0, tau = 0.0;
    int integrator = 0, n_step = 0, metropolis = 0, update_bias = 0, swap_mode = 0;
    bool operator==(const StepKey& o) const
    {
        return beta == o.beta && tau == o.tau && integrator == o.integrator && n_step == o.n_step &&
               metropolis == o.metropolis && update_bias == o.update_bias && swap_mode == o.swap_mode;
    }
};

template <typename T>
T* dev_alloc(size_t n)
{
    T* p = nullptr;
    if (n == 0) n = 1;
    if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) throw lt::LtError{LT_ENOMEM, "batch device allocation failed"};
    return p;
}

}  // namespace

struct lt_batch_s {
    lt_ctx ctx = nullptr;
    int n = 0, n_levels = 0;
    double rho = 0.0;
    std::vector<lt::BatchRep> reps;
    std::vector<int> pre;                  // lb | sb | kb | cb prefixes, n+1 each
    int total_lb = 0, total_sb = 0, total_kb = 0, total_cb = 0, nthreads = 0;
    long long total = 0;                   // complex doubles per link-layout field

    lt::BatchRep* d_reps = nullptr;
    int* d_pre = nullptr;
    double2 *U = nullptr, *Ucopy = nullptr, *P = nullptr, *F = nullptr, *clover = nullptr;
    double2* lambda[2] = {nullptr, nullptr};   // Lambda of alternate levels
    std::vector<double2*> levels;          // levels[0] aliases U
    std::vector<double2*> tapes;
    double* part = nullptr;

    char* state_mem = nullptr;
    lt::BatchState st{};
    double* strength = nullptr;
    int n_sub_cap = 0;

    std::vector<lt_field> views;

    lt::DevBias bias{};
    bool has_bias = false;
    int grid_cap = 0;

    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    StepKey key{};
    int last_n_sub = 0;

    lt::BatchGrid grid() const
    {
        const int k = n + 1;
        return lt::BatchGrid{d_reps, d_pre, d_pre + k, d_pre + 2 * k, d_pre + 3 * k, n, nthreads};
    }

    void drop_graph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }

    void release()
    {
        drop_graph();
        for (auto v : views) {
            lt::drop_update_graph(v);
            delete v;
        }
        views.clear();
        for (size_t k = 1; k < levels.size(); ++k) (void)hipFree(levels[k]);
        for (auto p : tapes) (void)hipFree(p);
        levels.clear();
        tapes.clear();
        for (void* p : {(void*)d_reps, (void*)d_pre, (void*)U, (void*)Ucopy, (void*)P, (void*)F, (void*)lambda[0], (void*)lambda[1],
                        (void*)clover, (void*)part, (void*)state_mem, (void*)st.samples,
                        (void*)bias.grid})
            if (p) (void)hipFree(p);
    }
};

namespace {

using lt::fail;
using lt::guarded;

int n_sub_of(int integrator, double tau, int n_step)
{
    int n_sub = 0;
    lt::metad_schedule(integrator, tau, n_step, [&](double, bool sub, int) { n_sub += sub ? 1 : 0; },
                       [](double) {});
    return n_sub;
}

void ensure_samples(lt_batch b, int n_sub)
{
    if (n_sub <= b->n_sub_cap) return;
    b->drop_graph();
    if (b->st.samples) LT_HIP_CHECK(hipFree(b->st.samples));
    b->st.samples = nullptr;
    b->st.samples = dev_alloc<double>((size_t)b->n * n_sub);
    b->n_sub_cap = n_sub;
}

// CV forward for all replicas: stout levels with tapes, clover planes, raw charge sums
void forward(lt_batch b, const lt::BatchGrid& bg)
{
    for (int k = 0; k < b->n_levels; ++k)
        lt::kb_stout_tape(b->ctx, bg, b->total_lb, b->levels[k], b->levels[k + 1], b->tapes[k], b->rho);
    lt::kb_clover_charge(b->ctx, bg, b->total_sb, b->levels[b->n_levels], b->clover, b->part, b->st.qraw);
}

// One PT-MetaD step (HMCMetaDKernel::operator(), hmc_metadynamics.hpp:552-655, per replica;
// then the replica exchange of tempering.hpp:57-72 generalised to a strength ladder).
void enqueue_step(lt_batch b, const StepKey& p, int n_sub)
{
    lt_ctx ctx = b->ctx;
    const lt::BatchGrid bg = b->grid();
    lt::BatchState st = b->st;
    st.n_sub = n_sub;
    const lt::DevBias& bias = b->bias;

    lt::k_copy(ctx, b->U, b->Ucopy, (size_t)b->total);
    lt::kb_momentum(ctx, bg, b->total_lb, b->P, st.counter);
    forward(b, bg);
    lt::kb_plaquette(ctx, bg, b->total_sb, b->U, b->part, st.plaq);
    lt::kb_kinetic(ctx, bg, b->total_kb, b->P, b->part, st.kin);
    lt::kb_h0(ctx, bg, st, bias, p.beta);
    int isub = 0;
    lt::metad_schedule(
        p.integrator, p.tau, p.n_step,
        [&](double e, bool sub, int kick) {
            if (kick & lt::KICK_BIAS) {
                // CalculateTopologicalForce (hmc_metadynamics.hpp:374-436)
                forward(b, bg);
                lt::kb_dvdq(ctx, bg, st, bias, sub ? isub++ : -1);
                const int n = b->n_levels;
                // the clover derivative goes on with the Lambda of the top level
                const lt::SigmaNext top{n > 0 ? b->levels[n - 1] : nullptr, n > 0 ? b->tapes[n - 1] : nullptr,
                                        b->lambda[0]};
                lt::kb_clover_deriv(ctx, bg, b->total_lb, b->levels[n], b->clover, b->F, st.dvdq, n > 0 ? &top : nullptr);
                // KICK_BOTH: UpdateMomenta applied by the last kernel of the force (level 0 = U)
                const lt::MomentumKick mk{b->P, e * p.beta / 6.0, e};
                const bool fused = kick == lt::KICK_BOTH && n > 0;
                if (n > 0)
                    lt::kb_stout_recursion_run(
                        ctx, bg, b->total_lb, [&](int j) -> const double2* { return b->levels[j]; },
                        [&](int j) -> const double2* { return b->tapes[j]; }, n, 0, b->F, b->lambda, b->rho,
                        fused ? &mk : nullptr, true);
                if (fused) return;
            }
            if (kick == lt::KICK_BOTH)   // UpdateMomenta (:441-459)
                lt::kb_force_update(ctx, bg, b->total_lb, b->U, b->P, b->F, e * p.beta / 6.0, e);
            else if (kick == lt::KICK_GAUGE)   // UpdateGaugeMomenta (:464-477)
                lt::kb_force_update(ctx, bg, b->total_lb, b->U, b->P, nullptr, e * p.beta / 6.0, 0.0);
            else   // UpdateMetadynamicsMomenta (:479-491)
                lt::k_axpy(ctx, b->F, b->P, e, (size_t)b->total);
        },
        [&](double e) { lt::kb_field_update(ctx, bg, b->total_lb, b->U, b->P, e); });
    forward(b, bg);
    lt::kb_plaquette(ctx, bg, b->total_sb, b->U, b->part, st.plaq);
    lt::kb_kinetic(ctx, bg, b->total_kb, b->P, b->part, st.kin);
    lt::kb_h1_accept(ctx, bg, st, bias, p.beta, p.metropolis);
    lt::kb_copy_links(ctx, bg, b->total_cb, b->Ucopy, b->U, st.accepted);
    if (p.update_bias) lt::kb_deposit(ctx, bg, st, bias, b->n);
    lt::kb_swap_advance(ctx, bg, st, bias, b->n, p.swap_mode);
}

void upload(void* dst, const void* src, size_t bytes)
{
    LT_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
}
void download(void* dst, const void* src, size_t bytes)
{
    LT_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
}

bool valid_rep(lt_batch b, int r) { return b && r >= 0 && r < b->n; }

}  // namespace

extern "C" {

int lt_batch_create(lt_ctx ctx, int n_rep, const int* shapes, int n_levels, double rho, lt_batch* out)
{
    if (!ctx || !shapes || !out || n_rep < 1 || n_rep > 1024 || n_levels < 0 || n_levels > 64)
        return fail(LT_EINVAL, "lt_batch_create: bad arguments");
    for (int r = 0; r < n_rep; ++r)
        for (int d = 0; d < 4; ++d) {
            const int L = shapes[4 * r + d];
            if (L < 2 || (L & 1)) return fail(LT_EINVAL, "lt_batch_create: extents must be even and >= 2");
        }
    return guarded([&] {
        LT_HIP_CHECK(hipSetDevice(ctx->device));
        auto* b = new lt_batch_s();
        b->ctx = ctx;
        b->n = n_rep;
        b->n_levels = n_levels;
        b->rho = rho;
        try {
            const int k = n_rep + 1;
            b->pre.assign((size_t)4 * k, 0);
            long long lb = 0, sb = 0, kb = 0, cb = 0;
            for (int r = 0; r < n_rep; ++r) {
                lt::BatchRep R{};
                R.g = lt::make_geom(shapes[4 * r], shapes[4 * r + 1], shapes[4 * r + 2], shapes[4 * r + 3]);
                R.m = lt::default_linkmap(R.g);
                R.off = b->total;
                R.seed = (uint64_t)r;
                b->total += 72LL * R.g.Vh;
                lb += 8LL * lt::map_blocks(R.m);
                sb += lt::nblocks(R.g.V, 256);
                kb += lt::nblocks(4LL * R.g.V, 256);
                cb += lt::nblocks(72LL * R.g.Vh, 256);
                if (std::max({lb, sb, kb, cb}) > 0x7FFFFFFFLL) throw lt::LtError{LT_EINVAL, "batch too large"};
                b->pre[r + 1] = (int)lb;
                b->pre[k + r + 1] = (int)sb;
                b->pre[2 * k + r + 1] = (int)kb;
                b->pre[3 * k + r + 1] = (int)cb;
                b->nthreads = std::max(b->nthreads, R.m.nthreads);
                b->reps.push_back(R);
            }
            b->total_lb = (int)lb;
            b->total_sb = (int)sb;
            b->total_kb = (int)kb;
            b->total_cb = (int)cb;

            b->d_reps = dev_alloc<lt::BatchRep>(n_rep);
            upload(b->d_reps, b->reps.data(), sizeof(lt::BatchRep) * n_rep);
            b->d_pre = dev_alloc<int>(b->pre.size());
            upload(b->d_pre, b->pre.data(), sizeof(int) * b->pre.size());

            const size_t nl = (size_t)b->total;
            b->U = dev_alloc<double2>(nl);
            b->Ucopy = dev_alloc<double2>(nl);
            b->P = dev_alloc<double2>(nl);
            b->F = dev_alloc<double2>(nl);
            b->lambda[0] = dev_alloc<double2>((size_t)lt::alg_offset(b->total));
            b->lambda[1] = dev_alloc<double2>((size_t)lt::alg_offset(b->total));
            b->clover = dev_alloc<double2>((size_t)lt::clover_offset(b->total));
            b->levels.push_back(b->U);
            for (int l = 0; l < n_levels; ++l) {
                b->levels.push_back(dev_alloc<double2>(nl));
                b->tapes.push_back(dev_alloc<double2>((size_t)lt::alg_offset(b->total)));
            }
            b->part = dev_alloc<double>((size_t)std::max(b->total_sb, b->total_kb));

            // per-replica scalar state, one allocation
            const size_t n = (size_t)n_rep;
            const size_t bytes = n * (8 + 4 + 8 * 10 + 4) + 2 * n * 8 + 3 * 8 + 64;
            b->state_mem = dev_alloc<char>(bytes);
            size_t o = 0;
            auto take = [&](size_t sz) {
                char* p = b->state_mem + o;
                o += (sz + 7) & ~(size_t)7;
                return (void*)p;
            };
            lt::BatchState& st = b->st;
            st.counter = (uint64_t*)take(8 * n);
            st.label = (int*)take(4 * n);
            b->strength = (double*)take(8 * n);
            st.strength = b->strength;
            st.qraw = (double*)take(8 * n);
            st.dvdq = (double*)take(8 * n);
            st.plaq = (double*)take(8 * n);
            st.kin = (double*)take(8 * n);
            st.H0 = (double*)take(8 * n);
            st.H1 = (double*)take(8 * n);
            st.cv_old = (double*)take(8 * n);
            st.cv_cur = (double*)take(8 * n);
            st.dH = (double*)take(8 * n);
            st.accepted = (int*)take(4 * n);
            st.swap_stats = (long long*)take(16 * n);
            st.pt_counter = (uint64_t*)take(8);
            st.exceeded = (unsigned long long*)take(8);
            st.bias_cv = (double*)take(8);
            if (o > bytes) throw lt::LtError{LT_ESTATE, "batch state layout overflow"};
            LT_HIP_CHECK(hipMemset(b->state_mem, 0, bytes));
            std::vector<uint64_t> ctr(n, 0);
            std::vector<int> lab(n);
            std::vector<double> s(n);
            for (size_t r = 0; r < n; ++r) {
                lab[r] = (int)r;
                s[r] = n > 1 ? (double)r / (double)(n - 1) : 1.0;
            }
            upload(st.counter, ctr.data(), 8 * n);
            upload(st.label, lab.data(), 4 * n);
            upload(b->strength, s.data(), 8 * n);

            for (int r = 0; r < n_rep; ++r) {
                auto* v = new lt_field_s();
                v->ctx = ctx;
                v->g = b->reps[r].g;
                v->map = b->reps[r].m;
                v->su3 = true;
                v->view = true;
                v->d = b->U + b->reps[r].off;
                v->seed = b->reps[r].seed;
                b->views.push_back(v);
                lt::k_identity(v);
            }
            LT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            b->release();
            delete b;
            throw;
        }
        ctx->live += 1;
        *out = b;
        return LT_OK;
    });
}

int lt_batch_destroy(lt_batch b)
{
    if (!b) return LT_OK;
    return guarded([&] {
        LT_HIP_CHECK(hipSetDevice(b->ctx->device));
        LT_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
        b->release();
        b->ctx->live -= 1;
        delete b;
        return LT_OK;
    });
}

int lt_batch_size(lt_batch b, int* n_rep)
{
    if (!b || !n_rep) return fail(LT_EINVAL, "null argument");
    *n_rep = b->n;
    return LT_OK;
}

int lt_batch_field(lt_batch b, int r, lt_field* view)
{
    if (!valid_rep(b, r) || !view) return fail(LT_EINVAL, "lt_batch_field: bad replica index");
    *view = b->views[r];
    return LT_OK;
}

int lt_batch_set_rng(lt_batch b, int r, uint64_t seed, uint64_t counter)
{
    if (!valid_rep(b, r)) return fail(LT_EINVAL, "lt_batch_set_rng: bad replica index");
    return guarded([&] {
        LT_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
        b->reps[r].seed = seed;
        upload(b->d_reps + r, &b->reps[r], sizeof(lt::BatchRep));
        upload(b->st.counter + r, &counter, sizeof(uint64_t));
        b->views[r]->seed = seed;
        b->views[r]->counter = counter;
        return LT_OK;
    });
}

int lt_batch_get_rng(lt_batch b, int r, uint64_t* seed, uint64_t* counter)
{
    if (!valid_rep(b, r)) return fail(LT_EINVAL, "lt_batch_get_rng: bad replica index");
    return guarded([&] {
        LT_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
        if (seed) *seed = b->reps[r].seed;
        if (counter) download(counter, b->st.counter + r, sizeof(uint64_t));
        return LT_OK;
    });
}

int lt_batch_set_ladder(lt_batch b, const double* strengths, const int* labels, uint64_t pt_seed, uint64_t pt_counter)
{
    if (!b || !strengths) return fail(LT_EINVAL, "lt_batch_set_ladder: null argument");
    std::vector<int> lab(b->n);
    std::vector<int> seen(b->n, 0);
    for (int r = 0; r < b->n; ++r) {
        lab[r] = labels ? labels[r] : r;
        if (lab[r] < 0 || lab[r] >= b->n || seen[lab[r]]++) return fail(LT_EINVAL, "lt_batch_set_ladder: labels must be a permutation");
        if (!std::isfinite(strengths[r])) return fail(LT_EINVAL, "lt_batch_set_ladder: non-finite strength");
    }
    return guarded([&] {
        LT_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
        b->drop_graph();
        upload(b->strength, strengths, sizeof(double) * b->n);
        upload(b->st.label, lab.data(), sizeof(int) * b->n);
        upload(b->st.pt_counter, &pt_counter, sizeof(uint64_t));
        b->st.pt_seed = pt_seed;
        return LT_OK;
    });
}

int lt_batch_set_bias(lt_batch b, lt_bias bias)
{
    if (!b || !bias) return fail(LT_EINVAL, "lt_batch_set_bias: null argument");
    return guarded([&] {
        LT_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
        const int ng = (int)bias->grid.size();
        if (ng > b->grid_cap) {
            b->drop_graph();
            if (b->bias.grid) LT_HIP_CHECK(hipFree(b->bias.grid));
            b->bias.grid = nullptr;
            b->bias.grid = dev_alloc<double>(ng);
            b->grid_cap = ng;
        }
        lt::DevBias d{};
        d.grid = b->bias.grid;
        d.cv_min = bias->cv_min;
        d.cv_max = bias->cv_max;
        d.bin_width = bias->bin_width;
        d.bin_width_inverse = bias->bin_width_inverse;
        d.weight = bias->weight;
        d.width = bias->width;
        d.wt_param = bias->wt_param;
        d.threshold = bias->threshold;
        d.bins = bias->bins;
        d.well_tempered = bias->well_tempered;
        if (!b->has_bias || std::memcmp(&d, &b->bias, sizeof(d)) != 0) b->drop_graph();
        b->bias = d;
        b->has_bias = true;
        upload(d.grid, bias->grid.data(), sizeof(double) * ng);
        upload(b->st.bias_cv, &bias->cv_current, sizeof(double));
        upload(b->st.exceeded, &bias->exceeded, sizeof(unsigned long long));
        return LT_OK;
    });
}

int lt_batch_get_bias(lt_batch b, lt_bias bias)
{
    if (!b || !bias) return fail(LT_EINVAL, "lt_batch_get_bias: null argument");
    if (!b->has_bias || bias->bins != b->bias.bins) return fail(LT_ESTATE, "lt_batch_get_bias: no bias set or bin count differs");
    return guarded([&] {
        LT_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
        download(bias->grid.data(), b->bias.grid, sizeof(double) * bias->grid.size());
        download(&bias->cv_current, b->st.bias_cv, sizeof(double));
        download(&bias->exceeded, b->st.exceeded, sizeof(unsigned long long));
        return LT_OK;
    });
}

// For each distinct replica extent, time the batched force kernels that dominate a
// MetaD step (a stout level with tape and one level of the stout-force recursion)
// with each candidate tile of the batch's block size applied to every replica of
// that extent.  The batch's fields are scratch here: U (level 0) is only read, the
// other buffers are recomputed by every trajectory.
int lt_batch_autotune(lt_batch b, const char* cache_path)
{
    if (!b) return fail(LT_EINVAL, "null batch");
    return guarded([&] {
        LT_HIP_CHECK(hipSetDevice(b->ctx->device));
        LT_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
        b->drop_graph();
        std::vector<int> done(b->n, 0);
        for (int r = 0; r < b->n; ++r) {
            if (done[r]) continue;
            const lt::Geom g = b->reps[r].g;
            std::vector<int> same;
            for (int q = r; q < b->n; ++q) {
                bool eq = true;
                for (int d = 0; d < 4; ++d) eq = eq && b->reps[q].g.L[d] == g.L[d];
                if (eq) {
                    same.push_back(q);
                    done[q] = 1;
                }
            }
            const std::string key = lt::tune_key(b->ctx->device, g, ("batch" + std::to_string(b->nthreads)).c_str());
            lt::LinkMap best{};
            double ms = 0.0;
            auto apply = [&](const lt::LinkMap& m) {
                for (int q : same) b->reps[q].m = m;
                upload(b->d_reps, b->reps.data(), sizeof(lt::BatchRep) * b->n);
            };
            // the replica's block range was sized for the map in use: a new map must need the
            // same number of blocks (fewer would leave links out)
            const lt::LinkMap cur = b->reps[same[0]].m;
            const unsigned nb = lt::map_blocks(cur);
            if (!lt::cached_tile(cache_path, key, g, best, ms) || lt::map_blocks(best) != nb) {
                const lt::BatchGrid bg = b->grid();
                auto run = [&](const lt::LinkMap& m) {
                    apply(m);
                    double2* tape = b->n_levels > 0 ? b->tapes[0] : b->lambda[1];
                    if (b->n_levels > 0)   // never writes level 0 (the replicas' fields)
                        lt::kb_stout_tape(b->ctx, bg, b->total_lb, b->levels[0], b->levels[1], tape, b->rho);
                    lt::kb_stout_recursion_run(
                        b->ctx, bg, b->total_lb, [&](int) -> const double2* { return b->U; },
                        [&](int) -> const double2* { return tape; }, 1, 0, b->F, b->lambda, b->rho);
                };
                std::vector<lt::LinkMap> cands;
                for (const lt::LinkMap& c : lt::tile_candidates(g, cur.nthreads))
                    if (lt::map_blocks(c) == nb) cands.push_back(c);
                if (cands.empty()) cands.push_back(cur);
                best = lt::tune(b->ctx, cands, cur, run, ms, 2);
                best = lt::tune(b->ctx, lt::order_candidates(best, false), best, run, ms, 2);
                lt::store_tile(cache_path, key, best, ms);
            }
            apply(best);
            for (int q : same) b->views[q]->map = best;
        }
        LT_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
        return LT_OK;
    });
}

int lt_batch_use_graph(lt_batch b, int enable)
{
    if (!b) return fail(LT_EINVAL, "null batch");
    return guarded([&] {
        LT_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
        b->drop_graph();
        b->use_graph = enable != 0;
        return LT_OK;
    });
}

int lt_batch_metad_hmc(lt_batch b, double beta, int integrator, double tau, int n_step, int metropolis,
                       int update_bias, int swap_mode, int n_traj)
{
    if (!b || !lt::valid_metad_integrator(integrator) || n_step < 1 || n_traj < 0 || swap_mode < 0 || swap_mode > 3)
        return fail(LT_EINVAL, "lt_batch_metad_hmc: bad arguments");
    if (!b->has_bias) return fail(LT_ESTATE, "lt_batch_metad_hmc: call lt_batch_set_bias first");
    return guarded([&] {
        LT_HIP_CHECK(hipSetDevice(b->ctx->device));
        StepKey key{beta, tau, integrator, n_step, metropolis ? 1 : 0, update_bias ? 1 : 0, swap_mode};
        const int n_sub = n_sub_of(integrator, tau, n_step);
        ensure_samples(b, std::max(n_sub, 1));
        b->last_n_sub = n_sub;
        if (!b->use_graph) {
            for (int t = 0; t < n_traj; ++t) enqueue_step(b, key, n_sub);
            return LT_OK;
        }
        if (!b->exec || !(b->key == key)) {
            b->drop_graph();
            lt_ctx ctx = b->ctx;
            const bool prof = ctx->profiling;
            ctx->profiling = false;
            LT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
            LT_HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            try {
                enqueue_step(b, key, n_sub);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(ctx->stream, &g);
                if (g) (void)hipGraphDestroy(g);
                ctx->profiling = prof;
                throw;
            }
            ctx->profiling = prof;
            LT_HIP_CHECK(hipStreamEndCapture(ctx->stream, &b->graph));
            LT_HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
            b->key = key;
        }
        for (int t = 0; t < n_traj; ++t) LT_HIP_CHECK(hipGraphLaunch(b->exec, b->ctx->stream));
        return LT_OK;
    });
}

int lt_batch_results(lt_batch b, double* dH, int* accepted, double* cv, int* labels)
{
    if (!b) return fail(LT_EINVAL, "null batch");
    return guarded([&] {
        LT_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
        const size_t n = (size_t)b->n;
        if (dH) download(dH, b->st.dH, 8 * n);
        if (accepted) download(accepted, b->st.accepted, 4 * n);
        if (cv) download(cv, b->st.cv_cur, 8 * n);
        if (labels) download(labels, b->st.label, 4 * n);
        return LT_OK;
    });
}

int lt_batch_samples(lt_batch b, double* samples, int n)
{
    if (!b || !samples || n != b->n * b->last_n_sub)
        return fail(LT_EINVAL, "lt_batch_samples: need n_rep * (samples per trajectory) entries");
    return guarded([&] {
        LT_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
        // rows of the last step's sample count; the buffer is strided by that count
        if (n) download(samples, b->st.samples, sizeof(double) * n);
        return LT_OK;
    });
}

int lt_batch_swap_stats(lt_batch b, long long* stats, int n)
{
    if (!b || !stats || n != 2 * b->n) return fail(LT_EINVAL, "lt_batch_swap_stats: need 2 * n_rep entries");
    return guarded([&] {
        LT_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
        download(stats, b->st.swap_stats, sizeof(long long) * n);
        return LT_OK;
    });
}

}  // extern "C"
