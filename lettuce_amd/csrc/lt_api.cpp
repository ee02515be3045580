// lt_api.cpp -- the C ABI (include/lettuce_hip.h): argument checking, object
// lifetimes, host<->device transfers and the host-side sequencing of the hot
// path (sweep schedules, smearing levels, flow integrators).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "lt_internal.hpp"

namespace lt {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

Geom make_geom(int nt, int nx, int ny, int nz)
{
    Geom g{};
    g.L[0] = nt;
    g.L[1] = nx;
    g.L[2] = ny;
    g.L[3] = nz;
    g.st[3] = 1;
    g.st[2] = nz;
    g.st[1] = ny * nz;
    g.st[0] = nx * ny * nz;
    g.V = nt * nx * ny * nz;
    g.Vh = g.V / 2;
    return g;
}

bool make_linkmap(const Geom& g, int tt, int tx, int ty, int tz, LinkMap& m)
{
    if (tt == 0 && tx == 0 && ty == 0) {   // linear map of tz sites per block
        if (tz < 64 || tz > LT_LINK_BS || tz % 64) return false;
        m = LinkMap{};
        m.lin = 1;
        m.nthreads = tz;
        m.T[0] = m.T[1] = m.T[2] = 1;
        m.T[3] = tz;
        m.N[0] = (g.Vh + tz - 1) / tz;
        m.N[1] = m.N[2] = m.N[3] = 1;
        m.tc = m.yc = m.ycs = 1;
        m.xr = LT_XCD_REMAP;
        return true;
    }
    const int ext[4] = {g.L[0], g.L[1], g.L[2], g.L[3] / 2};
    const int T[4] = {tt, tx, ty, tz};
    int n = 1;
    for (int i = 0; i < 4; ++i) {
        if (T[i] < 1 || ext[i] % T[i]) return false;
        n *= T[i];
    }
    if (n < 1 || n > LT_LINK_BS) return false;
    for (int i = 0; i < 4; ++i) {
        m.T[i] = T[i];
        m.N[i] = ext[i] / T[i];
    }
    m.nthreads = n;
    m.lin = 0;
    // one t-slab per XCD when the t tiles split evenly over the eight XCDs
    m.tc = 1;
#if LT_TCHUNK
    if (m.N[0] % 8 == 0) m.tc = m.N[0] / 8;
#endif
    // y chunks: the tiles in flight per x step (tc * yc of them) and the x-neighbour
    // tiles just finished share the XCD's L2
    m.yc = m.N[2];
#if LT_YCHUNK > 0
    if (m.N[2] % LT_YCHUNK == 0) m.yc = LT_YCHUNK;
#endif
    // the colour sweeps walk all y tiles of an x step in one chunk
    m.ycs = m.N[2];
    m.xr = LT_XCD_REMAP;
    return true;
}

// Default tile: 128 threads; z (checkerboard) fastest, filled up to 16 entries,
// then y, x, t (measured best at 32^4: whole z rows, 8 y rows).  When the box this
// gives holds fewer than 8 sites of a z row and not the whole row (z/2 with few
// factors of two: 12^4, 20^4, 24^4), a linear map of 128 consecutive checkerboard sites
// per block (measured at 20^4: overrelaxation colour launch 82 -> 25 us, HB + 4 OR update
// 3.3 -> 1.0 ms; 24^4: 1.72 -> 1.50 ms).
LinkMap default_linkmap(const Geom& g)
{
    LinkMap m{};
    const int ext[4] = {g.L[0], g.L[1], g.L[2], g.L[3] / 2};
    int T[4] = {1, 1, 1, 1};
    int want = 128;
    const int order[4] = {3, 2, 1, 0};
    for (int k = 0; k < 4 && want > 1; ++k) {
        int i = order[k];
        int lim = (i == 3) ? 16 : ext[i];
        while (want > 1 && T[i] * 2 <= lim && ext[i] % (T[i] * 2) == 0) {
            T[i] *= 2;
            want /= 2;
        }
    }
    for (int k = 0; k < 4 && want > 1; ++k) {   // second pass without the z/y caps
        int i = order[k];
        while (want > 1 && ext[i] % (T[i] * 2) == 0) {
            T[i] *= 2;
            want /= 2;
        }
    }
    // rows cut into pieces of 8 or more (z/2 = 24 at 48^4): a box of whole rows whose
    // thread count is a multiple of 64 (48^4: 1 x 1 x 8 x 24, HB + OR sweeps -4 %, RK3
    // stage -5 % against 1 x 1 x 16 x 8), most y rows first
    if (want == 1 && T[3] >= 8 && T[3] < ext[3] && ext[3] <= 32) {
        int bestn = 0, B[4] = {0, 0, 0, 0};
        for (int ty = ext[2]; ty >= 1; --ty) {
            if (ext[2] % ty) continue;
            for (int tx = 1; tx <= ext[1]; ++tx) {
                if (ext[1] % tx) continue;
                for (int tt = 1; tt <= ext[0]; ++tt) {
                    if (ext[0] % tt) continue;
                    const int n = tt * tx * ty * ext[3];
                    if (n % 64 == 0 && n <= 256 && n > bestn) {
                        bestn = n;
                        B[0] = tt;
                        B[1] = tx;
                        B[2] = ty;
                        B[3] = ext[3];
                    }
                }
            }
        }
        if (bestn >= 128 && make_linkmap(g, B[0], B[1], B[2], B[3], m)) return m;
    }
    if (want == 1 && (T[3] >= 8 || T[3] == ext[3]) && make_linkmap(g, T[0], T[1], T[2], T[3], m)) return m;
    if (make_linkmap(g, 0, 0, 0, 128, m)) return m;
    // extents without enough factors of two: the largest dividing tile of at most
    // 256 threads, z-major
    int best = 0, bt[4] = {1, 1, 1, 1};
    for (int tz = 1; tz <= ext[3]; ++tz) {
        if (ext[3] % tz) continue;
        for (int ty = 1; ty <= ext[2]; ++ty) {
            if (ext[2] % ty) continue;
            for (int tx = 1; tx <= ext[1]; ++tx) {
                if (ext[1] % tx) continue;
                for (int tt = 1; tt <= ext[0]; ++tt) {
                    if (ext[0] % tt) continue;
                    int n = tt * tx * ty * tz;
                    if (n <= LT_LINK_BS && n > best) {
                        best = n;
                        bt[0] = tt;
                        bt[1] = tx;
                        bt[2] = ty;
                        bt[3] = tz;
                    }
                }
            }
        }
    }
    make_linkmap(g, bt[0], bt[1], bt[2], bt[3], m);
    return m;
}

}  // namespace lt

double* lt_ctx_s::scratch_for(size_t n)
{
    if (n > scratch_doubles) {
        if (scratch) LT_HIP_CHECK(hipFree(scratch));
        scratch = nullptr;
        scratch_doubles = 0;
        if (hipMalloc(&scratch, n * sizeof(double)) != hipSuccess)
            throw lt::LtError{LT_ENOMEM, "scratch allocation failed"};
        scratch_doubles = n;
    }
    return scratch;
}

double* lt_ctx_s::host_for(size_t n)
{
    if (n > host_doubles) {
        if (host) LT_HIP_CHECK(hipHostFree(host));
        host = nullptr;
        host_doubles = 0;
        if (hipHostMalloc((void**)&host, n * sizeof(double), hipHostMallocDefault) != hipSuccess)
            throw lt::LtError{LT_ENOMEM, "pinned host allocation failed"};
        host_doubles = n;
    }
    return host;
}

void lt_ctx_s::collect_profile()
{
    if (pending.empty()) return;
    LT_HIP_CHECK(hipStreamSynchronize(stream));
    for (auto& p : pending) {
        float ms = 0.0f;
        LT_HIP_CHECK(hipEventElapsedTime(&ms, p.second.first, p.second.second));
        auto& e = prof[p.first];
        e.launches += 1;
        e.ms += ms;
        event_pool.push_back(p.second);
    }
    pending.clear();
}

namespace {

using lt::fail;
using lt::guarded;

void use_device(lt_ctx ctx) { LT_HIP_CHECK(hipSetDevice(ctx->device)); }

bool same_shape(lt_field a, lt_field b)
{
    for (int i = 0; i < 4; ++i)
        if (a->g.L[i] != b->g.L[i]) return false;
    return true;
}

double2* alloc_links(const lt::Geom& g)
{
    double2* p = nullptr;
    if (hipMalloc(&p, (size_t)72 * g.Vh * sizeof(double2)) != hipSuccess)
        throw lt::LtError{LT_ENOMEM, "link buffer allocation failed"};
    return p;
}

}  // namespace

extern "C" {

const char* lt_last_error(void) { return lt::g_last_error.c_str(); }

int lt_version(int* major, int* minor)
{
    if (major) *major = 0;
    if (minor) *minor = 1;
    return LT_OK;
}

int lt_ctx_create(int device, lt_ctx* out)
{
    return guarded([&] {
        if (!out) return fail(LT_EINVAL, "lt_ctx_create: null out");
        int n = 0;
        LT_HIP_CHECK(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) return fail(LT_EINVAL, "lt_ctx_create: no such device");
        LT_HIP_CHECK(hipSetDevice(device));
        lt::noalloc_setup(device);
        auto* c = new lt_ctx_s();
        c->device = device;
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            return fail(LT_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
        }
        *out = c;
        return LT_OK;
    });
}

int lt_ctx_destroy(lt_ctx ctx)
{
    if (!ctx) return LT_OK;
    if (ctx->live > 0)
        return fail(LT_ESTATE, "lt_ctx_destroy: " + std::to_string(ctx->live) +
                                   " field / flow / workspace / batch handle(s) of this context still exist");
    return guarded([&] {
        use_device(ctx);
        LT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        for (auto& p : ctx->pending) ctx->event_pool.push_back(p.second);
        for (auto& p : ctx->event_pool) {
            (void)hipEventDestroy(p.first);
            (void)hipEventDestroy(p.second);
        }
        if (ctx->scratch) (void)hipFree(ctx->scratch);
        if (ctx->host) (void)hipHostFree(ctx->host);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return LT_OK;
    });
}

int lt_ctx_synchronize(lt_ctx ctx)
{
    if (!ctx) return fail(LT_EINVAL, "null context");
    return guarded([&] {
        LT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return LT_OK;
    });
}

int lt_ctx_stream(lt_ctx ctx, void** stream)
{
    if (!ctx || !stream) return fail(LT_EINVAL, "null argument");
    *stream = (void*)ctx->stream;
    return LT_OK;
}

int lt_ctx_mem_info(lt_ctx ctx, size_t* free_bytes, size_t* total_bytes)
{
    if (!ctx) return fail(LT_EINVAL, "null context");
    return guarded([&] {
        use_device(ctx);
        size_t f = 0, t = 0;
        LT_HIP_CHECK(hipMemGetInfo(&f, &t));
        if (free_bytes) *free_bytes = f;
        if (total_bytes) *total_bytes = t;
        return LT_OK;
    });
}

int lt_ctx_read_rate(lt_ctx ctx, size_t footprint_bytes, double* bytes_per_s)
{
    if (!ctx || !bytes_per_s) return fail(LT_EINVAL, "null argument");
    if (footprint_bytes < ((size_t)1 << 16) || footprint_bytes > ((size_t)1 << 30) ||
        (footprint_bytes & (footprint_bytes - 1)))
        return fail(LT_EINVAL, "lt_ctx_read_rate: footprint must be a power of two in [64 KiB, 1 GiB]");
    return guarded([&] {
        use_device(ctx);
        const unsigned nb = 2048, lanes = nb * 256;
        const int iters = 128;
        double2 *buf = nullptr, *out = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        auto release = [&] {
            (void)hipStreamSynchronize(ctx->stream);
            if (buf) (void)hipFree(buf);
            if (out) (void)hipFree(out);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
        };
        try {
            LT_HIP_CHECK(hipMalloc(&buf, footprint_bytes));
            LT_HIP_CHECK(hipMalloc(&out, (size_t)lanes * sizeof(double2)));
            LT_HIP_CHECK(hipMemsetAsync(buf, 0, footprint_bytes, ctx->stream));
            LT_HIP_CHECK(hipEventCreate(&e0));
            LT_HIP_CHECK(hipEventCreate(&e1));
            const unsigned mask = (unsigned)(footprint_bytes / sizeof(double2)) - 1u;
            const bool prof = ctx->profiling;
            ctx->profiling = false;
            double best = 1e30;
            for (int r = 0; r < 6; ++r) {   // the first launch warms the caches and the clock
                LT_HIP_CHECK(hipEventRecord(e0, ctx->stream));
                lt::k_l2_read(ctx, buf, mask, iters, out, nb);
                LT_HIP_CHECK(hipEventRecord(e1, ctx->stream));
                LT_HIP_CHECK(hipEventSynchronize(e1));
                float ms = 0.0f;
                LT_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                if (r > 0) best = std::min(best, (double)ms);
            }
            ctx->profiling = prof;
            *bytes_per_s = (double)lanes * iters * sizeof(double2) / (best * 1e-3);
        } catch (...) {
            release();
            throw;
        }
        release();
        return LT_OK;
    });
}

int lt_ctx_profile(lt_ctx ctx, int enable)
{
    if (!ctx) return fail(LT_EINVAL, "null context");
    return guarded([&] {
        ctx->collect_profile();
        ctx->prof.clear();
        ctx->profiling = enable != 0;
        return LT_OK;
    });
}

int lt_ctx_profile_query(lt_ctx ctx, const char* kernel, int* launches, double* total_ms)
{
    if (!ctx || !kernel) return fail(LT_EINVAL, "null argument");
    return guarded([&] {
        ctx->collect_profile();
        auto it = ctx->prof.find(kernel);
        if (launches) *launches = it == ctx->prof.end() ? 0 : it->second.launches;
        if (total_ms) *total_ms = it == ctx->prof.end() ? 0.0 : it->second.ms;
        return LT_OK;
    });
}

int lt_field_create(lt_ctx ctx, int nt, int nx, int ny, int nz, lt_field* out)
{
    if (!ctx || !out) return fail(LT_EINVAL, "lt_field_create: null argument");
    if (nt < 2 || nx < 2 || ny < 2 || nz < 2 || (nt | nx | ny | nz) & 1)
        return fail(LT_EINVAL, "lt_field_create: extents must be even and >= 2");
    if ((double)nt * nx * ny * nz > 1.0e9) return fail(LT_EINVAL, "lt_field_create: volume too large");
    return guarded([&] {
        use_device(ctx);
        auto* f = new lt_field_s();
        f->ctx = ctx;
        f->g = lt::make_geom(nt, nx, ny, nz);
        f->map = lt::default_linkmap(f->g);
        try {
            f->d = alloc_links(f->g);
        } catch (...) {
            delete f;
            throw;
        }
        ctx->live += 1;
        *out = f;
        return LT_OK;
    });
}

int lt_field_destroy(lt_field f)
{
    if (!f) return LT_OK;
    if (f->view) return fail(LT_ESTATE, "lt_field_destroy: this field is a replica of a batch (destroy the batch)");
    return guarded([&] {
        use_device(f->ctx);
        LT_HIP_CHECK(hipStreamSynchronize(f->ctx->stream));
        lt::drop_update_graph(f);
        if (f->d) LT_HIP_CHECK(hipFree(f->d));
        f->ctx->live -= 1;
        delete f;
        return LT_OK;
    });
}

int lt_field_shape(lt_field f, int shape[4])
{
    if (!f || !shape) return fail(LT_EINVAL, "null argument");
    for (int i = 0; i < 4; ++i) shape[i] = f->g.L[i];
    return LT_OK;
}

int lt_field_upload(lt_field f, const double* aos)
{
    if (!f || !aos) return fail(LT_EINVAL, "lt_field_upload: null argument");
    return guarded([&] {
        use_device(f->ctx);
        size_t bytes = (size_t)4 * f->g.V * 18 * sizeof(double);
        double* stage = nullptr;
        if (hipMalloc(&stage, bytes) != hipSuccess) throw lt::LtError{LT_ENOMEM, "upload staging allocation failed"};
        hipError_t e = hipMemcpyAsync(stage, aos, bytes, hipMemcpyHostToDevice, f->ctx->stream);
        if (e == hipSuccess) {
            try {
                lt::k_pack(f, stage);
            } catch (...) {
                (void)hipStreamSynchronize(f->ctx->stream);
                (void)hipFree(stage);
                throw;
            }
            e = hipStreamSynchronize(f->ctx->stream);
        }
        (void)hipFree(stage);
        if (e != hipSuccess) throw lt::LtError{LT_EHIP, std::string("upload: ") + hipGetErrorString(e)};
        return LT_OK;
    });
}

int lt_field_download(lt_field f, double* aos)
{
    if (!f || !aos) return fail(LT_EINVAL, "lt_field_download: null argument");
    return guarded([&] {
        use_device(f->ctx);
        size_t bytes = (size_t)4 * f->g.V * 18 * sizeof(double);
        double* stage = nullptr;
        if (hipMalloc(&stage, bytes) != hipSuccess) throw lt::LtError{LT_ENOMEM, "download staging allocation failed"};
        hipError_t e = hipSuccess;
        try {
            lt::k_unpack(f, stage);
            e = hipMemcpyAsync(aos, stage, bytes, hipMemcpyDeviceToHost, f->ctx->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(f->ctx->stream);
        } catch (...) {
            (void)hipStreamSynchronize(f->ctx->stream);
            (void)hipFree(stage);
            throw;
        }
        (void)hipFree(stage);
        if (e != hipSuccess) throw lt::LtError{LT_EHIP, std::string("download: ") + hipGetErrorString(e)};
        return LT_OK;
    });
}

int lt_field_copy(lt_field dst, lt_field src)
{
    if (!dst || !src) return fail(LT_EINVAL, "lt_field_copy: null argument");
    if (!same_shape(dst, src)) return fail(LT_ESTATE, "lt_field_copy: shape mismatch");
    if (dst == src) return LT_OK;
    return guarded([&] {
        use_device(dst->ctx);
        lt::k_copy(dst->ctx, src->d, dst->d, (size_t)72 * dst->g.Vh);
        dst->su3 = src->su3;
        return LT_OK;
    });
}

int lt_field_swap(lt_field a, lt_field b)
{
    if (!a || !b) return fail(LT_EINVAL, "lt_field_swap: null argument");
    if (!same_shape(a, b) || a->ctx != b->ctx) return fail(LT_ESTATE, "lt_field_swap: shape/context mismatch");
    if (a->view || b->view) return fail(LT_ESTATE, "lt_field_swap: replicas of a batch cannot be swapped");
    std::swap(a->d, b->d);
    std::swap(a->su3, b->su3);
    return LT_OK;
}

int lt_field_set_identity(lt_field f)
{
    if (!f) return fail(LT_EINVAL, "null field");
    return guarded([&] {
        use_device(f->ctx);
        lt::k_identity(f);
        f->su3 = true;
        return LT_OK;
    });
}

int lt_field_set_haar(lt_field f, uint64_t seed)
{
    if (!f) return fail(LT_EINVAL, "null field");
    return guarded([&] {
        use_device(f->ctx);
        lt::k_haar(f, seed);
        f->su3 = true;
        return LT_OK;
    });
}

int lt_field_reunitarize(lt_field f)
{
    if (!f) return fail(LT_EINVAL, "null field");
    return guarded([&] {
        use_device(f->ctx);
        lt::k_reunitarize(f);
        return LT_OK;
    });
}

int lt_field_check(lt_field f, double* max_dev, long long* n_nonfinite)
{
    if (!f || !max_dev || !n_nonfinite) return fail(LT_EINVAL, "null argument");
    return guarded([&] {
        use_device(f->ctx);
        lt::k_check(f, max_dev, n_nonfinite);
        return LT_OK;
    });
}

int lt_field_set_rng(lt_field f, uint64_t seed, uint64_t counter)
{
    if (!f) return fail(LT_EINVAL, "null field");
    f->seed = seed;
    f->counter = counter;
    return LT_OK;
}

int lt_field_set_tile(lt_field f, int tt, int tx, int ty, int tz)
{
    if (!f) return fail(LT_EINVAL, "null field");
    lt::LinkMap m{};
    if (!lt::make_linkmap(f->g, tt, tx, ty, tz, m))
        return fail(LT_EINVAL, "lt_field_set_tile: tile must divide (Nt, Nx, Ny, Nz/2) with 1..1024 threads");
    f->map = m;
    return LT_OK;
}

int lt_field_get_tile(lt_field f, int tile[4])
{
    if (!f || !tile) return fail(LT_EINVAL, "null argument");
    for (int i = 0; i < 4; ++i) tile[i] = f->map.lin && i < 3 ? 0 : f->map.T[i];
    return LT_OK;
}

int lt_field_get_rng(lt_field f, uint64_t* seed, uint64_t* counter)
{
    if (!f) return fail(LT_EINVAL, "null field");
    if (seed) *seed = f->seed;
    if (counter) *counter = f->counter;
    return LT_OK;
}

int lt_staple_field(lt_field u, lt_field out)
{
    if (!u || !out) return fail(LT_EINVAL, "null field");
    if (!same_shape(u, out) || u == out) return fail(LT_ESTATE, "lt_staple_field: need a distinct field of equal shape");
    return guarded([&] {
        use_device(u->ctx);
        lt::k_staple_field(u, out);
        out->su3 = false;
        return LT_OK;
    });
}

static int check_schedule(lt_field u, int schedule)
{
    if (schedule != LT_CB2 && schedule != LT_CB4) return fail(LT_EINVAL, "schedule must be LT_CB2 or LT_CB4");
    if (schedule == LT_CB4)
        for (int i = 0; i < 4; ++i)
            if (u->g.L[i] % 4) return fail(LT_EINVAL, "LT_CB4 needs extents divisible by 4");
    return LT_OK;
}

int lt_or_sweep(lt_field u, int n_sweep, int schedule)
{
    if (!u || n_sweep < 0) return fail(LT_EINVAL, "lt_or_sweep: bad argument");
    int rc = check_schedule(u, schedule);
    if (rc) return rc;
    return guarded([&] {
        use_device(u->ctx);
        lt::k_or_sweep(u, n_sweep, schedule);
        return LT_OK;
    });
}

int lt_heatbath_sweep(lt_field u, double beta, int n_sweep, int schedule)
{
    if (!u || n_sweep < 0 || !(beta > 0.0)) return fail(LT_EINVAL, "lt_heatbath_sweep: bad argument");
    int rc = check_schedule(u, schedule);
    if (rc) return rc;
    return guarded([&] {
        use_device(u->ctx);
        lt::k_hb_sweep(u, beta, n_sweep, schedule);
        return LT_OK;
    });
}

int lt_update_sweeps(lt_field u, double beta, int n_hb, int n_or, int schedule, int n_update, double* plaq_sum)
{
    if (!u || n_hb < 0 || n_or < 0 || n_update < 0 || !(beta > 0.0))
        return fail(LT_EINVAL, "lt_update_sweeps: bad argument");
    int rc = check_schedule(u, schedule);
    if (rc) return rc;
    if (n_update == 0) return LT_OK;
    return guarded([&] {
        use_device(u->ctx);
        lt::k_update_graph(u, beta, n_hb, n_or, schedule, n_update, plaq_sum);
        return LT_OK;
    });
}

// Rectangular<2>(beta, c_plaq, c_rect): the rectangle staples reach two links along mu,
// so only the four-colour schedule keeps simultaneously updated links out of each
// other's staples (Checkerboard4, iterators.hpp:141-169)
static int check_action(lt_field u, int schedule, double c_plaq, double c_rect)
{
    if (!std::isfinite(c_plaq) || !std::isfinite(c_rect)) return fail(LT_EINVAL, "non-finite action coefficients");
    if (!u->su3) return fail(LT_ESTATE, "the field must hold SU(3) links");
    if (c_rect != 0.0 && schedule != LT_CB4) return fail(LT_EINVAL, "rectangle terms need the LT_CB4 schedule");
    return check_schedule(u, schedule);
}

int lt_staple_field_action(lt_field u, double c_plaq, double c_rect, lt_field out)
{
    if (!u || !out) return fail(LT_EINVAL, "null field");
    if (!same_shape(u, out) || u == out || u->ctx != out->ctx)
        return fail(LT_ESTATE, "lt_staple_field_action: need a distinct field of equal shape");
    if (!std::isfinite(c_plaq) || !std::isfinite(c_rect)) return fail(LT_EINVAL, "non-finite action coefficients");
    return guarded([&] {
        use_device(u->ctx);
        lt::k_staple_field_action(u, out, c_plaq, c_rect);
        out->su3 = false;
        return LT_OK;
    });
}

int lt_or_sweep_action(lt_field u, double c_plaq, double c_rect, int n_sweep, int schedule)
{
    if (!u || n_sweep < 0) return fail(LT_EINVAL, "lt_or_sweep_action: bad argument");
    int rc = check_action(u, schedule, c_plaq, c_rect);
    if (rc) return rc;
    if (c_rect == 0.0 && c_plaq == 1.0) return lt_or_sweep(u, n_sweep, schedule);
    if (schedule != LT_CB4) return fail(LT_EINVAL, "lt_or_sweep_action: scaled actions use LT_CB4");
    return guarded([&] {
        use_device(u->ctx);
        lt::k_or_sweep_action(u, n_sweep, c_plaq, c_rect);
        return LT_OK;
    });
}

int lt_heatbath_sweep_action(lt_field u, double beta, double c_plaq, double c_rect, int n_sweep, int schedule)
{
    if (!u || n_sweep < 0 || !(beta > 0.0)) return fail(LT_EINVAL, "lt_heatbath_sweep_action: bad argument");
    int rc = check_action(u, schedule, c_plaq, c_rect);
    if (rc) return rc;
    if (c_rect == 0.0 && c_plaq == 1.0) return lt_heatbath_sweep(u, beta, n_sweep, schedule);
    if (schedule != LT_CB4) return fail(LT_EINVAL, "lt_heatbath_sweep_action: scaled actions use LT_CB4");
    return guarded([&] {
        use_device(u->ctx);
        lt::k_heatbath_sweep_action(u, beta, n_sweep, c_plaq, c_rect);
        return LT_OK;
    });
}

// out: [plaq_sum, Q, E_plaq, E_clov, q_t[Nt], e_t[Nt]]
int lt_observables(lt_field u, double* out)
{
    if (!u || !out) return fail(LT_EINVAL, "null argument");
    return guarded([&] {
        use_device(u->ctx);
        const int nt = u->g.L[0];
        double* h = u->ctx->host_for(3 * (size_t)nt);
        lt::k_observables(u, h);
        const double V = (double)u->g.V;
        const double qnorm = -1.0 / (256.0 * M_PI * M_PI);
        double p = 0.0, q = 0.0, e = 0.0;
        for (int t = 0; t < nt; ++t) {
            double qt = qnorm * h[3 * t + 1];
            double et = h[3 * t + 2] / V;
            out[4 + t] = qt;
            out[4 + nt + t] = et;
            p += h[3 * t];
            q += qt;
            e += et;
        }
        out[0] = p;
        out[1] = q;
        out[2] = 2.0 * (18.0 - p / V);
        out[3] = e;
        return LT_OK;
    });
}

int lt_observables_noalloc(lt_field u, double* out)
{
    if (!u || !out) return fail(LT_EINVAL, "null argument");
    if (!u->su3) return fail(LT_ESTATE, "lt_observables_noalloc: the field must hold SU(3) links");
    return guarded([&] {
        use_device(u->ctx);
        lt::k_observables_noalloc(u, out);
        return LT_OK;
    });
}

int lt_parity_update(lt_field u, lt_field u_copy)
{
    if (!u || !u_copy) return fail(LT_EINVAL, "null argument");
    if (u == u_copy) return fail(LT_EINVAL, "lt_parity_update: the copy must be another field");
    if (!u->su3 || !u_copy->su3) return fail(LT_ESTATE, "lt_parity_update: SU(3) fields required");
    if (!same_shape(u, u_copy) || u->ctx != u_copy->ctx) return fail(LT_ESTATE, "lt_parity_update: shape/context mismatch");
    return guarded([&] {
        use_device(u->ctx);
        lt::k_copy(u->ctx, u->d, u_copy->d, (size_t)72 * u->g.Vh);
        lt::k_parity(u, u_copy->d);
        LT_HIP_CHECK(hipStreamSynchronize(u->ctx->stream));
        return LT_OK;
    });
}

int lt_action_rectangular(lt_field u, double c_plaq, double c_rect, double* action_normalized)
{
    if (!u || !action_normalized) return fail(LT_EINVAL, "null argument");
    if (!u->su3) return fail(LT_ESTATE, "lt_action_rectangular: the field must hold SU(3) links");
    return guarded([&] {
        use_device(u->ctx);
        double r[3];
        lt::k_rect_sums(u, r);
        const double V = (double)u->g.V;
        *action_normalized = (c_plaq * (6.0 * V - r[0] / 3.0) + c_rect * (12.0 * V - r[1] / 3.0)) / (6.0 * V);
        return LT_OK;
    });
}

int lt_topcharge_plaquette(lt_field u, double* q)
{
    if (!u || !q) return fail(LT_EINVAL, "null argument");
    if (!u->su3) return fail(LT_ESTATE, "lt_topcharge_plaquette: the field must hold SU(3) links");
    return guarded([&] {
        use_device(u->ctx);
        double r[3];
        lt::k_rect_sums(u, r);
        *q = r[2] / (16.0 * M_PI * M_PI);
        return LT_OK;
    });
}

int lt_wilson_loop(lt_field u, lt_field chain, int n_start, int n_end, int reset, double* w)
{
    if (!u || !chain || !w) return fail(LT_EINVAL, "null argument");
    if (n_start < 0 || n_end <= n_start) return fail(LT_EINVAL, "lt_wilson_loop: 0 <= n_start < n_end required");
    if (!u->su3 || chain->su3) return fail(LT_ESTATE, "lt_wilson_loop: SU(3) field and matrix-field chain required");
    if (!same_shape(u, chain) || u->ctx != chain->ctx) return fail(LT_ESTATE, "lt_wilson_loop: shape/context mismatch");
    return guarded([&] {
        use_device(u->ctx);
        const double s = lt::k_wilson_loop(u, chain, n_start, n_end, reset);
        *w = 1.0 - s / (18.0 * (double)u->g.V);
        return LT_OK;
    });
}

int lt_polyakov_loop(lt_field u, double* re, double* im)
{
    if (!u || !re || !im) return fail(LT_EINVAL, "null argument");
    if (!u->su3) return fail(LT_ESTATE, "lt_polyakov_loop: the field must hold SU(3) links");
    return guarded([&] {
        use_device(u->ctx);
        lt::k_polyakov_loop(u, re, im);
        return LT_OK;
    });
}

int lt_plaquette_sum(lt_field u, double* out)
{
    if (!u || !out) return fail(LT_EINVAL, "null argument");
    return guarded([&] {
        use_device(u->ctx);
        *out = lt::k_plaquette_planes(u);
        return LT_OK;
    });
}

int lt_wilson_action(lt_field u, double beta, double* out)
{
    if (!u || !out) return fail(LT_EINVAL, "null argument");
    double p = 0.0;
    int rc = lt_plaquette_sum(u, &p);
    if (rc == LT_OK) *out = beta * (6.0 * (double)u->g.V - p / 3.0);
    return rc;
}

int lt_topcharge_clover(lt_field u, double* Q, double* q_t)
{
    if (!u || !Q) return fail(LT_EINVAL, "null argument");
    const int nt = u->g.L[0];
    std::vector<double> o(4 + 2 * (size_t)nt);
    int rc = lt_observables(u, o.data());
    if (rc) return rc;
    *Q = o[1];
    if (q_t) std::memcpy(q_t, o.data() + 4, nt * sizeof(double));
    return LT_OK;
}

int lt_topcharge_clover_density(lt_field u, double* q_x)
{
    if (!u || !q_x) return fail(LT_EINVAL, "null argument");
    if (!u->su3) return fail(LT_ESTATE, "lt_topcharge_clover_density: the field must hold SU(3) links");
    return guarded([&] {
        use_device(u->ctx);
        lt::k_topcharge_density(u, q_x);
        return LT_OK;
    });
}

int lt_energy(lt_field u, double* e_plaq, double* e_clov, double* e_t)
{
    if (!u) return fail(LT_EINVAL, "null argument");
    const int nt = u->g.L[0];
    std::vector<double> o(4 + 2 * (size_t)nt);
    int rc = lt_observables(u, o.data());
    if (rc) return rc;
    if (e_plaq) *e_plaq = o[2];
    if (e_clov) *e_clov = o[3];
    if (e_t) std::memcpy(e_t, o.data() + 4 + nt, nt * sizeof(double));
    return LT_OK;
}

int lt_stout(lt_field in, lt_field out, double rho, int n_levels)
{
    if (!in || !out || n_levels < 0) return fail(LT_EINVAL, "lt_stout: bad argument");
    if (!same_shape(in, out) || in->ctx != out->ctx) return fail(LT_ESTATE, "lt_stout: shape/context mismatch");
    return guarded([&] {
        use_device(in->ctx);
        lt_ctx ctx = in->ctx;
        if (n_levels == 0) {
            if (in != out) lt::k_copy(ctx, in->d, out->d, (size_t)72 * in->g.Vh);
            return LT_OK;
        }
        // ping-pong so that the last level lands in out->d
        double2* tmp = alloc_links(in->g);
        const double2* src = in->d;
        for (int k = 0; k < n_levels; ++k) {
            const bool last = k == n_levels - 1;
            double2* dst;
            if (last) dst = (src == out->d) ? tmp : out->d;
            else dst = (src == tmp) ? out->d : tmp;
            if (dst == src) dst = (dst == tmp) ? out->d : tmp;
            lt::k_stout(ctx, in->g, in->map, src, dst, rho);
            src = dst;
        }
        if (src != out->d) lt::k_copy(ctx, src, out->d, (size_t)72 * in->g.Vh);
        LT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        LT_HIP_CHECK(hipFree(tmp));
        return LT_OK;
    });
}

int lt_flow_create(lt_field unflowed, lt_field flowed, int integrator, double eps, lt_flow* out)
{
    if (!unflowed || !flowed || !out) return fail(LT_EINVAL, "lt_flow_create: null argument");
    if (integrator != LT_FLOW_EULER && integrator != LT_FLOW_RK3) return fail(LT_EINVAL, "unknown integrator");
    if (!same_shape(unflowed, flowed) || unflowed->ctx != flowed->ctx || unflowed == flowed)
        return fail(LT_ESTATE, "lt_flow_create: need two distinct fields of equal shape");
    return guarded([&] {
        use_device(flowed->ctx);
        auto* fl = new lt_flow_s();
        fl->unflowed = unflowed;
        fl->flowed = flowed;
        fl->integrator = integrator;
        fl->eps = eps;
        try {
            // Z is an su(3) field: 8 reals per link (4 complex doubles)
            if (hipMalloc(&fl->z, (size_t)32 * flowed->g.Vh * sizeof(double2)) != hipSuccess) {
                fl->z = nullptr;
                throw lt::LtError{LT_ENOMEM, "flow Z allocation failed"};
            }
            fl->tmp = alloc_links(flowed->g);
        } catch (...) {
            if (fl->z) (void)hipFree(fl->z);
            delete fl;
            throw;
        }
        flowed->ctx->live += 1;
        *out = fl;
        return LT_OK;
    });
}

int lt_flow_destroy(lt_flow fl)
{
    if (!fl) return LT_OK;
    return guarded([&] {
        use_device(fl->flowed->ctx);
        LT_HIP_CHECK(hipStreamSynchronize(fl->flowed->ctx->stream));
        (void)hipFree(fl->z);
        (void)hipFree(fl->tmp);
        fl->flowed->ctx->live -= 1;
        delete fl;
        return LT_OK;
    });
}

// One integrator step on fl->flowed; each stage writes the other buffer and the
// buffers are exchanged, so the flowed field always owns the newest links.
static void flow_steps(lt_flow fl, int n)
{
    lt_field u = fl->flowed;
    lt_ctx ctx = u->ctx;
    const double e = fl->eps;
    auto stage = [&](double c_old, double c_new, double c_exp, int read_z, int write_z) {
        lt::k_flow_stage(ctx, u->g, u->map, u->d, fl->tmp, fl->z, c_old, c_new, c_exp, read_z, write_z);
        std::swap(u->d, fl->tmp);
    };
    for (int k = 0; k < n; ++k) {
        if (fl->integrator == LT_FLOW_EULER) {
            stage(0.0, e, 1.0, 0, 0);
        } else {
            stage(0.0, e, 0.25, 0, 1);
            stage(-17.0 / 36.0, 8.0 / 9.0 * e, 1.0, 1, 1);
            stage(-1.0, 3.0 / 4.0 * e, 1.0, 1, 0);
        }
    }
}

int lt_flow_run(lt_flow fl, int n_steps)
{
    if (!fl || n_steps < 0) return fail(LT_EINVAL, "lt_flow_run: bad argument");
    return guarded([&] {
        use_device(fl->flowed->ctx);
        lt::k_copy(fl->flowed->ctx, fl->unflowed->d, fl->flowed->d, (size_t)72 * fl->flowed->g.Vh);
        flow_steps(fl, n_steps);
        return LT_OK;
    });
}

int lt_flow_resume(lt_flow fl, int n_steps)
{
    if (!fl || n_steps < 0) return fail(LT_EINVAL, "lt_flow_resume: bad argument");
    return guarded([&] {
        use_device(fl->flowed->ctx);
        flow_steps(fl, n_steps);
        return LT_OK;
    });
}

}  // extern "C"
