// lt_field.hip -- layout conversion (reference AoS <-> even/odd SoA), field
// initialisation (identity, Haar-random), reunitarisation and health checks.
#include "lt_internal.hpp"

namespace lt {

namespace {

constexpr int BS = 256;

// one thread per (site, mu): AoS link ((s*4+mu)*18 doubles) <-> SoA slots
__global__ void __launch_bounds__(BS) pack_kernel(double2* __restrict__ U, const double* __restrict__ aos, Geom g)
{
    long long id = (long long)blockIdx.x * BS + threadIdx.x;
    if (id >= 4LL * g.V) return;
    int s = (int)(id >> 2), mu = (int)(id & 3);
    int c[4];
    coords_of(g, s, c);
    const double* src = aos + id * 18;
    M3 a;
#pragma unroll
    for (int k = 0; k < 9; ++k) a.m[k] = cd{src[2 * k], src[2 * k + 1]};
    st_link(U, g, s, mu, parity_of(c), a);
}

__global__ void __launch_bounds__(BS) unpack_kernel(const double2* __restrict__ U, double* __restrict__ aos, Geom g,
                                                    int su3)
{
    long long id = (long long)blockIdx.x * BS + threadIdx.x;
    if (id >= 4LL * g.V) return;
    int s = (int)(id >> 2), mu = (int)(id & 3);
    int c[4];
    coords_of(g, s, c);
    M3 a = su3 ? ldu(U, g, s, mu, parity_of(c)) : ld(U, g, s, mu, parity_of(c));
    double* dst = aos + id * 18;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        dst[2 * k] = a.m[k].re;
        dst[2 * k + 1] = a.m[k].im;
    }
}

// one thread per SoA slot group: (mu, parity, cb)
__global__ void __launch_bounds__(BS) identity_kernel(double2* __restrict__ U, Geom g)
{
    int cb = blockIdx.x * BS + threadIdx.x;
    if (cb >= g.Vh) return;
    int mp = blockIdx.y;   // mu*2 + parity
    double2* p = U + slot_base(g, mp, 9, cb);
#pragma unroll
    for (int c = 0; c < 9; ++c) p[(size_t)c * slot_stride(g)] = make_double2((c % 4 == 0) ? 1.0 : 0.0, 0.0);
}

// Haar SU(3): complex Gaussian rows 0 and 1 (Philox stream RNG_HAAR, draw block k
// gives entry k), Gram-Schmidt, row 2 = conj(row0 x row1)
__global__ void __launch_bounds__(BS) haar_kernel(double2* __restrict__ U, Geom g, uint64_t seed)
{
    int cb = blockIdx.x * BS + threadIdx.x;
    if (cb >= g.Vh) return;
    int mu = blockIdx.y >> 1, par = blockIdx.y & 1;
    int c[4];
    int s = site_of_cb(g, cb, par, c);
    M3 a = m3_zero();
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        double u0, u1, g0, g1;
        draw2(seed, (uint32_t)s, RNG_HAAR, (uint32_t)k, 0u, (uint32_t)mu, 0ull, u0, u1);
        gauss2(u0, u1, g0, g1);
        a.m[k] = cd{g0, g1};
    }
    gram_schmidt(a);
    stu(U, g, s, mu, par, a);
}

__global__ void __launch_bounds__(BS) reunit_kernel(double2* __restrict__ U, Geom g)
{
    int cb = blockIdx.x * BS + threadIdx.x;
    if (cb >= g.Vh) return;
    int mu = blockIdx.y >> 1, par = blockIdx.y & 1;
    int c[4];
    int s = site_of_cb(g, cb, par, c);
    M3 a = ld(U, g, s, mu, par);
    gram_schmidt(a);
    stu(U, g, s, mu, par, a);
}

// per-block max |U^+U - 1| (Frobenius, per link) and count of non-finite entries
__global__ void __launch_bounds__(BS) check_kernel(const double2* __restrict__ U, Geom g, int su3, double* part)
{
    int cb = blockIdx.x * BS + threadIdx.x;
    double dev = 0.0, bad = 0.0;
    if (cb < g.Vh) {
        int mu = blockIdx.y >> 1, par = blockIdx.y & 1;
        int c[4];
        int s = site_of_cb(g, cb, par, c);
        M3 a = su3 ? ldu(U, g, s, mu, par) : ld(U, g, s, mu, par);
#pragma unroll
        for (int k = 0; k < 9; ++k)
            if (!isfinite(a.m[k].re) || !isfinite(a.m[k].im)) bad += 1.0;
        M3 p = mul_dn(a, a);
        p.m[0].re -= 1.0;
        p.m[4].re -= 1.0;
        p.m[8].re -= 1.0;
        double f = 0.0;
#pragma unroll
        for (int k = 0; k < 9; ++k) f += p.m[k].re * p.m[k].re + p.m[k].im * p.m[k].im;
        dev = sqrt(f);
        if (!(dev == dev)) dev = 1.0e300;
    }
    double m[1] = {dev}, n[1] = {bad};
    block_reduce<BS>(m, OpMax{});
    block_reduce<BS>(n, OpAdd{});
    if (threadIdx.x == 0) {
        size_t b = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
        part[2 * b] = m[0];
        part[2 * b + 1] = n[0];
    }
}

// ParityUpdate (updates/parity_update.hpp:22-47): with P x = (t, Nx-1-x, Ny-1-y, Nz-1-z),
// U'_t(x) = U_t(P x) and U'_i(x) = U_i(P x - i)^+ for the spatial directions; one thread
// per (site, mu), reading the copy `src` and writing `dst`
__global__ void __launch_bounds__(BS) parity_kernel(const double2* __restrict__ src, double2* __restrict__ dst, Geom g)
{
    const long long id = (long long)blockIdx.x * BS + threadIdx.x;
    if (id >= 4LL * g.V) return;
    const int s = (int)(id >> 2), mu = (int)(id & 3);
    int c[4];
    coords_of(g, s, c);
    int r[4] = {c[0], g.L[1] - 1 - c[1], g.L[2] - 1 - c[2], g.L[3] - 1 - c[3]};
    if (mu > 0) r[mu] = r[mu] == 0 ? g.L[mu] - 1 : r[mu] - 1;
    const int sr = ((r[0] * g.L[1] + r[1]) * g.L[2] + r[2]) * g.L[3] + r[3];
    const M3 a = ldu(src, g, sr, mu, parity_of(r));
    stu(dst, g, s, mu, parity_of(c), mu == 0 ? a : dag(a));
}

// plain streaming copy of a link buffer (16-byte elements)
__global__ void __launch_bounds__(BS) copy_kernel(const double2* __restrict__ src, double2* __restrict__ dst, size_t n)
{
    size_t i = (size_t)blockIdx.x * BS + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * BS;
    for (; i < n; i += stride) dst[i] = src[i];
}

// Read rate of a buffer of mask + 1 16-byte elements (a power of two): every lane reads
// `iters` elements, a wavefront 1 KiB of consecutive elements per load, and successive
// loads of a lane jump 1021 KiB ahead (mod the buffer), so the L1 misses and a buffer
// smaller than an XCD's L2 is read from the L2.  One result per lane keeps the loads.
__global__ void __launch_bounds__(BS) l2_read_kernel(const double2* __restrict__ buf, unsigned mask, int iters,
                                                     double2* __restrict__ out)
{
    const unsigned tid = blockIdx.x * BS + threadIdx.x;
    unsigned i = tid;
    double ax = 0.0, ay = 0.0;
#pragma unroll 8
    for (int k = 0; k < iters; ++k) {
        const double2 v = buf[i & mask];
        ax += v.x;
        ay += v.y;
        i += 64u * 1021u;
    }
    out[tid] = make_double2(ax, ay);
}

__global__ void __launch_bounds__(BS) axpy_kernel(const double2* __restrict__ x, double2* __restrict__ y, double a,
                                                  size_t n)
{
    size_t i = (size_t)blockIdx.x * BS + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * BS;
    for (; i < n; i += stride) {
        const double2 u = x[i];
        double2 v = y[i];
        v.x += a * u.x;
        v.y += a * u.y;
        y[i] = v;
    }
}

__global__ void __launch_bounds__(BS) scale_kernel(double2* __restrict__ y, double a, size_t n)
{
    size_t i = (size_t)blockIdx.x * BS + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * BS;
    for (; i < n; i += stride) {
        double2 v = y[i];
        v.x *= a;
        v.y *= a;
        y[i] = v;
    }
}

}  // namespace

void k_scale(lt_ctx ctx, double2* y, double a, size_t n)
{
    unsigned nb = nblocks((long long)n, BS);
    if (nb > 65536u) nb = 65536u;
    launch(ctx, "scale_links", scale_kernel, dim3(nb), dim3(BS), 0, y, a, n);
}

void k_l2_read(lt_ctx ctx, const double2* buf, unsigned mask, int iters, double2* out, unsigned nb)
{
    launch(ctx, "l2_read", l2_read_kernel, dim3(nb), dim3(BS), 0, buf, mask, iters, out);
}

void k_copy(lt_ctx ctx, const double2* src, double2* dst, size_t n)
{
    unsigned nb = nblocks((long long)n, BS);
    if (nb > 65536u) nb = 65536u;
    launch(ctx, "copy_links", copy_kernel, dim3(nb), dim3(BS), 0, src, dst, n);
}

void k_axpy(lt_ctx ctx, const double2* x, double2* y, double a, size_t n)
{
    unsigned nb = nblocks((long long)n, BS);
    if (nb > 65536u) nb = 65536u;
    launch(ctx, "axpy_links", axpy_kernel, dim3(nb), dim3(BS), 0, x, y, a, n);
}

void k_pack(lt_field f, const double* d_aos)
{
    launch(f->ctx, "pack", pack_kernel, dim3(nblocks(4LL * f->g.V, BS)), dim3(BS), 0, f->d, d_aos, f->g);
}

void k_unpack(lt_field f, double* d_aos)
{
    launch(f->ctx, "unpack", unpack_kernel, dim3(nblocks(4LL * f->g.V, BS)), dim3(BS), 0, (const double2*)f->d, d_aos,
           f->g, (int)f->su3);
}

void k_parity(lt_field f, const double2* copy)
{
    launch(f->ctx, "parity_update", parity_kernel, dim3(nblocks(4LL * f->g.V, BS)), dim3(BS), 0, copy, f->d, f->g);
}

void k_identity(lt_field f)
{
    launch(f->ctx, "identity", identity_kernel, dim3(nblocks(f->g.Vh, BS), 8), dim3(BS), 0, f->d, f->g);
}

void k_haar(lt_field f, uint64_t seed)
{
    launch(f->ctx, "haar", haar_kernel, dim3(nblocks(f->g.Vh, BS), 8), dim3(BS), 0, f->d, f->g, seed);
}

void k_reunitarize(lt_field f)
{
    launch(f->ctx, "reunitarize", reunit_kernel, dim3(nblocks(f->g.Vh, BS), 8), dim3(BS), 0, f->d, f->g);
}

void k_check(lt_field f, double* max_dev, long long* nonfinite)
{
    lt_ctx ctx = f->ctx;
    unsigned nb = nblocks(f->g.Vh, BS);
    size_t np = (size_t)nb * 8;
    double* part = ctx->scratch_for(2 * np);
    launch(ctx, "check", check_kernel, dim3(nb, 8), dim3(BS), 0, (const double2*)f->d, f->g, (int)f->su3, part);
    double* h = ctx->host_for(2 * np);
    LT_HIP_CHECK(hipMemcpyAsync(h, part, 2 * np * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    LT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    double m = 0.0, n = 0.0;
    for (size_t i = 0; i < np; ++i) {
        m = h[2 * i] > m ? h[2 * i] : m;
        n += h[2 * i + 1];
    }
    *max_dev = m;
    *nonfinite = (long long)n;
}

}  // namespace lt
