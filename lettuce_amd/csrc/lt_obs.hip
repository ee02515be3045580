// lt_obs.hip -- fused observables pass: for every site the six clover terms C_munu
// (four leaves each) are built once and give the plaquette sum (leaf I is the
// plaquette), the clover topological-charge density and the clover energy
// density.  Reductions are deterministic: a fixed-shape block tree per timeslice
// chunk, then an ordered sum over chunks.
#include "lt_internal.hpp"

#include <mutex>

namespace lt {

namespace {

constexpr int BS = 256;

// F = C - C^+ (anti-hermitian)
__device__ __forceinline__ M3 ah2(const M3& C)
{
    M3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.m[3 * i + j] = C.m[3 * i + j] - conj(C.m[3 * j + i]);
    return r;
}

// Re tr(f f) with f = traceless(-i/8 F), F anti-hermitian
__device__ __forceinline__ double energy_term(const M3& F)
{
    M3 f;
#pragma unroll
    for (int k = 0; k < 9; ++k) f.m[k] = cd{0.125 * F.m[k].im, -0.125 * F.m[k].re};   // -i/8 F
    cd t = trace(f);
    t = (1.0 / 3.0) * t;
    f.m[0] -= t;
    f.m[4] -= t;
    f.m[8] -= t;
    return retr_mul(f, f);
}

__device__ __forceinline__ void block_sum3(double v0, double v1, double v2, double* out)
{
    double a[3] = {v0, v1, v2};
    block_reduce<BS>(a, OpAdd{});
    if (threadIdx.x == 0) {
        out[0] = a[0];
        out[1] = a[1];
        out[2] = a[2];
    }
}

// lds: two parking slots per lane (see park_store; entries e * BS + lane and, for F_a
// while F_b is formed, (9 + e) * BS + lane); LDS_BYTES per block
constexpr size_t LDS_BYTES = 18 * BS * sizeof(double2);
template <int A0, int A1, int B0, int B1>
__device__ __forceinline__ void pair_terms(const double2* __restrict__ U, const Geom& g, int s, const int c[4], int par,
                                           double& p, double& q, double& e, double2* lds)
{
    double pl;
    double2* slot = lds + threadIdx.x;
    M3 Fa = ah2(clover_plane(U, g, s, c, par, A0, A1, pl, slot, BS));
    p += pl;
    e += energy_term(Fa);
    if (lds) park_store(slot + 9 * BS, BS, Fa);
    M3 Fb = ah2(clover_plane(U, g, s, c, par, B0, B1, pl, slot, BS));
    p += pl;
    e += energy_term(Fb);
    if (lds) Fa = park_load(slot + 9 * BS, BS);
    q += retr_mul(Fa, Fb);
}

// grid: 3 plane pairs x chunks per timeslice x Nt blocks, pair fastest, spread
// over the XCDs so each XCD covers a contiguous range of timeslices with all three
// pairs of a chunk in neighbouring blocks (one L2-resident neighbourhood).  Partial
// sums of the plaquette, raw q and energy go to part[(pair, t, chunk)].  The pairs
// (01,23), (02,13), (03,12) enter q = tr F01F23 - tr F02F13 + tr F03F12; one pair
// per thread keeps registers low.
// two waves per SIMD: the parked clover planes fit 256 VGPRs (one wave and 330
// registers without parking)
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(2)))
obs_kernel(const double2* __restrict__ U, Geom g, int Vs, int nchunk, double* part)
{
    extern __shared__ double2 lds[];
    int b = xcd_block(blockIdx.x, gridDim.x);
    const int pair = b % 3;
    b /= 3;
    const int chunk = b % nchunk;
    const int t = b / nchunk;
    const int r = chunk * BS + threadIdx.x;
    double p = 0.0, q = 0.0, e = 0.0;
    if (r < Vs) {
        const int s = t * Vs + r;
        int c[4];
        coords_of(g, s, c);
        const int par = parity_of(c);
        if (pair == 0) {
            pair_terms<0, 1, 2, 3>(U, g, s, c, par, p, q, e, lds);
        } else if (pair == 1) {
            pair_terms<0, 2, 1, 3>(U, g, s, c, par, p, q, e, lds);
            q = -q;
        } else {
            pair_terms<0, 3, 1, 2>(U, g, s, c, par, p, q, e, lds);
        }
    }
    block_sum3(p, q, e, part + 3 * (((size_t)pair * g.L[0] + t) * nchunk + chunk));
}

// TopChargeClover's summand per site (topological_charge.hpp:184-199): one thread
// per site (lexicographic order) evaluates all three plane pairs and writes
// -1/(256 pi^2) (tr F01F23 - tr F02F13 + tr F03F12), F = C - C^+.
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(2)))
qdensity_kernel(const double2* __restrict__ U, Geom g, double* __restrict__ q_x)
{
    extern __shared__ double2 lds[];
    const int s = blockIdx.x * BS + threadIdx.x;
    if (s >= g.V) return;
    int c[4];
    coords_of(g, s, c);
    const int par = parity_of(c);
    double p = 0.0, e = 0.0, q01 = 0.0, q02 = 0.0, q03 = 0.0;
    pair_terms<0, 1, 2, 3>(U, g, s, c, par, p, q01, e, lds);
    pair_terms<0, 2, 1, 3>(U, g, s, c, par, p, q02, e, lds);
    pair_terms<0, 3, 1, 2>(U, g, s, c, par, p, q03, e, lds);
    q_x[s] = (-1.0 / (256.0 * M_PI * M_PI)) * (q01 - q02 + q03);
}

// one block per timeslice: lane i sums the (pair, chunk) entries e = i, i + BS, ... of
// its timeslice (e = pair * nchunk + chunk), then the block tree of block_sum3
__global__ void __launch_bounds__(BS) obs_finish_kernel(const double* part, int nchunk, int nt, double* out)
{
    const int t = blockIdx.x;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int e = threadIdx.x; e < 3 * nchunk; e += BS) {
        const int pr = e / nchunk, k = e - pr * nchunk;
        const double* p = part + 3 * (((size_t)pr * nt + t) * nchunk + k);
        a += p[0];
        b += p[1];
        c += p[2];
    }
    block_sum3(a, b, c, out + 3 * (size_t)t);
}

// ---- the rest of Observables() (SU3_main.cpp:125-136) --------------------------

// neighbour of lexicographic site s with coordinates c, n steps along mu (|n| < L)
__device__ __forceinline__ int shift(const Geom& g, int s, const int c[4], int mu, int n)
{
    int v = c[mu] + n;
    v = v >= g.L[mu] ? v - g.L[mu] : (v < 0 ? v + g.L[mu] : v);
    return s + (v - c[mu]) * g.st[mu];
}
__device__ __forceinline__ M3 link(const double2* __restrict__ U, const Geom& g, int s, int mu)
{
    int c[4];
    coords_of(g, s, c);
    return ldu(U, g, s, mu, parity_of(c));
}

// plaquette P_munu(x) (plaquette.hpp:20-24) and Re tr of the 1x2 and 2x1 loops
// (wilson_loop.hpp:155-175): each loop as Re tr(A B^+) of its two 3-link halves
__device__ __forceinline__ void plane_loops(const double2* __restrict__ U, const Geom& g, int s, const int c[4], int mu,
                                            int nu, M3& P, double& rect)
{
    const int smu = shift(g, s, c, mu, 1), snu = shift(g, s, c, nu, 1);
    const M3 Umu = link(U, g, s, mu), Unu = link(U, g, s, nu);
    const M3 Unu_mu = link(U, g, smu, nu), Umu_nu = link(U, g, snu, mu);
    P = mul_nd_r01(mul_nd_r01(mul_r01(Umu, Unu_mu), Umu_nu), Unu);
    third_row(P);
    int cm[4] = {c[0], c[1], c[2], c[3]}, cn[4] = {c[0], c[1], c[2], c[3]};
    cm[mu] = cm[mu] + 1 == g.L[mu] ? 0 : cm[mu] + 1;
    cn[nu] = cn[nu] + 1 == g.L[nu] ? 0 : cn[nu] + 1;
    // 1x2: U_mu(x) U_nu(x+mu) U_nu(x+mu+nu)  vs  U_nu(x) U_nu(x+nu) U_mu(x+2nu)
    {
        M3 a = mul_r01(Umu, Unu_mu);
        third_row(a);
        a = mul_r01(a, link(U, g, shift(g, smu, cm, nu, 1), nu));
        third_row(a);
        M3 b = mul_r01(Unu, link(U, g, snu, nu));
        third_row(b);
        b = mul_r01(b, link(U, g, shift(g, snu, cn, nu, 1), mu));
        third_row(b);
        rect = retr_mul(a, dag(b));
    }
    // 2x1: U_mu(x) U_mu(x+mu) U_nu(x+2mu)  vs  U_nu(x) U_mu(x+nu) U_mu(x+mu+nu)
    {
        M3 a = mul_r01(Umu, link(U, g, smu, mu));
        third_row(a);
        a = mul_r01(a, link(U, g, shift(g, smu, cm, mu, 1), nu));
        third_row(a);
        M3 b = mul_r01(Unu, Umu_nu);
        third_row(b);
        b = mul_r01(b, link(U, g, shift(g, snu, cn, mu, 1), mu));
        third_row(b);
        rect += retr_mul(a, dag(b));
    }
}

__device__ __forceinline__ void block_sum2(double v0, double v1, double* out)
{
    double a[2] = {v0, v1};
    block_reduce<BS>(a, OpAdd{});
    if (threadIdx.x == 0) {
        out[0] = a[0];
        out[1] = a[1];
    }
}

// grid: 3 plane pairs x ceil(V / BS) blocks, pair fastest, XCD-remapped.  Per block:
// sum Re tr P, sum Re tr (R12 + R21) and the plaquette charge density
// Re tr (F_a F_b) (signed as in q = F01F23 + F02F31 + F03F12, F = -i (P - P^+))
__global__ void __launch_bounds__(BS) rect_kernel(const double2* __restrict__ U, Geom g, double* part)
{
    int b = xcd_block(blockIdx.x, gridDim.x);
    const int pair = b % 3, chunk = b / 3;
    const int s = chunk * BS + threadIdx.x;
    double p = 0.0, r = 0.0, q = 0.0;
    if (s < g.V) {
        int c[4];
        coords_of(g, s, c);
        // pairs (01, 23), (02, 13) with F31 = -F13, (03, 12)
        const int a0 = 0, a1 = pair + 1;
        const int b0 = pair == 2 ? 1 : (pair == 0 ? 2 : 1), b1 = pair == 0 ? 3 : (pair == 1 ? 3 : 2);
        M3 Pa, Pb;
        double ra, rb;
        plane_loops(U, g, s, c, a0, a1, Pa, ra);
        plane_loops(U, g, s, c, b0, b1, Pb, rb);
        p = Pa.m[0].re + Pa.m[4].re + Pa.m[8].re + Pb.m[0].re + Pb.m[4].re + Pb.m[8].re;
        r = ra + rb;
        // F = -i (P - P^+); Re tr (Fa Fb) = -Re tr ((Pa - Pa^+)(Pb - Pb^+))
        const M3 Da = sub(Pa, dag(Pa)), Db = sub(Pb, dag(Pb));
        q = -retr_mul(Da, Db);
        if (pair == 1) q = -q;
    }
    block_sum3(p, r, q, part + 3 * (size_t)blockIdx.x);
}

// WilsonLoop chain update (wilson_loop.hpp:71-88): chain(x, mu) <- [1 or chain(x, mu)]
// * U_mu(x + n0 mu) ... U_mu(x + (n1 - 1) mu), one thread per (site, mu)
__global__ void __launch_bounds__(BS) wchain_kernel(const double2* __restrict__ U, double2* __restrict__ chain, Geom g,
                                                    int n0, int n1, int reset)
{
    const long long id = (long long)blockIdx.x * BS + threadIdx.x;
    if (id >= 4LL * g.V) return;
    const int s = (int)(id >> 2), mu = (int)(id & 3);
    int c[4];
    coords_of(g, s, c);
    const int par = parity_of(c);
    M3 a = reset ? m3_id() : ld(chain, g, s, mu, par);
    for (int n = n0; n < n1; ++n) {
        const int sn = shift(g, s, c, mu, n % g.L[mu]);
        a = mul(a, link(U, g, sn, mu));
    }
    st_link(chain, g, s, mu, par, a);
}

// the square loops of side L from the chains (wilson_loop.hpp:89-125), summed over
// the six planes: Re tr (chain_mu(x) chain_nu(x + L mu) (chain_nu(x) chain_mu(x + L nu))^+).
// Both products are products of links, so in SU(3): rows 0 and 1 of each are multiplied
// out and the trace is taken with retr_su3_nd_r01.
__global__ void __launch_bounds__(BS) wloop_kernel(const double2* __restrict__ chain, Geom g, int L, double* part)
{
    const int s = blockIdx.x * BS + threadIdx.x;
    double w = 0.0;
    if (s < g.V) {
        int c[4];
        coords_of(g, s, c);
        const int par = parity_of(c);
        auto ch = [&](int site, int mu) {
            int cc[4];
            coords_of(g, site, cc);
            return ld(chain, g, site, mu, parity_of(cc));
        };
        for (int nu = 1; nu < 4; ++nu)
            for (int mu = 0; mu < nu; ++mu) {
                const M3 A = mul_r01(ld(chain, g, s, mu, par), ch(shift(g, s, c, mu, L % g.L[mu]), nu));
                const M3 B = mul_r01(ld(chain, g, s, nu, par), ch(shift(g, s, c, nu, L % g.L[nu]), mu));
                w += retr_su3_nd_r01(A, B);
            }
    }
    block_sum2(w, 0.0, part + 2 * (size_t)blockIdx.x);
}

// PolyakovLoop (polyakov_loop.hpp:31-49): one thread per spatial site, tr prod_t U_t
__global__ void __launch_bounds__(BS) polyakov_kernel(const double2* __restrict__ U, Geom g, double* part)
{
    const int vs = g.L[1] * g.L[2] * g.L[3];
    const int r = blockIdx.x * BS + threadIdx.x;
    double re = 0.0, im = 0.0;
    if (r < vs) {
        M3 a = m3_id();
        for (int t = 0; t < g.L[0]; ++t) a = mul(a, link(U, g, t * vs + r, 0));
        re = a.m[0].re + a.m[4].re + a.m[8].re;
        im = a.m[0].im + a.m[4].im + a.m[8].im;
    }
    block_sum2(re, im, part + 2 * (size_t)blockIdx.x);
}

}  // namespace

// sums over sites of Re tr P, Re tr (R12 + R21) and the raw plaquette charge density
void k_rect_sums(lt_field u, double out[3])
{
    lt_ctx ctx = u->ctx;
    const unsigned nb = 3 * nblocks(u->g.V, BS);
    double* part = ctx->scratch_for(3 * (size_t)nb);
    launch(ctx, "rect", rect_kernel, dim3(nb), dim3(BS), 0, (const double2*)u->d, u->g, part);
    double* h = ctx->host_for(3 * (size_t)nb);
    LT_HIP_CHECK(hipMemcpyAsync(h, part, 3 * (size_t)nb * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    LT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    out[0] = out[1] = out[2] = 0.0;
    for (unsigned i = 0; i < nb; ++i)
        for (int k = 0; k < 3; ++k) out[k] += h[3 * i + k];
}

double k_wilson_loop(lt_field u, lt_field chain, int n0, int n1, int reset)
{
    lt_ctx ctx = u->ctx;
    const Geom& g = u->g;
    launch(ctx, "wilson_chain", wchain_kernel, dim3(nblocks(4LL * g.V, BS)), dim3(BS), 0, (const double2*)u->d,
           chain->d, g, n0, n1, reset);
    const unsigned nb = nblocks(g.V, BS);
    double* part = ctx->scratch_for(2 * (size_t)nb);
    launch(ctx, "wilson_loop", wloop_kernel, dim3(nb), dim3(BS), 0, (const double2*)chain->d, g, n1, part);
    double* h = ctx->host_for(2 * (size_t)nb);
    LT_HIP_CHECK(hipMemcpyAsync(h, part, 2 * (size_t)nb * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    LT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    double w = 0.0;
    for (unsigned i = 0; i < nb; ++i) w += h[2 * i];
    return w;
}

void k_polyakov_loop(lt_field u, double* re, double* im)
{
    lt_ctx ctx = u->ctx;
    const Geom& g = u->g;
    const int vs = g.L[1] * g.L[2] * g.L[3];
    const unsigned nb = nblocks(vs, BS);
    double* part = ctx->scratch_for(2 * (size_t)nb);
    launch(ctx, "polyakov_loop", polyakov_kernel, dim3(nb), dim3(BS), 0, (const double2*)u->d, g, part);
    double* h = ctx->host_for(2 * (size_t)nb);
    LT_HIP_CHECK(hipMemcpyAsync(h, part, 2 * (size_t)nb * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    LT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    double a = 0.0, b = 0.0;
    for (unsigned i = 0; i < nb; ++i) {
        a += h[2 * i];
        b += h[2 * i + 1];
    }
    *re = a / vs;
    *im = b / vs;
}

void k_observables(lt_field u, double* host_out)
{
    lt_ctx ctx = u->ctx;
    const Geom& g = u->g;
    const int nt = g.L[0];
    const int Vs = g.V / nt;
    const unsigned nchunk = nblocks(Vs, BS);
    double* part = ctx->scratch_for(9 * (size_t)nchunk * nt + 3 * (size_t)nt);
    double* fin = part + 9 * (size_t)nchunk * nt;
    launch(ctx, "observables", obs_kernel, dim3(3 * nchunk * nt), dim3(BS), LDS_BYTES, (const double2*)u->d, g, Vs,
           (int)nchunk, part);
    launch(ctx, "observables_finish", obs_finish_kernel, dim3(nt), dim3(BS), 0, (const double*)part, (int)nchunk, nt,
           fin);
    LT_HIP_CHECK(hipMemcpyAsync(host_out, fin, 3 * (size_t)nt * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    LT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

// ---- observables without allocation ---------------------------------------------
// lt_observables_noalloc keeps its partial sums in this fixed device array instead of
// the context's scratch.  Host code uses only the per-device address that
// hipGetSymbolAddress returns (the array's name is a host-side stand-in).
namespace {

constexpr size_t NOALLOC_DOUBLES = (size_t)1 << 20;   // 8 MiB per GPU
constexpr int NOALLOC_MAX_DEVICES = 64;
__device__ __align__(256) double scratch[NOALLOC_DOUBLES];

std::mutex noalloc_mutex;
double* noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// The per-timeslice sums -> [plaq_sum, Q, E_plaq, E_clov, q_t[Nt], e_t[Nt]] in one
// thread, operation for operation as lt_observables does it on the host (no
// contraction), so both entry points return the same bits.
__global__ void obs_final_kernel(const double* __restrict__ h, int nt, double V, double* __restrict__ out)
{
#pragma clang fp contract(off)
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const double qnorm = -1.0 / (256.0 * M_PI * M_PI);
    double p = 0.0, q = 0.0, e = 0.0;
    for (int t = 0; t < nt; ++t) {
        const double qt = qnorm * h[3 * t + 1];
        const double et = h[3 * t + 2] / V;
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
}

}  // namespace

void noalloc_setup(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES)
        throw LtError{LT_EINVAL, "device index beyond the fixed observables scratch table"};
    std::lock_guard<std::mutex> lock(noalloc_mutex);
    if (noalloc_ptr[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        throw LtError{LT_EHIP, std::string("hipGetSymbolAddress(observables scratch): ") + hipGetErrorString(e)};
    noalloc_ptr[device] = (double*)p;
}

size_t noalloc_observables_doubles(const Geom& g)
{
    const int nt = g.L[0];
    const size_t nchunk = nblocks(g.V / nt, BS);
    return 9 * nchunk * nt + 3 * (size_t)nt + 4 + 2 * (size_t)nt;
}

void k_observables_noalloc(lt_field u, double* host_out)
{
    lt_ctx ctx = u->ctx;
    const Geom& g = u->g;
    const size_t need = noalloc_observables_doubles(g);
    if (need > NOALLOC_DOUBLES)
        throw LtError{LT_EINVAL, "lt_observables_noalloc: the lattice needs " + std::to_string(need) +
                                     " doubles of scratch, the fixed array holds " + std::to_string(NOALLOC_DOUBLES)};
    double* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(noalloc_mutex);
        if (ctx->device >= 0 && ctx->device < NOALLOC_MAX_DEVICES) base = noalloc_ptr[ctx->device];
    }
    if (!base) throw LtError{LT_ESTATE, "lt_observables_noalloc: no scratch address for this device"};
    const int nt = g.L[0];
    const int Vs = g.V / nt;
    const unsigned nchunk = nblocks(Vs, BS);
    double* part = base;
    double* fin = part + 9 * (size_t)nchunk * nt;
    double* out = fin + 3 * (size_t)nt;
    // every partial and every per-timeslice sum is written before it is read: nothing
    // has to start at zero
    hipLaunchKernelGGL(obs_kernel, dim3(3 * nchunk * nt), dim3(BS), LDS_BYTES, ctx->stream, (const double2*)u->d, g, Vs,
                       (int)nchunk, part);
    LT_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(obs_finish_kernel, dim3(nt), dim3(BS), 0, ctx->stream, (const double*)part, (int)nchunk, nt,
                       fin);
    LT_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(obs_final_kernel, dim3(1), dim3(64), 0, ctx->stream, (const double*)fin, nt, (double)g.V, out);
    LT_HIP_CHECK(hipGetLastError());
    LT_HIP_CHECK(hipMemcpyAsync(host_out, out, (4 + 2 * (size_t)nt) * sizeof(double), hipMemcpyDeviceToHost,
                                ctx->stream));
    LT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

void k_topcharge_density(lt_field u, double* host_q)
{
    lt_ctx ctx = u->ctx;
    const Geom& g = u->g;
    double* d = ctx->scratch_for((size_t)g.V);
    launch(ctx, "topcharge_density", qdensity_kernel, dim3(nblocks(g.V, BS)), dim3(BS), LDS_BYTES, (const double2*)u->d, g, d);
    LT_HIP_CHECK(hipMemcpyAsync(host_q, d, (size_t)g.V * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    LT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

}  // namespace lt
