// lt_obs.hip -- fused observables pass: for every site the six clover terms C_munu
// (four leaves each) are built once and give the plaquette sum (leaf I is the
// plaquette), the clover topological-charge density and the clover energy
// density.  Reductions are deterministic: a fixed-shape block tree per timeslice
// chunk, then an ordered sum over chunks.
#include "lt_internal.hpp"

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
    __shared__ double sh[3][BS];
    sh[0][threadIdx.x] = v0;
    sh[1][threadIdx.x] = v1;
    sh[2][threadIdx.x] = v2;
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + w];
            sh[2][threadIdx.x] += sh[2][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sh[0][0];
        out[1] = sh[1][0];
        out[2] = sh[2][0];
    }
}

template <int A0, int A1, int B0, int B1>
__device__ __forceinline__ void pair_terms(const double2* __restrict__ U, const Geom& g, int s, const int c[4], int par,
                                           double& p, double& q, double& e)
{
    double pl;
    M3 Fa = ah2(clover_plane(U, g, s, c, par, A0, A1, pl));
    p += pl;
    e += energy_term(Fa);
    M3 Fb = ah2(clover_plane(U, g, s, c, par, B0, B1, pl));
    p += pl;
    e += energy_term(Fb);
    q += retr_mul(Fa, Fb);
}

// grid: 3 plane pairs x chunks per timeslice x Nt blocks, pair fastest, spread
// over the XCDs so each XCD covers a contiguous range of timeslices with all three
// pairs of a chunk in neighbouring blocks (one L2-resident neighbourhood).  Partial
// sums of the plaquette, raw q and energy go to part[(pair, t, chunk)].  The pairs
// (01,23), (02,13), (03,12) enter q = tr F01F23 - tr F02F13 + tr F03F12; one pair
// per thread keeps registers low.
__global__ void __launch_bounds__(BS) obs_kernel(const double2* __restrict__ U, Geom g, int Vs, int nchunk,
                                                 double* part)
{
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
            pair_terms<0, 1, 2, 3>(U, g, s, c, par, p, q, e);
        } else if (pair == 1) {
            pair_terms<0, 2, 1, 3>(U, g, s, c, par, p, q, e);
            q = -q;
        } else {
            pair_terms<0, 3, 1, 2>(U, g, s, c, par, p, q, e);
        }
    }
    block_sum3(p, q, e, part + 3 * (((size_t)pair * g.L[0] + t) * nchunk + chunk));
}

// one thread per timeslice: ordered sum over pairs and chunks
__global__ void obs_finish_kernel(const double* part, int nchunk, int nt, double* out)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int pr = 0; pr < 3; ++pr)
        for (int k = 0; k < nchunk; ++k) {
            const double* p = part + 3 * (((size_t)pr * nt + t) * nchunk + k);
            a += p[0];
            b += p[1];
            c += p[2];
        }
    out[3 * t] = a;
    out[3 * t + 1] = b;
    out[3 * t + 2] = c;
}

}  // namespace

void k_observables(lt_field u, double* host_out)
{
    lt_ctx ctx = u->ctx;
    const Geom& g = u->g;
    const int nt = g.L[0];
    const int Vs = g.V / nt;
    const unsigned nchunk = nblocks(Vs, BS);
    double* part = ctx->scratch_for(9 * (size_t)nchunk * nt + 3 * (size_t)nt);
    double* fin = part + 9 * (size_t)nchunk * nt;
    launch(ctx, "observables", obs_kernel, dim3(3 * nchunk * nt), dim3(BS), 0, (const double2*)u->d, g, Vs,
           (int)nchunk, part);
    launch(ctx, "observables_finish", obs_finish_kernel, dim3(nblocks(nt, 64)), dim3(64), 0, (const double*)part,
           (int)nchunk, nt, fin);
    LT_HIP_CHECK(hipMemcpyAsync(host_out, fin, 3 * (size_t)nt * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    LT_HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

}  // namespace lt
