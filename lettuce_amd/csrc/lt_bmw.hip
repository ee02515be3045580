// lt_bmw.hip -- BMW configuration format on the device: packing the even/odd SoA
// field into the file's byte order and back, with the format's 64-bit Adler
// checksum accumulated in the same pass.
//
// Format (LettuceGauge/IO/config_io/bmw_format.hpp:29-49): after a 4096-byte ASCII
// header the links follow with t, z, y, x from slowest to fastest and the four
// directions of a site in the order x, y, z, t; each link is rows 0 and 1 (the
// "#BMW" format, K = 12 doubles) or all three rows ("#18BMW", K = 18), row-major,
// (re, im) pairs, big-endian.
//
// Checksum (Checksum_Adler64, bmw_format.hpp:55-100): the 4K native 64-bit words of
// site n (BMW order) form block n of V blocks; with p = 4294967291 and
// r = (V - n) * 2 * 4K decreasing by 2 per word, a word w = (h, l) (high and low
// 32 bits) adds h + l to A and (2 + r h + (r - 1) l) mod p to B (64-bit wrapping
// products).  A and B are reduced mod p; the sums are order independent, so each
// site contributes its own pair and blocks reduce them with mod-p additions.
#include "lt_internal.hpp"

namespace lt {

namespace {

constexpr int BS = 256;
constexpr uint64_t ADLER_P = 4294967291ull;

__device__ __forceinline__ uint64_t bswap64(uint64_t v) { return __builtin_bswap64(v); }

__device__ __forceinline__ void adler_word(uint64_t w, uint64_t& r, uint64_t& a, uint64_t& b)
{
    const uint64_t h = w >> 32, l = w & 0xffffffffull;
    a += h + l;
    b += (2 + r * h + (r - 1) * l) % ADLER_P;
    r -= 2;
}

// mod-p sums of (a, b) over the block; thread 0 writes the pair
struct OpAddMod {
    __device__ __forceinline__ unsigned long long operator()(unsigned long long x, unsigned long long y) const
    {
        return (x + y) % ADLER_P;
    }
};
__device__ __forceinline__ void block_sum_mod(uint64_t a, uint64_t b, uint64_t* part)
{
    unsigned long long v[2] = {a % ADLER_P, b % ADLER_P};
    block_reduce<BS>(v, OpAddMod{});
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = v[0];
        part[2 * blockIdx.x + 1] = v[1];
    }
}

// lexicographic site (t, x, y, z) of BMW site n = ((t * Nz + z) * Ny + y) * Nx + x
__device__ __forceinline__ int site_of_bmw(const Geom& g, long long n, int c[4])
{
    c[1] = (int)(n % g.L[1]);
    n /= g.L[1];
    c[2] = (int)(n % g.L[2]);
    n /= g.L[2];
    c[3] = (int)(n % g.L[3]);
    c[0] = (int)(n / g.L[3]);
    return ((c[0] * g.L[1] + c[1]) * g.L[2] + c[2]) * g.L[3] + c[3];
}

// One thread per site of BMW sites [n0, n0 + count): out (may be null: checksum
// only) receives the site's 4K big-endian words at (n - n0) * 4K.
template <int K>
__global__ void __launch_bounds__(BS) bmw_pack_kernel(const double2* __restrict__ U, Geom g, long long n0,
                                                      long long count, uint64_t* __restrict__ out,
                                                      uint64_t* __restrict__ part)
{
    const long long i = (long long)blockIdx.x * BS + threadIdx.x;
    uint64_t a = 0, b = 0;
    if (i < count) {
        const long long n = n0 + i;
        int c[4];
        const int s = site_of_bmw(g, n, c);
        const int par = parity_of(c);
        uint64_t r = (uint64_t)(g.V - n) * (uint64_t)(2 * 4 * K);
        uint64_t* dst = out ? out + i * 4 * K : nullptr;
#pragma unroll
        for (int k = 0; k < 4; ++k) {   // BMW directions x, y, z, t
            const M3 m = ldu(U, g, s, (k + 1) & 3, par);
#pragma unroll
            for (int e = 0; e < K / 2; ++e) {
                const uint64_t wr = (uint64_t)__double_as_longlong(m.m[e].re);
                const uint64_t wi = (uint64_t)__double_as_longlong(m.m[e].im);
                adler_word(wr, r, a, b);
                adler_word(wi, r, a, b);
                if (dst) {
                    dst[k * K + 2 * e] = bswap64(wr);
                    dst[k * K + 2 * e + 1] = bswap64(wi);
                }
            }
        }
    }
    block_sum_mod(a, b, part);
}

// Inverse of bmw_pack_kernel: rows 0 and 1 of every link go to the field.  For
// K = 18 the stored row 2 is compared with conj(row0 x row1) (the field keeps rows
// 0 and 1 only) and the largest deviation per block goes to dev.
template <int K>
__global__ void __launch_bounds__(BS) bmw_unpack_kernel(double2* __restrict__ U, Geom g, long long n0,
                                                        long long count, const uint64_t* __restrict__ in,
                                                        uint64_t* __restrict__ part, double* __restrict__ dev)
{
    const long long i = (long long)blockIdx.x * BS + threadIdx.x;
    uint64_t a = 0, b = 0;
    double d = 0.0;
    if (i < count) {
        const long long n = n0 + i;
        int c[4];
        const int s = site_of_bmw(g, n, c);
        const int par = parity_of(c);
        uint64_t r = (uint64_t)(g.V - n) * (uint64_t)(2 * 4 * K);
        const uint64_t* src = in + i * 4 * K;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            M3 m;
#pragma unroll
            for (int e = 0; e < K / 2; ++e) {
                const uint64_t wr = bswap64(src[k * K + 2 * e]), wi = bswap64(src[k * K + 2 * e + 1]);
                adler_word(wr, r, a, b);
                adler_word(wi, r, a, b);
                m.m[e] = cd{__longlong_as_double((long long)wr), __longlong_as_double((long long)wi)};
            }
            if (K == 18) {
                M3 q = m;
                third_row(q);
#pragma unroll
                for (int e = 6; e < 9; ++e) d = fmax(d, fmax(fabs(q.m[e].re - m.m[e].re), fabs(q.m[e].im - m.m[e].im)));
            }
            stu(U, g, s, (k + 1) & 3, par, m);
        }
    }
    block_sum_mod(a, b, part);
    double dm[1] = {d};
    block_reduce<BS>(dm, OpMax{});
    if (threadIdx.x == 0) dev[blockIdx.x] = dm[0];
}

}  // namespace

void k_bmw_pack(lt_field f, int K, long long n0, long long count, uint64_t* d_out, uint64_t* d_part)
{
    const dim3 grid(nblocks(count, BS));
    if (K == 18)
        launch(f->ctx, "bmw_pack", bmw_pack_kernel<18>, grid, dim3(BS), 0, (const double2*)f->d, f->g, n0, count, d_out,
               d_part);
    else
        launch(f->ctx, "bmw_pack", bmw_pack_kernel<12>, grid, dim3(BS), 0, (const double2*)f->d, f->g, n0, count, d_out,
               d_part);
}

void k_bmw_unpack(lt_field f, int K, long long n0, long long count, const uint64_t* d_in, uint64_t* d_part,
                  double* d_dev)
{
    const dim3 grid(nblocks(count, BS));
    if (K == 18)
        launch(f->ctx, "bmw_unpack", bmw_unpack_kernel<18>, grid, dim3(BS), 0, f->d, f->g, n0, count, d_in, d_part,
               d_dev);
    else
        launch(f->ctx, "bmw_unpack", bmw_unpack_kernel<12>, grid, dim3(BS), 0, f->d, f->g, n0, count, d_in, d_part,
               d_dev);
}

}  // namespace lt
