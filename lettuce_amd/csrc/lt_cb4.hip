// lt_cb4.hip -- the four-colour (Checkerboard4) sweeps, iterators.hpp:141-169.
//
// A colour class of direction mu != z is every fourth z of a row, i.e. every other
// checkerboard index of one parity array.  In the plain cb order its loads and
// stores use half of every 128-byte line.  This translation unit is compiled with
// LT_SPLIT_CB = 1: its kernels address links in the order where every aligned group
// of 16 checkerboard indices holds its 8 even ones first (cbi in lt_device.hpp), in
// which a colour class reads and writes whole lines.  A four-colour sweep call
// reorders the field in place into that order (cb_layout_kernel), runs its colour
// launches and restores the plain order, so every other kernel of the library, all
// compiled with the plain order, sees the field as always.  The device code of each
// translation unit is its own code object; the inline device functions of the two
// orders never meet.
#define LT_SPLIT_CB 1
#include "lt_internal.hpp"
#include "lt_bodies.hpp"
#include "lt_sweep.hpp"

namespace lt {

namespace {

constexpr int BS = LT_LINK_BS;
constexpr int RBS = 256;

// One thread per link of the class (cb4_site), consecutive threads along z.
// Colours 1 and 3 walk the blocks backwards (xcd_block_dir): consecutive colour
// launches then start where the previous one ended.
template <bool HEATBATH>
__global__ void LT_LINK_KERNEL sweep_cb4_kernel(double2* __restrict__ U, Geom g, int mu, int colour,
                                                 double prefactor, uint64_t seed, uint64_t counter)
{
    int c[4], s;
    if (!cb4_site(g, mu, colour, xcd_block_dir(blockIdx.x, gridDim.x, colour & 1) * BS + threadIdx.x, c, s)) return;
    update_link<HEATBATH>(U, g, s, c, mu, parity_of(c), prefactor, seed, counter);
}

// heatbath of a captured update: Philox counter state[0] + k from device memory
__global__ void LT_LINK_KERNEL hb_cb4_graph_kernel(double2* __restrict__ U, Geom g, int mu, int colour, double prefactor,
                                                   uint64_t seed, const uint64_t* __restrict__ state, int k)
{
    int c[4], s;
    if (!cb4_site(g, mu, colour, xcd_block_dir(blockIdx.x, gridDim.x, colour & 1) * BS + threadIdx.x, c, s)) return;
    update_link<true>(U, g, s, c, mu, parity_of(c), prefactor, seed, state[0] + (uint64_t)k);
}

// Rectangular<R>(beta, c_plaq, c_rect) sweeps; the rectangle staple's sum waits in LDS
// between its terms (9 x 16 bytes per lane: two waves per SIMD without spills, against
// one wave and 22 spilled VGPRs when it stays in registers)
template <bool HEATBATH>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(2)))
sweep_action_kernel(double2* __restrict__ U, Geom g, int mu, int colour, double prefactor, uint64_t seed,
                    uint64_t counter, double c_plaq, double c_rect)
{
    extern __shared__ double2 lds[];
    int c[4], s;
    if (!cb4_site(g, mu, colour, xcd_block_dir(blockIdx.x, gridDim.x, colour & 1) * BS + threadIdx.x, c, s)) return;
    update_link_action<HEATBATH>(U, g, s, c, mu, parity_of(c), prefactor, seed, counter, c_plaq, c_rect,
                                 lds + threadIdx.x, BS);
}

// PlaquetteSum block partials of a captured four-colour update (the field is in the
// split order inside the graph); the same terms, launch and partials as the
// plain-order kernel
__global__ void LT_LINK_KERNEL cb4_plaq_kernel(const double2* __restrict__ U, Geom g, LinkMap m, double* part)
{
    plaq_link_partials(U, g, m, part);
}

// In-place reorder of rows 0 and 1 of every link between the plain cb order and the
// split order.  One lane per slot of a group of 16 (a group lies in one wavefront):
// each lane loads its slot, the wavefront's loads complete before its stores, and no
// other wavefront touches the group.  A last group of 8 (Vh not a multiple of 16)
// keeps its order in both layouts.
__global__ void __launch_bounds__(RBS) cb_layout_kernel(double2* __restrict__ U, Geom g, int to_split)
{
    const long long vh16 = g.Vh & ~15;
    const long long id = (long long)blockIdx.x * RBS + threadIdx.x;
    if (id >= 48 * vh16) return;
    const int a = (int)(id / vh16);            // (mu*2 + parity)*6 + component
    const int e = (int)(id - a * vh16);
    const int grp = e & ~15, l = e & 15;
    const int dst = to_split ? ((l & 1) << 3) | (l >> 1)    // cb l -> slot
                             : (l < 8 ? 2 * l : 2 * (l - 8) + 1);   // slot l -> cb
    double2* p = U + (size_t)(a % 6) * slot_stride(g);
    const double2 v = p[slot_base(g, a / 6, 9, grp + l)];
    p[slot_base(g, a / 6, 9, grp + dst)] = v;
}

}  // namespace

void k_cb_layout(lt_field u, int to_split)
{
    if (!u->su3) throw LtError{LT_EINVAL, "four-colour sweeps need an SU(3) gauge field"};
    const long long n = 48LL * (u->g.Vh & ~15);
    if (n) launch(u->ctx, "cb_layout", cb_layout_kernel, dim3(nblocks(n, RBS)), dim3(RBS), 0, u->d, u->g, to_split);
}

void k_cb4_colour(lt_field u, bool heatbath, int mu, int colour, double prefactor, uint64_t counter, const char* name)
{
    const dim3 grid(nblocks(u->g.V / 4, BS)), block(BS);
    if (heatbath)
        launch(u->ctx, name, sweep_cb4_kernel<true>, grid, block, 0, u->d, u->g, mu, colour, prefactor, u->seed, counter);
    else
        launch(u->ctx, name, sweep_cb4_kernel<false>, grid, block, 0, u->d, u->g, mu, colour, prefactor, u->seed,
               counter);
}

void k_cb4_colour_graph_hb(lt_field u, int mu, int colour, double prefactor, const uint64_t* state, int k)
{
    launch(u->ctx, "hb_cb4", hb_cb4_graph_kernel, dim3(nblocks(u->g.V / 4, BS)), dim3(BS), 0, u->d, u->g, mu, colour,
           prefactor, u->seed, state, k);
}

void k_cb4_colour_action(lt_field u, bool heatbath, int mu, int colour, double prefactor, uint64_t counter,
                         double c_plaq, double c_rect, const char* name)
{
    const dim3 grid(nblocks(u->g.V / 4, BS)), block(BS);
    if (heatbath)
        launch(u->ctx, name, sweep_action_kernel<true>, grid, block, 9 * BS * sizeof(double2), u->d, u->g, mu, colour,
               prefactor, u->seed, counter, c_plaq, c_rect);
    else
        launch(u->ctx, name, sweep_action_kernel<false>, grid, block, 9 * BS * sizeof(double2), u->d, u->g, mu, colour,
               prefactor, u->seed, counter, c_plaq, c_rect);
}

void k_cb4_plaq_partials(lt_field u, double* part)
{
    launch(u->ctx, "plaquette", cb4_plaq_kernel, dim3(8 * map_blocks(u->map)), dim3(u->map.nthreads), 0,
           (const double2*)u->d, u->g, u->map, part);
}

}  // namespace lt
