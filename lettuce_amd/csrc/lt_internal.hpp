// lt_internal.hpp -- host-side objects behind the C ABI: contexts (device, stream,
// scratch, per-kernel event timing) and fields (device link buffer + Philox state),
// plus the launch helper every kernel goes through.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <exception>
#include <functional>
#include <map>
#include <string>
#include <vector>

#include "lettuce_hip.h"
#include "lt_device.hpp"

namespace lt {

void set_error(const std::string& msg);

struct LtError {
    int code;
    std::string msg;
};

#define LT_HIP_CHECK(expr)                                                                          \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            throw ::lt::LtError{LT_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)};        \
    } while (0)

// C ABI bodies: exceptions become a status code plus the lt_last_error message.
template <typename F>
int guarded(F&& f)
{
    try {
        return f();
    } catch (const LtError& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::exception& e) {
        set_error(e.what());
        return LT_ESTATE;
    }
}

inline int fail(int code, const std::string& msg)
{
    set_error(msg);
    return code;
}

struct ProfEntry {
    int launches = 0;
    double ms = 0.0;
};

}  // namespace lt

struct lt_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    int live = 0;              // fields, flows, CV workspaces and batches still holding this context
    // device scratch for reduction partials, grown on demand
    double* scratch = nullptr;
    size_t scratch_doubles = 0;
    // pinned host staging for reduction results
    double* host = nullptr;
    size_t host_doubles = 0;
    bool profiling = false;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> event_pool;
    std::map<std::string, lt::ProfEntry> prof;

    double* scratch_for(size_t n);
    double* host_for(size_t n);
    void collect_profile();
};

namespace lt {
struct UpdateGraph;   // captured local-update loop body of a field (lt_update_sweeps)
}

struct lt_field_s {
    lt_ctx ctx = nullptr;
    lt::Geom g{};
    lt::LinkMap map{};         // thread -> site tiling of the per-parity link kernels
    bool su3 = true;           // rows 0,1 stored, row 2 implied (false: general 3x3 matrices)
    bool view = false;         // d belongs to a replica batch (not freed, not swappable)
    double2* d = nullptr;      // 4 * 2 * 9 * Vh complex doubles
    uint64_t seed = 0;
    uint64_t counter = 0;
    lt::UpdateGraph* ug = nullptr;
    size_t bytes() const { return (size_t)72 * g.Vh * sizeof(double2); }
};

// MetaBiasPotential (LettuceGauge/actions/bias_potential/metadynamics.hpp:29-507)
struct lt_bias_s {
    std::vector<double> grid;
    double cv_current = 0.0, cv_min = 0.0, cv_max = 0.0, bin_width = 0.0, bin_width_inverse = 0.0;
    double weight = 0.0, width = 0.0, wt_param = 0.0, threshold = 0.0;
    int bins = 0;
    int well_tempered = 0;
    unsigned long long exceeded = 0;

    int bin_of(double cv) const { return (int)std::floor((cv - cv_min) * bin_width_inverse); }
    double cv_of(int i) const { return cv_min + i * bin_width; }
    // UpdatePotential (metadynamics.hpp:103-124)
    void deposit(double cv)
    {
        if (cv < cv_min || cv >= cv_max) exceeded += 1;
        const double wis = 1.0 / (width * width);
        for (size_t b = 0; b < grid.size(); ++b) {
            const double dist = cv - cv_of((int)b);
            if (well_tempered)
                grid[b] += weight * std::exp(-0.5 * wis * dist * dist - grid[b] / wt_param);
            else
                grid[b] += weight * std::exp(-0.5 * wis * dist * dist);
        }
    }
    // ReturnPotential (metadynamics.hpp:151-179)
    double potential(double cv) const
    {
        const int i = bin_of(cv);
        if ((unsigned)i < (unsigned)(bins)) {
            const double ic = (cv - cv_of(i)) * bin_width_inverse;
            return grid[i] * (1.0 - ic) + ic * grid[i + 1];
        }
        if (i < 0) {
            const double d = cv - cv_min;
            return grid.front() + threshold * d * d;
        }
        const double d = cv - cv_max;
        return grid.back() + threshold * d * d;
    }
    // ReturnDerivative (metadynamics.hpp:181-187)
    double derivative(double cv) const
    {
        const double h = 0.5 * bin_width;
        return bin_width_inverse * (potential(cv + h) - potential(cv - h));
    }
};

struct lt_flow_s {
    lt_field unflowed = nullptr;
    lt_field flowed = nullptr;
    int integrator = LT_FLOW_RK3;
    double eps = 0.0;
    double2* z = nullptr;      // RK3 Z field: su(3), 8 reals per link (ld_alg8 / st_alg8 layout)
    double2* tmp = nullptr;    // second link buffer for the Jacobi stages
};

namespace lt {

// Kernel launch with optional event timing on the context stream.
template <typename Kernel, typename... Args>
void launch(lt_ctx ctx, const char* name, Kernel kernel, dim3 grid, dim3 block, size_t shmem, Args... args)
{
    if (grid.x == 0 || grid.y == 0 || grid.z == 0) return;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ctx->profiling) {
        if (!ctx->event_pool.empty()) {
            e0 = ctx->event_pool.back().first;
            e1 = ctx->event_pool.back().second;
            ctx->event_pool.pop_back();
        } else {
            LT_HIP_CHECK(hipEventCreate(&e0));
            LT_HIP_CHECK(hipEventCreate(&e1));
        }
        LT_HIP_CHECK(hipEventRecord(e0, ctx->stream));
    }
    hipLaunchKernelGGL(kernel, grid, block, shmem, ctx->stream, args...);
    LT_HIP_CHECK(hipGetLastError());
    if (ctx->profiling) {
        LT_HIP_CHECK(hipEventRecord(e1, ctx->stream));
        ctx->pending.push_back({name, {e0, e1}});
        if (ctx->pending.size() > 4096) ctx->collect_profile();
    }
}

inline unsigned nblocks(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }
// grid size of a tiled link launch; refuses maps the kernels were not compiled for
inline unsigned map_blocks(const LinkMap& m)
{
    if (m.nthreads < 1 || m.nthreads > LT_LINK_BS || m.T[0] * m.T[1] * m.T[2] * m.T[3] != m.nthreads || m.tc < 1
        || m.N[0] % m.tc || m.yc < 1 || m.N[2] % m.yc || m.ycs < 1 || m.N[2] % m.ycs)
        throw LtError{LT_ESTATE, "invalid thread-block tile for the link kernels"};
    return (unsigned)(m.N[0] * m.N[1] * m.N[2] * m.N[3]);
}

// The momentum / field update sequence of Integrators::HMC::{Leapfrog, OMF_2, OMF_4}
// (LettuceGauge/updates/hmc_gauge.hpp:48-75,155-190,241-295).  P(eps, sub) is a
// momentum update, sub marks the updates whose CV the MetaD path update samples.
void hmc_schedule(int kind, double tau, int n_step, const std::function<void(double, bool)>& P,
                  const std::function<void(double)>& U);
inline bool valid_integrator(int k)
{
    return k == LT_INT_LEAPFROG || k == LT_INT_OMF2 || k == LT_INT_OMF4 || k == LT_INT_LEAPFROG_SLOW ||
           k == LT_INT_OMF2_SLOW || k == LT_INT_OMF4_SLOW;
}

// Momentum kicks of a MetaD-HMC schedule: gauge force, bias force, or both in one update.
enum Kick { KICK_GAUGE = 1, KICK_BIAS = 2, KICK_BOTH = 3 };
// A KICK_BOTH momentum update P += a Algebra(st U^+) + bf F applied by the last kernel
// of the force computation instead of a separate pass over the stored force F.
struct MomentumKick {
    double2* P;
    double a, bf;
};
// Fields of a stout stack by level: level(j) is level j's field, tape(j) the exponent
// tape of the step from level j to level j + 1.
typedef std::function<const double2*(int)> LevelFn;
// The stout-force recursion (StoutForceRecursion, stout_smearing.hpp:122-234) from level
// hi down to level lo: the Lambda of level hi, then for k = hi..lo+1 the sigma update of
// level k, fused with the Lambda of level k-1 while k-1 > lo (the two Lambda buffers
// alternate); Sig holds the force, propagated in place.  kick: applied by the update of
// level 1 instead of storing the level-0 force (lo == 0 only).  top_lambda_done: the
// Lambda of level hi is already in lam[0] (the clover derivative made it).
void k_stout_recursion_run(lt_ctx ctx, const Geom& g, const LinkMap& m, const LevelFn& level, const LevelFn& tape,
                           int hi, int lo, double2* Sig, double2* const lam[2], double rho,
                           const MomentumKick* kick = nullptr, bool top_lambda_done = false);
// The level below in a stout-force recursion step fused with its Lambda: its input
// field, its exponent tape and the Lambda buffer to write.
struct SigmaNext {
    const double2* U;
    const double2* T;
    double2* Lam;
};
// The MetaD-HMC sequence: the single-scale integrators kick with both forces
// (HMCMetaDKernel::UpdateMomenta, hmc_metadynamics.hpp:441-459); the multiple-timescale
// ones (hmc_metadynamics.hpp:28-238) kick the bias force on the outer scale and the
// gauge force on the inner OMF4.  P(eps, sub, kick); sub marks the path samples.
void metad_schedule(int kind, double tau, int n_step, const std::function<void(double, bool, int)>& P,
                    const std::function<void(double)>& U);
inline bool valid_metad_integrator(int k)
{
    return valid_integrator(k) || k == LT_INT_LEAPFROG_OMF4 || k == LT_INT_OMF2_OMF4_SLOW || k == LT_INT_OMF2_OMF4;
}

Geom make_geom(int nt, int nx, int ny, int nz);
// tile map with extents (tt, tx, ty, tz) over (t, x, y, z/2); returns false if they do not divide
bool make_linkmap(const Geom& g, int tt, int tx, int ty, int tz, LinkMap& m);
LinkMap default_linkmap(const Geom& g);
// same block shape and site order
inline bool same_map(const LinkMap& a, const LinkMap& b)
{
    for (int i = 0; i < 4; ++i)
        if (a.T[i] != b.T[i] || a.N[i] != b.N[i]) return false;
    return a.nthreads == b.nthreads && a.tc == b.tc && a.yc == b.yc && a.lin == b.lin && a.ycs == b.ycs &&
           a.xr == b.xr;
}

// kernel launchers (defined in the .hip translation units)
void k_copy(lt_ctx ctx, const double2* src, double2* dst, size_t n);
// read-rate probe: nb blocks of 256 lanes, each reading iters 16-byte elements of buf
// (mask + 1 elements, a power of two) and writing one to out (nb * 256 elements)
void k_l2_read(lt_ctx ctx, const double2* buf, unsigned mask, int iters, double2* out, unsigned nb);
// y += a x over n complex entries (the bias-force kick P += eps F)
void k_axpy(lt_ctx ctx, const double2* x, double2* y, double a, size_t n);
// y *= a over n complex entries (ReverseMomenta with a = -1)
void k_scale(lt_ctx ctx, double2* y, double a, size_t n);
void k_pack(lt_field f, const double* d_aos);
void k_unpack(lt_field f, double* d_aos);
void k_identity(lt_field f);
void k_haar(lt_field f, uint64_t seed);
void k_reunitarize(lt_field f);
void k_check(lt_field f, double* max_dev, long long* nonfinite);
void k_staple_field(lt_field u, lt_field out);
void k_or_sweep(lt_field u, int n, int schedule);
void k_hb_sweep(lt_field u, double beta, int n, int schedule);
// four-colour sweeps (lt_cb4.hip): the field's rows 0, 1 reordered in place between the
// plain cb order and the split order (to_split 1 / 0), and single colour launches that
// expect the split order
void k_cb_layout(lt_field u, int to_split);
void k_cb4_colour(lt_field u, bool heatbath, int mu, int colour, double prefactor, uint64_t counter, const char* name);
void k_cb4_colour_graph_hb(lt_field u, int mu, int colour, double prefactor, const uint64_t* state, int k);
void k_cb4_colour_action(lt_field u, bool heatbath, int mu, int colour, double prefactor, uint64_t counter,
                         double c_plaq, double c_rect, const char* name);
void k_cb4_plaq_partials(lt_field u, double* part);   // one partial per block of u->map's launch
// n_update x (n_hb heatbath + n_or overrelaxation sweeps) as one graph replayed per
// update, PlaquetteSum after each into plaq_host (may be null); and its teardown
void k_update_graph(lt_field u, double beta, int n_hb, int n_or, int schedule, int n_update, double* plaq_host);
void drop_update_graph(lt_field f);
// fused observables: per-timeslice plaquette sum, raw q, E_clov numerator -> host[3*Nt]
// BMW sites [n0, n0 + count) <-> big-endian file words (4 * K per site, K = 12 or
// 18), with per-block checksum partials (A, B) and, on unpacking, row-2 deviations
void k_bmw_pack(lt_field f, int K, long long n0, long long count, uint64_t* d_out, uint64_t* d_part);
void k_bmw_unpack(lt_field f, int K, long long n0, long long count, const uint64_t* d_in, uint64_t* d_part,
                  double* d_dev);
void k_parity(lt_field f, const double2* copy);
void k_staple_field_action(lt_field u, lt_field out, double c_plaq, double c_rect);
void k_or_sweep_action(lt_field u, int n, double c_plaq, double c_rect);
void k_heatbath_sweep_action(lt_field u, double beta, int n, double c_plaq, double c_rect);
void k_observables(lt_field u, double* host_out);
// sum over sites and planes of Re tr P, the planes of a site spread over its links (the HMC
// Hamiltonian keeps k_plaquette_sum, one site per lane, which the replica batch repeats)
double k_plaquette_planes(lt_field u);
// per-site clover charge density (V doubles, lexicographic order) into host memory
void k_topcharge_density(lt_field u, double* host_q);
// the fixed observables scratch of lt_observables_noalloc: its address on `device`
// (which must be current), looked up once per device
void noalloc_setup(int device);
// lt_observables' sums through the fixed scratch, final values computed on the device
void k_observables_noalloc(lt_field u, double* host_out);
void k_rect_sums(lt_field u, double out[3]);
double k_wilson_loop(lt_field u, lt_field chain, int n0, int n1, int reset);
void k_polyakov_loop(lt_field u, double* re, double* im);
void k_stout(lt_ctx ctx, const Geom& g, const LinkMap& m, const double2* in, double2* out, double rho);
void k_flow_stage(lt_ctx ctx, const Geom& g, const LinkMap& m, const double2* in, double2* out, double2* z, double c_old,
                  double c_new, double c_exp, int read_z, int write_z);

}  // namespace lt
