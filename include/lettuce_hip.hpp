// lettuce_hip.hpp -- header-only C++ layer over the C ABI (lettuce_hip.h) with the
// reference's names, for C++ callers such as the reference's own driver
// (SU3_main.cpp).  RAII handles; a failing call throws lettuce_hip::Error.
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "lettuce_hip.h"

namespace lettuce_hip {

struct Error : std::runtime_error {
    int code;
    Error(const char* fn, int c) : std::runtime_error(std::string(fn) + ": " + lt_last_error()), code(c) {}
};

#define LETTUCE_HIP_CALL(expr)                        \
    do {                                              \
        int rc_ = (expr);                             \
        if (rc_ != LT_OK) throw Error(#expr, rc_);    \
    } while (0)

class Context {
  public:
    explicit Context(int device = 0) { LETTUCE_HIP_CALL(lt_ctx_create(device, &h_)); }
    ~Context() { lt_ctx_destroy(h_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    void Synchronize() { LETTUCE_HIP_CALL(lt_ctx_synchronize(h_)); }
    lt_ctx handle() const { return h_; }

  private:
    lt_ctx h_ = nullptr;
};

// GaugeField4D (LettuceGauge/lattice.hpp:87-394) with runtime extents
class GaugeField {
  public:
    GaugeField(Context& ctx, int nt, int nx, int ny, int nz) : shape_{nt, nx, ny, nz}
    {
        LETTUCE_HIP_CALL(lt_field_create(ctx.handle(), nt, nx, ny, nz, &h_));
    }
    ~GaugeField() { lt_field_destroy(h_); }
    GaugeField(const GaugeField&) = delete;
    GaugeField& operator=(const GaugeField& other)
    {
        LETTUCE_HIP_CALL(lt_field_copy(h_, other.h_));
        return *this;
    }
    friend void Swap(GaugeField& a, GaugeField& b) { LETTUCE_HIP_CALL(lt_field_swap(a.h_, b.h_)); }
    // aos: 18 doubles per link in the reference order ((((t*Nx+x)*Ny+y)*Nz+z)*4+mu)
    void Upload(const double* aos) { LETTUCE_HIP_CALL(lt_field_upload(h_, aos)); }
    void Download(double* aos) const { LETTUCE_HIP_CALL(lt_field_download(h_, aos)); }
    void SetToIdentity() { LETTUCE_HIP_CALL(lt_field_set_identity(h_)); }
    void SetToHaar(std::uint64_t seed) { LETTUCE_HIP_CALL(lt_field_set_haar(h_, seed)); }
    void SeedPRNG(std::uint64_t seed, std::uint64_t counter = 0) { LETTUCE_HIP_CALL(lt_field_set_rng(h_, seed, counter)); }
    std::size_t Volume() const { return (std::size_t)shape_[0] * shape_[1] * shape_[2] * shape_[3]; }
    std::array<int, 4> Shape() const { return shape_; }
    lt_field handle() const { return h_; }

  private:
    lt_field h_ = nullptr;
    std::array<int, 4> shape_;
};

// HeatbathKernel / OverrelaxationSubgroupKernel (updates/heatbath.hpp:106-108,
// updates/overrelaxation.hpp:87-89): whole sweeps, driven by Iterator below
struct HeatbathKernel {
    GaugeField& U;
    double beta;
    void Sweep(int n, int schedule) const { LETTUCE_HIP_CALL(lt_heatbath_sweep(U.handle(), beta, n, schedule)); }
};
struct OverrelaxationSubgroupKernel {
    GaugeField& U;
    void Sweep(int n, int schedule) const { LETTUCE_HIP_CALL(lt_or_sweep(U.handle(), n, schedule)); }
};

// Iterator::Checkerboard / Checkerboard4 (iterators/iterators.hpp:65-169)
namespace Iterator {
template <typename KernelT>
void Checkerboard(const KernelT& k, int n_sweep = 1) { k.Sweep(n_sweep, LT_CB2); }
template <typename KernelT>
void Checkerboard4(const KernelT& k, int n_sweep = 1) { k.Sweep(n_sweep, LT_CB4); }
}  // namespace Iterator

inline double PlaquetteSum(const GaugeField& U)
{
    double v = 0.0;
    LETTUCE_HIP_CALL(lt_plaquette_sum(U.handle(), &v));
    return v;
}
inline double TopChargeClover(const GaugeField& U)
{
    double q = 0.0;
    LETTUCE_HIP_CALL(lt_topcharge_clover(U.handle(), &q, nullptr));
    return q;
}
inline std::vector<double> TopChargeCloverTimeslice(const GaugeField& U)
{
    std::vector<double> qt(U.Shape()[0]);
    double q = 0.0;
    LETTUCE_HIP_CALL(lt_topcharge_clover(U.handle(), &q, qt.data()));
    return qt;
}
inline void StoutSmearing4D(const GaugeField& in, GaugeField& out, double rho = 0.12)
{
    LETTUCE_HIP_CALL(lt_stout(in.handle(), out.handle(), rho, 1));
}

// GradientFlowKernel<Integrator, Rectangular<1>> (smearing/gradient_flow.hpp:252-367)
class GradientFlowKernel {
  public:
    GradientFlowKernel(const GaugeField& unflowed, GaugeField& flowed, int integrator, double epsilon)
    {
        LETTUCE_HIP_CALL(lt_flow_create(unflowed.handle(), flowed.handle(), integrator, epsilon, &h_));
    }
    ~GradientFlowKernel() { lt_flow_destroy(h_); }
    GradientFlowKernel(const GradientFlowKernel&) = delete;
    void operator()(int n_step) { LETTUCE_HIP_CALL(lt_flow_run(h_, n_step)); }
    void Resume(int n_step) { LETTUCE_HIP_CALL(lt_flow_resume(h_, n_step)); }

  private:
    lt_flow h_ = nullptr;
};

}  // namespace lettuce_hip
