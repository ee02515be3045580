// lettuce_hip.hpp -- header-only C++ layer over the C ABI (lettuce_hip.h) with the
// reference's names, for C++ callers such as the reference's own driver
// (SU3_main.cpp).  RAII handles; a failing call throws lettuce_hip::Error.
#pragma once

#include <array>
#include <complex>
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <initializer_list>
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

// GaugeAction::Rectangular<2> (actions/gauge/rectangular_action.hpp:29-300): plaquette
// and 1x2 rectangle coefficients; c_rect = 0 is the Wilson action
struct RectangularAction {
    double beta;
    double c_plaq = 1.0, c_rect = 0.0;
};
inline RectangularAction WilsonGaugeAction(double beta) { return {beta, 1.0, 0.0}; }
inline RectangularAction LuescherWeiszAction(double beta) { return {beta, 1.0 + 8.0 / 12.0, -1.0 / 12.0}; }
inline RectangularAction IwasakiAction(double beta) { return {beta, 1.0 + 8.0 * 0.331, -0.331}; }
inline RectangularAction DBW2Action(double beta) { return {beta, 1.0 + 8.0 * 1.4088, -1.4088}; }

// HeatbathKernel / OverrelaxationSubgroupKernel (updates/heatbath.hpp:106-108,
// updates/overrelaxation.hpp:87-89): whole sweeps, driven by Iterator below; with an
// improved action (c_rect != 0) only Checkerboard4 applies
struct HeatbathKernel {
    GaugeField& U;
    double beta;
    RectangularAction action{0.0, 1.0, 0.0};
    void Sweep(int n, int schedule) const
    {
        if (action.c_rect == 0.0 && action.c_plaq == 1.0)
            LETTUCE_HIP_CALL(lt_heatbath_sweep(U.handle(), beta, n, schedule));
        else
            LETTUCE_HIP_CALL(
                lt_heatbath_sweep_action(U.handle(), beta, action.c_plaq, action.c_rect, n, schedule));
    }
};
struct OverrelaxationSubgroupKernel {
    GaugeField& U;
    RectangularAction action{0.0, 1.0, 0.0};
    void Sweep(int n, int schedule) const
    {
        if (action.c_rect == 0.0 && action.c_plaq == 1.0)
            LETTUCE_HIP_CALL(lt_or_sweep(U.handle(), n, schedule));
        else
            LETTUCE_HIP_CALL(lt_or_sweep_action(U.handle(), action.c_plaq, action.c_rect, n, schedule));
    }
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
inline void StoutSmearingN(const GaugeField& in, GaugeField& out, int n_smear, double rho = 0.12)
{
    LETTUCE_HIP_CALL(lt_stout(in.handle(), out.handle(), rho, n_smear));
}
inline double WilsonActionValue(const GaugeField& U, double beta)
{
    double s = 0.0;
    LETTUCE_HIP_CALL(lt_wilson_action(U.handle(), beta, &s));
    return s;
}
// Rectangular<2>::ActionNormalized, e.g. the tree-level Symanzik action of Observables()
inline double ActionNormalized(const GaugeField& U, const RectangularAction& a)
{
    double s = 0.0;
    LETTUCE_HIP_CALL(lt_action_rectangular(U.handle(), a.c_plaq, a.c_rect, &s));
    return s;
}
inline double TopChargePlaquette(const GaugeField& U)
{
    double q = 0.0;
    LETTUCE_HIP_CALL(lt_topcharge_plaquette(U.handle(), &q));
    return q;
}
// EnergyDensity::Plaquette / Clover / CloverTimeslice (field_strength_tensor.hpp:136-196)
struct EnergyDensities {
    double plaquette = 0.0, clover = 0.0;
    std::vector<double> clover_timeslice;
};
inline EnergyDensities EnergyDensity(const GaugeField& U)
{
    EnergyDensities e;
    e.clover_timeslice.resize(U.Shape()[0]);
    LETTUCE_HIP_CALL(lt_energy(U.handle(), &e.plaquette, &e.clover, e.clover_timeslice.data()));
    return e;
}
// PolyakovLoop (observables/polyakov_loop.hpp:31-49)
inline std::complex<double> PolyakovLoop(const GaugeField& U)
{
    double re = 0.0, im = 0.0;
    LETTUCE_HIP_CALL(lt_polyakov_loop(U.handle(), &re, &im));
    return {re, im};
}
// ParityUpdate (updates/parity_update.hpp:22-47); U_copy is overwritten
inline void ParityUpdate(GaugeField& U, GaugeField& U_copy) { LETTUCE_HIP_CALL(lt_parity_update(U.handle(), U_copy.handle())); }

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

// A field of general 3x3 matrices (HMC momenta, forces)
class MatrixField {
  public:
    MatrixField(Context& ctx, int nt, int nx, int ny, int nz)
    {
        LETTUCE_HIP_CALL(lt_matrix_field_create(ctx.handle(), nt, nx, ny, nz, &h_));
    }
    ~MatrixField() { lt_field_destroy(h_); }
    MatrixField(const MatrixField&) = delete;
    MatrixField& operator=(const MatrixField&) = delete;
    lt_field handle() const { return h_; }

  private:
    lt_field h_ = nullptr;
};

// WilsonLoop<n_start, n_end, reset>(U, U_chain) (observables/wilson_loop.hpp:63-128);
// U_chain keeps the straight link products between calls
inline double WilsonLoop(const GaugeField& U, MatrixField& U_chain, int n_start, int n_end, bool reset)
{
    double w = 0.0;
    LETTUCE_HIP_CALL(lt_wilson_loop(U.handle(), U_chain.handle(), n_start, n_end, reset ? 1 : 0, &w));
    return w;
}
// CalculateClover<1> + CloverDerivative on every link (observables/clover.hpp:59-89,238-321)
inline void CloverDerivative(const GaugeField& U, MatrixField& F, double scale = 1.0)
{
    LETTUCE_HIP_CALL(lt_clover_derivative(U.handle(), F.handle(), scale));
}

// The kicks and drifts an integrator issues: kind 'P' (both forces), 'G' (gauge force),
// 'M' (bias force) or 'U' (field update), step size, path-sample flag
struct ScheduleEvent {
    char kind;
    double eps;
    bool sub;
};
inline std::vector<ScheduleEvent> IntegratorSchedule(int integrator, double tau, int n_step)
{
    int n = 0;
    LETTUCE_HIP_CALL(lt_integrator_schedule(integrator, tau, n_step, nullptr, nullptr, nullptr, 0, &n));
    std::vector<char> k(n);
    std::vector<double> e(n);
    std::vector<int> s(n);
    LETTUCE_HIP_CALL(lt_integrator_schedule(integrator, tau, n_step, k.data(), e.data(), s.data(), n, &n));
    std::vector<ScheduleEvent> out;
    for (int i = 0; i < n; ++i) out.push_back({k[i], e[i], s[i] != 0});
    return out;
}

// HMCKernel(U, U_copy, P, Integrator, Action, prng, ..., trajectory_length)
// (updates/hmc_gauge.hpp:381-600); momenta from the field's Philox stream
class HMCKernel {
  public:
    HMCKernel(GaugeField& U, GaugeField& U_copy, MatrixField& P, int integrator, double beta, double tau)
        : U_(U), Uc_(U_copy), P_(P), integrator_(integrator), beta_(beta), tau_(tau)
    {
    }
    bool operator()(int n_step, bool metropolis_step = true)
    {
        int acc = 0;
        LETTUCE_HIP_CALL(lt_hmc_trajectory(U_.handle(), P_.handle(), Uc_.handle(), beta_, integrator_, tau_, n_step,
                                           metropolis_step ? 1 : 0, &delta_H, &acc));
        return acc != 0;
    }
    // integrate, reverse the momenta, integrate back; {H_old, H_new}
    std::array<double, 2> ReversibilityTest(int n_step)
    {
        std::array<double, 2> h{};
        LETTUCE_HIP_CALL(
            lt_hmc_reversibility_test(U_.handle(), P_.handle(), beta_, integrator_, tau_, n_step, &h[0], &h[1]));
        return h;
    }
    double delta_H = 0.0;

  private:
    GaugeField& U_;
    GaugeField& Uc_;
    MatrixField& P_;
    int integrator_;
    double beta_, tau_;
};

// MetaBiasPotential (actions/bias_potential/metadynamics.hpp:29-507); throws on bad ranges like the reference
class MetaBiasPotential {
  public:
    MetaBiasPotential(double cv_min, double cv_max, int bins, double weight, double width, double wt_param,
                      double threshold, bool well_tempered)
    {
        LETTUCE_HIP_CALL(lt_bias_create(cv_min, cv_max, bins, weight, width, well_tempered ? 1 : 0, wt_param,
                                        threshold, &h_));
    }
    ~MetaBiasPotential() { lt_bias_destroy(h_); }
    MetaBiasPotential(const MetaBiasPotential&) = delete;
    void UpdatePotential(double cv) { LETTUCE_HIP_CALL(lt_bias_deposit(h_, cv, 0)); }
    void UpdatePotentialSymmetric(double cv) { LETTUCE_HIP_CALL(lt_bias_deposit(h_, cv, 1)); }
    double ReturnPotential(double cv) const
    {
        double v = 0.0;
        LETTUCE_HIP_CALL(lt_bias_potential(h_, cv, &v));
        return v;
    }
    double ReturnDerivative(double cv) const
    {
        double v = 0.0;
        LETTUCE_HIP_CALL(lt_bias_derivative(h_, cv, &v));
        return v;
    }
    // SaveParameters / SavePotential / LoadPotential (metadynamics.hpp:304-505): the
    // reference's text files; saving appends unless overwrite, loading returns false on
    // a missing or foreign file
    void SaveParameters(const std::string& filename, bool overwrite = false) const
    {
        LETTUCE_HIP_CALL(lt_bias_save_parameters(h_, filename.c_str(), overwrite ? 1 : 0));
    }
    void SavePotential(const std::string& filename, bool overwrite = false) const
    {
        LETTUCE_HIP_CALL(lt_bias_save_potential(h_, filename.c_str(), overwrite ? 1 : 0));
    }
    bool LoadPotential(const std::string& filename) { return lt_bias_load_potential(h_, filename.c_str()) == LT_OK; }
    lt_bias handle() const { return h_; }

  private:
    lt_bias h_ = nullptr;
};

// The CV workspace of HMCMetaDData (updates/hmc_metadynamics.hpp:251-260):
// n stout levels (checkpoint interval 0 = all stored) and the adjoint buffers
class CollectiveVariable {
  public:
    CollectiveVariable(Context& ctx, int nt, int nx, int ny, int nz, int n_levels, double rho, int checkpoint = 0)
    {
        LETTUCE_HIP_CALL(lt_cv_create_checkpointed(ctx.handle(), nt, nx, ny, nz, n_levels, rho, checkpoint, &h_));
    }
    ~CollectiveVariable() { lt_cv_destroy(h_); }
    CollectiveVariable(const CollectiveVariable&) = delete;
    double Forward(const GaugeField& U)
    {
        double q = 0.0;
        LETTUCE_HIP_CALL(lt_cv_forward(h_, U.handle(), &q));
        return q;
    }
    void Backward(double dvdq, MatrixField& F) { LETTUCE_HIP_CALL(lt_cv_backward(h_, dvdq, F.handle())); }
    lt_cv handle() const { return h_; }

  private:
    lt_cv h_ = nullptr;
};

// HMCMetaDKernel(U, U_copy, P, bias, data, Integrator, Action, ..., tau, rho)
// (updates/hmc_metadynamics.hpp:264-675) with the hand-written adjoint force
class HMCMetaDKernel {
  public:
    // hmc_log_path: where each trajectory appends its CV path (LogCVData,
    // hmc_metadynamics.hpp:537-545); empty = no log
    HMCMetaDKernel(GaugeField& U, GaugeField& U_copy, MatrixField& P, MatrixField& F, MetaBiasPotential& bias,
                   CollectiveVariable& cv, int integrator, double beta, double tau, bool update_bias = true,
                   std::string hmc_log_path = {})
        : U_(U), Uc_(U_copy), P_(P), F_(F), bias_(bias), cv_(cv), integrator_(integrator), beta_(beta), tau_(tau),
          update_bias_(update_bias), log_(std::move(hmc_log_path))
    {
    }
    bool operator()(int n_step, bool metropolis_step = true)
    {
        int acc = 0;
        LETTUCE_HIP_CALL(lt_metad_hmc_trajectory(U_.handle(), P_.handle(), Uc_.handle(), F_.handle(), cv_.handle(),
                                                 bias_.handle(), beta_, integrator_, tau_, n_step,
                                                 metropolis_step ? 1 : 0, update_bias_ ? 1 : 0, &delta_H, &acc,
                                                 &cv_value));
        if (!log_.empty() && n_step > 0) {
            LETTUCE_HIP_CALL(lt_cv_log_path(cv_.handle(), 0, log_.c_str()));
            if (metropolis_step) LETTUCE_HIP_CALL(lt_cv_log_path(cv_.handle(), 1, log_.c_str()));
        }
        return acc != 0;
    }
    // integrate with the bias, reverse the momenta, integrate back; {H_old, H_new}
    std::array<double, 2> ReversibilityTest(int n_step)
    {
        std::array<double, 2> h{};
        LETTUCE_HIP_CALL(lt_metad_reversibility_test(U_.handle(), P_.handle(), F_.handle(), cv_.handle(),
                                                     bias_.handle(), beta_, integrator_, tau_, n_step, &h[0], &h[1]));
        return h;
    }
    double delta_H = 0.0, cv_value = 0.0;

  private:
    GaugeField& U_;
    GaugeField& Uc_;
    MatrixField& P_;
    MatrixField& F_;
    MetaBiasPotential& bias_;
    CollectiveVariable& cv_;
    int integrator_;
    double beta_, tau_;
    bool update_bias_;
    std::string log_;
};

// MetadynamicsTemperingKernel(U, U_temper, ..., bias, prng, ...) (updates/tempering.hpp:24-73):
// each call proposes the swap of the unbiased stream U and the MetaD stream U_temper,
// with the uniform from Philox stream 4 at (seed, counter), and advances the counter
class MetadynamicsTemperingKernel {
  public:
    MetadynamicsTemperingKernel(GaugeField& U, GaugeField& U_temper, MetaBiasPotential& bias, CollectiveVariable& cv,
                                std::uint64_t seed, std::uint64_t counter = 0)
        : U_(U), Ut_(U_temper), bias_(bias), cv_(cv), seed_(seed), counter_(counter)
    {
    }
    bool operator()()
    {
        int acc = 0;
        LETTUCE_HIP_CALL(lt_metad_tempering_swap(U_.handle(), Ut_.handle(), cv_.handle(), bias_.handle(), seed_,
                                                 counter_, &acc, &delta_V));
        ++counter_;
        return acc != 0;
    }
    double delta_V = 0.0;

  private:
    GaugeField& U_;
    GaugeField& Ut_;
    MetaBiasPotential& bias_;
    CollectiveVariable& cv_;
    std::uint64_t seed_, counter_;
};

// Replica batch: the reference's two-stream MetadynamicsTemperingKernel
// (updates/tempering.hpp:24-73) generalised to a ladder of bias strengths, all
// replicas advanced by the same launches (optionally one captured graph per step)
class ReplicaBatch {
  public:
    ReplicaBatch(Context& ctx, const std::vector<std::array<int, 4>>& shapes, int n_levels, double rho)
    {
        std::vector<int> flat;
        for (const auto& s : shapes) flat.insert(flat.end(), s.begin(), s.end());
        LETTUCE_HIP_CALL(lt_batch_create(ctx.handle(), (int)shapes.size(), flat.data(), n_levels, rho, &h_));
        n_ = (int)shapes.size();
    }
    ~ReplicaBatch() { lt_batch_destroy(h_); }
    ReplicaBatch(const ReplicaBatch&) = delete;
    // the replica's links in the reference AoS layout
    void Upload(int r, const double* aos) { LETTUCE_HIP_CALL(lt_field_upload(field(r), aos)); }
    void Download(int r, double* aos) const { LETTUCE_HIP_CALL(lt_field_download(field(r), aos)); }
    void SeedPRNG(int r, std::uint64_t seed, std::uint64_t counter = 0)
    {
        LETTUCE_HIP_CALL(lt_batch_set_rng(h_, r, seed, counter));
    }
    void SetLadder(const std::vector<double>& strengths, std::uint64_t pt_seed = 0)
    {
        LETTUCE_HIP_CALL(lt_batch_set_ladder(h_, strengths.data(), nullptr, pt_seed, 0));
    }
    void SetBias(const MetaBiasPotential& b) { LETTUCE_HIP_CALL(lt_batch_set_bias(h_, b.handle())); }
    void GetBias(MetaBiasPotential& b) { LETTUCE_HIP_CALL(lt_batch_get_bias(h_, b.handle())); }
    void UseGraph(bool on) { LETTUCE_HIP_CALL(lt_batch_use_graph(h_, on ? 1 : 0)); }
    void Autotune(const char* cache = nullptr) { LETTUCE_HIP_CALL(lt_batch_autotune(h_, cache)); }
    // n_traj PT-MetaD steps; swap_mode LT_SWAP_*
    void Step(int n_step, double beta, int integrator, double tau, bool update_bias, int swap_mode, int n_traj = 1)
    {
        LETTUCE_HIP_CALL(lt_batch_metad_hmc(h_, beta, integrator, tau, n_step, 1, update_bias ? 1 : 0, swap_mode,
                                            n_traj));
    }
    std::vector<int> Labels() const
    {
        std::vector<int> l(n_);
        LETTUCE_HIP_CALL(lt_batch_results(h_, nullptr, nullptr, nullptr, l.data()));
        return l;
    }
    int Size() const { return n_; }

  private:
    lt_field field(int r) const
    {
        lt_field f = nullptr;
        LETTUCE_HIP_CALL(lt_batch_field(h_, r, &f));
        return f;
    }
    lt_batch h_ = nullptr;
    int n_ = 0;
};

// SaveConfigBMW / SaveConfigBMWFull / LoadConfigBMW / LoadConfigBMWFull
// (IO/config_io/bmw_format.hpp): packed and unpacked on the device; bool results
// as in the reference (false: file exists without overwrite, missing or foreign
// file, other extents, links not in SU(3) within 1e-6)
inline bool SaveConfigBMW(const GaugeField& U, const std::string& filename, bool overwrite = false)
{
    return lt_field_save_bmw(U.handle(), filename.c_str(), 0, overwrite ? 1 : 0, nullptr, 0) == LT_OK;
}
inline bool SaveConfigBMWFull(const GaugeField& U, const std::string& filename, bool overwrite = false)
{
    return lt_field_save_bmw(U.handle(), filename.c_str(), 1, overwrite ? 1 : 0, nullptr, 0) == LT_OK;
}
inline bool LoadConfigBMW(GaugeField& U, const std::string& filename)
{
    int ok = 0;
    double dev = 0.0;
    return lt_field_load_bmw(U.handle(), filename.c_str(), 0, &ok, &dev) == LT_OK;
}
inline bool LoadConfigBMWFull(GaugeField& U, const std::string& filename)
{
    int ok = 0;
    double dev = 0.0;
    return lt_field_load_bmw(U.handle(), filename.c_str(), 1, &ok, &dev) == LT_OK;
}

// Generator state of a field: seed and counter of its Philox stream, in place of
// the PCG state files of PRNG4D::SaveState / LoadState; the second file (the
// normal distribution's cached value there) only carries a marker, the Gaussian
// draws here keep no state
struct PhiloxState {
    GaugeField& U;
    bool SaveState(const std::filesystem::path& prng_file, const std::filesystem::path& distribution_file,
                   bool overwrite = false) const
    {
        if (!overwrite && (std::filesystem::exists(prng_file) || std::filesystem::exists(distribution_file)))
            return false;
        std::uint64_t seed = 0, counter = 0;
        LETTUCE_HIP_CALL(lt_field_get_rng(U.handle(), &seed, &counter));
        std::ofstream(prng_file) << "philox4x32-10 " << seed << " " << counter << "\n";
        std::ofstream(distribution_file) << "philox4x32-10 stateless\n";
        return true;
    }
    bool LoadState(const std::filesystem::path& prng_file, const std::filesystem::path& = {}) const
    {
        std::ifstream in(prng_file);
        std::string tag;
        std::uint64_t seed = 0, counter = 0;
        if (!(in >> tag >> seed >> counter) || tag != "philox4x32-10") return false;
        LETTUCE_HIP_CALL(lt_field_set_rng(U.handle(), seed, counter));
        return true;
    }
};

// CheckpointManager (IO/config_io/checkpoint_manager.hpp:28-159): the newest
// checkpoint in the directory, older generations in backup1, backup2, ... (higher
// is older), n_checkpoint_files generations kept
class CheckpointManager {
  public:
    CheckpointManager(std::filesystem::path dir, int n_checkpoint_files, bool create_subdirectories)
        : dir_(std::move(dir)), n_(n_checkpoint_files)
    {
        if (create_subdirectories) CreateSubdirectories();
    }
    const std::filesystem::path& CheckpointDirectory() const { return dir_; }
    static std::string ReturnBackupSubdirectory(int k) { return k == 0 ? "" : "backup" + std::to_string(k); }
    void CreateSubdirectories() const
    {
        for (int k = 1; k < n_; ++k) std::filesystem::create_directories(dir_ / ReturnBackupSubdirectory(k));
    }
    static void RotateFiles(const std::filesystem::path& current, const std::filesystem::path& old, bool copy = true)
    {
        if (!std::filesystem::exists(current)) return;
        std::error_code err;
        if (copy)
            std::filesystem::copy_file(current, old, std::filesystem::copy_options::overwrite_existing, err);
        else
            std::filesystem::rename(current, old, err);
    }
    template <typename SaveT>
    void AlternatingConfigCheckpoints(SaveT&& save, const GaugeField& U, const std::string& config) const
    {
        Rotate({config});
        save(U, (dir_ / config).string(), true);
    }
    template <typename SaveT, typename PrngT>
    void AlternatingCheckpoints(SaveT&& save, PrngT& prng, const GaugeField& U, const std::string& config,
                                const std::string& prng_file, const std::string& distribution_file) const
    {
        Rotate({config, prng_file, distribution_file});
        save(U, (dir_ / config).string(), true);
        prng.SaveState(dir_ / prng_file, dir_ / distribution_file, true);
    }
    template <typename LoadT, typename PrngT>
    bool LoadCheckpoint(LoadT&& load, PrngT& prng, GaugeField& U, const std::string& config,
                        const std::string& prng_file, const std::string& distribution_file) const
    {
        const bool ok = load(U, (dir_ / config).string());
        return prng.LoadState(dir_ / prng_file, dir_ / distribution_file) && ok;
    }
    static void CreateCompletedFile(const std::filesystem::path& where) { std::ofstream(where / "completed_run"); }

  private:
    // generation k-1 -> k from the oldest down; the newest is moved, not copied
    void Rotate(std::initializer_list<std::string> names) const
    {
        for (int k = n_ - 1; k > 0; --k)
            for (const auto& name : names)
                RotateFiles(dir_ / ReturnBackupSubdirectory(k - 1) / name, dir_ / ReturnBackupSubdirectory(k) / name,
                            k != 1);
    }
    std::filesystem::path dir_;
    int n_;
};

}  // namespace lettuce_hip
