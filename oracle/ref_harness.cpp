// ref_harness.cpp -- drives the reference Lettuce code (compiled from its headers
// where they lie) on configurations written by this project's oracle, to produce
// golden vectors.  TEST INFRASTRUCTURE ONLY: it is built by oracle/Makefile into
// oracle/_ref/ and run only by tests/golden/make_golden.py and bench.py's optional
// reference CPU timing.  Lattice extents are the reference's compile-time
// constants (16^4, LettuceGauge/defines.hpp:34-37).
//
// usage: lettuce_ref <command> [args...]; fields are raw little-endian doubles in the
// reference AoS layout (LettuceGauge/lattice.hpp:366-382).

#include "LettuceGauge/defines.hpp"
#include "LettuceGauge/coords.hpp"
#include "LettuceGauge/lattice.hpp"
#include "LettuceGauge/actions/bias_potential/metadynamics.hpp"
#include "LettuceGauge/actions/gauge/rectangular_action.hpp"
#include "LettuceGauge/actions/gauge/wilson_action.hpp"
#include "LettuceGauge/iterators/iterators.hpp"
#include "LettuceGauge/math/su2.hpp"
#include "LettuceGauge/math/su3.hpp"
#include "LettuceGauge/math/su3_exp.hpp"
#include "LettuceGauge/observables/clover.hpp"
#include "LettuceGauge/observables/plaquette.hpp"
#include "LettuceGauge/observables/field_strength_tensor.hpp"
#include "LettuceGauge/observables/topological_charge.hpp"
#include "LettuceGauge/smearing/stout_smearing.hpp"
#include "LettuceGauge/smearing/gradient_flow.hpp"
#include "LettuceGauge/updates/heatbath.hpp"
#include "LettuceGauge/updates/overrelaxation.hpp"
#include "LettuceGauge/updates/hmc_gauge.hpp"
#include "LettuceGauge/updates/hmc_metadynamics.hpp"
#include "LettuceGauge/updates/tempering.hpp"
#include "LettuceGauge/IO/config_io/bmw_format.hpp"
#include "LettuceGauge/observables/wilson_loop.hpp"
#include "LettuceGauge/observables/polyakov_loop.hpp"
#include "LettuceGauge/updates/parity_update.hpp"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace {

constexpr std::size_t kLinks = 4ull * Nt * Nx * Ny * Nz;

void read_field(const char* path, GaugeField& U)
{
    FILE* f = std::fopen(path, "rb");
    if (!f) { std::perror(path); std::exit(2); }
    for (std::size_t i = 0; i < kLinks; ++i)
        if (std::fread(U(i).data(), sizeof(double), 18, f) != 18) { std::fprintf(stderr, "short read %s\n", path); std::exit(2); }
    std::fclose(f);
}

void write_field(const char* path, const GaugeField& U)
{
    FILE* f = std::fopen(path, "wb");
    if (!f) { std::perror(path); std::exit(2); }
    for (std::size_t i = 0; i < kLinks; ++i) std::fwrite(U(i).data(), sizeof(double), 18, f);
    std::fclose(f);
}

void write_values(const char* path, const std::vector<double>& v)
{
    FILE* f = std::fopen(path, "wb");
    if (!f) { std::perror(path); std::exit(2); }
    std::fwrite(v.data(), sizeof(double), v.size(), f);
    std::fclose(f);
}

void set_beta(double b)
{
    beta = b;                                  // global read by HMC forces (hmc_gauge.hpp:461)
    GaugeAction::WilsonAction.SetBeta(b);      // member read by the action (rectangular_action.hpp:44-47)
}

// Observables in the order used by the fixtures:
// [plaq_sum, wilson_action(beta), Q_U, Q_tensor, Q_timeslice_sum, E_plaq, E_clov, q_t[Nt], e_t[Nt]]
std::vector<double> observables(const GaugeField& U)
{
    static FullTensor Clov;
    static FullTensor F;
    std::vector<double> v;
    v.push_back(PlaquetteSum(U));
    v.push_back(WilsonAction::Action(U));
    v.push_back(TopChargeClover(U));
    CalculateClover<1>(U, Clov);
    v.push_back(TopChargeClover(Clov));
    FieldStrengthTensor::Clover(U, F);
    std::vector<double> qt(Nt), et(Nt);
    double qs = 0.0;
    for (int t = 0; t < Nt; ++t) { qt[t] = TopChargeCloverTimeslice(F, t); qs += qt[t]; }
    v.push_back(qs);
    v.push_back(EnergyDensity::Plaquette(U));
    FieldStrengthTensor::MakeComponentsTraceless(F);
    double es = 0.0;
    for (int t = 0; t < Nt; ++t) { et[t] = EnergyDensity::CloverTimeslice(F, t); es += et[t]; }
    v.push_back(es);
    v.insert(v.end(), qt.begin(), qt.end());
    v.insert(v.end(), et.begin(), et.end());
    return v;
}

template<typename F>
double time_it(F&& f)
{
    auto t0 = std::chrono::steady_clock::now();
    f();
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Friend integrator: OMF_2 sequence of Integrators::HMC::OMF_2 (hmc_gauge.hpp:155-190)
// run on injected momenta, recording H before and after.
struct InjectedOMF2
{
    static constexpr bool merged_momentum_updates = true;
    static constexpr bool path_update_compatible  = true;
    const GaugeField* P0 {nullptr};
    double H_old {0.0}, H_new {0.0};

    template<typename HMCFunctor>
    void operator()(HMCFunctor& HMC, const double trajectory_length, const int n_step)
    {
        HMC.Momentum = *P0;
        H_old = HMC.Hamiltonian();
        Integrators::HMC::OMF_2 omf;
        run(HMC, omf, trajectory_length, n_step);
        H_new = HMC.Hamiltonian();
    }
    template<typename HMCFunctor>
    static void run(HMCFunctor& HMC, Integrators::HMC::OMF_2&, const double trajectory_length, const int n_step)
    {
        floatT epsilon {static_cast<floatT>(trajectory_length)/n_step};
        const double alpha {0.1931833275037836 * epsilon};
        const double b     {0.5 * epsilon};
        const double gamma {(1.0 - 2.0 * 0.1931833275037836) * epsilon};
        HMC.UpdateMomenta(alpha);
        HMC.UpdateFields(b);
        HMC.UpdateMomenta(gamma);
        HMC.UpdateFields(b);
        for (int step_count = 0; step_count < n_step - 1; ++step_count)
        {
            HMC.UpdateMomenta(2.0 * alpha, true);
            HMC.UpdateFields(b);
            HMC.UpdateMomenta(gamma);
            HMC.UpdateFields(b);
        }
        HMC.UpdateMomenta(alpha, true);
    }
};

// Friend integrator for HMCMetaDKernel: same OMF_2 sequence on injected momenta;
// H includes the bias at the CV of MetaCharge (hmc_metadynamics.hpp:556-583).
struct InjectedOMF2MetaD
{
    static constexpr bool merged_momentum_updates = true;
    static constexpr bool path_update_compatible  = true;
    const GaugeField* P0 {nullptr};
    double H_old {0.0}, H_new {0.0}, CV_old {0.0}, CV_new {0.0};
    std::vector<double> samples;

    template<typename HMCFunctor>
    void operator()(HMCFunctor& HMC, const double trajectory_length, const int n_step)
    {
        HMC.Momentum = *P0;
        CV_old = HMC.MetaCharge();
        H_old = HMC.Hamiltonian() + HMC.Metapotential.ReturnPotential(CV_old);
        // OMF_2 (hmc_gauge.hpp:155-190), written out: only this struct is the kernel's friend
        floatT epsilon {static_cast<floatT>(trajectory_length)/n_step};
        const double alpha {0.1931833275037836 * epsilon};
        const double b     {0.5 * epsilon};
        const double gamma {(1.0 - 2.0 * 0.1931833275037836) * epsilon};
        HMC.UpdateMomenta(alpha);
        HMC.UpdateFields(b);
        HMC.UpdateMomenta(gamma);
        HMC.UpdateFields(b);
        for (int step_count = 0; step_count < n_step - 1; ++step_count)
        {
            HMC.UpdateMomenta(2.0 * alpha, true);
            HMC.UpdateFields(b);
            HMC.UpdateMomenta(gamma);
            HMC.UpdateFields(b);
        }
        HMC.UpdateMomenta(alpha, true);
        samples = HMC.cv_path_samples;
        CV_new = HMC.MetaCharge();
        H_new = HMC.Hamiltonian() + HMC.Metapotential.ReturnPotential(CV_new);
    }
};

// Stands in for the HMC functor of an integrator: records every momentum kick and
// field drift with its step size, so the integrator's own operator() yields its
// schedule (kind P = UpdateMomenta, G = UpdateGaugeMomenta, M = UpdateMetadynamicsMomenta,
// U = UpdateFields; the flag is UpdateMomenta's submeasurement argument).
struct ScheduleRecorder
{
    void UpdateMomenta(const floatT e, const bool sub = false) { std::printf("P %.17g %d\n", e, (int)sub); }
    void UpdateGaugeMomenta(const floatT e) { std::printf("G %.17g 0\n", e); }
    void UpdateMetadynamicsMomenta(const floatT e) { std::printf("M %.17g 0\n", e); }
    void UpdateFields(const floatT e) { std::printf("U %.17g 0\n", e); }
};

template<typename IntegratorT>
void print_schedule(const double tau, const int n_step)
{
    ScheduleRecorder rec;
    IntegratorT{}(rec, tau, n_step);
}

// Friend integrator that only sets the kernel's cv_path_samples, so that the
// kernel's own LogCVData writes them (hmc_metadynamics.hpp:537-545, 613).
struct SetPathSamples
{
    static constexpr bool merged_momentum_updates = true;
    static constexpr bool path_update_compatible  = true;
    std::vector<double> values;

    template<typename HMCFunctor>
    void operator()(HMCFunctor& HMC, const double, const int)
    {
        HMC.cv_path_samples = values;
    }
};

} // namespace

int main(int argc, char** argv)
{
    std::ios_base::sync_with_stdio(false);
    if (argc < 2) { std::fprintf(stderr, "usage: lettuce_ref <cmd> ...\n"); return 2; }
    std::string cmd = argv[1];
    auto U  = std::make_unique<GaugeField>();
    auto U2 = std::make_unique<GaugeField>();
    std::printf("extents %d %d %d %d\n", Nt, Nx, Ny, Nz);

    if (cmd == "staple") {                       // staple in out
        read_field(argv[2], *U);
        #pragma omp parallel for
        for (std::size_t s = 0; s < kLinks / 4; ++s)
            for (int mu = 0; mu < 4; ++mu) {
                int z = s % Nz, y = (s / Nz) % Ny, x = (s / Nz / Ny) % Nx, t = s / Nz / Ny / Nx;
                (*U2)({t, x, y, z, mu}) = GaugeAction::WilsonAction.Staple(*U, {t, x, y, z, mu});
            }
        write_field(argv[3], *U2);
    } else if (cmd == "or") {                    // or in out schedule(2|4) n
        read_field(argv[2], *U);
        set_beta(6.0);
        OverrelaxationSubgroupKernel ORk(*U, GaugeAction::WilsonAction);
        int n = std::atoi(argv[5]);
        if (std::atoi(argv[4]) == 4) Iterator::Checkerboard4(ORk, n);
        else Iterator::Checkerboard(ORk, n);
        write_field(argv[3], *U);
    } else if (cmd == "obs") {                   // obs in out beta
        read_field(argv[2], *U);
        set_beta(std::atof(argv[4]));
        write_values(argv[3], observables(*U));
    } else if (cmd == "stout") {                 // stout in out rho nlev (out = last level)
        read_field(argv[2], *U);
        set_beta(6.0);
        double rho = std::atof(argv[4]);
        int n = std::atoi(argv[5]);
        for (int k = 0; k < n; ++k) { StoutSmearing4D(*U, *U2, rho); std::swap(U, U2); }
        write_field(argv[3], *U);
    } else if (cmd == "flow") {                  // flow in out eps nsteps integrator(rk3|euler) obsout
        read_field(argv[2], *U);
        set_beta(6.0);
        double eps = std::atof(argv[4]);
        int n = std::atoi(argv[5]);
        std::string integ = argv[6];
        auto Force = std::make_unique<GaugeField>();
        std::vector<double> obs;
        if (integ == "rk3") {
            Integrators::GradientFlow::RK3 I(*Force);
            GradientFlowKernel Flow(*U, *U2, I, GaugeAction::WilsonAction, eps);
            Flow(n);
        } else {
            Integrators::GradientFlow::Euler I(*Force);
            GradientFlowKernel Flow(*U, *U2, I, GaugeAction::WilsonAction, eps);
            Flow(n);
        }
        write_field(argv[3], *U2);
        if (argc > 7) write_values(argv[7], observables(*U2));
    } else if (cmd == "clover_deriv") {          // clover_deriv in out
        read_field(argv[2], *U);
        auto Clov = std::make_unique<FullTensor>();
        CalculateClover<1>(*U, *Clov);
        #pragma omp parallel for
        for (std::size_t s = 0; s < kLinks / 4; ++s)
            for (int mu = 0; mu < 4; ++mu) {
                int z = s % Nz, y = (s / Nz) % Ny, x = (s / Nz / Ny) % Nx, t = s / Nz / Ny / Nx;
                (*U2)({t, x, y, z, mu}) = CloverDerivative(*U, *Clov, {t, x, y, z}, mu);
            }
        write_field(argv[3], *U2);
    } else if (cmd == "cv_force") {              // cv_force in out rho nlev  (dV/dQ = 1)
        read_field(argv[2], *U);
        double rho = std::atof(argv[4]);
        int n = std::atoi(argv[5]);
        GaugeFieldSmeared S(n + 1);
        GaugeField4DSmeared<Nt, Nx, Ny, Nz, SU3::ExpConstants> E(n > 0 ? n : 1);
        auto Clov = std::make_unique<FullTensor>();
        S[0] = *U;
        for (int k = 0; k < n; ++k) StoutSmearing4DWithConstants(S[k], S[k + 1], E[k], rho);
        CalculateClover<1>(S[n], *Clov);
        double Q = TopChargeClover(*Clov);
        #pragma omp parallel for
        for (std::size_t s = 0; s < kLinks / 4; ++s)
            for (int mu = 0; mu < 4; ++mu) {
                int z = s % Nz, y = (s / Nz) % Ny, x = (s / Nz / Ny) % Nx, t = s / Nz / Ny / Nx;
                (*U2)({t, x, y, z, mu}) = CloverDerivative(S[n], *Clov, {t, x, y, z}, mu);
            }
        for (int k = n; k > 0; --k) StoutForceRecursion(S[k - 1], S[k], *U2, E[k - 1], rho);
        write_field(argv[3], *U2);
        std::printf("Q %.17g\n", Q);
    } else if (cmd == "fd") {                    // fd in out rho nlev h nlinks idx...  (numerical_derivative.hpp:71-249)
        read_field(argv[2], *U);
        double rho = std::atof(argv[4]);
        int n = std::atoi(argv[5]);
        double h = std::atof(argv[6]);
        int nl = std::atoi(argv[7]);
        std::vector<double> out;
        auto W = std::make_unique<GaugeField>();
        auto A = std::make_unique<GaugeField>();
        auto B = std::make_unique<GaugeField>();
        auto smearQ = [&](const GaugeField& X) {
            *A = X;
            for (int k = 0; k < n; ++k) { StoutSmearing4D(*A, *B, rho); std::swap(A, B); }
            return TopChargeClover(*A);
        };
        for (int l = 0; l < nl; ++l) {
            std::size_t li = std::strtoull(argv[8 + l], nullptr, 10);
            for (int a = 1; a <= 8; ++a) {
                *W = *U;
                (*W)(li) = SU3::Generators::Exp_lambda(a, h) * (*U)(li);
                double qf = smearQ(*W);
                *W = *U;
                (*W)(li) = SU3::Generators::Exp_lambda(a, -h) * (*U)(li);
                double qb = smearQ(*W);
                out.push_back(0.5 * (qf - qb) / h);
            }
        }
        write_values(argv[3], out);
    } else if (cmd == "hmc") {                   // hmc in P out beta tau nstep  -> out field, prints H
        read_field(argv[2], *U);
        auto P = std::make_unique<GaugeField>();
        read_field(argv[3], *P);
        set_beta(std::atof(argv[5]));
        double tau = std::atof(argv[6]);
        int nstep = std::atoi(argv[7]);
        auto Ucopy = std::make_unique<GaugeField>();
        auto Mom = std::make_unique<GaugeField>();
        pcg64 g(1);
        PRNG4D<Nt, Nx, Ny, Nz, pcg64, floatT, int> prng(g);
        RunStatistics stats;
        std::ofstream devnull("/dev/null");
        InjectedOMF2 integ;
        integ.P0 = P.get();
        auto act = GaugeAction::WilsonAction;
        GaugeUpdates::HMCKernel HMC(*U, *Ucopy, *Mom, integ, act, prng, stats, devnull, tau);
        HMC(nstep, false);
        write_field(argv[4], *U);
        std::printf("H_old %.17g\nH_new %.17g\n", integ.H_old, integ.H_new);
    } else if (cmd == "metad_hmc") {             // metad_hmc in P out gridout beta tau nstep rho cvmin cvmax bins w width wt thr ndep dep...
        read_field(argv[2], *U);
        auto P = std::make_unique<GaugeField>();
        read_field(argv[3], *P);
        set_beta(std::atof(argv[6]));
        double tau = std::atof(argv[7]);
        int nstep = std::atoi(argv[8]);
        double rho = std::atof(argv[9]);
        MetaBiasPotential B(std::atof(argv[10]), std::atof(argv[11]), std::atoi(argv[12]), std::atof(argv[13]),
                            std::atof(argv[14]), std::atof(argv[15]), std::atof(argv[16]));
        int k = 17;
        int ndep = std::atoi(argv[k++]);
        std::vector<double> deps;
        for (int i = 0; i < ndep; ++i) deps.push_back(std::atof(argv[k++]));
        B.UpdatePotentialSymmetric(deps);
        auto Ucopy = std::make_unique<GaugeField>();
        auto Mom = std::make_unique<GaugeField>();
        pcg64 g(1);
        PRNG4D<Nt, Nx, Ny, Nz, pcg64, floatT, int> prng(g);
        RunStatistics stats;
        std::ofstream devnull("/dev/null");
        std::string logpath = "/dev/null";
        GaugeUpdates::HMCMetaDData data(n_smear_meta);
        InjectedOMF2MetaD integ;
        integ.P0 = P.get();
        auto act = GaugeAction::WilsonAction;
        GaugeUpdates::HMCMetaDKernel HMC(*U, *Ucopy, *Mom, B, data, integ, act, prng, stats, devnull, logpath, tau, rho);
        HMC(nstep, false);
        write_field(argv[4], *U);
        std::vector<double> grid;
        for (int i = 0; i <= std::atoi(argv[12]); ++i)
            grid.push_back(B.ReturnPotential(std::atof(argv[10]) + i * (std::atof(argv[11]) - std::atof(argv[10])) / std::atoi(argv[12])));
        write_values(argv[5], grid);
        std::printf("H_old %.17g\nH_new %.17g\nCV_old %.17g\nCV_new %.17g\n", integ.H_old, integ.H_new, integ.CV_old, integ.CV_new);
        for (double s : integ.samples) std::printf("sample %.17g\n", s);
    } else if (cmd == "bias") {                  // bias out cvmin cvmax bins weight width wt thr ndep dep... neval eval...
        // MetaBiasPotential (metadynamics.hpp:29-507); the reference compiles the
        // well-tempered rule in (defines.hpp:72)
        MetaBiasPotential B(std::atof(argv[3]), std::atof(argv[4]), std::atoi(argv[5]), std::atof(argv[6]),
                            std::atof(argv[7]), std::atof(argv[8]), std::atof(argv[9]));
        int k = 10;
        int ndep = std::atoi(argv[k++]);
        std::vector<double> deps;
        for (int i = 0; i < ndep; ++i) deps.push_back(std::atof(argv[k++]));
        B.UpdatePotentialSymmetric(deps);
        int neval = std::atoi(argv[k++]);
        std::vector<double> out;
        for (int i = 0; i < neval; ++i) {
            double cv = std::atof(argv[k++]);
            out.push_back(B.ReturnPotential(cv));
            out.push_back(B.ReturnDerivative(cv));
        }
        write_values(argv[2], out);
    } else if (cmd == "hb_stats") {              // hb_stats out beta n_update n_or seed [in]
        set_beta(std::atof(argv[3]));
        int nupd = std::atoi(argv[4]);
        int nor = std::atoi(argv[5]);
        pcg64 g(std::strtoull(argv[6], nullptr, 10));
        PRNG4D<Nt, Nx, Ny, Nz, pcg64, floatT, int> prng(g);
        if (argc > 7) read_field(argv[7], *U); else U->SetToIdentity();
        auto act = GaugeAction::WilsonAction;
        HeatbathKernel HB(*U, act, prng);
        OverrelaxationSubgroupKernel ORk(*U, act);
        std::vector<double> series;
        for (int k = 0; k < nupd; ++k) {
            Iterator::Checkerboard4(HB, 1);
            Iterator::Checkerboard4(ORk, nor);
            series.push_back(PlaquetteSum(*U) / (6.0 * 3.0 * U->Volume()));
        }
        write_values(argv[2], series);
    } else if (cmd == "hb_qstats") {             // hb_qstats out beta n_update n_or seed n_stout rho
        // As hb_stats, recording per update the average plaquette and the clover charge
        // after n_stout StoutSmearing4D levels (the metadynamics CV, hmc_metadynamics.hpp:
        // 302-310), from a cold start: pairs (P, Q) per update.
        set_beta(std::atof(argv[3]));
        int nupd = std::atoi(argv[4]);
        int nor = std::atoi(argv[5]);
        pcg64 g(std::strtoull(argv[6], nullptr, 10));
        int nst = std::atoi(argv[7]);
        double rho = std::atof(argv[8]);
        PRNG4D<Nt, Nx, Ny, Nz, pcg64, floatT, int> prng(g);
        U->SetToIdentity();
        auto act = GaugeAction::WilsonAction;
        HeatbathKernel HB(*U, act, prng);
        OverrelaxationSubgroupKernel ORk(*U, act);
        auto A = std::make_unique<GaugeField>();
        auto B = std::make_unique<GaugeField>();
        std::vector<double> series;
        for (int k = 0; k < nupd; ++k) {
            Iterator::Checkerboard4(HB, 1);
            Iterator::Checkerboard4(ORk, nor);
            series.push_back(PlaquetteSum(*U) / (6.0 * 3.0 * U->Volume()));
            *A = *U;
            for (int l = 0; l < nst; ++l) { StoutSmearing4D(*A, *B, rho); std::swap(A, B); }
            series.push_back(TopChargeClover(*A));
        }
        write_values(argv[2], series);
    } else if (cmd == "time") {                  // time in beta nrep: seconds per HB sweep, OR sweep (CB2), RK3 step
        read_field(argv[2], *U);
        set_beta(std::atof(argv[3]));
        int nrep = std::atoi(argv[4]);
        pcg64 g(7);
        PRNG4D<Nt, Nx, Ny, Nz, pcg64, floatT, int> prng(g);
        auto act = GaugeAction::WilsonAction;
        HeatbathKernel HB(*U, act, prng);
        OverrelaxationSubgroupKernel ORk(*U, act);
        double thb = time_it([&] { Iterator::Checkerboard(HB, nrep); }) / nrep;
        double tor = time_it([&] { Iterator::Checkerboard(ORk, nrep); }) / nrep;
        auto Force = std::make_unique<GaugeField>();
        Integrators::GradientFlow::RK3 I(*Force);
        GradientFlowKernel Flow(*U, *U2, I, GaugeAction::WilsonAction, 0.01);
        double tfl = time_it([&] { Flow(nrep); }) / nrep;
        std::printf("hb_sweep_s %.6g\nor_sweep_s %.6g\nrk3_step_s %.6g\nthreads %d\n", thb, tor, tfl, omp_get_max_threads());
    } else if (cmd == "obs2") {                  // obs2 in out beta: the rest of Observables() (SU3_main.cpp:125-136)
        read_field(argv[2], *U);
        const double beta = std::atof(argv[4]);
        GaugeAction::Rectangular<2> Symanzik(beta, 1.0 + 8.0 * 1.0 / 12.0, -1.0 / 12.0);
        std::vector<double> out;
        out.push_back(Symanzik.ActionNormalized(*U));
        out.push_back(TopChargePlaquette(*U));
        out.push_back(WilsonLoop<0, 2, true>(*U, *U2));
        out.push_back(WilsonLoop<2, 4, false>(*U, *U2));
        out.push_back(WilsonLoop<4, 8, false>(*U, *U2));
        const std::complex<double> P = PolyakovLoop(*U);
        out.push_back(std::real(P));
        out.push_back(std::imag(P));
        write_values(argv[3], out);
    } else if (cmd == "staple_action") {         // staple_action in out c_plaq c_rect: Rectangular<2>::Staple
        read_field(argv[2], *U);
        GaugeAction::Rectangular<2> act(6.0, std::atof(argv[4]), std::atof(argv[5]));
        #pragma omp parallel for
        for (std::size_t s = 0; s < kLinks / 4; ++s)
            for (int mu = 0; mu < 4; ++mu) {
                int z = s % Nz, y = (s / Nz) % Ny, x = (s / Nz / Ny) % Nx, t = s / Nz / Ny / Nx;
                (*U2)({t, x, y, z, mu}) = act.Staple(*U, {t, x, y, z, mu});
            }
        write_field(argv[3], *U2);
    } else if (cmd == "or_action") {             // or_action in out c_plaq c_rect n: Checkerboard4 OR, Rectangular<2>
        read_field(argv[2], *U);
        GaugeAction::Rectangular<2> act(6.0, std::atof(argv[4]), std::atof(argv[5]));
        OverrelaxationSubgroupKernel ORk(*U, act);
        Iterator::Checkerboard4(ORk, std::atoi(argv[6]));
        write_field(argv[3], *U);
    } else if (cmd == "parity") {                // parity in out: ParityUpdate (parity_update.hpp:22-47)
        read_field(argv[2], *U);
        ParityUpdate(*U, *U2);
        write_field(argv[3], *U);
    } else if (cmd == "bmw_save") {              // bmw_save in out full(0|1): the reference's BMW writer
        read_field(argv[2], *U);
        const bool ok = std::atoi(argv[4]) ? SaveConfigBMWFull(*U, argv[3], true) : SaveConfigBMW(*U, argv[3], true);
        if (!ok) return 3;
    } else if (cmd == "bmw_load") {              // bmw_load file out full(0|1): the reference's BMW reader
        const bool ok = std::atoi(argv[4]) ? LoadConfigBMWFull(*U, argv[2]) : LoadConfigBMW(*U, argv[2]);
        write_field(argv[3], *U);
        std::printf("load_ok %d\n", (int)ok);
    } else if (cmd == "bias_save") {             // bias_save out cvmin cvmax bins weight width wt thr ndep dep...
        MetaBiasPotential B(std::atof(argv[3]), std::atof(argv[4]), std::atoi(argv[5]), std::atof(argv[6]),
                            std::atof(argv[7]), std::atof(argv[8]), std::atof(argv[9]));
        int k = 10;
        int ndep = std::atoi(argv[k++]);
        std::vector<double> deps;
        for (int i = 0; i < ndep; ++i) deps.push_back(std::atof(argv[k++]));
        B.UpdatePotentialSymmetric(deps);
        B.SaveParameters(argv[2], true);
        B.SavePotential(argv[2]);
    } else if (cmd == "schedule") {              // schedule name tau n_step: kicks and drifts of an integrator
        const std::string name = argv[2];
        const double tau = std::atof(argv[3]);
        const int n = std::atoi(argv[4]);
        using namespace Integrators::HMC;
        if (name == "Leapfrog") print_schedule<Leapfrog>(tau, n);
        else if (name == "Leapfrog_slow") print_schedule<Leapfrog_slow>(tau, n);
        else if (name == "OMF_2") print_schedule<OMF_2>(tau, n);
        else if (name == "OMF_2_slow") print_schedule<OMF_2_slow>(tau, n);
        else if (name == "OMF_4") print_schedule<OMF_4>(tau, n);
        else if (name == "OMF_4_slow") print_schedule<OMF_4_slow>(tau, n);
        else if (name == "Leapfrog_OMF_4") print_schedule<Leapfrog_OMF_4>(tau, n);
        else if (name == "OMF_2_OMF_4_slow") print_schedule<OMF_2_OMF_4_slow>(tau, n);
        else if (name == "OMF_2_OMF_4") print_schedule<OMF_2_OMF_4>(tau, n);
        else { std::fprintf(stderr, "unknown integrator %s\n", name.c_str()); return 2; }
    } else if (cmd == "cvlog") {                 // cvlog logfile v...: one LogCVData line of the values v
        U->SetToIdentity();
        MetaBiasPotential B(-8.0, 8.0, 800, 0.05, 0.02, 10.0, 1000.0);
        auto Ucopy = std::make_unique<GaugeField>();
        auto Mom = std::make_unique<GaugeField>();
        pcg64 g(1);
        PRNG4D<Nt, Nx, Ny, Nz, pcg64, floatT, int> prng(g);
        RunStatistics stats;
        std::ofstream devnull("/dev/null");
        std::string logpath = argv[2];
        GaugeUpdates::HMCMetaDData data(n_smear_meta);
        SetPathSamples integ;
        for (int i = 3; i < argc; ++i) integ.values.push_back(std::strtod(argv[i], nullptr));
        auto act = GaugeAction::WilsonAction;
        GaugeUpdates::HMCMetaDKernel HMC(*U, *Ucopy, *Mom, B, data, integ, act, prng, stats, devnull, logpath, 1.0, 0.12);
        HMC(1, false);
    } else {
        std::fprintf(stderr, "unknown command %s\n", cmd.c_str());
        return 2;
    }
    return 0;
}
