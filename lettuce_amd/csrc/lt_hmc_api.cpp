// lt_hmc_api.cpp -- C ABI of the HMC chain: momenta, integrators (leapfrog, OMF2,
// OMF4), the Hamiltonian and Metropolis step, the collective-variable forward /
// adjoint pair, the metadynamics bias potential and the MetaD-HMC trajectory.
// All lattice work runs in HIP kernels; only scalars (H, Q, dV/dQ) touch the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <functional>
#include <string>
#include <vector>

#include "lt_internal.hpp"

namespace lt {
void k_random_momentum(lt_field P, uint64_t seed, uint64_t counter);
void k_force_update(lt_field U, lt_field P, const double2* F, double a, double b);
void k_field_update(lt_field U, lt_field P, double eps);
double k_plaquette_sum(lt_field U);
double k_kinetic(lt_field P);
void k_stout_tape(lt_ctx ctx, const Geom& g, const LinkMap& m, const double2* in, double2* out, double2* tape,
                  double rho);
double k_clover_charge(lt_ctx ctx, const Geom& g, const double2* U, double2* C);
// next: the force is the Sigma of the top stout level of a recursion; go on with its
// Lambda (the level below, its tape, the Lambda buffer) instead of only storing it
void k_clover_deriv(lt_ctx ctx, const Geom& g, const LinkMap& m, const double2* U, const double2* C, double2* F,
                    double scale_factor, const SigmaNext* next = nullptr);

}  // namespace lt

// Collective-variable workspace: stout levels with their exponent tapes, the six
// clover planes of the most smeared field and the Lambda buffer of the recursion.
//
// Full storage keeps all n+1 levels and n tapes.  With a checkpoint interval k
// only the levels k, 2k, ... (< n) and the top level n are kept; the backward pass
// recomputes each segment of at most k levels (and its tapes) from its checkpoint
// into k+1 segment buffers, costing n extra stout passes per backward pass.  In
// that mode level 0 is the input field itself, which must not change between
// lt_cv_forward and lt_cv_backward.
struct lt_cv_s {
    lt_ctx ctx = nullptr;
    lt::Geom g{};
    lt::LinkMap map{};
    int n_levels = 0;
    double rho = 0.0;
    int interval = 0;               // 0: full storage, else checkpoint interval k
    std::vector<double2*> levels;   // full: n+1 levels (levels[0] = copy of the input); else the checkpoints
    std::vector<double2*> tapes;    // full: n exponent buffers; else k segment tapes
    std::vector<double2*> seg;      // checkpointed: k+1 segment level buffers
    double2* top = nullptr;         // checkpointed: level n
    const double2* input = nullptr; // checkpointed: level 0 (the field of the last forward pass)
    double2* clover = nullptr;      // 6 planes per site
    double2* lambda[2] = {nullptr, nullptr};   // Lambda of alternate levels
    size_t bytes = 0;               // device memory held
    bool forward_done = false;
    std::vector<double> path;       // CVs visited by the last MetaD-HMC trajectory
    std::vector<double> path_used;  // the values it deposited (CV_old repeated if rejected)

    void release()
    {
        for (auto p : levels) (void)hipFree(p);
        for (auto p : tapes) (void)hipFree(p);
        for (auto p : seg) (void)hipFree(p);
        for (void* p : {(void*)top, (void*)clover, (void*)lambda[0], (void*)lambda[1]})
            if (p) (void)hipFree(p);
        levels.clear();
        tapes.clear();
        seg.clear();
    }
};

namespace {

using lt::fail;
using lt::guarded;
using lt::valid_integrator;

bool same_shape(const lt::Geom& a, const lt::Geom& b)
{
    for (int i = 0; i < 4; ++i)
        if (a.L[i] != b.L[i]) return false;
    return true;
}

double2* alloc_buf(size_t n_double2)
{
    double2* p = nullptr;
    if (hipMalloc(&p, n_double2 * sizeof(double2)) != hipSuccess)
        throw lt::LtError{LT_ENOMEM, "device allocation failed"};
    return p;
}

// Philox draw for scalar accept/reject decisions (stream 3), identical on host and device
double host_uniform(uint64_t seed, uint64_t counter)
{
    uint32_t c[4] = {0xFFFFFFFFu, 3u << 24, (uint32_t)counter, (uint32_t)(counter >> 32)};
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = (uint32_t)p1;
        c[2] = n2;
        c[3] = (uint32_t)p0;
    }
    return ((double)(c[0] >> 5) * 67108864.0 + (double)(c[1] >> 6)) * (1.0 / 9007199254740992.0);
}

}  // namespace

// Integrators::HMC::{Leapfrog, OMF_2, OMF_4} with merged momentum updates and
// {Leapfrog_slow, OMF_2_slow, OMF_4_slow} without (LettuceGauge/updates/hmc_gauge.hpp:21-75,
// 121-295), coefficients formed as there; momentum updates carry the "submeasurement"
// flag the MetaD path update uses.
void lt::hmc_schedule(int kind, double tau, int n_step, const std::function<void(double, bool)>& P,
                      const std::function<void(double)>& U)
{
    if (n_step <= 0) return;
    const double eps = tau / n_step;
    if (kind == LT_INT_LEAPFROG || kind == LT_INT_LEAPFROG_SLOW) {
        if (kind == LT_INT_LEAPFROG_SLOW) {
            for (int k = 0; k < n_step; ++k) {
                P(0.5 * eps, false);
                U(eps);
                P(0.5 * eps, true);
            }
            return;
        }
        P(0.5 * eps, false);
        for (int k = 0; k < n_step - 1; ++k) {
            U(eps);
            P(eps, true);
        }
        U(eps);
        P(0.5 * eps, true);
    } else if (kind == LT_INT_OMF2 || kind == LT_INT_OMF2_SLOW) {
        const double lam = 0.1931833275037836;
        const double alpha = lam * eps, beta = 0.5 * eps, gamma = (1.0 - 2.0 * lam) * eps;
        if (kind == LT_INT_OMF2_SLOW) {
            for (int k = 0; k < n_step; ++k) {
                P(alpha, false);
                U(beta);
                P(gamma, false);
                U(beta);
                P(alpha, true);
            }
            return;
        }
        P(alpha, false);
        U(beta);
        P(gamma, false);
        U(beta);
        for (int k = 0; k < n_step - 1; ++k) {
            P(2.0 * alpha, true);
            U(beta);
            P(gamma, false);
            U(beta);
        }
        P(alpha, true);
    } else {
        const double alpha = 0.08398315262876693 * eps, beta = 0.2539785108410595 * eps;
        const double gamma = 0.6822365335719091 * eps, delta = -0.03230286765269967 * eps;
        const double mu = (0.5 - 0.6822365335719091 - 0.08398315262876693) * eps;
        const double nu = (1.0 - 2.0 * -0.03230286765269967 - 2.0 * 0.2539785108410595) * eps;
        auto body = [&]() {
            U(beta);
            P(gamma, false);
            U(delta);
            P(mu, false);
            U(nu);
            P(mu, false);
            U(delta);
            P(gamma, false);
            U(beta);
        };
        if (kind == LT_INT_OMF4_SLOW) {
            for (int k = 0; k < n_step; ++k) {
                P(alpha, false);
                body();
                P(alpha, true);
            }
            return;
        }
        P(alpha, false);
        body();
        for (int k = 0; k < n_step - 1; ++k) {
            P(2.0 * alpha, true);
            body();
        }
        P(alpha, true);
    }
}

// Integrators::HMC::{Leapfrog_OMF_4, OMF_2_OMF_4_slow, OMF_2_OMF_4} (hmc_metadynamics.hpp:28-238)
// with the coefficients formed as there; the single-scale kinds kick both forces at once.
void lt::metad_schedule(int kind, double tau, int n_step, const std::function<void(double, bool, int)>& P,
                        const std::function<void(double)>& U)
{
    if (valid_integrator(kind)) {
        hmc_schedule(kind, tau, n_step, [&](double e, bool sub) { P(e, sub, KICK_BOTH); }, U);
        return;
    }
    if (n_step <= 0) return;
    const double eps = tau / n_step;
    // the inner OMF4 step; OMF_2_OMF_4 scales it by 1/2 (two inner steps per outer step)
    const double f = kind == LT_INT_OMF2_OMF4 ? 0.5 : 1.0;
    const double a4 = f * 0.08398315262876693 * eps, b4 = f * 0.2539785108410595 * eps;
    const double g4 = f * 0.6822365335719091 * eps, d4 = f * -0.03230286765269967 * eps;
    const double m4 = f * (0.5 - 0.6822365335719091 - 0.08398315262876693) * eps;
    const double n4 = f * (1.0 - 2.0 * -0.03230286765269967 - 2.0 * 0.2539785108410595) * eps;
    auto G = [&](double e) { P(e, false, KICK_GAUGE); };
    auto inner = [&]() {
        G(a4);
        U(b4);
        G(g4);
        U(d4);
        G(m4);
        U(n4);
        G(m4);
        U(d4);
        G(g4);
        U(b4);
        G(a4);
    };
    auto M = [&](double e, bool sub) { P(e, sub, KICK_BIAS); };
    if (kind == LT_INT_LEAPFROG_OMF4) {
        M(0.5 * eps, false);
        for (int k = 0; k < n_step - 1; ++k) {
            inner();
            M(eps, true);
        }
        inner();
        M(0.5 * eps, true);
        return;
    }
    const double a2 = 0.1931833275037836 * eps, b2 = (1.0 - 2.0 * 0.1931833275037836) * eps;
    if (kind == LT_INT_OMF2_OMF4_SLOW) {
        for (int k = 0; k < n_step; ++k) {
            M(a2, false);
            inner();
            M(b2, false);
            inner();
            M(a2, true);
        }
        return;
    }
    M(a2, false);
    for (int k = 0; k < n_step - 1; ++k) {
        inner();
        M(b2, false);
        inner();
        M(2.0 * a2, true);
    }
    inner();
    M(b2, false);
    inner();
    M(a2, true);
}

namespace {

void integrate(int kind, double tau, int n_step, const std::function<void(double, bool)>& P,
               const std::function<void(double)>& U)
{
    lt::hmc_schedule(kind, tau, n_step, P, U);
}

double hamiltonian(lt_field U, lt_field P, double beta)
{
    const double S = beta * (6.0 * (double)U->g.V - lt::k_plaquette_sum(U) / 3.0);
    return S + lt::k_kinetic(P);
}

void copy_field(lt_field dst, lt_field src)
{
    lt::k_copy(dst->ctx, src->d, dst->d, (size_t)72 * dst->g.Vh);
    dst->su3 = src->su3;
}

// CV forward: levels with tapes, clover planes, Q (FullTensor route, topological_charge.hpp:202-231)
// checkpointed mode: the buffer that holds level j after the forward pass, or null
const double2* ckpt_level(lt_cv cv, int j)
{
    if (j == 0) return cv->input;
    if (j == cv->n_levels) return cv->top;
    if (j % cv->interval == 0) return cv->levels[j / cv->interval - 1];
    return nullptr;
}

double cv_forward(lt_cv cv, lt_field U)
{
    const int n = cv->n_levels;
    if (cv->interval == 0) {
        lt::k_copy(cv->ctx, U->d, cv->levels[0], (size_t)72 * cv->g.Vh);
        for (int k = 0; k < n; ++k)
            lt::k_stout_tape(cv->ctx, cv->g, cv->map, cv->levels[k], cv->levels[k + 1], cv->tapes[k], cv->rho);
    } else {
        // keep checkpoints and the top level; other levels ping-pong in the segment buffers
        cv->input = U->d;
        const double2* prev = U->d;
        for (int j = 1; j <= n; ++j) {
            double2* out = const_cast<double2*>(ckpt_level(cv, j));
            if (!out) out = cv->seg[j & 1];
            lt::k_stout_tape(cv->ctx, cv->g, cv->map, prev, out, cv->tapes[0], cv->rho);
            prev = out;
        }
    }
    const double2* last = cv->interval == 0 ? cv->levels[n] : (n ? cv->top : U->d);
    double Q = lt::k_clover_charge(cv->ctx, cv->g, last, cv->clover);
    cv->forward_done = true;
    return Q;
}

// CV adjoint into F: dvdq * dQ/dU through the clover derivative and n recursions.
// With kick (n > 0 only) the last recursion step applies the momentum update to
// kick->P instead of storing the level-0 force, and F is left holding level 1's.
void cv_backward(lt_cv cv, double dvdq, double2* F, const lt::MomentumKick* kick = nullptr)
{
    const int n = cv->n_levels;
    if (cv->interval == 0) {
        if (n == 0) {
            lt::k_clover_deriv(cv->ctx, cv->g, cv->map, cv->levels[0], cv->clover, F, dvdq);
            return;
        }
        // the clover derivative goes on with the Lambda of the top level
        const lt::SigmaNext top{cv->levels[n - 1], cv->tapes[n - 1], cv->lambda[0]};
        lt::k_clover_deriv(cv->ctx, cv->g, cv->map, cv->levels[n], cv->clover, F, dvdq, &top);
        lt::k_stout_recursion_run(
            cv->ctx, cv->g, cv->map, [&](int j) -> const double2* { return cv->levels[j]; },
            [&](int j) -> const double2* { return cv->tapes[j]; }, n, 0, F, cv->lambda, cv->rho, kick, true);
        return;
    }
    lt::k_clover_deriv(cv->ctx, cv->g, cv->map, n ? cv->top : cv->input, cv->clover, F, dvdq);
    const int k = cv->interval;
    for (int a = ((n - 1) / k) * k; a >= 0 && n > 0; a -= k) {
        const int b = std::min(a + k, n);
        // recompute levels a+1..b and tapes a..b-1 of this segment from checkpoint a
        auto level = [&](int j) -> const double2* { return j == a ? ckpt_level(cv, a) : cv->seg[j - a]; };
        for (int j = a + 1; j <= b; ++j)
            lt::k_stout_tape(cv->ctx, cv->g, cv->map, level(j - 1), cv->seg[j - a], cv->tapes[j - a - 1], cv->rho);
        lt::k_stout_recursion_run(
            cv->ctx, cv->g, cv->map, level, [&](int j) -> const double2* { return cv->tapes[j - a]; }, b, a, F,
            cv->lambda, cv->rho, a == 0 ? kick : nullptr);
    }
}

// MetaCharge (hmc_metadynamics.hpp:302-310): n stout levels, clover Q
double cv_charge(lt_cv cv, lt_field U)
{
    return cv_forward(cv, U);
}

// The molecular dynamics of HMCMetaDKernel (Integrator(*this, ...), hmc_metadynamics.hpp:374-524)
// on the momenta in P; F is force scratch; the CVs of the sampling kicks are appended to samples.
void metad_integrate(lt_field U, lt_field P, lt_field F, lt_cv cv, lt_bias bias, double beta, int integrator,
                     double tau, int n_step, std::vector<double>& samples)
{
    lt::metad_schedule(
        integrator, tau, n_step,
        [&](double e, bool sub, int kick) {
            if (kick & lt::KICK_BIAS) {
                // CalculateTopologicalForce (hmc_metadynamics.hpp:374-436)
                const double q = cv_forward(cv, U);
                if (sub) samples.push_back(q);
                if (kick == lt::KICK_BOTH && cv->n_levels > 0) {
                    // UpdateMomenta applied by the last kernel of the force (on level 0 = U)
                    const lt::MomentumKick mk{P->d, e * beta / 6.0, e};
                    cv_backward(cv, bias->derivative(q), F->d, &mk);
                    return;
                }
                cv_backward(cv, bias->derivative(q), F->d);
            }
            if (kick == lt::KICK_BOTH)   // UpdateMomenta (:441-459)
                lt::k_force_update(U, P, F->d, e * beta / 6.0, e);
            else if (kick == lt::KICK_GAUGE)   // UpdateGaugeMomenta (:464-477)
                lt::k_force_update(U, P, nullptr, e * beta / 6.0, 0.0);
            else   // UpdateMetadynamicsMomenta (:479-491)
                lt::k_axpy(U->ctx, F->d, P->d, e, (size_t)72 * P->g.Vh);
        },
        [&](double e) { lt::k_field_update(U, P, e); });
}

}  // namespace

extern "C" {

int lt_matrix_field_create(lt_ctx ctx, int nt, int nx, int ny, int nz, lt_field* out)
{
    int rc = lt_field_create(ctx, nt, nx, ny, nz, out);
    if (rc == LT_OK) (*out)->su3 = false;
    return rc;
}

int lt_hmc_random_momentum(lt_field P, uint64_t seed, uint64_t counter)
{
    if (!P) return fail(LT_EINVAL, "null field");
    return guarded([&] {
        lt::k_random_momentum(P, seed, counter);
        P->su3 = false;
        return LT_OK;
    });
}

int lt_hmc_update_momenta(lt_field U, lt_field P, double eps, double beta)
{
    if (!U || !P || !same_shape(U->g, P->g)) return fail(LT_EINVAL, "lt_hmc_update_momenta: bad fields");
    return guarded([&] {
        lt::k_force_update(U, P, nullptr, eps * beta / 6.0, 0.0);
        return LT_OK;
    });
}

int lt_hmc_update_fields(lt_field U, lt_field P, double eps)
{
    if (!U || !P || !same_shape(U->g, P->g)) return fail(LT_EINVAL, "lt_hmc_update_fields: bad fields");
    return guarded([&] {
        lt::k_field_update(U, P, eps);
        return LT_OK;
    });
}

int lt_hmc_hamiltonian(lt_field U, lt_field P, double beta, double* H)
{
    if (!U || !P || !H || !same_shape(U->g, P->g)) return fail(LT_EINVAL, "lt_hmc_hamiltonian: bad arguments");
    return guarded([&] {
        *H = hamiltonian(U, P, beta);
        return LT_OK;
    });
}

int lt_hmc_integrate(lt_field U, lt_field P, double beta, int integrator, double tau, int n_step)
{
    if (!U || !P || !same_shape(U->g, P->g) || !valid_integrator(integrator) || n_step < 0)
        return fail(LT_EINVAL, "lt_hmc_integrate: bad arguments");
    return guarded([&] {
        integrate(
            integrator, tau, n_step, [&](double e, bool) { lt::k_force_update(U, P, nullptr, e * beta / 6.0, 0.0); },
            [&](double e) { lt::k_field_update(U, P, e); });
        return LT_OK;
    });
}

int lt_hmc_trajectory(lt_field U, lt_field P, lt_field U_copy, double beta, int integrator, double tau, int n_step,
                      int metropolis, double* dH, int* accepted)
{
    if (!U || !P || !U_copy || !same_shape(U->g, P->g) || !same_shape(U->g, U_copy->g) || U == U_copy ||
        !valid_integrator(integrator) || n_step < 0)
        return fail(LT_EINVAL, "lt_hmc_trajectory: bad arguments");
    return guarded([&] {
        // HMCKernel::operator() (hmc_gauge.hpp:549-583)
        const uint64_t counter = U->counter++;
        copy_field(U_copy, U);
        lt::k_random_momentum(P, U->seed, counter);
        P->su3 = false;
        const double H0 = hamiltonian(U, P, beta);
        integrate(
            integrator, tau, n_step, [&](double e, bool) { lt::k_force_update(U, P, nullptr, e * beta / 6.0, 0.0); },
            [&](double e) { lt::k_field_update(U, P, e); });
        const double H1 = hamiltonian(U, P, beta);
        const double d = H1 - H0;
        int acc = 1;
        if (metropolis) {
            const double q = host_uniform(U->seed, counter);
            acc = q <= std::exp(-d);
            if (!acc) copy_field(U, U_copy);
        }
        if (dH) *dH = d;
        if (accepted) *accepted = acc;
        return LT_OK;
    });
}

int lt_hmc_reversibility_test(lt_field U, lt_field P, double beta, int integrator, double tau, int n_step,
                              double* H_old, double* H_new)
{
    if (!U || !P || U == P || !same_shape(U->g, P->g) || !valid_integrator(integrator) || n_step < 0)
        return fail(LT_EINVAL, "lt_hmc_reversibility_test: bad arguments");
    return guarded([&] {
        // HMCKernel::ReversibilityTest (hmc_gauge.hpp:585-599)
        lt::k_random_momentum(P, U->seed, U->counter++);
        P->su3 = false;
        const double H0 = hamiltonian(U, P, beta);
        for (int pass = 0; pass < 2; ++pass) {
            if (pass) lt::k_scale(U->ctx, P->d, -1.0, (size_t)72 * P->g.Vh);   // ReverseMomenta
            integrate(
                integrator, tau, n_step,
                [&](double e, bool) { lt::k_force_update(U, P, nullptr, e * beta / 6.0, 0.0); },
                [&](double e) { lt::k_field_update(U, P, e); });
        }
        const double H1 = hamiltonian(U, P, beta);
        if (H_old) *H_old = H0;
        if (H_new) *H_new = H1;
        return LT_OK;
    });
}

int lt_cv_create_checkpointed(lt_ctx ctx, int nt, int nx, int ny, int nz, int n_levels, double rho, int interval,
                              lt_cv* out)
{
    if (!ctx || !out || n_levels < 0 || n_levels > 64 || interval < 0)
        return fail(LT_EINVAL, "lt_cv_create: bad arguments");
    if (nt < 2 || nx < 2 || ny < 2 || nz < 2 || (nt | nx | ny | nz) & 1)
        return fail(LT_EINVAL, "lt_cv_create: extents must be even and >= 2");
    return guarded([&] {
        LT_HIP_CHECK(hipSetDevice(ctx->device));
        auto* cv = new lt_cv_s();
        cv->ctx = ctx;
        cv->g = lt::make_geom(nt, nx, ny, nz);
        cv->map = lt::default_linkmap(cv->g);
        cv->n_levels = n_levels;
        cv->rho = rho;
        cv->interval = interval > 0 ? std::min(interval, std::max(n_levels, 1)) : 0;
        const size_t nl = (size_t)72 * cv->g.Vh, na = (size_t)32 * cv->g.Vh;
        auto take = [&](size_t n) {
            double2* p = alloc_buf(n);
            cv->bytes += n * sizeof(double2);
            return p;
        };
        try {
            if (cv->interval == 0) {
                for (int k = 0; k <= n_levels; ++k) cv->levels.push_back(take(nl));
                for (int k = 0; k < n_levels; ++k) cv->tapes.push_back(take(na));
            } else {
                const int k = cv->interval;
                for (int j = k; j < n_levels; j += k) cv->levels.push_back(take(nl));
                if (n_levels > 0) cv->top = take(nl);
                for (int j = 0; j <= k; ++j) cv->seg.push_back(take(nl));
                for (int j = 0; j < k; ++j) cv->tapes.push_back(take(na));
            }
            cv->clover = take((size_t)12 * lt::R_SLOTS * cv->g.Vh);
            cv->lambda[0] = take(na);
            cv->lambda[1] = take(na);
        } catch (...) {
            cv->release();
            delete cv;
            throw;
        }
        ctx->live += 1;
        *out = cv;
        return LT_OK;
    });
}

int lt_clover_derivative(lt_field U, lt_field F, double scale)
{
    if (!U || !F || U == F || !U->su3 || !same_shape(U->g, F->g))
        return fail(LT_EINVAL, "lt_clover_derivative: bad arguments");
    return guarded([&] {
        LT_HIP_CHECK(hipSetDevice(U->ctx->device));
        double2* clover = alloc_buf((size_t)12 * lt::R_SLOTS * U->g.Vh);
        try {
            (void)lt::k_clover_charge(U->ctx, U->g, U->d, clover);
            lt::k_clover_deriv(U->ctx, U->g, U->map, U->d, clover, F->d, scale);
            LT_HIP_CHECK(hipStreamSynchronize(U->ctx->stream));
        } catch (...) {
            (void)hipFree(clover);
            throw;
        }
        LT_HIP_CHECK(hipFree(clover));
        F->su3 = false;
        return LT_OK;
    });
}

int lt_cv_create(lt_ctx ctx, int nt, int nx, int ny, int nz, int n_levels, double rho, lt_cv* out)
{
    return lt_cv_create_checkpointed(ctx, nt, nx, ny, nz, n_levels, rho, 0, out);
}

int lt_cv_bytes(lt_cv cv, size_t* bytes)
{
    if (!cv || !bytes) return fail(LT_EINVAL, "null argument");
    *bytes = cv->bytes;
    return LT_OK;
}

int lt_cv_destroy(lt_cv cv)
{
    if (!cv) return LT_OK;
    return guarded([&] {
        LT_HIP_CHECK(hipStreamSynchronize(cv->ctx->stream));
        cv->release();
        cv->ctx->live -= 1;
        delete cv;
        return LT_OK;
    });
}

int lt_cv_forward(lt_cv cv, lt_field U, double* Q)
{
    if (!cv || !U || !Q || !same_shape(cv->g, U->g) || cv->ctx != U->ctx)
        return fail(LT_EINVAL, "lt_cv_forward: bad arguments");
    return guarded([&] {
        *Q = cv_forward(cv, U);
        return LT_OK;
    });
}

// MetadynamicsTemperingKernel::operator() (tempering.hpp:57-72): the unbiased stream's
// smeared clover charge against the bias's current CV (the biased stream's), a Philox
// uniform from stream 4 at site 0 (the draw of label 0 in a replica ladder), and a
// swap of the two fields' links on acceptance.
int lt_metad_tempering_swap(lt_field U, lt_field U_temper, lt_cv cv, lt_bias bias, uint64_t seed, uint64_t counter,
                            int* accepted, double* delta_v)
{
    if (!U || !U_temper || !cv || !bias) return fail(LT_EINVAL, "lt_metad_tempering_swap: null argument");
    if (!same_shape(U->g, U_temper->g) || !same_shape(cv->g, U->g) || U->ctx != U_temper->ctx || cv->ctx != U->ctx)
        return fail(LT_EINVAL, "lt_metad_tempering_swap: shape/context mismatch");
    if (!U->su3 || !U_temper->su3) return fail(LT_ESTATE, "lt_metad_tempering_swap: SU(3) fields required");
    if (U->view || U_temper->view) return fail(LT_ESTATE, "lt_metad_tempering_swap: replicas of a batch cannot be swapped");
    return guarded([&] {
        const double cv_old = bias->cv_current;
        const double cv_new = cv_forward(cv, U);
        const double dv = bias->potential(cv_new) - bias->potential(cv_old);
        double u0 = 0.0, u1 = 0.0;
        lt::draw2(seed, 0u, 4u, 0u, 0u, 0u, counter, u0, u1);
        const bool acc = u0 <= std::exp(-dv);
        if (acc) {
            std::swap(U->d, U_temper->d);
            bias->cv_current = cv_new;
        }
        if (accepted) *accepted = acc ? 1 : 0;
        if (delta_v) *delta_v = dv;
        return LT_OK;
    });
}

int lt_cv_backward(lt_cv cv, double dvdq, lt_field F)
{
    if (!cv || !F || !same_shape(cv->g, F->g) || cv->ctx != F->ctx) return fail(LT_EINVAL, "lt_cv_backward: bad arguments");
    if (!cv->forward_done) return fail(LT_ESTATE, "lt_cv_backward: call lt_cv_forward first");
    return guarded([&] {
        cv_backward(cv, dvdq, F->d);
        F->su3 = false;
        return LT_OK;
    });
}

int lt_bias_create(double cv_min, double cv_max, int bins, double weight, double width, int well_tempered,
                   double well_tempered_parameter, double threshold_weight, lt_bias* out)
{
    if (!out) return fail(LT_EINVAL, "lt_bias_create: null out");
    // MetaBiasPotential constructor checks (metadynamics.hpp:70-77)
    if (!(cv_min < cv_max)) return fail(LT_EINVAL, "CV_min < CV_max required");
    if (bins <= 0) return fail(LT_EINVAL, "bin_number > 0 required");
    auto* b = new lt_bias_s();
    b->cv_min = cv_min;
    b->cv_max = cv_max;
    b->bins = bins;
    b->bin_width = (cv_max - cv_min) / bins;
    b->bin_width_inverse = 1.0 / b->bin_width;
    b->weight = weight;
    b->width = width;
    b->well_tempered = well_tempered;
    b->wt_param = well_tempered_parameter;
    b->threshold = threshold_weight;
    b->grid.assign((size_t)bins + 1, 0.0);
    *out = b;
    return LT_OK;
}

int lt_bias_destroy(lt_bias b)
{
    delete b;
    return LT_OK;
}

int lt_bias_deposit(lt_bias b, double cv, int symmetric)
{
    if (!b) return fail(LT_EINVAL, "null bias");
    b->deposit(cv);
    if (symmetric) b->deposit(-cv);
    return LT_OK;
}

int lt_bias_potential(lt_bias b, double cv, double* v)
{
    if (!b || !v) return fail(LT_EINVAL, "null argument");
    *v = b->potential(cv);
    return LT_OK;
}

int lt_bias_derivative(lt_bias b, double cv, double* d)
{
    if (!b || !d) return fail(LT_EINVAL, "null argument");
    *d = b->derivative(cv);
    return LT_OK;
}

int lt_bias_grid(lt_bias b, double* values, int n, int set)
{
    if (!b || !values || n != (int)b->grid.size()) return fail(LT_EINVAL, "lt_bias_grid: size must be bins + 1");
    if (set)
        std::memcpy(b->grid.data(), values, sizeof(double) * n);
    else
        std::memcpy(values, b->grid.data(), sizeof(double) * n);
    return LT_OK;
}

int lt_bias_cv(lt_bias b, double* cv, int set)
{
    if (!b || !cv) return fail(LT_EINVAL, "null argument");
    if (set)
        b->cv_current = *cv;
    else
        *cv = b->cv_current;
    return LT_OK;
}

// HMCMetaDKernel::operator() (hmc_metadynamics.hpp:552-655) with path updates:
// each "submeasurement" momentum update records the CV it used; after the
// Metropolis step the samples (or CV_old repeated, if rejected) are deposited
// symmetrically when update_bias is set.
int lt_metad_hmc_trajectory(lt_field U, lt_field P, lt_field U_copy, lt_field F, lt_cv cv, lt_bias bias, double beta,
                            int integrator, double tau, int n_step, int metropolis, int update_bias, double* dH,
                            int* accepted, double* cv_new)
{
    if (!U || !P || !U_copy || !F || !cv || !bias || !lt::valid_metad_integrator(integrator) || n_step < 0 ||
        U == U_copy || !same_shape(U->g, P->g) || !same_shape(U->g, U_copy->g) || !same_shape(U->g, F->g) ||
        !same_shape(U->g, cv->g))
        return fail(LT_EINVAL, "lt_metad_hmc_trajectory: bad arguments");
    return guarded([&] {
        const uint64_t counter = U->counter++;
        copy_field(U_copy, U);
        lt::k_random_momentum(P, U->seed, counter);
        P->su3 = false;
        F->su3 = false;
        const double cv_old = cv_charge(cv, U);
        const double H0 = hamiltonian(U, P, beta) + bias->potential(cv_old);
        std::vector<double> samples;
        samples.reserve(n_step);
        metad_integrate(U, P, F, cv, bias, beta, integrator, tau, n_step, samples);
        cv->path = samples;
        const double cv1 = cv_charge(cv, U);
        const double H1 = hamiltonian(U, P, beta) + bias->potential(cv1);
        const double d = H1 - H0;
        int acc = 1;
        if (metropolis) {
            const double q = host_uniform(U->seed, counter);
            acc = q <= std::exp(-d);
        }
        if (acc) {
            bias->cv_current = cv1;
        } else {
            bias->cv_current = cv_old;
            std::fill(samples.begin(), samples.end(), cv_old);
            copy_field(U, U_copy);
        }
        cv->path_used = samples;
        if (update_bias)
            for (double s : samples) {
                bias->deposit(s);
                bias->deposit(-s);
            }
        if (dH) *dH = d;
        if (accepted) *accepted = acc;
        if (cv_new) *cv_new = acc ? cv1 : cv_old;
        return LT_OK;
    });
}

int lt_integrator_schedule(int integrator, double tau, int n_step, char* kinds, double* eps, int* sub, int cap,
                           int* n)
{
    if (!lt::valid_metad_integrator(integrator) || n_step < 0 || !n || cap < 0 ||
        (cap > 0 && (!kinds || !eps || !sub)))
        return fail(LT_EINVAL, "lt_integrator_schedule: bad arguments");
    int i = 0;
    auto put = [&](char k, double e, bool s) {
        if (i < cap) {
            kinds[i] = k;
            eps[i] = e;
            sub[i] = s ? 1 : 0;
        }
        ++i;
    };
    lt::metad_schedule(
        integrator, tau, n_step,
        [&](double e, bool s, int kick) { put(kick == lt::KICK_BOTH ? 'P' : kick == lt::KICK_GAUGE ? 'G' : 'M', e, s); },
        [&](double e) { put('U', e, false); });
    *n = i;
    return LT_OK;
}

int lt_metad_reversibility_test(lt_field U, lt_field P, lt_field F, lt_cv cv, lt_bias bias, double beta,
                                int integrator, double tau, int n_step, double* H_old, double* H_new)
{
    if (!U || !P || !F || !cv || !bias || !lt::valid_metad_integrator(integrator) || n_step < 0 || U == P || U == F ||
        P == F || !same_shape(U->g, P->g) || !same_shape(U->g, F->g) || !same_shape(U->g, cv->g))
        return fail(LT_EINVAL, "lt_metad_reversibility_test: bad arguments");
    return guarded([&] {
        // HMCMetaDKernel::ReversibilityTest (hmc_metadynamics.hpp:657-674)
        lt::k_random_momentum(P, U->seed, U->counter++);
        P->su3 = false;
        F->su3 = false;
        const double H0 = hamiltonian(U, P, beta) + bias->potential(cv_charge(cv, U));
        std::vector<double> samples;
        metad_integrate(U, P, F, cv, bias, beta, integrator, tau, n_step, samples);
        lt::k_scale(U->ctx, P->d, -1.0, (size_t)72 * P->g.Vh);   // ReverseMomenta
        metad_integrate(U, P, F, cv, bias, beta, integrator, tau, n_step, samples);
        const double H1 = hamiltonian(U, P, beta) + bias->potential(cv_charge(cv, U));
        if (H_old) *H_old = H0;
        if (H_new) *H_new = H1;
        return LT_OK;
    });
}

int lt_cv_path_samples(lt_cv cv, int deposited, double* out, int cap, int* n)
{
    if (!cv || !n || cap < 0 || (cap > 0 && !out)) return fail(LT_EINVAL, "lt_cv_path_samples: bad arguments");
    const std::vector<double>& s = deposited ? cv->path_used : cv->path;
    *n = (int)s.size();
    std::copy_n(s.begin(), std::min<size_t>(s.size(), (size_t)cap), out);
    return LT_OK;
}

int lt_cv_log_path(lt_cv cv, int deposited, const char* path)
{
    if (!cv || !path) return fail(LT_EINVAL, "lt_cv_log_path: bad arguments");
    const std::vector<double>& s = deposited ? cv->path_used : cv->path;
    if (s.empty()) return fail(LT_ESTATE, "lt_cv_log_path: no path samples (no MetaD-HMC trajectory yet)");
    std::ofstream log(path, std::ios::out | std::ios::app);
    if (!log) return fail(LT_EIO, std::string("lt_cv_log_path: cannot open ") + path);
    for (size_t i = 0; i + 1 < s.size(); ++i) log << s[i] << ' ';
    log << s.back() << '\n';
    return log ? LT_OK : fail(LT_EIO, std::string("lt_cv_log_path: write failed: ") + path);
}

}  // extern "C"
