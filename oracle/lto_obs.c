/*
 * lto_obs.c -- CPU oracle (test infrastructure only): plaquette, Wilson action,
 * clover term, topological charge, energy densities, stout smearing and the
 * Wilson (gradient) flow integrators.
 */
#include "lto_internal.h"
#include <omp.h>
#include <stdlib.h>

/* Plaquette (plaquette.hpp:20-24): U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+ */
static inline mat3 plaq(const double *U, lto_dims d, size_t s, int mu, int nu)
{
    size_t smu = lto_move(d, s, mu, 1), snu = lto_move(d, s, nu, 1);
    return m_mul_nd(m_mul_nd(m_mul(lto_link(U, s, mu), lto_link(U, smu, nu)), lto_link(U, snu, mu)),
                    lto_link(U, s, nu));
}

/* PlaquetteSum (plaquette.hpp:26-43) */
double lto_plaquette_sum(const double *U, lto_dims d)
{
    size_t V = lto_volume(d);
    double sum = 0.0;
    #pragma omp parallel for reduction(+:sum)
    for (size_t s = 0; s < V; ++s)
        for (int nu = 1; nu < 4; ++nu)
            for (int mu = 0; mu < nu; ++mu) sum += creal(m_tr(plaq(U, d, s, mu, nu)));
    return sum;
}

/* WilsonAction::Action (wilson_action.hpp:23-41) / Rectangular<1>::Action
 * (rectangular_action.hpp:61-94): beta * (6V - 1/3 * sum Re tr P) */
double lto_wilson_action(const double *U, lto_dims d, double beta)
{
    return beta * (6.0 * (double)lto_volume(d) - (1.0 / 3.0) * lto_plaquette_sum(U, d));
}

/* EnergyDensity::Plaquette (field_strength_tensor.hpp:140-145) */
double lto_energy_plaquette(const double *U, lto_dims d)
{
    return 2.0 * (18.0 - lto_plaquette_sum(U, d) / (double)lto_volume(d));
}

/* One clover component C_munu(x): sum of the four leaves in the orientation of
 * CalculateClover<1,1> (clover.hpp:94-152) and PlaquetteI..IV (plaquette.hpp:82-114) */
static mat3 clover_leaf_sum(const double *U, lto_dims d, size_t s, int mu, int nu)
{
    size_t smu = lto_move(d, s, mu, 1), snu = lto_move(d, s, nu, 1);
    size_t smd = lto_move(d, s, mu, -1), snd = lto_move(d, s, nu, -1);
    size_t smd_nu = lto_move(d, smd, nu, 1), smd_nd = lto_move(d, smd, nu, -1);
    size_t smu_nd = lto_move(d, snd, mu, 1);
    mat3 l1 = m_mul_nd(m_mul_nd(m_mul(lto_link(U, s, mu), lto_link(U, smu, nu)), lto_link(U, snu, mu)), lto_link(U, s, nu));
    mat3 l2 = m_mul(m_mul_nd(m_mul_nd(lto_link(U, s, nu), lto_link(U, smd_nu, mu)), lto_link(U, smd, nu)), lto_link(U, smd, mu));
    mat3 l3 = m_mul(m_mul(m_mul(m_dag(lto_link(U, smd, mu)), m_dag(lto_link(U, smd_nd, nu))), lto_link(U, smd_nd, mu)), lto_link(U, snd, nu));
    mat3 l4 = m_mul_nd(m_mul(m_mul_dn(lto_link(U, snd, nu), lto_link(U, snd, mu)), lto_link(U, smu_nd, nu)), lto_link(U, s, mu));
    return m_add(m_add(l1, l2), m_add(l3, l4));
}

static const int PL_MU[6] = {0, 0, 0, 1, 1, 2}, PL_NU[6] = {1, 2, 3, 2, 3, 3};

/* CalculateClover<1,1> (clover.hpp:94-152); only mu<nu stored (C_numu = C_munu^+) */
void lto_clover(const double *U, lto_dims d, double *C6)
{
    size_t V = lto_volume(d);
    #pragma omp parallel for
    for (size_t s = 0; s < V; ++s)
        for (int p = 0; p < 6; ++p)
            m_store(C6 + 18 * (6 * s + p), clover_leaf_sum(U, d, s, PL_MU[p], PL_NU[p]));
}

/* sum over sites of Re tr(F01 F23 - F02 F13 + F03 F12) with F = C - C^+ */
static double q_density_raw(const mat3 C[6])
{
    mat3 F[6];
    for (int p = 0; p < 6; ++p) F[p] = m_sub(C[p], m_dag(C[p]));
    return m_retr_mul(F[0], F[5]) - m_retr_mul(F[1], F[4]) + m_retr_mul(F[2], F[3]);
}

/* TopChargeClover(const GaugeField&) (topological_charge.hpp:97-200) */
double lto_topcharge_clover(const double *U, lto_dims d)
{
    size_t V = lto_volume(d);
    double Q = 0.0;
    #pragma omp parallel for reduction(+:Q)
    for (size_t s = 0; s < V; ++s) {
        mat3 C[6];
        for (int p = 0; p < 6; ++p) C[p] = clover_leaf_sum(U, d, s, PL_MU[p], PL_NU[p]);
        Q += q_density_raw(C);
    }
    return -1.0 / (64.0 * 4.0 * LTO_PI * LTO_PI) * Q;
}

/* The per-site summand of TopChargeClover (topological_charge.hpp:184-199): the
 * terms that lto_topcharge_clover adds up, one per site, with its normalisation. */
void lto_topcharge_density(const double *U, lto_dims d, double *q_x)
{
    size_t V = lto_volume(d);
    #pragma omp parallel for
    for (size_t s = 0; s < V; ++s) {
        mat3 C[6];
        for (int p = 0; p < 6; ++p) C[p] = clover_leaf_sum(U, d, s, PL_MU[p], PL_NU[p]);
        q_x[s] = -1.0 / (64.0 * 4.0 * LTO_PI * LTO_PI) * q_density_raw(C);
    }
}

/* TopChargeClover(const FullTensor&) (topological_charge.hpp:202-231) */
double lto_topcharge_clover_tensor(const double *C6, lto_dims d)
{
    size_t V = lto_volume(d);
    double Q = 0.0;
    #pragma omp parallel for reduction(+:Q)
    for (size_t s = 0; s < V; ++s) {
        mat3 C[6];
        for (int p = 0; p < 6; ++p) C[p] = m_load(C6 + 18 * (6 * s + p));
        Q += q_density_raw(C);
    }
    return -1.0 / (64.0 * 4.0 * LTO_PI * LTO_PI) * Q;
}

/* FieldStrengthTensor::Clover (field_strength_tensor.hpp:28-66) gives
 * F_munu = -i/4 Antihermitian(C) (hermitian), F_numu = -F_munu^+; then
 * TopChargeCloverTimeslice(F, t) (topological_charge.hpp:337-364) uses
 * F_munu - F_numu = 2 F_munu and the prefactor 1/(16 pi^2). */
void lto_topcharge_timeslice(const double *U, lto_dims d, double *q_t)
{
    size_t Vs = (size_t)d.nx * d.ny * d.nz;
    for (int t = 0; t < d.nt; ++t) {
        double Q = 0.0;
        #pragma omp parallel for reduction(+:Q)
        for (size_t r = 0; r < Vs; ++r) {
            size_t s = (size_t)t * Vs + r;
            mat3 F[6];
            for (int p = 0; p < 6; ++p) {
                mat3 h = m_scale(-I / 4.0, m_antiherm(clover_leaf_sum(U, d, s, PL_MU[p], PL_NU[p])));
                mat3 hn = m_rscale(-1.0, m_dag(h));
                F[p] = m_sub(h, hn);
            }
            Q += m_retr_mul(F[0], F[5]) - m_retr_mul(F[1], F[4]) + m_retr_mul(F[2], F[3]);
        }
        q_t[t] = 1.0 / (16.0 * LTO_PI * LTO_PI) * Q;
    }
}

/* MakeComponentsTraceless (field_strength_tensor.hpp:107-134) applied to
 * FieldStrengthTensor::Clover, then EnergyDensity::CloverTimeslice (:177-195):
 * E_t = (1/V) sum_{x in slice t} Re tr sum_{mu<nu} F_munu F_munu */
void lto_energy_clover_timeslice(const double *U, lto_dims d, double *e_t)
{
    size_t Vs = (size_t)d.nx * d.ny * d.nz;
    double V = (double)lto_volume(d);
    for (int t = 0; t < d.nt; ++t) {
        double E = 0.0;
        #pragma omp parallel for reduction(+:E)
        for (size_t r = 0; r < Vs; ++r) {
            size_t s = (size_t)t * Vs + r;
            mat3 acc = m_zero();
            for (int p = 0; p < 6; ++p) {
                mat3 f = m_traceless(m_scale(-I / 4.0, m_antiherm(clover_leaf_sum(U, d, s, PL_MU[p], PL_NU[p]))));
                acc = m_add(acc, m_mul(f, f));
            }
            E += creal(m_tr(acc));
        }
        e_t[t] = E / V;
    }
}

/* ------------------------------------------------------------------------- */
/* StoutSmearing4D (stout_smearing.hpp:29-57):
 *   A = st U^+, C = 1/2 (A - A^+) - 1/6 tr(A - A^+), U' = exp(-i rho C) U, GramSchmidt */
static inline mat3 stout_Q(const double *Uin, lto_dims d, size_t s, int mu, double rho)
{
    mat3 A = m_mul_nd(lto_staple_m(Uin, d, s, mu), lto_link(Uin, s, mu));
    mat3 B = m_sub(A, m_dag(A));
    cplx tr = m_tr(B);
    mat3 C = m_rscale(0.5, B);
    for (int i = 0; i < 3; ++i) C.m[4*i] -= (1.0/6.0) * tr;
    return m_scale(-I * rho, C);
}

void lto_stout(const double *Uin, double *Uout, lto_dims d, double rho)
{
    size_t V = lto_volume(d);
    #pragma omp parallel for
    for (size_t s = 0; s < V; ++s)
        for (int mu = 0; mu < 4; ++mu) {
            mat3 L = m_mul(lto_exp(stout_Q(Uin, d, s, mu, rho)), lto_link(Uin, s, mu));
            m_gram_schmidt(&L);
            lto_set_link(Uout, s, mu, L);
        }
}

/* StoutSmearing4DWithConstants (stout_smearing.hpp:92-119); the tape keeps the
 * hermitian exponent -i rho C from which ExpConstants are rebuilt on demand. */
void lto_stout_tape(const double *Uin, double *Uout, double *Qtape, lto_dims d, double rho)
{
    size_t V = lto_volume(d);
    #pragma omp parallel for
    for (size_t s = 0; s < V; ++s)
        for (int mu = 0; mu < 4; ++mu) {
            mat3 Q = stout_Q(Uin, d, s, mu, rho);
            lto_set_link(Qtape, s, mu, Q);
            mat3 L = m_mul(lto_exp(Q), lto_link(Uin, s, mu));
            m_gram_schmidt(&L);
            lto_set_link(Uout, s, mu, L);
        }
}

/* GradientFlowKernel::CalculateZ / UpdateZ (gradient_flow.hpp:298-332):
 *   Z <- c_old Z + c_new Algebra(st U^+)   (c_old = 0 means overwrite) */
static void flow_update_z(const double *U, double *Z, lto_dims d, double c_old, double c_new, int overwrite)
{
    size_t V = lto_volume(d);
    #pragma omp parallel for
    for (size_t s = 0; s < V; ++s)
        for (int mu = 0; mu < 4; ++mu) {
            mat3 F = m_algebra(m_mul_nd(lto_staple_m(U, d, s, mu), lto_link(U, s, mu)));
            mat3 z = overwrite ? m_rscale(c_new, F)
                               : m_add(m_rscale(c_old, lto_link(Z, s, mu)), m_rscale(c_new, F));
            lto_set_link(Z, s, mu, z);
        }
}

/* GradientFlowKernel::UpdateFields (gradient_flow.hpp:265-282): U <- exp(-i c Z) U, GS */
static void flow_update_fields(double *U, const double *Z, lto_dims d, double c)
{
    size_t n = 4 * lto_volume(d);
    #pragma omp parallel for
    for (size_t i = 0; i < n; ++i) {
        mat3 L = m_mul(lto_exp(m_scale(-I * c, m_load(Z + 18 * i))), m_load(U + 18 * i));
        m_gram_schmidt(&L);
        m_store(U + 18 * i, L);
    }
}

/* Integrators::GradientFlow::Euler (gradient_flow.hpp:31-56) */
void lto_flow_euler(double *U, double *Z, lto_dims d, double eps, int n_steps)
{
    for (int k = 0; k < n_steps; ++k) {
        flow_update_z(U, Z, d, 0.0, eps, 1);
        flow_update_fields(U, Z, d, 1.0);
    }
}

/* Integrators::GradientFlow::RK3 (gradient_flow.hpp:120-168) */
void lto_flow_rk3(double *U, double *Z, lto_dims d, double eps, int n_steps)
{
    for (int k = 0; k < n_steps; ++k) {
        flow_update_z(U, Z, d, 0.0, eps, 1);
        flow_update_fields(U, Z, d, 0.25);
        flow_update_z(U, Z, d, -17.0 / 36.0, 8.0 / 9.0 * eps, 0);
        flow_update_fields(U, Z, d, 1.0);
        flow_update_z(U, Z, d, -1.0, 3.0 / 4.0 * eps, 0);
        flow_update_fields(U, Z, d, 1.0);
    }
}

/* ---- the rest of Observables() (SU3_main.cpp:89-136) ---------------------------- */

/* 1x2 and 2x1 loops (wilson_loop.hpp:155-175): 2 links along mu (or nu) first */
static mat3 rect12(const double *U, lto_dims d, size_t s, int mu, int nu)
{
    size_t smu = lto_move(d, s, mu, 1), smu_nu = lto_move(d, smu, nu, 1);
    size_t snu = lto_move(d, s, nu, 1), snu2 = lto_move(d, snu, nu, 1);
    mat3 a = m_mul(m_mul(lto_link(U, s, mu), lto_link(U, smu, nu)), lto_link(U, smu_nu, nu));
    a = m_mul_nd(m_mul_nd(m_mul_nd(a, lto_link(U, snu2, mu)), lto_link(U, snu, nu)), lto_link(U, s, nu));
    return a;
}
static mat3 rect21(const double *U, lto_dims d, size_t s, int mu, int nu)
{
    size_t smu = lto_move(d, s, mu, 1), smu2 = lto_move(d, smu, mu, 1);
    size_t smu_nu = lto_move(d, smu, nu, 1), snu = lto_move(d, s, nu, 1);
    mat3 a = m_mul(m_mul(lto_link(U, s, mu), lto_link(U, smu, mu)), lto_link(U, smu2, nu));
    a = m_mul_nd(m_mul_nd(m_mul_nd(a, lto_link(U, smu_nu, mu)), lto_link(U, snu, mu)), lto_link(U, s, nu));
    return a;
}

/* Rectangular<2>::Action (rectangular_action.hpp:61-95): sums over x and mu < nu of
 * Re tr P_munu and Re tr (R12 + R21) */
void lto_rect_sums(const double *U, lto_dims d, double *plaq_sum, double *rect_sum)
{
    size_t V = lto_volume(d);
    double sp = 0.0, sr = 0.0;
    #pragma omp parallel for reduction(+:sp, sr)
    for (size_t s = 0; s < V; ++s)
        for (int nu = 1; nu < 4; ++nu)
            for (int mu = 0; mu < nu; ++mu) {
                sp += creal(m_tr(plaq(U, d, s, mu, nu)));
                sr += creal(m_tr(rect12(U, d, s, mu, nu))) + creal(m_tr(rect21(U, d, s, mu, nu)));
            }
    *plaq_sum = sp;
    *rect_sum = sr;
}

/* TopChargePlaquette (topological_charge.hpp:366-416): F_munu = -i (P_munu - P_munu^+)
 * with single plaquettes, Q = 1/(16 pi^2) sum_x Re tr (F01 F23 + F02 F31 + F03 F12) */
double lto_topcharge_plaquette(const double *U, lto_dims d)
{
    size_t V = lto_volume(d);
    double q = 0.0;
    #pragma omp parallel for reduction(+:q)
    for (size_t s = 0; s < V; ++s) {
        mat3 F[6];
        for (int p = 0; p < 6; ++p) {
            mat3 P = plaq(U, d, s, PL_MU[p], PL_NU[p]);
            F[p] = m_scale(-I, m_sub(P, m_dag(P)));
        }
        /* planes 01 02 03 12 13 23; F31 = -F13 */
        q += m_retr_mul(F[0], F[5]) - m_retr_mul(F[1], F[4]) + m_retr_mul(F[2], F[3]);
    }
    return q / (16.0 * LTO_PI * LTO_PI);
}

/* WilsonLoop<n0, n1, reset> (wilson_loop.hpp:65-126): chain(x, mu) is multiplied by the
 * links U_mu(x + n mu) for n0 <= n < n1 (after being reset to 1 when asked); then the
 * square loops of side n1 built from the chains, summed over x and the six planes:
 * returns 1 - sum Re tr / (18 V).  chain: V * 4 general matrices, kept between calls. */
double lto_wilson_loop(const double *U, double *chain, lto_dims d, int n0, int n1, int reset)
{
    size_t V = lto_volume(d);
    #pragma omp parallel for
    for (size_t s = 0; s < V; ++s)
        for (int mu = 0; mu < 4; ++mu) {
            mat3 c = reset ? m_id() : lto_link(chain, s, mu);
            for (int n = n0; n < n1; ++n) c = m_mul(c, lto_link(U, lto_move(d, s, mu, n), mu));
            lto_set_link(chain, s, mu, c);
        }
    double w = 0.0;
    #pragma omp parallel for reduction(+:w)
    for (size_t s = 0; s < V; ++s)
        for (int p = 0; p < 6; ++p) {
            int mu = PL_MU[p], nu = PL_NU[p];
            mat3 a = m_mul(lto_link(chain, s, mu), lto_link(chain, lto_move(d, s, mu, n1), nu));
            a = m_mul_nd(m_mul_nd(a, lto_link(chain, lto_move(d, s, nu, n1), mu)), lto_link(chain, s, nu));
            w += creal(m_tr(a));
        }
    return 1.0 - w / (18.0 * (double)V);
}

/* PolyakovLoop (polyakov_loop.hpp:31-49): (1/V_s) sum over spatial sites of
 * tr prod_t U_t(t, x) */
void lto_polyakov_loop(const double *U, lto_dims d, double *re, double *im)
{
    double sr = 0.0, si = 0.0;
    #pragma omp parallel for collapse(3) reduction(+:sr, si)
    for (int x = 0; x < d.nx; ++x)
        for (int y = 0; y < d.ny; ++y)
            for (int z = 0; z < d.nz; ++z) {
                mat3 c = m_id();
                for (int t = 0; t < d.nt; ++t) c = m_mul(c, lto_link(U, lto_site_index(d, t, x, y, z), 0));
                cplx tr = m_tr(c);
                sr += creal(tr);
                si += cimag(tr);
            }
    double vs = (double)d.nx * d.ny * d.nz;
    *re = sr / vs;
    *im = si / vs;
}

/* ParityUpdate (updates/parity_update.hpp:22-47): with P x = (t, Nx-1-x, Ny-1-y, Nz-1-z),
 * out_t(x) = U_t(P x), out_i(x) = U_i(P x - i)^+ (i spatial) */
void lto_parity_update(const double *U, double *out, lto_dims d)
{
    size_t V = lto_volume(d);
    #pragma omp parallel for
    for (size_t s = 0; s < V; ++s) {
        lto_site c = lto_site_coord(d, s);
        int r[4] = {c.c[0], d.nx - 1 - c.c[1], d.ny - 1 - c.c[2], d.nz - 1 - c.c[3]};
        for (int mu = 0; mu < 4; ++mu) {
            int q[4] = {r[0], r[1], r[2], r[3]};
            if (mu > 0) q[mu] = (q[mu] - 1 + lto_len(d, mu)) % lto_len(d, mu);
            mat3 a = lto_link(U, lto_site_index(d, q[0], q[1], q[2], q[3]), mu);
            lto_set_link(out, s, mu, mu == 0 ? a : m_dag(a));
        }
    }
}
