/*
 * lto.h -- CPU oracle for the Lettuce gauge-field hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C (C11 + OpenMP) restatement of the
 * reference algorithms, used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  The product library
 * (lettuce_amd/liblettuce_hip.so) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to the
 * reference repository root).  Field layout is the reference's AoS layout
 * (LettuceGauge/lattice.hpp:366-382): link index ((((t*Nx+x)*Ny+y)*Nz+z)*4+mu),
 * 18 doubles per link, row-major, interleaved re/im (Eigen
 * Matrix<std::complex<double>,3,3,RowMajor>, LettuceGauge/defines.hpp:122).
 *
 * What pins each function: the reference's own outputs at 16^4 (golden vectors in
 * tests/golden/, written from the compiled reference), and for what those leave
 * open, closed forms, invariants and brute force (tests/test_oracle_invariants.py).
 * The Philox streams and the Haar start have no reference counterpart.
 */
#ifndef LTO_H
#define LTO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { int nt, nx, ny, nz; } lto_dims;

/* schedules for local-update sweeps (LettuceGauge/iterators/iterators.hpp) */
enum { LTO_CB2 = 2, LTO_CB4 = 4 };

/* ---- RNG (Philox4x32-10, Salmon et al. SC'11) ---------------------------- */
void   lto_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double lto_u01(uint32_t a, uint32_t b);

/* ---- lattice ------------------------------------------------------------- */
size_t lto_volume(lto_dims d);
int    lto_threads(void);                                   /* OpenMP threads per region */
void   lto_set_identity(double *U, lto_dims d);
void   lto_haar(double *U, lto_dims d, uint64_t seed);
void   lto_reunitarize(double *U, lto_dims d);

/* ---- SU(3) helpers exposed for tests ------------------------------------- */
void   lto_su3_exp(const double *Q, double *out);           /* exp(iQ), Q traceless hermitian */
void   lto_gram_schmidt(double *M);
void   lto_third_row_fma(double *U, size_t nlinks);
void   lto_rect_sums(const double *U, lto_dims d, double *plaq_sum, double *rect_sum);
double lto_topcharge_plaquette(const double *U, lto_dims d);
double lto_wilson_loop(const double *U, double *chain, lto_dims d, int n0, int n1, int reset);
void   lto_polyakov_loop(const double *U, lto_dims d, double *re, double *im);
void   lto_parity_update(const double *U, double *out, lto_dims d);
void   lto_staple_action_field(const double *U, lto_dims d, double c_plaq, double c_rect, double *ST);
void   lto_or_sweep_action(double *U, lto_dims d, double c_plaq, double c_rect, int n_sweep, int schedule);
void   lto_hb_sweep_action(double *U, lto_dims d, double beta, double c_plaq, double c_rect, int n_sweep, int schedule,
                           uint64_t seed, uint64_t *counter);
void   lto_algebra(const double *M, double *out);           /* traceless anti-hermitian part */

/* ---- actions / updates --------------------------------------------------- */
void   lto_staple(const double *U, lto_dims d, size_t site, int mu, double *st);
void   lto_staple_field(const double *U, lto_dims d, double *ST);
void   lto_or_sweep(double *U, lto_dims d, int n_sweep, int schedule);
void   lto_hb_sweep(double *U, lto_dims d, double beta, int n_sweep, int schedule,
                    uint64_t seed, uint64_t *counter);
/* diagnostic: per-link heatbath with explicit rejection cap (heatbath.hpp:106-108) */
void   lto_hb_link(double *U, lto_dims d, double beta, size_t site, int mu, int max_iter,
                   uint64_t seed, uint64_t counter, int *attempts_out);

/* ---- observables --------------------------------------------------------- */
double lto_plaquette_sum(const double *U, lto_dims d);
double lto_wilson_action(const double *U, lto_dims d, double beta);
void   lto_clover(const double *U, lto_dims d, double *C6);       /* 6 planes per site: 01,02,03,12,13,23 */
double lto_topcharge_clover(const double *U, lto_dims d);          /* topological_charge.hpp:97-200  */
void   lto_topcharge_density(const double *U, lto_dims d, double *q_x);  /* per-site summand, :184-199 */
double lto_topcharge_clover_tensor(const double *C6, lto_dims d);  /* topological_charge.hpp:202-231 */
void   lto_topcharge_timeslice(const double *U, lto_dims d, double *q_t);   /* :337-364 via F */
void   lto_energy_clover_timeslice(const double *U, lto_dims d, double *e_t);
double lto_energy_plaquette(const double *U, lto_dims d);

/* ---- smearing ------------------------------------------------------------ */
void   lto_stout(const double *Uin, double *Uout, lto_dims d, double rho);
/* stout level with the algebra tape: Qtape[link] = -i*rho*C (18 doubles, hermitian) */
void   lto_stout_tape(const double *Uin, double *Uout, double *Qtape, lto_dims d, double rho);
void   lto_flow_euler(double *U, double *Z, lto_dims d, double eps, int n_steps);
void   lto_flow_rk3(double *U, double *Z, lto_dims d, double eps, int n_steps);

/* ---- adjoint (metadynamics force) ---------------------------------------- */
void   lto_clover_derivative(const double *U, const double *C6, lto_dims d, double *F);
void   lto_stout_force_recursion(const double *U, const double *Uprev, double *Sigma,
                                 const double *Qtape, lto_dims d, double rho, double *Lambda);
/* full CV force: n stout levels, clover Q, dV/dQ scale, recursions back to level 0 */
double lto_cv_force(const double *U, lto_dims d, double rho, int n_levels, double dvdq,
                    double *F);
double lto_cv_charge(const double *U, lto_dims d, double rho, int n_levels);

/* ---- HMC ----------------------------------------------------------------- */
void   lto_random_momentum(double *P, lto_dims d, uint64_t seed, uint64_t counter);
void   lto_gauge_force_update(const double *U, double *P, lto_dims d, double eps, double beta);
void   lto_field_update(double *U, const double *P, lto_dims d, double eps);
double lto_kinetic(const double *P, lto_dims d);

#ifdef __cplusplus
}
#endif
#endif
