/*
 * lettuce_hip.h -- C ABI of the MI355X (gfx950) gauge-field library.
 *
 * The reference (timo-eichhorn/Lettuce) is a header-only C++ code whose "API" is
 * templates over compile-time extents plus functors holding references to
 * caller-owned fields.  This ABI replaces that surface for the hot path with
 * opaque handles and runtime extents; the granularity is one sweep / pass /
 * trajectory per call, never one link.  Each entry point cites the reference
 * interface it replaces (paths relative to the reference repository root).
 *
 * Conventions
 *   - Every function returns 0 on success and a negative LT_E* code on failure;
 *     lt_last_error() returns a message for the calling thread's last failure.
 *   - Host buffers are in the reference AoS layout: link index
 *     ((((t*Nx+x)*Ny+y)*Nz+z)*4+mu), 18 doubles per link, row-major, interleaved
 *     re/im (LettuceGauge/lattice.hpp:366-382, defines.hpp:122).
 *   - A context owns one HIP stream; all work of its fields is ordered on it.
 *     Calls returning host scalars synchronise that stream; the others only
 *     enqueue.  One host thread per context at a time.
 *   - All arithmetic is FP64.
 */
#ifndef LETTUCE_HIP_H
#define LETTUCE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LT_OK            0
#define LT_EINVAL       -1   /* bad argument (shape, null pointer, schedule) */
#define LT_EHIP         -2   /* HIP runtime error */
#define LT_ENOMEM       -3   /* device allocation failed */
#define LT_ESTATE       -4   /* object in the wrong state / mismatched shapes */
#define LT_ENUMERIC     -5   /* non-finite or non-unitary links detected */
#define LT_EIO          -6   /* file missing, unreadable, malformed or not writable */

typedef struct lt_ctx_s   *lt_ctx;
typedef struct lt_field_s *lt_field;
typedef struct lt_flow_s  *lt_flow;

/* local-update schedules: Iterator::Checkerboard and Iterator::Checkerboard4
 * (LettuceGauge/iterators/iterators.hpp:65-87, 141-169) */
#define LT_CB2 2
#define LT_CB4 4

/* gradient-flow integrators (LettuceGauge/smearing/gradient_flow.hpp:31-56,120-168) */
#define LT_FLOW_EULER 1
#define LT_FLOW_RK3   3

const char *lt_last_error(void);
int lt_version(int *major, int *minor);

/* ---- context ------------------------------------------------------------- */
int lt_ctx_create(int device, lt_ctx *out);
int lt_ctx_destroy(lt_ctx ctx);
int lt_ctx_synchronize(lt_ctx ctx);
/* the context's HIP stream (a hipStream_t), for callers that time or order work */
int lt_ctx_stream(lt_ctx ctx, void **stream);
/* per-kernel timing with HIP events recorded around every launch on the context
 * stream; query returns launches and total milliseconds since enable/reset */
int lt_ctx_profile(lt_ctx ctx, int enable);
int lt_ctx_profile_query(lt_ctx ctx, const char *kernel, int *launches, double *total_ms);
/* free and total device memory of the context's GPU (hipMemGetInfo), for peak-memory reports */
int lt_ctx_mem_info(lt_ctx ctx, size_t *free_bytes, size_t *total_bytes);
/* Measurement helper with no counterpart in the reference: the rate (bytes/s) at which
 * 16-byte lane loads read a zeroed buffer of footprint_bytes (a power of two in
 * [64 KiB, 1 GiB]), each wavefront load 1 KiB of consecutive bytes, successive loads
 * of a lane far apart so that the L1 misses.  A footprint below one XCD's L2 gives the
 * L2 read ceiling that latency-bound link kernels are compared against.  Allocates and
 * frees its own buffers; best of five timed launches (HIP events, context stream). */
int lt_ctx_read_rate(lt_ctx ctx, size_t footprint_bytes, double *bytes_per_s);

/* ---- gauge fields ---------------------------------------------------------
 * GaugeField4D<Nt,Nx,Ny,Nz,Matrix_SU3> (lattice.hpp:87-394) with runtime
 * extents; all extents must be even (Checkerboard, iterators.hpp:68).          */
int lt_field_create(lt_ctx ctx, int nt, int nx, int ny, int nz, lt_field *out);
int lt_field_destroy(lt_field f);
int lt_field_shape(lt_field f, int shape[4]);
/* host AoS <-> device (GaugeField4D::operator() element access, lattice.hpp:140-168) */
int lt_field_upload(lt_field f, const double *aos);
int lt_field_download(lt_field f, double *aos);
/* copy assignment and pointer swap (lattice.hpp:110-127, 238-246) */
int lt_field_copy(lt_field dst, lt_field src);
int lt_field_swap(lt_field a, lt_field b);
/* SetToIdentity (lattice.hpp:170-183) */
int lt_field_set_identity(lt_field f);
/* Haar-random SU(3) links from Philox key = seed (no reference counterpart;
 * the reference starts from identity or loads a configuration) */
int lt_field_set_haar(lt_field f, uint64_t seed);
/* GramSchmidt on every link (su3.hpp:370-394) */
int lt_field_reunitarize(lt_field f);
/* max |U^+U - 1| and number of non-finite entries (SU3::Tests, su3.hpp:465-659) */
int lt_field_check(lt_field f, double *max_unitarity_dev, long long *n_nonfinite);

/* ---- BMW configuration files (IO/config_io/bmw_format.hpp) ----
 * SaveConfigBMW (full_rows = 0, "#BMW", rows 0 and 1) and SaveConfigBMWFull
 * (full_rows = 1, "#18BMW", all three rows) (bmw_format.hpp:315-402, 517-600): 4096-byte
 * header "#BMW Nx Ny Nz Nt <checksum>", links in t, z, y, x order with directions
 * x, y, z, t, big-endian doubles; an existing file is an error unless overwrite != 0.
 * checksum_hex (hex_len bytes, may be null) receives the header's checksum. */
int lt_field_save_bmw(lt_field f, const char *path, int full_rows, int overwrite, char *checksum_hex, int hex_len);
/* LoadConfigBMW / LoadConfigBMWFull (bmw_format.hpp:153-313, 425-515): the header's
 * extents must equal the field's; *checksum_ok tells whether the recomputed checksum
 * equals the header's (a mismatch is reported, not an error, as in the reference);
 * *max_deviation is max |1 - U^+U| (and, with three rows, the stored row 2 against
 * conj(row0 x row1)); above 1e-6 (SU3::Tests::IsGroupElement) the call returns
 * LT_ENUMERIC. */
int lt_field_load_bmw(lt_field f, const char *path, int full_rows, int *checksum_ok, double *max_deviation);
/* Checksum_Adler64 (bmw_format.hpp:55-100) of the field as it would be saved */
int lt_field_bmw_checksum(lt_field f, int full_rows, unsigned long long *checksum);
/* counter-based RNG state of the field (replaces PRNG4D, random/prng_wrapper.hpp:43-379) */
int lt_field_set_rng(lt_field f, uint64_t seed, uint64_t counter);
int lt_field_get_rng(lt_field f, uint64_t *seed, uint64_t *counter);
/* launch shape of the link kernels: thread blocks cover 4D tiles of extent
 * (tt, tx, ty, tz) in (t, x, y, z/2), which must divide the lattice, or, with
 * tt = tx = ty = 0, tz consecutive checkerboard sites (tz = 64, 128, 192 or 256);
 * results do not depend on it (no reference counterpart: the reference's OpenMP
 * schedule) */
int lt_field_set_tile(lt_field f, int tt, int tx, int ty, int tz);
int lt_field_get_tile(lt_field f, int tile[4]);

/* ---- actions and local updates ------------------------------------------- */
/* Rectangular<1>::Staple / WilsonAction::Staple for every link into out
 * (rectangular_action.hpp:103-184, wilson_action.hpp:174-244) */
int lt_staple_field(lt_field u, lt_field out);
/* n sweeps of OverrelaxationSubgroupKernel driven by Iterator::Checkerboard[4]
 * (overrelaxation.hpp:68-117, iterators.hpp:65-169) */
int lt_or_sweep(lt_field u, int n_sweep, int schedule);
/* Improved gauge actions Rectangular<2>(beta, c_plaq, c_rect) (rectangular_action.hpp:
 * 29-300; Luescher-Weisz c_rect = -1/12, Iwasaki -0.331, DBW2 -1.4088, c_plaq = 1 - 8 c_rect):
 * the staple c_plaq StaplePlaq + c_rect StapleRect into a matrix field, and overrelaxation /
 * heatbath sweeps with it.  c_rect != 0 needs schedule LT_CB4 (Checkerboard4, extents
 * divisible by 4); (1, 0) is the Wilson action of lt_or_sweep / lt_heatbath_sweep. */
int lt_staple_field_action(lt_field u, double c_plaq, double c_rect, lt_field out);
int lt_or_sweep_action(lt_field u, double c_plaq, double c_rect, int n_sweep, int schedule);
int lt_heatbath_sweep_action(lt_field u, double beta, double c_plaq, double c_rect, int n_sweep, int schedule);
/* ParityUpdate (updates/parity_update.hpp:22-47): the spatial reflection x -> N-1-x,
 * which keeps the action and maps the topological charge Q to -Q; u_copy (another SU(3)
 * field of the same extents) receives the old links */
int lt_parity_update(lt_field u, lt_field u_copy);
/* n sweeps of HeatbathKernel (heatbath.hpp:30-130) at coupling beta; draws come
 * from the field's Philox state, whose counter advances by one per sweep */
int lt_heatbath_sweep(lt_field u, double beta, int n_sweep, int schedule);
/* n_update repetitions of the local-update loop body of the reference's main program
 * (SU3_main.cpp:358-385: Iterator::Checkerboard[4](Heatbath, n_heatbath), then
 * Iterator::Checkerboard[4](OverrelaxationSubgroup, n_orelax)), with PlaquetteSum after
 * every update written to plaq_sum[n_update] (may be NULL).  One update is captured
 * once as a hipGraph (the Philox counter and the output slot live in device memory)
 * and replayed per update, for small lattices whose sweeps are launch-bound; the field
 * equals that of the same lt_heatbath_sweep / lt_or_sweep calls bit for bit, and the
 * field's counter advances by n_hb per update.  The plaquette sums are lt_plaquette_sum's
 * (the same terms and reduction tree, bit for bit). */
int lt_update_sweeps(lt_field u, double beta, int n_hb, int n_or, int schedule, int n_update, double *plaq_sum);

/* ---- observables ---------------------------------------------------------- */
/* PlaquetteSum (plaquette.hpp:26-43): sum over sites and the six planes of Re tr P, one
 * link (x, mu) per GPU lane with the planes (mu, mu+1 mod 4) and, for mu < 2, (mu, mu+2),
 * in the tile order of the field's link kernels; agrees with the plaquette of
 * lt_observables to rounding */
int lt_plaquette_sum(lt_field u, double *out);
/* WilsonAction::Action / Rectangular<1>::Action (wilson_action.hpp:23-41) */
int lt_wilson_action(lt_field u, double beta, double *out);
/* Rectangular<2>::ActionNormalized (rectangular_action.hpp:61-100): plaquettes and 1x2 +
 * 2x1 rectangles, (c_plaq (6V - sum Re tr P / 3) + c_rect (12V - sum Re tr R / 3)) / (6V);
 * tree-level Symanzik: c_plaq = 1 + 8/12, c_rect = -1/12 (SU3_main.cpp:119) */
int lt_action_rectangular(lt_field u, double c_plaq, double c_rect, double *action_normalized);
/* TopChargePlaquette (topological_charge.hpp:366-416): single-plaquette field strength */
int lt_topcharge_plaquette(lt_field u, double *q);
/* WilsonLoop<n_start, n_end, reset>(U, U_chain) (wilson_loop.hpp:65-126): the matrix
 * field `chain` holds the straight link products between calls (reset = 1 starts them
 * from the unit matrix); returns 1 - sum Re tr W / (18 V) over the six planes of the
 * n_end x n_end loops */
int lt_wilson_loop(lt_field u, lt_field chain, int n_start, int n_end, int reset, double *w);
/* PolyakovLoop (polyakov_loop.hpp:31-49): spatial average of tr prod_t U_t */
int lt_polyakov_loop(lt_field u, double *re, double *im);
/* TopChargeClover (topological_charge.hpp:97-231); q_t (Nt entries, may be NULL)
 * as TopChargeCloverTimeslice (:337-364) */
int lt_topcharge_clover(lt_field u, double *Q, double *q_t);
/* The per-site summand of TopChargeClover (topological_charge.hpp:184-199):
 * q_x[s] = -1/(256 pi^2) Re tr(F01 F23 - F02 F13 + F03 F12), F = C - C^+, for the V
 * sites in the reference's lexicographic order ((t*Nx+x)*Ny+y)*Nz+z.  Sums of
 * differences of two such densities are free of the rounding of a global sum
 * (finite-difference checks of the charge's derivative). */
int lt_topcharge_clover_density(lt_field u, double *q_x);
/* EnergyDensity::Clover / CloverTimeslice on traceless F (field_strength_tensor.hpp:
 * 107-134,154-195) and EnergyDensity::Plaquette (:140-145); e_t may be NULL */
int lt_energy(lt_field u, double *e_plaq, double *e_clov, double *e_t);
/* all of the above in one fused pass: out[0]=plaquette sum, out[1]=Q, out[2]=E_plaq,
 * out[3]=E_clov, then q_t[Nt], e_t[Nt]; out must hold 4 + 2*Nt doubles */
int lt_observables(lt_field u, double *out);
/* lt_observables without any allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device or pinned memory (tight measurement loops,
 * harnesses that forbid hipMalloc).  The partial sums live in a fixed device array of
 * 8 MiB per GPU (held until the process ends); the final values are computed on the
 * device and copied into out, which must hold 4 + 2*Nt doubles; the call synchronises
 * the context stream (it is not for graph capture).  Results equal lt_observables' bit
 * for bit.  Lattices whose partial sums exceed the
 * array (9*ceil(Nx*Ny*Nz/256)*Nt + 5*Nt + 4 doubles > 2^20: 72^4 fits, 76^4 does not)
 * return LT_EINVAL.  Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (other threads, other contexts' streams) race. */
int lt_observables_noalloc(lt_field u, double *out);

/* ---- smearing -------------------------------------------------------------- */
/* StoutSmearing4D applied n times (stout_smearing.hpp:29-57, StoutSmearingN :61-75);
 * in and out may be the same field */
int lt_stout(lt_field in, lt_field out, double rho, int n_levels);
/* GradientFlowKernel<Integrator, Rectangular<1>> (gradient_flow.hpp:252-367):
 * create binds (unflowed, flowed) like the reference functor; run copies
 * unflowed->flowed then integrates n steps, resume continues without copying */
int lt_flow_create(lt_field unflowed, lt_field flowed, int integrator, double eps, lt_flow *out);
int lt_flow_destroy(lt_flow fl);
int lt_flow_run(lt_flow fl, int n_steps);
int lt_flow_resume(lt_flow fl, int n_steps);

/* ---- HMC --------------------------------------------------------------------
 * Integrators::HMC::{Leapfrog, OMF_2, OMF_4} (LettuceGauge/updates/hmc_gauge.hpp:48-75,
 * 155-190, 241-295) */
#define LT_INT_LEAPFROG 1
#define LT_INT_OMF2     2
#define LT_INT_OMF4     4
/* Multiple-timescale MetaD integrators, Integrators::HMC::{Leapfrog_OMF_4,
 * OMF_2_OMF_4_slow, OMF_2_OMF_4} (hmc_metadynamics.hpp:28-238): the bias force
 * kicks on the outer scale (leapfrog or OMF2), the gauge force on an inner OMF4.
 * MetaD-HMC only (lt_metad_hmc_trajectory, lt_batch_metad_hmc).  Each outer kick
 * evaluates the bias force at the current field (UpdateMetadynamicsMomenta with a
 * fresh CalculateTopologicalForce; the reference's member reuses the last force
 * its UpdateMomenta computed, which these integrators never call).  The merged
 * outer kicks, and the last one, record the CV path samples, n_step per trajectory
 * (OMF_2_OMF_4_slow: the last kick of each step).  OMF_2_OMF_4_slow is the
 * reference's schedule as written: its inner OMF4 blocks advance the field by 2 eps
 * per step while its outer kicks add up to eps. */
#define LT_INT_LEAPFROG_OMF4  5
#define LT_INT_OMF2_OMF4_SLOW 6
#define LT_INT_OMF2_OMF4      7
/* Integrators::HMC::{Leapfrog_slow, OMF_2_slow, OMF_4_slow} (hmc_gauge.hpp:21-46,
 * 121-153, 195-237): the single-scale schedules without merged momentum updates */
#define LT_INT_LEAPFROG_SLOW  8
#define LT_INT_OMF2_SLOW      9
#define LT_INT_OMF4_SLOW      10

/* The kick / drift sequence of an integrator (no reference counterpart: there each
 * integrator's operator() issues it to the HMC kernel).  Event i: kinds[i] = 'P'
 * (momentum update with both forces, UpdateMomenta), 'G' (gauge force only), 'M'
 * (bias force only) or 'U' (UpdateFields), eps[i] its step size, sub[i] = 1 where the
 * MetaD path update samples the CV.  *n = the number of events; at most cap are
 * written (the arrays may be null when cap = 0).  Host only. */
int lt_integrator_schedule(int integrator, double tau, int n_step, char *kinds, double *eps, int *sub, int cap,
                           int *n);

/* a field of general 3x3 matrices (momenta, forces): all nine entries stored */
int lt_matrix_field_create(lt_ctx ctx, int nt, int nx, int ny, int nz, lt_field *out);
/* HMCKernel::RandomMomentum (hmc_gauge.hpp:400-428) with Philox Gaussians */
int lt_hmc_random_momentum(lt_field P, uint64_t seed, uint64_t counter);
/* HMCKernel::UpdateMomenta, P += eps*beta/6 * Algebra(st U^+) (hmc_gauge.hpp:450-464) */
int lt_hmc_update_momenta(lt_field U, lt_field P, double eps, double beta);
/* HMCKernel::UpdateFields, U <- exp(eps P) U, Gram-Schmidt (hmc_gauge.hpp:505-523) */
int lt_hmc_update_fields(lt_field U, lt_field P, double eps);
/* HMCKernel::Hamiltonian, Wilson action + kinetic -sum Re tr P^2 (hmc_gauge.hpp:525-541) */
int lt_hmc_hamiltonian(lt_field U, lt_field P, double beta, double *H);
/* molecular-dynamics part of a trajectory on given momenta */
int lt_hmc_integrate(lt_field U, lt_field P, double beta, int integrator, double tau, int n_step);
/* HMCKernel::operator() (hmc_gauge.hpp:549-583): copy, momenta from U's Philox state
 * (one counter value per trajectory), integrate, Metropolis (restores U on rejection) */
int lt_hmc_trajectory(lt_field U, lt_field P, lt_field U_copy, double beta, int integrator, double tau,
                      int n_step, int metropolis, double *dH, int *accepted);
/* HMCKernel::ReversibilityTest (hmc_gauge.hpp:585-599): momenta from U's Philox state,
 * integrate, reverse the momenta, integrate again; H before and after (the reference
 * prints them).  U and P are left at the end point. */
int lt_hmc_reversibility_test(lt_field U, lt_field P, double beta, int integrator, double tau, int n_step,
                              double *H_old, double *H_new);

/* ---- collective variable and its adjoint -----------------------------------
 * Workspace of HMCMetaDData (hmc_metadynamics.hpp:251-260): n stout levels with
 * their exponent tapes, the clover planes and the recursion's Lambda buffer. */
typedef struct lt_cv_s *lt_cv;
int lt_cv_create(lt_ctx ctx, int nt, int nx, int ny, int nz, int n_levels, double rho, lt_cv *out);
/* As lt_cv_create with per-level checkpointing (no reference counterpart; the
 * reference stores all n+1 levels and n ExpConstants fields, hmc_metadynamics.hpp:
 * 251-260): only levels interval, 2*interval, ... and the top level are kept, the
 * backward pass recomputes each segment from its checkpoint (n extra stout passes),
 * and level 0 is the forward pass's input field, which must stay unchanged until
 * lt_cv_backward.  interval = 0 is full storage; results are identical either way. */
int lt_cv_create_checkpointed(lt_ctx ctx, int nt, int nx, int ny, int nz, int n_levels, double rho,
                              int interval, lt_cv *out);
/* device memory held by the workspace */
int lt_cv_bytes(lt_cv cv, size_t *bytes);
int lt_cv_destroy(lt_cv cv);
/* StoutSmearing4DWithConstants x n, CalculateClover<1>, TopChargeClover(FullTensor)
 * (stout_smearing.hpp:92-119, clover.hpp:94-152, topological_charge.hpp:202-231) */
int lt_cv_forward(lt_cv cv, lt_field U, double *Q);
/* F = dvdq * dQ/dU: CloverDerivative on the last level then StoutForceRecursion down
 * to level 0 (clover.hpp:238-321, stout_smearing.hpp:122-234); F is a matrix field */
int lt_cv_backward(lt_cv cv, double dvdq, lt_field F);

/* ---- metadynamics bias potential (MetaBiasPotential, metadynamics.hpp:29-507) ---- */
typedef struct lt_bias_s *lt_bias;
int lt_bias_create(double cv_min, double cv_max, int bins, double weight, double width, int well_tempered,
                   double well_tempered_parameter, double threshold_weight, lt_bias *out);
int lt_bias_destroy(lt_bias b);
/* UpdatePotential / UpdatePotentialSymmetric (metadynamics.hpp:103-147) */
int lt_bias_deposit(lt_bias b, double cv, int symmetric);
/* ReturnPotential / ReturnDerivative (metadynamics.hpp:151-187) */
int lt_bias_potential(lt_bias b, double cv, double *v);
int lt_bias_derivative(lt_bias b, double cv, double *d);
/* read (set = 0) or write (set = 1) the bins + 1 grid values / the current CV */
int lt_bias_grid(lt_bias b, double *values, int n, int set);
int lt_bias_cv(lt_bias b, double *cv, int set);

/* MetadynamicsTemperingKernel::operator() (updates/tempering.hpp:57-72): two streams, U
 * (unbiased local updates) and U_temper (MetaD-HMC with `bias`).  CV_new = the clover charge
 * of U after the CV's stout levels, dV = V(CV_new) - V(bias CV_current); with q the Philox
 * uniform of stream 4, site 0, key seed, counter (the reference draws prng.UniformReal()),
 * q <= exp(-dV) swaps the links of U and U_temper and sets the bias CV to CV_new */
int lt_metad_tempering_swap(lt_field U, lt_field U_temper, lt_cv cv, lt_bias bias, uint64_t seed, uint64_t counter,
                            int *accepted, double *delta_v);

/* MetaBiasPotential::SaveParameters / SavePotential / LoadPotential
 * (metadynamics.hpp:304-505): the reference's text format (version line, parameter
 * lines up to END_METADYN_PARAMS, comma-separated grid, exceeded_count); overwrite = 0
 * appends.  Loading takes the parameters and the last grid snapshot of the file. */
int lt_bias_save_parameters(lt_bias b, const char *path, int overwrite);
int lt_bias_save_potential(lt_bias b, const char *path, int overwrite);
int lt_bias_load_potential(lt_bias b, const char *path);
/* the primary parameters (params[6] = CV_min, CV_max, weight, width,
 * well_tempered_parameter, threshold_weight), bin_number and exceeded_count;
 * any pointer may be null (no reference counterpart: its members are public) */
int lt_bias_info(lt_bias b, double *params, int *bins, unsigned long long *exceeded);

/* HMCMetaDKernel::operator() (hmc_metadynamics.hpp:552-655): MetaD-HMC trajectory with
 * the adjoint force dV/dQ * dQ/dU, path-sample bias updates when update_bias != 0;
 * F is a matrix field used as force scratch */
int lt_metad_hmc_trajectory(lt_field U, lt_field P, lt_field U_copy, lt_field F, lt_cv cv, lt_bias bias,
                            double beta, int integrator, double tau, int n_step, int metropolis,
                            int update_bias, double *dH, int *accepted, double *cv_new);
/* HMCMetaDKernel::ReversibilityTest (hmc_metadynamics.hpp:657-674): momenta from U's
 * Philox state, integrate, reverse the momenta, integrate again; the biased H before and
 * after (the reference prints them).  U and P are left at the end point, U close to its
 * start for a reversible integrator.  F is force scratch. */
int lt_metad_reversibility_test(lt_field U, lt_field P, lt_field F, lt_cv cv, lt_bias bias, double beta,
                                int integrator, double tau, int n_step, double *H_old, double *H_new);
/* CalculateClover<1> + CloverDerivative(U, Clover, site, mu) for every link
 * (observables/clover.hpp:59-89, 238-321), times scale: F = scale * dQ_clov/dU as
 * the reference's MetaD force uses it (hmc_metadynamics.hpp:420, scale = dV/dQ);
 * F is a matrix field */
int lt_clover_derivative(lt_field U, lt_field F, double scale);
/* cv_path_samples of the last lt_metad_hmc_trajectory on cv (hmc_metadynamics.hpp:398-404):
 * deposited = 0 the CVs visited on the path, else the values deposited (CV_old repeated
 * after a rejection, hmc_metadynamics.hpp:632-638); *n = their number, at most cap copied */
int lt_cv_path_samples(lt_cv cv, int deposited, double *out, int cap, int *n);
/* HMCMetaDKernel::LogCVData (hmc_metadynamics.hpp:537-545): appends those samples to
 * the file as one line, separated by blanks, in the default stream format; the
 * reference logs the visited values after every trajectory and, with the Metropolis
 * step, the deposited ones after it (hmc_metadynamics.hpp:613,642) */
int lt_cv_log_path(lt_cv cv, int deposited, const char *path);

/* ---- replica batches and parallel-tempering metadynamics ------------------ */
/* The reference tempers two streams, an unbiased one and one MetaD-HMC stream
 * with the full bias, swapping them with probability min(1, exp(-(V(Q_unbiased)
 * - V(Q_biased)))) (MetadynamicsTemperingKernel, updates/tempering.hpp:24-73,
 * driven from SU3_main.cpp:549-621).  A batch generalises this to n replicas of
 * equal or mixed extents run by the same launches: replica r runs MetaD-HMC with
 * bias strength[label[r]] * V(Q); adjacent labels k, k+1 held by replicas of
 * equal extents swap with probability min(1, exp(-(s_k - s_{k+1}) (V(Q_j) - V(Q_i)))),
 * which for s = (0, 1) is the reference rule.  The replica holding the top label
 * deposits its path samples (as HMCMetaDKernel does, hmc_metadynamics.hpp:615-650).
 * Defaults: replica r has seed r, counter 0, label r, strength r / (n - 1). */
typedef struct lt_batch_s *lt_batch;

/* swap_mode of lt_batch_metad_hmc */
#define LT_SWAP_NONE 0
#define LT_SWAP_EVEN 1   /* label pairs (0,1), (2,3), ... */
#define LT_SWAP_ODD  2   /* label pairs (1,2), (3,4), ... */
#define LT_SWAP_ALT  3   /* even and odd pairs on alternating steps */

/* shapes: n_rep * {Nt, Nx, Ny, Nz}; n_levels stout levels of parameter rho for the CV
 * (the reference's n_smear_meta and rho_stout_metadynamics); fields start at identity */
int lt_batch_create(lt_ctx ctx, int n_rep, const int *shapes, int n_levels, double rho, lt_batch *out);
int lt_batch_destroy(lt_batch b);
int lt_batch_size(lt_batch b, int *n_rep);
/* the gauge field of replica r as a field handle owned by the batch: every
 * single-lattice entry point accepts it (not lt_field_destroy / lt_field_swap) */
int lt_batch_field(lt_batch b, int r, lt_field *view);
/* Philox key and trajectory counter of replica r used by batched trajectories */
int lt_batch_set_rng(lt_batch b, int r, uint64_t seed, uint64_t counter);
int lt_batch_get_rng(lt_batch b, int r, uint64_t *seed, uint64_t *counter);
/* strengths[k] of label k (n_rep values), labels (NULL: identity), exchange RNG state */
int lt_batch_set_ladder(lt_batch b, const double *strengths, const int *labels, uint64_t pt_seed,
                        uint64_t pt_counter);
/* copy a host bias (parameters, grid, current CV) into the batch / the batch's grid back */
int lt_batch_set_bias(lt_batch b, lt_bias bias);
int lt_batch_get_bias(lt_batch b, lt_bias bias);
/* enable (1) or disable (0) replaying each step as one captured hipGraph */
int lt_batch_use_graph(lt_batch b, int enable);
/* n_traj PT-MetaD steps: one trajectory per replica each, then deposits (update_bias)
 * and exchanges (swap_mode); enqueues only */
int lt_batch_metad_hmc(lt_batch b, double beta, int integrator, double tau, int n_step, int metropolis,
                       int update_bias, int swap_mode, int n_traj);
/* last step, per replica (NULL entries skipped): dH, accepted, CV after the step, label */
int lt_batch_results(lt_batch b, double *dH, int *accepted, double *cv, int *labels);
/* launch-shape autotuning (no reference counterpart: the reference's OpenMP loops
 * have no launch shape).  Times the candidate thread-block tiles of the link
 * kernels on scratch copies and keeps the fastest; results are bit-identical for
 * every tile.  cache_path (may be NULL) is a JSON file of tuned tiles keyed by GPU
 * architecture, extents and kernel family; a cached entry is applied without timing. */
#define LT_TUNE_SWEEP 1   /* heatbath / overrelaxation colour kernels */
#define LT_TUNE_FLOW  2   /* stout / Wilson-flow stage kernels */
int lt_autotune_field(lt_field f, int family, const char *cache_path, int tile_out[4], double *best_ms);
/* per distinct replica extent: the tile of the batched MetaD force kernels */
int lt_batch_autotune(lt_batch b, const char *cache_path);

/* last step's CV path samples, n_rep rows of (samples per trajectory) */
int lt_batch_samples(lt_batch b, double *samples, int n);
/* cumulative exchange statistics: stats[2k] proposed, stats[2k+1] accepted for labels (k, k+1) */
int lt_batch_swap_stats(lt_batch b, long long *stats, int n);

#ifdef __cplusplus
}
#endif
#endif
