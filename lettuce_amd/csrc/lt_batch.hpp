// lt_batch.hpp -- replica batches: many lattices (equal or mixed extents) whose
// fields share one allocation per kind, so that one launch advances all of them.
// The per-replica table, the device-side bias potential and the launchers of the
// batched kernels (lt_batch.hip) used by the C ABI (lt_batch_api.cpp).
#pragma once

#include "lt_internal.hpp"

namespace lt {

// One replica of a batch.  Link-layout buffers of replica r start at off
// (in complex doubles, 72*Vh per field); the clover planes at clover_offset(off).
struct BatchRep {
    Geom g;
    LinkMap m;
    long long off;
    uint64_t seed;
};

// Block-range prefixes of the batched grids (n+1 entries each, last = total):
//   lb  link kernels: 8 * tiles blocks per replica
//   sb  site kernels: ceil(V / 256) blocks per replica
//   kb  per-link reductions: ceil(4V / 256) blocks per replica
//   cb  elementwise copies of one link field: ceil(72*Vh / 256) blocks per replica
struct BatchGrid {
    const BatchRep* reps;
    const int* lb;
    const int* sb;
    const int* kb;
    const int* cb;
    int n;
    int nthreads;   // block size of the link kernels: the largest tile of the batch
};

// MetaBiasPotential on the device (LettuceGauge/actions/bias_potential/metadynamics.hpp:29-507):
// the grid lives in device memory, the parameters are copied from the host object.
struct DevBias {
    double* grid;
    double cv_min, cv_max, bin_width, bin_width_inverse;
    double weight, width, wt_param, threshold;
    int bins, well_tempered;
};

// Device-resident per-replica state of a batched MetaD-HMC step.
struct BatchState {
    uint64_t* counter;     // per replica RNG counter (momenta: stream 1, accept: stream 3)
    int* label;            // label[r] = index into strength[] held by replica r
    const double* strength;
    double* qraw;          // raw clover sums of the last CV forward pass
    double* dvdq;          // strength * dV/dQ of the last force evaluation
    double* plaq;          // plaquette sums
    double* kin;           // kinetic terms
    double* H0;
    double* H1;
    double* cv_old;
    double* cv_cur;        // CV of the current configuration (after accept / reject)
    double* dH;
    int* accepted;
    double* samples;       // [n][n_sub] CV path samples
    int n_sub;
    long long* swap_stats; // [n_labels]: proposed, accepted per adjacent label pair (2 entries each)
    uint64_t* pt_counter;  // [1]
    uint64_t pt_seed;
    unsigned long long* exceeded;   // [1] deposits outside [cv_min, cv_max)
    double* bias_cv;       // [1] the bias' current CV (top-label replica)
};

void kb_copy_links(lt_ctx ctx, const BatchGrid& bg, int total_cb, const double2* src, double2* dst, const int* only_rejected);
void kb_momentum(lt_ctx ctx, const BatchGrid& bg, int total_lb, double2* P, const uint64_t* counter);
void kb_force_update(lt_ctx ctx, const BatchGrid& bg, int total_lb, const double2* U, double2* P, const double2* F,
                     double a, double b);
void kb_field_update(lt_ctx ctx, const BatchGrid& bg, int total_lb, double2* U, const double2* P, double eps);
void kb_stout_tape(lt_ctx ctx, const BatchGrid& bg, int total_lb, const double2* in, double2* out, double2* tape,
                   double rho);
// clover planes, charge-density partials and per-replica finish into qraw
void kb_clover_charge(lt_ctx ctx, const BatchGrid& bg, int total_sb, const double2* U, double2* C, double* part,
                      double* qraw);
// next: go on with the Lambda of the top stout level (see k_clover_deriv)
void kb_clover_deriv(lt_ctx ctx, const BatchGrid& bg, int total_lb, const double2* U, const double2* C, double2* F,
                     const double* scale, const SigmaNext* next = nullptr);
// the stout-force recursion of levels hi..lo+1 (see k_stout_recursion_run)
void kb_stout_recursion_run(lt_ctx ctx, const BatchGrid& bg, int total_lb, const LevelFn& level, const LevelFn& tape,
                            int hi, int lo, double2* Sig, double2* const lam[2], double rho,
                            const MomentumKick* kick = nullptr, bool top_lambda_done = false);
void kb_plaquette(lt_ctx ctx, const BatchGrid& bg, int total_sb, const double2* U, double* part, double* out);
void kb_kinetic(lt_ctx ctx, const BatchGrid& bg, int total_kb, const double2* P, double* part, double* out);
// scalar stages (one thread per replica)
void kb_dvdq(lt_ctx ctx, const BatchGrid& bg, const BatchState& st, const DevBias& bias, int sample_index);
void kb_h0(lt_ctx ctx, const BatchGrid& bg, const BatchState& st, const DevBias& bias, double beta);
void kb_h1_accept(lt_ctx ctx, const BatchGrid& bg, const BatchState& st, const DevBias& bias, double beta,
                  int metropolis);
void kb_deposit(lt_ctx ctx, const BatchGrid& bg, const BatchState& st, const DevBias& bias, int n_labels);
void kb_swap_advance(lt_ctx ctx, const BatchGrid& bg, const BatchState& st, const DevBias& bias, int n_labels,
                     int do_swap);

}  // namespace lt
