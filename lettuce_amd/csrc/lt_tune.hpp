// lt_tune.hpp -- launch-shape autotuner shared by fields and replica batches.
#pragma once

#include <functional>
#include <string>
#include <vector>

#include "lt_internal.hpp"

namespace lt {

// tiles of `threads` threads (0: 64, 128 and 256; whole-row boxes: any multiple of 64 up
// to 256) that divide the extents, z-major, linear maps, and the default map
std::vector<LinkMap> tile_candidates(const Geom& g, int threads);
// site orders (t / y chunking, XCD spread) of one block shape, for the colour sweeps
// (sweeps = true) or the Jacobi kernels
std::vector<LinkMap> order_candidates(const LinkMap& base, bool sweeps);
// "<arch>|<Nt>x<Nx>x<Ny>x<Nz>|<family>"
std::string tune_key(int device, const Geom& g, const char* family);
// cache lookup / update (no-op without a path)
bool cached_tile(const char* path, const std::string& key, const Geom& g, LinkMap& m, double& ms);
void store_tile(const char* path, const std::string& key, const LinkMap& m, double ms);
// times run(m) for every candidate with HIP events on the context stream, re-times the
// fastest few against `incumbent` (the map in use) and returns the winner; the incumbent
// stays unless a candidate beats it by more than 2 %
LinkMap tune(lt_ctx ctx, const std::vector<LinkMap>& cands, const LinkMap& incumbent,
             const std::function<void(const LinkMap&)>& run, double& best_ms, int reps);

}  // namespace lt
