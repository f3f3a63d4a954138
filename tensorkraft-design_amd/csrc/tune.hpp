// Autotuner for the tiling of grouped block contractions.
//
// A contraction list (contract.hpp) can run its outputs as 32 x 32 wavefront
// tiles, as 32 x 128 workgroup tiles sharing the A chunk (kWgWide), or as
// 64 x 64 workgroup tiles with both operand chunks in LDS (kWgTile64) from a
// size threshold on.  Which one is fastest depends on the block-size
// distribution and the precision (measured: the wide form wins on the L = 64
// Hubbard bonds, the 64 x 64 form on the FP32 cylinder bond), so each list
// is keyed by a coarse signature and the choice is measured once per
// signature with HIP events on the caller's own operands, then cached for
// the process (and optionally saved to / loaded from a text file).
//
// Reference: the contraction-path and batching decisions the reference
// leaves to its dimension-grouped batched GEMM (ARCHITECTURE.md:1196-1218,
// 2494-2503) with no measured choice.
#pragma once
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "contract.hpp"
#include "device.hpp"

namespace tk {

struct TileChoice {
  int wg_kind = kWgWide;
  int tile_min = 48;
};

// Coarse key of a list: precision, operand flags, log2 buckets of the largest
// output's rows and columns, of the mean inner dimension and of the output
// count, the rounded mean terms per output and sources per term, and
// half-octave buckets of the list's flops.
uint64_t tile_signature(const ContractList& l, int prec);

// The candidates a precision can run (split-BF16 has no 64 x 64 form).
std::vector<TileChoice> tile_candidates(int prec);

// Process-wide table signature -> choice (thread-safe).
bool tune_lookup(int device, uint64_t sig, TileChoice* out);
void tune_store(int device, uint64_t sig, const TileChoice& c, double us_best, double us_default);

struct TuneEntry {
  int device;
  uint64_t sig;
  TileChoice choice;
  double us_best, us_default;
};
std::vector<TuneEntry> tune_entries();
void tune_clear();
// text file: one entry per line "device sig wg_kind tile_min us_best us_default"
void tune_save(const std::string& path);
void tune_load(const std::string& path);

// Measures `run(i)` (which must enqueue candidate i's launch on ctx's stream,
// with its list already on the device) for every candidate: one untimed
// launch, then the best of three HIP-event timings.  Returns the fastest
// and stores it under (ctx's device, sig).
TileChoice tune_measure(Context* ctx, uint64_t sig, const std::vector<TileChoice>& cands,
                        const std::function<void(size_t)>& run);

}  // namespace tk
