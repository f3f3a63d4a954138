// Leg permutation and fuse/split of consecutive legs for block-sparse tensors
// on the device (blockops.hip).  Data are in FlatBlockStorage layout with
// every flux-allowed block present (make_block_table order).
//
// Reference: BlockSparseTensor::permute (tk-symmetry/src/block_sparse.rs:400-431),
// fuse_legs (block_sparse.rs:444-611) and split_leg (block_sparse.rs:622-793).
// The fused leg is Incoming and its sectors and per-combination offsets are
// those of make_fuse_map; each input block lands in its fused block as
// outer x (fuse_dim * trail) chunks at offset * trail, as in the reference.
#pragma once
#include <vector>

#include "device.hpp"
#include "sym.hpp"

namespace tk {

// out: the tensor with legs/dirs permuted (out leg k = in leg perm[k]), same flux.
void block_permute(Context* ctx, Sym sym, const std::vector<QIndex>& legs, const std::vector<Dir>& dirs, QN flux,
                   const std::vector<int>& perm, const double* in, int64_t in_n, double* out, int64_t out_n);
// Layout of legs [0, start), the fused leg (Incoming), legs [end, rank).
BlockTable fused_table(Sym sym, const std::vector<QIndex>& legs, const std::vector<Dir>& dirs, QN flux, int start,
                       int end);
// fuse_legs(start..end): in (layout of legs) -> out (layout of fused_table).
void block_fuse(Context* ctx, Sym sym, const std::vector<QIndex>& legs, const std::vector<Dir>& dirs, QN flux,
                int start, int end, const double* in, int64_t in_n, double* out, int64_t out_n);
// split_leg: the inverse, in (layout of fused_table) -> out (layout of legs).
void block_split(Context* ctx, Sym sym, const std::vector<QIndex>& legs, const std::vector<Dir>& dirs, QN flux,
                 int start, int end, const double* in, int64_t in_n, double* out, int64_t out_n);

}  // namespace tk
