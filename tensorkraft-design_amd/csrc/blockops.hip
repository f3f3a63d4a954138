// Block-sparse leg permutation and fuse/split on device data (see blockops.hpp).
#include "blockops.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>

#include "kernels.hpp"

namespace tk {
namespace {

constexpr int kMaxRank = 8;
void check_rank(const std::vector<QIndex>& legs, const std::vector<Dir>& dirs) {
  if (legs.empty() || legs.size() > size_t(kMaxRank) || dirs.size() != legs.size())
    throw ArgError("block op: rank outside 1..8 or legs/directions mismatch");
}

// Copy jobs between the original layout and the fused one (to_fused: scatter).
std::vector<CopyJob> fuse_jobs(Sym sym, const std::vector<QIndex>& legs, const std::vector<Dir>& dirs, QN flux,
                               int start, int end, bool to_fused, const BlockTable& tin, const BlockTable& tout) {
  const int rank = int(legs.size());
  const FuseMap fm = make_fuse_map(std::vector<QIndex>(legs.begin() + start, legs.begin() + end),
                                   std::vector<Dir>(dirs.begin() + start, dirs.begin() + end));
  std::vector<CopyJob> jobs;
  std::vector<int32_t> secs(size_t(rank), 0), osecs;
  for (int64_t b = 0; b < tin.n_blocks(); ++b) {
    for (int l = 0; l < rank; ++l) secs[l] = tin.s(b, l);
    int32_t F = -1;
    const int64_t off = fm.lookup(secs.data() + start, end - start, &F);
    if (off < 0 || F < 0) throw std::logic_error("fuse: combination missing from the fuse map");
    osecs.assign(secs.begin(), secs.begin() + start);
    osecs.push_back(F);
    osecs.insert(osecs.end(), secs.begin() + end, secs.end());
    const int64_t ob = tout.find(osecs.data());
    if (ob < 0) throw std::logic_error("fuse: fused block missing from the output table");
    int64_t outer = 1, fdim = 1, trail = 1;
    for (int l = 0; l < start; ++l) outer *= tin.dim(b, l);
    for (int l = start; l < end; ++l) fdim *= tin.dim(b, l);
    for (int l = end; l < rank; ++l) trail *= tin.dim(b, l);
    const int64_t ftot = fm.fused.dim[F];
    const int64_t chunk = fdim * trail, fstride = ftot * trail;
    if (chunk > INT32_MAX || fstride > INT32_MAX || outer > INT32_MAX)
      throw ArgError("fuse: block too large for the copy kernel");
    const int64_t orig = tin.offset[b], fused = tout.offset[ob] + off * trail;
    if (to_fused)
      jobs.push_back({orig, fused, int32_t(outer), int32_t(chunk), int32_t(chunk), int32_t(fstride), 0, 0});
    else
      jobs.push_back({fused, orig, int32_t(outer), int32_t(chunk), int32_t(fstride), int32_t(chunk), 0, 0});
  }
  (void)sym;
  (void)flux;
  return jobs;
}

}  // namespace

void block_permute(Context* ctx, Sym sym, const std::vector<QIndex>& legs, const std::vector<Dir>& dirs, QN flux,
                   const std::vector<int>& perm, const double* in, int64_t in_n, double* out, int64_t out_n) {
  check_rank(legs, dirs);
  const int rank = int(legs.size());
  if (int(perm.size()) != rank) throw ArgError("permute: permutation length differs from the rank");
  std::vector<char> seen(size_t(rank), 0);
  for (int p : perm) {
    if (p < 0 || p >= rank || seen[p]) throw ArgError("permute: not a permutation");
    seen[p] = 1;
  }
  std::vector<QIndex> pl;
  std::vector<Dir> pd;
  for (int p : perm) {
    pl.push_back(legs[p]);
    pd.push_back(dirs[p]);
  }
  const BlockTable tin = make_block_table(sym, legs, dirs, flux);
  const BlockTable tout = make_block_table(sym, pl, pd, flux);
  if (tin.total > in_n || tout.total > out_n) throw std::out_of_range("permute: buffers smaller than the tables");
  // Every block becomes a set of 2-D copies over two of its axes, the other
  // axes fixed (slabs).  When the input's innermost axis a is not the
  // output's innermost axis b, each slab over (a, b) is a transpose
  // out[a][b] = in[b][a] that the copy kernel stages through LDS in 32 x 32
  // tiles; otherwise the innermost axis stays innermost and each slab over
  // (r, innermost), r the longest other axis, is a copy of contiguous rows
  // with fixed row strides on both sides.  Reads and writes are coalesced
  // rows either way.
  std::vector<CopyJob> cjobs;
  std::vector<int32_t> osecs(static_cast<size_t>(rank));
  const int pa = [&] {  // output position of the input's innermost axis
    for (int k = 0; k < rank; ++k)
      if (perm[k] == rank - 1) return k;
    return rank - 1;
  }();
  for (int64_t b = 0; b < tin.n_blocks(); ++b) {
    for (int k = 0; k < rank; ++k) osecs[k] = tin.s(b, perm[k]);
    const int64_t ob = tout.find(osecs.data());
    if (ob < 0) throw std::logic_error("permute: block missing from the output table");
    const int64_t in_off = tin.offset[b], out_off = tout.offset[ob], size = tin.size(b);
    if (out_off + size > tout.total || in_off + size > tin.total)
      throw std::out_of_range("permute: block outside its buffer");
    int64_t st[kMaxRank], ost[kMaxRank], dout[kMaxRank], istr[kMaxRank];
    int64_t s = 1;
    for (int l = rank - 1; l >= 0; --l) {
      st[l] = s;
      s *= tin.dim(b, l);
    }
    for (int k = 0; k < rank; ++k) {
      dout[k] = tin.dim(b, perm[k]);
      istr[k] = st[perm[k]];
    }
    s = 1;
    for (int k = rank - 1; k >= 0; --k) {
      ost[k] = s;
      s *= dout[k];
    }
    // the two copy axes (output positions) and the job shape
    int ra, ca;
    bool trans;
    if (pa != rank - 1) {
      ra = pa;  // rows of the output tile: the input's innermost axis
      ca = rank - 1;
      trans = true;
    } else {
      ca = rank - 1;
      ra = -1;
      for (int k = 0; k < rank - 1; ++k)
        if (ra < 0 || dout[k] > dout[ra]) ra = k;
      trans = false;
    }
    const int64_t m = ra >= 0 ? dout[ra] : 1, n = dout[ca];
    const int64_t src_ld = trans ? istr[ca] : (ra >= 0 ? istr[ra] : n), dst_ld = ra >= 0 ? ost[ra] : n;
    if (m > INT32_MAX || n > INT32_MAX || src_ld > INT32_MAX || dst_ld > INT32_MAX)
      throw ArgError("permute: block too large for the copy kernel");
    std::vector<int> other;
    for (int k = 0; k < rank; ++k)
      if (k != ra && k != ca) other.push_back(k);
    const int64_t slabs = size / (m * n);
    for (int64_t sl = 0; sl < slabs; ++sl) {
      int64_t q = sl, bi = in_off, bo = out_off;
      for (int o = int(other.size()) - 1; o >= 0; --o) {
        const int k = other[size_t(o)];
        const int64_t i = q % dout[k];
        q /= dout[k];
        bi += i * istr[k];
        bo += i * ost[k];
      }
      cjobs.push_back({bi, bo, int32_t(m), int32_t(n), int32_t(src_ld), int32_t(dst_ld), trans ? 1 : 0, 0});
    }
  }
  if (!cjobs.empty()) launch_copies(ctx, cjobs, in, in_n, out, out_n);
  ctx->sync();
}

BlockTable fused_table(Sym sym, const std::vector<QIndex>& legs, const std::vector<Dir>& dirs, QN flux, int start,
                       int end) {
  check_rank(legs, dirs);
  const int rank = int(legs.size());
  if (start < 0 || end > rank || start >= end) throw ArgError("fuse: invalid leg range");
  if (end - start == 1) return make_block_table(sym, legs, dirs, flux);  // single leg: unchanged (bs:457-468)
  const FuseMap fm = make_fuse_map(std::vector<QIndex>(legs.begin() + start, legs.begin() + end),
                                   std::vector<Dir>(dirs.begin() + start, dirs.begin() + end));
  std::vector<QIndex> nl(legs.begin(), legs.begin() + start);
  std::vector<Dir> nd(dirs.begin(), dirs.begin() + start);
  nl.push_back(fm.fused);
  nd.push_back(Dir::In);
  nl.insert(nl.end(), legs.begin() + end, legs.end());
  nd.insert(nd.end(), dirs.begin() + end, dirs.end());
  return make_block_table(sym, nl, nd, flux);
}

void block_fuse(Context* ctx, Sym sym, const std::vector<QIndex>& legs, const std::vector<Dir>& dirs, QN flux,
                int start, int end, const double* in, int64_t in_n, double* out, int64_t out_n) {
  const BlockTable tin = make_block_table(sym, legs, dirs, flux);
  const BlockTable tout = fused_table(sym, legs, dirs, flux, start, end);
  if (tin.total > in_n || tout.total > out_n) throw std::out_of_range("fuse: buffers smaller than the tables");
  if (tout.total != tin.total) throw std::logic_error("fuse: element counts differ");
  if (end - start == 1) {
    TK_HIP(hipMemcpyAsync(out, in, size_t(tin.total) * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    ctx->sync();
    return;
  }
  launch_copies(ctx, fuse_jobs(sym, legs, dirs, flux, start, end, true, tin, tout), in, in_n, out, out_n);
}

void block_split(Context* ctx, Sym sym, const std::vector<QIndex>& legs, const std::vector<Dir>& dirs, QN flux,
                 int start, int end, const double* in, int64_t in_n, double* out, int64_t out_n) {
  const BlockTable torig = make_block_table(sym, legs, dirs, flux);
  const BlockTable tfused = fused_table(sym, legs, dirs, flux, start, end);
  if (tfused.total > in_n || torig.total > out_n) throw std::out_of_range("split: buffers smaller than the tables");
  if (tfused.total != torig.total) throw std::logic_error("split: element counts differ");
  if (end - start == 1) {
    TK_HIP(hipMemcpyAsync(out, in, size_t(torig.total) * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    ctx->sync();
    return;
  }
  launch_copies(ctx, fuse_jobs(sym, legs, dirs, flux, start, end, false, torig, tfused), in, in_n, out, out_n);
}

}  // namespace tk
