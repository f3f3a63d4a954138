// Block-sparse leg permutation and fuse/split on device data (see blockops.hpp).
#include "blockops.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>

#include "kernels.hpp"

namespace tk {
namespace {

constexpr int kMaxRank = 8;
constexpr int64_t kPermChunk = 4096;  // elements per work item

// out block = in block with axes permuted: element e of the out block (row-major
// over dims_out) reads sum_k idx_k * in_stride[k] of the in block.
struct PermJob {
  int64_t in_off, out_off, size;
  int64_t dims_out[kMaxRank];
  int64_t in_stride[kMaxRank];
  int32_t rank, pad;
};
struct PermItem {
  int64_t e0;
  int32_t job, pad;
};

__global__ __launch_bounds__(256) void permute_kernel(const PermItem* __restrict__ items,
                                                      const PermJob* __restrict__ jobs,
                                                      const double* __restrict__ in, double* __restrict__ out) {
  const PermItem it = items[blockIdx.x];
  const PermJob jb = jobs[it.job];
  const int64_t e1 = min(it.e0 + kPermChunk, jb.size);
  for (int64_t e = it.e0 + threadIdx.x; e < e1; e += blockDim.x) {
    int64_t rest = e, src = 0;
    for (int k = jb.rank - 1; k >= 0; --k) {
      const int64_t d = jb.dims_out[k];
      const int64_t i = rest % d;
      rest /= d;
      src += i * jb.in_stride[k];
    }
    out[jb.out_off + e] = in[jb.in_off + src];
  }
}

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
  std::vector<PermJob> jobs;
  std::vector<PermItem> items;
  std::vector<int32_t> osecs(static_cast<size_t>(rank));
  for (int64_t b = 0; b < tin.n_blocks(); ++b) {
    for (int k = 0; k < rank; ++k) osecs[k] = tin.s(b, perm[k]);
    const int64_t ob = tout.find(osecs.data());
    if (ob < 0) throw std::logic_error("permute: block missing from the output table");
    PermJob j{};
    j.in_off = tin.offset[b];
    j.out_off = tout.offset[ob];
    j.size = tin.size(b);
    j.rank = rank;
    int64_t st[kMaxRank];
    int64_t s = 1;
    for (int l = rank - 1; l >= 0; --l) {
      st[l] = s;
      s *= tin.dim(b, l);
    }
    for (int k = 0; k < rank; ++k) {
      j.dims_out[k] = tin.dim(b, perm[k]);
      j.in_stride[k] = st[perm[k]];
    }
    if (j.out_off + j.size > tout.total || j.in_off + j.size > tin.total)
      throw std::out_of_range("permute: block outside its buffer");
    for (int64_t e0 = 0; e0 < j.size; e0 += kPermChunk) items.push_back({e0, int32_t(jobs.size()), 0});
    jobs.push_back(j);
  }
  if (items.empty()) return;
  DBuf<PermJob> dj(ctx, jobs.size());
  DBuf<PermItem> di(ctx, items.size());
  dj.upload(jobs.data(), jobs.size());
  di.upload(items.data(), items.size());
  hipLaunchKernelGGL(permute_kernel, dim3(items.size()), dim3(256), 0, ctx->stream, di.get(), dj.get(), in, out);
  TK_HIP(hipGetLastError());
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
