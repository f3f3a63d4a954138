// Abelian quantum numbers, packed sector keys, per-leg bases and block tables.
//
// Host-side restatement of tk-symmetry's sector bookkeeping; every table built
// here is uploaded to the device as flat integer arrays.
//   U1 pack/unpack ............ tk-symmetry/src/builtins.rs:12-40
//   PackedSectorKey ........... tk-symmetry/src/sector_key.rs:31-55
//   QIndex .................... tk-symmetry/src/sector_key.rs:108-158
//   check_flux_rule ........... tk-symmetry/src/flux.rs:9-27
//   enumerate_valid_sectors ... tk-symmetry/src/flux.rs:43-68, 186-234
//   BlockSparseTensor order ... tk-symmetry/src/block_sparse.rs:40-84 (ascending u64 key)
//   fuse_legs offsets ......... tk-symmetry/src/block_sparse.rs:444-611
// U1U1 = (N_up, N_dn) has no reference counterpart; it is modelled on U1Z2
// (builtins.rs:86-113): 16 bits, pack = up | dn << 8, Ord = (up, dn).
#pragma once
#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace tk {

enum class Sym : int32_t { U1 = 1, U1U1 = 2 };
enum class Dir : int32_t { In = 0, Out = 1 };

struct QN {
  int32_t a = 0, b = 0;  // U1 uses `a` only; U1U1 = (up, dn)
  friend bool operator==(QN x, QN y) { return x.a == y.a && x.b == y.b; }
  friend bool operator!=(QN x, QN y) { return !(x == y); }
  friend bool operator<(QN x, QN y) { return x.a != y.a ? x.a < y.a : x.b < y.b; }
};
inline QN fuse(QN x, QN y) { return {x.a + y.a, x.b + y.b}; }
inline QN dual(QN x) { return {-x.a, -x.b}; }
inline QN effective(QN q, Dir d) { return d == Dir::In ? q : dual(q); }

struct SymError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline int bit_width(Sym s) { return s == Sym::U1 ? 8 : 16; }

// One quantum number -> low bits (two's complement bytes, builtins.rs:29-40).
inline uint64_t pack_qn(Sym s, QN q) {
  auto byte = [](int32_t v) -> uint64_t {
    if (v < -128 || v > 127) throw SymError("quantum number outside the 8-bit range: " + std::to_string(v));
    return uint64_t(uint8_t(int8_t(v)));
  };
  if (s == Sym::U1) {
    if (q.b != 0) throw SymError("U1 quantum number with a second component");
    return byte(q.a);
  }
  return byte(q.a) | (byte(q.b) << 8);
}
inline QN unpack_qn(Sym s, uint64_t bits) {
  QN q;
  q.a = int32_t(int8_t(uint8_t(bits & 0xFF)));
  q.b = s == Sym::U1 ? 0 : int32_t(int8_t(uint8_t((bits >> 8) & 0xFF)));
  return q;
}
inline uint64_t pack_key(Sym s, const QN* qs, int rank) {
  const int w = bit_width(s);
  if (rank * w > 64) throw SymError("sector rank exceeds the 64-bit key");
  uint64_t k = 0;
  for (int i = 0; i < rank; ++i) k |= pack_qn(s, qs[i]) << (i * w);
  return k;
}

// Per-leg basis: sectors sorted by QN, prefix offsets (sector_key.rs:108-140).
struct QIndex {
  std::vector<QN> q;
  std::vector<int64_t> dim;
  std::vector<int64_t> off;
  int64_t total = 0;

  QIndex() = default;
  QIndex(std::vector<std::pair<QN, int64_t>> s) {
    std::stable_sort(s.begin(), s.end(), [](auto& x, auto& y) { return x.first < y.first; });
    for (size_t i = 0; i < s.size(); ++i) {
      if (s[i].second <= 0) throw SymError("QIndex sector with non-positive dimension");
      if (i && !(s[i - 1].first < s[i].first)) throw SymError("QIndex with repeated quantum number");
      q.push_back(s[i].first);
      dim.push_back(s[i].second);
      off.push_back(total);
      total += s[i].second;
    }
  }
  int n() const { return int(q.size()); }
  // Sector position of `x`, or -1 (binary search over the sorted list).
  int find(QN x) const {
    auto it = std::lower_bound(q.begin(), q.end(), x);
    if (it == q.end() || *it != x) return -1;
    return int(it - q.begin());
  }
};

// Block table of a block-sparse tensor: every flux-allowed sector tuple, sorted
// by ascending packed key, with row-major blocks laid out back to back
// (FlatBlockStorage, tk-symmetry/src/flat_storage.rs:95-140).
struct BlockTable {
  Sym sym = Sym::U1;
  std::vector<QIndex> legs;
  std::vector<Dir> dirs;
  QN flux;
  std::vector<uint64_t> keys;
  std::vector<int32_t> sec;     // n_blocks * rank sector positions
  std::vector<int64_t> offset;  // element offset of each block
  int64_t total = 0;

  int rank() const { return int(legs.size()); }
  int64_t n_blocks() const { return int64_t(keys.size()); }
  int s(int64_t blk, int leg) const { return sec[size_t(blk) * rank() + leg]; }
  int64_t dim(int64_t blk, int leg) const { return legs[leg].dim[s(blk, leg)]; }
  QN qn(int64_t blk, int leg) const { return legs[leg].q[s(blk, leg)]; }
  int64_t size(int64_t blk) const {
    int64_t n = 1;
    for (int l = 0; l < rank(); ++l) n *= dim(blk, l);
    return n;
  }
  // Block position for a sector tuple, or -1.
  int64_t find(const int32_t* secs) const {
    QN qs[8];
    for (int l = 0; l < rank(); ++l) qs[l] = legs[l].q[secs[l]];
    uint64_t k = pack_key(sym, qs, rank());
    auto it = std::lower_bound(keys.begin(), keys.end(), k);
    if (it == keys.end() || *it != k) return -1;
    return int64_t(it - keys.begin());
  }
};

bool check_flux(const QN* qs, const Dir* dirs, int rank, QN flux);

// All flux-allowed sector tuples, sorted by packed key, with offsets.
BlockTable make_block_table(Sym sym, std::vector<QIndex> legs, std::vector<Dir> dirs, QN flux);

// Map of fuse_legs over legs [start, end): for every sub-sector tuple the fused
// QN and the offset inside the fused sector; the fused leg is Incoming.
struct FuseMap {
  QIndex fused;
  // combos in leg-by-leg enumeration order (later leg fastest)
  std::vector<std::vector<int32_t>> combo_secs;
  std::vector<int32_t> combo_fused_sector;
  std::vector<int64_t> combo_offset;
  std::vector<int64_t> combo_dim;
  std::vector<int32_t> radix;  // sector count of each sub-leg
  int64_t lookup(const int32_t* secs, int n, int32_t* fused_sector) const;
};
FuseMap make_fuse_map(const std::vector<QIndex>& legs, const std::vector<Dir>& dirs);

}  // namespace tk
