// Serialized device trees: a tree built once can be written to a byte buffer
// (or file) and brought back resident without rebuilding it, for restarts of
// a serving process and for checkpoints.  The reference has no counterpart
// (its trees live only in the Python objects that built them).
//
// Format (little-endian), version 2:
//   char[8]   magic "SBHTREE\0"
//   u32       version
//   i32 x5    dim, split_log, leaf_size, max_depth, moments64
//   u32 x2    n (points), num_nodes
//   i32       depth (deepest level)
//   f32       alpha
//   f32 x6    min_corner[3], max_corner[3]
//   u32       number of levels L, then u32 level_begin[L + 1]
//   sections  for each array, in the order of for_each_array:
//               u64 byte count, then the bytes
//   u64       FNV-1a 64 of every byte before it
// Array lengths follow from n and num_nodes; the loader checks every count,
// the structure sizes and the checksum before it allocates device memory.
// Version 2: the depth-first record (dfsC, dfsN) of a one-point leaf holds
// that point and its normal, not the leaf's moment.
#include <hip/hip_runtime.h>

#include <cstring>
#include <type_traits>
#include <vector>

#include "sbh_runtime.hpp"

namespace sbh {
namespace {

constexpr char kMagic[8] = {'S', 'B', 'H', 'T', 'R', 'E', 'E', '\0'};
constexpr uint32_t kVersion = 2;
constexpr uint32_t kEndHost = 0xFFFFFFFFu;      // rope terminator
constexpr uint32_t kLeafBitHost = 0x80000000u;  // leaf flag in a walk record

// FNV-1a, 64-bit: offset basis 0xcbf29ce484222325, prime 0x100000001b3.
uint64_t fnv1a(const uint8_t* p, size_t n) {
  uint64_t h = 0xcbf29ce484222325ull;
  for (size_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 0x100000001b3ull;
  }
  return h;
}

struct Writer {
  std::vector<uint8_t>& out;
  template <typename T>
  void pod(const T& v) {
    const uint8_t* p = reinterpret_cast<const uint8_t*>(&v);
    out.insert(out.end(), p, p + sizeof(T));
  }
  template <typename T>
  void device(const DevBuf<T>& b, size_t count, hipStream_t s) {
    const uint64_t bytes = count * sizeof(T);
    pod(bytes);
    const size_t at = out.size();
    out.resize(at + bytes);
    if (bytes) {
      SBH_REQUIRE(b.get() && b.size() >= count, "tree array shorter than its logical length");
      // completes before the next resize can move the buffer
      SBH_HIP(hipMemcpyAsync(out.data() + at, b.get(), bytes, hipMemcpyDeviceToHost, s));
      SBH_HIP(hipStreamSynchronize(s));
    }
  }
};

struct Reader {
  const uint8_t* p;
  size_t n;
  size_t at = 0;
  template <typename T>
  T pod() {
    SBH_REQUIRE(at + sizeof(T) <= n, "serialized tree is truncated");
    T v;
    std::memcpy(&v, p + at, sizeof(T));
    at += sizeof(T);
    return v;
  }
  // Position of the next section of count T values; advances past it.
  template <typename T>
  const uint8_t* section(size_t count) {
    const uint64_t bytes = pod<uint64_t>();
    SBH_REQUIRE(bytes == count * sizeof(T), "serialized tree array has the wrong length");
    SBH_REQUIRE(at + bytes <= n, "serialized tree is truncated");
    const uint8_t* q = p + at;
    at += bytes;
    return q;
  }
};

template <typename T>
std::vector<T> host_copy(const uint8_t* q, size_t count) {
  std::vector<T> v(count);
  if (count) std::memcpy(v.data(), q, count * sizeof(T));
  return v;
}

// The walks index node and point arrays through these records, so a
// corrupted or hostile buffer must not reach the device: every index they
// follow is checked to be in range and to move strictly forward (every walk
// then ends).
void validate_structure(const DeviceTree& t, const std::vector<const uint8_t*>& sec) {
  const uint32_t n = t.n, nn = t.num_nodes;
  const uint32_t max_children = 1u << (t.dim * t.split_log);
  // section order: see for_each_array
  const auto start = host_copy<uint32_t>(sec[3], nn);
  const auto end = host_copy<uint32_t>(sec[4], nn);
  const auto first = host_copy<uint32_t>(sec[6], nn);
  const auto nchild = host_copy<uint32_t>(sec[7], nn);
  const auto nodeL = host_copy<uint4>(sec[15], nn);
  const auto skip = host_copy<uint32_t>(sec[17], nn);
  const auto dfsL = host_copy<uint4>(sec[22], nn);
  const auto dfsC = host_copy<float4>(sec[20], nn);
  const auto dfsN = host_copy<float4>(sec[21], nn);
  const auto pts = host_copy<float4>(sec[0], n);
  const auto nrm = host_copy<float4>(sec[1], n);
  const auto perm = host_copy<uint32_t>(sec[2], n);
  for (uint32_t i = 0; i < n; ++i) SBH_REQUIRE(perm[i] < n, "serialized tree: bad leaf order");
  SBH_REQUIRE(start[0] == 0 && end[0] == n, "serialized tree: root must hold every point");
  for (uint32_t v = 0; v < nn; ++v) {
    SBH_REQUIRE(start[v] < end[v] && end[v] <= n, "serialized tree: bad point range");
    const uint4 l = nodeL[v];
    if (l.w & kLeafBitHost) {
      SBH_REQUIRE(nchild[v] == 0 && l.y == start[v] && (l.w & ~kLeafBitHost) == end[v] - start[v],
                  "serialized tree: bad leaf record");
    } else {
      SBH_REQUIRE(l.w == nchild[v] && l.w >= 1 && l.w <= max_children && l.y == first[v] &&
                      l.y > v && (uint64_t)l.y + l.w <= nn,
                  "serialized tree: bad child range");
      SBH_REQUIRE(start[l.y] == start[v] && end[l.y + l.w - 1] == end[v],
                  "serialized tree: children do not cover their parent");
      for (uint32_t c = l.y; c + 1 < l.y + l.w; ++c)
        SBH_REQUIRE(end[c] == start[c + 1], "serialized tree: children not contiguous");
    }
    SBH_REQUIRE(skip[v] >= v && skip[v] < nn && start[skip[v]] == start[v] &&
                    end[skip[v]] == end[v],
                "serialized tree: bad single-child chain target");
    const uint4 d = dfsL[v];
    SBH_REQUIRE(d.z == kEndHost || (d.z > v && d.z < nn), "serialized tree: bad rope");
    if (d.w & kLeafBitHost) {
      SBH_REQUIRE((uint64_t)d.y + (d.w & ~kLeafBitHost) <= n, "serialized tree: bad leaf record");
      if (d.w == (kLeafBitHost | 1u))
        SBH_REQUIRE(!std::memcmp(&dfsC[v], &pts[d.y], sizeof(float4)) &&
                        !std::memcmp(&dfsN[v], &nrm[d.y], sizeof(float4)),
                    "serialized tree: one-point leaf record does not hold its point");
    } else
      SBH_REQUIRE(d.y == v + 1 && d.y < nn && d.w >= 1 && d.w <= max_children,
                  "serialized tree: bad depth-first record");
  }
}

// Every device array with its logical length, in file order.
template <typename F>
void for_each_array(DeviceTree& t, F&& f) {
  const size_t n = t.n, nn = t.num_nodes;
  f(t.pts, n);
  f(t.nrm, n);
  f(t.perm, n);
  f(t.n_start, nn);
  f(t.n_end, nn);
  f(t.n_parent, nn);
  f(t.n_first, nn);
  f(t.n_nchild, nn);
  f(t.n_next, nn);
  f(t.n_depth, nn);
  f(t.n_lo, nn);
  f(t.n_hi, nn);
  f(t.n_diam, nn);
  f(t.nodeC, nn);
  f(t.nodeN, nn);
  f(t.nodeL, nn);
  f(t.n_m64, 8 * nn);
  f(t.n_skip, nn);
  f(t.n_size, nn);
  f(t.n_dfs, nn);
  f(t.dfsC, nn);
  f(t.dfsN, nn);
  f(t.dfsL, nn);
}

}  // namespace

void serialize_tree(const DeviceTree& tc, std::vector<uint8_t>& out) {
  DeviceTree& t = const_cast<DeviceTree&>(tc);  // arrays are only read
  out.clear();
  Writer w{out};
  out.insert(out.end(), kMagic, kMagic + 8);
  w.pod(kVersion);
  w.pod<int32_t>(t.dim);
  w.pod<int32_t>(t.split_log);
  w.pod<int32_t>(t.leaf_size);
  w.pod<int32_t>(t.max_depth);
  w.pod<int32_t>(t.moments64);
  w.pod<uint32_t>(t.n);
  w.pod<uint32_t>(t.num_nodes);
  w.pod<int32_t>(t.depth);
  w.pod<float>(t.alpha);
  for (int a = 0; a < 3; ++a) w.pod<float>(t.lo[a]);
  for (int a = 0; a < 3; ++a) w.pod<float>(t.hi[a]);
  SBH_REQUIRE(!t.level_begin.empty(), "tree has no levels");
  w.pod<uint32_t>((uint32_t)t.level_begin.size() - 1);
  for (uint32_t v : t.level_begin) w.pod(v);
  for_each_array(t, [&](auto& b, size_t count) { w.device(b, count, t.stream); });
  SBH_HIP(hipStreamSynchronize(t.stream));
  const uint64_t h = fnv1a(out.data(), out.size());
  w.pod(h);
}

void deserialize_tree(DeviceTree& t, const uint8_t* data, size_t size) {
  SBH_REQUIRE(data && size >= 8 + sizeof(uint64_t), "serialized tree is truncated");
  SBH_REQUIRE(std::memcmp(data, kMagic, 8) == 0, "not a serialized tree (bad magic)");
  uint64_t stored;
  std::memcpy(&stored, data + size - sizeof(uint64_t), sizeof(uint64_t));
  SBH_REQUIRE(fnv1a(data, size - sizeof(uint64_t)) == stored,
              "serialized tree checksum mismatch");
  Reader r{data, size - sizeof(uint64_t), 8};
  SBH_REQUIRE(r.pod<uint32_t>() == kVersion, "unsupported serialized tree version");
  t.dim = r.pod<int32_t>();
  t.split_log = r.pod<int32_t>();
  t.leaf_size = r.pod<int32_t>();
  t.max_depth = r.pod<int32_t>();
  t.moments64 = r.pod<int32_t>();
  t.n = r.pod<uint32_t>();
  t.num_nodes = r.pod<uint32_t>();
  t.depth = r.pod<int32_t>();
  t.alpha = r.pod<float>();
  for (int a = 0; a < 3; ++a) t.lo[a] = r.pod<float>();
  for (int a = 0; a < 3; ++a) t.hi[a] = r.pod<float>();
  SBH_REQUIRE(t.dim == 2 || t.dim == 3, "serialized tree: bad dim");
  SBH_REQUIRE(t.split_log >= 1 && t.split_log <= 3, "serialized tree: bad split_log");
  SBH_REQUIRE(t.leaf_size >= 1 && t.n >= 1 && t.num_nodes >= 1 && t.num_nodes < 0x7FFFFFFFu,
              "serialized tree: bad sizes");
  const uint32_t levels = r.pod<uint32_t>();
  SBH_REQUIRE(levels >= 1 && levels <= 1002 && (int)levels == t.depth + 1,
              "serialized tree: bad level count");
  t.level_begin.resize(levels + 1);
  for (auto& v : t.level_begin) v = r.pod<uint32_t>();
  SBH_REQUIRE(t.level_begin.front() == 0 && t.level_begin.back() == t.num_nodes,
              "serialized tree: bad level table");
  for (uint32_t l = 0; l < levels; ++l)
    SBH_REQUIRE(t.level_begin[l] < t.level_begin[l + 1], "serialized tree: bad level table");
  std::vector<const uint8_t*> sec;
  for_each_array(t, [&](auto& b, size_t count) {
    using T = typename std::remove_reference_t<decltype(b)>::value_type;
    sec.push_back(r.section<T>(count));
  });
  SBH_REQUIRE(r.at == r.n, "serialized tree has trailing bytes");
  validate_structure(t, sec);
  if (!t.stream) {
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0)
      throw Error(kErrNoDevice, "no GPU visible to the HIP runtime");
    SBH_HIP(hipStreamCreateWithFlags(&t.stream, hipStreamNonBlocking));
  }
  size_t i = 0;
  for_each_array(t, [&](auto& b, size_t count) {
    using T = typename std::remove_reference_t<decltype(b)>::value_type;
    b.alloc(count);
    if (count)
      SBH_HIP(hipMemcpyAsync(b.get(), sec[i], count * sizeof(T), hipMemcpyHostToDevice,
                             t.stream));
    ++i;
  });
  derive_child_sources(t);
  SBH_HIP(hipStreamSynchronize(t.stream));
}

}  // namespace sbh
