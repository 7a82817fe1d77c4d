// Host-side sparse-voxel octree builder.
//
// Produces the reference's flattened layout (fairnr/clib/src/octree.cpp:
// 93-123) without building a pointer tree: every point's octant path from the
// root is computed by descending the implicit node centres, the paths are
// sorted, and each level's nodes are the distinct path prefixes of that
// length.  Breadth-first order visits a level's nodes by (parent order, child
// slot), which is exactly the lexicographic order of their prefixes, so the
// numbering follows directly:
//   root = T-1; internal nodes count down from T-2 level by level in prefix
//   order; leaves keep their point index; children[i][8] = 1 << (depth + 1)
//   (1 for leaves); centers are the integer node centres.
// O(N log N) in total, against the reference's per-point tensor slicing.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include "common.h"

namespace nsvf {
namespace {

struct OctreeResult {
  int T = 0;
  std::vector<int32_t> centers;
  std::vector<int32_t> children;
};

}  // namespace
}  // namespace nsvf

using namespace nsvf;

extern "C" {

int nsvf_build_octree(const int64_t* center, const int64_t* points, int N,
                      int depth, void** handle, int* num_nodes) {
  clear_error();
  if (!handle || !num_nodes || !center || (N > 0 && !points))
    return set_error(NSVF_EINVAL, "build_octree: null pointer");
  *handle = nullptr;
  *num_nodes = 0;
  if (N < 0) return set_error(NSVF_EINVAL, "build_octree: negative N");
  if (depth < 0 || depth > 19)
    return set_error(NSVF_EINVAL, "build_octree: depth %d outside [0, 19]", depth);
  const int L = depth + 1;  // octant digits per path: D internal levels + leaf slot
  OctreeResult* res = new (std::nothrow) OctreeResult;
  if (!res) return set_error(NSVF_ENOMEM, "build_octree: out of memory");
  try {
    // Octant path of every point: digit i (most significant first) is the
    // octant x + 2y + 4z, x = (p > c), at the node of level i.
    std::vector<uint64_t> path(static_cast<size_t>(N));
    for (int k = 0; k < N; ++k) {
      int64_t c[3] = {center[0], center[1], center[2]};
      const int64_t* p = points + size_t(k) * 3;
      uint64_t key = 0;
      for (int lvl = 0; lvl < L; ++lvl) {
        const int d = depth - lvl;  // depth of the node at this level
        int bits[3];
        for (int a = 0; a < 3; ++a) bits[a] = p[a] > c[a] ? 1 : 0;
        key = (key << 3) | uint64_t(bits[0] + 2 * bits[1] + 4 * bits[2]);
        if (d > 0) {
          const int64_t len = int64_t(1) << (d - 1);
          for (int a = 0; a < 3; ++a) c[a] += (2 * bits[a] - 1) * len;
        }
      }
      path[size_t(k)] = key;
    }
    std::vector<int32_t> order(static_cast<size_t>(N));
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return path[size_t(a)] < path[size_t(b)]; });
    for (int i = 1; i < N; ++i) {
      if (path[size_t(order[i])] == path[size_t(order[i - 1])]) {
        delete res;
        return set_error(NSVF_EINVAL,
                         "build_octree: points %d and %d fall into the same leaf slot",
                         int(order[i - 1]), int(order[i]));
      }
    }
    // Distinct prefixes per level (level 0 = root).  A level-l prefix has l
    // digits: path >> 3*(L - l).
    std::vector<std::vector<uint64_t>> prefixes(static_cast<size_t>(L));
    for (int lvl = 0; lvl < L; ++lvl) {
      const int shift = 3 * (L - lvl);
      auto& v = prefixes[size_t(lvl)];
      for (int i = 0; i < N; ++i) {
        const uint64_t q = shift >= 64 ? 0 : (path[size_t(order[i])] >> shift);
        if (v.empty() || v.back() != q) v.push_back(q);
      }
      if (lvl == 0 && v.empty()) v.push_back(0);  // the root exists with no points
    }
    int64_t internal = 0;
    for (auto& v : prefixes) internal += int64_t(v.size());
    const int64_t T = internal + N;
    if (T > INT32_MAX / 9) {
      delete res;
      return set_error(NSVF_EINVAL, "build_octree: tree too large");
    }
    res->T = int(T);
    res->centers.assign(size_t(T) * 3, 0);
    res->children.assign(size_t(T) * 9, -1);
    // Node ids: level by level from the root, counting down from T-1.
    std::vector<std::vector<int32_t>> ids(static_cast<size_t>(L));
    int32_t next = int32_t(T - 1);
    for (int lvl = 0; lvl < L; ++lvl) {
      auto& v = ids[size_t(lvl)];
      v.resize(prefixes[size_t(lvl)].size());
      for (size_t i = 0; i < v.size(); ++i) v[i] = next--;
    }
    // Centres and scales of internal nodes, and links between levels.
    for (int lvl = 0; lvl < L; ++lvl) {
      const auto& pre = prefixes[size_t(lvl)];
      const auto& id = ids[size_t(lvl)];
      const int d = depth - lvl;
      for (size_t i = 0; i < pre.size(); ++i) {
        int64_t c[3] = {center[0], center[1], center[2]};
        for (int j = 0; j < lvl; ++j) {
          const int oct = int((pre[i] >> (3 * (lvl - 1 - j))) & 7);
          const int64_t len = int64_t(1) << (depth - j - 1);
          c[0] += (2 * (oct & 1) - 1) * len;
          c[1] += (2 * ((oct >> 1) & 1) - 1) * len;
          c[2] += (2 * ((oct >> 2) & 1) - 1) * len;
        }
        const size_t row = size_t(id[i]);
        for (int a = 0; a < 3; ++a) res->centers[row * 3 + a] = int32_t(c[a]);
        res->children[row * 9 + 8] = int32_t(1) << (d + 1);
      }
      if (lvl + 1 < L) {
        // Children of level lvl: the level lvl+1 prefixes, whose parent is
        // prefix >> 3; both lists are sorted, so walk them together.
        const auto& cpre = prefixes[size_t(lvl + 1)];
        const auto& cid = ids[size_t(lvl + 1)];
        size_t pi = 0;
        for (size_t i = 0; i < cpre.size(); ++i) {
          const uint64_t parent = cpre[i] >> 3;
          while (pre[pi] != parent) ++pi;
          res->children[size_t(id[pi]) * 9 + (cpre[i] & 7)] = cid[i];
        }
      }
    }
    // Leaves: point k is the child (slot = last digit) of its depth-0 node.
    {
      const auto& pre = prefixes[size_t(L - 1)];
      const auto& id = ids[size_t(L - 1)];
      size_t pi = 0;
      for (int i = 0; i < N; ++i) {
        const int32_t k = order[i];
        const uint64_t parent = path[size_t(k)] >> 3;
        while (pre[pi] != parent) ++pi;
        res->children[size_t(id[pi]) * 9 + (path[size_t(k)] & 7)] = k;
        for (int a = 0; a < 3; ++a)
          res->centers[size_t(k) * 3 + a] = int32_t(points[size_t(k) * 3 + a]);
        res->children[size_t(k) * 9 + 8] = 1;
      }
    }
  } catch (const std::bad_alloc&) {
    delete res;
    return set_error(NSVF_ENOMEM, "build_octree: out of memory");
  }
  *handle = res;
  *num_nodes = res->T;
  return NSVF_OK;
}

int nsvf_octree_take(void* handle, int32_t* centers, int32_t* children) {
  clear_error();
  if (!handle) return set_error(NSVF_EINVAL, "octree_take: null handle");
  OctreeResult* res = static_cast<OctreeResult*>(handle);
  if (centers) std::memcpy(centers, res->centers.data(), res->centers.size() * sizeof(int32_t));
  if (children) std::memcpy(children, res->children.data(), res->children.size() * sizeof(int32_t));
  delete res;
  return NSVF_OK;
}

}  // extern "C"
