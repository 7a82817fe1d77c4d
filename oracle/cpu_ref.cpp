// CPU restatement of NSVF's native ray-voxel ops, in FP32.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle: tests, the smoke check
// and the CPU-baseline leg of the benchmark call it to check or time the HIP
// library.  No product code path links or calls it.
//
// Every function follows the reference algorithm step by step, including the
// order of floating-point operations, so that an FP32 GPU kernel written to
// the same rules agrees bit for bit on indices and sample counts:
//
//   ray_aabb()                 fairnr/clib/src/intersect_gpu.cu:73-122
//   orc_aabb_intersect()       fairnr/clib/src/intersect_gpu.cu:125-167, intersect.cpp:49-75
//   orc_svo_intersect()        fairnr/clib/src/intersect_gpu.cu:170-237, intersect.cpp:84-112
//   orc_octree_*()             fairnr/clib/src/octree.cpp:13-136
//   orc_uniform_ray_sampling() fairnr/clib/src/sample_gpu.cu:15-106, sample.cpp:23-55
//   orc_inverse_cdf_sampling() fairnr/clib/src/sample_gpu.cu:108-202, sample.cpp:58-95
//
// Arithmetic rules (shared with the HIP kernels, which are built with the same
// rules):
//   * compiled with -ffp-contract=off: no implicit FMA.  The two sites where
//     the reference's nvcc build would contract a*b+c (sample_gpu.cu:63 and
//     :180) use an explicit fmaf() here and in the kernels.
//   * the slab test uses the IEEE reciprocal 1.0f/d where the reference calls
//     the approximate __fdividef (intersect_gpu.cu:86-90).
//   * `1.0 / steps[j]` is evaluated in double and rounded to float, as the
//     reference's double literal does (sample_gpu.cu:147).
//
// Undefined behaviour in the reference gets a defined result, reported through
// the `flags` counters:
//   * reads of pts_idx before the first or after the last element of the batch
//     buffer (sample_gpu.cu:73 with umin==0, :189 on the last ray) read -1;
//   * writes at slot s >= max_steps (no bound check at sample_gpu.cu:67-79,
//     :160-183, :189-191) are dropped;
//   * reads of uniform_noise past the end of the batch buffer read 0.5;
//   * the uniform merge loop stops when no branch applies (NaN depths), where
//     the reference would spin forever (sample_gpu.cu:48-81);
//   * an SVO stack deeper than 256 entries stops the ray (intersect_gpu.cu:209
//     asserts instead).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <queue>
#include <vector>

namespace {

// Slab test of one ray against one axis-aligned box.  Same comparisons,
// ternaries and NaN behaviour as intersect_gpu.cu:73-122: an axis whose
// bounds come out NaN fails every comparison and is skipped.
// (Templated on the arithmetic type: float is the restatement, double the
// FP64 check of which rays graze a face.)
template <class T>
inline bool ray_aabb(const T o[3], const T inv[3], const T c[3], T h, T* out_lo, T* out_hi) {
  T f_low = T(0);
  T f_high = T(100000);
  for (int a = 0; a < 3; ++a) {
    T lo = (c[a] - h - o[a]) * inv[a];
    T hi = (c[a] + h - o[a]) * inv[a];
    if (hi < lo) {
      T t = lo;
      lo = hi;
      hi = t;
    }
    if (hi < f_low) return false;
    if (lo > f_high) return false;
    f_low = (lo > f_low) ? lo : f_low;
    f_high = (hi < f_high) ? hi : f_high;
    if (f_low > f_high) return false;
  }
  *out_lo = f_low;
  *out_hi = f_high;
  return true;
}

struct Node {
  int64_t center[3];
  int depth;
  int index;
  Node* child[8];
};

struct Octree {
  std::vector<Node*> pool;
  Node* root = nullptr;
  int total = 0;
  int terminal = 0;
  std::vector<int32_t> centers;   // [T,3]
  std::vector<int32_t> children;  // [T,9]
  ~Octree() {
    for (Node* n : pool) delete n;
  }
  Node* make(const int64_t c[3], int depth, int index) {
    Node* n = new Node;
    for (int a = 0; a < 3; ++a) n->center[a] = c[a];
    n->depth = depth;
    n->index = index;
    for (int i = 0; i < 8; ++i) n->child[i] = nullptr;
    pool.push_back(n);
    return n;
  }
};

// octree.cpp:62-77.  The octant is x + 2y + 4z with x = (p > c).  Depth-0
// nodes hold leaves directly; a second point in the same octant of a depth-0
// node replaces the first, as the reference's pointer overwrite does.
void octree_insert(Octree& t, Node* p, const int64_t pt[3], int index) {
  for (;;) {
    int d[3];
    for (int a = 0; a < 3; ++a) d[a] = pt[a] > p->center[a] ? 1 : 0;
    int oct = d[0] + 2 * d[1] + 4 * d[2];
    if (p->depth == 0) {
      p->child[oct] = t.make(pt, -1, index);
      return;
    }
    if (p->child[oct] == nullptr) {
      int64_t len = int64_t(1) << (p->depth - 1);
      int64_t c[3];
      for (int a = 0; a < 3; ++a) c[a] = p->center[a] + (2 * d[a] - 1) * len;
      p->child[oct] = t.make(c, p->depth - 1, -1);
    }
    p = p->child[oct];
  }
}

// octree.cpp:79-91 (iterative): total nodes and leaves reachable from root.
void octree_count(Node* root, int* total, int* terminal) {
  int tot = 0, term = 0;
  std::vector<Node*> st{root};
  while (!st.empty()) {
    Node* n = st.back();
    st.pop_back();
    ++tot;
    if (n->depth == -1) ++term;
    for (int i = 0; i < 8; ++i)
      if (n->child[i]) st.push_back(n->child[i]);
  }
  *total = tot;
  *terminal = term;
}

// octree.cpp:93-123.  Breadth-first numbering: the root is T-1, internal
// nodes count down from there in BFS order, leaves keep their point index.
void octree_finalize(Octree& t) {
  octree_count(t.root, &t.total, &t.terminal);
  int T = t.total;
  t.centers.assign(size_t(T) * 3, 0);
  t.children.assign(size_t(T) * 9, -1);
  int node_idx = T - 1;
  t.root->index = node_idx;
  std::queue<Node*> q;
  q.push(t.root);
  while (!q.empty()) {
    Node* n = q.front();
    q.pop();
    for (int i = 0; i < 8; ++i) {
      Node* c = n->child[i];
      if (!c) continue;
      if (c->depth > -1) c->index = --node_idx;
      q.push(c);
      t.children[size_t(n->index) * 9 + i] = c->index;
    }
    t.children[size_t(n->index) * 9 + 8] = 1 << (n->depth + 1);
    for (int a = 0; a < 3; ++a)
      t.centers[size_t(n->index) * 3 + a] = int32_t(n->center[a]);
  }
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// aabb_intersect: brute force over all voxels, first n_max hits in voxel-index
// order.  Outputs are [B,M,n_max]; idx is -1 padded and the depths keep the
// caller's zero fill (intersect.cpp:64-69, intersect_gpu.cu:148-150).
// FP64 brute-force slab test of float32 rays against float32 box centres
// (the same comparisons in double): the first n_max hits of each ray in
// point order, for stating which FP32 hit sets differ from exact
// arithmetic.  Returns the largest hit count of a ray (may exceed n_max).
int orc_aabb_intersect_f64(int M, int N, const float* ray_start, const float* ray_dir,
                           const float* points, double half_voxel, int n_max, int32_t* idx,
                           double* min_depth, double* max_depth) {
  int most = 0;
  for (int r = 0; r < M; ++r) {
    const double o[3] = {ray_start[3 * r], ray_start[3 * r + 1], ray_start[3 * r + 2]};
    const double inv[3] = {1.0 / double(ray_dir[3 * r]), 1.0 / double(ray_dir[3 * r + 1]),
                           1.0 / double(ray_dir[3 * r + 2])};
    int k = 0;
    for (int n = 0; n < N; ++n) {
      const double c[3] = {points[3 * n], points[3 * n + 1], points[3 * n + 2]};
      double lo, hi;
      if (!ray_aabb<double>(o, inv, c, half_voxel, &lo, &hi)) continue;
      if (k < n_max) {
        idx[size_t(r) * n_max + k] = n;
        min_depth[size_t(r) * n_max + k] = lo;
        max_depth[size_t(r) * n_max + k] = hi;
      }
      ++k;
    }
    for (int j = k; j < n_max; ++j) {
      idx[size_t(r) * n_max + j] = -1;
      min_depth[size_t(r) * n_max + j] = 0.0;
      max_depth[size_t(r) * n_max + j] = 0.0;
    }
    if (k > most) most = k;
  }
  return most;
}

void orc_aabb_intersect(int B, int M, int N, const float* ray_start,
                        const float* ray_dir, const float* points,
                        float voxelsize, int n_max, int32_t* idx,
                        float* min_depth, float* max_depth) {
  const float half_voxel = voxelsize * 0.5f;
  for (int b = 0; b < B; ++b) {
    const float* pts = points + size_t(b) * N * 3;
    for (int j = 0; j < M; ++j) {
      size_t r = size_t(b) * M + j;
      const float* o = ray_start + r * 3;
      const float* d = ray_dir + r * 3;
      float inv[3] = {1.0f / d[0], 1.0f / d[1], 1.0f / d[2]};
      int32_t* oi = idx + r * n_max;
      float* omin = min_depth + r * n_max;
      float* omax = max_depth + r * n_max;
      for (int l = 0; l < n_max; ++l) oi[l] = -1;
      int cnt = 0;
      for (int k = 0; k < N && cnt < n_max; ++k) {
        float lo, hi;
        if (ray_aabb(o, inv, pts + size_t(k) * 3, half_voxel, &lo, &hi)) {
          oi[cnt] = k;
          omin[cnt] = lo;
          omax[cnt] = hi;
          ++cnt;
        }
      }
    }
  }
}

// svo_intersect: depth-first traversal from the root (the last node).  A
// node's box half-size is half_voxel * children[k][8].  Children are pushed in
// slot order 0..7, so slot 7 is visited first.  Leaves ([8] == 1) are appended
// in visit order until n_max.  Returns the number of rays whose stack would
// have exceeded 256 entries (those rays stop early).
int orc_svo_intersect(int B, int M, int T, const float* ray_start,
                      const float* ray_dir, const float* points,
                      const int32_t* children, float voxelsize, int n_max,
                      int32_t* idx, float* min_depth, float* max_depth) {
  const float half_voxel = voxelsize * 0.5f;
  int overflow = 0;
  int32_t stack[256];
  for (int b = 0; b < B; ++b) {
    const float* pts = points + size_t(b) * T * 3;
    const int32_t* ch = children + size_t(b) * T * 9;
    for (int j = 0; j < M; ++j) {
      size_t r = size_t(b) * M + j;
      const float* o = ray_start + r * 3;
      const float* d = ray_dir + r * 3;
      float inv[3] = {1.0f / d[0], 1.0f / d[1], 1.0f / d[2]};
      int32_t* oi = idx + r * n_max;
      float* omin = min_depth + r * n_max;
      float* omax = max_depth + r * n_max;
      for (int l = 0; l < n_max; ++l) oi[l] = -1;
      int ptr = 0, cnt = 0;
      stack[0] = T - 1;
      bool ovf = false;
      while (ptr > -1 && cnt < n_max) {
        int k = stack[ptr--];
        const int32_t* nk = ch + size_t(k) * 9;
        float lo, hi;
        if (!ray_aabb(o, inv, pts + size_t(k) * 3, half_voxel * float(nk[8]), &lo, &hi))
          continue;
        if (nk[8] == 1) {
          oi[cnt] = k;
          omin[cnt] = lo;
          omax[cnt] = hi;
          ++cnt;
          continue;
        }
        for (int u = 0; u < 8; ++u) {
          if (nk[u] > -1) {
            if (ptr + 1 >= 256) {
              ovf = true;
              break;
            }
            stack[++ptr] = nk[u];
          }
        }
        if (ovf) break;
      }
      overflow += ovf ? 1 : 0;
    }
  }
  return overflow;
}

// ---------------------------------------------------------------------------
// build_octree, as a handle: build, read the node count, copy out, free.
void* orc_octree_build(const int64_t* center, const int64_t* points, int N,
                       int depth) {
  Octree* t = new Octree;
  t->root = t->make(center, depth, -1);
  for (int k = 0; k < N; ++k) octree_insert(*t, t->root, points + size_t(k) * 3, k);
  octree_finalize(*t);
  return t;
}
int orc_octree_num_nodes(void* h) { return static_cast<Octree*>(h)->total; }
int orc_octree_num_leaves(void* h) { return static_cast<Octree*>(h)->terminal; }
void orc_octree_copy(void* h, int32_t* centers, int32_t* children) {
  Octree* t = static_cast<Octree*>(h);
  std::memcpy(centers, t->centers.data(), t->centers.size() * sizeof(int32_t));
  std::memcpy(children, t->children.data(), t->children.size() * sizeof(int32_t));
}
void orc_octree_free(void* h) { delete static_cast<Octree*>(h); }

// ---------------------------------------------------------------------------
// inverse_cdf_sampling.  Inputs [B,R,P] (steps [B,R], noise [B,R,S]); outputs
// [B,R,S] pre-filled by the caller with idx -1, depth 0, dists 0
// (sample.cpp:81-90).  `compat` = 1 keeps the reference's trailing-loop
// behaviour: the loop condition ignores `done` (sample_gpu.cu:188), so after
// the bins run out one more tail sample is written with the index stored one
// slot past the ray's last bin (the next ray's first index when the row is
// full), and the loop's stop test reads slot `curr_bin` of ray 0 of the batch
// row instead of this ray (sample_gpu.cu:194).  `compat` = 0 applies the
// evident intent: no trailing loop after `done`, and the stop test reads this
// ray's own slot.
//
// flags[0] += reads of pts_idx outside the batch buffer (defined as -1)
// flags[1] += dropped writes with s >= max_steps
// flags[2] += reads of noise outside the batch buffer (defined as 0.5)
void orc_inverse_cdf_sampling(int B, int R, int P, int S, float fixed_step_size,
                              const int32_t* pts_idx_all,
                              const float* min_depth_all,
                              const float* max_depth_all,
                              const float* noise_all, const float* probs_all,
                              const float* steps_all, int compat,
                              int32_t* out_idx, float* out_depth,
                              float* out_dists, int64_t* flags) {
  const size_t n_idx = size_t(B) * R * P;
  const size_t n_noise = size_t(B) * R * S;
  for (int b = 0; b < B; ++b) {
    const size_t base_h = size_t(b) * R * P;
    const size_t base_k = size_t(b) * R * S;
    const int32_t* pts_idx = pts_idx_all + base_h;
    const float* min_depth = min_depth_all + base_h;
    const float* max_depth = max_depth_all + base_h;
    const float* probs = probs_all + base_h;
    const float* steps = steps_all + size_t(b) * R;
    for (int j = 0; j < R; ++j) {
      const size_t H = size_t(j) * P, K = size_t(j) * S;
      auto idx_at = [&](size_t off) -> int32_t {  // off relative to batch base
        size_t g = base_h + off;
        if (g >= n_idx) {
          flags[0]++;
          return -1;
        }
        return pts_idx_all[g];
      };
      auto noise_at = [&](size_t off) -> float {
        size_t g = base_k + off;
        if (g >= n_noise) {
          flags[2]++;
          return 0.5f;
        }
        return noise_all[g];
      };
      auto emit = [&](int s, int32_t id, float dist, float depth) {
        if (s >= S) {
          flags[1]++;
          return;
        }
        out_idx[base_k + K + s] = id;
        out_dists[base_k + K + s] = dist;
        out_depth[base_k + K + s] = depth;
      };
      int curr_bin = 0, s = 0;
      float curr_min_depth = min_depth[H];
      float curr_max_depth = max_depth[H];
      float curr_min_cdf = 0.0f;
      float curr_max_cdf = probs[H];
      float step_size = float(1.0 / double(steps[j]));
      float z_low = curr_min_depth;
      int total_steps = int(std::ceil(steps[j]));
      bool done = false;
      if (fixed_step_size > 0.0f) step_size = fixed_step_size;

      for (int curr_step = 0; curr_step < total_steps; ++curr_step) {
        if (s >= S) break;  // every later write would be dropped
        float curr_cdf = (float(curr_step) + noise_at(K + curr_step)) * step_size;
        while (curr_cdf > curr_max_cdf) {
          emit(s, pts_idx[H + curr_bin], curr_max_depth - z_low,
               (curr_max_depth + z_low) * 0.5f);
          ++curr_bin;
          ++s;
          if (curr_bin >= P || pts_idx[H + curr_bin] == -1) {
            done = true;
            break;
          }
          curr_min_depth = min_depth[H + curr_bin];
          curr_max_depth = max_depth[H + curr_bin];
          curr_min_cdf = curr_max_cdf;
          curr_max_cdf = curr_max_cdf + probs[H + curr_bin];
          z_low = curr_min_depth;
        }
        if (done) break;
        float u = (curr_cdf - curr_min_cdf) / (curr_max_cdf - curr_min_cdf);
        float z = std::fmaf(u, curr_max_depth - curr_min_depth, curr_min_depth);
        emit(s, pts_idx[H + curr_bin], z - z_low, (z + z_low) * 0.5f);
        z_low = z;
        ++s;
      }
      if (done && !compat) continue;
      while (z_low < curr_max_depth) {
        emit(s, idx_at(H + curr_bin), curr_max_depth - z_low,
             (curr_max_depth + z_low) * 0.5f);
        ++curr_bin;
        ++s;
        if (curr_bin >= P) break;
        int32_t stop_probe = compat ? pts_idx[curr_bin] : pts_idx[H + curr_bin];
        if (stop_probe == -1) break;
        curr_min_depth = min_depth[H + curr_bin];
        curr_max_depth = max_depth[H + curr_bin];
        z_low = curr_min_depth;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// uniform_ray_sampling.  Inputs [B,R,P], noise [B,R,S]; outputs [B,R,S]
// pre-filled by the caller with idx -1, depth 0, dists 0 (sample.cpp:40-48).
// Phase 1 merges voxel exits, uniform steps from the first entry and voxel
// entries (sample_gpu.cu:48-81); phase 2 turns the merged depths into
// midpoints and gaps in place and keeps the in-voxel intervals
// (sample_gpu.cu:83-100); phase 3 marks the tail with -1 (:102-104).  Depth
// and dist values beyond the kept samples stay as phase 1/2 left them, as in
// the reference.
//
// flags[0] += pts_idx reads outside the batch buffer (defined as -1)
// flags[1] += dropped writes with s >= max_steps
// flags[2] += merge loops stopped because no branch applied (NaN input)
void orc_uniform_ray_sampling(int B, int R, int P, int S, float step_size,
                              const int32_t* pts_idx_all,
                              const float* min_depth_all,
                              const float* max_depth_all,
                              const float* noise_all, int32_t* out_idx,
                              float* out_depth, float* out_dists,
                              int64_t* flags) {
  const size_t n_idx = size_t(B) * R * P;
  for (int b = 0; b < B; ++b) {
    const size_t base_h = size_t(b) * R * P;
    const size_t base_k = size_t(b) * R * S;
    const int32_t* pts_idx = pts_idx_all + base_h;
    const float* min_depth = min_depth_all + base_h;
    const float* max_depth = max_depth_all + base_h;
    const float* noise = noise_all + base_k;
    int32_t* sidx = out_idx + base_k;
    float* sdep = out_depth + base_k;
    float* sdis = out_dists + base_k;
    for (int j = 0; j < R; ++j) {
      const int64_t H = int64_t(j) * P, K = int64_t(j) * S;
      auto idx_at = [&](int64_t off) -> int32_t {
        int64_t g = int64_t(base_h) + off;
        if (g < 0 || size_t(g) >= n_idx) {
          flags[0]++;
          return -1;
        }
        return pts_idx_all[g];
      };
      int s = 0, ucur = 0, umin = 0, umax = 0;
      float last_min_depth, last_max_depth, curr_depth = 0.0f;
      while (true) {
        if (umax == P || ucur == S || pts_idx[H + umax] == -1) break;
        if (s >= S) break;  // every later write would be dropped
        last_min_depth = (umin < P) ? min_depth[H + umin] : 10000.0f;
        last_max_depth = (umax < P) ? max_depth[H + umax] : 10000.0f;
        if (ucur < S)
          curr_depth = std::fmaf(float(ucur) + noise[K + ucur], step_size, min_depth[H]);
        if (last_max_depth <= curr_depth && last_max_depth <= last_min_depth) {
          if (s < S) {
            sdep[K + s] = last_max_depth;
            sidx[K + s] = pts_idx[H + umax];
          } else {
            flags[1]++;
          }
          umax++;
          s++;
          continue;
        }
        if (curr_depth <= last_min_depth && curr_depth <= last_max_depth) {
          if (s < S) {
            sdep[K + s] = curr_depth;
            sidx[K + s] = idx_at(H + umin - 1);
          } else {
            flags[1]++;
          }
          ucur++;
          s++;
          continue;
        }
        if (last_min_depth <= curr_depth && last_min_depth <= last_max_depth) {
          if (s < S) {
            sdep[K + s] = last_min_depth;
            sidx[K + s] = pts_idx[H + umin];
          } else {
            flags[1]++;
          }
          umin++;
          s++;
          continue;
        }
        flags[2]++;
        break;
      }

      int step = 0;
      umin = 0;
      umax = 0;
      for (ucur = 0; ucur < S - 1; ucur++) {
        if (sidx[K + ucur + 1] == -1) break;
        float l_depth = sdep[K + ucur];
        float r_depth = sdep[K + ucur + 1];
        sdep[K + ucur] = (l_depth + r_depth) * 0.5f;
        sdis[K + ucur] = r_depth - l_depth;
        if (umin < P && sdep[K + ucur] >= min_depth[H + umin] && pts_idx[H + umin] > -1) umin++;
        if (umax < P && sdep[K + ucur] >= max_depth[H + umax] && pts_idx[H + umax] > -1) umax++;
        if (umax == P || pts_idx[H + umax] == -1) break;
        if (umin - 1 == umax && sdis[K + ucur] > 0) {
          sdep[K + step] = sdep[K + ucur];
          sdis[K + step] = sdis[K + ucur];
          sidx[K + step] = sidx[K + ucur];
          step++;
        }
      }
      for (int t = step; t < S; t++) sidx[K + t] = -1;
    }
  }
}

}  // extern "C"
