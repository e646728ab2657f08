// scene_host.cpp -- host-side scene build (see scene_host.h).
#include "scene_host.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <exception>
#include <functional>
#include <limits>
#include <mutex>
#include <numeric>
#include <system_error>
#include <thread>

namespace tpt {

// Runs the tasks on up to 16 host threads (the calling thread among them), largest first as
// given; the first exception a task throws is rethrown here once every task has finished.  A
// thread that cannot be started leaves its share to the others.
static void run_parallel(const std::vector<std::function<void()>>& tasks) {
  const size_t n = tasks.size();
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t nt = std::min<size_t>(n, std::min(hw, 16u));
  std::atomic<size_t> next{0};
  std::exception_ptr err;
  std::mutex mu;
  auto work = [&] {
    for (size_t i; (i = next.fetch_add(1)) < n;) {
      try {
        tasks[i]();
      } catch (...) {
        std::lock_guard<std::mutex> lock(mu);
        if (!err) err = std::current_exception();
      }
    }
  };
  std::vector<std::thread> pool;
  for (size_t t = 1; t < nt; t++) {
    try {
      pool.emplace_back(work);
    } catch (const std::system_error&) {
      break;
    }
  }
  work();
  for (std::thread& th : pool) th.join();
  if (err) std::rethrow_exception(err);
}

static const double kPi = 3.14159265358979323846;
static const double kInf = std::numeric_limits<double>::infinity();

// ------------------------------------------------------------------ f64 helpers
static inline double fmx(double a, double b) { return std::fmax(a, b); }
static inline double fmn(double a, double b) { return std::fmin(a, b); }
static inline size_t sat_usize(double v) {  // Rust `as usize`
  if (!(v > 0.0)) return 0;
  if (v >= 18446744073709551615.0) return SIZE_MAX;
  return (size_t)v;
}

static DAabb empty_box() {
  DAabb b;
  for (int k = 0; k < 3; k++) { b.mn[k] = kInf; b.mx[k] = -kInf; }
  return b;
}
static DAabb unite(const DAabb& a, const DAabb& b) {
  DAabb r;
  for (int k = 0; k < 3; k++) { r.mn[k] = fmn(a.mn[k], b.mn[k]); r.mx[k] = fmx(a.mx[k], b.mx[k]); }
  return r;
}
static DAabb grow(const DAabb& a, const double* p) {
  DAabb r;
  for (int k = 0; k < 3; k++) { r.mn[k] = fmn(a.mn[k], p[k]); r.mx[k] = fmx(a.mx[k], p[k]); }
  return r;
}

// Aabb::from_points (geometry.rs:70-80) and Triangle::bounds (geometry.rs:221-225): for a
// static triangle the shutter-open and shutter-close boxes are equal, so their union is one.
DAabb triangle_bounds(const double* v) {
  DAabb b = grow(grow(grow(empty_box(), v), v + 3), v + 6);
  const double eps = 1e-9;
  for (int k = 0; k < 3; k++) { b.mn[k] = b.mn[k] - eps; b.mx[k] = b.mx[k] + eps; }
  return b;
}

// ------------------------------------------------------------------ reference BVH build
// The steps of Bvh::build (bvh.rs:51-242) over primitive boxes: pre-order node array, leaf if
// n <= leaf_size, split axis = largest centroid extent, 8-bin SAH when n >= 16 with a stable
// sort of the range by bin, otherwise the range is cut at n/2 in its current order.  The
// arithmetic is the reference's f64 arithmetic so the node array comes out identical.
namespace {
struct RefBuilder {
  // A primitive's box, centroid and index in the order of out.indices: the builder moves these
  // records, so each pass over a node's range reads memory in sequence.
  struct Prim {
    DAabb box;
    double c[3];
    int64_t idx;
  };
  std::vector<Prim> rec;
  std::vector<Prim> tmp;         // scatter target of the bin sort
  std::vector<uint8_t> bin_buf;  // bins of the current range
  size_t leaf_size;
  RefBvh out;
  static const size_t BINS = 8;

  RefBuilder(const std::vector<DAabb>& b, size_t ls) : leaf_size(ls) {
    rec.resize(b.size());
    for (size_t i = 0; i < b.size(); i++) {
      rec[i].box = b[i];
      for (int k = 0; k < 3; k++) rec[i].c[k] = (b[i].mn[k] + b[i].mx[k]) * 0.5;  // Aabb::centroid
      rec[i].idx = (int64_t)i;
    }
  }

  int32_t node(size_t start, size_t end, int depth) {
    out.depth = std::max(out.depth, depth);
    size_t n = end - start;
    DAabb bbox = empty_box(), cbox = empty_box();
    for (size_t k = start; k < end; k++) {
      bbox = unite(bbox, rec[k].box);
      cbox = grow(cbox, rec[k].c);
    }
    int32_t idx = (int32_t)out.nodes.size();
    out.nodes.push_back(RefNode{bbox, -1, -1, 0, 0});
    if (n <= leaf_size) {
      out.nodes[(size_t)idx].start = (uint32_t)start;
      out.nodes[(size_t)idx].count = (uint32_t)n;
      return idx;
    }
    double ex = cbox.mx[0] - cbox.mn[0], ey = cbox.mx[1] - cbox.mn[1], ez = cbox.mx[2] - cbox.mn[2];
    int axis = (ex >= ey && ex >= ez) ? 0 : (ey >= ez ? 1 : 2);
    size_t mid = start + n / 2;
    if (n >= BINS * 2) {
      double minc = cbox.mn[axis], maxc = cbox.mx[axis];
      double extent = maxc - minc;
      if (extent > 1e-12) {
        double inv_extent = 1.0 / extent;
        auto bin_of = [&](const Prim& q) {
          size_t b = sat_usize((q.c[axis] - minc) * inv_extent * (double)BINS);
          return b >= BINS ? BINS - 1 : b;
        };
        DAabb bb[BINS];
        size_t bc[BINS];
        for (size_t b = 0; b < BINS; b++) { bb[b] = empty_box(); bc[b] = 0; }
        bin_buf.resize(n);
        for (size_t k = start; k < end; k++) {
          const size_t b = bin_of(rec[k]);
          bin_buf[k - start] = (uint8_t)b;
          bb[b] = unite(bb[b], rec[k].box);
          bc[b] += 1;
        }
        DAabb lb[BINS], rb[BINS];
        size_t lc[BINS], rc[BINS];
        DAabb acc = empty_box();
        size_t an = 0;
        for (size_t b = 0; b < BINS; b++) {
          if (bc[b] > 0) { acc = unite(acc, bb[b]); an += bc[b]; }
          lb[b] = acc;
          lc[b] = an;
        }
        acc = empty_box();
        an = 0;
        for (size_t b = BINS; b-- > 0;) {
          if (bc[b] > 0) { acc = unite(acc, bb[b]); an += bc[b]; }
          rb[b] = acc;
          rc[b] = an;
        }
        double best_cost = kInf;
        size_t best_split = 0;
        for (size_t b = 0; b + 1 < BINS; b++) {
          if (lc[b] == 0 || rc[b + 1] == 0) continue;
          double lx = lb[b].mx[0] - lb[b].mn[0], ly = lb[b].mx[1] - lb[b].mn[1], lz = lb[b].mx[2] - lb[b].mn[2];
          double rx = rb[b + 1].mx[0] - rb[b + 1].mn[0], ry = rb[b + 1].mx[1] - rb[b + 1].mn[1],
                 rz = rb[b + 1].mx[2] - rb[b + 1].mn[2];
          double la = 2.0 * (lx * ly + ly * lz + lz * lx);
          double ra = 2.0 * (rx * ry + ry * rz + rz * rx);
          double cost = la * (double)lc[b] + ra * (double)rc[b + 1];
          if (cost < best_cost) { best_cost = cost; best_split = b + 1; }
        }
        if (std::isfinite(best_cost)) {
          // the reference's stable sort of the range by bin, as a stable counting sort (the
          // order of a stable sort by a key is the one order that keeps equal keys in place)
          size_t off[BINS];
          for (size_t b = 0, o = 0; b < BINS; b++) { off[b] = o; o += bc[b]; }
          tmp.resize(n);
          for (size_t k = start; k < end; k++) tmp[off[bin_buf[k - start]]++] = rec[k];
          std::copy(tmp.begin(), tmp.begin() + (std::ptrdiff_t)n, rec.begin() + (std::ptrdiff_t)start);
          size_t left_total = 0;
          for (size_t b = 0; b < best_split; b++) left_total += bc[b];
          if (left_total > 0 && left_total < n) mid = start + left_total;
        }
      }
    }
    int32_t l = node(start, mid, depth + 1);
    int32_t r = node(mid, end, depth + 1);
    out.nodes[(size_t)idx].left = l;
    out.nodes[(size_t)idx].right = r;
    return idx;
  }
};
}  // namespace

RefBvh build_ref_bvh(const std::vector<DAabb>& prim_bounds, size_t leaf_size) {
  RefBuilder b(prim_bounds, leaf_size);
  if (!prim_bounds.empty()) b.node(0, prim_bounds.size(), 1);
  b.out.indices.resize(prim_bounds.size());
  for (size_t i = 0; i < b.rec.size(); i++) b.out.indices[i] = b.rec[i].idx;
  return std::move(b.out);
}

// ------------------------------------------------------------------ SAH build for the FP32 trees (own design)
// The FP32 kernels need not walk the reference's tree (the plain FP64 form does): their trees
// come from a binned SAH build over all three axes with a cost-based leaf decision, which keeps
// leaves small where that saves triangle tests (the reference splits every node of fewer than
// 16 primitives at the middle of its input order).  Same RefBvh form, so the FP32 conversion
// below serves both.
namespace {
struct SahBuilder {
  static constexpr int kBins = 32;
  static constexpr double kTraverse = 1.0;   // cost of one binary node visit
  static constexpr double kIntersect = 1.0;  // cost of one triangle test
  // A primitive's box, box centre and index, kept in the order of the node ranges: the builder
  // partitions these records themselves, so every pass over a node's range reads memory in
  // sequence rather than through a permuted index list.
  struct Prim {
    DAabb box;
    double c[3];
    int64_t idx;
  };
  std::vector<Prim> rec;
  size_t max_leaf;
  RefBvh out;
  // Subtrees whose root lies at most this deep are built on threads of their own, each into its
  // own node array, and spliced in pre-order: the node array is the serial build's.
  static constexpr int kParallelDepth = 3;
  static constexpr size_t kParallelMin = 1u << 15;  // primitives worth a thread
  SahBuilder(const std::vector<DAabb>& b, size_t leaf) : max_leaf(leaf) {
    rec.resize(b.size());
    for (size_t i = 0; i < b.size(); i++) {
      rec[i].box = b[i];
      for (int k = 0; k < 3; k++) rec[i].c[k] = 0.5 * (b[i].mn[k] + b[i].mx[k]);
      rec[i].idx = (int64_t)i;
    }
  }
  static double half_area(const DAabb& b) {
    const double x = b.mx[0] - b.mn[0], y = b.mx[1] - b.mn[1], z = b.mx[2] - b.mn[2];
    return x * y + y * z + z * x;
  }
  // Builds the subtree of rec[lo, hi) into `nodes` (pre-order, child indices into `nodes`);
  // `right` is the partition's scratch, `max_depth` the deepest level reached.
  int32_t node(size_t lo, size_t hi, int depth, std::vector<RefNode>& nodes, std::vector<Prim>& right, int& max_depth) {
    max_depth = std::max(max_depth, depth);
    DAabb box = empty_box(), cb = empty_box();
    for (size_t i = lo; i < hi; i++) {
      box = unite(box, rec[i].box);
      cb = grow(cb, rec[i].c);
    }
    const int32_t me = (int32_t)nodes.size();
    nodes.push_back(RefNode{box, -1, -1, (uint32_t)lo, (uint32_t)(hi - lo)});
    const size_t n = hi - lo;
    if (n <= 1) return me;
    // binned SAH over the three axes, the bins of all three filled in one pass
    DAabb bb[3][kBins];
    size_t bc[3][kBins];
    double scale[3];
    bool use[3];
    for (int axis = 0; axis < 3; axis++) {
      const double ext = cb.mx[axis] - cb.mn[axis];
      use[axis] = ext > 0.0;
      scale[axis] = use[axis] ? kBins / ext : 0.0;
      for (int b = 0; b < kBins; b++) { bb[axis][b] = empty_box(); bc[axis][b] = 0; }
    }
    for (size_t i = lo; i < hi; i++) {
      const Prim& q = rec[i];
      for (int axis = 0; axis < 3; axis++) {
        if (!use[axis]) continue;
        int b = (int)((q.c[axis] - cb.mn[axis]) * scale[axis]);
        b = std::min(std::max(b, 0), kBins - 1);
        bb[axis][b] = unite(bb[axis][b], q.box);
        bc[axis][b]++;
      }
    }
    double best = kInf;
    int best_axis = -1, best_bin = 0;
    const double parent = half_area(box);
    for (int axis = 0; axis < 3; axis++) {
      if (!use[axis]) continue;
      double right_area[kBins];
      size_t right_n[kBins];
      DAabb acc = empty_box();
      size_t cnt = 0;
      for (int b = kBins - 1; b > 0; b--) {
        acc = unite(acc, bb[axis][b]);
        cnt += bc[axis][b];
        right_area[b] = half_area(acc);
        right_n[b] = cnt;
      }
      acc = empty_box();
      cnt = 0;
      for (int b = 1; b < kBins; b++) {
        acc = unite(acc, bb[axis][b - 1]);
        cnt += bc[axis][b - 1];
        if (cnt == 0 || right_n[b] == 0) continue;
        const double cost = kTraverse + kIntersect * (half_area(acc) * (double)cnt + right_area[b] * (double)right_n[b]) /
                                            std::max(parent, 1e-300);
        if (cost < best) { best = cost; best_axis = axis; best_bin = b; }
      }
    }
    const double leaf_cost = kIntersect * (double)n;
    if (n <= max_leaf && (best_axis < 0 || leaf_cost <= best)) return me;
    size_t mid;
    if (best_axis < 0) {  // every centre coincides: split the range in two
      mid = lo + n / 2;
    } else {
      // stable partition by the split bin: the left records move down in order, the right
      // ones go through the scratch buffer and follow them
      const double sc = scale[best_axis], c0 = cb.mn[best_axis];
      size_t w = lo, nr = 0;
      right.resize(std::max(right.size(), n));
      for (size_t i = lo; i < hi; i++) {
        int b = (int)((rec[i].c[best_axis] - c0) * sc);
        if (std::min(std::max(b, 0), kBins - 1) < best_bin) rec[w++] = rec[i];
        else right[nr++] = rec[i];
      }
      std::copy(right.begin(), right.begin() + (std::ptrdiff_t)nr, rec.begin() + (std::ptrdiff_t)w);
      mid = w;
      if (mid == lo || mid == hi) mid = lo + n / 2;
    }
    int32_t l, r;
    if (depth <= kParallelDepth && n >= kParallelMin) {
      // the right subtree on a thread of its own (into its own array, spliced after the left one)
      std::vector<RefNode> rnodes;
      std::vector<Prim> rright;
      int rdepth = 0;
      std::exception_ptr rerr;
      std::thread th;
      bool threaded = true;
      try {
        th = std::thread([&] {
          try {
            node(mid, hi, depth + 1, rnodes, rright, rdepth);
          } catch (...) {
            rerr = std::current_exception();
          }
        });
      } catch (const std::system_error&) {
        threaded = false;
      }
      if (!threaded) node(mid, hi, depth + 1, rnodes, rright, rdepth);
      std::exception_ptr lerr;
      try {
        l = node(lo, mid, depth + 1, nodes, right, max_depth);
      } catch (...) {
        lerr = std::current_exception();
      }
      if (threaded) th.join();
      if (lerr) std::rethrow_exception(lerr);
      if (rerr) std::rethrow_exception(rerr);
      const int32_t base = (int32_t)nodes.size();
      for (RefNode& x : rnodes) {
        if (x.left >= 0) x.left += base;
        if (x.right >= 0) x.right += base;
      }
      nodes.insert(nodes.end(), rnodes.begin(), rnodes.end());
      r = base;
      max_depth = std::max(max_depth, rdepth);
    } else {
      l = node(lo, mid, depth + 1, nodes, right, max_depth);
      r = node(mid, hi, depth + 1, nodes, right, max_depth);
    }
    nodes[(size_t)me].left = l;
    nodes[(size_t)me].right = r;
    nodes[(size_t)me].count = 0;
    return me;
  }
};
}  // namespace

// The FP32 triangle record of the vertices v[0..8] (see TriF32).
TriF32 tri_f32(const double* v) {
  double e1[3], e2[3], n[3];
  for (int c = 0; c < 3; c++) {
    e1[c] = v[3 + c] - v[c];
    e2[c] = v[6 + c] - v[c];
  }
  auto cross = [](const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
  };
  cross(e1, e2, n);
  const double nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
  double r0[3], r1[3];
  cross(e2, n, r0);
  cross(n, e1, r1);
  TriF32 t{};
  const bool ok = nn > 0.0 && std::isfinite(nn);
  for (int c = 0; c < 3; c++) {
    t.r0[c] = ok ? (float)(r0[c] / nn) : 0.f;
    t.r1[c] = ok ? (float)(r1[c] / nn) : 0.f;
    t.r2[c] = ok ? (float)(n[c] / nn) : 0.f;
  }
  t.r0[3] = (float)v[0];
  t.r1[3] = (float)v[1];
  t.r2[3] = (float)v[2];
  return t;
}

RefBvh build_sah_bvh(const std::vector<DAabb>& prim_bounds, size_t max_leaf) {
  SahBuilder b(prim_bounds, max_leaf);
  std::vector<SahBuilder::Prim> scratch;
  if (!prim_bounds.empty()) b.node(0, prim_bounds.size(), 1, b.out.nodes, scratch, b.out.depth);
  b.out.indices.resize(prim_bounds.size());
  for (size_t i = 0; i < b.rec.size(); i++) b.out.indices[i] = b.rec[i].idx;
  return std::move(b.out);
}

// ------------------------------------------------------------------ FP32 conversion
static inline float round_down(double x) {
  float f = (float)x;
  if ((double)f > x) f = std::nextafter(f, -std::numeric_limits<float>::infinity());
  return f;
}
static inline float round_up(double x) {
  float f = (float)x;
  if ((double)f < x) f = std::nextafter(f, std::numeric_limits<float>::infinity());
  return f;
}
static void put_box(float* b, const DAabb& x) {
  for (int k = 0; k < 3; k++) { b[k] = round_down(x.mn[k]); b[3 + k] = round_up(x.mx[k]); }
}

// Reference BVH -> Node2 array.  Returns the child ref of `id`.  Leaf slots are offset by
// `slot_base` (position of the mesh's first reordered primitive).
static int32_t to_node2(const RefBvh& r, int32_t id, int64_t slot_base, std::vector<Node2>& out, std::string& err) {
  const RefNode& n = r.nodes[(size_t)id];
  if (n.left == -1 && n.right == -1) {
    int64_t s = slot_base + (int64_t)n.start;
    if (n.count == 0 || n.count > 8 || s > (int64_t)TPT_LEAF_START_MASK) {
      err = "leaf does not fit the node encoding";
      return TPT_REF_EMPTY;
    }
    return (int32_t)(TPT_LEAF_BIT | ((uint32_t)(n.count - 1) << 28) | (uint32_t)s);
  }
  int32_t me = (int32_t)out.size();
  out.push_back(Node2{});
  Node2 tmp{};
  put_box(tmp.b, r.nodes[(size_t)n.left].bbox);
  put_box(tmp.b + 6, r.nodes[(size_t)n.right].bbox);
  tmp.child[0] = to_node2(r, n.left, slot_base, out, err);
  tmp.child[1] = to_node2(r, n.right, slot_base, out, err);
  out[(size_t)me] = tmp;
  return me;
}

// ------------------------------------------------------------------ four-wide collapse (own design)
// A Node2 subtree becomes one Node4 by opening, while fewer than four children are held, the
// inner child with the largest box area; leaves and empty refs pass through unchanged.
static int32_t collapse4(const std::vector<Node2>& n2, int32_t ref, std::vector<Node4>& out) {
  if (ref < 0 || ref == TPT_REF_EMPTY) return ref;
  struct Kid {
    int32_t ref;
    float b[6];
  };
  auto area = [](const float* b) {
    float x = b[3] - b[0], y = b[4] - b[1], z = b[5] - b[2];
    return x * y + y * z + z * x;
  };
  std::vector<Kid> kids;
  const Node2& root = n2[(size_t)ref];
  for (int c = 0; c < 2; c++) {
    Kid k;
    k.ref = root.child[c];
    std::copy(root.b + 6 * c, root.b + 6 * c + 6, k.b);
    kids.push_back(k);
  }
  while (kids.size() < 4) {
    int best = -1;
    float best_a = -1.f;
    for (size_t i = 0; i < kids.size(); i++)
      if (kids[i].ref >= 0 && kids[i].ref != TPT_REF_EMPTY && area(kids[i].b) > best_a) {
        best = (int)i;
        best_a = area(kids[i].b);
      }
    if (best < 0) break;
    const Node2& n = n2[(size_t)kids[(size_t)best].ref];
    Kid a, b;
    a.ref = n.child[0];
    std::copy(n.b, n.b + 6, a.b);
    b.ref = n.child[1];
    std::copy(n.b + 6, n.b + 12, b.b);
    kids[(size_t)best] = a;
    kids.insert(kids.begin() + best + 1, b);
  }
  const int32_t me = (int32_t)out.size();
  out.push_back(Node4{});
  Node4 tmp{};
  for (int i = 0; i < 4; i++) {
    if (i < (int)kids.size()) {
      const Kid& k = kids[(size_t)i];
      tmp.lo_x[i] = k.b[0]; tmp.lo_y[i] = k.b[1]; tmp.lo_z[i] = k.b[2];
      tmp.hi_x[i] = k.b[3]; tmp.hi_y[i] = k.b[4]; tmp.hi_z[i] = k.b[5];
      tmp.child[i] = collapse4(n2, k.ref, out);
    } else {
      tmp.lo_x[i] = tmp.lo_y[i] = tmp.lo_z[i] = 1.f;  // never tested: the traversal skips empty slots
      tmp.hi_x[i] = tmp.hi_y[i] = tmp.hi_z[i] = -1.f;
      tmp.child[i] = TPT_REF_EMPTY;
    }
  }
  out[(size_t)me] = tmp;
  return me;
}

// Worst-case traversal stack entries taken inside a Node4 tree: every child of a node hit, so
// count - 1 far children pushed at each node along the deepest root-to-leaf path.
static int push_need4(const std::vector<Node4>& n4, int32_t ref) {
  if (ref < 0 || ref == TPT_REF_EMPTY) return 0;
  const Node4& n = n4[(size_t)ref];
  int kids = 0, deepest = 0;
  for (int i = 0; i < 4; i++) {
    if (n.child[i] == TPT_REF_EMPTY) continue;
    kids++;
    deepest = std::max(deepest, push_need4(n4, n.child[i]));
  }
  return std::max(kids - 1, 0) + deepest;
}

// ------------------------------------------------------------------ TLAS (own design)
// Full-sweep SAH over item boxes, one item per leaf.  The reference has no top-level tree; it
// visits every instance and then every sphere for each ray (world.rs:86-118).
namespace {
struct TlasBuilder {
  const std::vector<DAabb>& box;
  std::vector<int32_t> order;
  std::vector<Node2>& nodes;
  int depth = 0;
  TlasBuilder(const std::vector<DAabb>& b, std::vector<Node2>& n) : box(b), nodes(n) {}
  static double area(const DAabb& b) {
    double x = b.mx[0] - b.mn[0], y = b.mx[1] - b.mn[1], z = b.mx[2] - b.mn[2];
    return 2.0 * (x * y + y * z + z * x);
  }
  DAabb range_box(size_t lo, size_t hi) const {
    DAabb b = empty_box();
    for (size_t i = lo; i < hi; i++) b = unite(b, box[(size_t)order[i]]);
    return b;
  }
  int32_t build(size_t lo, size_t hi, int d) {
    depth = std::max(depth, d);
    size_t n = hi - lo;
    if (n == 1) return (int32_t)(TPT_LEAF_BIT | (uint32_t)lo);
    double best = kInf;
    int best_axis = 0;
    size_t best_mid = lo + n / 2;
    std::vector<double> right_area(n + 1);
    for (int axis = 0; axis < 3; axis++) {
      std::stable_sort(order.begin() + (std::ptrdiff_t)lo, order.begin() + (std::ptrdiff_t)hi, [&](int32_t a, int32_t b) {
        return box[(size_t)a].mn[axis] + box[(size_t)a].mx[axis] < box[(size_t)b].mn[axis] + box[(size_t)b].mx[axis];
      });
      DAabb acc = empty_box();
      for (size_t k = n; k-- > 1;) {
        acc = unite(acc, box[(size_t)order[lo + k]]);
        right_area[k] = area(acc);
      }
      acc = empty_box();
      for (size_t k = 1; k < n; k++) {
        acc = unite(acc, box[(size_t)order[lo + k - 1]]);
        double cost = area(acc) * (double)k + right_area[k] * (double)(n - k);
        if (cost < best) { best = cost; best_axis = axis; best_mid = lo + k; }
      }
    }
    std::stable_sort(order.begin() + (std::ptrdiff_t)lo, order.begin() + (std::ptrdiff_t)hi, [&](int32_t a, int32_t b) {
      return box[(size_t)a].mn[best_axis] + box[(size_t)a].mx[best_axis] <
             box[(size_t)b].mn[best_axis] + box[(size_t)b].mx[best_axis];
    });
    int32_t me = (int32_t)nodes.size();
    nodes.push_back(Node2{});
    Node2 tmp{};
    put_box(tmp.b, range_box(lo, best_mid));
    put_box(tmp.b + 6, range_box(best_mid, hi));
    tmp.child[0] = build(lo, best_mid, d + 1);
    tmp.child[1] = build(best_mid, hi, d + 1);
    nodes[(size_t)me] = tmp;
    return me;
  }
};
}  // namespace

// ------------------------------------------------------------------ matrices (math.rs:100-170)
static void mat_invert(const double* m, double* r) {
  double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
  double inv_det = 1.0 / det;
  r[0] = (m[4] * m[8] - m[5] * m[7]) * inv_det;
  r[1] = (m[2] * m[7] - m[1] * m[8]) * inv_det;
  r[2] = (m[1] * m[5] - m[2] * m[4]) * inv_det;
  r[3] = (m[5] * m[6] - m[3] * m[8]) * inv_det;
  r[4] = (m[0] * m[8] - m[2] * m[6]) * inv_det;
  r[5] = (m[2] * m[3] - m[0] * m[5]) * inv_det;
  r[6] = (m[3] * m[7] - m[4] * m[6]) * inv_det;
  r[7] = (m[1] * m[6] - m[0] * m[7]) * inv_det;
  r[8] = (m[0] * m[4] - m[1] * m[3]) * inv_det;
}
static void mat_vec(const double* m, const double* v, double* out) {
  out[0] = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
  out[1] = m[3] * v[0] + m[4] * v[1] + m[5] * v[2];
  out[2] = m[6] * v[0] + m[7] * v[1] + m[8] * v[2];
}

static void tree_cut(const RefBvh& t, int32_t node, int depth, int max_depth, std::vector<DAabb>* out) {
  const RefNode& n = t.nodes[(size_t)node];
  if (depth >= max_depth || n.left < 0) {
    out->push_back(n.bbox);
    return;
  }
  tree_cut(t, n.left, depth + 1, max_depth, out);
  tree_cut(t, n.right, depth + 1, max_depth, out);
}

std::vector<DAabb> tree_cut_boxes(const RefBvh& t, int max_depth) {
  std::vector<DAabb> out;
  if (!t.nodes.empty()) tree_cut(t, 0, 0, max_depth, &out);
  return out;
}

DAabb instance_world_box(const std::vector<DAabb>& cut, const double* a, const double* t) {
  DAabb wb = empty_box();
  for (const DAabb& ob : cut) {
    for (int c = 0; c < 8; c++) {
      double p[3] = {(c & 1) ? ob.mx[0] : ob.mn[0], (c & 2) ? ob.mx[1] : ob.mn[1], (c & 4) ? ob.mx[2] : ob.mn[2]};
      double q[3];
      mat_vec(a, p, q);
      for (int k = 0; k < 3; k++) q[k] += t[k];
      wb = grow(wb, q);
    }
  }
  const double ext = std::max({wb.mx[0] - wb.mn[0], wb.mx[1] - wb.mn[1], wb.mx[2] - wb.mn[2], 1e-6});
  for (int k = 0; k < 3; k++) {
    const double pad = 1e-6 * ext + 1e-6 * std::max(std::fabs(wb.mn[k]), std::fabs(wb.mx[k]));
    wb.mn[k] -= pad;
    wb.mx[k] += pad;
  }
  return wb;
}

static double mat_det(const double* m) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
}
static double lum(const double* c) { return 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2]; }

// ------------------------------------------------------------------ build
bool HostScene::build(const tpt_scene_desc* d) {
  const int nm = d->n_materials, ns = d->n_spheres, nmesh = d->n_meshes, ni = d->n_instances;
  if (nm < 0 || ns < 0 || nmesh < 0 || ni < 0) { error = "negative counts in scene descriptor"; return false; }
  if ((nm && (!d->mat_kind || !d->mat_param)) || (ns && (!d->sphere_center || !d->sphere_radius || !d->sphere_mat)) ||
      (nmesh && !d->mesh_tri_begin) || (ni && (!d->inst_mesh || !d->inst_xform || !d->inst_mat_override))) {
    error = "null array in scene descriptor";
    return false;
  }
  // ---- camera: Camera::look_at_dof (ray.rs:59-86)
  {
    const tpt_camera_desc& c = d->camera;
    double theta = c.vfov_deg * (kPi / 180.0);
    cam.half_h = std::tan(theta * 0.5);
    cam.half_w = c.aspect * cam.half_h;
    double w[3] = {c.eye[0] - c.target[0], c.eye[1] - c.target[1], c.eye[2] - c.target[2]};
    double wl = fmx(std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]), 1e-30);
    for (int k = 0; k < 3; k++) cam.w[k] = w[k] / wl;
    double u[3] = {c.up[1] * cam.w[2] - c.up[2] * cam.w[1], c.up[2] * cam.w[0] - c.up[0] * cam.w[2],
                   c.up[0] * cam.w[1] - c.up[1] * cam.w[0]};
    double ul = fmx(std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]), 1e-30);
    for (int k = 0; k < 3; k++) cam.u[k] = u[k] / ul;
    cam.v[0] = cam.w[1] * cam.u[2] - cam.w[2] * cam.u[1];
    cam.v[1] = cam.w[2] * cam.u[0] - cam.w[0] * cam.u[2];
    cam.v[2] = cam.w[0] * cam.u[1] - cam.w[1] * cam.u[0];
    for (int k = 0; k < 3; k++) cam.o[k] = c.eye[k];
    cam.lens_radius = 0.5 * fmx(c.aperture, 0.0);
    cam.focus_dist = fmx(c.focus_dist, 1e-6);
    cam.shutter_open = 0.0;
    cam.shutter_close = 1.0;
  }
  // ---- copies + validation
  mat_kind.assign(d->mat_kind, d->mat_kind + nm);
  mat_param.assign(d->mat_param, d->mat_param + 4 * (size_t)nm);
  for (int i = 0; i < nm; i++)
    if (mat_kind[(size_t)i] < TPT_MAT_LAMBERT || mat_kind[(size_t)i] > TPT_MAT_DIFFUSE_LIGHT) {
      error = "unknown material kind";
      return false;
    }
  sph.resize(4 * (size_t)ns);
  sph_mat.assign(d->sphere_mat, d->sphere_mat + ns);
  for (int i = 0; i < ns; i++) {
    for (int k = 0; k < 3; k++) sph[4 * (size_t)i + k] = d->sphere_center[3 * i + k];
    sph[4 * (size_t)i + 3] = d->sphere_radius[i];
    if (sph_mat[(size_t)i] < 0 || sph_mat[(size_t)i] >= nm) { error = "sphere material out of range"; return false; }
  }
  mesh_tri_begin.assign(d->mesh_tri_begin, d->mesh_tri_begin + (nmesh ? nmesh + 1 : 0));
  int64_t ntri = nmesh ? mesh_tri_begin[(size_t)nmesh] : 0;
  if (nmesh && mesh_tri_begin[0] != 0) { error = "mesh_tri_begin[0] must be 0"; return false; }
  for (int m = 0; m < nmesh; m++)
    if (mesh_tri_begin[(size_t)m + 1] < mesh_tri_begin[(size_t)m]) { error = "mesh_tri_begin not monotone"; return false; }
  if (ntri > 0 && (!d->tri_vertices || !d->tri_mat)) { error = "null triangle arrays"; return false; }
  if (ntri > (int64_t)TPT_LEAF_START_MASK) { error = "too many triangles"; return false; }
  tri_v.assign(d->tri_vertices, d->tri_vertices + 9 * (size_t)ntri);
  tri_mat.assign(d->tri_mat, d->tri_mat + (size_t)ntri);
  for (int64_t t = 0; t < ntri; t++)
    if (tri_mat[(size_t)t] < 0 || tri_mat[(size_t)t] >= nm) { error = "triangle material out of range"; return false; }
  inst_mesh.assign(d->inst_mesh, d->inst_mesh + ni);
  inst_xform.assign(d->inst_xform, d->inst_xform + 12 * (size_t)ni);
  inst_override.assign(d->inst_mat_override, d->inst_mat_override + ni);
  for (int i = 0; i < ni; i++)
    if (inst_override[(size_t)i] < -1 || inst_override[(size_t)i] >= nm) { error = "instance material override out of range"; return false; }
  env_kind = d->env_kind;
  if (env_kind == TPT_ENV_IMAGE) {
    env_w = d->env_width;
    env_h = d->env_height;
    if (env_w <= 0 || env_h <= 0 || !d->env_rgb) { error = "image environment needs width, height and pixels"; return false; }
    env_rgb.assign(d->env_rgb, d->env_rgb + 3 * (size_t)env_w * (size_t)env_h);
  } else if (env_kind != TPT_ENV_SKY) {
    error = "unknown env_kind";
    return false;
  }

  // ---- per-mesh BLAS: reference build and layout (plain FP64 form); own SAH tree, reordered
  // FP32 triangles and Node2 (FP32 forms)
  std::vector<std::vector<DAabb>> mesh_cut((size_t)nmesh);  // bounding boxes for the instances' world boxes
  blas.resize((size_t)nmesh);
  meshes.resize((size_t)nmesh);
  mesh_root2.assign((size_t)nmesh, TPT_REF_EMPTY);
  tris32.resize((size_t)ntri);
  tri_orig.resize((size_t)ntri);
  tri_mat32.resize((size_t)ntri);
  // Both trees of every mesh are built on host threads: each build is deterministic and
  // independent of the others, and the results are laid out in mesh order below, so the scene
  // is the same whatever the threads do.  Largest meshes first.
  std::vector<int> by_size((size_t)nmesh);
  std::iota(by_size.begin(), by_size.end(), 0);
  std::stable_sort(by_size.begin(), by_size.end(), [&](int a, int b) {
    return mesh_tri_begin[(size_t)a + 1] - mesh_tri_begin[(size_t)a] > mesh_tri_begin[(size_t)b + 1] - mesh_tri_begin[(size_t)b];
  });
  std::vector<std::vector<DAabb>> bounds_of((size_t)nmesh);
  std::vector<RefBvh> sah_of((size_t)nmesh);
  {
    std::vector<std::function<void()>> tasks;
    for (int m : by_size)
      tasks.push_back([&, m] {
        const int64_t b0 = mesh_tri_begin[(size_t)m], b1 = mesh_tri_begin[(size_t)m + 1];
        std::vector<DAabb>& bounds = bounds_of[(size_t)m];
        bounds.resize((size_t)(b1 - b0));
        for (int64_t t = b0; t < b1; t++) bounds[(size_t)(t - b0)] = triangle_bounds(&tri_v[9 * (size_t)t]);
      });
    run_parallel(tasks);
    tasks.clear();
    for (int m : by_size) {
      tasks.push_back([&, m] { blas[(size_t)m] = build_ref_bvh(bounds_of[(size_t)m], 4); });  // LEAF_SIZE (constants.rs:21)
      tasks.push_back([&, m] { sah_of[(size_t)m] = build_sah_bvh(bounds_of[(size_t)m], 4); });
    }
    run_parallel(tasks);
  }
  for (int m = 0; m < nmesh; m++) {
    int64_t b0 = mesh_tri_begin[(size_t)m], b1 = mesh_tri_begin[(size_t)m + 1];
    std::vector<DAabb>().swap(bounds_of[(size_t)m]);
    const RefBvh& r = blas[(size_t)m];
    max_blas_depth = std::max(max_blas_depth, r.depth);
    MeshDev md{};
    md.node_off = (int32_t)ref_nodes.size();
    md.idx_off = (int32_t)ref_indices.size();
    md.tri_off = (int32_t)b0;
    md.n_nodes = (int32_t)r.nodes.size();
    md.n_tris = (int32_t)(b1 - b0);
    for (const RefNode& n : r.nodes) {
      RefNodeDev x{};
      for (int k = 0; k < 3; k++) { x.bbox[k] = n.bbox.mn[k]; x.bbox[3 + k] = n.bbox.mx[k]; }
      x.left = n.left; x.right = n.right; x.start = n.start; x.count = n.count;
      ref_nodes.push_back(x);
    }
    for (int64_t ix : r.indices) ref_indices.push_back((int32_t)ix);
    // the FP32 tree (own SAH build) and its reordered triangles: slot b0 + k holds triangle
    // indices[k] of that tree
    const RefBvh f = std::move(sah_of[(size_t)m]);
    for (size_t k = 0; k < f.indices.size(); k++) {
      int64_t t = b0 + f.indices[k];
      const double* v = &tri_v[9 * (size_t)t];
      tris32[(size_t)(b0 + (int64_t)k)] = tri_f32(v);
      tri_orig[(size_t)(b0 + (int64_t)k)] = (int32_t)f.indices[k];
      tri_mat32[(size_t)(b0 + (int64_t)k)] = tri_mat[(size_t)t];
    }
    if (!f.nodes.empty()) {
      mesh_root2[(size_t)m] = to_node2(f, 0, b0, nodes2, error);
      if (!error.empty()) return false;
      mesh_cut[(size_t)m] = tree_cut_boxes(f, kInstanceCutDepth);
    }
    meshes[(size_t)m] = md;
  }
  for (int m = 0; m < nmesh; m++) meshes[(size_t)m].pad = mesh_root2[(size_t)m];

  // ---- instances
  inst64.resize((size_t)ni);
  inst32.resize((size_t)ni);
  inst_c64.assign(3 * (size_t)ni, 0.0);
  std::vector<DAabb> item_box;
  std::vector<int32_t> item_code;
  for (int i = 0; i < ni; i++) {
    InstF64& x = inst64[(size_t)i];
    const double* src = &inst_xform[12 * (size_t)i];
    for (int k = 0; k < 9; k++) x.a[k] = src[k];
    for (int k = 0; k < 3; k++) x.t[k] = src[9 + k];
    mat_invert(x.a, x.a_inv);  // Transform::from_affine (transform.rs:24-27)
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) x.nrm[3 * r + c] = x.a_inv[3 * c + r];
    x.mesh = inst_mesh[(size_t)i];
    x.mat_override = inst_override[(size_t)i];
    InstF32& y = inst32[(size_t)i];
    double ait[3];
    mat_vec(x.a_inv, x.t, ait);
    for (int r = 0; r < 3; r++) {
      for (int c = 0; c < 3; c++) {
        y.w2o[4 * r + c] = (float)x.a_inv[3 * r + c];
        y.nrm[4 * r + c] = (float)x.nrm[3 * r + c];
      }
      y.w2o[4 * r + 3] = (float)(-ait[r]);
      y.nrm[4 * r + 3] = 0.f;
    }
    bool valid = x.mesh >= 0 && x.mesh < nmesh;  // world.rs:87-90 skips unknown meshes
    y.root = valid ? mesh_root2[(size_t)x.mesh] : TPT_REF_EMPTY;
    y.mat_override = x.mat_override;
    y.mesh = valid ? x.mesh : -1;
    y.pad = 0;
    if (!valid || blas[(size_t)x.mesh].nodes.empty()) continue;
    const DAabb wb = instance_world_box(mesh_cut[(size_t)x.mesh], x.a, x.t);
    for (int k = 0; k < 3; k++) inst_c64[3 * (size_t)i + k] = 0.5 * (wb.mn[k] + wb.mx[k]);
    item_box.push_back(wb);
    item_code.push_back(i);
  }
  // ---- small scenes: every instance's triangles in world space under one tree (own design).
  // The top level then holds this tree and the spheres, and a ray enters no instance: no
  // per-instance transform, no per-instance BLAS roots.  Hits still report the original
  // instance and triangle, and ties between instances keep the reference's order (the
  // triangle's instance gives its tie key).
  {
    int64_t total = 0;
    for (int i = 0; i < ni; i++) {
      const InstF64& x = inst64[(size_t)i];
      if (x.mesh >= 0 && x.mesh < nmesh) total += mesh_tri_begin[(size_t)x.mesh + 1] - mesh_tri_begin[(size_t)x.mesh];
    }
    flat = total > 0 && total <= kFlatMaxTris;
    if (flat) {
      std::vector<double> wv;  // 9 per flattened triangle, world space
      std::vector<int32_t> f_inst, f_orig, f_mat;
      wv.reserve(9 * (size_t)total);
      for (int i = 0; i < ni; i++) {
        const InstF64& x = inst64[(size_t)i];
        if (x.mesh < 0 || x.mesh >= nmesh) continue;
        const bool flip = mat_det(x.a) < 0.0;  // (A e1) x (A e2) = det(A) A^-T (e1 x e2)
        const int64_t b0 = mesh_tri_begin[(size_t)x.mesh], b1 = mesh_tri_begin[(size_t)x.mesh + 1];
        for (int64_t t = b0; t < b1; t++) {
          for (int v = 0; v < 3; v++) {
            double q[3];
            mat_vec(x.a, &tri_v[9 * (size_t)t + 3 * v], q);  // Transform::apply_point (transform.rs:75-78)
            for (int k = 0; k < 3; k++) wv.push_back(q[k] + x.t[k]);
          }
          f_inst.push_back(i | (flip ? (int32_t)TPT_TRI_FLIP : 0));
          f_orig.push_back((int32_t)(t - b0));
          f_mat.push_back(x.mat_override >= 0 ? x.mat_override : tri_mat[(size_t)t]);
        }
      }
      std::vector<DAabb> fb((size_t)total);
      for (int64_t t = 0; t < total; t++) fb[(size_t)t] = triangle_bounds(&wv[9 * (size_t)t]);
      RefBvh fr = build_sah_bvh(fb, 4);
      flat_depth = fr.depth;
      const int64_t base = (int64_t)tris32.size();
      tris32.resize((size_t)(base + total));
      tri_orig.resize((size_t)(base + total));
      tri_mat32.resize((size_t)(base + total));
      tri_inst32.assign((size_t)(base + total), 0);
      for (size_t k = 0; k < fr.indices.size(); k++) {
        const size_t t = (size_t)fr.indices[k];
        const double* v = &wv[9 * t];
        tris32[(size_t)base + k] = tri_f32(v);
        tri_orig[(size_t)base + k] = f_orig[t];
        tri_mat32[(size_t)base + k] = f_mat[t];
        tri_inst32[(size_t)base + k] = f_inst[t];
      }
      InstF32 y{};
      y.w2o[0] = y.w2o[5] = y.w2o[10] = 1.f;  // identity world->object rows
      y.nrm[0] = y.nrm[5] = y.nrm[10] = 1.f;
      y.root = to_node2(fr, 0, base, nodes2, error);
      if (!error.empty()) return false;
      y.mat_override = -1;
      y.mesh = -1;
      flat_inst = (int32_t)inst32.size();
      inst_c64.insert(inst_c64.end(), {0.0, 0.0, 0.0});  // the flattened tree is traversed in world space
      inst32.push_back(y);
      item_box.clear();
      item_code.clear();
      item_box.push_back(fr.nodes[0].bbox);
      item_code.push_back(flat_inst);
    }
  }

  sph32.resize(4 * (size_t)ns);
  for (int s = 0; s < ns; s++) {
    const double* c = &sph[4 * (size_t)s];
    double r = std::fabs(c[3]);
    DAabb b;
    for (int k = 0; k < 3; k++) {
      double pad = 1e-6 * (r + std::fabs(c[k]));
      b.mn[k] = c[k] - r - pad;
      b.mx[k] = c[k] + r + pad;
      sph32[4 * (size_t)s + k] = (float)c[k];
    }
    sph32[4 * (size_t)s + 3] = (float)c[3];
    item_box.push_back(b);
    item_code.push_back((int32_t)(TPT_ITEM_SPHERE | (uint32_t)s));
  }
  // ---- TLAS
  if (!item_box.empty()) {
    TlasBuilder tb(item_box, nodes2);
    tb.order.resize(item_box.size());
    std::iota(tb.order.begin(), tb.order.end(), 0);
    tlas_root = tb.build(0, item_box.size(), 1);
    max_tlas_depth = tb.depth;
    tlas_items.resize(item_box.size());
    for (size_t k = 0; k < tb.order.size(); k++) tlas_items[k] = item_code[(size_t)tb.order[k]];
  }
  if (nodes2.size() > (size_t)INT32_MAX / 2) { error = "too many nodes"; return false; }
  // ---- four-wide form of every tree the FP32 kernels enter
  nodes4.clear();
  nodes4.reserve(nodes2.size() / 2 + 16);
  for (int m = 0; m < nmesh; m++) mesh_root2[(size_t)m] = collapse4(nodes2, mesh_root2[(size_t)m], nodes4);
  for (int m = 0; m < nmesh; m++) meshes[(size_t)m].pad = mesh_root2[(size_t)m];
  for (InstF32& y : inst32) {
    if (y.mesh >= 0) y.root = mesh_root2[(size_t)y.mesh];
  }
  if (flat) inst32[(size_t)flat_inst].root = collapse4(nodes2, inst32[(size_t)flat_inst].root, nodes4);
  tlas_root = collapse4(nodes2, tlas_root, nodes4);
  if (nodes4.size() > ((size_t)1 << 25)) {  // the traversal addresses nodes by 32-bit byte offsets
    error = "scene too large: more than 2^25 four-wide BVH nodes";
    return false;
  }
  // stack bound of the FP32 traversal: top-level pushes, the instance sentinel, the deepest
  // bottom-level tree's pushes, one entry of slack
  int blas_push = 0;
  for (int m = 0; m < nmesh; m++) blas_push = std::max(blas_push, push_need4(nodes4, mesh_root2[(size_t)m]));
  if (flat) blas_push = std::max(blas_push, push_need4(nodes4, inst32[(size_t)flat_inst].root));
  stack4 = push_need4(nodes4, tlas_root) + 1 + blas_push + 1;

  // ---- environment: EnvMap::from_pixels (env.rs:61-89) + alias table over the same weights
  if (env_kind == TPT_ENV_IMAGE) {
    size_t w = (size_t)env_w, h = (size_t)env_h;
    env_row_cdf.assign(h + 1, 0.0);
    env_col_cdf.assign(h * (w + 1), 0.0);
    std::vector<double> cell(w * h);
    double total = 0.0;
    for (size_t y = 0; y < h; y++) {
      double theta0 = kPi * (double)y / (double)h;
      double theta1 = kPi * (double)(y + 1) / (double)h;
      double row_weight = fmx(std::cos(theta0) - std::cos(theta1), 0.0);
      size_t off = y * (w + 1);
      double row_sum = 0.0;
      for (size_t x = 0; x < w; x++) {
        double l = fmx(lum(&env_rgb[3 * (y * w + x)]), 0.0);
        cell[y * w + x] = l * row_weight;
        row_sum += l * row_weight;
        env_col_cdf[off + x + 1] = row_sum;
      }
      total += row_sum;
      env_row_cdf[y + 1] = total;
    }
    env_total = total;
    env_rgb32.resize(4 * w * h);
    for (size_t i = 0; i < w * h; i++) {
      for (int k = 0; k < 3; k++) env_rgb32[4 * i + (size_t)k] = (float)env_rgb[3 * i + (size_t)k];
      env_rgb32[4 * i + 3] = 0.f;
    }
    size_t N = w * h;
    alias_prob.assign(N, 1.0f);
    alias_idx.resize(N);
    cell_pdf.assign(N, 0.0f);
    for (size_t i = 0; i < N; i++) alias_idx[i] = (int32_t)i;
    if (total > 0.0) {
      // Vose's alias method (Vose 1991) in FP64.
      std::vector<double> q(N);
      std::vector<size_t> small, large;
      for (size_t i = 0; i < N; i++) {
        q[i] = cell[i] / total * (double)N;
        cell_pdf[i] = (float)((cell[i] / total) * (double)w * (double)h / (2.0 * kPi * kPi));
        (q[i] < 1.0 ? small : large).push_back(i);
      }
      while (!small.empty() && !large.empty()) {
        size_t s = small.back(); small.pop_back();
        size_t l = large.back(); large.pop_back();
        alias_prob[s] = (float)q[s];
        alias_idx[s] = (int32_t)l;
        q[l] = (q[l] + q[s]) - 1.0;
        (q[l] < 1.0 ? small : large).push_back(l);
      }
      for (size_t i : large) { alias_prob[i] = 1.0f; alias_idx[i] = (int32_t)i; }
      for (size_t i : small) {  // leftovers from rounding: keep only a positive-weight cell
        alias_prob[i] = 1.0f;
        alias_idx[i] = (int32_t)i;
      }
    }
  }
  rebuild_lights();
  return error.empty();
}

// World::build_lights (world.rs:127-167): spheres first, then every instance's triangles whose
// effective material emits; weight = area(t=0.5) * luminance(emit); cumulative CDF.
void HostScene::rebuild_lights() {
  lights.clear();
  light_cdf.clear();
  lights32.clear();
  light_cdf32.clear();
  double total = 0.0;
  auto emits = [&](int mid) { return mid >= 0 && mid < (int)mat_kind.size() && mat_kind[(size_t)mid] == TPT_MAT_DIFFUSE_LIGHT; };
  auto add = [&](LightF64 L, double area, const double* world_v) {
    const int mid = L.kind == 0 ? sph_mat[(size_t)L.idx]
                                : (inst_override[(size_t)L.inst] >= 0 ? inst_override[(size_t)L.inst]
                                                                      : tri_mat[(size_t)(mesh_tri_begin[(size_t)L.mesh] + L.tri)]);
    const double* e = &mat_param[4 * (size_t)mid];
    double weight = area * lum(e);
    if (!(weight > 0.0)) return;
    total += weight;
    for (int k = 0; k < 3; k++) L.emit[k] = e[k];
    L.weight = weight;
    lights.push_back(L);
    light_cdf.push_back(total);
    LightF32 f{};
    f.kind = L.kind;
    f.mat = mid;
    if (L.kind == 0) {
      for (int k = 0; k < 4; k++) f.p0[k] = (float)sph[4 * (size_t)L.idx + (size_t)k];
    } else {
      for (int k = 0; k < 3; k++) {
        f.p0[k] = (float)world_v[k];
        f.p1[k] = (float)world_v[3 + k];
        f.p2[k] = (float)world_v[6 + k];
      }
    }
    for (int k = 0; k < 3; k++) f.emit[k] = (float)e[k];
    f.emit[3] = (float)weight;
    lights32.push_back(f);
  };
  for (size_t s = 0; s < sph_mat.size(); s++) {
    if (!emits(sph_mat[s])) continue;
    double r = sph[4 * s + 3];
    LightF64 L{};
    L.kind = 0;
    L.idx = (int32_t)s;
    add(L, 4.0 * kPi * r * r, nullptr);
  }
  // per mesh, the triangles whose own material emits (instances without an override)
  std::vector<std::vector<int32_t>> emissive(meshes.size());
  for (size_t m = 0; m < meshes.size(); m++)
    for (int64_t t = mesh_tri_begin[m]; t < mesh_tri_begin[m + 1]; t++)
      if (emits(tri_mat[(size_t)t])) emissive[m].push_back((int32_t)(t - mesh_tri_begin[m]));
  for (size_t i = 0; i < inst64.size(); i++) {
    const InstF64& x = inst64[i];
    if (x.mesh < 0 || x.mesh >= (int)meshes.size()) continue;
    std::vector<int32_t> all;
    const std::vector<int32_t>* list = &emissive[(size_t)x.mesh];
    if (x.mat_override >= 0) {
      if (!emits(x.mat_override)) continue;
      all.resize((size_t)meshes[(size_t)x.mesh].n_tris);
      std::iota(all.begin(), all.end(), 0);
      list = &all;
    }
    for (int32_t t : *list) {
      // tri_world_verts at time 0.5 (world.rs:276-292): vertices_at(0.5) of a static triangle
      // is v*0.5 + v*0.5 = v exactly, then apply_point.
      const double* v = &tri_v[9 * (size_t)(mesh_tri_begin[(size_t)x.mesh] + t)];
      double wv[9];
      for (int j = 0; j < 3; j++) {
        double vt[3];
        for (int k = 0; k < 3; k++) vt[k] = v[3 * j + k] * (1.0 - 0.5) + v[3 * j + k] * 0.5;
        mat_vec(x.a, vt, wv + 3 * j);
        for (int k = 0; k < 3; k++) wv[3 * j + k] = wv[3 * j + k] + x.t[k];
      }
      double e1[3] = {wv[3] - wv[0], wv[4] - wv[1], wv[5] - wv[2]};
      double e2[3] = {wv[6] - wv[0], wv[7] - wv[1], wv[8] - wv[2]};
      double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2], cz = e1[0] * e2[1] - e1[1] * e2[0];
      double area = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
      LightF64 L{};
      L.kind = 1;
      L.mesh = x.mesh;
      L.tri = t;
      L.inst = (int32_t)i;
      add(L, area, wv);
    }
  }
  light_total = total;
  for (double c : light_cdf) light_cdf32.push_back((float)c);
}

}  // namespace tpt
