/*
 * Binned-SAH BVH over the ground triangles (see bvh.h for the layout).
 */
#include "bvh.h"

#include "htgx.h"

#include <algorithm>
#include <cfloat>
#include <cmath>

namespace htgx {
namespace {

constexpr int kBins = 16;
constexpr uint32_t kLeafMax = 4;

struct Box {
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
  float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  void grow(const float p[3])
  {
    for(int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], p[a]); hi[a] = std::max(hi[a], p[a]); }
  }
  void grow(const Box& b)
  {
    for(int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], b.lo[a]); hi[a] = std::max(hi[a], b.hi[a]); }
  }
  float area() const
  {
    const float d0 = hi[0] - lo[0], d1 = hi[1] - lo[1], d2 = hi[2] - lo[2];
    if(!(d0 >= 0 && d1 >= 0 && d2 >= 0)) return 0.f;
    return 2.f * (d0 * d1 + d1 * d2 + d2 * d0);
  }
};

/* Tree node of the build: box, children (inner) or triangle range (leaf) */
struct TreeNode {
  Box box;
  int32_t child[2] = {-1, -1};
  uint32_t tris = 0; /* leaf: first << 3 | count */
};

struct Builder {
  const float* verts;
  const uint32_t* tris;
  std::vector<Box> tbox;
  std::vector<float> cent; /* [ntris][3] */
  std::vector<uint32_t> order;
  std::vector<TreeNode> tree;
  BvhHost* out;

  void emit_leaf(uint32_t node, uint32_t begin, uint32_t end, const uint32_t* tri_mtl)
  {
    const uint32_t first = (uint32_t)out->tris.size();
    for(uint32_t i = begin; i < end; ++i) {
      const uint32_t t = order[i];
      const float* v0 = verts + 3 * (size_t)tris[3 * (size_t)t + 0];
      const float* v1 = verts + 3 * (size_t)tris[3 * (size_t)t + 1];
      const float* v2 = verts + 3 * (size_t)tris[3 * (size_t)t + 2];
      BvhTri tr;
      for(int a = 0; a < 3; ++a) {
        tr.v0[a] = v0[a];
        tr.e1[a] = v1[a] - v0[a];
        tr.e2[a] = v2[a] - v0[a];
      }
      out->tris.push_back(tr);
      out->prim.push_back(t);
      out->mtl.push_back(tri_mtl ? tri_mtl[t] : 0u);
    }
    tree[node].tris = (first << 3) | (end - begin);
  }

  uint32_t build(uint32_t begin, uint32_t end, uint32_t depth, const uint32_t* tri_mtl)
  {
    const uint32_t node = (uint32_t)tree.size();
    tree.push_back(TreeNode{});
    out->depth = std::max(out->depth, depth);
    Box bb, cb;
    for(uint32_t i = begin; i < end; ++i) {
      bb.grow(tbox[order[i]]);
      cb.grow(&cent[3 * (size_t)order[i]]);
    }
    tree[node].box = bb;
    const uint32_t n = end - begin;
    if(n <= kLeafMax) {
      emit_leaf(node, begin, end, tri_mtl);
      return node;
    }
    /* binned SAH on the widest centroid axis */
    int axis = 0;
    for(int a = 1; a < 3; ++a)
      if(cb.hi[a] - cb.lo[a] > cb.hi[axis] - cb.lo[axis]) axis = a;
    const float ext = cb.hi[axis] - cb.lo[axis];
    uint32_t mid = begin + n / 2;
    if(ext > 0.f) {
      Box bins[kBins];
      uint32_t cnt[kBins] = {0};
      const float scale = (float)kBins / ext;
      auto bin_of = [&](uint32_t t) {
        const int b = (int)((cent[3 * (size_t)t + axis] - cb.lo[axis]) * scale);
        return std::min(kBins - 1, std::max(0, b));
      };
      for(uint32_t i = begin; i < end; ++i) {
        const int b = bin_of(order[i]);
        bins[b].grow(tbox[order[i]]);
        cnt[b]++;
      }
      float right_area[kBins];
      uint32_t right_cnt[kBins];
      Box acc;
      uint32_t c = 0;
      for(int b = kBins - 1; b > 0; --b) {
        acc.grow(bins[b]);
        c += cnt[b];
        right_area[b] = acc.area();
        right_cnt[b] = c;
      }
      Box left;
      uint32_t lc = 0;
      float best = FLT_MAX;
      int best_b = -1;
      for(int b = 1; b < kBins; ++b) {
        left.grow(bins[b - 1]);
        lc += cnt[b - 1];
        if(!lc || !right_cnt[b]) continue;
        const float cost = left.area() * (float)lc + right_area[b] * (float)right_cnt[b];
        if(cost < best) { best = cost; best_b = b; }
      }
      if(best_b > 0) {
        const auto it = std::partition(order.begin() + begin, order.begin() + end,
                                       [&](uint32_t t) { return bin_of(t) < best_b; });
        mid = (uint32_t)(it - order.begin());
      }
    }
    if(mid == begin || mid == end) {
      /* all centroids in one bin: split at the median */
      mid = begin + n / 2;
      std::nth_element(order.begin() + begin, order.begin() + mid, order.begin() + end,
                       [&](uint32_t x, uint32_t y) {
                         const float cx = cent[3 * (size_t)x + axis], cy = cent[3 * (size_t)y + axis];
                         return cx < cy || (cx == cy && x < y);
                       });
    }
    /* child 0 holds the lower centroids along `axis' */
    const uint32_t l = build(begin, mid, depth + 1, tri_mtl);
    const uint32_t r = build(mid, end, depth + 1, tri_mtl);
    tree[node].child[0] = (int32_t)l;
    tree[node].child[1] = (int32_t)r;
    return node;
  }

  /* Depth-first layout with skip links, appended to out->nodes (iterative:
   * the tree can be deep) */
  void layout()
  {
    struct Item { uint32_t node; uint32_t slot; bool done; };
    std::vector<Item> stack;
    stack.push_back({0u, 0u, false});
    while(!stack.empty()) {
      const Item it = stack.back();
      stack.pop_back();
      if(it.done) {
        out->nodes[it.slot].skip = (uint32_t)out->nodes.size(); /* first node after the subtree */
        continue;
      }
      const TreeNode& T = tree[it.node];
      const uint32_t slot = (uint32_t)out->nodes.size();
      out->nodes.push_back(BvhNode{});
      BvhNode& N = out->nodes[slot];
      for(int a = 0; a < 3; ++a) { N.lo[a] = T.box.lo[a]; N.hi[a] = T.box.hi[a]; }
      N.tris = T.tris;
      stack.push_back({it.node, slot, true});
      if(!T.tris) {
        stack.push_back({(uint32_t)T.child[1], 0u, false});
        stack.push_back({(uint32_t)T.child[0], 0u, false});
      }
    }
  }
};

} /* namespace */

int bvh_build(const float* verts, uint32_t nverts, const uint32_t* tris, const uint32_t* tri_mtl,
              uint32_t ntris, BvhHost& out)
{
  if(!verts || !tris || !nverts || !ntris) return HTGX_BAD_ARG;
  if(ntris >= (1u << 29)) return HTGX_BAD_ARG;
  Builder b;
  b.verts = verts;
  b.tris = tris;
  b.out = &out;
  b.tbox.resize(ntris);
  b.cent.resize(3 * (size_t)ntris);
  b.order.resize(ntris);
  Box scene;
  for(uint32_t t = 0; t < ntris; ++t) {
    for(int k = 0; k < 3; ++k) {
      const uint32_t v = tris[3 * (size_t)t + k];
      if(v >= nverts) return HTGX_BAD_ARG;
      b.tbox[t].grow(verts + 3 * (size_t)v);
    }
    for(int a = 0; a < 3; ++a) b.cent[3 * (size_t)t + a] = 0.5f * (b.tbox[t].lo[a] + b.tbox[t].hi[a]);
    b.order[t] = t;
    scene.grow(b.tbox[t]);
  }
  out.nodes.clear();
  out.tris.clear();
  out.prim.clear();
  out.mtl.clear();
  out.tris.reserve(ntris);
  b.tree.reserve(ntris);
  b.build(0, ntris, 0, tri_mtl);
  out.nodes.reserve(b.tree.size());
  b.layout();
  for(int a = 0; a < 3; ++a) { out.lo[a] = scene.lo[a]; out.hi[a] = scene.hi[a]; }
  return HTGX_OK;
}

} /* namespace htgx */
