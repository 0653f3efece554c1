// bvh_build.cpp -- binned-SAH BVH build on the host (replaces the TinyEmbree /
// Embree Raytracer.CommitScene build, SeeSharp/Scene.cs:108-113).
#include "bvh_build.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <future>
#include <thread>

#include "scene_types.h"

namespace sspt {
namespace {

struct BNode {
    float lo[3], hi[3];
    int left, right;  // inner: child BNode indices
    int first, count; // leaf: range in `order`, count > 0
    int depth;
};

struct Builder {
    const float* tlo;  // per-triangle box
    const float* thi;
    const float* cen;  // per-triangle centroid
    std::vector<int> order;
    std::vector<BNode> nodes;
    std::atomic<int> next_node{0};
    std::atomic<int> max_depth{0};
    std::atomic<int> spawn_budget{0};
    int max_leaf = 4;

    int alloc2() { return next_node.fetch_add(2); }

    static float half_area(const float* l, const float* h) {
        float dx = h[0] - l[0], dy = h[1] - l[1], dz = h[2] - l[2];
        if (!(dx >= 0.0f)) return 0.0f;
        return dx * dy + dy * dz + dz * dx;
    }

    void build(int node, int begin, int end, int depth) {
        for (;;) {
            BNode& n = nodes[node];
            n.depth = depth;
            int cur = max_depth.load();
            while (depth > cur && !max_depth.compare_exchange_weak(cur, depth)) {
            }
            float cl[3] = {INFINITY, INFINITY, INFINITY}, ch[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (int k = 0; k < 3; ++k) n.lo[k] = INFINITY, n.hi[k] = -INFINITY;
            for (int i = begin; i < end; ++i) {
                int t = order[i];
                for (int k = 0; k < 3; ++k) {
                    n.lo[k] = std::min(n.lo[k], tlo[t * 3 + k]);
                    n.hi[k] = std::max(n.hi[k], thi[t * 3 + k]);
                    cl[k] = std::min(cl[k], cen[t * 3 + k]);
                    ch[k] = std::max(ch[k], cen[t * 3 + k]);
                }
            }
            int count = end - begin;
            if (count <= max_leaf || depth >= kMaxBvhDepth - 1) {
                n.first = begin;
                n.count = count;
                n.left = n.right = -1;
                return;
            }
            constexpr int B = kSahBins;
            float best_cost = INFINITY;
            int best_axis = -1, best_split = -1;
            float leaf_cost = half_area(n.lo, n.hi) * (float)count;
            for (int k = 0; k < 3; ++k) {
                float extent = ch[k] - cl[k];
                if (!(extent > 0.0f)) continue;
                float scale = (float)B / extent;
                int cnt[B] = {0};
                float bl[B][3], bh[B][3];
                for (int j = 0; j < B; ++j)
                    for (int q = 0; q < 3; ++q) bl[j][q] = INFINITY, bh[j][q] = -INFINITY;
                for (int i = begin; i < end; ++i) {
                    int t = order[i];
                    int bin = std::min(B - 1, (int)((cen[t * 3 + k] - cl[k]) * scale));
                    cnt[bin]++;
                    for (int q = 0; q < 3; ++q) {
                        bl[bin][q] = std::min(bl[bin][q], tlo[t * 3 + q]);
                        bh[bin][q] = std::max(bh[bin][q], thi[t * 3 + q]);
                    }
                }
                float ra[B];
                int rc[B];
                float rl[3] = {INFINITY, INFINITY, INFINITY}, rh[3] = {-INFINITY, -INFINITY, -INFINITY};
                int rcount = 0;
                for (int j = B - 1; j > 0; --j) {
                    rcount += cnt[j];
                    for (int q = 0; q < 3; ++q) rl[q] = std::min(rl[q], bl[j][q]), rh[q] = std::max(rh[q], bh[j][q]);
                    ra[j] = half_area(rl, rh);
                    rc[j] = rcount;
                }
                float ll[3] = {INFINITY, INFINITY, INFINITY}, lh[3] = {-INFINITY, -INFINITY, -INFINITY};
                int lcount = 0;
                for (int j = 0; j < B - 1; ++j) {
                    lcount += cnt[j];
                    for (int q = 0; q < 3; ++q) ll[q] = std::min(ll[q], bl[j][q]), lh[q] = std::max(lh[q], bh[j][q]);
                    if (lcount == 0 || rc[j + 1] == 0) continue;
                    float cost = half_area(ll, lh) * (float)lcount + ra[j + 1] * (float)rc[j + 1];
                    if (cost < best_cost) {
                        best_cost = cost;
                        best_axis = k;
                        best_split = j;
                    }
                }
            }
            (void)leaf_cost;
            int mid;
            if (best_axis < 0) {
                mid = begin + count / 2;  // all centroids coincide: object median
            } else {
                int k = best_axis;
                float scale = (float)B / (ch[k] - cl[k]);
                float c0 = cl[k];
                // stable: every node's range keeps ascending triangle ids, so the object
                // median of coincident centroids takes the lower half of the ids
                auto it = std::stable_partition(order.begin() + begin, order.begin() + end, [&](int t) {
                    return std::min(B - 1, (int)((cen[t * 3 + k] - c0) * scale)) <= best_split;
                });
                mid = (int)(it - order.begin());
                if (mid == begin || mid == end) mid = begin + count / 2;
            }
            int l = alloc2();
            n.left = l;
            n.right = l + 1;
            n.count = 0;
            n.first = -1;
            // Build the smaller half on another thread when the range is large.
            if (count > 65536 && spawn_budget.fetch_sub(1) > 0) {
                auto fut = std::async(std::launch::async, [=]() { build(l + 1, mid, end, depth + 1); });
                build(l, begin, mid, depth + 1);
                fut.get();
                return;
            }
            build(l + 1, mid, end, depth + 1);
            node = l;
            end = mid;
            depth = depth + 1;
        }
    }
};

}  // namespace

void build_bvh(const BvhInput& in, int max_leaf, int threads, BvhOutput* out) {
    int n = in.num_tris;
    std::vector<float> lo((size_t)n * 3), hi((size_t)n * 3), ce((size_t)n * 3);
    for (int i = 0; i < n; ++i) {
        const float* v = in.tri_vertices + (size_t)i * 9;
        for (int k = 0; k < 3; ++k) {
            float a = v[k], b = v[3 + k], c = v[6 + k];
            lo[i * 3 + k] = std::min(a, std::min(b, c));
            hi[i * 3 + k] = std::max(a, std::max(b, c));
            ce[i * 3 + k] = 0.5f * lo[i * 3 + k] + 0.5f * hi[i * 3 + k];
        }
    }
    Builder b;
    b.tlo = lo.data();
    b.thi = hi.data();
    b.cen = ce.data();
    b.max_leaf = max_leaf;
    b.order.resize(n);
    for (int i = 0; i < n; ++i) b.order[i] = i;
    b.nodes.resize(std::max(1, 2 * n));
    b.next_node = 1;
    b.spawn_budget = std::max(0, threads - 1);
    if (n > 0) b.build(0, 0, n, 0);
    int used = b.next_node.load();

    // Collapse into a 4-wide BVH: each device node adopts the children of its
    // largest-area inner children until it holds four, breadth-first so the top
    // of the tree is contiguous in memory.
    const bool root_is_leaf = n == 0 || b.nodes[0].count > 0;
    auto collapse = [&](int bn, int* kids) {
        int k = 2;
        kids[0] = b.nodes[bn].left;
        kids[1] = b.nodes[bn].right;
        while (k < 4) {
            int best = -1;
            float best_area = -1.0f;
            for (int i = 0; i < k; ++i) {
                const BNode& c = b.nodes[kids[i]];
                if (c.count > 0) continue;
                float a = Builder::half_area(c.lo, c.hi);
                if (a > best_area) best_area = a, best = i;
            }
            if (best < 0) break;
            int c = kids[best];
            kids[best] = b.nodes[c].left;
            kids[k++] = b.nodes[c].right;
        }
        return k;
    };
    std::vector<int> dev_of(used, -1);   // BNode -> device node
    std::vector<int> order4;             // device node -> BNode
    std::vector<int> depth4;             // device node depth
    std::vector<std::array<int, 4>> kids4;
    std::vector<int> nkids4;
    if (!root_is_leaf) {
        dev_of[0] = 0;
        order4.push_back(0);
        depth4.push_back(0);
    }
    for (size_t q = 0; q < order4.size(); ++q) {
        std::array<int, 4> kids{-1, -1, -1, -1};
        int k = collapse(order4[q], kids.data());
        for (int i = 0; i < k; ++i) {
            if (b.nodes[kids[i]].count > 0) continue;
            dev_of[kids[i]] = (int)order4.size();
            order4.push_back(kids[i]);
            depth4.push_back(depth4[q] + 1);
        }
        kids4.push_back(kids);
        nkids4.push_back(k);
    }
    const BNode& root = b.nodes[0];
    int dev_count = root_is_leaf ? 1 : (int)order4.size();
    out->nodes.assign((size_t)dev_count * 32, 0.0f);
    auto put_child = [&](float* node, int slot, const BNode* c) {
        // An absent child has count -1 (the traversal skips it) and the box
        // [+inf, +inf]^3. The box alone is not enough: with tmax = +inf and a
        // direction whose components are all positive every slab gives +inf.
        for (int a = 0; a < 3; ++a) {
            node[(2 * a) * 4 + slot] = c ? c->lo[a] : INFINITY;
            node[(2 * a + 1) * 4 + slot] = c ? c->hi[a] : INFINITY;
        }
        int idx = 0, cnt = c ? 0 : -1;
        if (c && c->count > 0) {
            idx = c->first, cnt = c->count;
        } else if (c) {
            idx = dev_of[(int)(c - b.nodes.data())];
        }
        int* ints = reinterpret_cast<int*>(node + 24);
        ints[slot] = idx;
        ints[4 + slot] = cnt;
    };
    if (root_is_leaf) {
        float* node = out->nodes.data();
        put_child(node, 0, n > 0 ? &root : nullptr);
        for (int i = 1; i < 4; ++i) put_child(node, i, nullptr);
        out->stack_need = 1;
        out->depth4 = 0;
    } else {
        for (size_t d = 0; d < order4.size(); ++d) {
            float* node = out->nodes.data() + d * 32;
            for (int i = 0; i < 4; ++i) put_child(node, i, i < nkids4[d] ? &b.nodes[kids4[d][i]] : nullptr);
        }
        // exact traversal stack bound: a node pushes all but one of its inner children
        std::vector<int> need(order4.size(), 0);
        for (size_t d = order4.size(); d-- > 0;) {
            int inner = 0, deepest = 0;
            for (int i = 0; i < nkids4[d]; ++i) {
                int c = dev_of[kids4[d][i]];
                if (b.nodes[kids4[d][i]].count > 0 || c < 0) continue;
                ++inner;
                deepest = std::max(deepest, need[c]);
            }
            need[d] = std::max(0, inner - 1) + deepest;
        }
        out->stack_need = need[0] + 1;
        out->depth4 = *std::max_element(depth4.begin(), depth4.end());
    }
    out->leaf_order = std::move(b.order);
    out->depth = b.max_depth.load();
    for (int k = 0; k < 3; ++k) {
        out->scene_lo[k] = n ? root.lo[k] : 0.0f;
        out->scene_hi[k] = n ? root.hi[k] : 0.0f;
    }
}

}  // namespace sspt
