// bvh_build.cpp -- binned-SAH BVH build on the host (replaces the TinyEmbree /
// Embree Raytracer.CommitScene build, SeeSharp/Scene.cs:108-113).
#include "bvh_build.h"

#include <algorithm>
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
            constexpr int B = 16;
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
                auto it = std::partition(order.begin() + begin, order.begin() + end, [&](int t) {
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

    // Convert: one device node per inner BNode holding both child boxes.
    std::vector<int> dev_index(used, -1);
    int dev_count = 0;
    // Breadth-first order keeps the top of the tree compact.
    const bool root_is_leaf = n == 0 || b.nodes[0].count > 0;
    std::vector<int> queue;
    if (!root_is_leaf) queue.push_back(0);
    for (size_t q = 0; q < queue.size(); ++q) {
        int bn = queue[q];
        if (b.nodes[bn].count > 0) continue;
        dev_index[bn] = dev_count++;
        queue.push_back(b.nodes[bn].left);
        queue.push_back(b.nodes[bn].right);
    }
    const BNode& root = b.nodes[0];
    bool root_leaf = root_is_leaf;
    if (root_leaf) dev_count = 1;
    out->nodes.assign((size_t)dev_count * 16, 0.0f);
    auto put_child = [&](float* node, int side, const BNode* c) {
        // side 0 = left, 1 = right
        // An absent child gets the box [+inf, +inf]^3: both slab distances are
        // +inf (or -inf) on every axis, so the slab test never accepts it.
        float lox = c ? c->lo[0] : INFINITY, hix = c ? c->hi[0] : INFINITY;
        float loy = c ? c->lo[1] : INFINITY, hiy = c ? c->hi[1] : INFINITY;
        float loz = c ? c->lo[2] : INFINITY, hiz = c ? c->hi[2] : INFINITY;
        float* row = node + side * 4;
        row[0] = lox, row[1] = hix, row[2] = loy, row[3] = hiy;
        node[8 + side * 2 + 0] = loz;
        node[8 + side * 2 + 1] = hiz;
        int idx, cnt;
        if (!c) {
            idx = 0, cnt = 0;
            // an empty box is never entered, so the index is never followed
        } else if (c->count > 0) {
            idx = c->first, cnt = c->count;
        } else {
            idx = dev_index[(int)(c - b.nodes.data())], cnt = 0;
        }
        int* ints = reinterpret_cast<int*>(node + 12);
        ints[side] = idx;
        ints[2 + side] = cnt;
    };
    if (root_leaf) {
        float* node = out->nodes.data();
        if (n > 0) put_child(node, 0, &root);
        else put_child(node, 0, nullptr);
        put_child(node, 1, nullptr);
    } else {
        for (int bn = 0; bn < used; ++bn) {
            int d = dev_index[bn];
            if (d < 0) continue;
            float* node = out->nodes.data() + (size_t)d * 16;
            put_child(node, 0, &b.nodes[b.nodes[bn].left]);
            put_child(node, 1, &b.nodes[b.nodes[bn].right]);
        }
    }
    out->leaf_order = std::move(b.order);
    out->depth = b.max_depth.load();
    for (int k = 0; k < 3; ++k) {
        out->scene_lo[k] = n ? root.lo[k] : 0.0f;
        out->scene_hi[k] = n ? root.hi[k] : 0.0f;
    }
}

}  // namespace sspt
