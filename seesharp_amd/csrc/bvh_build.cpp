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

namespace {

// fp16 helpers for traversal-node planes: values are non-negative, finite and
// below 2^16; only zero and normal fp16 numbers are produced.
uint16_t f16_bits(double v) {  // v: zero or an exactly representable normal fp16
    if (v == 0.0) return 0;
    int p;
    const double m = std::frexp(v, &p);  // v = m 2^p, m in [0.5, 1)
    const uint32_t mant = (uint32_t)((2.0 * m - 1.0) * 1024.0);
    return (uint16_t)(((uint32_t)(p - 1 + 15) << 10) | mant);
}
double f16_value(uint16_t b) {
    if (b == 0) return 0.0;
    return std::ldexp(1.0 + (double)(b & 1023u) / 1024.0, (int)(b >> 10) - 15);
}
// largest zero-or-normal fp16 <= y
uint16_t f16_floor(double y) {
    if (!(y >= 0x1p-14)) return 0;
    int p;
    std::frexp(y, &p);
    const double q = std::ldexp(1.0, p - 1 - 10);
    return f16_bits(std::floor(y / q) * q);
}
// smallest normal fp16 >= y (y > 0), 0 for y <= 0
uint16_t f16_ceil(double y) {
    if (y <= 0.0) return 0;
    if (y <= 0x1p-14) return f16_bits(0x1p-14);
    int p;
    std::frexp(y, &p);
    const double q = std::ldexp(1.0, p - 1 - 10);
    return f16_bits(std::ceil(y / q) * q);
}
// a - b as s + err exactly (Knuth's TwoSum in double)
double diff_exact(float a, float b, double* err) {
    const double x = (double)a, y = -(double)b;
    const double s = x + y, bb = s - x;
    *err = (x - (s - bb)) + (y - bb);
    return s;
}

}  // namespace

bool encode_traversal_nodes(const std::vector<float>& nodes, int threads, std::vector<uint32_t>* words,
                            std::vector<int>* big_leaves, std::string* error) {
    const size_t nn = nodes.size() / 32;
    auto count_of = [&](size_t n, int k) -> int {
        int32_t c;
        std::memcpy(&c, &nodes[n * 32 + 28 + k], 4);
        if (k == 0) c = (c << 8) >> 8;  // slot 0 keeps the slot masks in its top byte
        return c;
    };
    auto child_of = [&](size_t n, int k) -> int32_t {
        int32_t c;
        std::memcpy(&c, &nodes[n * 32 + 24 + k], 4);
        return c;
    };
    // leaves of more than four triangles: numbered in node order
    std::vector<int> big_index(nn * 4, -1);
    big_leaves->clear();
    for (size_t n = 0; n < nn; ++n)
        for (int k = 0; k < 4; ++k) {
            const int c = count_of(n, k);
            if (c > 4) {
                big_index[n * 4 + k] = (int)(big_leaves->size() / 2);
                big_leaves->push_back(child_of(n, k));
                big_leaves->push_back(c);
            } else if (c > 0 && (uint32_t)child_of(n, k) >= (1u << 30)) {
                if (error) *error = "a leaf's first triangle index exceeds 2^30";
                return false;
            }
        }
    words->assign(nn * kTravNodeWords, 0u);
    auto encode = [&](size_t n0, size_t n1) {
        for (size_t n = n0; n < n1; ++n) {
            const float* f = &nodes[n * 32];
            uint32_t* w = &(*words)[n * kTravNodeWords];
            int kind[4];
            uint32_t leafm = 0, innerm = 0;
            for (int k = 0; k < 4; ++k) {
                const int c = count_of(n, k);
                kind[k] = c < 0 ? -1 : c;  // -1 absent, 0 inner, > 0 leaf of c triangles
                if (c == 0) innerm |= 1u << k;
                if (c > 0 && c <= 4) leafm |= 1u << k;
                if (c > 4) {
                    leafm |= 1u << k;
                    innerm |= 1u << k;
                }
            }
            uint32_t meta = (leafm | (innerm << 4)) << 24;
            for (int a = 0; a < 3; ++a) {
                const float* lo = f + 4 * (2 * a);
                const float* hi = f + 4 * (2 * a + 1);
                float origin = 0.0f;
                bool any = false;
                for (int k = 0; k < 4; ++k)
                    if (kind[k] >= 0) {
                        origin = any ? std::min(origin, lo[k]) : lo[k];
                        any = true;
                    }
                double extent = 0.0;
                for (int k = 0; k < 4; ++k)
                    if (kind[k] >= 0) extent = std::max(extent, (double)hi[k] - (double)origin);
                int e = 0;
                if (extent > 0.0) {
                    int p;
                    std::frexp(extent, &p);  // extent in [2^(p-1), 2^p): offsets * 2^-e below 2^15
                    e = std::min(127, std::max(-128, p - 15));
                }
                std::memcpy(&w[a], &origin, 4);
                meta |= (uint32_t)(uint8_t)(int8_t)e << (8 * a);
                uint16_t hl[4] = {0, 0, 0, 0}, hh[4] = {0, 0, 0, 0};
                for (int k = 0; k < 4; ++k) {
                    if (kind[k] < 0) continue;
                    double err;
                    // lower plane rounded down: origin + h 2^e <= lo
                    double sl = diff_exact(lo[k], origin, &err);
                    double x = std::ldexp(sl, -e);
                    uint16_t h = f16_floor(x);
                    if (err < 0.0 && f16_value(h) == x) h = h > 0x0400 ? (uint16_t)(h - 1) : (uint16_t)0;
                    hl[k] = h;
                    // upper plane rounded up: origin + h 2^e >= hi
                    double sh = diff_exact(hi[k], origin, &err);
                    x = std::ldexp(sh, -e);
                    h = f16_ceil(x);
                    if (err > 0.0 && f16_value(h) == x) h = h == 0 ? (uint16_t)0x0400 : (uint16_t)(h + 1);
                    hh[k] = h;
                }
                w[4 + 4 * a] = (uint32_t)hl[0] | ((uint32_t)hl[1] << 16);
                w[5 + 4 * a] = (uint32_t)hl[2] | ((uint32_t)hl[3] << 16);
                w[6 + 4 * a] = (uint32_t)hh[0] | ((uint32_t)hh[1] << 16);
                w[7 + 4 * a] = (uint32_t)hh[2] | ((uint32_t)hh[3] << 16);
            }
            w[3] = meta;
            for (int k = 0; k < 4; ++k) {
                int32_t c;
                if (kind[k] < 0) c = -1;
                else if (kind[k] == 0) c = child_of(n, k);
                else if (kind[k] <= 4) c = (int32_t)((uint32_t)child_of(n, k) | ((uint32_t)(kind[k] - 1) << 30));
                else c = big_index[n * 4 + k];
                std::memcpy(&w[16 + k], &c, 4);
            }
        }
    };
    threads = std::max(1, std::min(threads, (int)(nn / 4096) + 1));
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back(encode, nn * t / threads, nn * (t + 1) / threads);
    for (auto& th : pool) th.join();
    return true;
}

}  // namespace sspt
