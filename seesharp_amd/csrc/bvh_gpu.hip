// bvh_gpu.hip -- binned-SAH BVH build on the device (replaces the TinyEmbree /
// Embree Raytracer.CommitScene build, SeeSharp/Scene.cs:108-113), producing the
// 4-wide node layout of scene_types.h.
//
// The build runs the host builder's algorithm (bvh_build.cpp) in three phases:
//   1. large nodes (more than kSmallMax triangles), level by level: chunks of
//      kChunk positions reduce the node's triangle and centroid bounds, bin
//      their triangles into per-block LDS bins that are merged with atomics,
//      one thread per node sweeps the 3 x kSahBins bins for the cheapest split,
//      and the chunks scatter their triangles to the two sides;
//   2. small nodes: one wave builds a node's whole subtree, depth first, with its
//      bins in LDS and the split sweep as wave prefix / suffix scans over the bins;
//   3. the collapse to 4-wide nodes, level by level in breadth-first order.
// Bounds, bins and counts are exact (min / max and integer sums do not depend on
// the order they are taken in) and every split cost is computed from them with
// the host's fp32 operations in the host's order, so the device picks the host's
// splits; both partitions are stable, so every node's range holds its triangles
// in ascending id order on both sides and the trees agree node for node and
// leaf for leaf (a bound may be -0 where the host has +0).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "bvh_build.h"
#include "scene_types.h"

// Split costs must round exactly like the host builder's: no fused multiply-adds.
#pragma clang fp contract(off)

namespace sspt {
namespace {

constexpr int kBins = kSahBins;
constexpr int kChunk = 2048;      // positions per block in the large-node levels
constexpr int kBlock = 256;       // threads per large-node block
constexpr int kSmallMax = 128;    // nodes of at most this many triangles go to phase 2
constexpr int kSmallStack = 48;   // depth-first stack of one phase-2 wave (kMaxBvhDepth + 1 suffices)
static_assert(kBins == 64, "the phase-2 split sweep maps one bin to one lane of a wave");
static_assert(kSmallStack > kMaxBvhDepth, "phase-2 stack");

// Binary node: inner nodes have count 0 and child indices, leaves count > 0.
struct GNode {
    float lo[3], hi[3];
    int left, right;
    int first, count;
    int depth, pad;
};
struct Task {  // a node and its range of `order`
    int node, begin, end, depth;
};
struct Chunk {
    int task, begin, end;
    int first;  // index of the task's first chunk (a task's chunks are consecutive)
};
struct TaskOut {
    int left, right, mid;
    int split;  // axis * 256 + bin, -1 leaf, -2 object median
};

// Accumulators of one large node: bounds (triangle lo, hi, centroid lo, hi: 12
// words) then kBins bins per axis of (count, lo xyz, hi xyz).
constexpr int kAccBounds = 12;
constexpr int kAccWords = kAccBounds + 3 * kBins * 7;

// Floats as uint32 keys whose unsigned order is the float order, so that
// atomicMin / atomicMax take exact float minima and maxima.
__device__ __forceinline__ uint32_t ord(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unord(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}
// std::min / std::max semantics (the host builder's)
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }

__device__ __forceinline__ float half_area(const float* l, const float* h) {
    float dx = h[0] - l[0], dy = h[1] - l[1], dz = h[2] - l[2];
    if (!(dx >= 0.0f)) return 0.0f;
    return dx * dy + dy * dz + dz * dx;
}

__device__ __forceinline__ int bin_of(float c, float cl, float scale) {
    return min(kBins - 1, (int)((c - cl) * scale));
}

// Per-triangle box and centroid, as the host computes them.
__global__ void k_prep(const float* __restrict__ verts, int n, float* tlo, float* thi, float* cen, int* order) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float* v = verts + (size_t)i * 9;
        for (int k = 0; k < 3; ++k) {
            const float a = v[k], b = v[3 + k], c = v[6 + k];
            const float lo = smin(a, smin(b, c)), hi = smax(a, smax(b, c));
            tlo[i * 3 + k] = lo;
            thi[i * 3 + k] = hi;
            cen[i * 3 + k] = 0.5f * lo + 0.5f * hi;
        }
        order[i] = i;
    }
}

__global__ void k_acc_init(uint32_t* acc, int ntasks) {
    const int total = ntasks * kAccWords;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const int w = i % kAccWords;
        uint32_t v;
        if (w < kAccBounds) {
            v = ((w / 3) & 1) ? ord(-INFINITY) : ord(INFINITY);  // lo, hi, clo, chi
        } else {
            const int f = (w - kAccBounds) % 7;
            v = f == 0 ? 0u : (f <= 3 ? ord(INFINITY) : ord(-INFINITY));
        }
        acc[i] = v;
    }
}

// Block-wide reduction of 12 bounds (min for slots 0-2 and 6-8, max otherwise).
__device__ void block_bounds(float* v, float* lds, uint32_t* acc) {
    const int t = threadIdx.x;
    for (int q = 0; q < 12; ++q) lds[q * kBlock + t] = v[q];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (t < s)
            for (int q = 0; q < 12; ++q) {
                const float a = lds[q * kBlock + t], b = lds[q * kBlock + t + s];
                lds[q * kBlock + t] = ((q / 3) & 1) ? smax(a, b) : smin(a, b);
            }
        __syncthreads();
    }
    if (t < 12) {
        const float r = lds[t * kBlock];
        if ((t / 3) & 1) atomicMax(&acc[t], ord(r));
        else atomicMin(&acc[t], ord(r));
    }
}

__global__ void __launch_bounds__(kBlock) k_large_bounds(const Chunk* chunks, const Task* tasks, const int* order,
                                                          const float* tlo, const float* thi, const float* cen,
                                                          uint32_t* acc) {
    __shared__ float lds[12 * kBlock];
    const Chunk c = chunks[blockIdx.x];
    float v[12];
    for (int q = 0; q < 12; ++q) v[q] = ((q / 3) & 1) ? -INFINITY : INFINITY;
    for (int p = c.begin + threadIdx.x; p < c.end; p += kBlock) {
        const int t = order[p];
        for (int k = 0; k < 3; ++k) {
            v[k] = smin(v[k], tlo[t * 3 + k]);
            v[3 + k] = smax(v[3 + k], thi[t * 3 + k]);
            v[6 + k] = smin(v[6 + k], cen[t * 3 + k]);
            v[9 + k] = smax(v[9 + k], cen[t * 3 + k]);
        }
    }
    block_bounds(v, lds, acc + (size_t)c.task * kAccWords);
}

// The node's centroid bounds and, per axis, whether it splits and the bin scale.
struct Axes {
    float cl[3], ch[3], scale[3];
    bool ok[3];
};
__device__ __forceinline__ Axes axes_of(const uint32_t* acc) {
    Axes a;
    for (int k = 0; k < 3; ++k) {
        a.cl[k] = unord(acc[6 + k]);
        a.ch[k] = unord(acc[9 + k]);
        const float extent = a.ch[k] - a.cl[k];
        a.ok[k] = extent > 0.0f;
        a.scale[k] = a.ok[k] ? (float)kBins / extent : 0.0f;
    }
    return a;
}

__global__ void __launch_bounds__(kBlock) k_large_bins(const Chunk* chunks, const int* order, const float* tlo,
                                                        const float* thi, const float* cen, uint32_t* acc) {
    __shared__ uint32_t bins[3 * kBins * 7];
    const Chunk c = chunks[blockIdx.x];
    uint32_t* A = acc + (size_t)c.task * kAccWords;
    for (int i = threadIdx.x; i < 3 * kBins * 7; i += kBlock) {
        const int f = i % 7;
        bins[i] = f == 0 ? 0u : (f <= 3 ? ord(INFINITY) : ord(-INFINITY));
    }
    const Axes ax = axes_of(A);
    __syncthreads();
    for (int p = c.begin + threadIdx.x; p < c.end; p += kBlock) {
        const int t = order[p];
        const uint32_t lo[3] = {ord(tlo[t * 3]), ord(tlo[t * 3 + 1]), ord(tlo[t * 3 + 2])};
        const uint32_t hi[3] = {ord(thi[t * 3]), ord(thi[t * 3 + 1]), ord(thi[t * 3 + 2])};
        for (int k = 0; k < 3; ++k) {
            if (!ax.ok[k]) continue;
            uint32_t* b = bins + (k * kBins + bin_of(cen[t * 3 + k], ax.cl[k], ax.scale[k])) * 7;
            atomicAdd(&b[0], 1u);
            for (int q = 0; q < 3; ++q) {
                atomicMin(&b[1 + q], lo[q]);
                atomicMax(&b[4 + q], hi[q]);
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 3 * kBins; i += kBlock) {
        const uint32_t* b = bins + i * 7;
        if (b[0] == 0) continue;
        uint32_t* g = A + kAccBounds + i * 7;
        atomicAdd(&g[0], b[0]);
        for (int q = 0; q < 3; ++q) {
            atomicMin(&g[1 + q], b[1 + q]);
            atomicMax(&g[4 + q], b[4 + q]);
        }
    }
}

// One axis of the binned SAH sweep (bvh_build.cpp, Builder::build) by one wave:
// `bins` holds the axis's kBins bins (count, then lo xyz and hi xyz as ord keys)
// and lane j takes bin j; the left side of split j is bins 0..j (an inclusive
// prefix scan), the right side bins j+1.. (the inclusive suffix scan of lane
// j + 1). Costs round as the host's (the same operations in the same order) and
// the choice is the host's: the first bin of the least cost, taken over the best
// split so far only when strictly cheaper. All values wave-uniform.
struct Split {
    int axis, bin, lcount;
};
__device__ __forceinline__ void wave_sweep_axis(const uint32_t* bins, int k, float* best_cost, Split* best) {
    const int lane = threadIdx.x & 63;
    const uint32_t* b = bins + lane * 7;
    const int c = (int)b[0];
    int pc = c, sc = c;
    float pl[3], ph[3], sl[3], sh[3];
    for (int q = 0; q < 3; ++q) pl[q] = sl[q] = unord(b[1 + q]), ph[q] = sh[q] = unord(b[4 + q]);
    for (int s = 1; s < 64; s <<= 1) {
        const int upc = __shfl_up(pc, s), dnc = __shfl_down(sc, s);
        float upl[3], uph[3], dnl[3], dnh[3];
        for (int q = 0; q < 3; ++q) {
            upl[q] = __shfl_up(pl[q], s), uph[q] = __shfl_up(ph[q], s);
            dnl[q] = __shfl_down(sl[q], s), dnh[q] = __shfl_down(sh[q], s);
        }
        if (lane >= s) {
            pc += upc;
            for (int q = 0; q < 3; ++q) pl[q] = smin(upl[q], pl[q]), ph[q] = smax(uph[q], ph[q]);
        }
        if (lane + s < 64) {
            sc += dnc;
            for (int q = 0; q < 3; ++q) sl[q] = smin(sl[q], dnl[q]), sh[q] = smax(sh[q], dnh[q]);
        }
    }
    // the right side of split `lane`: the suffix that starts at lane + 1
    const int rc = __shfl_down(sc, 1);
    float rl[3], rh[3];
    for (int q = 0; q < 3; ++q) rl[q] = __shfl_down(sl[q], 1), rh[q] = __shfl_down(sh[q], 1);
    float cost = INFINITY;
    if (lane < kBins - 1 && pc != 0 && rc != 0) cost = half_area(pl, ph) * (float)pc + half_area(rl, rh) * (float)rc;
    // the first minimum over the bins (the smallest bin among equal costs)
    float mc = cost;
    int mj = cost < INFINITY ? lane : 64;
    for (int s = 1; s < 64; s <<= 1) {
        const float oc = __shfl_xor(mc, s);
        const int oj = __shfl_xor(mj, s);
        if (oc < mc || (oc == mc && oj < mj)) mc = oc, mj = oj;
    }
    if (mj < 64 && mc < *best_cost) {
        *best_cost = mc;
        *best = Split{k, mj, __shfl(pc, mj)};
    }
}

// One wave per large node: leaf test, split choice over the node's bins, children.
__global__ void __launch_bounds__(64) k_large_split(const Task* tasks, int ntasks, const uint32_t* acc, GNode* nodes,
                                                    int max_nodes, TaskOut* outs, int* node_counter, int max_leaf,
                                                    int* max_depth) {
    const int i = blockIdx.x;
    if (i >= ntasks) return;
    const Task T = tasks[i];
    const uint32_t* A = acc + (size_t)i * kAccWords;
    const int count = T.end - T.begin;
    const bool leaf = count <= max_leaf || T.depth >= kMaxBvhDepth - 1;
    Split best{-1, -1, 0};
    if (!leaf) {
        const Axes ax = axes_of(A);
        float best_cost = INFINITY;
        for (int k = 0; k < 3; ++k)
            if (ax.ok[k]) wave_sweep_axis(A + kAccBounds + k * kBins * 7, k, &best_cost, &best);
    }
    if (threadIdx.x != 0) return;
    GNode& n = nodes[T.node];
    for (int k = 0; k < 3; ++k) {
        n.lo[k] = unord(A[k]);
        n.hi[k] = unord(A[3 + k]);
    }
    n.depth = T.depth;
    atomicMax(max_depth, T.depth);
    TaskOut o{-1, -1, T.begin, -1};
    const int l = leaf ? 0 : atomicAdd(node_counter, 2);
    if (leaf || l + 1 >= max_nodes) {  // the second: never for a valid count (a binary tree has < 2n nodes)
        n.first = T.begin;
        n.count = count;
        n.left = n.right = -1;
        outs[i] = o;
        return;
    }
    int mid = T.begin + count / 2;  // all centroids coincide: object median
    o.split = -2;
    if (best.axis >= 0) {
        mid = T.begin + best.lcount;
        o.split = best.axis * 256 + best.bin;
    }
    n.left = l;
    n.right = l + 1;
    n.first = -1;
    n.count = 0;
    nodes[l].first = T.begin;
    nodes[l].count = mid - T.begin;
    nodes[l].depth = T.depth + 1;
    nodes[l + 1].first = mid;
    nodes[l + 1].count = T.end - mid;
    nodes[l + 1].depth = T.depth + 1;
    o.left = l;
    o.right = l + 1;
    o.mid = mid;
    outs[i] = o;
}

// The sides of a chunk's triangles under its node's split: thread i holds the
// kPer consecutive positions from c.begin + i * kPer (so thread order is
// position order); returns the thread's count of left-side triangles.
constexpr int kPer = kChunk / kBlock;
__device__ __forceinline__ int chunk_sides(const Chunk& c, const TaskOut& o, const uint32_t* acc, const int* order,
                                           const float* cen, int* tri, bool* left) {
    const int axis = o.split >> 8, split = o.split & 255;
    const Axes ax = axes_of(acc + (size_t)c.task * kAccWords);
    int nl = 0;
    for (int j = 0; j < kPer; ++j) {
        const int p = c.begin + threadIdx.x * kPer + j;
        tri[j] = p < c.end ? order[p] : -1;
        left[j] = tri[j] >= 0 && bin_of(cen[tri[j] * 3 + axis], ax.cl[axis], ax.scale[axis]) <= split;
        nl += left[j];
    }
    return nl;
}

// Left-side triangles of each chunk (0 for leaves and median splits).
__global__ void __launch_bounds__(kBlock) k_large_count(const Chunk* chunks, const TaskOut* outs, const uint32_t* acc,
                                                         const int* order, const float* cen, int* chunk_left) {
    __shared__ int part[kBlock / 64];
    const Chunk c = chunks[blockIdx.x];
    const TaskOut o = outs[c.task];
    if (o.split < 0) {
        if (threadIdx.x == 0) chunk_left[blockIdx.x] = 0;
        return;
    }
    int tri[kPer];
    bool left[kPer];
    int nl = chunk_sides(c, o, acc, order, cen, tri, left);
    for (int s = 1; s < 64; s <<= 1) nl += __shfl_xor(nl, s);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = nl;
    __syncthreads();
    if (threadIdx.x == 0) {
        int sum = 0;
        for (int w = 0; w < kBlock / 64; ++w) sum += part[w];
        chunk_left[blockIdx.x] = sum;
    }
}

// Stable scatter of each chunk's triangles to the two sides of its node's split
// (order -> order2): the left side before the right, each in position order;
// `chunk_off` is the exclusive scan of the chunks' left counts. Leaves and
// median splits keep their positions.
__global__ void __launch_bounds__(kBlock) k_large_partition(const Chunk* chunks, const Task* tasks, const TaskOut* outs,
                                                             const uint32_t* acc, const int* order, int* order2,
                                                             const float* cen, const int* chunk_off) {
    __shared__ int scan_l[kBlock];
    const Chunk c = chunks[blockIdx.x];
    const TaskOut o = outs[c.task];
    if (o.split < 0) {
        for (int p = c.begin + threadIdx.x; p < c.end; p += kBlock) order2[p] = order[p];
        return;
    }
    const Task T = tasks[c.task];
    int tri[kPer];
    bool left[kPer];
    const int nl = chunk_sides(c, o, acc, order, cen, tri, left);
    scan_l[threadIdx.x] = nl;
    __syncthreads();
    for (int s = 1; s < kBlock; s <<= 1) {  // inclusive scan
        const int a = threadIdx.x >= s ? scan_l[threadIdx.x - s] : 0;
        __syncthreads();
        scan_l[threadIdx.x] += a;
        __syncthreads();
    }
    // triangles of this node before this thread's first position, and the left ones among them
    const int before = c.begin - T.begin + (int)threadIdx.x * kPer;
    const int left_before = chunk_off[blockIdx.x] - chunk_off[c.first] + scan_l[threadIdx.x] - nl;
    int wl = T.begin + left_before;
    int wr = o.mid + (min(before, c.end - T.begin) - left_before);
    for (int j = 0; j < kPer; ++j) {
        if (tri[j] < 0) continue;
        if (left[j]) order2[wl++] = tri[j];
        else order2[wr++] = tri[j];
    }
}

__global__ void k_copy_chunks(const Chunk* chunks, const int* src, int* dst) {
    const Chunk c = chunks[blockIdx.x];
    for (int p = c.begin + threadIdx.x; p < c.end; p += blockDim.x) dst[p] = src[p];
}

// ------------------------------------------------------------ phase 2 -------
template <class F>
__device__ __forceinline__ float wave_reduce(float v, F&& op) {
    for (int s = 1; s < 64; s <<= 1) v = op(v, __shfl_xor(v, s));
    return v;
}

// One wave per small node: its whole subtree, depth first, out of LDS. The
// node's triangles (boxes, centroids, ids) are staged once; node ranges are
// ranges of a permutation of them, partitioned stably; the subtree's node
// slots (at most 2 (n - 1) below its root) are reserved with one atomic, and
// the final triangle order is written back at the end.
__global__ void __launch_bounds__(64) k_small(const Task* tasks, int ntasks, GNode* nodes, int max_nodes, int* order,
                                              const float* tlo, const float* thi, const float* cen, int* node_counter,
                                              int max_leaf, int* max_depth) {
    __shared__ float s_box[kSmallMax * 9];  // per staged triangle: lo xyz, hi xyz, centroid xyz
    __shared__ int s_id[kSmallMax];         // its global triangle id
    __shared__ int s_perm[kSmallMax];       // position -> staged triangle
    __shared__ int s_tmp[kSmallMax];
    __shared__ uint32_t bins[kBins * 7];    // one axis at a time
    __shared__ int4 stack[kSmallStack];
    const int lane = threadIdx.x;
    if ((int)blockIdx.x >= ntasks) return;
    const Task T = tasks[blockIdx.x];
    const int base = T.begin, total = T.end - T.begin;  // total <= kSmallMax
    for (int i = lane; i < total; i += 64) {
        const int t = order[base + i];
        s_id[i] = t;
        s_perm[i] = i;
        for (int k = 0; k < 3; ++k) {
            s_box[i * 9 + k] = tlo[t * 3 + k];
            s_box[i * 9 + 3 + k] = thi[t * 3 + k];
            s_box[i * 9 + 6 + k] = cen[t * 3 + k];
        }
    }
    int next = 0;
    if (lane == 0 && total > 1) next = atomicAdd(node_counter, 2 * (total - 1));
    next = __shfl(next, 0);
    const int limit = min(next + 2 * (total - 1), max_nodes);
    int deepest = T.depth;
    if (lane == 0) stack[0] = make_int4(T.node, 0, total, T.depth);
    int sp = 1;
    __syncthreads();
    while (sp > 0) {
        const int4 e = stack[--sp];
        __syncthreads();
        const int node = e.x, b = e.y, end = e.z, depth = e.w;  // positions b..end of the staged triangles
        const int count = end - b;
        deepest = max(deepest, depth);
        float v[12];
        for (int q = 0; q < 12; ++q) v[q] = ((q / 3) & 1) ? -INFINITY : INFINITY;
        for (int p = b + lane; p < end; p += 64) {
            const float* t = s_box + s_perm[p] * 9;
            for (int k = 0; k < 3; ++k) {
                v[k] = smin(v[k], t[k]);
                v[3 + k] = smax(v[3 + k], t[3 + k]);
                v[6 + k] = smin(v[6 + k], t[6 + k]);
                v[9 + k] = smax(v[9 + k], t[6 + k]);
            }
        }
        for (int q = 0; q < 12; ++q)
            v[q] = ((q / 3) & 1) ? wave_reduce(v[q], [](float x, float y) { return smax(x, y); })
                                 : wave_reduce(v[q], [](float x, float y) { return smin(x, y); });
        if (lane == 0) {
            GNode& n = nodes[node];
            for (int k = 0; k < 3; ++k) n.lo[k] = v[k], n.hi[k] = v[3 + k];
            n.depth = depth;
        }
        // (next + 2 > limit never holds for a valid count: a binary tree of n leaves has 2n - 1 nodes)
        if (count <= max_leaf || depth >= kMaxBvhDepth - 1 || next + 2 > limit) {
            if (lane == 0) {
                nodes[node].first = base + b;
                nodes[node].count = count;
                nodes[node].left = nodes[node].right = -1;
            }
            __syncthreads();
            continue;
        }
        float best_cost = INFINITY;
        Split best{-1, -1, 0};
        float cl[3], scale[3];
        for (int k = 0; k < 3; ++k) {
            cl[k] = v[6 + k];
            const float extent = v[9 + k] - v[6 + k];
            if (!(extent > 0.0f)) continue;
            scale[k] = (float)kBins / extent;
            for (int i = lane; i < kBins * 7; i += 64) {
                const int f = i % 7;
                bins[i] = f == 0 ? 0u : (f <= 3 ? ord(INFINITY) : ord(-INFINITY));
            }
            __syncthreads();
            for (int p = b + lane; p < end; p += 64) {
                const float* t = s_box + s_perm[p] * 9;
                uint32_t* bin = bins + bin_of(t[6 + k], cl[k], scale[k]) * 7;
                atomicAdd(&bin[0], 1u);
                for (int q = 0; q < 3; ++q) {
                    atomicMin(&bin[1 + q], ord(t[q]));
                    atomicMax(&bin[4 + q], ord(t[3 + q]));
                }
            }
            __syncthreads();
            wave_sweep_axis(bins, k, &best_cost, &best);
            __syncthreads();
        }
        int mid = b + count / 2;  // all centroids coincide: object median
        if (best.axis >= 0) {
            const int k = best.axis;
            mid = b + best.lcount;
            // stable partition of the positions into s_tmp, then back
            int wl = b, wr = mid;
            for (int p0 = b; p0 < end; p0 += 64) {
                const int p = p0 + lane;
                const int t = p < end ? s_perm[p] : -1;
                const bool lf = t >= 0 && bin_of(s_box[t * 9 + 6 + k], cl[k], scale[k]) <= best.bin;
                const uint64_t ml = __ballot(lf), mr = __ballot(t >= 0 && !lf);
                const uint64_t below = (1ull << lane) - 1ull;
                if (t >= 0) {
                    if (lf) s_tmp[wl + __popcll(ml & below)] = t;
                    else s_tmp[wr + __popcll(mr & below)] = t;
                }
                wl += __popcll(ml);
                wr += __popcll(mr);
            }
            __syncthreads();
            for (int p = b + lane; p < end; p += 64) s_perm[p] = s_tmp[p];
        }
        const int l = next;
        next += 2;
        if (lane == 0) {
            nodes[node].left = l;
            nodes[node].right = l + 1;
            nodes[node].first = -1;
            nodes[node].count = 0;
            stack[sp] = make_int4(l + 1, mid, end, depth + 1);
            stack[sp + 1] = make_int4(l, b, mid, depth + 1);
        }
        // children are pushed only below depth kMaxBvhDepth - 1, which bounds sp by kMaxBvhDepth
        sp += 2;
        __syncthreads();
    }
    for (int i = lane; i < total; i += 64) order[base + i] = s_id[s_perm[i]];
    if (lane == 0) atomicMax(max_depth, deepest);
}

// ------------------------------------------------------------ phase 3 -------
// The host's collapse of one binary node into up to four children: expand the
// inner child of the largest half area (the first among equals) until four.
__global__ void k_collapse_expand(const int* level, int n, const GNode* nodes, int* kids, int* nkids, int* ninner) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int bn = level[i];
    int k4[4] = {nodes[bn].left, nodes[bn].right, -1, -1};
    int k = 2;
    while (k < 4) {
        int best = -1;
        float best_area = -1.0f;
        for (int j = 0; j < k; ++j) {
            const GNode& c = nodes[k4[j]];
            if (c.count > 0) continue;
            const float a = half_area(c.lo, c.hi);
            if (a > best_area) best_area = a, best = j;
        }
        if (best < 0) break;
        const int c = k4[best];
        k4[best] = nodes[c].left;
        k4[k++] = nodes[c].right;
    }
    int inner = 0;
    for (int j = 0; j < 4; ++j) {
        kids[i * 4 + j] = k4[j];
        if (j < k && nodes[k4[j]].count == 0) ++inner;
    }
    nkids[i] = k;
    ninner[i] = inner;
}

// Writes level node i (device node `base + i`) in the 4-wide layout; inner
// children get device nodes next_base + offset[i] + rank and enter the next level.
__global__ void k_collapse_write(int n, int base, int next_base, const int* kids, const int* nkids, const int* offset,
                                 const GNode* nodes, float* out, int* next_level, int* wide_kids) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float* node = out + (size_t)(base + i) * 32;
    int* ints = reinterpret_cast<int*>(node + 24);
    int r = 0;
    for (int s = 0; s < 4; ++s) {
        int idx = 0, cnt = -1, wide = -1;
        if (s < nkids[i]) {
            const GNode& c = nodes[kids[i * 4 + s]];
            for (int a = 0; a < 3; ++a) {
                node[(2 * a) * 4 + s] = c.lo[a];
                node[(2 * a + 1) * 4 + s] = c.hi[a];
            }
            if (c.count > 0) {
                idx = c.first;
                cnt = c.count;
            } else {
                wide = next_base + offset[i] + r;
                next_level[offset[i] + r] = kids[i * 4 + s];
                ++r;
                idx = wide;
                cnt = 0;
            }
        } else {
            for (int a = 0; a < 3; ++a) {
                node[(2 * a) * 4 + s] = INFINITY;
                node[(2 * a + 1) * 4 + s] = INFINITY;
            }
        }
        ints[s] = idx;
        ints[4 + s] = cnt;
        wide_kids[(size_t)(base + i) * 4 + s] = wide;
    }
}

// The exact traversal stack bound, bottom-up: a node pushes all but one of its
// inner children.
__global__ void k_stack_need(int base, int n, const int* wide_kids, int* need) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int w = base + i;
    int inner = 0, deepest = 0;
    for (int s = 0; s < 4; ++s) {
        const int c = wide_kids[(size_t)w * 4 + s];
        if (c < 0) continue;
        ++inner;
        deepest = max(deepest, need[c]);
    }
    need[w] = max(0, inner - 1) + deepest;
}

// Exclusive scan of n ints (n <= 1024 * 1024): per-block scans, a scan of the
// block sums, then the offsets added.
constexpr int kScanBlock = 1024;
__global__ void __launch_bounds__(kScanBlock) k_scan_blocks(const int* in, int n, int* out, int* sums) {
    __shared__ int s[kScanBlock];
    const int i = blockIdx.x * kScanBlock + threadIdx.x;
    const int v = i < n ? in[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < kScanBlock; d <<= 1) {
        const int a = threadIdx.x >= d ? s[threadIdx.x - d] : 0;
        __syncthreads();
        s[threadIdx.x] += a;
        __syncthreads();
    }
    if (i < n) out[i] = s[threadIdx.x] - v;
    if (threadIdx.x == kScanBlock - 1) sums[blockIdx.x] = s[threadIdx.x];
}
__global__ void __launch_bounds__(kScanBlock) k_scan_sums(int* sums, int nb, int* total) {
    __shared__ int s[kScanBlock];
    const int v = (int)threadIdx.x < nb ? sums[threadIdx.x] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < kScanBlock; d <<= 1) {
        const int a = threadIdx.x >= d ? s[threadIdx.x - d] : 0;
        __syncthreads();
        s[threadIdx.x] += a;
        __syncthreads();
    }
    if ((int)threadIdx.x < nb) sums[threadIdx.x] = s[threadIdx.x] - v;
    if (threadIdx.x == kScanBlock - 1) *total = s[threadIdx.x];
}
__global__ void k_scan_add(int* out, int n, const int* sums) {
    const int i = blockIdx.x * kScanBlock + threadIdx.x;
    if (i < n) out[i] += sums[blockIdx.x];
}

// ------------------------------------------------------------ host side -----
struct DeviceBuild {
    hipStream_t st;
    std::string* err;
    std::vector<void*> allocs;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    bool ok = true;

    ~DeviceBuild() {
        for (void* p : allocs) (void)hipFree(p);
        for (auto& e : events) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
    }
    bool check(hipError_t e, const char* what) {
        if (e == hipSuccess) return true;
        if (ok) *err = std::string(what) + " failed: " + hipGetErrorString(e);
        ok = false;
        return false;
    }
    template <class T>
    T* alloc(size_t count) {
        void* p = nullptr;
        if (!check(hipMalloc(&p, std::max<size_t>(1, count) * sizeof(T)), "hipMalloc")) return nullptr;
        allocs.push_back(p);
        return static_cast<T*>(p);
    }
    // Runs `launch` between two events: the build's kernel time is the sum.
    template <class F>
    void timed(F&& launch) {
        hipEvent_t a, b;
        if (!check(hipEventCreate(&a), "hipEventCreate") || !check(hipEventCreate(&b), "hipEventCreate")) return;
        events.push_back({a, b});
        (void)hipEventRecord(a, st);
        launch();
        (void)hipEventRecord(b, st);
        check(hipGetLastError(), "kernel launch");
    }
    double kernel_ms() {
        double sum = 0.0;
        for (auto& e : events) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, e.first, e.second) == hipSuccess) sum += ms;
        }
        return sum;
    }
    // exclusive scan of n ints into out; returns the total (synchronises)
    // exclusive scan of n ints into out, enqueued only (the total lands in *total)
    void scan_async(const int* in, int n, int* out, int* sums, int* total) {
        if (n <= 0) return;
        const int nb = (n + kScanBlock - 1) / kScanBlock;
        if (nb > kScanBlock) {
            check(hipErrorInvalidValue, "device BVH build: scan of more than 2^20 entries");
            return;
        }
        timed([&] { hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(kScanBlock), 0, st, in, n, out, sums); });
        timed([&] { hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanBlock), 0, st, sums, nb, total); });
        timed([&] { hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kScanBlock), 0, st, out, n, sums); });
    }
    int scan(const int* in, int n, int* out, int* sums, int* total) {
        if (n <= 0) return 0;
        scan_async(in, n, out, sums, total);
        if (!ok) return 0;
        int t = 0;
        check(hipMemcpyAsync(&t, total, sizeof(int), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
        check(hipStreamSynchronize(st), "hipStreamSynchronize");
        return t;
    }
};

static unsigned grid_for(size_t n, unsigned block) { return (unsigned)std::max<size_t>(1, (n + block - 1) / block); }

}  // namespace

bool build_bvh_device(void* stream, const BvhInput& in, int max_leaf, BvhOutput* out, double* device_ms,
                      std::string* error) {
    const int n = in.num_tris;
    DeviceBuild D;
    D.st = static_cast<hipStream_t>(stream);
    D.err = error;
    *device_ms = 0.0;
    out->nodes.clear();
    out->leaf_order.assign(n, 0);
    if (n <= 0) {  // an empty tree: one node of absent slots
        out->nodes.assign(32, 0.0f);
        float* node = out->nodes.data();
        int* ints = reinterpret_cast<int*>(node + 24);
        for (int s = 0; s < 4; ++s) {
            for (int a = 0; a < 6; ++a) node[a * 4 + s] = INFINITY;
            ints[s] = 0;
            ints[4 + s] = -1;
        }
        out->stack_need = 1;
        out->depth4 = 0;
        out->depth = 0;
        for (int k = 0; k < 3; ++k) out->scene_lo[k] = out->scene_hi[k] = 0.0f;
        return true;
    }
    const int max_nodes = std::max(1, 2 * n);
    float* d_verts = D.alloc<float>((size_t)n * 9);
    float* tlo = D.alloc<float>((size_t)n * 3);
    float* thi = D.alloc<float>((size_t)n * 3);
    float* cen = D.alloc<float>((size_t)n * 3);
    int* order = D.alloc<int>(n);
    int* order2 = D.alloc<int>(n);
    GNode* nodes = D.alloc<GNode>(max_nodes);
    int* counters = D.alloc<int>(4);  // node counter, max depth
    if (!D.ok) return false;
    D.check(hipMemcpyAsync(d_verts, in.tri_vertices, (size_t)n * 9 * sizeof(float), hipMemcpyHostToDevice, D.st),
            "hipMemcpyAsync");
    const int init_counters[4] = {1, 0, 0, 0};
    D.check(hipMemcpyAsync(counters, init_counters, sizeof(init_counters), hipMemcpyHostToDevice, D.st),
            "hipMemcpyAsync");
    D.timed([&] {
        hipLaunchKernelGGL(k_prep, dim3(std::min(grid_for(n, 256), 4096u)), dim3(256), 0, D.st, d_verts, n, tlo, thi,
                           cen, order);
    });

    // ---- phase 1: large nodes, level by level
    std::vector<Task> small;
    std::vector<Task> level;
    if (n > kSmallMax) level.push_back(Task{0, 0, n, 0});
    else small.push_back(Task{0, 0, n, 0});
    Task* d_tasks = nullptr;
    Chunk* d_chunks = nullptr;
    TaskOut* d_outs = nullptr;
    uint32_t* d_acc = nullptr;
    int* d_chunk_left = nullptr;
    int* d_chunk_off = nullptr;
    int* d_scan_sums = D.alloc<int>(kScanBlock);
    int* d_scan_total = D.alloc<int>(1);
    size_t task_cap = 0, chunk_cap = 0;
    while (!level.empty() && D.ok) {
        const int nt = (int)level.size();
        std::vector<Chunk> chunks;
        for (int t = 0, first = 0; t < nt; ++t, first = (int)chunks.size())
            for (int b = level[t].begin; b < level[t].end; b += kChunk)
                chunks.push_back(Chunk{t, b, std::min(level[t].end, b + kChunk), first});
        const int nc = (int)chunks.size();
        if ((size_t)nt > task_cap) {
            task_cap = std::max<size_t>(nt, task_cap * 2);
            d_tasks = D.alloc<Task>(task_cap);
            d_outs = D.alloc<TaskOut>(task_cap);
            d_acc = D.alloc<uint32_t>(task_cap * kAccWords);
        }
        if ((size_t)nc > chunk_cap) {
            chunk_cap = std::max<size_t>(nc, chunk_cap * 2);
            d_chunks = D.alloc<Chunk>(chunk_cap);
            d_chunk_left = D.alloc<int>(chunk_cap);
            d_chunk_off = D.alloc<int>(chunk_cap);
        }
        if (!D.ok) break;
        D.check(hipMemcpyAsync(d_tasks, level.data(), nt * sizeof(Task), hipMemcpyHostToDevice, D.st), "hipMemcpyAsync");
        D.check(hipMemcpyAsync(d_chunks, chunks.data(), nc * sizeof(Chunk), hipMemcpyHostToDevice, D.st),
                "hipMemcpyAsync");
        D.timed([&] {
            hipLaunchKernelGGL(k_acc_init, dim3(std::min(grid_for((size_t)nt * kAccWords, 256), 4096u)), dim3(256), 0,
                               D.st, d_acc, nt);
        });
        D.timed([&] {
            hipLaunchKernelGGL(k_large_bounds, dim3(nc), dim3(kBlock), 0, D.st, d_chunks, d_tasks, order, tlo, thi,
                               cen, d_acc);
        });
        D.timed([&] {
            hipLaunchKernelGGL(k_large_bins, dim3(nc), dim3(kBlock), 0, D.st, d_chunks, order, tlo, thi, cen, d_acc);
        });
        D.timed([&] {
            hipLaunchKernelGGL(k_large_split, dim3(nt), dim3(64), 0, D.st, d_tasks, nt, d_acc, nodes,
                               max_nodes, d_outs, counters, max_leaf, counters + 1);
        });
        D.timed([&] {
            hipLaunchKernelGGL(k_large_count, dim3(nc), dim3(kBlock), 0, D.st, d_chunks, d_outs, d_acc, order, cen,
                               d_chunk_left);
        });
        D.scan_async(d_chunk_left, nc, d_chunk_off, d_scan_sums, d_scan_total);
        D.timed([&] {
            hipLaunchKernelGGL(k_large_partition, dim3(nc), dim3(kBlock), 0, D.st, d_chunks, d_tasks, d_outs, d_acc,
                               order, order2, cen, d_chunk_off);
        });
        D.timed([&] { hipLaunchKernelGGL(k_copy_chunks, dim3(nc), dim3(kBlock), 0, D.st, d_chunks, order2, order); });
        std::vector<TaskOut> outs(nt);
        D.check(hipMemcpyAsync(outs.data(), d_outs, nt * sizeof(TaskOut), hipMemcpyDeviceToHost, D.st),
                "hipMemcpyAsync");
        D.check(hipStreamSynchronize(D.st), "hipStreamSynchronize");
        if (!D.ok) break;
        std::vector<Task> next;
        for (int t = 0; t < nt; ++t) {
            const TaskOut& o = outs[t];
            if (o.split == -1) continue;  // a leaf
            const Task kids[2] = {Task{o.left, level[t].begin, o.mid, level[t].depth + 1},
                                  Task{o.right, o.mid, level[t].end, level[t].depth + 1}};
            for (const Task& k : kids) (k.end - k.begin > kSmallMax ? next : small).push_back(k);
        }
        level.swap(next);
    }
    if (!D.ok) return false;

    // ---- phase 2: small nodes, one wave each
    if (!small.empty()) {
        Task* d_small = D.alloc<Task>(small.size());
        if (!D.ok) return false;
        D.check(hipMemcpyAsync(d_small, small.data(), small.size() * sizeof(Task), hipMemcpyHostToDevice, D.st),
                "hipMemcpyAsync");
        D.timed([&] {
            hipLaunchKernelGGL(k_small, dim3((unsigned)small.size()), dim3(64), 0, D.st, d_small, (int)small.size(),
                               nodes, max_nodes, order, tlo, thi, cen, counters, max_leaf, counters + 1);
        });
    }
    int host_counters[4] = {0, 0, 0, 0};
    D.check(hipMemcpyAsync(host_counters, counters, sizeof(host_counters), hipMemcpyDeviceToHost, D.st),
            "hipMemcpyAsync");
    GNode root{};
    D.check(hipMemcpyAsync(&root, nodes, sizeof(GNode), hipMemcpyDeviceToHost, D.st), "hipMemcpyAsync");
    D.check(hipStreamSynchronize(D.st), "hipStreamSynchronize");
    if (!D.ok) return false;
    const int used = host_counters[0];
    if (used > max_nodes) {
        *error = "device BVH build ran out of nodes";
        return false;
    }
    out->depth = host_counters[1];
    for (int k = 0; k < 3; ++k) out->scene_lo[k] = root.lo[k], out->scene_hi[k] = root.hi[k];

    // ---- phase 3: the 4-wide collapse, breadth first
    const bool root_is_leaf = root.count > 0;
    int* d_lvl[2] = {D.alloc<int>(used), D.alloc<int>(used)};
    int* d_kids = D.alloc<int>((size_t)used * 4);
    int* d_nkids = D.alloc<int>(used);
    int* d_ninner = D.alloc<int>(used);
    int* d_offset = D.alloc<int>(used);
    int* d_sums = D.alloc<int>(kScanBlock);
    int* d_total = D.alloc<int>(1);
    float* d_out = D.alloc<float>((size_t)used * 32);
    int* d_wide_kids = D.alloc<int>((size_t)used * 4);
    int* d_need = D.alloc<int>(used);
    if (!D.ok) return false;
    int dev_count = 1;
    std::vector<std::pair<int, int>> levels;  // (base, size) of each level of device nodes
    if (root_is_leaf) {
        std::vector<float> node(32, 0.0f);
        int* ints = reinterpret_cast<int*>(node.data() + 24);
        for (int s = 0; s < 4; ++s) {
            for (int a = 0; a < 3; ++a) {
                node[(2 * a) * 4 + s] = s == 0 ? root.lo[a] : INFINITY;
                node[(2 * a + 1) * 4 + s] = s == 0 ? root.hi[a] : INFINITY;
            }
            ints[s] = s == 0 ? root.first : 0;
            ints[4 + s] = s == 0 ? root.count : -1;
        }
        out->nodes = node;
        out->stack_need = 1;
        out->depth4 = 0;
    } else {
        const int zero = 0;
        D.check(hipMemcpyAsync(d_lvl[0], &zero, sizeof(int), hipMemcpyHostToDevice, D.st), "hipMemcpyAsync");
        int cur = 0, base = 0, size = 1;
        while (size > 0 && D.ok) {
            levels.push_back({base, size});
            D.timed([&] {
                hipLaunchKernelGGL(k_collapse_expand, dim3(grid_for(size, 256)), dim3(256), 0, D.st, d_lvl[cur], size,
                                   nodes, d_kids, d_nkids, d_ninner);
            });
            const int next = D.scan(d_ninner, size, d_offset, d_sums, d_total);
            if (!D.ok) return false;
            const int next_base = base + size;
            if (next_base + next > used) {
                *error = "device BVH collapse overflow";
                return false;
            }
            D.timed([&] {
                hipLaunchKernelGGL(k_collapse_write, dim3(grid_for(size, 256)), dim3(256), 0, D.st, size, base,
                                   next_base, d_kids, d_nkids, d_offset, nodes, d_out, d_lvl[cur ^ 1], d_wide_kids);
            });
            cur ^= 1;
            base = next_base;
            size = next;
        }
        dev_count = base;
        for (size_t L = levels.size(); L-- > 0;) {
            D.timed([&] {
                hipLaunchKernelGGL(k_stack_need, dim3(grid_for(levels[L].second, 256)), dim3(256), 0, D.st,
                                   levels[L].first, levels[L].second, d_wide_kids, d_need);
            });
        }
        int need0 = 0;
        out->nodes.resize((size_t)dev_count * 32);
        D.check(hipMemcpyAsync(out->nodes.data(), d_out, (size_t)dev_count * 32 * sizeof(float), hipMemcpyDeviceToHost,
                               D.st),
                "hipMemcpyAsync");
        D.check(hipMemcpyAsync(&need0, d_need, sizeof(int), hipMemcpyDeviceToHost, D.st), "hipMemcpyAsync");
        D.check(hipStreamSynchronize(D.st), "hipStreamSynchronize");
        out->stack_need = need0 + 1;
        out->depth4 = (int)levels.size() - 1;
    }
    D.check(hipMemcpyAsync(out->leaf_order.data(), order, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, D.st),
            "hipMemcpyAsync");
    D.check(hipStreamSynchronize(D.st), "hipStreamSynchronize");
    if (!D.ok) return false;
    *device_ms = D.kernel_ms();
    return true;
}

}  // namespace sspt
