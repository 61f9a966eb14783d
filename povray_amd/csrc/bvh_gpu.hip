// GPU BVH builder.
//
// The reference builds its bounding-slab tree serially on the host
// (Build_BBox_Tree / sort_and_split, core/bounding/boundingbox.cpp:262-430,
// 834-927: largest-extent axis, sort by box centre, prefix/suffix surface-area
// sweep).  Here the tree is built on the device as a linear BVH:
//   1. centroid bounds (block reduction, ordered-int atomics);
//   2. 63-bit Morton codes of the primitive centroids, radix-sorted (rocPRIM);
//   3. the binary radix tree over the sorted codes (Karras, "Maximizing
//      Parallelism in the Construction of BVHs, Octrees, and k-d Trees",
//      HPG 2012): one thread per internal node, equal codes split by index;
//   4. bottom-up bounds and heights in passes: a pass completes every node
//      whose children were completed by earlier passes (kernel boundaries
//      order the passes, so no atomics or fences are needed; an agent-scope
//      climb with acquire/release counters cost 5.3 ms of L2 write-backs on
//      1.3 M primitives, the passes about 0.5 ms);
//   5. every internal node covering at most max_leaf primitives becomes a
//      leaf over its (contiguous) range of sorted primitives; the remaining
//      nodes are compacted by an exclusive scan into the 64 B binary layout;
//   6. optionally the 4-wide layout: level by level from the root, each wide
//      node takes its binary node's two children and opens the inner child
//      of largest surface area until it holds four (the host collapse's
//      rule); children are numbered by a scan, so the layout is breadth-first
//      and the same on every run.
// Any tree over the primitives gives the same closest hits (ties are broken
// by object and primitive id), so the build changes speed, not results.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cstring>

#include "bvh_gpu.h"
#include "internal.h"

namespace pvx {
namespace {

constexpr int kB = 256;
constexpr uint32_t kLeafBit = 0x80000000u;

__device__ inline uint32_t ord_of(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float f_of(uint32_t u) { return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u); }

// 21 bits -> every third bit of 63.
__host__ __device__ inline uint64_t spread3(uint32_t a) {
    uint64_t x = a & 0x1FFFFFu;
    x = (x | x << 32) & 0x1F00000000FFFFull;
    x = (x | x << 16) & 0x1F0000FF0000FFull;
    x = (x | x << 8) & 0x100F00F00F00F00Full;
    x = (x | x << 4) & 0x10C30C30C30C30C3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

__device__ inline int32_t leaf_link(uint32_t start, uint32_t count) {
    return ~(int32_t)((start << 3) | (count - 1));
}

__global__ void k_bounds_init(uint32_t *b) {
    if (threadIdx.x < 6) b[threadIdx.x] = threadIdx.x < 3 ? 0xFFFFFFFFu : 0u;
}

__global__ void __launch_bounds__(kB) k_centroid_bounds(const float *lo, const float *hi, uint32_t n, uint32_t *b) {
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = (size_t)blockIdx.x * kB + threadIdx.x; i < n; i += (size_t)gridDim.x * kB)
        for (int a = 0; a < 3; a++) {
            const float c = 0.5f * (lo[3 * i + a] + hi[3 * i + a]);
            mn[a] = fminf(mn[a], c);
            mx[a] = fmaxf(mx[a], c);
        }
    for (int off = 32; off > 0; off >>= 1)
        for (int a = 0; a < 3; a++) {
            mn[a] = fminf(mn[a], __shfl_xor(mn[a], off, 64));
            mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], off, 64));
        }
    __shared__ float smn[3][kB / 64], smx[3][kB / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
        for (int a = 0; a < 3; a++) {
            smn[a][w] = mn[a];
            smx[a][w] = mx[a];
        }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int a = threadIdx.x;
        float vmn = smn[a][0], vmx = smx[a][0];
        for (int k = 1; k < kB / 64; k++) {
            vmn = fminf(vmn, smn[a][k]);
            vmx = fmaxf(vmx, smx[a][k]);
        }
        atomicMin(b + a, ord_of(vmn));
        atomicMax(b + 3 + a, ord_of(vmx));
    }
}

__global__ void __launch_bounds__(kB) k_morton(const float *lo, const float *hi, uint32_t n, const uint32_t *b,
                                               uint64_t *keys, uint32_t *vals) {
    const size_t i = (size_t)blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    uint32_t q[3];
    for (int a = 0; a < 3; a++) {
        const float cmin = f_of(b[a]), cmax = f_of(b[3 + a]);
        const float ext = cmax - cmin;
        const float c = 0.5f * (lo[3 * i + a] + hi[3 * i + a]);
        const float t = ext > 0.0f ? (c - cmin) / ext : 0.0f;
        q[a] = (uint32_t)fminf(fmaxf(t * 2097152.0f, 0.0f), 2097151.0f);
    }
    keys[i] = (spread3(q[0]) << 2) | (spread3(q[1]) << 1) | spread3(q[2]);
    vals[i] = (uint32_t)i;
}

__device__ inline int delta(const uint64_t *k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const uint64_t a = k[i], b = k[j];
    return a == b ? 64 + __clz(i ^ j) : __clzll((long long)(a ^ b));
}

// Internal node i of the radix tree over n >= 2 sorted codes (root = 0).
// child: the two children, leaf positions tagged with kLeafBit; range: the
// sorted positions the node covers.
__global__ void __launch_bounds__(kB) k_radix_tree(const uint64_t *keys, int n, int2 *child, int2 *range) {
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= n - 1) return;
    const int d = delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1) > 0 ? 1 : -1;
    const int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(keys, n, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + min(d, 0);
    const int first = min(i, j), last = max(i, j);
    const bool lleaf = first == gamma, rleaf = last == gamma + 1;
    child[i] = make_int2(lleaf ? (int)((uint32_t)gamma | kLeafBit) : gamma,
                         rleaf ? (int)((uint32_t)(gamma + 1) | kLeafBit) : gamma + 1);
    range[i] = make_int2(first, last);
}

__global__ void __launch_bounds__(kB) k_fill_i32(int *p, size_t n, int v) {
    const size_t i = (size_t)blockIdx.x * kB + threadIdx.x;
    if (i < n) p[i] = v;
}

// One bottom-up pass: every internal node not done yet whose internal
// children were done by an earlier pass gets its bounds and height (0 for a
// node that becomes a leaf, at most max_leaf primitives; else 1 + the larger
// child height).  done[i] = the pass that completed node i, -1 before; a
// child completed in this same pass is not used (its bounds may not be
// visible yet), it is in the next one.
__global__ void __launch_bounds__(kB) k_refit_pass(const float *lo, const float *hi, const uint32_t *vals, int n,
                                                   const int2 *child, const int2 *range, int max_leaf, float *ibox,
                                                   int *iheight, int *done, int pass) {
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= n - 1 || done[i] >= 0) return;
    const int2 c = child[i];
    for (int s = 0; s < 2; s++) {
        const int cc = s ? c.y : c.x;
        if (!((uint32_t)cc & kLeafBit)) {
            const int dp = done[cc];
            if (dp < 0 || dp == pass) return;
        }
    }
    float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    int h = 0;
    for (int s = 0; s < 2; s++) {
        const int cc = s ? c.y : c.x;
        if ((uint32_t)cc & kLeafBit) {
            const size_t pid = vals[(uint32_t)cc & ~kLeafBit];
            for (int a = 0; a < 3; a++) {
                b[a] = fminf(b[a], lo[3 * pid + a]);
                b[3 + a] = fmaxf(b[3 + a], hi[3 * pid + a]);
            }
        } else {
            for (int a = 0; a < 3; a++) {
                b[a] = fminf(b[a], ibox[6 * (size_t)cc + a]);
                b[3 + a] = fmaxf(b[3 + a], ibox[6 * (size_t)cc + 3 + a]);
            }
            h = max(h, iheight[cc]);
        }
    }
    for (int q = 0; q < 6; q++) ibox[6 * (size_t)i + q] = b[q];
    const int2 r = range[i];
    iheight[i] = r.y - r.x + 1 <= max_leaf ? 0 : h + 1;
    done[i] = pass;
}

// keep[i] = 1 for internal nodes that stay nodes; keep[n - 1] = 0 closes the
// scan so that its last entry is the kept count.
__global__ void __launch_bounds__(kB) k_keep(const int2 *range, int n, int max_leaf, uint32_t *keep) {
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    keep[i] = i < n - 1 && range[i].y - range[i].x + 1 > max_leaf ? 1u : 0u;
}

__global__ void __launch_bounds__(kB) k_emit_nodes(const int2 *child, const int2 *range, const float *ibox,
                                                   const float *lo, const float *hi, const uint32_t *vals,
                                                   const uint32_t *keep, const uint32_t *newidx, int n,
                                                   Node *nodes) {
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= n - 1 || !keep[i]) return;
    const int2 c = child[i];
    float bx[2][6];
    int32_t link[2];
    for (int s = 0; s < 2; s++) {
        const int cc = s ? c.y : c.x;
        if ((uint32_t)cc & kLeafBit) {
            const uint32_t pos = (uint32_t)cc & ~kLeafBit;
            const size_t pid = vals[pos];
            for (int a = 0; a < 3; a++) {
                bx[s][a] = lo[3 * pid + a];
                bx[s][3 + a] = hi[3 * pid + a];
            }
            link[s] = leaf_link(pos, 1);
        } else {
            for (int q = 0; q < 6; q++) bx[s][q] = ibox[6 * (size_t)cc + q];
            const int2 r = range[cc];
            link[s] = keep[cc] ? (int32_t)newidx[cc] : leaf_link((uint32_t)r.x, (uint32_t)(r.y - r.x + 1));
        }
    }
    Node o;
    o.lo0x = bx[0][0]; o.hi0x = bx[0][3]; o.lo0y = bx[0][1]; o.hi0y = bx[0][4];
    o.lo1x = bx[1][0]; o.hi1x = bx[1][3]; o.lo1y = bx[1][1]; o.hi1y = bx[1][4];
    o.lo0z = bx[0][2]; o.hi0z = bx[0][5]; o.lo1z = bx[1][2]; o.hi1z = bx[1][5];
    o.c0 = link[0];
    o.c1 = link[1];
    o.pad0 = 0;
    o.pad1 = 0;
    nodes[newidx[i]] = o;
}

__global__ void __launch_bounds__(kB) k_gather_prims(const PrimRec *in, const uint32_t *vals, uint32_t n,
                                                     PrimRec *out) {
    const size_t k = (size_t)blockIdx.x * kB + threadIdx.x;
    if (k >= n) return;
    const uint4 *src = reinterpret_cast<const uint4 *>(in + vals[k]);
    uint4 *dst = reinterpret_cast<uint4 *>(out + k);
    for (int q = 0; q < 4; q++) dst[q] = src[q];
}

__global__ void __launch_bounds__(kB) k_iota(uint32_t *v, uint32_t n) {
    const size_t k = (size_t)blockIdx.x * kB + threadIdx.x;
    if (k < n) v[k] = (uint32_t)k;
}

// ---- 4-wide collapse -----------------------------------------------------------
struct Cand {
    int32_t link;
    float lo[3], hi[3];
};

__device__ inline float cand_area(const Cand &c) {
    const float dx = fmaxf(0.0f, c.hi[0] - c.lo[0]), dy = fmaxf(0.0f, c.hi[1] - c.lo[1]),
                dz = fmaxf(0.0f, c.hi[2] - c.lo[2]);
    return dx * dy + dy * dz + dz * dx;
}

__device__ inline void node_cands(const Node &m, Cand &a, Cand &b) {
    a.link = m.c0;
    a.lo[0] = m.lo0x; a.lo[1] = m.lo0y; a.lo[2] = m.lo0z;
    a.hi[0] = m.hi0x; a.hi[1] = m.hi0y; a.hi[2] = m.hi0z;
    b.link = m.c1;
    b.lo[0] = m.lo1x; b.lo[1] = m.lo1y; b.lo[2] = m.lo1z;
    b.hi[0] = m.hi1x; b.hi[1] = m.hi1y; b.hi[2] = m.hi1z;
}

// The children of the wide node rooted at binary node `bn`.
__device__ inline int wide_children(const Node *nodes, int32_t bn, Cand c[4]) {
    node_cands(nodes[bn], c[0], c[1]);
    int cnt = 2;
    while (cnt < 4) {
        int best = -1;
        float ba = -1.0f;
        for (int k = 0; k < cnt; k++)
            if (c[k].link >= 0 && cand_area(c[k]) > ba) {
                ba = cand_area(c[k]);
                best = k;
            }
        if (best < 0) break;
        Cand a, b;
        node_cands(nodes[c[best].link], a, b);
        c[best] = a;
        c[cnt++] = b;
    }
    return cnt;
}

__global__ void __launch_bounds__(kB) k_wide_count(const Node *nodes, const int32_t *frontier, uint32_t m,
                                                   uint32_t *cnt) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i == 0) cnt[m] = 0;
    if (i >= m) return;
    Cand c[4];
    const int k = wide_children(nodes, frontier[i], c);
    uint32_t inner = 0;
    for (int q = 0; q < k; q++) inner += c[q].link >= 0 ? 1u : 0u;
    cnt[i] = inner;
}

// Wide node base + i from frontier[i]; its inner children become wide nodes
// base + m + offs[i] + j and entries offs[i] + j of the next frontier.
__global__ void __launch_bounds__(kB) k_wide_emit(const Node *nodes, const int32_t *frontier, uint32_t m,
                                                  const uint32_t *offs, uint32_t base, Node4 *nodes4,
                                                  int32_t *next) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= m) return;
    Cand c[4];
    const int k = wide_children(nodes, frontier[i], c);
    uint32_t o = offs[i];
    Node4 w;
    for (int q = 0; q < 4; q++) {
        const bool v = q < k;
        int32_t link = INT32_MIN;
        if (v) {
            if (c[q].link >= 0) {
                next[o] = c[q].link;
                link = (int32_t)(base + m + o);
                o++;
            } else {
                link = c[q].link;
            }
        }
        w.lox[q] = v ? c[q].lo[0] : 0.0f;
        w.loy[q] = v ? c[q].lo[1] : 0.0f;
        w.loz[q] = v ? c[q].lo[2] : 0.0f;
        w.hix[q] = v ? c[q].hi[0] : 0.0f;
        w.hiy[q] = v ? c[q].hi[1] : 0.0f;
        w.hiz[q] = v ? c[q].hi[2] : 0.0f;
        w.c[q] = link;
        w.pad[q] = 0;
    }
    nodes4[base + i] = w;
}

inline unsigned blocks(size_t n) { return (unsigned)((n + kB - 1) / kB); }

// Temporaries from the stream-ordered allocator (no device-wide wait on
// free), released on every return path after the work queued on `st`.
struct Temps {
    hipStream_t st;
    std::vector<void *> p;
    explicit Temps(hipStream_t s) : st(s) {}
    ~Temps() {
        for (void *q : p) (void)hipFreeAsync(q, st);
        (void)hipStreamSynchronize(st);
    }
    template <class T>
    int get(T **out, size_t count) {
        void *q = nullptr;
        HIP_TRY(hipMallocAsync(&q, std::max<size_t>(count, 1) * sizeof(T), st));
        p.push_back(q);
        *out = static_cast<T *>(q);
        return PVX_OK;
    }
};

} // namespace

int gpu_build_bvh(const float *d_lo, const float *d_hi, const PrimRec *d_prims_in, PrimRec *d_prims_out,
                  uint32_t n, int max_leaf, bool wide, hipStream_t st, GpuBvh &out, std::vector<void *> &allocs) {
    out = GpuBvh{};
    max_leaf = std::max(1, std::min(8, max_leaf));
    if (n >= (1u << 28)) return fail(PVX_EUNSUPPORTED, "too many primitives for the 28-bit leaf encoding");
    if (n == 0) {
        out.root4 = out.root = INT32_MIN;
        return PVX_OK;
    }
    Temps tmp(st);
    int r;
    uint32_t *vals = nullptr;
    if ((r = tmp.get(&vals, n))) return r;
    if (n <= (uint32_t)max_leaf) { // one leaf over the primitives in input order
        k_iota<<<blocks(n), kB, 0, st>>>(vals, n);
        k_gather_prims<<<blocks(n), kB, 0, st>>>(d_prims_in, vals, n, d_prims_out);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(st));
        out.root = ~(int32_t)(0u | (n - 1));
        out.root4 = out.root;
        return PVX_OK;
    }
    const int ni = (int)n;
    uint32_t *bounds = nullptr, *vals_in = nullptr, *keep = nullptr, *newidx = nullptr;
    int *done = nullptr;
    uint64_t *keys_in = nullptr, *keys = nullptr;
    int2 *child = nullptr, *range = nullptr;
    int *iheight = nullptr;
    float *ibox = nullptr;
    if ((r = tmp.get(&bounds, 6)) || (r = tmp.get(&vals_in, n)) || (r = tmp.get(&keys_in, n)) ||
        (r = tmp.get(&keys, n)) || (r = tmp.get(&child, n - 1)) || (r = tmp.get(&range, n - 1)) ||
        (r = tmp.get(&iheight, n - 1)) ||
        (r = tmp.get(&ibox, 6 * (size_t)(n - 1))) || (r = tmp.get(&done, n - 1)) || (r = tmp.get(&keep, n)) ||
        (r = tmp.get(&newidx, n)))
        return r;

    // scratch for rocPRIM: the larger of the sort and the scans
    size_t sort_bytes = 0, scan_bytes = 0;
    HIP_TRY(rocprim::radix_sort_pairs(nullptr, sort_bytes, keys_in, keys, vals_in, vals, ni, 0u, 63u, st));
    HIP_TRY(rocprim::exclusive_scan(nullptr, scan_bytes, keep, newidx, 0u, (size_t)ni, rocprim::plus<uint32_t>(), st));
    void *cub = nullptr;
    size_t cub_bytes = std::max(sort_bytes, scan_bytes);
    unsigned char *cub_p = nullptr;
    if ((r = tmp.get(&cub_p, cub_bytes))) return r;
    cub = cub_p;

    // 1-2: Morton codes of the centroids, sorted
    k_bounds_init<<<1, 64, 0, st>>>(bounds);
    k_centroid_bounds<<<std::min<unsigned>(blocks(n), 1024u), kB, 0, st>>>(d_lo, d_hi, n, bounds);
    k_morton<<<blocks(n), kB, 0, st>>>(d_lo, d_hi, n, bounds, keys_in, vals_in);
    HIP_TRY(hipGetLastError());
    HIP_TRY(rocprim::radix_sort_pairs(cub, cub_bytes, keys_in, keys, vals_in, vals, ni, 0u, 63u, st));
    // 3: radix tree
    k_fill_i32<<<blocks(n - 1), kB, 0, st>>>(done, n - 1, -1);
    k_radix_tree<<<blocks(n - 1), kB, 0, st>>>(keys, ni, child, range);
    // 4: bounds and heights, pass by pass until the root is done (a radix
    // tree over 63-bit codes split by 32-bit indices is at most 95 deep)
    for (int pass = 0;; pass++) {
        k_refit_pass<<<blocks(n - 1), kB, 0, st>>>(d_lo, d_hi, vals, ni, child, range, max_leaf, ibox, iheight, done,
                                                   pass);
        HIP_TRY(hipGetLastError());
        if (pass % 8 == 7) {
            int root_done = -1;
            HIP_TRY(hipMemcpyAsync(&root_done, done, 4, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            if (root_done >= 0) break;
            if (pass > 200) return fail(PVX_EDEVICE, "GPU BVH build: refit did not reach the root");
        }
    }
    // 5: compaction into the binary layout
    k_keep<<<blocks(n), kB, 0, st>>>(range, ni, max_leaf, keep);
    HIP_TRY(hipGetLastError());
    HIP_TRY(rocprim::exclusive_scan(cub, cub_bytes, keep, newidx, 0u, (size_t)ni, rocprim::plus<uint32_t>(), st));
    uint32_t n_nodes = 0;
    int32_t root_h = 0;
    HIP_TRY(hipMemcpyAsync(&n_nodes, newidx + (n - 1), 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(&root_h, iheight, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (n_nodes == 0 || n_nodes > n - 1 || root_h < 1)
        return fail(PVX_EDEVICE, "GPU BVH build: inconsistent tree");
    void *pn = nullptr;
    HIP_TRY(hipMalloc(&pn, (size_t)n_nodes * sizeof(Node)));
    allocs.push_back(pn);
    out.nodes = static_cast<Node *>(pn);
    out.n_nodes = n_nodes;
    out.device_bytes += (size_t)n_nodes * sizeof(Node);
    out.root = 0; // node 0 of the radix tree covers every primitive and is kept
    out.depth = (uint32_t)(root_h - 1);
    k_emit_nodes<<<blocks(n - 1), kB, 0, st>>>(child, range, ibox, d_lo, d_hi, vals, keep, newidx, ni, out.nodes);
    k_gather_prims<<<blocks(n), kB, 0, st>>>(d_prims_in, vals, n, d_prims_out);
    HIP_TRY(hipGetLastError());

    // 6: 4-wide layout, breadth-first
    if (wide) {
        Node4 *w4 = nullptr;
        int32_t *fa = nullptr, *fb = nullptr;
        uint32_t *cnt = nullptr, *offs = nullptr;
        if ((r = tmp.get(&w4, n_nodes)) || (r = tmp.get(&fa, n_nodes)) || (r = tmp.get(&fb, n_nodes)) ||
            (r = tmp.get(&cnt, n_nodes + 1)) || (r = tmp.get(&offs, n_nodes + 1)))
            return r;
        size_t wscan = 0;
        HIP_TRY(rocprim::exclusive_scan(nullptr, wscan, cnt, offs, 0u, (size_t)n_nodes + 1, rocprim::plus<uint32_t>(), st));
        if (wscan > cub_bytes) return fail(PVX_EDEVICE, "GPU BVH build: scan scratch too small");
        k_fill_i32<<<1, kB, 0, st>>>(fa, 1, 0); // the root
        uint32_t m = 1, base = 0, levels = 0;
        while (m > 0) {
            if (base + m > n_nodes) return fail(PVX_EDEVICE, "GPU BVH build: wide tree overflow");
            k_wide_count<<<blocks(m), kB, 0, st>>>(out.nodes, fa, m, cnt);
            HIP_TRY(hipGetLastError());
            HIP_TRY(rocprim::exclusive_scan(cub, cub_bytes, cnt, offs, 0u, (size_t)m + 1, rocprim::plus<uint32_t>(), st));
            uint32_t next = 0;
            HIP_TRY(hipMemcpyAsync(&next, offs + m, 4, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            if (base + m + next > n_nodes) return fail(PVX_EDEVICE, "GPU BVH build: wide tree overflow");
            k_wide_emit<<<blocks(m), kB, 0, st>>>(out.nodes, fa, m, offs, base, w4, fb);
            HIP_TRY(hipGetLastError());
            base += m;
            m = next;
            std::swap(fa, fb);
            levels++;
        }
        void *p4 = nullptr;
        HIP_TRY(hipMalloc(&p4, (size_t)base * sizeof(Node4)));
        allocs.push_back(p4);
        HIP_TRY(hipMemcpyAsync(p4, w4, (size_t)base * sizeof(Node4), hipMemcpyDeviceToDevice, st));
        out.nodes4 = static_cast<Node4 *>(p4);
        out.n_nodes4 = base;
        out.root4 = 0;
        out.depth4 = levels - 1;
        out.device_bytes += (size_t)base * sizeof(Node4);
    }
    HIP_TRY(hipStreamSynchronize(st));
    return PVX_OK;
}

} // namespace pvx
