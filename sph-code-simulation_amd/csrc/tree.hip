// tree.hip -- octree build on the GPU with the reference's exact topology.
//
// Reference: BHTree::make / BHNode::create_tree / assign / set_kernel
// (src/bhtree.cpp:59-115, 154-247, 249-275).  The reference inserts particles
// one by one into a pointer octree: the child of a node is chosen by the FP64
// comparisons pos[k] > center[k] (ties go to the lower child), child centres
// are centre + (+-1) * edge * 0.25, and a child with more than
// leafParticleNumber particles is split again while the parent's level is below
// maxTreeLevel.
//
// Here the same decisions are made for every particle independently:
//  1. key: the particle's octant digit at every level, from replaying the
//     FP64 centre recursion (bits = dim per level, most significant = level 1);
//  2. stable LSD radix sort of the keys (8-bit digits, hand-written);
//  3. the particle state is permuted into key order;
//  4. kdig[i] = first level at which key[i-1] and key[i] differ; a node at
//     level L is a maximal run of positions without a boundary of kdig <= L-1;
//  5. lam[i] = level of the leaf that holds particle i, from a scan of the
//     kdig window of +-leafParticleNumber positions (a node is a leaf when it
//     holds <= leaf particles or sits at level maxTreeLevel+1);
//  6. the nodes starting at position s are the levels max(2, kdig[s]+1) ..
//     lam[s]; an exclusive scan of their counts gives every node its index in
//     pre-order (sorted by start, then level), which is the reference's
//     depth-first child order;
//  7. per node: level, [start, end), skip (index of the next node outside its
//     subtree, for stackless walks), centre and edge replayed in FP64, and mass,
//     centre of mass and kernel_size from block-hierarchy range reductions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "ctx.hpp"

namespace sphgpu {

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;

struct TreeShape {
    int dim;
    int bits;        // bits per level = dim
    int ml;          // maxTreeLevel
    int leaf;        // leafParticleNumber
    int n;
};

// ---- root geometry (src/bhtree.cpp:20-32 periodic, :67-103 bounding box) --
__global__ void k_bbox_partial(const double *__restrict__ x, const double *__restrict__ y,
                               const double *__restrict__ z, int n, int dim, double *__restrict__ part)
{
    __shared__ double smin[3][256], smax[3][256];
    double mn[3] = {1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308};
    double mx[3] = {-1.7976931348623157e308, -1.7976931348623157e308, -1.7976931348623157e308};
    const double *p[3] = {x, y, z};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        for (int k = 0; k < dim; ++k) {
            const double v = p[k][i];
            mn[k] = v < mn[k] ? v : mn[k];
            mx[k] = v > mx[k] ? v : mx[k];
        }
    }
    for (int k = 0; k < 3; ++k) { smin[k][threadIdx.x] = mn[k]; smax[k][threadIdx.x] = mx[k]; }
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            for (int k = 0; k < 3; ++k) {
                const double a = smin[k][threadIdx.x + s], b = smax[k][threadIdx.x + s];
                if (a < smin[k][threadIdx.x]) smin[k][threadIdx.x] = a;
                if (b > smax[k][threadIdx.x]) smax[k][threadIdx.x] = b;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) {
            part[blockIdx.x * 6 + k] = smin[k][0];
            part[blockIdx.x * 6 + 3 + k] = smax[k][0];
        }
}

// final bounds over the block partials: 256 threads, then a tree in LDS
// (min and max are exact in any order)
__global__ __launch_bounds__(256) void k_bbox_final(const double *__restrict__ part, int nblocks, int dim,
                                                    double *__restrict__ root)
{
    __shared__ double smin[3][256], smax[3][256];
    double mn[3] = {1.7976931348623157e308, 1.7976931348623157e308, 1.7976931348623157e308};
    double mx[3] = {-1.7976931348623157e308, -1.7976931348623157e308, -1.7976931348623157e308};
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
        for (int k = 0; k < 3; ++k) {
            if (part[b * 6 + k] < mn[k]) mn[k] = part[b * 6 + k];
            if (part[b * 6 + 3 + k] > mx[k]) mx[k] = part[b * 6 + 3 + k];
        }
    for (int k = 0; k < 3; ++k) { smin[k][threadIdx.x] = mn[k]; smax[k][threadIdx.x] = mx[k]; }
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < 3; ++k) {
                if (smin[k][threadIdx.x + s] < smin[k][threadIdx.x]) smin[k][threadIdx.x] = smin[k][threadIdx.x + s];
                if (smax[k][threadIdx.x + s] > smax[k][threadIdx.x]) smax[k][threadIdx.x] = smax[k][threadIdx.x + s];
            }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    double l = 0.0;
    for (int k = 0; k < dim; ++k) {
        root[k] = (smax[k][0] + smin[k][0]) * 0.5;
        const double range = smax[k][0] - smin[k][0];
        if (l < range) l = range;
    }
    for (int k = dim; k < 3; ++k) root[k] = 0.0;
    root[3] = l;
}

// ---- 1. keys ----------------------------------------------------------------
__global__ void k_keys(const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
                       const double *__restrict__ root, TreeShape sh, unsigned long long *__restrict__ key,
                       int *__restrict__ perm, int *__restrict__ sync_words)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {  // node count, partial-sort check
        sync_words[0] = 0;
        sync_words[1] = 0;
    }
    if (i >= sh.n) return;
    double pos[3] = {x[i], sh.dim > 1 ? y[i] : 0.0, sh.dim > 2 ? z[i] : 0.0};
    double c[3] = {root[0], root[1], root[2]};
    double e = root[3];
    unsigned long long kk = 0;
    for (int lev = 0; lev < sh.ml; ++lev) {
        // child centre, src/bhtree.cpp:233-239: c + (+-1 * edge) * 0.25; the
        // products are exact (sign and power-of-two scalings), so c + q or
        // c - q with q = edge * 0.25 rounds identically
        const double q = e * 0.25;
        unsigned dig = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (k >= sh.dim) break;
            const bool up = pos[k] > c[k];  // src/bhtree.cpp:211-218
            dig |= (up ? 1u : 0u) << k;
            c[k] = c[k] + (up ? q : -q);
        }
        e = e * 0.5;
        kk = (kk << sh.bits) | dig;
    }
    key[i] = kk;
    perm[i] = i;
}


// ---- 2. radix sort -------------------------------------------------------------
// Partial sort check (in k_kdig): the keys were sorted on their bits above
// `cut` only (the low radix passes skipped), i.e. on the first D digits;
// that order is the full order as far as the tree goes when no cell of depth
// D holds more than `leaf` particles (no node deeper than level D + 1 exists,
// and every node boundary the build uses lies at depth <= D).  flag[1] gets 1
// if some cell of depth D holds more.

__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const unsigned long long *__restrict__ key, int n,
                                                            int shift, unsigned int *__restrict__ hist, int nblocks)
{
    __shared__ unsigned int cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kSortTile;
    for (int r = 0; r < kSortItems; ++r) {
        const int e = base + r * kSortThreads + threadIdx.x;
        if (e < n) atomicAdd(&cnt[(unsigned)(key[e] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

// ---- multi-block exclusive scan of unsigned counts -----------------------------
// Three passes: per-tile sums, a scan of the tile sums in one workgroup, and
// the per-tile scan with its offset.  Tiles of 2048 values are loaded
// coalesced and transposed through LDS so each thread scans 8 consecutive
// values.  In place (in == out) is allowed.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ unsigned block_excl_scan(unsigned v, unsigned &total, unsigned *wsum)
{
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    unsigned x = v;
    for (int off = 1; off < kWave; off <<= 1) {
        const unsigned y = __shfl_up(x, off, kWave);
        if (lane >= off) x += y;
    }
    if (lane == kWave - 1) wsum[w] = x;
    __syncthreads();
    unsigned base = 0;
    total = 0;
    for (int k = 0; k < kScanThreads / kWave; ++k) {
        if (k < w) base += wsum[k];
        total += wsum[k];
    }
    __syncthreads();
    return base + x - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(const unsigned *__restrict__ in, int n,
                                                             unsigned *__restrict__ tile_sum)
{
    __shared__ unsigned wsum[kScanThreads / kWave];
    const long long base = (long long)blockIdx.x * kScanTile;
    unsigned s = 0;
    for (int r = 0; r < kScanItems; ++r) {
        const long long e = base + r * kScanThreads + threadIdx.x;
        if (e < n) s += in[e];
    }
    unsigned total;
    block_excl_scan(s, total, wsum);
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

// exclusive scan of the tile sums (one workgroup); tile_sum[ntiles] = total
__global__ __launch_bounds__(kScanThreads) void k_scan_tile_sums(unsigned *__restrict__ tile_sum, int ntiles)
{
    __shared__ unsigned wsum[kScanThreads / kWave];
    unsigned carry = 0;
    for (int b = 0; b < ntiles; b += kScanThreads) {
        const int e = b + threadIdx.x;
        const unsigned v = e < ntiles ? tile_sum[e] : 0u;
        unsigned total;
        const unsigned ex = block_excl_scan(v, total, wsum);
        if (e < ntiles) tile_sum[e] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) tile_sum[ntiles] = carry;
}

// one workgroup scans a short array in place or out of place, tile by tile:
// each thread loads 8 consecutive values of a tile of 8192 (coalesced
// 16-byte loads), the thread sums are scanned across the workgroup, a carry
// runs from tile to tile, and each thread writes its 8 running values (a
// tile is read before it is written, so in == out is allowed); total ->
// *total_out (and *total_out2) when given
constexpr int kSingleThreads = 1024;
constexpr int kSingleItems = 8;
constexpr int kSingleTile = kSingleThreads * kSingleItems;
constexpr int kSingleMax = kSingleTile * 24;

__global__ __launch_bounds__(kSingleThreads) void k_scan_single(const unsigned *in, unsigned *out, int n,
                                                                unsigned *total_out, unsigned *total_out2)
{
    __shared__ unsigned wsum[kSingleThreads / kWave];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    unsigned carry = 0;
    // tile t + 1 is loaded while tile t is scanned
    auto load_tile = [&](int t0, unsigned v[kSingleItems]) {
        const int b = t0 + threadIdx.x * kSingleItems;
        if (b + kSingleItems <= n) {
            const uint4 a0 = *reinterpret_cast<const uint4 *>(in + b);
            const uint4 a1 = *reinterpret_cast<const uint4 *>(in + b + 4);
            v[0] = a0.x; v[1] = a0.y; v[2] = a0.z; v[3] = a0.w;
            v[4] = a1.x; v[5] = a1.y; v[6] = a1.z; v[7] = a1.w;
        } else {
#pragma unroll
            for (int k = 0; k < kSingleItems; ++k) v[k] = b + k < n ? in[b + k] : 0u;
        }
    };
    unsigned vn[kSingleItems];
    if (n > 0) load_tile(0, vn);
    for (int t0 = 0; t0 < n; t0 += kSingleTile) {
        const int b = t0 + threadIdx.x * kSingleItems;
        unsigned v[kSingleItems];
#pragma unroll
        for (int k = 0; k < kSingleItems; ++k) v[k] = vn[k];
        if (t0 + kSingleTile < n) load_tile(t0 + kSingleTile, vn);
        unsigned s = 0;
#pragma unroll
        for (int k = 0; k < kSingleItems; ++k) s += v[k];
        unsigned x = s;
        for (int off = 1; off < kWave; off <<= 1) {
            const unsigned y = __shfl_up(x, off, kWave);
            if (lane >= off) x += y;
        }
        if (lane == kWave - 1) wsum[w] = x;
        __syncthreads();
        unsigned run = carry + x - s, total = 0;
        for (int k = 0; k < kSingleThreads / kWave; ++k) {
            if (k < w) run += wsum[k];
            total += wsum[k];
        }
        __syncthreads();
        carry += total;
        unsigned o[kSingleItems];
#pragma unroll
        for (int k = 0; k < kSingleItems; ++k) {
            o[k] = run;
            run += v[k];
        }
        if (b + kSingleItems <= n) {
            *reinterpret_cast<uint4 *>(out + b) = make_uint4(o[0], o[1], o[2], o[3]);
            *reinterpret_cast<uint4 *>(out + b + 4) = make_uint4(o[4], o[5], o[6], o[7]);
        } else {
#pragma unroll
            for (int k = 0; k < kSingleItems; ++k)
                if (b + k < n) out[b + k] = o[k];
        }
    }
    if (threadIdx.x == 0 && total_out) *total_out = carry;
    if (threadIdx.x == 0 && total_out2) *total_out2 = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const unsigned *in, int n, const unsigned *__restrict__ tile_off,
                                                             unsigned *out, int ntiles, unsigned *total_out,
                                                             unsigned *total_out2)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (total_out) *total_out = tile_off[ntiles];
        if (total_out2) *total_out2 = tile_off[ntiles];
    }
    __shared__ unsigned tile[kScanTile];
    __shared__ unsigned wsum[kScanThreads / kWave];
    const long long base = (long long)blockIdx.x * kScanTile;
    for (int r = 0; r < kScanItems; ++r) {
        const int q = r * kScanThreads + threadIdx.x;
        tile[q] = (base + q < n) ? in[base + q] : 0u;
    }
    __syncthreads();
    unsigned v[kScanItems], s = 0;
    for (int r = 0; r < kScanItems; ++r) {
        v[r] = tile[threadIdx.x * kScanItems + r];
        s += v[r];
    }
    unsigned total;
    unsigned run = tile_off[blockIdx.x] + block_excl_scan(s, total, wsum);
    for (int r = 0; r < kScanItems; ++r) {
        tile[threadIdx.x * kScanItems + r] = run;
        run += v[r];
    }
    __syncthreads();
    for (int r = 0; r < kScanItems; ++r) {
        const int q = r * kScanThreads + threadIdx.x;
        if (base + q < n) out[base + q] = tile[q];
    }
}

__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(
    const unsigned long long *__restrict__ key_in, const int *__restrict__ val_in, unsigned long long *__restrict__ key_out,
    int *__restrict__ val_out, int n, int shift, const unsigned int *__restrict__ hist, int nblocks)
{
    constexpr int kWaves = kSortThreads / kWave;
    __shared__ unsigned int base[256];
    __shared__ unsigned int running[256];
    __shared__ unsigned int wcnt[kWaves][256];
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int w = tid / kWave;
    base[tid] = hist[tid * nblocks + blockIdx.x];
    running[tid] = 0;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int tile = blockIdx.x * kSortTile;
    for (int r = 0; r < kSortItems; ++r) {
        for (int q = 0; q < kWaves; ++q) wcnt[q][tid] = 0;
        __syncthreads();
        const int e = tile + r * kSortThreads + tid;
        const bool valid = e < n;
        unsigned long long k = 0;
        int v = 0;
        unsigned d = 256;  // invalid lanes take a digit no valid lane has
        if (valid) {
            k = key_in[e];
            v = val_in[e];
            d = (unsigned)(k >> shift) & 255u;
        }
        // lanes holding the same digit (9 ballots: 8 digit bits + validity)
        unsigned long long same = __ballot(1);
        for (int b = 0; b < 9; ++b) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long bal = __ballot(bit);
            same &= bit ? bal : ~bal;
        }
        const int rank = __popcll(same & lt);
        if (valid && rank == 0) wcnt[w][d] = __popcll(same);
        __syncthreads();
        if (valid) {
            unsigned int pre = 0;
            for (int q = 0; q < w; ++q) pre += wcnt[q][d];
            const unsigned int dst = base[d] + running[d] + pre + rank;
            key_out[dst] = k;
            val_out[dst] = v;
        }
        __syncthreads();
        unsigned int tot = 0;
        for (int q = 0; q < kWaves; ++q) tot += wcnt[q][tid];
        running[tid] += tot;
        __syncthreads();
    }
}

// ---- 3. permutation of the particle state -----------------------------------
struct GatherBatch {
    const double *src[16];
    double *dst[16];
    int count;
};

__global__ void k_gather_batch(GatherBatch b, const int *__restrict__ perm, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = perm[i];
    for (int f = 0; f < b.count; ++f) b.dst[f][i] = b.src[f][s];
}

__global__ void k_gather_int2(const int *__restrict__ a, int *__restrict__ ad, const int *__restrict__ b,
                              int *__restrict__ bd, const int *__restrict__ perm, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = perm[i];
    ad[i] = a[s];
    bd[i] = b[s];
}

// ---- 4. first differing level ---------------------------------------------------
// also the partial-sort check (cut >= 0): flag[1] = 1 when particles i and
// i + leaf share the sorted top digits (key >> cut), see k_prefix_runs
__global__ void k_kdig(const unsigned long long *__restrict__ key, TreeShape sh, unsigned char *__restrict__ kdig,
                       int cut, int *__restrict__ flag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (cut >= 0) {
        const bool bad = i + sh.leaf < sh.n && (key[i] >> cut) == (key[i + sh.leaf] >> cut);
        if (__ballot(bad) && (threadIdx.x & (kWave - 1)) == 0) atomicOr(flag + 1, 1);
    }
    if (i >= sh.n) return;
    if (i == 0) { kdig[0] = 0; return; }
    const unsigned long long x = key[i - 1] ^ key[i];
    if (x == 0) { kdig[i] = (unsigned char)(sh.ml + 1); return; }
    const int msb = 63 - __clzll((long long)x);
    kdig[i] = (unsigned char)(sh.ml - msb / sh.bits);
}

// ---- 5. leaf level and node counts ----------------------------------------------
constexpr int kMaxLevels = 64;

constexpr int kLambdaThreads = 128;

__global__ void k_lambda(const unsigned char *__restrict__ kdig, TreeShape sh, unsigned char *__restrict__ lam,
                         int *__restrict__ leaf_end, int *__restrict__ ncount)
{
    // per thread and level t: the nearest boundary with kdig <= t to the
    // left (offset <= 0) and to the right (offset >= 1) of i, in LDS (a
    // level-indexed table in registers would go to scratch memory)
    extern __shared__ short s_lr[];
    // kdig over [block start - leaf, block end + leaf) (0 outside [0, n))
    __shared__ unsigned char s_kd[kLambdaThreads + 2 * 64];
    const int B = blockDim.x;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int ML = sh.ml;
    if (sh.leaf <= 64) {
        const int wb = blockIdx.x * B - sh.leaf;
        for (int t = threadIdx.x; t < B + 2 * sh.leaf; t += B) {
            const int j = wb + t;
            s_kd[t] = (j < 0 || j >= sh.n) ? (unsigned char)0 : kdig[j];
        }
        __syncthreads();
        if (i >= sh.n) return;
        const unsigned char *const kw = s_kd + (i - wb);  // kw[x] = kdig[i + x]
        // With pmL(d) = min kdig over [i - d, i] and pmR(e) = min kdig over
        // (i, i + e] (kdig[n] = 0), the node of level t + 1 around i starts
        // within d and ends within e positions iff pmL(d) <= t and pmR(e) <=
        // t, so it holds <= leaf particles iff max(pmL(d), pmR(leaf - d)) <= t
        // for some d < leaf: the leaf level is 1 + min over d of that max.
        unsigned char *const pm = reinterpret_cast<unsigned char *>(s_lr) + threadIdx.x;
        const int leaf = sh.leaf;
        int m = 255;
        for (int e = 1; e <= leaf; ++e) {
            m = min(m, (int)kw[e]);  // kdig[n] = 0
            pm[(e - 1) * B] = (unsigned char)m;
        }
        int ml = 255, best = 255;
        for (int d = 0; d < leaf; ++d) {
            // the window's zeros left of 0 change nothing: kdig[0] = 0 came first
            ml = min(ml, (int)kw[-d]);
            best = min(best, max(ml, (int)pm[(leaf - d - 1) * B]));
        }
        int L = ML + 1, lend = -1;
        if (best + 1 <= ML) {
            L = max(2, best + 1);
            const int t = L - 1;
            for (int e = 1; e <= leaf; ++e)
                if (pm[(e - 1) * B] <= t) {
                    lend = i + e;
                    break;
                }
        }
        lam[i] = (unsigned char)L;
        leaf_end[i] = lend;
        const int kd = kw[0];
        ncount[i] = (i == 0 ? 1 : 0) + max(0, L - max(2, kd + 1) + 1);
        return;
    }
    if (i >= sh.n) return;
    short *const nl = s_lr + threadIdx.x;
    short *const nr = s_lr + (size_t)(sh.ml + 1) * B + threadIdx.x;
    constexpr short kNone = -32768, kNoneR = 32767;
    for (int t = 0; t <= ML; ++t) { nl[t * B] = kNone; nr[t * B] = kNoneR; }
    // nearest boundary to the left with kdig <= t (position of the group start)
    {
        int m = ML + 2;
        const int lo = max(0, i - sh.leaf + 1);
        for (int j = i; j >= lo && m > 0; --j) {
            const int kj = kdig[j];
            if (kj < m) {
                for (int t = kj; t < m && t <= ML; ++t) nl[t * B] = (short)(j - i);
                m = kj;
            }
        }
    }
    // nearest boundary to the right (position of the group end); N counts as one
    {
        int m = ML + 2;
        const int hi = min(sh.n, i + sh.leaf);
        for (int j = i + 1; j <= hi && m > 0; ++j) {
            const int kj = (j == sh.n) ? 0 : (int)kdig[j];
            if (kj < m) {
                for (int t = kj; t < m && t <= ML; ++t) nr[t * B] = (short)(j - i);
                m = kj;
            }
        }
    }
    int L = ML + 1;
    int lend = -1;
    for (int l = 2; l <= ML; ++l) {
        const int t = l - 1;
        const int a = nl[t * B], b = nr[t * B];
        if (a != kNone && b != kNoneR && b - a <= sh.leaf) {
            L = l;
            lend = i + b;
            break;
        }
    }
    lam[i] = (unsigned char)L;
    leaf_end[i] = lend;
    const int kd = kdig[i];
    const int lmin = max(2, kd + 1);
    ncount[i] = (i == 0 ? 1 : 0) + max(0, L - lmin + 1);
}

// ---- 6. node fill ------------------------------------------------------------------
// first j >= lo whose key prefix (key >> shift) differs from p, searching
// forward from lo with doubling steps, then bisecting the last step; nodes on
// one particle's path are nested, so each level starts from the end of the
// level below it and the search stays short
__device__ int gallop_end(const unsigned long long *__restrict__ key, int n, int lo, unsigned long long p, int shift)
{
    if (lo >= n || (key[lo] >> shift) != p) return lo;
    int step = 1;
    int hi = lo + 1;
    while (hi < n && (key[hi] >> shift) == p) {
        lo = hi;
        step <<= 1;
        hi = min(n, lo + step);
    }
    // key[lo] has the prefix; the first differing position is in (lo, hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if ((key[mid] >> shift) == p) lo = mid;
        else hi = mid;
    }
    return hi;
}

__global__ void k_fill(const unsigned long long *__restrict__ key, const unsigned char *__restrict__ kdig,
                       const unsigned char *__restrict__ lam, const int *__restrict__ leaf_end,
                       const int *__restrict__ nbase, const double *__restrict__ root, TreeShape sh,
                       const int *__restrict__ sync_words,
                       int *__restrict__ nstart, int *__restrict__ nend, int *__restrict__ nskip,
                       int *__restrict__ nlevel, double *__restrict__ ncenter, double *__restrict__ nedge, int cap,
                       long long *__restrict__ stats)
{
    // the node count comes from the device (the host reads it while this
    // runs); nothing is written when it exceeds the capacity or when the
    // partial-sort check failed (the build is then repeated)
    const int total = sync_words[0];
    if (total > cap || sync_words[1] != 0) return;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    int leaves = 0, depth = 0;
    const int idx0 = (s < sh.n) ? nbase[s] : 0;
    const int cnt = (s < sh.n) ? nbase[s + 1] - idx0 : 0;
    if (cnt > 0) {
        int idx = idx0;
        const int L1 = lam[s];
        const int ML = sh.ml;
        depth = L1;
        if (s == 0) {
            nstart[idx] = 0;
            nend[idx] = sh.n;
            nskip[idx] = total;
            nlevel[idx] = 1;
            for (int k = 0; k < 3; ++k) ncenter[(size_t)k * cap + idx] = root[k];
            nedge[idx] = root[3];
            ++idx;
        }
        const int lmin = max(2, (int)kdig[s] + 1);
        const unsigned long long kk = key[s];
        // replay the centre recursion down to level lmin
        double c[3] = {root[0], root[1], root[2]};
        double e = root[3];
        for (int d = 1; d < lmin; ++d) {
            const unsigned dig = (unsigned)(kk >> (sh.bits * (ML - d))) & ((1u << sh.bits) - 1u);
            for (int k = 0; k < sh.dim; ++k) {
                const double sg = ((dig >> k) & 1u) ? 1.0 : -1.0;
                c[k] = c[k] + sg * e * 0.25;
            }
            e = e * 0.5;
        }
        // top-down: start, level and geometry of the nodes on this path
        for (int L = lmin, q = idx; L <= L1; ++L, ++q) {
            const bool leaf = (L == L1);
            nstart[q] = s;
            nlevel[q] = leaf ? -L : L;
            for (int k = 0; k < 3; ++k) ncenter[(size_t)k * cap + q] = c[k];
            nedge[q] = e;
            leaves += leaf ? 1 : 0;
            // centre of the child on this particle's path
            const unsigned dig = (unsigned)(kk >> (sh.bits * (ML - L))) & ((1u << sh.bits) - 1u);
            for (int k = 0; k < sh.dim; ++k) {
                const double sg = ((dig >> k) & 1u) ? 1.0 : -1.0;
                c[k] = c[k] + sg * e * 0.25;
            }
            e = e * 0.5;
        }
        // bottom-up: the node at level L holds the particles sharing key's
        // first L-1 digits; its end is at or after the end of the level below
        int end_below = s + 1;
        for (int L = L1, q = idx + (L1 - lmin); L >= lmin; --L, --q) {
            int end;
            if (L == L1 && L1 <= ML) end = leaf_end[s];
            else {
                const int shift = sh.bits * (ML - (L - 1));
                end = (shift >= 64) ? sh.n : gallop_end(key, sh.n, end_below, kk >> shift, shift);
            }
            nend[q] = end;
            nskip[q] = (end < sh.n) ? nbase[end] : total;
            end_below = end;
        }
    }
    // counters: one atomic per wavefront
    for (int off = kWave / 2; off > 0; off >>= 1) {
        leaves += __shfl_xor(leaves, off, kWave);
        depth = max(depth, __shfl_xor(depth, off, kWave));
    }
    if ((threadIdx.x & (kWave - 1)) == 0) {
        if (leaves) atomicAdd(stat_ptr(stats, 11), (unsigned long long)leaves);
        if (depth) atomicMax(stat_ptr(stats, 12), (unsigned long long)depth);
    }
}

// ---- 7. block hierarchy and node reductions ------------------------------------------
__global__ void k_block_max(const double *__restrict__ in, int n, double *__restrict__ out)
{
    // one 64-lane wave per block of 64 values
    const int b = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int i = b * kWave + lane;
    if ((long long)b * kWave >= n) return;
    double v = (i < n) ? in[i] : 0.0;
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const double o = __shfl_down(v, off, kWave);
        v = o > v ? o : v;
    }
    if (lane == 0) out[b] = v;
}

__global__ void k_block_sum4(const double *__restrict__ m, const double *__restrict__ x, const double *__restrict__ y,
                             const double *__restrict__ z, int n, int nb, double *__restrict__ out, int stride_in,
                             int level)
{
    // level 0: sums of m, m x, m y, m z over 64 particles
    // level 1: sums of the level-0 partials over 64 blocks (inputs in out-layout)
    const int b = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    if (b >= nb) return;
    const int i = b * kWave + lane;
    double s[4] = {0, 0, 0, 0};
    if (i < n) {
        if (level == 0) {
            const double mm = m[i];
            s[0] = mm;
            s[1] = mm * x[i];
            s[2] = y ? mm * y[i] : 0.0;
            s[3] = z ? mm * z[i] : 0.0;
        } else {
            for (int q = 0; q < 4; ++q) s[q] = m[(size_t)q * stride_in + i];
        }
    }
    for (int off = kWave / 2; off > 0; off >>= 1)
        for (int q = 0; q < 4; ++q) s[q] += __shfl_down(s[q], off, kWave);
    if (lane == 0)
        for (int q = 0; q < 4; ++q) out[(size_t)q * nb + b] = s[q];
}

struct RangeTables {
    const double *h, *b1, *b2;
};

// Lanes per node in the node passes: a node's range splits into partial
// 64-blocks at both ends, partial 4096-blocks of whole 64-blocks, and whole
// 4096-blocks; each of the group's lanes takes every kNodeLanes-th entry of
// each piece, so the root costs tens of loads per lane instead of hundreds
// in one thread.
constexpr int kNodeLanes = 8;

// max of h over [s, e), all kNodeLanes lanes of the group taking part (max is
// exact, so the order does not matter)
__device__ double range_max_group(RangeTables t, int s, int e, int sub)
{
    double m = 0.0;
    const int s1 = min(e, (s + 63) & ~63), e1 = max(s1, e & ~63);  // whole 64-blocks [s1, e1)
    for (int j = s + sub; j < s1; j += kNodeLanes) m = fmax(m, t.h[j]);
    for (int j = e1 + sub; j < e; j += kNodeLanes) m = fmax(m, t.h[j]);
    const int b0 = s1 >> 6, b1 = e1 >> 6;
    const int c0 = min(b1, (b0 + 63) & ~63), c1 = max(c0, b1 & ~63);  // whole 4096-blocks
    for (int j = b0 + sub; j < c0; j += kNodeLanes) m = fmax(m, t.b1[j]);
    for (int j = c1 + sub; j < b1; j += kNodeLanes) m = fmax(m, t.b1[j]);
    for (int B = (c0 >> 6) + sub; B < (c1 >> 6); B += kNodeLanes) m = fmax(m, t.b2[B]);
    for (int off = kNodeLanes / 2; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, kNodeLanes));
    return m;
}

// positive float not below x (x >= 0)
__device__ float float_up(double x)
{
    float f = (float)x;
    if ((double)f < x) f = __int_as_float(__float_as_int(f) + 1);
    return f;
}

__global__ void k_node_ks(const int *__restrict__ nstart, const int *__restrict__ nend, int total, RangeTables t,
                          double *__restrict__ ks, int4 *__restrict__ nlink)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int idx = g / kNodeLanes, sub = g % kNodeLanes;
    const bool ok = idx < total;
    const double k = range_max_group(t, ok ? nstart[idx] : 0, ok ? nend[idx] : 0, sub);
    if (!ok || sub) return;
    ks[idx] = k;
    int4 l = nlink[idx];
    l.w = (l.w & (int)0x80000000) | __float_as_int(float_up(k));
    nlink[idx] = l;
}

// FP32 walk records of the neighbour search (kernel_size is added by
// k_node_ks).  Internal nodes keep their cell (centre rounded to nearest, half
// edge rounded up); a leaf gets the bounding box of its members, which can be
// much smaller than its cell.  FP32 rounding of the centre is covered by the
// search's reach padding.
__global__ void k_node_records(int total, int dim, const int *__restrict__ nstart, const int *__restrict__ nend,
                               const int *__restrict__ nskip, const int *__restrict__ nlevel,
                               const double *__restrict__ ncenter, const double *__restrict__ nedge, int cap,
                               const double *__restrict__ px, const double *__restrict__ py,
                               const double *__restrict__ pz, float4 *__restrict__ nbox, float2 *__restrict__ nhalf,
                               int4 *__restrict__ nlink)
{
    // kNodeLanes lanes per node share a leaf's member loop
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int idx = g / kNodeLanes, sub = g % kNodeLanes;
    if (idx >= total) return;  // whole groups leave together
    const int s = nstart[idx], e = nend[idx];
    const bool leaf = nlevel[idx] < 0;
    double c[3] = {ncenter[idx], ncenter[(size_t)cap + idx], ncenter[2 * (size_t)cap + idx]};
    double h[3] = {0.5 * nedge[idx], 0.5 * nedge[idx], 0.5 * nedge[idx]};
    if (leaf && e > s) {
        const double *p[3] = {px, py, pz};
        for (int k = 0; k < dim; ++k) {
            double lo = p[k][s], hi = lo;
            for (int j = s + 1 + sub; j < e; j += kNodeLanes) {
                const double v = p[k][j];
                lo = v < lo ? v : lo;
                hi = v > hi ? v : hi;
            }
            for (int off = kNodeLanes / 2; off > 0; off >>= 1) {
                const double ol = __shfl_xor(lo, off, kNodeLanes), oh = __shfl_xor(hi, off, kNodeLanes);
                lo = ol < lo ? ol : lo;
                hi = oh > hi ? oh : hi;
            }
            c[k] = 0.5 * (lo + hi);
            // half extent, and rounding of the centre in FP64
            h[k] = 0.5 * (hi - lo) + 4.0 * 1.1102230246251565e-16 * (fabs(lo) + fabs(hi));
        }
    }
    if (sub) return;
    nbox[idx] = make_float4((float)c[0], (float)c[1], (float)c[2], float_up(h[0]));
    nhalf[idx] = make_float2(float_up(h[1]), float_up(h[2]));
    nlink[idx] = make_int4(s, e, nskip[idx], leaf ? (int)0x80000000 : 0);
}

// child lists: count the direct children of each internal node (its
// sibling chain k+1, skip[k+1], ... below skip[k]) ...
__global__ void k_child_count(int total, const int *__restrict__ nskip, const int *__restrict__ nlevel,
                              unsigned int *__restrict__ ccount)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= total) return;
    unsigned int c = 0;
    if (nlevel[k] > 0)
        for (int ch = k + 1; ch < nskip[k]; ch = nskip[ch]) ++c;
    ccount[k] = c;
}

// ... and write them at the scanned offsets
__global__ void k_child_write(int total, const int *__restrict__ nskip, const int *__restrict__ nlevel,
                              const unsigned int *__restrict__ coff, int2 *__restrict__ nchild, int *__restrict__ child)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= total) return;
    const int off = (int)coff[k];
    int c = 0;
    if (nlevel[k] > 0)
        for (int ch = k + 1; ch < nskip[k]; ch = nskip[ch]) child[off + c++] = ch;
    nchild[k] = make_int2(off, c);
}

struct SumTables {
    const double *m, *x, *y, *z;   // particles
    const double *b1, *b2;         // [4][nb1], [4][nb2]
    int nb1, nb2;
};

__device__ void range_sum4(const SumTables &t, int s, int e, double out[4])
{
    double a[4] = {0, 0, 0, 0};
    auto addp = [&](int i) {
        const double mm = t.m[i];
        a[0] += mm;
        a[1] += mm * t.x[i];
        a[2] += t.y ? mm * t.y[i] : 0.0;
        a[3] += t.z ? mm * t.z[i] : 0.0;
    };
    while (s < e && (s & 63)) { addp(s); ++s; }
    while (e > s && (e & 63)) { --e; addp(e); }
    int b0 = s >> 6, b1 = e >> 6;
    while (b0 < b1 && (b0 & 63)) { for (int q = 0; q < 4; ++q) a[q] += t.b1[(size_t)q * t.nb1 + b0]; ++b0; }
    while (b1 > b0 && (b1 & 63)) { --b1; for (int q = 0; q < 4; ++q) a[q] += t.b1[(size_t)q * t.nb1 + b1]; }
    for (int B = b0 >> 6; B < (b1 >> 6); ++B)
        for (int q = 0; q < 4; ++q) a[q] += t.b2[(size_t)q * t.nb2 + B];
    for (int q = 0; q < 4; ++q) out[q] = a[q];
}

__global__ void k_node_mass(const int *__restrict__ nstart, const int *__restrict__ nend, int total, SumTables t,
                            double *__restrict__ mass, double *__restrict__ com, int cap)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    double a[4];
    range_sum4(t, nstart[idx], nend[idx], a);
    mass[idx] = a[0];
    for (int k = 0; k < 3; ++k) com[(size_t)k * cap + idx] = (a[0] > 0.0) ? a[1 + k] / a[0] : 0.0;
}

} // namespace

// exclusive scan of in[0..n) into out[0..n), with out[n] = total when
// write_total (out then has n + 1 entries) and the total also in *total_copy
// when given; tmp holds ceil(n / tile) + 1 values.  Short arrays take one
// workgroup (one launch), longer ones the three tile passes.
static void scan_counts(hipStream_t st, const unsigned *in, unsigned *out, int n, unsigned *tmp, bool write_total,
                        unsigned *total_copy = nullptr)
{
    if (n <= kSingleMax) {
        hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(kSingleThreads), 0, st, in, out, n,
                           write_total ? out + n : nullptr, total_copy);
        return;
    }
    const int nt = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(kScanThreads), 0, st, in, n, tmp);
    hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), dim3(kScanThreads), 0, st, tmp, nt);
    hipLaunchKernelGGL(k_scan_apply, dim3(nt), dim3(kScanThreads), 0, st, in, n, (const unsigned *)tmp, out, nt,
                       write_total ? out + n : nullptr, total_copy);
}

// ------------------------------------------------------------------------------
int tree_alloc(sphgpu_ctx *c)
{
    Tree &t = c->tree;
    const int n = c->n;
    t.node_cap = 5 * n + 1;  // BHTree::resize pool of 5N nodes plus the root
    auto al = [&](void **p, size_t b) { return hipMalloc(p, b) == hipSuccess && hipMemset(*p, 0, b) == hipSuccess; };
    const int nb1 = (n + 63) / 64, nb2 = (nb1 + 63) / 64;
    t.hist_blocks = (n + kSortTile - 1) / kSortTile;
    bool ok = al((void **)&t.key, 8ull * n) && al((void **)&t.key_alt, 8ull * n) && al((void **)&t.perm, 4ull * n) &&
              al((void **)&t.perm_alt, 4ull * n) && al((void **)&t.kdig, n) && al((void **)&t.lam, n) &&
              al((void **)&t.sync_words, 4ull * 2) &&
              al((void **)&t.leaf_end, 4ull * n) && al((void **)&t.ncount, 4ull * (n + 1)) &&
              al((void **)&t.nbase, 4ull * (n + 1)) && al((void **)&t.start, 4ull * t.node_cap) &&
              al((void **)&t.end, 4ull * t.node_cap) && al((void **)&t.skip, 4ull * t.node_cap) &&
              al((void **)&t.level, 4ull * t.node_cap) && al((void **)&t.center, 24ull * t.node_cap) &&
              al((void **)&t.edge, 8ull * t.node_cap) && al((void **)&t.ks, 8ull * t.node_cap) &&
              al((void **)&t.mass, 8ull * t.node_cap) && al((void **)&t.com, 24ull * t.node_cap) &&
              al((void **)&t.bmax1, 8ull * nb1) && al((void **)&t.bmax2, 8ull * nb2) &&
              al((void **)&t.bsum1, 32ull * nb1) && al((void **)&t.bsum2, 32ull * nb2) &&
              al((void **)&t.root, 8ull * 8) && al((void **)&t.hist, 4ull * 256 * t.hist_blocks) &&
              al((void **)&t.nbox, 16ull * t.node_cap) && al((void **)&t.nlink, 16ull * t.node_cap) &&
              al((void **)&t.nhalf, 8ull * t.node_cap) && al((void **)&t.nchild, 8ull * t.node_cap) &&
              al((void **)&t.child, 4ull * (t.node_cap + 8)) && al((void **)&t.ccount, 4ull * t.node_cap);
    const long long scan_max = std::max<long long>(std::max<long long>(n + 1, 256ll * t.hist_blocks), t.node_cap);
    ok = ok && al((void **)&t.scan_tmp, 4ull * ((scan_max + kScanTile - 1) / kScanTile + 1));
    ok = ok && hipHostMalloc((void **)&t.host_words, 4ull * 2, hipHostMallocDefault) == hipSuccess &&
         hipEventCreateWithFlags(&t.words_ready, hipEventDisableTiming) == hipSuccess;
    if (!ok) return SPHGPU_E_NOMEM;
    if (c->prm.periodic) {
        double r[8] = {0};
        double l = 0.0;
        for (int k = 0; k < c->dim; ++k) {
            r[k] = (c->prm.range_max[k] + c->prm.range_min[k]) * 0.5;
            const double range = c->prm.range_max[k] - c->prm.range_min[k];
            if (l < range) l = range;
        }
        r[3] = l;
        if (hipMemcpy(t.root, r, sizeof(r), hipMemcpyHostToDevice) != hipSuccess) return SPHGPU_E_HIP;
    }
    return 0;
}

void tree_free(sphgpu_ctx *c)
{
    Tree &t = c->tree;
    void *ps[] = {t.key, t.key_alt, t.perm, t.perm_alt, t.kdig, t.lam, t.sync_words, t.leaf_end, t.ncount, t.nbase, t.start,
                  t.end, t.skip, t.level, t.center, t.edge, t.ks, t.mass, t.com, t.bmax1, t.bmax2, t.bsum1,
                  t.bsum2, t.root, t.hist, t.nbox, t.nlink, t.scan_tmp, t.nhalf, t.nchild, t.child, t.ccount, t.gnode};
    for (void *p : ps)
        if (p) hipFree(p);
    if (t.host_words) hipHostFree(t.host_words);
    if (t.words_ready) hipEventDestroy(t.words_ready);
    t = Tree();
}

static int permute_state(sphgpu_ctx *c)
{
    const int n = c->n;
    const int *perm = c->tree.perm;
    std::vector<double **> fields;
    // inside a time step the fields that the step recomputes before anyone
    // reads them stay unpermuted: acc and du/dt (FluidForce assigns them)
    // and the MUSCL gradients (the second-order pre-interaction does)
    const bool light = c->in_step;
    const bool gsph = c->prm.sph_type == SPHGPU_GSPH || c->prm.sph_type == SPHGPU_GDISPH;
    for (int k = 0; k < c->dim; ++k) {
        fields.push_back(&c->pos[k]);
        fields.push_back(&c->vel[k]);
        fields.push_back(&c->vel_p[k]);
        if (!light) fields.push_back(&c->acc[k]);
    }
    for (int s = 0; s < S_NSCALAR; ++s)
        if (!(light && s == S_DENE)) fields.push_back(&c->sc[s]);
    if (gsph && !(light && c->prm.gsph_2nd_order)) {
        for (int k = 0; k < c->dim; ++k) {
            fields.push_back(&c->grad[k]);
            fields.push_back(&c->grad[3 + k]);
            for (int a = 0; a < c->dim; ++a) fields.push_back(&c->grad[6 + 3 * a + k]);
        }
    }
    for (size_t off = 0; off < fields.size(); off += 16) {
        GatherBatch b;
        b.count = (int)std::min<size_t>(16, fields.size() - off);
        for (int f = 0; f < b.count; ++f) {
            b.src[f] = *fields[off + f];
            b.dst[f] = c->alt[f];
        }
        hipLaunchKernelGGL(k_gather_batch, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, b, perm, n);
        for (int f = 0; f < b.count; ++f) std::swap(*fields[off + f], c->alt[f]);
    }
    hipLaunchKernelGGL(k_gather_int2, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, (const int *)c->id, c->id_alt,
                       (const int *)c->neighbor, c->tree.perm_alt, perm, n);
    std::swap(c->id, c->id_alt);
    std::swap(c->neighbor, c->tree.perm_alt);
    HIPCHK(c, hipGetLastError());
    return 0;
}

int tree_build(sphgpu_ctx *c)
{
    Tree &t = c->tree;
    const int n = c->n;
    TreeShape sh{c->dim, c->dim, c->prm.max_tree_level, c->prm.leaf_particle_num, n};
    const double *px = c->pos[0];
    const double *py = c->dim > 1 ? c->pos[1] : c->pos[0];
    const double *pz = c->dim > 2 ? c->pos[2] : c->pos[0];

    time_begin(c, "tree_keys");
    if (!c->prm.periodic) {
        const int nbb = std::min(c->red_blocks, grid_for(n, 256));
        hipLaunchKernelGGL(k_bbox_partial, dim3(nbb), dim3(256), 0, c->stream, px, py, pz, n, c->dim, c->red);
        hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(256), 0, c->stream, (const double *)c->red, nbb, c->dim, t.root);
    }
    hipLaunchKernelGGL(k_keys, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, px, py, pz, (const double *)t.root, sh,
                       t.key, t.perm, t.sync_words);
    time_end(c);

    // stable LSD radix sort over the dim*maxTreeLevel key bits.  Large
    // 2D/3D sets sort on the top t.sorted_digits digits only (8 at first;
    // the low passes are skipped) and check that no cell of that depth holds
    // more than leafParticleNumber particles; otherwise the tree is built
    // again with the full sort, and later builds sort 4 digits more.
    time_begin(c, "tree_sort");
    const int key_bits = sh.bits * sh.ml;
    const int passes = (key_bits + 7) / 8;
    const int nb = t.hist_blocks;
    int skip = 0, cut = 0;
    if (!t.full_sort && !t.full_once && n >= 65536 && c->dim >= 2) {
        const int keep = (c->dim * t.sorted_digits + 7) / 8;
        if (keep < passes) {
            skip = passes - keep;
            cut = key_bits - c->dim * ((key_bits - 8 * skip) / c->dim);
        }
    }
    for (int p = skip; p < passes; ++p) {
        const int shift = 8 * p;
        hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(kSortThreads), 0, c->stream, (const unsigned long long *)t.key, n,
                           shift, t.hist, nb);
        scan_counts(c->stream, t.hist, t.hist, 256 * nb, t.scan_tmp, false);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(kSortThreads), 0, c->stream,
                           (const unsigned long long *)t.key, (const int *)t.perm, t.key_alt, t.perm_alt, n, shift,
                           (const unsigned int *)t.hist, nb);
        std::swap(t.key, t.key_alt);
        std::swap(t.perm, t.perm_alt);
    }
    time_end(c);
    HIPCHK(c, hipGetLastError());

    time_begin(c, "tree_permute");
    int rc = permute_state(c);
    if (rc) return rc;
    time_end(c);

    time_begin(c, "tree_nodes");
    hipLaunchKernelGGL(k_kdig, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, (const unsigned long long *)t.key, sh,
                       t.kdig, skip ? cut : -1, t.sync_words);
    const size_t lam_lds = sh.leaf <= 64 ? (size_t)sh.leaf * kLambdaThreads : sizeof(short) * 2 * (size_t)(sh.ml + 1) * kLambdaThreads;
    hipLaunchKernelGGL(k_lambda, dim3(grid_for(n, kLambdaThreads)), dim3(kLambdaThreads), lam_lds,
                       c->stream, (const unsigned char *)t.kdig, sh,
                       t.lam, t.leaf_end, t.ncount);
    // the node count goes to nbase[n] and to sync_words[0] (read by the fill
    // on the device and copied to the host)
    scan_counts(c->stream, (const unsigned *)t.ncount, (unsigned *)t.nbase, n, t.scan_tmp, true,
                (unsigned *)t.sync_words);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(t.host_words, t.sync_words, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipEventRecord(t.words_ready, c->stream));
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 128)), dim3(128), 0, c->stream, (const unsigned long long *)t.key,
                       (const unsigned char *)t.kdig, (const unsigned char *)t.lam, (const int *)t.leaf_end,
                       (const int *)t.nbase, (const double *)t.root, sh, (const int *)t.sync_words, t.start, t.end,
                       t.skip, t.level, t.center, t.edge, t.node_cap, c->dstats);
    HIPCHK(c, hipGetLastError());
    time_end(c);
    // the pair records depend on the permuted particle state only: they are
    // refreshed while the host waits for the node count
    c->flist = 0;  // particles were reordered
    rc = refresh_mirrors(c);
    if (rc) return rc;
    HIPCHK(c, hipEventSynchronize(t.words_ready));
    const int total = t.host_words[0];
    if (skip) {
        if (t.host_words[1]) {
            // a cell of the sorted depth holds more than a leaf: this build
            // sorts fully (the state is already in the partial order; the
            // full build starts from it), later ones on 4 more digits
            t.sorted_digits += 4;
            if (t.sorted_digits >= sh.ml) t.full_sort = true;
            t.full_once = true;
            rc = tree_build(c);
            t.full_once = false;
            return rc;
        }
        ++t.partial_sorts;
    } else {
        ++t.full_sorts;
    }
    if (total > t.node_cap) {
        t.valid = false;
        return fail(c, SPHGPU_E_NO_FREE_NODE, "There is no free node.");
    }
    t.n_nodes = total;
    time_begin(c, "tree_node_records");
    hipLaunchKernelGGL(k_node_records, dim3(grid_for((long long)total * kNodeLanes, 256)), dim3(256), 0, c->stream, total, c->dim,
                       (const int *)t.start, (const int *)t.end, (const int *)t.skip, (const int *)t.level,
                       (const double *)t.center, (const double *)t.edge, t.node_cap, (const double *)c->pos[0],
                       (const double *)(c->dim > 1 ? c->pos[1] : c->pos[0]),
                       (const double *)(c->dim > 2 ? c->pos[2] : c->pos[0]), t.nbox, t.nhalf, t.nlink);
    hipLaunchKernelGGL(k_child_count, dim3(grid_for(total, 256)), dim3(256), 0, c->stream, total,
                       (const int *)t.skip, (const int *)t.level, t.ccount);
    scan_counts(c->stream, (const unsigned *)t.ccount, t.ccount, total, t.scan_tmp, false);
    hipLaunchKernelGGL(k_child_write, dim3(grid_for(total, 256)), dim3(256), 0, c->stream, total,
                       (const int *)t.skip, (const int *)t.level, (const unsigned *)t.ccount, t.nchild, t.child);
    HIPCHK(c, hipGetLastError());
    // mass and centre of mass per node (needed by the gravity walk); the
    // positions are read from the permuted (key-ordered) arrays
    px = c->pos[0];
    py = c->dim > 1 ? c->pos[1] : c->pos[0];
    pz = c->dim > 2 ? c->pos[2] : c->pos[0];
    if (c->prm.use_gravity) {
        const int nb1 = (n + 63) / 64, nb2 = (nb1 + 63) / 64;
        hipLaunchKernelGGL(k_block_sum4, dim3(grid_for((long long)nb1 * 64, 256)), dim3(256), 0, c->stream,
                           (const double *)c->sc[S_MASS], px, c->dim > 1 ? py : nullptr, c->dim > 2 ? pz : nullptr, n,
                           nb1, t.bsum1, 0, 0);
        hipLaunchKernelGGL(k_block_sum4, dim3(grid_for((long long)nb2 * 64, 256)), dim3(256), 0, c->stream,
                           (const double *)t.bsum1, nullptr, nullptr, nullptr, nb1, nb2, t.bsum2, nb1, 1);
        SumTables st{c->sc[S_MASS], px, c->dim > 1 ? py : nullptr, c->dim > 2 ? pz : nullptr, t.bsum1, t.bsum2, nb1, nb2};
        hipLaunchKernelGGL(k_node_mass, dim3(grid_for(total, 128)), dim3(128), 0, c->stream, (const int *)t.start,
                           (const int *)t.end, total, st, t.mass, t.com, t.node_cap);
    }
    time_end(c);
    HIPCHK(c, hipGetLastError());
    t.valid = true;
    c->sorted_valid = true;
    return 0;
}

// BHTree::set_kernel: kernel_size = max smoothing length in each subtree
int tree_set_kernel(sphgpu_ctx *c)
{
    Tree &t = c->tree;
    if (!t.valid) return fail(c, SPHGPU_E_STATE, "set_kernel without a tree");
    const int n = c->n;
    const int nb1 = (n + 63) / 64, nb2 = (nb1 + 63) / 64;
    time_begin(c, "tree_set_kernel");
    hipLaunchKernelGGL(k_block_max, dim3(grid_for((long long)nb1 * 64, 256)), dim3(256), 0, c->stream,
                       (const double *)c->sc[S_SML], n, t.bmax1);
    hipLaunchKernelGGL(k_block_max, dim3(grid_for((long long)nb2 * 64, 256)), dim3(256), 0, c->stream,
                       (const double *)t.bmax1, nb1, t.bmax2);
    RangeTables rt{c->sc[S_SML], t.bmax1, t.bmax2};
    hipLaunchKernelGGL(k_node_ks, dim3(grid_for((long long)t.n_nodes * kNodeLanes, 128)), dim3(128), 0, c->stream, (const int *)t.start,
                       (const int *)t.end, t.n_nodes, rt, t.ks, t.nlink);
    time_end(c);
    HIPCHK(c, hipGetLastError());
    return 0;
}

int tree_dump(sphgpu_ctx *c, int64_t *n_nodes, int32_t *level, int32_t *count, int32_t *first, int32_t *is_leaf,
              double *center, double *edge, double *mass, double *com, double *kernel_size, int32_t *members)
{
    Tree &t = c->tree;
    if (!t.valid) return fail(c, SPHGPU_E_STATE, "tree_dump without a tree");
    const int m = t.n_nodes;
    *n_nodes = m;
    if (!level) return 0;
    std::vector<int> st(m), en(m), lv(m);
    std::vector<double> ce(3ull * m), ed(m), ma(m), co(3ull * m), ks(m);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(st.data(), t.start, 4ull * m, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(en.data(), t.end, 4ull * m, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(lv.data(), t.level, 4ull * m, hipMemcpyDeviceToHost));
    for (int k = 0; k < 3; ++k) {
        HIPCHK(c, hipMemcpy(ce.data() + (size_t)k * m, t.center + (size_t)k * t.node_cap, 8ull * m, hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(co.data() + (size_t)k * m, t.com + (size_t)k * t.node_cap, 8ull * m, hipMemcpyDeviceToHost));
    }
    HIPCHK(c, hipMemcpy(ed.data(), t.edge, 8ull * m, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(ma.data(), t.mass, 8ull * m, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(ks.data(), t.ks, 8ull * m, hipMemcpyDeviceToHost));
    std::vector<int> ids(c->n);
    HIPCHK(c, hipMemcpy(ids.data(), c->id, 4ull * c->n, hipMemcpyDeviceToHost));
    for (int i = 0; i < m; ++i) {
        level[i] = lv[i] < 0 ? -lv[i] : lv[i];
        is_leaf[i] = lv[i] < 0 ? 1 : 0;
        count[i] = en[i] - st[i];
        first[i] = st[i];
        for (int k = 0; k < 3; ++k) {
            center[3 * i + k] = ce[(size_t)k * m + i];
            com[3 * i + k] = co[(size_t)k * m + i];
        }
        edge[i] = ed[i];
        mass[i] = ma[i];
        kernel_size[i] = ks[i];
    }
    for (int k = 0; k < c->n; ++k) members[k] = ids[k];
    return 0;
}

} // namespace sphgpu
