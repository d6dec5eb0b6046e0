// sph_cellsort.hip -- counting sort of the NAIVE formulation's cell keys for gfx950.
//
// The NAIVE path's keys are cell indices in [0, C) (clamped to the grid; bits 30-31 carry the
// particle's x slab in its cell), so the order the step needs -- by (cell, particle id), the
// order of the reference's stable sort of the voxel map (voxel_sph_strategy.py:81-89), or by
// (cell, slab, id) in the fused step -- is a counting sort by cell followed by that order
// within each cell.  The keys arrive in the state's current slot order (the previous step's cell
// order, see hash_particle), so neighbouring threads mostly hit the same or neighbouring
// cells.  Per step:
//   count    one thread per slot: rank = atomicAdd(&cnt[key], 1) (an arbitrary order among
//            the particles of a cell).  Between the steps of one multi-step call the previous
//            step's force pass has done this already as it wrote the keys (rank_in_cell), and
//            the sort starts at the scan
//   scan     exclusive prefix sum of cnt over all cells: tile sums (16-byte loads; a force
//            pass that counted with one lane per particle has added them up already, into
//            kTileSumCopies copies per tile), then each tile's offset is the sum of the tile
//            sums before it (added up by the tile's own workgroup when there are few tiles,
//            else by a one-workgroup scan first), then the tile is scanned; the result is
//            the cell start table cs[0..C] itself, and the tile scan resets cnt, and the
//            other array of tile sums, to zero for the next step
//   scatter  one thread per slot: tmp[cs[key] + rank] = slot
//   order    one thread per position: the particle there moves to cs[key] + (number of
//            particles of its cell with a smaller id), which is id order within the cell; the
//            sorted key is written there and either the slot (read by the separate reorder
//            pass of a staged step) or, in the fused step, the particle's gathered state.
//            With cells cut into x slabs (StepConsts::slab_shift) the fused step orders a
//            cell by (slab, id) instead, and the first particle of each cell stores the
//            offsets of its slabs in the high bits of the cell's table entry; the scan has
//            stored offset 0 there for the empty cells.  A staged step keeps the id order
//            (the reference's), and its table carries no offsets for occupied cells, which
//            the row walk then searches whole.
// The within-cell pass costs one comparison per particle of the cell, less than the density
// pass spends on the same cell.  The final order does not depend on the atomic order.
// The per-slot passes (count, scatter, order) map workgroups to slot ranges XCD by XCD as
// the density and force passes do (xcd_tile), so each XCD reads back from its own L2 the keys
// and state its force pass wrote (step 0.2245 -> 0.2226 ms on the 1M dam break).
#include "sph_device.h"

namespace sph {
namespace {

constexpr int CT = 256;                        // threads per workgroup
constexpr int CTILE = 1 << kScanTileShift;     // cells per scan tile
constexpr int CI = CTILE / CT;                 // cells per thread in the scan
static_assert(CI == 16, "the scan's loads assume 16 cells per thread");
constexpr int CWAVES = CT / 64;
// up to this many tiles, each tile's workgroup sums the tile sums before it (at most 32
// loads per thread)
constexpr int kMergedTiles = 8192 / kTileSumCopies;

// The slab layout of the cell table (StepConsts::slab_shift .. start_mask).
struct SlabTable {
    int shift;
    int start_bits;
    int bits;
    int count_pos;
    int count_bits;
    uint32_t start_mask;
    // the fields of a cell with all its slab offsets 0 (an empty cell)
    __host__ __device__ uint32_t empty() const {
        uint32_t f = 0;
        for (int j = 1; j < (1 << shift); ++j) f |= 1u << (start_bits + bits * (j - 1));
        return f;
    }
};

SlabTable slab_table(const StepConsts& S) {
    return SlabTable{S.slab_shift, S.start_bits, S.slab_bits, S.count_pos, S.count_bits, S.start_mask};
}

// One thread per slot; the ranks of rank_in_cell.
__global__ __launch_bounds__(CT) void cell_count_kernel(int n, const uint32_t* __restrict__ keyI,
                                                         uint32_t* __restrict__ cnt, uint32_t* __restrict__ rank) {
    const int i = xcd_tile(blockIdx.x, gridDim.x) * CT + threadIdx.x;
    const bool valid = i < n;
    const uint32_t r = rank_in_cell(valid ? keyI[i] & kCellKeyMask : 0xffffffffu, valid, cnt);
    if (valid) rank[i] = r;
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    return x;
}

// Exclusive scan of one value per thread across the workgroup; *total receives the sum.
__device__ __forceinline__ uint32_t block_excl(uint32_t v, uint32_t* lw, uint32_t* total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t x = wave_incl_scan(v);
    if (lane == 63) lw[w] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < CWAVES; ++k) {
        const uint32_t t = lw[k];
        off += k < w ? t : 0u;
        tot += t;
    }
    *total = tot;
    __syncthreads();
    return off + x - v;
}

constexpr int CV = CI / 4;  // 16-byte vectors per thread and tile

// Tile sums into part copy 0 (the scan that follows reads that copy only).
__global__ __launch_bounds__(CT) void cell_scan_reduce(const uint32_t* __restrict__ cnt, int m,
                                                        uint32_t* __restrict__ part) {
    __shared__ uint32_t lw[CWAVES];
    const int64_t base = (int64_t)blockIdx.x * CTILE;
    uint32_t s = 0;
    if (base + CTILE <= m) {  // base is a multiple of 4 cells: 16-byte aligned
        const uint4* c4 = reinterpret_cast<const uint4*>(cnt + base);
#pragma unroll
        for (int k = 0; k < CV; ++k) {
            const uint4 x = c4[k * CT + threadIdx.x];
            s += x.x + x.y + x.z + x.w;
        }
    } else {
        for (int k = 0; k < CI; ++k) {
            const int64_t c = base + k * CT + threadIdx.x;
            if (c < m) s += cnt[c];
        }
    }
    uint32_t tot;
    block_excl(s, lw, &tot);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// One workgroup: exclusive scan of the tile sums (over their first `copies` copies) into
// copy 0, in chunks of CT with a carry.
__global__ __launch_bounds__(CT) void cell_scan_partials(uint32_t* __restrict__ part, int nparts, int copies) {
    __shared__ uint32_t lw[CWAVES];
    uint32_t carry = 0;
    for (int base = 0; base < nparts; base += CT) {
        const int i = base + threadIdx.x;
        uint32_t v = 0u;
        if (i < nparts)
            for (int k = 0; k < copies; ++k) v += part[(size_t)k * nparts + i];
        uint32_t tot;
        const uint32_t ex = block_excl(v, lw, &tot);
        if (i < nparts) part[i] = carry + ex;
        carry += tot;
    }
}

// Sum of the tile sums before tile b, over their first `copies` copies (the tile's offset
// when the tile sums were not scanned first).
__device__ __forceinline__ uint32_t tiles_before(const uint32_t* __restrict__ part, int b, int tiles, int copies,
                                                 uint32_t* lw) {
    uint32_t s = 0;
    for (int j = threadIdx.x; j < b; j += CT) {  // the copies of a tile in flight together
        uint32_t v[kTileSumCopies];
#pragma unroll
        for (int k = 0; k < kTileSumCopies; ++k) v[k] = k < copies ? part[(size_t)k * tiles + j] : 0u;
#pragma unroll
        for (int k = 0; k < kTileSumCopies; ++k) s += v[k];
    }
    uint32_t tot;
    block_excl(s, lw, &tot);
    return tot;
}

// The tile is read (and cnt cleared) with coalesced 16-byte accesses into LDS, each thread
// scans CI consecutive cells there, and cs is written back coalesced.  LDS index i is padded
// by i / 32 so that the thread-contiguous pass does not hit one bank.  The last tile also
// writes cs[m] = n.  MERGED: the first `copies` copies of part hold tile sums (all of them
// when a force pass added them up, copy 0 when cell_scan_reduce did; the offset is summed
// here); otherwise copy 0 of part holds the tiles' exclusive prefix sums.  Each workgroup also
// clears its tile's copies in next, the tile sums that the next force pass adds into.

__device__ __forceinline__ int cpad(int i) { return i + (i >> 5); }

template <bool MERGED>
__global__ __launch_bounds__(CT) void cell_scan_down(uint32_t* __restrict__ cnt, int m,
                                                      const uint32_t* __restrict__ part, uint32_t* __restrict__ next,
                                                      uint32_t* __restrict__ cs, uint32_t n, uint32_t empty,
                                                      int copies) {
    __shared__ uint32_t lw[CWAVES];
    __shared__ uint32_t tile[CTILE + CTILE / 32];
    const int64_t base = (int64_t)blockIdx.x * CTILE;
    if (threadIdx.x < kTileSumCopies) next[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = 0u;
    const uint32_t offset = MERGED ? tiles_before(part, blockIdx.x, gridDim.x, copies, lw) : part[blockIdx.x];
    const bool full = base + CTILE <= m;  // base is a multiple of 4 cells: 16-byte aligned
    if (full) {
        uint4* c4 = reinterpret_cast<uint4*>(cnt + base);
#pragma unroll
        for (int k = 0; k < CV; ++k) {
            const int v = k * CT + threadIdx.x;
            const uint4 x = c4[v];
            c4[v] = make_uint4(0u, 0u, 0u, 0u);
            tile[cpad(4 * v)] = x.x;
            tile[cpad(4 * v + 1)] = x.y;
            tile[cpad(4 * v + 2)] = x.z;
            tile[cpad(4 * v + 3)] = x.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < CI; ++k) {
            const int j = k * CT + threadIdx.x;
            const int64_t c = base + j;
            uint32_t v = 0u;
            if (c < m) {
                v = cnt[c];
                cnt[c] = 0u;
            }
            tile[cpad(j)] = v;
        }
    }
    __syncthreads();
    const int j0 = threadIdx.x * CI;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < CI; ++k) s += tile[cpad(j0 + k)];
    uint32_t tot;
    uint32_t run = offset + block_excl(s, lw, &tot);  // synchronises
#pragma unroll
    for (int k = 0; k < CI; ++k) {
        const uint32_t v = tile[cpad(j0 + k)];
        tile[cpad(j0 + k)] = run | (v == 0u ? empty : 0u);
        run += v;
    }
    __syncthreads();
    if (full) {
        uint4* o4 = reinterpret_cast<uint4*>(cs + base);
#pragma unroll
        for (int k = 0; k < CV; ++k) {
            const int v = k * CT + threadIdx.x;
            o4[v] = make_uint4(tile[cpad(4 * v)], tile[cpad(4 * v + 1)], tile[cpad(4 * v + 2)], tile[cpad(4 * v + 3)]);
        }
    } else {
#pragma unroll
        for (int k = 0; k < CI; ++k) {
            const int j = k * CT + threadIdx.x;
            if (base + j < m) cs[base + j] = tile[cpad(j)];
        }
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) cs[m] = n;
}

__global__ __launch_bounds__(CT) void cell_scatter_kernel(int n, const uint32_t* __restrict__ keyI,
                                                           const uint32_t* __restrict__ rank,
                                                           const uint32_t* __restrict__ cs, uint32_t* __restrict__ tmp,
                                                           uint32_t start_mask) {
    const int i = xcd_tile(blockIdx.x, gridDim.x) * CT + threadIdx.x;
    if (i < n) tmp[(cs[keyI[i] & kCellKeyMask] & start_mask) + rank[i]] = (uint32_t)i;
}

template <bool GATHER>
__global__ __launch_bounds__(CT) void cell_order_kernel(int n, SlabTable T, const uint32_t* __restrict__ tmp,
                                                         const uint32_t* __restrict__ keyI, uint32_t* __restrict__ cs,
                                                         uint32_t* __restrict__ skey, uint32_t* __restrict__ sval,
                                                         const float4* __restrict__ posB,
                                                         const float4* __restrict__ velB,
                                                         const uint32_t* __restrict__ pidB, float4* __restrict__ posA,
                                                         float4* __restrict__ velA, uint32_t* __restrict__ pidA) {
    const int s = xcd_tile(blockIdx.x, gridDim.x) * CT + threadIdx.x;
    if (s >= n) return;
    const uint32_t slot = tmp[s];
    const uint32_t id = pidB[slot];
    const uint32_t kf = keyI[slot];
    const uint32_t k = kf & kCellKeyMask;
    // the first particle of cell k + 1 may be storing its slab offsets into cs[k + 1]
    const uint32_t a = cs[k] & T.start_mask, b = cs[k + 1] & T.start_mask;
    // the fused step's order within a cell is by (slab, id): ids are below 2^27, so the slab
    // bits of the key above the id order both
    const bool slabs = GATHER && T.shift > 0;
    const uint32_t own = slabs ? (kf & ~kCellKeyMask) | id : id;
    uint32_t r = 0, b1 = 0, b2 = 0, b3 = 0;  // b_j: particles of the cell in slabs below j
    for (uint32_t t = a; t < b; ++t) {
        const uint32_t st = tmp[t];
        if (slabs) {
            const uint32_t kt = keyI[st];
            const uint32_t q = kt >> kSlabKeyShift;
            r += ((kt & ~kCellKeyMask) | pidB[st]) < own ? 1u : 0u;
            b1 += q < 1u ? 1u : 0u;
            b2 += q < 2u ? 1u : 0u;
            b3 += q < 3u ? 1u : 0u;
        } else {
            r += pidB[st] < id ? 1u : 0u;
        }
    }
    const uint32_t f = a + r;
    if (slabs && r == 0u) {
        // offsets of slabs 1 .. S - 1, each stored + 1, none if the largest does not fit; the
        // count + 1 where it fits
        const bool four = T.shift == 2;
        const uint32_t top = four ? b3 : b1;
        uint32_t fields = 0u;
        if (top + 1u <= (1u << T.bits) - 1u) {
            fields = (b1 + 1u) << T.start_bits;
            if (four) fields |= ((b2 + 1u) << (T.start_bits + T.bits)) | ((b3 + 1u) << (T.start_bits + 2 * T.bits));
        }
        if (T.count_bits > 0 && (b - a) + 1u <= (1u << T.count_bits) - 1u) fields |= ((b - a) + 1u) << T.count_pos;
        if (fields) cs[k] = a | fields;
    }
    skey[f] = k;
    if (!GATHER) sval[f] = slot;
    if (GATHER) {
        posA[f] = posB[slot];
        velA[f] = velB[slot];
        pidA[f] = id;
    }
}

// Small problems (at most kLdsSortCells cells and kLdsSortMax particles): count, scan and
// scatter in one workgroup with the counts in LDS, in place of the four kernels above, whose
// launches dominate at these sizes.  The LDS atomics rank the particles of a cell in an
// arbitrary order, as the global ones do; cell_order_kernel makes the order.  The global
// per-cell counts are not used (they stay zero).  Measured on 4,096 particles: 20.2 against
// 23.1 us per step with 4,096 cells; with 32,768 cells (or 16,384 particles) the one
// workgroup's sweep costs more than the launches it saves.
constexpr int LT = 1024;                // threads of the one workgroup
constexpr int kLdsSortCells = 8192;
constexpr int kLdsSortMax = 8192;
constexpr int LWAVES = LT / 64;

__global__ __launch_bounds__(LT) void cell_sort_lds_kernel(int n, int m, const uint32_t* __restrict__ keyI,
                                                           uint32_t* __restrict__ rank, uint32_t* __restrict__ cs,
                                                           uint32_t* __restrict__ tmp, SlabTable T) {
    __shared__ uint32_t cnt[kLdsSortCells + kLdsSortCells / 32];  // index c at cpad(c)
    __shared__ uint32_t lw[LWAVES];
    for (int c = threadIdx.x; c < m; c += LT) cnt[cpad(c)] = 0u;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += LT) rank[i] = atomicAdd(&cnt[cpad((int)(keyI[i] & kCellKeyMask))], 1u);
    __syncthreads();
    // exclusive scan: thread t owns the cells [t per, (t + 1) per)
    const int per = (m + LT - 1) / LT;
    const int c0 = min((int)threadIdx.x * per, m), c1 = min(c0 + per, m);
    uint32_t sum = 0;
    for (int c = c0; c < c1; ++c) sum += cnt[cpad(c)];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t incl = wave_incl_scan(sum);
    if (lane == 63) lw[w] = incl;
    __syncthreads();
    uint32_t run = incl - sum;
    for (int k = 0; k < w; ++k) run += lw[k];
    const uint32_t empty = T.empty();
    for (int c = c0; c < c1; ++c) {
        const uint32_t v = cnt[cpad(c)];
        cnt[cpad(c)] = run | (v == 0u ? empty : 0u);
        run += v;
    }
    __syncthreads();
    for (int c = threadIdx.x; c < m; c += LT) cs[c] = cnt[cpad(c)];
    if (threadIdx.x == 0) cs[m] = (uint32_t)n;
    for (int i = threadIdx.x; i < n; i += LT)
        tmp[(cnt[cpad((int)(keyI[i] & kCellKeyMask))] & T.start_mask) + rank[i]] = (uint32_t)i;
}

inline int cblocks(int64_t n) { return (int)((n + CT - 1) / CT); }

}  // namespace

bool cell_sort_global(const sph_ctx* c) { return !(c->sc.cs_cells <= kLdsSortCells && c->n <= kLdsSortMax); }

int64_t cell_scan_tiles(int64_t cells) { return (cells + CTILE - 1) / CTILE; }

// Sorted keys and slots go to keyS[1] / valS[1]; the cell table to cs.  counted: the
// per-cell counts and the ranks (in keyS[0]) of the current keys are there already, and with
// tile_sums_in_force also the tile sums (cell_part array part_phase) (launch_force with
// count; only on the global path, see count_in_force).  The scan clears the other array and
// part_phase names it afterwards.
hipError_t launch_cell_sort(sph_ctx* c, bool gather, hipStream_t s, bool counted) {
    const int n = (int)c->n;
    const int m = c->sc.cs_cells;
    const Buffers& b = c->b;
    uint32_t* rank = b.keyS[0];
    uint32_t* tmp = b.valS[0];
    const int tiles = (int)cell_scan_tiles(m);
    const SlabTable T = slab_table(c->sc);
    if (!cell_sort_global(c)) {
        cell_sort_lds_kernel<<<1, LT, 0, s>>>(n, m, b.keyI, rank, b.cs, tmp, T);
        if (n > 0) {
            if (gather)
                cell_order_kernel<true><<<cblocks(n), CT, 0, s>>>(n, T, tmp, b.keyI, b.cs, b.keyS[1], b.valS[1],
                                                                   b.posB, b.velB, b.pidB, b.posA, b.velA, b.pidA);
            else
                cell_order_kernel<false><<<cblocks(n), CT, 0, s>>>(n, T, tmp, b.keyI, b.cs, b.keyS[1], b.valS[1],
                                                                    nullptr, nullptr, b.pidB, nullptr, nullptr, nullptr);
        }
        c->sorted_keys = b.keyS[1];
        c->sorted_vals = b.valS[1];
        return hipGetLastError();
    }
    const bool summed = counted && tile_sums_in_force(c);
    const int cur = summed ? c->part_phase : 0;
    uint32_t* part = b.cell_part + (size_t)cur * kTileSumCopies * tiles;
    uint32_t* next = b.cell_part + (size_t)(cur ^ 1) * kTileSumCopies * tiles;
    c->part_phase = cur ^ 1;
    if (n > 0 && !counted) cell_count_kernel<<<cblocks(n), CT, 0, s>>>(n, b.keyI, b.cell_cnt, rank);
    if (!summed) cell_scan_reduce<<<tiles, CT, 0, s>>>(b.cell_cnt, m, part);
    const int copies = summed ? kTileSumCopies : 1;
    if (tiles <= kMergedTiles) {
        cell_scan_down<true><<<tiles, CT, 0, s>>>(b.cell_cnt, m, part, next, b.cs, (uint32_t)n, T.empty(), copies);
    } else {
        cell_scan_partials<<<1, CT, 0, s>>>(part, tiles, copies);
        cell_scan_down<false><<<tiles, CT, 0, s>>>(b.cell_cnt, m, part, next, b.cs, (uint32_t)n, T.empty(), copies);
    }
    if (n > 0) {
        cell_scatter_kernel<<<cblocks(n), CT, 0, s>>>(n, b.keyI, rank, b.cs, tmp, T.start_mask);
        if (gather)
            cell_order_kernel<true><<<cblocks(n), CT, 0, s>>>(n, T, tmp, b.keyI, b.cs, b.keyS[1], b.valS[1], b.posB,
                                                               b.velB, b.pidB, b.posA, b.velA, b.pidA);
        else
            cell_order_kernel<false><<<cblocks(n), CT, 0, s>>>(n, T, tmp, b.keyI, b.cs, b.keyS[1], b.valS[1], nullptr,
                                                                nullptr, b.pidB, nullptr, nullptr, nullptr);
    }
    c->sorted_keys = b.keyS[1];
    c->sorted_vals = b.valS[1];
    return hipGetLastError();
}

}  // namespace sph
