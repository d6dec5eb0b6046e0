// sph_device.h -- device helpers shared by the step kernels (sph_kernels.hip) and the
// reverse-mode kernels (sph_adjoint.hip).  The adjoint recomputes the force pass with the
// same code as the step, so its recomputed forward is bitwise the step's own.
#pragma once
#include <math.h>

#include "sph_internal.h"

namespace sph {
namespace {

// Workgroup -> tile mapping that gives each of the 8 XCDs (one L2 each) a contiguous
// range of cell-ordered tiles, so neighbour reads of adjacent tiles hit the same L2.
// Workgroups are dispatched round-robin over XCDs: block b runs on XCD b % 8.
__device__ __forceinline__ int xcd_tile(int bid, int nb) {
    if (nb < 2 * kXcds) return bid;
    const int q = nb / kXcds, r = nb % kXcds;
    const int x = bid % kXcds, local = bid / kXcds;
    return x < r ? x * (q + 1) + local : r * (q + 1) + (x - r) * q + local;
}

// Constants of the instance that owns particle slot i.  A plain step has one instance (the
// kernel argument itself).  In the batched form every instance owns a contiguous range of
// particle ids and, because its cell keys form a contiguous range too, the same range of
// cell-ordered slots; the range is found by binary search over the instance offsets.
template <bool BATCH>
__device__ __forceinline__ const StepConsts& inst_consts(const StepConsts& G, int i) {
    if constexpr (!BATCH) {
        return G;
    } else {
        int lo = 0, hi = G.n_inst;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (G.inst_first[mid] <= i) lo = mid;
            else hi = mid;
        }
        return G.inst_table[lo];
    }
}


// Raw buffer loads with 32-bit byte offsets: no 64-bit address arithmetic per access, and
// an offset past num_records reads zero instead of faulting.  Word 3 of the descriptor is
// the gfx9 value for untyped 32-bit-element buffers.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float4 ld_f4(__amdgpu_buffer_rsrc_t r, uint32_t byte_off) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)byte_off, 0, 0);
    return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]), __uint_as_float(v[3]));
}
__device__ __forceinline__ uint32_t ld_u32(__amdgpu_buffer_rsrc_t r, uint32_t byte_off) {
    return __builtin_amdgcn_raw_buffer_load_b32(r, (int)byte_off, 0, 0);
}

// ---------------------------------------------------------------------------------------
// Voxel keys.  Reference semantics (voxel_kernels.py:9-17, 88-105): per axis
// int32(x / voxel) truncated toward zero, flattened as x + y*Dx + z*Dx*Dy with wrapping
// int32 arithmetic and no clamping.  Grid semantics: the same but floor and clamped to the
// grid, which keeps every pair within h in neighbouring cells for any position.  Both are
// equal for 0 <= x < Dx*voxel, and both are exact: with power-of-two voxels x * (1/voxel)
// is exact in FP32; otherwise the quotient is formed in FP64 as the reference does.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ int grid_coord(float x, int d, const StepConsts& C) {
    float q;
    if (C.pow2) q = floorf(x * C.inv_cell[d]);
    else q = (float)floor((double)x / C.dcell[d]);
    q = fminf(fmaxf(q, 0.0f), (float)(C.dims[d] - 1));  // NaN -> 0
    return (int)q;
}

__device__ __forceinline__ int32_t ref_coord(float x, int d, const StepConsts& C) {
    if (C.pow2) return __float2int_rz(x * C.inv_cell[d]);
    return __double2int_rz((double)x / C.dcell[d]);
}

__device__ __forceinline__ uint32_t sort_key(float x, float y, float z, const StepConsts& C) {
    if (C.ref_keys) {
        uint32_t vx = (uint32_t)ref_coord(x, 0, C), vy = (uint32_t)ref_coord(y, 1, C),
                 vz = (uint32_t)ref_coord(z, 2, C);
        uint32_t key = vx + vy * (uint32_t)C.dims[0] + vz * (uint32_t)C.dims[0] * (uint32_t)C.dims[1];
        return key ^ 0x80000000u;  // signed order as unsigned
    }
    const int cx = grid_coord(x, 0, C);
    uint32_t key = (uint32_t)(C.cell_off + cx + C.dims[0] * (grid_coord(y, 1, C) + C.dims[1] * grid_coord(z, 2, C)));
    if (C.slab_shift > 0) {
        // the x slab within the cell: floor(S x / edge) - S cx, clamped to the cell's slabs
        // (the edge cells' outer slabs extend to infinity, NaN -> 0); exact for power-of-two
        // edges, and monotone in x otherwise
        const float S = (float)(1 << C.slab_shift);
        const float q = floorf(x * C.inv_cell[0] * S) - S * (float)cx;
        key |= (uint32_t)fminf(fmaxf(q, 0.0f), S - 1.0f) << kSlabKeyShift;
    }
    return key;
}

// The start of a cell held by a cell-table entry, and the offset + 1 of slab k (1 .. S - 1)
// from it (0 for k = 0, and 0 when not known); see StepConsts::slab_shift.
__device__ __forceinline__ uint32_t table_start(uint32_t w, const StepConsts& C) { return w & C.start_mask; }
__device__ __forceinline__ uint32_t slab_field(uint32_t w, uint32_t k, const StepConsts& C) {
    return __builtin_amdgcn_ubfe(w, (uint32_t)(C.start_bits + C.slab_bits * ((int)k - 1)),
                                 k ? (uint32_t)C.slab_bits : 0u);
}

// bucket of a sort key in the cell table: -1 below the grid, n_cells above it
__device__ __forceinline__ int key_bucket(uint32_t ukey, const StepConsts& C) {
    if (!C.ref_keys) return (int)ukey;
    const int32_t s = (int32_t)(ukey ^ 0x80000000u);
    return s < 0 ? -1 : (s >= C.n_cells ? C.n_cells : s);
}

__device__ __forceinline__ bool finite3(float x, float y, float z) {
    return isfinite(x) && isfinite(y) && isfinite(z);
}

// Sort input of the next step.  SLOT_KEYS (the NAIVE formulation's counting sort): the key
// of the particle in posB slot `slot`, indexed by the slot, so the keys stay in the current
// cell order (coalesced stores; the sort's atomics and scatters touch neighbouring cells).
// Otherwise (the radix sort, which must see particle-id order to be stable by id): key and
// slot indexed by particle id.
// Returns the key.
template <bool SLOT_KEYS>
__device__ __forceinline__ uint32_t hash_particle(const StepConsts& C, float x, float y, float z, uint32_t pid,
                                                  uint32_t slot, uint32_t* __restrict__ keyI,
                                                  uint32_t* __restrict__ valI, uint32_t* __restrict__ flags) {
    const uint32_t key = sort_key(x, y, z, C);
    if constexpr (SLOT_KEYS) {
        keyI[slot] = key;
    } else {
        keyI[pid] = key;
        valI[pid] = slot;
    }
    if (!finite3(x, y, z)) atomicAdd(flags + kFlagNonFinite, 1u);
    return key;
}

// Rank of a key among the equal keys counted into cnt (the counting sort's per-cell
// counts): rank = atomicAdd(&cnt[key], 1), in an arbitrary order among the particles of a
// cell.  Consecutive slots are often in the same cell, so a wavefront's lanes come in runs
// of equal keys: the first lane of each run adds the run's length with one atomic and the
// run's lanes take consecutive ranks from the returned base.  The lanes that hold a key
// (valid) must be a prefix of the wavefront, and every active lane must call this together.
__device__ __forceinline__ uint32_t rank_in_cell(uint32_t key, bool valid, uint32_t* __restrict__ cnt) {
    const int lane = threadIdx.x & 63;
    const uint32_t prev = __shfl_up(key, 1);
    const bool head = valid && (lane == 0 || prev != key);
    const uint64_t heads = __ballot(head);
    const uint64_t valids = __ballot(valid);
    const uint64_t upto = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1ull);
    const int first = 63 - __clzll((long long)(heads & upto));  // first lane of this lane's run
    const uint64_t above = heads & ~upto;
    // first lane of the next run, or one past the last valid lane
    const int next = above ? __ffsll((unsigned long long)above) - 1 : 64 - __clzll((long long)valids);
    uint32_t base = 0;
    if (head) base = atomicAdd(&cnt[key], (uint32_t)(next - lane));
    base = __shfl(base, first);
    return base + (uint32_t)(lane - first);
}

// Adds each active lane's cell to the sum of its scan tile, the counting sort's tile sums:
// part[cell >> kScanTileShift] of the copy that part points to.  The lanes with the first
// remaining lane's tile add their number with one atomic, until none is left: the keys of a
// wavefront are nearly in cell order, so that is mostly one atomic per wavefront.  Every
// active lane must call this together (the loop is uniform across the wavefront).
__device__ __forceinline__ void add_to_tile_sums(uint32_t cell, uint32_t* __restrict__ part) {
    const uint32_t tile = cell >> kScanTileShift;
    bool pending = true;
    for (uint64_t left = __ballot(1); left != 0; left = __ballot(pending)) {
        const int lead = __ffsll((unsigned long long)left) - 1;
        const uint32_t t = (uint32_t)__builtin_amdgcn_readlane((int)tile, lead);
        const bool mine = pending && tile == t;
        const uint64_t same = __ballot(mine);
        if ((int)(threadIdx.x & 63) == lead) atomicAdd(&part[t], (uint32_t)__popcll(same));
        pending = pending && !mine;
    }
}

// VOXEL formulation: whether a finite position lies outside the voxel grid, i.e. its key
// (truncated voxel index per axis, voxel_kernels.py:9-17) leaves [0, dims) on some axis and
// so wraps onto the key of another voxel or off the table.  Positions in (-voxel, 0) truncate
// into voxel 0 and count as inside (the density pass treats voxel 0 as open below).
// Non-finite positions count as inside: no pair test ever accepts them.
__device__ __forceinline__ bool outside_voxel_grid(float x, float y, float z, const StepConsts& C) {
    if (!finite3(x, y, z)) return false;
    const int32_t vx = ref_coord(x, 0, C), vy = ref_coord(y, 1, C), vz = ref_coord(z, 2, C);
    return vx < 0 || vx >= C.dims[0] || vy < 0 || vy >= C.dims[1] || vz < 0 || vz >= C.dims[2];
}

// ---------------------------------------------------------------------------------------
// Neighbour candidates of the NAIVE formulation.  For each row of cells (y, z) whose slab
// comes within h of the particle, the x cells are those overlapping [x - e, x + e] with
// e = sqrt(h^2 - dy^2 - dz^2); cells of one row are consecutive keys, so the candidates
// of a row are one contiguous range cs[first]..cs[last+1].  Boundary cells extend to
// infinity (positions outside the grid are clamped into them).  The search radius carries
// a small margin; the exact test r^2 <= h^2 is applied to each candidate.
// ---------------------------------------------------------------------------------------
// The rows of cells that the sphere of radius h_search around a particle reaches, in
// increasing cell order (cz, then cy): a row is the run of cells along x that the sphere's
// chord through the nearest point of the row's (y, z) slab covers, taken in x slabs of the
// cells (StepConsts::slab_shift; whole cells when the cells are not cut), so that its
// candidates are one contiguous range of the cell order.  The position is clamped into the
// grid first; an edge cell holds the particles beyond it, so a particle outside the grid
// gets a superset of its rows.  The margin of h_search over h covers the rounding of this
// float arithmetic (a slab it adds or drops holds only particles beyond h).
struct RowGrid {
    float gx, gy, gz, fx, hs2, mx, sinv;  // fx, mx: x in slabs; sinv: slabs per unit length
    int ylo, yhi, zlo, zhi;

    __device__ __forceinline__ RowGrid(float px, float py, float pz, const StepConsts& C) {
        const float hs = C.h_search;
        hs2 = hs * hs;
        const float S = (float)(1 << C.slab_shift);
        const float my = (float)(C.dims[1] - 1), mz = (float)(C.dims[2] - 1);
        gx = fminf(fmaxf(px, 0.0f), (float)C.dims[0] * C.cell[0]);  // NaN -> 0
        gy = fminf(fmaxf(py, 0.0f), (my + 1.0f) * C.cell[1]);
        gz = fminf(fmaxf(pz, 0.0f), (mz + 1.0f) * C.cell[2]);
        ylo = (int)__builtin_amdgcn_fmed3f((gy - hs) * C.inv_cell[1], 0.0f, my);
        yhi = (int)__builtin_amdgcn_fmed3f((gy + hs) * C.inv_cell[1], 0.0f, my);
        zlo = (int)__builtin_amdgcn_fmed3f((gz - hs) * C.inv_cell[2], 0.0f, mz);
        zhi = (int)__builtin_amdgcn_fmed3f((gz + hs) * C.inv_cell[2], 0.0f, mz);
        sinv = C.inv_cell[0] * S;
        fx = gx * sinv;
        mx = (float)C.dims[0] * S - 1.0f;
    }

    // distance from the particle to the slab [c s, (c + 1) s), zero inside it
    __device__ __forceinline__ static float slab(float g, int c, float s) {
        const float f = (float)c;
        return fmaxf(fmaxf(fmaf(f, s, -g), fmaf(-f, s, g - s)), 0.0f);
    }

    // h_search^2 less the squared distance to the slab of cz
    __device__ __forceinline__ float rest_z(int cz, const StepConsts& C) const {
        const float ez = slab(gz, cz, C.cell[2]);
        return fmaf(-ez, ez, hs2);
    }

    // the row at cy of the slab of cz (rz: rest_z(cz)): its first x slab q0 and one past its
    // last, q1, counted from the row's first cell; false if the sphere misses the row's slab
    __device__ __forceinline__ bool row(int cy, float rz, const StepConsts& C, uint32_t& q0, uint32_t& q1) const {
        const float ey = slab(gy, cy, C.cell[1]);
        const float rem = fmaf(-ey, ey, rz);
        if (!(rem >= 0.0f)) return false;
        const float ex = __builtin_amdgcn_sqrtf(rem) * sinv;
        q0 = (uint32_t)(int)__builtin_amdgcn_fmed3f(fx - ex, 0.0f, mx);
        q1 = (uint32_t)(int)__builtin_amdgcn_fmed3f(fx + ex, 0.0f, mx) + 1u;
        return true;
    }

    // the candidates [k0, k1) of slabs q0 .. q1 - 1 of the row whose first cell is row0, from
    // the cell-table entries of the cell holding q0 and of the cell holding q1 (the next
    // cell's start when q1 is a cell's first slab), ld(c) reading entry c
    template <typename LD>
    __device__ __forceinline__ static void range(int row0, uint32_t q0, uint32_t q1, const StepConsts& C, LD&& ld,
                                                 uint32_t& k0, uint32_t& k1) {
        const int sh = C.slab_shift;
        const uint32_t m = (1u << sh) - 1u;
        const int e = row0 + (int)(q1 >> sh);
        const uint32_t a = ld(row0 + (int)(q0 >> sh)), b = ld(e);
        const uint32_t fa = slab_field(a, q0 & m, C), fb = slab_field(b, q1 & m, C);
        k0 = table_start(a, C) + (fa ? fa - 1u : 0u);
        k1 = table_start(b, C) + (fb ? fb - 1u : 0u);
        // the end lies inside a cell whose slab offsets are not known: the whole cell, from its
        // count if the entry keeps it, else from the next cell's start
        if ((q1 & m) != 0u && fb == 0u) {
            const uint32_t nb = C.count_bits ? __builtin_amdgcn_ubfe(b, (uint32_t)C.count_pos, (uint32_t)C.count_bits) : 0u;
            k1 = nb ? table_start(b, C) + nb - 1u : table_start(ld(e + 1), C);
        }
    }
};

template <typename LD, typename F>
__device__ __forceinline__ void for_each_row(float px, float py, float pz, const StepConsts& C, LD&& ld, F&& f) {
    const RowGrid R(px, py, pz, C);
    for (int cz = R.zlo; cz <= R.zhi; ++cz) {
        const float rz = R.rest_z(cz, C);
        int row0 = C.cell_off + C.dims[0] * (R.ylo + C.dims[1] * cz);
        for (int cy = R.ylo; cy <= R.yhi; ++cy, row0 += C.dims[0]) {
            uint32_t q0, q1;
            if (R.row(cy, rz, C, q0, q1)) {
                uint32_t k0, k1;
                RowGrid::range(row0, q0, q1, C, ld, k0, k1);
                f(k0, k1);
            }
        }
    }
}

// for_each_row for a group of L lanes sharing one particle (groups aligned in the
// wavefront, l the lane's index in its group): the particle's (cy, cz) pairs are numbered
// t = 0, 1, ... in for_each_row's order, lane l works out the rows of pairs l, l + L, ...
// (geometry and cell-table entries), and every lane of the group then gets f(k0, k1) for
// each row in order, the row's candidate range taken from the lane that owns it.  Rows the
// sphere misses are skipped; rows without candidates are passed on (k0 == k1).
template <int L, typename F>
__device__ __forceinline__ void for_each_row_range(float px, float py, float pz, const StepConsts& C,
                                                   __amdgpu_buffer_rsrc_t rcs, uint32_t l, F&& f) {
    const int gbase = (int)((threadIdx.x & 63u) & ~(uint32_t)(L - 1));
    const RowGrid R(px, py, pz, C);
    const int ny = R.yhi - R.ylo + 1;
    const int npairs = ny * (R.zhi - R.zlo + 1);
    const float inv_ny = 1.0f / (float)ny;
    for (int t0 = 0; t0 < npairs; t0 += L) {
        uint32_t a0 = 0xffffffffu, a1 = 0u;  // an empty range unless the row exists
        const int t = t0 + (int)l;
        if (t < npairs) {
            const int q = (int)(((float)t + 0.5f) * inv_ny);  // exact: t and ny are small
            const int cz = R.zlo + q, cy = R.ylo + (t - q * ny);
            const int row0 = C.cell_off + C.dims[0] * (cy + C.dims[1] * cz);
            uint32_t q0, q1;
            if (R.row(cy, R.rest_z(cz, C), C, q0, q1))
                RowGrid::range(row0, q0, q1, C, [&](int c) { return ld_u32(rcs, (uint32_t)c * 4u); }, a0, a1);
        }
        const int last = min(L, npairs - t0);
        for (int j = 0; j < last; ++j) {
            const uint32_t k0 = __shfl(a0, gbase + j), k1 = __shfl(a1, gbase + j);
            if (k0 <= k1) f(k0, k1);
        }
    }
}

template <typename F>
__device__ __forceinline__ void for_each_candidate(float px, float py, float pz, const StepConsts& C,
                                                   const uint32_t* __restrict__ cs, F&& f) {
    for_each_row(px, py, pz, C, [&](int c) { return cs[c]; }, [&](uint32_t k0, uint32_t k1) {
        for (uint32_t k = k0; k < k1; ++k) f(k);
    });
}

// ---------------------------------------------------------------------------------------
// Wall collision.  Box: base_kernels.py:75-98.  Pipe: base_kernels.py:56-72, 101-127 and
// util_kernels.py:38-204, in FP32, with the particle's own xoroshiro128+ stream.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ void collide_box(const StepConsts& C, float& x, float& v, int d) {
    bool bounced = false;
    if (x < 0.0f) { x = C.box_lo_target; bounced = true; }
    if (x > C.space[d]) { x = C.box_hi_target[d]; bounced = true; }
    if (bounced) { v *= -1.0f; v *= C.damp; }
}

struct Vec3 { float x, y, z; };
__device__ __forceinline__ float len3(Vec3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ float dot3(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

__device__ int pipe_find_segment(const PipeTable& T, int rows, float x) {
    for (int j = 0; j < rows - 1; ++j)
        if (T.row[j][0] <= x && x < T.row[j + 1][0]) return j;
    return -1;
}

__device__ float pipe_radius(const PipeTable& T, float x, int seg) {
    const float r0 = T.row[seg][3], r1 = T.row[seg + 1][3], len = T.row[seg][4];
    if (r0 == r1) return r0;
    if (r0 < r1) {
        const float delta = x - T.xbeg[seg];
        const float tl = len * r0 / (r1 - r0);
        return r0 * (1.0f + delta / tl);
    }
    const float delta = T.xbeg[seg] + len - x;
    const float tl = len * r1 / (r0 - r1);
    return r1 * (1.0f + delta / tl);
}

__device__ bool pipe_out(const PipeTable& T, Vec3 p, int seg) {
    const float yn = p.y - T.row[seg][1], zn = p.z - T.row[seg][2];
    return sqrtf(yn * yn + zn * zn) > pipe_radius(T, p.x, seg);
}

__device__ void pipe_solve_collision(const PipeTable& T, Vec3& p, Vec3& v, int seg) {
    const float sr = T.row[seg][3], er = T.row[seg + 1][3];
    const float yn = p.y - T.row[seg][1], zn = p.z - T.row[seg][2];
    const float hh = sqrtf(yn * yn + zn * zn);
    // calc_edge_vector (util_kernels.py:152-171)
    Vec3 first{T.row[seg][0], yn / hh * sr, zn / hh * sr};
    const Vec3 second{T.row[seg + 1][0], yn / hh * er, zn / hh * er};
    const Vec3 edge{second.x - first.x, second.y - first.y, second.z - first.z};
    first.y += T.row[seg][1];
    first.z += T.row[seg][2];
    // calc_dt (util_kernels.py:113-121) with calc_distance_to_pipe (:124-135)
    const float sl = len3(v), el = len3(edge);
    const Vec3 pv{first.x - p.x, first.y - p.y, first.z - p.z};
    const Vec3 cr{edge.y * pv.z - edge.z * pv.y, edge.z * pv.x - edge.x * pv.z, edge.x * pv.y - edge.y * pv.x};
    const float dist = len3(cr) / el;
    // sin of the angle between v and the wall line.  The reference writes sqrt(1 - cos^2)
    // (util_kernels.py:118); for a particle sliding along a conical wall cos is within 1e-8
    // of 1, which FP32 rounds above 1 (sqrt of a negative: NaN).  |v x edge| / (|v| |edge|)
    // is the same quantity and stays accurate at small angles.
    const Vec3 ve{v.y * edge.z - v.z * edge.y, v.z * edge.x - v.x * edge.z, v.x * edge.y - v.y * edge.x};
    const float sin_a = len3(ve) / (sl * el);
    const float dt = -(dist / (sin_a + 0.001f)) / sl;
    const Vec3 cp{p.x + v.x * dt, p.y + v.y * dt, p.z + v.z * dt};
    // reflection (util_kernels.py:191-199)
    if (sr == er) {
        v.y = -v.y;
        v.z = -v.z;
    } else {
        const float c = dot3(edge, v) / (el * sl);
        const Vec3 sv{2.0f * edge.x / el * c * sl, 2.0f * edge.y / el * c * sl, 2.0f * edge.z / el * c * sl};
        v = Vec3{sv.x - v.x, sv.y - v.y, sv.z - v.z};
    }
    const Vec3 dd{p.x - cp.x, p.y - cp.y, p.z - cp.z};
    const float way = len3(dd);
    const float svl = len3(v);
    p = Vec3{cp.x + v.x * way / svl, cp.y + v.y * way / svl, cp.z + v.z * way / svl};
}

// xoroshiro128+ (Blackman & Vigna 2016; rotations 55, 14, 36) and the 53-bit uniform of
// numba.cuda.random.xoroshiro128p_uniform_float64, rounded once to FP32.
__device__ __forceinline__ uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
__device__ __forceinline__ uint64_t xoro_next(uint64_t& s0, uint64_t& s1) {
    const uint64_t a = s0;
    uint64_t b = s1;
    const uint64_t result = a + b;
    b ^= a;
    s0 = rotl64(a, 55) ^ b ^ (b << 14);
    s1 = rotl64(b, 36);
    return result;
}
__device__ __forceinline__ float xoro_uniform(uint64_t& s0, uint64_t& s1) {
    return (float)(double)(xoro_next(s0, s1) >> 11) * 0x1p-53f;
}

__device__ void collide_pipe(const StepConsts& C, const PipeTable& T, Vec3& p, Vec3& v, uint64_t* __restrict__ rng,
                             uint32_t pid) {
    const int rows = C.pipe_rows;
    const int seg = pipe_find_segment(T, rows, p.x);
    if (seg == -1) {
        // put_particle_at_pipe_begin (base_kernels.py:115-127)
        if (p.x < 0.0f) {
            p.x = -p.x;
            v.x = -v.x;
            if (pipe_out(T, p, 0)) pipe_solve_collision(T, p, v, 0);
        } else {
            p.x = 0.0f;
            // rand_position_inside_pipe (base_kernels.py:101-112); a pipe that does not start
            // at x <= 0 has no segment here and the reference reads past its table, so this
            // uses segment 0 instead
            int s0g = pipe_find_segment(T, rows, p.x);
            if (s0g < 0) s0g = 0;
            const float R = pipe_radius(T, p.x, s0g);
            uint64_t a = rng[2 * (uint64_t)pid], b = rng[2 * (uint64_t)pid + 1];
            const float r = R * sqrtf(xoro_uniform(a, b));
            const float theta = xoro_uniform(a, b) * 2.0f * 3.14159265358979f;
            rng[2 * (uint64_t)pid] = a;
            rng[2 * (uint64_t)pid + 1] = b;
            p.y = T.row[0][1] + r * cosf(theta);
            p.z = T.row[0][2] + r * sinf(theta);
        }
    } else if (pipe_out(T, p, seg)) {
        pipe_solve_collision(T, p, v, seg);
    }
}

// Symplectic Euler (base_kernels.py:30-53) before the wall collision: result force f,
// new velocity v and new position x.
struct Integrated {
    Vec3 f, v, x;
};
__device__ __forceinline__ Integrated integrate_particle(const StepConsts& C, float4 p, float4 v, float rho, Vec3 P,
                                                         Vec3 V) {
    Integrated r;
    r.f = Vec3{C.ext[0] + -P.x + V.x, C.ext[1] + -P.y + V.y, C.ext[2] + -P.z + V.z};
    r.v = Vec3{v.x + r.f.x / rho * C.dt, v.y + r.f.y / rho * C.dt, v.z + r.f.z / rho * C.dt};
    r.x = Vec3{p.x + r.v.x * C.dt, p.y + r.v.y * C.dt, p.z + r.v.z * C.dt};
    return r;
}

// Whether collide_box moves component d of a position x (the same two tests).
__device__ __forceinline__ bool box_bounced(const StepConsts& C, float x, int d) {
    const bool below = x < 0.0f;
    return below || (below ? C.box_lo_target : x) > C.space[d];
}

// Symplectic Euler (base_kernels.py:30-53), collision, output in the same sorted slot of
// posB/velB, hash of the new position for the next step (SLOT_KEYS: see hash_particle);
// returns the key.
template <int MODE, bool SLOT_KEYS>
__device__ __forceinline__ uint32_t finish_particle(const StepConsts& C, int i, float4 p, float4 v, uint32_t pid,
                                                    float rho, Vec3 P, Vec3 V, float4* __restrict__ posOut,
                                                    float4* __restrict__ velOut, uint32_t* __restrict__ pidOut,
                                                    uint32_t* __restrict__ keyI, uint32_t* __restrict__ valI,
                                                    uint64_t* __restrict__ rng, const PipeTable* __restrict__ pipe,
                                                    float4* __restrict__ dbgP, float4* __restrict__ dbgV,
                                                    float4* __restrict__ dbgF, uint32_t* __restrict__ flags) {
    const Integrated I = integrate_particle(C, p, v, rho, P, V);
    const float fx = I.f.x, fy = I.f.y, fz = I.f.z;
    Vec3 vv = I.v, pp = I.x;
    if (MODE == SPH_MODE_PIPE) {
        collide_pipe(C, *pipe, pp, vv, rng, pid);
    } else {
        collide_box(C, pp.x, vv.x, 0);
        collide_box(C, pp.y, vv.y, 1);
        collide_box(C, pp.z, vv.z, 2);
    }
    posOut[i] = make_float4(pp.x, pp.y, pp.z, rho);
    velOut[i] = make_float4(vv.x, vv.y, vv.z, 0.0f);
    pidOut[i] = pid;
    if (dbgP) {
        dbgP[i] = make_float4(P.x, P.y, P.z, 0.0f);
        dbgV[i] = make_float4(V.x, V.y, V.z, 0.0f);
        dbgF[i] = make_float4(fx, fy, fz, 0.0f);
    }
    return hash_particle<SLOT_KEYS>(C, pp.x, pp.y, pp.z, pid, (uint32_t)i, keyI, valI, flags);
}

// Neighbour-list entries (ListEntry): the neighbour's index relative to the particle's own.
// Layout: slots t and t + 1 (t even) of particle i share the 32-bit word (t / 2) * n + i, so
// a lane reads two slots per load and consecutive particles' words are contiguous.
__device__ __forceinline__ ListEntry list_encode(uint32_t k, int i) { return (ListEntry)(uint16_t)((int)k - i); }
__device__ __forceinline__ uint32_t list_decode(ListEntry e, int i) { return (uint32_t)(i + (int)(int16_t)e); }
__device__ __forceinline__ size_t list_at(int n, int i, uint32_t t) {
    return ((size_t)(t >> 1) * (size_t)n + (size_t)i) * 2u + (t & 1u);
}

// Pressure and viscosity sums of one pair (naive_kernels.py:38-52, :70-77) without the
// constant factors.  q = (x, y, z, p) and vj = (vx, vy, vz, 1/rho) of the neighbour.
struct PairSums {
    float ax = 0.f, ay = 0.f, az = 0.f, bx = 0.f, by = 0.f, bz = 0.f;
    // self = the pair is the particle itself: 1/r is taken as 0, which makes both terms zero
    // (the pressure term carries a factor 1/r, the viscosity term a factor v_j - v_i = 0).
    // 1/r is the hardware reciprocal square root: rsqrtf gives the same for every r^2 above
    // the FP32 denormal range and only rescales below it (distances under 1e-19).
    __device__ __forceinline__ void add(const StepConsts& C, float4 p, float4 v, float4 q, float4 vj,
                                        bool self = false) {
        const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        const float rinv = self ? 0.0f : __builtin_amdgcn_rsqf(r2);
        const float hr = C.h - r2 * rinv;
        const float cp = (p.w + q.w) * vj.w * hr * hr * rinv;
        ax += cp * dx;
        ay += cp * dy;
        az += cp * dz;
        const float cv = hr * vj.w;
        bx += cv * (vj.x - v.x);
        by += cv * (vj.y - v.y);
        bz += cv * (vj.z - v.z);
    }
};

// List slots (and their neighbour gathers) in flight per lane in the force pass.
#ifndef SPH_FORCE_UNROLL
#define SPH_FORCE_UNROLL 4
#endif
constexpr int kForceUnroll = SPH_FORCE_UNROLL;
#ifndef SPH_FORCE_PREFETCH
#define SPH_FORCE_PREFETCH 1
#endif

// Pressure and viscosity sums of particle i over its neighbour list (or, for a particle
// with more neighbours than list slots, a rescan of its candidates).  The list holds the
// particle itself; PairSums gives that pair a zero contribution.
__device__ __forceinline__ PairSums naive_pair_sums(const StepConsts& C, int i, int n, float4 p, float4 v,
                                                    const float4* __restrict__ pos, const float4* __restrict__ vel,
                                                    const uint32_t* __restrict__ cs, const ListEntry* __restrict__ nl,
                                                    const uint32_t* __restrict__ nn) {
    const __amdgpu_buffer_rsrc_t rpos = make_rsrc(pos, (uint32_t)n * 16u);
    const __amdgpu_buffer_rsrc_t rvel = make_rsrc(vel, (uint32_t)n * 16u);
    PairSums S;
    const uint32_t cnt = nn[i];
    if (cnt <= (uint32_t)kListCap) {
        // two slots per 32-bit load (list_at); the next four slots are loaded while this
        // group's neighbours are gathered, so each group waits on one memory latency
        static_assert(kForceUnroll == 4, "the force pass walks the list four slots at a time");
        const uint32_t* words = reinterpret_cast<const uint32_t*>(nl) + i;
        const size_t stride = (size_t)n;
        uint32_t t = 0;
        uint32_t wn0 = 0u, wn1 = 0u;
        if (cnt >= 4u) {
            wn0 = words[0];
            wn1 = words[stride];
        }
        for (; t + 4u <= cnt; t += 4u) {
            const uint32_t w0 = wn0, w1 = wn1;
            words += 2 * stride;
            if (t + 8u <= cnt) {
                wn0 = words[0];
                wn1 = words[stride];
            }
            const uint32_t k[4] = {list_decode((ListEntry)w0, i), list_decode((ListEntry)(w0 >> 16), i),
                                   list_decode((ListEntry)w1, i), list_decode((ListEntry)(w1 >> 16), i)};
            float4 q[4], w[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                q[u] = ld_f4(rpos, k[u] * 16u);
                w[u] = ld_f4(rvel, k[u] * 16u);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) S.add(C, p, v, q[u], w[u], k[u] == (uint32_t)i);
        }
        for (; t < cnt; t += 2u) {  // the last one to three slots, a word at a time
            const uint32_t wd = words[0];
            words += stride;
            const uint32_t k0 = list_decode((ListEntry)wd, i);
            S.add(C, p, v, ld_f4(rpos, k0 * 16u), ld_f4(rvel, k0 * 16u), k0 == (uint32_t)i);
            if (t + 1u < cnt) {
                const uint32_t k1 = list_decode((ListEntry)(wd >> 16), i);
                S.add(C, p, v, ld_f4(rpos, k1 * 16u), ld_f4(rvel, k1 * 16u), k1 == (uint32_t)i);
            }
        }
    } else {
        for_each_candidate(p.x, p.y, p.z, C, cs, [&](uint32_t k) {
            const float4 q = pos[k];
            const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
            if (dx * dx + dy * dy + dz * dz <= C.h2 && k != (uint32_t)i) S.add(C, p, v, q, vel[k]);
        });
    }
    return S;
}

// The same sums with L lanes per particle (L = 1 is naive_pair_sums): lane l of the group
// walks list slots l, l + L, ... (or, past the list, every L-th candidate of each row), and
// the partial sums are added across the group; every lane of the group returns the total.
template <int L>
__device__ __forceinline__ PairSums naive_pair_sums_lanes(const StepConsts& C, int i, uint32_t l, int n, float4 p,
                                                          float4 v, const float4* __restrict__ pos,
                                                          const float4* __restrict__ vel,
                                                          const uint32_t* __restrict__ cs,
                                                          const ListEntry* __restrict__ nl,
                                                          const uint32_t* __restrict__ nn) {
    if constexpr (L == 1) {
        return naive_pair_sums(C, i, n, p, v, pos, vel, cs, nl, nn);
    } else {
        const __amdgpu_buffer_rsrc_t rpos = make_rsrc(pos, (uint32_t)n * 16u);
        const __amdgpu_buffer_rsrc_t rvel = make_rsrc(vel, (uint32_t)n * 16u);
        PairSums S;
        const uint32_t cnt = nn[i];
        if (cnt <= (uint32_t)kListCap) {
#if SPH_FORCE_PREFETCH
            // the lane's next list slot is loaded while this slot's neighbour is gathered
            uint32_t kn = l < cnt ? list_decode(nl[list_at(n, i, l)], i) : 0u;
            for (uint32_t t = l; t < cnt; t += L) {
                const uint32_t k = kn;
                if (t + L < cnt) kn = list_decode(nl[list_at(n, i, t + L)], i);
                S.add(C, p, v, ld_f4(rpos, k * 16u), ld_f4(rvel, k * 16u), k == (uint32_t)i);
            }
#else
            for (uint32_t t = l; t < cnt; t += L) {
                const uint32_t k = list_decode(nl[list_at(n, i, t)], i);
                S.add(C, p, v, ld_f4(rpos, k * 16u), ld_f4(rvel, k * 16u), k == (uint32_t)i);
            }
#endif
        } else {
            const __amdgpu_buffer_rsrc_t rcs = make_rsrc(cs, ((uint32_t)C.cell_off + (uint32_t)C.n_cells + 1u) * 4u);
            for_each_row_range<L>(p.x, p.y, p.z, C, rcs, l, [&](uint32_t k0, uint32_t k1) {
                for (uint32_t k = k0 + l; k < k1; k += L) {
                    const float4 q = ld_f4(rpos, k * 16u);
                    const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
                    if (dx * dx + dy * dy + dz * dz <= C.h2 && k != (uint32_t)i) S.add(C, p, v, q, ld_f4(rvel, k * 16u));
                }
            });
        }
#pragma unroll
        for (int o = 1; o < L; o <<= 1) {
            S.ax += __shfl_xor(S.ax, o);
            S.ay += __shfl_xor(S.ay, o);
            S.az += __shfl_xor(S.az, o);
            S.bx += __shfl_xor(S.bx, o);
            S.by += __shfl_xor(S.by, o);
            S.bz += __shfl_xor(S.bz, o);
        }
        return S;
    }
}

}  // namespace
}  // namespace sph
