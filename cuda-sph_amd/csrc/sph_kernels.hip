// sph_kernels.hip -- gfx950 kernels of one SPH step (everything but the radix sort).
//
// One thread per particle, particles in cell order.  A step is
//   sort (sph_radix.hip, last pass gathers posB/velB -> posA/velA in cell order)
//   cell_table   cs[c] = first sorted index with key >= c (dense, no empty markers)
//   density      rho, p = K (rho - rho0) into posA.w and 1/rho into velA.w, neighbour lists
//   force        pressure + viscosity sums, symplectic Euler, wall collision, then the
//                voxel hash of the new position for the next step's sort
// The NAIVE formulation (the reference default) sums over every j with |x_i - x_j|^2 <= h^2;
// the grid only skips pairs that are provably farther apart.  The VOXEL formulation
// reproduces voxel_kernels.py including its 32-neighbour cap and enumeration order.
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
    return (uint32_t)(C.cell_off + grid_coord(x, 0, C) +
                      C.dims[0] * (grid_coord(y, 1, C) + C.dims[1] * grid_coord(z, 2, C)));
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

__device__ __forceinline__ void hash_particle(const StepConsts& C, float x, float y, float z, uint32_t pid,
                                              uint32_t slot, uint32_t* __restrict__ keyI, uint32_t* __restrict__ valI,
                                              uint32_t* __restrict__ flags) {
    keyI[pid] = sort_key(x, y, z, C);
    valI[pid] = slot;
    if (!finite3(x, y, z)) atomicAdd(flags, 1u);
}

// ---------------------------------------------------------------------------------------
// Upload: float64 (N,3) AoS or float32 SoA in particle order -> posB/velB (slot = id), hash.
// ---------------------------------------------------------------------------------------
template <bool BATCH>
__global__ __launch_bounds__(256) void upload_f64_kernel(StepConsts G, const double* __restrict__ pos,
                                                          const double* __restrict__ vel, float4* __restrict__ posB,
                                                          float4* __restrict__ velB, uint32_t* __restrict__ pidB,
                                                          uint32_t* __restrict__ keyI, uint32_t* __restrict__ valI,
                                                          uint32_t* __restrict__ flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G.n) return;
    const StepConsts& C = inst_consts<BATCH>(G, i);
    const float x = (float)pos[3 * (int64_t)i], y = (float)pos[3 * (int64_t)i + 1], z = (float)pos[3 * (int64_t)i + 2];
    posB[i] = make_float4(x, y, z, 0.0f);
    velB[i] = make_float4((float)vel[3 * (int64_t)i], (float)vel[3 * (int64_t)i + 1], (float)vel[3 * (int64_t)i + 2],
                          0.0f);
    pidB[i] = (uint32_t)i;
    hash_particle(C, x, y, z, (uint32_t)i, (uint32_t)i, keyI, valI, flags);
}

template <bool BATCH>
__global__ __launch_bounds__(256) void upload_f32_kernel(StepConsts G, const float* __restrict__ px,
                                                          const float* __restrict__ py, const float* __restrict__ pz,
                                                          const float* __restrict__ vx, const float* __restrict__ vy,
                                                          const float* __restrict__ vz, float4* __restrict__ posB,
                                                          float4* __restrict__ velB, uint32_t* __restrict__ pidB,
                                                          uint32_t* __restrict__ keyI, uint32_t* __restrict__ valI,
                                                          uint32_t* __restrict__ flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G.n) return;
    const StepConsts& C = inst_consts<BATCH>(G, i);
    const float x = px[i], y = py[i], z = pz[i];
    posB[i] = make_float4(x, y, z, 0.0f);
    velB[i] = make_float4(vx[i], vy[i], vz[i], 0.0f);
    pidB[i] = (uint32_t)i;
    hash_particle(C, x, y, z, (uint32_t)i, (uint32_t)i, keyI, valI, flags);
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint32_t v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------------------
// Cell table from the sorted keys: cs[b] = j for every bucket b in (bucket(j-1), bucket(j)],
// with bucket(-1) = -1 and bucket(N) = n_cells, so each of cs[0..n_cells] is written once.
// A long run of empty cells is filled by the whole wavefront, 64 entries per store.
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void cell_table_kernel(StepConsts C, const uint32_t* __restrict__ skeys,
                                                          uint32_t* __restrict__ cs) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    int lo = 1, hi = 0;
    if (j <= C.n) {
        const int prev = j == 0 ? -1 : key_bucket(skeys[j - 1], C);
        const int cur = j == C.n ? C.n_cells : key_bucket(skeys[j], C);
        lo = max(prev + 1, 0);
        hi = min(cur, C.n_cells);
    }
    int len = hi - lo + 1;
    if (len > 0 && len <= 4) {
        for (int b = lo; b <= hi; ++b) cs[b] = (uint32_t)j;
        len = 0;
    }
    while (true) {
        const uint64_t m = __ballot(len > 0);
        if (m == 0) break;
        const int src = __ffsll((unsigned long long)m) - 1;
        const int slo = __shfl(lo, src), shi = __shfl(hi, src), sj = __shfl(j, src);
        for (int b = slo + lane; b <= shi; b += 64) cs[b] = (uint32_t)sj;
        if (lane == src) len = 0;
    }
}

__global__ __launch_bounds__(256) void reorder_kernel(int n, const uint32_t* __restrict__ svals,
                                                       const float4* __restrict__ posB, const float4* __restrict__ velB,
                                                       const uint32_t* __restrict__ pidB, float4* __restrict__ posA,
                                                       float4* __restrict__ velA, uint32_t* __restrict__ pidA) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t s = svals[j];
    posA[j] = posB[s];
    velA[j] = velB[s];
    pidA[j] = pidB[s];
}

// ---------------------------------------------------------------------------------------
// Neighbour candidates of the NAIVE formulation.  For each row of cells (y, z) whose slab
// comes within h of the particle, the x cells are those overlapping [x - e, x + e] with
// e = sqrt(h^2 - dy^2 - dz^2); cells of one row are consecutive keys, so the candidates
// of a row are one contiguous range cs[first]..cs[last+1].  Boundary cells extend to
// infinity (positions outside the grid are clamped into them).  The search radius carries
// a small margin; the exact test r^2 <= h^2 is applied to each candidate.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ float slab_dist(float p, int c, int d, const StepConsts& C) {
    const float lo = c == 0 ? -INFINITY : (float)c * C.cell[d];
    const float hi = c == C.dims[d] - 1 ? INFINITY : (float)(c + 1) * C.cell[d];
    return fmaxf(fmaxf(lo - p, p - hi), 0.0f);
}

template <typename F>
__device__ __forceinline__ void for_each_candidate(float px, float py, float pz, const StepConsts& C,
                                                   const uint32_t* __restrict__ cs, F&& f) {
    const float hs = C.h_search;
    const float hs2 = hs * hs;
    const int ylo = grid_coord(py - hs, 1, C), yhi = grid_coord(py + hs, 1, C);
    const int zlo = grid_coord(pz - hs, 2, C), zhi = grid_coord(pz + hs, 2, C);
    for (int cz = zlo; cz <= zhi; ++cz) {
        const float dz = slab_dist(pz, cz, 2, C);
        for (int cy = ylo; cy <= yhi; ++cy) {
            const float dy = slab_dist(py, cy, 1, C);
            const float rem = hs2 - dz * dz - dy * dy;
            if (rem < 0.0f) continue;
            const float e = sqrtf(rem);
            const int xlo = grid_coord(px - e, 0, C), xhi = grid_coord(px + e, 0, C);
            const int row = C.cell_off + C.dims[0] * (cy + C.dims[1] * cz);
            const uint32_t k0 = cs[row + xlo], k1 = cs[row + xhi + 1];
            for (uint32_t k = k0; k < k1; ++k) f(k);
        }
    }
}

// Writes pressure and 1/rho into the .w lanes of posA / velA (4-byte stores; the .w lanes
// are not read by this kernel, other threads read only x, y, z of their neighbours).
__device__ __forceinline__ void store_density(const StepConsts& C, int i, float acc, float4* pos, float4* vel,
                                              float* __restrict__ rhoA) {
    const float rho = C.w_norm * acc;
    rhoA[i] = rho;
    reinterpret_cast<float*>(pos + i)[3] = C.k * (rho - C.rho0);
    reinterpret_cast<float*>(vel + i)[3] = 1.0f / rho;
}

// Density (naive_kernels.py:7-21) and the particle's neighbour list: every k with
// r^2 <= h^2 in candidate order, the particle itself included (the force pass gives the
// self pair a zero contribution).  (1 - r^2/h^2)^3 is accumulated without a branch and is
// zero outside h.  Candidates are read two at a time through 32-bit buffer offsets.
constexpr int kDensityThreads = 256;
#ifndef SPH_DENSITY_UNROLL
#define SPH_DENSITY_UNROLL 2
#endif
constexpr int kDensityUnroll = SPH_DENSITY_UNROLL;

template <bool BATCH>
__global__ __launch_bounds__(kDensityThreads) void naive_density_kernel(StepConsts G, float4* pos, float4* vel,
                                                                         const uint32_t* __restrict__ cs,
                                                                         float* __restrict__ rhoA,
                                                                         uint32_t* __restrict__ nl,
                                                                         uint32_t* __restrict__ nn) {
    __shared__ uint32_t s_list[kLdsListCap][kDensityThreads];
    const int i = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= G.n) return;
    const StepConsts& C = inst_consts<BATCH>(G, i);
    const __amdgpu_buffer_rsrc_t rpos = make_rsrc(pos, (uint32_t)G.n * 16u);
    const __amdgpu_buffer_rsrc_t rcs = make_rsrc(cs, ((uint32_t)G.cs_cells + 1u) * 4u);
    const float4 p = pos[i];
    const size_t stride = (size_t)G.n;
    float acc = 0.0f;
    uint32_t cnt = 0;
    auto visit = [&](float4 q, uint32_t off) {
        const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        const float t = fmaxf(1.0f - r2 * C.inv_h2, 0.0f);
        acc += t * t * t;
        if (r2 <= C.h2) {
            const uint32_t k = off >> 4;
            if (cnt < (uint32_t)kLdsListCap) s_list[cnt][threadIdx.x] = k;
            else if (cnt < (uint32_t)kListCap) nl[cnt * stride + i] = k;
            ++cnt;
        }
    };
    const float hs = C.h_search;
    const float hs2 = hs * hs;
    const int ylo = grid_coord(p.y - hs, 1, C), yhi = grid_coord(p.y + hs, 1, C);
    const int zlo = grid_coord(p.z - hs, 2, C), zhi = grid_coord(p.z + hs, 2, C);
    for (int cz = zlo; cz <= zhi; ++cz) {
        const float ez = slab_dist(p.z, cz, 2, C);
        for (int cy = ylo; cy <= yhi; ++cy) {
            const float ey = slab_dist(p.y, cy, 1, C);
            const float rem = hs2 - ez * ez - ey * ey;
            if (rem < 0.0f) continue;
            const float e = __builtin_amdgcn_sqrtf(rem);  // within 1 ulp; hs carries the margin
            const int xlo = grid_coord(p.x - e, 0, C), xhi = grid_coord(p.x + e, 0, C);
            const uint32_t row = (uint32_t)(C.cell_off + C.dims[0] * (cy + C.dims[1] * cz));
            uint32_t off = ld_u32(rcs, (row + (uint32_t)xlo) * 4u) * 16u;
            const uint32_t end = ld_u32(rcs, (row + (uint32_t)xhi + 1u) * 4u) * 16u;
            for (; off + 16u * (kDensityUnroll - 1) < end; off += 16u * kDensityUnroll) {
                float4 q[kDensityUnroll];
#pragma unroll
                for (int u = 0; u < kDensityUnroll; ++u) q[u] = ld_f4(rpos, off + 16u * u);
#pragma unroll
                for (int u = 0; u < kDensityUnroll; ++u) visit(q[u], off + 16u * u);
            }
            for (; off < end; off += 16u) visit(ld_f4(rpos, off), off);
        }
    }
    nn[i] = cnt;
    const uint32_t staged = cnt < (uint32_t)kLdsListCap ? cnt : (uint32_t)kLdsListCap;
    for (uint32_t t = 0; t < staged; ++t) nl[t * stride + i] = s_list[t][threadIdx.x];
    store_density(C, i, acc, pos, vel, rhoA);
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
    const float cos_a = dot3(v, edge) / (sl * el);
    const Vec3 pv{first.x - p.x, first.y - p.y, first.z - p.z};
    const Vec3 cr{edge.y * pv.z - edge.z * pv.y, edge.z * pv.x - edge.x * pv.z, edge.x * pv.y - edge.y * pv.x};
    const float dist = len3(cr) / el;
    const float sin_a = sqrtf(1.0f - cos_a * cos_a);
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

// Symplectic Euler (base_kernels.py:30-53), collision, output in the same sorted slot of
// posB/velB, hash of the new position for the next step.
template <int MODE>
__device__ __forceinline__ void finish_particle(const StepConsts& C, int i, float4 p, float4 v, uint32_t pid,
                                                float rho, Vec3 P, Vec3 V, float4* __restrict__ posOut,
                                                float4* __restrict__ velOut, uint32_t* __restrict__ pidOut,
                                                uint32_t* __restrict__ keyI, uint32_t* __restrict__ valI,
                                                uint64_t* __restrict__ rng, const PipeTable* __restrict__ pipe,
                                                float4* __restrict__ dbgP, float4* __restrict__ dbgV,
                                                float4* __restrict__ dbgF, uint32_t* __restrict__ flags) {
    const float fx = C.ext[0] + -P.x + V.x;
    const float fy = C.ext[1] + -P.y + V.y;
    const float fz = C.ext[2] + -P.z + V.z;
    Vec3 vv{v.x + fx / rho * C.dt, v.y + fy / rho * C.dt, v.z + fz / rho * C.dt};
    Vec3 pp{p.x + vv.x * C.dt, p.y + vv.y * C.dt, p.z + vv.z * C.dt};
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
    hash_particle(C, pp.x, pp.y, pp.z, pid, (uint32_t)i, keyI, valI, flags);
}

// Pressure and viscosity sums of one pair (naive_kernels.py:38-52, :70-77) without the
// constant factors.  q = (x, y, z, p) and vj = (vx, vy, vz, 1/rho) of the neighbour.
struct PairSums {
    float ax = 0.f, ay = 0.f, az = 0.f, bx = 0.f, by = 0.f, bz = 0.f;
    // self = the pair is the particle itself: 1/r is taken as 0, which makes both terms zero
    // (the pressure term carries a factor 1/r, the viscosity term a factor v_j - v_i = 0)
    __device__ __forceinline__ void add(const StepConsts& C, float4 p, float4 v, float4 q, float4 vj,
                                        bool self = false) {
        const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        const float rinv = self ? 0.0f : rsqrtf(r2);
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

// Force pass (naive_kernels.py:24-80) fused with integration and collision: each lane
// walks its particle's neighbour list (one coalesced load per slot, kForceUnroll slots and
// their gathers in flight); a particle with more neighbours than list slots rescans its
// candidates instead.  MODE selects the wall collision at compile time.
#ifndef SPH_FORCE_UNROLL
#define SPH_FORCE_UNROLL 4
#endif
constexpr int kForceUnroll = SPH_FORCE_UNROLL;

template <int MODE, bool BATCH>
__global__ __launch_bounds__(256) void naive_force_kernel(
    StepConsts G, const float4* __restrict__ pos, const float4* __restrict__ vel, const uint32_t* __restrict__ pidA,
    const float* __restrict__ rhoA, const uint32_t* __restrict__ cs, const uint32_t* __restrict__ nl,
    const uint32_t* __restrict__ nn, float4* __restrict__ posOut, float4* __restrict__ velOut,
    uint32_t* __restrict__ pidOut, uint32_t* __restrict__ keyI, uint32_t* __restrict__ valI,
    uint64_t* __restrict__ rng, const PipeTable* __restrict__ pipe, float4* __restrict__ dbgP,
    float4* __restrict__ dbgV, float4* __restrict__ dbgF, uint32_t* __restrict__ flags) {
    const int i = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= G.n) return;
    const StepConsts& C = inst_consts<BATCH>(G, i);
    const __amdgpu_buffer_rsrc_t rpos = make_rsrc(pos, (uint32_t)G.n * 16u);
    const __amdgpu_buffer_rsrc_t rvel = make_rsrc(vel, (uint32_t)G.n * 16u);
    const float4 p = pos[i], v = vel[i];
    PairSums S;
    const uint32_t cnt = nn[i];
    if (cnt <= (uint32_t)kListCap) {
        const uint32_t* slot = nl + i;
        const size_t stride = (size_t)G.n;
        uint32_t t = 0;
        for (; t + kForceUnroll <= cnt; t += kForceUnroll) {
            uint32_t k[kForceUnroll];
            float4 q[kForceUnroll], w[kForceUnroll];
#pragma unroll
            for (int u = 0; u < kForceUnroll; ++u) k[u] = slot[u * stride];
            slot += kForceUnroll * stride;
#pragma unroll
            for (int u = 0; u < kForceUnroll; ++u) {
                q[u] = ld_f4(rpos, k[u] * 16u);
                w[u] = ld_f4(rvel, k[u] * 16u);
            }
#pragma unroll
            for (int u = 0; u < kForceUnroll; ++u) S.add(C, p, v, q[u], w[u], k[u] == (uint32_t)i);
        }
        for (; t < cnt; ++t) {
            const uint32_t k = slot[0];
            slot += stride;
            S.add(C, p, v, ld_f4(rpos, k * 16u), ld_f4(rvel, k * 16u), k == (uint32_t)i);
        }
    } else {
        for_each_candidate(p.x, p.y, p.z, C, cs, [&](uint32_t k) {
            const float4 q = pos[k];
            const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
            if (dx * dx + dy * dy + dz * dz <= C.h2 && k != (uint32_t)i) S.add(C, p, v, q, vel[k]);
        });
    }
    const float ps = 0.5f * C.mass * C.grad_c;   // naive_kernels.py:40-52: m * (pi+pj)/(2 rho_j) * grad W
    const float vs = C.mass * C.visc * C.lap_c;  // naive_kernels.py:71-77: m * mu * (vj-vi)/rho_j * lap W
    finish_particle<MODE>(C, i, p, v, pidA[i], rhoA[i], Vec3{ps * S.ax, ps * S.ay, ps * S.az},
                          Vec3{vs * S.bx, vs * S.by, vs * S.bz}, posOut, velOut, pidOut, keyI, valI, rng, pipe, dbgP,
                          dbgV, dbgF, flags);
}

// ---------------------------------------------------------------------------------------
// VOXEL formulation (voxel_kernels.py).  get_neighbours (:29-85): 27 cells in (dx, dy, dz)
// order with dz innermost, bounds from trunc(space/voxel), each cell's range runs to the
// next non-empty cell (or to N when none follows), accept sqrt(r^2) <= h, self included,
// stop at 32.  The list is built once per step and reused by the force pass.
// ---------------------------------------------------------------------------------------
constexpr int kNbrStride = 1 + kMaxNeighbours;

__global__ __launch_bounds__(256) void voxel_density_kernel(StepConsts C, float4* pos, float4* vel,
                                                             const uint32_t* __restrict__ cs,
                                                             float* __restrict__ rhoA, int32_t* __restrict__ nbr) {
    const int i = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= C.n) return;
    const float4 p = pos[i];
    const int vx = ref_coord(p.x, 0, C), vy = ref_coord(p.y, 1, C), vz = ref_coord(p.z, 2, C);
    int32_t* list = nbr + (int64_t)i * kNbrStride + 1;
    int cnt = 0;
    const uint32_t cs_end = cs[C.n_cells];
    for (int dx = -1; dx < 2 && cnt < kMaxNeighbours; ++dx)
        for (int dy = -1; dy < 2 && cnt < kMaxNeighbours; ++dy)
            for (int dz = -1; dz < 2 && cnt < kMaxNeighbours; ++dz) {
                const int nx = vx + dx, ny = vy + dy, nz = vz + dz;
                if (nx < 0 || nx >= C.tdims[0] || ny < 0 || ny >= C.tdims[1] || nz < 0 || nz >= C.tdims[2]) continue;
                const int ci = nx + ny * C.tdims[0] + nz * C.tdims[0] * C.tdims[1];
                const uint32_t start = cs[ci], next = cs[ci + 1];
                if (start == next) continue;  // voxel_begin == -1
                const uint32_t end = next < cs_end ? next : (uint32_t)C.n;
                for (uint32_t e = start; e < end && cnt < kMaxNeighbours; ++e) {
                    const float4 q = pos[e];
                    const float ddx = p.x - q.x, ddy = p.y - q.y, ddz = p.z - q.z;
                    if (sqrtf(ddx * ddx + ddy * ddy + ddz * ddz) <= C.h) list[cnt++] = (int32_t)e;
                }
            }
    nbr[(int64_t)i * kNbrStride] = cnt;
    float acc = 0.0f;
    for (int a = 0; a < cnt; ++a) {
        const int j = list[a];
        if (j == i) continue;
        const float4 q = pos[j];
        const float ddx = p.x - q.x, ddy = p.y - q.y, ddz = p.z - q.z;
        const float t = 1.0f - (ddx * ddx + ddy * ddy + ddz * ddz) * C.inv_h2;
        acc += t * t * t;
    }
    store_density(C, i, acc, pos, vel, rhoA);
}

template <int MODE>
__global__ __launch_bounds__(256) void voxel_force_kernel(
    StepConsts C, const float4* __restrict__ pos, const float4* __restrict__ vel, const uint32_t* __restrict__ pidA,
    const float* __restrict__ rhoA, const int32_t* __restrict__ nbr, float4* __restrict__ posOut,
    float4* __restrict__ velOut, uint32_t* __restrict__ pidOut, uint32_t* __restrict__ keyI,
    uint32_t* __restrict__ valI, uint64_t* __restrict__ rng,
    const PipeTable* __restrict__ pipe, float4* __restrict__ dbgP, float4* __restrict__ dbgV,
    float4* __restrict__ dbgF, uint32_t* __restrict__ flags) {
    const int i = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= C.n) return;
    const float4 p = pos[i], v = vel[i];
    const int32_t* list = nbr + (int64_t)i * kNbrStride;
    const int cnt = list[0];
    const float rho_i = rhoA[i];
    const float fi = p.w / (rho_i * rho_i);
    float ax = 0.f, ay = 0.f, az = 0.f, bx = 0.f, by = 0.f, bz = 0.f;
    for (int t = 0; t < cnt; ++t) {
        const int j = list[1 + t];
        if (j == i) continue;
        const float4 q = pos[j], vj = vel[j];
        const float rho_j = rhoA[j];
        const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
        const float r = sqrtf(dx * dx + dy * dy + dz * dz);
        const float hr = C.h - r;
        // voxel_kernels.py:162-168: (p_i/rho_i^2 + p_j/rho_j^2) * grad W
        const float cp = (fi + q.w / (rho_j * rho_j)) * C.grad_c * hr * hr / r;
        ax += cp * dx;
        ay += cp * dy;
        az += cp * dz;
        // voxel_kernels.py:203-208: (v_j - v_i) / rho_j * lap W * m * mu / rho_i
        const float lap = C.lap_c * hr;
        const float s = C.mass * C.visc / rho_i;
        bx += (vj.x - v.x) / rho_j * lap * s;
        by += (vj.y - v.y) / rho_j * lap * s;
        bz += (vj.z - v.z) / rho_j * lap * s;
    }
    // voxel_kernels.py:211 min(1, .)
    const Vec3 V{fminf(1.0f, bx), fminf(1.0f, by), fminf(1.0f, bz)};
    finish_particle<MODE>(C, i, p, v, pidA[i], rho_i, Vec3{ax, ay, az}, V, posOut, velOut, pidOut, keyI, valI, rng, pipe,
                    dbgP, dbgV, dbgF, flags);
}

// ---------------------------------------------------------------------------------------
// Boundary output: scatter from cell order back to particle order.
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void download_f64_kernel(int n, const float4* __restrict__ posB,
                                                            const float4* __restrict__ velB,
                                                            const uint32_t* __restrict__ pidB,
                                                            const float4* __restrict__ dbgF, double* __restrict__ pos,
                                                            double* __restrict__ vel, double* __restrict__ rho,
                                                            double* __restrict__ force) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = posB[i], v = velB[i];
    const int64_t id = pidB[i];
    if (pos) { pos[3 * id] = p.x; pos[3 * id + 1] = p.y; pos[3 * id + 2] = p.z; }
    if (vel) { vel[3 * id] = v.x; vel[3 * id + 1] = v.y; vel[3 * id + 2] = v.z; }
    if (rho) rho[id] = p.w;
    if (force) { const float4 f = dbgF[i]; force[3 * id] = f.x; force[3 * id + 1] = f.y; force[3 * id + 2] = f.z; }
}

__global__ __launch_bounds__(256) void download_f32_kernel(int n, const float4* __restrict__ posB,
                                                            const float4* __restrict__ velB,
                                                            const uint32_t* __restrict__ pidB, float* x, float* y,
                                                            float* z, float* vx, float* vy, float* vz, float* rho) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = posB[i], v = velB[i];
    const uint32_t id = pidB[i];
    if (x) x[id] = p.x;
    if (y) y[id] = p.y;
    if (z) z[id] = p.z;
    if (vx) vx[id] = v.x;
    if (vy) vy[id] = v.y;
    if (vz) vz[id] = v.z;
    if (rho) rho[id] = p.w;
}

// which: 1 keys, 2 perm, 3 sorted keys, 4 cell begin, 5 density, 6-8 terms, 10 neighbours
__global__ __launch_bounds__(256) void read_buffer_kernel(StepConsts C, int which, int perm_from_A,
                                                           const uint32_t* __restrict__ keyI,
                                                           const uint32_t* __restrict__ skeys,
                                                           const uint32_t* __restrict__ svals,
                                                           const uint32_t* __restrict__ cs,
                                                           const uint32_t* __restrict__ pidA,
                                                           const uint32_t* __restrict__ pidB,
                                                           const float* __restrict__ rhoA, const float4* __restrict__ dbg,
                                                           const int32_t* __restrict__ nbr, void* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (which == SPH_BUF_CELL_BEGIN) {
        if (i >= C.n_cells) return;
        const uint32_t s = cs[i], e = cs[i + 1];
        ((int32_t*)out)[i] = s < e ? (int32_t)s : -1;
        return;
    }
    if (i >= C.n) return;
    switch (which) {
        case SPH_BUF_KEYS: {
            const uint32_t u = keyI[i];
            ((int32_t*)out)[i] = C.ref_keys ? (int32_t)(u ^ 0x80000000u) : (int32_t)u;
            break;
        }
        case SPH_BUF_PERM: {
            ((int32_t*)out)[i] = (int32_t)(perm_from_A ? pidA[i] : pidB[svals[i]]);
            break;
        }
        case SPH_BUF_SORTED_KEYS: {
            const uint32_t u = skeys[i];
            ((int32_t*)out)[i] = C.ref_keys ? (int32_t)(u ^ 0x80000000u) : (int32_t)u;
            break;
        }
        case SPH_BUF_DENSITY:
            ((float*)out)[pidA[i]] = rhoA[i];
            break;
        case SPH_BUF_PRESSURE:
        case SPH_BUF_VISCOSITY:
        case SPH_BUF_FORCE: {
            const int64_t id = pidA[i];
            const float4 t = dbg[i];
            ((float*)out)[3 * id] = t.x;
            ((float*)out)[3 * id + 1] = t.y;
            ((float*)out)[3 * id + 2] = t.z;
            break;
        }
        case SPH_BUF_NEIGHBOURS: {
            const int64_t id = pidA[i];
            const int32_t* src = nbr + i * kNbrStride;
            int32_t* dst = (int32_t*)out + id * kNbrStride;
            const int cnt = src[0];
            dst[0] = cnt;
            for (int t = 0; t < kMaxNeighbours; ++t) dst[1 + t] = t < cnt ? (int32_t)pidA[src[1 + t]] : -1;
            break;
        }
        default:
            break;
    }
}

// ---------------------------------------------------------------------------------------
// FP64 diagnostics: per-block partial sums, then one block folds them in a fixed order,
// so the result is bitwise reproducible.
// partial layout: [sum vx, sum vy, sum vz, sum |v|^2, non-finite count, max |v|, sum rho, n]
// ---------------------------------------------------------------------------------------
constexpr int kDiagThreads = 256;
constexpr int kDiagVals = 8;

__global__ __launch_bounds__(kDiagThreads) void diag_partial_kernel(int n, const float4* __restrict__ posB,
                                                                     const float4* __restrict__ velB,
                                                                     double* __restrict__ partials) {
    __shared__ double red[kDiagVals][kDiagThreads];
    double acc[kDiagVals] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t i = (int64_t)blockIdx.x * kDiagThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kDiagThreads) {
        const float4 p = posB[i], v = velB[i];
        const double x = v.x, y = v.y, z = v.z;
        const double s2 = x * x + y * y + z * z;
        if (!(finite3(p.x, p.y, p.z) && finite3(v.x, v.y, v.z))) {
            acc[4] += 1.0;
            continue;
        }
        acc[0] += x;
        acc[1] += y;
        acc[2] += z;
        acc[3] += s2;
        acc[5] = fmax(acc[5], sqrt(s2));
        acc[6] += (double)p.w;
        acc[7] += 1.0;
    }
    for (int k = 0; k < kDiagVals; ++k) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int s = kDiagThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < kDiagVals; ++k)
                red[k][threadIdx.x] = k == 5 ? fmax(red[k][threadIdx.x], red[k][threadIdx.x + s])
                                             : red[k][threadIdx.x] + red[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < kDiagVals) partials[blockIdx.x * kDiagVals + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void diag_final_kernel(int nblocks, double mass, int n, const double* __restrict__ partials,
                                  double* __restrict__ out) {
    if (threadIdx.x != 0) return;
    double acc[kDiagVals] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = 0; b < nblocks; ++b)
        for (int k = 0; k < kDiagVals; ++k)
            acc[k] = k == 5 ? fmax(acc[k], partials[b * kDiagVals + k]) : acc[k] + partials[b * kDiagVals + k];
    out[0] = mass * n;
    out[1] = mass * acc[0];
    out[2] = mass * acc[1];
    out[3] = mass * acc[2];
    out[4] = 0.5 * mass * acc[3];
    out[5] = acc[4];
    out[6] = acc[5];
    out[7] = acc[7] > 0 ? acc[6] / acc[7] : 0.0;
}

inline int blocks_for(int64_t n, int t) { return (int)((n + t - 1) / t); }

}  // namespace

constexpr int kDiagBlocks = 512;

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s) {
    if (n > 0) fill_u32_kernel<<<blocks_for(n, 256), 256, 0, s>>>(p, v, n);
    return hipGetLastError();
}

hipError_t launch_upload_f64(sph_ctx* c, const double* pos, const double* vel, hipStream_t s) {
    if (c->n > 0)
        (c->sc.n_inst > 1 ? upload_f64_kernel<true> : upload_f64_kernel<false>)<<<blocks_for(c->n, 256), 256, 0, s>>>(c->sc, pos, vel, c->b.posB, c->b.velB, c->b.pidB,
                                                                  c->b.keyI, c->b.valI, c->b.flags);
    return hipGetLastError();
}

hipError_t launch_upload_f32(sph_ctx* c, const float* x, const float* y, const float* z, const float* vx,
                             const float* vy, const float* vz, hipStream_t s) {
    if (c->n > 0)
        (c->sc.n_inst > 1 ? upload_f32_kernel<true> : upload_f32_kernel<false>)<<<blocks_for(c->n, 256), 256, 0, s>>>(c->sc, x, y, z, vx, vy, vz, c->b.posB, c->b.velB,
                                                                  c->b.pidB, c->b.keyI, c->b.valI, c->b.flags);
    return hipGetLastError();
}

hipError_t launch_cell_table(sph_ctx* c, hipStream_t s) {
    cell_table_kernel<<<blocks_for(c->n + 1, 256), 256, 0, s>>>(c->sc, c->sorted_keys, c->b.cs);
    return hipGetLastError();
}

hipError_t launch_reorder(sph_ctx* c, hipStream_t s) {
    if (c->n > 0)
        reorder_kernel<<<blocks_for(c->n, 256), 256, 0, s>>>((int)c->n, c->sorted_vals, c->b.posB, c->b.velB,
                                                               c->b.pidB, c->b.posA, c->b.velA, c->b.pidA);
    return hipGetLastError();
}

hipError_t launch_density(sph_ctx* c, hipStream_t s) {
    if (c->n == 0) return hipGetLastError();
    const int bs = c->density_block;
    if (c->sc.formulation == SPH_FORM_VOXEL)
        voxel_density_kernel<<<blocks_for(c->n, bs), bs, 0, s>>>(c->sc, c->b.posA, c->b.velA, c->b.cs, c->b.rhoA,
                                                                   c->b.nbr);
    else
        (c->sc.n_inst > 1 ? naive_density_kernel<true> : naive_density_kernel<false>)<<<blocks_for(c->n, bs), bs, 0, s>>>(c->sc, c->b.posA, c->b.velA, c->b.cs, c->b.rhoA,
                                                                   c->b.nl, c->b.nn);
    return hipGetLastError();
}

template <int MODE>
void launch_force_mode(sph_ctx* c, hipStream_t s) {
    const int bs = c->force_block;
    float4* dP = c->keep_terms ? c->b.dbgP : nullptr;
    float4* dV = c->keep_terms ? c->b.dbgV : nullptr;
    float4* dF = c->keep_terms ? c->b.dbgF : nullptr;
    if (c->sc.formulation == SPH_FORM_VOXEL)
        voxel_force_kernel<MODE><<<blocks_for(c->n, bs), bs, 0, s>>>(
            c->sc, c->b.posA, c->b.velA, c->b.pidA, c->b.rhoA, c->b.nbr, c->b.posB, c->b.velB, c->b.pidB, c->b.keyI,
            c->b.valI, c->b.rng, c->b.pipe, dP, dV, dF, c->b.flags);
    else
        (c->sc.n_inst > 1 ? naive_force_kernel<MODE, true> : naive_force_kernel<MODE, false>)<<<blocks_for(c->n, bs), bs, 0, s>>>(
            c->sc, c->b.posA, c->b.velA, c->b.pidA, c->b.rhoA, c->b.cs, c->b.nl, c->b.nn, c->b.posB, c->b.velB,
            c->b.pidB, c->b.keyI, c->b.valI, c->b.rng, c->b.pipe, dP, dV, dF, c->b.flags);
}

hipError_t launch_force(sph_ctx* c, hipStream_t s) {
    if (c->n == 0) return hipGetLastError();
    if (c->sc.mode == SPH_MODE_PIPE) launch_force_mode<SPH_MODE_PIPE>(c, s);
    else launch_force_mode<SPH_MODE_BOX>(c, s);
    return hipGetLastError();
}

hipError_t launch_download_f64(sph_ctx* c, double* pos, double* vel, double* rho, double* force, hipStream_t s) {
    if (c->n > 0)
        download_f64_kernel<<<blocks_for(c->n, 256), 256, 0, s>>>((int)c->n, c->b.posB, c->b.velB, c->b.pidB,
                                                                    c->b.dbgF, pos, vel, rho, force);
    return hipGetLastError();
}

hipError_t launch_download_f32(sph_ctx* c, float* x, float* y, float* z, float* vx, float* vy, float* vz, float* rho,
                               hipStream_t s) {
    if (c->n > 0)
        download_f32_kernel<<<blocks_for(c->n, 256), 256, 0, s>>>((int)c->n, c->b.posB, c->b.velB, c->b.pidB, x, y, z,
                                                                    vx, vy, vz, rho);
    return hipGetLastError();
}

hipError_t launch_read_buffer(sph_ctx* c, int which, void* out, hipStream_t s) {
    const float4* dbg = which == SPH_BUF_PRESSURE ? c->b.dbgP : which == SPH_BUF_VISCOSITY ? c->b.dbgV : c->b.dbgF;
    const int64_t count = which == SPH_BUF_CELL_BEGIN ? c->sc.n_cells : c->n;
    const int perm_from_A = c->last_stage >= SPH_STAGE_REORDER ? 1 : 0;
    if (count > 0)
        read_buffer_kernel<<<blocks_for(count, 256), 256, 0, s>>>(c->sc, which, perm_from_A, c->b.keyI, c->sorted_keys,
                                                                    c->sorted_vals, c->b.cs, c->b.pidA, c->b.pidB,
                                                                    c->b.rhoA, dbg, c->b.nbr, out);
    return hipGetLastError();
}

hipError_t launch_diagnostics(sph_ctx* c, double* out8, hipStream_t s) {
    diag_partial_kernel<<<kDiagBlocks, kDiagThreads, 0, s>>>((int)c->n, c->b.posB, c->b.velB, c->b.diag_partials);
    diag_final_kernel<<<1, 64, 0, s>>>(kDiagBlocks, c->cst.mass, (int)c->n, c->b.diag_partials, out8);
    return hipGetLastError();
}

}  // namespace sph
