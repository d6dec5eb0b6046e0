// sph_kernels.hip -- gfx950 kernels of one SPH step (everything but the sorts).
//
// Particles in cell order; one lane (or a group of L lanes) per particle.  A step is
//   sort         NAIVE: counting sort by cell (sph_cellsort.hip), which also builds the cell
//                table; VOXEL: radix sort (sph_radix.hip) then cell_table.  The sort's last
//                pass gathers posB/velB -> posA/velA in cell order
//   cell_table   cs[c] = first sorted index with key >= c (dense, no empty markers)
//   density      rho, p = K (rho - rho0) into posA.w and 1/rho into velA.w, neighbour lists
//   force        pressure + viscosity sums, symplectic Euler, wall collision, then the
//                voxel hash of the new position for the next step's sort
// The NAIVE formulation (the reference default) sums over every j with |x_i - x_j|^2 <= h^2;
// the grid only skips pairs that are provably farther apart.  The VOXEL formulation
// reproduces voxel_kernels.py including its 32-neighbour cap and enumeration order.
#include <math.h>

#include <type_traits>

#include "sph_device.h"

namespace sph {
namespace {

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
    hash_particle<false>(C, x, y, z, (uint32_t)i, (uint32_t)i, keyI, valI, flags);  // slot = id: both orders
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
    hash_particle<false>(C, x, y, z, (uint32_t)i, (uint32_t)i, keyI, valI, flags);  // slot = id: both orders
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

// Gather of the state into cell order (the staged step's own pass; the fused step gathers in
// the sort's last pass).  VOXEL: reports a position outside the voxel grid in
// flags[kFlagOutOfGrid], which the sort cleared.
__global__ __launch_bounds__(256) void reorder_kernel(StepConsts C, const uint32_t* __restrict__ svals,
                                                       const float4* __restrict__ posB, const float4* __restrict__ velB,
                                                       const uint32_t* __restrict__ pidB, float4* __restrict__ posA,
                                                       float4* __restrict__ velA, uint32_t* __restrict__ pidA,
                                                       uint32_t* __restrict__ flags) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= C.n) return;
    const uint32_t s = svals[j];
    const float4 p = posB[s];
    posA[j] = p;
    velA[j] = velB[s];
    pidA[j] = pidB[s];
    if (C.formulation == SPH_FORM_VOXEL && outside_voxel_grid(p.x, p.y, p.z, C)) flags[kFlagOutOfGrid] = 1u;
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

// Keys of the resident state (NAIVE: indexed by posB slot) under the current cell layout,
// after sph_set_slabs changed it; the non-finite count was taken when the state was hashed.
__global__ __launch_bounds__(256) void rekey_kernel(StepConsts C, const float4* __restrict__ posB,
                                                     uint32_t* __restrict__ keyI) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C.n) return;
    const float4 p = posB[i];
    keyI[i] = sort_key(p.x, p.y, p.z, C);
}

// Density (naive_kernels.py:7-21) and the particle's neighbour list: every k with
// r^2 <= h^2 in candidate order, the particle itself included (the force pass gives the
// self pair a zero contribution).  (1 - r^2/h^2)^3 is accumulated for exactly the listed
// pairs (naive_kernels.py:18 skips a pair with norm_squared > h^2).
constexpr int kDensityThreads = 256;
#ifndef SPH_DENSITY_UNROLL
#define SPH_DENSITY_UNROLL 2
#endif
constexpr int kDensityUnroll = SPH_DENSITY_UNROLL;

// LDS staging of the list slots, as 16-bit entries (the list's own format).  Each wavefront
// owns a band of kBandSlots entries; with L lanes per particle the band holds 64 / L
// particles with 32 L slots each, so a particle with 4 or more lanes stages every one of its
// kListCap slots (with fewer lanes, slots past 32 L go straight to HBM: one scattered 16-bit
// store per hit).  Slots 2w and 2w + 1 of particle p of a band are the halves of 32-bit word
// w (64 / L) + p, so no two lanes of a store share a word (nor its bank), and the copy-out
// reads each list word with one load.
constexpr int kBandSlots = 64 * 2 * kLdsListCap;  // 2,048 entries (4 KiB) per wavefront
typedef ListEntry DensityList[kBandSlots];        // one band
template <int L>
constexpr uint32_t stage_cap() { return (uint32_t)(2 * kLdsListCap * L < kListCap ? 2 * kLdsListCap * L : kListCap); }

template <int L>
struct StageBand {
    ListEntry* base;  // the wavefront's band
    uint32_t p;       // the particle's place in it, 0 .. 64 / L - 1
    __device__ __forceinline__ ListEntry& at(uint32_t t) const { return base[((t >> 1) * (64 / L) + p) * 2 + (t & 1)]; }
    // slots t and t + 1 (t even) as one list word; the high half is zero past `staged` slots
    __device__ __forceinline__ uint32_t word(uint32_t t, uint32_t staged) const {
        uint32_t w;
        __builtin_memcpy(&w, &at(t), sizeof(w));
        return t + 1 < staged ? w : (w & 0xffffu);
    }
};

// Density of cell-ordered particle i with L lanes per particle.  L = 1: the lane walks every
// candidate row, two candidates in flight through 32-bit buffer offsets.  L = 2..16: a group
// of L consecutive lanes of one wavefront shares the particle, walks the same rows and takes
// each row's candidates L at a time (lane l tests candidate k0 + l + L m), so a particle with
// many candidates is spread over L lanes; hits are appended in candidate order (a group
// ballot ranks them), so the neighbour list is the one a single lane builds, and the density
// is summed per lane, then across the group.  G holds the totals (particle count, cell
// table length), C the particle's own instance; band is the particle's place in its
// wavefront's LDS staging band
// (unique in the workgroup), gbase its group's first lane in the wavefront.
template <int L>
__device__ __forceinline__ void density_particle(const StepConsts& G, const StepConsts& C, int i, uint32_t l,
                                                 StageBand<L> band, uint32_t gbase, float4* pos, float4* vel,
                                                 const uint32_t* __restrict__ cs, float* __restrict__ rhoA,
                                                 ListEntry* __restrict__ nl, uint32_t* __restrict__ nn) {
    constexpr uint32_t kCap = stage_cap<L>();
    const __amdgpu_buffer_rsrc_t rpos = make_rsrc(pos, (uint32_t)G.n * 16u);
    const __amdgpu_buffer_rsrc_t rcs = make_rsrc(cs, ((uint32_t)G.cs_cells + 1u) * 4u);
    const float4 p = pos[i];
    const size_t stride = (size_t)G.n;
    float acc = 0.0f;
    uint32_t cnt = 0;
    // first and one past the last candidate: every list entry fits an int16 offset when the
    // candidates do (checked once per particle, not per hit)
    uint32_t klo = 0xffffffffu, khi = 0u;
    // L = 1 appends its hits in slot order: the next staged slot is reached from the last by a
    // step that alternates between 1 (the other half of the word) and the band's row pitch
    // less 1, in place of working out its address from the count on every hit
    ListEntry* slot = &band.at(0);
    uint32_t slot_step = 1;
    auto stage = [&](ListEntry e) {
        *slot = e;
        slot += slot_step;
        slot_step = 2u * (64u / L) - slot_step;
    };
    // one candidate: a hit (r^2 <= h^2) adds (1 - r^2/h^2)^3 to the sum and is appended to the
    // list; CHECKED: the list may be full (staged slots, then HBM slots, then only counted)
    auto visit = [&](float4 q, uint32_t k, auto checked) {
        const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (r2 <= C.h2) {
            const float t = 1.0f - r2 * C.inv_h2;
            acc += t * t * t;
            if constexpr (decltype(checked)::value) {
                if (cnt < kCap) stage(list_encode(k, i));
                else if (cnt < (uint32_t)kListCap) nl[list_at(G.n, i, cnt)] = list_encode(k, i);
            } else {
                stage(list_encode(k, i));
            }
            ++cnt;
        }
    };
    // candidates k0 .. k1 - 1 of one row
    auto run_row = [&](uint32_t k0, uint32_t k1) {
        klo = min(klo, k0);
        khi = max(khi, k1);
        if constexpr (L == 1) {
            auto walk = [&](auto checked) {
                uint32_t off = k0 * 16u;
                const uint32_t end = k1 * 16u;
                for (; off + 16u * (kDensityUnroll - 1) < end; off += 16u * kDensityUnroll) {
                    float4 q[kDensityUnroll];
#pragma unroll
                    for (int u = 0; u < kDensityUnroll; ++u) q[u] = ld_f4(rpos, off + 16u * u);
#pragma unroll
                    for (int u = 0; u < kDensityUnroll; ++u) visit(q[u], (off >> 4) + (uint32_t)u, checked);
                }
                for (; off < end; off += 16u) visit(ld_f4(rpos, off), off >> 4, checked);
            };
            // while every active lane's list has room in LDS for all of the row's candidates, the
            // row's hits need no capacity checks (a wave-uniform choice)
            if (__ballot(cnt + (k1 - k0) > kCap) == 0) walk(std::false_type{});
            else walk(std::true_type{});
        } else {
            for (uint32_t base = k0; base < k1; base += L) {
                const uint32_t k = base + l;
                const bool valid = k < k1;
                const float4 q = ld_f4(rpos, (valid ? k : k0) * 16u);
                const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
                const float r2 = valid ? dx * dx + dy * dy + dz * dz : INFINITY;
                const bool hit = r2 <= C.h2;
                if (hit) {
                    const float t = 1.0f - r2 * C.inv_h2;
                    acc += t * t * t;
                }
                constexpr uint64_t group = L == 64 ? ~0ull : (1ull << L) - 1ull;
                const uint64_t gm = (__ballot(hit) >> gbase) & group;
                if (hit) {
                    const uint32_t slot = cnt + (uint32_t)__popcll(gm & ((1ull << l) - 1ull));
                    if (slot < kCap) band.at(slot) = list_encode(k, i);
                    else if (slot < (uint32_t)kListCap) nl[list_at(G.n, i, slot)] = list_encode(k, i);
                }
                cnt += (uint32_t)__popcll(gm);
            }
        }
    };
    if constexpr (L == 1) {
        for_each_row(p.x, p.y, p.z, C, [&](int c) { return ld_u32(rcs, (uint32_t)c * 4u); }, run_row);
    } else {
        // the group shares the row walk; rows without candidates still bound klo, khi
        for_each_row_range<L>(p.x, p.y, p.z, C, rcs, l, run_row);
    }
#pragma unroll
    for (int o = 1; o < L; o <<= 1) acc += __shfl_xor(acc, o);
    // the lanes of a group walk the same rows, so they agree on this
    const bool wide = klo < khi && (klo + 32768u < (uint32_t)i || khi > (uint32_t)i + 32768u);
    const uint32_t staged = cnt < kCap ? cnt : kCap;
    // the staged slots as 32-bit words (two slots each), the group's lanes taking turns
    uint32_t* words = reinterpret_cast<uint32_t*>(nl);
    for (uint32_t t = 2 * l; t < staged; t += 2 * L) {
        words[(size_t)(t >> 1) * stride + i] = band.word(t, staged);
    }
    if (l == 0) {
        // a particle with a far neighbour is rescanned by the force pass, as one past the cap
        nn[i] = wide && cnt <= (uint32_t)kListCap ? (uint32_t)kListCap + 1u : cnt;
        store_density(C, i, acc, pos, vel, rhoA);
    }
}

// Plain step: L lanes per particle (1, 2, 4, 8, 16, 32 or 64), particle gt / L.  Groups are aligned,
// so a whole group leaves together at the end of the range.
template <int L>
__global__ __launch_bounds__(kDensityThreads) void naive_density_kernel(StepConsts G, float4* pos, float4* vel,
                                                                         const uint32_t* __restrict__ cs,
                                                                         float* __restrict__ rhoA,
                                                                         ListEntry* __restrict__ nl,
                                                                         uint32_t* __restrict__ nn) {
    static_assert(L >= 1 && L <= 64 && (L & (L - 1)) == 0, "lanes per particle: 1, 2, 4, ..., 64");
    __shared__ __align__(16) DensityList s_stage[kDensityThreads / 64];
    const uint32_t gt = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int i = (int)(gt / L);
    if (i >= G.n) return;
    const uint32_t lane = threadIdx.x & 63u;
    const StageBand<L> band{s_stage[threadIdx.x >> 6], lane / L};
    density_particle<L>(G, G, i, gt & (L - 1), band, lane & ~(uint32_t)(L - 1), pos, vel, cs, rhoA, nl, nn);
}

// Batched form: each instance has its own lanes per particle and its own thread range, which
// starts at a multiple of 64, so a wavefront belongs to one instance and its lane count is
// uniform.  The instance is found once per wavefront, by a binary search over the ranges'
// first threads with the wavefront's first thread made scalar.
struct BatchLane {
    int b;       // instance
    int i;       // cell-ordered particle (total index), or -1 past the instance's particles
    uint32_t l;  // lane within the particle's group
    int lanes;
};

__device__ __forceinline__ BatchLane batch_lane(const StepConsts& G, uint32_t gt) {
    const int w0 = __builtin_amdgcn_readfirstlane((int)(gt & ~63u));
    int lo = 0, hi = G.n_inst;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (G.inst_table[mid].tfirst <= w0) lo = mid;
        else hi = mid;
    }
    const StepConsts& C = G.inst_table[lo];
    const uint32_t local = gt - (uint32_t)C.tfirst;
    const uint32_t p = local >> C.lane_shift;
    BatchLane r;
    r.b = lo;
    r.i = p < (uint32_t)C.n ? C.first + (int)p : -1;
    r.l = local & (uint32_t)(C.lanes - 1);
    r.lanes = C.lanes;
    return r;
}

__global__ __launch_bounds__(kDensityThreads) void naive_density_batch_kernel(StepConsts G, float4* pos, float4* vel,
                                                                               const uint32_t* __restrict__ cs,
                                                                               float* __restrict__ rhoA,
                                                                               ListEntry* __restrict__ nl,
                                                                               uint32_t* __restrict__ nn) {
    __shared__ __align__(16) DensityList s_stage[kDensityThreads / 64];
    const uint32_t gt = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const BatchLane t = batch_lane(G, gt);
    if (t.i < 0) return;
    const StepConsts& C = G.inst_table[t.b];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gbase = lane & ~(uint32_t)(t.lanes - 1);
    ListEntry* const base = s_stage[threadIdx.x >> 6];
    const uint32_t p = lane >> C.lane_shift;  // the particle's place in the wavefront's band
    switch (t.lanes) {
        case 1: density_particle<1>(G, C, t.i, 0u, StageBand<1>{base, p}, gbase, pos, vel, cs, rhoA, nl, nn); break;
        case 2: density_particle<2>(G, C, t.i, t.l, StageBand<2>{base, p}, gbase, pos, vel, cs, rhoA, nl, nn); break;
        case 4: density_particle<4>(G, C, t.i, t.l, StageBand<4>{base, p}, gbase, pos, vel, cs, rhoA, nl, nn); break;
        case 8: density_particle<8>(G, C, t.i, t.l, StageBand<8>{base, p}, gbase, pos, vel, cs, rhoA, nl, nn); break;
        case 16: density_particle<16>(G, C, t.i, t.l, StageBand<16>{base, p}, gbase, pos, vel, cs, rhoA, nl, nn); break;
        case 32: density_particle<32>(G, C, t.i, t.l, StageBand<32>{base, p}, gbase, pos, vel, cs, rhoA, nl, nn); break;
        default: density_particle<64>(G, C, t.i, t.l, StageBand<64>{base, p}, gbase, pos, vel, cs, rhoA, nl, nn); break;
    }
}

// Force pass (naive_kernels.py:24-80) fused with integration and collision: each lane
// walks its particle's neighbour list (one coalesced load per slot, kForceUnroll slots and
// their gathers in flight); a particle with more neighbours than list slots rescans its
// candidates instead.  With L lanes per particle the group splits the list (or the rescan)
// and lane 0 finishes the particle.  MODE selects the wall collision at compile time.
template <int MODE, int L = 1, bool COUNT = false>
__global__ __launch_bounds__(256) void naive_force_kernel(
    StepConsts G, const float4* __restrict__ pos, const float4* __restrict__ vel, const uint32_t* __restrict__ pidA,
    const float* __restrict__ rhoA, const uint32_t* __restrict__ cs, const ListEntry* __restrict__ nl,
    const uint32_t* __restrict__ nn, float4* __restrict__ posOut, float4* __restrict__ velOut,
    uint32_t* __restrict__ pidOut, uint32_t* __restrict__ keyI, uint32_t* __restrict__ valI,
    uint64_t* __restrict__ rng, const PipeTable* __restrict__ pipe, float4* __restrict__ dbgP,
    float4* __restrict__ dbgV, float4* __restrict__ dbgF, uint32_t* __restrict__ flags,
    float4* __restrict__ sums, uint32_t* __restrict__ cnt, uint32_t* __restrict__ rank,
    uint32_t* __restrict__ part) {
    const uint32_t gt = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int i = (int)(gt / L);
    if (i >= G.n) return;
    const StepConsts& C = G;
    const float4 p = pos[i], v = vel[i];
    const PairSums S = naive_pair_sums_lanes<L>(C, i, gt & (L - 1), G.n, p, v, pos, vel, cs, nl, nn);
    if (L > 1 && (gt & (L - 1)) != 0) return;
    if (sums) {  // reverse mode keeps the sums for the pullback of this step
        sums[i] = make_float4(S.ax, S.ay, S.az, 0.0f);
        sums[G.n + i] = make_float4(S.bx, S.by, S.bz, 0.0f);
    }
    const float ps = 0.5f * C.mass * C.grad_c;   // naive_kernels.py:40-52: m * (pi+pj)/(2 rho_j) * grad W
    const float vs = C.mass * C.visc * C.lap_c;  // naive_kernels.py:71-77: m * mu * (vj-vi)/rho_j * lap W
    const Vec3 P{ps * S.ax, ps * S.ay, ps * S.az}, V{vs * S.bx, vs * S.by, vs * S.bz};
    const uint32_t key = finish_particle<MODE, true>(C, i, p, v, pidA[i], rhoA[i], P, V, posOut, velOut, pidOut, keyI,
                                                     valI, rng, pipe, dbgP, dbgV, dbgF, flags);
    // the next sort's count (cell_count_kernel's work).  One lane per particle: the
    // wavefront's particles past n left at the top, so the lanes here are a prefix of it and
    // runs of equal keys share an atomic; with L lanes only the groups' first lanes are
    // here, at most 64 / L of them, and each counts its own particle.  With one lane per
    // particle also the tile sums (cell_scan_reduce's work; tile_sums_in_force), whose
    // atomics return nothing and go first, so that their latency overlaps the count's.  With
    // L lanes a wavefront holds 64 / L particles, and L times as many wavefronts would queue
    // on each tile sum (the 16k adjoint workload's backward at L = 32: 150M -> 112M
    // particle-updates/s), so the sort adds the tile sums up itself.
    if constexpr (COUNT) {
        const uint32_t cell = key & kCellKeyMask;
        if constexpr (L == 1) {
            const uint32_t tiles = (uint32_t)(G.cs_cells + (1 << kScanTileShift) - 1) >> kScanTileShift;
            add_to_tile_sums(cell, part + ((gt >> 6) & (kTileSumCopies - 1)) * tiles);
            rank[i] = rank_in_cell(cell, true, cnt);
        } else {
            rank[i] = atomicAdd(&cnt[cell], 1u);
        }
    }
}

// Batched form of the force pass (box walls): the thread map of naive_density_batch_kernel.
__global__ __launch_bounds__(256) void naive_force_batch_kernel(
    StepConsts G, const float4* __restrict__ pos, const float4* __restrict__ vel, const uint32_t* __restrict__ pidA,
    const float* __restrict__ rhoA, const uint32_t* __restrict__ cs, const ListEntry* __restrict__ nl,
    const uint32_t* __restrict__ nn, float4* __restrict__ posOut, float4* __restrict__ velOut,
    uint32_t* __restrict__ pidOut, uint32_t* __restrict__ keyI, uint32_t* __restrict__ valI,
    float4* __restrict__ dbgP, float4* __restrict__ dbgV, float4* __restrict__ dbgF, uint32_t* __restrict__ flags) {
    const uint32_t gt = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const BatchLane t = batch_lane(G, gt);
    if (t.i < 0) return;
    const StepConsts& C = G.inst_table[t.b];
    const int i = t.i;
    const float4 p = pos[i], v = vel[i];
    PairSums S;
    switch (t.lanes) {
        case 1: S = naive_pair_sums_lanes<1>(C, i, 0u, G.n, p, v, pos, vel, cs, nl, nn); break;
        case 2: S = naive_pair_sums_lanes<2>(C, i, t.l, G.n, p, v, pos, vel, cs, nl, nn); break;
        case 4: S = naive_pair_sums_lanes<4>(C, i, t.l, G.n, p, v, pos, vel, cs, nl, nn); break;
        case 8: S = naive_pair_sums_lanes<8>(C, i, t.l, G.n, p, v, pos, vel, cs, nl, nn); break;
        case 16: S = naive_pair_sums_lanes<16>(C, i, t.l, G.n, p, v, pos, vel, cs, nl, nn); break;
        case 32: S = naive_pair_sums_lanes<32>(C, i, t.l, G.n, p, v, pos, vel, cs, nl, nn); break;
        default: S = naive_pair_sums_lanes<64>(C, i, t.l, G.n, p, v, pos, vel, cs, nl, nn); break;
    }
    if (t.l != 0) return;
    const float ps = 0.5f * C.mass * C.grad_c;
    const float vs = C.mass * C.visc * C.lap_c;
    finish_particle<SPH_MODE_BOX, true>(C, i, p, v, pidA[i], rhoA[i], Vec3{ps * S.ax, ps * S.ay, ps * S.az},
                                  Vec3{vs * S.bx, vs * S.by, vs * S.bz}, posOut, velOut, pidOut, keyI, valI, nullptr,
                                  nullptr, dbgP, dbgV, dbgF, flags);
}

// ---------------------------------------------------------------------------------------
// VOXEL formulation (voxel_kernels.py).  get_neighbours (:29-85): 27 cells in (dx, dy, dz)
// order with dz innermost, bounds from trunc(space/voxel), each cell's range runs to the
// next non-empty cell (or to N when none follows), accept sqrt(r^2) <= h, self included,
// stop at 32.  The list is built once per step and reused by the force pass.  A cell whose
// nearest point lies beyond h_search holds no particle within h (h_search's margin covers
// the rounding of the distance), so skipping it leaves every list as it is; the density is
// summed while the list is built, in list order.  Lists are column-major: the count of
// cell-ordered particle i at nbr[i], its t-th neighbour at nbr[(1 + t) N + i].
// ---------------------------------------------------------------------------------------
constexpr int kNbrStride = 1 + kMaxNeighbours;

// squared distance from p to the key cells v - 1, v, v + 1 of one axis; cell 0 also holds
// the positions in (-voxel, 0), which trunc maps to it, so its slab is open below
__device__ __forceinline__ void voxel_slab2(float p, int v, float s, float e2[3]) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int c = v - 1 + d;
        const float lo = c == 0 ? -INFINITY : (float)c * s;
        const float e = fmaxf(fmaxf(lo - p, p - (float)(c + 1) * s), 0.0f);
        e2[d] = e * e;
    }
}

// The list is staged in LDS as 16-bit offsets from the particle's own index, two slots per
// 32-bit word of the lane's column (slot t of column col: half t & 1 of word (t / 2, col)),
// and copied out slot by slot at the end, so every list store is coalesced across the
// wavefront (a store per hit, straight to the column-major list, wrote each 64-byte segment
// many times over: 985 MB per launch on the 1M dam break for 134 MB of lists).  An entry
// farther than an int16 offset is written directly and its slot marked kVoxelFar.
constexpr uint16_t kVoxelFar = 0x8000u;
constexpr size_t kVoxelStageBytes = (size_t)kMaxNeighbours * sizeof(uint16_t);  // per thread

__device__ __forceinline__ uint16_t& voxel_stage_at(uint16_t* s, uint32_t t, uint32_t col) {
    return s[((t >> 1) * blockDim.x + col) * 2 + (t & 1)];
}

__global__ __launch_bounds__(256) void voxel_density_kernel(StepConsts C, float4* pos, float4* vel,
                                                             const uint32_t* __restrict__ cs,
                                                             float* __restrict__ rhoA, int32_t* __restrict__ nbr,
                                                             const uint32_t* __restrict__ flags) {
    extern __shared__ __align__(16) uint16_t s_vlist[];  // kVoxelStageBytes per thread of the block
    const int i = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= C.n) return;
    const uint32_t col = threadIdx.x;
    // a voxel beyond h may be skipped only if every particle it lists lies in its slab: the
    // lookup dims name the voxels the keys do and no position of the sorted state lies
    // outside the grid (a wrapped key lists such a particle under another voxel)
    const bool skip_ok = C.voxel_skip && __builtin_amdgcn_readfirstlane(flags[kFlagOutOfGrid]) == 0u;
    const size_t n = (size_t)C.n;
    const float4 p = pos[i];
    const int vx = ref_coord(p.x, 0, C), vy = ref_coord(p.y, 1, C), vz = ref_coord(p.z, 2, C);
    float ex[3], ey[3], ez[3];
    voxel_slab2(p.x, vx, C.cell[0], ex);
    voxel_slab2(p.y, vy, C.cell[1], ey);
    voxel_slab2(p.z, vz, C.cell[2], ez);
    const float hs2 = C.h_search * C.h_search;
    int cnt = 0;
    float acc = 0.0f;
    const uint32_t cs_end = cs[C.n_cells];
    for (int dx = -1; dx < 2 && cnt < kMaxNeighbours; ++dx)
        for (int dy = -1; dy < 2 && cnt < kMaxNeighbours; ++dy)
            for (int dz = -1; dz < 2 && cnt < kMaxNeighbours; ++dz) {
                const int nx = vx + dx, ny = vy + dy, nz = vz + dz;
                if (nx < 0 || nx >= C.tdims[0] || ny < 0 || ny >= C.tdims[1] || nz < 0 || nz >= C.tdims[2]) continue;
                if (skip_ok && ex[dx + 1] + ey[dy + 1] + ez[dz + 1] > hs2) continue;
                const int ci = nx + ny * C.tdims[0] + nz * C.tdims[0] * C.tdims[1];
                uint2 se;  // cs[ci] and cs[ci + 1] in one load
                __builtin_memcpy(&se, cs + ci, sizeof(se));
                const uint32_t start = se.x, next = se.y;
                if (start == next) continue;  // voxel_begin == -1
                const uint32_t end = next < cs_end ? next : (uint32_t)C.n;
                // four candidates loaded at a time (the stop at 32 would otherwise hold each
                // load until the previous test is done), then tested in order
                for (uint32_t e0 = start; e0 < end && cnt < kMaxNeighbours; e0 += 4) {
                    float4 q[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) q[u] = pos[min(e0 + (uint32_t)u, end - 1u)];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const uint32_t e = e0 + (uint32_t)u;
                        if (e >= end || cnt >= kMaxNeighbours) break;
                        const float ddx = p.x - q[u].x, ddy = p.y - q[u].y, ddz = p.z - q[u].z;
                        const float r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                        if (r2 <= C.r2_accept) {  // sqrtf(r2) <= h, exactly
                            const int off = (int)e - i;
                            uint16_t staged = kVoxelFar;
                            if (off >= -32767 && off <= 32767) staged = (uint16_t)(int16_t)off;
                            else nbr[(size_t)(1 + cnt) * n + i] = (int32_t)e;
                            voxel_stage_at(s_vlist, (uint32_t)cnt, col) = staged;
                            ++cnt;
                            if (e != (uint32_t)i) {  // voxel_kernels.py:127-130: rho skips the particle itself
                                const float t = 1.0f - r2 * C.inv_h2;
                                acc += t * t * t;
                            }
                        }
                    }
                }
            }
    nbr[i] = cnt;
    for (int t = 0; t < cnt; t += 2) {
        uint32_t w;
        __builtin_memcpy(&w, &voxel_stage_at(s_vlist, (uint32_t)t, col), sizeof(w));
        const uint16_t lo = (uint16_t)(w & 0xffffu), hi = (uint16_t)(w >> 16);
        if (lo != kVoxelFar) nbr[(size_t)(1 + t) * n + i] = i + (int)(int16_t)lo;
        if (t + 1 < cnt && hi != kVoxelFar) nbr[(size_t)(2 + t) * n + i] = i + (int)(int16_t)hi;
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
    const size_t n = (size_t)C.n;
    const float4 p = pos[i], v = vel[i];
    const int cnt = nbr[i];
    const float rho_i = rhoA[i];
    const float fi = p.w / (rho_i * rho_i);
    const float s = C.mass * C.visc / rho_i;
    float ax = 0.f, ay = 0.f, az = 0.f, bx = 0.f, by = 0.f, bz = 0.f;
    // four list entries and their gathers in flight, then summed in list order
    for (int t0 = 0; t0 < cnt; t0 += 4) {
        int jj[4];
        float4 qq[4], vv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) jj[u] = nbr[(size_t)(1 + min(t0 + u, cnt - 1)) * n + i];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            qq[u] = pos[jj[u]];  // q.w = p_j, vj.w = 1 / rho_j (store_density)
            vv[u] = vel[jj[u]];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (t0 + u >= cnt || jj[u] == i) continue;
            const float4 q = qq[u], vj = vv[u];
            const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
            const float r = sqrtf(dx * dx + dy * dy + dz * dz);
            const float hr = C.h - r;
            // voxel_kernels.py:162-168: (p_i/rho_i^2 + p_j/rho_j^2) * grad W
            const float cp = (fi + q.w * vj.w * vj.w) * C.grad_c * hr * hr * (1.0f / r);
            ax += cp * dx;
            ay += cp * dy;
            az += cp * dz;
            // voxel_kernels.py:203-208: (v_j - v_i) / rho_j * lap W * m * mu / rho_i
            const float cv = vj.w * (C.lap_c * hr) * s;
            bx += (vj.x - v.x) * cv;
            by += (vj.y - v.y) * cv;
            bz += (vj.z - v.z) * cv;
        }
    }
    // voxel_kernels.py:211 min(1, .)
    const Vec3 V{fminf(1.0f, bx), fminf(1.0f, by), fminf(1.0f, bz)};
    finish_particle<MODE, false>(C, i, p, v, pidA[i], rho_i, Vec3{ax, ay, az}, V, posOut, velOut, pidOut, keyI, valI, rng, pipe,
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
        const uint32_t s = table_start(cs[i], C), e = table_start(cs[i + 1], C);
        ((int32_t*)out)[i] = s < e ? (int32_t)s : -1;
        return;
    }
    if (i >= C.n) return;
    switch (which) {
        case SPH_BUF_KEYS: {  // NAIVE: keys in posB slot order (see hash_particle)
            const uint32_t u = keyI[i];
            const int64_t o = C.formulation == SPH_FORM_NAIVE ? (int64_t)pidB[i] : i;
            ((int32_t*)out)[o] = C.ref_keys ? (int32_t)(u ^ 0x80000000u) : (int32_t)(u & kCellKeyMask);
            break;
        }
        case SPH_BUF_PERM: {
            ((int32_t*)out)[i] = (int32_t)(perm_from_A ? pidA[i] : pidB[svals[i]]);
            break;
        }
        case SPH_BUF_SORTED_KEYS: {
            const uint32_t u = skeys[i];
            ((int32_t*)out)[i] = C.ref_keys ? (int32_t)(u ^ 0x80000000u) : (int32_t)(u & kCellKeyMask);
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
        case SPH_BUF_NEIGHBOURS: {  // row-major in particle order: count, then particle ids
            const int64_t id = pidA[i];
            int32_t* dst = (int32_t*)out + id * kNbrStride;
            const int cnt = nbr[i];
            dst[0] = cnt;
            for (int t = 0; t < kMaxNeighbours; ++t)
                dst[1 + t] = t < cnt ? (int32_t)pidA[nbr[(int64_t)(1 + t) * C.n + i]] : -1;
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
        if ((int)threadIdx.x < s)
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
        reorder_kernel<<<blocks_for(c->n, 256), 256, 0, s>>>(c->sc, c->sorted_vals, c->b.posB, c->b.velB,
                                                               c->b.pidB, c->b.posA, c->b.velA, c->b.pidA, c->b.flags);
    return hipGetLastError();
}

hipError_t launch_rekey(sph_ctx* c, hipStream_t s) {
    if (c->n > 0) rekey_kernel<<<blocks_for(c->n, 256), 256, 0, s>>>(c->sc, c->b.posB, c->b.keyI);
    return hipGetLastError();
}

template <int L>
void launch_density_lanes(sph_ctx* c, hipStream_t s) {
    const int bs = c->density_block;
    const Buffers& b = c->b;
    naive_density_kernel<L><<<blocks_for(c->n * L, bs), bs, 0, s>>>(c->sc, b.posA, b.velA, b.cs, b.rhoA, b.nl, b.nn);
}

hipError_t launch_density(sph_ctx* c, hipStream_t s) {
    if (c->n == 0) return hipGetLastError();
    const int bs = c->density_block;
    if (c->sc.formulation == SPH_FORM_VOXEL) {
        voxel_density_kernel<<<blocks_for(c->n, bs), bs, bs * kVoxelStageBytes, s>>>(c->sc, c->b.posA, c->b.velA,
                                                                                   c->b.cs, c->b.rhoA,
                                                                   c->b.nbr, c->b.flags);
        return hipGetLastError();
    }
    if (!c->inst.empty()) {
        const Buffers& b = c->b;
        naive_density_batch_kernel<<<blocks_for(c->sc.tfirst, bs), bs, 0, s>>>(c->sc, b.posA, b.velA, b.cs, b.rhoA,
                                                                              b.nl, b.nn);
        return hipGetLastError();
    }
    switch (c->lanes) {
        case 2: launch_density_lanes<2>(c, s); break;
        case 4: launch_density_lanes<4>(c, s); break;
        case 8: launch_density_lanes<8>(c, s); break;
        case 16: launch_density_lanes<16>(c, s); break;
        case 32: launch_density_lanes<32>(c, s); break;
        case 64: launch_density_lanes<64>(c, s); break;
        default: launch_density_lanes<1>(c, s); break;
    }
    return hipGetLastError();
}

template <int MODE, int L, bool COUNT = false>
void launch_force_lanes(sph_ctx* c, hipStream_t s, float4* dP, float4* dV, float4* dF) {
    const int bs = c->force_block;
    naive_force_kernel<MODE, L, COUNT><<<blocks_for(c->n * L, bs), bs, 0, s>>>(
        c->sc, c->b.posA, c->b.velA, c->b.pidA, c->b.rhoA, c->b.cs, c->b.nl, c->b.nn, c->b.posB, c->b.velB,
        c->b.pidB, c->b.keyI, c->b.valI, c->b.rng, c->b.pipe, dP, dV, dF, c->b.flags, c->b.sums,
        COUNT ? c->b.cell_cnt : nullptr, COUNT ? c->b.keyS[0] : nullptr,
        COUNT && L == 1 ? c->b.cell_part + (size_t)c->part_phase * kTileSumCopies * cell_scan_tiles(c->sc.cs_cells)
                        : nullptr);
}

template <int MODE>
void launch_force_mode(sph_ctx* c, hipStream_t s, bool count) {
    const int bs = c->force_block;
    float4* dP = c->keep_terms ? c->b.dbgP : nullptr;
    float4* dV = c->keep_terms ? c->b.dbgV : nullptr;
    float4* dF = c->keep_terms ? c->b.dbgF : nullptr;
    if (c->sc.formulation == SPH_FORM_VOXEL) {
        voxel_force_kernel<MODE><<<blocks_for(c->n, bs), bs, 0, s>>>(
            c->sc, c->b.posA, c->b.velA, c->b.pidA, c->b.rhoA, c->b.nbr, c->b.posB, c->b.velB, c->b.pidB, c->b.keyI,
            c->b.valI, c->b.rng, c->b.pipe, dP, dV, dF, c->b.flags);
        return;
    }
    if (!c->inst.empty()) {  // batched instances have box walls
        naive_force_batch_kernel<<<blocks_for(c->sc.tfirst, bs), bs, 0, s>>>(
            c->sc, c->b.posA, c->b.velA, c->b.pidA, c->b.rhoA, c->b.cs, c->b.nl, c->b.nn, c->b.posB, c->b.velB,
            c->b.pidB, c->b.keyI, c->b.valI, dP, dV, dF, c->b.flags);
        return;
    }
    if (count) {  // count_in_force
        switch (c->lanes) {
            case 2: launch_force_lanes<MODE, 2, true>(c, s, dP, dV, dF); break;
            case 4: launch_force_lanes<MODE, 4, true>(c, s, dP, dV, dF); break;
            case 8: launch_force_lanes<MODE, 8, true>(c, s, dP, dV, dF); break;
            case 16: launch_force_lanes<MODE, 16, true>(c, s, dP, dV, dF); break;
            case 32: launch_force_lanes<MODE, 32, true>(c, s, dP, dV, dF); break;
            case 64: launch_force_lanes<MODE, 64, true>(c, s, dP, dV, dF); break;
            default: launch_force_lanes<MODE, 1, true>(c, s, dP, dV, dF); break;
        }
        return;
    }
    switch (c->lanes) {
        case 2: launch_force_lanes<MODE, 2>(c, s, dP, dV, dF); break;
        case 4: launch_force_lanes<MODE, 4>(c, s, dP, dV, dF); break;
        case 8: launch_force_lanes<MODE, 8>(c, s, dP, dV, dF); break;
        case 16: launch_force_lanes<MODE, 16>(c, s, dP, dV, dF); break;
        case 32: launch_force_lanes<MODE, 32>(c, s, dP, dV, dF); break;
        case 64: launch_force_lanes<MODE, 64>(c, s, dP, dV, dF); break;
        default: launch_force_lanes<MODE, 1>(c, s, dP, dV, dF); break;
    }
}

bool count_in_force(const sph_ctx* c) {
    return c->cell_sort && c->inst.empty() && c->sc.formulation == SPH_FORM_NAIVE && cell_sort_global(c);
}

bool tile_sums_in_force(const sph_ctx* c) { return count_in_force(c) && c->lanes == 1; }

hipError_t launch_force(sph_ctx* c, hipStream_t s, bool count) {
    if (c->n == 0) return hipGetLastError();
    if (count && !count_in_force(c)) return hipErrorInvalidValue;
    if (c->sc.mode == SPH_MODE_PIPE) launch_force_mode<SPH_MODE_PIPE>(c, s, count);
    else launch_force_mode<SPH_MODE_BOX>(c, s, count);
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
