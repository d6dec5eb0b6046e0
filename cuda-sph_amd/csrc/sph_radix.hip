// sph_radix.hip -- stable LSD radix sort of (voxel key, slot) pairs for gfx950 (the VOXEL
// formulation, whose keys follow the reference's wrapping semantics and may lie outside the
// grid; the NAIVE formulation's clamped keys use the counting sort of sph_cellsort.hip).
//
// Replaces the host-side sort of the voxel map (voxel_sph_strategy.py:81-89), whose
// result is the lexicographic order on (voxel_id, particle_id).  The input pairs are
// laid out in particle-id order, so a stable sort on the key alone reproduces that order.
//
// Each pass sorts on up to 8 key bits with three launches:
//   upsweep   one workgroup per 2048-key tile counts digits in LDS (one histogram per
//             wavefront to spread LDS atomics) and writes counts digit-major;
//   scan      one workgroup per digit turns the counts of that digit into exclusive
//             tile offsets and a digit total;
//   downsweep each tile ranks its keys stably with wave64 ballots (the lanes holding the
//             same digit are found with one ballot per digit bit), places them in LDS in
//             digit order and writes each digit run contiguously.  On the last pass it can
//             also gather the particle state into cell order (posB/velB/pidB -> A), so
//             the separate reorder pass disappears.
#include "sph_device.h"

namespace sph {
namespace {

constexpr int RB = 8;
constexpr int RADIX = 1 << RB;   // 256 digits
constexpr int ST = 256;          // threads per workgroup, 4 wavefronts
constexpr int WAVES = ST / 64;
constexpr int SI = 8;            // keys per thread
constexpr int TILE = ST * SI;    // 2048 keys per tile
static_assert(RADIX == ST, "one digit per thread in the per-digit phases");

// Exclusive scan of one value per thread across the 256-thread workgroup.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* lds_wave, uint32_t* total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds_wave[w] = x;
    __syncthreads();
    uint32_t woff = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < WAVES; ++i) {
        uint32_t t = lds_wave[i];
        woff += i < w ? t : 0u;
        tot += t;
    }
    *total = tot;
    __syncthreads();
    return woff + x - v;
}

__global__ __launch_bounds__(ST) void radix_upsweep(const uint32_t* __restrict__ keys, int n, int shift,
                                                     uint32_t mask, uint32_t* __restrict__ hist, int ntiles,
                                                     uint32_t* __restrict__ clear_flag) {
    __shared__ uint32_t h[WAVES][RADIX];
    // the first pass clears the out-of-grid flag that this step's gather sets (the previous
    // step's density pass, its only reader, has finished)
    if (clear_flag && blockIdx.x == 0 && threadIdx.x == 0) *clear_flag = 0u;
    for (int i = threadIdx.x; i < WAVES * RADIX; i += ST) (&h[0][0])[i] = 0;
    __syncthreads();
    const int w = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * TILE;
#pragma unroll
    for (int k = 0; k < SI; ++k) {
        int64_t idx = base + k * ST + threadIdx.x;
        if (idx < n) atomicAdd(&h[w][(keys[idx] >> shift) & mask], 1u);
    }
    __syncthreads();
    const int d = threadIdx.x;
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < WAVES; ++i) s += h[i][d];
    hist[(int64_t)d * ntiles + blockIdx.x] = s;
}

__global__ __launch_bounds__(ST) void radix_scan_digits(uint32_t* __restrict__ hist, int ntiles,
                                                         uint32_t* __restrict__ digit_total) {
    __shared__ uint32_t lw[WAVES];
    uint32_t* row = hist + (int64_t)blockIdx.x * ntiles;
    uint32_t carry = 0;
    for (int base = 0; base < ntiles; base += ST) {
        const int i = base + threadIdx.x;
        uint32_t v = i < ntiles ? row[i] : 0u;
        uint32_t tot;
        uint32_t ex = block_excl_scan(v, lw, &tot);
        if (i < ntiles) row[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) digit_total[blockIdx.x] = carry;
}

// GATHER (the last pass of the fused step): also gathers the state into cell order and, for
// the VOXEL formulation (C.formulation), reports a position outside the voxel grid in
// flags[kFlagOutOfGrid].
template <bool GATHER>
__global__ __launch_bounds__(ST) void radix_downsweep(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                       uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, int n,
                                                       int shift, int bits, const uint32_t* __restrict__ hist,
                                                       const uint32_t* __restrict__ digit_total, int ntiles,
                                                       const float4* __restrict__ posB, const float4* __restrict__ velB,
                                                       const uint32_t* __restrict__ pidB, float4* __restrict__ posA,
                                                       float4* __restrict__ velA, uint32_t* __restrict__ pidA,
                                                       StepConsts C, uint32_t* __restrict__ flags) {
    __shared__ uint32_t s_keys[TILE];
    __shared__ uint32_t s_vals[TILE];
    __shared__ uint32_t s_whist[WAVES][RADIX];
    __shared__ uint32_t s_tstart[RADIX];
    __shared__ uint32_t s_gbase[RADIX];
    __shared__ uint32_t s_lw[WAVES];

    const uint32_t mask = (1u << bits) - 1u;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int64_t base = (int64_t)blockIdx.x * TILE;

    for (int i = tid; i < WAVES * RADIX; i += ST) (&s_whist[0][0])[i] = 0;
    {
        uint32_t tot;
        uint32_t ex = block_excl_scan(digit_total[tid], s_lw, &tot);  // synchronises: s_whist zeroed
        s_gbase[tid] = ex + hist[(int64_t)tid * ntiles + blockIdx.x];
    }

    uint32_t key[SI], val[SI], rank[SI];
    const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int k = 0; k < SI; ++k) {
        const int64_t idx = base + (int64_t)w * 64 * SI + k * 64 + lane;
        const bool valid = idx < n;
        key[k] = valid ? kin[idx] : 0u;
        val[k] = valid ? vin[idx] : 0u;
    }
#pragma unroll
    for (int k = 0; k < SI; ++k) {
        const int64_t idx = base + (int64_t)w * 64 * SI + k * 64 + lane;
        const bool valid = idx < n;
        const uint32_t d = (key[k] >> shift) & mask;
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < bits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        uint32_t r = 0;
        if (valid) {
            const uint32_t prefix = s_whist[w][d];
            const uint32_t below = (uint32_t)__popcll(peers & lt);
            r = prefix + below;
            if (below == 0) s_whist[w][d] = prefix + (uint32_t)__popcll(peers);
        }
        rank[k] = r;
    }
    __syncthreads();
    {
        const int d = tid;
        uint32_t run = 0;
#pragma unroll
        for (int i = 0; i < WAVES; ++i) {
            uint32_t t = s_whist[i][d];
            s_whist[i][d] = run;
            run += t;
        }
        uint32_t tot;
        s_tstart[d] = block_excl_scan(run, s_lw, &tot);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SI; ++k) {
        const int64_t idx = base + (int64_t)w * 64 * SI + k * 64 + lane;
        if (idx < n) {
            const uint32_t d = (key[k] >> shift) & mask;
            const uint32_t p = s_tstart[d] + s_whist[w][d] + rank[k];
            s_keys[p] = key[k];
            s_vals[p] = val[k];
        }
    }
    __syncthreads();
    const int tile_n = (int)((n - base) < TILE ? (n - base) : TILE);
    for (int i = tid; i < tile_n; i += ST) {
        const uint32_t kk = s_keys[i];
        const uint32_t vv = s_vals[i];
        const uint32_t d = (kk >> shift) & mask;
        const uint32_t g = s_gbase[d] + (uint32_t)i - s_tstart[d];
        kout[g] = kk;
        vout[g] = vv;
        if (GATHER) {
            const float4 p = posB[vv];
            posA[g] = p;
            velA[g] = velB[vv];
            pidA[g] = pidB[vv];
            if (C.formulation == SPH_FORM_VOXEL && outside_voxel_grid(p.x, p.y, p.z, C)) flags[kFlagOutOfGrid] = 1u;
        }
    }
}

}  // namespace

int radix_num_tiles(int64_t n) { return (int)((n + TILE - 1) / TILE); }

// Sorts (keyI, valI) into the radix ping-pong buffers.  When gather is set the last pass
// also writes posA/velA in sorted order; returns through c->sorted_keys/sorted_vals.
hipError_t launch_radix_sort_impl(sph_ctx* c, bool gather, hipStream_t s) {
    const int n = (int)c->n;
    const int T = c->num_tiles;
    const uint32_t* kin = c->b.keyI;
    const uint32_t* vin = c->b.valI;
    if (n == 0) {
        c->sorted_keys = c->b.keyI;
        c->sorted_vals = c->b.valI;
        return hipGetLastError();
    }
    for (int p = 0; p < c->passes; ++p) {
        uint32_t* kout = c->b.keyS[p & 1];
        uint32_t* vout = c->b.valS[p & 1];
        const int shift = c->pass_shift[p], bits = c->pass_bits[p];
        const uint32_t mask = (1u << bits) - 1u;
        radix_upsweep<<<T, ST, 0, s>>>(kin, n, shift, mask, c->b.tile_hist, T,
                                         p == 0 ? c->b.flags + kFlagOutOfGrid : nullptr);
        radix_scan_digits<<<RADIX, ST, 0, s>>>(c->b.tile_hist, T, c->b.digit_total);
        if (gather && p == c->passes - 1)
            radix_downsweep<true><<<T, ST, 0, s>>>(kin, vin, kout, vout, n, shift, bits, c->b.tile_hist,
                                                    c->b.digit_total, T, c->b.posB, c->b.velB, c->b.pidB, c->b.posA,
                                                    c->b.velA, c->b.pidA, c->sc, c->b.flags);
        else
            radix_downsweep<false><<<T, ST, 0, s>>>(kin, vin, kout, vout, n, shift, bits, c->b.tile_hist,
                                                     c->b.digit_total, T, nullptr, nullptr, nullptr, nullptr, nullptr,
                                                     nullptr, c->sc, nullptr);
        kin = kout;
        vin = vout;
    }
    c->sorted_keys = const_cast<uint32_t*>(kin);
    c->sorted_vals = const_cast<uint32_t*>(vin);
    return hipGetLastError();
}

}  // namespace sph
