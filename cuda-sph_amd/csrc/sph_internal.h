// sph_internal.h -- types shared by the gfx950 kernels and the host side of libsph_amd.so.
//
// Device layout (all arrays in one caller-owned workspace, 256-byte aligned slices):
//   posA/velA  float4[N]  state in cell order of the current step; the density pass puts
//                         the pressure p in posA.w and 1/rho in velA.w, so one neighbour costs
//                         two 16-byte loads: (x, y, z, p) and (vx, vy, vz, 1/rho)
//   pidA, rhoA u32[N], f32[N]  particle id and density of each cell-ordered slot
//   posB/velB  float4[N]  state after integration, still in the previous cell order, with
//   pidB       u32[N]     rho of the step in posB.w; this is the resident state between steps
//   keyI/valI  u32[N]     sort input written with the new positions.  NAIVE: key per posB slot
//                         (valI unused); VOXEL: key and posB slot per particle id
//   keyS/valS  u32[2][N]  radix ping-pong buffers
//   cs         u32[C+1]   cell start table: cs[c] = first cell-sorted index with key >= c
//                         (NAIVE: in the low start_bits bits, with the x-slab offsets of the
//                         cell above them, see StepConsts::slab_shift)
//   cell_cnt   u32[C]     NAIVE: per-cell counts of the counting sort (kept zero between steps)
//   cell_part  u32[2][K][T]  NAIVE: sums of the counts over each scan tile of 2^kScanTileShift
//                         cells, in K = kTileSumCopies copies; the scan reads one array while
//                         it clears the other, into which a counting force pass adds the next
//                         step's keys
//   rng        u64[2N]    xoroshiro128+ stream per particle id (PIPE mode)
//   cotx/cotv  float4[N]  reverse mode only: cotangent of the state in particle order
//   nl/nn      i16[96][N], u32[N]  NAIVE neighbour lists of the density pass, read by the
//                         force pass: each entry is the neighbour's index minus the particle's
//                         own.  Slots t and t + 1 (t even) of particle i form the 32-bit word
//                         (t / 2) * N + i, so consecutive particles' words are contiguous.
//                         The density pass appends a particle's first 32 L entries (L lanes
//                         per particle, at most 96) to its place in its wavefront's LDS band
//                         and copies them out slot by slot at the end (every store coalesced);
//                         hits beyond that go straight to HBM.  nn > 96 marks a
//                         particle whose neighbours the force pass rescans (more than 96, or
//                         one farther than an int16 offset).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sph_amd.h"

namespace sph {

constexpr int kMaxPipeRows = 64;
constexpr int kMaxNeighbours = 32;  // MAX_NEIGHBOURS, config.py:30
#ifndef SPH_LIST_CAP
#define SPH_LIST_CAP 96
#endif
constexpr int kListCap = SPH_LIST_CAP;  // NAIVE neighbour-list slots per particle; more falls back to a rescan
#ifndef SPH_LDS_LIST_CAP
#define SPH_LDS_LIST_CAP 16
#endif
constexpr int kLdsListCap = SPH_LDS_LIST_CAP;  // of which staged in LDS by the density pass
constexpr int kXcds = 8;            // MI355X: 8 accelerator complex dies, one L2 each

// Words of the flags buffer.
constexpr int kFlagNonFinite = 0;   // non-finite positions hashed since the last sph_check_finite
constexpr int kFlagOutOfGrid = 1;   // VOXEL: a finite position of this step's sorted state lies
                                    // outside the voxel grid (set by the gather into cell order,
                                    // cleared by the first pass of the sort)

// NAIVE neighbour-list entry: the neighbour's cell-ordered index minus the particle's own, as
// int16 (neighbours in space are near each other in cell order).  A particle with a
// neighbour farther away than that in cell order is marked for a rescan instead.
typedef uint16_t ListEntry;

// Everything a step kernel needs, passed by value (fits the kernel-argument segment).
struct StepConsts {
    int n;
    int mode;         // SPH_MODE_*
    int formulation;  // SPH_FORM_*
    int ref_keys;     // 1: reference key semantics (trunc, no clamp, signed); 0: clamped floor
    int pow2;         // 1: cell edges are powers of two, x * inv_cell is exact in FP32
    int dims[3];      // voxels per axis for keys: ceil(space / voxel) (voxel_sph_strategy.py:70-73)
    int tdims[3];     // voxels per axis for lookups: trunc(space / voxel) (voxel_sph_strategy.py:109-116)
    int stencil[3];   // cells searched on each side: ceil(h / voxel)
    int n_cells;      // dims[0] * dims[1] * dims[2]
    float inv_cell[3];
    float cell[3];
    double dcell[3];  // voxel edges as given (used when pow2 == 0)
    float h, h2, inv_h2;
    float r2_accept;  // VOXEL: the largest FP32 r^2 whose correctly rounded sqrt is <= h
    float h_search;   // h plus a rounding margin, for choosing cells (the pair test uses h2)
    float mass, rho0, k, visc, damp, dt;
    float w_norm;     // mass * 315 / (64 pi h^3): rho = w_norm * sum (1 - r^2/h^2)^3
    float grad_c;     // -45 / (pi h^6)           (config.py:28)
    float lap_c;      // 45 / (pi h^6)            (config.py:29)
    float wc9;        // 315 / (64 pi h^9)        (config.py:27), VOXEL density
    float ext[3];
    float space[3];
    float box_lo_target;     // 1e-3                  (base_kernels.py:91)
    float box_hi_target[3];  // space - 1e-3          (base_kernels.py:94)
    int pipe_rows;
    // VOXEL formulation: 1 when the lookup dims equal the key dims on every axis, so that
    // cell-table entry ci holds the particles of the voxel that ci names; only then may the
    // density pass skip a voxel whose slab lies beyond h (and only while no particle of the
    // sorted state lies outside the voxel grid, which the sort's gather reports in
    // flags[kFlagOutOfGrid])
    int voxel_skip;
    // batched form: this instance's first cell key and first particle (0 for a plain step);
    // the kernel argument holds the totals, the device table one entry per instance
    int cell_off;
    int first;
    int cs_cells;                          // cells covered by the cell table (all instances)
    int n_inst;                            // 1 for a plain step
    const struct StepConsts* inst_table;   // device, n_inst entries (batched form)
    const int* inst_first;                 // device, n_inst + 1 particle offsets (batched form)
    // batched form: lanes per particle of this instance in the NAIVE density and force passes
    // (lanes = 1 << lane_shift) and the first thread of its range in those passes, a multiple
    // of 64 so that no wavefront spans two instances; the totals entry holds the thread count
    int lanes;
    int lane_shift;
    int tfirst;
    // NAIVE grid keys: each cell is cut into S = 1 << slab_shift slabs along x (1: not cut).
    // A key carries its particle's slab in bits kSlabKeyShift.. (the cell below them); the
    // fused step's sort orders the particles of a cell by (slab, id), and the entry of an
    // occupied cell in the cell table then carries, above its start_bits low bits, the
    // offsets of the cell's slabs 1 .. S - 1 from its start, each as offset + 1 in a field of
    // slab_bits bits (0: not known, the cell is searched whole); an empty cell's fields hold
    // offset 0.  Where bits are left above them (count_bits from bit count_pos), the cell's
    // particle count + 1 (0: not kept), which gives the next cell's start without reading its
    // entry.  The same values in the totals and in every instance entry.
    int slab_shift;
    int start_bits;
    int slab_bits;
    int count_pos;
    int count_bits;
    uint32_t start_mask;  // the low start_bits bits
};

constexpr int kSlabKeyShift = 30;
constexpr uint32_t kCellKeyMask = (1u << kSlabKeyShift) - 1u;
// The counting sort's scan works in tiles of 2^kScanTileShift cells.  A tile's sum is kept
// in kTileSumCopies places (cell_part: copy k of tile t at k * tiles + t, the sum is over the
// copies), so that the force pass's wavefronts, which add into copy (wavefront index mod
// kTileSumCopies), do not all queue on one address.
constexpr int kScanTileShift = 12;
#ifndef SPH_TILE_COPIES
#define SPH_TILE_COPIES 4
#endif
constexpr int kTileSumCopies = SPH_TILE_COPIES;
static_assert((kTileSumCopies & (kTileSumCopies - 1)) == 0 && kTileSumCopies <= 64, "a power of two up to 64");

struct PipeTable {
    float row[kMaxPipeRows][5];   // Pipe.to_numpy() rows
    float xbeg[kMaxPipeRows];     // calc_x_at_segment_beginning of each row (util_kernels.py:38-43)
};

struct Buffers {
    float4 *posA, *velA, *posB, *velB;
    uint32_t *pidA, *pidB;
    float* rhoA;
    uint32_t *keyI, *valI;
    uint32_t *keyS[2], *valS[2];
    uint32_t* cs;
    uint32_t* tile_hist;   // radix digit counts per tile, digit-major
    uint32_t* digit_total;
    uint64_t* rng;
    float4 *dbgP, *dbgV, *dbgF;
    int32_t* nbr;          // VOXEL formulation neighbour lists, N x (1 + 32)
    ListEntry* nl;         // NAIVE neighbour lists, column-major [kListCap][N] (index offsets, self included)
    uint32_t* nn;          // NAIVE neighbour count per cell-ordered particle (may exceed kListCap)
    double* diag_partials;
    uint32_t* flags;       // kFlagNonFinite, kFlagOutOfGrid (16 words)
    PipeTable* pipe;
    StepConsts* inst_table;
    int* inst_first;
    // reverse mode (SPH_FLAG_ADJOINT): cotangent of the state in particle order, and per
    // cell-ordered particle the force cotangent (.w: rho cotangent of the integration),
    // position cotangent so far (.w: the particle's dL/dmu term), velocity cotangent of the
    // integration and the total rho cotangent
    uint32_t* cell_cnt;    // NAIVE counting sort: particles per cell (zero between steps)
    uint32_t* cell_part;   // NAIVE counting sort: two arrays of per-tile sums of the cell counts,
                           // the one the scan reads and the one the force pass adds into
    float4 *cotx, *cotv;
    float4 *adj_fb, *adj_gx, *adj_gv;
    float* adj_rt;
    float* mu_acc;
    float2* mu_part;       // dL/dmu per particle id, summed over the reversed steps (compensated FP32 pair)
    float4* sums;          // adjoint contexts: the force pass's pressure and viscosity sums
                           // per cell-ordered particle, [2][N] (null otherwise)
};

}  // namespace sph

#include <vector>

struct sph_ctx {
    sph_params_t prm;
    sph_constants_t cst;
    sph::StepConsts sc;
    sph::PipeTable pipe_host;
    int64_t n;
    int key_bits;
    int passes;
    int pass_shift[4];
    int pass_bits[4];
    int num_tiles;
    int density_block, force_block;
    int lanes;               // NAIVE formulation: lanes of a wavefront per particle in density / force (1..16)
    int keep_terms;
    int check_finite;        // SPH_FLAG_CHECK_FINITE: sph_step reads the non-finite count
    int adjoint;             // SPH_FLAG_ADJOINT: reverse-mode buffers are part of the workspace
    int cell_sort;           // NAIVE formulation: counting sort by cell instead of the radix sort
    int whole_cells;         // sph_set_slabs(ctx, 0): cells are not cut into x slabs
    int part_phase;          // which cell_part array a force pass adds the tile sums into (the
                             // other was the last sort's), set by launch_cell_sort
    bool adj_seeded;
    size_t ws_bytes;
    void* ws;
    sph::Buffers b;
    uint32_t* sorted_keys;   // final radix output (key) in cell order
    uint32_t* sorted_vals;   // final radix output (slot in posB)
    bool bound;
    bool have_state;
    int last_stage;          // last stage run in the current step (0 = hashed, ready)
    bool a_in_snapshot;      // the last step's sorted state, densities and table went to a snapshot
    uint64_t* rng_host;      // seeded on the host, uploaded at bind
    // batched form (empty for a plain step)
    std::vector<sph::StepConsts> inst;
    std::vector<int> inst_first;
    std::vector<sph_constants_t> inst_cst;
};

namespace sph {
// Launchers (sph_kernels.hip / sph_radix.hip).  All return hipError_t of the last launch.
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s);
hipError_t launch_upload_f64(sph_ctx* c, const double* pos, const double* vel, hipStream_t s);
hipError_t launch_upload_f32(sph_ctx* c, const float* x, const float* y, const float* z, const float* vx,
                             const float* vy, const float* vz, hipStream_t s);
hipError_t launch_radix_sort_impl(sph_ctx* c, bool gather, hipStream_t s);
// sph_cellsort.hip; also builds cs.  counted: the previous force pass counted the keys
hipError_t launch_cell_sort(sph_ctx* c, bool gather, hipStream_t s, bool counted = false);
// whether the counting sort runs on the global path (per-cell counts in cell_cnt), not in LDS
bool cell_sort_global(const sph_ctx* c);
// whether the force pass can count the keys it writes for the next step's counting sort:
// a plain (not batched) NAIVE step on the global sort path
bool count_in_force(const sph_ctx* c);
// whether such a force pass also adds the keys into the scan's tile sums (one lane per
// particle)
bool tile_sums_in_force(const sph_ctx* c);
int64_t cell_scan_tiles(int64_t cells);
hipError_t launch_cell_table(sph_ctx* c, hipStream_t s);
hipError_t launch_reorder(sph_ctx* c, hipStream_t s);
// keyI of the resident state (posB slots) again, under the context's current cell layout
hipError_t launch_rekey(sph_ctx* c, hipStream_t s);
hipError_t launch_density(sph_ctx* c, hipStream_t s);
// count: also rank the new keys into cell_cnt / keyS[0] (count_in_force) and add them to the
// tile sums in cell_part array part_phase (tile_sums_in_force), for the next sort
hipError_t launch_force(sph_ctx* c, hipStream_t s, bool count = false);
hipError_t launch_download_f64(sph_ctx* c, double* pos, double* vel, double* rho, double* force, hipStream_t s);
hipError_t launch_download_f32(sph_ctx* c, float* x, float* y, float* z, float* vx, float* vy, float* vz,
                               float* rho, hipStream_t s);
hipError_t launch_read_buffer(sph_ctx* c, int which, void* out, hipStream_t s);
hipError_t launch_diagnostics(sph_ctx* c, double* out8, hipStream_t s);
int radix_num_tiles(int64_t n);
// sph_adjoint.hip
hipError_t launch_adjoint_set(sph_ctx* c, const float* gx, const float* gv, hipStream_t s);
hipError_t launch_adjoint_get(sph_ctx* c, float* gx, float* gv, float* gmu, hipStream_t s);
// the pullback of the step whose sorted state, densities, lists and cell table b holds
// from_sums: take the force pass's sums from b.sums instead of walking the lists again
hipError_t launch_adjoint_pullback(sph_ctx* c, const Buffers& b, hipStream_t s, bool from_sums = false);
int64_t adjoint_snapshot_bytes(const sph_ctx* c);
// c's buffers with the eight a snapshot holds pointed at their places in snap
Buffers snapshot_view(const sph_ctx* c, void* snap);
hipError_t launch_adjoint_save(sph_ctx* c, void* snap, hipStream_t s);  // copy c's eight buffers into snap
}  // namespace sph
