// sph_capi.hip -- host side of libsph_amd.so: context planning, workspace layout,
// RNG seeding and the extern "C" entry points declared in include/sph_amd.h.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <new>

#include "sph_internal.h"

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return SPH_OK;
    return fail(SPH_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

bool is_pow2(double v) {
    int e;
    double m = frexp(v, &e);
    return v > 0 && m == 0.5;
}

int bit_length(uint64_t v) {
    int b = 0;
    while (v) { ++b; v >>= 1; }
    return b;
}

// xoroshiro128+ stream seeding with the scheme of numba.cuda.random.create_xoroshiro128p_states
// (used at abstract_sph_strategy.py:27): one SplitMix64 output in both state words for
// stream 0, each further stream is the previous one advanced 2^64 draws by the published
// jump polynomial.  Written from the published algorithms (Vigna, SplitMix64;
// Blackman & Vigna, xoroshiro128+ 2016).
uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
void xnext(uint64_t* s) {
    uint64_t a = s[0], b = s[1];
    b ^= a;
    s[0] = rotl(a, 55) ^ b ^ (b << 14);
    s[1] = rotl(b, 36);
}
void xjump(uint64_t* s) {
    const uint64_t J[2] = {0xbeac0467eba5facbull, 0xd86b048b86aa9922ull};
    uint64_t a = 0, b = 0;
    for (int i = 0; i < 2; ++i)
        for (int bit = 0; bit < 64; ++bit) {
            if (J[i] >> bit & 1ull) { a ^= s[0]; b ^= s[1]; }
            xnext(s);
        }
    s[0] = a;
    s[1] = b;
}
void seed_streams(uint64_t seed, int64_t n, uint64_t* out) {
    if (n <= 0) return;
    uint64_t z = seed + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    out[0] = out[1] = z;
    for (int64_t i = 1; i < n; ++i) {
        out[2 * i] = out[2 * i - 2];
        out[2 * i + 1] = out[2 * i - 1];
        xjump(out + 2 * i);
    }
}

size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

// Workspace slices, in order.
enum Slice {
    S_POSA, S_VELA, S_POSB, S_VELB, S_PIDA, S_PIDB, S_RHOA, S_KEYI, S_VALI, S_KEYS0, S_VALS0, S_KEYS1, S_VALS1, S_CS,
    S_HIST, S_DTOT, S_RNG, S_DBGP, S_DBGV, S_DBGF, S_NBR, S_NL, S_NN, S_DIAG, S_FLAGS, S_PIPE, S_INST, S_INSTF,
    S_COTX, S_COTV, S_AFB, S_AGX, S_AGV, S_ART, S_MU, S_MUP, S_SUMS, S_CCNT, S_CPART, S_COUNT
};

struct Layout {
    size_t off[S_COUNT];
    size_t total;
};

Layout plan_layout(const sph_ctx* c) {
    const size_t n = (size_t)c->n, nn = n > 0 ? n : 1;
    size_t sz[S_COUNT];
    sz[S_POSA] = sz[S_VELA] = sz[S_POSB] = sz[S_VELB] = 16 * nn;
    sz[S_PIDA] = sz[S_PIDB] = sz[S_RHOA] = 4 * nn;
    sz[S_KEYI] = sz[S_VALI] = sz[S_KEYS0] = sz[S_VALS0] = sz[S_KEYS1] = sz[S_VALS1] = 4 * nn;
    sz[S_CS] = 4 * ((size_t)c->sc.n_cells + 1);
    sz[S_HIST] = 4 * 256 * (size_t)(c->num_tiles > 0 ? c->num_tiles : 1);
    sz[S_DTOT] = 4 * 256;
    sz[S_RNG] = c->prm.mode == SPH_MODE_PIPE ? 16 * nn : 16;
    sz[S_DBGP] = sz[S_DBGV] = sz[S_DBGF] = c->keep_terms ? 16 * nn : 16;
    sz[S_NBR] = c->prm.formulation == SPH_FORM_VOXEL ? 4 * 33 * nn : 16;
    sz[S_NL] = c->prm.formulation == SPH_FORM_NAIVE ? sizeof(sph::ListEntry) * (size_t)sph::kListCap * nn : 16;
    sz[S_NN] = 4 * nn;
    sz[S_DIAG] = 8 * 8 * 512;
    sz[S_FLAGS] = 64;
    sz[S_PIPE] = sizeof(sph::PipeTable);
    sz[S_INST] = sizeof(sph::StepConsts) * (c->inst.empty() ? 1 : c->inst.size());
    sz[S_INSTF] = sizeof(int) * (c->inst_first.empty() ? 1 : c->inst_first.size());
    sz[S_COTX] = sz[S_COTV] = sz[S_AFB] = sz[S_AGX] = sz[S_AGV] = c->adjoint ? 16 * nn : 16;
    sz[S_ART] = c->adjoint ? 4 * nn : 16;
    sz[S_MU] = 64;
    sz[S_MUP] = c->adjoint ? 8 * nn : 16;
    sz[S_SUMS] = c->adjoint ? 32 * nn : 16;
    sz[S_CCNT] = c->cell_sort ? 4 * (size_t)(c->sc.cs_cells > 0 ? c->sc.cs_cells : 1) : 16;
    sz[S_CPART] = c->cell_sort ? 8 * (size_t)sph::kTileSumCopies *
                                     (size_t)sph::cell_scan_tiles(c->sc.cs_cells > 0 ? c->sc.cs_cells : 1)
                               : 16;
    Layout L;
    size_t o = 0;
    for (int i = 0; i < S_COUNT; ++i) {
        L.off[i] = o;
        o = align_up(o + sz[i]);
    }
    L.total = o;
    return L;
}

// Validates one (params, constants) pair and fills its kernel constants and pipe table.
int plan_instance(const sph_params_t& P, const sph_constants_t& K, sph::StepConsts& S, sph::PipeTable& pipe) {
    // 2^27 particles keep every byte offset of a 16-byte-per-particle array within 31 bits
    if (P.particle_count < 0 || P.particle_count > (int64_t(1) << 27))
        return fail(SPH_E_INVALID, "particle_count %lld outside [0, 2^27]", (long long)P.particle_count);
    if (!(K.h > 0) || !(K.mass > 0) || !isfinite(K.h) || !isfinite(K.mass))
        return fail(SPH_E_INVALID, "h and mass must be positive and finite");
    if (!(P.dt > 0) || !isfinite(P.dt)) return fail(SPH_E_INVALID, "dt must be positive and finite");
    if (P.mode != SPH_MODE_BOX && P.mode != SPH_MODE_PIPE) return fail(SPH_E_INVALID, "unknown mode %d", P.mode);
    if (P.formulation != SPH_FORM_NAIVE && P.formulation != SPH_FORM_VOXEL)
        return fail(SPH_E_INVALID, "unknown formulation %d", P.formulation);
    memset(&S, 0, sizeof(S));
    S.n = (int)P.particle_count;
    S.mode = P.mode;
    S.formulation = P.formulation;
    S.ref_keys = P.formulation == SPH_FORM_VOXEL ? 1 : 0;
    S.pow2 = 1;
    S.n_inst = 1;
    S.lanes = 1;
    int64_t cells = 1;
    for (int d = 0; d < 3; ++d) {
        const double sp = P.space_size[d], vx = P.voxel_size[d];
        if (!(sp > 0) || !(vx > 0) || !isfinite(sp) || !isfinite(vx))
            return fail(SPH_E_INVALID, "space_size and voxel_size must be positive and finite");
        const double q = sp / vx;
        if (q > 1 << 20) return fail(SPH_E_INVALID, "space/voxel = %g cells on axis %d is too many", q, d);
        S.dims[d] = (int)ceil(q);        // voxel_sph_strategy.py:70-73
        S.tdims[d] = (int)q;             // voxel_sph_strategy.py:109-116
        if (S.dims[d] < 1) S.dims[d] = 1;
        // cells searched on each side; exact for power-of-two voxels, with the search margin otherwise
        S.stencil[d] = (int)ceil((is_pow2(vx) ? K.h : K.h * (1.0 + 2e-5)) / vx);
        if (!is_pow2(vx)) S.pow2 = 0;
        S.dcell[d] = vx;
        S.cell[d] = (float)vx;
        S.inv_cell[d] = (float)(1.0 / vx);
        cells *= S.dims[d];
        S.space[d] = (float)sp;
        S.box_hi_target[d] = (float)(sp - 1e-3);
        S.ext[d] = (float)P.external_force[d];
    }
    if (cells > (int64_t(1) << 28)) return fail(SPH_E_INVALID, "voxel grid of %lld cells is too large", (long long)cells);
    S.voxel_skip = S.dims[0] == S.tdims[0] && S.dims[1] == S.tdims[1] && S.dims[2] == S.tdims[2] ? 1 : 0;
    S.n_cells = (int)cells;
    S.cs_cells = (int)cells;
    S.box_lo_target = 1e-3f;
    S.h = (float)K.h;
    S.h2 = (float)(K.h * K.h);
    S.inv_h2 = (float)(1.0 / (K.h * K.h));
    // sqrtf is correctly rounded and monotone, so sqrtf(r2) <= h holds exactly for
    // r2 <= r2_accept: the VOXEL pair test without a square root per candidate
    {
        float t = S.h * S.h;
        while (sqrtf(nextafterf(t, INFINITY)) <= S.h) t = nextafterf(t, INFINITY);
        while (sqrtf(t) > S.h) t = nextafterf(t, 0.0f);
        S.r2_accept = t;
    }
    double maxsp = fmax(P.space_size[0], fmax(P.space_size[1], P.space_size[2]));
    S.h_search = (float)(K.h * (1.0 + 1e-5) + 4e-7 * maxsp);
    S.mass = (float)K.mass;
    S.rho0 = (float)K.rho0;
    S.k = (float)K.k;
    S.visc = (float)K.visc;
    S.damp = (float)K.damp;
    S.dt = (float)P.dt;
    S.w_norm = (float)(K.mass * 315.0 / (64.0 * M_PI * K.h * K.h * K.h));
    S.grad_c = (float)(-45.0 / (M_PI * pow(K.h, 6.0)));
    S.lap_c = (float)(45.0 / (M_PI * pow(K.h, 6.0)));
    S.wc9 = (float)(315.0 / (64.0 * M_PI * pow(K.h, 9.0)));
    memset(&pipe, 0, sizeof(pipe));
    S.pipe_rows = 0;
    if (P.mode == SPH_MODE_PIPE) {
        if (P.pipe_rows < 2 || P.pipe_rows > sph::kMaxPipeRows || !P.pipe)
            return fail(SPH_E_INVALID, "PIPE mode needs a pipe table of 2..%d rows", sph::kMaxPipeRows);
        S.pipe_rows = P.pipe_rows;
        double xb = P.pipe[0];
        for (int r = 0; r < P.pipe_rows; ++r) {
            for (int k = 0; k < 5; ++k) pipe.row[r][k] = (float)P.pipe[r * 5 + k];
            pipe.xbeg[r] = (float)xb;  // util_kernels.py:38-43, summed in FP64
            xb += P.pipe[r * 5 + 4];
        }
    }
    return SPH_OK;
}

#ifndef SPH_SLAB_SHIFT
#define SPH_SLAB_SHIFT 2
#endif

// x slabs per cell of the NAIVE grid (StepConsts::slab_shift) and the layout of their
// offsets in the cell table: four slabs (SPH_SLAB_SHIFT = 2) while the start, which needs
// bit_length(n) bits, leaves room for three offset fields, else two.  The field width is
// fixed at 3 bits (offsets up to 6), so whether a cell's offsets are kept depends on its
// particles alone: a batch instance walks the same candidate ranges as the same simulation
// run alone, given the same slab count (batches of fewer than 2^23 particles).  The bits
// left above the fields (up to 16) keep the cell's count where it fits.  The VOXEL
// formulation keeps the reference's table.
void plan_slabs(sph_ctx* c) {
    constexpr int kField = 3;
    int shift = 0, start_bits = 32, field = 0;
    if (c->sc.formulation == SPH_FORM_NAIVE) {
        const int need = bit_length((uint64_t)(c->n > 0 ? c->n : 1));
        for (int sh = c->whole_cells ? 0 : SPH_SLAB_SHIFT; sh > 0; --sh) {
            if (need + kField * ((1 << sh) - 1) <= 32) {
                shift = sh;
                start_bits = need;
                field = kField;
                break;
            }
        }
    }
    const int count_pos = start_bits + field * ((1 << shift) - 1);
    const int count_bits = shift > 0 ? std::min(16, 32 - count_pos) : 0;
    auto apply = [&](sph::StepConsts& S) {
        S.slab_shift = shift;
        S.start_bits = start_bits;
        S.slab_bits = field;
        S.count_pos = count_pos;
        S.count_bits = count_bits;
        S.start_mask = start_bits >= 32 ? ~0u : (1u << start_bits) - 1u;
    };
    apply(c->sc);
    for (auto& S : c->inst) apply(S);
}

// Sort plan and workspace size from c->sc (totals) and c->n.
void plan_sort_and_layout(sph_ctx* c) {
    plan_slabs(c);
    const int64_t cells = c->sc.cs_cells;
    c->key_bits = c->sc.ref_keys ? 32 : bit_length((uint64_t)(cells > 1 ? cells - 1 : 0));
    c->passes = (c->key_bits + 7) / 8;
    int shift = 0;
    for (int p = 0; p < c->passes; ++p) {
        const int bits = (c->key_bits - shift + (c->passes - p) - 1) / (c->passes - p);
        c->pass_shift[p] = shift;
        c->pass_bits[p] = bits;
        shift += bits;
    }
    c->num_tiles = sph::radix_num_tiles(c->n);
    c->cell_sort = c->sc.formulation == SPH_FORM_NAIVE ? 1 : 0;
    if (c->density_block <= 0) c->density_block = 256;
    if (c->force_block <= 0) c->force_block = 256;
    if (c->lanes <= 0) c->lanes = 1;
    c->ws_bytes = plan_layout(c).total;
}

int plan(sph_ctx* c) {
    int rc = plan_instance(c->prm, c->cst, c->sc, c->pipe_host);
    if (rc) return rc;
    c->n = c->prm.particle_count;
    c->keep_terms = (c->prm.flags & SPH_FLAG_KEEP_TERMS) ? 1 : 0;
    c->check_finite = (c->prm.flags & SPH_FLAG_CHECK_FINITE) ? 1 : 0;
    c->adjoint = (c->prm.flags & SPH_FLAG_ADJOINT) ? 1 : 0;
    if (c->adjoint && (c->prm.formulation != SPH_FORM_NAIVE || c->prm.mode != SPH_MODE_BOX))
        return fail(SPH_E_INVALID, "SPH_FLAG_ADJOINT needs the NAIVE formulation in BOX mode");
    plan_sort_and_layout(c);
    return SPH_OK;
}

// Lanes of a wavefront per particle: a power of two that divides the wavefront.
bool valid_lanes(int L) { return L >= 1 && L <= 64 && (L & (L - 1)) == 0; }

// Thread ranges of the batched density and force passes: instance b takes n_b * lanes_b
// threads from tfirst_b on, rounded up to whole wavefronts; the totals entry gets the count.
int batch_thread_layout(sph_ctx* c, const std::vector<int>& lanes) {
    int64_t t = 0;
    for (size_t b = 0; b < c->inst.size(); ++b) {
        const int L = lanes[b];
        if (!valid_lanes(L))
            return fail(SPH_E_INVALID, "lanes per particle must be 1, 2, 4, 8, 16, 32 or 64 (instance %zu: %d)", b, L);
        t += (((int64_t)c->inst[b].n * L) + 63) & ~int64_t(63);
        if (t >= (int64_t(1) << 31))
            return fail(SPH_E_INVALID, "the batch's particles x lanes exceed 2^31 threads");
    }
    t = 0;
    for (size_t b = 0; b < c->inst.size(); ++b) {
        sph::StepConsts& S = c->inst[b];
        S.lanes = lanes[b];
        S.lane_shift = __builtin_ctz((unsigned)lanes[b]);
        S.tfirst = (int)t;
        t += (((int64_t)S.n * S.lanes) + 63) & ~int64_t(63);
    }
    c->sc.tfirst = (int)t;
    return SPH_OK;
}

// Instance table and particle offsets to the workspace (at bind, and when lanes change).
int upload_instances(sph_ctx* c, hipStream_t s) {
    sph::Buffers& b = c->b;
    for (auto& st : c->inst) {
        st.inst_table = b.inst_table;
        st.inst_first = b.inst_first;
    }
    int rc;
    if ((rc = hip_check(hipMemcpyAsync(b.inst_table, c->inst.data(), sizeof(sph::StepConsts) * c->inst.size(),
                                       hipMemcpyHostToDevice, s), "upload instance table")))
        return rc;
    return hip_check(hipMemcpyAsync(b.inst_first, c->inst_first.data(), sizeof(int) * c->inst_first.size(),
                                    hipMemcpyHostToDevice, s), "upload instance offsets");
}

// Batched form: instances are concatenated in the given order; instance b owns particle ids
// [first_b, first_b + N_b) and cell keys [cell_off_b, cell_off_b + C_b).  All instances use
// the NAIVE formulation and box walls.
int plan_batch(sph_ctx* c, const sph_params_t* params, const sph_constants_t* consts, int count) {
    if (count < 1 || count > 4096) return fail(SPH_E_INVALID, "batch of %d instances (1..4096)", count);
    c->inst.resize(count);
    c->inst_first.resize(count + 1);
    c->inst_cst.assign(consts, consts + count);
    int64_t n = 0, cells = 0;
    for (int b = 0; b < count; ++b) {
        if (params[b].formulation != SPH_FORM_NAIVE || params[b].mode != SPH_MODE_BOX)
            return fail(SPH_E_INVALID, "batched instances use the NAIVE formulation in BOX mode (instance %d)", b);
        if (params[b].flags & SPH_FLAG_ADJOINT)
            return fail(SPH_E_INVALID, "SPH_FLAG_ADJOINT is for plain contexts (instance %d)", b);
        sph::PipeTable unused;
        int rc = plan_instance(params[b], consts[b], c->inst[b], unused);
        if (rc) {
            char why[sizeof(g_err)];
            snprintf(why, sizeof(why), "%s", g_err);
            return fail(rc, "instance %d: %s", b, why);
        }
        c->inst[b].first = (int)n;
        c->inst[b].cell_off = (int)cells;
        c->inst_first[b] = (int)n;
        n += params[b].particle_count;
        cells += c->inst[b].n_cells;
        if (n > (int64_t(1) << 27)) return fail(SPH_E_INVALID, "batch of %lld particles is too large", (long long)n);
        if (cells > (int64_t(1) << 28)) return fail(SPH_E_INVALID, "batch grid of %lld cells is too large", (long long)cells);
    }
    c->inst_first[count] = (int)n;
    c->prm = params[0];
    c->prm.pipe = nullptr;
    c->prm.particle_count = n;
    c->cst = consts[0];
    c->sc = c->inst[0];
    c->sc.n = (int)n;
    c->sc.n_cells = (int)cells;
    c->sc.cs_cells = (int)cells;
    c->sc.cell_off = 0;
    c->sc.first = 0;
    c->sc.n_inst = count;
    c->n = n;
    if (int rc = batch_thread_layout(c, std::vector<int>(count, 1))) return rc;
    c->keep_terms = (params[0].flags & SPH_FLAG_KEEP_TERMS) ? 1 : 0;
    c->check_finite = (params[0].flags & SPH_FLAG_CHECK_FINITE) ? 1 : 0;
    memset(&c->pipe_host, 0, sizeof(c->pipe_host));
    plan_sort_and_layout(c);
    return SPH_OK;
}

// Records ev (timing of sph_step_timed) on s, if given.
int mark(hipEvent_t ev, hipStream_t s) { return ev ? hip_check(hipEventRecord(ev, s), "record event") : SPH_OK; }

// Sort (with the gather into cell order), cell table and density of a step.  counted: the
// previous step's force pass, in the same call, counted the keys (see run_steps).  ev: if
// given, events recorded before the sort (ev[0]) and before the density pass (ev[1]).
int run_density_stages(sph_ctx* c, hipStream_t s, bool fused, bool counted = false, hipEvent_t* ev = nullptr) {
    int rc;
    if (ev && (rc = mark(ev[0], s))) return rc;
    if (c->cell_sort) {  // the counting sort builds the cell table itself
        if ((rc = hip_check(sph::launch_cell_sort(c, fused, s, counted), "cell sort"))) return rc;
        if (!fused && (rc = hip_check(sph::launch_reorder(c, s), "reorder"))) return rc;
        if (ev && (rc = mark(ev[1], s))) return rc;
        return hip_check(sph::launch_density(c, s), "density");
    }
    if (fused && c->passes > 0) {
        if ((rc = hip_check(sph::launch_radix_sort_impl(c, true, s), "sort"))) return rc;
    } else {
        if ((rc = hip_check(sph::launch_radix_sort_impl(c, false, s), "sort"))) return rc;
        if ((rc = hip_check(sph::launch_reorder(c, s), "reorder"))) return rc;
    }
    if ((rc = hip_check(sph::launch_cell_table(c, s), "cell table"))) return rc;
    if (ev && (rc = mark(ev[1], s))) return rc;
    return hip_check(sph::launch_density(c, s), "density");
}

// ev: if given, four events: before the sort, the density and the force pass, and after it
int run_step_stages(sph_ctx* c, hipStream_t s, bool fused, bool counted = false, bool count = false,
                    hipEvent_t* ev = nullptr) {
    int rc;
    if ((rc = run_density_stages(c, s, fused, counted, ev))) return rc;
    if (ev && (rc = mark(ev[2], s))) return rc;
    if ((rc = hip_check(sph::launch_force(c, s, count), "force"))) return rc;
    if (ev && (rc = mark(ev[3], s))) return rc;
    c->last_stage = SPH_STAGE_FORCE;
    return SPH_OK;
}

// nsteps fused steps.  Where the force pass can count the keys it writes (count_in_force),
// every step but the last counts them for the next step's sort, which then starts at the
// scan; the first step of a call counts in its sort, so between calls the per-cell counts
// stay zero and every other path (upload, staged steps, snapshots, the pullbacks) keeps
// its own sort.
// ev: if given, four events per step (see run_step_stages).
int run_steps(sph_ctx* c, hipStream_t s, int32_t nsteps, hipEvent_t* ev = nullptr) {
    const bool fuse = sph::count_in_force(c);
    for (int32_t k = 0; k < nsteps; ++k) {
        if (int rc = run_step_stages(c, s, true, fuse && k > 0, fuse && k + 1 < nsteps, ev ? ev + 4 * k : nullptr)) {
            // counts left by a force pass whose sort did not run would be counted twice by
            // the next sort (and scatter past the end): clear them
            if (fuse && k > 0) (void)sph::launch_fill_u32(c->b.cell_cnt, 0u, c->sc.cs_cells, s);
            return rc;
        }
        c->a_in_snapshot = false;
    }
    return SPH_OK;
}

// Reads and clears the count of non-finite positions hashed since the last read (waits for
// the stream).  Returns SPH_E_NONFINITE if it was not zero.
int read_nonfinite(sph_ctx* c, hipStream_t s, int64_t* count) {
    uint32_t host = 0;
    int rc;
    if ((rc = hip_check(hipMemcpyAsync(&host, c->b.flags + sph::kFlagNonFinite, sizeof(host), hipMemcpyDeviceToHost, s),
                        "read non-finite count")))
        return rc;
    if ((rc = hip_check(sph::launch_fill_u32(c->b.flags + sph::kFlagNonFinite, 0u, 1, s), "clear non-finite count")))
        return rc;
    if ((rc = hip_check(hipStreamSynchronize(s), "check finite"))) return rc;
    if (count) *count = host;
    if (host != 0)
        return fail(SPH_E_NONFINITE, "%u particle positions were NaN or infinite when hashed (since the last check)",
                    host);
    return SPH_OK;
}

void assign_buffers(sph_ctx* c, void* dev);

int ready(const sph_ctx* c, bool need_state) {
    if (!c) return fail(SPH_E_INVALID, "null context");
    if (!c->bound) return fail(SPH_E_STATE, "no workspace bound (call sph_bind_workspace)");
    if (need_state && !c->have_state) return fail(SPH_E_STATE, "no state uploaded (call sph_upload_state)");
    return SPH_OK;
}

}  // namespace

extern "C" {

int sph_abi_version(void) { return SPH_ABI_VERSION; }

int sph_last_error(char* buf, size_t len) {
    if (!buf || len == 0) return SPH_E_INVALID;
    snprintf(buf, len, "%s", g_err);
    return SPH_OK;
}

int sph_create(const sph_params_t* params, const sph_constants_t* consts, sph_ctx** out) {
    if (!params || !consts || !out) return fail(SPH_E_INVALID, "null argument");
    *out = nullptr;
    sph_ctx* c = new (std::nothrow) sph_ctx();
    if (!c) return fail(SPH_E_INVALID, "out of host memory");
    c->prm = *params;
    c->cst = *consts;
    int rc = plan(c);          // copies the pipe table out of params->pipe
    c->prm.pipe = nullptr;     // the caller's table is not referenced after this call
    if (rc == SPH_OK && params->mode == SPH_MODE_PIPE) {
        c->rng_host = new (std::nothrow) uint64_t[2 * (size_t)(c->n > 0 ? c->n : 1)];
        if (!c->rng_host) rc = fail(SPH_E_INVALID, "out of host memory");
        else seed_streams(params->seed, c->n, c->rng_host);
    }
    if (rc != SPH_OK) {
        delete[] c->rng_host;
        delete c;
        return rc;
    }
    *out = c;
    return SPH_OK;
}

int sph_create_batch(const sph_params_t* params, const sph_constants_t* consts, int32_t count, sph_ctx** out) {
    if (!params || !consts || !out) return fail(SPH_E_INVALID, "null argument");
    *out = nullptr;
    sph_ctx* c = new (std::nothrow) sph_ctx();
    if (!c) return fail(SPH_E_INVALID, "out of host memory");
    int rc = plan_batch(c, params, consts, count);
    if (rc != SPH_OK) {
        delete c;
        return rc;
    }
    *out = c;
    return SPH_OK;
}

int sph_batch_info(const sph_ctx* c, int32_t* count, int32_t* first) {
    if (!c) return fail(SPH_E_INVALID, "null context");
    const int nb = c->inst.empty() ? 1 : (int)c->inst.size();
    if (count) *count = nb;
    if (first) {
        if (c->inst.empty()) { first[0] = 0; first[1] = (int)c->n; }
        else for (int b = 0; b <= nb; ++b) first[b] = c->inst_first[b];
    }
    return SPH_OK;
}

int sph_destroy(sph_ctx* ctx) {
    if (!ctx) return SPH_OK;
    delete[] ctx->rng_host;
    delete ctx;
    return SPH_OK;
}

size_t sph_workspace_bytes(const sph_ctx* ctx) { return ctx ? ctx->ws_bytes : 0; }

int sph_bind_workspace(sph_ctx* c, void* dev, size_t bytes, void* stream) {
    if (!c || !dev) return fail(SPH_E_INVALID, "null argument");
    if (bytes < c->ws_bytes) return fail(SPH_E_INVALID, "workspace of %zu bytes, need %zu", bytes, c->ws_bytes);
    if (((uintptr_t)dev & 255) != 0) return fail(SPH_E_INVALID, "workspace must be 256-byte aligned");
    hipStream_t s = (hipStream_t)stream;
    assign_buffers(c, dev);
    sph::Buffers& b = c->b;
    int rc;
    if ((rc = hip_check(sph::launch_fill_u32(b.flags, 0u, 16, s), "clear flags"))) return rc;
    if (c->cell_sort && (rc = hip_check(sph::launch_fill_u32(b.cell_cnt, 0u, c->sc.cs_cells, s), "clear cell counts")))
        return rc;
    if ((rc = hip_check(hipMemcpyAsync(b.pipe, &c->pipe_host, sizeof(sph::PipeTable), hipMemcpyHostToDevice, s),
                        "upload pipe table")))
        return rc;
    if (c->prm.mode == SPH_MODE_PIPE && c->n > 0)
        if ((rc = hip_check(hipMemcpyAsync(b.rng, c->rng_host, 16 * (size_t)c->n, hipMemcpyHostToDevice, s),
                            "upload rng streams")))
            return rc;
    if (!c->inst.empty()) {
        c->sc.inst_table = b.inst_table;
        c->sc.inst_first = b.inst_first;
        if ((rc = upload_instances(c, s))) return rc;
    }
    if ((rc = hip_check(hipStreamSynchronize(s), "bind"))) return rc;
    c->bound = true;
    c->have_state = false;
    c->adj_seeded = false;
    return SPH_OK;
}

}  // extern "C"

namespace {

// Points the context's buffers at their slices of the device memory at dev (256-byte
// aligned, at least c->ws_bytes).
void assign_buffers(sph_ctx* c, void* dev) {
    Layout L = plan_layout(c);
    char* base = (char*)dev;
    sph::Buffers& b = c->b;
    b.posA = (float4*)(base + L.off[S_POSA]);
    b.velA = (float4*)(base + L.off[S_VELA]);
    b.posB = (float4*)(base + L.off[S_POSB]);
    b.velB = (float4*)(base + L.off[S_VELB]);
    b.pidA = (uint32_t*)(base + L.off[S_PIDA]);
    b.pidB = (uint32_t*)(base + L.off[S_PIDB]);
    b.rhoA = (float*)(base + L.off[S_RHOA]);
    b.keyI = (uint32_t*)(base + L.off[S_KEYI]);
    b.valI = (uint32_t*)(base + L.off[S_VALI]);
    b.keyS[0] = (uint32_t*)(base + L.off[S_KEYS0]);
    b.valS[0] = (uint32_t*)(base + L.off[S_VALS0]);
    b.keyS[1] = (uint32_t*)(base + L.off[S_KEYS1]);
    b.valS[1] = (uint32_t*)(base + L.off[S_VALS1]);
    b.cs = (uint32_t*)(base + L.off[S_CS]);
    b.tile_hist = (uint32_t*)(base + L.off[S_HIST]);
    b.digit_total = (uint32_t*)(base + L.off[S_DTOT]);
    b.rng = (uint64_t*)(base + L.off[S_RNG]);
    b.dbgP = (float4*)(base + L.off[S_DBGP]);
    b.dbgV = (float4*)(base + L.off[S_DBGV]);
    b.dbgF = (float4*)(base + L.off[S_DBGF]);
    b.nbr = (int32_t*)(base + L.off[S_NBR]);
    b.nl = (sph::ListEntry*)(base + L.off[S_NL]);
    b.nn = (uint32_t*)(base + L.off[S_NN]);
    b.diag_partials = (double*)(base + L.off[S_DIAG]);
    b.flags = (uint32_t*)(base + L.off[S_FLAGS]);
    b.pipe = (sph::PipeTable*)(base + L.off[S_PIPE]);
    b.inst_table = (sph::StepConsts*)(base + L.off[S_INST]);
    b.inst_first = (int*)(base + L.off[S_INSTF]);
    b.cotx = (float4*)(base + L.off[S_COTX]);
    b.cotv = (float4*)(base + L.off[S_COTV]);
    b.adj_fb = (float4*)(base + L.off[S_AFB]);
    b.adj_gx = (float4*)(base + L.off[S_AGX]);
    b.adj_gv = (float4*)(base + L.off[S_AGV]);
    b.adj_rt = (float*)(base + L.off[S_ART]);
    b.mu_acc = (float*)(base + L.off[S_MU]);
    b.mu_part = (float2*)(base + L.off[S_MUP]);
    b.sums = c->adjoint ? (float4*)(base + L.off[S_SUMS]) : nullptr;
    b.cell_cnt = (uint32_t*)(base + L.off[S_CCNT]);
    b.cell_part = (uint32_t*)(base + L.off[S_CPART]);
    c->ws = dev;
    c->sorted_keys = b.keyI;
    c->sorted_vals = b.valI;
}

}  // namespace

extern "C" {

int sph_upload_state(sph_ctx* c, const double* pos, const double* vel, void* stream) {
    int rc = ready(c, false);
    if (rc) return rc;
    if (c->n > 0 && (!pos || !vel)) return fail(SPH_E_INVALID, "null state pointer");
    if ((rc = hip_check(sph::launch_upload_f64(c, pos, vel, (hipStream_t)stream), "upload"))) return rc;
    c->have_state = true;
    c->last_stage = 0;
    c->a_in_snapshot = false;
    return SPH_OK;
}

int sph_upload_state_f32(sph_ctx* c, const float* x, const float* y, const float* z, const float* vx, const float* vy,
                         const float* vz, void* stream) {
    int rc = ready(c, false);
    if (rc) return rc;
    if (c->n > 0 && (!x || !y || !z || !vx || !vy || !vz)) return fail(SPH_E_INVALID, "null state pointer");
    if ((rc = hip_check(sph::launch_upload_f32(c, x, y, z, vx, vy, vz, (hipStream_t)stream), "upload"))) return rc;
    c->have_state = true;
    c->last_stage = 0;
    c->a_in_snapshot = false;
    return SPH_OK;
}

int sph_step(sph_ctx* c, int32_t nsteps, void* stream) {
    int rc = ready(c, true);
    if (rc) return rc;
    if (nsteps < 0) return fail(SPH_E_INVALID, "nsteps < 0");
    if (c->last_stage != 0 && c->last_stage != SPH_STAGE_FORCE)
        return fail(SPH_E_STATE, "a staged step is in progress (run it to SPH_STAGE_FORCE first)");
    if ((rc = run_steps(c, (hipStream_t)stream, nsteps))) return rc;
    if (c->check_finite && nsteps > 0) {
        // a stream being captured into a graph cannot be waited on: the check is left to the
        // next call outside the capture (the count keeps accumulating)
        hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
        if ((rc = hip_check(hipStreamIsCapturing((hipStream_t)stream, &st), "capture status"))) return rc;
        if (st == hipStreamCaptureStatusNone) return read_nonfinite(c, (hipStream_t)stream, nullptr);
    }
    return SPH_OK;
}

int sph_step_timed(sph_ctx* c, int32_t nsteps, double* stage_ms, void* stream) {
    int rc = ready(c, true);
    if (rc) return rc;
    if (!stage_ms) return fail(SPH_E_INVALID, "null stage_ms");
    if (nsteps < 1 || nsteps > 100000) return fail(SPH_E_INVALID, "nsteps %d outside [1, 100000]", nsteps);
    if (c->last_stage != 0 && c->last_stage != SPH_STAGE_FORCE)
        return fail(SPH_E_STATE, "a staged step is in progress (run it to SPH_STAGE_FORCE first)");
    hipStream_t s = (hipStream_t)stream;
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if ((rc = hip_check(hipStreamIsCapturing(s, &st), "capture status"))) return rc;
    if (st != hipStreamCaptureStatusNone) return fail(SPH_E_STATE, "sph_step_timed waits for its events: not under capture");
    std::vector<hipEvent_t> ev(4 * (size_t)nsteps, nullptr);
    for (auto& e : ev)
        if ((rc = hip_check(hipEventCreate(&e), "create event"))) break;
    if (rc == SPH_OK) rc = run_steps(c, s, nsteps, ev.data());
    if (rc == SPH_OK) rc = hip_check(hipEventSynchronize(ev.back()), "wait for the timed steps");
    double acc[3] = {0.0, 0.0, 0.0};
    for (int32_t k = 0; k < nsteps && rc == SPH_OK; ++k)
        for (int j = 0; j < 3 && rc == SPH_OK; ++j) {
            float ms = 0.0f;
            rc = hip_check(hipEventElapsedTime(&ms, ev[4 * k + j], ev[4 * k + j + 1]), "stage time");
            acc[j] += ms;
        }
    for (auto e : ev)
        if (e) (void)hipEventDestroy(e);
    if (rc) return rc;
    for (int j = 0; j < 3; ++j) stage_ms[j] = acc[j] / nsteps;
    return SPH_OK;
}

int sph_check_finite(sph_ctx* c, int64_t* count, void* stream) {
    if (count) *count = 0;
    int rc = ready(c, false);
    if (rc) return rc;
    return read_nonfinite(c, (hipStream_t)stream, count);
}

int sph_run_stage(sph_ctx* c, int32_t stage, void* stream) {
    int rc = ready(c, true);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    const int expected = c->last_stage == SPH_STAGE_FORCE ? SPH_STAGE_SORT : c->last_stage + 1;
    if (stage != expected) return fail(SPH_E_STATE, "stage %d out of order (expected %d)", stage, expected);
    switch (stage) {
        case SPH_STAGE_SORT:
            rc = c->cell_sort ? hip_check(sph::launch_cell_sort(c, false, s), "cell sort")
                              : hip_check(sph::launch_radix_sort_impl(c, false, s), "sort");
            break;
        case SPH_STAGE_TABLE:  // the counting sort has built the table already
            rc = c->cell_sort ? SPH_OK : hip_check(sph::launch_cell_table(c, s), "cell table");
            break;
        case SPH_STAGE_REORDER: rc = hip_check(sph::launch_reorder(c, s), "reorder"); break;
        case SPH_STAGE_DENSITY: rc = hip_check(sph::launch_density(c, s), "density"); break;
        case SPH_STAGE_FORCE: rc = hip_check(sph::launch_force(c, s), "force"); break;
        default: return fail(SPH_E_INVALID, "unknown stage %d", stage);
    }
    if (rc) return rc;
    c->last_stage = stage;
    c->a_in_snapshot = false;
    return SPH_OK;
}

int sph_download_state(sph_ctx* c, double* pos, double* vel, double* rho, double* force, void* stream) {
    int rc = ready(c, true);
    if (rc) return rc;
    if (force && !c->keep_terms) return fail(SPH_E_STATE, "result force needs SPH_FLAG_KEEP_TERMS");
    return hip_check(sph::launch_download_f64(c, pos, vel, rho, force, (hipStream_t)stream), "download");
}

int sph_download_state_f32(sph_ctx* c, float* x, float* y, float* z, float* vx, float* vy, float* vz, float* rho,
                           void* stream) {
    int rc = ready(c, true);
    if (rc) return rc;
    return hip_check(sph::launch_download_f32(c, x, y, z, vx, vy, vz, rho, (hipStream_t)stream), "download");
}

int64_t sph_buffer_elems(const sph_ctx* c, int32_t which) {
    if (!c) return -1;
    switch (which) {
        case SPH_BUF_KEYS:
        case SPH_BUF_PERM:
        case SPH_BUF_SORTED_KEYS:
        case SPH_BUF_DENSITY: return c->n;
        case SPH_BUF_CELL_BEGIN: return c->sc.n_cells;
        case SPH_BUF_PRESSURE:
        case SPH_BUF_VISCOSITY:
        case SPH_BUF_FORCE: return 3 * c->n;
        case SPH_BUF_RNG: return 2 * c->n;
        case SPH_BUF_NEIGHBOURS: return 33 * c->n;
        default: return -1;
    }
}

int sph_read_buffer(sph_ctx* c, int32_t which, void* out, void* stream) {
    int rc = ready(c, true);
    if (rc) return rc;
    if (!out) return fail(SPH_E_INVALID, "null output");
    hipStream_t s = (hipStream_t)stream;
    if (c->a_in_snapshot && which != SPH_BUF_KEYS && which != SPH_BUF_RNG && which != SPH_BUF_SORTED_KEYS)
        return fail(SPH_E_STATE, "the last step was run by sph_step_save: its buffers are in the snapshot");
    switch (which) {
        case SPH_BUF_PRESSURE:
        case SPH_BUF_VISCOSITY:
        case SPH_BUF_FORCE:
            if (!c->keep_terms) return fail(SPH_E_STATE, "terms need SPH_FLAG_KEEP_TERMS");
            if (c->last_stage != SPH_STAGE_FORCE) return fail(SPH_E_STATE, "terms exist after the force stage");
            break;
        case SPH_BUF_DENSITY:
            if (c->last_stage < SPH_STAGE_DENSITY) return fail(SPH_E_STATE, "density exists after the density stage");
            break;
        case SPH_BUF_NEIGHBOURS:
            if (c->prm.formulation != SPH_FORM_VOXEL) return fail(SPH_E_STATE, "neighbour lists exist in VOXEL only");
            if (c->last_stage < SPH_STAGE_DENSITY) return fail(SPH_E_STATE, "lists exist after the density stage");
            break;
        case SPH_BUF_PERM:
        case SPH_BUF_SORTED_KEYS:
            if (c->last_stage < SPH_STAGE_SORT) return fail(SPH_E_STATE, "sorted order exists after the sort stage");
            break;
        case SPH_BUF_CELL_BEGIN:
            if (c->last_stage < SPH_STAGE_TABLE) return fail(SPH_E_STATE, "cell table exists after the table stage");
            break;
        case SPH_BUF_KEYS:
            break;
        case SPH_BUF_RNG:
            if (c->prm.mode != SPH_MODE_PIPE) return fail(SPH_E_STATE, "RNG streams exist in PIPE mode only");
            return hip_check(hipMemcpyAsync(out, c->b.rng, 16 * (size_t)c->n, hipMemcpyDeviceToDevice, s), "read rng");
        default:
            return fail(SPH_E_INVALID, "unknown buffer %d", which);
    }
    return hip_check(sph::launch_read_buffer(c, which, out, s), "read buffer");
}

int sph_diagnostics(sph_ctx* c, double* out8, void* stream) {
    int rc = ready(c, true);
    if (rc) return rc;
    if (!out8) return fail(SPH_E_INVALID, "null output");
    if (!c->inst.empty()) return fail(SPH_E_STATE, "diagnostics are per simulation; a batch mixes particle masses");
    return hip_check(sph::launch_diagnostics(c, out8, (hipStream_t)stream), "diagnostics");
}

int sph_grid_info(const sph_ctx* c, int32_t* dims3, int32_t* stencil3, int32_t* key_bits, int32_t* passes) {
    if (!c) return fail(SPH_E_INVALID, "null context");
    for (int d = 0; d < 3; ++d) {
        if (dims3) dims3[d] = c->sc.dims[d];
        if (stencil3) stencil3[d] = c->sc.stencil[d];
    }
    if (key_bits) *key_bits = c->key_bits;
    if (passes) *passes = c->passes;
    return SPH_OK;
}

int sph_set_launch(sph_ctx* c, int32_t density_block, int32_t force_block) {
    if (!c) return fail(SPH_E_INVALID, "null context");
    for (int32_t b : {density_block, force_block})
        if (b != 64 && b != 128 && b != 256) return fail(SPH_E_INVALID, "block size must be 64, 128 or 256");
    c->density_block = density_block;
    c->force_block = force_block;
    return SPH_OK;
}

int sph_set_lanes(sph_ctx* c, int32_t lanes) {
    if (!c) return fail(SPH_E_INVALID, "null context");
    if (!valid_lanes(lanes))
        return fail(SPH_E_INVALID, "lanes per particle must be 1, 2, 4, 8, 16, 32 or 64");
    if (lanes > 1 && c->prm.formulation != SPH_FORM_NAIVE)
        return fail(SPH_E_INVALID, "the VOXEL formulation walks its capped list with one lane per particle");
    if (c->n * (int64_t)lanes >= (int64_t(1) << 31))
        return fail(SPH_E_INVALID, "%lld particles x %d lanes exceed 2^31 threads", (long long)c->n, lanes);
    if (!c->inst.empty()) {
        std::vector<int32_t> all(c->inst.size(), lanes);
        return sph_set_batch_lanes(c, all.data(), (int32_t)all.size(), nullptr);
    }
    c->lanes = lanes;
    return SPH_OK;
}

int sph_set_slabs(sph_ctx* c, int32_t enable, void* stream) {
    if (!c) return fail(SPH_E_INVALID, "null context");
    if (!c->inst.empty()) return fail(SPH_E_STATE, "a batch's cells are always cut into slabs");
    if (c->prm.formulation != SPH_FORM_NAIVE)
        return fail(SPH_E_INVALID, "the VOXEL formulation keeps the reference's cell table");
    if (c->adjoint)
        return fail(SPH_E_STATE, "an SPH_FLAG_ADJOINT context keeps one cell layout (its snapshots hold tables)");
    if (c->last_stage != 0 && c->last_stage != SPH_STAGE_FORCE)
        return fail(SPH_E_STATE, "a staged step is in progress (run it to SPH_STAGE_FORCE first)");
    c->whole_cells = enable ? 0 : 1;
    plan_slabs(c);
    if (c->bound && c->have_state) {
        // the next sort reads keys of the new layout; the last step's table is not readable
        int rc;
        if ((rc = hip_check(sph::launch_rekey(c, (hipStream_t)stream), "rekey"))) return rc;
        c->last_stage = 0;
        c->a_in_snapshot = false;
    }
    return SPH_OK;
}

int sph_set_batch_lanes(sph_ctx* c, const int32_t* lanes, int32_t count, void* stream) {
    if (!c || !lanes) return fail(SPH_E_INVALID, "null argument");
    if (c->inst.empty()) return fail(SPH_E_STATE, "not a batched context (use sph_set_lanes)");
    if (count != (int32_t)c->inst.size())
        return fail(SPH_E_INVALID, "%d lane counts for %zu instances", count, c->inst.size());
    const std::vector<sph::StepConsts> keep = c->inst;
    const int keep_total = c->sc.tfirst;
    int rc = batch_thread_layout(c, std::vector<int>(lanes, lanes + count));
    if (rc == SPH_OK && c->bound) {
        hipStream_t s = (hipStream_t)stream;
        rc = upload_instances(c, s);
        if (rc == SPH_OK) rc = hip_check(hipStreamSynchronize(s), "set lanes");
    }
    if (rc != SPH_OK) {
        c->inst = keep;
        c->sc.tfirst = keep_total;
        return rc;
    }
    c->lanes = lanes[0];
    return SPH_OK;
}

int sph_set_constants(sph_ctx* c, const sph_constants_t* k) {
    if (!c || !k) return fail(SPH_E_INVALID, "null argument");
    if (!c->inst.empty()) return fail(SPH_E_STATE, "constants of a batch are set per instance at creation");
    if (k->h != c->cst.h) return fail(SPH_E_INVALID, "h %g differs from the context's %g (it fixes the grid)", k->h, c->cst.h);
    if (!(k->mass > 0) || !isfinite(k->mass)) return fail(SPH_E_INVALID, "mass must be positive and finite");
    for (double v : {k->rho0, k->visc, k->k, k->damp})
        if (!isfinite(v)) return fail(SPH_E_INVALID, "constants must be finite");
    c->cst = *k;
    sph::StepConsts& S = c->sc;
    S.mass = (float)k->mass;
    S.rho0 = (float)k->rho0;
    S.k = (float)k->k;
    S.visc = (float)k->visc;
    S.damp = (float)k->damp;
    S.w_norm = (float)(k->mass * 315.0 / (64.0 * M_PI * k->h * k->h * k->h));
    return SPH_OK;
}

int sph_adjoint_seed(sph_ctx* c, const float* gx, const float* gv, void* stream) {
    int rc = ready(c, false);
    if (rc) return rc;
    if (!c->adjoint) return fail(SPH_E_STATE, "context created without SPH_FLAG_ADJOINT");
    if ((rc = hip_check(sph::launch_adjoint_set(c, gx, gv, (hipStream_t)stream), "adjoint seed"))) return rc;
    c->adj_seeded = true;
    return SPH_OK;
}

int sph_adjoint_step(sph_ctx* c, void* stream) {
    int rc = ready(c, true);
    if (rc) return rc;
    if (!c->adjoint) return fail(SPH_E_STATE, "context created without SPH_FLAG_ADJOINT");
    if (!c->adj_seeded) return fail(SPH_E_STATE, "no cotangent (call sph_adjoint_seed)");
    if (c->last_stage != 0 && c->last_stage != SPH_STAGE_FORCE)
        return fail(SPH_E_STATE, "a staged step is in progress (run it to SPH_STAGE_FORCE first)");
    hipStream_t s = (hipStream_t)stream;
    if ((rc = run_density_stages(c, s, true))) return rc;
    c->a_in_snapshot = false;
    if ((rc = hip_check(sph::launch_adjoint_pullback(c, c->b, s), "adjoint"))) return rc;
    c->last_stage = SPH_STAGE_FORCE;
    return SPH_OK;
}

int64_t sph_adjoint_snapshot_bytes(const sph_ctx* c) { return c ? sph::adjoint_snapshot_bytes(c) : 0; }

static int check_snapshot(const sph_ctx* c, const void* snapshot, size_t bytes) {
    if (!snapshot || ((uintptr_t)snapshot & 255u)) return fail(SPH_E_INVALID, "snapshot must be 256-byte aligned device memory");
    const int64_t need = sph::adjoint_snapshot_bytes(c);
    if ((int64_t)bytes < need) return fail(SPH_E_INVALID, "snapshot of %zu bytes, %lld needed", bytes, (long long)need);
    return SPH_OK;
}

int sph_adjoint_save(sph_ctx* c, void* snapshot, size_t bytes, void* stream) {
    int rc = ready(c, true);
    if (rc) return rc;
    if (sph::adjoint_snapshot_bytes(c) == 0)
        return fail(SPH_E_STATE, "snapshots need a NAIVE context created with SPH_FLAG_ADJOINT");
    if (c->last_stage != SPH_STAGE_FORCE) return fail(SPH_E_STATE, "no completed step to save (run sph_step first)");
    if (c->a_in_snapshot) return fail(SPH_E_STATE, "the last step was run by sph_step_save: its snapshot is written");
    if ((rc = check_snapshot(c, snapshot, bytes))) return rc;
    return hip_check(sph::launch_adjoint_save(c, snapshot, (hipStream_t)stream), "adjoint save");
}

int sph_step_save(sph_ctx* c, void* snapshot, size_t bytes, void* stream) {
    int rc = ready(c, true);
    if (rc) return rc;
    if (sph::adjoint_snapshot_bytes(c) == 0)
        return fail(SPH_E_STATE, "snapshots need a NAIVE context created with SPH_FLAG_ADJOINT");
    if (c->last_stage != 0 && c->last_stage != SPH_STAGE_FORCE)
        return fail(SPH_E_STATE, "a staged step is in progress (run it to SPH_STAGE_FORCE first)");
    if ((rc = check_snapshot(c, snapshot, bytes))) return rc;
    // the step's sorted state, densities, lists and cell table are written into the
    // snapshot instead of the context's own buffers; everything else is the context's
    const sph::Buffers own = c->b;
    c->b = sph::snapshot_view(c, snapshot);
    rc = run_step_stages(c, (hipStream_t)stream, true);
    c->b = own;
    if (rc) return rc;
    c->a_in_snapshot = true;
    return SPH_OK;
}

int sph_adjoint_step_from(sph_ctx* c, const void* snapshot, size_t bytes, void* stream) {
    int rc = ready(c, false);
    if (rc) return rc;
    if (sph::adjoint_snapshot_bytes(c) == 0)
        return fail(SPH_E_STATE, "snapshots need a NAIVE context created with SPH_FLAG_ADJOINT");
    if (!c->adj_seeded) return fail(SPH_E_STATE, "no cotangent (call sph_adjoint_seed)");
    if (c->last_stage != 0 && c->last_stage != SPH_STAGE_FORCE)
        return fail(SPH_E_STATE, "a staged step is in progress (run it to SPH_STAGE_FORCE first)");
    if ((rc = check_snapshot(c, snapshot, bytes))) return rc;
    hipStream_t s = (hipStream_t)stream;
    if ((rc = hip_check(sph::launch_adjoint_pullback(c, sph::snapshot_view(c, const_cast<void*>(snapshot)), s, true),
                        "adjoint")))
        return rc;
    c->last_stage = SPH_STAGE_FORCE;
    return SPH_OK;
}

int sph_adjoint_read(sph_ctx* c, float* gx, float* gv, float* gmu, void* stream) {
    int rc = ready(c, false);
    if (rc) return rc;
    if (!c->adjoint) return fail(SPH_E_STATE, "context created without SPH_FLAG_ADJOINT");
    if (!c->adj_seeded) return fail(SPH_E_STATE, "no cotangent (call sph_adjoint_seed)");
    return hip_check(sph::launch_adjoint_get(c, gx, gv, gmu, (hipStream_t)stream), "adjoint read");
}
}  // extern "C"

// ---------------------------------------------------------------------------------------
// The step without any device allocation (sph_step_noalloc): the context's workspace is a
// file-scope device array, one per GPU, looked up once per device with
// hipGetSymbolAddress.  Host code never passes the array's name to a kernel or a copy (on
// the host the name stands in for the symbol, it is not a device address); only the
// looked-up address is used.
// ---------------------------------------------------------------------------------------
#include <mutex>

namespace {

constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;  // 64 MiB per GPU
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of this GPU's scratch array (looked up on the first call per device,
// with that device current).
int scratch_address(void** out) {
    int dev = -1;
    int rc;
    if ((rc = hip_check(hipGetDevice(&dev), "current device"))) return rc;
    if (dev < 0 || dev >= kMaxDevices) return fail(SPH_E_INVALID, "device %d beyond the %d supported", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if ((rc = hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "scratch address"))) return rc;
        if (!p || ((uintptr_t)p & 255u)) return fail(SPH_E_HIP, "scratch address %p is not 256-byte aligned", p);
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return SPH_OK;
}

}  // namespace

extern "C" {

size_t sph_noalloc_scratch_bytes(void) { return kNoAllocScratchBytes; }

int sph_step_noalloc(const sph_params_t* params, const sph_constants_t* consts, const double* dev_pos_in,
                     const double* dev_vel_in, double* dev_pos_out, double* dev_vel_out, double* dev_rho_out,
                     int32_t nsteps, void* stream) {
    if (!params || !consts) return fail(SPH_E_INVALID, "null argument");
    if (params->mode != SPH_MODE_BOX)
        return fail(SPH_E_INVALID, "sph_step_noalloc runs BOX mode only (PIPE mode keeps RNG streams across calls)");
    if (params->flags & (SPH_FLAG_ADJOINT | SPH_FLAG_KEEP_TERMS))
        return fail(SPH_E_INVALID, "sph_step_noalloc takes neither SPH_FLAG_ADJOINT nor SPH_FLAG_KEEP_TERMS");
    if (nsteps < 0) return fail(SPH_E_INVALID, "nsteps < 0");
    if (params->particle_count > 0 && (!dev_pos_in || !dev_vel_in || !dev_pos_out || !dev_vel_out))
        return fail(SPH_E_INVALID, "null state pointer");
    // a plain context planned on the stack: a BOX-mode plan allocates no host memory either
    sph_ctx c{};
    c.prm = *params;
    c.prm.pipe = nullptr;
    c.prm.pipe_rows = 0;
    c.cst = *consts;
    int rc = plan(&c);
    if (rc) return rc;
    if (c.ws_bytes > kNoAllocScratchBytes)
        return fail(SPH_E_INVALID,
                    "%lld particles on %d cells need %zu bytes of scratch; sph_step_noalloc has %zu (use sph_step "
                    "with a caller-owned workspace)",
                    (long long)c.n, c.sc.cs_cells, c.ws_bytes, kNoAllocScratchBytes);
    void* base = nullptr;
    if ((rc = scratch_address(&base))) return rc;
    hipStream_t s = (hipStream_t)stream;
    assign_buffers(&c, base);
    // the scratch keeps the previous call's contents: clear what the step needs at zero
    if ((rc = hip_check(sph::launch_fill_u32(c.b.flags, 0u, 16, s), "clear flags"))) return rc;
    if (c.cell_sort && (rc = hip_check(sph::launch_fill_u32(c.b.cell_cnt, 0u, c.sc.cs_cells, s), "clear cell counts")))
        return rc;
    c.bound = true;
    if ((rc = hip_check(sph::launch_upload_f64(&c, dev_pos_in, dev_vel_in, s), "upload"))) return rc;
    c.have_state = true;
    c.last_stage = 0;
    if ((rc = run_steps(&c, s, nsteps))) return rc;
    if ((rc = hip_check(sph::launch_download_f64(&c, dev_pos_out, dev_vel_out, dev_rho_out, nullptr, s), "download")))
        return rc;
    if (c.check_finite && nsteps > 0) {
        hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
        if ((rc = hip_check(hipStreamIsCapturing(s, &st), "capture status"))) return rc;
        if (st == hipStreamCaptureStatusNone) return read_nonfinite(&c, s, nullptr);
    }
    return SPH_OK;
}

}  // extern "C"
