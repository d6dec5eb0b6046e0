// sph_adjoint.hip -- reverse mode of one NAIVE-formulation step in a box (gfx950).
//
// The step (sph_kernels.hip), with rho_i = w_norm sum_j u_ij^3, u = 1 - r^2/h^2,
// p = K (rho - rho0), iota = 1/rho and d = x_i - x_j:
//   P_i = ps sum_{j != i} (p_i + p_j) iota_j g(r) d,   g(r) = (h - r)^2 / r,  ps = m Gc / 2
//   V_i = vs sum_j (h - r) iota_j (v_j - v_i),                              vs = m mu Lc
//   F = f - P + V,  v' = v + F / rho dt,  x' = x + v' dt,  then the box collision.
// Given the cotangent (gx', gv') of (x', v'), the pullback to (x, v) is, per particle k:
//   collision   a bounced component's position is constant, its velocity is scaled by
//               -damp:  gxp = b ? 0 : gx',  gvp = (b ? -damp gv' : gv') + gxp dt
//   integration Fb = gvp dt iota (the cotangent of F, so Pb = -Fb, Vb = Fb),
//               rho_b = -dt iota^2 (gvp . F),  gv = gvp + (viscosity), gx = gxp + ...
// and the pair terms pull back as sums over the same neighbour set, which is symmetric
// (j is a neighbour of k exactly when k is one of j: |d|^2 is computed bitwise alike), so
// every transposed sum is a gather over k's own neighbours and no atomics are needed:
//   gv_k   += vs sum_j (h - r) (iota_k Fb_j - iota_j Fb_k)
//   pbar_k  = ps sum_j g (iota_k Fb_j.d - iota_j Fb_k.d)
//   rho_b_k += -vs iota_k^2 sum_j (h - r) Fb_j.(v_k - v_j) - ps iota_k^2 sum_j (p_j + p_k) g Fb_j.d
//   gx_k   += -vs sum_j (iota_k Fb_j.w - iota_j Fb_k.w) d / r,     w = v_k - v_j
//             + ps sum_j J(d) (p_k + p_j) (iota_k Fb_j - iota_j Fb_k),
//               J(d) = g I + g'(r) / r d d^T,  g'(r) = -(h - r)(h + r) / r^2
//   rt_k    = rho_b_k + K pbar_k                  (total cotangent of rho_k)
//   gx_k   += -6 w_norm / h^2 sum_j (rt_k + rt_j) u^2 d       (density pass)
//   dL/dmu += sum_k Fb_k . V_k / mu
// Kernels, in order, all in the cell order of the step:
//   adjoint_seed_kernel     the force pass itself (same code, so the step's output is
//                           written bitwise as sph_step writes it) plus the collision and
//                           integration pullback; each particle's dL/dmu term is added to
//                           its running sum, an FP32 pair (sum, rounding error) kept exact
//                           to about 2^-48 by compensated addition (indexed by particle id)
//   adjoint_pair_kernel     pressure / viscosity pullback: gv final, rt, gx partial
//   adjoint_density_kernel  density pullback: gx final
// and when the gradient is read, adjoint_mu_total_kernel sums the running sums in a fixed
// order (deterministic).
// The derivation is checked against an FP64 restatement and central finite differences.
#include "sph_device.h"

namespace sph {
namespace {

constexpr int kAdjThreads = 256;

// Compensated FP32 sums (Knuth's two-sum): a value is the unevaluated sum hi + lo of an FP32
// pair, so a long running sum keeps about twice FP32's precision without FP64 arithmetic.
// Nothing here may be reassociated (the build uses no fast-math flags).
__device__ __forceinline__ float2 two_sum(float a, float b) {
    const float s = a + b;
    const float bb = s - a;
    const float err = (a - (s - bb)) + (b - bb);
    return make_float2(s, err);
}
__device__ __forceinline__ float2 two_float_add(float2 acc, float x) {
    const float2 t = two_sum(acc.x, x);
    return two_sum(t.x, acc.y + t.y);  // renormalised
}
__device__ __forceinline__ float2 two_float_sum(float2 a, float2 b) {
    const float2 t = two_sum(a.x, b.x);
    return two_sum(t.x, a.y + b.y + t.y);  // renormalised
}

// Neighbours of cell-ordered particle i (the particle itself included), from its list or,
// for a particle with more neighbours than list slots, by rescanning its candidates.
template <typename F>
__device__ __forceinline__ void walk_neighbours(const StepConsts& C, int i, float4 p, const float4* __restrict__ pos,
                                                const uint32_t* __restrict__ cs, const ListEntry* __restrict__ nl,
                                                const uint32_t* __restrict__ nn, F&& f) {
    const uint32_t cnt = nn[i];
    if (cnt <= (uint32_t)kListCap) {
        for (uint32_t t = 0; t < cnt; ++t) f(list_decode(nl[list_at(C.n, i, t)], i));
    } else {
        for_each_candidate(p.x, p.y, p.z, C, cs, [&](uint32_t k) {
            const float4 q = pos[k];
            const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
            if (dx * dx + dy * dy + dz * dz <= C.h2) f(k);
        });
    }
}

// The same walk with L lanes per particle: lane l visits list slots l, l + L, ... (or every
// L-th candidate of each row past the list); the caller adds the lanes' sums.
template <int L, typename F>
__device__ __forceinline__ void walk_neighbours_lanes(const StepConsts& C, int i, uint32_t l, float4 p,
                                                      const float4* __restrict__ pos, const uint32_t* __restrict__ cs,
                                                      const ListEntry* __restrict__ nl, const uint32_t* __restrict__ nn,
                                                      F&& f) {
    if constexpr (L == 1) {
        walk_neighbours(C, i, p, pos, cs, nl, nn, f);
    } else {
        const uint32_t cnt = nn[i];
        if (cnt <= (uint32_t)kListCap) {
            for (uint32_t t = l; t < cnt; t += L) f(list_decode(nl[list_at(C.n, i, t)], i));
        } else {
            const __amdgpu_buffer_rsrc_t rcs = make_rsrc(cs, ((uint32_t)C.cell_off + (uint32_t)C.n_cells + 1u) * 4u);
            for_each_row_range<L>(p.x, p.y, p.z, C, rcs, l, [&](uint32_t k0, uint32_t k1) {
                for (uint32_t k = k0 + l; k < k1; k += L) {
                    const float4 q = pos[k];
                    const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
                    if (dx * dx + dy * dy + dz * dz <= C.h2) f(k);
                }
            });
        }
    }
}

template <int L>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
    for (int o = 1; o < L; o <<= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ float comp(const Vec3& a, int d) { return d == 0 ? a.x : (d == 1 ? a.y : a.z); }
__device__ __forceinline__ float comp(const float4& a, int d) { return d == 0 ? a.x : (d == 1 ? a.y : a.z); }

// Everything of the seed after the force pass's sums: the step's output (bitwise sph_step's)
// and the collision / integration pullback of particle i.
__device__ __forceinline__ void seed_particle(
    const StepConsts& C, int i, float4 p, float4 v, const PairSums& S, const uint32_t* __restrict__ pidA,
    const float* __restrict__ rhoA, float4* __restrict__ posOut, float4* __restrict__ velOut,
    uint32_t* __restrict__ pidOut, uint32_t* __restrict__ keyI, uint32_t* __restrict__ valI,
    uint32_t* __restrict__ flags, const float4* __restrict__ cotx, const float4* __restrict__ cotv,
    float4* __restrict__ fb, float4* __restrict__ gxs, float4* __restrict__ gvs, float2* __restrict__ mu_part) {
    const float ps = 0.5f * C.mass * C.grad_c;
    const float vs = C.mass * C.visc * C.lap_c;
    const Vec3 P{ps * S.ax, ps * S.ay, ps * S.az}, V{vs * S.bx, vs * S.by, vs * S.bz};
    const float rho = rhoA[i];
    const uint32_t pid = pidA[i];
    finish_particle<SPH_MODE_BOX, true>(C, i, p, v, pid, rho, P, V, posOut, velOut, pidOut, keyI, valI, nullptr,
                                        nullptr, nullptr, nullptr, nullptr, flags);
    const Integrated I = integrate_particle(C, p, v, rho, P, V);
    const float4 gx1 = cotx[pid], gv1 = cotv[pid];
    float gxp[3], gvp[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const bool b = box_bounced(C, comp(I.x, d), d);
        gxp[d] = b ? 0.0f : comp(gx1, d);
        gvp[d] = (b ? -C.damp * comp(gv1, d) : comp(gv1, d)) + gxp[d] * C.dt;
    }
    const float iota = v.w;  // 1/rho from the density pass
    const float s = C.dt * iota;
    const float fbx = gvp[0] * s, fby = gvp[1] * s, fbz = gvp[2] * s;
    const float rb = -s * iota * (gvp[0] * I.f.x + gvp[1] * I.f.y + gvp[2] * I.f.z);
    // dV/dmu = m Lc sum_j (h - r) iota_j (v_j - v_i)
    const float mu_term = C.mass * C.lap_c * (fbx * S.bx + fby * S.by + fbz * S.bz);
    fb[i] = make_float4(fbx, fby, fbz, rb);
    mu_part[pid] = two_float_add(mu_part[pid], mu_term);  // one writer per particle id and step
    gxs[i] = make_float4(gxp[0], gxp[1], gxp[2], 0.0f);
    gvs[i] = make_float4(gvp[0], gvp[1], gvp[2], 0.0f);
}

template <int L>
__global__ __launch_bounds__(kAdjThreads) void adjoint_seed_kernel(
    StepConsts C, const float4* __restrict__ pos, const float4* __restrict__ vel, const uint32_t* __restrict__ pidA,
    const float* __restrict__ rhoA, const uint32_t* __restrict__ cs, const ListEntry* __restrict__ nl,
    const uint32_t* __restrict__ nn, float4* __restrict__ posOut, float4* __restrict__ velOut,
    uint32_t* __restrict__ pidOut, uint32_t* __restrict__ keyI, uint32_t* __restrict__ valI,
    uint32_t* __restrict__ flags, const float4* __restrict__ cotx, const float4* __restrict__ cotv,
    float4* __restrict__ fb, float4* __restrict__ gxs, float4* __restrict__ gvs, float2* __restrict__ mu_part,
    float4* __restrict__ sums) {
    // L lanes per particle, as the step's force pass uses, so the recomputed sums are its own
    const uint32_t gt = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int i = (int)(gt / L);
    if (i >= C.n) return;
    const float4 p = pos[i], v = vel[i];
    const PairSums S = naive_pair_sums_lanes<L>(C, i, gt & (L - 1), C.n, p, v, pos, vel, cs, nl, nn);
    if (L > 1 && (gt & (L - 1)) != 0) return;
    // kept as the force pass keeps them, so that a snapshot saved after this step is whole
    sums[i] = make_float4(S.ax, S.ay, S.az, 0.0f);
    sums[C.n + i] = make_float4(S.bx, S.by, S.bz, 0.0f);
    seed_particle(C, i, p, v, S, pidA, rhoA, posOut, velOut, pidOut, keyI, valI, flags, cotx, cotv, fb, gxs, gvs,
                  mu_part);
}

// The seed from the sums the step's own force pass kept (sums[i], sums[n + i]): the same
// values, without the neighbour walk; one thread per particle.
__global__ __launch_bounds__(kAdjThreads) void adjoint_seed_sums_kernel(
    StepConsts C, const float4* __restrict__ pos, const float4* __restrict__ vel, const uint32_t* __restrict__ pidA,
    const float* __restrict__ rhoA, const float4* __restrict__ sums, float4* __restrict__ posOut,
    float4* __restrict__ velOut, uint32_t* __restrict__ pidOut, uint32_t* __restrict__ keyI,
    uint32_t* __restrict__ valI, uint32_t* __restrict__ flags, const float4* __restrict__ cotx,
    const float4* __restrict__ cotv, float4* __restrict__ fb, float4* __restrict__ gxs, float4* __restrict__ gvs,
    float2* __restrict__ mu_part) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= C.n) return;
    const float4 a = sums[i], b = sums[C.n + i];
    PairSums S;
    S.ax = a.x;
    S.ay = a.y;
    S.az = a.z;
    S.bx = b.x;
    S.by = b.y;
    S.bz = b.z;
    seed_particle(C, i, pos[i], vel[i], S, pidA, rhoA, posOut, velOut, pidOut, keyI, valI, flags, cotx, cotv, fb, gxs,
                  gvs, mu_part);
}

template <int L>
__global__ __launch_bounds__(kAdjThreads) void adjoint_pair_kernel(
    StepConsts C, const float4* __restrict__ pos, const float4* __restrict__ vel, const uint32_t* __restrict__ pidA,
    const uint32_t* __restrict__ cs, const ListEntry* __restrict__ nl, const uint32_t* __restrict__ nn,
    const float4* __restrict__ fb, float4* __restrict__ gxs, const float4* __restrict__ gvs, float* __restrict__ rt,
    float4* __restrict__ cotv) {
    const uint32_t gt = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int k = (int)(gt / L);
    const uint32_t lane = gt & (L - 1);
    if (k >= C.n) return;
    const float4 p = pos[k], v = vel[k], fk = fb[k];
    const float ik = v.w, h = C.h;
    float gvx = 0.f, gvy = 0.f, gvz = 0.f;  // x vs
    float gxx = 0.f, gxy = 0.f, gxz = 0.f;  // final units
    float rbv = 0.f, rbp = 0.f, pb = 0.f;
    const float ps = 0.5f * C.mass * C.grad_c;
    const float vs = C.mass * C.visc * C.lap_c;
    walk_neighbours_lanes<L>(C, k, lane, p, pos, cs, nl, nn, [&](uint32_t j) {
        if (j == (uint32_t)k) return;
        const float4 q = pos[j], w = vel[j], fj = fb[j];
        const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        const float rinv = rsqrtf(r2);
        const float r = r2 * rinv;
        const float hr = h - r;
        const float ij = w.w;
        // viscosity
        gvx += hr * (ik * fj.x - ij * fk.x);
        gvy += hr * (ik * fj.y - ij * fk.y);
        gvz += hr * (ik * fj.z - ij * fk.z);
        const float wx = v.x - w.x, wy = v.y - w.y, wz = v.z - w.z;
        const float fjw = fj.x * wx + fj.y * wy + fj.z * wz;
        const float fkw = fk.x * wx + fk.y * wy + fk.z * wz;
        rbv += hr * fjw;
        const float sv = -vs * (ik * fjw - ij * fkw) * rinv;
        // pressure
        const float g = hr * hr * rinv;
        const float fjd = fj.x * dx + fj.y * dy + fj.z * dz;
        const float fkd = fk.x * dx + fk.y * dy + fk.z * dz;
        const float ppk = p.w + q.w;
        pb += g * (ik * fjd - ij * fkd);
        rbp += ppk * g * fjd;
        const float bx = ppk * (ik * fj.x - ij * fk.x);
        const float by = ppk * (ik * fj.y - ij * fk.y);
        const float bz = ppk * (ik * fj.z - ij * fk.z);
        const float db = dx * bx + dy * by + dz * bz;
        const float gp = -hr * (h + r) * rinv * rinv * rinv * db;
        gxx += sv * dx + ps * (g * bx + gp * dx);
        gxy += sv * dy + ps * (g * by + gp * dy);
        gxz += sv * dz + ps * (g * bz + gp * dz);
    });
    if constexpr (L > 1) {
        gvx = group_sum<L>(gvx);
        gvy = group_sum<L>(gvy);
        gvz = group_sum<L>(gvz);
        gxx = group_sum<L>(gxx);
        gxy = group_sum<L>(gxy);
        gxz = group_sum<L>(gxz);
        rbv = group_sum<L>(rbv);
        rbp = group_sum<L>(rbp);
        pb = group_sum<L>(pb);
        if (lane != 0) return;
    }
    rt[k] = fk.w - vs * ik * ik * rbv - ps * ik * ik * rbp + C.k * ps * pb;
    const float4 gx = gxs[k], gv = gvs[k];
    gxs[k] = make_float4(gx.x + gxx, gx.y + gxy, gx.z + gxz, gx.w);
    cotv[pidA[k]] = make_float4(gv.x + vs * gvx, gv.y + vs * gvy, gv.z + vs * gvz, 0.0f);
}

template <int L>
__global__ __launch_bounds__(kAdjThreads) void adjoint_density_kernel(
    StepConsts C, const float4* __restrict__ pos, const uint32_t* __restrict__ pidA, const uint32_t* __restrict__ cs,
    const ListEntry* __restrict__ nl, const uint32_t* __restrict__ nn, const float* __restrict__ rt,
    const float4* __restrict__ gxs, float4* __restrict__ cotx) {
    const uint32_t gt = (uint32_t)xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int k = (int)(gt / L);
    const uint32_t lane = gt & (L - 1);
    if (k >= C.n) return;
    const float4 p = pos[k];
    const float rk = rt[k];
    float ax = 0.f, ay = 0.f, az = 0.f;
    walk_neighbours_lanes<L>(C, k, lane, p, pos, cs, nl, nn, [&](uint32_t j) {
        const float4 q = pos[j];
        const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
        const float u = fmaxf(1.0f - (dx * dx + dy * dy + dz * dz) * C.inv_h2, 0.0f);
        const float c = (rk + rt[j]) * u * u;
        ax += c * dx;
        ay += c * dy;
        az += c * dz;
    });
    if constexpr (L > 1) {
        ax = group_sum<L>(ax);
        ay = group_sum<L>(ay);
        az = group_sum<L>(az);
        if (lane != 0) return;
    }
    const float s = -6.0f * C.w_norm * C.inv_h2;
    const float4 gx = gxs[k];
    cotx[pidA[k]] = make_float4(gx.x + s * ax, gx.y + s * ay, gx.z + s * az, 0.0f);
}

// dL/dmu: the per-particle sums over the reversed steps (mu_part, compensated FP32 pairs),
// summed by one workgroup in a fixed order (each lane a strided range in order, then a fixed
// tree), every addition compensated, into acc[0] when the gradient is read.
__global__ __launch_bounds__(kAdjThreads) void adjoint_mu_total_kernel(int n, const float2* __restrict__ mu_part,
                                                                        float* __restrict__ acc) {
    __shared__ float2 part[kAdjThreads];
    float2 s = make_float2(0.0f, 0.0f);
    for (int i = threadIdx.x; i < n; i += kAdjThreads) s = two_float_sum(s, mu_part[i]);
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = kAdjThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] = two_float_sum(part[threadIdx.x], part[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) acc[0] = part[0].x + part[0].y;
}

__global__ __launch_bounds__(kAdjThreads) void adjoint_set_kernel(int n, const float* __restrict__ gx,
                                                                   const float* __restrict__ gv,
                                                                   float4* __restrict__ cotx,
                                                                   float4* __restrict__ cotv, float* __restrict__ acc,
                                                                   float2* __restrict__ mu_part) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) acc[0] = 0.0f;
    if (i >= n) return;
    mu_part[i] = make_float2(0.0f, 0.0f);
    const int64_t o = 3 * (int64_t)i;
    cotx[i] = gx ? make_float4(gx[o], gx[o + 1], gx[o + 2], 0.0f) : make_float4(0.f, 0.f, 0.f, 0.f);
    cotv[i] = gv ? make_float4(gv[o], gv[o + 1], gv[o + 2], 0.0f) : make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ __launch_bounds__(kAdjThreads) void adjoint_get_kernel(int n, const float4* __restrict__ cotx,
                                                                   const float4* __restrict__ cotv,
                                                                   const float* __restrict__ acc, float* __restrict__ gx,
                                                                   float* __restrict__ gv, float* __restrict__ gmu) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && gmu) gmu[0] = acc[0];
    if (i >= n) return;
    const int64_t o = 3 * (int64_t)i;
    if (gx) { const float4 a = cotx[i]; gx[o] = a.x; gx[o + 1] = a.y; gx[o + 2] = a.z; }
    if (gv) { const float4 a = cotv[i]; gv[o] = a.x; gv[o + 1] = a.y; gv[o + 2] = a.z; }
}

inline int adj_blocks(int64_t n) { return (int)((n + kAdjThreads - 1) / kAdjThreads); }

}  // namespace

hipError_t launch_adjoint_set(sph_ctx* c, const float* gx, const float* gv, hipStream_t s) {
    const int64_t n = c->n;
    adjoint_set_kernel<<<adj_blocks(n > 0 ? n : 1), kAdjThreads, 0, s>>>((int)n, gx, gv, c->b.cotx, c->b.cotv,
                                                                         c->b.mu_acc, c->b.mu_part);
    return hipGetLastError();
}

hipError_t launch_adjoint_get(sph_ctx* c, float* gx, float* gv, float* gmu, hipStream_t s) {
    const int64_t n = c->n;
    adjoint_mu_total_kernel<<<1, kAdjThreads, 0, s>>>((int)n, c->b.mu_part, c->b.mu_acc);
    adjoint_get_kernel<<<adj_blocks(n > 0 ? n : 1), kAdjThreads, 0, s>>>((int)n, c->b.cotx, c->b.cotv, c->b.mu_acc,
                                                                         gx, gv, gmu);
    return hipGetLastError();
}

// The seed, pair and density pullbacks with L lanes per particle (the step's own count).
template <int L>
void launch_pullback_lanes(sph_ctx* c, const Buffers& b, hipStream_t s, bool from_sums) {
    const int nb = adj_blocks(c->n * L);
    if (from_sums)
        adjoint_seed_sums_kernel<<<adj_blocks(c->n), kAdjThreads, 0, s>>>(
            c->sc, b.posA, b.velA, b.pidA, b.rhoA, b.sums, b.posB, b.velB, b.pidB, b.keyI, b.valI, b.flags, b.cotx,
            b.cotv, b.adj_fb, b.adj_gx, b.adj_gv, b.mu_part);
    else
        adjoint_seed_kernel<L><<<nb, kAdjThreads, 0, s>>>(c->sc, b.posA, b.velA, b.pidA, b.rhoA, b.cs, b.nl, b.nn, b.posB,
                                                       b.velB, b.pidB, b.keyI, b.valI, b.flags, b.cotx, b.cotv,
                                                       b.adj_fb, b.adj_gx, b.adj_gv, b.mu_part, b.sums);
    adjoint_pair_kernel<L><<<nb, kAdjThreads, 0, s>>>(c->sc, b.posA, b.velA, b.pidA, b.cs, b.nl, b.nn, b.adj_fb,
                                                       b.adj_gx, b.adj_gv, b.adj_rt, b.cotv);
    adjoint_density_kernel<L><<<nb, kAdjThreads, 0, s>>>(c->sc, b.posA, b.pidA, b.cs, b.nl, b.nn, b.adj_rt, b.adj_gx,
                                                          b.cotx);
}

namespace {

// Step snapshots: the buffers the pullback reads of the step's forward pass (cell-ordered
// state with p and 1/rho in .w, pid, rho, counts, lists, cell table, the force pass's
// sums), each at a 256-byte
// aligned offset of the snapshot.  A step run with its buffers pointed into a snapshot
// (snapshot_view) fills it directly; sph_adjoint_save copies them there in one launch after
// an ordinary step; the pullback reads them in place.
struct CopyRegions {
    int count;
    uint32_t* dst[8];
    const uint32_t* src[8];
    uint32_t words[8];
};

__global__ __launch_bounds__(256) void copy_regions_kernel(CopyRegions R) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (int r = 0; r < R.count; ++r) {
        const uint32_t w = R.words[r];
        const uint32_t q = w / 4u;  // 16-byte pieces (every region starts 256-byte aligned)
        uint4* d4 = reinterpret_cast<uint4*>(R.dst[r]);
        const uint4* s4 = reinterpret_cast<const uint4*>(R.src[r]);
        for (uint32_t k = t; k < q; k += stride) d4[k] = s4[k];
        for (uint32_t k = 4u * q + t; k < w; k += stride) R.dst[r][k] = R.src[r][k];
    }
}

// Snapshot layout: the eight buffers in this order, each at a 256-byte aligned offset.
constexpr int kSnapParts = 8;

inline size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

void snapshot_sizes(const sph_ctx* c, size_t* bytes) {
    const size_t n = (size_t)c->n;
    bytes[0] = 16 * n;                                     // posA (x, y, z, p)
    bytes[1] = 16 * n;                                     // velA (vx, vy, vz, 1/rho)
    bytes[2] = 4 * n;                                      // pidA
    bytes[3] = 4 * n;                                      // rhoA
    bytes[4] = 4 * n;                                      // nn
    bytes[5] = sizeof(ListEntry) * (size_t)kListCap * n;   // nl
    bytes[6] = 4 * ((size_t)c->sc.cs_cells + 1);           // cs
    bytes[7] = 32 * n;                                     // sums (the force pass's, [2][N])
}

// The eight buffers of b, in layout order.
void snapshot_buffers(const Buffers& b, void** out) {
    out[0] = b.posA;
    out[1] = b.velA;
    out[2] = b.pidA;
    out[3] = b.rhoA;
    out[4] = b.nn;
    out[5] = b.nl;
    out[6] = b.cs;
    out[7] = b.sums;
}

}  // namespace

int64_t adjoint_snapshot_bytes(const sph_ctx* c) {
    if (!c->adjoint || !c->cell_sort) return 0;
    size_t bytes[kSnapParts], total = 0;
    snapshot_sizes(c, bytes);
    for (int i = 0; i < kSnapParts; ++i) total += align256(bytes[i]);
    return (int64_t)total;
}

Buffers snapshot_view(const sph_ctx* c, void* snap) {
    size_t bytes[kSnapParts], off[kSnapParts], at = 0;
    snapshot_sizes(c, bytes);
    for (int i = 0; i < kSnapParts; ++i) {
        off[i] = at;
        at += align256(bytes[i]);
    }
    char* p = static_cast<char*>(snap);
    Buffers v = c->b;
    v.posA = reinterpret_cast<float4*>(p + off[0]);
    v.velA = reinterpret_cast<float4*>(p + off[1]);
    v.pidA = reinterpret_cast<uint32_t*>(p + off[2]);
    v.rhoA = reinterpret_cast<float*>(p + off[3]);
    v.nn = reinterpret_cast<uint32_t*>(p + off[4]);
    v.nl = reinterpret_cast<ListEntry*>(p + off[5]);
    v.cs = reinterpret_cast<uint32_t*>(p + off[6]);
    v.sums = reinterpret_cast<float4*>(p + off[7]);
    return v;
}

hipError_t launch_adjoint_save(sph_ctx* c, void* snap, hipStream_t s) {
    size_t bytes[kSnapParts];
    snapshot_sizes(c, bytes);
    void* own[kSnapParts];
    void* in_snap[kSnapParts];
    snapshot_buffers(c->b, own);
    snapshot_buffers(snapshot_view(c, snap), in_snap);
    CopyRegions R{};
    R.count = kSnapParts;
    size_t most = 0;
    for (int i = 0; i < kSnapParts; ++i) {
        R.dst[i] = static_cast<uint32_t*>(in_snap[i]);
        R.src[i] = static_cast<const uint32_t*>(own[i]);
        R.words[i] = (uint32_t)(bytes[i] / 4);
        most = bytes[i] > most ? bytes[i] : most;
    }
    const size_t want = (most / 16 + 255) / 256 + 1;
    const int blocks = (int)(want < 2048 ? want : 2048);
    copy_regions_kernel<<<blocks, 256, 0, s>>>(R);
    return hipGetLastError();
}

// After sort, cell table and density of the step: the seed, pair, density and mu kernels.
hipError_t launch_adjoint_pullback(sph_ctx* c, const Buffers& b, hipStream_t s, bool from_sums) {
    const int64_t n = c->n;
    if (n == 0) return hipGetLastError();
    switch (c->lanes) {
        case 2: launch_pullback_lanes<2>(c, b, s, from_sums); break;
        case 4: launch_pullback_lanes<4>(c, b, s, from_sums); break;
        case 8: launch_pullback_lanes<8>(c, b, s, from_sums); break;
        case 16: launch_pullback_lanes<16>(c, b, s, from_sums); break;
        case 32: launch_pullback_lanes<32>(c, b, s, from_sums); break;
        case 64: launch_pullback_lanes<64>(c, b, s, from_sums); break;
        default: launch_pullback_lanes<1>(c, b, s, from_sums); break;
    }
    return hipGetLastError();
}

}  // namespace sph
