/*
 * sph_oracle.c -- CPU FP64 restatement of the reference SPH step (iwoplaza/cuda-sph).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by nothing else.
 * The product path (cuda-sph_amd/) never links, loads or calls it, and never falls
 * back to it.
 *
 * Every function restates the behaviour of one reference function, cited as
 * file:line relative to the reference repository root.  Arithmetic is FP64 with the
 * reference's operation order, so that results equal the reference's up to the
 * rounding of transcendental calls.  Nothing here is tuned for speed except the
 * optional grid search (oc_grid_*), which computes the NAIVE formulation's sums over
 * exactly the same neighbour set as the O(N^2) loops and is used as the CPU
 * throughput baseline at large N.
 *
 * Pinning (what checks each function, besides the code itself):
 *   geometry helpers        known answers of sim/tests/test_collisions.py:90-182
 *   pipe table              known answers of common/tests/test_pipe.py:12-34
 *   solve_collision         post-condition of sim/tests/test_collisions.py:184-209, closed-form
 *                           cylinder reflection, and on cones (tests/test_oracle_pins.py): the
 *                           backtrack d / (sin a + 0.001) with d from plane geometry, |v'| = |v|,
 *                           v' + v along the wall line, re-advance by the backtracked distance
 *   xoroshiro128+/splitmix  published SplitMix64 vectors and a GF(2) matrix-power check
 *                           of the jump polynomial (tests/test_oracle_rng.py)
 *   kernels W, grad W, lap W  normalisation integral, finite differences of the spiky and
 *                           viscosity kernels (Mueller et al. 2003) (tests/test_oracle_physics.py)
 *   naive rho / P / V       independent NumPy brute force, pair identities, invariances
 *   voxel keys / sort       closed form (trunc, NumPy stable argsort)
 *   voxel formulation       a hand-worked three-particle chain (rho without self, pressure with
 *                           no mass, clipped viscosity), pair antisymmetry (sum P = sum V = 0
 *                           with complete lists), mass scaling, an isolated particle going
 *                           non-finite, and the 32-cap order on a hand-built layout whose list
 *                           is written out literally (tests/test_oracle_pins.py); a second
 *                           restatement of the enumeration in plain Python on random inputs
 *   rollout adjoint         directional central differences of this file's own forward map
 *                           (tests/test_oracle_adjoint.py), wall bounces replayed
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define OC_API __attribute__((visibility("default")))

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct {
    double mass;  /* config.py:18 MASS  */
    double rho0;  /* config.py:19 RHO_0 */
    double h;     /* config.py:20 INF_R */
    double visc;  /* config.py:21 VISC  */
    double k;     /* config.py:22 K     */
    double damp;  /* config.py:23 DAMP  */
} oc_consts_t;

typedef struct {
    double h2, w_const, grad_w_const, lap_w_const;
} oc_derived_t;

/* config.py:24-29: INF_R_2, INF_R_6, INF_R_9, W_CONST, GRAD_W_CONST, LAP_W_CONST */
static oc_derived_t derive(const oc_consts_t* c) {
    oc_derived_t d;
    double h6 = pow(c->h, 6.0), h9 = pow(c->h, 9.0);
    d.h2 = pow(c->h, 2.0);
    d.w_const = 315.0 / (64.0 * M_PI * h9);
    d.grad_w_const = -45.0 / (M_PI * h6);
    d.lap_w_const = 45.0 / (M_PI * h6);
    return d;
}

OC_API void oc_derived_constants(const oc_consts_t* c, double* out4) {
    oc_derived_t d = derive(c);
    out4[0] = d.h2; out4[1] = d.w_const; out4[2] = d.grad_w_const; out4[3] = d.lap_w_const;
}

/* ------------------------------------------------------------------------------------
 * Vector helpers: sim/src/sph/kernels/util_kernels.py
 * ---------------------------------------------------------------------------------- */

/* util_kernels.py:21-28 norm_squared: sequential sum of squared differences */
static inline double norm_squared(const double* a, const double* b) {
    double res = 0.0;
    for (int d = 0; d < 3; ++d) { double t = a[d] - b[d]; res += t * t; }
    return res;
}

/* util_kernels.py:31-35 norm */
static inline double norm3(const double* a, const double* b) { return sqrt(norm_squared(a, b)); }

/* util_kernels.py:97-102 calc_vector_length */
OC_API double oc_vector_length(const double* v) {
    double len = 0.0;
    for (int i = 0; i < 3; ++i) len += v[i] * v[i];
    return sqrt(len);
}

/* util_kernels.py:105-110 calc_distance_between_points */
OC_API double oc_distance_between_points(const double* a, const double* b) {
    double len = 0.0;
    for (int i = 0; i < 3; ++i) { double t = a[i] - b[i]; len += t * t; }
    return sqrt(len);
}

/* util_kernels.py:80-85 calc_cos_between_vectors */
static double cos_between(const double* a, const double* b, double la, double lb) {
    double s = 0.0;
    for (int d = 0; d < 3; ++d) s = s + a[d] * b[d];
    return s / (la * lb);
}

/* util_kernels.py:138-142 calc_vector_scalar (a cross product, result = first x second) */
static void cross3(double* r, const double* a, const double* b) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

/* ------------------------------------------------------------------------------------
 * Pipe geometry: util_kernels.py:38-204.  The pipe table has one row per segment
 * [x, y, z, start_radius, length] plus a closing row (common/data_classes.py:34-44).
 * ---------------------------------------------------------------------------------- */
#define P(r, c) pipe[(size_t)(r) * 5 + (c)]

/* util_kernels.py:38-43 calc_x_at_segment_beginning */
OC_API double oc_x_at_segment_beginning(const double* pipe, int seg) {
    double x = P(0, 0);
    for (int i = 0; i < seg; ++i) x += P(i, 4);
    return x;
}

/* util_kernels.py:56-69 calc_radius_in_position: linear between the start radii of
 * rows seg and seg+1 (written through a truncated-cone length). */
OC_API double oc_radius_in_position(const double* pos, const double* pipe, int seg) {
    double r0 = P(seg, 3), r1 = P(seg + 1, 3);
    if (r0 == r1) return r0;
    if (r0 < r1) {
        double delta = pos[0] - oc_x_at_segment_beginning(pipe, seg);
        double tl = P(seg, 4) * r0 / (r1 - r0);
        return r0 * (1.0 + delta / tl);
    }
    double delta = oc_x_at_segment_beginning(pipe, seg) + P(seg, 4) - pos[0];
    double tl = P(seg, 4) * r1 / (r0 - r1);
    return r1 * (1.0 + delta / tl);
}

/* util_kernels.py:46-53 is_out_of_pipe */
OC_API int oc_is_out_of_pipe(const double* pos, const double* pipe, int seg) {
    double yn = pos[1] - P(seg, 1), zn = pos[2] - P(seg, 2);
    double hh = sqrt(yn * yn + zn * zn);
    return hh > oc_radius_in_position(pos, pipe, seg);
}

/* util_kernels.py:72-77 find_segment: row j with pipe[j][0] <= x < pipe[j+1][0], else -1 */
OC_API int oc_find_segment(const double* pos, const double* pipe, int rows) {
    for (int j = 0; j < rows - 1; ++j)
        if (P(j, 0) <= pos[0] && pos[0] < P(j + 1, 0)) return j;
    return -1;
}

/* util_kernels.py:152-171 calc_edge_vector */
static void edge_vector(double* edge, double* first_l, const double* pos, const double* pipe, int seg) {
    double sr = P(seg, 3), er = P(seg + 1, 3);
    double yn = pos[1] - P(seg, 1), zn = pos[2] - P(seg, 2);
    double hh = sqrt(yn * yn + zn * zn);
    double second_l[3];
    first_l[0] = P(seg, 0);
    second_l[0] = P(seg + 1, 0);
    for (int d = 1; d < 3; ++d) {
        first_l[d] = (pos[d] - P(seg, d)) / hh * sr;
        second_l[d] = (pos[d] - P(seg, d)) / hh * er;
    }
    for (int d = 0; d < 3; ++d) edge[d] = second_l[d] - first_l[d];
    for (int d = 1; d < 3; ++d) first_l[d] = first_l[d] + P(seg, d);
}

/* util_kernels.py:124-135 calc_distance_to_pipe */
static double distance_to_pipe(const double* point, const double* edge, const double* l_point) {
    double pv[3], vs[3];
    for (int d = 0; d < 3; ++d) pv[d] = l_point[d] - point[d];
    cross3(vs, edge, pv);
    return oc_vector_length(vs) / oc_vector_length(edge);
}

/* util_kernels.py:113-121 calc_dt */
static double calc_dt(const double* pos, const double* speed, const double* edge, const double* l_point) {
    double sl = oc_vector_length(speed);
    double el = oc_vector_length(edge);
    double cos_a = cos_between(speed, edge, sl, el);
    double d = distance_to_pipe(pos, edge, l_point);
    double sin_a = sqrt(1.0 - cos_a * cos_a);
    double dtc = d / (sin_a + 0.001);
    return -dtc / sl;
}

/* util_kernels.py:88-94 calc_summary_vector */
static void summary_vector(const double* l, const double* p, double* out) {
    double ll = oc_vector_length(l), pl = oc_vector_length(p);
    double c = cos_between(l, p, ll, pl);
    for (int d = 0; d < 3; ++d) out[d] = 2.0 * l[d] / ll * c * pl;
}

/* util_kernels.py:182-204 solve_collision (with calc_collision_point :174-179) */
OC_API void oc_solve_collision(double* pos, double* speed, const double* pipe, int seg) {
    double edge[3], first_l[3], cp[3];
    edge_vector(edge, first_l, pos, pipe, seg);
    double dt = calc_dt(pos, speed, edge, first_l);
    for (int d = 0; d < 3; ++d) cp[d] = pos[d] + speed[d] * dt;
    if (P(seg, 3) == P(seg + 1, 3)) {
        for (int d = 1; d < 3; ++d) speed[d] = -speed[d];
    } else {
        double sv[3];
        summary_vector(edge, speed, sv);
        for (int d = 0; d < 3; ++d) speed[d] = sv[d] - speed[d];
    }
    double way = oc_distance_between_points(pos, cp);
    double svl = oc_vector_length(speed);
    for (int d = 0; d < 3; ++d) pos[d] = cp[d] + speed[d] * way / svl;
}

/* ------------------------------------------------------------------------------------
 * xoroshiro128+ as used by the reference's collision RNG (abstract_sph_strategy.py:27,
 * base_kernels.py:101-112).  The third-party implementation is numba 0.54.1
 * (sim/requirements.txt:2), numba/cuda/random.py; it follows the published
 * Blackman-Vigna xoroshiro128+ (rotations 55/14/36, jump polynomial
 * 0xbeac0467eba5facb, 0xd86b048b86aa9922) seeded through one SplitMix64 step with
 * s0 = s1, stream i = stream i-1 jumped by 2^64, uniform = (x >> 11) * 2^-53.
 * ---------------------------------------------------------------------------------- */
static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

OC_API uint64_t oc_splitmix64(uint64_t* x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

OC_API uint64_t oc_rng_next(uint64_t* st) {
    uint64_t s0 = st[0], s1 = st[1];
    uint64_t result = s0 + s1;
    s1 ^= s0;
    st[0] = rotl64(s0, 55) ^ s1 ^ (s1 << 14);
    st[1] = rotl64(s1, 36);
    return result;
}

OC_API void oc_rng_jump(uint64_t* st) {
    static const uint64_t J[2] = {0xbeac0467eba5facbull, 0xd86b048b86aa9922ull};
    uint64_t s0 = 0, s1 = 0;
    for (int i = 0; i < 2; ++i)
        for (int b = 0; b < 64; ++b) {
            if (J[i] & (1ull << b)) { s0 ^= st[0]; s1 ^= st[1]; }
            oc_rng_next(st);
        }
    st[0] = s0; st[1] = s1;
}

OC_API double oc_rng_uniform(uint64_t* st) {
    return (double)(oc_rng_next(st) >> 11) * (1.0 / 9007199254740992.0);
}

/* n states (2 x u64 each): state 0 = splitmix(seed) in both words, state i = jump(state i-1) */
OC_API void oc_rng_init(uint64_t seed, int64_t n, uint64_t* states) {
    if (n <= 0) return;
    uint64_t x = seed;
    uint64_t z = oc_splitmix64(&x);
    states[0] = z; states[1] = z;
    for (int64_t i = 1; i < n; ++i) {
        states[2 * i] = states[2 * i - 2];
        states[2 * i + 1] = states[2 * i - 1];
        oc_rng_jump(states + 2 * i);
    }
}

/* base_kernels.py:101-112 rand_position_inside_pipe */
static void rand_position_inside_pipe(double* pos, const double* pipe, int rows, uint64_t* st) {
    int seg = oc_find_segment(pos, pipe, rows);
    /* x = 0 lies in no segment only when the pipe starts at x > 0; the reference then
     * indexes pipe[-1] and pipe[rows] (past the table).  Segment 0 is used instead. */
    if (seg < 0) seg = 0;
    double R = oc_radius_in_position(pos, pipe, seg);
    double r = R * sqrt(oc_rng_uniform(st));
    double theta = oc_rng_uniform(st) * 2.0 * M_PI;
    pos[1] = P(0, 1) + r * cos(theta);
    pos[2] = P(0, 2) + r * sin(theta);
}

/* base_kernels.py:115-127 put_particle_at_pipe_begin */
static void put_particle_at_pipe_begin(double* pos, double* vel, const double* pipe, int rows, uint64_t* st) {
    if (pos[0] < 0) {
        pos[0] = -pos[0];
        vel[0] = -vel[0];
        if (oc_is_out_of_pipe(pos, pipe, 0)) oc_solve_collision(pos, vel, pipe, 0);
    } else {
        pos[0] = 0.0;
        rand_position_inside_pipe(pos, pipe, rows, st);
    }
}
#undef P

/* base_kernels.py:56-72 collision_kernel, one thread per particle, state i uses RNG stream i */
OC_API void oc_collide_pipe(double* pos, double* vel, int64_t n, const double* pipe, int rows, uint64_t* states) {
    for (int64_t i = 0; i < n; ++i) {
        double* p = pos + 3 * i;
        double* v = vel + 3 * i;
        int seg = oc_find_segment(p, pipe, rows);
        if (seg == -1) put_particle_at_pipe_begin(p, v, pipe, rows, states + 2 * i);
        else if (oc_is_out_of_pipe(p, pipe, seg)) oc_solve_collision(p, v, pipe, seg);
    }
}

/* base_kernels.py:75-98 collision_kernel_box */
OC_API void oc_collide_box(double* pos, double* vel, int64_t n, const double* space, double damp) {
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            double* x = pos + 3 * i + d;
            int bounced = 0;
            if (*x < 0) { *x = 1e-3; bounced = 1; }
            if (*x > space[d]) { *x = space[d] - 1e-3; bounced = 1; }
            if (bounced) { vel[3 * i + d] *= -1; vel[3 * i + d] *= damp; }
        }
}

/* base_kernels.py:30-53 integrating_kernel (symplectic Euler, f_ext is a force density) */
OC_API void oc_integrate(double* force, double* pos, double* vel, const double* ext, const double* rho,
                         const double* Pt, const double* Vt, int64_t n, double dt) {
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            size_t k = 3 * i + d;
            force[k] = ext[d] + -Pt[k] + Vt[k];
            vel[k] += force[k] / rho[i] * dt;
            pos[k] += vel[k] * dt;
        }
}

/* ------------------------------------------------------------------------------------
 * SPH kernels: base_kernels.py:6-27.  Callers filter r^2 <= h^2 before calling.
 * ---------------------------------------------------------------------------------- */
static inline double compute_w(const oc_derived_t* D, const double* pi, const double* pj) {
    double t = D->h2 - norm_squared(pi, pj);
    return D->w_const * (t * t * t);
}
static inline void compute_grad_w(const oc_consts_t* C, const oc_derived_t* D, const double* pi,
                                  const double* pj, double* out) {
    double dist = norm3(pi, pj);
    double hr = C->h - dist;
    double factor = D->grad_w_const * (hr * hr);
    for (int d = 0; d < 3; ++d) out[d] = factor * (pi[d] - pj[d]) / dist;
}
static inline double compute_lap_w(const oc_consts_t* C, const oc_derived_t* D, const double* pi, const double* pj) {
    return D->lap_w_const * (C->h - norm3(pi, pj));
}

OC_API double oc_w(const oc_consts_t* c, const double* pi, const double* pj) {
    oc_derived_t D = derive(c);
    return compute_w(&D, pi, pj);
}
OC_API void oc_grad_w(const oc_consts_t* c, const double* pi, const double* pj, double* out) {
    oc_derived_t D = derive(c);
    compute_grad_w(c, &D, pi, pj, out);
}
OC_API double oc_lap_w(const oc_consts_t* c, const double* pi, const double* pj) {
    oc_derived_t D = derive(c);
    return compute_lap_w(c, &D, pi, pj);
}

/* ------------------------------------------------------------------------------------
 * NAIVE formulation, O(N^2): sim/src/sph/kernels/naive_kernels.py
 * ---------------------------------------------------------------------------------- */

/* naive_kernels.py:7-21 density_kernel: all j with r^2 <= h^2, self included */
static double naive_density_i(const double* pos, int64_t n, int64_t i, const oc_consts_t* C, const oc_derived_t* D) {
    double acc = 0.0;
    for (int64_t j = 0; j < n; ++j) {
        if (norm_squared(pos + 3 * i, pos + 3 * j) > D->h2) continue;
        acc += compute_w(D, pos + 3 * i, pos + 3 * j);
    }
    return acc * C->mass;
}

OC_API void oc_naive_density(const double* pos, int64_t n, const oc_consts_t* c, double* rho) {
    oc_derived_t D = derive(c);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) rho[i] = naive_density_i(pos, n, i, c, &D);
}

/* naive_kernels.py:24-52 pressure_kernel */
static void naive_pressure_i(const double* pos, const double* rho, int64_t n, int64_t i, const oc_consts_t* C,
                             const oc_derived_t* D, double* out) {
    double acc[3] = {0.0, 0.0, 0.0};
    for (int64_t j = 0; j < n; ++j) {
        if (j == i || norm_squared(pos + 3 * i, pos + 3 * j) > D->h2) continue;
        double pi = C->k * (rho[i] - C->rho0);
        double pj = C->k * (rho[j] - C->rho0);
        double factor = (pi + pj) / (2 * rho[j]);
        double g[3];
        compute_grad_w(C, D, pos + 3 * i, pos + 3 * j, g);
        for (int d = 0; d < 3; ++d) acc[d] += factor * g[d];
    }
    for (int d = 0; d < 3; ++d) out[d] = acc[d] * C->mass;
}

OC_API void oc_naive_pressure(const double* pos, const double* rho, int64_t n, const oc_consts_t* c, double* out) {
    oc_derived_t D = derive(c);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) naive_pressure_i(pos, rho, n, i, c, &D, out + 3 * i);
}

/* naive_kernels.py:55-80 viscosity_kernel */
static void naive_viscosity_i(const double* pos, const double* vel, const double* rho, int64_t n, int64_t i,
                              const oc_consts_t* C, const oc_derived_t* D, double* out) {
    double acc[3] = {0.0, 0.0, 0.0};
    for (int64_t j = 0; j < n; ++j) {
        if (j == i || norm_squared(pos + 3 * i, pos + 3 * j) > D->h2) continue;
        double lap = compute_lap_w(C, D, pos + 3 * i, pos + 3 * j);
        for (int d = 0; d < 3; ++d) {
            double vt = (vel[3 * j + d] - vel[3 * i + d]) / rho[j] * lap * C->visc;
            acc[d] += vt * C->mass;
        }
    }
    for (int d = 0; d < 3; ++d) out[d] = acc[d];
}

OC_API void oc_naive_viscosity(const double* pos, const double* vel, const double* rho, int64_t n,
                               const oc_consts_t* c, double* out) {
    oc_derived_t D = derive(c);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) naive_viscosity_i(pos, vel, rho, n, i, c, &D, out + 3 * i);
}

/* ------------------------------------------------------------------------------------
 * Voxel keys, sorted map and voxel_begin: voxel_kernels.py:9-17, 88-105 and
 * voxel_sph_strategy.py:64-116.
 * ---------------------------------------------------------------------------------- */

/* numba's np.int32(float64) truncates toward zero; out-of-range and NaN are undefined in
 * the reference.  Here they saturate (NaN -> 0), which is what the gfx950 cvt does. */
static inline int32_t trunc_i32(double q) {
    if (!(q == q)) return 0;
    if (q >= 2147483647.0) return 2147483647;
    if (q <= -2147483648.0) return (int32_t)0x80000000u;
    return (int32_t)q;
}

/* voxel_sph_strategy.py:70-73 (ceil) and :109-116 (trunc) */
OC_API void oc_space_dims(const double* space, const double* voxel, int use_ceil, int32_t* dims) {
    for (int d = 0; d < 3; ++d) {
        double q = space[d] / voxel[d];
        dims[d] = use_ceil ? (int32_t)ceil(q) : trunc_i32(q);
    }
}

/* voxel_kernels.py:9-17 + :88-105 assign_voxels_to_particles_kernel (wraps like int32) */
OC_API void oc_voxel_keys(const double* pos, int64_t n, const double* voxel, const int32_t* dims, int32_t* keys) {
    for (int64_t i = 0; i < n; ++i) {
        int32_t v[3];
        for (int d = 0; d < 3; ++d) v[d] = trunc_i32(pos[3 * i + d] / voxel[d]);
        uint32_t k = (uint32_t)v[0] + (uint32_t)v[1] * (uint32_t)dims[0] +
                     (uint32_t)v[2] * (uint32_t)dims[0] * (uint32_t)dims[1];
        keys[i] = (int32_t)k;
    }
}

/* voxel_sph_strategy.py:84-88: structured sort on voxel_id, ties by particle_id.
 * Counting by key value is not possible for arbitrary int32 keys, so this is a
 * stable merge sort of indices. */
static void merge_sort_idx(int32_t* idx, int32_t* tmp, const int32_t* keys, int64_t n) {
    for (int64_t w = 1; w < n; w *= 2) {
        for (int64_t lo = 0; lo < n; lo += 2 * w) {
            int64_t mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
            int64_t a = lo, b = mid, o = lo;
            while (a < mid && b < hi) {
                if (keys[idx[b]] < keys[idx[a]]) tmp[o++] = idx[b++];
                else tmp[o++] = idx[a++];
            }
            while (a < mid) tmp[o++] = idx[a++];
            while (b < hi) tmp[o++] = idx[b++];
        }
        memcpy(idx, tmp, sizeof(int32_t) * (size_t)n);
    }
}

OC_API void oc_sort_map(const int32_t* keys, int64_t n, int32_t* perm) {
    int32_t* tmp = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; ++i) perm[i] = (int32_t)i;
    merge_sort_idx(perm, tmp, keys, n);
    free(tmp);
}

/* voxel_sph_strategy.py:92-107 __populate_voxel_begins: first map index per voxel, -1 if empty */
OC_API void oc_voxel_begin(const int32_t* sorted_keys, int64_t n, int64_t n_voxels, int32_t* begin) {
    for (int64_t v = 0; v < n_voxels; ++v) begin[v] = -1;
    int64_t m = 0;
    for (int64_t v = 0; v < n_voxels; ++v) {
        while (m < n && sorted_keys[m] < v) ++m;
        if (m >= n) return;
        if (sorted_keys[m] == v) begin[v] = (int32_t)m;
    }
}

/* ------------------------------------------------------------------------------------
 * VOXEL formulation: voxel_kernels.py:20-85 and :108-211.  The neighbour list is the
 * first 32 accepted candidates (self included) over the 27 cells in (dx, dy, dz) order
 * with dz innermost, each cell in sorted-map order.  Out-of-range neighbour cells are
 * skipped (the reference leaves those slots of neigh_voxels uninitialised).
 * ---------------------------------------------------------------------------------- */
#define OC_MAX_NEIGHBOURS 32

typedef struct {
    const double* pos;
    const int32_t* begin;   /* n_voxels entries (ceil dims product) */
    const int32_t* map_pid; /* particle id per sorted map entry */
    int64_t n, n_voxels;
    double voxel[3];
    int32_t dims[3];        /* trunc dims used for lookups */
    double h;
} oc_voxel_ctx;

static int voxel_neighbours(const oc_voxel_ctx* X, int64_t p, int32_t* out) {
    int32_t v[3];
    for (int d = 0; d < 3; ++d) v[d] = trunc_i32(X->pos[3 * p + d] / X->voxel[d]);
    int64_t cells[27];
    int nc = 0;
    for (int dx = -1; dx < 2; ++dx)
        for (int dy = -1; dy < 2; ++dy)
            for (int dz = -1; dz < 2; ++dz) {
                int32_t nv[3] = {v[0] + dx, v[1] + dy, v[2] + dz};
                int in = 1;
                for (int d = 0; d < 3; ++d)
                    if (nv[d] < 0 || nv[d] >= X->dims[d]) { in = 0; break; }
                if (in) cells[nc++] = (int64_t)nv[0] + (int64_t)nv[1] * X->dims[0] +
                                      (int64_t)nv[2] * X->dims[0] * X->dims[1];
            }
    int k = 0;
    for (int c = 0; c < nc; ++c) {
        int64_t ci = cells[c];
        if (ci < 0 || ci >= X->n_voxels || X->begin[ci] == -1) continue;
        int64_t start = X->begin[ci];
        int64_t nx = ci + 1;
        while (nx < X->n_voxels && X->begin[nx] == -1) ++nx;
        int64_t end = nx == X->n_voxels ? X->n : X->begin[nx];
        for (int64_t e = start; e < end; ++e) {
            int32_t q = X->map_pid[e];
            const double* a = X->pos + 3 * p;
            const double* b = X->pos + 3 * (int64_t)q;
            double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
            if (sqrt(dx * dx + dy * dy + dz * dz) <= X->h) {
                out[k++] = q;
                if (k >= OC_MAX_NEIGHBOURS) return k;
            }
        }
    }
    return k;
}

OC_API int oc_voxel_neighbours(const double* pos, int64_t n, const int32_t* begin, int64_t n_voxels,
                               const int32_t* map_pid, const double* voxel, const int32_t* dims, double h,
                               int64_t p, int32_t* out) {
    oc_voxel_ctx X = {pos, begin, map_pid, n, n_voxels, {voxel[0], voxel[1], voxel[2]}, {dims[0], dims[1], dims[2]}, h};
    return voxel_neighbours(&X, p, out);
}

/* voxel_kernels.py:108-211: density (self skipped), pressure (pi/rhoi^2 + pj/rhoj^2, no
 * mass), viscosity (m mu / rho_i, then min(1, .) per component).  nb/nn are the
 * per-particle neighbour lists (N x 32) and counts. */
OC_API void oc_voxel_forces(const double* pos, const double* vel, int64_t n, const int32_t* nb, const int32_t* nn,
                            const oc_consts_t* C, double* rho, double* Pt, double* Vt) {
    oc_derived_t D = derive(C);
    for (int64_t i = 0; i < n; ++i) {
        double acc = 0.0;
        for (int a = 0; a < nn[i]; ++a) {
            int32_t j = nb[i * OC_MAX_NEIGHBOURS + a];
            if (j == i) continue;
            acc += compute_w(&D, pos + 3 * i, pos + 3 * (int64_t)j);
        }
        rho[i] = acc * C->mass;
    }
    for (int64_t i = 0; i < n; ++i) {
        double acc[3] = {0, 0, 0};
        for (int a = 0; a < nn[i]; ++a) {
            int32_t j = nb[i * OC_MAX_NEIGHBOURS + a];
            if (j == i) continue;
            double pi = C->k * (rho[i] - C->rho0);
            double pj = C->k * (rho[j] - C->rho0);
            double factor = pi / (rho[i] * rho[i]) + pj / (rho[j] * rho[j]);
            double g[3];
            compute_grad_w(C, &D, pos + 3 * i, pos + 3 * (int64_t)j, g);
            for (int d = 0; d < 3; ++d) acc[d] += factor * g[d];
        }
        for (int d = 0; d < 3; ++d) Pt[3 * i + d] = acc[d];
    }
    for (int64_t i = 0; i < n; ++i) {
        double acc[3] = {0, 0, 0};
        for (int a = 0; a < nn[i]; ++a) {
            int32_t j = nb[i * OC_MAX_NEIGHBOURS + a];
            if (j == i) continue;
            double lap = compute_lap_w(C, &D, pos + 3 * i, pos + 3 * (int64_t)j);
            for (int d = 0; d < 3; ++d) {
                double t = (vel[3 * (int64_t)j + d] - vel[3 * i + d]) / rho[j] * lap;
                acc[d] += t * C->mass * C->visc / rho[i];
            }
        }
        for (int d = 0; d < 3; ++d) Vt[3 * i + d] = acc[d] < 1.0 ? acc[d] : 1.0;
    }
}

/* ------------------------------------------------------------------------------------
 * Full step: abstract_sph_strategy.py:31-115 with naive_sph_strategy.py:16-45 or
 * voxel_sph_strategy.py:19-116.  The state (pos, vel) is advanced in place; rho_out
 * receives the density of the step (computed from the positions at the start of the
 * step, as the reference returns it) and force_out the result_force.
 * ---------------------------------------------------------------------------------- */
typedef struct {
    int64_t n;
    double ext[3];
    double dt;
    double space[3];
    double voxel[3];
    int32_t mode;        /* 0 = BOX, 1 = PIPE (config.py:13) */
    int32_t formulation; /* 0 = NAIVE, 1 = VOXEL (config.py:14) */
    int32_t pipe_rows;
    int32_t use_grid;    /* NAIVE only: 1 = grid-accelerated search with identical neighbour sets */
    const double* pipe;
} oc_params_t;

static void grid_naive_terms(const double* pos, const double* vel, int64_t n, const oc_consts_t* C,
                             double* rho, double* Pt, double* Vt);

OC_API int oc_step(const oc_params_t* prm, const oc_consts_t* C, double* pos, double* vel, double* rho_out,
                   double* force_out, uint64_t* rng_states) {
    int64_t n = prm->n;
    double* Pt = (double*)calloc((size_t)(3 * n + 3), sizeof(double));
    double* Vt = (double*)calloc((size_t)(3 * n + 3), sizeof(double));
    if (!Pt || !Vt) { free(Pt); free(Vt); return 1; }
    if (prm->formulation == 0) {
        if (prm->use_grid) {
            grid_naive_terms(pos, vel, n, C, rho_out, Pt, Vt);
        } else {
            oc_naive_density(pos, n, C, rho_out);
            oc_naive_pressure(pos, rho_out, n, C, Pt);
            oc_naive_viscosity(pos, vel, rho_out, n, C, Vt);
        }
    } else {
        int32_t cdims[3], tdims[3];
        oc_space_dims(prm->space, prm->voxel, 1, cdims);
        oc_space_dims(prm->space, prm->voxel, 0, tdims);
        int64_t nv = (int64_t)cdims[0] * cdims[1] * cdims[2];
        int32_t* keys = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n + 1));
        int32_t* perm = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n + 1));
        int32_t* skeys = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n + 1));
        int32_t* begin = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nv + 1));
        int32_t* nb = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n * OC_MAX_NEIGHBOURS + 1));
        int32_t* nn = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n + 1));
        oc_voxel_keys(pos, n, prm->voxel, cdims, keys);
        oc_sort_map(keys, n, perm);
        for (int64_t i = 0; i < n; ++i) skeys[i] = keys[perm[i]];
        oc_voxel_begin(skeys, n, nv, begin);
        oc_voxel_ctx X = {pos, begin, perm, n, nv, {prm->voxel[0], prm->voxel[1], prm->voxel[2]},
                          {tdims[0], tdims[1], tdims[2]}, C->h};
        for (int64_t i = 0; i < n; ++i) nn[i] = voxel_neighbours(&X, i, nb + i * OC_MAX_NEIGHBOURS);
        oc_voxel_forces(pos, vel, n, nb, nn, C, rho_out, Pt, Vt);
        free(keys); free(perm); free(skeys); free(begin); free(nb); free(nn);
    }
    double* F = force_out ? force_out : (double*)malloc(sizeof(double) * (size_t)(3 * n + 3));
    oc_integrate(F, pos, vel, prm->ext, rho_out, Pt, Vt, n, prm->dt);
    if (prm->mode == 1) oc_collide_pipe(pos, vel, n, prm->pipe, prm->pipe_rows, rng_states);
    else oc_collide_box(pos, vel, n, prm->space, C->damp);
    if (!force_out) free(F);
    free(Pt); free(Vt);
    return 0;
}

/* Separate-term variant of the naive stages for stage-wise parity: given rho (any source),
 * compute the pressure and viscosity terms with the grid search or O(N^2). */
OC_API void oc_naive_terms(const double* pos, const double* vel, const double* rho, int64_t n,
                           const oc_consts_t* c, double* Pt, double* Vt) {
    oc_naive_pressure(pos, rho, n, c, Pt);
    oc_naive_viscosity(pos, vel, rho, n, c, Vt);
}

/* ------------------------------------------------------------------------------------
 * Grid-accelerated NAIVE formulation (CPU throughput baseline).  The neighbour set of
 * particle i is {j : |x_i - x_j|^2 <= h^2}, exactly as in naive_kernels.py; the grid
 * only avoids testing pairs that are provably farther apart than h.  Cells have edge
 * >= h and clamp out-of-range coordinates, which keeps every pair within h in adjacent
 * cells.  Sums run over j in ascending order within each cell, so they equal the O(N^2)
 * sums up to summation order.
 * ---------------------------------------------------------------------------------- */
typedef struct {
    double lo[3], inv[3];
    int64_t dims[3];
    int64_t* start; /* n_cells + 1 */
    int64_t* idx;   /* particle ids sorted by cell, ascending id within cell */
} oc_grid_t;

static int64_t grid_coord(const oc_grid_t* G, double x, int d) {
    double q = floor((x - G->lo[d]) * G->inv[d]);
    if (!(q >= 0)) return 0;
    if (q >= (double)(G->dims[d] - 1)) return G->dims[d] - 1;
    return (int64_t)q;
}

static void grid_build(oc_grid_t* G, const double* pos, int64_t n, double h) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            double x = pos[3 * i + d];
            if (x == x && fabs(x) < 1e300) { if (x < lo[d]) lo[d] = x; if (x > hi[d]) hi[d] = x; }
        }
    int64_t total = 1;
    for (int d = 0; d < 3; ++d) {
        if (!(lo[d] <= hi[d])) { lo[d] = 0; hi[d] = 0; }
        double ext = hi[d] - lo[d];
        int64_t nd = (int64_t)floor(ext / h);
        if (nd < 1) nd = 1;
        if (nd > 1024) nd = 1024;
        G->lo[d] = lo[d];
        G->dims[d] = nd;
        G->inv[d] = (double)nd / (ext > 0 ? ext : 1.0); /* cell edge = ext / nd >= h */
        if (ext / nd < h) G->inv[d] = 1.0 / h;
        total *= nd;
    }
    G->start = (int64_t*)calloc((size_t)(total + 1), sizeof(int64_t));
    G->idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    int64_t* cell = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    for (int64_t i = 0; i < n; ++i) {
        int64_t c = grid_coord(G, pos[3 * i], 0) +
                    G->dims[0] * (grid_coord(G, pos[3 * i + 1], 1) + G->dims[1] * grid_coord(G, pos[3 * i + 2], 2));
        cell[i] = c;
        G->start[c + 1]++;
    }
    for (int64_t c = 0; c < total; ++c) G->start[c + 1] += G->start[c];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(total + 1));
    memcpy(fill, G->start, sizeof(int64_t) * (size_t)(total + 1));
    for (int64_t i = 0; i < n; ++i) G->idx[fill[cell[i]]++] = i;
    free(fill);
    free(cell);
}

static void grid_free(oc_grid_t* G) { free(G->start); free(G->idx); }

/* visit every j with |x_i - x_j|^2 <= h^2 (self included) in (cell, id) order */
#define GRID_FOR_NEIGHBOURS(G, pos, i, h2, BODY)                                                  \
    do {                                                                                          \
        int64_t cx = grid_coord(G, pos[3 * (i)], 0), cy = grid_coord(G, pos[3 * (i) + 1], 1),     \
                cz = grid_coord(G, pos[3 * (i) + 2], 2);                                          \
        for (int64_t z = cz - 1; z <= cz + 1; ++z) {                                              \
            if (z < 0 || z >= (G)->dims[2]) continue;                                             \
            for (int64_t y = cy - 1; y <= cy + 1; ++y) {                                          \
                if (y < 0 || y >= (G)->dims[1]) continue;                                         \
                for (int64_t x = cx - 1; x <= cx + 1; ++x) {                                      \
                    if (x < 0 || x >= (G)->dims[0]) continue;                                     \
                    int64_t c = x + (G)->dims[0] * (y + (G)->dims[1] * z);                        \
                    for (int64_t e = (G)->start[c]; e < (G)->start[c + 1]; ++e) {                 \
                        int64_t j = (G)->idx[e];                                                  \
                        if (norm_squared(pos + 3 * (i), pos + 3 * j) > (h2)) continue;            \
                        BODY                                                                      \
                    }                                                                             \
                }                                                                                 \
            }                                                                                     \
        }                                                                                         \
    } while (0)

static void grid_naive_terms(const double* pos, const double* vel, int64_t n, const oc_consts_t* C,
                             double* rho, double* Pt, double* Vt) {
    oc_derived_t D = derive(C);
    oc_grid_t G;
    grid_build(&G, pos, n, C->h);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
        double acc = 0.0;
        GRID_FOR_NEIGHBOURS(&G, pos, i, D.h2, { acc += compute_w(&D, pos + 3 * i, pos + 3 * j); });
        rho[i] = acc * C->mass;
    }
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
        double ap[3] = {0, 0, 0}, av[3] = {0, 0, 0};
        GRID_FOR_NEIGHBOURS(&G, pos, i, D.h2, {
            if (j != i) {
                double pi = C->k * (rho[i] - C->rho0);
                double pj = C->k * (rho[j] - C->rho0);
                double factor = (pi + pj) / (2 * rho[j]);
                double g[3];
                compute_grad_w(C, &D, pos + 3 * i, pos + 3 * j, g);
                double lap = compute_lap_w(C, &D, pos + 3 * i, pos + 3 * j);
                for (int d = 0; d < 3; ++d) {
                    ap[d] += factor * g[d];
                    double vt = (vel[3 * j + d] - vel[3 * i + d]) / rho[j] * lap * C->visc;
                    av[d] += vt * C->mass;
                }
            }
        });
        for (int d = 0; d < 3; ++d) { Pt[3 * i + d] = ap[d] * C->mass; Vt[3 * i + d] = av[d]; }
    }
    grid_free(&G);
}

OC_API void oc_grid_naive_terms(const double* pos, const double* vel, int64_t n, const oc_consts_t* c,
                                double* rho, double* Pt, double* Vt) {
    grid_naive_terms(pos, vel, n, c, rho, Pt, Vt);
}

OC_API int oc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

OC_API void oc_set_num_threads(int t) {
#ifdef _OPENMP
    if (t > 0) omp_set_num_threads(t);
#else
    (void)t;
#endif
}

/* ------------------------------------------------------------------------------------
 * Rollout and its adjoint (NAIVE formulation, box walls), FP64.  No reference
 * counterpart: the reference has no gradients.  This is the reference gradient for the
 * GPU adjoint kernels and is itself checked against central finite differences.
 *
 * Loss = (1/N) sum_i |x_T,i - target_i|^2; gradients with respect to the initial
 * velocities and the viscosity mu.  The step is the one of oc_step (NAIVE, BOX):
 *   rho_i = m sum_j W_ij (self included), p = K (rho - rho0),
 *   P_i = m sum_j (p_i + p_j) / (2 rho_j) G(x_i - x_j), G(d) = Gc (h - r)^2 d / r,
 *   V_i = m mu sum_j (v_j - v_i) / rho_j L(r_ij),   L(r) = Lc (h - r),
 *   F = f - P + V, v' = v + F / rho dt, x' = x + v' dt, then the box collision.
 * Every transposed sum of the adjoint is a sum over the (symmetric) neighbour set.
 * ---------------------------------------------------------------------------------- */
typedef struct {
    int64_t* start; /* n + 1 */
    int64_t* idx;
} oc_nlist_t;

static void build_nlist(const double* pos, int64_t n, double h, oc_nlist_t* L) {
    oc_grid_t G;
    grid_build(&G, pos, n, h);
    double h2 = h * h;
    int64_t* cnt = (int64_t*)calloc((size_t)(n + 1), sizeof(int64_t));
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
        int64_t found = 0; /* the macro's own locals are x, y, z, c, e, j */
        GRID_FOR_NEIGHBOURS(&G, pos, i, h2, { if (j != i) ++found; });
        cnt[i] = found;
    }
    L->start = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    L->start[0] = 0;
    for (int64_t i = 0; i < n; ++i) L->start[i + 1] = L->start[i] + cnt[i];
    L->idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(L->start[n] + 1));
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
        int64_t o = L->start[i];
        GRID_FOR_NEIGHBOURS(&G, pos, i, h2, { if (j != i) L->idx[o++] = j; });
    }
    free(cnt);
    grid_free(&G);
}

static void free_nlist(oc_nlist_t* L) { free(L->start); free(L->idx); }

typedef struct {
    double *rho, *P, *V, *F; /* n, 3n, 3n, 3n */
    oc_nlist_t nl;
} oc_fwd_t;

/* Box collision of base_kernels.py:75-98 with its branch per component exposed:
 * 0 = no wall, 1 = below 0 (x -> 1e-3), 2 = above space (x -> space - 1e-3).  bmode 0 decides
 * as the reference does, 1 decides and records, 2 replays recorded decisions (the same
 * smooth piece of the rollout map, for finite differences that do not step across a
 * change of bounce step). */
static void collide_box_branch(double* pos, double* vel, int64_t n, const double* space, double damp, uint8_t* br,
                               int bmode) {
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            size_t k = 3 * (size_t)i + d;
            double x = pos[k];
            uint8_t b;
            if (bmode == 2) b = br[k];
            else b = x < 0 ? (1e-3 > space[d] ? 2 : 1) : (x > space[d] ? 2 : 0);
            if (bmode == 1) br[k] = b;
            if (b == 1) pos[k] = 1e-3;
            if (b == 2) pos[k] = space[d] - 1e-3;
            if (b) { vel[k] *= -1; vel[k] *= damp; }
        }
}

/* one step with intermediates; x1/v1 receive the new state */
static void fwd_step(const oc_params_t* prm, const oc_consts_t* C, const double* x, const double* v, double* x1,
                     double* v1, oc_fwd_t* S, uint8_t* br, int bmode) {
    int64_t n = prm->n;
    oc_derived_t D = derive(C);
    build_nlist(x, n, C->h, &S->nl);
    double w0 = D.w_const * D.h2 * D.h2 * D.h2;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
        double acc = w0;
        for (int64_t e = S->nl.start[i]; e < S->nl.start[i + 1]; ++e)
            acc += compute_w(&D, x + 3 * i, x + 3 * S->nl.idx[e]);
        S->rho[i] = acc * C->mass;
    }
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
        double ap[3] = {0, 0, 0}, av[3] = {0, 0, 0};
        double pi = C->k * (S->rho[i] - C->rho0);
        for (int64_t e = S->nl.start[i]; e < S->nl.start[i + 1]; ++e) {
            int64_t j = S->nl.idx[e];
            double pj = C->k * (S->rho[j] - C->rho0);
            double g[3];
            compute_grad_w(C, &D, x + 3 * i, x + 3 * j, g);
            double lap = compute_lap_w(C, &D, x + 3 * i, x + 3 * j);
            for (int d = 0; d < 3; ++d) {
                ap[d] += (pi + pj) / (2 * S->rho[j]) * g[d];
                av[d] += (v[3 * j + d] - v[3 * i + d]) / S->rho[j] * lap * C->visc * C->mass;
            }
        }
        for (int d = 0; d < 3; ++d) { S->P[3 * i + d] = ap[d] * C->mass; S->V[3 * i + d] = av[d]; }
    }
    memcpy(x1, x, sizeof(double) * 3 * (size_t)n);
    memcpy(v1, v, sizeof(double) * 3 * (size_t)n);
    oc_integrate(S->F, x1, v1, prm->ext, S->rho, S->P, S->V, n, prm->dt);
    collide_box_branch(x1, v1, n, prm->space, C->damp, br, bmode);
}

/* adjoint of one step: (gx1, gv1) on the output -> (gx0, gv0) on the input; adds dL/dmu */
static void adj_step(const oc_params_t* prm, const oc_consts_t* C, const double* x, const double* v,
                     const oc_fwd_t* S, const double* gx1, const double* gv1, double* gx0, double* gv0, double* gmu) {
    int64_t n = prm->n;
    oc_derived_t D = derive(C);
    double dt = prm->dt, m = C->mass, mu = C->visc;
    double* Pb = (double*)malloc(sizeof(double) * 3 * (size_t)(n + 1));
    double* Vb = (double*)malloc(sizeof(double) * 3 * (size_t)(n + 1));
    double* rb = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    double mu_acc = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        rb[i] = 0.0;
        for (int d = 0; d < 3; ++d) {
            size_t k = 3 * (size_t)i + d;
            /* recompute the pre-collision state (base_kernels.py:46-53) */
            double vpre = v[k] + S->F[k] / S->rho[i] * dt;
            double xpre = x[k] + vpre * dt;
            int bounced = (xpre < 0) || ((xpre < 0 ? 1e-3 : xpre) > prm->space[d]);
            double gxp = bounced ? 0.0 : gx1[k];
            double gvp = (bounced ? -C->damp * gv1[k] : gv1[k]) + gxp * dt;
            gx0[k] = gxp;
            gv0[k] = gvp;
            double Fb = gvp * dt / S->rho[i];
            rb[i] += -gvp * S->F[k] * dt / (S->rho[i] * S->rho[i]);
            Pb[k] = -Fb;
            Vb[k] = Fb;
            mu_acc += Fb * S->V[k] / mu;
        }
    }
    *gmu += mu_acc;
    double* rtot = (double*)malloc(sizeof(double) * (size_t)(n + 1));
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t k = 0; k < n; ++k) {
        double pk = C->k * (S->rho[k] - C->rho0), rk = S->rho[k];
        double pbar = 0.0, rbar = rb[k];
        double gx[3] = {0, 0, 0}, gv[3] = {0, 0, 0};
        for (int64_t e = S->nl.start[k]; e < S->nl.start[k + 1]; ++e) {
            int64_t j = S->nl.idx[e];
            double pj = C->k * (S->rho[j] - C->rho0), rj = S->rho[j];
            double d3[3], r2 = 0.0;
            for (int d = 0; d < 3; ++d) { d3[d] = x[3 * k + d] - x[3 * j + d]; r2 += d3[d] * d3[d]; }
            double r = sqrt(r2), hr = C->h - r;
            double L = D.lap_w_const * hr;
            const double *Vk = Vb + 3 * k, *Vj = Vb + 3 * j, *Pk = Pb + 3 * k, *Pj = Pb + 3 * j;
            /* viscosity */
            double akj = 0.0, ajk = 0.0, gvk = 0.0;
            for (int d = 0; d < 3; ++d) {
                gv[d] += m * mu * (L / rk * Vj[d] - L / rj * Vk[d]);
                akj += Vk[d] * (v[3 * j + d] - v[3 * k + d]);
                ajk += Vj[d] * (v[3 * k + d] - v[3 * j + d]);
            }
            gvk = ajk;
            rbar += -(m * mu / (rk * rk)) * gvk * L;
            akj *= m * mu / rj;
            ajk *= m * mu / rk;
            for (int d = 0; d < 3; ++d) gx[d] += -D.lap_w_const * (akj + ajk) * d3[d] / r;
            /* pressure */
            double g = hr * hr / r, gp = -hr * (C->h + r) / (r * r);
            double PkG = 0.0, PjG = 0.0;
            for (int d = 0; d < 3; ++d) { PkG += Pk[d] * D.grad_w_const * g * d3[d]; PjG += Pj[d] * D.grad_w_const * g * (-d3[d]); }
            pbar += m * (PkG / (2 * rj) + PjG / (2 * rk));
            rbar += -m / (2 * rk * rk) * (pj + pk) * PjG;
            double ckj = (pk + pj) / (2 * rj), cjk = (pj + pk) / (2 * rk);
            double b[3], db = 0.0;
            for (int d = 0; d < 3; ++d) { b[d] = m * (ckj * Pk[d] - cjk * Pj[d]); db += d3[d] * b[d]; }
            for (int d = 0; d < 3; ++d) gx[d] += D.grad_w_const * (g * b[d] + gp * db * d3[d] / r);
        }
        rtot[k] = rbar + C->k * pbar;
        for (int d = 0; d < 3; ++d) { gx0[3 * k + d] += gx[d]; gv0[3 * k + d] += gv[d]; }
    }
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t k = 0; k < n; ++k) {
        double gx[3] = {0, 0, 0};
        for (int64_t e = S->nl.start[k]; e < S->nl.start[k + 1]; ++e) {
            int64_t j = S->nl.idx[e];
            double d3[3], r2 = 0.0;
            for (int d = 0; d < 3; ++d) { d3[d] = x[3 * k + d] - x[3 * j + d]; r2 += d3[d] * d3[d]; }
            double t = D.h2 - r2;
            double s = m * (rtot[k] + rtot[j]) * (-6.0 * D.w_const * t * t);
            for (int d = 0; d < 3; ++d) gx[d] += s * d3[d];
        }
        for (int d = 0; d < 3; ++d) gx0[3 * k + d] += gx[d];
    }
    free(Pb); free(Vb); free(rb); free(rtot);
}

/* x_T of a rollout (NAIVE, BOX).  branches (steps x 3n bytes) with bmode 1 records the
 * bounce decisions, with bmode 2 replays them; NULL / 0 for the plain rollout. */
OC_API int oc_rollout(const oc_params_t* prm, const oc_consts_t* C, const double* x0, const double* v0, int steps,
                      double* xT, double* vT, uint8_t* branches, int bmode) {
    int64_t n = prm->n;
    double* x = (double*)malloc(sizeof(double) * 3 * (size_t)(n + 1));
    double* v = (double*)malloc(sizeof(double) * 3 * (size_t)(n + 1));
    double* x1 = (double*)malloc(sizeof(double) * 3 * (size_t)(n + 1));
    double* v1 = (double*)malloc(sizeof(double) * 3 * (size_t)(n + 1));
    oc_fwd_t S;
    S.rho = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    S.P = (double*)malloc(sizeof(double) * 3 * (size_t)(n + 1));
    S.V = (double*)malloc(sizeof(double) * 3 * (size_t)(n + 1));
    S.F = (double*)malloc(sizeof(double) * 3 * (size_t)(n + 1));
    memcpy(x, x0, sizeof(double) * 3 * (size_t)n);
    memcpy(v, v0, sizeof(double) * 3 * (size_t)n);
    if (!branches) bmode = 0;
    for (int t = 0; t < steps; ++t) {
        fwd_step(prm, C, x, v, x1, v1, &S, bmode ? branches + (size_t)t * 3 * (size_t)n : NULL, bmode);
        free_nlist(&S.nl);
        double* tmp = x; x = x1; x1 = tmp;
        tmp = v; v = v1; v1 = tmp;
    }
    memcpy(xT, x, sizeof(double) * 3 * (size_t)n);
    if (vT) memcpy(vT, v, sizeof(double) * 3 * (size_t)n);
    free(x); free(v); free(x1); free(v1); free(S.rho); free(S.P); free(S.V); free(S.F);
    return 0;
}

/* loss and gradients; states are kept every `every` steps and recomputed in between */
OC_API int oc_rollout_adjoint(const oc_params_t* prm, const oc_consts_t* C, const double* x0, const double* v0,
                              int steps, const double* target, int every, double* loss, double* g_v0, double* g_mu) {
    int64_t n = prm->n;
    size_t sz = sizeof(double) * 3 * (size_t)(n + 1);
    if (every < 1) every = 1;
    int nck = (steps + every - 1) / every + 1;
    double* ckx = (double*)malloc(sz * (size_t)nck);
    double* ckv = (double*)malloc(sz * (size_t)nck);
    double* xs = (double*)malloc(sz * (size_t)(every + 1));
    double* vs = (double*)malloc(sz * (size_t)(every + 1));
    double* gx = (double*)malloc(sz);
    double* gv = (double*)malloc(sz);
    double* gx0 = (double*)malloc(sz);
    double* gv0 = (double*)malloc(sz);
    oc_fwd_t S;
    S.rho = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    S.P = (double*)malloc(sz);
    S.V = (double*)malloc(sz);
    S.F = (double*)malloc(sz);
    /* forward, keeping checkpoints */
    memcpy(xs, x0, sizeof(double) * 3 * (size_t)n);
    memcpy(vs, v0, sizeof(double) * 3 * (size_t)n);
    memcpy(ckx, x0, sizeof(double) * 3 * (size_t)n);
    memcpy(ckv, v0, sizeof(double) * 3 * (size_t)n);
    for (int t = 0; t < steps; ++t) {
        fwd_step(prm, C, xs, vs, xs + 3 * (n + 1), vs + 3 * (n + 1), &S, NULL, 0);
        free_nlist(&S.nl);
        memcpy(xs, xs + 3 * (n + 1), sizeof(double) * 3 * (size_t)n);
        memcpy(vs, vs + 3 * (n + 1), sizeof(double) * 3 * (size_t)n);
        if ((t + 1) % every == 0) {
            memcpy(ckx + 3 * (n + 1) * ((t + 1) / every), xs, sizeof(double) * 3 * (size_t)n);
            memcpy(ckv + 3 * (n + 1) * ((t + 1) / every), vs, sizeof(double) * 3 * (size_t)n);
        }
    }
    double l = 0.0;
    for (int64_t i = 0; i < 3 * n; ++i) {
        double d = xs[i] - target[i];
        l += d * d;
        gx[i] = 2.0 * d / (double)n;
        gv[i] = 0.0;
    }
    *loss = l / (double)n;
    double gm = 0.0;
    /* backward over segments */
    for (int seg = (steps - 1) / every; seg >= 0; --seg) {
        int t0 = seg * every, t1 = t0 + every < steps ? t0 + every : steps;
        memcpy(xs, ckx + 3 * (n + 1) * seg, sizeof(double) * 3 * (size_t)n);
        memcpy(vs, ckv + 3 * (n + 1) * seg, sizeof(double) * 3 * (size_t)n);
        for (int t = t0; t < t1; ++t) {
            int a = t - t0;
            fwd_step(prm, C, xs + 3 * (n + 1) * a, vs + 3 * (n + 1) * a, xs + 3 * (n + 1) * (a + 1),
                     vs + 3 * (n + 1) * (a + 1), &S, NULL, 0);
            free_nlist(&S.nl);
        }
        for (int t = t1 - 1; t >= t0; --t) {
            int a = t - t0;
            fwd_step(prm, C, xs + 3 * (n + 1) * a, vs + 3 * (n + 1) * a, gx0, gv0, &S, NULL, 0); /* intermediates */
            adj_step(prm, C, xs + 3 * (n + 1) * a, vs + 3 * (n + 1) * a, &S, gx, gv, gx0, gv0, &gm);
            free_nlist(&S.nl);
            memcpy(gx, gx0, sizeof(double) * 3 * (size_t)n);
            memcpy(gv, gv0, sizeof(double) * 3 * (size_t)n);
        }
    }
    memcpy(g_v0, gv, sizeof(double) * 3 * (size_t)n);
    *g_mu = gm;
    free(ckx); free(ckv); free(xs); free(vs); free(gx); free(gv); free(gx0); free(gv0);
    free(S.rho); free(S.P); free(S.V); free(S.F);
    return 0;
}

/* adjoint of one step at state (x, v): (gx1, gv1) on the step's output -> (gx0, gv0) on its
 * input, and the step's dL/dmu in *gmu */
OC_API int oc_adjoint_step(const oc_params_t* prm, const oc_consts_t* C, const double* x, const double* v,
                           const double* gx1, const double* gv1, double* gx0, double* gv0, double* gmu) {
    int64_t n = prm->n;
    size_t sz = sizeof(double) * 3 * (size_t)(n + 1);
    double* x1 = (double*)malloc(sz);
    double* v1 = (double*)malloc(sz);
    oc_fwd_t S;
    S.rho = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    S.P = (double*)malloc(sz);
    S.V = (double*)malloc(sz);
    S.F = (double*)malloc(sz);
    fwd_step(prm, C, x, v, x1, v1, &S, NULL, 0);
    *gmu = 0.0;
    adj_step(prm, C, x, v, &S, gx1, gv1, gx0, gv0, gmu);
    free_nlist(&S.nl);
    free(x1); free(v1); free(S.rho); free(S.P); free(S.V); free(S.F);
    return 0;
}
