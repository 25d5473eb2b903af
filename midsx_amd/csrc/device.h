// Device-side building blocks of the photon-transport path (FP32 transport,
// FP64 tallies).  Each function cites the reference code whose behaviour it
// reproduces.
#pragma once
#include <hip/hip_runtime.h>

#include "mdx_types.h"

namespace mdxd {

// include/MIDSX/Core/constants.h:13-23, evaluated in FP64 and rounded once.
constexpr float kPiF = 3.14159265358979323846f;
constexpr float kMc2 = 510998.95f;
constexpr float kInvMc2 = (float)(1.0 / 510998.95);
// ALPHA = mc^2 / (sqrt(2) h c 1e8) in 1/Angstrom
constexpr double kAlphaD = 0.51099895000e6 / (1.41421356237309504880 * 4.135667662e-15 * 299792458e2 * 1e8);
constexpr float kAlpha = (float)kAlphaD;
constexpr float kAlphaSqrt2 = (float)(kAlphaD * 1.41421356237309504880);
constexpr float kEps = 1e-9f;  // constants.h:14

// ------------------------------------------------------------------ RNG
// Philox-4x32-10 and Philox-2x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC'11): ten rounds of the multiply /
// xor-with-key bijection, the key bumped by the Weyl constants after each.
// Output = the block of counter (x0..x3) under key (k0, k1).
__device__ __forceinline__ void philox4x32_10(uint32_t x0, uint32_t x1, uint32_t x2, uint32_t x3, uint32_t a,
                                              uint32_t b, uint32_t& o0, uint32_t& o1, uint32_t& o2, uint32_t& o3) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t h0 = __umulhi(0xD2511F53u, x0), l0 = 0xD2511F53u * x0;
    const uint32_t h1 = __umulhi(0xCD9E8D57u, x2), l1 = 0xCD9E8D57u * x2;
    x0 = h1 ^ x1 ^ a;
    x1 = l1;
    x2 = h0 ^ x3 ^ b;
    x3 = l0;
    a += 0x9E3779B9u;
    b += 0xBB67AE85u;
  }
  o0 = x0;
  o1 = x1;
  o2 = x2;
  o3 = x3;
}

__device__ __forceinline__ void philox2x32_10(uint32_t x0, uint32_t x1, uint32_t k, uint32_t& o0, uint32_t& o1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi = __umulhi(0xD256D193u, x0), lo = 0xD256D193u * x0;
    x0 = hi ^ k ^ x1;
    x1 = lo;
    k += 0x9E3779B9u;
  }
  o0 = x0;
  o1 = x1;
}

// Per-history stream of Philox-4x32-10 blocks, keyed by (seed, job) with the
// history index in the counter.
struct Philox {
  uint32_t k0, k1;
  uint32_t c0, c1, c2, c3;
  uint32_t b0, b1, b2, b3;  // unused outputs of the last block, consumed b0 first
  int left;

  __device__ void init(uint64_t seed, uint32_t job, uint64_t history) {
    k0 = (uint32_t)seed;
    k1 = (uint32_t)(seed >> 32) ^ (job * 0x9E3779B9u);
    c0 = 0;
    c1 = (uint32_t)history;
    c2 = (uint32_t)(history >> 32);
    c3 = 0x4D445831u;
    left = 0;
  }
  __device__ void refill() {
    philox4x32_10(c0, c1, c2, c3, k0, k1, b0, b1, b2, b3);
    left = 4;
    ++c0;
  }
  // Uniform in (0, 1] with 24 random bits (never 0, so -log(u) is finite).
  // The block's words are used in order b0, b1, b2, b3 (a register shift,
  // no indexed access).
  __device__ float uniform() {
    if (left == 0) refill();
    uint32_t x = b0;
    b0 = b1;
    b1 = b2;
    b2 = b3;
    --left;
    return (float)((x >> 8) + 1u) * (1.0f / 16777216.0f);
  }
};

// Philox-2x32-10 for the Woodcock step: one block per step gives the step's
// two uniforms (free path, delta test), keyed by (seed, job, history high
// word) with counter (history low word, step index).  Every lane draws the
// same amount per step, so no lane waits on another's buffer refill.
__device__ __forceinline__ uint32_t mix32(uint32_t a, uint32_t b) {
  uint32_t h = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u + (a << 6) + (a >> 2));
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}

struct StepRng {
  uint32_t key, hist, step;
  __device__ void init(uint64_t seed, uint32_t job, uint64_t history) {
    key = mix32(mix32((uint32_t)seed, (uint32_t)(seed >> 32)), mix32(job, (uint32_t)(history >> 32)));
    hist = (uint32_t)history;
    step = 0;
  }
  __device__ void next(float& u1, float& u2) {
    uint32_t x0, x1;
    philox2x32_10(hist, step, key, x0, x1);
    ++step;
    u1 = (float)((x0 >> 8) + 1u) * (1.0f / 16777216.0f);
    u2 = (float)((x1 >> 8) + 1u) * (1.0f / 16777216.0f);
  }
};

// ------------------------------------------------------------------ pointers
// A pointer read from memory is a generic one, for which the compiler emits
// flat memory instructions (address-space checks, 64-bit VGPR addresses, and
// waits that also cover LDS traffic).  Every table, grid and tally buffer
// that the kernels reach through such a pointer is device global memory, so
// their loads and atomics go through address space 1.
#define MDX_GLOBAL __attribute__((address_space(1)))

template <class T>
__device__ __forceinline__ T gld(const T* p) {
  return *(const MDX_GLOBAL T*)p;
}

// Descriptors that no kernel writes while it runs (a job's domain, source and
// tallies), reached through a pointer that is the same in every lane of the
// wave: made a wave-uniform pointer into the constant address space, so that
// once inlined their loads are scalar loads into SGPRs through the scalar
// cache rather than per-lane vector loads into VGPRs.  The empty asm keeps the
// pointer opaque, so that the loads stay where they are used instead of being
// hoisted out of a loop and held in registers across it.
#define MDX_CONST __attribute__((address_space(4)))

template <class T>
__device__ __forceinline__ const T* wave_const(const T* p) {
  uint64_t b = (uint64_t)(uintptr_t)p;
  uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)b);
  uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(b >> 32));
  asm volatile("" : "+s"(lo), "+s"(hi));
  return (const T*)(const MDX_CONST T*)(((uint64_t)hi << 32) | lo);
}

typedef float mdx_v4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ mdx_f4 gld(const mdx_f4* p) {
  const mdx_v4 a = *(const MDX_GLOBAL mdx_v4*)p;
  mdx_f4 r;
  r.x = a.x; r.y = a.y; r.z = a.z; r.w = a.w;
  return r;
}

typedef float mdx_v2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ mdx_node gld(const mdx_node* p) {  // the padding is not read
  const mdx_v4 a = *(const MDX_GLOBAL mdx_v4*)p;
  const mdx_v2 b = *(const MDX_GLOBAL mdx_v2*)((const MDX_GLOBAL float*)p + 4);
  mdx_node r;
  r.a_hi = a.x; r.a_lo = a.y; r.c1 = a.z; r.c2 = a.w;
  r.c3 = b.x; r.x = b.y; r.pad0 = 0.0f; r.pad1 = 0.0f;
  return r;
}

template <class T>
__device__ __forceinline__ T gadd(T* p, T v) {  // atomicAdd on global memory; returns the old value
  return __hip_atomic_fetch_add((MDX_GLOBAL T*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The hardware reciprocal (v_rcp_f32, 1 ulp): quotients that feed sampling
// and geometry, where a last-place difference from the IEEE quotient is
// immaterial, are taken as a * frcp(b).
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

// ------------------------------------------------------------------ curves
// ln(X / xk) for X, xk > 0: r approximates the quotient within a few ulp,
// e = X - r xk is its residual (one fma), and ln(X / xk) = ln r + ln(1 + e /
// (r xk)) = ln r + e / X to well below an ulp of the result, so the
// approximate quotient costs no accuracy.
__device__ __forceinline__ float log_ratio(float X, float xk) {
  const float r = X * frcp(xk);
  const float e = __builtin_fmaf(-r, xk, X);
  return logf(r) + e * frcp(X);
}

// A curve lookup reads one bin record: the hash bin of ln X holds the record
// of the largest node index k with x_k at or below the bin's lower edge,
// clamped to [0, n-2], together with x_{k+1} and k.  That node brackets X
// unless the bin holds further nodes, in which case a short walk over the
// abscissae follows.  The result is the largest k in [0, n-2] with x_k <= X
// (0 below x_0), the same index as upper_bound - 1 clamped to the last span.
struct CurveAt {
  int k;
  mdx_node r;  // node k
  float xn;    // abscissa of node k + 1
};

// The bin comes from the hardware log2: a point within its error of a bin
// edge may land in the neighbouring bin, whose node curve_fix then corrects.
__device__ __forceinline__ CurveAt curve_bin(const mdx_tables_dev* T, const mdx_curve& c, float X) {
  constexpr float kLn2 = 0.69314718055994531f;
  int b = (int)((__log2f(X) * kLn2 - c.lnx0) * c.inv_dlnx);
  b = b < 0 ? 0 : (b >= c.nh ? c.nh - 1 : b);
  const MDX_GLOBAL float* p = (const MDX_GLOBAL float*)(T->bins + c.hoff + b);
  const mdx_v4 a = *(const MDX_GLOBAL mdx_v4*)p;
  const mdx_v4 q = *(const MDX_GLOBAL mdx_v4*)(p + 4);
  CurveAt r;
  r.k = __float_as_int(q.w);
  r.r.a_hi = a.x; r.r.a_lo = a.y; r.r.c1 = a.z; r.r.c2 = a.w;
  r.r.c3 = q.x; r.r.x = q.y; r.r.pad0 = 0.0f; r.r.pad1 = 0.0f;
  r.xn = q.z;
  return r;
}

__device__ __forceinline__ CurveAt curve_nodes(const mdx_tables_dev* T, const mdx_curve& c, int k) {
  const mdx_node* nd = T->nodes + c.off;
  return CurveAt{k, gld(nd + k), gld(&nd[k + 1].x)};
}

__device__ __forceinline__ void curve_fix(const mdx_tables_dev* T, const mdx_curve& c, float X, CurveAt& a) {
  const int kmax = c.n - 2;
  if ((X < a.r.x && a.k > 0) || (X >= a.xn && a.k < kmax)) {
    const float* xs = T->xs + c.off;
    int k = a.k;
    while (k > 0 && gld(xs + k) > X) --k;
    while (k < kmax && gld(xs + k + 1) <= X) ++k;
    a = curve_nodes(T, c, k);
  }
}

// Linear curves clamp outside their node range (interpolators.cpp:48-54);
// spline curves extrapolate with their end spans (Eigen Spline::Span).
__device__ __forceinline__ float curve_finish(const mdx_tables_dev* T, const mdx_curve& c, float X, CurveAt a) {
  float t;
  if (c.kind == 0) {
    if (X <= a.r.x) return a.r.a_hi + a.r.a_lo;  // X <= x_0 (or exactly on node k)
    if (X >= a.xn) {                             // X >= x_{n-1}
      const mdx_node e = gld(T->nodes + c.off + a.k + 1);
      return e.a_hi + e.a_lo;
    }
    t = log_ratio(X, a.r.x);
  } else {
    t = X >= a.r.x ? log_ratio(X, a.r.x) : -log_ratio(a.r.x, X);
  }
  const mdx_node& r = a.r;
  const float p = t * (r.c1 + t * (r.c2 + t * r.c3));
  const float em1 = expm1f(p);
  return __builtin_fmaf(r.a_hi, em1, __builtin_fmaf(r.a_lo, em1, r.a_lo)) + r.a_hi;
}

// Evaluates a packed curve (see mdx_types.h).
__device__ __forceinline__ float curve_eval(const mdx_tables_dev* T, const mdx_curve& c, float X) {
  CurveAt a = curve_bin(T, c, X);
  curve_fix(T, c, X, a);
  return curve_finish(T, c, X, a);
}

// Two curves at once, their loads issued side by side.
__device__ __forceinline__ void curve_eval2(const mdx_tables_dev* T, const mdx_curve& c1, float X1,
                                            const mdx_curve& c2, float X2, float& v1, float& v2) {
  CurveAt a1 = curve_bin(T, c1, X1), a2 = curve_bin(T, c2, X2);
  curve_fix(T, c1, X1, a1);
  curve_fix(T, c2, X2, a2);
  v1 = curve_finish(T, c1, X1, a1);
  v2 = curve_finish(T, c2, X2, a2);
}

// Three curves with their table loads issued together (one chain of
// dependent loads instead of three).
__device__ __forceinline__ void curve_eval3(const mdx_tables_dev* T, const mdx_curve& c1, float X1,
                                            const mdx_curve& c2, float X2, const mdx_curve& c3, float X3,
                                            float& v1, float& v2, float& v3) {
  CurveAt a1 = curve_bin(T, c1, X1), a2 = curve_bin(T, c2, X2), a3 = curve_bin(T, c3, X3);
  curve_fix(T, c1, X1, a1);
  curve_fix(T, c2, X2, a2);
  curve_fix(T, c3, X3, a3);
  v1 = curve_finish(T, c1, X1, a1);
  v2 = curve_finish(T, c2, X2, a2);
  v3 = curve_finish(T, c3, X3, a3);
}

// ------------------------------------------------------------------ geometry
// computational_domain.cpp:16-21 (inclusive bounds)
__device__ __forceinline__ bool in_domain(const mdx_domain_dev* D, float x, float y, float z) {
  return x >= 0.0f && x <= D->dim_space[0] && y >= 0.0f && y <= D->dim_space[1] && z >= 0.0f &&
         z <= D->dim_space[2];
}

// trunc(fl(q / s)) for q >= 0, s > 0, as the reference's
// (position / spacing).cast<int>() (voxel_grid.cpp:55-59), without a
// division in the common case: r = q * fl(1/s) is within 3 * 2^-24 relative of
// fl(q / s), so when r is farther than that from an integer both truncate to
// the same value; otherwise the IEEE quotient is taken.
__device__ __forceinline__ int trunc_quotient(float q, float s, float inv_s) {
  float r = q * inv_s;
  float f = floorf(r);
  float frac = r - f;  // exact: f and r share the binade or f = 0
  float tol = r * 3.0e-7f;
  if (frac > tol && frac < 1.0f - tol) return (int)f;
  return (int)(q / s);
}

// computational_domain.cpp:23-31, voxel_grid.cpp:55-59,119-128: first grid
// containing the point wins; index = truncation of the IEEE quotient
// (position - origin) / spacing.  A point exactly on a grid's far face maps to
// the last voxel (the reference's at() throws there).  Returns the grid, or -1
// for the background voxel.
// The voxel of grid G holding a point of its box (offsets qx, qy, qz >= 0
// from its origin): linear index, and optionally the voxel's macro cell.
__device__ __forceinline__ uint64_t grid_voxel(const mdx_grid_dev& G, float qx, float qy, float qz,
                                               uint32_t* cell = nullptr) {
  int i = trunc_quotient(qx, G.spacing[0], G.inv_spacing[0]);
  int j = trunc_quotient(qy, G.spacing[1], G.inv_spacing[1]);
  int k = trunc_quotient(qz, G.spacing[2], G.inv_spacing[2]);
  i = i >= G.n[0] ? G.n[0] - 1 : i;
  j = j >= G.n[1] ? G.n[1] - 1 : j;
  k = k >= G.n[2] ? G.n[2] - 1 : k;
  if (cell)  // the voxel's macro cell (local majorants)
    *cell = (uint32_t)(i >> MDX_MACRO_SHIFT) +
            (uint32_t)G.mcx * ((uint32_t)(j >> MDX_MACRO_SHIFT) + (uint32_t)G.mcy * (uint32_t)(k >> MDX_MACRO_SHIFT));
  return (uint64_t)i + (uint64_t)G.n[0] * ((uint64_t)j + (uint64_t)G.n[1] * (uint64_t)k);
}

__device__ __forceinline__ int voxel_lookup(const mdx_domain_dev* D, float x, float y, float z, uint64_t& lin,
                                            uint32_t* cell = nullptr) {
  for (int g = 0; g < D->n_grids; ++g) {
    const mdx_grid_dev& G = D->grids[g];
    float qx = x - G.origin[0], qy = y - G.origin[1], qz = z - G.origin[2];
    if (qx >= 0.0f && qy >= 0.0f && qz >= 0.0f && qx <= G.dim[0] && qy <= G.dim[1] && qz <= G.dim[2]) {
      lin = grid_voxel(G, qx, qy, qz, cell);
      return g;
    }
  }
  return -1;
}

// physics_engine.cpp:180-183: a step is a delta (virtual) collision when
// u < 1 - sigma_tot / sigma_max; the threshold is formed with the IEEE
// quotient and difference, in that order, by both transport kernels.
__device__ __forceinline__ float delta_threshold(float s_tot, float s_max) { return 1.0f - s_tot / s_max; }
__device__ __forceinline__ bool real_collision(float u, float thr) { return !(u < thr); }

// physics_engine.cpp:82-109: u < p_coh -> Rayleigh (1); u < p_coh + p_inc ->
// Compton (2); otherwise photoelectric (0).  p = sigma / sigma_tot with IEEE
// division, compared in the reference's order.
__device__ __forceinline__ int select_type(float s_coh, float s_inc, float s_tot, float u) {
  float pc = s_coh / s_tot;
  float pi = s_inc / s_tot;
  if (u < pc) return 1;
  if (u < pc + pi) return 2;
  return 0;
}

// particle.cpp:7-45: d' = R(phi about d) R(theta about n) d with
// n = normalize(e_min x d), e_min the axis of the smallest |d_i|.  The
// quaternion product reduces to d' = cos(t) d + sin(t) (cos(p) n x d + sin(p) n).
__device__ __forceinline__ void rotate(float& dx, float& dy, float& dz, float mu, float phi) {
  float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
  float nx, ny, nz;
  if (ax <= ay && ax <= az) {  // e_x x d
    nx = 0.0f; ny = -dz; nz = dy;
  } else if (ay <= az) {       // e_y x d
    nx = dz; ny = 0.0f; nz = -dx;
  } else {                     // e_z x d
    nx = -dy; ny = dx; nz = 0.0f;
  }
  float inv = rsqrtf(nx * nx + ny * ny + nz * nz);
  nx *= inv; ny *= inv; nz *= inv;
  float cx = ny * dz - nz * dy, cy = nz * dx - nx * dz, cz = nx * dy - ny * dx;  // n x d
  float st = sqrtf(fmaxf(0.0f, 1.0f - mu * mu));
  const float sp = __sinf(phi), cp = __cosf(phi);  // hardware sine and cosine, phi in [0, 2 pi)
  float ox = mu * dx + st * (cp * cx + sp * nx);
  float oy = mu * dy + st * (cp * cy + sp * ny);
  float oz = mu * dz + st * (cp * cz + sp * nz);
  float r = rsqrtf(ox * ox + oy * oy + oz * oz);
  dx = ox * r; dy = oy * r; dz = oz * r;
}

// ------------------------------------------------------------------ RITA
// probability_dist.cpp:90-119 on one energy row, in w = 1 + x: the largest
// node i with y_i < y (lower_bound - 1, the CDF values taken in FP64), then
// the rational interpolation on [y_i, y_{i+1}].  The row's guide entry for y
// gives a start node with y_i < y; the next node's CDF value, loaded beside
// the start node, tells whether it brackets y, and a short walk follows when
// it does not.
__device__ __forceinline__ mdx_rita_row gld(const mdx_rita_row* p) {
  const mdx_v4 a = *(const MDX_GLOBAL mdx_v4*)p;
  const mdx_v4 b = *(const MDX_GLOBAL mdx_v4*)((const MDX_GLOBAL float*)p + 4);
  mdx_rita_row r;
  r.w_hi = a.x; r.w_lo = a.y; r.y_hi = a.z; r.y_lo = a.w;
  r.a = b.x; r.b = b.y; r.dw = b.z; r.d = b.w;
  return r;
}

// y > y_k for the FP64 CDF value y_k = yh + yl (yh = y_k rounded to
// nearest): exact for an FP32 y, since no FP32 value lies strictly between
// yh and y_k.
__device__ __forceinline__ bool rita_above(float y, float yh, float yl) { return y > yh || (y == yh && yl < 0.0f); }

struct RitaAt {
  int i;
  mdx_rita_row r;  // node i (clamped to the row)
  float yn_hi, yn_lo;  // CDF value of node i + 1 (clamped)
};

__device__ __forceinline__ int rita_guide(const mdx_tables_dev* T, int grow, float y) {
  int j = (int)(y * (float)MDX_RITA_GUIDE);  // exact scaling: the guide's bins are [j / G, (j + 1) / G)
  j = j < 0 ? 0 : (j > MDX_RITA_GUIDE - 1 ? MDX_RITA_GUIDE - 1 : j);
  return (int)gld(T->rita_guide + (size_t)grow * MDX_RITA_GUIDE + j) - 1;
}

__device__ __forceinline__ RitaAt rita_nodes(const mdx_tables_dev* T, int row0, int m, int i) {
  const mdx_rita_row* R = T->rita_rows + row0;
  const int ia = i < 0 ? 0 : i, ib = i + 1 > m - 1 ? m - 1 : i + 1;
  const mdx_v2 yn = *(const MDX_GLOBAL mdx_v2*)(&R[ib].y_hi);
  return RitaAt{i, gld(R + ia), yn.x, yn.y};
}

// Returns w as an unevaluated pair hi + lo (TwoSum of the last addition).
__device__ __forceinline__ mdx_v2 rita_finish(const mdx_tables_dev* T, int row0, int m, float y, RitaAt r) {
  if (r.i + 1 < m && rita_above(y, r.yn_hi, r.yn_lo)) {  // y lies beyond node i + 1: walk on
    const mdx_rita_row* R = T->rita_rows + row0;
    int i = r.i + 1;
    while (i + 1 < m && rita_above(y, gld(&R[i + 1].y_hi), gld(&R[i + 1].y_lo))) ++i;
    r = rita_nodes(T, row0, m, i);
  }
  if (r.i == m - 1) return mdx_v2{2.0f, 0.0f};  // above the last CDF value: x_max (probability_dist.cpp:97-100)
  if (r.i < 0) return mdx_v2{0.0f, 0.0f};       // x_min
  // nu = y - y_i with y_i's low part, rho = nu / d; the interval's fraction
  // (1 + a + b) d nu / (d^2 + a d nu + b nu^2) of dw, divided through by d^2
  // (the CDF widths of high-energy rows are ~1e-26: d^2 would underflow),
  // then w_i + fraction * dw with w_i's low part added before the last rounding
  const mdx_rita_row& q = r.r;
  const float rho = ((y - q.y_hi) - q.y_lo) / q.d;
  const float t = (1.0f + q.a + q.b) * rho / (1.0f + rho * (q.a + q.b * rho));
  const float c = __builtin_fmaf(t, q.dw, q.w_lo);
  const float sum = __fadd_rn(q.w_hi, c);  // TwoSum: sum + err = w_hi + c exactly
  const float bb = __fsub_rn(sum, q.w_hi);
  const float err = __fadd_rn(__fsub_rn(q.w_hi, __fsub_rn(sum, bb)), __fsub_rn(c, bb));
  return mdx_v2{sum, err};
}

// probability_dist.cpp:73-88: energy bracket by lower_bound - 1 on the
// incoherent-CS grid (the largest i with E_i < E, clamped to the last
// interval), x at both energies from the same y, then log-log interpolation
// of (x + 1) in E, here w = w_b + w_b expm1(g ln(w_u / w_l)) from the
// nearer energy b (g = f or f - 1) with w_l and w_u as hi + lo pairs (their
// low parts in the ratio and in the sum), so the result takes one rounding
// beyond the small correction's.  The stored energies
// are rounded down, so E > stored <=> E > E_i and the bracket is the FP64
// one; f = (ln(E / stored) + ln(stored / E_i)) / ln(E_{i+1} / E_i).  The
// energy hash gives the bracket, which holds unless its bin contains more
// grid energies; everything that depends only on the bracket (the two
// energies, the row offsets, both rows' guide entries) is loaded at once on
// that assumption and reloaded when the check fails.
__device__ __forceinline__ float rita_sample(const mdx_tables_dev* T, const mdx_material_dev& M, float E, float y) {
  const float* Eg = T->rita_E + M.rita_eoff;
  const int* ro = T->rita_row_off + M.rita_roff;
  const int ne = M.rita_ne;
  int b = (int)((logf(E) - M.rita_lnE0) * M.rita_hinv);
  b = b < 0 ? 0 : (b >= M.rita_nh ? M.rita_nh - 1 : b);
  int il = gld(T->rita_ehash + M.rita_hoff + b);
  il = il < 0 ? 0 : (il > ne - 2 ? ne - 2 : il);
  float e0 = gld(Eg + il), e1 = gld(Eg + il + 1);
  int r0 = gld(ro + il), r1 = gld(ro + il + 1), r2 = gld(ro + il + 2);
  int jl = rita_guide(T, M.rita_eoff + il, y), ju = rita_guide(T, M.rita_eoff + il + 1, y);
  if (!((e0 < E || il == 0) && (E <= e1 || il == ne - 2))) {  // the hash bin holds more energies
    int i = il;
    while (i >= 0 && gld(Eg + i) >= E) --i;
    while (i + 1 < ne && gld(Eg + i + 1) < E) ++i;
    il = i < 0 ? 0 : (i > ne - 2 ? ne - 2 : i);
    e0 = gld(Eg + il);
    r0 = gld(ro + il);
    r1 = gld(ro + il + 1);
    r2 = gld(ro + il + 2);
    jl = rita_guide(T, M.rita_eoff + il, y);
    ju = rita_guide(T, M.rita_eoff + il + 1, y);
  }
  const int ml = r1 - r0, mu = r2 - r1;
  const mdx_v2 wl = rita_finish(T, r0, ml, y, rita_nodes(T, r0, ml, jl));
  const mdx_v2 wu = rita_finish(T, r1, mu, y, rita_nodes(T, r1, mu, ju));
  if (!(wl.x > 0.0f && wu.x > 0.0f)) return wl.x - 1.0f;
  const mdx_v2 es = *(const MDX_GLOBAL mdx_v2*)(T->rita_escale + 2 * (M.rita_eoff + il));
  const float f = (log_ratio(E, e0) + es.y) * es.x;
  const float lr = log_ratio(wu.x, wl.x) + (wu.y * frcp(wu.x) - wl.y * frcp(wl.x));  // ln(w_u / w_l)
  // from the nearer energy's w: w = w_u exp((f - 1) ln(w_u / w_l)) above the
  // middle (f - 1 exact there), so that expm1's argument stays within half
  // the interval's ln ratio and its last-place error stays small against w
  const bool upper = f > 0.5f;
  const mdx_v2 wb = upper ? wu : wl;
  const float e = expm1f((upper ? f - 1.0f : f) * lr);
  const float w = wb.x + __builtin_fmaf(wb.x, e, __builtin_fmaf(wb.y, e, wb.y));
  return w - 1.0f;
}

}  // namespace mdxd
