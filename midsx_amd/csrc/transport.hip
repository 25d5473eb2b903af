// Photon-transport kernels for gfx950.
//
// k_transport_plain: one history per thread (grid-stride), the Woodcock
// delta-tracking loop of src/MIDSX/Core/physics_engine.cpp:17-79 in FP32 with
// each step's tallies processed as the step is taken (the reference records
// steps and replays them after the history, physics_engine.cpp:192-210; the
// per-step measurements are independent, so the order does not matter).
// Scalar tallies are pre-reduced in LDS and flushed once per block; voxel
// dose, images and spectra go straight to FP64 global atomics.
#include <hip/hip_runtime.h>

#include <mutex>

#include "device.h"
#include "kernels.h"

using namespace mdxd;

namespace {

constexpr int kBlock = 256;
// Bound on Woodcock steps per history: far above any physical history (the
// configurations here take tens of steps), it guarantees every lane exits.
constexpr uint32_t kMaxSteps = 1u << 20;
constexpr int kAccPerTally = MDX_NQ * MDX_NBUCKET * MDX_NFIELD;

__device__ __forceinline__ int bucket_of(int nc, int ni) {
  // *_quantity.cpp measure(): total > 1 -> multiple; coherent == 1; incoherent == 1; else primary
  return (nc + ni) > 1 ? 3 : (nc == 1 ? 2 : (ni == 1 ? 1 : 0));
}

__device__ __forceinline__ void lds_event(double* acc, int t, int q, int b, double v) {
  double* p = acc + ((t * MDX_NQ + q) * MDX_NBUCKET + b) * MDX_NFIELD;
  atomicAdd(p, 1.0);
  atomicAdd(p + 1, v);
  atomicAdd(p + 2, v * v);
}

struct Step {
  float px, py, pz;  // initial position
  float dx, dy, dz;  // initial direction
  float s;           // free path
  float E;           // initial energy
  int nc0, ni0;      // scatter counts of the initial photon
  int ncf, nif;      // scatter counts of the final photon (0 unless interacted)
  float dep;
  bool interact;
};

// LDS budget for block-level IncidentEnergy histograms (doubles); tallies whose
// histograms do not fit go straight to global atomics.
constexpr int kLdsHist = 2048;       // plain kernel
constexpr int kQueueLdsHist = 512;   // queue kernel (its LDS also holds descriptors and cross-section caches)
// The queue kernel keeps, per lane, the total cross sections of material
// slots 0 .. kSigmaCache - 2 and the global majorant (entry kSigmaCache - 1)
// at the photon's current energy once evaluated.
constexpr int kSigmaCache = 9;

// Offsets of each tally's histogram in the block's LDS histogram of cap
// doubles (-1: global).
__device__ __forceinline__ void hist_layout(const mdx_tallies_dev* TL, int* off, int cap = kLdsHist) {
  int used = 0;
  for (int t = 0; t < TL->n; ++t) {
    const mdx_tally_dev& T = TL->t[t];
    int need = (T.kind != MDX_TALLY_AACUBOID && T.hist_bins > 0) ? MDX_NBUCKET * T.hist_bins : 0;
    if (need > 0 && used + need <= cap) {
      off[t] = used;
      used += need;
    } else {
      off[t] = -1;
    }
  }
}

__device__ void flush_hist(const mdx_tallies_dev* TL, const int* off, double* lhist) {
  for (int t = 0; t < TL->n; ++t) {
    if (off[t] < 0) continue;
    const mdx_tally_dev& T = TL->t[t];
    for (int i = threadIdx.x; i < MDX_NBUCKET * T.hist_bins; i += blockDim.x) {
      double v = lhist[off[t] + i];
      if (v != 0.0) gadd(T.hist + i, v);
    }
  }
}

// surface_tally.cpp:16-34 (+ DiscSurfaceTally :63-69, RectangularSurfaceTally
// :100-112) and volume_tally.cpp:5-91 for one recorded step.
// FP32 atomic add that keeps the rounding error: the atomic returns the old
// value, so the stored sum is fl(old + v) and TwoSum gives the exact error,
// which goes to p[1].  hi + lo then carries the sum to the precision of the
// error sum lo, itself an FP32 sum of values below half an ulp of hi.
// GLOBAL: p is in global memory (else in LDS).
template <bool GLOBAL>
__device__ __forceinline__ void compensated_add(float* p, float v) {
  const float old = GLOBAL ? gadd(p, v) : atomicAdd(p, v);
  const float s = __fadd_rn(old, v);
  const float bb = __fsub_rn(s, old);
  const float err = __fadd_rn(__fsub_rn(old, __fsub_rn(s, bb)), __fsub_rn(v, bb));
  if (err != 0.0f) {
    if (GLOBAL) gadd(p + 1, err);
    else atomicAdd(p + 1, err);
  }
}

// One energy deposit in voxel lin of grid G (physics_engine.cpp:70-75,
// 160-168) in the grid's accumulation precision.
__device__ __forceinline__ void voxel_deposit(const mdx_grid_dev& G, uint64_t lin, float dep,
                                              float* ldose = nullptr) {
  if (G.dose_mode == MDX_DOSE_F32) {
    if (G.dose32) {
      gadd(G.dose32 + 2 * lin, dep);
      gadd(G.dose32 + 2 * lin + 1, dep * dep);
    }
  } else if (G.dose_mode == MDX_DOSE_F32C || G.dose_mode == MDX_DOSE_F32C_LDS) {
    if (G.dose32) {
      if (ldose) compensated_add<false>(ldose + 2 * lin, dep);  // this block's sum, in LDS
      else compensated_add<true>(G.dose32 + 4 * lin, dep);
      const float q = __fmul_rn(dep, dep);  // dep^2 = q + ql exactly (never contracted into an fma)
      const float ql = __builtin_fmaf(dep, dep, -q);
      compensated_add<true>(G.dose32 + 4 * lin + 2, q);
      if (ql != 0.0f) gadd(G.dose32 + 4 * lin + 3, ql);
    }
  } else if (G.dose) {
    double* r = G.dose + MDX_DOSE_REC * lin;
    gadd(r, (double)dep);
    gadd(r + 1, (double)dep * (double)dep);
    gadd(reinterpret_cast<unsigned long long*>(r + 2), 1ull);
    return;
  }
  if (G.dose_n) gadd(G.dose_n + lin, 1ull);
}

__device__ void process_tallies(const mdx_tallies_dev* TL, double* acc, const Step& r, uint32_t& surf_hits,
                                double* lhist, const int* hoff) {
  for (int t = 0; t < TL->n; ++t) {
    const mdx_tally_dev& T = TL->t[t];
    if (T.kind == MDX_TALLY_AACUBOID) {
      float pos[3] = {r.px, r.py, r.pz}, dir[3] = {r.dx, r.dy, r.dz};
      float enter = -INFINITY, exit = INFINITY;
      for (int i = 0; i < 3; ++i) {
        if (fabsf(dir[i]) < kEps) continue;
        float a = (T.p0[i] - pos[i]) / dir[i];
        float b = (T.p1[i] - pos[i]) / dir[i];
        if (a > b) { float tmp = a; a = b; b = tmp; }
        enter = fmaxf(enter, a);
        exit = fminf(exit, b);
      }
      if (enter > exit || exit < 0.0f || enter > r.s) continue;  // MISSES
      bool counted = false;
      float dep = r.dep;
      if (enter <= 0.0f) {
        counted = true;                  // STARTS_INSIDE_*
        if (!(exit >= r.s)) dep = 0.0f;  // ..._EXITS
      } else if (exit < r.s) {
        dep = 0.0f;                      // PASSES_THROUGH
      }                                  // else LANDS_INSIDE
      int b = bucket_of(r.ncf, r.nif);   // volume quantities use the final photon
      if (dep != 0.0f) lds_event(acc, t, MDX_VQ_ENERGY_DEPOSITION, b, (double)dep);
      if (!counted) {
        if (r.E != 0.0f) lds_event(acc, t, MDX_VQ_INCIDENT_ENERGY, b, (double)r.E);
        lds_event(acc, t, MDX_VQ_NUMBER_OF_PHOTONS, b, 1.0);
      }
      if (r.interact) lds_event(acc, t, MDX_VQ_NUMBER_OF_INTERACTIONS, b, 1.0);
    } else {
      float cosv = r.dx * T.nrm[0] + r.dy * T.nrm[1] + r.dz * T.nrm[2];
      if (fabsf(cosv) < kEps || cosv < 0.0f) continue;
      float tt = ((T.p0[0] - r.px) * T.nrm[0] + (T.p0[1] - r.py) * T.nrm[1] + (T.p0[2] - r.pz) * T.nrm[2]) / cosv;
      if (tt > r.s || tt < 0.0f) continue;
      float ix = r.px + tt * r.dx - T.p0[0];
      float iy = r.py + tt * r.dy - T.p0[1];
      float iz = r.pz + tt * r.dz - T.p0[2];
      float a = 0.0f, bb = 0.0f;
      if (T.kind == MDX_TALLY_DISC) {
        if (sqrtf(ix * ix + iy * iy + iz * iz) > T.radius) continue;
      } else {
        a = T.p1[0] * ix + T.p1[1] * iy + T.p1[2] * iz;
        bb = T.p2[0] * ix + T.p2[1] * iy + T.p2[2] * iz;
        if (!(a >= 0.0f && a <= T.e1e1 && bb >= 0.0f && bb <= T.e2e2)) continue;
      }
      ++surf_hits;
      int b = bucket_of(r.nc0, r.ni0);  // surface quantities use the initial photon
      if (T.kind == MDX_TALLY_IMAGE) {
        int px = (int)(a / T.e1e1 * (float)T.nx);
        int py = (int)(bb / T.e2e2 * (float)T.ny);
        px = px >= T.nx ? T.nx - 1 : (px < 0 ? 0 : px);
        py = py >= T.ny ? T.ny - 1 : (py < 0 ? 0 : py);
        double* im = T.image + 3 * ((size_t)py * T.nx + px);
        gadd(im, (double)r.E);
        gadd(im + 1, (double)r.E * (double)r.E);
        gadd(im + 2, 1.0);
      }
      if (r.E != 0.0f) lds_event(acc, t, MDX_SQ_INCIDENT_ENERGY, b, (double)r.E);
      lds_event(acc, t, MDX_SQ_ENTRANCE_COSINE, b, (double)cosv);
      lds_event(acc, t, MDX_SQ_NUMBER_OF_PHOTONS, b, 1.0);
      if (T.hist_bins > 0) {
        float fb = (r.E - T.hist_lo) * T.hist_inv_w;
        if (fb >= 0.0f && fb < (float)T.hist_bins) {
          const int bin = b * T.hist_bins + (int)fb;
          if (hoff[t] >= 0) atomicAdd(lhist + hoff[t] + bin, 1.0);
          else gadd(T.hist + bin, 1.0);
        }
      }
    }
  }
}

// source.cpp:150-155 and the directionality / spectrum samplers it calls.
// RNG order: position (point: none), direction, energy.
__device__ void generate(const mdx_source_dev* S, Philox& rng, float& px, float& py, float& pz, float& dx,
                         float& dy, float& dz, float& E) {
  px = S->pos[0];
  py = S->pos[1];
  pz = S->pos[2];
  float vx, vy, vz;
  switch (S->dir_kind) {
    case MDX_DIR_ISOTROPIC: {  // source.cpp:64-70 (theta ~ U[0, pi])
      float th = kPiF * rng.uniform();
      float ph = 2.0f * kPiF * rng.uniform();
      float st, ct, sp, cp;
      sincosf(th, &st, &ct);
      sincosf(ph, &sp, &cp);
      vx = st * cp; vy = st * sp; vz = ct;
      break;
    }
    case MDX_DIR_RECT_ISOTROPIC: {  // source.cpp:85-90
      float a1 = rng.uniform(), a2 = rng.uniform();
      vx = S->a[0] + a1 * S->b[0] + a2 * S->c[0] - px;
      vy = S->a[1] + a1 * S->b[1] + a2 * S->c[1] - py;
      vz = S->a[2] + a1 * S->b[2] + a2 * S->c[2] - pz;
      break;
    }
    case MDX_DIR_DISC_ISOTROPIC: {  // source.cpp:108-115 (r ~ U[0, R])
      float th = 2.0f * kPiF * rng.uniform();
      float rr = S->radius * rng.uniform();
      float st, ct;
      sincosf(th, &st, &ct);
      vx = S->a[0] + rr * (S->b[0] * ct + S->c[0] * st) - px;
      vy = S->a[1] + rr * (S->b[1] * ct + S->c[1] * st) - py;
      vz = S->a[2] + rr * (S->b[2] * ct + S->c[2] * st) - pz;
      break;
    }
    default:  // beam, source.cpp:74-77
      vx = S->a[0] - px; vy = S->a[1] - py; vz = S->a[2] - pz;
  }
  float inv = rsqrtf(vx * vx + vy * vy + vz * vz);
  dx = vx * inv; dy = vy * inv; dz = vz * inv;
  if (S->spec_kind == MDX_SPEC_POLY) {  // probability_dist.cpp:18-28
    // lower_bound of u in the CDF, searched between the guide entries of
    // u's bin [j / G, (j + 1) / G): at least guide[j] values lie below u and
    // at most guide[j + 1] (all n for the last bin)
    const float u = rng.uniform();
    int j = (int)(u * (float)MDX_SPEC_GUIDE);
    j = j > MDX_SPEC_GUIDE - 1 ? MDX_SPEC_GUIDE - 1 : j;
    const int n = S->n_bins;
    int lo = S->guide[j], hi = j + 1 < MDX_SPEC_GUIDE ? S->guide[j + 1] : n;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (gld(S->cdf + mid) < u) lo = mid + 1; else hi = mid;
    }
    lo = lo >= n ? n - 1 : lo;
    E = gld(S->energies + lo);
  } else {
    E = S->energy;
  }
}

// photon_interactions.cpp:62-141 (Kahn's method + S(x) rejection).  Returns
// the scattered energy and sets mu.  S(x_max) depends only on E, so it is
// evaluated once per event instead of once per trial.  The loop is bounded:
// after kMaxKahnTrials rejections (probability far below 1e-100 for any
// acceptance rate above 0.1) the last candidate is taken, so a corrupt table
// cannot hang the kernel.
constexpr uint32_t kMaxKahnTrials = 4096;
// The argument of S(x_max) in Kahn's S-test at energy E.
__device__ __forceinline__ float compton_xmax(float E) { return kAlphaSqrt2 * (E * kInvMc2); }

// s_max: S(compton_xmax(E)) when the caller has it (evaluated beside other
// table lookups), else a negative value.
__device__ float compton(const mdx_tables_dev* T, const mdx_material_dev& M, Philox& rng, float E, float& mu,
                         uint32_t& trials, uint32_t& s_tests, float s_max = -1.0f) {
  const float k = E * kInvMc2;
  const float ik = frcp(k), k2 = 1.0f + 2.0f * k;
  const float r1p = k2 * frcp(9.0f + 2.0f * k);
  const float is_max = frcp(s_max >= 0.0f ? s_max : curve_eval(T, M.sf, compton_xmax(E)));
  // Trial it draws (r1, r2, r3) for Kahn's test and, once that passes (or
  // the trial cap is reached), r4 for the S(x)/S_max test.  The Kahn draws
  // loop on their own so that a wave runs the table lookup of the S-test
  // once per round of S-tests rather than once per Kahn trial.
  uint32_t it = 0;
  for (;;) {
    float chi;
    for (;; ++it) {
      ++trials;
      const float r1 = rng.uniform(), r2 = rng.uniform(), r3 = rng.uniform();
      if (r1 <= r1p) {
        chi = 1.0f + 2.0f * k * r2;
        const float ic = frcp(chi);
        if (r3 <= 4.0f * (ic - ic * ic) || it >= kMaxKahnTrials) break;
      } else {
        chi = k2 * frcp(1.0f + 2.0f * k * r2);
        const float q = (1.0f - chi) * ik + 1.0f;
        if (2.0f * r3 <= q * q + frcp(chi) || it >= kMaxKahnTrials) break;
      }
    }
    ++s_tests;
    const float one_minus_mu = fminf((chi - 1.0f) * ik, 2.0f);  // 1 - mu, mu = 1 + 1/k - 1/k'
    const float x = kAlpha * k * sqrtf(one_minus_mu);
    const float sx = curve_eval(T, M.sf, x);
    if (rng.uniform() <= sx * is_max || it >= kMaxKahnTrials) {
      mu = 1.0f - one_minus_mu;
      return E * frcp(chi);
    }
    ++it;
  }
}

__global__ void __launch_bounds__(kBlock)
k_transport_plain(const mdx_tables_dev* __restrict__ T, const mdx_domain_dev* __restrict__ D,
                  const mdx_source_dev* __restrict__ S, const mdx_tallies_dev* __restrict__ TL, uint64_t first,
                  uint64_t n, uint64_t seed, uint32_t job, unsigned long long* __restrict__ counters) {
  __shared__ double acc[MDX_MAX_TALLIES * kAccPerTally];
  __shared__ double lhist[kLdsHist];
  __shared__ unsigned long long cnt[MDX_NCOUNTER];
  int hoff[MDX_MAX_TALLIES];
  hist_layout(TL, hoff);
  const int ntl = TL->n;
  for (int i = threadIdx.x; i < ntl * kAccPerTally; i += blockDim.x) acc[i] = 0.0;
  for (int i = threadIdx.x; i < kLdsHist; i += blockDim.x) lhist[i] = 0.0;
  if (threadIdx.x < MDX_NCOUNTER) cnt[threadIdx.x] = 0;
  __syncthreads();

  uint32_t c_hist = 0, c_step = 0, c_real = 0, c_ray = 0, c_comp = 0, c_photo = 0, c_trial = 0, c_stest = 0,
           c_dep = 0, c_surf = 0, c_esc = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; h < n; h += stride) {
    Philox rng;
    rng.init(seed, job, first + h);
    float px, py, pz, dx, dy, dz, E;
    generate(S, rng, px, py, pz, dx, dy, dz, E);
    ++c_hist;
    int nc = 0, ni = 0;
    float smax = curve_eval(T, T->majorant, E);
    int cslot = -1;
    float stot = 0.0f;
    for (uint32_t it = 0; it < kMaxSteps; ++it) {
      ++c_step;
      Step r;
      r.px = px; r.py = py; r.pz = pz;
      r.dx = dx; r.dy = dy; r.dz = dz;
      r.E = E;
      r.nc0 = nc; r.ni0 = ni;
      r.ncf = 0; r.nif = 0;
      r.dep = 0.0f;
      r.interact = false;
      // physics_engine.cpp:38-44: s = -ln(u) / sigma_max, move
      float s = -logf(rng.uniform()) / smax;
      r.s = s;
      px = px + s * dx;
      py = py + s * dy;
      pz = pz + s * dz;
      if (!in_domain(D, px, py, pz)) {  // :47-52
        ++c_esc;
        process_tallies(TL, acc, r, c_surf, lhist, hoff);
        break;
      }
      uint64_t lin = 0;
      int g = voxel_lookup(D, px, py, pz, lin);
      int slot = g < 0 ? D->bg_slot : (int)T->id_to_slot[gld(D->grids[g].ids + lin)];
      const mdx_material_dev& M = T->mat[slot];
      if (slot != cslot) {
        stot = curve_eval(T, M.total, E);
        cslot = slot;
      }
      bool alive = true;
      // :180-183 delta scatter if u < 1 - sigma_tot / sigma_max
      if (real_collision(rng.uniform(), delta_threshold(stot, smax))) {
        ++c_real;
        r.interact = true;
        float sc, si;
        curve_eval2(T, M.coh, E, M.incoh, E, sc, si);
        int type = select_type(sc, si, stot, rng.uniform());
        float dep = 0.0f;
        if (type == 1) {  // photon_interactions.cpp:50-60
          ++c_ray;
          ++nc;
          float mu = rita_sample(T, M, E, rng.uniform());
          rotate(dx, dy, dz, mu, 2.0f * kPiF * rng.uniform());
        } else if (type == 2) {  // :62-88
          ++c_comp;
          ++ni;
          float mu;
          float Ep = compton(T, M, rng, E, mu, c_trial, c_stest);
          rotate(dx, dy, dz, mu, 2.0f * kPiF * rng.uniform());
          dep = E - Ep;
          E = Ep;
          smax = curve_eval(T, T->majorant, E);
          cslot = -1;
        } else {  // :43-48
          ++c_photo;
          dep = E;
          alive = false;
        }
        r.ncf = nc;
        r.nif = ni;
        r.dep = dep;
        if (dep != 0.0f) {  // physics_engine.cpp:70-75, 205-208
          ++c_dep;
          if (g >= 0) {
            voxel_deposit(D->grids[g], lin, dep);
          } else if (D->bg_dose) {
            gadd(D->bg_dose, (double)dep);
            gadd(D->bg_dose + 1, (double)dep * (double)dep);
            gadd(D->bg_dose + 2, 1.0);
          }
        }
      }
      process_tallies(TL, acc, r, c_surf, lhist, hoff);
      if (!alive) break;
    }
  }
  atomicAdd(&cnt[MDX_C_HIST], (unsigned long long)c_hist);
  atomicAdd(&cnt[MDX_C_STEP], (unsigned long long)c_step);
  atomicAdd(&cnt[MDX_C_REAL], (unsigned long long)c_real);
  atomicAdd(&cnt[MDX_C_RAYLEIGH], (unsigned long long)c_ray);
  atomicAdd(&cnt[MDX_C_COMPTON], (unsigned long long)c_comp);
  atomicAdd(&cnt[MDX_C_PHOTO], (unsigned long long)c_photo);
  atomicAdd(&cnt[MDX_C_KAHN_TRIAL], (unsigned long long)c_trial);
  atomicAdd(&cnt[MDX_C_KAHN_S], (unsigned long long)c_stest);
  atomicAdd(&cnt[MDX_C_DEPOSIT], (unsigned long long)c_dep);
  atomicAdd(&cnt[MDX_C_SURFHIT], (unsigned long long)c_surf);
  atomicAdd(&cnt[MDX_C_ESCAPE], (unsigned long long)c_esc);
  __syncthreads();
  for (int i = threadIdx.x; i < ntl * kAccPerTally; i += blockDim.x) {
    double v = acc[i];
    if (v != 0.0) {
      int t = i / kAccPerTally;
      gadd(TL->t[t].acc + (blockIdx.x % MDX_ACC_COPIES) * kAccPerTally + (i - t * kAccPerTally), v);
    }
  }
  flush_hist(TL, hoff, lhist);
  if (threadIdx.x < MDX_NCOUNTER && cnt[threadIdx.x]) atomicAdd(counters + threadIdx.x, cnt[threadIdx.x]);
}

// ------------------------------------------------------------------ queue kernel
// Persistent transport with per-lane refill: every lane carries one photon;
// when its history ends it takes the next history index from a global work
// counter (one wave-aggregated atomic per refill), so a wave never waits for
// its slowest history and the only tail is the launch's last histories.
// MULTI: the work space is the concatenation of several jobs' history slices
// (prefix[] holds their offsets); each job has its own scene and tallies and
// shares the tables.  Scalar tallies are pre-reduced in LDS for one job and
// go to global FP64 atomics for several.
// A lane's value read by the whole wave (v_readlane: a scalar result, no LDS
// round trip as a shuffle takes); src must be wave-uniform and active.
// The value of one lane (a wave-uniform lane index), known to the compiler
// as wave-uniform.
__device__ __forceinline__ unsigned long long readlane_u64(unsigned long long v, int src) {
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, src);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), src);
  return ((unsigned long long)hi << 32) | lo;
}

__device__ __forceinline__ double readlane_f64(double v, int src) {
  const unsigned long long b = __double_as_longlong(v);
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)b, src);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(b >> 32), src);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ double* readlane_ptr(double* p, int src) {
  const uintptr_t b = (uintptr_t)p;
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)b, src);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)((unsigned long long)b >> 32), src);
  return (double*)(((uintptr_t)hi << 32) | lo);
}

// One scalar-tally event (count 1, value v, v^2) at acc[q][b].  The active
// lanes of the wave that hit the same record are summed first (their values
// are read with shuffles from active lanes only) and one of them adds the
// sums: a job's tally records are few, so without this the atomics of a
// whole wave would queue on one L2 line.
__device__ __forceinline__ void event_global(double* acc, int q, int b, double v) {
  double* p = acc + (blockIdx.x % MDX_ACC_COPIES) * kAccPerTally + (q * MDX_NBUCKET + b) * MDX_NFIELD;
  const int lane = threadIdx.x & 63;
  const double v2 = v * v;
  unsigned long long todo = __ballot(true);
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;  // wave-uniform, as every source lane below
    double* const lp = readlane_ptr(p, leader);
    const unsigned long long same = __ballot(p == lp);
    todo &= ~same;
    double s = 0.0, s2 = 0.0;
    for (unsigned long long m = same; m; m &= m - 1) {
      const int src = __ffsll((long long)m) - 1;
      s += readlane_f64(v, src);
      s2 += readlane_f64(v2, src);
    }
    if (lane == leader) {
      gadd(p, (double)__popcll(same));
      gadd(p + 1, s);
      gadd(p + 2, s2);
    }
  }
}

// atomicAdd(p, v) with the active lanes that hit the same address summed
// first (spectrum bins: a monoenergetic beam puts every primary in one bin).
__device__ __forceinline__ void wave_add(double* p, double v) {
  const int lane = threadIdx.x & 63;
  unsigned long long todo = __ballot(true);
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    double* const lp = readlane_ptr(p, leader);
    const unsigned long long same = __ballot(p == lp);
    todo &= ~same;
    double s = 0.0;
    for (unsigned long long m = same; m; m &= m - 1) s += readlane_f64(v, __ffsll((long long)m) - 1);
    if (lane == leader) gadd(p, s);
  }
}

template <bool MULTI>
__device__ void tallies_step(const mdx_tallies_dev* TL, double* acc, const Step& r, uint32_t& surf_hits,
                             double* lhist, const int* hoff) {
  if (!MULTI) {
    process_tallies(TL, acc, r, surf_hits, lhist, hoff);
    return;
  }
  for (int t = 0; t < TL->n; ++t) {
    const mdx_tally_dev& T = TL->t[t];
    if (T.kind == MDX_TALLY_AACUBOID) {
      float pos[3] = {r.px, r.py, r.pz}, dir[3] = {r.dx, r.dy, r.dz};
      float enter = -INFINITY, exit = INFINITY;
      for (int i = 0; i < 3; ++i) {
        if (fabsf(dir[i]) < kEps) continue;
        float a = (T.p0[i] - pos[i]) / dir[i];
        float b = (T.p1[i] - pos[i]) / dir[i];
        if (a > b) { float tmp = a; a = b; b = tmp; }
        enter = fmaxf(enter, a);
        exit = fminf(exit, b);
      }
      if (enter > exit || exit < 0.0f || enter > r.s) continue;
      bool counted = false;
      float dep = r.dep;
      if (enter <= 0.0f) {
        counted = true;
        if (!(exit >= r.s)) dep = 0.0f;
      } else if (exit < r.s) {
        dep = 0.0f;
      }
      int b = bucket_of(r.ncf, r.nif);
      if (dep != 0.0f) event_global(T.acc, MDX_VQ_ENERGY_DEPOSITION, b, (double)dep);
      if (!counted) {
        if (r.E != 0.0f) event_global(T.acc, MDX_VQ_INCIDENT_ENERGY, b, (double)r.E);
        event_global(T.acc, MDX_VQ_NUMBER_OF_PHOTONS, b, 1.0);
      }
      if (r.interact) event_global(T.acc, MDX_VQ_NUMBER_OF_INTERACTIONS, b, 1.0);
    } else {
      float cosv = r.dx * T.nrm[0] + r.dy * T.nrm[1] + r.dz * T.nrm[2];
      if (fabsf(cosv) < kEps || cosv < 0.0f) continue;
      float tt = ((T.p0[0] - r.px) * T.nrm[0] + (T.p0[1] - r.py) * T.nrm[1] + (T.p0[2] - r.pz) * T.nrm[2]) / cosv;
      if (tt > r.s || tt < 0.0f) continue;
      float ix = r.px + tt * r.dx - T.p0[0];
      float iy = r.py + tt * r.dy - T.p0[1];
      float iz = r.pz + tt * r.dz - T.p0[2];
      float a = 0.0f, bb = 0.0f;
      if (T.kind == MDX_TALLY_DISC) {
        if (sqrtf(ix * ix + iy * iy + iz * iz) > T.radius) continue;
      } else {
        a = T.p1[0] * ix + T.p1[1] * iy + T.p1[2] * iz;
        bb = T.p2[0] * ix + T.p2[1] * iy + T.p2[2] * iz;
        if (!(a >= 0.0f && a <= T.e1e1 && bb >= 0.0f && bb <= T.e2e2)) continue;
      }
      ++surf_hits;
      int b = bucket_of(r.nc0, r.ni0);
      if (T.kind == MDX_TALLY_IMAGE) {
        int px = (int)(a / T.e1e1 * (float)T.nx);
        int py = (int)(bb / T.e2e2 * (float)T.ny);
        px = px >= T.nx ? T.nx - 1 : (px < 0 ? 0 : px);
        py = py >= T.ny ? T.ny - 1 : (py < 0 ? 0 : py);
        double* im = T.image + 3 * ((size_t)py * T.nx + px);
        gadd(im, (double)r.E);
        gadd(im + 1, (double)r.E * (double)r.E);
        gadd(im + 2, 1.0);
      }
      if (r.E != 0.0f) event_global(T.acc, MDX_SQ_INCIDENT_ENERGY, b, (double)r.E);
      event_global(T.acc, MDX_SQ_ENTRANCE_COSINE, b, (double)cosv);
      event_global(T.acc, MDX_SQ_NUMBER_OF_PHOTONS, b, 1.0);
      if (T.hist_bins > 0) {
        float fb = (r.E - T.hist_lo) * T.hist_inv_w;
        if (fb >= 0.0f && fb < (float)T.hist_bins) wave_add(T.hist + b * T.hist_bins + (int)fb, 1.0);
      }
    }
  }
}

// Distance along the flight line from (x, y, z) in direction d to the first
// point where a step without an interaction can score in any tally, less a
// margin that covers the FP32 drift of the stepped position from that line;
// +inf if there is none.  Such a step scores only where its segment crosses a
// surface tally's plane in the direction of the normal (surface_tally.cpp:
// 16-34) or enters a volume tally's box from outside (volume_tally.cpp:
// 38-60): starting inside a box it carries no deposit and no interaction, and
// a straight line enters a box once.  Steps short of this distance skip the
// tally replay; the step that reaches it is replayed in full.
__device__ float next_tally_event(const mdx_tallies_dev* TL, float x, float y, float z, float dx, float dy,
                                  float dz) {
  float e = INFINITY;
  for (int t = 0; t < TL->n; ++t) {
    const mdx_tally_dev& T = TL->t[t];
    if (T.kind == MDX_TALLY_AACUBOID) {
      const float pos[3] = {x, y, z}, dir[3] = {dx, dy, dz};
      float enter = -INFINITY, exit = INFINITY;
      for (int i = 0; i < 3; ++i) {
        if (fabsf(dir[i]) < kEps) continue;
        const float id = frcp(dir[i]);
        float a = (T.p0[i] - pos[i]) * id;
        float b = (T.p1[i] - pos[i]) * id;
        if (a > b) { float tmp = a; a = b; b = tmp; }
        enter = fmaxf(enter, a);
        exit = fminf(exit, b);
      }
      if (enter > exit || exit < 0.0f || enter <= 0.0f) continue;
      e = fminf(e, enter);
    } else {
      const float cosv = dx * T.nrm[0] + dy * T.nrm[1] + dz * T.nrm[2];
      if (fabsf(cosv) < kEps || cosv < 0.0f) continue;
      const float tt = ((T.p0[0] - x) * T.nrm[0] + (T.p0[1] - y) * T.nrm[1] + (T.p0[2] - z) * T.nrm[2]) * frcp(cosv);
      if (tt < -0.01f) continue;
      e = fminf(e, tt);
    }
  }
  return e - (0.01f + 1.0e-4f * fabsf(e));
}

// From a point (x, y, z) in the domain along d: the distance to the first
// face of a grid box the line enters (0 when the point already lies in a box,
// +inf when it meets none) and to the domain's boundary.
__device__ __forceinline__ void flight_bounds(const mdx_domain_dev* D, float x, float y, float z, float dx, float dy,
                                              float dz, float& dgrid, float& dexit) {
  const float pos[3] = {x, y, z}, dir[3] = {dx, dy, dz};
  float inv[3];
  dexit = INFINITY;
  for (int i = 0; i < 3; ++i) {
    inv[i] = frcp(dir[i]);  // +-inf for a zero component, as 1 / 0
    if (dir[i] > 0.0f) dexit = fminf(dexit, (D->dim_space[i] - pos[i]) * inv[i]);
    else if (dir[i] < 0.0f) dexit = fminf(dexit, -pos[i] * inv[i]);
  }
  dexit = fmaxf(dexit, 0.0f);
  dgrid = INFINITY;
  for (int g = 0; g < D->n_grids; ++g) {
    const mdx_grid_dev& G = D->grids[g];
    float enter = -INFINITY, exit = INFINITY;
    bool miss = false;
    for (int i = 0; i < 3; ++i) {
      const float lo = G.origin[i] - pos[i], hi = G.origin[i] + G.dim[i] - pos[i];
      if (dir[i] == 0.0f) {
        miss = miss || lo > 0.0f || hi < 0.0f;
        continue;
      }
      float a = lo * inv[i], b = hi * inv[i];
      if (a > b) { float t = a; a = b; b = t; }
      enter = fmaxf(enter, a);
      exit = fminf(exit, b);
    }
    if (!miss && enter <= exit && exit >= 0.0f) dgrid = fminf(dgrid, fmaxf(enter, 0.0f));
  }
}

// Waves per SIMD the queue kernel is compiled for (register budget).
#ifndef MDX_QUEUE_WAVES_PER_EU
#define MDX_QUEUE_WAVES_PER_EU 4
#endif

static_assert(sizeof(mdx_material_dev) % 4 == 0 && sizeof(mdx_domain_dev) % 4 == 0 &&
                  sizeof(mdx_source_dev) % 4 == 0 && sizeof(mdx_tallies_dev) % 4 == 0,
              "descriptors are copied to LDS in 4-byte words");
// An upper bound of the one-job queue kernel's static LDS (its shared arrays
// plus 1 KB for alignment), for hosts that cannot ask the runtime.
constexpr size_t kQueueStaticLdsBound =
    sizeof(double) * (MDX_MAX_TALLIES * kAccPerTally + kQueueLdsHist) + sizeof(unsigned long long) * MDX_NCOUNTER +
    sizeof(float) * kSigmaCache * kBlock + sizeof(mdx_curve) +
    sizeof(mdx_material_dev) * MDX_MAX_MATERIALS + 256 + sizeof(mdx_domain_dev) + sizeof(mdx_source_dev) +
    sizeof(mdx_tallies_dev) + 1024;

// Copies a descriptor (a multiple of 4 bytes) from global memory to LDS with
// the whole block; the caller synchronises the block before reading it.
__device__ __forceinline__ void block_copy(void* dst, const void* src, int bytes) {
  unsigned char* d = (unsigned char*)dst;
  const unsigned char* s = (const unsigned char*)src;
  for (int i = threadIdx.x; i < bytes / 4; i += blockDim.x) __builtin_memcpy(d + 4 * i, s + 4 * i, 4);
}

// The same with one (converged) wave; its lanes read the copy afterwards.
__device__ __forceinline__ void wave_copy(void* dst, const void* src, int bytes) {
  unsigned char* d = (unsigned char*)dst;
  const unsigned char* s = (const unsigned char*)src;
  for (int i = threadIdx.x & 63; i < bytes / 4; i += 64) __builtin_memcpy(d + 4 * i, s + 4 * i, 4);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A job's scene descriptors as the kernel keeps them in LDS.
struct JobDesc {
  mdx_domain_dev D;
  mdx_source_dev S;
  mdx_tallies_dev TL;
};

__device__ __forceinline__ void load_job(JobDesc* dst, const mdx_job_dev& j) {
  wave_copy(&dst->D, j.D, (int)sizeof(mdx_domain_dev));
  wave_copy(&dst->S, j.S, (int)sizeof(mdx_source_dev));
  wave_copy(&dst->TL, j.TL, (int)sizeof(mdx_tallies_dev));
}

constexpr int kWaves = kBlock / 64;

// Work distribution.  One job: lanes claim histories from one counter.
// Several jobs: each wave works on one job at a time, claiming from that
// job's counter (work[j]); the wave starts on the job its index falls in when
// waves are spread over the jobs in proportion to their histories, and when
// its job runs out and its last photons of it are done it moves to the next
// job with histories left.  A wave's photons thus always belong to the job
// whose descriptors it holds, whose address is the same for every lane: the
// descriptors are read as wave-uniform data.  (The wait for a job's last
// photons costs little: a wave runs thousands of histories per launch.)
template <bool MULTI>
__device__ __forceinline__ void transport_queue_body(const mdx_tables_dev* __restrict__ T,
                                                     const mdx_job_dev* __restrict__ jobs,
                                                     const uint64_t* __restrict__ prefix, int n_jobs, uint64_t total,
                                                     unsigned long long* __restrict__ work,
                                                     unsigned long long* __restrict__ counters, int refill_min,
                                                     int interact_min, uint32_t lds_voxels) {
  __shared__ double acc[MULTI ? 1 : MDX_MAX_TALLIES * kAccPerTally];
  __shared__ double lhist[MULTI ? 1 : kQueueLdsHist];
  __shared__ float scache[kWaves][kSigmaCache][64];  // per lane: cached total cross sections
  __shared__ unsigned long long cnt[MDX_NCOUNTER];
  // Per-block copies of what every step and interaction reads: the material
  // descriptors and the ID -> slot map, and the scene descriptors of the
  // job(s) being run (one for the block in a one-job launch, two per wave
  // otherwise).  Divergent reads of these then cost an LDS access instead of a
  // vector-memory round trip, and uniform ones are not re-read from memory
  // after every atomic.
  __shared__ mdx_material_dev lmat[MDX_MAX_MATERIALS];
  __shared__ mdx_curve lmaj;  // the global majorant
  __shared__ uint8_t lslot[256];
  __shared__ JobDesc desc[MULTI ? kWaves : 1];
  extern __shared__ float ldose_mem[];  // MDX_DOSE_F32C_LDS: (sum, error) per voxel of grid 0
  int hoff[MDX_MAX_TALLIES];
  float* ldose = nullptr;
  if (!MULTI && lds_voxels > 0) {
    ldose = ldose_mem;
    for (uint32_t i = threadIdx.x; i < 2 * lds_voxels; i += blockDim.x) ldose[i] = 0.0f;
  }
  block_copy(lmat, T->mat, T->n_mat * (int)sizeof(mdx_material_dev));
  block_copy(lslot, T->id_to_slot, 256);
  block_copy(&lmaj, &T->majorant, (int)sizeof(mdx_curve));
  if (!MULTI) {
    block_copy(&desc[0].D, jobs[0].D, (int)sizeof(mdx_domain_dev));
    block_copy(&desc[0].S, jobs[0].S, (int)sizeof(mdx_source_dev));
    block_copy(&desc[0].TL, jobs[0].TL, (int)sizeof(mdx_tallies_dev));
    hist_layout(jobs[0].TL, hoff, kQueueLdsHist);
    const int ntl = jobs[0].TL->n;
    for (int i = threadIdx.x; i < ntl * kAccPerTally; i += blockDim.x) acc[i] = 0.0;
    for (int i = threadIdx.x; i < kQueueLdsHist; i += blockDim.x) lhist[i] = 0.0;
  }
  if (threadIdx.x < MDX_NCOUNTER) cnt[threadIdx.x] = 0;
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wsl = __builtin_amdgcn_readfirstlane(wave);  // the same, known to be wave-uniform
  uint32_t c_hist = 0, c_step = 0, c_real = 0, c_ray = 0, c_comp = 0, c_photo = 0, c_trial = 0, c_stest = 0,
           c_dep = 0, c_surf = 0, c_esc = 0;
  // lane state: alive = carries a photon; pend = its last step ended in a
  // real interaction that waits for the wave's interaction phase
  bool alive = false, pend = false, exhausted = false;
  // several jobs (uniform per wave): the current job, the last one that ran
  // out, and whether every job has run out
  int jc = -1, jc_last = -1;
  bool done_all = false;
  if (MULTI) {
    const uint64_t nw = (uint64_t)gridDim.x * kWaves;
    const uint64_t gw = (uint64_t)blockIdx.x * kWaves + (uint64_t)wave;
    const uint64_t h = (uint64_t)(((double)gw + 0.5) / (double)nw * (double)total);
    int lo = 0, hi = n_jobs - 1;  // last job with prefix[j] <= h
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (prefix[mid] <= h) lo = mid; else hi = mid - 1;
    }
    jc = lo;
    load_job(&desc[wsl], jobs[jc]);
  }
  const mdx_domain_dev* D = &desc[0].D;
  const mdx_source_dev* S = &desc[0].S;
  const mdx_tallies_dev* TL = &desc[0].TL;
  Philox rng;     // source and interaction sampling
  StepRng srng;   // Woodcock steps
  float px = 0, py = 0, pz = 0, dx = 0, dy = 0, dz = 1, E = 1;
  int nc = 0, ni = 0;
  // The majorant of the next step: the global one (mcls 255) or, inside a
  // grid with local majorants, the total cross section of material mcls,
  // which bounds every cross section within dsafe of the photon.
  float smax = 1, step_scale = 1;  // step_scale = ln 2 / majorant
  int mcls = 255;
  float dsafe = INFINITY;
  uint32_t cvalid = 0;  // cache entries valid at the current energy
  int cslot = -1;
  float stot = 0, thr = 0;         // thr = 1 - sigma_tot / majorant
  // the pending step: start point, length, voxel
  float qx = 0, qy = 0, qz = 0, qs = 0;
  int qg = -1;
  uint64_t qlin = 0;
  uint32_t nstep = 0;
  float tfl = 0;  // flight distance left before a step can score in a tally
  // free flights: the job allows them / the photon sits in the background
  bool flights = false, fly = false;
  for (;;) {
    {
      // The descriptors are re-read from LDS in every iteration: an offset
      // the compiler cannot see through keeps it from holding them in
      // registers across the loop, which would spill.
      int o = 0;
      asm volatile("" : "+s"(o));
      const JobDesc* jd = MULTI ? &desc[wsl] + o : &desc[0] + o;
      D = &jd->D;
      S = &jd->S;
      TL = &jd->TL;
    }
    // ---- refill: claim histories for free lanes, in batches
    const bool free_lane = !alive && !exhausted;
    const unsigned long long fmask = __ballot(free_lane);
    const bool steppable = __ballot(alive && !pend) != 0ull;
    if (fmask && (__popcll(fmask) >= refill_min || !steppable) && !done_all) {
      const int leader = __ffsll((long long)fmask) - 1;
      const uint32_t rank = __popcll(fmask & ((1ull << lane) - 1ull));
      bool fresh = false;  // this lane claimed a history just now
      if (!MULTI) {
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(work, (unsigned long long)__popcll(fmask));
        base = __shfl(base, leader);
        if (free_lane) {
          if (base + rank < total) {
            const uint64_t h = jobs[0].first + base + rank;
            rng.init(jobs[0].seed, jobs[0].stream, h);
            srng.init(jobs[0].seed, jobs[0].stream, h);
            flights = jobs[0].flights != 0;
            fresh = true;
          } else {
            exhausted = true;
          }
        }
      } else {
        if (jc < 0) {  // the current job ran out: once its photons are done, the next one with histories left
          if (__ballot(alive) == 0ull) {  // else wait for them
            int jn = -1;
            for (int k = 0; k < n_jobs; ++k) {
              const int j = (jc_last + 1 + k) % n_jobs;
              const int left = __hip_atomic_load(work + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < jobs[j].n;
              if (__builtin_amdgcn_readfirstlane(left)) {  // the same counter in every lane: wave-uniform
                jn = j;
                break;
              }
            }
            if (jn < 0) {
              done_all = true;
            } else {
              load_job(&desc[wsl], jobs[jn]);
              jc = jn;
            }
          }
        }
        if (jc >= 0) {
          const uint64_t nj = jobs[jc].n;
          unsigned long long base = 0;
          if (lane == leader) base = atomicAdd(work + jc, (unsigned long long)__popcll(fmask));
          base = readlane_u64(base, leader);  // the leader's claim, wave-uniform
          if (base >= nj) {
            jc_last = jc;
            jc = -1;
          } else if (free_lane && base + rank < nj) {
            const uint64_t h = jobs[jc].first + base + rank;
            rng.init(jobs[jc].seed, jobs[jc].stream, h);
            srng.init(jobs[jc].seed, jobs[jc].stream, h);
            flights = jobs[jc].flights != 0;
            fresh = true;
          }
        }
      }
      if (fresh) {
        generate(S, rng, px, py, pz, dx, dy, dz, E);
        ++c_hist;
        nc = ni = 0;
        smax = curve_eval(T, lmaj, E);
        scache[wave][kSigmaCache - 1][lane] = smax;
        cvalid = 1u << (kSigmaCache - 1);
        step_scale = 0.69314718055994531f * frcp(smax);
        mcls = 255;
        dsafe = INFINITY;
        cslot = -1;
        nstep = 0;
        tfl = next_tally_event(TL, px, py, pz, dx, dy, dz);
        // a source photon in the background starts with a free flight
        uint64_t lin0;
        fly = flights && in_domain(D, px, py, pz) && voxel_lookup(D, px, py, pz, lin0) < 0;
        alive = true;
      }
    }
    if (__ballot(alive) == 0ull && (!MULTI || done_all)) break;

    // ---- one Woodcock step for every lane that can step (physics_engine.cpp:38-66)
    // Written so that the costly parts (a cross-section evaluation, a tally
    // replay, the next tally event) each have one site that all lanes
    // needing it run together, whatever branch brought them there.
    if (alive && !pend) {
      ++c_step;
      ++nstep;
      float u1, u2;
      srng.next(u1, u2);
      const float x0 = px, y0 = py, z0 = pz;
      bool escaped = false;
      int g = -1;
      uint64_t lin = 0;
      float s = 0.0f;
      int slot = D->bg_slot;  // free flights: the background's cross section
      const bool flew = fly;
      bool limited = false;  // the step stopped where the local majorant's bound ends
      uint32_t mword = 255u;  // the new position's macro cell (class | radius << 8)
      if (!fly) {
        s = -__log2f(u1) * step_scale;
        if (s > dsafe) {
          s = dsafe;
          limited = true;
        }
        px = px + s * dx;
        py = py + s * dy;
        pz = pz + s * dz;
        escaped = !in_domain(D, px, py, pz);
        if (!escaped) {
          uint32_t cell = 0;
          g = voxel_lookup(D, px, py, pz, lin, &cell);
          const uint16_t* mc = D->grids[0].macro;
          if (g == 0 && mc) mword = gld(mc + cell);
          const uint32_t c8 = mword & 255u;
          if (g < 0)
            slot = D->bg_slot;
          else if (c8 != 255u && (c8 & MDX_MACRO_UNIFORM))  // a one-material macro cell: its class
            slot = (int)(c8 & 0x7Fu);
          else
            slot = (int)lslot[gld(D->grids[g].ids + lin)];
        }
      }
      if (!escaped && slot != cslot) {
        if (slot < kSigmaCache - 1 && (cvalid >> slot & 1u)) {
          stot = scache[wave][slot][lane];
        } else {
          stot = curve_eval(T, lmat[slot].total, E);
          if (slot < kSigmaCache - 1) {
            scache[wave][slot][lane] = stot;
            cvalid |= 1u << slot;
          }
        }
        thr = delta_threshold(stot, smax);
        cslot = slot;
      }
      // replay this step's segment for the tallies: an escaping step only
      // when it reaches the next tally event (a shorter one scores nothing)
      bool replay = escaped && tfl - s <= 0.0f;
      float seg = s;
      if (escaped) {
        ++c_esc;
        alive = false;
      } else if (fly) {
        // background free flight: distance at the background's own cross
        // section, cut at the next grid face or the domain's boundary
        const float si = -__log2f(u1) * (0.69314718055994531f * frcp(stot));
        float dgrid, dexit;
        flight_bounds(D, x0, y0, z0, dx, dy, dz, dgrid, dexit);
        if (si < dgrid && si < dexit) {  // real interaction in the background
          px = x0 + si * dx;
          py = y0 + si * dy;
          pz = z0 + si * dz;
          pend = true;
          qx = x0; qy = y0; qz = z0; qs = si;
          qg = -1;
          qlin = 0;
        } else if (dgrid <= dexit) {  // up to the grid's face; majorant steps from there
          px = x0 + dgrid * dx;
          py = y0 + dgrid * dy;
          pz = z0 + dgrid * dz;
          fly = false;
          tfl -= dgrid;
          replay = tfl <= 0.0f;
          seg = dgrid;
          if (nstep >= kMaxSteps) alive = false;
        } else {  // out of the domain, overshooting by a majorant step
          seg = dexit - __log2f(u2) * step_scale;
          px = x0 + seg * dx;
          py = y0 + seg * dy;
          pz = z0 + seg * dz;
          ++c_esc;
          replay = tfl - seg <= 0.0f;  // an escaping step short of the next tally event scores nothing
          alive = false;
        }
      } else if (!limited && real_collision(u2, thr)) {  // real interaction: park it for the interaction phase
        pend = true;
        qx = x0; qy = y0; qz = z0; qs = s;
        qg = g;
        qlin = lin;
      } else {  // a rejected step, or one cut at the local majorant's bound (memoryless restart)
        tfl -= s;
        replay = tfl <= 0.0f;  // the step may reach a tally
        if (nstep >= kMaxSteps) alive = false;
        fly = flights && g < 0;
      }
      if (!escaped && !flew) {  // the majorant for the step from the new position
        int ncls = (mword & 255u) == 255u ? 255 : (int)(mword & 0x7Fu);
        const uint32_t nr = mword >> 8;
        if (nr == 0u || E < 1.0e3f) ncls = 255;
        if (ncls != mcls) {
          mcls = ncls;
          const int ce = ncls == 255 ? kSigmaCache - 1 : ncls;  // cache entry
          if (ce < kSigmaCache && (cvalid >> ce & 1u)) {
            smax = scache[wave][ce][lane];
          } else {
            smax = curve_eval(T, ncls == 255 ? lmaj : lmat[ncls].total, E);
            if (ce < kSigmaCache) {
              scache[wave][ce][lane] = smax;
              cvalid |= 1u << ce;
            }
          }
          step_scale = 0.69314718055994531f * frcp(smax);
          thr = delta_threshold(stot, smax);
        }
        dsafe = ncls == 255 ? INFINITY : (float)nr * D->grids[0].mcell;
      }
      if (replay) {
        Step r{x0, y0, z0, dx, dy, dz, seg, E, nc, ni, 0, 0, 0.0f, false};
        tallies_step<MULTI>(TL, acc, r, c_surf, lhist, hoff);
        if (alive) tfl = next_tally_event(TL, px, py, pz, dx, dy, dz);
      }
    }

    // ---- interaction phase: the parked lanes together (physics_engine.cpp:62-78)
    const unsigned long long pmask = __ballot(pend);
    // once interact_min lanes are parked, or at least as many as can still
    // step (their photons would otherwise wait on a thinning step phase)
    if (pmask && (__popcll(pmask) >= interact_min || __popcll(pmask) >= __popcll(__ballot(alive && !pend)))) {
      if (pend) {
        pend = false;
        ++c_real;
        const mdx_material_dev& M = T->mat[cslot];
        Step r{qx, qy, qz, dx, dy, dz, qs, E, nc, ni, 0, 0, 0.0f, true};
        float sc, si;
        curve_eval2(T, M.coh, E, M.incoh, E, sc, si);
        const int type = select_type(sc, si, stot, rng.uniform());
        float dep = 0.0f;
        if (type == 1) {
          ++c_ray;
          ++nc;
          const float mu = rita_sample(T, M, E, rng.uniform());
          rotate(dx, dy, dz, mu, 2.0f * kPiF * rng.uniform());
        } else if (type == 2) {
          ++c_comp;
          ++ni;
          float mu;
          const float Ep = compton(T, M, rng, E, mu, c_trial, c_stest);
          rotate(dx, dy, dz, mu, 2.0f * kPiF * rng.uniform());
          dep = E - Ep;
          E = Ep;
          if (E < 1.0e3f) {
            mcls = 255;
            dsafe = INFINITY;
          }
          smax = curve_eval(T, mcls == 255 ? lmaj : lmat[mcls].total, E);
          {
            const int ce = mcls == 255 ? kSigmaCache - 1 : mcls;
            cvalid = 0;  // a new energy
            if (ce < kSigmaCache) {
              scache[wave][ce][lane] = smax;
              cvalid = 1u << ce;
            }
          }
          step_scale = 0.69314718055994531f * frcp(smax);
          cslot = -1;
        } else {
          ++c_photo;
          dep = E;
          alive = false;
        }
        r.ncf = nc;
        r.nif = ni;
        r.dep = dep;
        if (dep != 0.0f) {
          ++c_dep;
          if (qg >= 0) {
            voxel_deposit(D->grids[qg], qlin, dep, qg == 0 ? ldose : nullptr);
          } else if (D->bg_dose) {
            gadd(D->bg_dose, (double)dep);
            gadd(D->bg_dose + 1, (double)dep * (double)dep);
            gadd(D->bg_dose + 2, 1.0);
          }
        }
        tallies_step<MULTI>(TL, acc, r, c_surf, lhist, hoff);
        if (nstep >= kMaxSteps) alive = false;
        if (alive) tfl = next_tally_event(TL, px, py, pz, dx, dy, dz);
      }
    }
  }
  atomicAdd(&cnt[MDX_C_HIST], (unsigned long long)c_hist);
  atomicAdd(&cnt[MDX_C_STEP], (unsigned long long)c_step);
  atomicAdd(&cnt[MDX_C_REAL], (unsigned long long)c_real);
  atomicAdd(&cnt[MDX_C_RAYLEIGH], (unsigned long long)c_ray);
  atomicAdd(&cnt[MDX_C_COMPTON], (unsigned long long)c_comp);
  atomicAdd(&cnt[MDX_C_PHOTO], (unsigned long long)c_photo);
  atomicAdd(&cnt[MDX_C_KAHN_TRIAL], (unsigned long long)c_trial);
  atomicAdd(&cnt[MDX_C_KAHN_S], (unsigned long long)c_stest);
  atomicAdd(&cnt[MDX_C_DEPOSIT], (unsigned long long)c_dep);
  atomicAdd(&cnt[MDX_C_SURFHIT], (unsigned long long)c_surf);
  atomicAdd(&cnt[MDX_C_ESCAPE], (unsigned long long)c_esc);
  __syncthreads();
  if (!MULTI) {
    const int ntl = jobs[0].TL->n;
    for (int i = threadIdx.x; i < ntl * kAccPerTally; i += blockDim.x) {
      double v = acc[i];
      if (v != 0.0) {
        int t = i / kAccPerTally;
        gadd(jobs[0].TL->t[t].acc + (blockIdx.x % MDX_ACC_COPIES) * kAccPerTally + (i - t * kAccPerTally), v);
      }
    }
    flush_hist(jobs[0].TL, hoff, lhist);
    if (ldose) {  // the block's sums of grid 0 into the global compensated sums
      float* g = jobs[0].D->grids[0].dose32;
      for (uint32_t i = threadIdx.x; i < lds_voxels; i += blockDim.x) {
        const float hi = ldose[2 * i], lo = ldose[2 * i + 1];
        if (hi != 0.0f) compensated_add<true>(g + 4 * i, hi);
        if (lo != 0.0f) gadd(g + 4 * i + 1, lo);
      }
    }
  }
  if (threadIdx.x < MDX_NCOUNTER && cnt[threadIdx.x]) atomicAdd(counters + threadIdx.x, cnt[threadIdx.x]);
}

template <bool MULTI>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(MDX_QUEUE_WAVES_PER_EU)))
k_transport_queue(const mdx_tables_dev* __restrict__ T, const mdx_job_dev* __restrict__ jobs,
                  const uint64_t* __restrict__ prefix, int n_jobs, uint64_t total,
                  unsigned long long* __restrict__ work, unsigned long long* __restrict__ counters,
                  int refill_min, int interact_min, uint32_t lds_voxels) {
  transport_queue_body<MULTI>(T, jobs, prefix, n_jobs, total, work, counters, refill_min, interact_min,
                              lds_voxels);
}

// Graph form: job count and total come from the round buffer's header
// (uint64 n_jobs, uint64 total, uint64 prefix[1025], mdx_job_dev jobs[1024]).
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(MDX_QUEUE_WAVES_PER_EU)))
k_transport_queue_round(const mdx_tables_dev* __restrict__ T, const char* __restrict__ round,
                        unsigned long long* __restrict__ work, unsigned long long* __restrict__ counters,
                        int refill_min, int interact_min) {
  const uint64_t* hdr = reinterpret_cast<const uint64_t*>(round);
  const int n_jobs = (int)hdr[0];
  const uint64_t total = hdr[1];
  const uint64_t* prefix = hdr + 2;
  const mdx_job_dev* jobs = reinterpret_cast<const mdx_job_dev*>(round + 2 * sizeof(uint64_t) + 1025 * sizeof(uint64_t));
  if (n_jobs <= 0 || total == 0) return;
  transport_queue_body<true>(T, jobs, prefix, n_jobs, total, work, counters, refill_min, interact_min, 0u);
}

// ------------------------------------------------------------------ bank kernel
// The event form with photon banks.  Every wave keeps a bank of B photons in
// LDS (structure of arrays, 60 B each) and a list of bank slots per next
// event: a Woodcock step inside a grid, a free flight through the
// background, a parked real interaction, or a free slot.  Each iteration the
// wave takes up to 64 photons from one list (the fullest step or flight
// list, the interaction list once it holds interact_min photons, new source
// photons once refill_min slots are free), loads their state into its lanes,
// runs that one event for all of them, stores the state back and appends
// each slot to the list of its photon's next event: ballot per event type,
// the lane's rank among the lanes of its type (popcount of the lower lanes)
// and one count per type, all wave-uniform.  So every event runs on lanes
// that all do the same kind of work, whatever mix of events the bank holds,
// instead of the lanes a wave's resident photons happen to leave for it.
// The physics of each event is that of transport_queue_body; the
// differences are bookkeeping: the interaction stream's buffered Philox
// words are dropped at the end of every event (its block counter lives in
// the bank; no counter is reused), the start of an interaction's step is
// recomputed from its end and length for the tally replay, and a photon
// that interacts in the background resumes with a free flight.
// A wave holds the photons of one job at a time (MULTI: the next job with
// histories left once its own has run out and its bank is empty).
enum { kEvStep = 0, kEvFly = 1, kEvInt = 2, kEvFree = 3, kEvTypes = 4 };

// A photon's slot is 60 bytes, so that a wave's bank holds 160 of them at
// four blocks per CU (banks of 96 or 128 leave LDS unused); nothing that a
// load from the tables or the grid can give again is kept: the safe
// distance of the local majorant is its radius (8 bits of the flags) times
// the job's macro-cell edge, and no cross section is cached beside the two
// below.  The distance to the next tally event is dead while a photon waits
// for its interaction (the interaction recomputes it), so that word holds
// the length of the step that ended there meanwhile.  The high word of the
// history index shares a word with the step counter, which limits the bank
// kernel to history indices below 2^42.
constexpr uint32_t kStepBits = 22;  // step counter bits: kMaxSteps = 2^20 fits
constexpr uint32_t kStepMask = (1u << kStepBits) - 1u;
static_assert(kMaxSteps <= kStepMask, "the step counter fits its bits");
static_assert(MDX_BANK_MAX_HISTORY == 1ull << (32 + (32 - kStepBits)), "history index bits of a bank slot");

template <int B>
struct WaveBank {
  mdx_v4 pe[B];   // position, energy
  mdx_v4 dm[B];   // direction, majorant of the next step
  // sigma_tot of the cached slot, distance to the next tally event (while
  // parked for an interaction: length of the step that ended there),
  // history index (low word) and Philox block counter (as their bits)
  mdx_v4 aux[B];
  uint32_t hn[B];     // step counter | high word of the history index << kStepBits
  uint32_t flags[B];  // see fl_get / fl_set
  uint8_t list[kEvTypes][B];
};

// flags: bits 0-1 coherent scatters (saturating at 3), 2-3 incoherent, 4-11
// majorant class (255: global), 12-19 slot of the cached sigma_tot (255:
// none), 20-23 grid of the pending interaction + 1 (0: background), 24-31
// safe radius of the class in macro cells (the safe distance when the class
// is not the global majorant's).
__device__ __forceinline__ uint32_t fl_get(uint32_t f, int sh, uint32_t m) { return (f >> sh) & m; }
__device__ __forceinline__ uint32_t fl_set(uint32_t f, int sh, uint32_t m, uint32_t v) {
  return (f & ~(m << sh)) | ((v & m) << sh);
}

// The step stream's key of a history (StepRng::init).
__device__ __forceinline__ uint32_t step_key(uint64_t seed, uint32_t job, uint32_t hhi) {
  return mix32(mix32((uint32_t)seed, (uint32_t)(seed >> 32)), mix32(job, hhi));
}

__device__ __forceinline__ void step_uniforms(uint32_t key, uint32_t hlo, uint32_t step, float& u1, float& u2) {
  uint32_t x0, x1;
  philox2x32_10(hlo, step, key, x0, x1);
  u1 = (float)((x0 >> 8) + 1u) * (1.0f / 16777216.0f);
  u2 = (float)((x1 >> 8) + 1u) * (1.0f / 16777216.0f);
}

// One photon in a lane while an event runs on it.
struct Ph {
  float px, py, pz, E, dx, dy, dz, smax, dsafe, stot, tfl, qs;
  uint32_t hlo, hhi, c0, nstep, flags;
};

// mcell: the job's macro-cell edge (the safe distance's unit).
template <int B>
__device__ __forceinline__ Ph bank_load(const WaveBank<B>* W, int s, float mcell) {
  Ph p;
  const mdx_v4 a = W->pe[s], b = W->dm[s], c = W->aux[s];
  p.px = a.x; p.py = a.y; p.pz = a.z; p.E = a.w;
  p.dx = b.x; p.dy = b.y; p.dz = b.z; p.smax = b.w;
  p.stot = c.x; p.tfl = c.y; p.qs = c.y; p.hlo = __float_as_uint(c.z); p.c0 = __float_as_uint(c.w);
  const uint32_t hn = W->hn[s];
  p.nstep = hn & kStepMask;
  p.hhi = hn >> kStepBits;
  p.flags = W->flags[s];
  p.dsafe = fl_get(p.flags, 4, 255u) == 255u ? INFINITY : (float)fl_get(p.flags, 24, 255u) * mcell;
  return p;
}

template <int B>
__device__ __forceinline__ void bank_store(WaveBank<B>* W, int s, const Ph& p) {
  W->pe[s] = mdx_v4{p.px, p.py, p.pz, p.E};
  W->dm[s] = mdx_v4{p.dx, p.dy, p.dz, p.smax};
  W->aux[s] = mdx_v4{p.stot, p.tfl, __uint_as_float(p.hlo), __uint_as_float(p.c0)};
  W->hn[s] = p.nstep | (p.hhi << kStepBits);
  W->flags[s] = p.flags;
}

// A wave's own LDS writes made visible to its other lanes before they read.
__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Total cross section of material slot (or the global majorant, slot 255)
// at energy E.
__device__ __forceinline__ float bank_sigma(const mdx_tables_dev* T, const mdx_curve* ltot, const mdx_curve& maj,
                                            float E, int slot) {
  return curve_eval(T, slot == 255 ? maj : ltot[slot], E);
}

template <bool MULTI, int B>
__device__ __forceinline__ void transport_bank_body(const mdx_tables_dev* __restrict__ T,
                                                    const mdx_job_dev* __restrict__ jobs,
                                                    const uint64_t* __restrict__ prefix, int n_jobs, uint64_t total,
                                                    unsigned long long* __restrict__ work,
                                                    unsigned long long* __restrict__ counters, int refill_min,
                                                    int interact_min) {
  // scalar tallies and spectra go to global atomics with the lanes of a wave
  // that hit one record summed first (no LDS pre-reduction: the LDS holds
  // photon banks)
  __shared__ unsigned long long cnt[MDX_NCOUNTER];
  __shared__ mdx_curve ltot[MDX_MAX_MATERIALS];  // the materials' total cross sections
  __shared__ uint8_t lslot[256];
  __shared__ WaveBank<B> banks[kWaves];
  for (int m = threadIdx.x; m < T->n_mat; m += blockDim.x) ltot[m] = T->mat[m].total;
  block_copy(lslot, T->id_to_slot, 256);
  if (threadIdx.x < MDX_NCOUNTER) cnt[threadIdx.x] = 0;
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int wsl = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  WaveBank<B>* W = &banks[wsl];
  uint32_t c_hist = 0, c_step = 0, c_real = 0, c_ray = 0, c_comp = 0, c_photo = 0, c_trial = 0, c_stest = 0,
           c_dep = 0, c_surf = 0, c_esc = 0;
  // every slot starts free; list counts are wave-uniform
  for (int s = lane; s < B; s += 64) W->list[kEvFree][s] = (uint8_t)s;
  int n_step = 0, n_fly = 0, n_int = 0, n_free = B;
  bool exhausted = false;  // the current job has no histories left to claim
  int jc = 0;
  bool done_all = false;
  // the current job's record, wave-uniform
  uint64_t jseed = 0, jfirst = 0, jn = total;
  uint32_t jstream = 0;
  bool jflights = false;
  // the current job's descriptors (wave-uniform pointers, read through the
  // scalar cache)
  const mdx_domain_dev* jdom = nullptr;
  const mdx_source_dev* jsrc = nullptr;
  const mdx_tallies_dev* jtl = nullptr;
  auto take_job = [&](int j) {
    jdom = jobs[j].D;
    jsrc = jobs[j].S;
    jtl = jobs[j].TL;
    jseed = jobs[j].seed;
    jfirst = jobs[j].first;
    jn = MULTI ? jobs[j].n : total;
    jstream = jobs[j].stream;
    jflights = jobs[j].flights != 0;
  };
  if (MULTI) {
    const uint64_t nw = (uint64_t)gridDim.x * kWaves;
    const uint64_t gw = (uint64_t)blockIdx.x * kWaves + (uint64_t)wsl;
    const uint64_t h = (uint64_t)(((double)gw + 0.5) / (double)nw * (double)total);
    int lo = 0, hi = n_jobs - 1;  // last job with prefix[j] <= h
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (prefix[mid] <= h) lo = mid; else hi = mid - 1;
    }
    jc = lo;
  }
  take_job(jc);
  wave_sync_lds();
  refill_min = refill_min > B ? B : refill_min;
  for (;;) {
    const mdx_domain_dev* D = wave_const(jdom);
    const mdx_source_dev* S = wave_const(jsrc);
    const mdx_tallies_dev* TL = wave_const(jtl);
    const int alive = n_step + n_fly + n_int;
    if (MULTI && exhausted && alive == 0) {  // the wave's job is done here: the next one with histories left
      int jnext = -1;
      for (int k = 0; k < n_jobs; ++k) {
        const int j = (jc + 1 + k) % n_jobs;
        // every lane reads the same counter; its first lane's answer is the
        // wave's (kept wave-uniform, so that the loop stays scalar)
        const int left = __hip_atomic_load(work + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < jobs[j].n;
        if (__builtin_amdgcn_readfirstlane(left)) {
          jnext = j;
          break;
        }
      }
      if (jnext < 0) {
        done_all = true;
      } else {
        jc = jnext;
        take_job(jc);
        exhausted = false;
      }
      if (done_all) break;
      continue;
    }
    if (!MULTI && exhausted && alive == 0) break;
    // the event to run: new photons once refill_min slots are free, else the
    // fullest of the step and flight lists, the interactions once
    // interact_min are parked or they are the most numerous
    int ev;
    if (!exhausted && n_free >= refill_min) ev = kEvFree;
    else if (n_int > 0 && (n_int >= interact_min || (n_int >= n_step && n_int >= n_fly))) ev = kEvInt;
    else if (n_step >= n_fly && n_step > 0) ev = kEvStep;
    else if (n_fly > 0) ev = kEvFly;
    else if (n_int > 0) ev = kEvInt;
    else ev = kEvFree;  // nothing alive, slots free, histories left
    int avail = ev == kEvStep ? n_step : ev == kEvFly ? n_fly : ev == kEvInt ? n_int : n_free;
    const int k = avail < 64 ? avail : 64;
    avail -= k;
    if (ev == kEvStep) n_step = avail;
    else if (ev == kEvFly) n_fly = avail;
    else if (ev == kEvInt) n_int = avail;
    else n_free = avail;
    const bool act = lane < k;
    const int slot = act ? (int)W->list[ev][avail + lane] : 0;
    int next = kEvFree;
    if (ev == kEvFree) {
      // ---- new source photons for the k free slots
      unsigned long long base = 0;
      if (lane == 0) base = atomicAdd(MULTI ? work + jc : work, (unsigned long long)k);
      base = readlane_u64(base, 0);  // lane 0's claim, wave-uniform
      if (base + (uint64_t)k >= jn) exhausted = true;
      if (act && base + (uint64_t)lane < jn) {
        const uint64_t h = jfirst + base + (uint64_t)lane;
        Philox rng;
        rng.init(jseed, jstream, h);
        Ph p;
        generate(S, rng, p.px, p.py, p.pz, p.dx, p.dy, p.dz, p.E);
        ++c_hist;
        p.hlo = (uint32_t)h;
        p.hhi = (uint32_t)(h >> 32);
        p.c0 = rng.c0;
        p.nstep = 0;
        p.flags = fl_set(fl_set(0u, 4, 255u, 255u), 12, 255u, 255u);
        p.smax = curve_eval(T, wave_const(T)->majorant, p.E);
        p.dsafe = INFINITY;
        p.stot = 0.0f;
        p.tfl = next_tally_event(TL, p.px, p.py, p.pz, p.dx, p.dy, p.dz);
        uint64_t lin0;
        const bool fly = jflights && in_domain(D, p.px, p.py, p.pz) && voxel_lookup(D, p.px, p.py, p.pz, lin0) < 0;
        next = fly ? kEvFly : kEvStep;
        bank_store(W, slot, p);
      }
    } else if (ev == kEvStep || ev == kEvFly) {
      // ---- one Woodcock step or one background free flight (physics_engine.cpp:38-66)
      if (act) {
        Ph p = bank_load(W, slot, D->grids[0].mcell);
        ++c_step;
        float u1, u2;
        step_uniforms(step_key(jseed, jstream, p.hhi), p.hlo, p.nstep, u1, u2);
        ++p.nstep;
        const float x0 = p.px, y0 = p.py, z0 = p.pz;
        const int nc = (int)fl_get(p.flags, 0, 3u), ni = (int)fl_get(p.flags, 2, 3u);
        int cslot = (int)fl_get(p.flags, 12, 255u);
        bool escaped = false, replay = false, alive_ = true;
        float seg = 0.0f;
        int g = -1;
        uint64_t lin = 0;
        uint32_t mword = 255u;
        if (ev == kEvStep) {
          float s = -__log2f(u1) * (0.69314718055994531f * frcp(p.smax));
          bool limited = false;
          if (s > p.dsafe) {
            s = p.dsafe;
            limited = true;
          }
          p.px = x0 + s * p.dx;
          p.py = y0 + s * p.dy;
          p.pz = z0 + s * p.dz;
          seg = s;
          escaped = !in_domain(D, p.px, p.py, p.pz);
          if (escaped) {
            ++c_esc;
            replay = p.tfl - s <= 0.0f;  // an escaping step short of the next tally event scores nothing
            alive_ = false;
          } else {
            uint32_t cell = 0;
            g = voxel_lookup(D, p.px, p.py, p.pz, lin, &cell);
            const uint16_t* mc = D->grids[0].macro;
            if (g == 0 && mc) mword = gld(mc + cell);
            const uint32_t c8 = mword & 255u;
            int sl;
            if (g < 0) sl = D->bg_slot;
            else if (c8 != 255u && (c8 & MDX_MACRO_UNIFORM)) sl = (int)(c8 & 0x7Fu);
            else sl = (int)lslot[gld(D->grids[g].ids + lin)];
            if (sl != cslot) {
              // a local majorant of class k is material k's own total cross
              // section at this energy: the same value, no second lookup
              p.stot = sl == (int)fl_get(p.flags, 4, 255u) ? p.smax : bank_sigma(T, ltot, wave_const(T)->majorant, p.E, sl);
              cslot = sl;
            }
            if (!limited && real_collision(u2, delta_threshold(p.stot, p.smax))) {
              next = kEvInt;
              p.tfl = s;  // parked: the word holds the step's length (loaded as qs)
              p.flags = fl_set(p.flags, 20, 15u, (uint32_t)(g + 1));
            } else {
              p.tfl -= s;
              replay = p.tfl <= 0.0f;
              if (p.nstep >= kMaxSteps) alive_ = false;
              next = (jflights && g < 0) ? kEvFly : kEvStep;
            }
            // the majorant for the step from the new position
            int ncls = (mword & 255u) == 255u ? 255 : (int)(mword & 0x7Fu);
            const uint32_t nr = mword >> 8;
            if (nr == 0u || p.E < 1.0e3f) ncls = 255;
            if (ncls != (int)fl_get(p.flags, 4, 255u)) {
              p.flags = fl_set(p.flags, 4, 255u, (uint32_t)ncls);
              p.smax = ncls == cslot ? p.stot : bank_sigma(T, ltot, wave_const(T)->majorant, p.E, ncls);
            }
            p.dsafe = ncls == 255 ? INFINITY : (float)nr * D->grids[0].mcell;
            p.flags = fl_set(p.flags, 24, 255u, nr);
          }
        } else {
          // background free flight at the background's own cross section,
          // cut at the next grid face or the domain's boundary
          const int sl = D->bg_slot;
          if (sl != cslot) {
            p.stot = bank_sigma(T, ltot, wave_const(T)->majorant, p.E, sl);
            cslot = sl;
          }
          const float si = -__log2f(u1) * (0.69314718055994531f * frcp(p.stot));
          float dgrid, dexit;
          flight_bounds(D, x0, y0, z0, p.dx, p.dy, p.dz, dgrid, dexit);
          if (si < dgrid && si < dexit) {  // a real interaction in the background
            p.px = x0 + si * p.dx;
            p.py = y0 + si * p.dy;
            p.pz = z0 + si * p.dz;
            next = kEvInt;
            p.tfl = si;
            p.flags = fl_set(p.flags, 20, 15u, 0u);
          } else if (dgrid <= dexit) {  // up to the grid's face; majorant steps from there
            p.px = x0 + dgrid * p.dx;
            p.py = y0 + dgrid * p.dy;
            p.pz = z0 + dgrid * p.dz;
            p.tfl -= dgrid;
            replay = p.tfl <= 0.0f;
            seg = dgrid;
            next = kEvStep;
            if (p.nstep >= kMaxSteps) alive_ = false;
          } else {  // out of the domain, overshooting by a majorant step
            seg = dexit - __log2f(u2) * (0.69314718055994531f * frcp(p.smax));
            p.px = x0 + seg * p.dx;
            p.py = y0 + seg * p.dy;
            p.pz = z0 + seg * p.dz;
            ++c_esc;
            replay = p.tfl - seg <= 0.0f;
            alive_ = false;
          }
        }
        p.flags = fl_set(p.flags, 12, 255u, (uint32_t)cslot);
        if (replay) {
          Step r{x0, y0, z0, p.dx, p.dy, p.dz, seg, p.E, nc, ni, 0, 0, 0.0f, false};
          tallies_step<true>(TL, nullptr, r, c_surf, nullptr, nullptr);
          if (alive_) p.tfl = next_tally_event(TL, p.px, p.py, p.pz, p.dx, p.dy, p.dz);
        }
        if (!alive_) next = kEvFree;
        else bank_store(W, slot, p);
      }
    } else {
      // ---- the parked real interactions (physics_engine.cpp:62-78)
      if (act) {
        Ph p = bank_load(W, slot, D->grids[0].mcell);
        ++c_real;
        const int cslot = (int)fl_get(p.flags, 12, 255u);
        const mdx_material_dev& M = T->mat[cslot];
        int nc = (int)fl_get(p.flags, 0, 3u), ni = (int)fl_get(p.flags, 2, 3u);
        Step r{p.px - p.qs * p.dx, p.py - p.qs * p.dy, p.pz - p.qs * p.dz, p.dx, p.dy, p.dz, p.qs, p.E,
               nc, ni, 0, 0, 0.0f, true};
        Philox rng;
        rng.init(jseed, jstream, (uint64_t)p.hlo | ((uint64_t)p.hhi << 32));
        rng.c0 = p.c0;
        // the Compton sampler's S(x_max) is looked up beside the partial
        // cross sections (it depends only on the energy and the material)
        float sc, si, smx;
        curve_eval3(T, M.coh, p.E, M.incoh, p.E, M.sf, compton_xmax(p.E), sc, si, smx);
        const int type = select_type(sc, si, p.stot, rng.uniform());
        float dep = 0.0f;
        bool alive_ = true;
        // both scatters draw the azimuth after their polar sampler, so one
        // rotation below serves both (one copy of it in the divergent code)
        float mu = 1.0f;
        if (type == 1) {
          ++c_ray;
          nc = nc < 3 ? nc + 1 : 3;
          mu = rita_sample(T, M, p.E, rng.uniform());
        } else if (type == 2) {
          ++c_comp;
          ni = ni < 3 ? ni + 1 : 3;
          const float Ep = compton(T, M, rng, p.E, mu, c_trial, c_stest, smx);
          dep = p.E - Ep;
          p.E = Ep;
          int mcls = (int)fl_get(p.flags, 4, 255u);
          if (p.E < 1.0e3f) {
            mcls = 255;
            p.dsafe = INFINITY;
          }
          p.flags = fl_set(p.flags, 4, 255u, (uint32_t)mcls);
          p.smax = bank_sigma(T, ltot, wave_const(T)->majorant, p.E, mcls);
          p.flags = fl_set(p.flags, 12, 255u, 255u);
        } else {
          ++c_photo;
          dep = p.E;
          alive_ = false;
        }
        if (alive_) rotate(p.dx, p.dy, p.dz, mu, 2.0f * kPiF * rng.uniform());
        p.c0 = rng.c0;
        p.flags = fl_set(fl_set(p.flags, 0, 3u, (uint32_t)nc), 2, 3u, (uint32_t)ni);
        r.ncf = nc;
        r.nif = ni;
        r.dep = dep;
        const int qg = (int)fl_get(p.flags, 20, 15u) - 1;
        if (dep != 0.0f) {
          ++c_dep;
          if (qg >= 0) {
            // the voxel of the interaction point (the step that parked it
            // found the same one: the position is unchanged since)
            const mdx_grid_dev& G = D->grids[qg];
            voxel_deposit(G, grid_voxel(G, p.px - G.origin[0], p.py - G.origin[1], p.pz - G.origin[2]), dep);
          } else if (D->bg_dose) {
            gadd(D->bg_dose, (double)dep);
            gadd(D->bg_dose + 1, (double)dep * (double)dep);
            gadd(D->bg_dose + 2, 1.0);
          }
        }
        tallies_step<true>(TL, nullptr, r, c_surf, nullptr, nullptr);
        if (p.nstep >= kMaxSteps) alive_ = false;
        if (alive_) {
          p.tfl = next_tally_event(TL, p.px, p.py, p.pz, p.dx, p.dy, p.dz);
          next = (jflights && qg < 0) ? kEvFly : kEvStep;
          bank_store(W, slot, p);
        }
      }
    }
    // ---- each processed slot to the list of its photon's next event
    wave_sync_lds();
    {
      const unsigned long long below = (1ull << lane) - 1ull;
      const unsigned long long m0 = __ballot(act && next == kEvStep), m1 = __ballot(act && next == kEvFly),
                               m2 = __ballot(act && next == kEvInt), m3 = __ballot(act && next == kEvFree);
      const unsigned long long mine = next == kEvStep ? m0 : next == kEvFly ? m1 : next == kEvInt ? m2 : m3;
      const int basep = next == kEvStep ? n_step : next == kEvFly ? n_fly : next == kEvInt ? n_int : n_free;
      if (act) W->list[next][basep + __popcll(mine & below)] = (uint8_t)slot;
      n_step += __popcll(m0);
      n_fly += __popcll(m1);
      n_int += __popcll(m2);
      n_free += __popcll(m3);
    }
    wave_sync_lds();
  }
  atomicAdd(&cnt[MDX_C_HIST], (unsigned long long)c_hist);
  atomicAdd(&cnt[MDX_C_STEP], (unsigned long long)c_step);
  atomicAdd(&cnt[MDX_C_REAL], (unsigned long long)c_real);
  atomicAdd(&cnt[MDX_C_RAYLEIGH], (unsigned long long)c_ray);
  atomicAdd(&cnt[MDX_C_COMPTON], (unsigned long long)c_comp);
  atomicAdd(&cnt[MDX_C_PHOTO], (unsigned long long)c_photo);
  atomicAdd(&cnt[MDX_C_KAHN_TRIAL], (unsigned long long)c_trial);
  atomicAdd(&cnt[MDX_C_KAHN_S], (unsigned long long)c_stest);
  atomicAdd(&cnt[MDX_C_DEPOSIT], (unsigned long long)c_dep);
  atomicAdd(&cnt[MDX_C_SURFHIT], (unsigned long long)c_surf);
  atomicAdd(&cnt[MDX_C_ESCAPE], (unsigned long long)c_esc);
  __syncthreads();
  if (threadIdx.x < MDX_NCOUNTER && cnt[threadIdx.x]) atomicAdd(counters + threadIdx.x, cnt[threadIdx.x]);
}

// Blocks of four waves with banks of 96, 128 or 160 photons each fit four to a
// CU's LDS; the registers are held to four waves per SIMD to match.
template <bool MULTI, int B>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4)))
k_transport_bank(const mdx_tables_dev* __restrict__ T, const mdx_job_dev* __restrict__ jobs,
                 const uint64_t* __restrict__ prefix, int n_jobs, uint64_t total, unsigned long long* __restrict__ work,
                 unsigned long long* __restrict__ counters, int refill_min, int interact_min) {
  transport_bank_body<MULTI, B>(T, jobs, prefix, n_jobs, total, work, counters, refill_min, interact_min);
}

// Graph form (round buffer as for k_transport_queue_round).
template <int B>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4)))
k_transport_bank_round(const mdx_tables_dev* __restrict__ T, const char* __restrict__ round,
                       unsigned long long* __restrict__ work, unsigned long long* __restrict__ counters,
                       int refill_min, int interact_min) {
  const uint64_t* hdr = reinterpret_cast<const uint64_t*>(round);
  const int n_jobs = (int)hdr[0];
  const uint64_t total = hdr[1];
  const uint64_t* prefix = hdr + 2;
  const mdx_job_dev* jobs = reinterpret_cast<const mdx_job_dev*>(round + 2 * sizeof(uint64_t) + 1025 * sizeof(uint64_t));
  if (n_jobs <= 0 || total == 0) return;
  transport_bank_body<true, B>(T, jobs, prefix, n_jobs, total, work, counters, refill_min, interact_min);
}

// ------------------------------------------------------------------ RITA table build (FP64)
// The adaptive PENELOPE RITA tables of the Rayleigh angular distribution
// (probability_dist.cpp:62-71,121-175,267-287), one wave per energy of the
// incoherent-CS grid: the pdf (1 + mu^2) FF(x)^2 with FF log-log linear
// (interpolators.cpp:71-76 over log10), normalised by 1001-point Simpson;
// from 10 equal intervals, every interval whose error (51-point rule)
// exceeds 1e-4 is halved until none does.  Lanes take intervals; the
// running sums (normalisation, CDF) are added by one lane in the order the
// CPU path adds them.
constexpr int kRitaCap = 256;  // grid points per energy

struct RitaFF {
  const double* lx;  // log10 x of the FF table
  const double* ly;  // log10 FF
  int n;
  double alpha;      // x = alpha * (E / mc^2) * sqrt(1 - mu), 1/Angstrom
  double inv_mc2;
};

__device__ double ff_loglog(const RitaFF& F, double x) {
#pragma clang fp contract(off)
  const double l = log10(x);
  if (l <= F.lx[0]) return pow(10.0, F.ly[0]);
  if (l >= F.lx[F.n - 1]) return pow(10.0, F.ly[F.n - 1]);
  int lo = 0, hi = F.n;  // upper_bound
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (F.lx[mid] <= l) lo = mid + 1; else hi = mid;
  }
  const double x1 = F.lx[lo - 1], x2 = F.lx[lo], y1 = F.ly[lo - 1], y2 = F.ly[lo];
  const double slope = (y2 - y1) / (x2 - x1);
  return pow(10.0, y1 + slope * (l - x1));
}

__device__ __forceinline__ double rayleigh_pdf(const RitaFF& F, double E, double mu) {
#pragma clang fp contract(off)
  const double x = F.alpha * (E * F.inv_mc2) * sqrt(1.0 - mu);
  const double ff = ff_loglog(F, x);
  return (1.0 + mu * mu) * ff * ff;
}

__device__ __forceinline__ double rita_pnum(double x, double x0, double x1, double y0, double y1, double a,
                                           double b) {
#pragma clang fp contract(off)
  const double tau = (x - x0) / (x1 - x0);
  double eta = 0.0;
  if (tau != 0.0) {
    const double c = 1.0 + a + b - a * tau;
    eta = c / (2.0 * b * tau) * (1.0 - sqrt(1.0 - 4.0 * b * tau * tau / (c * c)));
  }
  const double q = 1.0 + a * eta + b * eta * eta;
  return q * q / ((1.0 + a + b) * (1.0 - b * eta * eta)) * (y1 - y0) / (x1 - x0);
}

__global__ void __launch_bounds__(64) k_rita_build(RitaFF F, const double* __restrict__ energies,
                                                    double* __restrict__ xs, double* __restrict__ ys,
                                                    double* __restrict__ as, double* __restrict__ bs,
                                                    int* __restrict__ count) {
#pragma clang fp contract(off)  // the CPU path's operation order, no fused multiply-adds
  __shared__ double g[kRitaCap], c[kRitaCap], inc[kRitaCap], ng[kRitaCap];
  __shared__ double terms[1001];
  __shared__ int split[kRitaCap];
  __shared__ int n_sh, done_sh;
  __shared__ double norm_sh;
  const int lane = threadIdx.x;
  const int ie = blockIdx.x;
  const double E = energies[ie];
  // probability_dist.cpp:267-287: Simpson normalisation with 1001 points
  const double h0 = 2.0 / 1000;
  for (int k = lane; k <= 1000; k += 64) {
    const double f = rayleigh_pdf(F, E, -1.0 + k * h0);
    terms[k] = (k == 0 || k == 1000) ? f : ((k % 2 == 0) ? 2 * f : 4 * f);
  }
  __syncthreads();
  if (lane == 0) {
    double s = 0;
    for (int k = 0; k <= 1000; ++k) s += terms[k];
    norm_sh = h0 / 3 * s;
    for (int i = 0; i < 10; ++i) g[i] = -1.0 + i * (2.0 / 9.0);
    g[9] = 1.0;
    n_sh = 10;
  }
  __syncthreads();
  for (int iter = 0; iter < 64; ++iter) {
    const int n = n_sh, m = n - 1;
    // CDF increments, 51-point Simpson per interval
    for (int j = lane; j < m; j += 64) {
      const double h = (g[j + 1] - g[j]) / 50.0;
      double s = 0;
      for (int k = 0; k < 51; ++k) {
        const double f = rayleigh_pdf(F, E, g[j] + k * h) / norm_sh;
        s += (k == 0 || k == 50) ? f : ((k % 2 == 0) ? 2 * f : 4 * f);
      }
      inc[j] = h / 3 * s;
    }
    __syncthreads();
    if (lane == 0) {
      c[0] = 0.0;
      for (int j = 0; j < m; ++j) c[j + 1] = c[j] + inc[j];
    }
    __syncthreads();
    // interpolation parameters and the interval error (probability_dist.cpp:138-175)
    for (int j = lane; j < m; j += 64) {
      const double p0 = rayleigh_pdf(F, E, g[j]) / norm_sh, p1 = rayleigh_pdf(F, E, g[j + 1]) / norm_sh;
      const double sl = (c[j + 1] - c[j]) / (g[j + 1] - g[j]);
      const double b = 1 - sl * sl * (1 / (p0 * p1));
      const double a = sl * (1 / p0) - b - 1;
      const double x0 = g[j], x1 = g[j + 1], y0 = c[j], y1 = c[j + 1];
      const int N = 51;
      const double d = (x1 - x0) / N;
      double s = 0;
      for (int i = 1; i < N; ++i) {
        const double xk = x0 + i * d;
        s += fabs(rayleigh_pdf(F, E, xk) / norm_sh - rita_pnum(xk, x0, x1, y0, y1, a, b));
      }
      const double err = d * (s + (fabs(p1 - rita_pnum(x1, x0, x1, y0, y1, a, b)) +
                                   fabs(p0 - rita_pnum(x0, x0, x1, y0, y1, a, b))) / 2);
      split[j] = err > 1e-4;
      xs[(size_t)ie * kRitaCap + j] = g[j];
      as[(size_t)ie * kRitaCap + j] = a;
      bs[(size_t)ie * kRitaCap + j] = b;
    }
    __syncthreads();
    if (lane == 0) {
      int k = 0;
      bool ok = true;
      for (int j = 0; j < m && k < kRitaCap - 1; ++j) {
        ng[k++] = g[j];
        if (split[j] && k < kRitaCap - 1) {
          ng[k++] = (g[j] + g[j + 1]) / 2;
          ok = false;
        }
      }
      ng[k++] = g[m];
      done_sh = ok ? 1 : (k >= kRitaCap - 1 ? -1 : 0);
      if (ok) {
        for (int j = 0; j < n; ++j) {
          ys[(size_t)ie * kRitaCap + j] = c[j];
        }
        xs[(size_t)ie * kRitaCap + m] = g[m];
        count[ie] = n;
      } else {
        n_sh = k;
      }
    }
    __syncthreads();
    if (done_sh != 0) break;
    for (int j = lane; j < n_sh; j += 64) g[j] = ng[j];
    __syncthreads();
  }
  if (lane == 0 && done_sh != 1) count[ie] = -1;  // did not converge within the cap
}

// ------------------------------------------------------------------ inspection kernels
__global__ void k_eval_curve(const mdx_tables_dev* T, int slot, int which, int n, const float* x, float* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const mdx_curve* c;
  if (slot < 0) c = &T->majorant;
  else {
    const mdx_material_dev& M = T->mat[slot];
    c = which == 0 ? &M.total : which == 1 ? &M.coh : which == 2 ? &M.incoh : which == 3 ? &M.photo : &M.sf;
  }
  out[i] = curve_eval(T, *c, x[i]);
}

__global__ void k_rita(const mdx_tables_dev* T, int slot, int n, const float* E, const float* y, float* mu) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  mu[i] = rita_sample(T, T->mat[slot], E[i], y[i]);
}

__global__ void k_voxel(const mdx_tables_dev* T, const mdx_domain_dev* D, int n, const float* pos, int* grid,
                        long long* lin, int* mat) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t l = 0;
  int g = voxel_lookup(D, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], l);
  grid[i] = g;
  lin[i] = g < 0 ? -1 : (long long)l;
  int slot = g < 0 ? D->bg_slot : (int)T->id_to_slot[D->grids[g].ids[l]];
  mat[i] = T->mat[slot].material_id;
}

__global__ void k_select(int n, const float* sc, const float* si, const float* st, const float* u, int* type) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  type[i] = select_type(sc[i], si[i], st[i], u[i]);
}

// The direction change every scattering event of the transport kernels makes.
__global__ void k_rotate(int n, const float* dir, const float* mu, const float* phi, float* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float dx = dir[3 * i], dy = dir[3 * i + 1], dz = dir[3 * i + 2];
  rotate(dx, dy, dz, mu[i], phi[i]);
  out[3 * i] = dx;
  out[3 * i + 1] = dy;
  out[3 * i + 2] = dz;
}

// The delta-scatter decision of both transport kernels (limited: the step
// was cut where a local majorant's bound ends, which takes no collision
// test) and the raw Philox blocks, for known-answer tests.
__global__ void k_delta(int n, const float* st, const float* sm, const float* u, const int* limited, int* real) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  real[i] = (!limited[i] && real_collision(u[i], delta_threshold(st[i], sm[i]))) ? 1 : 0;
}

// Source photons of histories [0, n) as the transport kernels draw them.
__global__ void k_source(const mdx_source_dev* S, int n, uint64_t seed, float* dir, float* E) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Philox rng;
  rng.init(seed, 0u, (uint64_t)i);
  float px, py, pz, dx, dy, dz, e;
  generate(S, rng, px, py, pz, dx, dy, dz, e);
  dir[3 * i] = dx;
  dir[3 * i + 1] = dy;
  dir[3 * i + 2] = dz;
  E[i] = e;
}

__global__ void k_philox(int n, int words, const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (words == 4) {
    philox4x32_10(ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3], key[2 * i], key[2 * i + 1], out[4 * i],
                  out[4 * i + 1], out[4 * i + 2], out[4 * i + 3]);
  } else {
    philox2x32_10(ctr[2 * i], ctr[2 * i + 1], key[i], out[2 * i], out[2 * i + 1]);
  }
}

__global__ void k_kahn(const mdx_tables_dev* T, int slot, int n, const float* E, uint64_t seed, float* mu,
                       float* Ep) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Philox rng;
  rng.init(seed, 0x4B41u, (uint64_t)i);
  uint32_t a = 0, b = 0;
  float m;
  Ep[i] = compton(T, T->mat[slot], rng, E[i], m, a, b);
  mu[i] = m;
}

__global__ void k_zero(double* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// Clears the ranges of a list {count, (address, bytes) x count} read from
// device memory (a round buffer), so that one captured launch serves any
// list.  Addresses are 16-byte aligned and sizes multiples of 4 bytes.
__global__ void k_zero_list(const uint64_t* __restrict__ list) {
  const uint64_t n = list[0];
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (uint64_t r = 0; r < n; ++r) {
    char* base = reinterpret_cast<char*>(list[1 + 2 * r]);
    const size_t bytes = list[2 + 2 * r];
    uint4* q = reinterpret_cast<uint4*>(base);
    const size_t n16 = bytes / 16;
    for (size_t i = tid; i < n16; i += stride) q[i] = make_uint4(0u, 0u, 0u, 0u);
    const size_t tail = (bytes - n16 * 16) / 4;
    if (tid < tail) reinterpret_cast<uint32_t*>(base + n16 * 16)[tid] = 0u;
  }
}

__global__ void k_zero_u32(uint32_t* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

__global__ void k_zero_u64(unsigned long long* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// Per voxel: the mean deposit per history m = S / N and the relative
// standard error of that mean, sqrt((S2 / N - m^2) / (N - 1)) / m (0 where
// nothing was deposited), from the (S, S2) of the FP64 voxel records, as
// FP32 pairs.
__global__ void k_dose_f32(const double* __restrict__ rec, size_t nv, double n, float2* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const double2 v = *reinterpret_cast<const double2*>(rec + MDX_DOSE_REC * i);
  const double m = v.x / n;
  const double var = fmax(v.y / n - m * m, 0.0) / fmax(n - 1.0, 1.0);
  out[i] = make_float2((float)m, m > 0.0 ? (float)(sqrt(var) / m) : 0.0f);
}

// Sums (deposits, sum, sum of squares) of the FP64 voxel records over each
// requested box, one block per box: out[3 b .. 3 b + 2] = (n, sum, sum^2).
__global__ void __launch_bounds__(256) k_roi_stats(const mdx_roi_req* __restrict__ reqs, double* __restrict__ out) {
  const mdx_roi_req q = reqs[blockIdx.x];
  const int sx = q.hi[0] - q.lo[0], sy = q.hi[1] - q.lo[1], sz = q.hi[2] - q.lo[2];
  const int nv = sx * sy * sz;
  double n = 0.0, s = 0.0, s2 = 0.0;
  for (int v = threadIdx.x; v < nv; v += blockDim.x) {
    const int i = q.lo[0] + v % sx, j = q.lo[1] + (v / sx) % sy, k = q.lo[2] + v / (sx * sy);
    const double* r = q.rec + MDX_DOSE_REC * ((size_t)i + (size_t)q.nx * ((size_t)j + (size_t)q.ny * (size_t)k));
    s += r[0];
    s2 += r[1];
    n += (double)*reinterpret_cast<const unsigned long long*>(r + 2);
  }
  __shared__ double red[3][256];
  red[0][threadIdx.x] = n;
  red[1][threadIdx.x] = s;
  red[2][threadIdx.x] = s2;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int f = 0; f < 3; ++f) red[f][threadIdx.x] += red[f][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 3) out[3 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// A few 64-bit words from device memory to pinned host memory (the round's
// event counters), written by the kernel instead of a copy engine.
__global__ void k_copy_u64(const unsigned long long* __restrict__ src, unsigned long long* __restrict__ dst, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

// Plain streaming copy (16 B per lane per iteration): the HBM reference for
// scaling traffic counters.
__global__ void k_copy16(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// Scattered deposits: deposit i adds (v, v^2, 1) to the 32-byte record of
// a hashed index among n_rec, with the three device-scope atomics of a voxel
// deposit.  The reference for the atomic rate of the transport kernels.
__global__ void k_atomic_probe(double* __restrict__ rec, uint64_t n_rec, uint64_t n_dep, uint64_t seed) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_dep;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t h = (i + seed) * 0x9E3779B97F4A7C15ull;  // splitmix64 finaliser
    h = (h ^ (h >> 30)) * 0xBF58476D1CE4E5B9ull;
    h = (h ^ (h >> 27)) * 0x94D049BB133111EBull;
    h ^= h >> 31;
    double* p = rec + 4 * __umul64hi(h, n_rec);  // index in [0, n_rec)
    const double v = 1.0 + (double)(h & 1023u) * (1.0 / 1024.0);
    gadd(p, v);
    gadd(p + 1, v * v);
    gadd(reinterpret_cast<unsigned long long*>(p + 2), 1ull);
  }
}

inline unsigned blocks_for(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

// Static scratch of mdx_run_jobs_noalloc: the launch's event counters, the
// jobs' work counters, prefix offsets and job records, at the offsets of
// mdx_launch::ScratchLayout (each region 256-byte aligned).  One copy per
// device, held until the process ends; calls that overlap on one device
// share it and race.
constexpr size_t kAlign = 256;
constexpr size_t round_up(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }
constexpr size_t kScratchCounters = 0;
constexpr size_t kScratchWork = round_up(MDX_NCOUNTER * sizeof(unsigned long long));
constexpr size_t kScratchPrefix = kScratchWork + round_up(MDX_NOALLOC_MAX_JOBS * sizeof(unsigned long long));
constexpr size_t kScratchJobs = kScratchPrefix + round_up((MDX_NOALLOC_MAX_JOBS + 1) * sizeof(uint64_t));
constexpr size_t kScratchBytes = kScratchJobs + round_up(MDX_NOALLOC_MAX_JOBS * sizeof(mdx_job_dev));

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

}  // namespace

namespace mdx_launch {

ScratchLayout scratch_layout() {
  return ScratchLayout{kScratchCounters, kScratchWork, kScratchPrefix, kScratchJobs, kScratchBytes};
}

// The device address of this device's scratch array, looked up once per
// device (with that device current) and kept: in host code the array's name
// is a host-side stand-in, never a device address.
hipError_t scratch_address(int device, void** out) {
  static std::mutex mu;
  static void* addr[64] = {nullptr};
  if (device < 0 || device >= 64) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[device] = p;
  }
  *out = addr[device];
  return hipSuccess;
}

hipError_t transport_plain(const mdx_tables_dev* T, const mdx_domain_dev* D, const mdx_source_dev* S,
                           const mdx_tallies_dev* TL, uint64_t first, uint64_t n, uint64_t seed, uint32_t job,
                           unsigned long long* counters, int grid_blocks, hipStream_t st) {
  unsigned blocks = (unsigned)grid_blocks;
  uint64_t need = (n + kBlock - 1) / kBlock;
  if (need < blocks) blocks = (unsigned)(need ? need : 1);
  hipLaunchKernelGGL(k_transport_plain, dim3(blocks), dim3(kBlock), 0, st, T, D, S, TL, first, n, seed, job,
                     counters);
  return hipGetLastError();
}

hipError_t transport_queue(const mdx_tables_dev* T, const mdx_job_dev* jobs, const uint64_t* prefix, int n_jobs,
                           uint64_t total, unsigned long long* work, unsigned long long* counters, int grid_blocks,
                           int refill_min, int interact_min, uint32_t lds_voxels, hipStream_t st) {
  unsigned blocks = (unsigned)grid_blocks;
  uint64_t need = (total + kBlock - 1) / kBlock;
  if (need < blocks) blocks = (unsigned)(need ? need : 1);
  refill_min = refill_min < 1 ? 1 : (refill_min > 64 ? 64 : refill_min);
  interact_min = interact_min < 1 ? 1 : (interact_min > 64 ? 64 : interact_min);
  if (n_jobs == 1) {
    const size_t lds = (size_t)lds_voxels * 2 * sizeof(float);
    if (lds > 0) {
      hipError_t e = hipFuncSetAttribute((const void*)k_transport_queue<false>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_transport_queue<false>, dim3(blocks), dim3(kBlock), lds, st, T, jobs, prefix, n_jobs, total,
                       work, counters, refill_min, interact_min, lds_voxels);
  } else {
    hipLaunchKernelGGL(k_transport_queue<true>, dim3(blocks), dim3(kBlock), 0, st, T, jobs, prefix, n_jobs, total,
                       work, counters, refill_min, interact_min, 0u);
  }
  return hipGetLastError();
}

template <bool MULTI>
static hipError_t launch_bank(int bank, unsigned blocks, const mdx_tables_dev* T, const mdx_job_dev* jobs,
                              const uint64_t* prefix, int n_jobs, uint64_t total, unsigned long long* work,
                              unsigned long long* counters, int refill_min, int interact_min, hipStream_t st) {
  if (bank == 96)
    hipLaunchKernelGGL((k_transport_bank<MULTI, 96>), dim3(blocks), dim3(kBlock), 0, st, T, jobs, prefix, n_jobs,
                       total, work, counters, refill_min, interact_min);
  else if (bank == 128)
    hipLaunchKernelGGL((k_transport_bank<MULTI, 128>), dim3(blocks), dim3(kBlock), 0, st, T, jobs, prefix, n_jobs,
                       total, work, counters, refill_min, interact_min);
  else if (bank == 160)
    hipLaunchKernelGGL((k_transport_bank<MULTI, 160>), dim3(blocks), dim3(kBlock), 0, st, T, jobs, prefix, n_jobs,
                       total, work, counters, refill_min, interact_min);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t transport_bank(const mdx_tables_dev* T, const mdx_job_dev* jobs, const uint64_t* prefix, int n_jobs,
                          uint64_t total, unsigned long long* work, unsigned long long* counters, int grid_blocks,
                          int refill_min, int interact_min, int bank, hipStream_t st) {
  unsigned blocks = (unsigned)grid_blocks;
  const uint64_t need = (total + kBlock - 1) / kBlock;
  if (need < blocks) blocks = (unsigned)(need ? need : 1);
  refill_min = refill_min < 1 ? 1 : (refill_min > 64 ? 64 : refill_min);
  interact_min = interact_min < 1 ? 1 : (interact_min > 64 ? 64 : interact_min);
  if (n_jobs == 1)
    return launch_bank<false>(bank, blocks, T, jobs, prefix, n_jobs, total, work, counters, refill_min, interact_min,
                              st);
  return launch_bank<true>(bank, blocks, T, jobs, prefix, n_jobs, total, work, counters, refill_min, interact_min, st);
}

hipError_t transport_bank_round(const mdx_tables_dev* T, const char* round, unsigned long long* work,
                                unsigned long long* counters, int grid_blocks, int refill_min, int interact_min,
                                int bank, hipStream_t st) {
  if (bank == 96)
    hipLaunchKernelGGL(k_transport_bank_round<96>, dim3((unsigned)grid_blocks), dim3(kBlock), 0, st, T, round, work,
                       counters, refill_min, interact_min);
  else if (bank == 128)
    hipLaunchKernelGGL(k_transport_bank_round<128>, dim3((unsigned)grid_blocks), dim3(kBlock), 0, st, T, round, work,
                       counters, refill_min, interact_min);
  else if (bank == 160)
    hipLaunchKernelGGL(k_transport_bank_round<160>, dim3((unsigned)grid_blocks), dim3(kBlock), 0, st, T, round, work,
                       counters, refill_min, interact_min);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

size_t transport_queue_static_lds(bool on_device) {
  if (on_device) {
    hipFuncAttributes a;
    if (hipFuncGetAttributes(&a, (const void*)k_transport_queue<false>) == hipSuccess) return a.sharedSizeBytes;
  }
  return kQueueStaticLdsBound;
}

hipError_t transport_queue_round(const mdx_tables_dev* T, const char* round, unsigned long long* work,
                                 unsigned long long* counters, int grid_blocks, int refill_min, int interact_min,
                                 hipStream_t st) {
  hipLaunchKernelGGL(k_transport_queue_round, dim3((unsigned)grid_blocks), dim3(kBlock), 0, st, T, round, work,
                     counters, refill_min, interact_min);
  return hipGetLastError();
}

hipError_t eval_curve(const mdx_tables_dev* T, int slot, int which, int n, const float* x, float* out,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_eval_curve, dim3(blocks_for(n, 256)), dim3(256), 0, st, T, slot, which, n, x, out);
  return hipGetLastError();
}

hipError_t rita(const mdx_tables_dev* T, int slot, int n, const float* E, const float* y, float* mu,
                hipStream_t st) {
  hipLaunchKernelGGL(k_rita, dim3(blocks_for(n, 256)), dim3(256), 0, st, T, slot, n, E, y, mu);
  return hipGetLastError();
}

hipError_t voxel(const mdx_tables_dev* T, const mdx_domain_dev* D, int n, const float* pos, int* grid,
                 long long* lin, int* mat, hipStream_t st) {
  hipLaunchKernelGGL(k_voxel, dim3(blocks_for(n, 256)), dim3(256), 0, st, T, D, n, pos, grid, lin, mat);
  return hipGetLastError();
}

hipError_t select(int n, const float* sc, const float* si, const float* stt, const float* u, int* type,
                  hipStream_t st) {
  hipLaunchKernelGGL(k_select, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, sc, si, stt, u, type);
  return hipGetLastError();
}

hipError_t rotate(int n, const float* dir, const float* mu, const float* phi, float* out, hipStream_t st) {
  hipLaunchKernelGGL(k_rotate, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, dir, mu, phi, out);
  return hipGetLastError();
}

hipError_t delta(int n, const float* stt, const float* sm, const float* u, const int* limited, int* real,
                 hipStream_t st) {
  hipLaunchKernelGGL(k_delta, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, stt, sm, u, limited, real);
  return hipGetLastError();
}

hipError_t source(const mdx_source_dev* S, int n, uint64_t seed, float* dir, float* E, hipStream_t st) {
  hipLaunchKernelGGL(k_source, dim3(blocks_for(n, 256)), dim3(256), 0, st, S, n, seed, dir, E);
  return hipGetLastError();
}

hipError_t philox(int n, int words, const uint32_t* ctr, const uint32_t* key, uint32_t* out, hipStream_t st) {
  hipLaunchKernelGGL(k_philox, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, words, ctr, key, out);
  return hipGetLastError();
}

hipError_t kahn(const mdx_tables_dev* T, int slot, int n, const float* E, uint64_t seed, float* mu, float* Ep,
                hipStream_t st) {
  hipLaunchKernelGGL(k_kahn, dim3(blocks_for(n, 256)), dim3(256), 0, st, T, slot, n, E, seed, mu, Ep);
  return hipGetLastError();
}

hipError_t zero(double* p, size_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero, dim3(blocks_for(n, 256)), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

hipError_t zero_list(const uint64_t* list, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_list, dim3(2048), dim3(256), 0, st, list);
  return hipGetLastError();
}

hipError_t zero_u32(uint32_t* p, size_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_u32, dim3(blocks_for(n, 256)), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

hipError_t dose_f32(const double* ss, size_t nv, double n, float* out, hipStream_t st) {
  if (nv == 0) return hipSuccess;
  hipLaunchKernelGGL(k_dose_f32, dim3(blocks_for(nv, 256)), dim3(256), 0, st, ss, nv, n,
                     reinterpret_cast<float2*>(out));
  return hipGetLastError();
}

hipError_t rita_build(const double* lx, const double* ly, int n_ff, double alpha, double inv_mc2,
                      const double* energies, int ne, double* xs, double* ys, double* as, double* bs, int* count,
                      hipStream_t st) {
  if (ne == 0) return hipSuccess;
  RitaFF F{lx, ly, n_ff, alpha, inv_mc2};
  hipLaunchKernelGGL(k_rita_build, dim3((unsigned)ne), dim3(64), 0, st, F, energies, xs, ys, as, bs, count);
  return hipGetLastError();
}

int rita_cap() { return kRitaCap; }

hipError_t roi_stats(const mdx_roi_req* reqs, int n, double* out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_roi_stats, dim3((unsigned)n), dim3(256), 0, st, reqs, out);
  return hipGetLastError();
}

hipError_t copy16(const void* src, void* dst, size_t n16, int blocks, hipStream_t st) {
  if (n16 == 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(256), 0, st, static_cast<const float4*>(src),
                     static_cast<float4*>(dst), n16);
  return hipGetLastError();
}

hipError_t atomic_probe(double* rec, uint64_t n_rec, uint64_t n_dep, uint64_t seed, int blocks, hipStream_t st) {
  if (n_dep == 0 || n_rec == 0) return hipSuccess;
  hipLaunchKernelGGL(k_atomic_probe, dim3((unsigned)blocks), dim3(256), 0, st, rec, n_rec, n_dep, seed);
  return hipGetLastError();
}

hipError_t copy_u64(const unsigned long long* src, unsigned long long* dst, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_u64, dim3(1), dim3(64), 0, st, src, dst, n);
  return hipGetLastError();
}

hipError_t zero_u64(unsigned long long* p, size_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_u64, dim3(blocks_for(n, 256)), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

}  // namespace mdx_launch
