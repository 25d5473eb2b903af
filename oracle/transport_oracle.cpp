// ORACLE (test infrastructure only) -- FP64 C++/OpenMP restatement of the
// reference's photon-transport path.  Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg load it, as the checker or the timed CPU
// baseline; the library never calls it.
//
// It follows, step for step (paths relative to the reference tree):
//   runSimulation                 src/MIDSX/Core/run_simulation.cpp:6-41
//   PhysicsEngine::transportPhoton / transportPhotonOneStep / setInteractionType /
//     getFreePath / isDeltaScatter / processTallies      physics_engine.cpp:17-210
//   PhotoelectricEffect / CoherentScattering / IncoherentScattering
//                                 photon_interactions.cpp:43-141
//   ParticleHelpers::rotateDirection (quaternions)       particle.cpp:7-45
//   Interpolator::LogLogLinear / LogLogSpline            interpolators.cpp:3-76
//   ContinuousInversion::sample / getXFromY              probability_dist.cpp:73-119
//   DiscreteInversion::sample                            probability_dist.cpp:18-28
//   PhotonSource and the directionalities                source.cpp:44-155
//   ComputationalDomain::isInComputationalDomain / getVoxel
//                                 computational_domain.cpp:16-31, voxel_grid.cpp:43-128
//   Surface / volume tallies and quantities              surface_tally.cpp, volume_tally.cpp,
//                                 surface_quantity.cpp, volume_quantity.cpp
// with a thread_local std::mt19937 per OpenMP thread and
// std::uniform_real_distribution<double>, as probability_dist.h:23-55.
//
// Differences, by design: event values are accumulated as (n, sum, sum^2) per
// quantity and bucket instead of stored one by one (the statistics of
// quantity.cpp:27-94 follow from these); the per-call copies of RITA rows
// (probability_dist.cpp:92-93) and of the spectrum CDF (:20) are not made;
// the per-thread RNG is seeded deterministically from (seed, thread index).
// The tables (nodes, spline knots and control points, RITA rows) are inputs.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <random>
#include <vector>

namespace {

constexpr double PI = 3.14159265358979323846;
constexpr double EPSILON = 1e-9;
constexpr double INF = std::numeric_limits<double>::infinity();
constexpr double MC2 = 0.51099895000E6;
const double ALPHA = MC2 / (std::sqrt(2.0) * 4.135667662e-15 * 299792458E2 * 1E8);

struct V3 {
  double x = 0, y = 0, z = 0;
  V3() = default;
  V3(double a, double b, double c) : x(a), y(b), z(c) {}
  V3 operator+(const V3& o) const { return {x + o.x, y + o.y, z + o.z}; }
  V3 operator-(const V3& o) const { return {x - o.x, y - o.y, z - o.z}; }
  V3 operator*(double s) const { return {x * s, y * s, z * s}; }
  double dot(const V3& o) const { return x * o.x + y * o.y + z * o.z; }
  V3 cross(const V3& o) const { return {y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x}; }
  double norm() const { return std::sqrt(dot(*this)); }
  V3 normalized() const { double n = norm(); return {x / n, y / n, z / n}; }
  double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

struct LogLogLinear {  // interpolators.cpp:37-76
  std::vector<double> lx, ly;
  void set(int n, const double* x, const double* y) {
    lx.resize(n);
    ly.resize(n);
    for (int i = 0; i < n; ++i) {
      lx[i] = std::log10(x[i]);
      ly[i] = std::log10(y[i]);
    }
  }
  double operator()(double x) const {
    double l = std::log10(x);
    double v;
    if (l <= lx.front()) v = ly.front();
    else if (l >= lx.back()) v = ly.back();
    else {
      int idx = (int)(std::upper_bound(lx.begin(), lx.end(), l) - lx.begin());
      double x1 = lx[idx - 1], x2 = lx[idx], y1 = ly[idx - 1], y2 = ly[idx];
      v = y1 + (y2 - y1) / (x2 - x1) * (l - x1);
    }
    return std::pow(10, v);
  }
};

struct LogLogSpline {  // interpolators.cpp:3-31, Eigen Spline::operator() (span + Cox-de Boor)
  std::vector<double> t, c;
  double xmin = 0, xmax = 1;
  int p = 3;
  double operator()(double x) const {
    double u = (std::log10(x) - xmin) / (xmax - xmin);
    int span;
    if (u <= t[0]) span = p;
    else span = (int)(std::upper_bound(t.begin() + p - 1, t.end() - p - 1, u) - t.begin()) - 1;
    double N[8], left[8], right[8];
    N[0] = 1;
    for (int j = 1; j <= p; ++j) {
      left[j] = u - t[span + 1 - j];
      right[j] = t[span + j] - u;
      double saved = 0;
      for (int r = 0; r < j; ++r) {
        double tmp = N[r] / (right[r + 1] + left[j - r]);
        N[r] = saved + right[r + 1] * tmp;
        saved = left[j - r] * tmp;
      }
      N[j] = saved;
    }
    double v = 0;
    for (int r = 0; r <= p; ++r) v += N[r] * c[span - p + r];
    return std::pow(10, v);
  }
};

struct Rita {  // probability_dist.cpp:73-119
  std::vector<double> E;
  std::vector<std::vector<double>> x, y, a, b;
  static int lower_minus_one(const std::vector<double>& v, double q) {
    return (int)(std::lower_bound(v.begin(), v.end(), q) - v.begin()) - 1;
  }
  double x_from_y(int ie, double yy) const {
    const auto& Y = y[ie];
    int i = lower_minus_one(Y, yy);
    if (i == (int)Y.size() - 1) return 1.0;
    if (i < 0) return -1.0;
    double yi = Y[i], yi1 = Y[i + 1], xi = x[ie][i], xi1 = x[ie][i + 1], ai = a[ie][i], bi = b[ie][i];
    double nu = yy - yi, d = yi1 - yi;
    return xi + (1 + ai + bi) * d * nu / (d * d + ai * d * nu + bi * nu * nu) * (xi1 - xi);
  }
  double sample(double e, double yy) const {
    // lower_bound - 1 clamped to the end intervals: the reference reads out of
    // its arrays for energies at or beyond the grid's ends (probability_dist.cpp:76)
    int lo = std::max(0, std::min(lower_minus_one(E, e), (int)E.size() - 2)), up = lo + 1;
    double xl = x_from_y(lo, yy), xu = x_from_y(up, yy);
    return std::exp(std::log(xl + 1) + (std::log(xu + 1) - std::log(xl + 1)) * (std::log(e) - std::log(E[lo])) /
                                           (std::log(E[up]) - std::log(E[lo]))) -
           1.0;
  }
};

struct Material {
  LogLogLinear total, sf;
  LogLogSpline coh, incoh;
  Rita rita;
};

struct Grid {
  std::vector<uint8_t> ids;
  int n[3];
  V3 spacing, origin, dim;
};

// Tally description: same field order as mdx_tally in include/midsx_amd.h.
struct TallyDesc {
  int kind;
  double p0[3], p1[3], p2[3], normal[3];
  double radius;
  int nx, ny, hist_bins;
  double hist_lo, hist_width;
};

struct Photon {  // photon.h:25-104, particle.h:36-147
  V3 pos, dir;
  double E = 0;
  bool terminated = false;
  int coh = 0, inc = 0;
  int total() const { return coh + inc; }
};

struct TempSurface {  // tally_data.h:24-31
  double free_path = 0;
  bool isInteract = false;
  double entrance_cosine = 0;
  Photon initial;
};
struct TempVolume {  // tally_data.h:33-44
  double energy_deposited = 0, free_path = 0;
  bool isInteract = false, already_counted = false;
  Photon initial, final_;
};
struct TempVoxel {
  int grid;  // -1 background
  int64_t lin;
  double dep;
};

struct TallyAcc {
  std::vector<double> acc;    // [4][4][3]
  std::vector<double> hist;   // [4][bins]
  std::vector<double> image;  // [ny*nx][3]
};

int bucket(const Photon& p) {  // *_quantity.cpp measure()
  if (p.total() > 1) return 3;
  if (p.coh == 1) return 2;
  if (p.inc == 1) return 1;
  return 0;
}

struct Oracle {
  std::map<int, Material> mats;
  LogLogLinear majorant;
  V3 dim_space;
  int bg_id = 0;
  std::vector<Grid> grids;
  bool voxel_dose = false;
  // source
  int dir_kind = 0, spec_kind = 0;
  V3 pos, sa, sb, sc;
  double radius = 0, energy = 0;
  std::vector<double> cdf, bin_e;
  std::vector<TallyDesc> tallies;
  std::vector<TallyAcc> acc;
  std::vector<std::vector<double>> dose;  // per grid: sum, sum^2 interleaved
  std::vector<std::vector<uint32_t>> dose_n;
  double bg_dose[3] = {0, 0, 0};
  uint64_t counters[12] = {0};
};

struct Rng {
  std::mt19937 gen;
  std::uniform_real_distribution<double> u01{0.0, 1.0};
  double operator()() { return u01(gen); }
};

V3 rotate(const V3& v, double theta, double phi) {  // particle.cpp:7-45 (quaternion form)
  V3 perp;
  if (std::abs(v.x) <= std::abs(v.y) && std::abs(v.x) <= std::abs(v.z)) perp = V3(1, 0, 0).cross(v);
  else if (std::abs(v.y) <= std::abs(v.z)) perp = V3(0, 1, 0).cross(v);
  else perp = V3(0, 0, 1).cross(v);
  perp = perp.normalized();
  struct Q { double w, x, y, z; };
  auto mul = [](const Q& a, const Q& b) {
    return Q{a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
             a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x, a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w};
  };
  double ct = std::cos(theta / 2), st = std::sin(theta / 2), cp = std::cos(phi / 2), sp = std::sin(phi / 2);
  Q qt{ct, perp.x * st, perp.y * st, perp.z * st};
  Q qp{cp, v.x * sp, v.y * sp, v.z * sp};
  Q q = mul(qp, qt);
  double n2 = q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z;
  Q qi{q.w / n2, -q.x / n2, -q.y / n2, -q.z / n2};
  Q r = mul(mul(q, Q{0, v.x, v.y, v.z}), qi);
  return {r.x, r.y, r.z};
}

// AACuboidVolumeTally::determineVolumeTraversal (volume_tally.cpp:38-91):
// the slab lengths along the step's line to the box's planes (an axis along
// which the direction is below EPSILON constrains nothing), entering length
// = the largest of the near ones, exiting = the smallest of the far ones.
enum { kMisses = 0, kLandsInside = 1, kPassesThrough = 2, kStartsInsideExits = 3, kStartsInsideStays = 4 };
int aacuboid_traversal(const double* mn, const double* mx, const V3& pos, const V3& dir, double fp) {
  double lo[3], hi[3];
  for (int i = 0; i < 3; ++i) {
    const double di = dir[i];
    if (std::abs(di) < EPSILON) {
      lo[i] = -INF;
      hi[i] = INF;
      continue;
    }
    lo[i] = (mn[i] - pos[i]) / di;
    hi[i] = (mx[i] - pos[i]) / di;
    if (lo[i] > hi[i]) std::swap(lo[i], hi[i]);
  }
  const double enter = std::max(lo[0], std::max(lo[1], lo[2]));
  const double exit = std::min(hi[0], std::min(hi[1], hi[2]));
  if (enter > exit || exit < 0 || enter > fp) return kMisses;
  if (enter <= 0) return exit >= fp ? kStartsInsideStays : kStartsInsideExits;
  return exit < fp ? kPassesThrough : kLandsInside;
}

struct Engine {
  Oracle& O;
  explicit Engine(Oracle& o) : O(o) {}

  Photon generate(Rng& rng) {  // source.cpp:150-155
    Photon p;
    p.pos = O.pos;
    V3 d;
    switch (O.dir_kind) {
      case 1: {  // IsotropicDirectionality
        double th = PI * rng(), ph = 2 * PI * rng();
        d = V3(std::sin(th) * std::cos(ph), std::sin(th) * std::sin(ph), std::cos(th));
        break;
      }
      case 2: {  // RectangularIsotropicDirectionality
        double a1 = rng(), a2 = rng();
        d = (O.sa + O.sb * a1 + O.sc * a2 - p.pos).normalized();
        break;
      }
      case 3: {  // DiscIsotropicDirectionality (sb = u, sc = v)
        double th = 2 * PI * rng(), r = O.radius * rng();
        d = (O.sa + (O.sb * std::cos(th) + O.sc * std::sin(th)) * r - p.pos).normalized();
        break;
      }
      default:
        d = (O.sa - p.pos).normalized();
    }
    p.dir = d;
    if (O.spec_kind == 1) {
      double u = rng();
      int idx = (int)(std::lower_bound(O.cdf.begin(), O.cdf.end(), u) - O.cdf.begin());
      if (idx >= (int)O.cdf.size()) idx = (int)O.cdf.size() - 1;
      p.E = O.bin_e[idx];
    } else {
      p.E = O.energy;
    }
    return p;
  }

  bool in_domain(const V3& p) const {
    return p.x >= 0 && p.x <= O.dim_space.x && p.y >= 0 && p.y <= O.dim_space.y && p.z >= 0 && p.z <= O.dim_space.z;
  }

  int voxel(const V3& p, int64_t& lin, int& mid) const {
    for (size_t g = 0; g < O.grids.size(); ++g) {
      const Grid& G = O.grids[g];
      V3 q = p - G.origin;
      if (q.x >= 0 && q.y >= 0 && q.z >= 0 && q.x <= G.dim.x && q.y <= G.dim.y && q.z <= G.dim.z) {
        int i = (int)(q.x / G.spacing.x), j = (int)(q.y / G.spacing.y), k = (int)(q.z / G.spacing.z);
        i = std::min(i, G.n[0] - 1);
        j = std::min(j, G.n[1] - 1);
        k = std::min(k, G.n[2] - 1);
        lin = i + (int64_t)j * G.n[0] + (int64_t)k * G.n[0] * G.n[1];
        mid = G.ids[lin];
        return (int)g;
      }
    }
    lin = -1;
    mid = O.bg_id;
    return -1;
  }

  double compton(const Material& m, Photon& ph, Rng& rng, uint64_t* c) {  // photon_interactions.cpp:62-141
    double k = ph.E / MC2, chi, mu, kp;
    for (;;) {
      ++c[6];
      double r1 = rng(), r2 = rng(), r3 = rng();
      if (r1 <= (1 + 2 * k) / (9 + 2 * k)) {
        chi = 1 + 2 * k * r2;
        if (!(r3 <= 4 * (1 / chi - 1 / (chi * chi)))) continue;
      } else {
        chi = (1 + 2 * k) / (1 + 2 * k * r2);
        if (!(2 * r3 <= std::pow(1 / k - chi / k + 1, 2) + 1 / chi)) continue;
      }
      kp = k / chi;
      mu = 1 + 1 / k - 1 / kp;
      ++c[7];
      double x = ALPHA * k * std::sqrt(1 - mu), xm = ALPHA * k * std::sqrt(2.0);
      double accept = m.sf(x) / m.sf(xm);
      if (rng() <= accept) break;
    }
    double phi = 2 * PI * rng();
    ph.dir = rotate(ph.dir, std::acos(mu), phi);
    ph.E = kp * MC2;
    return MC2 * (k - kp);
  }

  void step(Photon& ph, Rng& rng, std::vector<TempSurface>& ts, std::vector<TempVolume>& tv,
            std::vector<TempVoxel>& tx, uint64_t* c) {  // physics_engine.cpp:28-79
    ++c[1];
    TempSurface s;
    TempVolume v;
    s.initial = v.initial = ph;
    double smax = O.majorant(ph.E);
    double fp = -std::log(rng()) / smax;
    s.free_path = v.free_path = fp;
    ph.pos = ph.pos + ph.dir * fp;
    if (!in_domain(ph.pos)) {
      ++c[10];
      ts.push_back(s);
      tv.push_back(v);
      ph.terminated = true;
      return;
    }
    int64_t lin;
    int mid;
    int g = voxel(ph.pos, lin, mid);
    const Material& m = O.mats.at(mid);
    double stot = m.total(ph.E);
    if (!(rng() < 1 - stot / smax)) {
      ++c[2];
      s.isInteract = v.isInteract = true;
      double pc = m.coh(ph.E) / stot, pi = m.incoh(ph.E) / stot;
      double u = rng();
      double dep = 0;
      if (u < pc) {
        ++c[3];
        ph.coh++;
        double mu = m.rita.sample(ph.E, rng());
        double phi = 2 * PI * rng();
        ph.dir = rotate(ph.dir, std::acos(mu), phi);
      } else if (u < pc + pi) {
        ++c[4];
        ph.inc++;
        dep = compton(m, ph, rng, c);
      } else {
        ++c[5];
        ph.terminated = true;
        dep = ph.E;
      }
      tx.push_back({g, lin, dep});
      v.energy_deposited = dep;
      v.final_ = ph;
    }
    ts.push_back(s);
    tv.push_back(v);
  }

  static void event(std::vector<double>& a, int q, int b, double val) {
    double* p = &a[(q * 4 + b) * 3];
    p[0] += 1;
    p[1] += val;
    p[2] += val * val;
  }

  void surface(const TallyDesc& T, TallyAcc& A, TempSurface& d, uint64_t* c) {  // surface_tally.cpp:5-34
    V3 n(T.normal[0], T.normal[1], T.normal[2]);
    double cosv = d.initial.dir.dot(n);
    if (std::abs(cosv) < EPSILON || cosv < 0) return;
    V3 ref(T.p0[0], T.p0[1], T.p0[2]);
    double t = (ref - d.initial.pos).dot(n) / cosv;
    if (t > d.free_path || t < 0) return;
    V3 ip = d.initial.pos + d.initial.dir * t;
    double a = 0, b2 = 0;
    if (T.kind == 1) {
      if ((ip - ref).norm() > T.radius) return;
    } else {
      V3 e1(T.p1[0], T.p1[1], T.p1[2]), e2(T.p2[0], T.p2[1], T.p2[2]);
      V3 ci = ip - ref;
      a = e1.dot(ci);
      b2 = e2.dot(ci);
      if (a < 0 || a > e1.dot(e1) || b2 < 0 || b2 > e2.dot(e2)) return;
    }
    ++c[9];
    d.entrance_cosine = cosv;
    int b = bucket(d.initial);
    double E = d.initial.E;
    if (E != 0) event(A.acc, 0, b, E);
    if (cosv != 0) event(A.acc, 1, b, cosv);
    event(A.acc, 2, b, 1.0);
    if (T.kind == 3) {
      V3 e1(T.p1[0], T.p1[1], T.p1[2]), e2(T.p2[0], T.p2[1], T.p2[2]);
      int px = std::min(T.nx - 1, (int)(a / e1.dot(e1) * T.nx));
      int py = std::min(T.ny - 1, (int)(b2 / e2.dot(e2) * T.ny));
      double* im = &A.image[3 * ((size_t)py * T.nx + px)];
      im[0] += E;
      im[1] += E * E;
      im[2] += 1;
    }
    if (T.hist_bins > 0) {
      double fb = (E - T.hist_lo) / T.hist_width;
      if (fb >= 0 && fb < T.hist_bins) A.hist[b * T.hist_bins + (int)fb] += 1;
    }
  }

  void volume(const TallyDesc& T, TallyAcc& A, TempVolume d) {  // volume_tally.cpp:5-27
    const int tr = aacuboid_traversal(T.p0, T.p1, d.initial.pos, d.initial.dir, d.free_path);
    if (tr == kMisses) return;
    if (tr == kStartsInsideExits || tr == kStartsInsideStays) d.already_counted = true;
    if (tr == kPassesThrough || tr == kStartsInsideExits) d.energy_deposited = 0;
    int b = bucket(d.final_);
    if (d.energy_deposited != 0) event(A.acc, 0, b, d.energy_deposited);
    if (!d.already_counted && d.initial.E != 0) event(A.acc, 1, b, d.initial.E);
    if (!d.already_counted) event(A.acc, 2, b, 1.0);
    if (d.isInteract) event(A.acc, 3, b, 1.0);
  }
};

}  // namespace

extern "C" {

void* orc_create() { return new Oracle(); }
void orc_destroy(void* h) { delete static_cast<Oracle*>(h); }

void orc_set_majorant(void* h, int n, const double* x, const double* y) {
  static_cast<Oracle*>(h)->majorant.set(n, x, y);
}

// Spline arguments: knot count, knots, control points (knot count - degree - 1), xmin, xmax, degree.
void orc_add_material(void* h, int id, int n_tot, const double* tx, const double* ty, int n_sf, const double* sx,
                      const double* sy, int nk_coh, const double* k_coh, const double* c_coh, double xmin_coh,
                      double xmax_coh, int deg_coh, int nk_inc, const double* k_inc, const double* c_inc,
                      double xmin_inc, double xmax_inc, int deg_inc, int ne, const double* rE, const int* row_off,
                      const double* rx, const double* ry, const double* ra, const double* rb) {
  Oracle& O = *static_cast<Oracle*>(h);
  Material m;
  m.total.set(n_tot, tx, ty);
  m.sf.set(n_sf, sx, sy);
  m.coh.t.assign(k_coh, k_coh + nk_coh);
  m.coh.c.assign(c_coh, c_coh + (nk_coh - deg_coh - 1));
  m.coh.xmin = xmin_coh;
  m.coh.xmax = xmax_coh;
  m.coh.p = deg_coh;
  m.incoh.t.assign(k_inc, k_inc + nk_inc);
  m.incoh.c.assign(c_inc, c_inc + (nk_inc - deg_inc - 1));
  m.incoh.xmin = xmin_inc;
  m.incoh.xmax = xmax_inc;
  m.incoh.p = deg_inc;
  m.rita.E.assign(rE, rE + ne);
  for (int i = 0; i < ne; ++i) {
    int r0 = row_off[i], r1 = row_off[i + 1];
    m.rita.x.emplace_back(rx + r0, rx + r1);
    m.rita.y.emplace_back(ry + r0, ry + r1);
    m.rita.a.emplace_back(ra + r0, ra + r1 - 1);
    m.rita.b.emplace_back(rb + r0, rb + r1 - 1);
  }
  O.mats[id] = std::move(m);
}

void orc_set_domain(void* h, const double* dim, int bg) {
  Oracle& O = *static_cast<Oracle*>(h);
  O.dim_space = V3(dim[0], dim[1], dim[2]);
  O.bg_id = bg;
  O.grids.clear();
}

void orc_add_grid(void* h, const uint8_t* ids, const int* dims, const double* spacing, const double* origin) {
  Oracle& O = *static_cast<Oracle*>(h);
  Grid g;
  size_t nv = (size_t)dims[0] * dims[1] * dims[2];
  g.ids.assign(ids, ids + nv);
  for (int k = 0; k < 3; ++k) g.n[k] = dims[k];
  g.spacing = V3(spacing[0], spacing[1], spacing[2]);
  g.origin = V3(origin[0], origin[1], origin[2]);
  g.dim = V3(dims[0] * spacing[0], dims[1] * spacing[1], dims[2] * spacing[2]);
  O.grids.push_back(std::move(g));
}

void orc_set_source(void* h, int dir_kind, int spec_kind, const double* pos, const double* a, const double* b,
                    const double* c, double radius, double energy, int n_bins, const double* be, const double* bw) {
  Oracle& O = *static_cast<Oracle*>(h);
  O.dir_kind = dir_kind;
  O.spec_kind = spec_kind;
  O.pos = V3(pos[0], pos[1], pos[2]);
  O.sa = V3(a[0], a[1], a[2]);
  O.sb = V3(b[0], b[1], b[2]);
  O.sc = V3(c[0], c[1], c[2]);
  O.radius = radius;
  O.energy = energy;
  if (dir_kind == 3) {  // source.cpp:117-131 (b holds the disc normal)
    V3 n = O.sb, u;
    if (std::abs(n.dot(V3(1, 0, 0))) < EPSILON) u = n.cross(V3(0, 1, 0)).normalized();
    else u = n.cross(V3(1, 0, 0)).normalized();
    O.sb = u;
    O.sc = n.cross(u).normalized();
  }
  O.cdf.clear();
  O.bin_e.clear();
  if (spec_kind == 1) {
    double sum = 0;
    for (int i = 0; i < n_bins; ++i) sum += bw[i];
    double acc = 0;
    for (int i = 0; i < n_bins; ++i) {
      acc += sum != 1.0 ? bw[i] / sum : bw[i];
      O.cdf.push_back(acc);
      O.bin_e.push_back(be[i]);
    }
  }
}

void orc_set_tallies(void* h, int n, const void* desc) {
  Oracle& O = *static_cast<Oracle*>(h);
  const TallyDesc* d = static_cast<const TallyDesc*>(desc);
  O.tallies.assign(d, d + n);
  O.acc.assign(n, TallyAcc());
  for (int i = 0; i < n; ++i) {
    TallyDesc& t = O.tallies[i];
    if (t.kind == 2 || t.kind == 3) {
      V3 e1(t.p1[0], t.p1[1], t.p1[2]), e2(t.p2[0], t.p2[1], t.p2[2]);
      V3 nn = e1.cross(e2).normalized();
      t.normal[0] = nn.x;
      t.normal[1] = nn.y;
      t.normal[2] = nn.z;
    }
    O.acc[i].acc.assign(48, 0.0);
    O.acc[i].hist.assign(t.kind != 0 && t.hist_bins > 0 ? 4 * t.hist_bins : 0, 0.0);
    O.acc[i].image.assign(t.kind == 3 ? 3 * (size_t)t.nx * t.ny : 0, 0.0);
  }
}

void orc_enable_voxel_dose(void* h, int on) {
  Oracle& O = *static_cast<Oracle*>(h);
  O.voxel_dose = on != 0;
  O.dose.clear();
  O.dose_n.clear();
  for (auto& g : O.grids) {
    O.dose.emplace_back(on ? 2 * g.ids.size() : 0, 0.0);
    O.dose_n.emplace_back(on ? g.ids.size() : 0, 0u);
  }
}

// runSimulation (run_simulation.cpp:6-41): OpenMP over histories, per-thread
// tallies merged after the loop, then the serial voxel-deposit fold.
// Returns the wall time of the parallel region in seconds.
double orc_run(void* h, uint64_t n, uint64_t seed, int threads) {
  Oracle& O = *static_cast<Oracle*>(h);
  if (threads > 0) omp_set_num_threads(threads);
  const int nt = threads > 0 ? threads : omp_get_max_threads();
  std::vector<std::vector<TallyAcc>> local(nt, O.acc);
  for (auto& l : local)
    for (auto& a : l) {
      std::fill(a.acc.begin(), a.acc.end(), 0.0);
      std::fill(a.hist.begin(), a.hist.end(), 0.0);
      std::fill(a.image.begin(), a.image.end(), 0.0);
    }
  std::vector<std::vector<TempVoxel>> deposits(nt);
  std::vector<std::vector<uint64_t>> cnt(nt, std::vector<uint64_t>(12, 0));
  double t0 = omp_get_wtime();
#pragma omp parallel num_threads(nt)
  {
    int tid = omp_get_thread_num();
    Rng rng;
    rng.gen.seed((uint32_t)(seed * 1000003ull + (uint64_t)tid * 7919ull + 1ull));
    Engine eng(O);
    std::vector<TempSurface> ts;
    std::vector<TempVolume> tv;
    std::vector<TempVoxel> tx;
    uint64_t* c = cnt[tid].data();
#pragma omp for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
      Photon ph = eng.generate(rng);
      ++c[0];
      ts.clear();
      tv.clear();
      tx.clear();
      while (!ph.terminated) eng.step(ph, rng, ts, tv, tx, c);
      // processTallies (physics_engine.cpp:192-210)
      for (size_t t = 0; t < O.tallies.size(); ++t) {
        const TallyDesc& T = O.tallies[t];
        if (T.kind == 0)
          for (auto& d : tv) eng.volume(T, local[tid][t], d);
        else
          for (auto& d : ts) eng.surface(T, local[tid][t], d, c);
      }
      for (auto& d : tx)
        if (d.dep != 0.0) {
          deposits[tid].push_back(d);
          ++c[8];
        }
    }
  }
  double t1 = omp_get_wtime();
  for (int t = 0; t < nt; ++t) {
    for (size_t k = 0; k < O.acc.size(); ++k) {
      for (size_t j = 0; j < O.acc[k].acc.size(); ++j) O.acc[k].acc[j] += local[t][k].acc[j];
      for (size_t j = 0; j < O.acc[k].hist.size(); ++j) O.acc[k].hist[j] += local[t][k].hist[j];
      for (size_t j = 0; j < O.acc[k].image.size(); ++j) O.acc[k].image[j] += local[t][k].image[j];
    }
    for (int j = 0; j < 12; ++j) O.counters[j] += cnt[t][j];
  }
  // addVoxelDataToComputationalDomain (physics_engine.cpp:160-168)
  for (auto& dl : deposits)
    for (auto& d : dl) {
      if (d.grid < 0) {
        O.bg_dose[0] += d.dep;
        O.bg_dose[1] += d.dep * d.dep;
        O.bg_dose[2] += 1;
      } else if (O.voxel_dose) {
        O.dose[d.grid][2 * d.lin] += d.dep;
        O.dose[d.grid][2 * d.lin + 1] += d.dep * d.dep;
        O.dose_n[d.grid][d.lin] += 1;
      }
    }
  return t1 - t0;
}

void orc_get_tally(void* h, int i, double* acc, double* hist, double* image) {
  Oracle& O = *static_cast<Oracle*>(h);
  const TallyAcc& a = O.acc[i];
  if (acc) std::memcpy(acc, a.acc.data(), a.acc.size() * 8);
  if (hist && !a.hist.empty()) std::memcpy(hist, a.hist.data(), a.hist.size() * 8);
  if (image && !a.image.empty()) std::memcpy(image, a.image.data(), a.image.size() * 8);
}

void orc_get_voxel_dose(void* h, int g, double* ss, uint32_t* n) {
  Oracle& O = *static_cast<Oracle*>(h);
  if (ss) std::memcpy(ss, O.dose[g].data(), O.dose[g].size() * 8);
  if (n) std::memcpy(n, O.dose_n[g].data(), O.dose_n[g].size() * 4);
}

void orc_get_background_dose(void* h, double* out) {
  Oracle& O = *static_cast<Oracle*>(h);
  std::memcpy(out, O.bg_dose, 24);
}

void orc_get_counters(void* h, uint64_t* out) { std::memcpy(out, static_cast<Oracle*>(h)->counters, 96); }

void orc_reset(void* h) {
  Oracle& O = *static_cast<Oracle*>(h);
  for (auto& a : O.acc) {
    std::fill(a.acc.begin(), a.acc.end(), 0.0);
    std::fill(a.hist.begin(), a.hist.end(), 0.0);
    std::fill(a.image.begin(), a.image.end(), 0.0);
  }
  for (auto& d : O.dose) std::fill(d.begin(), d.end(), 0.0);
  for (auto& d : O.dose_n) std::fill(d.begin(), d.end(), 0u);
  std::memset(O.bg_dose, 0, sizeof O.bg_dose);
  std::memset(O.counters, 0, sizeof O.counters);
}

// Single-function probes for unit tests (FP64, same code as the run).
double orc_rita_sample(void* h, int id, double E, double y) {
  return static_cast<Oracle*>(h)->mats.at(id).rita.sample(E, y);
}
double orc_interp(void* h, int id, int which, double x) {
  Oracle& O = *static_cast<Oracle*>(h);
  if (id < 0) return O.majorant(x);
  const Material& m = O.mats.at(id);
  switch (which) {
    case 0: return m.total(x);
    case 1: return m.coh(x);
    case 2: return m.incoh(x);
    default: return m.sf(x);
  }
}
// IncoherentScattering::interact alone: n events at energy E in material id;
// writes mu and the scattered energy of each (photon_interactions.cpp:62-141).
void orc_kahn_sample(void* h, int id, double E, int n, uint64_t seed, double* mu, double* Ep) {
  Oracle& O = *static_cast<Oracle*>(h);
  Engine eng(O);
  Rng rng;
  rng.gen.seed((uint32_t)seed);
  uint64_t c[12] = {0};
  const Material& m = O.mats.at(id);
  for (int i = 0; i < n; ++i) {
    Photon ph;
    ph.dir = V3(0, 0, 1);
    ph.E = E;
    eng.compton(m, ph, rng, c);
    mu[i] = ph.dir.z;
    Ep[i] = ph.E;
  }
}

// The traversal class of n steps (position, direction, free path) against
// the box [mn, mx]: 0 misses, 1 lands inside, 2 passes through, 3 starts
// inside and exits, 4 starts inside and stays.
void orc_aacuboid_traversal(const double* mn, const double* mx, int n, const double* pos, const double* dir,
                            const double* fp, int* out) {
  for (int i = 0; i < n; ++i)
    out[i] = aacuboid_traversal(mn, mx, V3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]),
                                V3(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]), fp[i]);
}

// n source photons (PhotonSource::generatePhoton, source.cpp:150-155) from
// one generator seeded with seed: directions (3 per photon) and energies.
void orc_sample_source(void* h, int n, uint64_t seed, double* dir, double* E) {
  Oracle& O = *static_cast<Oracle*>(h);
  Rng rng;
  rng.gen.seed((uint32_t)(seed * 1000003ull + 1ull));
  Engine eng(O);
  for (int i = 0; i < n; ++i) {
    const Photon p = eng.generate(rng);
    dir[3 * i] = p.dir.x;
    dir[3 * i + 1] = p.dir.y;
    dir[3 * i + 2] = p.dir.z;
    E[i] = p.E;
  }
}

// The OpenMP team size a run with threads = 0 uses.
int orc_max_threads() { return omp_get_max_threads(); }

void orc_rotate(const double* d, double theta, double phi, double* out) {
  V3 r = rotate(V3(d[0], d[1], d[2]), theta, phi);
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
}

}  // extern "C"
