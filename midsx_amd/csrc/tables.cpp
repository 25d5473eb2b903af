// Host FP64 physics-table builder.  Follows the reference's table chain
// (src/MIDSX/Core/material_data.cpp, material_properties.cpp,
// interaction_data.cpp, interpolators.cpp, probability_dist.cpp) with the
// SQLite store replaced by the EPDL ENDF-6 files and a CSV registry.
#include "tables.h"

#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <sstream>
#include <stdexcept>
#include <thread>

namespace mdx {

double alpha_inv_angstrom() {
  return kElectronRestMass / (std::sqrt(2.0) * kPlanck * kSpeedOfLight * 1e8);
}

// ---------------------------------------------------------------- ENDF-6 reader
static double endf_number(const char* f, int w) {
  char buf[32];
  int n = 0;
  for (int i = 0; i < w && f[i] && f[i] != '\n'; ++i)
    if (f[i] != ' ') buf[n++] = f[i];
  buf[n] = 0;
  if (n == 0) return 0.0;
  // ENDF short exponent form "1.2345-5": insert the 'e'.
  char fixed[40];
  int m = 0;
  for (int i = 0; i < n; ++i) {
    if ((buf[i] == '+' || buf[i] == '-') && i > 0 && buf[i - 1] != 'e' && buf[i - 1] != 'E')
      fixed[m++] = 'e';
    fixed[m++] = buf[i];
  }
  fixed[m] = 0;
  return std::strtod(fixed, nullptr);
}

static bool read_lines(const std::string& path, std::vector<std::string>& lines) {
  gzFile g = gzopen(path.c_str(), "rb");  // reads plain files too
  if (!g) return false;
  char buf[512];
  while (gzgets(g, buf, sizeof buf)) lines.emplace_back(buf);
  gzclose(g);
  return true;
}

Element load_element(const std::string& dir, int z) {
  char name[64];
  std::vector<std::string> lines;
  std::snprintf(name, sizeof name, "/%03d.endf.gz", z);
  if (!read_lines(dir + name, lines)) {
    std::snprintf(name, sizeof name, "/%03d.txt", z);
    if (!read_lines(dir + name, lines))
      throw std::runtime_error("EPDL data for Z=" + std::to_string(z) + " not found in " + dir);
  }
  Element e;
  e.z = z;
  std::map<int, std::vector<const std::string*>> sec;
  double awr = -1;
  for (const auto& ln : lines) {
    if (ln.size() < 75) continue;
    int mf = std::atoi(ln.substr(70, 2).c_str());
    int mt = std::atoi(ln.substr(72, 3).c_str());
    if (awr < 0 && mf == 1 && mt == 451) awr = endf_number(ln.c_str() + 11, 11);
    sec[mf * 1000 + mt].push_back(&ln);
  }
  if (awr <= 0) throw std::runtime_error("EPDL file for Z=" + std::to_string(z) + " lacks a HEAD record");
  e.atomic_weight = awr * kNeutronMassU;
  auto tab1 = [&](int key, Table& t, bool drop_zero_x) {
    auto it = sec.find(key);
    if (it == sec.end() || it->second.size() < 3)
      throw std::runtime_error("EPDL section MF/MT " + std::to_string(key) + " missing for Z=" + std::to_string(z));
    const auto& ls = it->second;
    int nr = (int)endf_number(ls[1]->c_str() + 44, 11);
    int np = (int)endf_number(ls[1]->c_str() + 55, 11);
    size_t first = 2 + (2 * nr + 5) / 6;
    std::vector<double> v;
    v.reserve(2 * np);
    for (size_t i = first; i < ls.size() && (int)v.size() < 2 * np; ++i)
      for (int k = 0; k < 6 && (int)v.size() < 2 * np; ++k) {
        const char* f = ls[i]->c_str() + 11 * k;
        bool blank = true;
        for (int c = 0; c < 11; ++c)
          if (f[c] != ' ') blank = false;
        if (!blank) v.push_back(endf_number(f, 11));
      }
    for (int i = 0; i + 1 < (int)v.size(); i += 2) {
      // MF27 tables start at x = 0, whose log10 is -inf: dropped, so values
      // below the first positive node clamp to it.
      if (drop_zero_x && v[i] <= 0.0) continue;
      t.x.push_back(v[i]);
      t.y.push_back(v[i + 1]);
    }
  };
  tab1(23502, e.coh, false);
  tab1(23504, e.incoh, false);
  tab1(23522, e.photo, false);
  tab1(27502, e.ff, true);
  tab1(27504, e.sf, true);
  return e;
}

// ---------------------------------------------------------------- registry
std::map<int, MaterialDef> load_registry(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("material registry not found: " + path);
  std::map<int, MaterialDef> reg;
  std::string ln;
  while (std::getline(f, ln)) {
    if (ln.empty() || ln[0] == '#') continue;
    size_t a = ln.find('"'), b = ln.find('"', a + 1);
    if (a == std::string::npos || b == std::string::npos) continue;
    MaterialDef d;
    d.id = std::atoi(ln.substr(0, a).c_str());
    d.name = ln.substr(a + 1, b - a - 1);
    std::string rest = ln.substr(b + 2);
    size_t c = rest.find(',');
    d.density = std::strtod(rest.substr(0, c).c_str(), nullptr);
    std::stringstream comp(rest.substr(c + 1));
    std::string item;
    while (std::getline(comp, item, ';')) {
      size_t k = item.find(':');
      if (k == std::string::npos) continue;
      d.composition.emplace_back(std::atoi(item.substr(0, k).c_str()),
                                 std::strtod(item.substr(k + 1).c_str(), nullptr));
    }
    std::sort(d.composition.begin(), d.composition.end());
    reg[d.id] = d;
  }
  return reg;
}

// ---------------------------------------------------------------- interpolators
LogLogLinear::LogLogLinear(const std::vector<double>& x, const std::vector<double>& y) {
  lx.resize(x.size());
  ly.resize(y.size());
  for (size_t i = 0; i < x.size(); ++i) {
    lx[i] = std::log10(x[i]);
    ly[i] = std::log10(y[i]);
  }
}

double LogLogLinear::operator()(double x) const {
  double l = std::log10(x);
  if (l <= lx.front()) return std::pow(10.0, ly.front());
  if (l >= lx.back()) return std::pow(10.0, ly.back());
  size_t idx = std::upper_bound(lx.begin(), lx.end(), l) - lx.begin();
  double x1 = lx[idx - 1], x2 = lx[idx], y1 = ly[idx - 1], y2 = ly[idx];
  double slope = (y2 - y1) / (x2 - x1);
  return std::pow(10.0, y1 + slope * (l - x1));
}

// B-spline basis (Piegl & Tiller, The NURBS Book, Alg. A2.2).
static void basis(double u, int span, int p, const std::vector<double>& t, double* N) {
  double left[8], right[8];
  N[0] = 1.0;
  for (int j = 1; j <= p; ++j) {
    left[j] = u - t[span + 1 - j];
    right[j] = t[span + j] - u;
    double saved = 0.0;
    for (int r = 0; r < j; ++r) {
      double tmp = N[r] / (right[r + 1] + left[j - r]);
      N[r] = saved + right[r + 1] * tmp;
      saved = left[j - r] * tmp;
    }
    N[j] = saved;
  }
}

int LogLogSpline::span(double u) const {
  // Knot span of u (The NURBS Book Alg. A2.1; the clamped form Eigen's
  // Spline::Span uses): ends of the parameter range map to the end spans.
  if (u <= knots[0]) return degree;
  auto b = knots.begin() + (degree - 1);
  auto e = knots.end() - degree - 1;
  return (int)(std::upper_bound(b, e, u) - knots.begin()) - 1;
}

double LogLogSpline::eval_u(double u) const {
  int s = span(u);
  double N[8];
  basis(u, s, degree, knots, N);
  double v = 0;
  for (int r = 0; r <= degree; ++r) v += N[r] * ctrl[s - degree + r];
  return v;
}

LogLogSpline::LogLogSpline(const std::vector<double>& x, const std::vector<double>& y) {
  const int n = (int)x.size();
  if (n < 2) throw std::invalid_argument("LogLogSpline needs at least two nodes");
  std::vector<double> lx(n), ly(n), u(n);
  for (int i = 0; i < n; ++i) {
    lx[i] = std::log10(x[i]);
    ly[i] = std::log10(y[i]);
  }
  xmin = *std::min_element(lx.begin(), lx.end());
  xmax = *std::max_element(lx.begin(), lx.end());
  for (int i = 0; i < n; ++i) u[i] = (lx[i] - xmin) / (xmax - xmin);
  degree = std::min(n - 1, 3);
  const int p = degree;
  // Knot averaging (The NURBS Book eq. 9.8).
  knots.assign(n + p + 1, 0.0);
  for (int j = 1; j < n - p; ++j) {
    double s = 0;
    for (int i = j; i < j + p; ++i) s += u[i];
    knots[j + p] = s / p;
  }
  for (int i = n; i < n + p + 1; ++i) knots[i] = 1.0;
  // Collocation system A c = ly; rows 0 and n-1 are the interpolated end
  // points.  A is banded and totally nonnegative, so banded Gaussian
  // elimination without pivoting is stable (de Boor 1976).
  std::vector<int> c0(n);
  std::vector<std::array<double, 4>> rows(n);
  int kl = 0, ku = 0;
  for (int i = 0; i < n; ++i) {
    rows[i].fill(0.0);
    if (i == 0) {
      c0[i] = 0;
      rows[i][0] = 1.0;
    } else if (i == n - 1) {
      c0[i] = n - 1 - p;
      rows[i][p] = 1.0;
    } else {
      int s = span(u[i]);
      double N[8];
      basis(u[i], s, p, knots, N);
      c0[i] = s - p;
      for (int r = 0; r <= p; ++r) rows[i][r] = N[r];
    }
    kl = std::max(kl, i - c0[i]);
    ku = std::max(ku, c0[i] + p - i);
  }
  const int W = kl + ku + 1;
  std::vector<double> band((size_t)n * W, 0.0);
  auto at = [&](int i, int j) -> double& { return band[(size_t)i * W + (j - i + kl)]; };
  for (int i = 0; i < n; ++i)
    for (int r = 0; r <= p; ++r) {
      int j = c0[i] + r;
      if (j >= 0 && j < n && rows[i][r] != 0.0) at(i, j) = rows[i][r];
    }
  std::vector<double> b = ly;
  for (int k = 0; k < n; ++k) {
    double piv = at(k, k);
    if (piv == 0.0) throw std::runtime_error("singular spline collocation system");
    for (int i = k + 1; i <= std::min(n - 1, k + kl); ++i) {
      double f = at(i, k) / piv;
      if (f == 0.0) continue;
      for (int j = k; j <= std::min(n - 1, k + ku); ++j) at(i, j) -= f * at(k, j);
      b[i] -= f * b[k];
    }
  }
  ctrl.assign(n, 0.0);
  for (int k = n - 1; k >= 0; --k) {
    double s = b[k];
    for (int j = k + 1; j <= std::min(n - 1, k + ku); ++j) s -= at(k, j) * ctrl[j];
    ctrl[k] = s / at(k, k);
  }
}

double LogLogSpline::operator()(double x) const {
  double u = (std::log10(x) - xmin) / (xmax - xmin);
  return std::pow(10.0, eval_u(u));
}

std::vector<double> merge_grids(const std::vector<const std::vector<double>*>& grids) {
  std::vector<double> m;
  for (auto* g : grids) m.insert(m.end(), g->begin(), g->end());
  std::sort(m.begin(), m.end());
  m.erase(std::unique(m.begin(), m.end()), m.end());
  return m;
}

// ---------------------------------------------------------------- RITA
double Rita::x_from_y(int ie, double yu) const {
  const auto& X = x[ie];
  const auto& Y = y[ie];
  int i = (int)(std::lower_bound(Y.begin(), Y.end(), yu) - Y.begin()) - 1;
  if (i == (int)Y.size() - 1) return 1.0;
  if (i < 0) return -1.0;
  double nu = yu - Y[i];
  double d = Y[i + 1] - Y[i];
  double ai = a[ie][i], bi = b[ie][i];
  return X[i] + (1 + ai + bi) * d * nu / (d * d + ai * d * nu + bi * nu * nu) * (X[i + 1] - X[i]);
}

// probability_dist.cpp:73-88.  The bracket lower_bound - 1 is clamped to the
// first / last interval (the reference reads outside its arrays for an energy
// at or below the grid's first energy or above its last, as its comment at
// :76 warns; here the end interval's log-log line is extended, as the device
// does).
double Rita::sample(double E, double yu) const {
  int lo = (int)(std::lower_bound(energies.begin(), energies.end(), E) - energies.begin()) - 1;
  lo = std::max(0, std::min(lo, (int)energies.size() - 2));
  int up = lo + 1;
  double xl = x_from_y(lo, yu), xu = x_from_y(up, yu);
  return std::exp(std::log(xl + 1) + (std::log(xu + 1) - std::log(xl + 1)) * (std::log(E) - std::log(energies[lo])) /
                                         (std::log(energies[up]) - std::log(energies[lo]))) -
         1.0;
}

namespace {
struct RitaBuilder {
  std::function<double(double)> pdf;  // normalised PDF at fixed energy
  static constexpr double kXmin = -1.0, kXmax = 1.0;

  void cdf(const std::vector<double>& g, std::vector<double>& c) const {
    c.assign(g.size(), 0.0);
    for (size_t j = 0; j + 1 < g.size(); ++j) {
      double h = (g[j + 1] - g[j]) / 50.0;
      double s = 0;
      for (int k = 0; k < 51; ++k) {
        double f = pdf(g[j] + k * h);
        s += (k == 0 || k == 50) ? f : ((k % 2 == 0) ? 2 * f : 4 * f);
      }
      c[j + 1] = c[j] + h / 3 * s;
    }
  }
  void params(const std::vector<double>& g, const std::vector<double>& c, std::vector<double>& a,
              std::vector<double>& b) const {
    size_t m = g.size() - 1;
    a.resize(m);
    b.resize(m);
    for (size_t i = 0; i < m; ++i) {
      double sl = (c[i + 1] - c[i]) / (g[i + 1] - g[i]);
      b[i] = 1 - sl * sl * (1 / (pdf(g[i]) * pdf(g[i + 1])));
      a[i] = sl * (1 / pdf(g[i])) - b[i] - 1;
    }
  }
  static double pnum(double x, double x0, double x1, double y0, double y1, double a, double b) {
    double tau = (x - x0) / (x1 - x0);
    double eta = 0;
    if (tau != 0) {
      double c = 1 + a + b - a * tau;
      eta = c / (2 * b * tau) * (1 - std::sqrt(1 - 4 * b * tau * tau / (c * c)));
    }
    double q = 1 + a * eta + b * eta * eta;
    return q * q / ((1 + a + b) * (1 - b * eta * eta)) * (y1 - y0) / (x1 - x0);
  }
  double err(double x0, double x1, double y0, double y1, double a, double b) const {
    const int N = 51;
    double d = (x1 - x0) / N;
    double s = 0;
    for (int i = 1; i < N; ++i) {
      double xk = x0 + i * d;
      s += std::fabs(pdf(xk) - pnum(xk, x0, x1, y0, y1, a, b));
    }
    return d * (s + (std::fabs(pdf(x1) - pnum(x1, x0, x1, y0, y1, a, b)) +
                     std::fabs(pdf(x0) - pnum(x0, x0, x1, y0, y1, a, b))) /
                        2);
  }
};
}  // namespace

double MaterialTables::rayleigh_pdf(double mu, double E) const {
  double k = E / kElectronRestMass;
  double x = alpha_inv_angstrom() * k * std::sqrt(1 - mu);
  double ff = ff_i(x);
  return (1 + mu * mu) * ff * ff;
}

void build_rita(MaterialTables& m, int threads) {
  Rita& R = m.rita;
  R.energies = m.incoh.x;  // material_data.cpp:70: the incoherent-CS grid
  const size_t ne = R.energies.size();
  R.x.assign(ne, {});
  R.y.assign(ne, {});
  R.a.assign(ne, {});
  R.b.assign(ne, {});
  std::atomic<size_t> next{0};
  auto work = [&]() {
    for (;;) {
      size_t ie = next.fetch_add(1);
      if (ie >= ne) return;
      const double E = R.energies[ie];
      // probability_dist.cpp:267-287: Simpson normalisation with 1001 points
      const double h = 2.0 / 1000;
      double s = 0;
      for (int k = 0; k <= 1000; ++k) {
        double f = m.rayleigh_pdf(-1.0 + k * h, E);
        s += (k == 0 || k == 1000) ? f : ((k % 2 == 0) ? 2 * f : 4 * f);
      }
      const double integral = h / 3 * s;
      RitaBuilder rb;
      rb.pdf = [&](double mu) { return m.rayleigh_pdf(mu, E) / integral; };
      // probability_dist.cpp:138-175: bisect intervals whose error exceeds 1e-4
      std::vector<double> g(10), c, a, b;
      for (int i = 0; i < 10; ++i) g[i] = -1.0 + i * (2.0 / 9.0);
      g[9] = 1.0;
      for (;;) {
        rb.cdf(g, c);
        rb.params(g, c, a, b);
        std::vector<double> ng;
        bool ok = true;
        for (size_t i = 0; i + 1 < g.size(); ++i) {
          ng.push_back(g[i]);
          if (rb.err(g[i], g[i + 1], c[i], c[i + 1], a[i], b[i]) > 1e-4) {
            ng.push_back((g[i] + g[i + 1]) / 2);
            ok = false;
          }
        }
        ng.push_back(g.back());
        g.swap(ng);
        if (ok) break;
      }
      rb.cdf(g, c);
      rb.params(g, c, a, b);
      R.x[ie] = g;
      R.y[ie] = c;
      R.a[ie] = a;
      R.b[ie] = b;
    }
  };
  int nt = std::max(1, threads);
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
}

// ---------------------------------------------------------------- materials
MaterialTables build_material(const MaterialDef& def, const std::string& epdl_dir, bool with_rita, int threads) {
  MaterialTables m;
  m.def = def;
  std::vector<Element> els;
  for (auto& zc : def.composition) els.push_back(load_element(epdl_dir, zc.first));
  // material_properties.cpp:66-73
  std::vector<double> nd(els.size()), wf(els.size());
  for (size_t i = 0; i < els.size(); ++i) {
    wf[i] = def.composition[i].second;
    nd[i] = kAvogadro * def.density * wf[i] / els[i].atomic_weight;
  }
  // material_data.cpp:87-114
  auto weighted = [&](Table Element::*tab, bool macroscopic, bool spline) {
    std::vector<const std::vector<double>*> grids;
    for (auto& e : els) grids.push_back(&(e.*tab).x);
    Table out;
    out.x = merge_grids(grids);
    out.y.assign(out.x.size(), 0.0);
    for (size_t i = 0; i < els.size(); ++i) {
      const Table& t = els[i].*tab;
      double w = macroscopic ? nd[i] : wf[i];
      if (spline) {
        LogLogSpline f(t.x, t.y);
        for (size_t k = 0; k < out.x.size(); ++k) out.y[k] += w * f(out.x[k]);
      } else {
        LogLogLinear f(t.x, t.y);
        for (size_t k = 0; k < out.x.size(); ++k) out.y[k] += w * f(out.x[k]);
      }
    }
    if (macroscopic)
      for (auto& v : out.y) v *= 1e-24;  // barn -> cm^2
    return out;
  };
  m.incoh = weighted(&Element::incoh, true, true);
  m.coh = weighted(&Element::coh, true, true);
  m.photo = weighted(&Element::photo, true, false);
  m.incoh_i = LogLogSpline(m.incoh.x, m.incoh.y);
  m.coh_i = LogLogSpline(m.coh.x, m.coh.y);
  m.photo_i = LogLogLinear(m.photo.x, m.photo.y);
  // material_data.cpp:75-85, 125-134
  m.total.x = merge_grids({&m.incoh.x, &m.coh.x, &m.photo.x});
  m.total.y.resize(m.total.x.size());
  for (size_t k = 0; k < m.total.x.size(); ++k) {
    double E = m.total.x[k];
    m.total.y[k] = m.incoh_i(E) + m.coh_i(E) + m.photo_i(E);
  }
  m.total_i = LogLogLinear(m.total.x, m.total.y);
  m.sf = weighted(&Element::sf, false, false);
  m.ff = weighted(&Element::ff, false, false);
  m.sf_i = LogLogLinear(m.sf.x, m.sf.y);
  m.ff_i = LogLogLinear(m.ff.x, m.ff.y);
  if (with_rita) build_rita(m, threads);
  return m;
}

Table build_majorant(const std::vector<const MaterialTables*>& mats) {
  // interaction_data.cpp:42-73
  std::vector<const std::vector<double>*> grids;
  for (auto* m : mats) grids.push_back(&m->total.x);
  Table t;
  t.x = merge_grids(grids);
  t.y.resize(t.x.size());
  for (size_t k = 0; k < t.x.size(); ++k) {
    double v = -1;
    for (auto* m : mats) v = std::max(v, m->total_i(t.x[k]));
    t.y[k] = v;
  }
  return t;
}

// ---------------------------------------------------------------- device packing
static void split(double a, float& hi, float& lo) {
  hi = (float)a;
  lo = (float)(a - (double)hi);
}

static void build_hash(Packed& p, mdx_curve& c, double lo, double hi) {
  // the device reads the node pair (k, k + 1) of every lookup
  if (c.n < 2) throw std::runtime_error("a packed curve needs at least two nodes");
  const float* xs = p.xs.data() + c.off;
  lo = std::max(lo, (double)xs[0]);
  hi = std::min(hi, (double)xs[c.n - 1]);
  if (!(hi > lo)) {
    lo = xs[0];
    hi = xs[c.n - 1] > xs[0] ? xs[c.n - 1] : xs[0] * 2;
  }
  int in_range = 0;
  for (int k = 0; k < c.n; ++k)
    if (xs[k] >= lo && xs[k] <= hi) ++in_range;
  int nh = 256;
  while (nh < 2 * in_range && nh < 16384) nh *= 2;
  c.nh = nh;
  c.lnx0 = (float)std::log(lo);
  c.inv_dlnx = (float)(nh / (std::log(hi) - std::log(lo)));
  c.hoff = (int)p.bins.size();
  int k = 0;
  for (int b = 0; b < nh; ++b) {
    double xb = std::exp(std::log(lo) + b / (double)c.inv_dlnx);
    while (k + 1 < c.n && xs[k + 1] <= xb) ++k;
    // the bin's record: node k (clamped to the last span) with the next
    // node's abscissa and k itself in its padding, so that a lookup whose
    // X lies in [x_k, x_{k+1}) reads one record
    const int kc = k > c.n - 2 ? c.n - 2 : k;
    mdx_node r = p.nodes[c.off + kc];
    r.pad0 = xs[kc + 1];
    std::memcpy(&r.pad1, &kc, sizeof(float));
    p.bins.push_back(r);
  }
}

mdx_curve pack_linear(Packed& p, const Table& t, double hash_lo, double hash_hi, int) {
  mdx_curve c{};
  c.kind = 0;
  c.off = (int)p.xs.size();
  const size_t n = t.x.size();
  for (size_t k = 0; k < n; ++k) {
    mdx_node r{};
    // Abscissae are rounded up, so that for every FP32 input X the test
    // X >= xs[k] is the same as X >= x_k in FP64 and the device brackets X
    // exactly as the FP64 upper_bound does (interpolators.cpp:56-59); this
    // matters on the near-vertical ramps across absorption edges.
    float xk = (float)t.x[k];
    if ((double)xk < t.x[k]) xk = std::nextafter(xk, INFINITY);
    double slope = 0, a = t.y[k];
    if (k + 1 < n) {
      slope = std::log(t.y[k + 1] / t.y[k]) / std::log(t.x[k + 1] / t.x[k]);
      // value of the FP64 segment at the rounded node (exact interpolant)
      a = t.y[k] * std::pow((double)xk / t.x[k], slope);
    }
    split(a, r.a_hi, r.a_lo);
    r.c1 = (float)slope;
    r.x = xk;
    p.xs.push_back(xk);
    p.nodes.push_back(r);
  }
  c.n = (int)n;
  build_hash(p, c, hash_lo, hash_hi);
  return c;
}

mdx_curve pack_spline(Packed& p, const LogLogSpline& s, double hash_lo, double hash_hi, int) {
  // One node per non-empty knot span: log10 f is a cubic in u on a span and u
  // is linear in ln E, so log f is a cubic in t = ln(E / E_span) there.
  mdx_curve c{};
  c.kind = 1;
  c.off = (int)p.xs.size();
  const double W = s.xmax - s.xmin;
  const double ln10 = std::log(10.0);
  const int nk = (int)s.knots.size();
  std::vector<int> spans;
  for (int j = s.degree; j < nk - s.degree - 1; ++j)
    if (s.knots[j + 1] > s.knots[j]) spans.push_back(j);
  auto eval_span = [&](int j, double u) {
    double N[8];
    basis(u, j, s.degree, s.knots, N);
    double v = 0;
    for (int r = 0; r <= s.degree; ++r) v += N[r] * s.ctrl[j - s.degree + r];
    return v;
  };
  for (size_t q = 0; q < spans.size(); ++q) {
    int j = spans[q];
    float Ej = (float)std::pow(10.0, s.xmin + s.knots[j] * W);
    double Ej1 = std::pow(10.0, s.xmin + s.knots[j + 1] * W);
    double T = std::log(Ej1 / (double)Ej);
    double uj = (std::log10((double)Ej) - s.xmin) / W;
    double tt[4], vv[4];
    for (int m = 0; m < 4; ++m) {
      tt[m] = T * m / 3.0;
      vv[m] = eval_span(j, uj + tt[m] / (ln10 * W));
    }
    // exact cubic through 4 points: Newton divided differences -> monomials
    double d0 = vv[0];
    double d1 = (vv[1] - vv[0]) / (tt[1] - tt[0]);
    double d12 = (vv[2] - vv[1]) / (tt[2] - tt[1]);
    double d23 = (vv[3] - vv[2]) / (tt[3] - tt[2]);
    double d2 = (d12 - d1) / (tt[2] - tt[0]);
    double d2b = (d23 - d12) / (tt[3] - tt[1]);
    double d3 = (d2b - d2) / (tt[3] - tt[0]);
    // P(t) = d0 + d1 t + d2 t (t - t1) + d3 t (t - t1)(t - t2)
    double q1 = d1 - d2 * tt[1] + d3 * tt[1] * tt[2];
    double q2 = d2 - d3 * (tt[1] + tt[2]);
    double q3 = d3;
    mdx_node r{};
    split(std::pow(10.0, d0), r.a_hi, r.a_lo);
    r.c1 = (float)(ln10 * q1);
    r.c2 = (float)(ln10 * q2);
    r.c3 = (float)(ln10 * q3);
    r.x = Ej;
    p.xs.push_back(Ej);
    p.nodes.push_back(r);
  }
  // closing node (end of the last span); its record is never evaluated
  mdx_node r{};
  float Eend = (float)std::pow(10.0, s.xmax);
  r.x = Eend;
  split(std::pow(10.0, s.eval_u(1.0)), r.a_hi, r.a_lo);
  p.xs.push_back(Eend);
  p.nodes.push_back(r);
  c.n = (int)spans.size() + 1;
  build_hash(p, c, hash_lo, hash_hi);
  return c;
}

void pack_all(Packed& p, const std::vector<const MaterialTables*>& mats, const Table& majorant, double e_lo,
              double e_hi) {
  p = Packed();
  std::memset(p.id_to_slot, 255, sizeof p.id_to_slot);
  const double k_hi = e_hi / kElectronRestMass;
  const double x_hi = alpha_inv_angstrom() * k_hi * std::sqrt(2.0) * 1.01;
  p.majorant = pack_linear(p, majorant, e_lo, e_hi, 0);
  for (size_t s = 0; s < mats.size(); ++s) {
    const MaterialTables& m = *mats[s];
    mdx_material_dev d{};
    d.material_id = m.def.id;
    d.total = pack_linear(p, m.total, e_lo, e_hi, 0);
    d.coh = pack_spline(p, m.coh_i, e_lo, e_hi, 0);
    d.incoh = pack_spline(p, m.incoh_i, e_lo, e_hi, 0);
    d.photo = pack_linear(p, m.photo, e_lo, e_hi, 0);
    d.sf = pack_linear(p, m.sf, 1e-4, x_hi, 0);
    const Rita& R = m.rita;
    d.rita_ne = (int)R.energies.size();
    d.rita_eoff = (int)p.rita_E.size();
    d.rita_roff = (int)p.rita_row_off.size();
    for (size_t i = 0; i < R.energies.size(); ++i) {
      // energies rounded toward -inf: for an FP32 energy E, E > stored <=>
      // E > E_i, so the device brackets as the FP64 lower_bound - 1 does;
      // ln(stored / E_i) corrects the interpolation weight for the rounding
      const double e = R.energies[i];
      float e32 = (float)e;
      if ((double)e32 > e) e32 = std::nextafter(e32, 0.0f);
      p.rita_E.push_back(e32);
      p.rita_escale.push_back(i + 1 < R.energies.size() ? (float)(1.0 / std::log(R.energies[i + 1] / e)) : 0.0f);
      p.rita_escale.push_back((float)std::log((double)e32 / e));
      p.rita_row_off.push_back((int)p.rita_rows.size());
      if (R.x[i].size() > 255) throw std::runtime_error("a RITA row has more than 255 points");
      const size_t m = R.x[i].size();
      for (size_t r = 0; r < m; ++r) {
        mdx_rita_row q{};
        const double w = 1.0 + R.x[i][r], y = R.y[i][r];
        q.w_hi = (float)w;
        q.w_lo = (float)(w - (double)q.w_hi);
        q.y_hi = (float)y;
        q.y_lo = (float)(y - (double)q.y_hi);
        if (r + 1 < m) {
          q.a = r < R.a[i].size() ? (float)R.a[i][r] : 0.0f;
          q.b = r < R.b[i].size() ? (float)R.b[i][r] : 0.0f;
          q.dw = (float)((1.0 + R.x[i][r + 1]) - w);
          q.d = (float)(R.y[i][r + 1] - y);
        }
        p.rita_rows.push_back(q);
      }
      // guide: entry j = number of CDF values (FP64) below j / G; the
      // device scales y by G (a power of two, so exactly) to find j, and the
      // node before the entry has y_i < y
      for (int j = 0; j < MDX_RITA_GUIDE; ++j) {
        const double t = (double)j / (double)MDX_RITA_GUIDE;
        int cnt = 0;
        while (cnt < (int)m && R.y[i][cnt] < t) ++cnt;
        p.rita_guide.push_back((uint8_t)cnt);
      }
    }
    p.rita_row_off.push_back((int)p.rita_rows.size());
    // energy hash over [e_lo, e_hi]: entry b = largest index whose (stored)
    // energy lies below the bin's start, -1 if none
    {
      const int nh = 512;
      const double l0 = std::log(e_lo), l1 = std::log(e_hi);
      d.rita_hoff = (int)p.rita_ehash.size();
      d.rita_nh = nh;
      d.rita_lnE0 = (float)l0;
      d.rita_hinv = (float)(nh / (l1 - l0));
      int k = -1;
      for (int b = 0; b < nh; ++b) {
        const double eb = std::exp(l0 + b / (double)d.rita_hinv);
        while (k + 1 < d.rita_ne && (double)p.rita_E[d.rita_eoff + k + 1] < eb) ++k;
        p.rita_ehash.push_back(k);
      }
    }
    p.mats.push_back(d);
    p.id_to_slot[m.def.id & 255] = (uint8_t)s;
  }
}

}  // namespace mdx
