// Host FP64 physics-table builder: EPDL ENDF-6 reader, material registry,
// per-material interaction tables, majorant and RITA tables, and the packing
// of all of them into the FP32 device layout of mdx_types.h.
#pragma once
#include <map>
#include <string>
#include <vector>

#include "mdx_types.h"

namespace mdx {

// include/MIDSX/Core/constants.h:13-23
constexpr double kPi = 3.14159265358979323846;
constexpr double kElectronRestMass = 0.51099895000e6;  // eV
constexpr double kSpeedOfLight = 299792458e2;          // cm/s
constexpr double kPlanck = 4.135667662e-15;            // eV s
constexpr double kAvogadro = 6.022140857e23;
constexpr double kNeutronMassU = 1.00866491595;        // ENDF AWR unit
double alpha_inv_angstrom();                           // ALPHA, constants.h:23

struct Table {
  std::vector<double> x, y;
};

// EPDL element: atomic weight and the five tables the transport path reads.
struct Element {
  int z = 0;
  double atomic_weight = 0;
  Table coh, incoh, photo, ff, sf;
};
// Reads <dir>/NNN.endf.gz (trimmed) or <dir>/NNN.txt (full EPDL file).
Element load_element(const std::string& dir, int z);

struct MaterialDef {
  int id = 0;
  std::string name;
  double density = 0;
  std::vector<std::pair<int, double>> composition;  // (Z, mass fraction)
};
std::map<int, MaterialDef> load_registry(const std::string& csv_path);

// interpolators.cpp:37-76
struct LogLogLinear {
  std::vector<double> lx, ly;
  LogLogLinear() = default;
  LogLogLinear(const std::vector<double>& x, const std::vector<double>& y);
  double operator()(double x) const;
};

// interpolators.cpp:3-31 with the fit of Eigen SplineFitting::Interpolate
// (knot averaging, collocation at the parameters) solved as a banded system.
struct LogLogSpline {
  double xmin = 0, xmax = 0;
  int degree = 0;
  std::vector<double> knots, ctrl;
  LogLogSpline() = default;
  LogLogSpline(const std::vector<double>& x, const std::vector<double>& y);
  double eval_u(double u) const;  // spline value (log10 y) at parameter u
  double operator()(double x) const;
  int span(double u) const;
};

struct Rita {
  std::vector<double> energies;
  // per energy: x_i, y_i (CDF), a_i, b_i; a and b have one entry less
  std::vector<std::vector<double>> x, y, a, b;
  double sample(double E, double yu) const;  // probability_dist.cpp:73-88
  double x_from_y(int ie, double yu) const;  // probability_dist.cpp:90-119
};

struct MaterialTables {
  MaterialDef def;
  Table coh, incoh, photo, total, sf, ff;
  LogLogSpline coh_i, incoh_i;
  LogLogLinear photo_i, total_i, sf_i, ff_i;
  Rita rita;
  double rayleigh_pdf(double mu, double E) const;  // material_data.cpp:64-68
};

MaterialTables build_material(const MaterialDef& def, const std::string& epdl_dir, bool with_rita,
                              int threads);
Table build_majorant(const std::vector<const MaterialTables*>& mats);
void build_rita(MaterialTables& m, int threads);

std::vector<double> merge_grids(const std::vector<const std::vector<double>*>& grids);

// ---------------------------------------------------------------- device packing
struct Packed {
  std::vector<float> xs;
  std::vector<mdx_node> nodes;
  std::vector<mdx_node> bins;  // per hash bin: node k of the bin with x_{k+1} and k (see build_hash)
  std::vector<float> rita_E, rita_escale;
  std::vector<int32_t> rita_row_off;
  std::vector<mdx_rita_row> rita_rows;
  std::vector<int32_t> rita_ehash;
  std::vector<uint8_t> rita_guide;
  mdx_curve majorant{};
  std::vector<mdx_material_dev> mats;
  uint8_t id_to_slot[256];
};
mdx_curve pack_linear(Packed& p, const Table& t, double hash_lo, double hash_hi, int nh);
mdx_curve pack_spline(Packed& p, const LogLogSpline& s, double hash_lo, double hash_hi, int nh);
void pack_all(Packed& p, const std::vector<const MaterialTables*>& mats, const Table& majorant,
              double e_lo, double e_hi);

}  // namespace mdx
