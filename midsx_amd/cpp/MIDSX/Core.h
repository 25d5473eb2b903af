// C++ host API of the MI355X photon-transport library, with the class names,
// constructors and result accessors of the reference's include/MIDSX/Core.h
// (include/MIDSX/Core/*.h), implemented over the C-ABI of libmidsx_amd.so
// (include/midsx_amd.h).  A driver written against the reference
// (e.g. cpp_simulations/half_value_layer/hvl_30keV.cpp:9-80) compiles against
// this header with its Eigen vectors, or with braced {x, y, z} triples:
//
//   ComputationalDomain comp_domain("hvl.json");
//   InteractionData interaction_data = comp_domain.getInteractionData();
//   PhysicsEngine physics_engine(comp_domain, interaction_data);
//   PhotonSource source(std::make_unique<MonoenergeticSpectrum>(30E3), ...);
//   runSimulation(source, physics_engine, surface_init, volume_init, N);  // or ..., N, run_time)
//   auto results = physics_engine.getSurfaceQuantityContainers();
//
// Differences from the reference, by design:
//  * the transport runs on the GPU through one context per PhysicsEngine;
//    the tally factories are called once (not once per OpenMP thread);
//  * VectorValue keeps (count, sum, sum of squares) instead of every value, so
//    getVector() is not available; getSum / getVariance / getSumSTD / getMean /
//    getMeanSTD / getCount return what the reference computes from its list
//    (the variance from the moments, clamped at 0);
//  * vectors are Vec3 (or anything with x(), y(), z(), Eigen's included), and
//    SourceHelpers::readCSV and DerivedQuantity take std::vector where the
//    reference takes Eigen matrices;
//  * only the closed set of tally, spectrum, directionality and geometry
//    classes below is accepted (user subclasses throw std::invalid_argument).
#pragma once

#include <chrono>
#include <cmath>
#include <fstream>
#include <functional>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../../include/midsx_amd.h"

namespace midsx_amd_detail {

inline void check(int status) {
  if (status == MDX_OK) return;
  std::string msg = mdx_last_error();
  switch (status) {
    case MDX_ERR_INVALID: throw std::invalid_argument(msg);
    case MDX_ERR_RANGE: throw std::out_of_range(msg);
    default: throw std::runtime_error(msg);
  }
}

// Owning handle of an mdx_ctx.
struct Ctx {
  mdx_ctx* p = nullptr;
  explicit Ctx(int device) { check(mdx_create(device, &p)); }
  ~Ctx() {
    if (p) mdx_destroy(p);
  }
  Ctx(const Ctx&) = delete;
  Ctx& operator=(const Ctx&) = delete;
};

}  // namespace midsx_amd_detail

// 3-vector accepted wherever the reference takes an Eigen::Vector3d; converts
// from any type with x(), y(), z() accessors (Eigen::Vector3d included).
struct Vec3 {
  double v[3] = {0, 0, 0};
  Vec3() = default;
  Vec3(double x, double y, double z) : v{x, y, z} {}
  template <class V, class = decltype(std::declval<const V&>().x())>
  Vec3(const V& o) : v{(double)o.x(), (double)o.y(), (double)o.z()} {}  // NOLINT: implicit on purpose
  double x() const { return v[0]; }
  double y() const { return v[1]; }
  double z() const { return v[2]; }
  double operator[](int i) const { return v[i]; }
  // and back to a caller's 3-vector type constructible from three doubles
  // (where the reference returns an Eigen::Vector3d)
  template <class V, class = decltype(V(0.0, 0.0, 0.0)), class = decltype(std::declval<const V&>().x())>
  operator V() const {  // NOLINT: implicit on purpose
    return V(v[0], v[1], v[2]);
  }
};

// ------------------------------------------------------------------ quantities
// quantity.h:14-51
class VectorValue {
 public:
  double getSum() const { return sum_; }
  double getCount() const { return (double)n_; }
  double getCountSTD() const { return std::sqrt((double)n_); }
  double getMean() const { return sum_ / (double)n_; }
  // Sample variance (quantity.cpp:80-94) from the moments: sum of squares
  // less sum * mean over n - 1, clamped at 0 (the two-pass form of the
  // reference never goes below 0; from the moments, rounding can for
  // near-constant values).  No values: 0 (the reference divides 0 by
  // size() - 1 in size_t, a huge number); one value: 0 / 0, NaN, as there.
  double getVariance() const {
    if (n_ == 0) return 0.0;
    if (n_ == 1) return std::nan("");
    const double mean = sum_ / (double)n_;
    const double v = (sum2_ - mean * sum_) / (double)(n_ - 1);
    return v > 0.0 ? v : 0.0;
  }
  double getSumSTD() const { return std::sqrt((double)n_ * getVariance()); }
  double getMeanSTD() const { return std::sqrt(getVariance() / (double)n_); }
  // quantity.cpp:15-21 (the values are folded into the moments)
  void addValue(double v) { addMoments(1.0, v, v * v); }
  void addValues(const std::vector<double>& vs) {
    for (double v : vs) addValue(v);
  }
  void addMoments(double n, double s, double s2) {
    n_ += (long long)std::llround(n);
    sum_ += s;
    sum2_ += s2;
  }
  VectorValue operator+(const VectorValue& o) const {
    VectorValue r = *this;
    r.addMoments((double)o.n_, o.sum_, o.sum2_);
    return r;
  }

 private:
  long long n_ = 0;
  double sum_ = 0, sum2_ = 0;
};

class CountValue {
 public:
  int getCount() const { return count_; }
  double getCountSTD() const { return std::sqrt((double)count_); }
  void addCount() { ++count_; }  // quantity.cpp:104-106
  void addCounts(int c) { count_ += c; }
  CountValue operator+(const CountValue& o) const {
    CountValue r;
    r.count_ = count_ + o.count_;
    return r;
  }
  int count_ = 0;
};

enum class VectorVolumeQuantityType { EnergyDeposition, IncidentEnergy };
enum class CountVolumeQuantityType { NumberOfPhotons, NumberOfInteractions };
enum class VectorSurfaceQuantityType { IncidentEnergy, EntranceCosine };
enum class CountSurfaceQuantityType { NumberOfPhotons };

// The quantity names drivers print (volume_quantity.h:18-33,
// surface_quantity.h:19-33): each enumerator's own name.
namespace VectorVolumeQuantityHelper {
inline std::string toString(VectorVolumeQuantityType type) {
  return type == VectorVolumeQuantityType::EnergyDeposition ? "EnergyDeposition" : "IncidentEnergy";
}
}  // namespace VectorVolumeQuantityHelper
namespace CountVolumeQuantityHelper {
inline std::string toString(CountVolumeQuantityType type) {
  return type == CountVolumeQuantityType::NumberOfPhotons ? "NumberOfPhotons" : "NumberOfInteractions";
}
}  // namespace CountVolumeQuantityHelper
namespace VectorSurfaceQuantityHelper {
inline std::string toString(VectorSurfaceQuantityType type) {
  return type == VectorSurfaceQuantityType::IncidentEnergy ? "IncidentEnergy" : "EntranceCosine";
}
}  // namespace VectorSurfaceQuantityHelper
namespace CountSurfaceQuantityHelper {
inline std::string toString(CountSurfaceQuantityType) { return "NumberOfPhotons"; }
}  // namespace CountSurfaceQuantityHelper

// Buckets of *_quantity.cpp measure(): primary, single incoherent, single
// coherent, multiple (the order of the device result layout).
template <class V>
class BucketedQuantity {
 public:
  V& getPrimaryValues() { return b_[0]; }
  V& getSingleIncoherentScatterValues() { return b_[1]; }
  V& getSingleCoherentScatterValues() { return b_[2]; }
  V& getMultipleScatterValues() { return b_[3]; }
  V getTotalValues() const { return b_[0] + b_[1] + b_[2] + b_[3]; }
  V b_[4];
};

class VectorVolumeQuantity : public BucketedQuantity<VectorValue> {
 public:
  explicit VectorVolumeQuantity(VectorVolumeQuantityType t) : type_(t) {}
  VectorVolumeQuantityType getType() const { return type_; }

 private:
  VectorVolumeQuantityType type_;
};
class CountVolumeQuantity : public BucketedQuantity<CountValue> {
 public:
  explicit CountVolumeQuantity(CountVolumeQuantityType t) : type_(t) {}
  CountVolumeQuantityType getType() const { return type_; }

 private:
  CountVolumeQuantityType type_;
};
class VectorSurfaceQuantity : public BucketedQuantity<VectorValue> {
 public:
  explicit VectorSurfaceQuantity(VectorSurfaceQuantityType t) : type_(t) {}
  VectorSurfaceQuantityType getType() const { return type_; }

 private:
  VectorSurfaceQuantityType type_;
};
class CountSurfaceQuantity : public BucketedQuantity<CountValue> {
 public:
  explicit CountSurfaceQuantity(CountSurfaceQuantityType t) : type_(t) {}
  CountSurfaceQuantityType getType() const { return type_; }

 private:
  CountSurfaceQuantityType type_;
};

// volume_quantity_container.h
class VolumeQuantityContainer {
 public:
  void addVectorQuantity(const VectorVolumeQuantity& q) { vq_.emplace(q.getType(), q); }
  void addCountQuantity(const CountVolumeQuantity& q) { cq_.emplace(q.getType(), q); }
  std::unordered_map<VectorVolumeQuantityType, VectorVolumeQuantity>& getVectorQuantities() { return vq_; }
  std::unordered_map<CountVolumeQuantityType, CountVolumeQuantity>& getCountQuantities() { return cq_; }
  void setVolume(double v) { volume_ = v; }
  double getVolume() const { return volume_; }
  // Fills the requested quantities from a device result block
  // [quantity][bucket][n, sum, sum^2].
  void load(const double* acc) {
    for (auto& kv : vq_) {
      int q = kv.first == VectorVolumeQuantityType::EnergyDeposition ? 0 : 1;
      for (int b = 0; b < 4; ++b) {
        const double* f = acc + (q * 4 + b) * 3;
        kv.second.b_[b] = VectorValue();
        kv.second.b_[b].addMoments(f[0], f[1], f[2]);
      }
    }
    for (auto& kv : cq_) {
      int q = kv.first == CountVolumeQuantityType::NumberOfPhotons ? 2 : 3;
      for (int b = 0; b < 4; ++b) {
        kv.second.b_[b] = CountValue();
        kv.second.b_[b].addCounts((int)std::llround(acc[(q * 4 + b) * 3]));
      }
    }
  }

 private:
  std::unordered_map<VectorVolumeQuantityType, VectorVolumeQuantity> vq_;
  std::unordered_map<CountVolumeQuantityType, CountVolumeQuantity> cq_;
  double volume_ = 0;
};

class SurfaceQuantityContainer {
 public:
  void addVectorQuantity(const VectorSurfaceQuantity& q) { vq_.emplace(q.getType(), q); }
  void addCountQuantity(const CountSurfaceQuantity& q) { cq_.emplace(q.getType(), q); }
  std::unordered_map<VectorSurfaceQuantityType, VectorSurfaceQuantity>& getVectorQuantities() { return vq_; }
  std::unordered_map<CountSurfaceQuantityType, CountSurfaceQuantity>& getCountQuantities() { return cq_; }
  void setArea(double a) { area_ = a; }
  double getArea() const { return area_; }
  void load(const double* acc) {
    for (auto& kv : vq_) {
      int q = kv.first == VectorSurfaceQuantityType::IncidentEnergy ? 0 : 1;
      for (int b = 0; b < 4; ++b) {
        const double* f = acc + (q * 4 + b) * 3;
        kv.second.b_[b] = VectorValue();
        kv.second.b_[b].addMoments(f[0], f[1], f[2]);
      }
    }
    for (auto& kv : cq_)
      for (int b = 0; b < 4; ++b) {
        kv.second.b_[b] = CountValue();
        kv.second.b_[b].addCounts((int)std::llround(acc[(2 * 4 + b) * 3]));
      }
  }
  std::vector<double> spectrum;  // IncidentEnergy histogram [bucket][bin] when requested
  double spectrum_lo = 0, spectrum_width = 0;  // its first bin's lower edge and bin width (eV)
  int spectrum_bins = 0;

 private:
  std::unordered_map<VectorSurfaceQuantityType, VectorSurfaceQuantity> vq_;
  std::unordered_map<CountSurfaceQuantityType, CountSurfaceQuantity> cq_;
  double area_ = 0;
};

struct VolumeQuantityContainerFactory {  // volume_quantity_container.cpp:66-85
  static VolumeQuantityContainer AllQuantities() {
    VolumeQuantityContainer c;
    c.addVectorQuantity(VectorVolumeQuantity(VectorVolumeQuantityType::EnergyDeposition));
    c.addVectorQuantity(VectorVolumeQuantity(VectorVolumeQuantityType::IncidentEnergy));
    c.addCountQuantity(CountVolumeQuantity(CountVolumeQuantityType::NumberOfPhotons));
    c.addCountQuantity(CountVolumeQuantity(CountVolumeQuantityType::NumberOfInteractions));
    return c;
  }
  static VolumeQuantityContainer EnergyDeposition() {
    VolumeQuantityContainer c;
    c.addVectorQuantity(VectorVolumeQuantity(VectorVolumeQuantityType::EnergyDeposition));
    return c;
  }
};

struct SurfaceQuantityContainerFactory {  // surface_quantity_container.cpp:66-82
  static SurfaceQuantityContainer AllQuantities() {
    SurfaceQuantityContainer c;
    c.addVectorQuantity(VectorSurfaceQuantity(VectorSurfaceQuantityType::IncidentEnergy));
    c.addVectorQuantity(VectorSurfaceQuantity(VectorSurfaceQuantityType::EntranceCosine));
    c.addCountQuantity(CountSurfaceQuantity(CountSurfaceQuantityType::NumberOfPhotons));
    return c;
  }
};

// ------------------------------------------------------------------ tallies
// surface_tally.h:12-100, volume_tally.h:25-71
class SurfaceTally {
 public:
  virtual ~SurfaceTally() = default;
  SurfaceQuantityContainer& getSurfaceQuantityContainer() { return quantities_; }
  // IncidentEnergy histogram of this surface (device extension; off = 0 bins)
  void setSpectrum(int bins, double lo_eV, double width_eV) {
    desc_.hist_bins = bins;
    desc_.hist_lo = lo_eV;
    desc_.hist_width = width_eV;
  }
  const mdx_tally& descriptor() const { return desc_; }

 protected:
  mdx_tally desc_{};
  SurfaceQuantityContainer quantities_;
};

class DiscSurfaceTally : public SurfaceTally {
 public:
  DiscSurfaceTally(Vec3 center, Vec3 norm, double radius, const SurfaceQuantityContainer& quantities) {
    quantities_ = quantities;
    quantities_.setArea(3.14159265358979323846 * radius * radius);
    desc_.kind = MDX_TALLY_DISC;
    for (int k = 0; k < 3; ++k) {
      desc_.p0[k] = center[k];
      desc_.normal[k] = norm[k];
    }
    desc_.radius = radius;
  }
};

class RectangularSurfaceTally : public SurfaceTally {
 public:
  RectangularSurfaceTally(Vec3 corner, Vec3 edge1, Vec3 edge2, const SurfaceQuantityContainer& quantities) {
    quantities_ = quantities;
    if (std::abs(edge1[0] * edge2[0] + edge1[1] * edge2[1] + edge1[2] * edge2[2]) >= 1e-9)
      throw std::invalid_argument("Edges are not orthogonal");
    double cx = edge1[1] * edge2[2] - edge1[2] * edge2[1], cy = edge1[2] * edge2[0] - edge1[0] * edge2[2],
           cz = edge1[0] * edge2[1] - edge1[1] * edge2[0];
    quantities_.setArea(std::sqrt(cx * cx + cy * cy + cz * cz));
    desc_.kind = MDX_TALLY_RECT;
    for (int k = 0; k < 3; ++k) {
      desc_.p0[k] = corner[k];
      desc_.p1[k] = edge1[k];
      desc_.p2[k] = edge2[k];
    }
  }
};

// nx x ny pixels on a rectangle; per-pixel incident-energy sums
// (no reference class: equivalent to nx * ny RectangularSurfaceTally pixels).
class ImageSurfaceTally : public SurfaceTally {
 public:
  ImageSurfaceTally(Vec3 corner, Vec3 edge1, Vec3 edge2, int nx, int ny, const SurfaceQuantityContainer& quantities) {
    quantities_ = quantities;
    desc_.kind = MDX_TALLY_IMAGE;
    for (int k = 0; k < 3; ++k) {
      desc_.p0[k] = corner[k];
      desc_.p1[k] = edge1[k];
      desc_.p2[k] = edge2[k];
    }
    desc_.nx = nx;
    desc_.ny = ny;
  }
  std::vector<double> image;  // [ny][nx][sum E, sum E^2, count]
};

class VolumeTally {
 public:
  virtual ~VolumeTally() = default;
  VolumeQuantityContainer& getVolumeQuantityContainer() { return quantities_; }
  const mdx_tally& descriptor() const { return desc_; }

 protected:
  mdx_tally desc_{};
  VolumeQuantityContainer quantities_;
};

class AACuboidVolumeTally : public VolumeTally {
 public:
  AACuboidVolumeTally(Vec3 min_corner, Vec3 max_corner, const VolumeQuantityContainer& quantities) {
    quantities_ = quantities;
    desc_.kind = MDX_TALLY_AACUBOID;
    for (int k = 0; k < 3; ++k) {
      desc_.p0[k] = min_corner[k];
      desc_.p1[k] = max_corner[k];
    }
  }
};

// ------------------------------------------------------------------ source
// source.h:14-264
class EnergySpectrum {
 public:
  virtual ~EnergySpectrum() = default;
  virtual void describe(mdx_source& s, std::vector<double>& e, std::vector<double>& w) const = 0;
};
class MonoenergeticSpectrum : public EnergySpectrum {
 public:
  explicit MonoenergeticSpectrum(double energy) : energy_(energy) {}
  void describe(mdx_source& s, std::vector<double>&, std::vector<double>&) const override {
    s.spectrum = MDX_SPEC_MONO;
    s.energy = energy_;
  }

 private:
  double energy_;
};
// probabilities: rows of (energy eV, weight), as the reference's Nx2 matrix
class PolyenergeticSpectrum : public EnergySpectrum {
 public:
  explicit PolyenergeticSpectrum(std::vector<std::pair<double, double>> probabilities) : p_(std::move(probabilities)) {}
  // source.cpp:42-50 takes an Eigen::MatrixXd of (energy, probability) rows:
  // any matrix with rows() and operator()(row, col) is read the same way
  template <class M, class = decltype(std::declval<const M&>().rows()),
            class = decltype(std::declval<const M&>()(0, 0))>
  explicit PolyenergeticSpectrum(const M& probabilities) {
    for (long i = 0; i < (long)probabilities.rows(); ++i)
      p_.emplace_back((double)probabilities(i, 0), (double)probabilities(i, 1));
  }
  void describe(mdx_source& s, std::vector<double>& e, std::vector<double>& w) const override {
    e.clear();
    w.clear();
    for (auto& r : p_) {
      e.push_back(r.first);
      w.push_back(r.second);
    }
    s.spectrum = MDX_SPEC_POLY;
    s.n_bins = (int)e.size();
    s.bin_energy = e.data();
    s.bin_weight = w.data();
  }
  double getExpectationValue() const {
    double sw = 0, se = 0;
    for (auto& r : p_) {
      sw += r.second;
      se += r.first * r.second;
    }
    return se / sw;
  }

 private:
  std::vector<std::pair<double, double>> p_;
};

class Directionality {
 public:
  virtual ~Directionality() = default;
  virtual void describe(mdx_source& s) const = 0;
};
class IsotropicDirectionality : public Directionality {
 public:
  void describe(mdx_source& s) const override { s.direction = MDX_DIR_ISOTROPIC; }
};
class BeamDirectionality : public Directionality {
 public:
  explicit BeamDirectionality(Vec3 pass_through_point) : p_(pass_through_point) {}
  void describe(mdx_source& s) const override {
    s.direction = MDX_DIR_BEAM;
    for (int k = 0; k < 3; ++k) s.a[k] = p_[k];
  }

 private:
  Vec3 p_;
};
class RectangularIsotropicDirectionality : public Directionality {
 public:
  RectangularIsotropicDirectionality(Vec3 corner, Vec3 edge1, Vec3 edge2) : c_(corner), e1_(edge1), e2_(edge2) {
    if (std::abs(edge1[0] * edge2[0] + edge1[1] * edge2[1] + edge1[2] * edge2[2]) >= 1e-9)
      throw std::invalid_argument("Edges are not orthogonal");
  }
  void describe(mdx_source& s) const override {
    s.direction = MDX_DIR_RECT_ISOTROPIC;
    for (int k = 0; k < 3; ++k) {
      s.a[k] = c_[k];
      s.b[k] = e1_[k];
      s.c[k] = e2_[k];
    }
  }

 private:
  Vec3 c_, e1_, e2_;
};
class DiscIsotropicDirectionality : public Directionality {
 public:
  DiscIsotropicDirectionality(Vec3 center, Vec3 normal, double radius) : c_(center), n_(normal), r_(radius) {}
  void describe(mdx_source& s) const override {
    s.direction = MDX_DIR_DISC_ISOTROPIC;
    for (int k = 0; k < 3; ++k) {
      s.a[k] = c_[k];
      s.b[k] = n_[k];
    }
    s.radius = r_;
  }

 private:
  Vec3 c_, n_;
  double r_;
};

class SourceGeometry {
 public:
  explicit SourceGeometry(Vec3 position) : position_(position) {}
  virtual ~SourceGeometry() = default;
  Vec3 position_;
};
class PointGeometry : public SourceGeometry {
 public:
  explicit PointGeometry(const Vec3& position) : SourceGeometry(position) {}
};

class PhotonSource {
 public:
  PhotonSource(std::unique_ptr<EnergySpectrum> spectrum, std::unique_ptr<Directionality> directionality,
               std::unique_ptr<SourceGeometry> geometry)
      : spectrum_(std::move(spectrum)), dir_(std::move(directionality)), geom_(std::move(geometry)) {}
  void upload(mdx_ctx* c) const {
    mdx_source s{};
    std::vector<double> e, w;
    for (int k = 0; k < 3; ++k) s.position[k] = geom_->position_[k];
    dir_->describe(s);
    spectrum_->describe(s, e, w);
    midsx_amd_detail::check(mdx_set_source(c, &s));
  }

 private:
  std::unique_ptr<EnergySpectrum> spectrum_;
  std::unique_ptr<Directionality> dir_;
  std::unique_ptr<SourceGeometry> geom_;
};

namespace SourceHelpers {
// The rows of a two-column CSV file: a vector of (first, second) pairs that
// also converts to any matrix type constructible from (rows, cols) with
// operator()(row, col), e.g. the Eigen::MatrixXd the reference returns.
struct CSVRows : std::vector<std::pair<double, double>> {
  template <class M, class = decltype(std::declval<M&>()(0, 0) = 0.0)>
  operator M() const {
    M m((long)size(), 2);
    for (size_t i = 0; i < size(); ++i) {
      m((long)i, 0) = (*this)[i].first;
      m((long)i, 1) = (*this)[i].second;
    }
    return m;
  }
};

// source.cpp:10-40: comma-separated rows
inline CSVRows readCSV(const std::string& file) {
  std::ifstream in(file);
  if (!in) throw std::invalid_argument("File does not exist");
  CSVRows rows;
  std::string line;
  while (std::getline(in, line)) {
    std::stringstream ls(line);
    std::string a, b;
    if (!std::getline(ls, a, ',') || !std::getline(ls, b, ',')) continue;
    rows.emplace_back(std::stod(a), std::stod(b));
  }
  return rows;
}
}  // namespace SourceHelpers

// ------------------------------------------------------------------ domain and tables
struct Voxel {  // voxel.h:12-15
  uint8_t materialID = 0;
  VectorValue dose;
};

class VoxelGrid {
 public:
  int dims[3] = {0, 0, 0};
  double spacing[3] = {0, 0, 0};
  std::vector<Voxel> voxels;
  Voxel& getVoxel(int i, int j, int k) { return voxels.at((size_t)i + (size_t)dims[0] * (j + (size_t)dims[1] * k)); }
  double getTotalEnergyDeposited() const {
    double s = 0;
    for (auto& v : voxels) s += v.dose.getSum();
    return s;
  }
  std::unordered_map<int, VectorValue> getEnergyDepositedInMaterials() const {
    std::unordered_map<int, VectorValue> m;
    for (auto& v : voxels) m[v.materialID] = m[v.materialID] + v.dose;
    return m;
  }
  // voxel_grid.h:116-118: the grid's extent in cm (voxels x spacing per axis)
  Vec3 getDimSpace() const { return Vec3(dims[0] * spacing[0], dims[1] * spacing[1], dims[2] * spacing[2]); }
};

class InteractionData;

class ComputationalDomain {
 public:
  ComputationalDomain() = default;
  explicit ComputationalDomain(const std::string& json_file_path, bool = false) : path_(json_file_path), host_(-1) {
    midsx_amd_detail::check(mdx_load_scene(host_.p, json_file_path.c_str()));
    int n = 0;
    midsx_amd_detail::check(mdx_num_grids(host_.p, &n));
    grids_.resize(n);
    for (int g = 0; g < n; ++g) {
      double org[3];
      midsx_amd_detail::check(mdx_grid_info(host_.p, g, grids_[g].dims, grids_[g].spacing, org));
      origins_.emplace_back(org[0], org[1], org[2]);
      size_t nv = (size_t)grids_[g].dims[0] * grids_[g].dims[1] * grids_[g].dims[2];
      std::vector<uint8_t> ids(nv);
      midsx_amd_detail::check(mdx_get_grid_ids(host_.p, g, ids.data()));
      grids_[g].voxels.resize(nv);
      for (size_t i = 0; i < nv; ++i) grids_[g].voxels[i].materialID = ids[i];
    }
  }
  inline InteractionData getInteractionData() const;
  int getNumVoxelGrids() const { return (int)grids_.size(); }
  VoxelGrid& getVoxelGridN(int N) { return grids_.at(N); }
  Vec3 getVoxelGridOriginN(int N) const { return origins_.at(N); }
  const std::string& path() const { return path_; }
  mdx_ctx* host() const { return host_.p; }

 private:
  std::string path_;
  midsx_amd_detail::Ctx host_{-1};
  std::vector<VoxelGrid> grids_;
  std::vector<Vec3> origins_;
};

class InteractionData {
 public:
  explicit InteractionData(const std::vector<std::string>& material_names) : host_(std::make_shared<midsx_amd_detail::Ctx>(-1)) {
    for (auto& n : material_names) {
      int id = 0;
      midsx_amd_detail::check(mdx_material_id(host_->p, n.c_str(), &id));
      ids_.push_back(id);
    }
    if (!ids_.empty()) midsx_amd_detail::check(mdx_build_tables(host_->p, ids_.data(), (int)ids_.size()));
  }
  // material_data.h:95-151 on the FP64 tables
  double interpolate(int material_id, int which, double x) const {
    double out = 0;
    midsx_amd_detail::check(mdx_interpolate_host(host_->p, material_id, which, 1, &x, &out));
    return out;
  }
  double interpolateMaxTotalCrossSection(double energy) const { return interpolate(-1, MDX_XS_TOTAL, energy); }
  std::string getAnyMaterialNameFromId(int id) const {
    char buf[256];
    midsx_amd_detail::check(mdx_material_name(host_->p, id, buf, sizeof buf));
    return buf;
  }
  uint8_t getAnyMaterialIdFromName(const std::string& name) const {
    int id = 0;
    midsx_amd_detail::check(mdx_material_id(host_->p, name.c_str(), &id));
    return (uint8_t)id;
  }
  const std::vector<int>& ids() const { return ids_; }
  // mu_en / rho of dry air (cm^2/g), TG-195 table (derived_quantities.cpp:53-57)
  double interpolateAirMassEnergyAbsorptionCoefficient(double energy) const {
    double out = 0;
    midsx_amd_detail::check(mdx_air_muen(host_->p, 1, &energy, &out));
    return out;
  }

 private:
  std::shared_ptr<midsx_amd_detail::Ctx> host_;
  std::vector<int> ids_;
};

// derived_quantities.cpp:3-68 from a surface tally's IncidentEnergy
// histogram (SurfaceTally::setSpectrum): the reference keeps every incident
// energy, this API keeps their histogram, so a fluence "at E +- w/2" counts
// the primary photons of the bins whose centres lie in [E - w/2, E + w/2].
// With bins centred on the source spectrum's energies that is the
// reference's count for primary photons.  Without a histogram, the fluence
// counts every primary photon (a monoenergetic source's, as the reference's
// monoenergetic drivers use it).  Cosine weighting is not available.
struct DerivedQuantity {
  static double getPrimaryFluence(SurfaceQuantityContainer& c, double energy, double energy_width,
                                  bool is_cosine_weighted) {
    if (is_cosine_weighted) throw std::invalid_argument("cosine-weighted fluence is not available");
    if (c.getArea() <= 0) throw std::runtime_error("surface area unknown");
    double total = 0;
    if (c.spectrum_bins > 0 && !c.spectrum.empty()) {
      for (int i = 0; i < c.spectrum_bins; ++i) {
        const double centre = c.spectrum_lo + (i + 0.5) * c.spectrum_width;
        if (centre >= energy - energy_width / 2 && centre <= energy + energy_width / 2) total += c.spectrum[i];
      }
    } else {
      auto& q = c.getCountQuantities().at(CountSurfaceQuantityType::NumberOfPhotons);
      total = q.getPrimaryValues().getCount();
    }
    return total / c.getArea();
  }
  static double getPrimaryAirKerma(SurfaceQuantityContainer& c, InteractionData& data, double energy,
                                   double energy_width, bool is_cosine_weighted) {
    const double mu_en = data.interpolateAirMassEnergyAbsorptionCoefficient(energy);
    return energy * getPrimaryFluence(c, energy, energy_width, is_cosine_weighted) * mu_en;
  }
  // energy_spectrum: evenly spaced bin centres (eV); any type with size() and operator[]
  template <class V>
  static double getPrimarySpectrumAirKerma(SurfaceQuantityContainer& c, InteractionData& data,
                                           const V& energy_spectrum, bool is_cosine_weighted) {
    const double w = energy_spectrum[1] - energy_spectrum[0];
    double total = 0;
    for (size_t i = 0; i < (size_t)energy_spectrum.size(); ++i)
      total += getPrimaryAirKerma(c, data, energy_spectrum[i], w, is_cosine_weighted);
    return total;
  }
};

InteractionData ComputationalDomain::getInteractionData() const {
  // computational_domain.cpp:109-123: background first, then each grid's materials
  std::vector<std::string> names;
  auto add = [&](int id) {
    char buf[256];
    midsx_amd_detail::check(mdx_material_name(host_.p, id, buf, sizeof buf));
    std::string s(buf);
    for (auto& n : names)
      if (n == s) return;
    names.push_back(s);
  };
  std::ifstream f(path_);
  std::string src((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  size_t k = src.find("background_material_name");
  size_t a = src.find('"', src.find(':', k) + 1), b = src.find('"', a + 1);
  int bg = 0;
  midsx_amd_detail::check(mdx_material_id(host_.p, src.substr(a + 1, b - a - 1).c_str(), &bg));
  add(bg);
  for (auto& g : grids_) {
    bool seen[256] = {false};
    for (auto& v : g.voxels) seen[v.materialID] = true;
    for (int i = 0; i < 256; ++i)
      if (seen[i]) add(i);
  }
  return InteractionData(names);
}

// ------------------------------------------------------------------ engine and driver
class PhysicsEngine {
 public:
  PhysicsEngine(ComputationalDomain& comp_domain, InteractionData& interaction_data, int device = 0)
      : domain_(comp_domain), data_(interaction_data), ctx_(device) {
    midsx_amd_detail::check(mdx_load_scene(ctx_.p, comp_domain.path().c_str()));
    const auto& ids = interaction_data.ids();
    midsx_amd_detail::check(mdx_build_tables(ctx_.p, ids.data(), (int)ids.size()));
    midsx_amd_detail::check(mdx_enable_voxel_dose(ctx_.p, 1));
  }

  void setTallies(std::vector<std::unique_ptr<SurfaceTally>> s, std::vector<std::unique_ptr<VolumeTally>> v) {
    bool same = s.size() == surface_.size() && v.size() == volume_.size();
    surface_ = std::move(s);
    volume_ = std::move(v);
    std::vector<mdx_tally> d;
    for (auto& t : volume_) d.push_back(t->descriptor());
    for (auto& t : surface_) d.push_back(t->descriptor());
    midsx_amd_detail::check(mdx_set_tallies(ctx_.p, d.data(), (int)d.size()));
    if (!same) {
      acc_.assign(d.size(), std::vector<double>(MDX_TALLY_ACC_SIZE, 0.0));
      hist_.assign(d.size(), {});
      img_.assign(d.size(), {});
    }
  }

  // Transports n histories and adds the results to the engine's totals:
  // results accumulate over runs, as the reference's do when runSimulation is
  // called again on one engine (physics_engine.cpp:111-122).
  void run(const PhotonSource& source, uint64_t n) {
    source.upload(ctx_.p);
    mdx_stats st;
    midsx_amd_detail::check(mdx_run(ctx_.p, n, seed, histories_done_, kernel, &st));
    histories_done_ += n;
    last_ = st;
    std::vector<double> acc(MDX_TALLY_ACC_SIZE);
    for (size_t i = 0; i < acc_.size(); ++i) {
      const mdx_tally& d = i < volume_.size() ? volume_[i]->descriptor() : surface_[i - volume_.size()]->descriptor();
      std::vector<double> h(d.kind != MDX_TALLY_AACUBOID && d.hist_bins > 0 ? 4 * (size_t)d.hist_bins : 0);
      std::vector<double> im(d.kind == MDX_TALLY_IMAGE ? 3 * (size_t)d.nx * d.ny : 0);
      midsx_amd_detail::check(mdx_get_tally(ctx_.p, (int)i, acc.data(), h.empty() ? nullptr : h.data(),
                                            im.empty() ? nullptr : im.data()));
      add(acc_[i], acc);
      add(hist_[i], h);
      add(img_[i], im);
    }
    for (int g = 0; g < domain_.getNumVoxelGrids(); ++g) {
      VoxelGrid& G = domain_.getVoxelGridN(g);
      std::vector<double> ss(2 * G.voxels.size());
      std::vector<uint64_t> cnt(G.voxels.size());
      midsx_amd_detail::check(mdx_get_voxel_dose(ctx_.p, g, ss.data(), cnt.data()));
      for (size_t i = 0; i < G.voxels.size(); ++i)
        if (cnt[i]) G.voxels[i].dose.addMoments((double)cnt[i], ss[2 * i], ss[2 * i + 1]);
    }
    midsx_amd_detail::check(mdx_reset_tallies(ctx_.p));
  }

  std::vector<VolumeQuantityContainer> getVolumeQuantityContainers() {
    std::vector<VolumeQuantityContainer> out;
    for (size_t i = 0; i < volume_.size(); ++i) {
      VolumeQuantityContainer c = volume_[i]->getVolumeQuantityContainer();
      c.load(acc_[i].data());
      out.push_back(c);
    }
    return out;
  }
  std::vector<SurfaceQuantityContainer> getSurfaceQuantityContainers() {
    std::vector<SurfaceQuantityContainer> out;
    for (size_t i = 0; i < surface_.size(); ++i) {
      size_t k = volume_.size() + i;
      SurfaceQuantityContainer c = surface_[i]->getSurfaceQuantityContainer();
      c.load(acc_[k].data());
      c.spectrum = hist_[k];
      const mdx_tally& d = surface_[i]->descriptor();
      c.spectrum_bins = d.hist_bins;
      c.spectrum_lo = d.hist_lo;
      c.spectrum_width = d.hist_width;
      if (auto* im = dynamic_cast<ImageSurfaceTally*>(surface_[i].get())) im->image = img_[k];
      out.push_back(c);
    }
    return out;
  }
  // Voxel deposits are folded into the domain's voxels after every run
  // (physics_engine.cpp:160-168); kept for drivers that call it explicitly.
  void addVoxelDataToComputationalDomain() {}
  const mdx_stats& lastStats() const { return last_; }
  mdx_ctx* context() const { return ctx_.p; }
  uint64_t seed = 1;
  // MDX_KERNEL_QUEUE (default: the event-regrouped kernel) or MDX_KERNEL_PLAIN
  int kernel = MDX_KERNEL_QUEUE;

 private:
  static void add(std::vector<double>& a, const std::vector<double>& b) {
    if (a.size() < b.size()) a.resize(b.size(), 0.0);
    for (size_t i = 0; i < b.size(); ++i) a[i] += b[i];
  }
  ComputationalDomain& domain_;
  InteractionData& data_;
  midsx_amd_detail::Ctx ctx_;
  std::vector<std::unique_ptr<SurfaceTally>> surface_;
  std::vector<std::unique_ptr<VolumeTally>> volume_;
  std::vector<std::vector<double>> acc_, hist_, img_;
  uint64_t histories_done_ = 0;
  mdx_stats last_{};
};

// run_simulation.h:24-27.  run_time is the wall time of the transport.
inline void runSimulation(PhotonSource& source, PhysicsEngine& physics_engine,
                          std::function<std::vector<std::unique_ptr<SurfaceTally>>()> surface_tally_init,
                          std::function<std::vector<std::unique_ptr<VolumeTally>>()> volume_tally_init,
                          int N_photons, double& run_time) {
  if (N_photons < 0) throw std::invalid_argument("N_photons must be non-negative");
  physics_engine.setTallies(surface_tally_init(), volume_tally_init());
  auto t0 = std::chrono::steady_clock::now();
  physics_engine.run(source, (uint64_t)N_photons);
  auto t1 = std::chrono::steady_clock::now();
  run_time = std::chrono::duration<double>(t1 - t0).count();
}

// The reference defaults run_time (run_simulation.h:27), so drivers call the
// five-argument form (hvl_30kVp.cpp:70, CT_w_voxelized_solid.cpp:166,
// radiography_15_degrees.cpp:326); the run time is then not reported.
inline void runSimulation(PhotonSource& source, PhysicsEngine& physics_engine,
                          std::function<std::vector<std::unique_ptr<SurfaceTally>>()> surface_tally_init,
                          std::function<std::vector<std::unique_ptr<VolumeTally>>()> volume_tally_init,
                          int N_photons) {
  double run_time = 0.0;
  runSimulation(source, physics_engine, std::move(surface_tally_init), std::move(volume_tally_init), N_photons,
                run_time);
}
