// Call forms of the reference's drivers that pass or receive Eigen types
// (readCSV into a MatrixXd, PolyenergeticSpectrum from it, getDimSpace into a
// Vector3d, the quantity-name helpers), written against minimal matrix and
// vector types with the members those drivers use, so that the drop-in
// header is checked without Eigen.  Host-only: no device work.
#include <MIDSX/Core.h>

#include <cassert>
#include <cstdio>
#include <vector>

struct Vector3 {  // the members of Eigen::Vector3d the drivers use
  double v[3];
  Vector3(double a, double b, double c) : v{a, b, c} {}
  double x() const { return v[0]; }
  double y() const { return v[1]; }
  double z() const { return v[2]; }
};

struct Matrix {  // the members of Eigen::MatrixXd the drivers use
  long r, c;
  std::vector<double> d;
  Matrix(long rows, long cols) : r(rows), c(cols), d(rows * cols, 0.0) {}
  long rows() const { return r; }
  long cols() const { return c; }
  double& operator()(long i, long j) { return d[i * c + j]; }
  double operator()(long i, long j) const { return d[i * c + j]; }
};

int main(int argc, char** argv) {
  assert(argc == 2);
  Matrix m = SourceHelpers::readCSV(argv[1]);  // hvl_30kVp.cpp:21
  assert(m.rows() == 3 && m.cols() == 2 && m(2, 0) == 30000.0 && m(1, 1) == 0.5);
  std::vector<std::pair<double, double>> rows = SourceHelpers::readCSV(argv[1]);
  assert(rows.size() == 3 && rows[0].first == 10000.0);
  PolyenergeticSpectrum spec(m);  // hvl_30kVp.cpp:29
  const double mean = (10000.0 * 0.25 + 20000.0 * 0.5 + 30000.0 * 0.25) / 1.0;
  assert(std::fabs(spec.getExpectationValue() - mean) < 1e-9);
  VoxelGrid g;
  g.dims[0] = 4; g.dims[1] = 5; g.dims[2] = 6;
  g.spacing[0] = 0.5; g.spacing[1] = 0.25; g.spacing[2] = 2.0;
  Vector3 ext = g.getDimSpace();  // lead_attenuation.cpp:176
  assert(ext.x() == 2.0 && ext.y() == 1.25 && ext.z() == 12.0);
  assert(VectorVolumeQuantityHelper::toString(VectorVolumeQuantityType::EnergyDeposition) == "EnergyDeposition");
  assert(VectorVolumeQuantityHelper::toString(VectorVolumeQuantityType::IncidentEnergy) == "IncidentEnergy");
  assert(CountVolumeQuantityHelper::toString(CountVolumeQuantityType::NumberOfPhotons) == "NumberOfPhotons");
  assert(CountVolumeQuantityHelper::toString(CountVolumeQuantityType::NumberOfInteractions) == "NumberOfInteractions");
  assert(VectorSurfaceQuantityHelper::toString(VectorSurfaceQuantityType::IncidentEnergy) == "IncidentEnergy");
  assert(VectorSurfaceQuantityHelper::toString(VectorSurfaceQuantityType::EntranceCosine) == "EntranceCosine");
  assert(CountSurfaceQuantityHelper::toString(CountSurfaceQuantityType::NumberOfPhotons) == "NumberOfPhotons");
  std::cout << "matrix call forms ok" << std::endl;  // <iostream> comes with the header, as in the reference
  return 0;
}
