// The polyenergetic half-value-layer driver written against the reference's
// API (cpp_simulations/half_value_layer/hvl_30kVp.cpp:9-80): a spectrum read
// from CSV (keV, weight), a disc surface tally, runSimulation and the primary
// air kerma from DerivedQuantity::getPrimarySpectrumAirKerma.  The tally's
// IncidentEnergy histogram is given the spectrum's bins (setSpectrum), which
// is what the derived quantity reads here.
//
// usage: hvl_kvp <scene.json> <spectrum.csv> <n_photons>
#include <MIDSX/Core.h>

#include <cstdio>
#include <cstdlib>

static std::vector<std::pair<double, double>> spectrum_eV;

std::vector<std::unique_ptr<SurfaceTally>> initializeSurfaceTallies() {
  std::vector<std::unique_ptr<SurfaceTally>> tallies;
  auto disc = std::make_unique<DiscSurfaceTally>(Vec3(2, 2, 100), Vec3(0, 0, 1), 1.0,
                                                 SurfaceQuantityContainerFactory::AllQuantities());
  const double w = spectrum_eV[1].first - spectrum_eV[0].first;
  disc->setSpectrum((int)spectrum_eV.size(), spectrum_eV[0].first - w / 2, w);
  tallies.emplace_back(std::move(disc));
  return tallies;
}

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s scene.json spectrum.csv n_photons\n", argv[0]);
    return 2;
  }
  spectrum_eV = SourceHelpers::readCSV(argv[2]);
  for (auto& r : spectrum_eV) r.first *= 1000;  // keV -> eV
  ComputationalDomain comp_domain(argv[1]);
  InteractionData interaction_data = comp_domain.getInteractionData();
  PhysicsEngine physics_engine(comp_domain, interaction_data);
  PhotonSource source(std::make_unique<PolyenergeticSpectrum>(spectrum_eV),
                      std::make_unique<DiscIsotropicDirectionality>(Vec3(2, 2, 10), Vec3(0, 0, 1), 0.1),
                      std::make_unique<PointGeometry>(Vec3(2, 2, 0)));
  int n = std::atoi(argv[3]);
  // the five-argument call of hvl_30kVp.cpp:70 (run_time defaulted)
  runSimulation(source, physics_engine, initializeSurfaceTallies,
                []() { return std::vector<std::unique_ptr<VolumeTally>>(); }, n);
  auto surface = physics_engine.getSurfaceQuantityContainers();
  std::vector<double> energies;
  for (auto& r : spectrum_eV) energies.push_back(r.first);
  double ak = DerivedQuantity::getPrimarySpectrumAirKerma(surface.at(0), interaction_data, energies, false);
  auto& photons = surface.at(0).getCountQuantities().at(CountSurfaceQuantityType::NumberOfPhotons);
  std::printf("primary_photons %d\n", photons.getPrimaryValues().getCount());
  std::printf("primary_air_kerma %.9e\n", ak);
  return 0;
}
