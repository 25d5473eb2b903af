// A half-value-layer style driver written against the reference's API
// (cpp_simulations/half_value_layer/hvl_30keV.cpp:9-80): scene JSON with a
// NIfTI grid, a disc surface tally, a volume tally, runSimulation, results
// through the quantity containers.
//
// usage: hvl_slab <scene.json> <n_photons> [energy_eV]
#include <MIDSX/Core.h>

#include <cstdio>
#include <cstdlib>

std::vector<std::unique_ptr<SurfaceTally>> initializeSurfaceTallies() {
  std::vector<std::unique_ptr<SurfaceTally>> tallies;
  tallies.emplace_back(std::make_unique<DiscSurfaceTally>(Vec3(2, 2, 100), Vec3(0, 0, 1), 1.0,
                                                          SurfaceQuantityContainerFactory::AllQuantities()));
  return tallies;
}

std::vector<std::unique_ptr<VolumeTally>> initializeVolumeTallies() {
  std::vector<std::unique_ptr<VolumeTally>> tallies;
  tallies.emplace_back(std::make_unique<AACuboidVolumeTally>(Vec3(0, 0, 10), Vec3(4, 4, 11),
                                                             VolumeQuantityContainerFactory::AllQuantities()));
  return tallies;
}

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s scene.json n_photons [energy_eV]\n", argv[0]);
    return 2;
  }
  double energy = argc > 3 ? std::atof(argv[3]) : 30e3;
  ComputationalDomain comp_domain(argv[1]);
  InteractionData interaction_data = comp_domain.getInteractionData();
  PhysicsEngine physics_engine(comp_domain, interaction_data);
  PhotonSource source(std::make_unique<MonoenergeticSpectrum>(energy),
                      std::make_unique<DiscIsotropicDirectionality>(Vec3(2, 2, 10), Vec3(0, 0, 1), 0.1),
                      std::make_unique<PointGeometry>(Vec3(2, 2, 0)));
  int n = std::atoi(argv[2]);
  double run_time = 0;
  runSimulation(source, physics_engine, initializeSurfaceTallies, initializeVolumeTallies, n, run_time);
  auto surface = physics_engine.getSurfaceQuantityContainers();
  auto volume = physics_engine.getVolumeQuantityContainers();
  auto& photons = surface.at(0).getCountQuantities().at(CountSurfaceQuantityType::NumberOfPhotons);
  auto& energies = surface.at(0).getVectorQuantities().at(VectorSurfaceQuantityType::IncidentEnergy);
  auto& edep = volume.at(0).getVectorQuantities().at(VectorVolumeQuantityType::EnergyDeposition);
  std::printf("run_time %.6f\n", run_time);
  std::printf("primary_photons %d\n", photons.getPrimaryValues().getCount());
  std::printf("total_photons %d\n", photons.getTotalValues().getCount());
  std::printf("primary_energy_sum %.6f\n", energies.getPrimaryValues().getSum());
  std::printf("slab_energy_deposit %.6f %.6f\n", edep.getTotalValues().getSum(), edep.getTotalValues().getSumSTD());
  double grid_dep = 0;
  for (int g = 0; g < comp_domain.getNumVoxelGrids(); ++g) grid_dep += comp_domain.getVoxelGridN(g).getTotalEnergyDeposited();
  std::printf("grid_energy_deposit %.6f\n", grid_dep);
  return 0;
}
