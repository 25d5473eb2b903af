// Scene I/O: the scene JSON of ComputationalDomain and NIfTI-1 material grids.
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace mdx {

struct NiftiGrid {
  std::array<int, 3> dims{};
  std::array<double, 3> spacing{};  // cm
  std::vector<uint8_t> ids;         // x-fastest
};

// voxel_grid.cpp:109-181 without nibabel: header dims, pixdim x xyzt units
// (converted to cm, voxel_grid.cpp:12-22), data scaled by scl_slope /
// scl_inter and cast to uint8, x-fastest.
NiftiGrid read_nifti(const std::string& path);

struct SceneGrid {
  std::string path;
  std::array<double, 3> origin{};
};

struct Scene {
  std::array<double, 3> dim_space{};
  std::string background_material_name;
  std::vector<SceneGrid> grids;
};

// computational_domain.cpp:50-107, 133-159: keys dim_space,
// background_material_name, voxel_grids[{file_path, origin}]; relative NIfTI
// paths resolve against the JSON's directory.
Scene read_scene(const std::string& json_path);

}  // namespace mdx
