// Uniform MAC grid description — the reference's UniformGrid / FaceBC types
// (core/include/nabla/flow/grid.hpp:9-73), same names and meaning.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "nbl_api.h"

namespace nabla::flow {

struct V3 {
  double x{0.0};
  double y{0.0};
  double z{0.0};
};

enum class BoundaryType { Wall = NBL_BC_WALL, Inlet = NBL_BC_INLET, Outlet = NBL_BC_OUTLET, Periodic = NBL_BC_PERIODIC };
enum class InletProfile { Uniform = NBL_INLET_UNIFORM, Parabolic = NBL_INLET_PARABOLIC };

struct FaceBC {
  BoundaryType type = BoundaryType::Wall;
  V3 wallVelocity{};
  InletProfile profile = InletProfile::Uniform;
  double inletSpeed = 0.0;
  int inletAxis = 0;
};

// Faces 0=-x 1=+x 2=-y 3=+y 4=-z 5=+z; compact cell index i + nx*(j + ny*k).
struct UniformGrid {
  int nx = 1, ny = 1, nz = 1;
  V3 origin{};
  double dx = 1.0, dy = 1.0, dz = 1.0;
  FaceBC face[6];
  std::vector<uint8_t> solid;

  void allocate() { solid.assign(cellCount(), 0); }
  [[nodiscard]] std::size_t cellCount() const {
    return std::size_t(nx) * std::size_t(ny) * std::size_t(nz);
  }
  [[nodiscard]] std::size_t cidx(int i, int j, int k) const {
    return std::size_t(i) + std::size_t(nx) * (std::size_t(j) + std::size_t(ny) * std::size_t(k));
  }
  [[nodiscard]] bool isSolid(int i, int j, int k) const {
    if (i < 0 || i >= nx || j < 0 || j >= ny || k < 0 || k >= nz) return false;
    return solid[cidx(i, j, k)] != 0;
  }
  [[nodiscard]] V3 cellCenter(int i, int j, int k) const {
    return {origin.x + (i + 0.5) * dx, origin.y + (j + 0.5) * dy, origin.z + (k + 0.5) * dz};
  }
  [[nodiscard]] double cellVolume() const { return dx * dy * dz; }
  [[nodiscard]] bool periodic(int axis) const { return face[2 * axis].type == BoundaryType::Periodic; }

  // C-ABI descriptor view (valid while this grid is alive and unchanged).
  [[nodiscard]] nbl_grid_desc desc() const {
    nbl_grid_desc d{};
    d.nx = nx;
    d.ny = ny;
    d.nz = nz;
    d.origin[0] = origin.x;
    d.origin[1] = origin.y;
    d.origin[2] = origin.z;
    d.dx = dx;
    d.dy = dy;
    d.dz = dz;
    for (int f = 0; f < 6; ++f) {
      d.face[f].type = static_cast<int>(face[f].type);
      d.face[f].wall_velocity[0] = face[f].wallVelocity.x;
      d.face[f].wall_velocity[1] = face[f].wallVelocity.y;
      d.face[f].wall_velocity[2] = face[f].wallVelocity.z;
      d.face[f].profile = static_cast<int>(face[f].profile);
      d.face[f].inlet_speed = face[f].inletSpeed;
      d.face[f].inlet_axis = face[f].inletAxis;
    }
    d.solid = solid.empty() ? nullptr : solid.data();
    return d;
  }
};

}  // namespace nabla::flow
