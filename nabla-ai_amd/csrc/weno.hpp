// weno.hpp — the Jiang-Shu WENO-5 face reconstruction shared by the MAC
// right-hand side (ns.hip) and the octree flow step (octflow.hip).
#pragma once

namespace nbl {

// Jiang-Shu WENO-5 at the right face of v2 (ns_solver.cpp:15-31), same
// expression order as the reference.
__device__ __forceinline__ double weno5(double v0, double v1, double v2, double v3, double v4) {
  const double e = 1e-12;
  const double s0 = 13.0 / 12.0 * (v0 - 2 * v1 + v2) * (v0 - 2 * v1 + v2) +
                    0.25 * (v0 - 4 * v1 + 3 * v2) * (v0 - 4 * v1 + 3 * v2);
  const double s1 = 13.0 / 12.0 * (v1 - 2 * v2 + v3) * (v1 - 2 * v2 + v3) + 0.25 * (v1 - v3) * (v1 - v3);
  const double s2 = 13.0 / 12.0 * (v2 - 2 * v3 + v4) * (v2 - 2 * v3 + v4) +
                    0.25 * (3 * v2 - 4 * v3 + v4) * (3 * v2 - 4 * v3 + v4);
  const double a0 = 0.1 / ((e + s0) * (e + s0));
  const double a1 = 0.6 / ((e + s1) * (e + s1));
  const double a2 = 0.3 / ((e + s2) * (e + s2));
  const double sw = a0 + a1 + a2;
  const double q0 = (2 * v0 - 7 * v1 + 11 * v2) / 6.0;
  const double q1 = (-v1 + 5 * v2 + 2 * v3) / 6.0;
  const double q2 = (2 * v2 + 5 * v3 - v4) / 6.0;
  return (a0 * q0 + a1 * q1 + a2 * q2) / sw;
}

}  // namespace nbl
