// Particle state and run parameters -- same names and meaning as the
// reference's src/simulation.hpp:4-28, on HIP vector types.
#pragma once
#include <hip/hip_runtime.h>

enum class ForceMethod { kNaive, kBarnesHut };
enum class BhTraversal { kThread, kWarp };

// Simulation parameters shared across host and device code
// (simulation.hpp:8-15).  The trailing fields are additive knobs; their
// defaults reproduce the reference (63-bit keys, 32-lane group walk).
struct SimParams {
  float dt;                                     // timestep
  float epsilon;                                // softening length
  float G = 1.0f;                               // gravitational constant
  float theta = 0.5f;                           // BH opening angle; 0 = exact
  ForceMethod force = ForceMethod::kNaive;      // which force module runs
  BhTraversal traverse = BhTraversal::kThread;  // BH tree-walk variant
  int key_bits = 63;                            // Morton key width: 63 or 30
  int group_width = 32;                         // lanes per kWarp group: 16, 32, 64
  bool compensated = false;                     // Kahan tile sums (all-pairs)
  bool far16 = false;                           // FP16 far field (BH wave64 walk)
};

// Device-resident SoA state (simulation.hpp:17-28): pos = (x, y, z, m),
// vel = (vx, vy, vz, _).
struct ParticleSystem {
  float4* d_pos = nullptr;
  float4* d_vel = nullptr;
  int n = 0;

  void allocate(int count);
  void release();
  void upload(const float4* h_pos, const float4* h_vel);
  void download_positions(float4* h_pos) const;
};
