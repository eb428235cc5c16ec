// Leapfrog integrator and energy diagnostic -- same names and meaning as the
// reference's src/integrator.cuh:4-21.
#pragma once
#include "simulation.hpp"

struct Energy {
  double total;
  double kinetic;
  double potential;
};

// Kick-drift-kick (integrator.cuh:12-19): on entry d_acc holds a(x); on exit
// it holds a(x_new).  The first kick and the drift run as one kernel and the
// closing kick is fused into the force kernel's epilogue.
void leapfrog_step(ParticleSystem& sys, float4* d_acc, const SimParams& p);

// KE + PE reduced in FP64 on the device (integrator.cu:93-110).
Energy total_energy(const ParticleSystem& sys, const SimParams& p);

// Additive: the same energy at O(n log n) cost for large n -- PE from a
// wave64 tree walk at theta_e in (0, 0.577) (see nb_total_energy_tree).
Energy total_energy_tree(const ParticleSystem& sys, const SimParams& p, float theta_e);
