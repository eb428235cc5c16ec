// Swappable force-computation interface -- same names and meaning as the
// reference's src/forces.cuh:8-25.  Errors abort the process with the HIP
// error text, as the reference's CUDA_CHECK does (cuda_check.cuh:7-15).
#pragma once
#include "simulation.hpp"

// Fills d_acc (device, sys.n float4, w unused) with G * acceleration, in the
// original particle order, using the module p.force selects.  Blocks until
// done and adds the elapsed time to the force timer (forces.cu:68-86).
void compute_forces(const ParticleSystem& sys, float4* d_acc, const SimParams& p);

// The implementations behind the dispatcher (forces.cuh:11-12).
void forces_naive(const ParticleSystem& sys, float4* d_acc, const SimParams& p);
void forces_barnes_hut(const ParticleSystem& sys, float4* d_acc, const SimParams& p);

// Force timing accumulated by the dispatcher (forces.cuh:17-19).
double force_kernel_ms_total();
long force_kernel_calls();
void reset_force_timing();

// Barnes-Hut per-phase timing table and scratch cleanup (forces.cuh:22-23).
void bh_report_phase_timing();
void bh_release();

constexpr int kForceBlock = 256;
