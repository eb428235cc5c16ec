// galaxy_sim -- command-line driver with the reference's flags, formats and
// report lines (src/main.cu:18-215), over the C++ API of libnbody.so.
//
// Same flags and defaults as main.cu:18-70 (--ic --out --energy-log --steps
// --dump-every --dt --eps --G --theta --compare-forces --force naive|bh
// --traverse thread|warp), same ic.bin / frame / energy CSV formats
// (docs/FORMATS.md; main.cu:114-179) and the same stdout lines, so scripts
// that parse the reference's output (gpu_bench.sh:31,41-44) work unchanged.
// Additive flags: --keys 30|63, --group-width 16|32|64, --compensated, --far16,
// --energy-theta f (energy log from the tree potential at that theta instead
// of the O(n^2) sum: the cost policy for large n), --checkpoint path (write
// the final state, positions and velocities, in the ic.bin format, so that
// --ic path resumes the run).
//
// Frame files are written by a second host thread while the steps go on: a
// dump copies the positions to one of two host buffers (the copy the
// reference makes too) and hands the buffer to the writer, which the next
// dump waits for before reusing the other buffer.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <future>
#include <string>
#include <vector>

#include "nbody/forces.hpp"
#include "nbody/integrator.hpp"
#include "nbody/simulation.hpp"

namespace fs = std::filesystem;

namespace {

struct Args {
  std::string ic_path;
  std::string out_dir = "frames";
  std::string energy_log = "benchmarks/energy_gpu.csv";
  int steps = 1000;
  int dump_every = 5;
  bool compare = false;
  float energy_theta = 0.f;  // 0: exact O(n^2) energy
  std::string checkpoint;    // empty: none
  SimParams params{/*dt=*/0.01f, /*epsilon=*/0.05f, /*G=*/1.0f};
};

[[noreturn]] void die(const std::string& msg) {
  std::fprintf(stderr, "%s\n", msg.c_str());
  std::exit(EXIT_FAILURE);
}

int to_int(const std::string& k, const std::string& v) {
  try {
    size_t used = 0;
    int x = std::stoi(v, &used);
    if (used != v.size()) throw std::invalid_argument(v);
    return x;
  } catch (const std::exception&) {
    die("bad value for " + k + ": " + v);
  }
}

float to_float(const std::string& k, const std::string& v) {
  try {
    size_t used = 0;
    float x = std::stof(v, &used);
    if (used != v.size()) throw std::invalid_argument(v);
    return x;
  } catch (const std::exception&) {
    die("bad value for " + k + ": " + v);
  }
}

Args parse_args(int argc, char** argv) {
  Args a;
  for (int i = 1; i < argc; ++i) {
    std::string k = argv[i];
    auto next = [&]() -> std::string {
      if (i + 1 >= argc) die("missing value for " + k);
      return argv[++i];
    };
    if (k == "--ic") a.ic_path = next();
    else if (k == "--out") a.out_dir = next();
    else if (k == "--energy-log") a.energy_log = next();
    else if (k == "--steps") a.steps = to_int(k, next());
    else if (k == "--dump-every") a.dump_every = to_int(k, next());
    else if (k == "--dt") a.params.dt = to_float(k, next());
    else if (k == "--eps") a.params.epsilon = to_float(k, next());
    else if (k == "--G") a.params.G = to_float(k, next());
    else if (k == "--theta") a.params.theta = to_float(k, next());
    else if (k == "--compare-forces") a.compare = true;
    else if (k == "--force") {
      std::string v = next();
      if (v == "naive") a.params.force = ForceMethod::kNaive;
      else if (v == "bh") a.params.force = ForceMethod::kBarnesHut;
      else die("--force must be naive or bh, got: " + v);
    } else if (k == "--traverse") {
      std::string v = next();
      if (v == "thread") a.params.traverse = BhTraversal::kThread;
      else if (v == "warp") a.params.traverse = BhTraversal::kWarp;
      else die("--traverse must be thread or warp, got: " + v);
    } else if (k == "--keys") {
      std::string v = next();
      if (v == "30") a.params.key_bits = 30;
      else if (v == "63") a.params.key_bits = 63;
      else die("--keys must be 30 or 63, got: " + v);
    } else if (k == "--group-width") {
      std::string v = next();
      if (v == "16" || v == "32" || v == "64") a.params.group_width = to_int(k, v);
      else die("--group-width must be 16, 32 or 64, got: " + v);
    } else if (k == "--compensated") {
      a.params.compensated = true;
    } else if (k == "--far16") {
      a.params.far16 = true;
    } else if (k == "--checkpoint") {
      a.checkpoint = next();
    } else if (k == "--energy-theta") {
      a.energy_theta = to_float(k, next());
      if (!(a.energy_theta >= 0.f && a.energy_theta < 0.577f))
        die("--energy-theta must be in [0, 0.577)");
    } else {
      die("unknown argument: " + k);
    }
  }
  if (a.ic_path.empty())
    die("usage: galaxy_sim --ic ic.bin [--steps N] [--dt f] [--eps f] [--dump-every N] "
        "[--out dir] [--force naive|bh] [--theta f] [--traverse thread|warp] "
        "[--compare-forces] [--keys 30|63] [--group-width 16|32|64] [--compensated] "
        "[--far16] [--energy-theta f] [--checkpoint path]");
  if (a.dump_every <= 0) die("--dump-every must be positive");
  return a;
}

// main.cu:75-112: naive vs Barnes-Hut on one state; FP64 host error report.
int compare_forces(ParticleSystem& sys, const SimParams& base) {
  int n = sys.n;
  float4 *d_a = nullptr, *d_b = nullptr;
  if (hipMalloc(reinterpret_cast<void**>(&d_a), n * sizeof(float4)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&d_b), n * sizeof(float4)) != hipSuccess)
    die("device allocation failed in compare_forces");
  SimParams pn = base, pb = base;
  pn.force = ForceMethod::kNaive;
  pb.force = ForceMethod::kBarnesHut;
  compute_forces(sys, d_a, pn);
  compute_forces(sys, d_b, pb);
  std::vector<float4> a(n), b(n);
  if (hipMemcpy(a.data(), d_a, n * sizeof(float4), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(b.data(), d_b, n * sizeof(float4), hipMemcpyDeviceToHost) != hipSuccess)
    die("device copy failed in compare_forces");
  double max_rel = 0.0, sum_rel = 0.0;
  for (int i = 0; i < n; ++i) {
    double dx = a[i].x - b[i].x, dy = a[i].y - b[i].y, dz = a[i].z - b[i].z;
    double an = std::sqrt((double)a[i].x * a[i].x + (double)a[i].y * a[i].y +
                          (double)a[i].z * a[i].z);
    double rel = std::sqrt(dx * dx + dy * dy + dz * dz) / (an + 1e-30);
    sum_rel += rel;
    if (rel > max_rel) max_rel = rel;
  }
  std::printf("naive vs barnes-hut, theta=%.3g, traverse=%s, n=%d\n", base.theta,
              base.traverse == BhTraversal::kWarp ? "warp" : "thread", n);
  std::printf("relative force error: max %.3e, mean %.3e\n", max_rel, sum_rel / n);
  (void)hipFree(d_a);
  (void)hipFree(d_b);
  bh_release();
  return 0;
}

// main.cu:115-128
int load_ic(const std::string& path, std::vector<float4>& pos, std::vector<float4>& vel) {
  std::ifstream f(path, std::ios::binary);
  if (!f) die("cannot open IC file: " + path);
  std::int32_t n = 0;
  f.read(reinterpret_cast<char*>(&n), sizeof(n));
  if (!f || n <= 0) die("bad particle count in " + path);
  pos.resize(n);
  vel.resize(n);
  f.read(reinterpret_cast<char*>(pos.data()), n * sizeof(float4));
  f.read(reinterpret_cast<char*>(vel.data()), n * sizeof(float4));
  if (!f) die("truncated IC file: " + path);
  return n;
}

// main.cu:130-138, on the writer thread (unchecked writes, as there).
void dump_frame(const std::vector<float4>& pos, const std::string& dir, int frame) {
  char name[64];
  std::snprintf(name, sizeof(name), "frame_%05d.bin", frame);
  std::ofstream f(fs::path(dir) / name, std::ios::binary);
  std::int32_t n = static_cast<std::int32_t>(pos.size());
  f.write(reinterpret_cast<const char*>(&n), sizeof(n));
  f.write(reinterpret_cast<const char*>(pos.data()), n * sizeof(float4));
}

// The full state in the ic.bin layout (docs/FORMATS.md:7-18): int32 n, then
// float4 pos[n], then float4 vel[n].
void write_checkpoint(const ParticleSystem& sys, const std::string& path) {
  std::vector<float4> pos(sys.n), vel(sys.n);
  sys.download_positions(pos.data());
  if (hipMemcpy(vel.data(), sys.d_vel, sys.n * sizeof(float4), hipMemcpyDeviceToHost) !=
      hipSuccess)
    die("device copy failed for the checkpoint velocities");
  if (fs::path(path).has_parent_path()) fs::create_directories(fs::path(path).parent_path());
  std::ofstream f(path, std::ios::binary);
  std::int32_t n = sys.n;
  f.write(reinterpret_cast<const char*>(&n), sizeof(n));
  f.write(reinterpret_cast<const char*>(pos.data()), n * sizeof(float4));
  f.write(reinterpret_cast<const char*>(vel.data()), n * sizeof(float4));
  if (!f) die("cannot write checkpoint: " + path);
}

}  // namespace

int main(int argc, char** argv) {
  Args args = parse_args(argc, argv);
  std::vector<float4> h_pos, h_vel;
  int n = load_ic(args.ic_path, h_pos, h_vel);

  ParticleSystem sys;
  sys.allocate(n);
  sys.upload(h_pos.data(), h_vel.data());
  if (args.compare) return compare_forces(sys, args.params);

  float4* d_acc = nullptr;
  if (hipMalloc(reinterpret_cast<void**>(&d_acc), n * sizeof(float4)) != hipSuccess)
    die("device allocation failed for the acceleration array");

  fs::create_directories(args.out_dir);
  if (fs::path(args.energy_log).has_parent_path())
    fs::create_directories(fs::path(args.energy_log).parent_path());
  std::ofstream elog(args.energy_log);
  elog << "step,time,energy,kinetic,potential,rel_error\n";

  bool bh = args.params.force == ForceMethod::kBarnesHut;
  auto energy = [&]() {
    return args.energy_theta > 0.f ? total_energy_tree(sys, args.params, args.energy_theta)
                                   : total_energy(sys, args.params);
  };
  Energy e0 = energy();
  std::printf("n=%d dt=%.4g eps=%.4g steps=%d force=%s", n, args.params.dt, args.params.epsilon,
              args.steps, bh ? "barnes-hut" : "naive");
  if (bh)
    std::printf(" theta=%.3g traverse=%s", args.params.theta,
                args.params.traverse == BhTraversal::kWarp ? "warp" : "thread");
  std::printf("\n");
  std::printf("E0=%.6e KE0=%.6e PE0=%.6e\n", e0.total, e0.kinetic, e0.potential);

  auto log_energy = [&](int step) {
    Energy e = energy();
    double rel = (e.total - e0.total) / std::abs(e0.total);
    char line[256];
    std::snprintf(line, sizeof(line), "%d,%.5f,%.8e,%.8e,%.8e,%.3e\n", step,
                  step * args.params.dt, e.total, e.kinetic, e.potential, rel);
    elog << line;
  };

  std::vector<float4> frame_buf[2] = {std::vector<float4>(n), std::vector<float4>(n)};
  std::future<void> writing;  // the frame in flight on the writer thread
  auto wait_writer = [&]() {
    if (writing.valid()) writing.get();
  };
  int frame = 0;
  reset_force_timing();
  compute_forces(sys, d_acc, args.params);  // a(x0) for the first kick
  for (int step = 0; step < args.steps; ++step) {
    leapfrog_step(sys, d_acc, args.params);
    if (step % args.dump_every == 0) {
      std::vector<float4>& buf = frame_buf[frame & 1];
      wait_writer();  // the other buffer's frame; this one was written two dumps ago
      sys.download_positions(buf.data());
      writing = std::async(std::launch::async, dump_frame, std::cref(buf),
                           std::cref(args.out_dir), frame++);
      log_energy(step);
    }
  }
  log_energy(args.steps);
  wait_writer();
  if (!args.checkpoint.empty()) write_checkpoint(sys, args.checkpoint);

  double ms = force_kernel_ms_total();
  long calls = force_kernel_calls();
  std::printf("force pipeline: %ld calls, %.1f ms total, %.3f ms/call\n", calls, ms, ms / calls);
  if (bh) {
    bh_report_phase_timing();
  } else {
    double interactions = static_cast<double>(n) * n * calls;
    double gips = interactions / (ms * 1e-3) / 1e9;
    std::printf("throughput: %.2f Ginteractions/s (~%.1f GFLOP/s at 20 flop/pair)\n", gips,
                gips * 20.0);
  }
  std::printf("wrote %d frames to %s\n", frame, args.out_dir.c_str());
  (void)hipFree(d_acc);
  bh_release();
  sys.release();
  return 0;
}
