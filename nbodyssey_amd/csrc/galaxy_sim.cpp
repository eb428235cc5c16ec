// galaxy_sim -- the simulation command line over the C++ API of libnbody.so.
//
// Interface kept identical to the reference CLI (src/main.cu:18-215) so that
// its scripts drive this binary unchanged (gpu_bench.sh:31,41-44 parse the
// output): the flags and defaults (--ic --out --energy-log --steps
// --dump-every --dt --eps --G --theta --compare-forces --force naive|bh
// --traverse thread|warp), the error messages and exit status 1, the ic.bin /
// frame / energy-CSV formats (docs/FORMATS.md, main.cu:114-179), the stdout
// report lines, the force-call count (a(x0) plus one per step) and the log
// cadence (energy row k written after step k, every dump_every steps, plus a
// final row).
//
// Additive flags: --keys 30|63, --group-width 16|32|64, --compensated,
// --far16, --energy-theta f (energy from the tree potential at that theta:
// the O(n log n) cost policy for large n), --checkpoint path (the final
// positions and velocities in the ic.bin layout; --ic path resumes).
//
// Structure: a flag table drives the parser; one snapshot layer reads and
// writes every particle file (ic.bin, frames, checkpoints); frames go to disk
// on a writer thread with two host buffers while the steps continue; the
// step loop runs from one dump point to the next.
#include <cctype>
#include <cerrno>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <initializer_list>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

#include "nbody/forces.hpp"
#include "nbody/integrator.hpp"
#include "nbody/simulation.hpp"

namespace fs = std::filesystem;

namespace {

[[noreturn]] void fail(const std::string& msg) {
  std::fprintf(stderr, "%s\n", msg.c_str());
  std::exit(EXIT_FAILURE);
}

void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) fail(std::string(what) + ": " + hipGetErrorString(e));
}

// ------------------------------------------------------------------ options

struct RunConfig {
  std::string ic_path;
  std::string out_dir = "frames";
  std::string energy_log = "benchmarks/energy_gpu.csv";
  std::string checkpoint;  // empty: none
  int steps = 1000;
  int dump_every = 5;
  bool compare = false;
  float energy_theta = 0.f;  // 0: exact O(n^2) energy
  SimParams params{/*dt=*/0.01f, /*epsilon=*/0.05f, /*G=*/1.0f};
};

const char kUsage[] =
    "usage: galaxy_sim --ic ic.bin [--steps N] [--dt f] [--eps f] [--dump-every N] "
    "[--out dir] [--force naive|bh] [--theta f] [--traverse thread|warp] "
    "[--compare-forces] [--keys 30|63] [--group-width 16|32|64] [--compensated] "
    "[--far16] [--energy-theta f] [--checkpoint path]";

// Whole-string numeric conversions; anything left over is an error.
template <typename T>
T number(const std::string& flag, const std::string& text) {
  const char* s = text.c_str();
  char* end = nullptr;
  errno = 0;
  T v{};
  if (std::is_same<T, int>::value) {
    const long x = std::strtol(s, &end, 10);
    if (x < INT32_MIN || x > INT32_MAX) errno = ERANGE;
    v = static_cast<T>(x);
  } else {
    v = static_cast<T>(std::strtof(s, &end));
  }
  if (text.empty() || errno == ERANGE || end != s + text.size() ||
      std::isspace(static_cast<unsigned char>(text[0])))
    fail("bad value for " + flag + ": " + text);
  return v;
}

// One of a fixed set of spellings; `spelled` lists them for the message.
template <typename T, size_t N>
T choice(const std::string& flag, const std::string& text,
         const std::pair<const char*, T> (&options)[N], const char* spelled) {
  for (const auto& o : options)
    if (text == o.first) return o.second;
  fail(flag + " must be " + spelled + ", got: " + text);
}

using Apply = void (*)(RunConfig&, const std::string& flag, const std::string& value);

struct Flag {
  const char* name;
  bool takes_value;
  Apply apply;
};

const std::pair<const char*, ForceMethod> kForces[] = {{"naive", ForceMethod::kNaive},
                                                       {"bh", ForceMethod::kBarnesHut}};
const std::pair<const char*, BhTraversal> kWalks[] = {{"thread", BhTraversal::kThread},
                                                      {"warp", BhTraversal::kWarp}};
const std::pair<const char*, int> kKeyWidths[] = {{"30", 30}, {"63", 63}};
const std::pair<const char*, int> kGroupWidths[] = {{"16", 16}, {"32", 32}, {"64", 64}};

const Flag kFlags[] = {
    {"--ic", true, [](RunConfig& c, const std::string&, const std::string& v) { c.ic_path = v; }},
    {"--out", true, [](RunConfig& c, const std::string&, const std::string& v) { c.out_dir = v; }},
    {"--energy-log", true,
     [](RunConfig& c, const std::string&, const std::string& v) { c.energy_log = v; }},
    {"--steps", true,
     [](RunConfig& c, const std::string& f, const std::string& v) { c.steps = number<int>(f, v); }},
    {"--dump-every", true,
     [](RunConfig& c, const std::string& f, const std::string& v) {
       c.dump_every = number<int>(f, v);
     }},
    {"--dt", true,
     [](RunConfig& c, const std::string& f, const std::string& v) {
       c.params.dt = number<float>(f, v);
     }},
    {"--eps", true,
     [](RunConfig& c, const std::string& f, const std::string& v) {
       c.params.epsilon = number<float>(f, v);
     }},
    {"--G", true,
     [](RunConfig& c, const std::string& f, const std::string& v) {
       c.params.G = number<float>(f, v);
     }},
    {"--theta", true,
     [](RunConfig& c, const std::string& f, const std::string& v) {
       c.params.theta = number<float>(f, v);
     }},
    {"--compare-forces", false,
     [](RunConfig& c, const std::string&, const std::string&) { c.compare = true; }},
    {"--force", true,
     [](RunConfig& c, const std::string& f, const std::string& v) {
       c.params.force = choice(f, v, kForces, "naive or bh");
     }},
    {"--traverse", true,
     [](RunConfig& c, const std::string& f, const std::string& v) {
       c.params.traverse = choice(f, v, kWalks, "thread or warp");
     }},
    {"--keys", true,
     [](RunConfig& c, const std::string& f, const std::string& v) {
       c.params.key_bits = choice(f, v, kKeyWidths, "30 or 63");
     }},
    {"--group-width", true,
     [](RunConfig& c, const std::string& f, const std::string& v) {
       c.params.group_width = choice(f, v, kGroupWidths, "16, 32 or 64");
     }},
    {"--compensated", false,
     [](RunConfig& c, const std::string&, const std::string&) { c.params.compensated = true; }},
    {"--far16", false,
     [](RunConfig& c, const std::string&, const std::string&) { c.params.far16 = true; }},
    {"--checkpoint", true,
     [](RunConfig& c, const std::string&, const std::string& v) { c.checkpoint = v; }},
    {"--energy-theta", true,
     [](RunConfig& c, const std::string& f, const std::string& v) {
       c.energy_theta = number<float>(f, v);
       if (!(c.energy_theta >= 0.f && c.energy_theta < 0.577f))
         fail("--energy-theta must be in [0, 0.577)");
     }},
};

RunConfig read_command_line(int argc, char** argv) {
  RunConfig c;
  for (int i = 1; i < argc; ++i) {
    const std::string flag = argv[i];
    const Flag* f = nullptr;
    for (const Flag& k : kFlags)
      if (flag == k.name) f = &k;
    if (!f) fail("unknown argument: " + flag);
    std::string value;
    if (f->takes_value) {
      if (i + 1 >= argc) fail("missing value for " + flag);
      value = argv[++i];
    }
    f->apply(c, flag, value);
  }
  if (c.ic_path.empty()) fail(kUsage);
  if (c.dump_every <= 0) fail("--dump-every must be positive");
  return c;
}

// ---------------------------------------------------------- particle files
// docs/FORMATS.md: little-endian int32 n, then n float4 positions and, in
// ic.bin and checkpoints, n float4 velocities.  Frames hold positions only.

struct Snapshot {
  std::vector<float4> pos, vel;
};

Snapshot read_snapshot(const std::string& path) {
  std::ifstream in(path, std::ios::binary | std::ios::ate);
  if (!in) fail("cannot open IC file: " + path);
  const std::streamoff bytes = in.tellg();
  in.seekg(0);
  std::int32_t n = 0;
  if (bytes < (std::streamoff)sizeof n || !in.read(reinterpret_cast<char*>(&n), sizeof n) ||
      n <= 0)
    fail("bad particle count in " + path);
  const std::streamoff body = 2 * (std::streamoff)n * (std::streamoff)sizeof(float4);
  if (bytes - (std::streamoff)sizeof n < body) fail("truncated IC file: " + path);
  Snapshot s;
  s.pos.resize(n);
  s.vel.resize(n);
  in.read(reinterpret_cast<char*>(s.pos.data()), n * sizeof(float4));
  in.read(reinterpret_cast<char*>(s.vel.data()), n * sizeof(float4));
  if (!in) fail("truncated IC file: " + path);
  return s;
}

// Writes the count and then each array whole; false if the file could not
// be written.
bool write_snapshot(const fs::path& path, std::int32_t n,
                    std::initializer_list<const float4*> arrays) {
  std::ofstream out(path, std::ios::binary);
  out.write(reinterpret_cast<const char*>(&n), sizeof n);
  for (const float4* a : arrays) out.write(reinterpret_cast<const char*>(a), n * sizeof(float4));
  return static_cast<bool>(out);
}

fs::path frame_path(const std::string& dir, int frame) {
  char name[32];
  std::snprintf(name, sizeof name, "frame_%05d.bin", frame);
  return fs::path(dir) / name;
}

// Frames are written on a worker thread.  The step loop copies positions
// into the idle one of two host buffers and hands it over; before reusing a
// buffer it waits until the worker has written the frame it held.
class FrameWriter {
 public:
  FrameWriter(std::string dir, int n) : dir_(std::move(dir)), worker_([this] { run(); }) {
    buf_[0].resize(n);
    buf_[1].resize(n);
  }
  ~FrameWriter() {
    {
      std::lock_guard<std::mutex> lock(m_);
      stop_ = true;
    }
    cv_.notify_all();
    worker_.join();
  }
  // Snapshot of the current positions as the next frame.
  void dump(const ParticleSystem& sys) {
    const int slot = frames_ & 1;
    std::unique_lock<std::mutex> lock(m_);
    cv_.wait(lock, [&] { return !pending_[slot]; });
    lock.unlock();
    sys.download_positions(buf_[slot].data());
    lock.lock();
    pending_[slot] = true;
    index_[slot] = frames_++;
    cv_.notify_all();
  }
  void finish() {
    std::unique_lock<std::mutex> lock(m_);
    cv_.wait(lock, [&] { return !pending_[0] && !pending_[1]; });
  }
  int frames() const { return frames_; }

 private:
  void run() {
    int next = 0;  // buffers are written in the order they were filled
    std::unique_lock<std::mutex> lock(m_);
    for (;;) {
      cv_.wait(lock, [&] { return stop_ || pending_[next]; });
      if (!pending_[next]) return;  // stopping with nothing left to write
      lock.unlock();
      const std::vector<float4>& b = buf_[next];
      write_snapshot(frame_path(dir_, index_[next]), static_cast<std::int32_t>(b.size()),
                     {b.data()});  // unchecked, like the reference's frame writes
      lock.lock();
      pending_[next] = false;
      next ^= 1;
      cv_.notify_all();
    }
  }
  std::string dir_;
  std::vector<float4> buf_[2];
  bool pending_[2] = {false, false};
  int index_[2] = {0, 0};
  int frames_ = 0;
  bool stop_ = false;
  std::mutex m_;
  std::condition_variable cv_;
  std::thread worker_;
};

// ------------------------------------------------------------- energy log
// CSV "step,time,energy,kinetic,potential,rel_error" (main.cu:160,176),
// rel_error against the energy measured before the first step.
class EnergyLog {
 public:
  explicit EnergyLog(const std::string& path) : out_(path) {
    out_ << "step,time,energy,kinetic,potential,rel_error\n";
  }
  void row(int step, float dt, const Energy& e, const Energy& e0) {
    char line[256];
    const double rel = (e.total - e0.total) / std::abs(e0.total);
    std::snprintf(line, sizeof line, "%d,%.5f,%.8e,%.8e,%.8e,%.3e\n", step, step * dt, e.total,
                  e.kinetic, e.potential, rel);
    out_ << line;
  }

 private:
  std::ofstream out_;
};

// ---------------------------------------------------------- force gate
struct RelError {
  double max = 0.0, mean = 0.0;
};

// |b - a| / (|a| + 1e-30) per particle, a the reference (main.cu:94-99).
RelError relative_error(const float4* a, const float4* b, int n) {
  RelError r;
  for (int i = 0; i < n; ++i) {
    const double d[3] = {(double)a[i].x - b[i].x, (double)a[i].y - b[i].y,
                         (double)a[i].z - b[i].z};
    const double ref = std::sqrt((double)a[i].x * a[i].x + (double)a[i].y * a[i].y +
                                 (double)a[i].z * a[i].z);
    const double e = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) / (ref + 1e-30);
    r.mean += e;
    r.max = e > r.max ? e : r.max;
  }
  r.mean /= n;
  return r;
}

// --compare-forces: both force modules on the loaded state, the naive one as
// the reference, into one device array of 2n accelerations.
int compare_force_modules(const ParticleSystem& sys, const SimParams& base) {
  const int n = sys.n;
  float4* d_acc = nullptr;
  check_hip(hipMalloc(reinterpret_cast<void**>(&d_acc), 2 * (size_t)n * sizeof(float4)),
            "allocating the comparison accelerations");
  SimParams p = base;
  p.force = ForceMethod::kNaive;
  compute_forces(sys, d_acc, p);
  p.force = ForceMethod::kBarnesHut;
  compute_forces(sys, d_acc + n, p);
  std::vector<float4> both(2 * (size_t)n);
  check_hip(hipMemcpy(both.data(), d_acc, both.size() * sizeof(float4), hipMemcpyDeviceToHost),
            "copying the comparison accelerations");
  (void)hipFree(d_acc);
  const RelError err = relative_error(both.data(), both.data() + n, n);
  std::printf("naive vs barnes-hut, theta=%.3g, traverse=%s, n=%d\n", base.theta,
              base.traverse == BhTraversal::kWarp ? "warp" : "thread", n);
  std::printf("relative force error: max %.3e, mean %.3e\n", err.max, err.mean);
  bh_release();
  return 0;
}

void print_run_header(const RunConfig& c, int n, const Energy& e0) {
  const SimParams& p = c.params;
  const bool bh = p.force == ForceMethod::kBarnesHut;
  std::printf("n=%d dt=%.4g eps=%.4g steps=%d force=%s", n, p.dt, p.epsilon, c.steps,
              bh ? "barnes-hut" : "naive");
  if (bh)
    std::printf(" theta=%.3g traverse=%s", p.theta,
                p.traverse == BhTraversal::kWarp ? "warp" : "thread");
  std::printf("\nE0=%.6e KE0=%.6e PE0=%.6e\n", e0.total, e0.kinetic, e0.potential);
}

void print_force_report(const SimParams& p, int n) {
  const double ms = force_kernel_ms_total();
  const long calls = force_kernel_calls();
  std::printf("force pipeline: %ld calls, %.1f ms total, %.3f ms/call\n", calls, ms, ms / calls);
  if (p.force == ForceMethod::kBarnesHut) {
    bh_report_phase_timing();
    return;
  }
  const double gint = (double)n * n * calls / (ms * 1e-3) / 1e9;  // billion interactions/s
  std::printf("throughput: %.2f Ginteractions/s (~%.1f GFLOP/s at 20 flop/pair)\n", gint,
              gint * 20.0);
}

}  // namespace

int main(int argc, char** argv) {
  const RunConfig cfg = read_command_line(argc, argv);
  Snapshot ic = read_snapshot(cfg.ic_path);
  const int n = static_cast<int>(ic.pos.size());

  ParticleSystem sys;
  sys.allocate(n);
  sys.upload(ic.pos.data(), ic.vel.data());
  if (cfg.compare) return compare_force_modules(sys, cfg.params);

  float4* d_acc = nullptr;
  check_hip(hipMalloc(reinterpret_cast<void**>(&d_acc), n * sizeof(float4)),
            "allocating the acceleration array");
  fs::create_directories(cfg.out_dir);
  const fs::path log_path(cfg.energy_log);
  if (log_path.has_parent_path()) fs::create_directories(log_path.parent_path());
  EnergyLog elog(cfg.energy_log);

  const SimParams& p = cfg.params;
  auto measure = [&] {
    return cfg.energy_theta > 0.f ? total_energy_tree(sys, p, cfg.energy_theta)
                                  : total_energy(sys, p);
  };
  const Energy e0 = measure();
  print_run_header(cfg, n, e0);

  {
    FrameWriter frames(cfg.out_dir, n);
    reset_force_timing();
    compute_forces(sys, d_acc, p);  // a(x0) for the first kick
    // Dump points are the steps s = 0, D, 2D, ... < steps: after step s has
    // run, its frame and energy row are written.
    int done = 0;
    auto run_until = [&](int count) {
      for (; done < count; ++done) leapfrog_step(sys, d_acc, p);
    };
    for (int s = 0; s < cfg.steps; s += cfg.dump_every) {
      run_until(s + 1);
      frames.dump(sys);
      elog.row(s, p.dt, measure(), e0);
      if (s > cfg.steps - cfg.dump_every) break;  // the next dump point would overflow
    }
    run_until(cfg.steps);
    elog.row(cfg.steps, p.dt, measure(), e0);
    frames.finish();
    if (!cfg.checkpoint.empty()) {
      Snapshot last;
      last.pos.resize(n);
      last.vel.resize(n);
      sys.download_positions(last.pos.data());
      check_hip(hipMemcpy(last.vel.data(), sys.d_vel, n * sizeof(float4), hipMemcpyDeviceToHost),
                "copying the checkpoint velocities");
      const fs::path ck(cfg.checkpoint);
      if (ck.has_parent_path()) fs::create_directories(ck.parent_path());
      if (!write_snapshot(ck, n, {last.pos.data(), last.vel.data()}))
        fail("cannot write checkpoint: " + cfg.checkpoint);
    }
    print_force_report(p, n);
    std::printf("wrote %d frames to %s\n", frames.frames(), cfg.out_dir.c_str());
  }
  (void)hipFree(d_acc);
  bh_release();
  sys.release();
  return 0;
}
