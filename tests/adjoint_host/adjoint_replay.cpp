// Host replay of the GPU adjoint's reverse sweep (wd_step_adjoint), for
// checking its launch glue on the CPU: the tape layout and views
// (tape_slots, make_view), the order and grid sizes of the reverse launches
// (reverse_sequence) and every kernel's per-thread index map (rev_thread),
// run for every thread of every block, tail threads included, exactly as
// wrfdyn_adjoint.hip launches them.  The only difference from the GPU is the
// sum of shared bars (a plain add, serial, instead of atomic adds).
//
// Each tape array is its own heap allocation, so a build with
// AddressSanitizer reports any access outside any of them.  Test
// infrastructure: tests/adjoint_host/replay.py fills the tape from the
// oracle's traced step and compares the result with the gradient oracle.
//
//   adjoint_replay layout                     print the tape's array order
//   adjoint_replay run <in> <out> <scalars>   replay one large step's reverse
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "wrfdyn_adjoint_glue.h"

using namespace wd;
using namespace wd::adjg;

namespace {

struct PlainAdd {
  void operator()(double* p, double v) const { *p += v; }
};

struct HostLaunch {
  long launches = 0;
  bool operator()(AK k, unsigned nb, const RevArgs& a) {
    for (size_t t = 0; t < (size_t)nb * kBs; ++t) rev_thread(k, a, t, PlainAdd{});
    ++launches;
    return true;
  }
};

[[noreturn]] void die(const char* msg) {
  std::fprintf(stderr, "adjoint_replay: %s\n", msg);
  std::exit(2);
}

std::vector<double> read_all(const char* path) {
  FILE* f = std::fopen(path, "rb");
  if (!f) die("cannot open the input file");
  std::fseek(f, 0, SEEK_END);
  const long bytes = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  std::vector<double> v((size_t)bytes / sizeof(double));
  if (std::fread(v.data(), sizeof(double), v.size(), f) != v.size()) die("short read");
  std::fclose(f);
  return v;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc >= 2 && std::strcmp(argv[1], "layout") == 0) {
    std::printf("stage_state");
    for (int a : kStageState) std::printf(" %d", a);
    std::printf("\nstage_const");
    for (int a : kStageConst) std::printf(" %d", a);
    std::printf("\ncarried");
    for (int a : kCarried) std::printf(" %d", a);
    std::printf("\ncount %d\n", (int)A_count);
    return 0;
  }
  if (argc != 20 || std::strcmp(argv[1], "run") != 0)
    die("usage: run <in> <out> nx ny nz ns dx dy dt epssm top_lid w_damping damp_opt dampcoef zdamp cf1 cf2 cf3");
  const int nx = std::atoi(argv[4]), ny = std::atoi(argv[5]), nz = std::atoi(argv[6]), ns = std::atoi(argv[7]);
  DGrid g{};
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.dx = std::atof(argv[8]);
  g.dy = std::atof(argv[9]);
  g.rdx = 1.0 / g.dx;
  g.rdy = 1.0 / g.dy;
  DNamelist nl{};
  nl.ns = ns;
  nl.dt = std::atof(argv[10]);
  nl.epssm = std::atof(argv[11]);
  nl.top_lid = std::atoi(argv[12]);
  nl.w_damping = std::atoi(argv[13]);
  nl.damp_opt = std::atoi(argv[14]);
  nl.dampcoef = std::atof(argv[15]);
  nl.zdamp = std::atof(argv[16]);
  g.cf1 = std::atof(argv[17]);
  g.cf2 = std::atof(argv[18]);
  g.cf3 = std::atof(argv[19]);
  const std::vector<double> in = read_all(argv[2]);
  size_t pos = 0;
  std::vector<std::vector<double>> keep;
  auto take = [&](size_t n) -> double* {
    if (pos + n > in.size()) die("input too short");
    keep.emplace_back(in.begin() + (long)pos, in.begin() + (long)(pos + n));
    pos += n;
    return keep.back().data();
  };
  keep.reserve(4096);
  const size_t C = (size_t)ny * nx, CU = (size_t)ny * (nx + 1), CV = (size_t)(ny + 1) * nx;
  g.c1h = take(nz);
  g.c2h = take(nz);
  g.c1f = take(nz + 1);
  g.c2f = take(nz + 1);
  g.dnw = take(nz);
  g.rdnw = take(nz);
  g.dn = take(nz);
  g.rdn = take(nz);
  g.fnm = take(nz);
  g.fnp = take(nz);
  g.msftx = take(C);
  g.msfty = take(C);
  g.msfux = take(CU);
  g.msfuy = take(CU);
  g.msfvx = take(CV);
  g.msfvy = take(CV);
  const double dn_top = g.dn[nz - 1];
  const double dn_safe = std::fabs(dn_top) > 1.0e-30 ? dn_top : 1.0;
  g.cfn = (0.5 * g.dnw[nz - 1] + g.dn[nz - 1]) / dn_safe;
  g.cfn1 = -0.5 * g.dnw[nz - 1] / dn_safe;
  // the tape: every array its own allocation, filled in tape_slots order
  TapePtrs t;
  const std::vector<TapeSlot> slots = tape_slots(g, nl, &t);
  std::vector<double*> owned;
  for (const TapeSlot& s : slots) {
    double* p = static_cast<double*>(std::calloc(s.n ? s.n : 1, sizeof(double)));
    if (!p) die("out of memory");
    owned.push_back(p);
    *s.p = p;
  }
  int nsub[3];
  const int nb = n_boundaries(nl, nsub);
  auto fill = [&](double* dst, int arr) {
    const size_t n = ext_len(g, kExtent[arr]);
    std::memcpy(dst, take(n), n * sizeof(double));
  };
  for (int s = 0; s < 3; ++s) {
    for (int a = 0; a < kNStageState; ++a) fill(t.stage_state[s][a], kStageState[a]);
    for (int a = 0; a < kNStageConst; ++a) fill(t.stage_const[s][a], kStageConst[a]);
  }
  for (int b = 0; b < nb; ++b)
    for (int a = 0; a < kNCarried; ++a) fill(t.carried[b][a], kCarried[a]);
  for (int a = 0; a <= A_mudf; ++a) fill(t.bar[a], a);
  if (pos != in.size()) die("input has trailing data");
  HostLaunch L;
  if (!reverse_sequence(g, nl, t, L)) die("reverse_sequence stopped");
  FILE* f = std::fopen(argv[3], "wb");
  if (!f) die("cannot open the output file");
  for (int a = 0; a <= A_mudf; ++a) {
    const size_t n = ext_len(g, kExtent[a]);
    if (std::fwrite(t.bar[a], sizeof(double), n, f) != n) die("short write");
  }
  std::fclose(f);
  std::printf("launches %ld\n", L.launches);
  for (double* p : owned) std::free(p);
  return 0;
}
