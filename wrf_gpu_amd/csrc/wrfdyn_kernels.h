// Launchers for the wrfdyn kernels and the
// optional per-kernel event timing used for benchmarking.
#pragma once

#include <string>
#include <utility>
#include <vector>

#include "wrfdyn_common.h"

namespace wd {

// Kernel ids for per-kernel timing (wd_profile_enable / wd_profile_read).
enum KernelId {
  KID_ORIGIN_COPY = 0,
  KID_RUV,
  KID_ROM,
  KID_PREP_COL,
  KID_TEND_U,
  KID_TEND_V,
  KID_TEND_W,
  KID_TEND_T,
  KID_PREP_2D,
  KID_PREP_FACES,
  KID_PREP_U,
  KID_PREP_V,
  KID_COL_SETUP,
  KID_UV,
  KID_COLUMN,     // LDS-tiled column kernel
  KID_COLUMN_MU,  // column march, advance_mu_t kernel
  KID_COLUMN_W,   // column march, advance_w + calc_p_rho kernel
  KID_FIN_UV,
  KID_FIN_COL,
  KID_DIFF6,
  KID_SMAG,
  KID_LIMITER,
  KID_CONSTK,
  KID_CORIOLIS,
  KID_MOIST,         // the moisture factors and the species' tendencies
  KID_MOIST_UPDATE,  // the species' large-step update after a stage
  KID_COUNT
};

const char* kernel_name(int id);

// Brackets selected launches with HIP events on the launch stream; the
// elapsed times are summed when read.  Disabled (mask 0) costs nothing.
struct Prof {
  unsigned long long mask = 0;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<std::pair<int, size_t>> recs;  // kernel id, index of the start event
  bool failed = false;

  bool on(int id) const { return (mask >> id) & 1ull; }
  hipEvent_t next() {
    if (used == pool.size()) {
      hipEvent_t e = nullptr;
      if (hipEventCreate(&e) != hipSuccess) {
        failed = true;
        return nullptr;
      }
      pool.push_back(e);
    }
    return pool[used++];
  }
  // returns a token (start index + 1) or 0 when not recording
  size_t begin(int id, hipStream_t st) {
    if (!on(id) || failed) return 0;
    size_t idx = used;
    hipEvent_t a = next();
    hipEvent_t b = next();
    if (!a || !b) return 0;
    if (hipEventRecord(a, st) != hipSuccess) {
      failed = true;
      return 0;
    }
    recs.emplace_back(id, idx);
    return idx + 1;
  }
  void end(size_t tok, hipStream_t st) {
    if (tok && hipEventRecord(pool[tok], st) != hipSuccess) failed = true;
  }
  ~Prof() {
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
  }
};

#define WD_TIMED(prof, id, st, launch)                      \
  do {                                                      \
    size_t _tok = (prof) ? (prof)->begin((id), (st)) : 0;  \
    launch;                                                 \
    if (prof) (prof)->end(_tok, (st));                      \
  } while (0)

// RK-stage setup: transport velocities, tendencies, small_step_prep,
// calc_p_rho step 0, rw_tend/ph_tend and calc_coef_w.
hipError_t launch_stage_setup(const DGrid& g, const DNamelist& nl, const DState& s, const DWork& w,
                              int rk_step, double dts, hipStream_t st, Prof* prof);

// Column-kernel tile shapes: TX columns x KT threads per column; each thread
// owns up to kMaxLevelsPerThread of the nz+1 levels, so a shape serves
// nz + 1 <= kt * kMaxLevelsPerThread.
struct ColumnTile {
  int tx, kt;
};
constexpr int kMaxLevelsPerThread = 8;
// column_tile value that selects the column-march kernel instead of a tile shape
constexpr int kColumnMarch = -1;
extern const ColumnTile kColumnTiles[];
extern const int kNumColumnTiles;

// One acoustic substep: advance_uv, then the fused column kernel with tile
// shape kColumnTiles[column_tile].  `first` marks a stage's first substep,
// which starts from t_2ave = 0 and pm1 = the step-0 pressure.
hipError_t launch_acoustic_substep(const DGrid& g, const DNamelist& nl, const DState& s, const DWork& w,
                                   double dts, int first, int column_tile, hipStream_t st, Prof* prof);

// The moisture's large-step update after the acoustic loop of a stage with
// transport on (moist_adv_opt != 0): q = (m_origin q0 + dt_rk qt) / m_new.
hipError_t launch_moist_update(const DGrid& g, const DState& s, const DWork& w, double dt_rk, hipStream_t st,
                               Prof* prof);

// small_step_finish + grid%p refresh.
hipError_t launch_stage_finish(const DGrid& g, const DState& s, const DWork& w, hipStream_t st, Prof* prof);

// One kernel of a plain-FP64 acoustic substep with the column-march pair:
// part 0 advance_uv (k_uv), 1 advance_mu_t + theta (k_column_mu), 2 advance_w
// + calc_p_rho (k_column_w).  The three in order are launch_acoustic_substep.
hipError_t launch_substep_part(int part, const DGrid& g, const DNamelist& nl, const DState& s, const DWork& w,
                               double dts, int first, hipStream_t st);

// ---- The persistent ensemble step (wrfdyn_persistent.h) ----
// One large step of every member of an ensemble from a single launch: a
// program of phases, each a set of kernel bodies that read nothing another
// body of the phase writes, run for every member, with a grid barrier
// between phases.  A segment is one kernel body for one member; its virtual
// blocks are the blocks of that kernel's standalone launch.
struct PMember {
  DGrid g;
  DNamelist nl;
  DState s;
  DWork w;
};
struct PSeg {
  int kind;      // PKind
  int member;
  int rk_step;   // k_prep_2d
  int first;     // the column kernels: a stage's first substep
  double dts;    // the stage's acoustic step
  unsigned vb0;  // the segment's first virtual block in its phase (a multiple of 8)
  unsigned nvb;  // its virtual blocks (256-thread workgroups)
  unsigned gx, gy;  // the standalone launch's grid
  unsigned aux;     // k_coriolis: its u-face block count
};
struct PPhase {
  unsigned seg0, nseg;  // segments [seg0, seg0 + nseg), sorted by vb0
  unsigned total;       // virtual blocks of the phase
};
// The grid barrier's state (device memory, zeroed at creation): arrivals,
// the generation every arrival waits to see change, the abort flag a timed
// out wait raises, and the bound on a wait in polls.
struct PBarrier {
  unsigned count, gen, abort, spin_limit;
};
// The program of one large step of the members (all of one precision form
// and one substep count; no optional term but rotation).  column_tiles holds
// each member's acoustic column kernel (kColumnMarch or a kColumnTiles index):
// the 16 x 16 tile runs as such for up to 63 levels in the FP64 form, any
// other choice as the column march, which gives the same bits.  Returns false
// with a message for a member it cannot run.
bool persistent_program(const std::vector<PMember>& members, const std::vector<int>& column_tiles,
                        std::vector<PSeg>* segs, std::vector<PPhase>* phases, std::string* why);
// Levels per thread with which the persistent step runs a member's column
// tile, 0 where it runs the column march.
int persistent_column_levels(const DGrid& g, const DNamelist& nl, int column_tile);
// Workgroups of the persistent kernel one CU holds at once.  `tile`: the
// FP64 kernel that also carries the column tile, for a program with a
// PK_COLUMN_TILE segment.
hipError_t persistent_blocks_per_cu(bool mixed, bool tile, int* out);
hipError_t launch_persistent_step(bool mixed, bool tile, unsigned grid, const PMember* members, const PSeg* segs,
                                  const PPhase* phases, int n_phases, PBarrier* bar, hipStream_t st);

// Adjoint of one plain-FP64 large step (wrfdyn_adjoint.hip).  The tape is
// allocated on first use and kept by the caller; the forward recompute runs
// in w and the state scratch of the tape.
struct AdjTape;
hipError_t adjoint_tape_create(const DGrid& g, const DNamelist& nl, AdjTape** out, size_t* bytes);
void adjoint_tape_destroy(AdjTape* t);
hipError_t adjoint_large_step(const DGrid& g, const DNamelist& nl, const DWork& w, AdjTape* t, const DState& start,
                              const DState& bar, hipStream_t st);

}  // namespace wd
