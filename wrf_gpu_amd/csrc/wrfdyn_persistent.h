// The persistent ensemble step: one large step of every member of an
// ensemble from a single launch (wd_batch_create_persistent).
//
// The host builds a program for one large step: phases in wd_step's order,
// each a set of kernel bodies that read nothing another body of the phase
// writes, run for every member.  k_step_persistent is launched with as many
// 256-thread workgroups as the GPU holds at once.  In each phase the
// workgroups take the phase's virtual blocks in turn; a virtual block is one
// block of the standalone launch of one kernel for one member, and the
// workgroup runs that kernel's body on it with the block index, grid size
// and thread layout the standalone launch would have given.  So every value
// is computed by the same code in the same order as wd_step, and the result
// is the same bits.  Between phases the workgroups meet at a grid barrier.
//
// Included once, after wrfdyn_stage.hip and wrfdyn_acoustic.hip, by
// wrfdyn_device.hip: the bodies live in those files.
#pragma once

#include <algorithm>
#include <type_traits>

namespace wd {

enum PKind : int {
  PK_ORIGIN,
  PK_RUV,
  PK_ROM,
  PK_PREP_2D,
  PK_PREP_FACES,
  PK_PREP_COL,
  PK_PREP_U,
  PK_PREP_V,
  PK_TEND_U,
  PK_TEND_V,
  PK_TEND_W,
  PK_TEND_T,
  PK_CORIOLIS,
  PK_COL_SETUP,
  PK_UV,
  PK_COLUMN_MU,
  PK_COLUMN_W,
  PK_FIN_UV,
  PK_FIN_COL,
  PK_COLUMN_TILE,
};

namespace {

// The LDS-tiled acoustic column kernel the persistent step carries: the
// 16 x 16 shape (16 columns, 16 threads per column, 256 threads), with up to
// kPersistTileLevels levels per thread, so for nz + 1 <= 64 w levels.  Its LDS
// is six (nz + 1) x 16 planes and three rows of 16.
constexpr int kPersistTile = 0;  // index of the 16 x 16 shape in kColumnTiles
constexpr int kPersistTileTx = 16, kPersistTileKt = 16, kPersistTileLevels = 4;
constexpr int kPersistTileLds = (6 * kPersistTileKt * kPersistTileLevels + 3) * kPersistTileTx;

// The tile bodies share one LDS buffer, the size of the largest (k_tend_v's);
// the kernel that also carries the column tile sizes it for that tile's
// deepest grid.
constexpr int kPersistLds = 2 * kTendURows * kTendVFields * 64;
constexpr int kPersistLdsTile = std::max(kPersistLds, kPersistTileLds);
static_assert(kPersistTileTx * kPersistTileKt == 256, "the column tile runs on a 256-thread workgroup");
static_assert(2 * kTendURows * 2 * 64 <= kPersistLds, "k_tend_u / k_tend_t tile");
static_assert(2 * kColRows * 2 * 64 <= kPersistLds, "k_col_setup tile");
static_assert(2 * kMuRows * 2 * 64 <= kPersistLds, "k_column_mu tile");
static_assert(2 * kUvBy * 7 * kUvBx <= kPersistLds, "k_uv tile");
static_assert(64 * kTendURows == 256 && 64 * kColRows == 256 && 64 * kMuRows == 256 && kUvBx * kUvBy == 256,
              "every tile body runs on a 256-thread workgroup");

// wd_step's step-origin copies (u0 = u, v0 = v, w0 = w, th0 = theta,
// ph0 = ph', mu0 = mu'): a virtual block copies kOriginChunk runs of 256
// elements of each array.
constexpr unsigned kOriginChunk = 8;
__device__ inline void origin_copy(const PMember& M, size_t e) {
  const DGrid& g = M.g;
  const size_t C = (size_t)g.ny * g.nx, Mn = C * g.nz, Wn = C * (g.nz + 1);
  const size_t U = (size_t)g.nz * g.ny * (g.nx + 1), V = (size_t)g.nz * (g.ny + 1) * g.nx;
  if (e < U) M.w.u0[e] = M.s.u[e];
  if (e < V) M.w.v0[e] = M.s.v[e];
  if (e < Wn) {
    M.w.w0[e] = M.s.w[e];
    M.w.ph0[e] = M.s.ph_pert[e];
  }
  if (e < Mn) M.w.th0[e] = M.s.theta[e];
  if (e < C) M.w.mu0[e] = M.s.mu_pert[e];
}

// Virtual block `local` of segment S, on the thread `tid` of a 256-thread
// workgroup.
template <bool MX, bool TILE>
__device__ __forceinline__ void run_segment(const PSeg& S, const PMember& M, unsigned local, unsigned tid,
                                            double* lds) {
  using TT = typename std::conditional<MX, float, double>::type;
  // 1-D and 2-D grids of 256-thread blocks, and 64 x 4 tiles
  const Blk flat{local % S.gx, local / S.gx, S.gx, tid, 0u, 256u};
  const Blk tile{local, 0u, S.gx, tid & 63u, tid >> 6, 64u};
  switch (S.kind) {
    case PK_ORIGIN:
      for (unsigned r = 0; r < kOriginChunk; ++r) origin_copy(M, ((size_t)local * kOriginChunk + r) * 256u + tid);
      break;
    case PK_RUV:
      b_ruv(flat, M.g, M.s.u, M.s.v, M.s.mu_total, M.w.ru, M.w.rv);
      break;
    case PK_ROM: {
      // 64-thread blocks: each wave of the workgroup is one
      const Blk wave{4u * local + (tid >> 6), 0u, S.gx, tid & 63u, 0u, 64u};
      if (wave.x < S.gx) b_rom(wave, M.g, M.w.ru, M.w.rv, M.w.rom);
      break;
    }
    case PK_PREP_2D:
      b_prep_2d(flat, M.g, M.s, M.w, S.rk_step);
      break;
    case PK_PREP_FACES:
      b_prep_faces(flat, M.g, M.w);
      break;
    case PK_PREP_COL:
      b_prep_col(flat, M.g, M.s, M.w);
      break;
    case PK_PREP_U:
      b_prep_u(flat, M.g, M.s, M.w);
      break;
    case PK_PREP_V:
      b_prep_v(flat, M.g, M.s, M.w);
      break;
    case PK_TEND_U:
      b_tend_u<TT, false>(tile, lds, M.g, M.s, M.w);
      break;
    case PK_TEND_V:
      b_tend_v<TT, false>(tile, lds, M.g, M.s, M.w);
      break;
    case PK_TEND_W:
      b_tend_w<TT>(flat, M.g, M.s, M.w);
      break;
    case PK_TEND_T:
      b_tend_t<TT>(tile, lds, M.g, M.w, M.s.theta, kT0, M.w.tht, M.w.d6t, M.w.smt);
      break;
    case PK_CORIOLIS:
      b_coriolis<TT>(flat, M.g, M.s, M.w, S.aux);
      break;
    case PK_COL_SETUP:
      b_col_setup<TT, false>(tile, lds, M.g, M.nl, M.s, M.w, S.dts);
      break;
    case PK_UV:
      b_uv<kUvBx, kUvBy, TT, false>(tile, lds, M.g, M.s, M.w, S.dts);
      break;
    case PK_COLUMN_MU:
      b_column_mu<TT>(tile, lds, M.g, M.nl, M.s, M.w, S.dts, S.first);
      break;
    case PK_COLUMN_W:
      b_column_w<TT>(flat, M.g, M.nl, M.s, M.w, S.dts, S.first);
      break;
    case PK_FIN_UV:
      b_fin_uv(flat, M.g, M.s, M.w);
      break;
    case PK_FIN_COL:
      b_fin_col(flat, M.g, M.s, M.w);
      break;
    case PK_COLUMN_TILE:
      // FP64 only (the host never emits it for the mixed form); aux = levels per thread
      if constexpr (!MX && TILE) {
        constexpr int TX = kPersistTileTx, KT = kPersistTileKt;
        const int tx = (int)(tid % TX), kt = (int)(tid / TX);
        const unsigned t = tile_id(M.g, flat);
        switch (S.aux) {
          case 1: b_column<TX, KT, 1>(tx, kt, t, lds, M.g, M.nl, M.s, M.w, S.dts, S.first); break;
          case 2: b_column<TX, KT, 2>(tx, kt, t, lds, M.g, M.nl, M.s, M.w, S.dts, S.first); break;
          case 3: b_column<TX, KT, 3>(tx, kt, t, lds, M.g, M.nl, M.s, M.w, S.dts, S.first); break;
          case 4: b_column<TX, KT, 4>(tx, kt, t, lds, M.g, M.nl, M.s, M.w, S.dts, S.first); break;
          default: break;
        }
      }
      break;
    default:
      break;
  }
}

// s_waitcnt vmcnt(0) in the gfx9 encoding (vmcnt 0, expcnt 7, lgkmcnt 15):
// every vector memory access the wave issued, stores included, has completed.
constexpr int kWaitVmcnt0 = 0x0F70;

// The grid barrier.  Each wave waits for its own stores of the phase; the
// workgroup's first thread then arrives with a device-scope release, which
// writes its XCD's L2 back (each XCD has its own L2), and the last to
// arrive resets the count and advances the generation the others poll.
// Polls are relaxed device-scope loads; one device-scope acquire after the
// wait drops the stale lines from the CU's L1 and the XCD's L2.  A wait is
// bounded (spin_limit polls): one that runs out raises the abort flag, which
// every other wait and the next launches see, so each workgroup returns and
// the host reports the failure; no wait lasts for ever.
__device__ bool grid_barrier(PBarrier* bar, int* ok) {
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
  __syncthreads();
  if (threadIdx.x == 0) {
    int good = 1;
    const unsigned gen = __hip_atomic_load(&bar->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned arrived = __hip_atomic_fetch_add(&bar->count, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    if (arrived + 1u == gridDim.x) {
      __hip_atomic_store(&bar->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&bar->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      // the budget is checked before each poll, so a budget of 0 gives up at once
      const unsigned limit = __hip_atomic_load(&bar->spin_limit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      for (unsigned polls = 0;; ++polls) {
        if (__hip_atomic_load(&bar->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u || polls >= limit) {
          good = 0;
          break;
        }
        if (__hip_atomic_load(&bar->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != gen) break;
        __builtin_amdgcn_s_sleep(2);
      }
      if (!good) __hip_atomic_store(&bar->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    *ok = good;
  }
  __syncthreads();
  return *ok != 0;
}

// One large step: the phases in order, the workgroup taking every
// gridDim.x-th virtual block of each, with a grid barrier between phases.
template <bool MX, bool TILE>
__device__ __forceinline__ void step_phases(const PMember* __restrict__ members, const PSeg* __restrict__ segs,
                                            const PPhase* __restrict__ phases, int n_phases, PBarrier* bar,
                                            double* lds, int* ok) {
  if (threadIdx.x == 0)
    *ok = __hip_atomic_load(&bar->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u;
  __syncthreads();
  if (!*ok) return;  // an earlier launch timed out: the host must see it first
  const unsigned tid = threadIdx.x;
  for (int p = 0; p < n_phases; ++p) {
    const PPhase P = phases[p];
    for (unsigned v = blockIdx.x; v < P.total; v += gridDim.x) {
      // the segment holding v: the last whose first block is <= v
      unsigned lo = P.seg0, hi = P.seg0 + P.nseg;
      while (hi - lo > 1u) {
        const unsigned mid = (lo + hi) >> 1;
        if (segs[mid].vb0 <= v) lo = mid;
        else hi = mid;
      }
      const PSeg& S = segs[lo];
      const unsigned local = v - S.vb0;
      if (local >= S.nvb) continue;  // the padding that aligns the next segment
      __syncthreads();               // the previous body's LDS reads are done
      run_segment<MX, TILE>(S, members[S.member], local, tid, lds);
    }
    if (p + 1 < n_phases && !grid_barrier(bar, ok)) return;
  }
}

// The step without the column tile (every member on the column march).
template <bool MX>
__global__ void __launch_bounds__(256) k_step_persistent(const PMember* __restrict__ members,
                                                        const PSeg* __restrict__ segs,
                                                        const PPhase* __restrict__ phases, int n_phases,
                                                        PBarrier* bar) {
  __shared__ double lds[kPersistLds];
  __shared__ int ok;
  step_phases<MX, false>(members, segs, phases, n_phases, bar, lds, &ok);
}

// The FP64 step with the 16 x 16 column tile among its bodies, held to two
// waves per SIMD: the tile's body would otherwise take the whole register
// file and leave one.  The extra registers and LDS slow the other bodies, so
// a batch runs this kernel only when a member runs the tile.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
    k_step_persistent_tile(const PMember* __restrict__ members, const PSeg* __restrict__ segs,
                           const PPhase* __restrict__ phases, int n_phases, PBarrier* bar) {
  __shared__ double lds[kPersistLdsTile];
  __shared__ int ok;
  step_phases<false, true>(members, segs, phases, n_phases, bar, lds, &ok);
}

}  // namespace

int persistent_column_levels(const DGrid& g, const DNamelist& nl, int column_tile) {
  if (column_tile != kPersistTile || nl.mixed) return 0;
  const int r = (g.nz + 1 + kPersistTileKt - 1) / kPersistTileKt;
  return r <= kPersistTileLevels ? r : 0;
}

bool persistent_program(const std::vector<PMember>& members, const std::vector<int>& column_tiles,
                        std::vector<PSeg>* segs, std::vector<PPhase>* phases, std::string* why) {
  segs->clear();
  phases->clear();
  if (members.empty() || column_tiles.size() != members.size()) {
    *why = members.empty() ? "no members" : "one column kernel choice per member";
    return false;
  }
  const int nm = (int)members.size();
  // levels per thread of the member's LDS column tile, 0 for the column march
  std::vector<int> tile_levels(nm);
  for (int m = 0; m < nm; ++m) tile_levels[m] = persistent_column_levels(members[m].g, members[m].nl, column_tiles[m]);
  const bool any_march = std::count(tile_levels.begin(), tile_levels.end(), 0) > 0;
  const DNamelist& n0 = members[0].nl;
  for (int m = 0; m < nm; ++m) {
    const DNamelist& nl = members[m].nl;
    const std::string who = "member " + std::to_string(m) + ": ";
    if (nl.mixed != n0.mixed) {
      *why = who + "every member must use the same precision form";
      return false;
    }
    if (nl.ns != n0.ns) {
      *why = who + "every member must have the same number of acoustic substeps";
      return false;
    }
    if (nl.diff_6th_opt != 0 || nl.smag || nl.const_nu > 0.0 || nl.scalar_adv_opt != 0 || nl.moist) {
      *why = who + "the persistent step runs the dry dynamics without explicit diffusion or limiter "
                   "(diff_6th_opt, diff_opt, const_nu, scalar_adv_opt and moist must be 0)";
      return false;
    }
  }
  // one phase: its (kind, per-member geometry) segments, each aligned to a
  // multiple of 8 virtual blocks so that a segment's block b runs on XCD
  // b % 8, as in a standalone launch
  struct Job {
    int kind;
    int rk_step, first;
  };
  auto add_phase = [&](std::initializer_list<Job> jobs) {
    PPhase P{(unsigned)segs->size(), 0u, 0u};
    unsigned vb = 0;
    for (const Job& jb : jobs)
      for (int m = 0; m < nm; ++m) {
        const DGrid& g = members[m].g;
        const DNamelist& nl = members[m].nl;
        if (jb.kind == PK_CORIOLIS && !g.rot) continue;
        const bool march_kind = jb.kind == PK_COLUMN_MU || jb.kind == PK_COLUMN_W;
        if ((march_kind && tile_levels[m] > 0) || (jb.kind == PK_COLUMN_TILE && tile_levels[m] == 0)) continue;
        const size_t C = (size_t)g.ny * g.nx, CU = (size_t)g.ny * (g.nx + 1), CV = (size_t)(g.ny + 1) * g.nx;
        PSeg S{};
        S.kind = jb.kind;
        S.member = m;
        S.rk_step = jb.rk_step;
        S.first = jb.first;
        S.dts = jb.rk_step == 1 ? nl.dt / 3.0 : nl.dt / (double)nl.ns;
        S.gy = 1;
        switch (jb.kind) {
          case PK_ORIGIN: {
            const size_t U = (size_t)g.nz * CU, V = (size_t)g.nz * CV, Wn = (size_t)(g.nz + 1) * C;
            S.gx = blocks_for(std::max(U, std::max(V, Wn)), 256 * kOriginChunk);
            break;
          }
          case PK_RUV: S.gx = blocks_for(C, 256); S.gy = level_chunks(g); break;
          case PK_ROM: S.gx = rom_blocks(g); break;
          case PK_PREP_2D: S.gx = blocks_for(C, 256); break;
          case PK_PREP_FACES: S.gx = blocks_for(CU + CV, 256); break;
          case PK_PREP_COL: S.gx = blocks_for(C, 256); break;
          case PK_PREP_U: S.gx = blocks_for(CU, 256); S.gy = level_chunks(g); break;
          case PK_PREP_V: S.gx = blocks_for(CV, 256); S.gy = level_chunks(g); break;
          case PK_TEND_U: S.gx = tend_u_tiles(g); break;
          case PK_TEND_V: S.gx = tend_v_tiles(g); break;
          case PK_TEND_W: S.gx = blocks_for(C, 256); break;
          case PK_TEND_T: S.gx = tend_t_tiles(g); break;
          case PK_CORIOLIS: S.gx = coriolis_blocks(g); S.aux = coriolis_u_blocks(g); break;
          case PK_COL_SETUP: S.gx = col_setup_tiles(g); break;
          case PK_UV: S.gx = uv_tiles(g); break;
          case PK_COLUMN_MU: S.gx = column_mu_tiles(g); break;
          case PK_COLUMN_W: S.gx = blocks_for(C, 256); break;
          case PK_FIN_UV: S.gx = blocks_for(CU + CV, 256); S.gy = fin_chunks(g); break;
          case PK_FIN_COL: S.gx = blocks_for(C, 256); break;
          case PK_COLUMN_TILE:
            S.gx = (unsigned)((g.nx + kPersistTileTx - 1) / kPersistTileTx) * (unsigned)g.ny;
            S.aux = (unsigned)tile_levels[m];
            break;
          default: break;
        }
        S.nvb = jb.kind == PK_ROM ? (S.gx + 3u) / 4u : S.gx * S.gy;
        S.vb0 = vb;
        vb += (S.nvb + kNumXcd - 1) / kNumXcd * kNumXcd;
        segs->push_back(S);
        ++P.nseg;
      }
    P.total = vb;
    phases->push_back(P);
  };
  const int nsub[3] = {1, n0.ns / 2 > 1 ? n0.ns / 2 : 1, n0.ns};
  // Phases in wd_step's order.  The bodies grouped in one phase touch
  // disjoint arrays: k_ruv (ru, rv) and k_prep_2d (the 2-D mu work); k_rom
  // (rom) and k_prep_faces (face mu); k_prep_col (the column diagnostics)
  // and k_prep_u/v (u_work, v_work); the four tendencies (ut, vt, wt, tht);
  // the Coriolis terms (ut, vt) and k_col_setup (rw_tend, ph_tend and the
  // implicit-w coefficients); one member's k_column_mu and another's column
  // tile; the two finish kernels (u, v and the rest of the state).
  add_phase({{PK_ORIGIN, 1, 0}});
  for (int st = 0; st < 3; ++st) {
    const int rk = st + 1;
    add_phase({{PK_RUV, rk, 0}, {PK_PREP_2D, rk, 0}});
    add_phase({{PK_ROM, rk, 0}, {PK_PREP_FACES, rk, 0}});
    add_phase({{PK_PREP_COL, rk, 0}, {PK_PREP_U, rk, 0}, {PK_PREP_V, rk, 0}});
    add_phase({{PK_TEND_U, rk, 0}, {PK_TEND_V, rk, 0}, {PK_TEND_W, rk, 0}, {PK_TEND_T, rk, 0}});
    add_phase({{PK_CORIOLIS, rk, 0}, {PK_COL_SETUP, rk, 0}});
    for (int it = 0; it < nsub[st]; ++it) {
      const int first = it == 0 ? 1 : 0;
      add_phase({{PK_UV, rk, first}});
      // a member runs either the column-march pair or the LDS column tile
      add_phase({{PK_COLUMN_MU, rk, first}, {PK_COLUMN_TILE, rk, first}});
      if (any_march) add_phase({{PK_COLUMN_W, rk, first}});
    }
    add_phase({{PK_FIN_UV, rk, 0}, {PK_FIN_COL, rk, 0}});
  }
  return true;
}

hipError_t persistent_blocks_per_cu(bool mixed, bool tile, int* out) {
  if (tile)
    return mixed ? hipErrorInvalidValue
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(out, k_step_persistent_tile, 256, 0);
  return mixed ? hipOccupancyMaxActiveBlocksPerMultiprocessor(out, k_step_persistent<true>, 256, 0)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(out, k_step_persistent<false>, 256, 0);
}

hipError_t launch_persistent_step(bool mixed, bool tile, unsigned grid, const PMember* members, const PSeg* segs,
                                  const PPhase* phases, int n_phases, PBarrier* bar, hipStream_t st) {
  if (tile) {
    if (mixed) return hipErrorInvalidValue;
    k_step_persistent_tile<<<grid, 256, 0, st>>>(members, segs, phases, n_phases, bar);
  } else if (mixed) {
    k_step_persistent<true><<<grid, 256, 0, st>>>(members, segs, phases, n_phases, bar);
  } else {
    k_step_persistent<false><<<grid, 256, 0, st>>>(members, segs, phases, n_phases, bar);
  }
  return hipGetLastError();
}

}  // namespace wd
