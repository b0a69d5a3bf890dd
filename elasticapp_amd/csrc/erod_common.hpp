// Shared definitions of the device-resident Cosserat-rod block: field table,
// ghost values, batch state and launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/erod.h"

namespace erod {
  constexpr uint32_t kNoBudget = 0xFFFFFFFFu;
  // one rod copied between batches: slots, active length, rod ids
  struct CopyPair {
    uint64_t src_start, dst_start;
    uint32_t n, src_rod, dst_rod, pad;
  };
}

namespace erod {

  struct FieldInfo {
    int ncomp;
    int stagger;   // erod_stagger
    double ghost;  // value of ghost slots; < -1 marks the director ghost
  };

  // Ghost values of the reference variables: scalars {} = 0
  // (common/Traits/DataType/ScalarTag.hpp:85), vectors 0 (MatrixTag.hpp:102),
  // lengths / volume / dimension / mass 1 (CosseratRodSpanwiseGeometry.hpp:97,153,179,
  // CosseratRodCrossSectionInterface.hpp:95,121, WithRodKinematics.hpp:101),
  // director {{.8,.6,0},{0,0,1},{.6,-.8,0}} (RodSpanwiseGeometry.hpp:96-101).
  inline const FieldInfo& field_info(int f) {
    static const FieldInfo t[EROD_N_FIELDS] = {
        {3, EROD_NODE, 0.0},     {3, EROD_NODE, 0.0},    {3, EROD_NODE, 0.0},
        {1, EROD_NODE, 1.0},     {1, EROD_NODE, 0.0},    {3, EROD_NODE, 0.0},
        {3, EROD_NODE, 0.0},     {1, EROD_NODE, 0.0},
        {9, EROD_ELEMENT, -2.0}, {3, EROD_ELEMENT, 0.0}, {3, EROD_ELEMENT, 0.0},
        {1, EROD_ELEMENT, 1.0},  {1, EROD_ELEMENT, 1.0}, {1, EROD_ELEMENT, 0.0},
        {3, EROD_ELEMENT, 0.0},  {1, EROD_ELEMENT, 1.0}, {1, EROD_ELEMENT, 1.0},
        {3, EROD_ELEMENT, 0.0},  {3, EROD_ELEMENT, 0.0}, {3, EROD_ELEMENT, 0.0},
        {3, EROD_ELEMENT, 0.0},  {3, EROD_ELEMENT, 0.0}, {3, EROD_ELEMENT, 0.0},
        {3, EROD_ELEMENT, 0.0},  {3, EROD_ELEMENT, 0.0}, {1, EROD_ELEMENT, 0.0},
        {3, EROD_ELEMENT, 0.0},  {3, EROD_ELEMENT, 0.0},
        {1, EROD_VORONOI, 1.0},  {1, EROD_VORONOI, 0.0}, {1, EROD_VORONOI, 0.0},
        {3, EROD_VORONOI, 0.0},  {3, EROD_VORONOI, 0.0}, {3, EROD_VORONOI, 0.0},
        {3, EROD_VORONOI, 0.0},  {3, EROD_VORONOI, 0.0},
    };
    return t[f];
  }

  // Rod descriptor used by the fused kernels and the hooks (device copy).
  struct RodInfo {
    uint64_t start;   // slot of node 0 / element 0 / voronoi 0
    uint32_t n;       // elements
    uint32_t flags;   // bit0 clamp 0, bit1 end force, bit2 damping rates that vary along the rod, bit3 clamp 1
    double dt;        // timestep of this rod (fused path)
    double gamma_f;   // force damping rate (when the same along the rod)
    double gamma_t;   // torque damping rate (when the same along the rod)
    // up to two clamps: node clamp_idx[k] held at clamp[k].x, and element
    // clamp_idx[k] (when below n) at director clamp[k].q
    struct Pose {
      double x[3];
      double q[9];
    } clamp[2];
    uint32_t clamp_idx[2];
    double force[3];
    int ramp;
    double ramp_start, ramp_interval;
  };
  // flag bit of clamp k
  constexpr uint32_t clamp_bit(int k) { return k == 0 ? 1u : 8u; }
  constexpr uint32_t kClampBits = 1u | 8u;

  struct Env {
    int gravity;
    double g[3];
    int contact;
    double k, c, mu, ct, z0, offset;
    double n[3] = {0.0, 0.0, 1.0};  // unit normal of the contact plane n . x = z0
  };

  // A tile: the rods one workgroup of the fused kernel owns.
  struct Tile {
    uint32_t first;   // index into the tile rod list
    uint32_t count;   // rods in this tile
    uint32_t lanes;   // elements (lanes) used
    uint32_t pad;
  };

  // A piece of a rod that one tile holds: elements [off, off + n) of rod
  // `rod`.  The tile writes back only the elements [off + w0, off + w1) it
  // owns (and their nodes); the others are a halo, the neighbouring pieces'
  // elements, recomputed from the state they had when the launch began.  In
  // one step a node or element depends on one neighbour before it and two
  // after it, so a halo of h elements on each side keeps the owned elements
  // exact for h / 2 steps: a launch over pieces takes at most that many.
  // `ends`: kSegFirst if the piece starts at the rod's first element,
  // kSegLast if it ends at the rod's last (the rod end rules, clamp and end
  // force apply only at such ends).
  struct Seg {
    uint32_t rod, off, n, w0, w1, ends;
  };
  constexpr uint32_t kSegFirst = 1u, kSegLast = 2u;

  // One group of tiles of a fused plan, all of one shape.
  struct TileGroup {
    int shape = -1;                    // index into fused::kShapes (-1: no tiles)
    uint64_t n_tiles = 0;
    Tile* d_tiles = nullptr;
    uint32_t* d_tile_rods = nullptr;   // rod ids (or, with pieces, piece ids) in tile order
    uint32_t* d_tile_lane0 = nullptr;  // first lane of each within its tile
    uint64_t capacity = 0;             // entries allocated in each of the three arrays
    // pieces of long rods (nullptr: the entries are whole rods)
    Seg* d_segs = nullptr;
    uint64_t n_segs = 0, seg_capacity = 0;
    uint32_t* d_seg_rods = nullptr;    // the rods cut into the pieces
    uint64_t n_seg_rods = 0, seg_rod_capacity = 0;
    int max_steps = 0;                 // steps per launch at most (0: no limit)
  };

  struct FusedPlan {
    int lanes_per_tile = 0;       // requested elements per tile (0 = any)
    int elements_per_thread = 0;  // requested elements per thread (0 = any)
    int steps_per_launch = 0;     // requested (0 = 100)
    int tuned_shape = -1;         // the autotuner's choice (-1: none), preferred while feasible
    // optional second group: rods of at most split_lanes elements that find no
    // room beside longer rods run in tiles of that shape, in a second launch
    int split_lanes = 0, split_epl = 0;  // requested (0 = no second group)
    // rods longer than the main tile, or than seg_lanes when it is set, run
    // as overlapping pieces of at most seg_lanes (0: the main tile's) lanes
    // with `halo` elements recomputed on each side
    int seg_lanes = 0, seg_epl = 0, halo = 16;
    // groups: 0 main, 1 second (short rods), 2 pieces of long rods
    TileGroup g[3];
    bool valid = false;
  };

}  // namespace erod

struct erod_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t n_rods = 0;
  uint64_t E = 0;  // element slots incl. ghosts
  bool damped = true;
  std::vector<uint64_t> n_elems, starts, ghost_nodes;  // n_elems: the slot capacity of each rod
  // active length of each rod (<= its capacity; 0 = empty slot).  Equal to
  // n_elems unless erod_set_rod_active_length changed it; the op-by-op path
  // needs every rod at full length.
  std::vector<uint64_t> n_active;
  bool all_full = true;
  // per-rod step budgets of the fused path (kNoBudget = unlimited)
  std::vector<uint32_t> steps_left;
  uint32_t* d_steps_left = nullptr;
  bool budgets = false, budgets_dirty = false;
  double* d_field[EROD_N_FIELDS] = {};
  uint64_t* d_ghost_nodes = nullptr;
  uint8_t* d_node_is_ghost = nullptr;  // N flags: 1 on the ghost node between two rods
  // rods, hooks, clocks
  std::vector<erod::RodInfo> rods;
  erod::RodInfo* d_rods = nullptr;
  double* d_time = nullptr;  // per-rod clock
  bool rods_dirty = true;                // every rod record must be uploaded
  std::vector<uint32_t> dirty_rods;      // otherwise: the records changed since the last upload
  erod::RodInfo* d_rod_stage = nullptr;  // staging of a scattered record upload
  uint32_t* d_rod_stage_idx = nullptr;
  uint64_t rod_stage_capacity = 0;
  bool damping_dirty = true;  // per-rod damping rates must be re-read for the fused path
  bool has_clamps = false, has_forces = false;
  erod::Env env{};
  // FP32 copies of the constants for the FP32-bulk path
  float* d_const32 = nullptr;
  uint64_t const32_floats = 0;
  bool const32_dirty = true;
  erod::FusedPlan plan;
  // graph cache
  hipGraphExec_t graph_exec = nullptr;
  uint64_t graph_launches = 0;  // kernel launches in one replay of graph_exec
  bool graph_one_node = false;  // graph_exec is a single kernel node
  int graph_precision = -1;
  uint64_t graph_chunk = 0;
  uint64_t graph_reortho[2] = {0, 0};  // interval and phase the graph was captured at
  uint64_t graph_phase_after = 0;      // phase after one replay
  // scheduled re-orthonormalisation of the directors on the fused paths:
  // after the launch during which the batch's step count reaches a multiple
  // of reortho_every (0: never); reortho_phase = steps since the last one
  uint64_t reortho_every = 0, reortho_phase = 0;
  double* d_time_next = nullptr;  // rod clocks written by launches over pieces, copied back after each
  // second copy of x, v, Q, w that launches over pieces alternate with (same layout as the fields)
  double* d_shadow[4] = {};
  // timing of the dominant kernel
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // a second stream, and the events that fork a plan's tile groups onto it and
  // join them back: the groups hold disjoint rods, so the whole-rod tiles run
  // beside the chain of launches over rod pieces
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  uint64_t last_launches = 0;
  // scratch
  int* d_flags = nullptr;
  erod::CopyPair* d_copy_pairs = nullptr;  // grow-only staging of erod_copy_rods
  uint64_t copy_capacity = 0;

  uint64_t slots(int stagger) const { return stagger == EROD_NODE ? E + 1 : (stagger == EROD_ELEMENT ? E : E - 1); }
  uint64_t field_slots(int f) const { return slots(erod::field_info(f).stagger); }
};

namespace erod {

  void set_error(const std::string& msg);

  inline int hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
      set_error(std::string(what) + ": " + hipGetErrorString(e));
      return EROD_ERR_HIP;
    }
    return EROD_OK;
  }

#define EROD_HIP(call)                                   \
  do {                                                   \
    int _rc = ::erod::hip_check((call), #call);          \
    if (_rc != EROD_OK) return _rc;                      \
  } while (0)

  constexpr int kThreads = 256;
  inline unsigned grid_for(uint64_t n) {
    uint64_t g = (n + kThreads - 1) / kThreads;
    if (g == 0) g = 1;
    if (g > 65535ull * 16) g = 65535ull * 16;
    return static_cast<unsigned>(g);
  }

  // op-by-op path (erod_ops.hip)
  int ops_kinematic_update(erod_batch* b, double dt);
  int ops_kinematic_update_by(erod_batch* b, double c);
  int ops_fill_field(erod_batch* b, int field, double v);
  int ops_compute_internal_dynamics(erod_batch* b);
  int ops_dynamic_update(erod_batch* b, double dt);
  int ops_fill_ghosts(erod_batch* b, int field);
  int ops_mark_ghost_nodes(erod_batch* b);
  int ops_set_rod_clocks(erod_batch* b, double t);
  int ops_apply_post_kinematic(erod_batch* b);
  int ops_apply_pre_dynamic(erod_batch* b, double time);
  int ops_apply_post_dynamic(erod_batch* b);
  int ops_zero_external(erod_batch* b);
  // fused path (erod_fused.hip)
  int fused_prepare(erod_batch* b, int precision);
  int fused_launch(erod_batch* b, int precision, int steps, bool timed);
  int fused_reorthonormalize(erod_batch* b);
  // the directors of every rod with steps left to take (no budget, or one
  // not spent), after a launch on the batch's stream
  int fused_reorthonormalize_live(erod_batch* b);
  int fused_check_finite(erod_batch* b, uint8_t* flags);
  int fused_check_finite_noalloc(erod_batch* b, uint8_t* flags);
  int fused_autotune(erod_batch* b, int precision, int trial_steps, double* best_ms);
  // offset (in floats) of a field in the FP32 constant copy, or ~0 if the field
  // has none; valid while the copy is allocated and not dirty
  uint64_t fused_const32_offset(const erod_batch* b, int field);
  // device copy of whole rods between batches (admission of a request)
  int serve_copy_rods(erod_batch* dst, const erod_batch* src, const std::vector<CopyPair>& pairs);
  int fused_shape_lanes(int shape);
  int fused_shape_epl(int shape);
  int sync_rods(erod_batch* b);
  // note a changed rod record (uploaded by the next sync_rods)
  inline void mark_rod(erod_batch* b, uint64_t r) {
    if (!b->rods_dirty) b->dirty_rods.push_back(static_cast<uint32_t>(r));
  }

}  // namespace erod
