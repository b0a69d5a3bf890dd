// Host side of the fused path: tile planning (rod -> workgroup bin packing),
// the FP32 copies of the constants, launches, and the director
// re-orthonormalisation and finiteness checks.
#include <algorithm>
#include <mutex>
#include <vector>

#include "erod_fused_kernel.hpp"

namespace erod {

  // FP32 constants of the FP32-bulk path: whole fields, in this order, packed
  // back to back (each keeps the ghost-padded slot indexing of its stagger)
  static const int kConst32Fields[] = {
      EROD_MASS, EROD_INV_MASS, EROD_REFERENCE_ELEMENT_LENGTH, EROD_REFERENCE_SHEAR_STRETCH_STRAIN,
      EROD_SHEAR_STRETCH_RIGIDITY, EROD_MASS_SECOND_MOMENT_OF_INERTIA, EROD_INV_MASS_SECOND_MOMENT_OF_INERTIA,
      EROD_REFERENCE_VORONOI_LENGTH, EROD_REFERENCE_CURVATURE, EROD_BENDING_TWIST_RIGIDITY,
      EROD_FORCE_DAMPING_RATE, EROD_TORQUE_DAMPING_RATE};
  static constexpr int kNConst32 = sizeof(kConst32Fields) / sizeof(kConst32Fields[0]);

  // offset (in floats) of constant i; const32_offset(b, kNConst32) is the total
  static uint64_t const32_offset(const erod_batch* b, int i) {
    uint64_t off = 0;
    for (int k = 0; k < i; ++k) off += uint64_t(field_info(kConst32Fields[k]).ncomp) * b->field_slots(kConst32Fields[k]);
    return off;
  }

  uint64_t fused_const32_offset(const erod_batch* b, int field) {
    for (int i = 0; i < kNConst32; ++i)
      if (kConst32Fields[i] == field) return const32_offset(b, i);
    return ~uint64_t(0);
  }

  using fused::Consts;

  namespace {
    // FP64 constants -> FP32 copies (same slot indexing)
    __global__ void k_to_f32(float* __restrict__ o, const double* __restrict__ a, uint64_t n) {
      for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        o[i] = static_cast<float>(a[i]);
    }

    // one Newton-Schulz polar step Q <- Q (3 I - Q^T Q) / 2 on element slot k
    __device__ __forceinline__ void reortho_slot(double* __restrict__ Q, uint64_t E, uint64_t k) {
      double q[9], g[9], o[9];
      for (int m = 0; m < 9; ++m) q[m] = Q[m * E + k];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) g[3 * a + b] = q[a] * q[b] + q[3 + a] * q[3 + b] + q[6 + a] * q[6 + b];
      for (int p = 0; p < 3; ++p)
        for (int r = 0; r < 3; ++r) {
          double s = 0.0;
          for (int c = 0; c < 3; ++c) s += q[3 * p + c] * ((c == r ? 3.0 : 0.0) - g[3 * c + r]);
          o[3 * p + r] = 0.5 * s;
        }
      for (int m = 0; m < 9; ++m) Q[m * E + k] = o[m];
    }

    // every element slot
    __global__ void k_reortho(double* __restrict__ Q, uint64_t E) {
      for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < E; k += (uint64_t)gridDim.x * blockDim.x)
        reortho_slot(Q, E, k);
    }

    // the active elements of every rod with steps left (steps_left null: all
    // rods); a rod whose budget is spent keeps its state bit for bit
    __global__ void k_reortho_live(double* __restrict__ Q, uint64_t E, const RodInfo* __restrict__ rods, uint64_t n_rods,
                                   const uint32_t* __restrict__ steps_left) {
      for (uint64_t r = blockIdx.x; r < n_rods; r += gridDim.x) {
        if (steps_left && steps_left[r] == 0) continue;
        const RodInfo ri = rods[r];
        for (uint32_t k = threadIdx.x; k < ri.n; k += blockDim.x) reortho_slot(Q, E, ri.start + k);
      }
    }

    template <typename Flag>
    __global__ void k_check_finite(const double* __restrict__ X, const double* __restrict__ Vel, const double* __restrict__ Qd,
                                   const double* __restrict__ W, const RodInfo* __restrict__ rods, uint64_t n_rods,
                                   uint64_t E, Flag* __restrict__ flags) {
      const uint64_t N = E + 1;
      for (uint64_t r = blockIdx.x; r < n_rods; r += gridDim.x) {
        const RodInfo ri = rods[r];
        int bad = 0;
        for (uint32_t k = threadIdx.x; k <= ri.n; k += blockDim.x) {
          const uint64_t s = ri.start + k;
          for (int c = 0; c < 3; ++c) bad |= int(!isfinite(X[c * N + s])) | int(!isfinite(Vel[c * N + s]));
          if (k < ri.n) {
            for (int c = 0; c < 3; ++c) bad |= !isfinite(W[c * E + s]);
            for (int m = 0; m < 9; ++m) bad |= !isfinite(Qd[m * E + s]);
          }
        }
        if (bad) flags[r] = Flag(1);
      }
    }

    __global__ void k_zero_u8(uint8_t* __restrict__ p, uint64_t n) {
      for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
    }

  }  // namespace

  // ---- planning: bin-pack rods into tiles --------------------------------------

  int fused_shape_lanes(int si) { return fused::kShapes[si].P * fused::kShapes[si].NT; }
  int fused_shape_epl(int si) { return fused::kShapes[si].P; }

  // rods the fused kernels advance: non-empty slots with steps left
  static bool planned(const erod_batch* b, uint64_t r) {
    return b->n_active[r] != 0 && (!b->budgets || b->steps_left[r] != 0);
  }

  // the shape of the pieces of long rods: the one of seg_lanes lanes (and
  // seg_epl elements per thread when set), or the main shape; -2 if none
  static int piece_shape(const erod_batch* b, int main_shape) {
    if (b->plan.seg_lanes <= 0) return main_shape;
    for (int si : {0, 1, 2, 4, 5, 3, 7, 6})  // the default order of preference
      if (fused_shape_lanes(si) == b->plan.seg_lanes && (b->plan.seg_epl <= 0 || fused_shape_epl(si) == b->plan.seg_epl))
        return si;
    return -2;
  }

  // A rod runs as pieces when it is longer than the main tile, or longer
  // than the pieces' tile when seg_lanes is set.
  static bool cut_rod(const erod_batch* b, uint64_t n, int main_shape) {
    const uint64_t cap = uint64_t(fused_shape_lanes(main_shape));
    return n > cap || (b->plan.seg_lanes > 0 && n > uint64_t(b->plan.seg_lanes));
  }

  // Feasible main shapes: its lanes and elements per thread are the ones
  // requested, and every rod that is not cut has at least P elements (a
  // thread then holds at most one rod end).
  static bool shape_feasible(const erod_batch* b, int si) {
    const fused::Shape sh = fused::kShapes[si];
    const int lanes = b->plan.lanes_per_tile, epl = b->plan.elements_per_thread;
    if (lanes && sh.P * sh.NT != lanes) return false;
    if (epl && sh.P != epl) return false;
    for (uint64_t r = 0; r < b->n_rods; ++r)
      if (planned(b, r) && !cut_rod(b, b->n_active[r], si) && b->n_active[r] < uint64_t(sh.P)) return false;
    return true;
  }

  // Default shape without measurement: the smallest capacity that holds the
  // longest rod; one element per thread up to 256 lanes, two elements per
  // thread for 512 and 1024 lanes; rods longer than 1024 elements run as
  // pieces.  (FP32-bulk on ragged batches: 512 lanes at two elements per
  // thread, two 4-wave workgroups per CU, run 27% faster than one element
  // per thread in 8-wave workgroups; four elements per thread spill and run
  // one wave per SIMD.)  A shape the autotuner chose is taken again while it
  // stays feasible (a replan after rods changed length keeps it, with its
  // second group).
  static int default_shape(const erod_batch* b) {
    static const int pref[] = {0, 1, 2, 4, 5, 3, 7, 6};
    if (b->plan.tuned_shape >= 0 && shape_feasible(b, b->plan.tuned_shape)) return b->plan.tuned_shape;
    uint64_t maxn = 0;
    for (uint64_t r = 0; r < b->n_rods; ++r)
      if (planned(b, r)) maxn = std::max<uint64_t>(maxn, b->n_active[r]);
    for (int si : pref)
      if (uint64_t(fused_shape_lanes(si)) >= maxn && shape_feasible(b, si)) return si;
    // longer rods than any tile: the largest tiles, the longer rods in pieces
    for (int si : {5, 4, 3, 2, 1, 0, 7, 6})
      if (shape_feasible(b, si)) return si;
    return -1;
  }

  // One group of tiles on the host: the tiles, their entries (rods or
  // pieces) in order and each entry's first lane; rods sent on to a second group.
  struct Packed {
    std::vector<Tile> tiles;
    std::vector<uint32_t> trods, tlane0, overflow;
  };

  // Best-fit-decreasing bin packing of entries into tiles of P * NT elements:
  // each entry, longest first, goes to the tile with the least free room that
  // holds it (ties: the tile that reached that room first).  A workgroup costs
  // about the same whether its tile is full or not, so the tile count sets the
  // launch time; on the 4096-rod ragged batch this packs to the lower bound
  // (972 tiles of 1024 lanes, against 1170 for first fit over the 64 newest).
  // Lengths and rooms are at most P * NT, so the sort is a counting sort and
  // the open tiles sit in one FIFO list per room, found through a bitset: a
  // serving pool replans its ~30,000 slots in about 0.5 ms of host time.
  // Here `order` holds the entries, longest first, len[e] their lengths, and
  // cap is P * NT.  With spill_cap > 0, an entry of at most spill_cap
  // elements that finds no room in an open tile goes to `overflow` instead of
  // opening one, provided it has at least spill_min elements (the second
  // group's elements per thread; a shorter rod stays in the main group).
  static void pack(const std::vector<uint32_t>& len, uint64_t cap, const std::vector<uint32_t>& order,
                   uint64_t spill_cap, uint64_t spill_min, Packed& out) {
    // open tiles by free room: FIFO lists (head, tail, next) and a bitset of non-empty rooms
    std::vector<int32_t> head(cap + 1, -1), tail(cap + 1, -1), next;
    std::vector<uint64_t> nonempty(cap / 64 + 1, 0);
    auto open = [&](uint64_t room, uint32_t t) {
      next[t] = -1;
      if (tail[room] < 0) head[room] = int32_t(t);
      else next[tail[room]] = int32_t(t);
      tail[room] = int32_t(t);
      nonempty[room >> 6] |= uint64_t(1) << (room & 63);
    };
    std::vector<uint32_t> placed, tile_of, tile_rods_count;
    for (const uint32_t r : order) {
      const uint64_t n = len[r];
      int64_t room = -1;  // the smallest room >= n that has a tile
      for (uint64_t w = n >> 6; w < nonempty.size() && room < 0; ++w) {
        uint64_t m = nonempty[w];
        if (w == (n >> 6)) m &= ~uint64_t(0) << (n & 63);
        if (m) room = int64_t(w * 64 + uint64_t(__builtin_ctzll(m)));
      }
      uint32_t t;
      if (room < 0) {
        if (n <= spill_cap && n >= spill_min) {
          out.overflow.push_back(r);
          continue;
        }
        t = static_cast<uint32_t>(tile_rods_count.size());
        tile_rods_count.push_back(0);
        next.push_back(-1);
        if (cap > n) open(cap - n, t);
      } else {
        t = static_cast<uint32_t>(head[room]);
        head[room] = next[t];
        if (head[room] < 0) {
          tail[room] = -1;
          nonempty[room >> 6] &= ~(uint64_t(1) << (room & 63));
        }
        if (const uint64_t left = uint64_t(room) - n) open(left, t);
      }
      placed.push_back(r);
      tile_of.push_back(t);
      ++tile_rods_count[t];
    }
    // tiles in creation order, each with its entries in the order they were placed
    const size_t n_tiles = tile_rods_count.size(), n_placed = placed.size();
    out.tiles.assign(n_tiles, Tile{});
    out.trods.assign(n_placed, 0);
    out.tlane0.assign(n_placed, 0);
    std::vector<uint32_t> fill(n_tiles), lane(n_tiles, 0);
    for (uint32_t t = 0, first = 0; t < n_tiles; ++t) {
      out.tiles[t].first = first;
      out.tiles[t].count = tile_rods_count[t];
      out.tiles[t].pad = 0;
      fill[t] = first;
      first += tile_rods_count[t];
    }
    for (size_t i = 0; i < n_placed; ++i) {
      const uint32_t t = tile_of[i], k = fill[t]++;
      out.trods[k] = placed[i];
      out.tlane0[k] = lane[t];
      lane[t] += len[placed[i]];
    }
    for (size_t t = 0; t < n_tiles; ++t) out.tiles[t].lanes = lane[t];
  }

  // entries longest first (stable), by a counting sort over lengths <= cap
  static std::vector<uint32_t> longest_first(const std::vector<uint32_t>& len, const std::vector<uint32_t>& ids,
                                             uint64_t cap) {
    std::vector<uint32_t> count(cap + 2, 0), next_pos(cap + 2, 0), order(ids.size());
    for (const uint32_t e : ids) ++count[len[e]];
    uint32_t k = 0;
    for (uint64_t n = cap + 1; n >= 1; --n) {
      next_pos[n] = k;
      k += count[n];
    }
    for (const uint32_t e : ids) order[next_pos[len[e]]++] = e;
    return order;
  }

  // The pieces of a rod of n elements for tiles of `cap` lanes: owned parts
  // of nearly equal length, each piece loading `halo` more elements on each
  // side that has a neighbouring piece; every piece at most cap long.
  static void cut(uint32_t rod, uint64_t n, uint64_t cap, uint64_t halo, std::vector<Seg>& out) {
    uint64_t m = 2;
    while (true) {
      // the longest piece of an even split into m: a middle part plus two halos
      const uint64_t part = (n + m - 1) / m;
      if (part + 2 * halo <= cap || part <= 1) break;
      ++m;
    }
    for (uint64_t i = 0; i < m; ++i) {
      const uint64_t a = i * n / m, z = (i + 1) * n / m;  // owned [a, z)
      const uint64_t lo = i == 0 ? 0 : a - halo, hi = i + 1 == m ? n : z + halo;
      Seg sg{};
      sg.rod = rod;
      sg.off = uint32_t(lo);
      sg.n = uint32_t(hi - lo);
      sg.w0 = uint32_t(a - lo);
      sg.w1 = uint32_t(z - lo);
      sg.ends = (i == 0 ? kSegFirst : 0u) | (i + 1 == m ? kSegLast : 0u);
      out.push_back(sg);
    }
  }

  // Device copy of one group.  The arrays grow to the largest plan seen and
  // are reused when the rod set changes.
  template <typename T>
  static int grow(T*& p, uint64_t& cap, uint64_t need, hipStream_t stream) {
    if (cap >= need) return EROD_OK;
    EROD_HIP(hipStreamSynchronize(stream));
    if (p) EROD_HIP(hipFree(p));
    p = nullptr;
    cap = 0;
    EROD_HIP(hipMalloc(&p, need * sizeof(T)));
    cap = need;
    return EROD_OK;
  }

  static int upload_group(erod_batch* b, const Packed& pk, TileGroup& g) {
    const uint64_t need = std::max<uint64_t>({pk.tiles.size(), pk.trods.size(), 1});
    if (g.capacity < need) {
      uint64_t c1 = g.capacity, c2 = g.capacity, c3 = g.capacity;
      if (int rc = grow(g.d_tiles, c1, need, b->stream)) return rc;
      if (int rc = grow(g.d_tile_rods, c2, need, b->stream)) return rc;
      if (int rc = grow(g.d_tile_lane0, c3, need, b->stream)) return rc;
      g.capacity = need;
    }
    if (!pk.tiles.empty()) {
      EROD_HIP(hipMemcpyAsync(g.d_tiles, pk.tiles.data(), pk.tiles.size() * sizeof(Tile), hipMemcpyHostToDevice, b->stream));
      EROD_HIP(hipMemcpyAsync(g.d_tile_rods, pk.trods.data(), pk.trods.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                              b->stream));
      EROD_HIP(hipMemcpyAsync(g.d_tile_lane0, pk.tlane0.data(), pk.tlane0.size() * sizeof(uint32_t),
                              hipMemcpyHostToDevice, b->stream));
    }
    return EROD_OK;
  }

  // the shape a requested second group uses, or -1 (none requested)
  static int split_shape(const erod_batch* b) {
    if (b->plan.split_lanes <= 0) return -1;
    for (int si : {0, 1, 2, 4, 5, 3, 7, 6})  // the default order of preference
      if (fused::kShapes[si].P * fused::kShapes[si].NT == b->plan.split_lanes &&
          (b->plan.split_epl <= 0 || fused::kShapes[si].P == b->plan.split_epl))
        return si;
    return -2;
  }

  static int plan_tiles(erod_batch* b, int si) {
    const uint64_t cap = uint64_t(fused_shape_lanes(si));
    const int ps = piece_shape(b, si);
    if (ps == -2) {
      set_error("plan_tiles: no tile shape of the requested piece lanes and elements per thread");
      return EROD_ERR_INVALID;
    }
    const uint64_t pcap = uint64_t(fused_shape_lanes(ps));
    const uint64_t halo = uint64_t(b->plan.halo);
    if (halo < 2 || 2 * halo + 2 > pcap) {
      set_error("plan_tiles: the halo of rod pieces must be at least 2 and leave room in the pieces' tile");
      return EROD_ERR_INVALID;
    }
    // whole rods of the main (and second) group, and pieces of long rods
    std::vector<uint32_t> len(b->n_rods), whole;
    std::vector<uint32_t> seg_rods;
    std::vector<Seg> segs;
    for (uint32_t i = 0; i < b->n_rods; ++i) {
      if (!planned(b, i)) continue;
      len[i] = static_cast<uint32_t>(b->n_active[i]);
      if (cut_rod(b, b->n_active[i], si)) {
        if (b->n_active[i] > (uint64_t(1) << 31)) {
          set_error("plan_tiles: a rod longer than 2^31 elements");
          return EROD_ERR_INVALID;
        }
        seg_rods.push_back(i);
        cut(i, b->n_active[i], pcap, halo, segs);
      } else {
        whole.push_back(i);
      }
    }
    const int s2 = split_shape(b);
    if (s2 == -2) {
      set_error("plan_tiles: no tile shape of the requested split lanes and elements per thread");
      return EROD_ERR_INVALID;
    }
    // a second group only of tiles smaller than the main ones (a requested
    // split at least as large as the main tile is ignored); it takes rods of
    // at most its lanes and at least its elements per thread
    uint64_t cap2 = s2 >= 0 ? uint64_t(fused_shape_lanes(s2)) : 0;
    if (cap2 >= cap) cap2 = 0;
    const uint64_t min2 = cap2 ? uint64_t(fused_shape_epl(s2)) : 0;
    // each entry takes a whole number of its group's threads: its length
    // rounded up to the elements per thread (the kernel puts the idle pad
    // lanes before the entry's elements, so that its last element is a
    // thread's last)
    auto padded = [](uint64_t n, uint64_t p) { return static_cast<uint32_t>((n + p - 1) / p * p); };
    std::vector<uint32_t> len1(len.size());
    for (const uint32_t i : whole) len1[i] = padded(len[i], uint64_t(fused_shape_epl(si)));
    Packed g1, g2, g3;
    pack(len1, cap, longest_first(len1, whole, cap), cap2, min2, g1);
    if (!g1.overflow.empty()) {
      std::vector<uint32_t> len2(len.size());
      for (const uint32_t i : g1.overflow) len2[i] = padded(len[i], uint64_t(fused_shape_epl(s2)));
      pack(len2, cap2, g1.overflow, 0, 0, g2);
    }
    std::vector<uint32_t> slen(segs.size()), sids(segs.size());
    for (uint32_t k = 0; k < segs.size(); ++k) {
      slen[k] = padded(segs[k].n, uint64_t(fused_shape_epl(ps)));
      sids[k] = k;
    }
    if (!segs.empty()) pack(slen, pcap, longest_first(slen, sids, pcap), 0, 0, g3);
    // the kernel stages a tile's entry table in NT entries of LDS (rods have
    // at least two elements, and at least P: a feasible shape never exceeds it)
    const int shapes[3] = {si, s2 >= 0 ? s2 : si, ps};
    const Packed* pks[3] = {&g1, &g2, &g3};
    for (int k = 0; k < 3; ++k)
      for (const Tile& t : pks[k]->tiles)
        if (t.count > uint32_t(fused::kShapes[shapes[k]].NT)) {
          set_error("plan_tiles: more rods in a tile than threads");
          return EROD_ERR_INVALID;
        }
    FusedPlan& pl = b->plan;
    for (int k = 0; k < 3; ++k)
      if (!pks[k]->tiles.empty())
        if (int rc = upload_group(b, *pks[k], pl.g[k])) return rc;
    TileGroup& gs = pl.g[2];
    if (!segs.empty()) {
      if (int rc = grow(gs.d_segs, gs.seg_capacity, segs.size(), b->stream)) return rc;
      if (int rc = grow(gs.d_seg_rods, gs.seg_rod_capacity, seg_rods.size(), b->stream)) return rc;
      if (!b->d_time_next) EROD_HIP(hipMalloc(&b->d_time_next, sizeof(double) * b->n_rods));
      const int shadow_fields[4] = {EROD_POSITION, EROD_VELOCITY, EROD_DIRECTOR, EROD_ANGULAR_VELOCITY};
      for (int k = 0; k < 4; ++k)
        if (!b->d_shadow[k])
          EROD_HIP(hipMalloc(&b->d_shadow[k], sizeof(double) * field_info(shadow_fields[k]).ncomp *
                                                  b->field_slots(shadow_fields[k])));
      EROD_HIP(hipMemcpyAsync(gs.d_segs, segs.data(), segs.size() * sizeof(Seg), hipMemcpyHostToDevice, b->stream));
      EROD_HIP(hipMemcpyAsync(gs.d_seg_rods, seg_rods.data(), seg_rods.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                              b->stream));
    }
    EROD_HIP(hipStreamSynchronize(b->stream));
    for (int k = 0; k < 3; ++k) {
      pl.g[k].n_tiles = pks[k]->tiles.size();
      pl.g[k].shape = pks[k]->tiles.empty() ? -1 : shapes[k];
      pl.g[k].max_steps = 0;
    }
    gs.n_segs = segs.size();
    gs.n_seg_rods = seg_rods.size();
    gs.max_steps = int(halo / 2);
    if (segs.empty()) gs.shape = -1;
    pl.valid = true;
    b->graph_precision = -1;
    return EROD_OK;
  }

  int fused_prepare(erod_batch* b, int precision) {
    int rc = sync_rods(b);
    if (rc) return rc;
    if (!b->plan.valid) {
      const int si = default_shape(b);
      if (si < 0) {
        set_error("no fused tile shape holds this batch (a requested shape with more elements per thread than "
                  "a rod has)");
        return EROD_ERR_UNSUPPORTED;
      }
      if ((rc = plan_tiles(b, si))) return rc;
    }
    if (precision == EROD_FP32_BULK && b->const32_dirty) {
      const uint64_t total = const32_offset(b, kNConst32);
      if (!b->d_const32 || b->const32_floats != total) {
        if (b->d_const32) EROD_HIP(hipFree(b->d_const32));
        b->d_const32 = nullptr;
        EROD_HIP(hipMalloc(&b->d_const32, sizeof(float) * total));
        b->const32_floats = total;
      }
      for (int i = 0; i < kNConst32; ++i) {
        const int f = kConst32Fields[i];
        const uint64_t n = uint64_t(field_info(f).ncomp) * b->field_slots(f);
        hipLaunchKernelGGL(k_to_f32, dim3(grid_for(n)), dim3(kThreads), 0, b->stream, b->d_const32 + const32_offset(b, i),
                           b->d_field[f], n);
      }
      EROD_HIP(hipGetLastError());
      b->const32_dirty = false;
    }
    return EROD_OK;
  }

  template <typename R>
  static Consts<R> make_consts(erod_batch* b);

  template <>
  Consts<double> make_consts<double>(erod_batch* b) {
    Consts<double> C;
    C.m = b->d_field[EROD_MASS];
    C.invm = b->d_field[EROD_INV_MASS];
    C.l0 = b->d_field[EROD_REFERENCE_ELEMENT_LENGTH];
    C.sig0 = b->d_field[EROD_REFERENCE_SHEAR_STRETCH_STRAIN];
    C.S = b->d_field[EROD_SHEAR_STRETCH_RIGIDITY];
    C.J = b->d_field[EROD_MASS_SECOND_MOMENT_OF_INERTIA];
    C.iJ = b->d_field[EROD_INV_MASS_SECOND_MOMENT_OF_INERTIA];
    C.D0 = b->d_field[EROD_REFERENCE_VORONOI_LENGTH];
    C.kap0 = b->d_field[EROD_REFERENCE_CURVATURE];
    C.B = b->d_field[EROD_BENDING_TWIST_RIGIDITY];
    C.gf = b->d_field[EROD_FORCE_DAMPING_RATE];
    C.gt = b->d_field[EROD_TORQUE_DAMPING_RATE];
    return C;
  }

  template <>
  Consts<float> make_consts<float>(erod_batch* b) {
    const float* p = b->d_const32;
    Consts<float> C;
    C.m = p + const32_offset(b, 0);
    C.invm = p + const32_offset(b, 1);
    C.l0 = p + const32_offset(b, 2);
    C.sig0 = p + const32_offset(b, 3);
    C.S = p + const32_offset(b, 4);
    C.J = p + const32_offset(b, 5);
    C.iJ = p + const32_offset(b, 6);
    C.D0 = p + const32_offset(b, 7);
    C.kap0 = p + const32_offset(b, 8);
    C.B = p + const32_offset(b, 9);
    C.gf = p + const32_offset(b, 10);
    C.gt = p + const32_offset(b, 11);
    return C;
  }

  // Measured choice of the tile shape: every feasible shape advances the
  // batch `trial_steps` steps from the same snapshot of x, v, Q, w and the rod
  // clocks (restored afterwards), timed with HIP events; the fastest is kept.
  // Unless the caller set a second group, each smaller shape is then tried as
  // one for the rods that find no room beside longer ones, and kept if faster.
  int fused_autotune(erod_batch* b, int precision, int trial_steps, double* best_ms) {
    int rc = fused_prepare(b, precision);
    if (rc) return rc;
    const int snap_fields[4] = {EROD_POSITION, EROD_VELOCITY, EROD_DIRECTOR, EROD_ANGULAR_VELOCITY};
    // snapshot buffers and the plan settings the trials change, released and
    // restored on every return
    struct Guard {
      erod_batch* b;
      double* snap[5] = {};
      int lanes, epl, split_lanes, split_epl;
      bool budgets;
      bool keep_split = false;  // a second group chosen by the trials stays
      ~Guard() {
        for (double* p : snap)
          if (p) (void)hipFree(p);
        b->plan.lanes_per_tile = lanes;
        b->plan.elements_per_thread = epl;
        if (!keep_split) {
          b->plan.split_lanes = split_lanes;
          b->plan.split_epl = split_epl;
        }
        b->budgets = budgets;
        b->last_launches = 0;  // the trial launches are not a fused call's timing
      }
    } g{b, {}, b->plan.lanes_per_tile, b->plan.elements_per_thread, b->plan.split_lanes, b->plan.split_epl, b->budgets};
    size_t bytes[5];
    for (int i = 0; i < 4; ++i) bytes[i] = sizeof(double) * field_info(snap_fields[i]).ncomp * b->field_slots(snap_fields[i]);
    bytes[4] = sizeof(double) * b->n_rods;
    auto dev = [&](int i) { return i < 4 ? b->d_field[snap_fields[i]] : b->d_time; };
    for (int i = 0; i < 5; ++i) {
      EROD_HIP(hipMalloc(&g.snap[i], bytes[i]));
      EROD_HIP(hipMemcpyAsync(g.snap[i], dev(i), bytes[i], hipMemcpyDeviceToDevice, b->stream));
    }
    auto restore = [&]() -> int {
      for (int i = 0; i < 5; ++i) EROD_HIP(hipMemcpyAsync(dev(i), g.snap[i], bytes[i], hipMemcpyDeviceToDevice, b->stream));
      return EROD_OK;
    };
    b->plan.lanes_per_tile = 0;
    b->plan.elements_per_thread = 0;
    b->budgets = false;  // trial launches leave the step budgets alone
    int best = -1;
    double best_t = 0.0;
    for (int si = 0; si < fused::kNShapes && rc == EROD_OK; ++si) {
      if (!shape_feasible(b, si)) continue;
      if ((rc = plan_tiles(b, si))) break;
      if ((rc = fused_launch(b, precision, trial_steps, false))) break;  // warm
      if ((rc = fused_launch(b, precision, trial_steps, true))) break;
      float ms = 0.0f;
      if ((rc = hip_check(hipEventSynchronize(b->ev1), "autotune sync"))) break;
      if ((rc = hip_check(hipEventElapsedTime(&ms, b->ev0, b->ev1), "autotune time"))) break;
      if (best < 0 || ms < best_t) {
        best = si;
        best_t = ms;
      }
      if ((rc = restore())) break;
    }
    // a second group of smaller tiles, when the caller has not set one
    int best_split = -1;
    if (rc == EROD_OK && best >= 0 && g.split_lanes == 0) {
      const int cap = fused_shape_lanes(best);
      for (int s2 = 0; s2 < fused::kNShapes && rc == EROD_OK; ++s2) {
        if (fused_shape_lanes(s2) >= cap) continue;
        b->plan.split_lanes = fused_shape_lanes(s2);
        b->plan.split_epl = fused_shape_epl(s2);
        const int prc = plan_tiles(b, best);
        if (prc == EROD_ERR_HIP) {  // a device failure ends the search; an infeasible shape does not
          rc = prc;
          break;
        }
        if (prc != EROD_OK || b->plan.g[1].shape < 0) continue;  // infeasible here, or no second group
        if ((rc = fused_launch(b, precision, trial_steps, false))) break;  // warm
        if ((rc = fused_launch(b, precision, trial_steps, true))) break;
        float ms = 0.0f;
        if ((rc = hip_check(hipEventSynchronize(b->ev1), "autotune sync"))) break;
        if ((rc = hip_check(hipEventElapsedTime(&ms, b->ev0, b->ev1), "autotune time"))) break;
        if (ms < best_t) {
          best_split = s2;
          best_t = ms;
        }
        if ((rc = restore())) break;
      }
    }
    // the chosen plan is made with the caller's settings and budgets; the
    // chosen main shape is kept for later replans while it stays feasible
    b->plan.tuned_shape = best;
    b->plan.lanes_per_tile = g.lanes;
    b->plan.elements_per_thread = g.epl;
    b->plan.split_lanes = best_split >= 0 ? fused_shape_lanes(best_split) : g.split_lanes;
    b->plan.split_epl = best_split >= 0 ? fused_shape_epl(best_split) : g.split_epl;
    g.keep_split = true;
    b->budgets = g.budgets;
    if (rc == EROD_OK && best >= 0) rc = plan_tiles(b, best);
    if (rc == EROD_OK) rc = restore();
    if (rc == EROD_OK) rc = hip_check(hipStreamSynchronize(b->stream), "autotune");
    if (rc == EROD_OK && best < 0) {
      set_error("no feasible fused tile shape");
      rc = EROD_ERR_UNSUPPORTED;
    }
    if (best_ms) *best_ms = best_t / trial_steps;
    return rc;
  }

  int fused_launch(erod_batch* b, int precision, int steps, bool timed) {
    if (timed) EROD_HIP(hipEventRecord(b->ev0, b->stream));
    const int rc = precision == EROD_FP64 ? fused_launch_fp64(b, make_consts<double>(b), steps)
                                          : fused_launch_fp32(b, make_consts<float>(b), steps);
    if (rc) return rc;
    if (timed) EROD_HIP(hipEventRecord(b->ev1, b->stream));
    return EROD_OK;
  }

  int fused_reorthonormalize(erod_batch* b) {
    hipLaunchKernelGGL(k_reortho, dim3(grid_for(b->E)), dim3(kThreads), 0, b->stream, b->d_field[EROD_DIRECTOR], b->E);
    return hip_check(hipGetLastError(), "reorthonormalize");
  }

  int fused_reorthonormalize_live(erod_batch* b) {
    const unsigned g = static_cast<unsigned>(b->n_rods < 65535 ? b->n_rods : 65535);
    hipLaunchKernelGGL(k_reortho_live, dim3(g), dim3(256), 0, b->stream, b->d_field[EROD_DIRECTOR], b->E, b->d_rods,
                       b->n_rods, static_cast<const uint32_t*>(b->d_steps_left));
    return hip_check(hipGetLastError(), "scheduled reorthonormalization");
  }

  int fused_check_finite(erod_batch* b, uint8_t* flags) {
    int rc = sync_rods(b);
    if (rc) return rc;
    EROD_HIP(hipMemsetAsync(b->d_flags, 0, sizeof(int) * b->n_rods, b->stream));
    const unsigned g = static_cast<unsigned>(b->n_rods < 65535 ? b->n_rods : 65535);
    hipLaunchKernelGGL(k_check_finite<int>, dim3(g), dim3(256), 0, b->stream, b->d_field[EROD_POSITION],
                       b->d_field[EROD_VELOCITY], b->d_field[EROD_DIRECTOR], b->d_field[EROD_ANGULAR_VELOCITY], b->d_rods,
                       b->n_rods, b->E, b->d_flags);
    EROD_HIP(hipGetLastError());
    std::vector<int> h(b->n_rods);
    EROD_HIP(hipMemcpyAsync(h.data(), b->d_flags, sizeof(int) * b->n_rods, hipMemcpyDeviceToHost, b->stream));
    EROD_HIP(hipStreamSynchronize(b->stream));
    for (uint64_t i = 0; i < b->n_rods; ++i) flags[i] = static_cast<uint8_t>(h[i] != 0);
    return EROD_OK;
  }

  // ---- finiteness check that allocates nothing -----------------------------------
  //
  // The per-rod flags live in `scratch`, a file-scope device array of
  // kScratchRods bytes that every GPU the library runs on holds until the
  // process ends (each GPU its own copy).  Its device address is looked up
  // once per device and kept; the array's name is never used as an address
  // in host code.  The flags the kernel needs cleared are cleared by a small
  // kernel before each check (no memset node), and the result is copied
  // straight into the caller's array.  Calls that overlap on one GPU share
  // the array and race.

  constexpr uint64_t kScratchRods = uint64_t(1) << 20;  // 1 MiB: one byte per rod
  static __device__ __align__(256) uint8_t scratch[kScratchRods];

  namespace {
    constexpr int kMaxDevices = 64;
    std::mutex g_scratch_mutex;
    uint8_t* g_scratch_ptr[kMaxDevices] = {};

    // the address of `scratch` on `device`, looked up on the first call there
    int scratch_on(int device, uint8_t** out) {
      if (device < 0 || device >= kMaxDevices) {
        set_error("check_finite_noalloc: device ordinal out of range");
        return EROD_ERR_INVALID;
      }
      std::lock_guard<std::mutex> lock(g_scratch_mutex);
      if (!g_scratch_ptr[device]) {
        int current = 0;
        EROD_HIP(hipGetDevice(&current));
        if (current != device) EROD_HIP(hipSetDevice(device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (current != device) EROD_HIP(hipSetDevice(current));
        EROD_HIP(e);
        g_scratch_ptr[device] = static_cast<uint8_t*>(p);
      }
      *out = g_scratch_ptr[device];
      return EROD_OK;
    }
  }  // namespace

  int fused_check_finite_noalloc(erod_batch* b, uint8_t* flags) {
    if (b->n_rods > kScratchRods) {
      set_error("erod_check_finite_noalloc: at most 1,048,576 rods (the size of its device scratch array)");
      return EROD_ERR_INVALID;
    }
    uint8_t* d = nullptr;
    if (int rc = scratch_on(b->device, &d)) return rc;
    // rod records changed on the host go up whole (the staged upload of a
    // few records would allocate its staging buffer)
    if (b->rods_dirty || !b->dirty_rods.empty()) {
      EROD_HIP(hipMemcpyAsync(b->d_rods, b->rods.data(), sizeof(RodInfo) * b->n_rods, hipMemcpyHostToDevice, b->stream));
      b->rods_dirty = false;
      b->dirty_rods.clear();
    }
    hipLaunchKernelGGL(k_zero_u8, dim3(grid_for(b->n_rods)), dim3(kThreads), 0, b->stream, d, b->n_rods);
    const unsigned g = static_cast<unsigned>(b->n_rods < 65535 ? b->n_rods : 65535);
    hipLaunchKernelGGL(k_check_finite<uint8_t>, dim3(g), dim3(256), 0, b->stream, b->d_field[EROD_POSITION],
                       b->d_field[EROD_VELOCITY], b->d_field[EROD_DIRECTOR], b->d_field[EROD_ANGULAR_VELOCITY], b->d_rods,
                       b->n_rods, b->E, d);
    EROD_HIP(hipGetLastError());
    EROD_HIP(hipMemcpyAsync(flags, d, b->n_rods, hipMemcpyDeviceToHost, b->stream));
    EROD_HIP(hipStreamSynchronize(b->stream));
    return EROD_OK;
  }

}  // namespace erod
