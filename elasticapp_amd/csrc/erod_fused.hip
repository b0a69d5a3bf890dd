// Host side of the fused path: tile planning (rod -> workgroup bin packing),
// the FP32 copies of the constants, launches, and the director
// re-orthonormalisation and finiteness checks.
#include <algorithm>
#include <vector>

#include "erod_fused_kernel.hpp"

namespace erod {

  // FP32 constants of the FP32-bulk path: whole fields, in this order, packed
  // back to back (each keeps the ghost-padded slot indexing of its stagger)
  static const int kConst32Fields[] = {
      EROD_MASS, EROD_INV_MASS, EROD_REFERENCE_ELEMENT_LENGTH, EROD_REFERENCE_SHEAR_STRETCH_STRAIN,
      EROD_SHEAR_STRETCH_RIGIDITY, EROD_MASS_SECOND_MOMENT_OF_INERTIA, EROD_INV_MASS_SECOND_MOMENT_OF_INERTIA,
      EROD_REFERENCE_VORONOI_LENGTH, EROD_REFERENCE_CURVATURE, EROD_BENDING_TWIST_RIGIDITY};
  static constexpr int kNConst32 = sizeof(kConst32Fields) / sizeof(kConst32Fields[0]);

  // offset (in floats) of constant i; const32_offset(b, kNConst32) is the total
  static uint64_t const32_offset(const erod_batch* b, int i) {
    uint64_t off = 0;
    for (int k = 0; k < i; ++k) off += uint64_t(field_info(kConst32Fields[k]).ncomp) * b->field_slots(kConst32Fields[k]);
    return off;
  }

  using fused::Consts;

  namespace {
    // FP64 constants -> FP32 copies (same slot indexing)
    __global__ void k_to_f32(float* __restrict__ o, const double* __restrict__ a, uint64_t n) {
      for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        o[i] = static_cast<float>(a[i]);
    }

    // one Newton-Schulz polar step Q <- Q (3 I - Q^T Q) / 2 on every element slot
    __global__ void k_reortho(double* __restrict__ Q, uint64_t E) {
      for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < E; k += (uint64_t)gridDim.x * blockDim.x) {
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
    }

    __global__ void k_check_finite(const double* __restrict__ X, const double* __restrict__ Vel, const double* __restrict__ Qd,
                                   const double* __restrict__ W, const RodInfo* __restrict__ rods, uint64_t n_rods,
                                   uint64_t E, int* __restrict__ flags) {
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
        if (bad) flags[r] = 1;
      }
    }

  }  // namespace

  // ---- planning: bin-pack rods into tiles --------------------------------------

  int fused_prepare(erod_batch* b, int precision) {
    int rc = sync_rods(b);
    if (rc) return rc;
    if (!b->plan.valid) {
      // tile capacity (elements) and elements per thread
      int cap = b->plan.lanes_per_tile;
      uint64_t maxn = 0;
      for (auto n : b->n_elems) maxn = n > maxn ? n : maxn;
      if (cap == 0) cap = maxn > 512 ? 1024 : (maxn > 256 ? 512 : 256);
      if (cap != 64 && cap != 128 && cap != 256 && cap != 512 && cap != 1024) {
        set_error("lanes_per_tile must be one of 64, 128, 256, 512, 1024");
        return EROD_ERR_INVALID;
      }
      if (maxn > uint64_t(cap)) {
        set_error("a rod is longer than the tile capacity of the fused path");
        return EROD_ERR_UNSUPPORTED;
      }
      b->plan.epl = cap == 1024 ? 2 : 1;
      b->plan.chosen_lanes = cap;
      // first-fit decreasing by length
      std::vector<uint32_t> order(b->n_rods);
      for (uint32_t i = 0; i < b->n_rods; ++i) order[i] = i;
      std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t c) { return b->n_elems[a] > b->n_elems[c]; });
      std::vector<std::vector<uint32_t>> bins;
      std::vector<uint64_t> fill;
      // bins sorted by remaining space would be faster; rods per bin are few
      for (uint32_t r : order) {
        const uint64_t n = b->n_elems[r];
        size_t bi = bins.size();
        for (size_t i = (bins.size() > 64 ? bins.size() - 64 : 0); i < bins.size(); ++i)
          if (fill[i] + n <= uint64_t(cap)) { bi = i; break; }
        if (bi == bins.size()) { bins.emplace_back(); fill.push_back(0); }
        bins[bi].push_back(r);
        fill[bi] += n;
      }
      std::vector<Tile> tiles(bins.size());
      std::vector<uint32_t> trods, tlane0;
      for (size_t i = 0; i < bins.size(); ++i) {
        tiles[i].first = static_cast<uint32_t>(trods.size());
        tiles[i].count = static_cast<uint32_t>(bins[i].size());
        uint32_t lane = 0;
        for (uint32_t r : bins[i]) {
          trods.push_back(r);
          tlane0.push_back(lane);
          lane += static_cast<uint32_t>(b->n_elems[r]);
        }
        tiles[i].lanes = lane;
        tiles[i].pad = 0;
      }
      hipFree(b->plan.d_tiles);
      hipFree(b->plan.d_tile_rods);
      hipFree(b->plan.d_tile_lane0);
      EROD_HIP(hipMalloc(&b->plan.d_tiles, tiles.size() * sizeof(Tile)));
      EROD_HIP(hipMalloc(&b->plan.d_tile_rods, trods.size() * sizeof(uint32_t)));
      EROD_HIP(hipMalloc(&b->plan.d_tile_lane0, tlane0.size() * sizeof(uint32_t)));
      EROD_HIP(hipMemcpyAsync(b->plan.d_tiles, tiles.data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice, b->stream));
      EROD_HIP(hipMemcpyAsync(b->plan.d_tile_rods, trods.data(), trods.size() * sizeof(uint32_t), hipMemcpyHostToDevice, b->stream));
      EROD_HIP(hipMemcpyAsync(b->plan.d_tile_lane0, tlane0.data(), tlane0.size() * sizeof(uint32_t), hipMemcpyHostToDevice, b->stream));
      EROD_HIP(hipStreamSynchronize(b->stream));
      b->plan.n_tiles = tiles.size();
      b->plan.chosen_steps = b->plan.steps_per_launch > 0 ? b->plan.steps_per_launch : 100;
      b->plan.valid = true;
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
    return C;
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

  int fused_check_finite(erod_batch* b, uint8_t* flags) {
    int rc = sync_rods(b);
    if (rc) return rc;
    EROD_HIP(hipMemsetAsync(b->d_flags, 0, sizeof(int) * b->n_rods, b->stream));
    const unsigned g = static_cast<unsigned>(b->n_rods < 65535 ? b->n_rods : 65535);
    hipLaunchKernelGGL(k_check_finite, dim3(g), dim3(256), 0, b->stream, b->d_field[EROD_POSITION], b->d_field[EROD_VELOCITY],
                       b->d_field[EROD_DIRECTOR], b->d_field[EROD_ANGULAR_VELOCITY], b->d_rods, b->n_rods, b->E, b->d_flags);
    EROD_HIP(hipGetLastError());
    std::vector<int> h(b->n_rods);
    EROD_HIP(hipMemcpyAsync(h.data(), b->d_flags, sizeof(int) * b->n_rods, hipMemcpyDeviceToHost, b->stream));
    EROD_HIP(hipStreamSynchronize(b->stream));
    for (uint64_t i = 0; i < b->n_rods; ++i) flags[i] = static_cast<uint8_t>(h[i] != 0);
    return EROD_OK;
  }

}  // namespace erod
