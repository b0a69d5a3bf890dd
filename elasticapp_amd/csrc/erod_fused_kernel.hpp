// Fused, persistent position-Verlet step kernel (device code, instantiated
// for FP64 in erod_fused64.hip and for FP32-bulk in erod_fused32.hip).
//
// One workgroup owns a tile of whole rods (or of pieces of long rods), each
// starting on a thread's first lane and ending on a thread's last (idle pad
// lanes before an entry whose length is not a multiple of P).  Each thread
// holds P contiguous lanes of the tile (element k, node k and, for the last
// element of a rod, the rod's last node n) with the state x, v, Q, w in FP64
// registers across
// `nsteps` steps.  The only cross-thread data of a step are the +1 neighbour
// node values and director and the next element's node difference (first
// exchange) and the -1 neighbour element force / couple terms (second
// exchange), passed through LDS, so a step costs two workgroup barriers and
// no HBM traffic.  P > 1 gives each thread several independent elements (ILP)
// and lets rods of up to P * NT elements share one workgroup.
//
// Arithmetic of one step follows the reference block timestep
// (PositionVerlet.hpp:49-88, SymplecticSteppers.hpp:152-216) with explicit rod
// ends in place of the ghost slots (zero ghosts make the block-wide difference
// and quadrature operators reduce to the per-rod boundary rules used here):
//   geometry/strain   CosseratRodSpanwiseGeometry.hpp:578-664
//   loads             WithDiagonalLinearHyperElasticModel.hpp:197-238
//   torques           WithDiagonalLinearHyperElasticModel.hpp:246-391
//   damping           ExplicitDampingAdapterFacade.hpp:177-235
//   accelerations     WithRodKinematics.hpp:683-714
//   director update   SO3PrimitiveAddAssign/Scalar.hpp:25-96
//
// Precision
//   R = double: the reference formulas (acos-based log map with its 1e-10 /
//   1e-14 offsets, sin / cos in the director update; acos, sin and cos are
//   the library's own, erod_math.hpp), its divisions
//   and its order of evaluation; built without FMA contraction, it rounds
//   like the op-by-op port (and equals it bitwise on the test batches).
//   R = float (FP32-bulk): FP64 state, FP64 accumulation of x, v, w and an
//   FP64 exponential map (the series form below, which needs no sqrt or
//   division); strains, stresses, couples, loads and torques in FP32 with the
//   cancellation-prone parts (node differences and the relative rotation of
//   neighbouring directors) formed in FP64 before rounding.  The log map is
//   the reference's acos(cos(theta) - 1e-10) rewritten without cancellation
//   as theta = 2 asin(sqrt(sin^2(theta/2) + 5e-11)) (a series for small
//   angles; past a quarter turn from the cosine side, with cos^2(theta/2)
//   also from the FP64 trace, so it stays accurate up to a half turn);
//   element lengths come
//   from node differences rounded to FP32, which the neighbour that needs the
//   same length receives instead of recomputing it.  The FP32 constants are
//   held in an LDS table for the whole launch, loaded straight from global
//   memory into LDS beside the state loads (1/m, 1/J, 1/l0, 1/D0 and the
//   damping products formed once per launch into registers), and inside a
//   launch the closing and opening half steps of consecutive steps are one
//   kinematic update of dt.
#pragma once

#include "erod_common.hpp"
#include "erod_math.hpp"

namespace erod {
  namespace fused {

    template <typename R>
    struct Consts {
      const R* m;     // node
      const R* invm;  // node
      const R* l0;    // element
      const R* sig0;  // element, 3 lines
      const R* S;     // element, 3 lines
      const R* J;     // element, 3 lines
      const R* iJ;    // element, 3 lines
      const R* D0;    // voronoi
      const R* kap0;  // voronoi, 3 lines
      const R* B;     // voronoi, 3 lines
      const R* gf;    // node: force damping rate (read for rods whose rates vary along them)
      const R* gt;    // element: torque damping rate (same)
    };

    // the state arrays a launch reads (in) and writes (out): the batch's
    // fields for whole rods; for pieces of rods, which read their halo from
    // neighbouring pieces, alternately the fields and a second copy, so that
    // no launch reads what it writes
    struct StatePtrs {
      double* x;  // node, 3 lines
      double* v;  // node, 3 lines
      double* q;  // element, 9 lines
      double* w;  // element, 3 lines
    };

    // ---- small helpers ---------------------------------------------------------

    __device__ __forceinline__ float rcp_(float x) { return __builtin_amdgcn_rcpf(x); }
    __device__ __forceinline__ double rcp_(double x) { return 1.0 / x; }

    // one word per lane from global memory straight into LDS, without a
    // register (global_load_lds_dword): lane l of the wave writes wave_dst[l],
    // wave_dst being the same for every lane of the wave; lanes that are off
    // write nothing
    __device__ __forceinline__ void load_to_lds(const float* src, float* wave_dst) {
      // (C-style casts: they convert between address spaces)
      __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                       (__attribute__((address_space(3))) void*)wave_dst, 4, 0, 0);
    }
    __device__ __forceinline__ float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
    __device__ __forceinline__ double sqrt_(double x) { return sqrt(x); }

    // a * b + c: one fused multiply-add in the FP32-bulk path, a separately
    // rounded product and sum in the FP64 path (which keeps the reference's
    // rounding).  The units are built without automatic contraction, so every
    // FMA is one of these and the arithmetic of an element does not depend on
    // the surrounding code (tile shape, position within a thread).
    template <typename R>
    __device__ __forceinline__ double maddd(double a, double b, double c) {
      if constexpr (sizeof(R) == 8) return a * b + c;
      else return __builtin_fma(a, b, c);
    }
    __device__ __forceinline__ float madd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
    __device__ __forceinline__ double madd(double a, double b, double c) { return a * b + c; }

    // element length from an FP64 node difference; both neighbours of a
    // Voronoi region evaluate it the same way
    template <typename R>
    __device__ __forceinline__ R len_(double d0, double d1, double d2) {
      if constexpr (sizeof(R) == 8) return sqrt(d0 * d0 + d1 * d1 + d2 * d2);
      else return sqrt_(static_cast<float>(__builtin_fma(d2, d2, __builtin_fma(d1, d1, d0 * d0))));
    }

    // FP32 length of an FP32 difference
    __device__ __forceinline__ float len32(float a, float b, float c) { return sqrt_(madd(c, c, madd(b, b, a * a))); }

    template <typename R>
    __device__ __forceinline__ void cross3(const R a[3], const R b[3], R o[3]) {
      o[0] = madd(a[1], b[2], -(a[2] * b[1]));
      o[1] = madd(a[2], b[0], -(a[0] * b[2]));
      o[2] = madd(a[0], b[1], -(a[1] * b[0]));
    }

    __device__ __forceinline__ double ramp_value(int kind, double t, double start, double interval) {
      if (kind == EROD_RAMP_NONE) return 1.0;
      if (t < start) return 0.0;
      if (t > start + interval) return 1.0;
      if (kind == EROD_RAMP_LINEAR) return (t - start) / interval;
      const double s = 2.0 * ((t - start) / interval) - 1.0;
      return 0.5 * (s + (1.0 + (math::sin(M_PI * s) / M_PI)));
    }

    // ---- director update Q <- exp(-[u]x) Q, u = c w, in FP64 ----------------------

    template <typename R>
    __device__ __forceinline__ void rodrigues(double q[9], const double w[3], double cdt);

    // the reference (SO3AddAssign<scalar>, Scalar.hpp:25-96)
    template <>
    __device__ __forceinline__ void rodrigues<double>(double q[9], const double w[3], double cdt) {
      double u0 = cdt * w[0], u1 = cdt * w[1], u2 = cdt * w[2];
      const double theta = sqrt(u0 * u0 + u1 * u1 + u2 * u2);
      double beta, c_alpha;
      math::sincos(theta, &beta, &c_alpha);
      const double alpha = 1.0 - c_alpha;
      const double den = theta + 1e-14;
      u0 /= den;
      u1 /= den;
      u2 /= den;
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double a = q[j], b = q[3 + j], cc = q[6 + j];
        const double com = a * u0 + b * u1 + cc * u2;
        q[j] = c_alpha * a + alpha * u0 * com + beta * (b * u2 - cc * u1);
        q[3 + j] = c_alpha * b + alpha * u1 * com + beta * (cc * u0 - a * u2);
        q[6 + j] = c_alpha * cc + alpha * u2 * com + beta * (a * u1 - b * u0);
      }
    }

    // FP32-bulk: the same map written in u itself,
    //   q' = q - th^2 B q + B (u . q) u + A (q x u),
    // A = sin(th)/th, B = (1 - cos(th))/th^2, both even series in th^2 = u . u
    // (exact to rounding for th < 1e-2), so no sqrt or division per step;
    // larger rotations per update (on the 4,096-rod ragged batch at each rod's
    // stable dt, |w| dt stays below 5e-3) take sin and cos.  Each column c of
    // q (the components of d1, d2, d3 along one axis) maps as c' = M c with
    // M = (1 - th^2 B) I + B u u^T + A [c -> c x u]: M is formed once (15
    // operations) and applied with 27 FMAs, against 57 operations for the
    // three columns taken one at a time.
    template <>
    __device__ __forceinline__ void rodrigues<float>(double q[9], const double w[3], double cdt) {
      const double u0 = cdt * w[0], u1 = cdt * w[1], u2 = cdt * w[2];
      const double t2 = __builtin_fma(u2, u2, __builtin_fma(u1, u1, u0 * u0));
      double A, B;
      if (t2 < 1.0e-4) {
        // A = 1 - t2/6 + t2^2/120 - t2^3/5040, B = 1/2 - t2/24 + t2^2/720 - t2^3/40320
        A = __builtin_fma(t2, __builtin_fma(t2, __builtin_fma(t2, -1.0 / 5040.0, 1.0 / 120.0), -1.0 / 6.0), 1.0);
        B = __builtin_fma(t2, __builtin_fma(t2, __builtin_fma(t2, -1.0 / 40320.0, 1.0 / 720.0), -1.0 / 24.0), 0.5);
      } else {
        const double th = sqrt(t2), r = 1.0 / t2;
        double s, c;
        math::sincos(th, &s, &c);
        A = s * th * r;
        B = (1.0 - c) * r;
      }
      const double cq = __builtin_fma(-t2, B, 1.0);
      const double b0 = B * u0, b1 = B * u1, b2 = B * u2, a0 = A * u0, a1 = A * u1, a2 = A * u2;
      const double m00 = __builtin_fma(b0, u0, cq), m11 = __builtin_fma(b1, u1, cq), m22 = __builtin_fma(b2, u2, cq);
      const double m01 = __builtin_fma(b0, u1, a2), m10 = __builtin_fma(b1, u0, -a2);
      const double m02 = __builtin_fma(b0, u2, -a1), m20 = __builtin_fma(b2, u0, a1);
      const double m12 = __builtin_fma(b1, u2, a0), m21 = __builtin_fma(b2, u1, -a0);
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double a = q[j], b = q[3 + j], cc = q[6 + j];
        q[j] = __builtin_fma(m02, cc, __builtin_fma(m01, b, m00 * a));
        q[3 + j] = __builtin_fma(m12, cc, __builtin_fma(m11, b, m10 * a));
        q[6 + j] = __builtin_fma(m22, cc, __builtin_fma(m21, b, m20 * a));
      }
    }

    // ---- curvature of the Voronoi region between directors qc and qn / D0 --------

    template <typename R>
    __device__ __forceinline__ void curvature(const double qc[9], const double qn[9], R D0, R kap[3]);

    // reference formula (InvRotateDivide/Scalar.hpp:76-103)
    template <>
    __device__ __forceinline__ void curvature<double>(const double qc[9], const double qn[9], double D0, double kap[3]) {
      const double tr = qc[0] * qn[0] + qc[1] * qn[1] + qc[2] * qn[2] + qc[3] * qn[3] + qc[4] * qn[4] +
                        qc[5] * qn[5] + qc[6] * qn[6] + qc[7] * qn[7] + qc[8] * qn[8];
      const double theta = math::acos(0.5 * tr - (0.5 + 1e-10));
      const double st = 0.5 * theta / math::sin(theta + 1e-14) / D0;
      kap[0] = st * (-qc[3] * qn[6] - qc[4] * qn[7] - qc[5] * qn[8] + qc[6] * qn[3] + qc[7] * qn[4] + qc[8] * qn[5]);
      kap[1] = st * (qc[0] * qn[6] + qc[1] * qn[7] + qc[2] * qn[8] - qc[6] * qn[0] - qc[7] * qn[1] - qc[8] * qn[2]);
      kap[2] = st * (-qc[0] * qn[3] - qc[1] * qn[4] - qc[2] * qn[5] + qc[3] * qn[0] + qc[4] * qn[1] + qc[5] * qn[2]);
    }

    // FP32-bulk: skew part and trace in FP64, prefactor in FP32.  With
    // a = sin^2(theta/2) (3 - tr)/4 + 5e-11, the reference prefactor
    // theta / (2 sin theta) equals asin(sqrt a) / (2 sqrt(a) sqrt(1 - a)),
    // taken as a series for a < 0.02 (theta < 16 degrees) and with library
    // asin above, 1 - a coming from the FP64 trace.
    __device__ __forceinline__ double dot3_fma(const double* a, const double* b) {
      return __builtin_fma(a[2], b[2], __builtin_fma(a[1], b[1], a[0] * b[0]));
    }
    // sum_j a_j b_j - c_j d_j over three j, with FMAs
    __device__ __forceinline__ double skew3_fma(const double* a, const double* b, const double* c, const double* d) {
      double s = a[0] * b[0];
      s = __builtin_fma(a[1], b[1], s);
      s = __builtin_fma(a[2], b[2], s);
      s = __builtin_fma(-c[0], d[0], s);
      s = __builtin_fma(-c[1], d[1], s);
      return __builtin_fma(-c[2], d[2], s);
    }

    template <>
    __device__ __forceinline__ void curvature<float>(const double qc[9], const double qn[9], float iD0, float kap[3]) {
      const double tr = dot3_fma(qc, qn) + dot3_fma(qc + 3, qn + 3) + dot3_fma(qc + 6, qn + 6);
      // rows: d1 = q[0..2], d2 = q[3..5], d3 = q[6..8]
      const double s0 = skew3_fma(qc + 6, qn + 3, qc + 3, qn + 6);
      const double s1 = skew3_fma(qc + 0, qn + 6, qc + 6, qn + 0);
      const double s2 = skew3_fma(qc + 3, qn + 0, qc + 0, qn + 3);
      float a = static_cast<float>((3.0 - tr) * 0.25) + 5e-11f;
      a = fminf(fmaxf(a, 0.0f), 1.0f);
      // st = (theta / 2) / sin(theta) / D0 with a = sin^2(theta / 2):
      // theta / (2 sin theta) = g(a) / 2, g(a) = asin(sqrt a) / (sqrt(a) sqrt(1 - a))
      //                      = sum_n 4^n (n!)^2 / (2n + 1)! a^n = 1 + 2a/3 + 8a^2/15 + 16a^3/35 + 128a^4/315 + ...
      // (the next term, 0.37 a^5, is below 1.2e-9 of g for a < 0.02)
      float g;
      if (a < 0.02f) {
        g = madd(a, madd(a, madd(a, madd(a, 128.0f / 315.0f, 16.0f / 35.0f), 8.0f / 15.0f), 2.0f / 3.0f), 1.0f);
      } else {
        // 1 - a = cos^2(theta / 2) - 5e-11 = (1 + tr) / 4 - 5e-11 from the FP64
        // trace, and theta / 2 = asin(sqrt a) up to a quarter turn, pi/2 -
        // asin(sqrt(1 - a)) beyond: accurate up to a relative rotation of pi,
        // where 1 - a formed in FP32 would cancel (the floor only keeps an
        // exact half turn, where the axis is undefined, finite)
        const float b = fmaxf(static_cast<float>((1.0 + tr) * 0.25) - 5e-11f, 1e-30f);
        const float sh = sqrt_(a), cb = sqrt_(b);
        const float half = a <= 0.5f ? asinf(sh) : 1.57079632679489662f - asinf(cb);
        g = half / (sh * cb);
      }
      const float st = 0.5f * g * iD0;
      kap[0] = st * static_cast<float>(s0);
      kap[1] = st * static_cast<float>(s1);
      kap[2] = st * static_cast<float>(s2);
    }

    // ---- per element constants ------------------------------------------------

    template <typename R>
    struct ElemConst {
      R m, invm, mn;  // node k mass and inverse, node k+1 mass
      R l0, sig0[3], S[3], J[3], iJ[3];
      R D0, kap0[3], B[3];  // Voronoi k (unused at the rod end)
      R il0, iD0;           // 1 / l0, 1 / D0 (FP32-bulk only)
    };

    template <typename R>
    __device__ __forceinline__ void load_const(const Consts<R>& C, uint64_t s, uint64_t E, bool has_vor, ElemConst<R>& K) {
      const uint64_t V = E - 1;
      K.m = C.m[s];
      K.invm = C.invm[s];
      K.mn = C.m[s + 1];
      K.l0 = C.l0[s];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        K.sig0[c] = C.sig0[c * E + s];
        K.S[c] = C.S[c * E + s];
        K.J[c] = C.J[c * E + s];
        K.iJ[c] = C.iJ[c * E + s];
      }
      // (loads from slot 0 and selects at a rod's last element, whose Voronoi
      // slot may lie past the arrays' end: no branch, and no stack copy of
      // the constants that a branch over them leads the compiler to make)
      const uint64_t sv = has_vor ? s : 0;
      K.D0 = has_vor ? C.D0[sv] : R(1);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const R k0 = C.kap0[c * V + sv], b = C.B[c * V + sv];
        K.kap0[c] = has_vor ? k0 : R(0);
        K.B[c] = has_vor ? b : R(0);
      }
      if constexpr (sizeof(R) == 4) {
        K.il0 = 1.0f / K.l0;
        K.iD0 = 1.0f / K.D0;
      }
    }

    // per-element values carried from the first to the second half of a step
    template <typename R>
    struct Mid {
      R c[3];      // Q^T n / e
      R P[3];      // M / eps^3 (Voronoi k; 0 at the rod end)
      R Rv[3];     // (kappa x M) D0 / eps^3 (Voronoi k; 0 at the rod end)
      R shear[3];  // (Q t x n) l0
      R lag[3];    // (J w / e) x w
      R unst[3];   // (J w / e) (t . dv) / (e l0)
      R e;         // dilatation
    };

    // Element k (and Voronoi k when it exists) of a rod: everything that needs
    // only the +1 / +2 neighbours.  qr is the director used for Q t and Q^T n
    // (the FP64 state for R = double, its FP32 copy for R = float).
    template <typename R>
    __device__ __forceinline__ void element_forces(const double xk[3], const double xk1[3], const double xk2[3],
                                                   const float dn[3], const double vk[3], const double vk1[3],
                                                   const double qk[9], const R qr[9], const double qk1[9],
                                                   const double wk[3], bool has_vor, const ElemConst<R>& K,
                                                   Mid<R>& out, const float* geo = nullptr) {
      // spanwise geometry.  FP64: the reference's node difference, length and
      // tangent.  FP32-bulk: the node difference in FP64 rounded to FP32, then
      // everything in FP32; an element's length is a function of that rounded
      // difference alone, so the neighbour that needs it (for the Voronoi
      // length) gets the same bits from the same difference (dn).  `geo`, when
      // given (FP32-bulk), holds that rounded difference, its length and the
      // next element's length, formed once per element and step by the caller.
      R l, t[3];
      if constexpr (sizeof(R) == 8) {
        const double d0 = xk1[0] - xk[0], d1 = xk1[1] - xk[1], d2 = xk1[2] - xk[2];
        l = len_<R>(d0, d1, d2);
        t[0] = d0 / l;
        t[1] = d1 / l;
        t[2] = d2 / l;
      } else {
        const float d0 = geo ? geo[0] : static_cast<float>(xk1[0] - xk[0]),
                    d1 = geo ? geo[1] : static_cast<float>(xk1[1] - xk[1]),
                    d2 = geo ? geo[2] : static_cast<float>(xk1[2] - xk[2]);
        l = geo ? geo[3] : len32(d0, d1, d2);
        const float il = rcp_(l);
        t[0] = d0 * il;
        t[1] = d1 * il;
        t[2] = d2 * il;
      }
      R e, ie;
      if constexpr (sizeof(R) == 8) {
        e = l / K.l0;
        ie = R(0);
      } else {
        e = l * K.il0;
        ie = rcp_(e);
      }
      // shear / stretch strain sigma = e (Q t) - z
      R qt[3];
#pragma unroll
      for (int p = 0; p < 3; ++p) qt[p] = madd(qr[3 * p + 2], t[2], madd(qr[3 * p + 1], t[1], qr[3 * p] * t[0]));
      R n[3];
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        R sig = p == 2 ? madd(qt[p], e, R(-1)) : qt[p] * e;
        n[p] = K.S[p] * (sig - K.sig0[p]);
      }
      // Q^T n / e
      R qn[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) qn[j] = madd(qr[6 + j], n[2], madd(qr[3 + j], n[1], qr[j] * n[0]));
      if constexpr (sizeof(R) == 8) {
#pragma unroll
        for (int j = 0; j < 3; ++j) out.c[j] = qn[j] / e;
      } else {
#pragma unroll
        for (int j = 0; j < 3; ++j) out.c[j] = qn[j] * ie;
      }
      out.e = e;
      // local torque terms: (Q t x n) l0 + (J w / e) x w + (J w / e) (t . dv) / (e l0)
      R sh[3];
      cross3(qt, n, sh);
      R w[3] = {static_cast<R>(wk[0]), static_cast<R>(wk[1]), static_cast<R>(wk[2])};
      R jw[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if constexpr (sizeof(R) == 8) jw[c] = (K.J[c] * w[c]) / e;
        else jw[c] = (K.J[c] * w[c]) * ie;
      }
      R lag[3];
      cross3(jw, w, lag);
      const R dv0 = static_cast<R>(vk1[0] - vk[0]), dv1 = static_cast<R>(vk1[1] - vk[1]), dv2 = static_cast<R>(vk1[2] - vk[2]);
      const R tdv = madd(t[2], dv2, madd(t[1], dv1, t[0] * dv0));
      R rate;
      if constexpr (sizeof(R) == 8) rate = tdv / (e * K.l0);
      else rate = tdv * (ie * K.il0);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        out.shear[c] = sh[c] * K.l0;
        out.lag[c] = lag[c];
        out.unst[c] = jw[c] * rate;
      }
      // Voronoi k.  A rod's last element has none.  FP32-bulk computes the
      // terms anyway: its constants hold B = 0, kappa0 = 0 and D0 = 1, its
      // neighbour inputs are finite values of the next lane and its log map
      // clamps its argument, so the couple M = B (kappa - kappa0) and both
      // terms below are exactly zero, with no branch (which every wave with a
      // rod end in it would take both sides of).  The FP64 path keeps the
      // reference's acos(tr / 2 - 1/2 - 1e-10), which is not defined for a
      // relative rotation of pi (a rod's last director against an idle lane's
      // identity), so it branches.
      if (sizeof(R) == 4 || has_vor) {
        R l1;
        if constexpr (sizeof(R) == 8) {
          const double f0 = xk2[0] - xk1[0], f1 = xk2[1] - xk1[1], f2 = xk2[2] - xk1[2];
          l1 = len_<R>(f0, f1, f2);
        } else {
          l1 = geo ? geo[4] : len32(dn[0], dn[1], dn[2]);
        }
        const R D = R(0.5) * (l + l1);
        R eps;
        R kap[3];
        if constexpr (sizeof(R) == 8) {
          eps = D / K.D0;
          curvature<R>(qk, qk1, K.D0, kap);
        } else {
          eps = D * K.iD0;
          curvature<R>(qk, qk1, K.iD0, kap);
        }
        R M[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) M[c] = K.B[c] * (kap[c] - K.kap0[c]);
        const R ic = rcp_(eps * eps * eps);
        R km[3];
        cross3(kap, M, km);
        const R icd = ic * K.D0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          out.P[c] = M[c] * ic;
          out.Rv[c] = km[c] * icd;
        }
      } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          out.P[c] = R(0);
          out.Rv[c] = R(0);
        }
      }
    }

    // external load on a node: gravity and per-node plane contact.  kLevel:
    // the plane z = z0 (normal (0, 0, 1)), where the general form reduces
    // exactly, up to the sign of zeros, to the terms kept here
    template <bool kLevel>
    __device__ __forceinline__ void node_environment(const Env& env, double m, const double x[3], const double v[3], double fe[3]) {
      if (env.gravity)
        for (int c = 0; c < 3; ++c) fe[c] += env.g[c] * m;
      if (!env.contact) return;
      double xn, vnv;  // height and velocity along the normal
      if constexpr (kLevel) {
        xn = x[2];
        vnv = v[2];
      } else {
        xn = (env.n[0] * x[0] + env.n[1] * x[1]) + env.n[2] * x[2];
        vnv = (env.n[0] * v[0] + env.n[1] * v[1]) + env.n[2] * v[2];
      }
      const double gap = xn - env.z0 - env.offset;
      if (gap < 0.0) {
        const double delta = -gap;
        const double vn = -vnv;
        const double fn = fmax(0.0, env.k * delta + env.c * vn);
        double vt[3], vtn;  // slip in the plane
        if constexpr (kLevel) {
          vt[0] = v[0];
          vt[1] = v[1];
          vt[2] = 0.0;
          vtn = sqrt(vt[0] * vt[0] + vt[1] * vt[1]);
          fe[2] += fn;
        } else {
#pragma unroll
          for (int c = 0; c < 3; ++c) vt[c] = v[c] - vnv * env.n[c];
          vtn = sqrt((vt[0] * vt[0] + vt[1] * vt[1]) + vt[2] * vt[2]);
#pragma unroll
          for (int c = 0; c < 3; ++c) fe[c] += fn * env.n[c];
        }
        if (vtn > 0.0) {
          const double ft = fmin(env.ct * vtn, env.mu * fn);
          const double s = ft / vtn;  // f_T along the unit slip: one division
#pragma unroll
          for (int c = 0; c < (kLevel ? 2 : 3); ++c) fe[c] -= s * vt[c];
        }
      }
    }

    // FP32-bulk: the same law, with the node's height above the plane formed
    // in FP64 (a resting node's gap is a small difference of heights) and
    // the rest in FP32 like the path's other loads: no FP64 root or division
    template <bool kLevel>
    __device__ __forceinline__ void node_environment32(const Env& env, float m, const double x[3], const double v[3],
                                                       float fe[3]) {
      if (env.gravity)
        for (int c = 0; c < 3; ++c) fe[c] = madd(static_cast<float>(env.g[c]), m, fe[c]);
      if (!env.contact) return;
      double xn;
      if constexpr (kLevel) xn = x[2];
      else xn = (env.n[0] * x[0] + env.n[1] * x[1]) + env.n[2] * x[2];
      const double gap = xn - env.z0 - env.offset;
      if (gap < 0.0) {
        const float delta = static_cast<float>(-gap);
        const float vv[3] = {static_cast<float>(v[0]), static_cast<float>(v[1]), static_cast<float>(v[2])};
        const float nn[3] = {static_cast<float>(env.n[0]), static_cast<float>(env.n[1]), static_cast<float>(env.n[2])};
        const float vnv = kLevel ? vv[2] : madd(nn[2], vv[2], madd(nn[1], vv[1], nn[0] * vv[0]));
        const float fn = fmaxf(0.0f, madd(static_cast<float>(env.k), delta, static_cast<float>(env.c) * -vnv));
        float vt[3], vtn;
        if constexpr (kLevel) {
          vt[0] = vv[0];
          vt[1] = vv[1];
          vt[2] = 0.0f;
          vtn = sqrt_(madd(vt[1], vt[1], vt[0] * vt[0]));
          fe[2] += fn;
        } else {
#pragma unroll
          for (int c = 0; c < 3; ++c) vt[c] = madd(-vnv, nn[c], vv[c]);
          vtn = sqrt_(madd(vt[2], vt[2], madd(vt[1], vt[1], vt[0] * vt[0])));
#pragma unroll
          for (int c = 0; c < 3; ++c) fe[c] = madd(fn, nn[c], fe[c]);
        }
        if (vtn > 0.0f) {
          const float s = fminf(static_cast<float>(env.ct) * vtn, static_cast<float>(env.mu) * fn) * rcp_(vtn);
#pragma unroll
          for (int c = 0; c < (kLevel ? 2 : 3); ++c) fe[c] = madd(-s, vt[c], fe[c]);
        }
      }
    }

    // ---- the kernel ------------------------------------------------------------------

    // kEnv: 0 no environment, 1 gravity and / or contact with a level plane,
    // 2 contact with a plane of any normal
    //
    // The entries of a tile are whole rods, or, with `segs`, pieces of rods
    // (Seg): a piece runs as a rod of its own length whose elements outside
    // [w0, w1) are not written back, and whose cut ends take the rod-end
    // rules with neither clamp nor end force (its halo absorbs the error).
    // Rod clocks are read from time_in and written to time_out, step budgets
    // read from steps_in and, when steps_out is set, written there: for whole
    // rods the same arrays; over pieces, which share their rod's clock and
    // budget across tiles, the clock goes to a second array (from the piece
    // holding the rod's last element) and the budget is spent after the launch.
    template <typename R, int P, int NT, int kEnv>
    __global__ void __launch_bounds__(NT) k_fused(StatePtrs in, StatePtrs out, Consts<R> C, const RodInfo* __restrict__ rods,
                                                  const Tile* __restrict__ tiles, const uint32_t* __restrict__ tile_rods,
                                                  const uint32_t* __restrict__ tile_lane0, const Seg* __restrict__ segs,
                                                  const double* time_in, double* time_out, const uint32_t* steps_in,
                                                  uint32_t* steps_out, Env env, int nsteps, uint64_t E) {
      constexpr bool kCache = sizeof(R) == 4;  // FP32 constants stay in registers for the launch
      // FP32-bulk: the closing half step of one step and the opening half step
      // of the next (same v and w) are taken as one update of dt within a launch
      constexpr bool kMerge = sizeof(R) == 4;
      const uint64_t N = E + 1;
      const Tile tile = tiles[blockIdx.x];
      const int tid = threadIdx.x;

      // exchange 1: per thread, its first element's node x, v and director, and
      // the thread's second node (FP64), or the FP32 difference of its second
      // and first node (FP32-bulk; for P = 1 only when the element is the rod's
      // last, the second node being the rod's tail)
      constexpr int kS1 = kCache ? 15 : 18;
      __shared__ double s1[kS1][NT];
      __shared__ float s5[kCache ? 4 : 1][NT];
      // exchange 2: per thread, its last element's c, P, R, in column tid + 1
      // (column 0 stays zero: the -1 terms of the tile's first element)
      __shared__ R s3[9][NT + 1];
      // FP32-bulk: 18 constants of every lane of the tile, for the whole launch:
      //   0 m, 1 l0, 2-4 sig0, 5-7 S, 8-10 J, 11 D0, 12-14 kap0, 15-17 B
      // (1/m, 1/J, 1/l0, 1/D0 are formed from it once per launch; the defaults of an element
      // without a Voronoi region are stored for it).  Lane el = t P + j is kept
      // at j NT + t, so that a wave reading element j of its threads reads
      // consecutive words.
      __shared__ float sc[kCache ? 18 : 1][kCache ? P * NT : 1];

      // ---- per element: rod, local index, slots, state -----------------------
      int rodi[P];
      int kk[P], nn[P];
      bool valid[P];
      uint64_t slot[P];
      double x[P][3], v[P][3], q[P][9], w[P][3];
      double tm[P], dt[P];
      // per element: the rod's flags (bit1 end force, bit2 damping per node)
      // with 1 + the clamp (0 or 1; 0: none) holding the element's node in
      // bits 8-9, the element itself in bits 12-13, and, on the element
      // holding the thread's tail node, that node in bits 10-11.  A clamp at
      // index i holds node i and element min(i, n - 1).  Clamps are matched
      // by the rod's own indices, so a piece applies those of its halo too.
      uint32_t rflags[P];
      uint32_t lim[P];  // steps this launch may take for the element's rod (its step budget)
      bool own[P];      // written back (false in a piece's halo)
      bool clockw[P];   // writes its rod's clock and budget
      bool tailw = false;  // the thread's tail node is its rod's last node (written back)
      // damping rates of node and element j; FP32-bulk folds them, once the
      // constants are in LDS, into m gf and (m + m_next)/2 gt (tail_gf into
      // tail_m tail_gf)
      R gfj[P], gtj[P];
      R tail_m = R(0), tail_im = R(0), tail_gf = R(0);
      // FP32-bulk: reciprocals of the element constants, formed once per launch
      R inv_m[kCache ? P : 1], inv_l0[kCache ? P : 1], inv_D0[kCache ? P : 1], inv_J[kCache ? P : 1][3];
      double xt[3] = {0, 0, 0}, vt[3] = {0, 0, 0};
      int tailj = -1;  // P - 1 when the thread holds its rod's last element, else -1
      // FP32-bulk without environment: the end force on the rod's last node
      // is kept in registers for the launch (the ramp's times are read only
      // while a ramp is set), instead of two dependent rod-record loads per
      // step; the other instantiations have no registers to spare for it
      constexpr bool kTailForce = kCache && kEnv == 0;
      double tforce[3] = {0, 0, 0};
      int tramp = EROD_RAMP_NONE;
      // the tile's rod table (first lane, rod) in LDS, in the exchange-1
      // buffer until the first step: a tile holds at most NT rods (each rod
      // has at least two elements, and at least P), so one pass loads it and
      // the lookups below are LDS searches, not chains of dependent global loads
      uint32_t* const t_lane0 = reinterpret_cast<uint32_t*>(&s1[0][0]);
      uint32_t* const t_rod = t_lane0 + NT;
      if (uint32_t(tid) < tile.count) {
        t_lane0[tid] = tile_lane0[tile.first + tid];
        t_rod[tid] = tile_rods[tile.first + tid];
      }
      __syncthreads();
      // the table entry of lane el: the last with first lane <= el
      auto find = [&](uint32_t el) {
        uint32_t lo = 0, hi = tile.count;
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (t_lane0[mid] <= el) lo = mid;
          else hi = mid;
        }
        return lo;
      };
      // pass 1: which rod and slot each of the thread's elements is.  Every
      // entry of a tile takes a whole number of threads: its first lane is a
      // thread's first, and (P - n mod P) mod P idle pad lanes come before its
      // elements, so that its last element is a thread's last (lane P - 1).
      // A pad lane computes on stand-in state (zero node state, the entry's
      // first director) with the entry's first-element constants and no
      // Voronoi region (zero couple), like an idle lane.
      bool pad[P];
      // the fields of the rod records first, for the P elements together, so
      // that their loads go out before any of them is used (one round trip,
      // not a chain of loads each consumed under a branch)
      uint32_t e_[P], r_n[P], r_fl[P], r_ci[P][2], s_off[P], s_n[P], s_w0[P], s_w1[P], s_ends[P];
      uint64_t r_start[P];
      double r_gf[P], r_gt[P];
#pragma unroll
      for (int j = 0; j < P; ++j) {
        e_[j] = find(uint32_t(tid) * P + j);
        const uint32_t ent = t_rod[e_[j]];
        uint32_t r;
        if (segs) {
          const Seg sg = segs[ent];
          r = sg.rod;
          s_off[j] = sg.off;
          s_n[j] = sg.n;
          s_w0[j] = sg.w0;
          s_w1[j] = sg.w1;
          s_ends[j] = sg.ends;
        } else {
          r = ent;
          s_off[j] = 0;
          s_w0[j] = 0;
          s_ends[j] = kSegFirst | kSegLast;
        }
        rodi[j] = int(r);
        const RodInfo* rp = rods + r;
        r_start[j] = rp->start;
        r_n[j] = rp->n;
        r_fl[j] = rp->flags;
        dt[j] = rp->dt;
        r_gf[j] = rp->gamma_f;
        r_gt[j] = rp->gamma_t;
        r_ci[j][0] = rp->clamp_idx[0];
        r_ci[j][1] = rp->clamp_idx[1];
        tm[j] = time_in[r];
        lim[j] = steps_in ? steps_in[r] : kNoBudget;
      }
#pragma unroll
      for (int j = 0; j < P; ++j) {
        const uint32_t el = uint32_t(tid) * P + j, e = e_[j];
        const uint32_t segn = segs ? s_n[j] : r_n[j], off = s_off[j], ends = s_ends[j];
        const uint32_t w0 = s_w0[j], w1 = segs ? s_w1[j] : segn;
        const uint32_t padn = (uint32_t(P) - segn % uint32_t(P)) % uint32_t(P), raw = el - t_lane0[e];
        pad[j] = el < tile.lanes && raw < padn;
        valid[j] = el < tile.lanes && raw >= padn;
        kk[j] = valid[j] ? int(raw - padn) : 0;
        nn[j] = int(segn);
        slot[j] = r_start[j] + uint64_t(off) + uint64_t(kk[j]);
        // the end force acts at the rod's last node
        rflags[j] = r_fl[j] & ~kClampBits & ((ends & kSegLast) ? ~0u : ~2u);
        const uint32_t g = off + uint32_t(kk[j]);
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          if (!valid[j] || !(r_fl[j] & clamp_bit(k))) continue;
          const uint32_t i = r_ci[j][k];
          if (i == g) rflags[j] |= uint32_t(k + 1) << 8;
          if (kk[j] == int(segn) - 1 && i == off + segn) rflags[j] |= uint32_t(k + 1) << 10;
          if ((i < r_n[j] ? i : r_n[j] - 1) == g) rflags[j] |= uint32_t(k + 1) << 12;
        }
        own[j] = valid[j] && uint32_t(kk[j]) >= w0 && uint32_t(kk[j]) < w1;
        clockw[j] = valid[j] && kk[j] == 0 && (segs == nullptr || (ends & kSegLast) != 0);
        if (j == P - 1 && valid[j] && kk[j] == int(segn) - 1) tailw = (ends & kSegLast) != 0;
        // damping rates: one per rod, or (flag bit 2) per node and element
        const bool per_node = (r_fl[j] & 4u) != 0 && valid[j];
        gfj[j] = per_node ? C.gf[slot[j]] : static_cast<R>(r_gf[j]);
        gtj[j] = per_node ? C.gt[slot[j]] : static_cast<R>(r_gt[j]);
      }
      // pass 2: the state, one field line at a time over the P elements (their
      // slots are contiguous within a rod, so the P loads of a line share cache
      // lines and are issued back to back)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int j = 0; j < P; ++j) x[j][c] = valid[j] ? in.x[c * N + slot[j]] : 0.0;
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int j = 0; j < P; ++j) v[j][c] = valid[j] ? in.v[c * N + slot[j]] : 0.0;
      }
      // (a pad lane takes its entry's first director: the relative rotation
      // between them is zero, so the pad's log map stays on its series branch)
#pragma unroll
      for (int m = 0; m < 9; ++m) {
#pragma unroll
        for (int j = 0; j < P; ++j)
          q[j][m] = valid[j] || pad[j] ? in.q[m * E + slot[j]] : ((m % 4 == 0) ? 1.0 : 0.0);
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int j = 0; j < P; ++j) w[j][c] = valid[j] ? in.w[c * E + slot[j]] : 0.0;
      }
      // FP32-bulk: the thread's constants into the LDS table, from the slots
      // found above, loaded straight into LDS right after the state loads:
      // nothing waits between them, so both are in flight together, and the
      // table takes no registers.  An element without a Voronoi region gets
      // the defaults written instead (other entries: the two writes never meet)
      if constexpr (kCache) {
        const uint64_t V = E - 1;
        const uint32_t wave0 = uint32_t(tid) & ~63u;
#pragma unroll
        for (int j = 0; j < P; ++j) {
          const uint64_t sl = slot[j];  // (a pad lane: the entry's first element)
          const uint32_t cw = uint32_t(j) * NT + wave0;  // lane (wave0 + l) P + j is entry cw + l
          if (valid[j] || pad[j]) {
            load_to_lds(C.m + sl, &sc[0][cw]);
            load_to_lds(C.l0 + sl, &sc[1][cw]);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              load_to_lds(C.sig0 + c * E + sl, &sc[2 + c][cw]);
              load_to_lds(C.S + c * E + sl, &sc[5 + c][cw]);
              load_to_lds(C.J + c * E + sl, &sc[8 + c][cw]);
            }
          }
          if (valid[j] && kk[j] + 1 < nn[j]) {
            load_to_lds(C.D0 + sl, &sc[11][cw]);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              load_to_lds(C.kap0 + c * V + sl, &sc[12 + c][cw]);
              load_to_lds(C.B + c * V + sl, &sc[15 + c][cw]);
            }
          } else {
            const uint32_t ce = uint32_t(j) * NT + tid;  // lane tid P + j
            sc[11][ce] = 1.0f;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              sc[12 + c][ce] = 0.0f;
              sc[15 + c][ce] = 0.0f;
            }
          }
        }
      }
#pragma unroll
      for (int j = 0; j < P; ++j) {
        if (j == P - 1 && valid[j] && kk[j] == nn[j] - 1) {
          tailj = j;
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            xt[c] = in.x[c * N + slot[j] + 1];
            vt[c] = in.v[c * N + slot[j] + 1];
          }
          tail_m = C.m[slot[j] + 1];
          tail_im = C.invm[slot[j] + 1];
          tail_gf = (rods[rodi[j]].flags & 4u) ? C.gf[slot[j] + 1] : gfj[j];
          if (kTailForce && (rflags[j] & 2u)) {
            const RodInfo& ri = rods[rodi[j]];
#pragma unroll
            for (int c = 0; c < 3; ++c) tforce[c] = ri.force[c];
            tramp = ri.ramp;
          }
        }
      }
      if constexpr (kCache) {
        __syncthreads();  // the constant table is written
        // per element, for the launch: 1/m, 1/J, 1/l0, 1/D0 and the damping
        // products m gf and (m + m_next)/2 gt (the same FP32 operations as
        // forming them every step, so the same bits)
#pragma unroll
        for (int j = 0; j < P; ++j) {
          const uint32_t el = uint32_t(j) * NT + tid;                   // lane tid P + j
          const uint32_t eln = j + 1 < P ? el + NT : uint32_t(tid) + 1;  // lane tid P + j + 1
          const R mk = sc[0][el];
          const R mkn = j == P - 1 && j == tailj ? tail_m : sc[0][eln];
          inv_m[j] = rcp_(mk);
          inv_l0[j] = rcp_(sc[1][el]);
          inv_D0[j] = rcp_(sc[11][el]);
#pragma unroll
          for (int c = 0; c < 3; ++c) inv_J[j][c] = rcp_(sc[8 + c][el]);
          gfj[j] = mk * gfj[j];
          gtj[j] = (R(0.5) * (mk + mkn)) * gtj[j];
        }
        tail_gf = tail_m * tail_gf;
      }
      if constexpr (!kCache) __syncthreads();  // every lookup in the rod table done before s1 is written
      if (tid < 9) s3[tid][0] = R(0);  // read only after the barriers of the first step
      // whether a lane of this wave holds a clamped node or element: a
      // wave-uniform value, so that the waves without clamps (most of a
      // batch) branch over the clamp updates instead of selecting through
      // them every step
      uint32_t cbits = 0;
#pragma unroll
      for (int j = 0; j < P; ++j) cbits |= rflags[j];
      const bool wclamp = __ballot((cbits & (0x3fu << 8)) != 0u) != 0ull;
      // the write-back predicates as bits of one register (not lane masks the
      // step loop would carry in scalar registers): bit j element j owned,
      // bit 8 + j element j writes its rod's clock, bit 16 the tail node;
      // an element that is no rod's gets no steps
      uint32_t wb = tailw ? 1u << 16 : 0u;
#pragma unroll
      for (int j = 0; j < P; ++j) {
        wb |= (own[j] ? 1u : 0u) << j;
        wb |= (clockw[j] ? 1u : 0u) << (8 + j);
        lim[j] = valid[j] ? lim[j] : 0u;
      }
      asm volatile("" : "+v"(wb));
      for (int step = 0; step < nsteps; ++step) {
        // a rod whose budget is spent keeps its state (its element forces are
        // still evaluated, since the tile moves in lockstep, but never applied)
        bool live[P];
#pragma unroll
        for (int j = 0; j < P; ++j) live[j] = uint32_t(step) < lim[j];
        // ---- kinematic half step, time, post-kinematic clamp --------------------
        // (FP32-bulk: after the first step of a launch this half step was taken
        // together with the closing half step of the previous one)
#pragma unroll
        for (int j = 0; j < P; ++j) {
          if (!live[j]) continue;
          if (!kMerge || step == 0) {
            const double ch = 0.5 * dt[j];
#pragma unroll
            for (int c = 0; c < 3; ++c) x[j][c] = maddd<R>(ch, v[j][c], x[j][c]);
            rodrigues<R>(q[j], w[j], ch);
            if (j == P - 1 && j == tailj)
#pragma unroll
              for (int c = 0; c < 3; ++c) xt[c] = maddd<R>(ch, vt[c], xt[c]);
          }
          tm[j] += 0.5 * dt[j];
          if (!wclamp) continue;
          if (rflags[j] & (3u << 8)) {
            const RodInfo& ri = rods[rodi[j]];
            int k = int((rflags[j] >> 8) & 3u) - 1;
            // (k opaque here, so the pose's address is formed in this rarely
            // taken branch instead of being kept in registers for the launch)
            asm volatile("" : "+v"(k));
#pragma unroll
            for (int c = 0; c < 3; ++c) x[j][c] = ri.clamp[k].x[c];
          }
          if (rflags[j] & (3u << 12)) {
            const RodInfo& ri = rods[rodi[j]];
            int k = int((rflags[j] >> 12) & 3u) - 1;
            asm volatile("" : "+v"(k));
#pragma unroll
            for (int m = 0; m < 9; ++m) q[j][m] = ri.clamp[k].q[m];
          }
          if (rflags[j] & (3u << 10)) {
            const RodInfo& ri = rods[rodi[j]];
            int k = int((rflags[j] >> 10) & 3u) - 1;
            asm volatile("" : "+v"(k));
#pragma unroll
            for (int c = 0; c < 3; ++c) xt[c] = ri.clamp[k].x[c];
          }
        }
        // ---- exchange 1 -----------------------------------------------------------
        // FP32-bulk with P > 1: per element, the FP32-rounded node difference
        // and its length (geo[j] = {d, |d|, next element's |d|}), the first
        // element's before the exchange, which publishes both
        constexpr bool kGeo = kCache && P > 1;
        float geo[kGeo ? P : 1][5];
        {
          const bool last0 = P == 1 && valid[0] && (kk[0] == nn[0] - 1);
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            s1[c][tid] = x[0][c];
            s1[3 + c][tid] = v[0][c];
            const double xs = last0 ? xt[c] : (P > 1 ? x[P > 1 ? 1 : 0][c] : 0.0);
            if constexpr (!kCache) s1[kCache ? 0 : 15 + c][tid] = xs;
            else {
              geo[0][c] = static_cast<float>(xs - x[0][c]);
              s5[kCache ? c : 0][tid] = geo[0][c];
            }
          }
          if constexpr (kCache) {
            geo[0][3] = len32(geo[0][0], geo[0][1], geo[0][2]);
            s5[kCache ? 3 : 0][tid] = geo[0][3];
          }
#pragma unroll
          for (int m = 0; m < 9; ++m) s1[6 + m][tid] = q[0][m];
        }
        __syncthreads();
        const int tn = tid + 1 < NT ? tid + 1 : tid;
        const int tn2 = tid + 2 < NT ? tid + 2 : tn;
        if constexpr (kGeo) {
#pragma unroll
          for (int j = 1; j < P; ++j) {
            const bool lastj = j == P - 1 && kk[j] == nn[j] - 1;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              const double xn = lastj ? xt[c] : (j + 1 < P ? x[j + 1 < P ? j + 1 : j][c] : s1[c][tn]);
              geo[kGeo ? j : 0][c] = static_cast<float>(xn - x[j][c]);
            }
            geo[kGeo ? j : 0][3] = len32(geo[kGeo ? j : 0][0], geo[kGeo ? j : 0][1], geo[kGeo ? j : 0][2]);
          }
#pragma unroll
          for (int j = 0; j < P; ++j) geo[kGeo ? j : 0][4] = j + 1 < P ? geo[kGeo ? (j + 1 < P ? j + 1 : j) : 0][3] : s5[kCache ? 3 : 0][tn];
        }
        // Idle lanes (past the tile's last entry) and pad lanes compute on
        // their stand-in state and constants too.  A tile's entries end on a
        // thread's last lane, so an idle lane's terms reach only idle lanes;
        // a pad lane's reach only its entry's first element, which masks the
        // force term and reads couple terms that are zero (the pad has no
        // Voronoi region).  None are applied; a branch that zeroed them would
        // cost every wave its moves each step.
        Mid<R> mid[P];
#pragma unroll
        for (int j = 0; j < P; ++j) {
          if (kEnv == 2 && !valid[j]) {
            // (the tilted plane's kernel keeps the branch: without it, the
            // compiler's schedule for it needs more than 256 registers)
#pragma unroll
            for (int c = 0; c < 3; ++c)
              mid[j].c[c] = mid[j].P[c] = mid[j].Rv[c] = mid[j].shear[c] = mid[j].lag[c] = mid[j].unst[c] = R(0);
            mid[j].e = R(1);
            continue;
          }
          const bool last = j == P - 1 && kk[j] == nn[j] - 1;
          double xk1[3], xk2[3] = {0, 0, 0}, vk1[3], qk1[9];
          float dn[3] = {0.0f, 0.0f, 0.0f};
          if (last) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              xk1[c] = xt[c];
              vk1[c] = vt[c];
            }
          } else if (j + 1 < P) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              xk1[c] = x[j + 1 < P ? j + 1 : j][c];
              vk1[c] = v[j + 1 < P ? j + 1 : j][c];
            }
          } else {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              xk1[c] = s1[c][tn];
              vk1[c] = s1[3 + c][tn];
            }
          }
          const bool has_vor = !last && !pad[j];
          // the next lane's director and node difference; for a rod's last
          // element they belong to the next rod (or to an idle lane: identity
          // director, zero state) and only meet its zero Voronoi constants
          // (FP32-bulk) or are not used (FP64)
          {
            if (j + 1 < P) {
#pragma unroll
              for (int m = 0; m < 9; ++m) qk1[m] = q[j + 1 < P ? j + 1 : j][m];
            } else {
#pragma unroll
              for (int m = 0; m < 9; ++m) qk1[m] = s1[6 + m][tn];
            }
            const bool next_last = (j + 1) % P == P - 1 && kk[j] + 1 == nn[j] - 1;
            if (kCache && j == P - 1 && (P > 1 || next_last)) {
              // FP32-bulk: the next element's rounded difference, from its thread
#pragma unroll
              for (int c = 0; c < 3; ++c) dn[c] = s5[kCache ? c : 0][tn];
            } else {
              if (next_last) {
                if (j + 1 < P) {
#pragma unroll
                  for (int c = 0; c < 3; ++c) xk2[c] = xt[c];
                } else {
#pragma unroll
                  for (int c = 0; c < 3; ++c) xk2[c] = s1[kCache ? 0 : 15 + c][tn];
                }
              } else if (j + 2 < P) {
#pragma unroll
                for (int c = 0; c < 3; ++c) xk2[c] = x[j + 2 < P ? j + 2 : j][c];
              } else if (j + 2 == P) {
#pragma unroll
                for (int c = 0; c < 3; ++c) xk2[c] = s1[c][tn];
              } else if (P > 1) {
#pragma unroll
                for (int c = 0; c < 3; ++c) xk2[c] = s1[kCache ? 0 : 15 + c][tn];
              } else {
#pragma unroll
                for (int c = 0; c < 3; ++c) xk2[c] = s1[c][tn2];
              }
#pragma unroll
              for (int c = 0; c < 3; ++c) dn[c] = static_cast<float>(xk2[c] - xk1[c]);
            }
          }
          ElemConst<R> K;
          if constexpr (kCache) {
            const uint32_t el = uint32_t(j) * NT + tid;  // lane tid P + j
            K.l0 = sc[1][el];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              K.sig0[c] = sc[kCache ? 2 + c : 0][kCache ? el : 0];
              K.S[c] = sc[kCache ? 5 + c : 0][kCache ? el : 0];
              K.J[c] = sc[kCache ? 8 + c : 0][kCache ? el : 0];
              K.kap0[c] = sc[kCache ? 12 + c : 0][kCache ? el : 0];
              K.B[c] = sc[kCache ? 15 + c : 0][kCache ? el : 0];
            }
            K.D0 = sc[kCache ? 11 : 0][kCache ? el : 0];
            K.il0 = inv_l0[kCache ? j : 0];
            K.iD0 = inv_D0[kCache ? j : 0];
          } else {
            load_const<R>(C, slot[j], E, has_vor, K);
          }
          R qr[9];
#pragma unroll
          for (int m = 0; m < 9; ++m) qr[m] = static_cast<R>(q[j][m]);
          element_forces<R>(x[j], xk1, xk2, dn, v[j], vk1, q[j], qr, qk1, w[j], has_vor, K, mid[j],
                            kGeo ? geo[kGeo ? j : 0] : nullptr);
        }
        // ---- exchange 2 -----------------------------------------------------------
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          s3[c][tid + 1] = mid[P - 1].c[c];
          s3[3 + c][tid + 1] = mid[P - 1].P[c];
          s3[6 + c][tid + 1] = mid[P - 1].Rv[c];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < P; ++j) {
          if (!live[j]) continue;
          const int k = kk[j];
          const bool first = k == 0, last = j == P - 1 && k == nn[j] - 1;
          // the -1 element's terms (the previous thread's from exchange 2 for
          // j = 0).  Before a rod's first element sits the previous rod's last
          // element, whose P and R are zero (no Voronoi region), or the zero
          // column; only its c is masked.
          R cm[3], Pm[3], Rm[3];
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            cm[c] = j > 0 ? mid[j > 0 ? j - 1 : 0].c[c] : s3[c][tid];
            Pm[c] = j > 0 ? mid[j > 0 ? j - 1 : 0].P[c] : s3[3 + c][tid];
            Rm[c] = j > 0 ? mid[j > 0 ? j - 1 : 0].Rv[c] : s3[6 + c][tid];
            cm[c] = first ? R(0) : cm[c];
          }
          const uint64_t s = slot[j];
          const Mid<R>& md = mid[j];
          R mk, im, iJ[3], mg, eg;
          if constexpr (kCache) {
            // (the mass itself only for the environment's loads)
            mk = kEnv != 0 ? sc[0][uint32_t(j) * NT + tid] : R(0);
            im = inv_m[kCache ? j : 0];
#pragma unroll
            for (int c = 0; c < 3; ++c) iJ[c] = inv_J[kCache ? j : 0][c];
            mg = gfj[j];
            eg = gtj[j];
          } else {
            mk = C.m[s];
            const R mkn = C.m[s + 1];
            im = C.invm[s];
#pragma unroll
            for (int c = 0; c < 3; ++c) iJ[c] = C.iJ[c * E + s];
            mg = mk * gfj[j];
            eg = (R(0.5) * (mk + mkn)) * gtj[j];
          }
          // internal force on node k: difference of c, then damping
          // The rod-end cases of the difference and quadrature operators need no
          // branches: the -1 neighbour terms are zero at the first element and
          // the Voronoi terms (P, Rv) are zero at the last one, and x - 0, 0 - x
          // and 0 + x are exact, as with the reference's zero ghosts.
          R f[3];
          // (madd is one FMA for FP32-bulk and, for FP64, a product and a sum
          // rounded on their own: the values of the reference's expressions)
          R vk[3] = {static_cast<R>(v[j][0]), static_cast<R>(v[j][1]), static_cast<R>(v[j][2])};
#pragma unroll
          for (int c = 0; c < 3; ++c) f[c] = madd(-mg, vk[c], md.c[c] - cm[c]);
          // internal torque on element k
          R tau[3];
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const R dq = madd(R(0.5), md.Rv[c] + Rm[c], md.P[c] - Pm[c]);
            tau[c] = madd(-eg, static_cast<R>(w[j][c]), ((dq + md.shear[c]) + md.lag[c]) + md.unst[c]);
          }
          // external loads of node k (forcing, then environment)
          const double dtt = dt[j];
          if constexpr (kCache) {
            float fe[3] = {0.0f, 0.0f, 0.0f};
            if constexpr (kEnv != 0) node_environment32<kEnv == 1>(env, mk, x[j], v[j], fe);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              // (without an environment the external load is zero)
              const R a = (kEnv != 0 ? f[c] + fe[c] : f[c]) * im;
              v[j][c] = maddd<R>(dtt, static_cast<double>(a), v[j][c]);
            }
          } else {
            double fe[3] = {0.0, 0.0, 0.0};
            if constexpr (kEnv != 0) node_environment<kEnv == 1>(env, static_cast<double>(mk), x[j], v[j], fe);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              const R a = (kEnv != 0 ? f[c] + static_cast<R>(fe[c]) : f[c]) * im;
              v[j][c] = maddd<R>(dtt, static_cast<double>(a), v[j][c]);
            }
          }
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const R al = (iJ[c] * tau[c]) * md.e;
            w[j][c] = maddd<R>(dtt, static_cast<double>(al), w[j][c]);
          }
          // last node of the rod
          if (last) {
            const R mgn = kCache ? tail_gf : tail_m * tail_gf;
            double fen[3] = {0.0, 0.0, 0.0};
            if (rflags[j] & 2u) {
              if constexpr (kTailForce) {
                double rv = 1.0;
                if (tramp != EROD_RAMP_NONE) {
                  const RodInfo& ri = rods[rodi[j]];
                  rv = ramp_value(tramp, tm[j], ri.ramp_start, ri.ramp_interval);
                }
#pragma unroll
                for (int c = 0; c < 3; ++c) fen[c] = tforce[c] * rv;
              } else {
                const RodInfo& ri = rods[rodi[j]];
                const double rv = ramp_value(ri.ramp, tm[j], ri.ramp_start, ri.ramp_interval);
#pragma unroll
                for (int c = 0; c < 3; ++c) fen[c] = ri.force[c] * rv;
              }
            }
            if constexpr (kCache) {
              float fenv[3] = {0.0f, 0.0f, 0.0f};
              if constexpr (kEnv != 0) node_environment32<kEnv == 1>(env, tail_m, xt, vt, fenv);
#pragma unroll
              for (int c = 0; c < 3; ++c) {
                const R fn = madd(-mgn, static_cast<R>(vt[c]), -md.c[c]);
                const R a = (kEnv != 0 ? (fn + static_cast<R>(fen[c])) + fenv[c] : fn + static_cast<R>(fen[c])) * tail_im;
                vt[c] = maddd<R>(dtt, static_cast<double>(a), vt[c]);
              }
            } else {
              if constexpr (kEnv != 0) node_environment<kEnv == 1>(env, static_cast<double>(tail_m), xt, vt, fen);
#pragma unroll
              for (int c = 0; c < 3; ++c) {
                const R fn = madd(-mgn, static_cast<R>(vt[c]), -md.c[c]);
                const R a = (fn + static_cast<R>(fen[c])) * tail_im;
                vt[c] = maddd<R>(dtt, static_cast<double>(a), vt[c]);
              }
            }
            if (wclamp && (rflags[j] & (3u << 10)))
#pragma unroll
              for (int c = 0; c < 3; ++c) vt[c] = 0.0;
          }
          // post-dynamic clamp
          if (wclamp) {
            if (rflags[j] & (3u << 8))
#pragma unroll
              for (int c = 0; c < 3; ++c) v[j][c] = 0.0;
            if (rflags[j] & (3u << 12))
#pragma unroll
              for (int c = 0; c < 3; ++c) w[j][c] = 0.0;
          }
        }
        // ---- time, kinematic half step ------------------------------------------
        // FP32-bulk: unless this is the rod's last step of the launch, also take
        // the next step's opening half step (same v and w: one update of dt)
#pragma unroll
        for (int j = 0; j < P; ++j) {
          if (!live[j]) continue;
          tm[j] += 0.5 * dt[j];
          const bool both = kMerge && uint32_t(step + 1) < lim[j] && step + 1 < nsteps;
          const double ch = both ? dt[j] : 0.5 * dt[j];
#pragma unroll
          for (int c = 0; c < 3; ++c) x[j][c] = maddd<R>(ch, v[j][c], x[j][c]);
          rodrigues<R>(q[j], w[j], ch);
          if (j == P - 1 && j == tailj)
#pragma unroll
            for (int c = 0; c < 3; ++c) xt[c] = maddd<R>(ch, vt[c], xt[c]);
        }
      }

      // ---- write back (one field line at a time over the P elements) --------------
#pragma unroll
      for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int j = 0; j < P; ++j)
          if ((wb >> j) & 1u) out.x[c * N + slot[j]] = x[j][c];
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int j = 0; j < P; ++j)
          if ((wb >> j) & 1u) out.v[c * N + slot[j]] = v[j][c];
      }
#pragma unroll
      for (int m = 0; m < 9; ++m) {
#pragma unroll
        for (int j = 0; j < P; ++j)
          if ((wb >> j) & 1u) out.q[m * E + slot[j]] = q[j][m];
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int j = 0; j < P; ++j)
          if ((wb >> j) & 1u) out.w[c * E + slot[j]] = w[j][c];
      }
#pragma unroll
      for (int j = 0; j < P; ++j) {
        const uint64_t s = slot[j];
        if (j == P - 1 && j == tailj && ((wb >> 16) & 1u)) {
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            out.x[c * N + s + 1] = xt[c];
            out.v[c * N + s + 1] = vt[c];
          }
        }
        if ((wb >> (8 + j)) & 1u) {
          time_out[rodi[j]] = tm[j];
          if (steps_out && lim[j] != kNoBudget) steps_out[rodi[j]] = lim[j] > uint32_t(nsteps) ? lim[j] - uint32_t(nsteps) : 0u;
        }
      }
    }

    // the state of the rods cut into pieces, from the second copy back into
    // the fields (after an odd number of launches over pieces); one
    // workgroup per rod
    static __global__ void k_copy_cut_state(const uint32_t* __restrict__ seg_rods, const RodInfo* __restrict__ rods,
                                            StatePtrs from, StatePtrs to, uint64_t E) {
      const uint64_t N = E + 1;
      const RodInfo ri = rods[seg_rods[blockIdx.x]];
      for (uint32_t k = threadIdx.x; k <= ri.n; k += blockDim.x) {
        const uint64_t s = ri.start + k;
        for (int c = 0; c < 3; ++c) {
          to.x[c * N + s] = from.x[c * N + s];
          to.v[c * N + s] = from.v[c * N + s];
        }
        if (k < ri.n) {
          for (int m = 0; m < 9; ++m) to.q[m * E + s] = from.q[m * E + s];
          for (int c = 0; c < 3; ++c) to.w[c * E + s] = from.w[c * E + s];
        }
      }
    }

    // after a launch over pieces: the rods' clocks (written by the pieces
    // holding their last element) and budgets (spent by the launch's steps)
    static __global__ void k_seg_clocks(const uint32_t* __restrict__ seg_rods, uint64_t n, const double* __restrict__ time_next,
                                 double* __restrict__ rod_time, uint32_t* __restrict__ steps_left, uint32_t nsteps) {
      const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
      if (i >= n) return;
      const uint32_t r = seg_rods[i];
      rod_time[r] = time_next[r];
      if (steps_left) {
        const uint32_t l = steps_left[r];
        if (l != kNoBudget) steps_left[r] = l > nsteps ? l - nsteps : 0u;
      }
    }

    inline StatePtrs field_state(erod_batch* b) {
      return StatePtrs{b->d_field[EROD_POSITION], b->d_field[EROD_VELOCITY], b->d_field[EROD_DIRECTOR],
                       b->d_field[EROD_ANGULAR_VELOCITY]};
    }
    inline StatePtrs shadow_state(erod_batch* b) {
      return StatePtrs{b->d_shadow[0], b->d_shadow[1], b->d_shadow[2], b->d_shadow[3]};
    }

    template <typename R, int P, int NT, int kEnv>
    inline void launch_env(erod_batch* b, const Consts<R>& C, int steps, const TileGroup& g, const StatePtrs& in,
                           const StatePtrs& out, hipStream_t st) {
      uint32_t* budgets = b->budgets ? b->d_steps_left : nullptr;
      if (g.d_segs) {
        // pieces: clocks to the second array, budgets spent afterwards
        hipLaunchKernelGGL((k_fused<R, P, NT, kEnv>), dim3(static_cast<unsigned>(g.n_tiles)), dim3(NT), 0, st, in,
                           out, C, b->d_rods, g.d_tiles, g.d_tile_rods, g.d_tile_lane0, g.d_segs, b->d_time, b->d_time_next,
                           budgets, nullptr, b->env, steps, b->E);
        hipLaunchKernelGGL(k_seg_clocks, dim3(static_cast<unsigned>((g.n_seg_rods + 255) / 256)), dim3(256), 0, st,
                           g.d_seg_rods, g.n_seg_rods, b->d_time_next, b->d_time, budgets, uint32_t(steps));
      } else {
        hipLaunchKernelGGL((k_fused<R, P, NT, kEnv>), dim3(static_cast<unsigned>(g.n_tiles)), dim3(NT), 0, st, in,
                           out, C, b->d_rods, g.d_tiles, g.d_tile_rods, g.d_tile_lane0, nullptr, b->d_time, b->d_time,
                           budgets, budgets, b->env, steps, b->E);
      }
    }

    // batches without gravity or contact run a kernel without environment code
    // (and without the registers its parameters would hold); a level contact
    // plane runs one without the arithmetic of a general normal
    template <typename R, int P, int NT>
    inline int launch_tiles(erod_batch* b, const Consts<R>& C, int steps, const TileGroup& g, const StatePtrs& in,
                            const StatePtrs& out, hipStream_t st) {
      if (g.n_tiles == 0) return EROD_OK;
      const Env& e = b->env;
      const bool level = e.n[0] == 0.0 && e.n[1] == 0.0 && e.n[2] == 1.0;
      if (e.contact && !level) launch_env<R, P, NT, 2>(b, C, steps, g, in, out, st);
      else if (e.gravity || e.contact) launch_env<R, P, NT, 1>(b, C, steps, g, in, out, st);
      else launch_env<R, P, NT, 0>(b, C, steps, g, in, out, st);
      return hip_check(hipGetLastError(), "k_fused");
    }

    // the tile shapes the planner can choose: elements per thread x threads
    struct Shape {
      int P, NT;
    };
    constexpr Shape kShapes[] = {{1, 64}, {1, 128}, {1, 256}, {1, 512}, {2, 256}, {2, 512}, {4, 128}, {4, 256}};
    constexpr int kNShapes = sizeof(kShapes) / sizeof(kShapes[0]);

    template <typename R>
    inline int launch_shape(erod_batch* b, const Consts<R>& C, int steps, int shape, const TileGroup& g,
                            const StatePtrs& in, const StatePtrs& out, hipStream_t st) {
      switch (shape) {
        case 0: return launch_tiles<R, 1, 64>(b, C, steps, g, in, out, st);
        case 1: return launch_tiles<R, 1, 128>(b, C, steps, g, in, out, st);
        case 2: return launch_tiles<R, 1, 256>(b, C, steps, g, in, out, st);
        case 3: return launch_tiles<R, 1, 512>(b, C, steps, g, in, out, st);
        case 4: return launch_tiles<R, 2, 256>(b, C, steps, g, in, out, st);
        case 5: return launch_tiles<R, 2, 512>(b, C, steps, g, in, out, st);
        case 6: return launch_tiles<R, 4, 128>(b, C, steps, g, in, out, st);
        case 7: return launch_tiles<R, 4, 256>(b, C, steps, g, in, out, st);
        default: set_error("bad tile shape"); return EROD_ERR_INVALID;
      }
    }

    // one group of tiles advanced `steps` steps: one launch, or, for a group
    // with a limit on its steps per launch (pieces of rods), launches of at
    // most that many
    template <typename R>
    inline int launch_group(erod_batch* b, const Consts<R>& C, int steps, const TileGroup& g, hipStream_t st) {
      if (g.shape < 0 || g.n_tiles == 0) return EROD_OK;
      const StatePtrs fields = field_state(b);
      if (!g.d_segs) return launch_shape(b, C, steps, g.shape, g, fields, fields, st);
      // pieces: launches of at most max_steps steps, each reading the state
      // the previous one wrote and writing the other copy; the state ends in
      // the fields
      const StatePtrs shadow = shadow_state(b);
      const int cap = g.max_steps > 0 ? g.max_steps : steps;
      bool in_fields = true;
      for (int done = 0; done < steps;) {
        const int s = steps - done < cap ? steps - done : cap;
        if (int rc = launch_shape(b, C, s, g.shape, g, in_fields ? fields : shadow, in_fields ? shadow : fields, st))
          return rc;
        in_fields = !in_fields;
        done += s;
      }
      if (!in_fields) {
        hipLaunchKernelGGL(k_copy_cut_state, dim3(static_cast<unsigned>(g.n_seg_rods)), dim3(256), 0, st, g.d_seg_rods,
                           b->d_rods, shadow, fields, b->E);
        return hip_check(hipGetLastError(), "k_copy_cut_state");
      }
      return EROD_OK;
    }

    // `steps` steps over every tile of the plan: its main group, its second
    // group, the pieces of its long rods.  Rods never interact, so the groups
    // run side by side: with pieces, the whole-rod groups go to the batch's
    // second stream beside the chain of piece launches; without, the second
    // group goes there beside the main one.  Both streams join back before
    // the call returns (in a captured graph: two branches and a join).
    template <typename R>
    inline int launch_all(erod_batch* b, const Consts<R>& C, int steps) {
      const TileGroup* g = b->plan.g;
      auto busy = [](const TileGroup& t) { return t.shape >= 0 && t.n_tiles > 0; };
      const bool pieces = busy(g[2]);
      const int side = (pieces ? busy(g[0]) : 0) + busy(g[1]);
      if (side == 0 || (!pieces && !busy(g[0]))) {
        for (int k = 0; k < 3; ++k)
          if (int rc = launch_group(b, C, steps, g[k], b->stream)) return rc;
        return EROD_OK;
      }
      EROD_HIP(hipEventRecord(b->ev_fork, b->stream));
      EROD_HIP(hipStreamWaitEvent(b->stream2, b->ev_fork, 0));
      int rc = EROD_OK;
      if (pieces) rc = launch_group(b, C, steps, g[0], b->stream2);
      if (rc == EROD_OK) rc = launch_group(b, C, steps, g[1], b->stream2);
      int rc2 = launch_group(b, C, steps, pieces ? g[2] : g[0], b->stream);
      // the join is recorded whatever failed, so that a capture ends whole
      EROD_HIP(hipEventRecord(b->ev_join, b->stream2));
      EROD_HIP(hipStreamWaitEvent(b->stream, b->ev_join, 0));
      return rc ? rc : rc2;
    }

  }  // namespace fused

  int fused_launch_fp64(erod_batch* b, const fused::Consts<double>& C, int steps);
  int fused_launch_fp32(erod_batch* b, const fused::Consts<float>& C, int steps);

}  // namespace erod
