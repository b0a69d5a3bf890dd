// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A scalar C++20 CPU restatement of Elastica++'s FP64 Cosserat-rod block
// timestep.  It is the checker that every GPU path of this repository is
// compared against, and the CPU baseline timed beside the GPU.  Only the
// test suite, the smoke check and the benchmark's CPU-baseline leg may load
// it; the product library never links, loads or calls it.
//
// Layout: the reference's ghost-padded multi-rod Block
// (elastica/Systems/CosseratRods/Block.hpp:401-443,568-611,658-732).  Rod i
// starts at the same offset s_i in the node, element and Voronoi arrays with
// s_{i+1} = s_i + n_i + 2; each gap holds 1 ghost node, 2 ghost elements and
// 3 ghost Voronoi slots.  Every field is component-major: component c of
// slot k lives at data[c * slots + k]; a director is 9 lines indexed
// (page p, row r) -> line 3p + r, the layout of a blaze::DynamicTensor(3,3,X)
// (InvRotateDivide/Scalar.hpp:36-57 uses the same q[3k+j] = Q(k,j,.) order).
//
// Operation order follows the reference's scalar backend so that the
// restatement is correctly rounded in the same places (build with
// -ffp-contract=off so that no FMA contraction is introduced).
//
// Pinned by (tests/test_oracle_*.py):
//   * the PyElastica golden vectors of FullRodTest (0/1/2/3/5 steps, 1 and 2
//     rods, damped and undamped, one force evaluation from the straight-rod
//     initializer) at the reference's own tolerances;
//   * closed forms: two-point difference and quadrature against their
//     zero-padded / trapezoid definitions, InvRotateDivide on
//     rotate_via_axis_angle directors (kappa = (pi/5)(1,0,1)), the decimal
//     Rodrigues matrices of Test_SO3PrimitiveAddAssign, and an independent
//     matrix log / exp (scipy Rotation) on random directors;
//   * physics: telescoping sum of undamped internal forces, zero loads on a
//     straight unstressed rod, axial stretching tip displacement
//     F L / (E A - F), first-order convergence of the Timoshenko cantilever
//     deflection, exact free fall under gravity, a resting contact carrying
//     the weight, sliding friction decelerating at mu g (damped closed form),
//     and contact on a turned plane equal to the turned level case (the plane
//     normal in the gap, approach speed, slip and both forces);
//   * a rod stepped inside a multi-rod block equals the same rod alone, bitwise.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#if defined(ORACLE_DEVICE_MATH)
#include "../elasticapp_amd/csrc/erod_math.hpp"
#endif

namespace oracle {

  // The transcendental functions of the FP64 step (acos and sin in
  // InvRotateDivide, sin and cos in SO3AddAssign, sin in the sin ramp).  The
  // default build calls the C library's, as the reference does (std::acos,
  // std::sin, std::cos).  Two other builds of this file exist for the tests
  // (oracle/Makefile):
  //   ORACLE_DEVICE_MATH: the device library's own functions
  //     (elasticapp_amd/csrc/erod_math.hpp, themselves pinned to the C
  //     library within 1 ulp), so that an FP64 device run can be compared
  //     with the restatement bit for bit: the only operations whose results
  //     differ between a C library and the device are these functions;
  //   ORACLE_PERTURBED_MATH: the C library's results, a fixed quarter of
  //     them (chosen by a hash of the argument) moved one ulp up or down: a
  //     stand-in for another C library, to measure how far a trajectory
  //     carries last-bit differences of these functions.
  namespace tmath {
#if defined(ORACLE_DEVICE_MATH)
    inline double acos(double x) { return erod::math::acos(x); }
    inline double sin(double x) { return erod::math::sin(x); }
    inline double cos(double x) { return erod::math::cos(x); }
#elif defined(ORACLE_PERTURBED_MATH)
    inline double nudge(double x, double y) {
      uint64_t h;
      std::memcpy(&h, &x, sizeof(h));
      h ^= h >> 33;
      h *= 0xff51afd7ed558ccdULL;
      h ^= h >> 33;
      h *= 0xc4ceb9fe1a85ec53ULL;
      h ^= h >> 33;
      switch (h & 7) {
        case 0: return std::nextafter(y, -INFINITY);
        case 1: return std::nextafter(y, INFINITY);
        default: return y;
      }
    }
    inline double acos(double x) { return nudge(x, std::acos(x)); }
    inline double sin(double x) { return nudge(x, std::sin(x)); }
    inline double cos(double x) { return nudge(-x, std::cos(x)); }
#else
    inline double acos(double x) { return std::acos(x); }
    inline double sin(double x) { return std::sin(x); }
    inline double cos(double x) { return std::cos(x); }
#endif
  }  // namespace tmath

  enum class Stagger { Node, Element, Voronoi };

  struct FieldInfo {
    const char* name;
    Stagger stagger;
    int ncomp;
    double ghost;  // value written into ghost slots by fill_ghosts
  };

  // Ghost values: ScalarTag {} = 0 (common/Traits/DataType/ScalarTag.hpp:85),
  // vectors 0 (BlazeBackend/MatrixTag.hpp:102), lengths/masses/dimensions 1
  // (CosseratRodSpanwiseGeometry.hpp:97,153,179; CrossSectionInterface.hpp:95,121;
  // WithRodKinematics.hpp:101), director {{.8,.6,0},{0,0,1},{.6,-.8,0}}
  // (RodSpanwiseGeometry.hpp:96-101).  ghost = -1 marks "special" (director).
  enum F : int {
    // nodes
    Position, Velocity, Acceleration, Mass, InvMass, InternalLoads, ExternalLoads,
    ForceDampingRate,
    // elements
    Director, AngularVelocity, AngularAcceleration, ReferenceElementLength,
    ElementLength, ElementDilatation, Tangent, ElementDimension, ElementVolume,
    ShearStretchRigidityMatrix, MassSecondMomentOfInertia,
    InvMassSecondMomentOfInertia, ShearStretchStrain, ReferenceShearStretchStrain,
    InternalStress, InternalTorques, ExternalTorques, TorqueDampingRate,
    DummyElementVector, DummyElementVector2,
    // voronoi
    ReferenceVoronoiLength, VoronoiLength, VoronoiDilatation, Curvature,
    ReferenceCurvature, BendingTwistRigidityMatrix, InternalCouple,
    DummyVoronoiVector,
    NFields
  };

  inline const FieldInfo* field_table() {
    static const FieldInfo t[NFields] = {
        {"Position", Stagger::Node, 3, 0.0},
        {"Velocity", Stagger::Node, 3, 0.0},
        {"Acceleration", Stagger::Node, 3, 0.0},
        {"Mass", Stagger::Node, 1, 1.0},
        {"InvMass", Stagger::Node, 1, 0.0},
        {"InternalLoads", Stagger::Node, 3, 0.0},
        {"ExternalLoads", Stagger::Node, 3, 0.0},
        {"ForceDampingRate", Stagger::Node, 1, 0.0},
        {"Director", Stagger::Element, 9, -1.0},
        {"AngularVelocity", Stagger::Element, 3, 0.0},
        {"AngularAcceleration", Stagger::Element, 3, 0.0},
        {"ReferenceElementLength", Stagger::Element, 1, 1.0},
        {"ElementLength", Stagger::Element, 1, 1.0},
        {"ElementDilatation", Stagger::Element, 1, 0.0},
        {"Tangent", Stagger::Element, 3, 0.0},
        {"ElementDimension", Stagger::Element, 1, 1.0},
        {"ElementVolume", Stagger::Element, 1, 1.0},
        {"ShearStretchRigidityMatrix", Stagger::Element, 3, 0.0},
        {"MassSecondMomentOfInertia", Stagger::Element, 3, 0.0},
        {"InvMassSecondMomentOfInertia", Stagger::Element, 3, 0.0},
        {"ShearStretchStrain", Stagger::Element, 3, 0.0},
        {"ReferenceShearStretchStrain", Stagger::Element, 3, 0.0},
        {"InternalStress", Stagger::Element, 3, 0.0},
        {"InternalTorques", Stagger::Element, 3, 0.0},
        {"ExternalTorques", Stagger::Element, 3, 0.0},
        {"TorqueDampingRate", Stagger::Element, 1, 0.0},
        {"_DummyElementVector", Stagger::Element, 3, 0.0},
        {"_DummyElementVector2", Stagger::Element, 3, 0.0},
        {"ReferenceVoronoiLength", Stagger::Voronoi, 1, 1.0},
        {"VoronoiLength", Stagger::Voronoi, 1, 0.0},
        {"VoronoiDilatation", Stagger::Voronoi, 1, 0.0},
        {"Curvature", Stagger::Voronoi, 3, 0.0},
        {"ReferenceCurvature", Stagger::Voronoi, 3, 0.0},
        {"BendingTwistRigidityMatrix", Stagger::Voronoi, 3, 0.0},
        {"InternalCouple", Stagger::Voronoi, 3, 0.0},
        {"_DummyVoronoiVector", Stagger::Voronoi, 3, 0.0},
    };
    return t;
  }

  inline int field_id(const char* name) {
    for (int i = 0; i < NFields; ++i)
      if (std::strcmp(field_table()[i].name, name) == 0) return i;
    return -1;
  }

  // One ghost-padded block of rods (elastica/Systems/CosseratRods/Block.hpp).
  template <typename T>
  struct Block {
    std::vector<std::size_t> n_elems;      // elements per rod
    std::vector<std::size_t> starts;       // s_i (shared by all staggers)
    std::vector<std::size_t> ghost_nodes;  // ghost node index after each rod but the last
    std::size_t E = 0;                     // element slots incl. ghosts
    bool damped = true;                    // CosseratRod vs CosseratRodWithoutDamping
    std::vector<T> data[NFields];

    std::size_t slots(Stagger s) const {
      return s == Stagger::Node ? E + 1 : (s == Stagger::Element ? E : E - 1);
    }
    std::size_t slots(int f) const { return slots(field_table()[f].stagger); }
    T* get(int f) { return data[f].data(); }
    const T* get(int f) const { return data[f].data(); }

    // prepare_for_resize + resize_internal + fill_ghosts (Block.hpp:401-443,568-599)
    explicit Block(std::vector<std::size_t> const& rods, bool damped_ = true)
        : n_elems(rods), damped(damped_) {
      if (rods.empty()) throw std::invalid_argument("a block needs at least one rod");
      for (std::size_t n : rods)
        if (n < 2) throw std::invalid_argument("Cannot have rods with < 2 elements!");
      // ghost node index after rod i = total_n_nodes() before rod i+1 is added
      std::size_t s = 0;
      for (std::size_t i = 0; i < rods.size(); ++i) {
        starts.push_back(s);
        if (i + 1 < rods.size()) ghost_nodes.push_back(s + rods[i] + 1);
        s += rods[i] + 2;
      }
      E = s - 2;
      for (int f = 0; f < NFields; ++f)
        data[f].assign(std::size_t(field_table()[f].ncomp) * slots(f), T(0));
      fill_all_ghosts();
    }

    // fill_ghosts_for (Block.hpp:658-690): for ghost node index g, with deficit
    // d = n_ghosts(stagger) (node 1, element 2, voronoi 3), write g, g-1, ..., g-d+1.
    void fill_ghosts_for(int f) {
      auto const& info = field_table()[f];
      const std::size_t deficit = info.stagger == Stagger::Node ? 1 : (info.stagger == Stagger::Element ? 2 : 3);
      const std::size_t n = slots(f);
      T* p = get(f);
      for (std::size_t g : ghost_nodes) {
        for (std::size_t idx = g; idx > g - deficit; --idx) {
          if (info.ghost < 0.0) {
            static const double gd[9] = {0.8, 0.6, 0.0, 0.0, 0.0, 1.0, 0.6, -0.8, 0.0};
            for (int c = 0; c < 9; ++c) p[c * n + idx] = T(gd[c]);
          } else {
            for (int c = 0; c < info.ncomp; ++c) p[c * n + idx] = T(info.ghost);
          }
        }
      }
    }
    void fill_all_ghosts() {
      for (int f = 0; f < NFields; ++f) fill_ghosts_for(f);
    }
  };

  // ---------------------------------------------------------------------------
  // Primitive operators (OpsTraits.hpp:88-346, BlazeDetail/*.hpp)
  // ---------------------------------------------------------------------------

  // batch_norm (BlazeLinearAlgebra.hpp:236-242): sqrt of the column sum of squares
  template <typename T>
  inline T norm3(T a, T b, T c) { return std::sqrt(a * a + b * b + c * c); }

  // matvec: out_p = Q(p,0) v0 + Q(p,1) v1 + Q(p,2) v2
  // (LinearAlgebra/KernelGenerators/MatVec/Scalar.hpp:35-48, MatVec/Operation.hpp)
  template <typename T>
  inline void matvec(const T* Q, std::size_t n, std::size_t k, const T* v, std::size_t vn, std::size_t vk, T out[3]) {
    const T v0 = v[0 * vn + vk], v1 = v[1 * vn + vk], v2 = v[2 * vn + vk];
    for (int p = 0; p < 3; ++p)
      out[p] = Q[(3 * p + 0) * n + k] * v0 + Q[(3 * p + 1) * n + k] * v1 + Q[(3 * p + 2) * n + k] * v2;
  }

  // mattranspvec: out_j = sum_p Q(p,j) v_p  (OpsTraits.hpp:162-188, rowslice sums)
  template <typename T>
  inline void mattranspvec(const T* Q, std::size_t n, std::size_t k, const T v[3], T out[3]) {
    for (int j = 0; j < 3; ++j)
      out[j] = Q[(0 + j) * n + k] * v[0] + Q[(3 + j) * n + k] * v[1] + Q[(6 + j) * n + k] * v[2];
  }

  // cross product in the reference component order
  // (LinearAlgebra/CrossProduct/Operation.hpp: m2 v3 - m3 v2, m3 v1 - m1 v3, m1 v2 - m2 v1)
  template <typename T>
  inline void cross(const T a[3], const T b[3], T out[3]) {
    out[0] = a[1] * b[2] - a[2] * b[1];
    out[1] = a[2] * b[0] - a[0] * b[2];
    out[2] = a[0] * b[1] - a[1] * b[0];
  }

  // two_point_difference_kernel (BlazeCalculus.hpp:29-45): X -> X+1 columns
  template <typename T>
  inline void two_point_difference(const T* in, std::size_t nin, T* out, std::size_t nout) {
    for (int c = 0; c < 3; ++c) {
      const T* a = in + c * nin;
      T* o = out + c * nout;
      o[0] = a[0];
      o[nin] = -a[nin - 1];
      for (std::size_t i = 1; i < nin; ++i) o[i] = a[i] - a[i - 1];
    }
  }

  // quadrature_kernel (BlazeCalculus.hpp:67-86)
  template <typename T>
  inline void quadrature(const T* in, std::size_t nin, T* out, std::size_t nout) {
    for (int c = 0; c < 3; ++c) {
      const T* a = in + c * nin;
      T* o = out + c * nout;
      o[0] = T(0.5) * a[0];
      o[nin] = T(0.5) * a[nin - 1];
      for (std::size_t i = 1; i < nin; ++i) o[i] = T(0.5) * (a[i] + a[i - 1]);
    }
  }

  // InvRotateDivideOp<scalar> (InvRotateDivide/Scalar.hpp:26-110)
  template <typename T>
  inline void inv_rotate_divide(const T* Q, std::size_t ne, const T* D0, T* kappa, std::size_t nv) {
    for (std::size_t i = 0; i < nv; ++i) {
      T qc[9], qn[9];
      for (int m = 0; m < 9; ++m) {
        qc[m] = Q[m * ne + i];
        qn[m] = Q[m * ne + i + 1];
      }
      const T tr = qc[0] * qn[0] + qc[1] * qn[1] + qc[2] * qn[2] + qc[3] * qn[3] + qc[4] * qn[4] +
                   qc[5] * qn[5] + qc[6] * qn[6] + qc[7] * qn[7] + qc[8] * qn[8];
      constexpr T rest(T(0.5) + T(1e-10));
      const T theta = tmath::acos(T(0.5) * tr - rest);
      const T st = T(0.5) * theta / tmath::sin(theta + T(1e-14)) / D0[i];
      kappa[0 * nv + i] = st * (-qc[3] * qn[6] - qc[4] * qn[7] - qc[5] * qn[8] + qc[6] * qn[3] +
                                qc[7] * qn[4] + qc[8] * qn[5]);
      kappa[1 * nv + i] = st * (qc[0] * qn[6] + qc[1] * qn[7] + qc[2] * qn[8] - qc[6] * qn[0] -
                                qc[7] * qn[1] - qc[8] * qn[2]);
      kappa[2 * nv + i] = st * (-qc[0] * qn[3] - qc[1] * qn[4] - qc[2] * qn[5] + qc[3] * qn[0] +
                                qc[4] * qn[1] + qc[5] * qn[2]);
    }
  }

  // SO3AddAssign<scalar> (SO3PrimitiveAddAssign/Scalar.hpp:25-96):
  // Q <- exp(-[u]x) Q with u the rotation vector of this slot.
  template <typename T>
  inline void so3_add_assign_one(T* Q, std::size_t ne, std::size_t i, const T uin[3]) {
    T q[9], u[3] = {uin[0], uin[1], uin[2]};
    for (int m = 0; m < 9; ++m) q[m] = Q[m * ne + i];
    const T theta = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const T c_alpha = tmath::cos(theta);
    const T alpha = T(1.0) - c_alpha;
    const T beta = tmath::sin(theta);
    for (int d = 0; d < 3; ++d) u[d] /= (theta + T(1e-14));
    for (int j = 0; j < 3; ++j) {
      const T com = q[j] * u[0] + q[3 + j] * u[1] + q[6 + j] * u[2];
      Q[(0 + j) * ne + i] = c_alpha * q[j] + alpha * u[0] * com + beta * (q[3 + j] * u[2] - q[6 + j] * u[1]);
      Q[(3 + j) * ne + i] = c_alpha * q[3 + j] + alpha * u[1] * com + beta * (q[6 + j] * u[0] - q[j] * u[2]);
      Q[(6 + j) * ne + i] = c_alpha * q[6 + j] + alpha * u[2] * com + beta * (q[j] * u[1] - q[3 + j] * u[0]);
    }
  }

  // ---------------------------------------------------------------------------
  // Component kernels
  // ---------------------------------------------------------------------------

  // update_spanwise_variables (CosseratRodSpanwiseGeometry.hpp:578-622)
  template <typename T>
  void update_spanwise_variables(Block<T>& b) {
    const std::size_t E = b.E, N = E + 1, V = E - 1;
    const T* x = b.get(Position);
    T* l = b.get(ElementLength);
    for (std::size_t k = 0; k < E; ++k) {
      const T d0 = x[0 * N + k + 1] - x[0 * N + k];
      const T d1 = x[1 * N + k + 1] - x[1 * N + k];
      const T d2 = x[2 * N + k + 1] - x[2 * N + k];
      l[k] = norm3(d0, d1, d2);
    }
    b.fill_ghosts_for(ElementLength);
    T* t = b.get(Tangent);
    T* e = b.get(ElementDilatation);
    const T* l0 = b.get(ReferenceElementLength);
    for (std::size_t k = 0; k < E; ++k) {
      for (int c = 0; c < 3; ++c) t[c * E + k] = (x[c * N + k + 1] - x[c * N + k]) / l[k];
      e[k] = l[k] / l0[k];
    }
    T* D = b.get(VoronoiLength);
    T* eps = b.get(VoronoiDilatation);
    const T* D0 = b.get(ReferenceVoronoiLength);
    for (std::size_t v = 0; v < V; ++v) {
      D[v] = T(0.5) * (l[v] + l[v + 1]);
      eps[v] = D[v] / D0[v];
    }
  }

  // compute_cross_sectional_variables + update_dimension
  // (CosseratRodCrossSectionInterface.hpp:445-456, CircleCrossSectionOperations.hpp:109-117)
  template <typename T>
  void update_dimension(Block<T>& b) {
    const T inv_pi = static_cast<T>(M_1_PI);
    T* r = b.get(ElementDimension);
    const T* vol = b.get(ElementVolume);
    const T* l = b.get(ElementLength);
    for (std::size_t k = 0; k < b.E; ++k) r[k] = std::sqrt(inv_pi * vol[k] / l[k]);
  }

  // update_shear_stretch_strain (CosseratRodSpanwiseGeometry.hpp:629-644)
  template <typename T>
  void update_shear_stretch_strain(Block<T>& b) {
    const std::size_t E = b.E;
    const T* Q = b.get(Director);
    const T* t = b.get(Tangent);
    const T* e = b.get(ElementDilatation);
    T* s = b.get(ShearStretchStrain);
    for (std::size_t k = 0; k < E; ++k) {
      T o[3];
      matvec(Q, E, k, t, E, k, o);
      for (int c = 0; c < 3; ++c) s[c * E + k] = o[c] * e[k];
      s[2 * E + k] -= T(1.0);
    }
  }

  template <typename T>
  void compute_shear_strains(Block<T>& b) {
    update_spanwise_variables(b);
    update_dimension(b);
    update_shear_stretch_strain(b);
  }

  // compute_internal_modeled_loads_impl (WithDiagonalLinearHyperElasticModel.hpp:197-238)
  // and its damping wrapper (ExplicitDampingAdapterFacade.hpp:177-200)
  template <typename T>
  void compute_internal_forces(Block<T>& b) {
    const std::size_t E = b.E, N = E + 1;
    compute_shear_strains(b);
    const T* sig = b.get(ShearStretchStrain);
    const T* sig0 = b.get(ReferenceShearStretchStrain);
    const T* S = b.get(ShearStretchRigidityMatrix);
    T* n = b.get(InternalStress);
    for (std::size_t k = 0; k < E; ++k)
      for (int c = 0; c < 3; ++c) n[c * E + k] = S[c * E + k] * (sig[c * E + k] - sig0[c * E + k]);
    const T* Q = b.get(Director);
    const T* e = b.get(ElementDilatation);
    T* tmp = b.get(DummyElementVector2);
    T* cis = b.get(DummyElementVector);
    for (std::size_t k = 0; k < E; ++k) {
      const T nv[3] = {n[0 * E + k], n[1 * E + k], n[2 * E + k]};
      T o[3];
      mattranspvec(Q, E, k, nv, o);
      for (int c = 0; c < 3; ++c) tmp[c * E + k] = o[c];
    }
    for (std::size_t k = 0; k < E; ++k)
      for (int c = 0; c < 3; ++c) cis[c * E + k] = tmp[c * E + k] / e[k];
    b.fill_ghosts_for(DummyElementVector);
    T* f = b.get(InternalLoads);
    two_point_difference(cis, E, f, N);
    if (b.damped) {
      const T* m = b.get(Mass);
      const T* gf = b.get(ForceDampingRate);
      const T* v = b.get(Velocity);
      for (std::size_t i = 0; i < N; ++i)
        for (int c = 0; c < 3; ++c) f[c * N + i] -= (m[i] * gf[i]) * v[c * N + i];
    }
  }

  // compute_internal_modeled_torques_impl (WithDiagonalLinearHyperElasticModel.hpp:246-391)
  // and its damping wrapper (ExplicitDampingAdapterFacade.hpp:207-235)
  template <typename T>
  void compute_internal_torques(Block<T>& b) {
    const std::size_t E = b.E, N = E + 1, V = E - 1;
    const T* Q = b.get(Director);
    const T* w = b.get(AngularVelocity);
    const T* vel = b.get(Velocity);
    const T* t = b.get(Tangent);
    const T* e = b.get(ElementDilatation);
    const T* eps = b.get(VoronoiDilatation);
    const T* B = b.get(BendingTwistRigidityMatrix);
    const T* k0 = b.get(ReferenceCurvature);
    const T* D0 = b.get(ReferenceVoronoiLength);
    const T* l0 = b.get(ReferenceElementLength);
    const T* J = b.get(MassSecondMomentOfInertia);
    const T* n = b.get(InternalStress);
    T* M = b.get(InternalCouple);
    T* tau = b.get(InternalTorques);
    T* l = b.get(ElementLength);
    T* D = b.get(VoronoiLength);
    T* d1 = b.get(DummyElementVector);
    T* d2 = b.get(DummyElementVector2);
    T* dv = b.get(DummyVoronoiVector);
    T* kap = b.get(Curvature);

    inv_rotate_divide(Q, E, D0, kap, V);
    b.fill_ghosts_for(Curvature);
    for (std::size_t i = 0; i < V; ++i)
      for (int c = 0; c < 3; ++c) M[c * V + i] = B[c * V + i] * (kap[c * V + i] - k0[c * V + i]);
    // voronoi_length used as storage for 1/eps^3
    for (std::size_t i = 0; i < V; ++i) D[i] = T(1.0) / (eps[i] * eps[i] * eps[i]);
    for (std::size_t i = 0; i < V; ++i)
      for (int c = 0; c < 3; ++c) dv[c * V + i] = M[c * V + i] * D[i];
    two_point_difference(dv, V, tau, E);
    for (std::size_t i = 0; i < V; ++i) {
      const T a[3] = {kap[i], kap[V + i], kap[2 * V + i]};
      const T m3[3] = {M[i], M[V + i], M[2 * V + i]};
      T o[3];
      cross(a, m3, o);
      for (int c = 0; c < 3; ++c) dv[c * V + i] = o[c];
    }
    for (std::size_t i = 0; i < V; ++i) D[i] *= D0[i];
    // quadrature of (kappa x M) * (D0 / eps^3); the scaled input is a lazy
    // expression in the reference, so scale on the fly
    {
      std::vector<T> tmp(3 * V);
      for (std::size_t i = 0; i < V; ++i)
        for (int c = 0; c < 3; ++c) tmp[c * V + i] = dv[c * V + i] * D[i];
      quadrature(tmp.data(), V, d1, E);
    }
    for (std::size_t k = 0; k < E; ++k)
      for (int c = 0; c < 3; ++c) tau[c * E + k] += d1[c * E + k];
    for (std::size_t k = 0; k < E; ++k) {
      T qt[3], o[3];
      matvec(Q, E, k, t, E, k, qt);
      for (int c = 0; c < 3; ++c) d1[c * E + k] = qt[c];
      const T nv[3] = {n[k], n[E + k], n[2 * E + k]};
      cross(qt, nv, o);
      for (int c = 0; c < 3; ++c) d2[c * E + k] = o[c];
    }
    for (std::size_t k = 0; k < E; ++k)
      for (int c = 0; c < 3; ++c) tau[c * E + k] += d2[c * E + k] * l0[k];
    // J omega / e (J diagonal: the reference's full 3x3 matvec adds exact zeros)
    for (std::size_t k = 0; k < E; ++k)
      for (int c = 0; c < 3; ++c) d2[c * E + k] = J[c * E + k] * w[c * E + k];
    for (std::size_t k = 0; k < E; ++k)
      for (int c = 0; c < 3; ++c) d1[c * E + k] = d2[c * E + k] / e[k];
    for (std::size_t k = 0; k < E; ++k) {
      const T a[3] = {d1[k], d1[E + k], d1[2 * E + k]};
      const T om[3] = {w[k], w[E + k], w[2 * E + k]};
      T o[3];
      cross(a, om, o);
      for (int c = 0; c < 3; ++c) d2[c * E + k] = o[c];
    }
    for (std::size_t k = 0; k < E; ++k)
      for (int c = 0; c < 3; ++c) tau[c * E + k] += d2[c * E + k];
    for (std::size_t k = 0; k < E; ++k)
      for (int c = 0; c < 3; ++c) d2[c * E + k] = vel[c * N + k + 1] - vel[c * N + k];
    for (std::size_t k = 0; k < E; ++k)
      l[k] = (t[k] * d2[k] + t[E + k] * d2[E + k] + t[2 * E + k] * d2[2 * E + k]) / (e[k] * l0[k]);
    for (std::size_t k = 0; k < E; ++k)
      for (int c = 0; c < 3; ++c) tau[c * E + k] += d1[c * E + k] * l[k];
    for (std::size_t i = 0; i < V; ++i) D[i] = eps[i] * D0[i];
    for (std::size_t k = 0; k < E; ++k) l[k] = e[k] * l0[k];
    if (b.damped) {
      const T* m = b.get(Mass);
      const T* gt = b.get(TorqueDampingRate);
      for (std::size_t k = 0; k < E; ++k) {
        const T em = T(0.5) * (m[k] + m[k + 1]);
        for (int c = 0; c < 3; ++c) tau[c * E + k] -= (em * gt[k]) * w[c * E + k];
      }
    }
  }

  // update_acceleration (WithRodKinematics.hpp:683-714)
  template <typename T>
  void update_acceleration(Block<T>& b) {
    const std::size_t E = b.E, N = E + 1;
    const T* im = b.get(InvMass);
    const T* e = b.get(ElementDilatation);
    const T* iJ = b.get(InvMassSecondMomentOfInertia);
    const T* fi = b.get(InternalLoads);
    const T* fe = b.get(ExternalLoads);
    const T* ti = b.get(InternalTorques);
    const T* te = b.get(ExternalTorques);
    T* a = b.get(Acceleration);
    T* al = b.get(AngularAcceleration);
    for (std::size_t i = 0; i < N; ++i)
      for (int c = 0; c < 3; ++c) a[c * N + i] = (fi[c * N + i] + fe[c * N + i]) * im[i];
    for (std::size_t k = 0; k < E; ++k)
      for (int c = 0; c < 3; ++c) al[c * E + k] = (iJ[c * E + k] * (ti[c * E + k] + te[c * E + k])) * e[k];
  }

  // PositionVerlet::KinematicSubStep (PositionVerlet.hpp:60-69): x += (dt/2) v,
  // Q <- exp(-(dt/2)[w]x) Q (SE3.hpp:302-306, SO3.hpp:342-348)
  template <typename T>
  void kinematic_update(Block<T>& b, T dt) {
    const T c = T(0.5) * dt;
    const std::size_t E = b.E, N = E + 1;
    T* x = b.get(Position);
    const T* v = b.get(Velocity);
    for (std::size_t i = 0; i < 3 * N; ++i) x[i] += c * v[i];
    T* Q = b.get(Director);
    const T* w = b.get(AngularVelocity);
    for (std::size_t k = 0; k < E; ++k) {
      const T u[3] = {c * w[k], c * w[E + k], c * w[2 * E + k]};
      so3_add_assign_one(Q, E, k, u);
    }
  }

  // PositionVerlet::DynamicSubStep (PositionVerlet.hpp:71-77) with the rates
  // from update_acceleration (SymplecticStepperAdapter.hpp:357-366)
  template <typename T>
  void dynamic_update(Block<T>& b, T dt) {
    update_acceleration(b);
    const std::size_t E = b.E, N = E + 1;
    T* v = b.get(Velocity);
    const T* a = b.get(Acceleration);
    for (std::size_t i = 0; i < 3 * N; ++i) v[i] += dt * a[i];
    T* w = b.get(AngularVelocity);
    const T* al = b.get(AngularAcceleration);
    for (std::size_t i = 0; i < 3 * E; ++i) w[i] += dt * al[i];
  }

  template <typename T>
  void compute_internal_dynamics(Block<T>& b) {
    std::fill(b.data[ExternalLoads].begin(), b.data[ExternalLoads].end(), T(0));
    std::fill(b.data[ExternalTorques].begin(), b.data[ExternalTorques].end(), T(0));
    compute_internal_forces(b);
    compute_internal_torques(b);
  }

  // ---------------------------------------------------------------------------
  // Module hooks restated as descriptors (Constraints / Forcing / Environments)
  // ---------------------------------------------------------------------------
  // value + rate constraint of a fixed clamp at index i of a rod: node i and
  // element min(i, n - 1) (a clamp at the last node, i = n, holds the last
  // element with it, as a clamp at i = 0 holds the first).  The constraint's
  // per-index callbacks
  // constrain_primitives / constrain_rates
  // (elastica/Simulator/Constraints/Constraint.hpp:290-313,
  // elastica/Simulator/Interface/Constraints/ExecutionPolicy.hpp:591-593).
  struct Clamp {
    std::size_t rod, index;
    double x[3], Q[9];
  };
  struct EndForce {  // ExternalLoads(:, last node) = F * ramp(t)
    std::size_t rod;
    double F[3];
    int ramp;  // 0 none, 1 linear, 2 sin  (elastica/Utilities/Rampup.hpp:95-150)
    double start, interval;
  };
  struct PlaneContact {  // per-node plane contact with regularised Coulomb friction
    double k, c, mu, ct, z0, offset;
    double n[3] = {0.0, 0.0, 1.0};  // unit normal of the plane n . x = z0
  };

  inline double ramp_value(int kind, double t, double start, double interval) {
    if (kind == 0) return 1.0;
    if (t < start) return 0.0;
    if (t > start + interval) return 1.0;
    if (kind == 1) return (t - start) / interval;
    const double s = 2.0 * ((t - start) / interval) - 1.0;
    return 0.5 * (s + (1.0 + (tmath::sin(M_PI * s) / M_PI)));
  }

  template <typename T>
  struct System {
    std::vector<Block<T>> blocks;
    // rod -> (block, index in block)
    std::vector<std::pair<std::size_t, std::size_t>> where;
    std::vector<Clamp> clamps;
    std::vector<EndForce> forces;
    bool gravity = false;
    double g[3] = {0, 0, 0};
    bool contact = false;
    PlaneContact plane{};
    double time = 0.0;
  };

  template <typename T>
  void apply_post_kinematic(System<T>& s) {
    for (auto const& c : s.clamps) {
      auto [bi, ri] = s.where[c.rod];
      auto& b = s.blocks[bi];
      const std::size_t N = b.E + 1, E = b.E;
      const std::size_t st = b.starts[ri];
      const std::size_t e = std::min(c.index, b.n_elems[ri] - 1);
      for (int d = 0; d < 3; ++d) b.get(Position)[d * N + st + c.index] = T(c.x[d]);
      for (int m = 0; m < 9; ++m) b.get(Director)[m * E + st + e] = T(c.Q[m]);
    }
  }

  template <typename T>
  void apply_post_dynamic(System<T>& s) {
    for (auto const& c : s.clamps) {
      auto [bi, ri] = s.where[c.rod];
      auto& b = s.blocks[bi];
      const std::size_t N = b.E + 1, E = b.E;
      const std::size_t st = b.starts[ri];
      const std::size_t e = std::min(c.index, b.n_elems[ri] - 1);
      for (int d = 0; d < 3; ++d) b.get(Velocity)[d * N + st + c.index] = T(0);
      for (int d = 0; d < 3; ++d) b.get(AngularVelocity)[d * E + st + e] = T(0);
    }
  }

  // GravityField resolve (GravityField.hpp:239-246): f_ext += g (x) m;
  // per-node plane contact after LinearSpringDashpot.cpp:49-71.
  template <typename T>
  void apply_environment(Block<T>& b, System<T> const& s) {
    const std::size_t N = b.E + 1;
    T* fe = b.get(ExternalLoads);
    const T* m = b.get(Mass);
    if (s.gravity)
      for (int c = 0; c < 3; ++c)
        for (std::size_t i = 0; i < N; ++i) fe[c * N + i] += T(s.g[c]) * m[i];
    if (s.contact) {
      const T* x = b.get(Position);
      const T* v = b.get(Velocity);
      for (std::size_t r = 0; r < b.n_elems.size(); ++r) {
        const std::size_t st = b.starts[r];
        for (std::size_t i = st; i <= st + b.n_elems[r]; ++i) {
          // the plane's normal and contact law as in CapsulePlane.hpp:73-80
          // (distance along the plane normal) and LinearSpringDashpot.cpp:49-71
          // (normal and tangential relative velocity, friction capped at mu fN)
          const T nx = T(s.plane.n[0]), ny = T(s.plane.n[1]), nz = T(s.plane.n[2]);
          const T gap = ((nx * x[i] + ny * x[N + i]) + nz * x[2 * N + i]) - T(s.plane.z0) - T(s.plane.offset);
          if (!(gap < T(0))) continue;
          const T delta = -gap;
          const T vnv = (nx * v[i] + ny * v[N + i]) + nz * v[2 * N + i];
          const T vn = -vnv;  // approach speed along the plane normal
          const T fn = std::max(T(0), T(s.plane.k) * delta + T(s.plane.c) * vn);
          const T vt[3] = {v[i] - vnv * nx, v[N + i] - vnv * ny, v[2 * N + i] - vnv * nz};
          const T vtn = std::sqrt((vt[0] * vt[0] + vt[1] * vt[1]) + vt[2] * vt[2]);
          const T nv[3] = {nx, ny, nz};
          for (int c = 0; c < 3; ++c) fe[c * N + i] += fn * nv[c];
          if (vtn > T(0)) {
            const T ft = std::min(T(s.plane.ct) * vtn, T(s.plane.mu) * fn);
            const T sc = ft / vtn;  // f_T along the unit slip
            for (int c = 0; c < 3; ++c) fe[c * N + i] -= sc * vt[c];
          }
        }
      }
    }
  }

  // Module order Forcing, then Environments (make_simulator<Forcing, ...,
  // Environments>; Simulator.hpp:284-285 appends each module's steps in order).
  template <typename T>
  void apply_pre_dynamic(System<T>& s, double t) {
    for (auto const& f : s.forces) {
      auto [bi, ri] = s.where[f.rod];
      auto& b = s.blocks[bi];
      const std::size_t N = b.E + 1;
      const std::size_t node = b.starts[ri] + b.n_elems[ri];
      const double r = ramp_value(f.ramp, t, f.start, f.interval);
      for (int d = 0; d < 3; ++d) b.get(ExternalLoads)[d * N + node] = T(f.F[d] * r);
    }
    for (auto& b : s.blocks) apply_environment(b, s);
  }

  // One position-Verlet step over all blocks
  // (SymplecticSteppers.hpp:152-216; SymplecticStepper.hpp:255-273)
  template <typename T, typename ForEach>
  void step(System<T>& s, double dt, ForEach&& for_each_block) {
    const T tdt = T(dt);
    for_each_block([&](Block<T>& b) { kinematic_update(b, tdt); });
    s.time += 0.5 * dt;
    apply_post_kinematic(s);
    for_each_block([&](Block<T>& b) { compute_internal_dynamics(b); });
    apply_pre_dynamic(s, s.time);
    for_each_block([&](Block<T>& b) { dynamic_update(b, tdt); });
    apply_post_dynamic(s);
    s.time += 0.5 * dt;
    for_each_block([&](Block<T>& b) { kinematic_update(b, tdt); });
  }

  // ---------------------------------------------------------------------------
  // Straight-rod initializer (A14): Initializers/detail/Geometry.hpp:110-214,
  // Orthogonal.cpp:33-116, CosseratRodSpanwiseGeometry.hpp:351-374,525-545,
  // CosseratRodCrossSectionInterface.hpp:266-307, CircleCrossSectionOperations.hpp:83-132,
  // LinearHyperElasticityFacade.hpp:260-354, Initializers/detail/Elasticity.hpp:54-76,
  // WithRodKinematics.hpp:330-436.
  // ---------------------------------------------------------------------------
  struct StraightRodParams {
    double radius, length, origin[3], direction[3], normal[3];
    double density, youngs, shear, force_damping, torque_damping;
  };

  inline void normalize3(const double a[3], double o[3]) {
    const double len = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    const double inv = 1.0 / len;
    for (int i = 0; i < 3; ++i) o[i] = a[i] * inv;
  }

  template <typename T>
  void initialize_straight_rod(Block<T>& b, std::size_t rod, StraightRodParams const& p) {
    const std::size_t n = b.n_elems[rod], st = b.starts[rod];
    const std::size_t E = b.E, N = E + 1, V = E - 1;
    double dir[3], nor[3];
    normalize3(p.direction, dir);
    normalize3(p.normal, nor);
    // make_orthogonal_bases_from_normal_and_tangent(normal, direction)
    double v1[3], v2[3], v3[3], tmp[3];
    normalize3(nor, v1);
    cross(dir, nor, v2);  // proposed_tangent % proposed_normal
    cross(v1, v2, v3);
    cross(v3, v1, tmp);
    double d2[3], d3[3];
    normalize3(tmp, d2);
    normalize3(v3, d3);
    const double rows[9] = {v1[0], v1[1], v1[2], d2[0], d2[1], d2[2], d3[0], d3[1], d3[2]};
    // positions: linspace(origin, origin + L dir, n + 1)
    double end[3], step3[3];
    for (int c = 0; c < 3; ++c) end[c] = p.origin[c] + p.length * dir[c];
    for (int c = 0; c < 3; ++c) step3[c] = (end[c] - p.origin[c]) / double(n);
    T* x = b.get(Position);
    for (std::size_t i = 0; i <= n; ++i)
      for (int c = 0; c < 3; ++c) x[c * N + st + i] = T(p.origin[c] + double(i) * step3[c]);
    T* Q = b.get(Director);
    for (std::size_t k = 0; k < n; ++k)
      for (int m = 0; m < 9; ++m) Q[m * E + st + k] = T(rows[m]);
    // defaults: reference curvature / strain 0, dilatation 1
    T* e = b.get(ElementDilatation);
    for (std::size_t k = 0; k < n; ++k) e[st + k] = T(1.0);
    for (int c = 0; c < 3; ++c) {
      for (std::size_t k = 0; k < n; ++k) b.get(ReferenceShearStretchStrain)[c * E + st + k] = T(0);
      for (std::size_t k = 0; k + 1 < n; ++k) b.get(ReferenceCurvature)[c * V + st + k] = T(0);
    }
    // compute_reference_geometry: l0 = |dx| / e; D0 = avg(l0)
    T* l0 = b.get(ReferenceElementLength);
    for (std::size_t k = 0; k < n; ++k) {
      const std::size_t j = st + k;
      l0[j] = norm3(x[j + 1] - x[j], x[N + j + 1] - x[N + j], x[2 * N + j + 1] - x[2 * N + j]) / e[j];
    }
    T* D0 = b.get(ReferenceVoronoiLength);
    for (std::size_t k = 0; k + 1 < n; ++k) D0[st + k] = T(0.5) * (l0[st + k] + l0[st + k + 1]);
    // radius, then volume = A(r) l0, then radius from volume and current length
    T* r = b.get(ElementDimension);
    T* vol = b.get(ElementVolume);
    for (std::size_t k = 0; k < n; ++k) r[st + k] = T(p.radius);
    update_spanwise_variables(b);  // sets l, t, e, D, eps (valid everywhere)
    for (std::size_t k = 0; k < n; ++k) {
      const T rr = r[st + k];
      vol[st + k] = (static_cast<T>(M_PI) * (rr * rr)) * l0[st + k];
    }
    {
      const T inv_pi = static_cast<T>(M_1_PI);
      const T* l = b.get(ElementLength);
      for (std::size_t k = 0; k < n; ++k) r[st + k] = std::sqrt(inv_pi * vol[st + k] / l[st + k]);
    }
    update_shear_stretch_strain(b);
    inv_rotate_divide(b.get(Director), E, b.get(ReferenceVoronoiLength), b.get(Curvature), V);
    auto area = [&](std::size_t j) { const T rr = r[j]; return static_cast<T>(M_PI) * (rr * rr); };
    auto moi = [&](std::size_t j, T out[3]) {
      const T A = area(j);
      const T I = static_cast<T>(0.25 * M_1_PI) * (A * A);
      out[0] = I;
      out[1] = I;
      out[2] = T(2.0) * I;
    };
    // bending: length-weighted voronoi average of (E,E,G) I
    const T EEG[3] = {T(p.youngs), T(p.youngs), T(p.shear)};
    T* B = b.get(BendingTwistRigidityMatrix);
    for (std::size_t k = 0; k + 1 < n; ++k) {
      const std::size_t j = st + k;
      T I0[3], I1[3];
      moi(j, I0);
      moi(j + 1, I1);
      for (int c = 0; c < 3; ++c)
        B[c * V + j] = ((l0[j] * EEG[c]) * I0[c] + (l0[j + 1] * EEG[c]) * I1[c]) / (l0[j] + l0[j + 1]);
    }
    // shear/stretch: (G, G, E) * (alpha A, alpha A, A), alpha = 13.5/14
    const T alpha = static_cast<T>(13.5) / static_cast<T>(14.);
    T* S = b.get(ShearStretchRigidityMatrix);
    for (std::size_t k = 0; k < n; ++k) {
      const std::size_t j = st + k;
      const T A = area(j);
      S[0 * E + j] = T(p.shear) * (A * alpha);
      S[1 * E + j] = T(p.shear) * (A * alpha);
      S[2 * E + j] = T(p.youngs) * A;
    }
    // damping rates
    for (std::size_t i = 0; i <= n; ++i) b.get(ForceDampingRate)[st + i] = T(p.force_damping);
    for (std::size_t k = 0; k < n; ++k) b.get(TorqueDampingRate)[st + k] = T(p.torque_damping);
    // kinematics: J = I rho l0 (diagonal), 1/J, lumped mass, 1/m
    T* J = b.get(MassSecondMomentOfInertia);
    T* iJ = b.get(InvMassSecondMomentOfInertia);
    for (std::size_t k = 0; k < n; ++k) {
      const std::size_t j = st + k;
      T I[3];
      moi(j, I);
      for (int c = 0; c < 3; ++c) {
        J[c * E + j] = (I[c] * T(p.density)) * l0[j];
        iJ[c * E + j] = T(1.0) / ((I[c] * T(p.density)) * l0[j]);
      }
    }
    T* m = b.get(Mass);
    for (std::size_t i = 0; i <= n; ++i) m[st + i] = T(0);
    for (std::size_t k = 0; k < n; ++k) {
      const T me = T(p.density) * vol[st + k];
      m[st + k] += T(0.5) * me;
      m[st + k + 1] += T(0.5) * me;
    }
    for (std::size_t i = 0; i <= n; ++i) b.get(InvMass)[st + i] = T(1.0) / m[st + i];
    for (int c = 0; c < 3; ++c) {
      for (std::size_t i = 0; i <= n; ++i) {
        b.get(Velocity)[c * N + st + i] = T(0);
        b.get(ExternalLoads)[c * N + st + i] = T(0);
      }
      for (std::size_t k = 0; k < n; ++k) {
        b.get(AngularVelocity)[c * E + st + k] = T(0);
        b.get(ExternalTorques)[c * E + st + k] = T(0);
      }
    }
  }

}  // namespace oracle
