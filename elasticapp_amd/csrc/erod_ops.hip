// The plain FP64 path: one kernel per operator of the reference's block
// timestep, each a grid-stride loop over every slot of the ghost-padded block
// (ghosts included), in the reference's evaluation order.  This file is built
// with -ffp-contract=off so that every product and sum is rounded on its own,
// as in the CPU restatement it is checked against.  acos, sin and cos are the
// library's own (erod_math.hpp), so that these kernels and a CPU build of the
// restatement with the same functions compute the same bits.
//
// Reference operators (paths under elastica/Systems/CosseratRods/):
//   spanwise geometry   Components/Geometry/detail/CosseratRodSpanwiseGeometry.hpp:578-664
//   radius              Components/Geometry/detail/CosseratRodCrossSectionInterface.hpp:445-456
//   internal loads      Components/Elasticity/WithDiagonalLinearHyperElasticModel.hpp:197-238
//   internal torques    Components/Elasticity/WithDiagonalLinearHyperElasticModel.hpp:246-391
//   damping             Components/Elasticity/detail/ExplicitDampingAdapterFacade.hpp:177-235
//   accelerations       Components/Kinematics/WithRodKinematics.hpp:683-714
//   operators           Traits/Operations/BlazeBackend/OpsTraits.hpp:88-346
//   director update     elastica/Systems/States/Expressions/backends/blaze/SO3PrimitiveAddAssign/Scalar.hpp:25-96
#include "erod_common.hpp"
#include "erod_math.hpp"

namespace erod {
  namespace {

#define GRID_STRIDE(i, n) \
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * blockDim.x)

    __global__ void k_axpy(double* __restrict__ x, const double* __restrict__ v, double c, uint64_t n) {
      GRID_STRIDE(i, n) x[i] += c * v[i];
    }

    __global__ void k_fill(double* __restrict__ p, double val, uint64_t n) {
      GRID_STRIDE(i, n) p[i] = val;
    }

    // fill_ghosts_for (Block.hpp:658-690): ghost node g, deficit d -> slots g, g-1, ..., g-d+1
    __global__ void k_fill_ghosts(double* __restrict__ p, const uint64_t* __restrict__ gn, uint64_t n_gn,
                                  int deficit, uint64_t slots, int ncomp, double ghost) {
      GRID_STRIDE(i, n_gn * (uint64_t)deficit) {
        const uint64_t g = gn[i / deficit];
        const uint64_t idx = g - (i % deficit);
        if (ghost < -1.0) {
          const double gd[9] = {0.8, 0.6, 0.0, 0.0, 0.0, 1.0, 0.6, -0.8, 0.0};
          for (int c = 0; c < 9; ++c) p[c * slots + idx] = gd[c];
        } else {
          for (int c = 0; c < ncomp; ++c) p[c * slots + idx] = ghost;
        }
      }
    }

    // SO3AddAssign<scalar> with u = c * w
    __global__ void k_so3(double* __restrict__ Q, const double* __restrict__ w, double c, uint64_t E) {
      GRID_STRIDE(i, E) {
        double q[9], u[3];
        for (int m = 0; m < 9; ++m) q[m] = Q[m * E + i];
        for (int d = 0; d < 3; ++d) u[d] = c * w[d * E + i];
        const double theta = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        const double c_alpha = math::cos(theta);
        const double alpha = 1.0 - c_alpha;
        const double beta = math::sin(theta);
        for (int d = 0; d < 3; ++d) u[d] /= (theta + 1e-14);
        for (int j = 0; j < 3; ++j) {
          const double com = q[j] * u[0] + q[3 + j] * u[1] + q[6 + j] * u[2];
          Q[(0 + j) * E + i] = c_alpha * q[j] + alpha * u[0] * com + beta * (q[3 + j] * u[2] - q[6 + j] * u[1]);
          Q[(3 + j) * E + i] = c_alpha * q[3 + j] + alpha * u[1] * com + beta * (q[6 + j] * u[0] - q[j] * u[2]);
          Q[(6 + j) * E + i] = c_alpha * q[6 + j] + alpha * u[2] * com + beta * (q[j] * u[1] - q[3 + j] * u[0]);
        }
      }
    }

    // element_length = batch_norm(difference_kernel(position))
    __global__ void k_norm_diff(double* __restrict__ l, const double* __restrict__ x, uint64_t E) {
      const uint64_t N = E + 1;
      GRID_STRIDE(k, E) {
        const double d0 = x[k + 1] - x[k];
        const double d1 = x[N + k + 1] - x[N + k];
        const double d2 = x[2 * N + k + 1] - x[2 * N + k];
        l[k] = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
      }
    }

    // batch_division_matvec(tangent, position_diff, element_length)
    __global__ void k_tangent(double* __restrict__ t, const double* __restrict__ x, const double* __restrict__ l, uint64_t E) {
      const uint64_t N = E + 1;
      GRID_STRIDE(k, E) {
        for (int c = 0; c < 3; ++c) t[c * E + k] = (x[c * N + k + 1] - x[c * N + k]) / l[k];
      }
    }

    __global__ void k_div(double* __restrict__ o, const double* __restrict__ a, const double* __restrict__ b, uint64_t n) {
      GRID_STRIDE(i, n) o[i] = a[i] / b[i];
    }

    __global__ void k_mul(double* __restrict__ o, const double* __restrict__ a, const double* __restrict__ b, uint64_t n) {
      GRID_STRIDE(i, n) o[i] = a[i] * b[i];
    }

    __global__ void k_mul_inplace(double* __restrict__ o, const double* __restrict__ b, uint64_t n) {
      GRID_STRIDE(i, n) o[i] *= b[i];
    }

    // average_kernel (OpsTraits.hpp:261-269): 0.5 * (a[i] + a[i+1]) over n - 1 outputs
    __global__ void k_avg(double* __restrict__ o, const double* __restrict__ a, uint64_t nout) {
      GRID_STRIDE(i, nout) o[i] = 0.5 * (a[i] + a[i + 1]);
    }

    // update_dimension: sqrt(inv_pi * volume / length)
    __global__ void k_radius(double* __restrict__ r, const double* __restrict__ vol, const double* __restrict__ l, uint64_t E) {
      const double inv_pi = M_1_PI;
      GRID_STRIDE(k, E) r[k] = sqrt(inv_pi * vol[k] / l[k]);
    }

    // MatVec (KernelGenerators/MatVec/Scalar.hpp:35-48): out_p = Q(p,0) v0 + Q(p,1) v1 + Q(p,2) v2
    __global__ void k_matvec(double* __restrict__ o, const double* __restrict__ Q, const double* __restrict__ v, uint64_t E) {
      GRID_STRIDE(k, E) {
        const double v0 = v[k], v1 = v[E + k], v2 = v[2 * E + k];
        for (int p = 0; p < 3; ++p)
          o[p * E + k] = Q[(3 * p + 0) * E + k] * v0 + Q[(3 * p + 1) * E + k] * v1 + Q[(3 * p + 2) * E + k] * v2;
      }
    }

    // mattranspvec (OpsTraits.hpp:162-188): out_j = sum_p Q(p,j) v_p
    __global__ void k_mattranspvec(double* __restrict__ o, const double* __restrict__ Q, const double* __restrict__ v, uint64_t E) {
      GRID_STRIDE(k, E) {
        const double v0 = v[k], v1 = v[E + k], v2 = v[2 * E + k];
        for (int j = 0; j < 3; ++j)
          o[j * E + k] = Q[(0 + j) * E + k] * v0 + Q[(3 + j) * E + k] * v1 + Q[(6 + j) * E + k] * v2;
      }
    }

    // a (3 lines) *= s (1 line)
    __global__ void k_scale3(double* __restrict__ a, const double* __restrict__ s, uint64_t n) {
      GRID_STRIDE(k, n) {
        for (int c = 0; c < 3; ++c) a[c * n + k] *= s[k];
      }
    }

    // batch_subtract_z_unit: row 2 -= 1
    __global__ void k_sub_z(double* __restrict__ a, uint64_t n) {
      GRID_STRIDE(k, n) a[2 * n + k] -= 1.0;
    }

    // diagonal matvec: o = S % (a - b) (OpsTraits.hpp:145-156)
    __global__ void k_diag_mul_diff(double* __restrict__ o, const double* __restrict__ S, const double* __restrict__ a,
                                    const double* __restrict__ b, uint64_t n3) {
      GRID_STRIDE(i, n3) o[i] = S[i] * (a[i] - b[i]);
    }

    // batch_division_matvec: o_c = a_c / s
    __global__ void k_div3(double* __restrict__ o, const double* __restrict__ a, const double* __restrict__ s, uint64_t n) {
      GRID_STRIDE(k, n) {
        for (int c = 0; c < 3; ++c) o[c * n + k] = a[c * n + k] / s[k];
      }
    }

    // two_point_difference_kernel (BlazeCalculus.hpp:29-45): nin -> nin + 1
    __global__ void k_two_point_diff(double* __restrict__ o, const double* __restrict__ a, uint64_t nin) {
      const uint64_t nout = nin + 1;
      GRID_STRIDE(i, nout) {
        for (int c = 0; c < 3; ++c) {
          const double* ac = a + c * nin;
          double r;
          if (i == 0) r = ac[0];
          else if (i == nin) r = -ac[nin - 1];
          else r = ac[i] - ac[i - 1];
          o[c * nout + i] = r;
        }
      }
    }

    // two_point_difference of (a_c * s) (the lazily scaled argument of the torque difference)
    __global__ void k_two_point_diff_scaled(double* __restrict__ o, const double* __restrict__ a,
                                            const double* __restrict__ s, uint64_t nin) {
      const uint64_t nout = nin + 1;
      GRID_STRIDE(i, nout) {
        for (int c = 0; c < 3; ++c) {
          const double* ac = a + c * nin;
          double r;
          if (i == 0) r = ac[0] * s[0];
          else if (i == nin) r = -(ac[nin - 1] * s[nin - 1]);
          else r = ac[i] * s[i] - ac[i - 1] * s[i - 1];
          o[c * nout + i] = r;
        }
      }
    }

    // quadrature_kernel (BlazeCalculus.hpp:67-86) of (a_c * s)
    __global__ void k_quad_scaled(double* __restrict__ o, const double* __restrict__ a, const double* __restrict__ s, uint64_t nin) {
      const uint64_t nout = nin + 1;
      GRID_STRIDE(i, nout) {
        for (int c = 0; c < 3; ++c) {
          const double* ac = a + c * nin;
          double r;
          if (i == 0) r = 0.5 * (ac[0] * s[0]);
          else if (i == nin) r = 0.5 * (ac[nin - 1] * s[nin - 1]);
          else r = 0.5 * (ac[i] * s[i] + ac[i - 1] * s[i - 1]);
          o[c * nout + i] = r;
        }
      }
    }

    // internal_force -= (mass * dissipation) % velocity
    __global__ void k_damp_nodes(double* __restrict__ f, const double* __restrict__ m, const double* __restrict__ g,
                                 const double* __restrict__ v, uint64_t N) {
      GRID_STRIDE(i, N) {
        const double mg = m[i] * g[i];
        for (int c = 0; c < 3; ++c) f[c * N + i] -= mg * v[c * N + i];
      }
    }

    // InvRotateDivideOp<scalar> (InvRotateDivide/Scalar.hpp:26-110)
    __global__ void k_inv_rotate_divide(double* __restrict__ kap, const double* __restrict__ Q,
                                        const double* __restrict__ D0, uint64_t E) {
      const uint64_t V = E - 1;
      GRID_STRIDE(i, V) {
        double qc[9], qn[9];
        for (int m = 0; m < 9; ++m) {
          qc[m] = Q[m * E + i];
          qn[m] = Q[m * E + i + 1];
        }
        const double tr = qc[0] * qn[0] + qc[1] * qn[1] + qc[2] * qn[2] + qc[3] * qn[3] + qc[4] * qn[4] +
                          qc[5] * qn[5] + qc[6] * qn[6] + qc[7] * qn[7] + qc[8] * qn[8];
        const double rest = 0.5 + 1e-10;
        const double theta = math::acos(0.5 * tr - rest);
        const double st = 0.5 * theta / math::sin(theta + 1e-14) / D0[i];
        kap[i] = st * (-qc[3] * qn[6] - qc[4] * qn[7] - qc[5] * qn[8] + qc[6] * qn[3] + qc[7] * qn[4] + qc[8] * qn[5]);
        kap[V + i] = st * (qc[0] * qn[6] + qc[1] * qn[7] + qc[2] * qn[8] - qc[6] * qn[0] - qc[7] * qn[1] - qc[8] * qn[2]);
        kap[2 * V + i] = st * (-qc[0] * qn[3] - qc[1] * qn[4] - qc[2] * qn[5] + qc[3] * qn[0] + qc[4] * qn[1] + qc[5] * qn[2]);
      }
    }

    // inverse_cube (OpsTraits.hpp:98-105): 1 / (v * v * v)
    __global__ void k_inv_cube(double* __restrict__ o, const double* __restrict__ a, uint64_t n) {
      GRID_STRIDE(i, n) o[i] = 1.0 / (a[i] * a[i] * a[i]);
    }

    // batch_cross (CrossProduct/Operation.hpp)
    __global__ void k_cross(double* __restrict__ o, const double* __restrict__ a, const double* __restrict__ b, uint64_t n) {
      GRID_STRIDE(k, n) {
        const double a0 = a[k], a1 = a[n + k], a2 = a[2 * n + k];
        const double b0 = b[k], b1 = b[n + k], b2 = b[2 * n + k];
        o[k] = a1 * b2 - a2 * b1;
        o[n + k] = a2 * b0 - a0 * b2;
        o[2 * n + k] = a0 * b1 - a1 * b0;
      }
    }

    __global__ void k_add(double* __restrict__ o, const double* __restrict__ a, uint64_t n) {
      GRID_STRIDE(i, n) o[i] += a[i];
    }

    // o_c += a_c * s
    __global__ void k_add_scaled3(double* __restrict__ o, const double* __restrict__ a, const double* __restrict__ s, uint64_t n) {
      GRID_STRIDE(k, n) {
        for (int c = 0; c < 3; ++c) o[c * n + k] += a[c * n + k] * s[k];
      }
    }

    // difference_kernel of nodal velocity: v_{k+1} - v_k
    __global__ void k_diff_nodes(double* __restrict__ o, const double* __restrict__ v, uint64_t E) {
      const uint64_t N = E + 1;
      GRID_STRIDE(k, E) {
        for (int c = 0; c < 3; ++c) o[c * E + k] = v[c * N + k + 1] - v[c * N + k];
      }
    }

    // dilatation_rate / dilatation = dot(t, dv) / (e * l0)
    __global__ void k_rate(double* __restrict__ o, const double* __restrict__ t, const double* __restrict__ dv,
                           const double* __restrict__ e, const double* __restrict__ l0, uint64_t E) {
      GRID_STRIDE(k, E) {
        o[k] = (t[k] * dv[k] + t[E + k] * dv[E + k] + t[2 * E + k] * dv[2 * E + k]) / (e[k] * l0[k]);
      }
    }

    // internal_torque -= ((0.5 (m_k + m_{k+1})) * dissipation) % omega
    __global__ void k_damp_torque(double* __restrict__ tau, const double* __restrict__ m, const double* __restrict__ g,
                                  const double* __restrict__ w, uint64_t E) {
      GRID_STRIDE(k, E) {
        const double em = 0.5 * (m[k] + m[k + 1]);
        const double eg = em * g[k];
        for (int c = 0; c < 3; ++c) tau[c * E + k] -= eg * w[c * E + k];
      }
    }

    // acceleration = (f_int + f_ext) * inv_mass
    __global__ void k_accel(double* __restrict__ a, const double* __restrict__ fi, const double* __restrict__ fe,
                            const double* __restrict__ im, uint64_t N) {
      GRID_STRIDE(i, N) {
        for (int c = 0; c < 3; ++c) a[c * N + i] = (fi[c * N + i] + fe[c * N + i]) * im[i];
      }
    }

    // angular acceleration = (J^-1 % (tau_int + tau_ext)) * e
    __global__ void k_ang_accel(double* __restrict__ al, const double* __restrict__ iJ, const double* __restrict__ ti,
                                const double* __restrict__ te, const double* __restrict__ e, uint64_t E) {
      GRID_STRIDE(k, E) {
        for (int c = 0; c < 3; ++c) al[c * E + k] = (iJ[c * E + k] * (ti[c * E + k] + te[c * E + k])) * e[k];
      }
    }

    // ---- hooks ---------------------------------------------------------------

    // ConstrainPrimitives of a fixed clamp at index i: node i's position and
    // element min(i, n - 1)'s director are held
    __global__ void k_clamp_values(double* __restrict__ x, double* __restrict__ Q, const RodInfo* __restrict__ rods,
                                   uint64_t n_rods, uint64_t E) {
      const uint64_t N = E + 1;
      GRID_STRIDE(r, n_rods) {
        const RodInfo& ri = rods[r];
        for (int k = 0; k < 2; ++k) {
          const uint32_t i = ri.clamp_idx[k];
          if (!(ri.flags & clamp_bit(k)) || i > ri.n) continue;
          const uint64_t s = ri.start + i;
          const uint64_t se = ri.start + (i < ri.n ? i : ri.n - 1);
          for (int d = 0; d < 3; ++d) x[d * N + s] = ri.clamp[k].x[d];
          for (int m = 0; m < 9; ++m) Q[m * E + se] = ri.clamp[k].q[m];
        }
      }
    }

    // ConstrainRates of the same clamp: its node's velocity and its
    // element's angular velocity are zeroed
    __global__ void k_clamp_rates(double* __restrict__ v, double* __restrict__ w, const RodInfo* __restrict__ rods,
                                  uint64_t n_rods, uint64_t E) {
      const uint64_t N = E + 1;
      GRID_STRIDE(r, n_rods) {
        const RodInfo& ri = rods[r];
        for (int k = 0; k < 2; ++k) {
          const uint32_t i = ri.clamp_idx[k];
          if (!(ri.flags & clamp_bit(k)) || i > ri.n) continue;
          const uint64_t s = ri.start + i;
          const uint64_t se = ri.start + (i < ri.n ? i : ri.n - 1);
          for (int d = 0; d < 3; ++d) v[d * N + s] = 0.0;
          for (int d = 0; d < 3; ++d) w[d * E + se] = 0.0;
        }
      }
    }

    __device__ inline double ramp_value(int kind, double t, double start, double interval) {
      if (kind == EROD_RAMP_NONE) return 1.0;
      if (t < start) return 0.0;
      if (t > start + interval) return 1.0;
      if (kind == EROD_RAMP_LINEAR) return (t - start) / interval;
      const double s = 2.0 * ((t - start) / interval) - 1.0;
      return 0.5 * (s + (1.0 + (math::sin(M_PI * s) / M_PI)));
    }

    __global__ void k_end_force(double* __restrict__ fe, const RodInfo* __restrict__ rods, uint64_t n_rods, uint64_t E, double t) {
      const uint64_t N = E + 1;
      GRID_STRIDE(r, n_rods) {
        const RodInfo& ri = rods[r];
        if (!(ri.flags & 2u)) continue;
        const uint64_t node = ri.start + ri.n;
        const double rv = ramp_value(ri.ramp, t, ri.ramp_start, ri.ramp_interval);
        for (int d = 0; d < 3; ++d) fe[d * N + node] = ri.force[d] * rv;
      }
    }

    // flags the ghost node between two rods
    __global__ void k_mark_ghost_nodes(uint8_t* __restrict__ flag, const uint64_t* __restrict__ ghost, uint64_t n) {
      GRID_STRIDE(i, n) flag[ghost[i]] = 1;
    }

    // gravity on every node (the block-wide update of GravityField.hpp:239-246);
    // plane contact on the rods' own nodes only
    __global__ void k_environment(double* __restrict__ fe, const double* __restrict__ m, const double* __restrict__ x,
                                  const double* __restrict__ v, const uint8_t* __restrict__ is_ghost, uint64_t E,
                                  Env env) {
      const uint64_t N = E + 1;
      GRID_STRIDE(i, N) {
        if (env.gravity)
          for (int c = 0; c < 3; ++c) fe[c * N + i] += env.g[c] * m[i];
        if (env.contact) {
          // the plane n . x = z0; the node's height above it, less the offset
          const double gap = ((env.n[0] * x[i] + env.n[1] * x[N + i]) + env.n[2] * x[2 * N + i]) - env.z0 - env.offset;
          if (gap < 0.0 && !is_ghost[i]) {
            const double delta = -gap;
            const double vnv = (env.n[0] * v[i] + env.n[1] * v[N + i]) + env.n[2] * v[2 * N + i];
            const double vn = -vnv;  // approach speed along the normal
            const double fn = fmax(0.0, env.k * delta + env.c * vn);
            double vt[3];
            for (int c = 0; c < 3; ++c) vt[c] = v[c * N + i] - vnv * env.n[c];
            const double vtn = sqrt((vt[0] * vt[0] + vt[1] * vt[1]) + vt[2] * vt[2]);
            for (int c = 0; c < 3; ++c) fe[c * N + i] += fn * env.n[c];
            if (vtn > 0.0) {
              const double ft = fmin(env.ct * vtn, env.mu * fn);
              const double s = ft / vtn;  // f_T along the unit slip
              for (int c = 0; c < 3; ++c) fe[c * N + i] -= s * vt[c];
            }
          }
        }
      }
    }

  }  // namespace

#define L(kernel, n, ...) hipLaunchKernelGGL(kernel, dim3(grid_for(n)), dim3(kThreads), 0, b->stream, __VA_ARGS__)
#define F(id) (b->d_field[id])

  int ops_set_rod_clocks(erod_batch* b, double t) {
    L(k_fill, b->n_rods, b->d_time, t, b->n_rods);
    return hip_check(hipGetLastError(), "set rod clocks");
  }

  int ops_mark_ghost_nodes(erod_batch* b) {
    EROD_HIP(hipMemsetAsync(b->d_node_is_ghost, 0, b->E + 1, b->stream));
    if (!b->ghost_nodes.empty()) L(k_mark_ghost_nodes, b->ghost_nodes.size(), b->d_node_is_ghost, b->d_ghost_nodes,
                                   static_cast<uint64_t>(b->ghost_nodes.size()));
    return hip_check(hipGetLastError(), "mark ghost nodes");
  }

  int ops_fill_ghosts(erod_batch* b, int field) {
    if (b->ghost_nodes.empty()) return EROD_OK;
    const FieldInfo& fi = field_info(field);
    const int deficit = fi.stagger == EROD_NODE ? 1 : (fi.stagger == EROD_ELEMENT ? 2 : 3);
    const uint64_t n = b->ghost_nodes.size();
    L(k_fill_ghosts, n * deficit, F(field), b->d_ghost_nodes, n, deficit, b->field_slots(field), fi.ncomp, fi.ghost);
    return hip_check(hipGetLastError(), "fill_ghosts");
  }

  // kinematic_states() += c * kinematic_rates(): x += c v, Q <- exp(-c [w]x) Q
  int ops_kinematic_update_by(erod_batch* b, double c) {
    const uint64_t E = b->E, N = E + 1;
    L(k_axpy, 3 * N, F(EROD_POSITION), F(EROD_VELOCITY), c, 3 * N);
    L(k_so3, E, F(EROD_DIRECTOR), F(EROD_ANGULAR_VELOCITY), c, E);
    return hip_check(hipGetLastError(), "kinematic_update");
  }

  // PositionVerlet's kinematic substep: the coefficient (1/2) dt
  int ops_kinematic_update(erod_batch* b, double dt) { return ops_kinematic_update_by(b, 0.5 * dt); }

  // every slot of a field (ghosts included) set to one value: the block-wide
  // assignment of a scalar to a variable
  int ops_fill_field(erod_batch* b, int field, double v) {
    const uint64_t n = uint64_t(field_info(field).ncomp) * b->field_slots(field);
    L(k_fill, n, b->d_field[field], v, n);
    return hip_check(hipGetLastError(), "fill_field");
  }

  int ops_zero_external(erod_batch* b) {
    const uint64_t E = b->E, N = E + 1;
    L(k_fill, 3 * N, F(EROD_EXTERNAL_LOADS), 0.0, 3 * N);
    L(k_fill, 3 * E, F(EROD_EXTERNAL_TORQUES), 0.0, 3 * E);
    return hip_check(hipGetLastError(), "zero_external");
  }

  static int compute_internal_forces(erod_batch* b) {
    const uint64_t E = b->E, N = E + 1, V = E - 1;
    // update_spanwise_variables
    L(k_norm_diff, E, F(EROD_ELEMENT_LENGTH), F(EROD_POSITION), E);
    int rc = ops_fill_ghosts(b, EROD_ELEMENT_LENGTH);
    if (rc) return rc;
    L(k_tangent, E, F(EROD_TANGENT), F(EROD_POSITION), F(EROD_ELEMENT_LENGTH), E);
    L(k_div, E, F(EROD_ELEMENT_DILATATION), F(EROD_ELEMENT_LENGTH), F(EROD_REFERENCE_ELEMENT_LENGTH), E);
    L(k_avg, V, F(EROD_VORONOI_LENGTH), F(EROD_ELEMENT_LENGTH), V);
    L(k_div, V, F(EROD_VORONOI_DILATATION), F(EROD_VORONOI_LENGTH), F(EROD_REFERENCE_VORONOI_LENGTH), V);
    // compute_cross_sectional_variables
    L(k_radius, E, F(EROD_ELEMENT_DIMENSION), F(EROD_ELEMENT_VOLUME), F(EROD_ELEMENT_LENGTH), E);
    // update_shear_stretch_strain: sigma = (Q t) e - z
    L(k_matvec, E, F(EROD_SHEAR_STRETCH_STRAIN), F(EROD_DIRECTOR), F(EROD_TANGENT), E);
    L(k_scale3, E, F(EROD_SHEAR_STRETCH_STRAIN), F(EROD_ELEMENT_DILATATION), E);
    L(k_sub_z, E, F(EROD_SHEAR_STRETCH_STRAIN), E);
    // n = S % (sigma - sigma0)
    L(k_diag_mul_diff, 3 * E, F(EROD_INTERNAL_STRESS), F(EROD_SHEAR_STRETCH_RIGIDITY), F(EROD_SHEAR_STRETCH_STRAIN),
      F(EROD_REFERENCE_SHEAR_STRETCH_STRAIN), 3 * E);
    // f = diff(Q^T n / e) with zeroed ghosts
    L(k_mattranspvec, E, F(EROD_DUMMY_ELEMENT_VECTOR2), F(EROD_DIRECTOR), F(EROD_INTERNAL_STRESS), E);
    L(k_div3, E, F(EROD_DUMMY_ELEMENT_VECTOR), F(EROD_DUMMY_ELEMENT_VECTOR2), F(EROD_ELEMENT_DILATATION), E);
    rc = ops_fill_ghosts(b, EROD_DUMMY_ELEMENT_VECTOR);
    if (rc) return rc;
    L(k_two_point_diff, N, F(EROD_INTERNAL_LOADS), F(EROD_DUMMY_ELEMENT_VECTOR), E);
    if (b->damped)
      L(k_damp_nodes, N, F(EROD_INTERNAL_LOADS), F(EROD_MASS), F(EROD_FORCE_DAMPING_RATE), F(EROD_VELOCITY), N);
    return hip_check(hipGetLastError(), "compute_internal_forces");
  }

  static int compute_internal_torques(erod_batch* b) {
    const uint64_t E = b->E, V = E - 1;
    double* d1 = F(EROD_DUMMY_ELEMENT_VECTOR);
    double* d2 = F(EROD_DUMMY_ELEMENT_VECTOR2);
    double* dv = F(EROD_DUMMY_VORONOI_VECTOR);
    double* D = F(EROD_VORONOI_LENGTH);
    double* l = F(EROD_ELEMENT_LENGTH);
    double* tau = F(EROD_INTERNAL_TORQUES);
    L(k_inv_rotate_divide, V, F(EROD_CURVATURE), F(EROD_DIRECTOR), F(EROD_REFERENCE_VORONOI_LENGTH), E);
    int rc = ops_fill_ghosts(b, EROD_CURVATURE);
    if (rc) return rc;
    L(k_diag_mul_diff, 3 * V, F(EROD_INTERNAL_COUPLE), F(EROD_BENDING_TWIST_RIGIDITY), F(EROD_CURVATURE),
      F(EROD_REFERENCE_CURVATURE), 3 * V);
    L(k_inv_cube, V, D, F(EROD_VORONOI_DILATATION), V);
    L(k_two_point_diff_scaled, E, tau, F(EROD_INTERNAL_COUPLE), D, V);
    L(k_cross, V, dv, F(EROD_CURVATURE), F(EROD_INTERNAL_COUPLE), V);
    L(k_mul_inplace, V, D, F(EROD_REFERENCE_VORONOI_LENGTH), V);
    L(k_quad_scaled, E, d1, dv, D, V);
    L(k_add, 3 * E, tau, d1, 3 * E);
    L(k_matvec, E, d1, F(EROD_DIRECTOR), F(EROD_TANGENT), E);
    L(k_cross, E, d2, d1, F(EROD_INTERNAL_STRESS), E);
    L(k_add_scaled3, E, tau, d2, F(EROD_REFERENCE_ELEMENT_LENGTH), E);
    L(k_mul, 3 * E, d2, F(EROD_MASS_SECOND_MOMENT_OF_INERTIA), F(EROD_ANGULAR_VELOCITY), 3 * E);
    L(k_div3, E, d1, d2, F(EROD_ELEMENT_DILATATION), E);
    L(k_cross, E, d2, d1, F(EROD_ANGULAR_VELOCITY), E);
    L(k_add, 3 * E, tau, d2, 3 * E);
    L(k_diff_nodes, E, d2, F(EROD_VELOCITY), E);
    L(k_rate, E, l, F(EROD_TANGENT), d2, F(EROD_ELEMENT_DILATATION), F(EROD_REFERENCE_ELEMENT_LENGTH), E);
    L(k_add_scaled3, E, tau, d1, l, E);
    L(k_mul, V, D, F(EROD_VORONOI_DILATATION), F(EROD_REFERENCE_VORONOI_LENGTH), V);
    L(k_mul, E, l, F(EROD_ELEMENT_DILATATION), F(EROD_REFERENCE_ELEMENT_LENGTH), E);
    if (b->damped)
      L(k_damp_torque, E, tau, F(EROD_MASS), F(EROD_TORQUE_DAMPING_RATE), F(EROD_ANGULAR_VELOCITY), E);
    return hip_check(hipGetLastError(), "compute_internal_torques");
  }

  int ops_compute_internal_dynamics(erod_batch* b) {
    int rc = ops_zero_external(b);
    if (rc) return rc;
    rc = compute_internal_forces(b);
    if (rc) return rc;
    return compute_internal_torques(b);
  }

  int ops_dynamic_update(erod_batch* b, double dt) {
    const uint64_t E = b->E, N = E + 1;
    L(k_accel, N, F(EROD_ACCELERATION), F(EROD_INTERNAL_LOADS), F(EROD_EXTERNAL_LOADS), F(EROD_INV_MASS), N);
    L(k_ang_accel, E, F(EROD_ANGULAR_ACCELERATION), F(EROD_INV_MASS_SECOND_MOMENT_OF_INERTIA),
      F(EROD_INTERNAL_TORQUES), F(EROD_EXTERNAL_TORQUES), F(EROD_ELEMENT_DILATATION), E);
    L(k_axpy, 3 * N, F(EROD_VELOCITY), F(EROD_ACCELERATION), dt, 3 * N);
    L(k_axpy, 3 * E, F(EROD_ANGULAR_VELOCITY), F(EROD_ANGULAR_ACCELERATION), dt, 3 * E);
    return hip_check(hipGetLastError(), "dynamic_update");
  }

  int ops_apply_post_kinematic(erod_batch* b) {
    if (!b->has_clamps) return EROD_OK;
    L(k_clamp_values, b->n_rods, F(EROD_POSITION), F(EROD_DIRECTOR), b->d_rods, b->n_rods, b->E);
    return hip_check(hipGetLastError(), "post_kinematic");
  }

  // Forcing first, then environments (the module order of the oracle).
  int ops_apply_pre_dynamic(erod_batch* b, double time) {
    const uint64_t N = b->E + 1;
    if (b->has_forces) L(k_end_force, b->n_rods, F(EROD_EXTERNAL_LOADS), b->d_rods, b->n_rods, b->E, time);
    if (b->env.gravity || b->env.contact)
      L(k_environment, N, F(EROD_EXTERNAL_LOADS), F(EROD_MASS), F(EROD_POSITION), F(EROD_VELOCITY), b->d_node_is_ghost,
        b->E, b->env);
    return hip_check(hipGetLastError(), "pre_dynamic");
  }

  int ops_apply_post_dynamic(erod_batch* b) {
    if (!b->has_clamps) return EROD_OK;
    L(k_clamp_rates, b->n_rods, F(EROD_VELOCITY), F(EROD_ANGULAR_VELOCITY), b->d_rods, b->n_rods, b->E);
    return hip_check(hipGetLastError(), "post_dynamic");
  }

#undef L
#undef F

}  // namespace erod
