// update.hpp -- the state update of one point (state_update_mod.h:92-283,
// with primitive_to_conserved :18-33 and conserved_vector_Ubar :35-90)
// templated on its scalar type, for the derivative kernels: on Dual<N> it
// carries N directional derivatives through the same operations the FP64
// update kernel performs.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "physics.hpp"

namespace lskum {

// state_update_mod.h:18-33
template <typename T>
__device__ __forceinline__ void prim_to_cons_t(const T p[4], T nx, T ny, T U[4]) {
  const T rho = p[0];
  U[0] = rho;
  const T temp1 = rho * p[1];
  const T temp2 = rho * p[2];
  U[3] = T(2.5) * p[3] + T(0.5) * (temp1 * temp1 + temp2 * temp2) / rho;
  U[1] = temp1 * ny - temp2 * nx;
  U[2] = temp1 * nx + temp2 * ny;
}

// state_update_mod.h:35-90 (the outer boundary's split state)
template <typename T>
__device__ __forceinline__ void ubar_t(const T p[4], T nx, T ny, const FlowConsts &fc,
                                       T Ubar[4]) {
  const double rho_inf = fc.rho_inf, pr_inf = fc.pr_inf, gamma_new = 1.4;
  const T tx = ny, ty = -nx;
  const T u1_inf_rot = T(fc.u1_inf) * tx + T(fc.u2_inf) * ty;
  const T u2_inf_rot = T(fc.u1_inf) * nx + T(fc.u2_inf) * ny;
  T temp1 = u1_inf_rot * u1_inf_rot + u2_inf_rot * u2_inf_rot;
  const T e_inf = T(pr_inf / (rho_inf * (gamma_new - 1.0))) + T(0.5) * temp1;
  const double beta_inf = (0.5 * rho_inf) / pr_inf;
  T S2 = u2_inf_rot * T(sqrt(beta_inf));
  const T B2_inf = exp(-S2 * S2) / T(2.0 * sqrt(kPi * beta_inf));
  const T A2n_inf = T(0.5) * (T(1.0) - erf(S2));
  const T rho = p[0], u1 = p[1], u2 = p[2], pr = p[3];
  const T u1_rot = u1 * tx + u2 * ty;
  const T u2_rot = u1 * nx + u2 * ny;
  temp1 = u1_rot * u1_rot + u2_rot * u2_rot;
  const T e = pr / (rho * T(gamma_new - 1.0)) + T(0.5) * temp1;
  const T beta = rho / (T(2.0) * pr);
  S2 = u2_rot * sqrt(beta);
  const T B2 = exp(-S2 * S2) / (T(2.0) * sqrt(T(kPi) * beta));
  const T A2p = T(0.5) * (T(1.0) + erf(S2));
  Ubar[0] = (T(rho_inf) * A2n_inf) + (rho * A2p);
  Ubar[1] = (T(rho_inf) * u1_inf_rot * A2n_inf) + (rho * u1_rot * A2p);
  temp1 = T(rho_inf) * (u2_inf_rot * A2n_inf - B2_inf);
  T temp2 = rho * (u2_rot * A2p + B2);
  Ubar[2] = temp1 + temp2;
  temp1 = (T(rho_inf) * A2n_inf * e_inf - T(0.5 * rho_inf) * u2_inf_rot * B2_inf);
  temp2 = (rho * A2p * e + T(0.5) * rho * u2_rot * B2);
  Ubar[3] = temp1 + temp2;
}

// One SSPRK stage at one point: p1 the stage's state, p2 the state at the
// start of the iteration, R its flux residual -> the new primitive state.
template <typename T>
__device__ __forceinline__ void su_point(int rk, int type, const T p1[4], const T p2[4],
                                         const T R[4], T nx, T ny, const FlowConsts &fc,
                                         T out[4]) {
  T U[4], Uo[4];
  if (type != 2) {
    prim_to_cons_t(p1, nx, ny, U);
    prim_to_cons_t(p2, nx, ny, Uo);
  } else {
    ubar_t(p1, nx, ny, fc, U);
    ubar_t(p2, nx, ny, fc, Uo);
  }
  if (rk != 3) {
    for (int j = 0; j < 4; j++) U[j] = U[j] - T(0.5) * R[j];
  } else {
    for (int j = 0; j < 4; j++)
      U[j] = T(2.0 / 3.0) * Uo[j] + T(1.0 / 3.0) * (U[j] - T(0.5) * R[j]);
  }
  if (type == 0) U[2] = T(0.0);
  const T U2_rot = U[1], U3_rot = U[2];
  U[1] = U2_rot * ny + U3_rot * nx;
  U[2] = U3_rot * ny - U2_rot * nx;
  const T temp = T(1.0) / U[0];
  out[0] = U[0];
  out[1] = U[1] * temp;
  out[2] = U[2] * temp;
  out[3] = T(0.4) * U[3] - (T(0.2) * temp) * (U[1] * U[1] + U[2] * U[2]);
}

}  // namespace lskum
