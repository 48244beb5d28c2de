// LSKUM stencil kernels, templated on the state scalar S, and their launchers.
//
// Two translation units include this header.  mfcfd.hip instantiates the FP64 path and
// is compiled with -ffp-contract=off, so every FP64 expression keeps the reference's
// operation order.  mfcfd_f32.hip instantiates the mixed path (S = float) and is
// compiled with FMA contraction: that path is held to |dCl|, |dCd| <= 1e-4 of the FP64
// path, not to rounding parity, and a fused multiply-add is one VALU instruction where
// a multiply and an add are two (the flux kernel is VALU bound).  The *_any launchers
// route each call to the unit that owns its scalar type.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "lskum_device.h"

using namespace mf;

// Slots per batch of a stencil walk (walk_stencil): a lane's share of a 24-slot stencil
// in two to four batches, with the batch's neighbour records kept within the register
// budget of the kernel's occupancy.
template <class S, int LPP>
constexpr int walk_batch() {
  return LPP >= 4 ? 3 : (sizeof(S) == 4 ? 4 : 2);
}

// q_variables.F90:8-37 over all points, plus each point's wave speed for func_delta
template <class S>
__global__ __launch_bounds__(BLOCK) void k_q_variables(int n, const V4<S>* __restrict__ prim,
                                                       V4<S>* __restrict__ q,
                                                       double* __restrict__ cspeed) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const V4<S> p = prim[i];
  q[i] = q_of_prim(p);
  cspeed[i] = wave_speed(p);
}

// Thread-to-point map of a stencil kernel.  Global walk: point_id() over BLOCK-thread
// blocks, LPP lanes per point, threads past the last point leave.  Staged walk: block b
// (XCD-remapped) holds points b*STAGE_P.. with LPP lanes each; every thread stays for
// the block's barrier, `live` tells whether its point exists.
template <int LPP, bool STAGED, int P = STAGE_P>
struct PointMap {
  int b, i, lane;
  bool live;
  __device__ PointMap(int n) {
    if constexpr (STAGED) {
      b = xcd_block(blockIdx.x, gridDim.x);
      i = b * P + (int)threadIdx.x / LPP;
      lane = (int)threadIdx.x % LPP;
    } else {
      const int t = point_id();
      b = 0;
      i = t / LPP;
      lane = t % LPP;
    }
    live = i < n;
  }
};

// LDS bytes of the staged kernels: per halo entry the records the kernel stages, plus the
// block's slot words (uint8 [W][P] in 64-point blocks, uint16 [W][P] in the flux kernel's)
inline size_t stage_lds_lsq(int hcap, int width, size_t ssz, bool delta) {
  return (size_t)hcap * (4 * ssz + sizeof(D2) + (delta ? sizeof(double) : 0)) + (size_t)width * STAGE_P;
}
inline size_t stage_lds_inner(int hcap, int width, size_t ssz, bool sums) {
  return (size_t)hcap * ((sums ? 8 : 12) * ssz + sizeof(D2)) + (size_t)width * STAGE_P;
}
// the resident inner sweeps (mixed path): halo dq and xy, slot words, and (more than
// one lane per point) the block's first-pass sums and solve factors
inline size_t sweeps_lds(int hcap, int width, int lanes) {
  return (size_t)hcap * (2 * sizeof(V4<float>) + sizeof(D2)) + (size_t)width * STAGE_P +
         (lanes > 1 ? 3 * STAGE_P * sizeof(V4<float>) : 0);
}
// the adjoint's LSQ reverse (FP64): q and xy per halo entry
inline size_t stage_lds_lsq_rev(int hcap, int width) {
  return (size_t)hcap * (4 * sizeof(double) + sizeof(D2)) + (size_t)width * STAGE_P;
}
// the adjoint's inner reverse (FP64): q, dq_x, dq_y and xy per halo entry
inline size_t stage_lds_inner_rev(int hcap, int width) {
  return (size_t)hcap * (3 * 4 * sizeof(double) + sizeof(D2)) + (size_t)width * STAGE_P;
}
inline size_t stage_lds_flux(int hcap, int width, size_t ssz) {
  return (size_t)hcap * (21 * ssz + sizeof(D2)) + (size_t)width * STAGE_PF * 2;
}

// The mixed path's 2x2 solve with the point's FP32 factors f = (sum dy^2, sum dx dy,
// sum dx^2) / det
__device__ inline void lsq_apply_f(const V4<float>& f, const double sx[4], const double sy[4], V4<float>& ox,
                                   V4<float>& oy) {
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    ox[r] = (float)(sx[r] * f.x - sy[r] * f.y);
    oy[r] = (float)(sy[r] * f.z - sx[r] * f.y);
  }
}

// The 2x2 least-squares solve of point i (q_variables.F90:104-110 and :198-208):
// d/dx = (Sx sum dy^2 - Sy sum dx dy) / det, d/dy = (Sy sum dx^2 - Sx sum dx dy) / det
// from the edge sums Sx, Sy.  FP64: the reference's operations; mixed path: the FP32
// products with 1/det of lsq_geo_f.
template <class S>
using LsqGeo = std::conditional_t<sizeof(S) == 4, V4<float>, V4d>;

template <class S>
__device__ inline LsqGeo<S> lsq_geo_of(const Geom& g, int i) {
  if constexpr (sizeof(S) == 4) return g.lsq_geo_f[i];
  else return g.lsq_geo[i];
}

template <class S>
__device__ inline void lsq_apply(const LsqGeo<S>& gg, const double sx[4], const double sy[4], V4<S>& ox,
                                 V4<S>& oy) {
  if constexpr (sizeof(S) == 4) {
    lsq_apply_f(gg, sx, sy, ox, oy);
  } else {
    // gg = (sum dx^2, sum dy^2, sum dxdy, 1/det)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      ox[r] = (sx[r] * gg.y - sy[r] * gg.z) * gg.w;
      oy[r] = (sy[r] * gg.x - sx[r] * gg.z) * gg.w;
    }
  }
}

template <class S>
__device__ inline void lsq_solve(const Geom& g, int i, const double sx[4], const double sy[4], V4<S>& ox,
                                 V4<S>& oy) {
  lsq_apply<S>(lsq_geo_of<S>(g, i), sx, sy, ox, oy);
}

// q_variables.F90:41-112 eval_q_derivatives; with DELTA also func_delta
// (state_update.F90:190-236), which gathers over the same stencil.
// qsum (mixed path): the first-pass sums sum w (dx, dy) (q_k - q_i) are also stored, for
// the inner iterations (q does not change between them).
template <class S, bool POWER0, bool DELTA, int LPP, bool STAGED>
__global__ __launch_bounds__(BLOCK) void k_lsq(Geom g, const V4<S>* __restrict__ q,
                                               V4<S>* __restrict__ dq, V4<S>* __restrict__ qm,
                                               const double* __restrict__ cspeed,
                                               double* __restrict__ delta, double cfl,
                                               double power, const Ctl* __restrict__ ctl,
                                               V4<float>* __restrict__ qsum) {
  // read now, tested before the stores: the load does not delay the gathers
  const int32_t halted = ctl ? ctl->halt : 0;
  const PointMap<LPP, STAGED> pm(g.n);
  if (!STAGED && !pm.live) return;  // whole groups: n * LPP threads are a multiple of LPP
  const int i = pm.live ? pm.i : g.n - 1, lane = pm.lane;
  const D2 pi = g.xy[i];
  const V4<S> qi = q[i];
  V4<S> qmax = qi, qmin = qi;
  double sxq[4] = {0.0, 0.0, 0.0, 0.0}, syq[4] = {0.0, 0.0, 0.0, 0.0};
  double min_delt = 1.0;
  const int nb = pm.live ? g.nbhs[i] : 0;
  // one edge; a slot past nb (global walk only) carries the point's own record with
  // valid = false: neutral for the strict max/min, masked for the sums
  auto edge = [&](const V4<S>& qk, const D2& pk, double ck, bool valid) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (qk[r] > qmax[r]) qmax[r] = qk[r];
      if (qk[r] < qmin[r]) qmin[r] = qk[r];
    }
    const double delx = pk.x - pi.x, dely = pk.y - pi.y;
    double w = 1.0;
    // the mixed path takes func_delta's time step in FP32 (it scales the pseudo-time
    // update, not the discrete equations the iteration converges to)
    constexpr bool DT32 = DELTA && POWER0 && sizeof(S) == 4;
    if (DT32) {
      const float dist = __builtin_amdgcn_sqrtf((float)(delx * delx + dely * dely));
      const double dt = (double)((float)cfl * (dist * __builtin_amdgcn_rcpf((float)ck)));
      if (valid && min_delt > dt) min_delt = dt;
    } else if (!POWER0 || DELTA) {
      const double dist = sqrt(delx * delx + dely * dely);
      if (!POWER0) w = pow(dist, power);
      if (DELTA) {
        double dt = dist / ck;
        dt = cfl * dt;
        if (valid && min_delt > dt) min_delt = dt;
      }
    }
    const double wx = w * delx, wy = w * dely;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const double d = (double)qk[r] - (double)qi[r];
      sxq[r] = valid ? sxq[r] + wx * d : sxq[r];
      syq[r] = valid ? syq[r] + wy * d : syq[r];
    }
  };
  if constexpr (STAGED) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int H = g.hl[0].hcap;
    V4<S>* s_q = (V4<S>*)smem;
    D2* s_xy = (D2*)(s_q + H);
    double* s_c = (double*)(s_xy + H);
    uint8_t* s_cl = (uint8_t*)(s_c + (DELTA ? H : 0));
    stage_and_walk<STAGE_P, LPP>(
        g, g.hl[0], pm.b, nb, s_cl,
        [&](int j, int k) {
          s_q[j] = q[k];
          s_xy[j] = g.xy[k];
          if (DELTA) s_c[j] = cspeed[k];
        },
        [&](int j, int) { edge(s_q[j], s_xy[j], DELTA ? s_c[j] : 0.0, true); },
        [&](int k, int) { edge(q[k], g.xy[k], DELTA ? cspeed[k] : 0.0, true); });
    if (!pm.live) return;
  } else {
    struct Nbr {
      V4<S> q;
      D2 p;
      double c;
    };
    walk_stencil<walk_batch<S, LPP>(), LPP>(
        g.conn + i, g.n, nb, lane, (uint32_t)i << 4,
        [&](uint32_t e) {
          const int k = (int)(e >> 4);
          return Nbr{q[k], g.xy[k], DELTA ? cspeed[k] : 0.0};
        },
        [&](const Nbr& nk, uint32_t, bool valid) { edge(nk.q, nk.p, nk.c, valid); });
  }
  if (LPP > 1) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      sxq[r] = group_sum<LPP>(sxq[r]);
      syq[r] = group_sum<LPP>(syq[r]);
      qmax[r] = group_max<LPP>(qmax[r]);
      qmin[r] = group_min<LPP>(qmin[r]);
    }
    if (DELTA) min_delt = group_min<LPP>(min_delt);
    if (lane > 1) return;
  }
  if (halted) return;
  V4<S> ox, oy;
  lsq_solve(g, i, sxq, syq, ox, oy);
  if constexpr (LPP == 1) {
    dq[2 * i] = ox;
    dq[2 * i + 1] = oy;
    qm[2 * i] = qmax;
    qm[2 * i + 1] = qmin;
    if (DELTA) delta[i] = min_delt;
    if (qsum) {
      qsum[2 * i] = V4<float>{(float)sxq[0], (float)sxq[1], (float)sxq[2], (float)sxq[3]};
      qsum[2 * i + 1] = V4<float>{(float)syq[0], (float)syq[1], (float)syq[2], (float)syq[3]};
    }
  } else {
    // every lane of the group holds the sums: lane 0 stores the x (max) half of each
    // pair, lane 1 the y (min) half, so one store covers a point's pair contiguously
    // (selects per component: a select of whole records would put them in scratch)
    const bool hy = lane == 1;
    V4<S> dv, mv;
    V4<float> sv;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      dv[r] = hy ? oy[r] : ox[r];
      mv[r] = hy ? qmin[r] : qmax[r];
      sv[r] = (float)(hy ? syq[r] : sxq[r]);
    }
    dq[2 * i + lane] = dv;
    qm[2 * i + lane] = mv;
    if (DELTA && !hy) delta[i] = min_delt;
    if (qsum) qsum[2 * i + lane] = sv;
  }
}

// One edge of an inner sweep in the sums form (mixed path): the terms
// -0.5 w (dx, dy) [dx (dq_x,k - dq_x,i) + dy (dq_y,k - dq_y,i)] of the edge sums, whose
// (q_k - q_i) part the LSQ kernel's first pass has already summed.  Shared by the
// per-sweep and the resident kernels so that both round alike.
template <class S, bool POWER0>
__device__ __forceinline__ void inner_edge_sums(const D2& pi, const V4<S>& dxi, const V4<S>& dyi, const D2& pk,
                                                const V4<S>& dxk, const V4<S>& dyk, double power, bool valid,
                                                double sxq[4], double syq[4]) {
  const double delx = pk.x - pi.x, dely = pk.y - pi.y;
  double w = 1.0;
  if (!POWER0) w = pow(sqrt(delx * delx + dely * dely), power);
  const double hx = 0.5 * w * delx, hy = 0.5 * w * dely;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const double d = delx * (double)(dxk[r] - dxi[r]) + dely * (double)(dyk[r] - dyi[r]);
    sxq[r] = valid ? sxq[r] - hx * d : sxq[r];
    syq[r] = valid ? syq[r] - hy * d : syq[r];
  }
}

// q_variables.F90:115-227 eval_q_inner_loop + eval_update_innerloop (Jacobi: reads
// dq_in, writes dq_out).  SUMS (mixed path): the edge sums of w (dx, dy) ((q_k - q_i)
// - 0.5 [dx (dq_x,k - dq_x,i) + dy (dq_y,k - dq_y,i)]) start from the LSQ kernel's
// first-pass sums, so only the neighbours' dq and coordinates are gathered; the FP64
// path keeps the reference's per-edge form.
template <class S, bool POWER0, int LPP, bool SUMS, bool STAGED>
__global__ __launch_bounds__(BLOCK) void k_inner(Geom g, const V4<S>* __restrict__ q,
                                                 const V4<S>* __restrict__ dq_in,
                                                 V4<S>* __restrict__ dq_out, double power,
                                                 const Ctl* __restrict__ ctl,
                                                 const V4<float>* __restrict__ qsum) {
  // read now, tested before the stores: the load does not delay the gathers
  const int32_t halted = ctl ? ctl->halt : 0;
  const PointMap<LPP, STAGED> pm(g.n);
  if (!STAGED && !pm.live) return;
  const int i = pm.live ? pm.i : g.n - 1, lane = pm.lane;
  // the point's own coordinates and dq: from LDS after the staged walk's barrier (own
  // points lead the halo), else from global memory
  D2 pi{};
  V4<S> dxi{}, dyi{};
  if constexpr (!STAGED) {
    pi = g.xy[i];
    dxi = dq_in[2 * i];
    dyi = dq_in[2 * i + 1];
  }
  double sxq[4] = {0.0, 0.0, 0.0, 0.0}, syq[4] = {0.0, 0.0, 0.0, 0.0};
  const int nb = pm.live ? g.nbhs[i] : 0;
  V4<S> qi{};
  if constexpr (!SUMS) {
    qi = q[i];
  }
  // one edge (qk unused in the sums form); valid as in k_lsq
  auto edge = [&](const D2& pk, const V4<S>& qk, const V4<S>& dxk, const V4<S>& dyk, bool valid) {
    if constexpr (SUMS) {
      inner_edge_sums<S, POWER0>(pi, dxi, dyi, pk, dxk, dyk, power, valid, sxq, syq);
    } else {
      const double delx = pk.x - pi.x, dely = pk.y - pi.y;
      double w = 1.0;
      if (!POWER0) w = pow(sqrt(delx * delx + dely * dely), power);
      const double wx = w * delx, wy = w * dely;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double temp1 = (double)qi[r] - 0.5 * (delx * (double)dxi[r] + dely * (double)dyi[r]);
        const double temp2 = (double)qk[r] - 0.5 * (delx * (double)dxk[r] + dely * (double)dyk[r]);
        sxq[r] = valid ? sxq[r] + (wx * (temp2 - temp1)) : sxq[r];
        syq[r] = valid ? syq[r] + (wy * (temp2 - temp1)) : syq[r];
      }
    }
  };
  if constexpr (STAGED) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int H = g.hl[0].hcap;
    V4<S>* s_dx = (V4<S>*)smem;
    V4<S>* s_dy = s_dx + H;
    V4<S>* s_q = s_dy + H;  // FP64 form only
    D2* s_xy = (D2*)(s_q + (SUMS ? 0 : H));
    uint8_t* s_cl = (uint8_t*)(s_xy + H);
    stage_and_walk<STAGE_P, LPP>(
        g, g.hl[0], pm.b, nb, s_cl,
        [&](int j, int k) {
          s_dx[j] = dq_in[2 * k];
          s_dy[j] = dq_in[2 * k + 1];
          if (!SUMS) s_q[j] = q[k];
          s_xy[j] = g.xy[k];
        },
        [&](int j, int) { edge(s_xy[j], SUMS ? qi : s_q[j], s_dx[j], s_dy[j], true); },
        [&](int k, int) { edge(g.xy[k], SUMS ? qi : q[k], dq_in[2 * k], dq_in[2 * k + 1], true); },
        [&] {
          const int li = (int)threadIdx.x / LPP;
          pi = s_xy[li];
          dxi = s_dx[li];
          dyi = s_dy[li];
        });
    if (!pm.live) return;
  } else if constexpr (SUMS) {
    struct Nbr {
      D2 p;
      V4<S> dx, dy;
    };
    walk_stencil<walk_batch<S, LPP>(), LPP>(
        g.conn + i, g.n, nb, lane, (uint32_t)i << 4,
        [&](uint32_t e) {
          const int k = (int)(e >> 4);
          return Nbr{g.xy[k], dq_in[2 * k], dq_in[2 * k + 1]};
        },
        [&](const Nbr& nk, uint32_t, bool valid) { edge(nk.p, qi, nk.dx, nk.dy, valid); });
  } else {
    struct Nbr {
      D2 p;
      V4<S> q, dx, dy;
    };
    walk_stencil<walk_batch<S, LPP>(), LPP>(
        g.conn + i, g.n, nb, lane, (uint32_t)i << 4,
        [&](uint32_t e) {
          const int k = (int)(e >> 4);
          return Nbr{g.xy[k], q[k], dq_in[2 * k], dq_in[2 * k + 1]};
        },
        [&](const Nbr& nk, uint32_t, bool valid) { edge(nk.p, nk.q, nk.dx, nk.dy, valid); });
  }
  if (LPP > 1) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      sxq[r] = group_sum<LPP>(sxq[r]);
      syq[r] = group_sum<LPP>(syq[r]);
    }
    if (lane > 1) return;
  }
  if (halted) return;
  if constexpr (SUMS) {
    // the first-pass sums join after the edge terms
    const V4<float> sa = qsum[2 * i], sb = qsum[2 * i + 1];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      sxq[r] = (double)sa[r] + sxq[r];
      syq[r] = (double)sb[r] + syq[r];
    }
  }
  V4<S> ox, oy;
  lsq_solve(g, i, sxq, syq, ox, oy);
  if constexpr (LPP == 1) {
    dq_out[2 * i] = ox;
    dq_out[2 * i + 1] = oy;
  } else {
    // lanes 0 and 1 hold the same sums: each stores one half of the point's pair
    V4<S> dv;
#pragma unroll
    for (int r = 0; r < 4; ++r) dv[r] = lane == 1 ? oy[r] : ox[r];
    dq_out[2 * i + lane] = dv;
  }
}

template <class S, bool POWER0, int LPP>
__device__ inline void flux_dispatch(const Geom& g, const State<S>& st, int i, double delta,
                                     double power, double fr[4], int lane) {
  if (i < g.nw) flux_point<WALL, S, POWER0, LPP>(g, st, i, delta, power, fr, lane);
  else if (i < g.nw + g.no) flux_point<OUTER, S, POWER0, LPP>(g, st, i, delta, power, fr, lane);
  else flux_point<INTERIOR, S, POWER0, LPP>(g, st, i, delta, power, fr, lane);
  if (LPP > 1) {
#pragma unroll
    for (int r = 0; r < 4; ++r) fr[r] = group_sum<LPP>(fr[r]);
  }
}

// flux_residual.F90:14-73 cal_flux_residual (unfused: stores flux_res)
template <class S, bool POWER0>
__global__ __launch_bounds__(BLOCK, 2) void k_flux(Geom g, State<S> st,
                                                const double* __restrict__ delta,
                                                V4d* __restrict__ flux_res, double power) {
  const int i = point_id();
  if (i >= g.n) return;
  double fr[4];
  flux_dispatch<S, POWER0, 1>(g, st, i, delta[i], power, fr, 0);
  flux_res[i] = V4d{fr[0], fr[1], fr[2], fr[3]};
}

template <class S>
__device__ inline V4<S> update_dispatch(const Geom& g, int i, const V4<S>& p, const V4<S>& po,
                                        const double fr[4], int rk, double euler,
                                        const Freestream& fs, double* rs) {
  const D2 nn = g.nrm[i];
  if (i < g.nw) return update_point<WALL, S>(p, po, fr, nn.x, nn.y, rk, euler, fs, rs);
  if (i < g.nw + g.no) {
    // the far-field state of the point's own case
    const Freestream f = g.fs_tab ? g.fs_tab[g.case_id[i]] : fs;
    return update_point<OUTER, S>(p, po, fr, nn.x, nn.y, rk, euler, f, rs);
  }
  return update_point<INTERIOR, S>(p, po, fr, nn.x, nn.y, rk, euler, fs, rs);
}

// state_update.F90:10-141 state_update(rk) (unfused: in place, partial residual sums)
template <class S>
__global__ __launch_bounds__(BLOCK) void k_update(Geom g, V4<S>* __restrict__ prim,
                                                  const V4<S>* __restrict__ prim_old,
                                                  const V4d* __restrict__ flux_res,
                                                  double* __restrict__ partial, int rk,
                                                  double euler, Freestream fs) {
  const int i = point_id();
  double rs = 0.0;
  if (i < g.n) {
    const V4d f = flux_res[i];
    const double fr[4] = {f.x, f.y, f.z, f.w};
    prim[i] = update_dispatch<S>(g, i, prim[i], prim_old[i], fr, rk, euler, fs, &rs);
  }
  const double t = block_sum(rs);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// Fused cal_flux_residual + state_update(rk) + eval_q_variables of the new state.
// Reads the stage's (q, dq, qm, prim), writes the next (prim, q) buffers; stage 1 also
// saves prim_old (fpi_solver.F90:20-22), the last stage the residual partial sums and
// the wave speeds for the next iteration's func_delta.
template <class S, bool POWER0, int LPP, bool STAGED>
__global__ __launch_bounds__(BLOCK, 2) void k_flux_update(
    Geom g, State<S> st, const V4<S>* __restrict__ prim, V4<S>* __restrict__ prim_old,
    V4<S>* __restrict__ prim_out, V4<S>* __restrict__ q_out, double* __restrict__ cspeed,
    const double* __restrict__ delta, double* __restrict__ partial, const Ctl* __restrict__ ctl,
    int rk, int last, double euler, double power, Freestream fs, double* __restrict__ rs_pt) {
  if (ctl->halt) return;
  const PointMap<LPP, STAGED, STAGE_PF> pm(g.n);
  const int i = pm.i, lane = pm.lane;
  double rs = 0.0;
  double fr[4] = {0.0, 0.0, 0.0, 0.0};
  if constexpr (STAGED) {
    // the block's halo in LDS, then the points' stencils from it (every thread reaches
    // the barrier and the group sums)
    extern __shared__ __align__(16) unsigned char smem[];
    HaloRecs<S> r;
    uint16_t* s_cl;
    const bool staged = stage_flux_halo<S, LPP, true>(g, st, pm.b, smem, r, s_cl);
    if (pm.live) {
      const double dl = delta[i];
      if (staged) {
        const StagedNbrs<S, STAGE_PF, LPP, true> nbrs{g, r, s_cl, (int)threadIdx.x / LPP, g.nbhs[i], lane};
        if (i < g.nw) flux_point_frame<WALL, S, POWER0, LPP>(g, st, i, dl, power, fr, nbrs);
        else if (i < g.nw + g.no) flux_point_frame<OUTER, S, POWER0, LPP>(g, st, i, dl, power, fr, nbrs);
        else flux_point_frame<INTERIOR, S, POWER0, LPP>(g, st, i, dl, power, fr, nbrs);
      } else {
        const GlobalNbrs<S> nbrs{g, st, i, g.nbhs[i], lane, LPP};
        if (i < g.nw) flux_point_frame<WALL, S, POWER0, LPP>(g, st, i, dl, power, fr, nbrs);
        else if (i < g.nw + g.no) flux_point_frame<OUTER, S, POWER0, LPP>(g, st, i, dl, power, fr, nbrs);
        else flux_point_frame<INTERIOR, S, POWER0, LPP>(g, st, i, dl, power, fr, nbrs);
      }
    }
    if (LPP > 1) {
#pragma unroll
      for (int r2 = 0; r2 < 4; ++r2) fr[r2] = group_sum<LPP>(fr[r2]);
    }
  } else {
    // every lane reaches the group sums (no early exit before them)
    if (pm.live) flux_dispatch<S, POWER0, LPP>(g, st, i, delta[i], power, fr, lane);
    else if (LPP > 1) {
#pragma unroll
      for (int r = 0; r < 4; ++r) fr[r] = group_sum<LPP>(fr[r]);
    }
  }
  if (pm.live && lane == 0) {
    const V4<S> p = prim[i];
    if (rk == 1) prim_old[i] = p;
    const V4<S> po = (rk == 3) ? prim_old[i] : p;
    const V4<S> pn = update_dispatch<S>(g, i, p, po, fr, rk, euler, fs, &rs);
    prim_out[i] = pn;
    q_out[i] = q_of_prim(pn);
    if (last) {
      cspeed[i] = wave_speed(pn);
      if (rs_pt) rs_pt[i] = rs;
    }
  }
  if (last) {
    const double tot = block_sum(rs);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
  }
}

// state_update.F90:190-236 func_delta (unfused)
template <class S>
__global__ __launch_bounds__(BLOCK) void k_func_delta(Geom g, const V4<S>* __restrict__ prim,
                                                      double* __restrict__ delta, double cfl) {
  const int i = point_id();
  if (i >= g.n) return;
  const D2 pi = g.xy[i];
  double min_delt = 1.0;
  const int nb = g.nbhs[i];
  for (int s = 0; s < nb; ++s) {
    const int k = (int)(g.conn[(size_t)s * g.n + i] >> 4);
    const D2 pk = g.xy[k];
    double dist = (pk.x - pi.x) * (pk.x - pi.x) + (pk.y - pi.y) * (pk.y - pi.y);
    dist = sqrt(dist);
    double dt = dist / wave_speed(prim[k]);
    dt = cfl * dt;
    if (min_delt > dt) min_delt = dt;
  }
  delta[i] = min_delt;
}


// compute_force_coeffs.F90:11-102, one block per shape id (a batch numbers the shapes
// of case c as c * shapes + s).  Shape points are grouped by shape id (sh_ptr), each
// with its (nx, ny, ds, mx - 0.25, my) in sgeo and its ordinal among the shape points
// in original order (sh_pos, the cp output's order); shape_par per shape id holds
// (0.5 rho_inf M^2, cos theta, sin theta) of its case.  Cl/Cd/Cm into clcdcm[3 s + k].
template <class S>
__global__ __launch_bounds__(BLOCK) void k_forces(const V4<S>* __restrict__ prim,
                                                  const int32_t* __restrict__ sh_m,
                                                  const int32_t* __restrict__ sh_pos,
                                                  const double* __restrict__ sgeo,
                                                  const int32_t* __restrict__ sh_ptr,
                                                  const double* __restrict__ shape_par, int nsid,
                                                  double* __restrict__ cp, double* __restrict__ clcdcm,
                                                  double* __restrict__ cd_hist,
                                                  const Ctl* __restrict__ ctl) {
  if (ctl && ctl->halt) return;
  const double pr_inf = 1.0 / 1.4;
  const int s = blockIdx.x;
  const double cp_scale = shape_par[3 * s], ct = shape_par[3 * s + 1], st = shape_par[3 * s + 2];
  double H = 0.0, V = 0.0, M = 0.0;
  const int j1 = sh_ptr[s + 1];
  auto point = [&](int j, double pw) {
    const double c0 = pw - pr_inf;
    const double c = -c0 / cp_scale;
    cp[sh_pos[j]] = c;
    const double nx = sgeo[5 * j], ny = sgeo[5 * j + 1], ds = sgeo[5 * j + 2];
    H += c * nx * ds;
    V += c * ny * ds;
    M += (-c * ny * ds * sgeo[5 * j + 3] + c * nx * ds * sgeo[5 * j + 4]);
  };
  // four points' gathers (index, then pressure) issued together, then used in order
  const int bs = (int)blockDim.x;
  int j = sh_ptr[s] + (int)threadIdx.x;
  for (; j + 3 * bs < j1; j += 4 * bs) {
    int m[4];
    double pw[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) m[u] = sh_m[j + u * bs];
#pragma unroll
    for (int u = 0; u < 4; ++u) pw[u] = (double)prim[m[u]].w;
#pragma unroll
    for (int u = 0; u < 4; ++u) point(j + u * bs, pw[u]);
  }
  for (; j < j1; j += bs) point(j, (double)prim[sh_m[j]].w);
  H = block_sum(H);
  __syncthreads();
  V = block_sum(V);
  __syncthreads();
  M = block_sum(M);
  if (threadIdx.x == 0) {
    const double cl = V * ct - H * st, cd = H * ct + V * st;
    clcdcm[3 * s] = cl;
    clcdcm[3 * s + 1] = cd;
    clcdcm[3 * s + 2] = M;
    if (ctl && cd_hist) cd_hist[(size_t)ctl->it * nsid + s] = cd;
  }
}


// ---------------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------------

// What a launcher needs of a context
struct LaunchEnv {
  hipStream_t stream;
  Geom g;
  int n;
  bool power0;
  double power, cfl, euler;
  Freestream fs;
};

inline dim3 lane_grid(int n, int lpp) { return dim3((unsigned)(((int64_t)n * lpp + BLOCK - 1) / BLOCK)); }

// LPP dispatch: f.template operator()<L>() for the runtime lanes-per-point value
template <class F>
inline void with_lpp(int lpp, F&& f) {
  switch (lpp) {
    case 2: f.template operator()<2>(); break;
    case 4: f.template operator()<4>(); break;
    case 8: f.template operator()<8>(); break;
    default: f.template operator()<1>(); break;
  }
}

// Staged shapes: 1, 2 or 4 lanes per point (a block of STAGE_P points is at most BLOCK
// threads)
template <class F>
inline void with_lpp_staged(int lpp, F&& f) {
  switch (lpp) {
    case 2: f.template operator()<2>(); break;
    case 4: f.template operator()<4>(); break;
    default: f.template operator()<1>(); break;
  }
}

inline dim3 stage_grid(int n) { return dim3((unsigned)((n + STAGE_P - 1) / STAGE_P)); }

// eval_q_derivatives (+ func_delta) of q into dq, qm (and delta).  lpp: lanes per point,
// with LPP_STAGED set for the staged walk (the caller has checked that it fits in LDS).
template <class S>
void launch_lsq_env(const LaunchEnv& e, const V4<S>* q, V4<S>* dq, V4<S>* qm, const double* cspeed,
                    double* delta, bool with_delta, const Ctl* ctl, int lpp, V4<float>* qsum) {
  const double cfl = e.cfl, pw = e.power;
  if (lpp & LPP_STAGED) {
    const dim3 gr = stage_grid(e.n);
    const size_t lds = stage_lds_lsq(e.g.hl[0].hcap, e.g.width, sizeof(S), with_delta);
    with_lpp_staged(lpp & ~LPP_STAGED, [&]<int L>() {
      const dim3 bl(STAGE_P * L);
      if (e.power0 && with_delta)
        hipLaunchKernelGGL((k_lsq<S, true, true, L, true>), gr, bl, lds, e.stream, e.g, q, dq, qm, cspeed, delta, cfl, pw, ctl, qsum);
      else if (e.power0)
        hipLaunchKernelGGL((k_lsq<S, true, false, L, true>), gr, bl, lds, e.stream, e.g, q, dq, qm, cspeed, delta, cfl, pw, ctl, qsum);
      else if (with_delta)
        hipLaunchKernelGGL((k_lsq<S, false, true, L, true>), gr, bl, lds, e.stream, e.g, q, dq, qm, cspeed, delta, cfl, pw, ctl, qsum);
      else
        hipLaunchKernelGGL((k_lsq<S, false, false, L, true>), gr, bl, lds, e.stream, e.g, q, dq, qm, cspeed, delta, cfl, pw, ctl, qsum);
    });
    return;
  }
  const dim3 gr = lane_grid(e.n, lpp), bl(BLOCK);
  with_lpp(lpp, [&]<int L>() {
    if (e.power0 && with_delta)
      hipLaunchKernelGGL((k_lsq<S, true, true, L, false>), gr, bl, 0, e.stream, e.g, q, dq, qm, cspeed, delta, cfl, pw, ctl, qsum);
    else if (e.power0)
      hipLaunchKernelGGL((k_lsq<S, true, false, L, false>), gr, bl, 0, e.stream, e.g, q, dq, qm, cspeed, delta, cfl, pw, ctl, qsum);
    else if (with_delta)
      hipLaunchKernelGGL((k_lsq<S, false, true, L, false>), gr, bl, 0, e.stream, e.g, q, dq, qm, cspeed, delta, cfl, pw, ctl, qsum);
    else
      hipLaunchKernelGGL((k_lsq<S, false, false, L, false>), gr, bl, 0, e.stream, e.g, q, dq, qm, cspeed, delta, cfl, pw, ctl, qsum);
  });
}

// one inner iteration: dq_in (a) -> dq_out (b); the mixed path gathers only dq (sums
// form, first-pass sums from the LSQ kernel).  lpp as for launch_lsq_env.
template <class S>
void launch_inner_env(const LaunchEnv& e, const V4<S>* q, const V4<S>* a, V4<S>* b, const Ctl* ctl, int lpp,
                      const V4<float>* qsum) {
  const double pw = e.power;
  constexpr bool SUMS = std::is_same<S, float>::value;
  if (lpp & LPP_STAGED) {
    const dim3 gr = stage_grid(e.n);
    const size_t lds = stage_lds_inner(e.g.hl[0].hcap, e.g.width, sizeof(S), SUMS);
    with_lpp_staged(lpp & ~LPP_STAGED, [&]<int L>() {
      const dim3 bl(STAGE_P * L);
      if (e.power0) hipLaunchKernelGGL((k_inner<S, true, L, SUMS, true>), gr, bl, lds, e.stream, e.g, q, a, b, pw, ctl, qsum);
      else hipLaunchKernelGGL((k_inner<S, false, L, SUMS, true>), gr, bl, lds, e.stream, e.g, q, a, b, pw, ctl, qsum);
    });
    return;
  }
  const dim3 gr = lane_grid(e.n, lpp), bl(BLOCK);
  with_lpp(lpp, [&]<int L>() {
    if (e.power0) hipLaunchKernelGGL((k_inner<S, true, L, SUMS, false>), gr, bl, 0, e.stream, e.g, q, a, b, pw, ctl, qsum);
    else hipLaunchKernelGGL((k_inner<S, false, L, SUMS, false>), gr, bl, 0, e.stream, e.g, q, a, b, pw, ctl, qsum);
  });
}

// fused flux residual + state update (rk) + q_variables of the new state.  lpp: lanes
// per point, with LPP_STAGED set for the staged walk (2, 4 or 8 lanes, blocks of
// STAGE_PF points)
template <class S>
void launch_flux_update_env(const LaunchEnv& e, const State<S>& st, const V4<S>* prim, V4<S>* pold, V4<S>* pout,
                            V4<S>* qout, double* cspeed, const double* delta, double* partial, const Ctl* ctl,
                            int rk, bool last, int lpp, double* rs_pt) {
  const int lst = last ? 1 : 0;
  const double eu = e.euler, pw = e.power;
  const Freestream fs = e.fs;
  if (lpp & LPP_STAGED) {
    const dim3 gr((unsigned)((e.n + STAGE_PF - 1) / STAGE_PF));
    const size_t lds = stage_lds_flux(e.g.hl[1].hcap, e.g.width, sizeof(S));
    auto go = [&]<int L>() {
      const dim3 bl(STAGE_PF * L);
      if (e.power0)
        hipLaunchKernelGGL((k_flux_update<S, true, L, true>), gr, bl, lds, e.stream, e.g, st, prim, pold, pout, qout,
                           cspeed, delta, partial, ctl, rk, lst, eu, pw, fs, rs_pt);
      else
        hipLaunchKernelGGL((k_flux_update<S, false, L, true>), gr, bl, lds, e.stream, e.g, st, prim, pold, pout,
                           qout, cspeed, delta, partial, ctl, rk, lst, eu, pw, fs, rs_pt);
    };
    switch (lpp & ~LPP_STAGED) {
      case 2: go.template operator()<2>(); break;
      case 4: go.template operator()<4>(); break;
      default: go.template operator()<8>(); break;
    }
    return;
  }
  const dim3 gr = lane_grid(e.n, lpp), bl(BLOCK);
  with_lpp(lpp, [&]<int L>() {
    if (e.power0)
      hipLaunchKernelGGL((k_flux_update<S, true, L, false>), gr, bl, 0, e.stream, e.g, st, prim, pold, pout, qout,
                         cspeed, delta, partial, ctl, rk, lst, eu, pw, fs, rs_pt);
    else
      hipLaunchKernelGGL((k_flux_update<S, false, L, false>), gr, bl, 0, e.stream, e.g, st, prim, pold, pout, qout,
                         cspeed, delta, partial, ctl, rk, lst, eu, pw, fs, rs_pt);
  });
}

// Residual partials one flux+update launch writes (one per block)
inline int flux_blocks(int n, int lpp) {
  return (lpp & LPP_STAGED) ? (n + STAGE_PF - 1) / STAGE_PF : (int)lane_grid(n, lpp).x;
}

// Cl, Cd, Cm (and cp) per shape id
struct ForcesArgs {
  const int32_t *sh_m, *sh_pos;
  const double* sgeo;
  const int32_t* sh_ptr;
  const double* shape_par;
  int nsid;
  double *cp, *clcdcm;
};

template <class S>
void launch_forces_env(const LaunchEnv& e, const V4<S>* prim, const ForcesArgs& f, double* cd_hist,
                       const Ctl* ctl) {
  hipLaunchKernelGGL((k_forces<S>), dim3(f.nsid), dim3(BLOCK), 0, e.stream, prim, f.sh_m, f.sh_pos, f.sgeo,
                     f.sh_ptr, f.shape_par, f.nsid, f.cp, f.clcdcm, cd_hist, ctl);
}

template <class S>
void launch_q_variables_env(const LaunchEnv& e, const V4<S>* prim, V4<S>* q, double* cspeed) {
  hipLaunchKernelGGL((k_q_variables<S>), dim3((e.n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, e.stream, e.n, prim,
                     q, cspeed);
}

template <class S>
void launch_func_delta_env(const LaunchEnv& e, const V4<S>* prim, double* delta) {
  hipLaunchKernelGGL((k_func_delta<S>), dim3((e.n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, e.stream, e.g, prim,
                     delta, e.cfl);
}

template <class S>
void launch_flux_env(const LaunchEnv& e, const State<S>& st, const double* delta, V4d* flux_res) {
  const dim3 gr((e.n + BLOCK - 1) / BLOCK), bl(BLOCK);
  if (e.power0) hipLaunchKernelGGL((k_flux<S, true>), gr, bl, 0, e.stream, e.g, st, delta, flux_res, e.power);
  else hipLaunchKernelGGL((k_flux<S, false>), gr, bl, 0, e.stream, e.g, st, delta, flux_res, e.power);
}

template <class S>
void launch_update_env(const LaunchEnv& e, V4<S>* prim, const V4<S>* pold, const V4d* flux_res, double* partial,
                       int rk) {
  hipLaunchKernelGGL((k_update<S>), dim3((e.n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, e.stream, e.g, prim, pold,
                     flux_res, partial, rk, e.euler, e.fs);
}

// The mixed path's launchers, defined in mfcfd_f32.hip (launch_lsq_f32 in mfcfd_lsq32.hip)
using V4f = V4<float>;
void launch_lsq_f32(const LaunchEnv& e, const V4f* q, V4f* dq, V4f* qm, const double* cspeed, double* delta,
                    bool with_delta, const Ctl* ctl, int lpp, V4f* qsum);
void launch_inner_f32(const LaunchEnv& e, const V4f* q, const V4f* a, V4f* b, const Ctl* ctl, int lpp,
                      const V4f* qsum);
void launch_flux_update_f32(const LaunchEnv& e, const State<float>& st, const V4f* prim, V4f* pold, V4f* pout,
                            V4f* qout, double* cspeed, const double* delta, double* partial, const Ctl* ctl,
                            int rk, bool last, int lpp, double* rs_pt);
// the resident inner sweeps (lskum_sweeps.h): launch, and how many blocks fit at once
void launch_sweeps_f32(const LaunchEnv& e, V4f* dq0, V4f* dq1, int sweeps, Ctl* ctl, const V4f* qsum, SweepBar* bar,
                       uint32_t spin_limit, int lpp);
int sweeps_capacity_f32(int lpp, bool power0, size_t lds);
void launch_forces_f32(const LaunchEnv& e, const V4f* prim, const ForcesArgs& f, double* cd_hist, const Ctl* ctl);
void launch_q_variables_f32(const LaunchEnv& e, const V4f* prim, V4f* q, double* cspeed);
void launch_func_delta_f32(const LaunchEnv& e, const V4f* prim, double* delta);
void launch_flux_f32(const LaunchEnv& e, const State<float>& st, const double* delta, V4d* flux_res);
void launch_update_f32(const LaunchEnv& e, V4f* prim, const V4f* pold, const V4d* flux_res, double* partial,
                       int rk);

// Scalar-type dispatch: float to mfcfd_f32.hip, double instantiated by the caller's unit
template <class S>
inline constexpr bool is_f32 = std::is_same<S, float>::value;

template <class S>
inline void launch_lsq_any(const LaunchEnv& e, const V4<S>* q, V4<S>* dq, V4<S>* qm, const double* cspeed,
                           double* delta, bool with_delta, const Ctl* ctl, int lpp, V4<float>* qsum) {
  if constexpr (is_f32<S>) launch_lsq_f32(e, q, dq, qm, cspeed, delta, with_delta, ctl, lpp, qsum);
  else launch_lsq_env<S>(e, q, dq, qm, cspeed, delta, with_delta, ctl, lpp, qsum);
}
template <class S>
inline void launch_inner_any(const LaunchEnv& e, const V4<S>* q, const V4<S>* a, V4<S>* b, const Ctl* ctl, int lpp,
                             const V4<float>* qsum) {
  if constexpr (is_f32<S>) launch_inner_f32(e, q, a, b, ctl, lpp, qsum);
  else launch_inner_env<S>(e, q, a, b, ctl, lpp, qsum);
}
template <class S>
inline void launch_flux_update_any(const LaunchEnv& e, const State<S>& st, const V4<S>* prim, V4<S>* pold,
                                   V4<S>* pout, V4<S>* qout, double* cspeed, const double* delta, double* partial,
                                   const Ctl* ctl, int rk, bool last, int lpp, double* rs_pt) {
  if constexpr (is_f32<S>)
    launch_flux_update_f32(e, st, prim, pold, pout, qout, cspeed, delta, partial, ctl, rk, last, lpp, rs_pt);
  else
    launch_flux_update_env<S>(e, st, prim, pold, pout, qout, cspeed, delta, partial, ctl, rk, last, lpp, rs_pt);
}
template <class S>
inline void launch_forces_any(const LaunchEnv& e, const V4<S>* prim, const ForcesArgs& f, double* cd_hist,
                              const Ctl* ctl) {
  if constexpr (is_f32<S>) launch_forces_f32(e, prim, f, cd_hist, ctl);
  else launch_forces_env<S>(e, prim, f, cd_hist, ctl);
}
template <class S>
inline void launch_q_variables_any(const LaunchEnv& e, const V4<S>* prim, V4<S>* q, double* cspeed) {
  if constexpr (is_f32<S>) launch_q_variables_f32(e, prim, q, cspeed);
  else launch_q_variables_env<S>(e, prim, q, cspeed);
}
template <class S>
inline void launch_func_delta_any(const LaunchEnv& e, const V4<S>* prim, double* delta) {
  if constexpr (is_f32<S>) launch_func_delta_f32(e, prim, delta);
  else launch_func_delta_env<S>(e, prim, delta);
}
template <class S>
inline void launch_flux_any(const LaunchEnv& e, const State<S>& st, const double* delta, V4d* flux_res) {
  if constexpr (is_f32<S>) launch_flux_f32(e, st, delta, flux_res);
  else launch_flux_env<S>(e, st, delta, flux_res);
}
template <class S>
inline void launch_update_any(const LaunchEnv& e, V4<S>* prim, const V4<S>* pold, const V4d* flux_res,
                              double* partial, int rk) {
  if constexpr (is_f32<S>) launch_update_f32(e, prim, pold, flux_res, partial, rk);
  else launch_update_env<S>(e, prim, pold, flux_res, partial, rk);
}
