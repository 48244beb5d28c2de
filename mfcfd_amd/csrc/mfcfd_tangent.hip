// Tangent (forward) mode of the LSKUM iteration on the device: mf_tangent.
//
// Replaces the Tapenade forward mode of src_mpi_forward (q_lskum_d.F90:28-60 seeds
// pointd%x/y, fpi_solver_d.F90:27-85 carries the tangents through every routine of an
// iteration, objective_function_d the cost's).  Every quantity is a dual number Dd
// (value, directional derivative along the seeded coordinate direction (vx, vy)); the
// kernels follow the reference routines, with the device header's generic point and
// edge routines (reconstruction and limiter, q-tilde -> primitive, the kinetic split
// fluxes, the conserved states) instantiated for Dd.  The stencil kernels (inner
// iteration, flux residual) split each point's edges over lanes, and the flux residual
// takes the primal kernels' linear form (each edge's two ends evaluated once, its flux
// differences times its half-stencils' solve coefficients, here with their tangents).
// The stencil split is the context's (fixed, as the reference's generate_connectivity
// is not differentiated); min_dist and so the limiter's epsilon are constants.
//
// Included by mfcfd.hip (it uses the context).  FP64 arithmetic throughout; a mixed
// context's state is widened at the start.

namespace mf {

struct Dd {
  double v, d;
  __host__ __device__ constexpr Dd(double a = 0.0, double b = 0.0) : v(a), d(b) {}
};
__host__ __device__ inline Dd operator+(Dd a, Dd b) { return Dd(a.v + b.v, a.d + b.d); }
__host__ __device__ inline Dd operator-(Dd a, Dd b) { return Dd(a.v - b.v, a.d - b.d); }
__host__ __device__ inline Dd operator-(Dd a) { return Dd(-a.v, -a.d); }
__host__ __device__ inline Dd operator*(Dd a, Dd b) { return Dd(a.v * b.v, a.d * b.v + a.v * b.d); }
__host__ __device__ inline Dd operator/(Dd a, Dd b) {
  const double q = a.v / b.v;
  return Dd(q, (a.d - q * b.d) / b.v);
}
__host__ __device__ inline Dd operator+(Dd a, double b) { return Dd(a.v + b, a.d); }
__host__ __device__ inline Dd operator+(double a, Dd b) { return Dd(a + b.v, b.d); }
__host__ __device__ inline Dd operator-(Dd a, double b) { return Dd(a.v - b, a.d); }
__host__ __device__ inline Dd operator-(double a, Dd b) { return Dd(a - b.v, -b.d); }
__host__ __device__ inline Dd operator*(Dd a, double b) { return Dd(a.v * b, a.d * b); }
__host__ __device__ inline Dd operator*(double a, Dd b) { return Dd(a * b.v, a * b.d); }
__host__ __device__ inline Dd operator/(Dd a, double b) { return Dd(a.v / b, a.d / b); }
__host__ __device__ inline Dd operator/(double a, Dd b) {
  const double q = a / b.v;
  return Dd(q, -q * b.d / b.v);
}
// branches follow the values, as the reference's differentiated code does
__host__ __device__ inline bool operator<(Dd a, Dd b) { return a.v < b.v; }
__host__ __device__ inline bool operator>(Dd a, Dd b) { return a.v > b.v; }
__host__ __device__ inline bool operator<=(Dd a, Dd b) { return a.v <= b.v; }
__host__ __device__ inline bool operator>=(Dd a, Dd b) { return a.v >= b.v; }

__host__ __device__ inline Dd m_exp(Dd a) {
  const double e = exp(a.v);
  return Dd(e, e * a.d);
}
__host__ __device__ inline Dd m_log(Dd a) { return Dd(log(a.v), a.d / a.v); }
__host__ __device__ inline Dd m_sqrt(Dd a) {
  const double s = sqrt(a.v);
  return Dd(s, a.d / (2.0 * s));
}
// d erf(s)/ds = 2/sqrt(pi) exp(-s^2) (the closed form the reference's Tapenade code
// was patched to use, split_fluxes_b.F90:82-83)
__host__ __device__ inline Dd m_erf(Dd a) {
  return Dd(erf(a.v), 1.1283791670955126 * exp(-a.v * a.v) * a.d);
}
__host__ __device__ inline Dd m_abs(Dd a) { return a.v < 0.0 ? -a : a; }
__host__ __device__ inline Dd m_pow(Dd a, double p) {
  if (p == 0.0) return Dd(pow(a.v, p), 0.0);
  return Dd(pow(a.v, p), p * pow(a.v, p - 1.0) * a.d);
}

template <> struct Consts<Dd> {
  static constexpr double PI = 3.141592653589793;
  static constexpr double THRESH = (double)10e-6f;
  static constexpr double GAMMA_M1 = 1.4 - 1.0;
};

using V4t = V4<Dd>;
struct alignas(32) D2t {
  Dd x, y;
};

struct TanArgs {
  const D2t* xy;
  const D2t* nrm;
  const int32_t* nbhs;
  const uint32_t* conn;  // (neighbour << 4) | half-stencil bits (xpos 1, xneg 2, ypos 4, yneg 8)
  const double* eps;     // (VL_CONST min_dist)^3
  int n, nw, no;
  double power;
  int first_order;
};

__device__ inline int tan_point() { return (int)(blockIdx.x * blockDim.x + threadIdx.x); }

// point_normals.F90:8-105 for the wall and outer points (their left/right neighbours)
__global__ __launch_bounds__(BLOCK) void k_tn_normals(int nb, const int32_t* __restrict__ bm,
                                                      const int32_t* __restrict__ bl,
                                                      const int32_t* __restrict__ br, const D2t* __restrict__ xy,
                                                      D2t* __restrict__ nrm) {
  const int j = tan_point();
  if (j >= nb) return;
  const int m = bm[j], l = bl[j], r = br[j];
  const Dd nx1 = xy[m].y - xy[l].y, nx2 = xy[r].y - xy[m].y;
  const Dd ny1 = xy[m].x - xy[l].x, ny2 = xy[r].x - xy[m].x;
  const Dd vx = 0.5 * (nx1 + nx2), vy = 0.5 * (ny1 + ny2);
  const Dd det = m_sqrt(vx * vx + vy * vy);
  nrm[m] = D2t{-vx / det, vy / det};
}

template <class S>
__global__ __launch_bounds__(BLOCK) void k_tn_load(int n, const V4<S>* __restrict__ p, V4t* __restrict__ out) {
  const int i = tan_point();
  if (i >= n) return;
  const V4<S> v = p[i];
  out[i] = V4t{Dd((double)v.x), Dd((double)v.y), Dd((double)v.z), Dd((double)v.w)};
}

__global__ __launch_bounds__(BLOCK) void k_tn_copy(int n, const V4t* __restrict__ a, V4t* __restrict__ b) {
  const int i = tan_point();
  if (i < n) b[i] = a[i];
}

// q_variables.F90:8-37
__global__ __launch_bounds__(BLOCK) void k_tn_qvar(int n, const V4t* __restrict__ prim, V4t* __restrict__ q) {
  const int i = tan_point();
  if (i < n) q[i] = q_of_prim<Dd>(prim[i]);
}

// state_update.F90:190-236 func_delta: the smallest neighbour time step, capped at 1
__global__ __launch_bounds__(BLOCK) void k_tn_delta(TanArgs a, const V4t* __restrict__ prim, double cfl,
                                                    Dd* __restrict__ delta) {
  const int i = tan_point();
  if (i >= a.n) return;
  Dd min_delt(1.0);
  const D2t pi = a.xy[i];
  for (int s = 0; s < a.nbhs[i]; ++s) {
    const int k = (int)(a.conn[(size_t)s * a.n + i] >> 4);
    const V4t p = prim[k];
    const Dd dx = a.xy[k].x - pi.x, dy = a.xy[k].y - pi.y;
    const Dd dist = m_sqrt(dx * dx + dy * dy);
    const Dd mod_u = m_sqrt(p.y * p.y + p.z * p.z);
    Dd dt = dist / (mod_u + 3.0 * m_sqrt(p.w / p.x));
    dt = cfl * dt;
    if (min_delt > dt) min_delt = dt;
  }
  delta[i] = min_delt;
}

// LSQ weight dist^power of an edge (q_variables.F90:89-91); POWER0: power = 0, weight 1
// with no pow evaluation (the library's pow is long and register-hungry)
template <bool POWER0>
__device__ inline Dd tn_weight(const Dd& r2, double power) {
  if constexpr (POWER0) return Dd(1.0);
  else return m_pow(m_sqrt(r2), power);
}

// sum of a dual number over a point's LPP lanes
template <int LPP>
__device__ inline Dd group_sum_d(Dd v) {
  return Dd(group_sum<LPP>(v.v), group_sum<LPP>(v.d));
}

// q_variables.F90:41-112 eval_q_derivatives: qm over {i} + stencil (strict > and <),
// distance-weighted least squares, Cramer's rule; LPP lanes per point, their extremes
// combined as the sequential walk selects them (the larger value; on ties the lower
// slot, the point itself first), so each qm carries the selected point's tangent
template <int LPP, bool POWER0>
__global__ __launch_bounds__(BLOCK) void k_tn_lsq(TanArgs a, const V4t* __restrict__ q, V4t* __restrict__ dq,
                                                  V4t* __restrict__ qm) {
  const int t = tan_point();
  const int i = t / LPP, lane = t % LPP;
  if (i >= a.n) return;  // whole lane groups
  const D2t pi = a.xy[i];
  const V4t qi = q[i];
  V4t mx = qi, mn = qi;
  int smx[4] = {-1, -1, -1, -1}, smn[4] = {-1, -1, -1, -1};
  Dd sxx, syy, sxy, sxq[4], syq[4];
#pragma unroll 1
  for (int s = lane; s < a.nbhs[i]; s += LPP) {
    const int k = (int)(a.conn[(size_t)s * a.n + i] >> 4);
    const V4t qk = q[k];
    for (int r = 0; r < 4; ++r) {
      if (qk[r] > mx[r]) {
        mx[r] = qk[r];
        smx[r] = s;
      }
      if (qk[r] < mn[r]) {
        mn[r] = qk[r];
        smn[r] = s;
      }
    }
    const Dd delx = a.xy[k].x - pi.x, dely = a.xy[k].y - pi.y;
    const Dd w = tn_weight<POWER0>(delx * delx + dely * dely, a.power);
    sxx = sxx + delx * delx * w;
    syy = syy + dely * dely * w;
    sxy = sxy + delx * dely * w;
    for (int r = 0; r < 4; ++r) {
      sxq[r] = sxq[r] + w * delx * (qk[r] - qi[r]);
      syq[r] = syq[r] + w * dely * (qk[r] - qi[r]);
    }
  }
  if (LPP > 1) {
    sxx = group_sum_d<LPP>(sxx);
    syy = group_sum_d<LPP>(syy);
    sxy = group_sum_d<LPP>(sxy);
    for (int r = 0; r < 4; ++r) {
      sxq[r] = group_sum_d<LPP>(sxq[r]);
      syq[r] = group_sum_d<LPP>(syq[r]);
      for (int off = LPP / 2; off > 0; off >>= 1) {
        const Dd ov(__shfl_xor(mx[r].v, off, 64), __shfl_xor(mx[r].d, off, 64));
        const int os = __shfl_xor(smx[r], off, 64);
        if (ov.v > mx[r].v || (ov.v == mx[r].v && os < smx[r])) {
          mx[r] = ov;
          smx[r] = os;
        }
        const Dd nv(__shfl_xor(mn[r].v, off, 64), __shfl_xor(mn[r].d, off, 64));
        const int ns = __shfl_xor(smn[r], off, 64);
        if (nv.v < mn[r].v || (nv.v == mn[r].v && ns < smn[r])) {
          mn[r] = nv;
          smn[r] = ns;
        }
      }
    }
    if (lane != 0) return;
  }
  const Dd obd = 1.0 / (sxx * syy - sxy * sxy);
  V4t tx, ty;
  for (int r = 0; r < 4; ++r) {
    tx[r] = (sxq[r] * syy - syq[r] * sxy) * obd;
    ty[r] = (syq[r] * sxx - sxq[r] * sxy) * obd;
  }
  dq[2 * i] = tx;
  dq[2 * i + 1] = ty;
  qm[2 * i] = mx;
  qm[2 * i + 1] = mn;
}

// q_variables.F90:115-227: one Jacobi sweep of the defect correction, din -> dout,
// LPP lanes per point (their sums combined before the solve)
template <int LPP, bool POWER0>
__global__ __launch_bounds__(BLOCK) void k_tn_inner(TanArgs a, const V4t* __restrict__ q,
                                                    const V4t* __restrict__ din, V4t* __restrict__ dout) {
  const int t = tan_point();
  const int i = t / LPP, lane = t % LPP;
  if (i >= a.n) return;  // whole lane groups
  const D2t pi = a.xy[i];
  const V4t qi = q[i], dxi = din[2 * i], dyi = din[2 * i + 1];
  Dd sxx, syy, sxy, sxq[4], syq[4];
  // one edge at a time: its neighbour records (24 doubles) live only for that edge
#pragma unroll 1
  for (int s = lane; s < a.nbhs[i]; s += LPP) {
    const int k = (int)(a.conn[(size_t)s * a.n + i] >> 4);
    const V4t qk = q[k], dxk = din[2 * k], dyk = din[2 * k + 1];
    const Dd delx = a.xy[k].x - pi.x, dely = a.xy[k].y - pi.y;
    const Dd w = tn_weight<POWER0>(delx * delx + dely * dely, a.power);
    sxx = sxx + delx * delx * w;
    syy = syy + dely * dely * w;
    sxy = sxy + delx * dely * w;
    for (int r = 0; r < 4; ++r) {
      const Dd t1 = qi[r] - 0.5 * (delx * dxi[r] + dely * dyi[r]);
      const Dd t2 = qk[r] - 0.5 * (delx * dxk[r] + dely * dyk[r]);
      sxq[r] = sxq[r] + (w * delx * (t2 - t1));
      syq[r] = syq[r] + (w * dely * (t2 - t1));
    }
  }
  if (LPP > 1) {
    sxx = group_sum_d<LPP>(sxx);
    syy = group_sum_d<LPP>(syy);
    sxy = group_sum_d<LPP>(sxy);
    for (int r = 0; r < 4; ++r) {
      sxq[r] = group_sum_d<LPP>(sxq[r]);
      syq[r] = group_sum_d<LPP>(syq[r]);
    }
    if (lane != 0) return;
  }
  const Dd obd = 1.0 / (sxx * syy - sxy * sxy);
  V4t tx, ty;
  for (int r = 0; r < 4; ++r) {
    tx[r] = obd * (sxq[r] * syy - syq[r] * sxy);
    ty[r] = obd * (syq[r] * sxx - sxq[r] * sxy);
  }
  dout[2 * i] = tx;
  dout[2 * i + 1] = ty;
}

// Flux residual of one point (flux_residual.F90:14-73 over wall_fluxes.F90:20-292,
// outer_fluxes.F90:18-300, interior_fluxes.F90:19-404), this lane's share of its edges.
// The reference solves each half-stencil's weighted 2x2 least squares after summing
// its edges; the solve is linear in the flux sums, so, as in the primal kernels, each
// edge adds its end-flux difference times its own coefficient, here carrying the
// coefficients' tangents: a first walk sums each half's geometry, then each edge's two
// ends are evaluated once (limited reconstruction, q-tilde -> primitive) with the x and
// y split fluxes of the halves the neighbour is in (both halves on a tie).
template <int CLS, int LPP, bool POWER0>
__device__ inline void tn_flux_point(const TanArgs& a, int i, int lane, const V4t* __restrict__ q,
                                     const V4t* __restrict__ dq, const V4t* __restrict__ qm, Dd acc[4]) {
  const D2t pi = a.xy[i], ni = a.nrm[i];
  const Dd nx = ni.x, ny = ni.y, tx = ny, ty = -nx;
  const int nb = a.nbhs[i];
  auto offset = [&](int k, Dd& delx, Dd& dely, Dd& dels, Dd& deln, Dd& dsw, Dd& dnw) {
    delx = a.xy[k].x - pi.x;
    dely = a.xy[k].y - pi.y;
    dels = delx * tx + dely * ty;
    deln = delx * nx + dely * ny;
    const Dd w = tn_weight<POWER0>(dels * dels + deln * deln, a.power);
    dsw = dels * w;
    dnw = deln * w;
  };
  // per half (xpos, xneg, ypos, yneg): sum ds^2 w, dn^2 w, ds dn w
  Dd hxx[4], hyy[4], hxy[4];
  for (int s = lane; s < nb; s += LPP) {
    const uint32_t word = a.conn[(size_t)s * a.n + i];
    Dd delx, dely, dels, deln, dsw, dnw;
    offset((int)(word >> 4), delx, dely, dels, deln, dsw, dnw);
    for (int d = 0; d < 4; ++d) {
      if (!(word & (1u << d))) continue;
      hxx[d] = hxx[d] + dels * dsw;
      hyy[d] = hyy[d] + deln * dnw;
      hxy[d] = hxy[d] + dels * dnw;
    }
  }
  Dd c1[4], obd[4];
  for (int d = 0; d < 4; ++d) {
    if (LPP > 1) {
      hxx[d] = group_sum_d<LPP>(hxx[d]);
      hyy[d] = group_sum_d<LPP>(hyy[d]);
      hxy[d] = group_sum_d<LPP>(hxy[d]);
    }
    c1[d] = d < 2 ? hyy[d] : hxx[d];
    obd[d] = 1.0 / (hxx[d] * hyy[d] - hxy[d] * hxy[d]);
  }
  // the t-derivative coefficient of x half 0 or 1 (first: xpos) and the n-derivative
  // coefficient of y half 2 or 3 (first: ypos); selects, not a runtime array index, so
  // the half sums stay in registers
  auto coef_x = [&](bool first, const Dd& dsw, const Dd& dnw) {
    return (dsw * (first ? c1[0] : c1[1]) - dnw * (first ? hxy[0] : hxy[1])) * (first ? obd[0] : obd[1]);
  };
  auto coef_y = [&](bool first, const Dd& dsw, const Dd& dnw) {
    return (dnw * (first ? c1[2] : c1[3]) - dsw * (first ? hxy[2] : hxy[3])) * (first ? obd[2] : obd[3]);
  };
  const V4t qi = q[i], dxi = dq[2 * i], dyi = dq[2 * i + 1], mxi = qm[2 * i], mni = qm[2 * i + 1];
  const Dd ei(a.eps[i]);
  for (int s = lane; s < nb; s += LPP) {
    const uint32_t word = a.conn[(size_t)s * a.n + i];
    const int k = (int)(word >> 4);
    const unsigned m = word & 15u;
    Dd delx, dely, dels, deln, dsw, dnw;
    offset(k, delx, dely, dels, deln, dsw, dnw);
    const bool xp = (m & XPOS) != 0u;
    const bool tie_x = xp && (m & XNEG) != 0u;
    // interior: y half from the mask; wall: yneg (every neighbour); outer: ypos
    const bool yp = (CLS == INTERIOR) ? (m & YPOS) != 0u : (CLS == OUTER);
    const bool tie_y = (CLS == INTERIOR) && yp && (m & YNEG) != 0u;
    const Dd cx = coef_x(xp, dsw, dnw), cy = coef_y(yp, dsw, dnw);
    const Dd cx2 = tie_x ? coef_x(false, dsw, dnw) : Dd(), cy2 = tie_y ? coef_y(false, dsw, dnw) : Dd();
    // the ends one after the other (own: -, neighbour: +)
#pragma unroll 1
    for (int e = 0; e < 2; ++e) {
      Dd qt[4];
      if (a.first_order) {
        const V4t qe = e ? q[k] : qi;
        for (int r = 0; r < 4; ++r) qt[r] = qe[r];
      } else if (e) {
        reconstruct<Dd>(q[k], dq[2 * k], dq[2 * k + 1], qm[2 * k], qm[2 * k + 1], Dd(a.eps[k]), delx, dely, qt);
      } else {
        reconstruct<Dd>(qi, dxi, dyi, mxi, mni, ei, delx, dely, qt);
      }
      const Kin<Dd> kn = kin_of_qtilde<Dd>(qt, nx, ny);
      const double sg = e ? 1.0 : -1.0;
      Dd Gx[4], Gy[4];
      end_fluxes<CLS, Dd>(kn, xp, yp, Gx, Gy);
      for (int r = 0; r < 4; ++r) acc[r] = acc[r] + sg * (Gx[r] * cx + Gy[r] * cy);
      if (tie_x || tie_y) {
        // the neighbour also belongs to the xneg (yneg) half, whose flux has the
        // opposite split
        end_fluxes<CLS, Dd>(kn, false, false, Gx, Gy);
        for (int r = 0; r < 4; ++r) acc[r] = acc[r] + sg * (Gx[r] * cx2 + Gy[r] * cy2);
      }
    }
  }
  if (LPP > 1)
    for (int r = 0; r < 4; ++r) acc[r] = group_sum_d<LPP>(acc[r]);
}

// flux_residual.F90:14-73 then state_update.F90:10-141 (stage rk) of every point, LPP
// lanes per point; the new prim replaces the old in place (no point reads another's
// prim here)
template <int LPP, bool POWER0>
__global__ __launch_bounds__(BLOCK) void k_tn_flux_update(TanArgs a, const V4t* __restrict__ q,
                                                          const V4t* __restrict__ dq, const V4t* __restrict__ qm,
                                                          const Dd* __restrict__ delta, V4t* __restrict__ prim,
                                                          const V4t* __restrict__ pold, int rk, double euler,
                                                          Freestream fs) {
  const int t = tan_point();
  const int i = t / LPP, lane = t % LPP;
  if (i >= a.n) return;  // whole lane groups
  const int cls = i < a.nw ? WALL : (i < a.nw + a.no ? OUTER : INTERIOR);
  Dd acc[4];
  if (cls == WALL) tn_flux_point<WALL, LPP, POWER0>(a, i, lane, q, dq, qm, acc);
  else if (cls == OUTER) tn_flux_point<OUTER, LPP, POWER0>(a, i, lane, q, dq, qm, acc);
  else tn_flux_point<INTERIOR, LPP, POWER0>(a, i, lane, q, dq, qm, acc);
  if (lane != 0) return;
  // flux_residual.F90:36 (wall: 2 delta), :50 and :68 (delta)
  const Dd scale = cls == WALL ? 2.0 * delta[i] : delta[i];
  Dd fr[4];
  for (int r = 0; r < 4; ++r) fr[r] = scale * acc[r];
  const Dd nx = a.nrm[i].x, ny = a.nrm[i].y;
  Dd U[4], Uo[4];
  if (cls == OUTER) {
    cons_ubar<Dd>(prim[i], nx, ny, fs, U);
    if (rk == 3) cons_ubar<Dd>(pold[i], nx, ny, fs, Uo);
  } else {
    prim_to_cons<Dd>(prim[i], nx, ny, U);
    if (rk == 3) prim_to_cons<Dd>(pold[i], nx, ny, Uo);
  }
  const double obt = 1.0 / 3.0, tbt = 2.0 / 3.0;
  for (int r = 0; r < 4; ++r) {
    if (rk != 3) U[r] = U[r] - ((0.5 * euler) * fr[r]);
    else U[r] = tbt * Uo[r] + obt * (U[r] - 0.5 * fr[r]);
  }
  if (cls == WALL) U[2] = Dd(0.0);
  const Dd u2r = U[1], u3r = U[2];
  U[1] = u2r * ny + u3r * nx;
  U[2] = u3r * ny - u2r * nx;
  V4t out;
  out.x = U[0];
  const Dd tt = 1.0 / U[0];
  out.y = U[1] * tt;
  out.z = U[2] * tt;
  out.w = 0.4 * (U[3] - (0.5 * tt) * (U[1] * U[1] + U[2] * U[2]));
  prim[i] = out;
}

// compute_force_coeffs.F90:37-77: one block per shape id, (H, V) of that shape as dual
// numbers into hv[4 s .. 4 s + 3] = (H.v, H.d, V.v, V.d)
__global__ __launch_bounds__(BLOCK) void k_tn_forces(int nshape, const int32_t* __restrict__ sh_m,
                                                     const int32_t* __restrict__ sh_id,
                                                     const int32_t* __restrict__ sh_l,
                                                     const int32_t* __restrict__ sh_r, const D2t* __restrict__ xy,
                                                     const D2t* __restrict__ nrm, const V4t* __restrict__ prim,
                                                     const double* __restrict__ shape_par, double* __restrict__ hv) {
  const int s = blockIdx.x;
  const double pr_inf = 1.0 / 1.4, cps = shape_par[3 * s];
  Dd H, V;
  for (int j = threadIdx.x; j < nshape; j += blockDim.x) {
    if (sh_id[j] != s) continue;
    const int m = sh_m[j], l = sh_l[j], r = sh_r[j];
    const D2t pl = xy[l], pm = xy[m], pr = xy[r];
    const Dd ds1 = m_sqrt((pm.x - pl.x) * (pm.x - pl.x) + (pm.y - pl.y) * (pm.y - pl.y));
    const Dd ds2 = m_sqrt((pr.x - pm.x) * (pr.x - pm.x) + (pr.y - pm.y) * (pr.y - pm.y));
    const Dd ds = 0.5 * (ds1 + ds2);
    Dd c = prim[m].w - pr_inf;
    c = -c / cps;
    H = H + c * nrm[m].x * ds;
    V = V + c * nrm[m].y * ds;
  }
  const double hv_[4] = {H.v, H.d, V.v, V.d};
  for (int k = 0; k < 4; ++k) {
    const double t = block_sum(hv_[k]);
    if (threadIdx.x == 0) hv[4 * s + k] = t;
    __syncthreads();
  }
}

}  // namespace mf

// ---------------------------------------------------------------------------------
// Host

struct TanState {
  D2t* xy = nullptr;
  D2t* nrm = nullptr;
  V4t *prim = nullptr, *pold = nullptr, *q = nullptr, *qm = nullptr;
  V4t* dq[2] = {nullptr, nullptr};
  Dd* delta = nullptr;
  int32_t *bd_m = nullptr, *bd_l = nullptr, *bd_r = nullptr;
  int nbd = 0;
  int32_t *sh_l = nullptr, *sh_r = nullptr;
  double* hv = nullptr;
};

static void tan_release(mf_ctx* c) {
  delete c->tan;  // its device arrays are the context's allocations
  c->tan = nullptr;
}

static int tan_alloc(mf_ctx* c, TanState& T) {
  const int n = c->n;
  int rc;
  if ((rc = dev_alloc(c, &T.xy, n)) || (rc = dev_alloc(c, &T.nrm, n)) || (rc = dev_alloc(c, &T.prim, n)) ||
      (rc = dev_alloc(c, &T.pold, n)) || (rc = dev_alloc(c, &T.q, n)) || (rc = dev_alloc(c, &T.qm, 2 * (size_t)n)) ||
      (rc = dev_alloc(c, &T.dq[0], 2 * (size_t)n)) || (rc = dev_alloc(c, &T.dq[1], 2 * (size_t)n)) ||
      (rc = dev_alloc(c, &T.delta, n)) || (rc = dev_alloc(c, &T.hv, 4 * (size_t)std::max(c->nsid, 1))))
    return rc;
  std::vector<int32_t> bm, bl, br, sl, sr;
  for (int j = 0; j < n; ++j) {
    if (c->left_i[j] < 0) continue;
    bm.push_back(j);
    bl.push_back(c->left_i[j]);
    br.push_back(c->right_i[j]);
  }
  for (int j = 0; j < c->nshape; ++j) {
    sl.push_back(c->left_i[c->h_sh_m[j]]);
    sr.push_back(c->right_i[c->h_sh_m[j]]);
  }
  T.nbd = (int)bm.size();
  if (T.nbd > 0 && ((rc = upload(c, &T.bd_m, bm)) || (rc = upload(c, &T.bd_l, bl)) || (rc = upload(c, &T.bd_r, br))))
    return rc;
  if (c->nshape > 0 && ((rc = upload(c, &T.sh_l, sl)) || (rc = upload(c, &T.sh_r, sr)))) return rc;
  return MF_OK;
}

extern "C" int mf_tangent(mf_ctx* c, int32_t iters, const double* vx, const double* vy, double* J, double* dJ,
                          double* prim_d) {
  if (!c || iters < 0 || !vx || !vy) return set_err(MF_ERR_INVALID, "mf_tangent: bad argument");
  if (c->ncases > 0) return set_err(MF_ERR_STATE, "mf_tangent runs on single-case contexts");
  if (c->nshape == 0) return set_err(MF_ERR_STATE, "J = Cd needs shape points");
  const int n = c->n;
  int rc;
  if (!c->tan) {
    c->tan = new TanState();
    if ((rc = tan_alloc(c, *c->tan))) return rc;
  }
  TanState& T = *c->tan;
  // seeded coordinates (the reference's pointd%x, pointd%y) and the normals: interior
  // ones fixed, boundary ones from the seeded coordinates
  {
    std::vector<D2> hxy(n), hn(n);
    HIPCHK(hipMemcpy(hxy.data(), c->g.xy, n * sizeof(D2), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hn.data(), c->g.nrm, n * sizeof(D2), hipMemcpyDeviceToHost));
    std::vector<D2t> txy(n), tn(n);
    for (int j = 0; j < n; ++j) {
      const int o = c->perm[j];
      txy[j] = D2t{Dd(hxy[j].x, vx[o]), Dd(hxy[j].y, vy[o])};
      tn[j] = D2t{Dd(hn[j].x), Dd(hn[j].y)};
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(T.xy, txy.data(), n * sizeof(D2t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(T.nrm, tn.data(), n * sizeof(D2t), hipMemcpyHostToDevice));
  }
  // the stencil kernels split each point's edges over kTanLanes lanes
  constexpr int kTanLanes = 4;
  const dim3 bl(BLOCK), gr((n + BLOCK - 1) / BLOCK);
  const dim3 gr_l((unsigned)(((int64_t)n * kTanLanes + BLOCK - 1) / BLOCK));
  if (T.nbd > 0)
    hipLaunchKernelGGL(k_tn_normals, dim3((T.nbd + BLOCK - 1) / BLOCK), bl, 0, c->stream, T.nbd, T.bd_m, T.bd_l,
                       T.bd_r, (const D2t*)T.xy, T.nrm);
  // the context's current state, with zero tangent
  if (c->ssz == 8)
    hipLaunchKernelGGL((k_tn_load<double>), gr, bl, 0, c->stream, n, (const V4d*)c->prim[c->cur], T.prim);
  else
    hipLaunchKernelGGL((k_tn_load<float>), gr, bl, 0, c->stream, n, (const V4<float>*)c->prim[c->cur], T.prim);
  const TanArgs a{T.xy, T.nrm, c->g.nbhs, c->g.conn, c->eps64, n, c->nw, c->no, c->p.power, c->p.first_order};
  const int S = c->p.rks, I = c->p.inner_iterations;
  for (int it = 0; it < iters; ++it) {
    // fpi_solver.F90:20-56
    hipLaunchKernelGGL(k_tn_copy, gr, bl, 0, c->stream, n, (const V4t*)T.prim, T.pold);
    hipLaunchKernelGGL(k_tn_delta, gr, bl, 0, c->stream, a, (const V4t*)T.prim, c->p.cfl, T.delta);
    for (int rk = 1; rk <= S; ++rk) {
      hipLaunchKernelGGL(k_tn_qvar, gr, bl, 0, c->stream, n, (const V4t*)T.prim, T.q);
      int d = 0;
      auto stencils = [&]<bool P0>() {
        hipLaunchKernelGGL((k_tn_lsq<kTanLanes, P0>), gr_l, bl, 0, c->stream, a, (const V4t*)T.q, T.dq[0], T.qm);
        for (int j = 0; j < I; ++j, d ^= 1)
          hipLaunchKernelGGL((k_tn_inner<kTanLanes, P0>), gr_l, bl, 0, c->stream, a, (const V4t*)T.q,
                             (const V4t*)T.dq[d], T.dq[d ^ 1]);
        hipLaunchKernelGGL((k_tn_flux_update<kTanLanes, P0>), gr_l, bl, 0, c->stream, a, (const V4t*)T.q,
                           (const V4t*)T.dq[d], (const V4t*)T.qm, (const Dd*)T.delta, T.prim, (const V4t*)T.pold, rk,
                           c->p.euler, c->fs);
      };
      if (c->p.power == 0.0) stencils.template operator()<true>();
      else stencils.template operator()<false>();
    }
    HIPCHK(hipGetLastError());
  }
  // objective_function_d: J = sum over shapes of Cd
  hipLaunchKernelGGL(k_tn_forces, dim3(c->nsid), bl, 0, c->stream, c->nshape, c->sh_m, c->sh_id, T.sh_l, T.sh_r,
                     (const D2t*)T.xy, (const D2t*)T.nrm, (const V4t*)T.prim, (const double*)c->shape_par, T.hv);
  HIPCHK(hipGetLastError());
  std::vector<double> hv(4 * (size_t)c->nsid);
  HIPCHK(hipMemcpyAsync(hv.data(), T.hv, hv.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  double jv = 0.0, jd = 0.0;
  for (int s = 0; s < c->nsid; ++s) {
    const double ct = c->h_shape_par[3 * s + 1], st = c->h_shape_par[3 * s + 2];
    jv += hv[4 * s] * ct + hv[4 * s + 2] * st;
    jd += hv[4 * s + 1] * ct + hv[4 * s + 3] * st;
  }
  if (J) *J = jv;
  if (dJ) *dJ = jd;
  if (prim_d) {
    std::vector<V4t> hp(n);
    HIPCHK(hipMemcpy(hp.data(), T.prim, n * sizeof(V4t), hipMemcpyDeviceToHost));
    for (int j = 0; j < n; ++j)
      for (int r = 0; r < 4; ++r) prim_d[(size_t)4 * c->perm[j] + r] = hp[j][r].d;
  }
  return MF_OK;
}
