// Adjoint of the LSKUM iteration: reverse kernels and the checkpointed reverse sweep.
//
// Included at the end of mfcfd.hip (one translation unit: it shares the context and
// the forward kernels).  J = sum over shapes of Cd after the last iteration, as the
// reference's adjoint seeds it (q_lskum_chkpts_b.F90:95, objective_function_b.F90:22-23);
// dJ/dx and dJ/dy at every point, through the LSQ geometry, the time steps, the
// normals and the force integral (compute_force_coeffs_b.F90, point_normals_b.F90).
//
// Per reversed iteration the stage intermediates are recomputed from that iteration's
// prim (one taped forward iteration, the forward kernels writing into tape buffers),
// then reversed stage by stage.  A point's own adjoints are summed in registers and
// written once; contributions to a neighbour's adjoints are written per edge into a
// slot-major edge buffer and gathered by the neighbour over the transposed
// connectivity, so the stencil reverses use no atomics (the argmin of func_delta, the
// force integral and the normals, a few contributions per point, still do).
// Iterates are checkpointed with a binomial (revolve-style) schedule of prim snapshots
// in HBM (q_lskum_chkpts_b.F90:116-279 drives Griewank's REVOLV the same way).
#include "lskum_adjoint.h"

namespace mf {

// reverse kernels: multiply-adds contracted into FMAs, as in lskum_adjoint.h
#pragma clang fp contract(fast)

__device__ inline void aadd(double* p, double v) {
  if (v != 0.0) unsafeAtomicAdd(p, v);
}
__device__ inline void aadd4(V4d* p, const double v[4]) {
  double* d = &p->x;
  for (int r = 0; r < 4; ++r) aadd(d + r, v[r]);
}
__device__ inline void aadd2(D2* p, double x, double y) {
  aadd(&p->x, x);
  aadd(&p->y, y);
}

// Per-edge adjoint contributions to the neighbour's record.  Each edge's fields are
// contiguous, and the edges are stored in the order of the transposed connectivity
// (grouped by the neighbour that receives them), so a gather reads one contiguous
// block per point; epos[s * N + i] is the position of the edge in slot s of point i.
// The flux reverse writes 14 fields (q(4), d(4), m(4), delx, dely; d being the
// reconstruction increment delx dq_x + dely dq_y, m the adjoint of the one of qmax or
// qmin the limiter read for that component, which esel's bit r names: 1 qmax, 0 qmin),
// the LSQ reverses 6 (q(4), delx, dely).
constexpr int EF_FLUX = 14;
constexpr int EF_LSQ = 6;

struct AdjOut {
  V4d* qb;      // adjoint of the stage's q (own contributions)
  V4d* dqb;     // adjoint of the dq the kernel reads [2N]
  V4d* qmb;     // adjoint of qm [2N]
  double* deltab;
  D2* xyb;      // adjoint of the coordinates
  D2* nrmb;     // adjoint of the normals
  double* eb;            // edge buffer
  const int32_t* epos;   // edge positions [W * N]
  uint8_t* esel;         // per edge: bit r set when field m(r) is qmax's adjoint [E]
};

// Half-stencil geometry sums (sum ds^2, dn^2, ds dn per half), kept for the adjoint
struct HalfSums {
  const V4d* hxx;
  const V4d* hyy;
  const V4d* hxy;
};

// Transposed connectivity: the edges whose neighbour is k sit at positions
// ptr[k] .. ptr[k+1]-1 of the edge buffer; own[j] is the point that owns edge j
struct Trans {
  const int32_t* ptr;  // [N+1]
  const int32_t* own;  // [E]
};

// Halo of the pull-form inner reverse in blocks of STAGE_P points: each block's own
// points, then once each (ascending) every other point that is a neighbour of one of
// them or names one of them as a neighbour.  loc8 [W][stride]: each slot's position in
// its block's halo; tloc8 [E]: the position of each transposed entry's owner in the
// halo of the block holding the entry's neighbour; 255 for a position past hcap.
struct PullHalo {
  const int32_t* ptr = nullptr;  // [nblk+1]
  const int32_t* ids = nullptr;
  const uint8_t* loc8 = nullptr;
  const uint8_t* tloc8 = nullptr;
  int32_t hcap = 0;    // entries held in LDS: at least STAGE_P, a multiple of 8, <= 248
  int32_t stride = 0;  // N rounded up to a multiple of STAGE_P
};

// Least-squares weight w = dist^power of an edge with offset (a, b) (q_variables.F90:89-91,
// interior_fluxes.F90:89-90) and f = power dist^(power - 2), so dw/da = f a and
// dw/db = f b; power = 0 gives w = 1 and f = 0 exactly
// POWER0: the kernels' unweighted specialisation (power = 0), which carries no pow
// evaluation (long and register-hungry in FP64)
template <bool POWER0>
__device__ inline void lsq_weight(double a, double b, double power, double& w, double& f) {
  if (POWER0 || power == 0.0) {
    w = 1.0;
    f = 0.0;
    return;
  }
  const double d2 = a * a + b * b;
  w = pow(sqrt(d2), power);
  f = power * w / d2;
}

// q-tilde of one end of an edge (the point's own q when first order); p indexes the
// records of nbrs (a point id, or a halo position)
template <class Nbrs, class S = typename Nbrs::scalar>
__device__ inline void end_qtilde(const Nbrs& nbrs, int p, S delx, S dely, int first_order, S qt[4]) {
  const V4<S> q = nbrs.q(p);
  if (first_order) {
    for (int r = 0; r < 4; ++r) qt[r] = q[r];
  } else {
    reconstruct(q, nbrs.dqx(p), nbrs.dqy(p), nbrs.qmax(p), nbrs.qmin(p), nbrs.eps(p), delx, dely, qt);
  }
}

// Reverse of one point's flux_res (wall/outer/interior_dGx_pos..neg with the edge
// coefficients).  Per edge, one end at a time (its q-tilde and kinetic intermediates
// live only while that end is processed): the end's fluxes feed the adjoints of the
// edge's coefficients and of delta, and the end's flux adjoints are reversed into its
// record.  nbrs gives the records of both ends (the taped state in global memory, or a
// block's halo staged in LDS) and the point's slots.
template <int CLS, int LPP, bool POWER0, class Nbrs>
__device__ void flux_point_rev(const Geom& g, const HalfSums& hs, const Nbrs& nbrs, int i, double delta,
                               const V4d& frbv, const AdjOut& o, int lane, double power) {
  // the records' precision: the arithmetic of the ends' reconstruction, kinetic state
  // and fluxes and of their reverse; coefficients and adjoint sums stay FP64
  using S = typename Nbrs::scalar;
  const D2 pi = g.xy[i];
  const D2 ni = g.nrm[i];
  const double nx = ni.x, ny = ni.y, tx = ny, ty = -nx;
  // the half coefficients as the products A = c1 obd, B = sxy obd, and their adjoints:
  // 8 values each live across the walk instead of 12
  HalfAB hc, hcb;
  {
    const V4d a = g.dg_c1[i], b = g.dg_sxy[i], c = g.dg_obd[i];
    for (int d = 0; d < 4; ++d) {
      hc.A[d] = a[d] * c[d];
      hc.B[d] = b[d] * c[d];
      hcb.A[d] = hcb.B[d] = 0.0;
    }
  }
  const double scale = (CLS == WALL) ? 2.0 * delta : delta;
  double accb[4];
  for (int r = 0; r < 4; ++r) accb[r] = scale * frbv[r];
  double dsum = 0.0;  // accb . (flux_res / scale) = scale times delta's adjoint
  EndD own;
  endd_zero(own);
  double odx[4] = {0, 0, 0, 0}, ody[4] = {0, 0, 0, 0};  // own dq adjoints
  double xib = 0.0, yib = 0.0, nxb = 0.0, nyb = 0.0;
  const int nb = g.nbhs[i];
  for (int s = lane; s < nb; s += LPP) {
    int k;
    unsigned m;
    nbrs.slot(s, k, m);
    const D2 pkk = nbrs.xy(k);
    const double delx = pkk.x - pi.x, dely = pkk.y - pi.y;
    const double dels = delx * tx + dely * ty;
    const double deln = delx * nx + dely * ny;
    double w, wf;
    lsq_weight<POWER0>(dels, deln, power, w, wf);
    const double dsw = dels * w, dnw = deln * w;
    const bool xp = (m & XPOS) != 0u;
    const bool tie_x = xp && (m & XNEG) != 0u;
    const bool yp = (CLS == INTERIOR) ? (m & YPOS) != 0u : (CLS == OUTER);
    const bool tie_y = (CLS == INTERIOR) && yp && (m & YNEG) != 0u;
    const bool tie = tie_x || tie_y;
    const double cx = coef_x(hc, xp, dsw, dnw);
    const double cy = coef_y(hc, yp, dsw, dnw);
    const double cx2 = tie_x ? coef_x(hc, false, dsw, dnw) : 0.0;
    const double cy2 = tie_y ? coef_y(hc, false, dsw, dnw) : 0.0;
    // per end (own, then neighbour): its q-tilde and kinetic intermediates once, the
    // forward fluxes for the coefficients' adjoints, then the reverse of the fluxes,
    // the intermediates and the reconstruction
    double cxb = 0.0, cyb = 0.0, cx2b = 0.0, cy2b = 0.0;
    double delxb = 0.0, delyb = 0.0;
    const size_t epj = (size_t)o.epos[(size_t)s * g.n + i];
    D2* ep = reinterpret_cast<D2*>(o.eb + epj * EF_FLUX);
    auto end_pass = [&](const int e) {
      const int p = e ? k : nbrs.own();
      S qt[4];
      end_qtilde(nbrs, p, (S)delx, (S)dely, g.first_order, qt);
      KinFT<S> f;
      kin_fwd<S>(qt, (S)nx, (S)ny, f);
      const double sg = e ? 1.0 : -1.0;
      KinBT<S> kb;
      kinb_zero(kb);
      {
        S Gx[4], Gy[4];
        end_fluxes_f<CLS>(f, xp, yp, Gx, Gy);
        double ax = 0.0, ay = 0.0, fd = 0.0;
        for (int r = 0; r < 4; ++r) {
          ax += accb[r] * (double)Gx[r];
          ay += accb[r] * (double)Gy[r];
          fd += accb[r] * ((double)Gx[r] * cx + (double)Gy[r] * cy);
          Gx[r] = (S)(sg * accb[r] * cx);
          Gy[r] = (S)(sg * accb[r] * cy);
        }
        cxb += sg * ax;
        cyb += sg * ay;
        dsum += sg * fd;
        end_fluxes_rev<CLS>(f, xp, yp, Gx, Gy, kb);
      }
      if (tie) {
        S Hx[4], Hy[4];
        end_fluxes_f<CLS>(f, false, false, Hx, Hy);
        double hx = 0.0, hy = 0.0, fd = 0.0;
        for (int r = 0; r < 4; ++r) {
          hx += accb[r] * (double)Hx[r];
          hy += accb[r] * (double)Hy[r];
          fd += accb[r] * ((double)Hx[r] * cx2 + (double)Hy[r] * cy2);
          Hx[r] = (S)(sg * accb[r] * cx2);
          Hy[r] = (S)(sg * accb[r] * cy2);
        }
        cx2b += sg * hx;
        cy2b += sg * hy;
        dsum += sg * fd;
        end_fluxes_rev<CLS>(f, false, false, Hx, Hy, kb);
      }
      S qtb[4] = {(S)0, (S)0, (S)0, (S)0};
      kin_rev<S>(f, kb, (S)nx, (S)ny, qtb, nxb, nyb);
      EndD d;
      endd_zero(d);
      if (g.first_order) {
        for (int r = 0; r < 4; ++r) d.q[r] = (double)qtb[r];
      } else {
        reconstruct_rev_d<S>(nbrs.q(p), nbrs.dqx(p), nbrs.dqy(p), nbrs.qmax(p), nbrs.qmin(p), nbrs.eps(p),
                             (S)delx, (S)dely, qtb, d, delxb, delyb);
      }
      if (e) {
        ep[0] = D2{d.q[0], d.q[1]};
        ep[1] = D2{d.q[2], d.q[3]};
        ep[2] = D2{d.d[0], d.d[1]};
        ep[3] = D2{d.d[2], d.d[3]};
        // the limiter reads qmax or qmin per component (venkat_rev), so at most one
        // of the two adjoints is nonzero
        double m[4];
        unsigned sel = 0u;
        for (int r = 0; r < 4; ++r) {
          const bool mx = d.qmax[r] != 0.0;
          m[r] = mx ? d.qmax[r] : d.qmin[r];
          sel |= mx ? 1u << r : 0u;
        }
        ep[4] = D2{m[0], m[1]};
        ep[5] = D2{m[2], m[3]};
        o.esel[epj] = (uint8_t)sel;
      } else {
        for (int r = 0; r < 4; ++r) {
          own.q[r] += d.q[r];
          own.qmax[r] += d.qmax[r];
          own.qmin[r] += d.qmin[r];
          // d = delx dq_x + dely dq_y of this edge: its adjoint splits on delx, dely
          odx[r] += d.d[r] * delx;
          ody[r] += d.d[r] * dely;
        }
      }
    };
    if constexpr (sizeof(S) == 8) {
      // FP64: both ends in straight-line code, so that the one wave a SIMD holds can
      // overlap one end's library transcendentals with the other's
      end_pass(0);
      end_pass(1);
    } else {
      // FP32: one end at a time, which keeps the kernel at 2 waves per SIMD
#pragma unroll 1
      for (int e = 0; e < 2; ++e) end_pass(e);
    }
    double dswb = 0.0, dnwb = 0.0;
    coef_rev_ab(hc, xp ? 0 : 1, false, dsw, dnw, cxb, hcb, dswb, dnwb);
    coef_rev_ab(hc, yp ? 2 : 3, true, dsw, dnw, cyb, hcb, dswb, dnwb);
    if (tie_x) coef_rev_ab(hc, 1, false, dsw, dnw, cx2b, hcb, dswb, dnwb);
    if (tie_y) coef_rev_ab(hc, 3, true, dsw, dnw, cy2b, hcb, dswb, dnwb);
    // dsw = dels w, dnw = deln w; dels = delx ny - dely nx, deln = delx nx + dely ny
    double delsb = dswb * w, delnb = dnwb * w;
    if (wf != 0.0) {
      const double wb = dswb * dels + dnwb * deln;
      delsb += wb * wf * dels;
      delnb += wb * wf * deln;
    }
    delxb += delsb * ny + delnb * nx;
    delyb += -delsb * nx + delnb * ny;
    nxb += -delsb * dely + delnb * delx;
    nyb += delsb * delx + delnb * dely;
    xib -= delxb;
    yib -= delyb;
    ep[6] = D2{delxb, delyb};
  }
  // half-stencil geometry: c1 = sum dn^2 (x halves) or sum ds^2 (y halves), sxy = sum
  // ds dn, obd = 1/(sum ds^2 sum dn^2 - sxy^2); every lane needs the point's totals
  if (LPP > 1) {
    for (int d = 0; d < 4; ++d) {
      hcb.A[d] = group_sum<LPP>(hcb.A[d]);
      hcb.B[d] = group_sum<LPP>(hcb.B[d]);
    }
  }
  double hxxb[4], hyyb[4], hxyb[4];
  {
    const V4d hxx = hs.hxx[i], hyy = hs.hyy[i], hxy = hs.hxy[i];
    const V4d c1v = g.dg_c1[i], sxyv = g.dg_sxy[i], obdv = g.dg_obd[i];
    for (int d = 0; d < 4; ++d) {
      // A = c1 obd, B = sxy obd; a half without edges (obd infinite) has no adjoint
      const double obd = obdv[d];
      const bool used = hcb.A[d] != 0.0 || hcb.B[d] != 0.0;
      const double c1b = used ? hcb.A[d] * obd : 0.0;
      const double obdb = used ? hcb.A[d] * c1v[d] + hcb.B[d] * sxyv[d] : 0.0;
      hxxb[d] = (d >= 2) ? c1b : 0.0;
      hyyb[d] = (d < 2) ? c1b : 0.0;
      hxyb[d] = used ? hcb.B[d] * obd : 0.0;
      const double detb = -obdb * obd * obd;
      if (obdb != 0.0) {
        hxxb[d] += detb * hyy[d];
        hyyb[d] += detb * hxx[d];
        hxyb[d] += -2.0 * detb * hxy[d];
      }
    }
  }
  for (int s = lane; s < nb; s += LPP) {
    int k;
    unsigned m;
    nbrs.slot(s, k, m);
    const D2 pkk = nbrs.xy(k);
    const double delx = pkk.x - pi.x, dely = pkk.y - pi.y;
    const double dels = delx * tx + dely * ty;
    const double deln = delx * nx + dely * ny;
    double w, wf;
    lsq_weight<POWER0>(dels, deln, power, w, wf);
    // the half sums dels^2 w, deln^2 w, dels deln w
    double delsb = 0.0, delnb = 0.0, wb = 0.0;
    for (int d = 0; d < 4; ++d) {
      if (!(m & (1u << d))) continue;
      delsb += w * (2.0 * dels * hxxb[d] + deln * hxyb[d]);
      delnb += w * (2.0 * deln * hyyb[d] + dels * hxyb[d]);
      wb += dels * dels * hxxb[d] + deln * deln * hyyb[d] + dels * deln * hxyb[d];
    }
    if (wf != 0.0) {
      delsb += wb * wf * dels;
      delnb += wb * wf * deln;
    }
    const double delxb = delsb * ny + delnb * nx;
    const double delyb = -delsb * nx + delnb * ny;
    nxb += -delsb * dely + delnb * delx;
    nyb += delsb * delx + delnb * dely;
    xib -= delxb;
    yib -= delyb;
    D2* ep = reinterpret_cast<D2*>(o.eb + (size_t)o.epos[(size_t)s * g.n + i] * EF_FLUX);
    const D2 t = ep[6];
    ep[6] = D2{t.x + delxb, t.y + delyb};
  }
  // own contributions, summed over the point's lanes and written (q, dq, qm) or
  // added (xy, normals, delta) by its first lane alone; the edges' contributions to
  // their neighbours are gathered by k_gather_flux
  if (LPP > 1) {
    for (int r = 0; r < 4; ++r) {
      own.q[r] = group_sum<LPP>(own.q[r]);
      own.qmax[r] = group_sum<LPP>(own.qmax[r]);
      own.qmin[r] = group_sum<LPP>(own.qmin[r]);
      odx[r] = group_sum<LPP>(odx[r]);
      ody[r] = group_sum<LPP>(ody[r]);
    }
    xib = group_sum<LPP>(xib);
    yib = group_sum<LPP>(yib);
    nxb = group_sum<LPP>(nxb);
    nyb = group_sum<LPP>(nyb);
    dsum = group_sum<LPP>(dsum);
    if (lane != 0) return;
  }
  {
    const D2 t = o.xyb[i];
    o.xyb[i] = D2{t.x + xib, t.y + yib};
    o.qb[i] = V4d{own.q[0], own.q[1], own.q[2], own.q[3]};
    o.dqb[2 * i] = V4d{odx[0], odx[1], odx[2], odx[3]};
    o.dqb[2 * i + 1] = V4d{ody[0], ody[1], ody[2], ody[3]};
    o.qmb[2 * i] = V4d{own.qmax[0], own.qmax[1], own.qmax[2], own.qmax[3]};
    o.qmb[2 * i + 1] = V4d{own.qmin[0], own.qmin[1], own.qmin[2], own.qmin[3]};
  }
  o.nrmb[i].x += nxb;
  o.nrmb[i].y += nyb;
  o.deltab[i] += dsum / delta;  // delta enters every stage (scale = 2 delta at a wall point)
}

// S: the taped records' precision (double; float for a mixed-precision context, whose
// records are the mixed primal's own FP32 ones)
template <int LPP, bool STAGED, class S, bool POWER0>
__global__ __launch_bounds__(BLOCK, 1) void k_flux_rev(Geom g, HalfSums hs, State<S> st,
                                                       const double* __restrict__ delta,
                                                       const V4d* __restrict__ frb, AdjOut o, double power) {
  if constexpr (STAGED) {
    // the block's halo of the taped stage state in LDS (blocks of STAGE_PF points)
    extern __shared__ __align__(16) unsigned char smem[];
    const int b = xcd_block(blockIdx.x, gridDim.x);
    HaloRecs<S> r;
    uint16_t* s_cl;
    const bool staged = stage_flux_halo<S, LPP>(g, st, b, smem, r, s_cl);
    const int li = threadIdx.x / LPP, lane = threadIdx.x % LPP;
    const int i = b * STAGE_PF + li;
    if (i >= g.n) return;  // whole lane groups
    if (staged) {
      const StagedNbrs<S, STAGE_PF, LPP> nb{g, r, s_cl, li, g.nbhs[i], lane};
      if (i < g.nw) flux_point_rev<WALL, LPP, POWER0>(g, hs, nb, i, delta[i], frb[i], o, lane, power);
      else if (i < g.nw + g.no) flux_point_rev<OUTER, LPP, POWER0>(g, hs, nb, i, delta[i], frb[i], o, lane, power);
      else flux_point_rev<INTERIOR, LPP, POWER0>(g, hs, nb, i, delta[i], frb[i], o, lane, power);
    } else {
      const GlobalNbrs<S> nb{g, st, i, g.nbhs[i], lane, LPP};
      if (i < g.nw) flux_point_rev<WALL, LPP, POWER0>(g, hs, nb, i, delta[i], frb[i], o, lane, power);
      else if (i < g.nw + g.no) flux_point_rev<OUTER, LPP, POWER0>(g, hs, nb, i, delta[i], frb[i], o, lane, power);
      else flux_point_rev<INTERIOR, LPP, POWER0>(g, hs, nb, i, delta[i], frb[i], o, lane, power);
    }
  } else {
    const int t = point_id();
    const int i = t / LPP, lane = t % LPP;
    if (i >= g.n) return;  // whole lane groups
    const GlobalNbrs<S> nb{g, st, i, g.nbhs[i], lane, LPP};
    if (i < g.nw) flux_point_rev<WALL, LPP, POWER0>(g, hs, nb, i, delta[i], frb[i], o, lane, power);
    else if (i < g.nw + g.no) flux_point_rev<OUTER, LPP, POWER0>(g, hs, nb, i, delta[i], frb[i], o, lane, power);
    else flux_point_rev<INTERIOR, LPP, POWER0>(g, hs, nb, i, delta[i], frb[i], o, lane, power);
  }
}

// reverse of the state update of stage rk: from the adjoint of the stage's new prim
// (pnb) to the adjoints of its input prim (pb, written), prim_old (poldb, added),
// flux_res (frb, written) and the normals.  The conserved state after the update is
// rebuilt from the stored new prim.
__global__ __launch_bounds__(BLOCK) void k_update_rev(Geom g, const V4d* __restrict__ prim,
                                                      const V4d* __restrict__ prim_old,
                                                      const V4d* __restrict__ prim_new,
                                                      const V4d* __restrict__ pnb, V4d* __restrict__ pb,
                                                      V4d* __restrict__ poldb, V4d* __restrict__ frb,
                                                      D2* __restrict__ nrmb, int rk, double euler,
                                                      Freestream fs) {
  const int i = point_id();
  if (i >= g.n) return;
  const D2 nn = g.nrm[i];
  const V4d pn = prim_new[i];
  double pbv[4] = {0, 0, 0, 0}, poldbv[4] = {0, 0, 0, 0}, frbv[4] = {0, 0, 0, 0};
  double nxb = 0.0, nyb = 0.0;
  const V4d p = prim[i], po = prim_old[i], b = pnb[i];
  if (i < g.nw) update_rev<WALL>(p, po, pn, nn.x, nn.y, rk, euler, fs, b, pbv, poldbv, frbv, nxb, nyb);
  else if (i < g.nw + g.no) {
    const Freestream f = g.fs_tab ? g.fs_tab[g.case_id[i]] : fs;
    update_rev<OUTER>(p, po, pn, nn.x, nn.y, rk, euler, f, b, pbv, poldbv, frbv, nxb, nyb);
  }
  else update_rev<INTERIOR>(p, po, pn, nn.x, nn.y, rk, euler, fs, b, pbv, poldbv, frbv, nxb, nyb);
  pb[i] = V4d{pbv[0], pbv[1], pbv[2], pbv[3]};
  if (rk == 3) {
    V4d t = poldb[i];
    poldb[i] = V4d{t.x + poldbv[0], t.y + poldbv[1], t.z + poldbv[2], t.w + poldbv[3]};
  }
  frb[i] = V4d{frbv[0], frbv[1], frbv[2], frbv[3]};
  nrmb[i].x += nxb;
  nrmb[i].y += nyb;
}

// Reverse of one point's inner-iteration solve (q_variables.F90:197-220): from the
// adjoints (bx, by) of its output dq = ((sum_x syy - sum_y sxy) obd, (sum_y sxx - sum_x
// sxy) obd) and that output (tx, ty), the adjoints of the sums sum_x, sum_y per
// component and of the geometry sums sxx, syy, sxy (also through obd = 1/det)
struct InnerLam {
  double sqb[4], syqb[4];
  double sxxb, syyb, sxyb;
};

__device__ inline void inner_lam(const V4d& gg, const V4d& tx, const V4d& ty, const V4d& bx, const V4d& by,
                                 InnerLam& L) {
  const double sxx = gg.x, syy = gg.y, sxy = gg.z, obd = gg.w;
  double sxxb = 0.0, syyb = 0.0, sxyb = 0.0, obdb = 0.0;
  for (int r = 0; r < 4; ++r) {
    const double sxq = sxx * tx[r] + sxy * ty[r];
    const double syq = sxy * tx[r] + syy * ty[r];
    L.sqb[r] = (bx[r] * syy - by[r] * sxy) * obd;
    L.syqb[r] = (by[r] * sxx - bx[r] * sxy) * obd;
    obdb += bx[r] * (sxq * syy - syq * sxy) + by[r] * (syq * sxx - sxq * sxy);
    syyb += bx[r] * sxq * obd;
    sxxb += by[r] * syq * obd;
    sxyb += -(bx[r] * syq + by[r] * sxq) * obd;
  }
  const double detb = -obdb * obd * obd;
  L.sxxb = sxxb + detb * syy;
  L.syyb = syyb + detb * sxx;
  L.sxyb = sxyb - 2.0 * detb * sxy;
}

// Reverse of one edge's terms of an inner iteration at its owner i, the neighbour k at
// offset (dx, dy) = x_k - x_i (q_variables.F90:164-189): with the owner's solve adjoint
// L, the adjoint Db of the edge's difference D = (q_k - dq_k.d/2) - (q_i - dq_i.d/2) per
// component (the neighbour's q takes +Db, the owner's -Db; the dq records -/+ Db d/2)
// and the adjoint (dxb, dyb) of the offset (the neighbour's coordinates take +, the
// owner's -).  The owner (push form) and the neighbour (pull form) evaluate it alike.
template <bool POWER0>
__device__ inline void inner_edge_rev(double dx, double dy, double power, const V4d& qi, const V4d& dxi,
                                      const V4d& dyi, const V4d& qk, const V4d& dxk, const V4d& dyk,
                                      const InnerLam& L, double Db[4], double& dxb, double& dyb) {
  double w, wf;
  lsq_weight<POWER0>(dx, dy, power, w, wf);
  // the sums of w dx^2, w dy^2, w dx dy and w (dx, dy) D
  dxb = w * (2.0 * L.sxxb * dx + L.sxyb * dy);
  dyb = w * (2.0 * L.syyb * dy + L.sxyb * dx);
  double wb = L.sxxb * dx * dx + L.syyb * dy * dy + L.sxyb * dx * dy;
  for (int r = 0; r < 4; ++r) {
    const double t1 = qi[r] - 0.5 * (dx * dxi[r] + dy * dyi[r]);
    const double t2 = qk[r] - 0.5 * (dx * dxk[r] + dy * dyk[r]);
    const double D = t2 - t1;
    Db[r] = w * (L.sqb[r] * dx + L.syqb[r] * dy);
    dxb += w * L.sqb[r] * D - 0.5 * Db[r] * (dxk[r] - dxi[r]);
    dyb += w * L.syqb[r] * D - 0.5 * Db[r] * (dyk[r] - dyi[r]);
    wb += D * (L.sqb[r] * dx + L.syqb[r] * dy);
  }
  if (wf != 0.0) {
    dxb += wb * wf * dx;
    dyb += wb * wf * dy;
  }
}

// Reverse of one edge's terms of the LSQ derivatives at its owner i, the neighbour k at
// offset (dx, dy) (q_variables.F90:89-103): with the owner's solve adjoint L, the
// adjoint Db of D = q_k - q_i per component (+ to the neighbour's q, - to the owner's)
// and the adjoint (dxb, dyb) of the offset.  The owner (push form) and the neighbour
// (pull form) evaluate it alike.
template <bool POWER0>
__device__ inline void lsq_edge_rev(double dx, double dy, double power, const V4d& qi, const V4d& qk,
                                    const InnerLam& L, double Db[4], double& dxb, double& dyb) {
  double w, wf;
  lsq_weight<POWER0>(dx, dy, power, w, wf);
  // the sums of w dx^2, w dy^2, w dx dy and w (dx, dy) D
  dxb = w * (2.0 * L.sxxb * dx + L.sxyb * dy);
  dyb = w * (2.0 * L.syyb * dy + L.sxyb * dx);
  double wb = L.sxxb * dx * dx + L.syyb * dy * dy + L.sxyb * dx * dy;
  for (int r = 0; r < 4; ++r) {
    const double D = qk[r] - qi[r];
    Db[r] = w * (L.sqb[r] * dx + L.syqb[r] * dy);
    dxb += w * L.sqb[r] * D;
    dyb += w * L.syqb[r] * D;
    wb += D * (L.sqb[r] * dx + L.syqb[r] * dy);
  }
  if (wf != 0.0) {
    dxb += wb * wf * dx;
    dyb += wb * wf * dy;
  }
}

// reverse of k_inner (one inner iteration): dq_out adjoint -> dq_in, q, xy adjoints.
// Own contributions: dqb_in written, qb and xyb added; per edge, the neighbour's q
// adjoint and the edge's (delx, dely) adjoint go to the edge buffer (fields 0-5), the
// neighbour's dq adjoint being -0.5 (delx, dely) times its q adjoint (k_gather_lsq).
template <int LPP, bool STAGED, bool POWER0>
__global__ __launch_bounds__(BLOCK) void k_inner_rev(Geom g, const V4d* __restrict__ q,
                                                     const V4d* __restrict__ dq_in,
                                                     const V4d* __restrict__ dq_out,
                                                     const V4d* __restrict__ dqb_out,
                                                     V4d* __restrict__ dqb_in, V4d* __restrict__ qb,
                                                     D2* __restrict__ xyb, double* __restrict__ eb,
                                                     const int32_t* __restrict__ epos, double power) {
  const PointMap<LPP, STAGED> pm(g.n);
  if (!STAGED && !pm.live) return;
  const int i = pm.live ? pm.i : g.n - 1, lane = pm.lane;
  const D2 pi = g.xy[i];
  InnerLam L;
  inner_lam(g.lsq_geo[i], dq_out[2 * i], dq_out[2 * i + 1], dqb_out[2 * i], dqb_out[2 * i + 1], L);
  const V4d qi = q[i], dxi = dq_in[2 * i], dyi = dq_in[2 * i + 1];
  double qib[4] = {0, 0, 0, 0}, dxib[4] = {0, 0, 0, 0}, dyib[4] = {0, 0, 0, 0};
  double xib = 0.0, yib = 0.0;
  const int nb = pm.live ? g.nbhs[i] : 0;
  // edge in slot s to the neighbour with coordinates pk and records qk, dxk, dyk
  auto edge = [&](int s, const D2& pk, const V4d& qk, const V4d& dxk, const V4d& dyk) {
    const double dx = pk.x - pi.x, dy = pk.y - pi.y;
    double Dbv[4], dxb, dyb;
    inner_edge_rev<POWER0>(dx, dy, power, qi, dxi, dyi, qk, dxk, dyk, L, Dbv, dxb, dyb);
    for (int r = 0; r < 4; ++r) {
      qib[r] -= Dbv[r];
      dxib[r] += 0.5 * dx * Dbv[r];
      dyib[r] += 0.5 * dy * Dbv[r];
    }
    D2* ep = reinterpret_cast<D2*>(eb + (size_t)epos[(size_t)s * g.n + i] * EF_LSQ);
    ep[0] = D2{Dbv[0], Dbv[1]};
    ep[1] = D2{Dbv[2], Dbv[3]};
    ep[2] = D2{dxb, dyb};
    xib -= dxb;
    yib -= dyb;
  };
  if constexpr (STAGED) {
    // the block's halo (q, dq_in, xy) in LDS
    extern __shared__ __align__(16) unsigned char smem[];
    const int H = g.hl[0].hcap;
    V4d* s_q = (V4d*)smem;
    V4d* s_dx = s_q + H;
    V4d* s_dy = s_dx + H;
    D2* s_xy = (D2*)(s_dy + H);
    uint8_t* s_cl = (uint8_t*)(s_xy + H);
    stage_and_walk<STAGE_P, LPP>(
        g, g.hl[0], pm.b, nb, s_cl,
        [&](int j, int k) {
          s_q[j] = q[k];
          s_dx[j] = dq_in[2 * k];
          s_dy[j] = dq_in[2 * k + 1];
          s_xy[j] = g.xy[k];
        },
        [&](int j, int s) { edge(s, s_xy[j], s_q[j], s_dx[j], s_dy[j]); },
        [&](int k, int s) { edge(s, g.xy[k], q[k], dq_in[2 * k], dq_in[2 * k + 1]); });
    if (!pm.live) return;
  } else {
    for (int s = lane; s < nb; s += LPP) {
      const int k = (int)(g.conn[(size_t)s * g.n + i] >> 4);
      edge(s, g.xy[k], q[k], dq_in[2 * k], dq_in[2 * k + 1]);
    }
  }
  if (LPP > 1) {
    for (int r = 0; r < 4; ++r) {
      qib[r] = group_sum<LPP>(qib[r]);
      dxib[r] = group_sum<LPP>(dxib[r]);
      dyib[r] = group_sum<LPP>(dyib[r]);
    }
    xib = group_sum<LPP>(xib);
    yib = group_sum<LPP>(yib);
    if (lane != 0) return;
  }
  const V4d a = qb[i];
  qb[i] = V4d{a.x + qib[0], a.y + qib[1], a.z + qib[2], a.w + qib[3]};
  dqb_in[2 * i] = V4d{dxib[0], dxib[1], dxib[2], dxib[3]};
  dqb_in[2 * i + 1] = V4d{dyib[0], dyib[1], dyib[2], dyib[3]};
  const D2 pxy = xyb[i];
  xyb[i] = D2{pxy.x + xib, pxy.y + yib};
}

// LDS of the pull-form inner reverse per halo entry: q, dq_x, dq_y, the solve adjoint's
// sqb and syqb (V4d each), xy, and sxxb, syyb, sxyb; then the block's slot words
inline size_t stage_lds_inner_pull(int hcap, int width) {
  return (size_t)hcap * (5 * sizeof(V4d) + sizeof(D2) + 3 * sizeof(double)) + (size_t)width * STAGE_P;
}

// A point's solve adjoint (InnerLam) as three records: sqb, syqb, (sxxb, syyb, sxyb, 0)
__device__ inline void lam_store(V4d* __restrict__ lam, int i, const InnerLam& L) {
  lam[3 * i] = V4d{L.sqb[0], L.sqb[1], L.sqb[2], L.sqb[3]};
  lam[3 * i + 1] = V4d{L.syqb[0], L.syqb[1], L.syqb[2], L.syqb[3]};
  lam[3 * i + 2] = V4d{L.sxxb, L.syyb, L.sxyb, 0.0};
}

// Solve adjoints of every point for the first reversed inner iteration of a stage, from
// the adjoint of that iteration's output (dqb_out) and the output itself (dq_out)
__global__ __launch_bounds__(BLOCK) void k_inner_lam(Geom g, const V4d* __restrict__ dq_out,
                                                     const V4d* __restrict__ dqb_out, V4d* __restrict__ lam) {
  const int i = point_id();
  if (i >= g.n) return;
  InnerLam L;
  inner_lam(g.lsq_geo[i], dq_out[2 * i], dq_out[2 * i + 1], dqb_out[2 * i], dqb_out[2 * i + 1], L);
  lam_store(lam, i, L);
}

// Pull form of k_inner_rev + k_gather_lsq<true> in one launch, without the edge
// buffer.  The inner iteration is linear in dq, and the adjoint a point receives over an
// edge it is the neighbour of depends only on the owner's solve adjoint (lam), the two
// ends' records and the offset, so each point gathers it over the transposed
// connectivity instead of the owner writing it out.  Blocks of STAGE_P points stage
// their PullHalo (own points, neighbours and owners: q, dq_in, xy and lam) in LDS; then
// every point walks its own slots (owner terms) and its transposed entries (neighbour
// terms), both in the push form's orders, and its first lane writes dqb_in, adds into
// qb and xyb and, for the next reversed inner iteration (lam_next non-null), writes its
// solve adjoint from that dqb_in (dq_in being the next iteration's output).  Entries past
// the LDS capacity (255) are read from global memory.
template <int LPP, bool POWER0>
__global__ __launch_bounds__(BLOCK) void k_inner_pull(Geom g, PullHalo ph, Trans tr, const V4d* __restrict__ q,
                                                      const V4d* __restrict__ dq_in, const V4d* __restrict__ lam,
                                                      V4d* __restrict__ dqb_in, V4d* __restrict__ lam_next,
                                                      V4d* __restrict__ qb, D2* __restrict__ xyb, double power) {
  const PointMap<LPP, true> pm(g.n);
  const int i = pm.live ? pm.i : g.n - 1, lane = pm.lane;
  extern __shared__ __align__(16) unsigned char smem[];
  const int H = ph.hcap;
  V4d* s_q = (V4d*)smem;
  V4d* s_dx = s_q + H;
  V4d* s_dy = s_dx + H;
  V4d* s_sqb = s_dy + H;
  V4d* s_syqb = s_sqb + H;
  D2* s_xy = (D2*)(s_syqb + H);
  double* s_gxx = (double*)(s_xy + H);
  double* s_gyy = s_gxx + H;
  double* s_gxy = s_gyy + H;
  uint8_t* s_cl = (uint8_t*)(s_gxy + H);
  const int b = pm.b;
  const int h0 = ph.ptr[b], hn = ph.ptr[b + 1] - h0;
  const int hl = hn < H ? hn : H;
  for (int j = threadIdx.x; j < hl; j += STAGE_P * LPP) {
    const int k = ph.ids[h0 + j];
    s_q[j] = q[k];
    s_dx[j] = dq_in[2 * k];
    s_dy[j] = dq_in[2 * k + 1];
    s_xy[j] = g.xy[k];
    s_sqb[j] = lam[3 * k];
    s_syqb[j] = lam[3 * k + 1];
    const V4d gb = lam[3 * k + 2];
    s_gxx[j] = gb.x;
    s_gyy[j] = gb.y;
    s_gxy[j] = gb.z;
  }
  // the block's slot words: per slot a STAGE_P-byte row segment in 16-byte loads
  constexpr int V = STAGE_P / 16;
  const uint4* src = (const uint4*)(ph.loc8 + (size_t)b * STAGE_P);
  for (int j = threadIdx.x; j < g.width * V; j += STAGE_P * LPP) {
    const int s = j / V;
    ((uint4*)s_cl)[j] = src[(size_t)s * (ph.stride / 16) + (j - s * V)];
  }
  __syncthreads();
  if (!pm.live) return;
  const int li = (int)threadIdx.x / LPP;  // the block's own points lead its halo
  const D2 pi = s_xy[li];
  const V4d qi = s_q[li], dxi = s_dx[li], dyi = s_dy[li];
  auto lam_lds = [&](int j, InnerLam& L) {
    const V4d a = s_sqb[j], c = s_syqb[j];
    for (int r = 0; r < 4; ++r) {
      L.sqb[r] = a[r];
      L.syqb[r] = c[r];
    }
    L.sxxb = s_gxx[j];
    L.syyb = s_gyy[j];
    L.sxyb = s_gxy[j];
  };
  auto lam_glb = [&](int k, InnerLam& L) {
    const V4d a = lam[3 * k], c = lam[3 * k + 1], gb = lam[3 * k + 2];
    for (int r = 0; r < 4; ++r) {
      L.sqb[r] = a[r];
      L.syqb[r] = c[r];
    }
    L.sxxb = gb.x;
    L.syyb = gb.y;
    L.sxyb = gb.z;
  };
  double aq[4] = {0, 0, 0, 0}, ax[4] = {0, 0, 0, 0}, ay[4] = {0, 0, 0, 0};
  double xb = 0.0, yb = 0.0;
  // owner terms: the point's own slots (k_inner_rev's walk)
  {
    InnerLam Li;
    lam_lds(li, Li);
    const int nb = g.nbhs[i];
    // the next slot's halo position is read while this edge is evaluated
    int jn = lane < nb ? s_cl[lane * STAGE_P + li] : 0;
    for (int s = lane; s < nb; s += LPP) {
      const int j = jn;
      if (s + LPP < nb) jn = s_cl[(s + LPP) * STAGE_P + li];
      D2 pk;
      V4d qk, dxk, dyk;
      if (j < H) {
        pk = s_xy[j];
        qk = s_q[j];
        dxk = s_dx[j];
        dyk = s_dy[j];
      } else {
        const int k = (int)(g.conn[(size_t)s * g.n + i] >> 4);
        pk = g.xy[k];
        qk = q[k];
        dxk = dq_in[2 * k];
        dyk = dq_in[2 * k + 1];
      }
      const double dx = pk.x - pi.x, dy = pk.y - pi.y;
      double Db[4], dxb, dyb;
      inner_edge_rev<POWER0>(dx, dy, power, qi, dxi, dyi, qk, dxk, dyk, Li, Db, dxb, dyb);
      for (int r = 0; r < 4; ++r) {
        aq[r] -= Db[r];
        ax[r] += 0.5 * dx * Db[r];
        ay[r] += 0.5 * dy * Db[r];
      }
      xb -= dxb;
      yb -= dyb;
    }
  }
  // neighbour terms: the edges that name the point (k_gather_lsq's walk)
  const int e0 = tr.ptr[i], e1 = tr.ptr[i + 1];
  // the next entry's halo position (a global byte) is loaded while this edge is evaluated
  int jn = e0 + lane < e1 ? ph.tloc8[e0 + lane] : 0;
  for (int e = e0 + lane; e < e1; e += LPP) {
    const int j = jn;
    if (e + LPP < e1) jn = ph.tloc8[e + LPP];
    D2 po;
    V4d qo, dxo, dyo;
    InnerLam Lo;
    if (j < H) {
      po = s_xy[j];
      qo = s_q[j];
      dxo = s_dx[j];
      dyo = s_dy[j];
      lam_lds(j, Lo);
    } else {
      const int o = tr.own[e];
      po = g.xy[o];
      qo = q[o];
      dxo = dq_in[2 * o];
      dyo = dq_in[2 * o + 1];
      lam_glb(o, Lo);
    }
    const double dx = pi.x - po.x, dy = pi.y - po.y;
    double Db[4], dxb, dyb;
    inner_edge_rev<POWER0>(dx, dy, power, qo, dxo, dyo, qi, dxi, dyi, Lo, Db, dxb, dyb);
    for (int r = 0; r < 4; ++r) {
      aq[r] += Db[r];
      ax[r] += -0.5 * dx * Db[r];
      ay[r] += -0.5 * dy * Db[r];
    }
    xb += dxb;
    yb += dyb;
  }
  if (LPP > 1) {
    for (int r = 0; r < 4; ++r) {
      aq[r] = group_sum<LPP>(aq[r]);
      ax[r] = group_sum<LPP>(ax[r]);
      ay[r] = group_sum<LPP>(ay[r]);
    }
    xb = group_sum<LPP>(xb);
    yb = group_sum<LPP>(yb);
    if (lane != 0) return;
  }
  const V4d a = qb[i];
  qb[i] = V4d{a.x + aq[0], a.y + aq[1], a.z + aq[2], a.w + aq[3]};
  const V4d bx{ax[0], ax[1], ax[2], ax[3]}, by{ay[0], ay[1], ay[2], ay[3]};
  dqb_in[2 * i] = bx;
  dqb_in[2 * i + 1] = by;
  const D2 pxy = xyb[i];
  xyb[i] = D2{pxy.x + xb, pxy.y + yb};
  if (lam_next) {
    InnerLam L;
    inner_lam(g.lsq_geo[i], dxi, dyi, bx, by, L);
    lam_store(lam_next, i, L);
  }
}

// reverse of k_lsq (eval_q_derivatives): dq and qm adjoints -> q, xy adjoints.  Own
// contributions added to qb, xyb; per edge, the neighbour's q adjoint (with the qm
// adjoint of the components it maximises or minimises) and the edge's (delx, dely)
// adjoint go to the edge buffer (fields 0-5, k_gather_lsq)
template <int LPP, bool STAGED, bool POWER0>
__global__ __launch_bounds__(BLOCK) void k_lsq_rev(Geom g, const V4d* __restrict__ q,
                                                   const V4d* __restrict__ dq,
                                                   const V4d* __restrict__ dqb,
                                                   const V4d* __restrict__ qmb, V4d* __restrict__ qb,
                                                   D2* __restrict__ xyb, double* __restrict__ eb,
                                                   const int32_t* __restrict__ epos, double power) {
  const PointMap<LPP, STAGED> pm(g.n);
  if (!STAGED && !pm.live) return;
  const int i = pm.live ? pm.i : g.n - 1, lane = pm.lane;
  const D2 pi = g.xy[i];
  const V4d gg = g.lsq_geo[i];
  const double sxx = gg.x, syy = gg.y, sxy = gg.z, obd = gg.w;
  const V4d tx = dq[2 * i], ty = dq[2 * i + 1];
  const V4d bx = dqb[2 * i], by = dqb[2 * i + 1];
  double sqb[4], syqb[4];
  double sxxb = 0.0, syyb = 0.0, sxyb = 0.0, obdb = 0.0;
  for (int r = 0; r < 4; ++r) {
    const double sxq = sxx * tx[r] + sxy * ty[r];
    const double syq = sxy * tx[r] + syy * ty[r];
    sqb[r] = (bx[r] * syy - by[r] * sxy) * obd;
    syqb[r] = (by[r] * sxx - bx[r] * sxy) * obd;
    obdb += bx[r] * (sxq * syy - syq * sxy) + by[r] * (syq * sxx - sxq * sxy);
    syyb += bx[r] * sxq * obd;
    sxxb += by[r] * syq * obd;
    sxyb += -(bx[r] * syq + by[r] * sxq) * obd;
  }
  const double detb = -obdb * obd * obd;
  sxxb += detb * syy;
  syyb += detb * sxx;
  sxyb += -2.0 * detb * sxy;
  const V4d qi = q[i];
  const V4d mxb = qmb[2 * i], mnb = qmb[2 * i + 1];
  // qm (q_variables.F90:68-82): the max/min over {i} + stencil, strict > and <, the
  // adjoint going to the selected point (slot -1: the point itself)
  V4d qmax = qi, qmin = qi;
  int smax[4] = {-1, -1, -1, -1}, smin[4] = {-1, -1, -1, -1};
  double qib[4] = {0, 0, 0, 0};
  double xib = 0.0, yib = 0.0;
  const int nb = pm.live ? g.nbhs[i] : 0;
  const InnerLam L{{sqb[0], sqb[1], sqb[2], sqb[3]}, {syqb[0], syqb[1], syqb[2], syqb[3]}, sxxb, syyb, sxyb};
  // edge in slot s to the neighbour with coordinates pk and q qk
  auto edge = [&](int s, const D2& pk, const V4d& qk) {
    const double dx = pk.x - pi.x, dy = pk.y - pi.y;
    for (int r = 0; r < 4; ++r) {
      if (qk[r] > qmax[r]) {
        qmax[r] = qk[r];
        smax[r] = s;
      }
      if (qk[r] < qmin[r]) {
        qmin[r] = qk[r];
        smin[r] = s;
      }
    }
    double Dbv[4], dxb, dyb;
    lsq_edge_rev<POWER0>(dx, dy, power, qi, qk, L, Dbv, dxb, dyb);
    for (int r = 0; r < 4; ++r) qib[r] -= Dbv[r];
    D2* ep = reinterpret_cast<D2*>(eb + (size_t)epos[(size_t)s * g.n + i] * EF_LSQ);
    ep[0] = D2{Dbv[0], Dbv[1]};
    ep[1] = D2{Dbv[2], Dbv[3]};
    ep[2] = D2{dxb, dyb};
    xib -= dxb;
    yib -= dyb;
  };
  if constexpr (STAGED) {
    // the block's halo (q, xy) in LDS
    extern __shared__ __align__(16) unsigned char smem[];
    const int H = g.hl[0].hcap;
    V4d* s_q = (V4d*)smem;
    D2* s_xy = (D2*)(s_q + H);
    uint8_t* s_cl = (uint8_t*)(s_xy + H);
    stage_and_walk<STAGE_P, LPP>(
        g, g.hl[0], pm.b, nb, s_cl,
        [&](int j, int k) {
          s_q[j] = q[k];
          s_xy[j] = g.xy[k];
        },
        [&](int j, int s) { edge(s, s_xy[j], s_q[j]); },
        [&](int k, int s) { edge(s, g.xy[k], q[k]); });
    if (!pm.live) return;
  } else {
    for (int s = lane; s < nb; s += LPP) {
      const int k = (int)(g.conn[(size_t)s * g.n + i] >> 4);
      edge(s, g.xy[k], q[k]);
    }
  }
  if (LPP > 1) {
    // the point's extremes over its lanes: the larger (smaller) value, on ties the
    // lower slot, the point itself (slot -1) first, as the sequential strict > and <
    for (int r = 0; r < 4; ++r) {
      for (int off = LPP / 2; off > 0; off >>= 1) {
        const double ov = __shfl_xor(qmax[r], off, 64);
        const int os = __shfl_xor(smax[r], off, 64);
        if (ov > qmax[r] || (ov == qmax[r] && os < smax[r])) {
          qmax[r] = ov;
          smax[r] = os;
        }
        const double nv = __shfl_xor(qmin[r], off, 64);
        const int ns = __shfl_xor(smin[r], off, 64);
        if (nv < qmin[r] || (nv == qmin[r] && ns < smin[r])) {
          qmin[r] = nv;
          smin[r] = ns;
        }
      }
    }
  }
  // the selected neighbour's entry is updated by the lane that wrote it
  for (int r = 0; r < 4; ++r) {
    if (smax[r] < 0) {
      if (lane == 0) qib[r] += mxb[r];
    } else if (smax[r] % LPP == lane) {
      eb[(size_t)epos[(size_t)smax[r] * g.n + i] * EF_LSQ + r] += mxb[r];
    }
    if (smin[r] < 0) {
      if (lane == 0) qib[r] += mnb[r];
    } else if (smin[r] % LPP == lane) {
      eb[(size_t)epos[(size_t)smin[r] * g.n + i] * EF_LSQ + r] += mnb[r];
    }
  }
  if (LPP > 1) {
    for (int r = 0; r < 4; ++r) qib[r] = group_sum<LPP>(qib[r]);
    xib = group_sum<LPP>(xib);
    yib = group_sum<LPP>(yib);
    if (lane != 0) return;
  }
  const V4d a = qb[i];
  qb[i] = V4d{a.x + qib[0], a.y + qib[1], a.z + qib[2], a.w + qib[3]};
  const D2 pxy = xyb[i];
  xyb[i] = D2{pxy.x + xib, pxy.y + yib};
}

// LDS of the pull-form LSQ reverse per halo entry: q, sqb, syqb (V4d each), xy, and
// sxxb, syyb, sxyb; then the block's slot words
inline size_t stage_lds_lsq_pull(int hcap, int width) {
  return (size_t)hcap * (3 * sizeof(V4d) + sizeof(D2) + 3 * sizeof(double)) + (size_t)width * STAGE_P;
}

// Pull form of k_lsq_rev + k_gather_lsq<false>: as k_inner_pull, each point adds the
// owner terms of its own slots and gathers the neighbour terms of the edges that name it
// (D = q_k - q_i, so no dq records), from the PullHalo staged with q, xy and the solve
// adjoints (lam).  The qm adjoints (q_variables.F90:68-82: the max and min over the
// point and its stencil, strict > and <, the adjoint going to the selected point) are
// resolved by the owner's walk as in k_lsq_rev; those of a selected neighbour, at most 8
// per point, are added atomically into qbx (summed into the stage's q adjoint and
// cleared by k_qvar_rev), the others into the point's own q adjoint.
template <int LPP, bool POWER0>
__global__ __launch_bounds__(BLOCK) void k_lsq_pull(Geom g, PullHalo ph, Trans tr, const V4d* __restrict__ q,
                                                    const V4d* __restrict__ lam, const V4d* __restrict__ qmb,
                                                    V4d* __restrict__ qb, V4d* __restrict__ qbx,
                                                    D2* __restrict__ xyb, double power) {
  const PointMap<LPP, true> pm(g.n);
  const int i = pm.live ? pm.i : g.n - 1, lane = pm.lane;
  extern __shared__ __align__(16) unsigned char smem[];
  const int H = ph.hcap;
  V4d* s_q = (V4d*)smem;
  V4d* s_sqb = s_q + H;
  V4d* s_syqb = s_sqb + H;
  D2* s_xy = (D2*)(s_syqb + H);
  double* s_gxx = (double*)(s_xy + H);
  double* s_gyy = s_gxx + H;
  double* s_gxy = s_gyy + H;
  uint8_t* s_cl = (uint8_t*)(s_gxy + H);
  const int b = pm.b;
  const int h0 = ph.ptr[b], hn = ph.ptr[b + 1] - h0;
  const int hl = hn < H ? hn : H;
  for (int j = threadIdx.x; j < hl; j += STAGE_P * LPP) {
    const int k = ph.ids[h0 + j];
    s_q[j] = q[k];
    s_xy[j] = g.xy[k];
    s_sqb[j] = lam[3 * k];
    s_syqb[j] = lam[3 * k + 1];
    const V4d gb = lam[3 * k + 2];
    s_gxx[j] = gb.x;
    s_gyy[j] = gb.y;
    s_gxy[j] = gb.z;
  }
  constexpr int V = STAGE_P / 16;
  const uint4* src = (const uint4*)(ph.loc8 + (size_t)b * STAGE_P);
  for (int j = threadIdx.x; j < g.width * V; j += STAGE_P * LPP) {
    const int s = j / V;
    ((uint4*)s_cl)[j] = src[(size_t)s * (ph.stride / 16) + (j - s * V)];
  }
  __syncthreads();
  if (!pm.live) return;
  const int li = (int)threadIdx.x / LPP;  // the block's own points lead its halo
  const D2 pi = s_xy[li];
  const V4d qi = s_q[li];
  auto lam_lds = [&](int j, InnerLam& L) {
    const V4d a = s_sqb[j], c = s_syqb[j];
    for (int r = 0; r < 4; ++r) {
      L.sqb[r] = a[r];
      L.syqb[r] = c[r];
    }
    L.sxxb = s_gxx[j];
    L.syyb = s_gyy[j];
    L.sxyb = s_gxy[j];
  };
  auto lam_glb = [&](int k, InnerLam& L) {
    const V4d a = lam[3 * k], c = lam[3 * k + 1], gb = lam[3 * k + 2];
    for (int r = 0; r < 4; ++r) {
      L.sqb[r] = a[r];
      L.syqb[r] = c[r];
    }
    L.sxxb = gb.x;
    L.syyb = gb.y;
    L.sxyb = gb.z;
  };
  double aq[4] = {0, 0, 0, 0};
  double xb = 0.0, yb = 0.0;
  V4d qmax = qi, qmin = qi;
  int smax[4] = {-1, -1, -1, -1}, smin[4] = {-1, -1, -1, -1};
  const int nb = g.nbhs[i];
  // owner terms and the qm extremes: the point's own slots (k_lsq_rev's walk)
  {
    InnerLam Li;
    lam_lds(li, Li);
    int jn = lane < nb ? s_cl[lane * STAGE_P + li] : 0;
    for (int s = lane; s < nb; s += LPP) {
      const int j = jn;
      if (s + LPP < nb) jn = s_cl[(s + LPP) * STAGE_P + li];
      D2 pk;
      V4d qk;
      if (j < H) {
        pk = s_xy[j];
        qk = s_q[j];
      } else {
        const int k = (int)(g.conn[(size_t)s * g.n + i] >> 4);
        pk = g.xy[k];
        qk = q[k];
      }
      for (int r = 0; r < 4; ++r) {
        if (qk[r] > qmax[r]) {
          qmax[r] = qk[r];
          smax[r] = s;
        }
        if (qk[r] < qmin[r]) {
          qmin[r] = qk[r];
          smin[r] = s;
        }
      }
      double Db[4], dxb, dyb;
      lsq_edge_rev<POWER0>(pk.x - pi.x, pk.y - pi.y, power, qi, qk, Li, Db, dxb, dyb);
      for (int r = 0; r < 4; ++r) aq[r] -= Db[r];
      xb -= dxb;
      yb -= dyb;
    }
  }
  // neighbour terms: the edges that name the point (k_gather_lsq's walk)
  const int e0 = tr.ptr[i], e1 = tr.ptr[i + 1];
  int jn = e0 + lane < e1 ? ph.tloc8[e0 + lane] : 0;
  for (int e = e0 + lane; e < e1; e += LPP) {
    const int j = jn;
    if (e + LPP < e1) jn = ph.tloc8[e + LPP];
    D2 po;
    V4d qo;
    InnerLam Lo;
    if (j < H) {
      po = s_xy[j];
      qo = s_q[j];
      lam_lds(j, Lo);
    } else {
      const int o = tr.own[e];
      po = g.xy[o];
      qo = q[o];
      lam_glb(o, Lo);
    }
    double Db[4], dxb, dyb;
    lsq_edge_rev<POWER0>(pi.x - po.x, pi.y - po.y, power, qo, qi, Lo, Db, dxb, dyb);
    for (int r = 0; r < 4; ++r) aq[r] += Db[r];
    xb += dxb;
    yb += dyb;
  }
  if (LPP > 1) {
    // the point's extremes over its lanes: the larger (smaller) value, on ties the
    // lower slot, the point itself (slot -1) first, as the sequential strict > and <
    for (int r = 0; r < 4; ++r) {
      for (int off = LPP / 2; off > 0; off >>= 1) {
        const double ov = __shfl_xor(qmax[r], off, 64);
        const int os = __shfl_xor(smax[r], off, 64);
        if (ov > qmax[r] || (ov == qmax[r] && os < smax[r])) {
          qmax[r] = ov;
          smax[r] = os;
        }
        const double nv = __shfl_xor(qmin[r], off, 64);
        const int ns = __shfl_xor(smin[r], off, 64);
        if (nv < qmin[r] || (nv == qmin[r] && ns < smin[r])) {
          qmin[r] = nv;
          smin[r] = ns;
        }
      }
    }
  }
  // qm adjoints: to the point itself (its first lane), or by the lane that walked the
  // selected slot to that neighbour's qbx
  const V4d mxb = qmb[2 * i], mnb = qmb[2 * i + 1];
  for (int r = 0; r < 4; ++r) {
    if (smax[r] < 0) {
      if (lane == 0) aq[r] += mxb[r];
    } else if (smax[r] % LPP == lane) {
      aadd(&qbx[(int)(g.conn[(size_t)smax[r] * g.n + i] >> 4)].x + r, mxb[r]);
    }
    if (smin[r] < 0) {
      if (lane == 0) aq[r] += mnb[r];
    } else if (smin[r] % LPP == lane) {
      aadd(&qbx[(int)(g.conn[(size_t)smin[r] * g.n + i] >> 4)].x + r, mnb[r]);
    }
  }
  if (LPP > 1) {
    for (int r = 0; r < 4; ++r) aq[r] = group_sum<LPP>(aq[r]);
    xb = group_sum<LPP>(xb);
    yb = group_sum<LPP>(yb);
    if (lane != 0) return;
  }
  const V4d a = qb[i];
  qb[i] = V4d{a.x + aq[0], a.y + aq[1], a.z + aq[2], a.w + aq[3]};
  const D2 pxy = xyb[i];
  xyb[i] = D2{pxy.x + xb, pxy.y + yb};
}

// Gathers of the edge buffer into the neighbour's adjoints, over the transposed
// connectivity (each point reads the edges that name it; no atomics)
template <int LPP>
// With lam non-null it also writes each point's solve adjoint (k_inner_lam's) from its
// completed dq adjoint and the stage's last dq (dq_last), for a pull-form reverse next.
__global__ __launch_bounds__(BLOCK) void k_gather_flux(Geom g, Trans tr, const double* __restrict__ eb,
                                                       const uint8_t* __restrict__ esel,
                                                       V4d* __restrict__ qb, V4d* __restrict__ dqb,
                                                       V4d* __restrict__ qmb, D2* __restrict__ xyb,
                                                       const V4d* __restrict__ dq_last, V4d* __restrict__ lam) {
  const int t = point_id();
  const int k = t / LPP, lane = t % LPP;
  if (k >= g.n) return;
  const D2 pk = g.xy[k];
  double aq[4] = {0, 0, 0, 0}, ax[4] = {0, 0, 0, 0}, ay[4] = {0, 0, 0, 0};
  double amx[4] = {0, 0, 0, 0}, amn[4] = {0, 0, 0, 0};
  double xb = 0.0, yb = 0.0;
  const int j1 = tr.ptr[k + 1];
  for (int j = tr.ptr[k] + lane; j < j1; j += LPP) {
    const D2* ep = reinterpret_cast<const D2*>(eb + (size_t)j * EF_FLUX);
    const D2 q01 = ep[0], q23 = ep[1], w = ep[6];
    aq[0] += q01.x;
    aq[1] += q01.y;
    aq[2] += q23.x;
    aq[3] += q23.y;
    xb += w.x;
    yb += w.y;
    if (!g.first_order) {
      const D2 pi = g.xy[tr.own[j]];
      const double delx = pk.x - pi.x, dely = pk.y - pi.y;
      const D2 d01 = ep[2], d23 = ep[3];
      const double d[4] = {d01.x, d01.y, d23.x, d23.y};
      for (int r = 0; r < 4; ++r) {
        ax[r] += d[r] * delx;
        ay[r] += d[r] * dely;
      }
      const D2 m01 = ep[4], m23 = ep[5];
      const double m[4] = {m01.x, m01.y, m23.x, m23.y};
      const unsigned sel = esel[j];
      for (int r = 0; r < 4; ++r) {
        const bool mx = (sel >> r) & 1u;
        amx[r] += mx ? m[r] : 0.0;
        amn[r] += mx ? 0.0 : m[r];
      }
    }
  }
  if (LPP > 1) {
    for (int r = 0; r < 4; ++r) {
      aq[r] = group_sum<LPP>(aq[r]);
      ax[r] = group_sum<LPP>(ax[r]);
      ay[r] = group_sum<LPP>(ay[r]);
      amx[r] = group_sum<LPP>(amx[r]);
      amn[r] = group_sum<LPP>(amn[r]);
    }
    xb = group_sum<LPP>(xb);
    yb = group_sum<LPP>(yb);
    if (lane != 0) return;
  }
  V4d t4 = qb[k];
  qb[k] = V4d{t4.x + aq[0], t4.y + aq[1], t4.z + aq[2], t4.w + aq[3]};
  V4d bx = dqb[2 * k], by = dqb[2 * k + 1];
  if (!g.first_order) {
    bx = V4d{bx.x + ax[0], bx.y + ax[1], bx.z + ax[2], bx.w + ax[3]};
    by = V4d{by.x + ay[0], by.y + ay[1], by.z + ay[2], by.w + ay[3]};
    dqb[2 * k] = bx;
    dqb[2 * k + 1] = by;
    V4d t = qmb[2 * k];
    qmb[2 * k] = V4d{t.x + amx[0], t.y + amx[1], t.z + amx[2], t.w + amx[3]};
    t = qmb[2 * k + 1];
    qmb[2 * k + 1] = V4d{t.x + amn[0], t.y + amn[1], t.z + amn[2], t.w + amn[3]};
  }
  const D2 p = xyb[k];
  xyb[k] = D2{p.x + xb, p.y + yb};
  if (lam) {
    InnerLam L;
    inner_lam(g.lsq_geo[k], dq_last[2 * k], dq_last[2 * k + 1], bx, by, L);
    lam_store(lam, k, L);
  }
}

// WITH_DQ: the inner-loop payload (dq adjoint = -0.5 (delx, dely) q adjoint)
template <bool WITH_DQ, int LPP>
__global__ __launch_bounds__(BLOCK) void k_gather_lsq(Geom g, Trans tr, const double* __restrict__ eb,
                                                      V4d* __restrict__ qb, V4d* __restrict__ dqb,
                                                      D2* __restrict__ xyb) {
  const int t = point_id();
  const int k = t / LPP, lane = t % LPP;
  if (k >= g.n) return;
  const D2 pk = g.xy[k];
  double aq[4] = {0, 0, 0, 0}, ax[4] = {0, 0, 0, 0}, ay[4] = {0, 0, 0, 0};
  double xb = 0.0, yb = 0.0;
  const int j1 = tr.ptr[k + 1];
  for (int j = tr.ptr[k] + lane; j < j1; j += LPP) {
    const D2* ep = reinterpret_cast<const D2*>(eb + (size_t)j * EF_LSQ);
    const D2 a = ep[0], b = ep[1], w = ep[2];
    const double v[4] = {a.x, a.y, b.x, b.y};
    for (int r = 0; r < 4; ++r) aq[r] += v[r];
    xb += w.x;
    yb += w.y;
    if (WITH_DQ) {
      const D2 pi = g.xy[tr.own[j]];
      const double delx = pk.x - pi.x, dely = pk.y - pi.y;
      for (int r = 0; r < 4; ++r) {
        ax[r] += -0.5 * delx * v[r];
        ay[r] += -0.5 * dely * v[r];
      }
    }
  }
  if (LPP > 1) {
    for (int r = 0; r < 4; ++r) {
      aq[r] = group_sum<LPP>(aq[r]);
      if (WITH_DQ) {
        ax[r] = group_sum<LPP>(ax[r]);
        ay[r] = group_sum<LPP>(ay[r]);
      }
    }
    xb = group_sum<LPP>(xb);
    yb = group_sum<LPP>(yb);
    if (lane != 0) return;
  }
  V4d t4 = qb[k];
  qb[k] = V4d{t4.x + aq[0], t4.y + aq[1], t4.z + aq[2], t4.w + aq[3]};
  if (WITH_DQ) {
    V4d t = dqb[2 * k];
    dqb[2 * k] = V4d{t.x + ax[0], t.y + ax[1], t.z + ax[2], t.w + ax[3]};
    t = dqb[2 * k + 1];
    dqb[2 * k + 1] = V4d{t.x + ay[0], t.y + ay[1], t.z + ay[2], t.w + ay[3]};
  }
  const D2 p = xyb[k];
  xyb[k] = D2{p.x + xb, p.y + yb};
}

// reverse of eval_q_variables: pb += dq/dprim^T qb (own point)
// qbx (non-null after the pull-form LSQ reverse): the qm adjoints it delivered, added
// to qb here and cleared for the next stage
__global__ __launch_bounds__(BLOCK) void k_qvar_rev(int n, const V4d* __restrict__ prim,
                                                    const V4d* __restrict__ qb, V4d* __restrict__ qbx,
                                                    V4d* __restrict__ pb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  V4d b = qb[i];
  if (qbx) {
    const V4d x = qbx[i];
    b = V4d{b.x + x.x, b.y + x.y, b.z + x.z, b.w + x.w};
    qbx[i] = V4d{0.0, 0.0, 0.0, 0.0};
  }
  double pv[4] = {0, 0, 0, 0};
  const double qbv[4] = {b.x, b.y, b.z, b.w};
  q_of_prim_rev(prim[i], qbv, pv);
  const V4d t = pb[i];
  pb[i] = V4d{t.x + pv[0], t.y + pv[1], t.z + pv[2], t.w + pv[3]};
}

// reverse of func_delta (state_update.F90:190-236): delta_b -> the minimising
// neighbour's prim and the edge's coordinates
__global__ __launch_bounds__(BLOCK) void k_delta_rev(Geom g, const V4d* __restrict__ prim,
                                                     const double* __restrict__ deltab, double cfl,
                                                     V4d* __restrict__ pb, D2* __restrict__ xyb) {
  const int i = point_id();
  if (i >= g.n) return;
  const double b = deltab[i];
  if (b == 0.0) return;
  const D2 pi = g.xy[i];
  double min_delt = 1.0, dist_s = 0.0, c_s = 0.0, dx_s = 0.0, dy_s = 0.0;
  int ks = -1;
  const int nb = g.nbhs[i];
  for (int s = 0; s < nb; ++s) {
    const int k = (int)(g.conn[(size_t)s * g.n + i] >> 4);
    const D2 pk = g.xy[k];
    const double dx = pk.x - pi.x, dy = pk.y - pi.y;
    const double dist = sqrt(dx * dx + dy * dy);
    const double c = wave_speed(prim[k]);
    double dt = dist / c;
    dt = cfl * dt;
    if (min_delt > dt) {
      min_delt = dt;
      ks = k;
      dist_s = dist;
      c_s = c;
      dx_s = dx;
      dy_s = dy;
    }
  }
  if (ks < 0) return;
  const double distb = b * cfl / c_s;
  const double cb = -b * cfl * dist_s / (c_s * c_s);
  double pv[4] = {0, 0, 0, 0};
  wave_speed_rev(prim[ks], cb, pv);
  aadd4(&pb[ks], pv);
  const double dxb = distb * dx_s / dist_s, dyb = distb * dy_s / dist_s;
  aadd2(&xyb[ks], dxb, dyb);
  aadd2(&xyb[i], -dxb, -dyb);
}

// reverse of compute_cl_cd_cm for J = sum over shapes of Cd (compute_force_coeffs_b.F90)
__global__ __launch_bounds__(BLOCK) void k_forces_rev(const V4d* __restrict__ prim,
                                                      const int32_t* __restrict__ sh_m,
                                                      const int32_t* __restrict__ sh_id,
                                                      const int32_t* __restrict__ sh_l,
                                                      const int32_t* __restrict__ sh_r,
                                                      const double* __restrict__ sgeo, int nsh,
                                                      const double* __restrict__ shape_par,
                                                      const D2* __restrict__ xy, V4d* __restrict__ pb,
                                                      D2* __restrict__ xyb, D2* __restrict__ nrmb) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nsh) return;
  const double* sp = shape_par + 3 * sh_id[j];
  const double cp_scale = sp[0], ct = sp[1], st = sp[2];
  const double pr_inf = 1.0 / 1.4;
  const int m = sh_m[j], l = sh_l[j], r = sh_r[j];
  const double nx = sgeo[5 * j], ny = sgeo[5 * j + 1], ds = sgeo[5 * j + 2];
  const double cp = -(prim[m].w - pr_inf) / cp_scale;
  const double Hb = ct, Vb = st;  // Cd = H cos(theta) + V sin(theta)
  const double cpb = Hb * nx * ds + Vb * ny * ds;
  nrmb[m].x += Hb * cp * ds;
  nrmb[m].y += Vb * cp * ds;
  const double dsb = Hb * cp * nx + Vb * cp * ny;
  pb[m].w += -cpb / cp_scale;
  // ds = 0.5 (|m - l| + |r - m|)
  const D2 pm_ = xy[m], pl = xy[l], pr = xy[r];
  const double d1x = pm_.x - pl.x, d1y = pm_.y - pl.y, d2x = pr.x - pm_.x, d2y = pr.y - pm_.y;
  const double ds1 = sqrt(d1x * d1x + d1y * d1y), ds2 = sqrt(d2x * d2x + d2y * d2y);
  const double a1 = 0.5 * dsb / ds1, a2 = 0.5 * dsb / ds2;
  aadd2(&xyb[m], a1 * d1x - a2 * d2x, a1 * d1y - a2 * d2y);
  aadd2(&xyb[l], -a1 * d1x, -a1 * d1y);
  aadd2(&xyb[r], a2 * d2x, a2 * d2y);
}

// reverse of compute_normals (point_normals.F90:8-105) for wall and outer points
__global__ __launch_bounds__(BLOCK) void k_normals_rev(int nb_pts, const int32_t* __restrict__ bm,
                                                       const int32_t* __restrict__ bl,
                                                       const int32_t* __restrict__ br,
                                                       const D2* __restrict__ xy,
                                                       const D2* __restrict__ nrmb,
                                                       D2* __restrict__ xyb) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nb_pts) return;
  const int m = bm[j], l = bl[j], r = br[j];
  const D2 pm_ = xy[m], pl = xy[l], pr = xy[r];
  const double vx = 0.5 * ((pm_.y - pl.y) + (pr.y - pm_.y));
  const double vy = 0.5 * ((pm_.x - pl.x) + (pr.x - pm_.x));
  const double det = sqrt(vx * vx + vy * vy);
  const double nx = -vx / det, ny = vy / det;
  const D2 b = nrmb[m];
  const double detb = -(b.x * nx + b.y * ny) / det;
  const double vxb = -b.x / det + detb * vx / det;
  const double vyb = b.y / det + detb * vy / det;
  aadd2(&xyb[l], -0.5 * vyb, -0.5 * vxb);
  aadd2(&xyb[r], 0.5 * vyb, 0.5 * vxb);
}

__global__ void k_axpy(double* __restrict__ y, const double* __restrict__ x, int64_t n) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x)
    y[j] += x[j];
}

// FP32 records into FP64 (the mixed path's tape, read by the FP64 reverse kernels)
__global__ void k_widen(const V4<float>* __restrict__ a, V4d* __restrict__ b, int64_t n) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    const V4<float> v = a[j];
    b[j] = V4d{(double)v.x, (double)v.y, (double)v.z, (double)v.w};
  }
}

// Several arrays widened in one launch (blockIdx.y picks the array): the taped
// iteration's records are about 20 arrays, each too short to fill the GPU alone
constexpr int kWidenMax = 32;
struct WidenJobs {
  const V4<float>* src[kWidenMax];
  V4d* dst[kWidenMax];
  int64_t n[kWidenMax];
};
__global__ void k_widen_jobs(WidenJobs w) {
  const int k = blockIdx.y;
  const V4<float>* __restrict__ a = w.src[k];
  V4d* __restrict__ b = w.dst[k];
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < w.n[k]; j += (int64_t)gridDim.x * blockDim.x) {
    const V4<float> v = a[j];
    b[j] = V4d{(double)v.x, (double)v.y, (double)v.z, (double)v.w};
  }
}

__global__ void k_zero(double* __restrict__ p, int64_t n) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x)
    p[j] = 0.0;
}

#pragma clang fp contract(off)

}  // namespace mf

// =================================================================================
// Host: tape, checkpoints, reverse sweep
// =================================================================================

struct AdjState {
  bool ready = false;
  int S = 0, I = 0, snaps = 0;
  // tape of one iteration
  std::vector<V4d*> tprim, tq;  // S+1
  std::vector<V4d*> tdq;        // S*(I+1), 2N each
  std::vector<V4d*> tqm;        // S, 2N each
  V4d* tpold = nullptr;
  double* tdelta = nullptr;
  double* tcspeed = nullptr;
  double* tpartial = nullptr;
  Ctl* tctl = nullptr;
  // adjoints
  V4d *Pb = nullptr, *Pstage = nullptr, *poldb = nullptr, *frb = nullptr, *qb = nullptr;
  V4d* dqb[2] = {nullptr, nullptr};
  V4d* qmb = nullptr;
  double* deltab = nullptr;
  D2 *xyb = nullptr, *nrmb = nullptr;
  // geometry for the reverse
  V4d *hxx = nullptr, *hyy = nullptr, *hxy = nullptr;
  int32_t *sh_l = nullptr, *sh_r = nullptr;
  int32_t *bd_m = nullptr, *bd_l = nullptr, *bd_r = nullptr;
  int nbd = 0;
  // edge buffer (EF_FLUX doubles per edge, transposed order), edge positions and the
  // transposed connectivity
  double* eb = nullptr;
  uint8_t* esel = nullptr;
  int32_t *epos = nullptr, *t_ptr = nullptr, *t_own = nullptr;
  // halo of the pull-form inner reverse (built with the transposed connectivity) and
  // its solve adjoints, 3 records per point, double-buffered over the inner iterations
  PullHalo ph;
  V4d* lam[2] = {nullptr, nullptr};
  // qm adjoints the pull-form LSQ reverse delivers to selected neighbours (atomics)
  V4d* qbx = nullptr;
  // checkpoints
  std::vector<V4d*> snap;       // prim snapshots
  std::vector<int> snap_t;      // iteration index held by each slot (-1 free)
  int first_it = 1, iters = 0;
  double J = 0.0;
  int64_t fwd_steps = 0, taped = 0;
  int peak_snaps = 0;
  // progress lines on stderr (environment MFCFD_PROGRESS=1): about 20 per sweep
  int64_t progress_every = 0;
  // lanes per point of the reverse kernels, timed once per context (adj_tune)
  bool tuned = false;
  int lpp_frev = 1, lpp_lrev = 1, lpp_irev = 1, lpp_gath = 1;
  float primal_ms = 0.f;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  // Every iteration taped during the primal sweep (snaps = MF_ADJ_TAPE_ALL): the
  // reverse then re-runs nothing.  tapes[t] points into tape_buf; bind_tape() points
  // the one-iteration tape above (tprim, tq, tdq, tqm, tpold, tdelta) at iteration t,
  // whose input prim is iteration t-1's output (base.prim0 for t = 0).
  struct Tape {
    std::vector<V4d*> prim, q, dq, qm;  // prim: stage outputs 1..S; q: stages 0..S-1
    V4d* pold = nullptr;
    double* delta = nullptr;
  };
  bool tape_all = false;
  void* tape_buf = nullptr;
  size_t tape_cap = 0;
  std::vector<Tape> tapes;
  // the one-iteration tape's own buffers (restored when a checkpointed schedule runs)
  std::vector<V4d*> base_prim, base_q, base_dq, base_qm;
  V4d* base_pold = nullptr;
  double* base_delta = nullptr;
  // Mixed-precision contexts: the taped re-run runs the mixed path's FP32 kernels into
  // FP32 records (fprim, fq, fdq, fqm, fpold, laid out as tprim ... tpold), which are
  // then widened into the FP64 tape above; the reverse kernels and every adjoint sum
  // stay FP64.  With every iteration taped, the kept tapes are FP32 (half the memory)
  // and each is widened just before its reverse.
  bool mixed = false;
  std::vector<V4f*> fprim, fq, fdq, fqm;
  V4f* fpold = nullptr;
  std::vector<V4f*> fbase_prim, fbase_q, fbase_dq, fbase_qm;
  V4f* fbase_pold = nullptr;
};

// Input and output state records of the taped iteration in the context's precision
static void* tape_in(AdjState& A) { return A.mixed ? (void*)A.fprim[0] : (void*)A.tprim[0]; }
static void* tape_out(AdjState& A) { return A.mixed ? (void*)A.fprim[A.S] : (void*)A.tprim[A.S]; }

static int adj_zero(mf_ctx* c, void* p, int64_t n_doubles) {
  hipLaunchKernelGGL(k_zero, dim3(1024), dim3(BLOCK), 0, c->stream, (double*)p, n_doubles);
  HIPCHK(hipGetLastError());
  return MF_OK;
}

static void adj_widen_n(mf_ctx* c, const V4f* a, V4d* b, int64_t n) {
  hipLaunchKernelGGL(k_widen, dim3((unsigned)std::min<int64_t>((n + BLOCK - 1) / BLOCK, 4096)), dim3(BLOCK), 0,
                     c->stream, a, b, n);
}

// the FP32 records of the taped iteration into the FP64 tape, as few launches of
// k_widen_jobs as the records need.  The qm records are not widened: the flux reverse
// reads the FP32 ones, and the LSQ reverses recompute the extremes from q.
static int adj_widen(mf_ctx* c, AdjState& A) {
  const int64_t n = c->n;
  WidenJobs w{};
  int cnt = 0;
  int64_t nmax = 0;
  auto flush = [&] {
    if (cnt == 0) return;
    const dim3 gr((unsigned)std::min<int64_t>((nmax + BLOCK - 1) / BLOCK, 1024), (unsigned)cnt);
    hipLaunchKernelGGL(k_widen_jobs, gr, dim3(BLOCK), 0, c->stream, w);
    cnt = 0;
    nmax = 0;
  };
  auto add = [&](const V4f* a, V4d* b, int64_t m) {
    w.src[cnt] = a;
    w.dst[cnt] = b;
    w.n[cnt] = m;
    nmax = std::max(nmax, m);
    if (++cnt == kWidenMax) flush();
  };
  for (int s = 0; s <= A.S; ++s) add(A.fprim[s], A.tprim[s], n);
  for (int s = 0; s < A.S; ++s) add(A.fq[s], A.tq[s], n);
  for (size_t k = 0; k < A.fdq.size(); ++k) add(A.fdq[k], A.tdq[k], 2 * n);
  add(A.fpold, A.tpold, n);
  flush();
  HIPCHK(hipGetLastError());
  return MF_OK;
}

// one taped forward iteration of a mixed-precision context from A.fprim[0], by the
// mixed path's own kernels (the same launches as an iteration of the context), into
// the FP32 records; widened into the FP64 tape when `widen`
static int adj_record_mixed(mf_ctx* c, AdjState& A, bool widen) {
  // q and the wave speeds by the mixed path's own unit (compiled with contraction, as
  // the kernels that produce them inside a graph-replayed iteration)
  launch_q_variables_any<float>(env_of(c), (const V4f*)A.fprim[0], A.fq[0], A.tcspeed);
  V4<float>* sums = A.I > 0 ? c->qsum : nullptr;
  for (int s = 0; s < A.S; ++s) {
    V4f* dq0 = A.fdq[s * (A.I + 1)];
    launch_lsq_on<float>(c, (const V4f*)A.fq[s], dq0, A.fqm[s], A.tcspeed, A.tdelta, s == 0, nullptr, c->lpp_lsq,
                         sums);
    for (int j = 1; j <= A.I; ++j)
      launch_inner_on<float>(c, (const V4f*)A.fq[s], (const V4f*)A.fdq[s * (A.I + 1) + j - 1],
                             A.fdq[s * (A.I + 1) + j], nullptr, sweep_shape(c->lpp_inner), sums);
    State<float> st{A.fq[s], A.fdq[s * (A.I + 1) + A.I], A.fqm[s], (const float*)c->eps};
    launch_flux_update_on<float>(c, st, (const V4f*)A.fprim[s], A.fpold, A.fprim[s + 1], A.fq[s + 1], A.tcspeed,
                                 (const double*)A.tdelta, A.tpartial, (const Ctl*)A.tctl, s + 1, s + 1 == A.S,
                                 c->lpp_flux);
  }
  HIPCHK(hipGetLastError());
  return widen ? adj_widen(c, A) : MF_OK;
}

// one taped forward iteration from the tape's input record
static int adj_record(mf_ctx* c, AdjState& A) {
  if (A.mixed) return adj_record_mixed(c, A, true);
  const int n = c->n;
  const dim3 bl(BLOCK);
  hipLaunchKernelGGL((k_q_variables<double>), dim3(nblocks(c)), bl, 0, c->stream, n,
                     (const V4d*)A.tprim[0], A.tq[0], A.tcspeed);
  for (int s = 0; s < A.S; ++s) {
    V4d* dq0 = A.tdq[s * (A.I + 1)];
    launch_lsq_on<double>(c, (const V4d*)A.tq[s], dq0, A.tqm[s], A.tcspeed, A.tdelta, s == 0, nullptr, c->lpp_lsq);
    for (int j = 1; j <= A.I; ++j)
      launch_inner_on<double>(c, (const V4d*)A.tq[s], (const V4d*)A.tdq[s * (A.I + 1) + j - 1],
                              A.tdq[s * (A.I + 1) + j], nullptr, sweep_shape(c->lpp_inner));
    State<double> st{A.tq[s], A.tdq[s * (A.I + 1) + A.I], A.tqm[s], c->eps64};
    launch_flux_update_on<double>(c, st, (const V4d*)A.tprim[s], A.tpold, A.tprim[s + 1], A.tq[s + 1], A.tcspeed,
                                  (const double*)A.tdelta, A.tpartial, (const Ctl*)A.tctl, s + 1, s + 1 == A.S,
                                  c->lpp_flux);
  }
  HIPCHK(hipGetLastError());
  return MF_OK;
}

// f.template operator()<P0>() with P0 = (power == 0): the reverse kernels' unweighted
// specialisation
template <class F>
static void with_power0(const mf_ctx* c, F&& f) {
  if (c->p.power == 0.0) f.template operator()<true>();
  else f.template operator()<false>();
}

// launchers of the reverse kernels at a given lanes-per-point shape
template <class S>
static void launch_flux_rev_s(mf_ctx* c, AdjState& A, const State<S>& st, V4d* dqb_last, int lpp) {
  const HalfSums hs{A.hxx, A.hyy, A.hxy};
  const AdjOut o{A.qb, dqb_last, A.qmb, A.deltab, A.xyb, A.nrmb, A.eb, A.epos, A.esel};
  if (lpp & LPP_STAGED) {
    // staged: blocks of STAGE_PF points, 2, 4 or 8 lanes each
    const dim3 gr((unsigned)((c->n + STAGE_PF - 1) / STAGE_PF));
    const size_t lds = stage_lds_flux(c->g.hl[1].hcap, c->g.width, sizeof(S));
    auto go = [&]<int L>() {
      with_power0(c, [&]<bool P0>() {
        hipLaunchKernelGGL((k_flux_rev<L, true, S, P0>), gr, dim3(STAGE_PF * L), lds, c->stream, c->g, hs, st,
                           (const double*)A.tdelta, (const V4d*)A.frb, o, c->p.power);
      });
    };
    switch (lpp & ~LPP_STAGED) {
      case 2: go.template operator()<2>(); break;
      case 4: go.template operator()<4>(); break;
      default: go.template operator()<8>(); break;
    }
    return;
  }
  with_lpp(lpp, [&]<int L>() {
    with_power0(c, [&]<bool P0>() {
      hipLaunchKernelGGL((k_flux_rev<L, false, S, P0>), lane_blocks(c, L), dim3(BLOCK), 0, c->stream, c->g, hs,
                         st, (const double*)A.tdelta, (const V4d*)A.frb, o, c->p.power);
    });
  });
}

// the flux reverse of stage s: on a mixed-precision context in FP32 from the taped
// iteration's FP32 records (as the mixed primal evaluates the fluxes), else in FP64
static void launch_flux_rev(mf_ctx* c, AdjState& A, int s, V4d* dqb_last, int lpp) {
  if (A.mixed) {
    const State<float> st{A.fq[s], A.fdq[s * (A.I + 1) + A.I], A.fqm[s], (const float*)c->eps};
    launch_flux_rev_s<float>(c, A, st, dqb_last, lpp);
  } else {
    const State<double> st{A.tq[s], A.tdq[s * (A.I + 1) + A.I], A.tqm[s], c->eps64};
    launch_flux_rev_s<double>(c, A, st, dqb_last, lpp);
  }
}

// the flux gather of stage s; with_lam: also the solve adjoints of the stage's last dq
// (into A.lam[I & 1]) for the pull-form reverse that follows
static void launch_gather_flux(mf_ctx* c, AdjState& A, int s, V4d* dqb_last, int lpp, bool with_lam) {
  const Trans tr{A.t_ptr, A.t_own};
  const V4d* dq_last = (const V4d*)A.tdq[s * (A.I + 1) + A.I];
  V4d* lam = with_lam ? A.lam[A.I & 1] : nullptr;
  with_lpp(lpp, [&]<int L>() {
    hipLaunchKernelGGL(k_gather_flux<L>, lane_blocks(c, L), dim3(BLOCK), 0, c->stream, c->g, tr,
                       (const double*)A.eb, (const uint8_t*)A.esel, A.qb, dqb_last, A.qmb, A.xyb, dq_last, lam);
  });
}

// inner iteration j of stage s: dq^(j) adjoint (dqb[j & 1]) -> dq^(j-1)
static void launch_inner_rev(mf_ctx* c, AdjState& A, int s, int j, int lpp) {
  const int cur = j & 1, prv = (j - 1) & 1;
  const V4d* q = (const V4d*)A.tq[s];
  const V4d* din = (const V4d*)A.tdq[s * (A.I + 1) + j - 1];
  const V4d* dout = (const V4d*)A.tdq[s * (A.I + 1) + j];
  if (lpp & LPP_STAGED) {
    const size_t lds = stage_lds_inner_rev(c->g.hl[0].hcap, c->g.width);
    with_lpp_staged(lpp & ~LPP_STAGED, [&]<int L>() {
      with_power0(c, [&]<bool P0>() {
        hipLaunchKernelGGL((k_inner_rev<L, true, P0>), stage_grid(c->n), dim3(STAGE_P * L), lds, c->stream, c->g, q,
                           din, dout, (const V4d*)A.dqb[cur], A.dqb[prv], A.qb, A.xyb, A.eb, (const int32_t*)A.epos,
                           c->p.power);
      });
    });
    return;
  }
  with_lpp(lpp, [&]<int L>() {
    with_power0(c, [&]<bool P0>() {
      hipLaunchKernelGGL((k_inner_rev<L, false, P0>), lane_blocks(c, L), dim3(BLOCK), 0, c->stream, c->g, q, din,
                         dout, (const V4d*)A.dqb[cur], A.dqb[prv], A.qb, A.xyb, A.eb, (const int32_t*)A.epos,
                         c->p.power);
    });
  });
}

// inner iteration j of stage s in the pull form (MF_LPP_PULL | lanes): one launch, no
// edge buffer and no gather.  It reads the solve adjoints of iteration j (A.lam[j & 1]:
// for j = I the flux gather's, or k_inner_lam's here when need_lam; else the previous
// pull launch's) and, for j > 1, writes those of iteration j - 1.
static void launch_inner_pull(mf_ctx* c, AdjState& A, int s, int j, int lpp, bool need_lam) {
  const int cur = j & 1, prv = (j - 1) & 1;
  const V4d* q = (const V4d*)A.tq[s];
  const V4d* din = (const V4d*)A.tdq[s * (A.I + 1) + j - 1];
  if (j == A.I && need_lam) {
    const V4d* dout = (const V4d*)A.tdq[s * (A.I + 1) + j];
    hipLaunchKernelGGL(k_inner_lam, dim3(nblocks(c)), dim3(BLOCK), 0, c->stream, c->g, dout,
                       (const V4d*)A.dqb[cur], A.lam[cur]);
  }
  const Trans tr{A.t_ptr, A.t_own};
  const size_t lds = stage_lds_inner_pull(A.ph.hcap, c->g.width);
  // the next reversed step's solve adjoints: inner iteration j - 1, or the LSQ
  // derivatives (same solve) when their reverse is in the pull form too
  V4d* lam_next = (j > 1 || (A.lpp_lrev & LPP_PULL)) ? A.lam[prv] : nullptr;
  with_lpp_staged(lpp & ~LPP_PULL, [&]<int L>() {
    with_power0(c, [&]<bool P0>() {
      hipLaunchKernelGGL((k_inner_pull<L, P0>), stage_grid(c->n), dim3(STAGE_P * L), lds, c->stream, c->g, A.ph, tr,
                         q, din, (const V4d*)A.lam[cur], A.dqb[prv], lam_next, A.qb, A.xyb, c->p.power);
    });
  });
}

template <bool WITH_DQ>
static void launch_gather_lsq(mf_ctx* c, AdjState& A, V4d* dqb, int lpp) {
  const Trans tr{A.t_ptr, A.t_own};
  with_lpp(lpp, [&]<int L>() {
    hipLaunchKernelGGL((k_gather_lsq<WITH_DQ, L>), lane_blocks(c, L), dim3(BLOCK), 0, c->stream, c->g, tr,
                       (const double*)A.eb, A.qb, dqb, A.xyb);
  });
}

// the LSQ reverse of stage s in the pull form (MF_LPP_PULL | lanes).  Its solve adjoints
// (A.lam[0]) come from the last pull-form inner reverse, else (need_lam) from k_inner_lam
// here.
static void launch_lsq_pull(mf_ctx* c, AdjState& A, int s, int lpp, bool need_lam) {
  const V4d* q = (const V4d*)A.tq[s];
  if (need_lam)
    hipLaunchKernelGGL(k_inner_lam, dim3(nblocks(c)), dim3(BLOCK), 0, c->stream, c->g,
                       (const V4d*)A.tdq[s * (A.I + 1)], (const V4d*)A.dqb[0], A.lam[0]);
  const Trans tr{A.t_ptr, A.t_own};
  const size_t lds = stage_lds_lsq_pull(A.ph.hcap, c->g.width);
  with_lpp_staged(lpp & ~LPP_PULL, [&]<int L>() {
    with_power0(c, [&]<bool P0>() {
      hipLaunchKernelGGL((k_lsq_pull<L, P0>), stage_grid(c->n), dim3(STAGE_P * L), lds, c->stream, c->g, A.ph, tr, q,
                         (const V4d*)A.lam[0], (const V4d*)A.qmb, A.qb, A.qbx, A.xyb, c->p.power);
    });
  });
}

static void launch_lsq_rev(mf_ctx* c, AdjState& A, int s, int lpp) {
  const V4d* q = (const V4d*)A.tq[s];
  const V4d* dq0 = (const V4d*)A.tdq[s * (A.I + 1)];
  if (lpp & LPP_STAGED) {
    const size_t lds = stage_lds_lsq_rev(c->g.hl[0].hcap, c->g.width);
    with_lpp_staged(lpp & ~LPP_STAGED, [&]<int L>() {
      with_power0(c, [&]<bool P0>() {
        hipLaunchKernelGGL((k_lsq_rev<L, true, P0>), stage_grid(c->n), dim3(STAGE_P * L), lds, c->stream, c->g, q,
                           dq0, (const V4d*)A.dqb[0], (const V4d*)A.qmb, A.qb, A.xyb, A.eb, (const int32_t*)A.epos,
                           c->p.power);
      });
    });
    return;
  }
  with_lpp(lpp, [&]<int L>() {
    with_power0(c, [&]<bool P0>() {
      hipLaunchKernelGGL((k_lsq_rev<L, false, P0>), lane_blocks(c, L), dim3(BLOCK), 0, c->stream, c->g, q, dq0,
                         (const V4d*)A.dqb[0], (const V4d*)A.qmb, A.qb, A.xyb, A.eb, (const int32_t*)A.epos,
                         c->p.power);
    });
  });
}

// reverse of the taped iteration: A.Pb holds the adjoint of its output prim and
// receives the adjoint of its input prim
static int adj_reverse_iteration(mf_ctx* c, AdjState& A) {
  const int n = c->n;
  const dim3 gr(nblocks(c)), bl(BLOCK);
  int rc;
  if ((rc = adj_zero(c, A.poldb, 4 * (int64_t)n)) || (rc = adj_zero(c, A.deltab, n))) return rc;
  for (int s = A.S - 1; s >= 0; --s) {
    // update: Pb (adjoint of prim_{s+1}) -> Pstage (prim_s), poldb, frb, nrmb
    hipLaunchKernelGGL(k_update_rev, gr, bl, 0, c->stream, c->g, (const V4d*)A.tprim[s], (const V4d*)A.tpold,
                       (const V4d*)A.tprim[s + 1], (const V4d*)A.Pb, A.Pstage, A.poldb, A.frb, A.nrmb, s + 1,
                       c->p.euler, c->fs);
    V4d* dqb_last = A.dqb[A.I & 1];
    // flux: frb -> q, dq, qm (written), delta, xy, nrm; neighbours' parts gathered
    launch_flux_rev(c, A, s, dqb_last, A.lpp_frev);
    // the pull-form reverse that follows the gather takes its solve adjoints from it
    const bool inner_pull = A.I > 0 && (A.lpp_irev & LPP_PULL);
    const bool lsq_pull = (A.lpp_lrev & LPP_PULL) != 0;
    launch_gather_flux(c, A, s, dqb_last, A.lpp_gath, inner_pull || (A.I == 0 && lsq_pull));
    // inner iterations, last first: dq^(j) adjoint -> dq^(j-1)
    for (int j = A.I; j >= 1; --j) {
      if (A.lpp_irev & LPP_PULL) {
        launch_inner_pull(c, A, s, j, A.lpp_irev, false);
      } else {
        launch_inner_rev(c, A, s, j, A.lpp_irev);
        launch_gather_lsq<true>(c, A, A.dqb[(j - 1) & 1], A.lpp_gath);
      }
    }
    // LSQ: dq^(0), qm adjoints -> q, xy
    if (lsq_pull) {
      launch_lsq_pull(c, A, s, A.lpp_lrev, A.I > 0 && !inner_pull);
    } else {
      launch_lsq_rev(c, A, s, A.lpp_lrev);
      launch_gather_lsq<false>(c, A, (V4d*)nullptr, A.lpp_gath);
    }
    // q = q(prim_s): into the stage-input adjoint
    hipLaunchKernelGGL(k_qvar_rev, dim3(nblocks(c)), bl, 0, c->stream, n, (const V4d*)A.tprim[s],
                       (const V4d*)A.qb, lsq_pull ? A.qbx : (V4d*)nullptr, A.Pstage);
    std::swap(A.Pb, A.Pstage);
  }
  // func_delta on the iteration input, then prim_old = prim (fpi_solver.F90:20-24)
  hipLaunchKernelGGL(k_delta_rev, gr, bl, 0, c->stream, c->g, (const V4d*)A.tprim[0], (const double*)A.deltab,
                     c->p.cfl, A.Pb, A.xyb);
  hipLaunchKernelGGL(k_axpy, dim3(1024), bl, 0, c->stream, (double*)A.Pb, (const double*)A.poldb, (int64_t)4 * n);
  HIPCHK(hipGetLastError());
  return MF_OK;
}

static int adj_copy(mf_ctx* c, void* dst, const void* src);

static void adj_forget_tuning(mf_ctx* c) {
  if (c->adj) c->adj->tuned = false;
}

// Picks the lanes per point of each reverse kernel by timing every shape on a tape of
// the context's current state.  The timed launches write only buffers that the sweep
// zeroes or overwrites before use (edge buffer, stage adjoints, xy/normal adjoints).
static int adj_tune(mf_ctx* c, AdjState& A) {
  int rc;
  const int n = c->n;
  if ((rc = adj_copy(c, tape_in(A), c->prim[c->cur])) || (rc = adj_record(c, A))) return rc;
  if ((rc = adj_zero(c, A.frb, 4 * (int64_t)n)) || (rc = adj_zero(c, A.dqb[0], 8 * (int64_t)n)) ||
      (rc = adj_zero(c, A.dqb[1], 8 * (int64_t)n)) || (rc = adj_zero(c, A.qmb, 8 * (int64_t)n)) ||
      (rc = adj_zero(c, A.qb, 4 * (int64_t)n)) || (rc = adj_zero(c, A.deltab, n)) ||
      (rc = adj_zero(c, A.xyb, 2 * (int64_t)n)) || (rc = adj_zero(c, A.nrmb, 2 * (int64_t)n)))
    return rc;
  const int reps = 5, cand[4] = {1, 2, 4, 8};
  double best[4] = {1e300, 1e300, 1e300, 1e300};
  int pick[4] = {1, 1, 1, 1};
  // the LSQ reverse tries its staged form (64-point blocks, 1, 2 or 4 lanes)
  for (int l : {1, 2, 4}) {
    const int lpp = LPP_STAGED | l;
    if (!valid_rev_staged(c, 1, lpp)) continue;
    double us = 0;
    if ((rc = time_launches(c, reps, [&] { launch_lsq_rev(c, A, 0, lpp); }, &us))) return rc;
    if (us < best[1]) best[1] = us, pick[1] = lpp;
  }
  // the inner reverse tries its staged form (64-point blocks, 1, 2 or 4 lanes)
  if (A.I > 0)
    for (int l : {1, 2, 4}) {
      const int lpp = LPP_STAGED | l;
      if (!valid_rev_staged(c, 2, lpp)) continue;
      double us = 0;
      if ((rc = time_launches(c, reps, [&] { launch_inner_rev(c, A, 0, 1, lpp); }, &us))) return rc;
      if (us < best[2]) best[2] = us, pick[2] = lpp;
    }
  // the flux reverse also tries its staged form
  for (int l : {2, 4, 8}) {
    const int lpp = LPP_STAGED | l;
    if (!valid_shape(c, 2, lpp)) continue;
    double us = 0;
    if ((rc = time_launches(c, reps, [&] { launch_flux_rev(c, A, 0, A.dqb[A.I & 1], lpp); }, &us))) return rc;
    if (us < best[0]) best[0] = us, pick[0] = lpp;
  }
  for (int lpp : cand) {
    double us[4] = {0, 0, 0, 0};
    if ((rc = time_launches(c, reps, [&] { launch_flux_rev(c, A, 0, A.dqb[A.I & 1], lpp); }, &us[0])) ||
        (rc = time_launches(c, reps, [&] { launch_lsq_rev(c, A, 0, lpp); }, &us[1])) ||
        (rc = time_launches(c, reps, [&] { launch_gather_flux(c, A, 0, A.dqb[A.I & 1], lpp, false); }, &us[3])))
      return rc;
    if (A.I > 0 && (rc = time_launches(c, reps, [&] { launch_inner_rev(c, A, 0, 1, lpp); }, &us[2]))) return rc;
    for (int k = 0; k < 4; ++k)
      if (us[k] < best[k]) best[k] = us[k], pick[k] = lpp;
  }
  // the inner reverse: a stage's I reversed inner iterations in the push form (kernel
  // and gather per iteration) against the pull form (its solve-adjoint kernel, then one
  // launch per iteration)
  if (A.I > 0) {
    double us_g = 0;
    if ((rc = time_launches(c, reps, [&] { launch_gather_lsq<true>(c, A, A.dqb[0], pick[3]); }, &us_g))) return rc;
    double push = A.I * (best[2] + us_g);
    for (int l : {1, 2, 4}) {
      double us = 0;
      if (A.ph.hcap == 0) break;
      auto stage = [&] {
        for (int j = A.I; j >= 1; --j) launch_inner_pull(c, A, 0, j, LPP_PULL | l, true);
      };
      if ((rc = time_launches(c, reps, stage, &us))) return rc;
      if (us < push) push = us, pick[2] = LPP_PULL | l;
    }
  }
  // the LSQ reverse: its push form's kernel and gather against the pull form (with the
  // solve-adjoint kernel, which it needs unless a pull-form inner reverse precedes it)
  {
    double us_g = 0;
    if ((rc = time_launches(c, reps, [&] { launch_gather_lsq<false>(c, A, (V4d*)nullptr, pick[3]); }, &us_g)))
      return rc;
    double push = best[1] + us_g;
    for (int l : {1, 2, 4}) {
      double us = 0;
      if (A.ph.hcap == 0) break;
      auto step = [&] { launch_lsq_pull(c, A, 0, LPP_PULL | l, true); };
      if ((rc = time_launches(c, reps, step, &us))) return rc;
      if (us < push) push = us, pick[1] = LPP_PULL | l;
    }
    // the timed pull launches delivered qm adjoints into qbx
    if ((rc = adj_zero(c, A.qbx, 4 * (int64_t)n))) return rc;
  }
  A.lpp_frev = pick[0];
  A.lpp_lrev = pick[1];
  A.lpp_irev = pick[2];
  A.lpp_gath = pick[3];
  A.tuned = true;
  return MF_OK;
}

// C(s + r, s), saturating: the number of iterations a binomial schedule with s free
// snapshots reverses when each is advanced at most r times
static int64_t eta(int s, int r) {
  double v = 1.0;
  for (int j = 1; j <= s; ++j) {
    v = v * (double)(r + j) / (double)j;
    if (v > 4e18) return (int64_t)4e18;
  }
  return (int64_t)(v + 0.5);
}

// split of [t0, t1) for `free_slots` free snapshots: returns the checkpoint position
static int binomial_mid(int t0, int t1, int free_slots) {
  const int m = t1 - t0;
  int r = 1;
  while (eta(free_slots, r) < m) ++r;
  const int64_t right = std::min<int64_t>(eta(free_slots - 1, r), (int64_t)m - 1);
  return t1 - (int)right;
}

// The pull-form inner reverse's halo (PullHalo) from the connectivity (hc [W][n], hn)
// and its transpose (ptr, own).  LDS holds the 99th percentile of the blocks' halo
// sizes, at most 248 entries and what kMaxStageLds allows, at least the STAGE_P own
// points; entries past it are marked 255 and read from global memory.
static int build_pull_halo(mf_ctx* c, AdjState& A, const std::vector<uint32_t>& hc, const std::vector<int32_t>& hn,
                           const std::vector<int32_t>& ptr, const std::vector<int32_t>& own) {
  const int n = c->n, W = c->W, P = STAGE_P;
  const int nblk = (n + P - 1) / P, stride = nblk * P;
  std::vector<int32_t> hptr(nblk + 1, 0), ids, other, sizes;
  ids.reserve((size_t)n * 4);
  sizes.reserve(nblk);
  std::vector<uint16_t> loc((size_t)W * stride, 0), tloc(std::max<size_t>(ptr[n], 1), 0);
  for (int b = 0; b < nblk; ++b) {
    const int p0 = b * P, p1 = std::min(n, p0 + P);
    other.clear();
    for (int i = p0; i < p1; ++i) {
      for (int s = 0; s < hn[i]; ++s) {
        const int k = (int)(hc[(size_t)s * n + i] >> 4);
        if (k < p0 || k >= p1) other.push_back(k);
      }
      for (int e = ptr[i]; e < ptr[i + 1]; ++e)
        if (own[e] < p0 || own[e] >= p1) other.push_back(own[e]);
    }
    std::sort(other.begin(), other.end());
    other.erase(std::unique(other.begin(), other.end()), other.end());
    const int nown = p1 - p0;
    auto pos_of = [&](int k) {
      return (k >= p0 && k < p1) ? k - p0
                                 : nown + (int)(std::lower_bound(other.begin(), other.end(), k) - other.begin());
    };
    sizes.push_back(nown + (int)other.size());
    for (int i = p0; i < p1; ++i) ids.push_back(i);
    ids.insert(ids.end(), other.begin(), other.end());
    for (int i = p0; i < p1; ++i) {
      for (int s = 0; s < hn[i]; ++s)
        loc[(size_t)s * stride + i] = (uint16_t)std::min(pos_of((int)(hc[(size_t)s * n + i] >> 4)), 65535);
      for (int e = ptr[i]; e < ptr[i + 1]; ++e) tloc[e] = (uint16_t)std::min(pos_of(own[e]), 65535);
    }
    hptr[b + 1] = (int32_t)ids.size();
  }
  std::sort(sizes.begin(), sizes.end());
  int cap = sizes[std::min(nblk - 1, (int)((int64_t)nblk * 99 / 100))];
  const int lds_cap = (int)((kMaxStageLds - std::min(kMaxStageLds, (size_t)W * P)) / stage_lds_inner_pull(1, 0));
  if (lds_cap < P) return MF_OK;  // stencils too wide: the pull form is not offered (hcap 0)
  cap = std::max(P, std::min({cap, 248, lds_cap}));
  const int hcap = std::min((cap + 7) / 8 * 8, std::max(P, lds_cap / 8 * 8));
  std::vector<uint8_t> loc8(loc.size()), tloc8(tloc.size());
  for (size_t e = 0; e < loc.size(); ++e) loc8[e] = (uint8_t)(loc[e] < hcap ? loc[e] : 255);
  for (size_t e = 0; e < tloc.size(); ++e) tloc8[e] = (uint8_t)(tloc[e] < hcap ? tloc[e] : 255);
  int32_t *d_ptr, *d_ids;
  uint8_t *d_loc8, *d_tloc8;
  int rc;
  if ((rc = upload(c, &d_ptr, hptr)) || (rc = upload(c, &d_ids, ids)) || (rc = upload(c, &d_loc8, loc8)) ||
      (rc = upload(c, &d_tloc8, tloc8)))
    return rc;
  A.ph = PullHalo{d_ptr, d_ids, d_loc8, d_tloc8, hcap, stride};
  return MF_OK;
}

static int adj_alloc(mf_ctx* c, AdjState& A, int snaps) {
  const int n = c->n;
  const int S = c->p.rks, I = c->p.inner_iterations;
  int rc;
  if (A.ready && (A.S != S || A.I != I))
    return set_err(MF_ERR_STATE, "adjoint buffers already sized for a different scheme");
  // snapshots hold the context's state records (FP64, or FP32 for a mixed context)
  const size_t snap_v4 = c->ssz == 8 ? (size_t)n : ((size_t)n + 1) / 2;
  if (A.ready) {
    // a schedule with more snapshots than any before on this context: add slots
    while ((int)A.snap.size() < snaps) {
      V4d* p = nullptr;
      if ((rc = dev_alloc(c, &p, snap_v4))) return rc;
      A.snap.push_back(p);
      A.snap_t.push_back(-1);
    }
    return MF_OK;
  }
  A.mixed = c->ssz == 4;
  if (A.mixed) {
    A.fprim.assign(S + 1, nullptr);
    A.fq.assign(S + 1, nullptr);
    A.fdq.assign((size_t)S * (I + 1), nullptr);
    A.fqm.assign(S, nullptr);
    for (auto& p : A.fprim) if ((rc = dev_alloc(c, &p, n))) return rc;
    for (auto& p : A.fq) if ((rc = dev_alloc(c, &p, n))) return rc;
    for (auto& p : A.fdq) if ((rc = dev_alloc(c, &p, 2 * (size_t)n))) return rc;
    for (auto& p : A.fqm) if ((rc = dev_alloc(c, &p, 2 * (size_t)n))) return rc;
    if ((rc = dev_alloc(c, &A.fpold, n))) return rc;
    A.fbase_prim = A.fprim;
    A.fbase_q = A.fq;
    A.fbase_dq = A.fdq;
    A.fbase_qm = A.fqm;
    A.fbase_pold = A.fpold;
  }
  A.S = S;
  A.I = I;
  A.tprim.assign(S + 1, nullptr);
  A.tq.assign(S + 1, nullptr);
  A.tdq.assign((size_t)S * (I + 1), nullptr);
  A.tqm.assign(S, nullptr);
  for (auto& p : A.tprim) if ((rc = dev_alloc(c, &p, n))) return rc;
  for (auto& p : A.tq) if ((rc = dev_alloc(c, &p, n))) return rc;
  for (auto& p : A.tdq) if ((rc = dev_alloc(c, &p, 2 * (size_t)n))) return rc;
  for (auto& p : A.tqm) if ((rc = dev_alloc(c, &p, 2 * (size_t)n))) return rc;
  if ((rc = dev_alloc(c, &A.tpold, n)) || (rc = dev_alloc(c, &A.tdelta, n))) return rc;
  A.base_prim = A.tprim;
  A.base_q = A.tq;
  A.base_dq = A.tdq;
  A.base_qm = A.tqm;
  A.base_pold = A.tpold;
  A.base_delta = A.tdelta;
  if ((rc = dev_alloc(c, &A.tcspeed, n)) || (rc = dev_alloc(c, &A.tpartial, c->npartial)) ||
      (rc = dev_alloc(c, &A.tctl, 1)) || (rc = dev_alloc(c, &A.Pb, n)) ||
      (rc = dev_alloc(c, &A.Pstage, n)) || (rc = dev_alloc(c, &A.poldb, n)) ||
      (rc = dev_alloc(c, &A.frb, n)) || (rc = dev_alloc(c, &A.qb, n)) ||
      (rc = dev_alloc(c, &A.dqb[0], 2 * (size_t)n)) || (rc = dev_alloc(c, &A.dqb[1], 2 * (size_t)n)) ||
      (rc = dev_alloc(c, &A.qmb, 2 * (size_t)n)) || (rc = dev_alloc(c, &A.deltab, n)) ||
      (rc = dev_alloc(c, &A.xyb, n)) || (rc = dev_alloc(c, &A.nrmb, n)))
    return rc;
  Ctl z{0, 0, 1.0};
  HIPCHK(hipMemcpy(A.tctl, &z, sizeof(Ctl), hipMemcpyHostToDevice));
  if ((rc = upload(c, &A.hxx, c->h_hxx)) || (rc = upload(c, &A.hyy, c->h_hyy)) ||
      (rc = upload(c, &A.hxy, c->h_hxy)))
    return rc;
  std::vector<int32_t> sl, sr, bm, bl, br;
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
  A.nbd = (int)bm.size();
  if ((rc = upload(c, &A.bd_m, bm)) || (rc = upload(c, &A.bd_l, bl)) || (rc = upload(c, &A.bd_r, br)))
    return rc;
  if (c->nshape > 0 && ((rc = upload(c, &A.sh_l, sl)) || (rc = upload(c, &A.sh_r, sr)))) return rc;
  // transposed connectivity: the edges that name each point, in increasing owner
  // order, and each edge's position in that order (the edge buffer's layout)
  {
    const int W = c->W;
    if ((int64_t)W * n > INT32_MAX) return set_err(MF_ERR_INVALID, "edge ids exceed int32");
    std::vector<uint32_t> hc((size_t)W * n);
    std::vector<int32_t> hn(n);
    HIPCHK(hipMemcpy(hc.data(), c->g.conn, hc.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hn.data(), c->g.nbhs, hn.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::vector<int32_t> ptr(n + 1, 0);
    for (int i = 0; i < n; ++i)
      for (int s = 0; s < hn[i]; ++s) ++ptr[(hc[(size_t)s * n + i] >> 4) + 1];
    for (int k = 0; k < n; ++k) ptr[k + 1] += ptr[k];
    std::vector<int32_t> own(ptr[n]), pos((size_t)W * n, 0), cur(ptr.begin(), ptr.end() - 1);
    for (int i = 0; i < n; ++i)
      for (int s = 0; s < hn[i]; ++s) {
        const int k = (int)(hc[(size_t)s * n + i] >> 4);
        pos[(size_t)s * n + i] = cur[k];
        own[cur[k]] = i;
        ++cur[k];
      }
    if ((rc = upload(c, &A.t_ptr, ptr)) || (rc = upload(c, &A.t_own, own)) || (rc = upload(c, &A.epos, pos)))
      return rc;
    if ((rc = dev_alloc(c, &A.eb, (size_t)EF_FLUX * std::max(ptr[n], 1))) ||
        (rc = dev_alloc(c, &A.esel, (size_t)std::max(ptr[n], 1))))
      return rc;
    if ((rc = build_pull_halo(c, A, hc, hn, ptr, own)) || (rc = dev_alloc(c, &A.lam[0], 3 * (size_t)n)) ||
        (rc = dev_alloc(c, &A.lam[1], 3 * (size_t)n)) || (rc = dev_alloc(c, &A.qbx, (size_t)n)) ||
        (rc = adj_zero(c, A.qbx, 4 * (int64_t)n)))
      return rc;
  }
  A.snap.assign(snaps, nullptr);
  for (auto& p : A.snap) if ((rc = dev_alloc(c, &p, snap_v4))) return rc;
  A.snap_t.assign(snaps, -1);
  HIPCHK(hipEventCreate(&A.e0));
  HIPCHK(hipEventCreate(&A.e1));
  A.ready = true;
  return MF_OK;
}

// The one-iteration tape back on its own buffers (checkpointed schedules)
static void bind_base(AdjState& A) {
  A.tprim = A.base_prim;
  A.tq = A.base_q;
  A.tdq = A.base_dq;
  A.tqm = A.base_qm;
  A.tpold = A.base_pold;
  A.tdelta = A.base_delta;
  if (A.mixed) {
    A.fprim = A.fbase_prim;
    A.fq = A.fbase_q;
    A.fdq = A.fbase_dq;
    A.fqm = A.fbase_qm;
    A.fpold = A.fbase_pold;
  }
}

// The one-iteration tape pointed at iteration t of the all-iterations tape
static void bind_tape(AdjState& A, int t) {
  const AdjState::Tape& T = A.tapes[t];
  if (A.mixed) {
    // the kept records are FP32: point the FP32 records at them; the FP64 tape stays
    // on its own buffers and is filled by widening
    A.fprim[0] = t == 0 ? A.fbase_prim[0] : (V4f*)A.tapes[t - 1].prim[A.S - 1];
    for (int s = 0; s < A.S; ++s) {
      A.fprim[s + 1] = (V4f*)T.prim[s];
      A.fq[s] = (V4f*)T.q[s];
      A.fqm[s] = (V4f*)T.qm[s];
    }
    A.fq[A.S] = A.fbase_q[A.S];
    for (size_t k = 0; k < T.dq.size(); ++k) A.fdq[k] = (V4f*)T.dq[k];
    A.fpold = (V4f*)T.pold;
    A.tdelta = T.delta;
    return;
  }
  A.tprim[0] = t == 0 ? A.base_prim[0] : A.tapes[t - 1].prim[A.S - 1];
  for (int s = 0; s < A.S; ++s) {
    A.tprim[s + 1] = T.prim[s];
    A.tq[s] = T.q[s];
    A.tqm[s] = T.qm[s];
  }
  A.tq[A.S] = A.base_q[A.S];  // the last stage's q is not read by the reverse
  for (size_t k = 0; k < T.dq.size(); ++k) A.tdq[k] = T.dq[k];
  A.tpold = T.pold;
  A.tdelta = T.delta;
}

// Room for `iters` taped iterations (one allocation, reused while large enough);
// refuses a tape larger than the device's free memory less 1 GiB
static int tape_alloc(mf_ctx* c, AdjState& A, int iters) {
  const size_t n = (size_t)c->n, S = (size_t)A.S, I = (size_t)A.I;
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  // records in the context's precision (FP32 for a mixed context)
  const size_t rec = 4 * (size_t)c->ssz;
  const size_t v4 = up(n * rec), v8 = up(2 * n * rec), dl = up(n * sizeof(double));
  const size_t per = S * 2 * v4 + S * (I + 1) * v8 + S * v8 + v4 + dl;
  const size_t need = per * (size_t)iters;
  if (need > A.tape_cap) {
    if (A.tape_buf) {
      HIPCHK(hipFree(A.tape_buf));
      A.tape_buf = nullptr;
      A.tape_cap = 0;
    }
    size_t free_b = 0, total_b = 0;
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    if (need + ((size_t)1 << 30) > free_b)
      return set_err(MF_ERR_INVALID, "taping all %d iterations needs %.1f GB; %.1f GB of device memory is free",
                     iters, need / 1e9, free_b / 1e9);
    HIPCHK(hipMalloc(&A.tape_buf, need));
    A.tape_cap = need;
  }
  A.tapes.assign(iters, AdjState::Tape{});
  char* p = (char*)A.tape_buf;
  auto take = [&](size_t b) {
    char* r = p;
    p += b;
    return r;
  };
  for (auto& T : A.tapes) {
    T.prim.resize(S);
    T.q.resize(S);
    T.qm.resize(S);
    T.dq.resize(S * (I + 1));
    for (size_t s = 0; s < S; ++s) {
      T.prim[s] = (V4d*)take(v4);
      T.q[s] = (V4d*)take(v4);
      T.qm[s] = (V4d*)take(v8);
    }
    for (auto& d : T.dq) d = (V4d*)take(v8);
    T.pold = (V4d*)take(v4);
    T.delta = (double*)take(dl);
  }
  return MF_OK;
}

// copy of one state record array in the context's precision (prim, a snapshot, the
// taped iteration's input or output)
static int adj_copy(mf_ctx* c, void* dst, const void* src) {
  HIPCHK(hipMemcpyAsync(dst, src, (size_t)c->n * 4 * c->ssz, hipMemcpyDeviceToDevice, c->stream));
  return MF_OK;
}

// advance the context's state `steps` iterations (fused graph path), counting them
static int adj_advance(mf_ctx* c, AdjState& A, int t_from, int steps) {
  if (steps <= 0) return MF_OK;
  A.fwd_steps += steps;
  return mf_run_async(c, A.first_it + t_from, steps);
}

static int adj_taped_reverse(mf_ctx* c, AdjState& A) {
  int rc;
  if ((rc = adj_record(c, A))) return rc;
  A.taped += 1;
  if ((rc = adj_reverse_iteration(c, A))) return rc;
  if (A.progress_every > 0 && A.taped % A.progress_every == 0) {
    HIPCHK(hipStreamSynchronize(c->stream));
    std::fprintf(stderr, "[mfcfd] adjoint: %lld of %d iterations reversed\n", (long long)A.taped, A.iters);
    std::fflush(stderr);
  }
  return MF_OK;
}

// Reverse iterations [t0, t1); snapshot `slot0` holds prim(t0); `free_slots` more free
static int adj_rev_range(mf_ctx* c, AdjState& A, int t0, int t1, int slot0, int free_slots) {
  int rc;
  if (t1 - t0 == 1) {
    if ((rc = adj_copy(c, tape_in(A), A.snap[slot0]))) return rc;
    return adj_taped_reverse(c, A);
  }
  if (free_slots == 0) {
    for (int t = t1 - 1; t >= t0; --t) {
      if ((rc = adj_copy(c, c->prim[c->cur], A.snap[slot0])) || (rc = adj_advance(c, A, t0, t - t0)) ||
          (rc = adj_copy(c, tape_in(A), c->prim[c->cur])) || (rc = adj_taped_reverse(c, A)))
        return rc;
    }
    return MF_OK;
  }
  const int mid = binomial_mid(t0, t1, free_slots);
  int slot = -1;
  for (size_t j = 0; j < A.snap_t.size(); ++j)
    if (A.snap_t[j] == mid) slot = (int)j;
  if (slot < 0) {
    for (size_t j = 0; j < A.snap_t.size(); ++j)
      if (A.snap_t[j] < 0) {
        slot = (int)j;
        break;
      }
    if (slot < 0) return set_err(MF_ERR_STATE, "checkpoint schedule ran out of snapshots");
    if ((rc = adj_copy(c, c->prim[c->cur], A.snap[slot0])) || (rc = adj_advance(c, A, t0, mid - t0)) ||
        (rc = adj_copy(c, A.snap[slot], c->prim[c->cur])))
      return rc;
    A.snap_t[slot] = mid;
  }
  int used = 0;
  for (int t : A.snap_t) used += (t >= 0);
  A.peak_snaps = std::max(A.peak_snaps, used);
  if ((rc = adj_rev_range(c, A, mid, t1, slot, free_slots - 1))) return rc;
  A.snap_t[slot] = -1;
  return adj_rev_range(c, A, t0, mid, slot0, free_slots);
}

extern "C" {

int mf_adjoint_forward(mf_ctx* c, int32_t first_it, int32_t iters, int32_t snaps, double* J) {
  if (!c) return set_err(MF_ERR_INVALID, "null context");
  if (iters < 1 || (snaps < 1 && snaps != MF_ADJ_TAPE_ALL))
    return set_err(MF_ERR_INVALID, "iters must be >= 1 and snaps >= 1 or MF_ADJ_TAPE_ALL");
  if (c->nshape == 0) return set_err(MF_ERR_STATE, "J = Cd needs shape points");
  if (!c->adj) c->adj = new AdjState();
  AdjState& A = *c->adj;
  int rc;
  const bool tape_all = snaps == MF_ADJ_TAPE_ALL;
  if ((rc = adj_alloc(c, A, tape_all ? 1 : snaps))) return rc;
  bind_base(A);
  const bool pinned = c->lpp_rev[0] && c->lpp_rev[1] && c->lpp_rev[2] && c->lpp_rev[3];
  if (!A.tuned && !pinned && (rc = adj_tune(c, A))) return rc;
  // shapes pinned with mf_set_tuning override the timed ones
  if (c->lpp_rev[0]) A.lpp_frev = c->lpp_rev[0];
  if (c->lpp_rev[1]) A.lpp_lrev = c->lpp_rev[1];
  if (c->lpp_rev[2]) A.lpp_irev = c->lpp_rev[2];
  if (c->lpp_rev[3]) A.lpp_gath = c->lpp_rev[3];
  if (((A.lpp_irev & LPP_PULL) && A.I > 0 && A.ph.hcap == 0) || ((A.lpp_lrev & LPP_PULL) && A.ph.hcap == 0))
    return set_err(MF_ERR_INVALID, "the pull-form reverses do not fit in LDS at this stencil width");
  A.first_it = first_it;
  A.iters = iters;
  A.fwd_steps = A.taped = 0;
  A.peak_snaps = 1;
  A.tape_all = tape_all;
  for (auto& t : A.snap_t) t = -1;
  if (tape_all) {
    // the primal sweep as taped iterations, each kept for the reverse
    if ((rc = tape_alloc(c, A, iters))) return rc;
    HIPCHK(hipEventRecord(A.e0, c->stream));
    if ((rc = adj_copy(c, A.mixed ? (void*)A.fbase_prim[0] : (void*)A.base_prim[0], c->prim[c->cur]))) return rc;
    for (int t = 0; t < iters; ++t) {
      bind_tape(A, t);
      // a mixed context keeps the FP32 records; each is widened before its reverse
      if ((rc = A.mixed ? adj_record_mixed(c, A, false) : adj_record(c, A))) return rc;
    }
    if ((rc = adj_copy(c, c->prim[c->cur], tape_out(A))) || (rc = q_variables(c))) return rc;
    HIPCHK(hipEventRecord(A.e1, c->stream));
    HIPCHK(hipEventSynchronize(A.e1));
    HIPCHK(hipEventElapsedTime(&A.primal_ms, A.e0, A.e1));
    A.fwd_steps = iters;
    A.peak_snaps = 0;
    std::vector<double> cl(c->nsid), cd(c->nsid), cm(c->nsid);
    if ((rc = mf_compute_cl_cd_cm(c, cl.data(), cd.data(), cm.data(), nullptr))) return rc;
    A.J = 0.0;
    for (double v : cd) A.J += v;
    if (J) *J = A.J;
    return MF_OK;
  }
  // snapshot of the initial state, then the primal sweep, storing the checkpoints the
  // reverse recursion's right spine will ask for
  HIPCHK(hipEventRecord(A.e0, c->stream));
  if ((rc = adj_copy(c, A.snap[0], c->prim[c->cur]))) return rc;
  A.snap_t[0] = 0;
  std::vector<int> spine;
  {
    int t0 = 0, fs = snaps - 1;
    while (iters - t0 > 1 && fs > 0) {
      const int mid = binomial_mid(t0, iters, fs);
      spine.push_back(mid);
      t0 = mid;
      --fs;
    }
  }
  int t = 0;
  for (size_t j = 0; j < spine.size(); ++j) {
    if ((rc = adj_advance(c, A, t, spine[j] - t))) return rc;
    t = spine[j];
    if ((rc = adj_copy(c, A.snap[j + 1], c->prim[c->cur]))) return rc;
    A.snap_t[j + 1] = spine[j];
  }
  if ((rc = adj_advance(c, A, t, iters - t))) return rc;
  HIPCHK(hipEventRecord(A.e1, c->stream));
  HIPCHK(hipEventSynchronize(A.e1));
  HIPCHK(hipEventElapsedTime(&A.primal_ms, A.e0, A.e1));
  if ((rc = check_sweeps(c))) return rc;
  A.peak_snaps = (int)spine.size() + 1;
  std::vector<double> cl(c->nsid), cd(c->nsid), cm(c->nsid);
  if ((rc = mf_compute_cl_cd_cm(c, cl.data(), cd.data(), cm.data(), nullptr))) return rc;
  A.J = 0.0;
  for (double v : cd) A.J += v;
  if (J) *J = A.J;
  return MF_OK;
}

int mf_adjoint_backward(mf_ctx* c, double* dJdx, double* dJdy, mf_adjoint_stats* stats) {
  if (!c || !c->adj || !c->adj->ready || c->adj->iters < 1)
    return set_err(MF_ERR_STATE, "mf_adjoint_backward needs a preceding mf_adjoint_forward");
  AdjState& A = *c->adj;
  const int n = c->n;
  int rc;
  {
    const char* e = std::getenv("MFCFD_PROGRESS");
    A.progress_every = (e && e[0] == '1') ? std::max<int64_t>(1, A.iters / 20) : 0;
  }
  A.taped = 0;
  hipEvent_t b0, b1;
  HIPCHK(hipEventCreate(&b0));
  HIPCHK(hipEventCreate(&b1));
  HIPCHK(hipEventRecord(b0, c->stream));
  // keep the final state; the recursion reuses the context's prim as its work buffer
  const size_t state_bytes = (size_t)n * 4 * c->ssz;
  std::vector<char> final_state(state_bytes);
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(final_state.data(), c->prim[c->cur], state_bytes, hipMemcpyDeviceToHost));
  // the final state in FP64 for the force integral's reverse (a mixed context's is
  // widened into the tape's first record, free until the sweep starts)
  const V4d* fin = (const V4d*)c->prim[c->cur];
  if (A.mixed) {
    adj_widen_n(c, (const V4f*)c->prim[c->cur], A.base_prim[0], n);
    fin = A.base_prim[0];
  }
  const int cur0 = c->cur;
  if ((rc = adj_zero(c, A.xyb, 2 * (int64_t)n)) || (rc = adj_zero(c, A.nrmb, 2 * (int64_t)n)) ||
      (rc = adj_zero(c, A.Pb, 4 * (int64_t)n)))
    return rc;
  // seed: J = sum of Cd over shapes on the final state (objective_function_b.F90:22-23)
  hipLaunchKernelGGL(k_forces_rev, dim3((c->nshape + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, c->stream,
                     fin, c->sh_m, c->sh_id, A.sh_l, A.sh_r, c->sgeo, c->nshape,
                     (const double*)c->shape_par, c->g.xy, A.Pb, A.xyb, A.nrmb);
  HIPCHK(hipGetLastError());
  if (A.tape_all) {
    for (int t = A.iters - 1; t >= 0; --t) {
      bind_tape(A, t);
      if (A.mixed && (rc = adj_widen(c, A))) return rc;
      if ((rc = adj_reverse_iteration(c, A))) return rc;
      A.taped += 1;
      if (A.progress_every > 0 && A.taped % A.progress_every == 0) {
        HIPCHK(hipStreamSynchronize(c->stream));
        std::fprintf(stderr, "[mfcfd] adjoint: %lld of %d iterations reversed\n", (long long)A.taped, A.iters);
        std::fflush(stderr);
      }
    }
    bind_base(A);
    // the tape is as large as the run: give it back now
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipFree(A.tape_buf));
    A.tape_buf = nullptr;
    A.tape_cap = 0;
    A.tapes.clear();
  } else if ((rc = adj_rev_range(c, A, 0, A.iters, 0, (int)A.snap.size() - 1))) {
    return rc;
  }
  // normals depend on x, y of the boundary neighbours (point_normals_b.F90)
  if (A.nbd > 0)
    hipLaunchKernelGGL(k_normals_rev, dim3((A.nbd + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, c->stream, A.nbd,
                       A.bd_m, A.bd_l, A.bd_r, c->g.xy, (const D2*)A.nrmb, A.xyb);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(b1, c->stream));
  HIPCHK(hipEventSynchronize(b1));
  float bms = 0.f;
  HIPCHK(hipEventElapsedTime(&bms, b0, b1));
  (void)hipEventDestroy(b0);
  (void)hipEventDestroy(b1);
  // the recomputed primal stretches ran the iteration's graphs
  if ((rc = check_sweeps(c))) return rc;
  std::vector<D2> h(n);
  HIPCHK(hipMemcpy(h.data(), A.xyb, n * sizeof(D2), hipMemcpyDeviceToHost));
  for (int j = 0; j < n; ++j) {
    if (dJdx) dJdx[c->perm[j]] = h[j].x;
    if (dJdy) dJdy[c->perm[j]] = h[j].y;
  }
  c->cur = cur0;
  HIPCHK(hipMemcpy(c->prim[c->cur], final_state.data(), state_bytes, hipMemcpyHostToDevice));
  if (q_variables(c)) return MF_ERR_HIP;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (stats) {
    stats->forward_steps = A.fwd_steps;
    stats->taped_steps = A.taped;
    stats->snapshots = A.peak_snaps;
    stats->reserved = 0;
    stats->primal_ms = A.primal_ms;
    stats->adjoint_ms = bms;
    stats->lpp_flux_rev = A.lpp_frev;
    stats->lpp_lsq_rev = A.lpp_lrev;
    stats->lpp_inner_rev = A.lpp_irev;
    stats->lpp_gather = A.lpp_gath;
  }
  return MF_OK;
}

int mf_adjoint(mf_ctx* c, int32_t first_it, int32_t iters, int32_t snaps, double* J, double* dJdx,
               double* dJdy, mf_adjoint_stats* stats) {
  int rc = mf_adjoint_forward(c, first_it, iters, snaps, J);
  if (rc) return rc;
  return mf_adjoint_backward(c, dJdx, dJdy, stats);
}

}  // extern "C"
