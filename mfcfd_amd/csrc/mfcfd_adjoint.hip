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
// then reversed stage by stage.  Contributions to a neighbour's adjoints are added with
// hardware FP64 atomics; a point's own adjoints are summed in registers and added once.
// Iterates are checkpointed with a binomial (revolve-style) schedule of prim snapshots
// in HBM (q_lskum_chkpts_b.F90:116-279 drives Griewank's REVOLV the same way).
#include "lskum_adjoint.h"

namespace mf {

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

struct AdjOut {
  V4d* qb;      // adjoint of the stage's q
  V4d* dqb;     // adjoint of the dq the kernel reads [2N]
  V4d* qmb;     // adjoint of qm [2N]
  double* deltab;
  D2* xyb;      // adjoint of the coordinates
  D2* nrmb;     // adjoint of the normals
};

// Half-stencil geometry sums (sum ds^2, dn^2, ds dn per half), kept for the adjoint
struct HalfSums {
  const V4d* hxx;
  const V4d* hyy;
  const V4d* hxy;
};

template <int CLS>
__device__ void flux_point_rev(const Geom& g, const HalfSums& hs, const State<double>& st, int i,
                               double delta, const V4d& frbv, const AdjOut& o) {
  const D2 pi = g.xy[i];
  const D2 ni = g.nrm[i];
  const double nx = ni.x, ny = ni.y, tx = ny, ty = -nx;
  const V4d qi = st.q[i], dxi = st.dq[2 * i], dyi = st.dq[2 * i + 1];
  const V4d qmaxi = st.qm[2 * i], qmini = st.qm[2 * i + 1];
  const double epsi = st.eps[i];
  HalfCoef hc;
  {
    const V4d a = g.dg_c1[i], b = g.dg_sxy[i], c = g.dg_obd[i];
    for (int d = 0; d < 4; ++d) {
      hc.c1[d] = a[d];
      hc.sxy[d] = b[d];
      hc.obd[d] = c[d];
    }
  }
  const double scale = (CLS == WALL) ? 2.0 * delta : delta;
  double accb[4];
  for (int r = 0; r < 4; ++r) accb[r] = scale * frbv[r];
  double acc[4] = {0, 0, 0, 0};
  HalfCoefB hcb;
  for (int d = 0; d < 4; ++d) hcb.c1[d] = hcb.sxy[d] = hcb.obd[d] = 0.0;
  EndB own;
  endb_zero(own);
  double xib = 0.0, yib = 0.0, nxb = 0.0, nyb = 0.0;
  const int nb = g.nbhs[i];
  const uint32_t* cp = g.conn + i;
  for (int s = 0; s < nb; ++s) {
    const uint32_t pk = cp[(size_t)s * g.n];
    const int k = (int)(pk >> 4);
    const unsigned m = pk & 15u;
    const D2 pkk = g.xy[k];
    const double delx = pkk.x - pi.x, dely = pkk.y - pi.y;
    const double dels = delx * tx + dely * ty;
    const double deln = delx * nx + dely * ny;
    const double dsw = dels, dnw = deln;
    const bool xp = (m & XPOS) != 0u;
    const bool tie_x = xp && (m & XNEG) != 0u;
    const bool yp = (CLS == INTERIOR) ? (m & YPOS) != 0u : (CLS == OUTER);
    const bool tie_y = (CLS == INTERIOR) && yp && (m & YNEG) != 0u;
    const double cx = coef_x(hc, xp, dsw, dnw);
    const double cy = coef_y(hc, yp, dsw, dnw);
    const V4d qk = st.q[k], dxk = st.dq[2 * k], dyk = st.dq[2 * k + 1];
    const V4d qmaxk = st.qm[2 * k], qmink = st.qm[2 * k + 1];
    const double epsk = st.eps[k];
    double qti[4], qtk[4];
    if (g.first_order) {
      for (int r = 0; r < 4; ++r) {
        qti[r] = qi[r];
        qtk[r] = qk[r];
      }
    } else {
      reconstruct(qi, dxi, dyi, qmaxi, qmini, epsi, delx, dely, qti);
      reconstruct(qk, dxk, dyk, qmaxk, qmink, epsk, delx, dely, qtk);
    }
    KinF fi, fk;
    kin_fwd(qti, nx, ny, fi);
    kin_fwd(qtk, nx, ny, fk);
    double Gxi[4], Gyi[4], Gxk[4], Gyk[4];
    end_fluxes_f<CLS>(fi, xp, yp, Gxi, Gyi);
    end_fluxes_f<CLS>(fk, xp, yp, Gxk, Gyk);
    double cxb = 0.0, cyb = 0.0, Gxb[4], Gyb[4], Gxbn[4], Gybn[4];
    for (int r = 0; r < 4; ++r) {
      const double dgx = Gxk[r] - Gxi[r], dgy = Gyk[r] - Gyi[r];
      acc[r] += dgx * cx + dgy * cy;
      cxb += accb[r] * dgx;
      cyb += accb[r] * dgy;
      Gxb[r] = accb[r] * cx;
      Gyb[r] = accb[r] * cy;
      Gxbn[r] = -Gxb[r];
      Gybn[r] = -Gyb[r];
    }
    KinB kbi, kbk;
    kinb_zero(kbi);
    kinb_zero(kbk);
    end_fluxes_rev<CLS>(fk, xp, yp, Gxb, Gyb, kbk);
    end_fluxes_rev<CLS>(fi, xp, yp, Gxbn, Gybn, kbi);
    double dswb = 0.0, dnwb = 0.0;
    coef_rev(hc, xp ? 0 : 1, false, dsw, dnw, cxb, hcb, dswb, dnwb);
    coef_rev(hc, yp ? 2 : 3, true, dsw, dnw, cyb, hcb, dswb, dnwb);
    if (tie_x || tie_y) {
      // the neighbour is also in the xneg (yneg) half, with the opposite split
      double Hxi[4], Hyi[4], Hxk[4], Hyk[4];
      end_fluxes_f<CLS>(fi, false, false, Hxi, Hyi);
      end_fluxes_f<CLS>(fk, false, false, Hxk, Hyk);
      const double cx2 = coef_x(hc, false, dsw, dnw), cy2 = coef_y(hc, false, dsw, dnw);
      double cx2b = 0.0, cy2b = 0.0, Hxb[4], Hyb[4], Hxbn[4], Hybn[4];
      for (int r = 0; r < 4; ++r) {
        const double dx2 = tie_x ? Hxk[r] - Hxi[r] : 0.0, dy2 = tie_y ? Hyk[r] - Hyi[r] : 0.0;
        acc[r] += dx2 * cx2 + dy2 * cy2;
        cx2b += accb[r] * dx2;
        cy2b += accb[r] * dy2;
        Hxb[r] = tie_x ? accb[r] * cx2 : 0.0;
        Hyb[r] = tie_y ? accb[r] * cy2 : 0.0;
        Hxbn[r] = -Hxb[r];
        Hybn[r] = -Hyb[r];
      }
      end_fluxes_rev<CLS>(fk, false, false, Hxb, Hyb, kbk);
      end_fluxes_rev<CLS>(fi, false, false, Hxbn, Hybn, kbi);
      if (tie_x) coef_rev(hc, 1, false, dsw, dnw, cx2b, hcb, dswb, dnwb);
      if (tie_y) coef_rev(hc, 3, true, dsw, dnw, cy2b, hcb, dswb, dnwb);
    }
    double qtib[4] = {0, 0, 0, 0}, qtkb[4] = {0, 0, 0, 0};
    kin_rev(fi, kbi, nx, ny, qtib, nxb, nyb);
    kin_rev(fk, kbk, nx, ny, qtkb, nxb, nyb);
    double delxb = 0.0, delyb = 0.0;
    EndB nbr;
    endb_zero(nbr);
    if (g.first_order) {
      for (int r = 0; r < 4; ++r) {
        own.q[r] += qtib[r];
        nbr.q[r] += qtkb[r];
      }
    } else {
      reconstruct_rev(qi, dxi, dyi, qmaxi, qmini, epsi, delx, dely, qtib, own, delxb, delyb);
      reconstruct_rev(qk, dxk, dyk, qmaxk, qmink, epsk, delx, dely, qtkb, nbr, delxb, delyb);
    }
    // dels = delx ny - dely nx, deln = delx nx + dely ny
    delxb += dswb * ny + dnwb * nx;
    delyb += -dswb * nx + dnwb * ny;
    nxb += -dswb * dely + dnwb * delx;
    nyb += dswb * delx + dnwb * dely;
    xib -= delxb;
    yib -= delyb;
    aadd2(&o.xyb[k], delxb, delyb);
    aadd4(&o.qb[k], nbr.q);
    if (!g.first_order) {
      aadd4(&o.dqb[2 * k], nbr.dx);
      aadd4(&o.dqb[2 * k + 1], nbr.dy);
      aadd4(&o.qmb[2 * k], nbr.qmax);
      aadd4(&o.qmb[2 * k + 1], nbr.qmin);
    }
  }
  // half-stencil geometry: c1 = sum dn^2 (x halves) or sum ds^2 (y halves), sxy = sum
  // ds dn, obd = 1/(sum ds^2 sum dn^2 - sxy^2)
  double hxxb[4], hyyb[4], hxyb[4];
  {
    const V4d hxx = hs.hxx[i], hyy = hs.hyy[i], hxy = hs.hxy[i];
    for (int d = 0; d < 4; ++d) {
      hxxb[d] = (d >= 2) ? hcb.c1[d] : 0.0;
      hyyb[d] = (d < 2) ? hcb.c1[d] : 0.0;
      hxyb[d] = hcb.sxy[d];
      const double obd = hc.obd[d];
      const double detb = -hcb.obd[d] * obd * obd;
      if (hcb.obd[d] != 0.0) {
        hxxb[d] += detb * hyy[d];
        hyyb[d] += detb * hxx[d];
        hxyb[d] += -2.0 * detb * hxy[d];
      }
    }
  }
  for (int s = 0; s < nb; ++s) {
    const uint32_t pk = cp[(size_t)s * g.n];
    const int k = (int)(pk >> 4);
    const unsigned m = pk & 15u;
    const D2 pkk = g.xy[k];
    const double delx = pkk.x - pi.x, dely = pkk.y - pi.y;
    const double dels = delx * tx + dely * ty;
    const double deln = delx * nx + dely * ny;
    double delsb = 0.0, delnb = 0.0;
    for (int d = 0; d < 4; ++d) {
      if (!(m & (1u << d))) continue;
      delsb += 2.0 * dels * hxxb[d] + deln * hxyb[d];
      delnb += 2.0 * deln * hyyb[d] + dels * hxyb[d];
    }
    const double delxb = delsb * ny + delnb * nx;
    const double delyb = -delsb * nx + delnb * ny;
    nxb += -delsb * dely + delnb * delx;
    nyb += delsb * delx + delnb * dely;
    xib -= delxb;
    yib -= delyb;
    aadd2(&o.xyb[k], delxb, delyb);
  }
  aadd2(&o.xyb[i], xib, yib);
  aadd4(&o.qb[i], own.q);
  if (!g.first_order) {
    aadd4(&o.dqb[2 * i], own.dx);
    aadd4(&o.dqb[2 * i + 1], own.dy);
    aadd4(&o.qmb[2 * i], own.qmax);
    aadd4(&o.qmb[2 * i + 1], own.qmin);
  }
  o.nrmb[i].x += nxb;
  o.nrmb[i].y += nyb;
  double db = 0.0;
  for (int r = 0; r < 4; ++r) db += frbv[r] * acc[r];
  o.deltab[i] += (CLS == WALL) ? 2.0 * db : db;  // delta enters every stage
}

__global__ __launch_bounds__(BLOCK, 1) void k_flux_rev(Geom g, HalfSums hs, State<double> st,
                                                       const double* __restrict__ delta,
                                                       const V4d* __restrict__ frb, AdjOut o) {
  const int i = point_id();
  if (i >= g.n) return;
  if (i < g.nw) flux_point_rev<WALL>(g, hs, st, i, delta[i], frb[i], o);
  else if (i < g.nw + g.no) flux_point_rev<OUTER>(g, hs, st, i, delta[i], frb[i], o);
  else flux_point_rev<INTERIOR>(g, hs, st, i, delta[i], frb[i], o);
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
  else if (i < g.nw + g.no) update_rev<OUTER>(p, po, pn, nn.x, nn.y, rk, euler, fs, b, pbv, poldbv, frbv, nxb, nyb);
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

// reverse of k_inner (one inner iteration): dq_out adjoint -> dq_in, q, xy adjoints
__global__ __launch_bounds__(BLOCK) void k_inner_rev(Geom g, const V4d* __restrict__ q,
                                                     const V4d* __restrict__ dq_in,
                                                     const V4d* __restrict__ dq_out,
                                                     const V4d* __restrict__ dqb_out,
                                                     V4d* __restrict__ dqb_in, V4d* __restrict__ qb,
                                                     D2* __restrict__ xyb) {
  const int i = point_id();
  if (i >= g.n) return;
  const D2 pi = g.xy[i];
  const V4d gg = g.lsq_geo[i];
  const double sxx = gg.x, syy = gg.y, sxy = gg.z, obd = gg.w;
  const V4d tx = dq_out[2 * i], ty = dq_out[2 * i + 1];
  const V4d bx = dqb_out[2 * i], by = dqb_out[2 * i + 1];
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
  const V4d qi = q[i], dxi = dq_in[2 * i], dyi = dq_in[2 * i + 1];
  double qib[4] = {0, 0, 0, 0}, dxib[4] = {0, 0, 0, 0}, dyib[4] = {0, 0, 0, 0};
  double xib = 0.0, yib = 0.0;
  const int nb = g.nbhs[i];
  for (int s = 0; s < nb; ++s) {
    const int k = (int)(g.conn[(size_t)s * g.n + i] >> 4);
    const D2 pk = g.xy[k];
    const double dx = pk.x - pi.x, dy = pk.y - pi.y;
    const V4d qk = q[k], dxk = dq_in[2 * k], dyk = dq_in[2 * k + 1];
    double dxb = 2.0 * sxxb * dx + sxyb * dy, dyb = 2.0 * syyb * dy + sxyb * dx;
    double qkb[4], dxkb[4], dykb[4];
    for (int r = 0; r < 4; ++r) {
      const double t1 = qi[r] - 0.5 * (dx * dxi[r] + dy * dyi[r]);
      const double t2 = qk[r] - 0.5 * (dx * dxk[r] + dy * dyk[r]);
      const double D = t2 - t1;
      const double Db = sqb[r] * dx + syqb[r] * dy;
      dxb += sqb[r] * D - 0.5 * Db * (dxk[r] - dxi[r]);
      dyb += syqb[r] * D - 0.5 * Db * (dyk[r] - dyi[r]);
      qkb[r] = Db;
      dxkb[r] = -0.5 * dx * Db;
      dykb[r] = -0.5 * dy * Db;
      qib[r] -= Db;
      dxib[r] += 0.5 * dx * Db;
      dyib[r] += 0.5 * dy * Db;
    }
    aadd4(&qb[k], qkb);
    aadd4(&dqb_in[2 * k], dxkb);
    aadd4(&dqb_in[2 * k + 1], dykb);
    aadd2(&xyb[k], dxb, dyb);
    xib -= dxb;
    yib -= dyb;
  }
  aadd4(&qb[i], qib);
  aadd4(&dqb_in[2 * i], dxib);
  aadd4(&dqb_in[2 * i + 1], dyib);
  aadd2(&xyb[i], xib, yib);
}

// reverse of k_lsq (eval_q_derivatives): dq and qm adjoints -> q, xy adjoints
__global__ __launch_bounds__(BLOCK) void k_lsq_rev(Geom g, const V4d* __restrict__ q,
                                                   const V4d* __restrict__ dq,
                                                   const V4d* __restrict__ dqb,
                                                   const V4d* __restrict__ qmb, V4d* __restrict__ qb,
                                                   D2* __restrict__ xyb) {
  const int i = point_id();
  if (i >= g.n) return;
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
  // adjoint going to the selected point
  V4d qmax = qi, qmin = qi;
  int amax[4] = {i, i, i, i}, amin[4] = {i, i, i, i};
  double qib[4] = {0, 0, 0, 0};
  double xib = 0.0, yib = 0.0;
  const int nb = g.nbhs[i];
  for (int s = 0; s < nb; ++s) {
    const int k = (int)(g.conn[(size_t)s * g.n + i] >> 4);
    const D2 pk = g.xy[k];
    const double dx = pk.x - pi.x, dy = pk.y - pi.y;
    const V4d qk = q[k];
    double dxb = 2.0 * sxxb * dx + sxyb * dy, dyb = 2.0 * syyb * dy + sxyb * dx;
    double qkb[4];
    for (int r = 0; r < 4; ++r) {
      if (qk[r] > qmax[r]) {
        qmax[r] = qk[r];
        amax[r] = k;
      }
      if (qk[r] < qmin[r]) {
        qmin[r] = qk[r];
        amin[r] = k;
      }
      const double D = qk[r] - qi[r];
      const double Db = sqb[r] * dx + syqb[r] * dy;
      dxb += sqb[r] * D;
      dyb += syqb[r] * D;
      qkb[r] = Db;
      qib[r] -= Db;
    }
    aadd4(&qb[k], qkb);
    aadd2(&xyb[k], dxb, dyb);
    xib -= dxb;
    yib -= dyb;
  }
  for (int r = 0; r < 4; ++r) {
    if (amax[r] == i) qib[r] += mxb[r];
    else aadd(&qb[amax[r]].x + r, mxb[r]);
    if (amin[r] == i) qib[r] += mnb[r];
    else aadd(&qb[amin[r]].x + r, mnb[r]);
  }
  aadd4(&qb[i], qib);
  aadd2(&xyb[i], xib, yib);
}

// reverse of eval_q_variables: pb += dq/dprim^T qb (own point)
__global__ __launch_bounds__(BLOCK) void k_qvar_rev(int n, const V4d* __restrict__ prim,
                                                    const V4d* __restrict__ qb, V4d* __restrict__ pb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const V4d b = qb[i];
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
                                                      const int32_t* __restrict__ sh_l,
                                                      const int32_t* __restrict__ sh_r,
                                                      const double* __restrict__ sgeo, int nsh,
                                                      double cp_scale, double ct, double st,
                                                      const D2* __restrict__ xy, V4d* __restrict__ pb,
                                                      D2* __restrict__ xyb, D2* __restrict__ nrmb) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nsh) return;
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

__global__ void k_zero(double* __restrict__ p, int64_t n) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x)
    p[j] = 0.0;
}

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
  // checkpoints
  std::vector<V4d*> snap;       // prim snapshots
  std::vector<int> snap_t;      // iteration index held by each slot (-1 free)
  int first_it = 1, iters = 0;
  double J = 0.0;
  int64_t fwd_steps = 0, taped = 0;
  int peak_snaps = 0;
  float primal_ms = 0.f;
  hipEvent_t e0 = nullptr, e1 = nullptr;
};

static int adj_zero(mf_ctx* c, void* p, int64_t n_doubles) {
  hipLaunchKernelGGL(k_zero, dim3(1024), dim3(BLOCK), 0, c->stream, (double*)p, n_doubles);
  HIPCHK(hipGetLastError());
  return MF_OK;
}

// one taped forward iteration from A.tprim[0]
static int adj_record(mf_ctx* c, AdjState& A) {
  const int n = c->n;
  const dim3 gr(nblocks(c)), bl(BLOCK);
  hipLaunchKernelGGL((k_q_variables<double>), dim3(nblocks(c)), bl, 0, c->stream, n,
                     (const V4d*)A.tprim[0], A.tq[0], A.tcspeed);
  for (int s = 0; s < A.S; ++s) {
    V4d* dq0 = A.tdq[s * (A.I + 1)];
    if (s == 0)
      hipLaunchKernelGGL((k_lsq<double, true, true>), gr, bl, 0, c->stream, c->g, (const V4d*)A.tq[s], dq0, A.tqm[s],
                         A.tcspeed, A.tdelta, c->p.cfl, 0.0, nullptr);
    else
      hipLaunchKernelGGL((k_lsq<double, true, false>), gr, bl, 0, c->stream, c->g, (const V4d*)A.tq[s], dq0, A.tqm[s],
                         A.tcspeed, A.tdelta, c->p.cfl, 0.0, nullptr);
    for (int j = 1; j <= A.I; ++j)
      hipLaunchKernelGGL((k_inner<double, true>), gr, bl, 0, c->stream, c->g, (const V4d*)A.tq[s],
                         (const V4d*)A.tdq[s * (A.I + 1) + j - 1], A.tdq[s * (A.I + 1) + j], 0.0, nullptr);
    State<double> st{A.tq[s], A.tdq[s * (A.I + 1) + A.I], A.tqm[s], (const double*)c->eps};
    hipLaunchKernelGGL((k_flux_update<double, true>), gr, bl, 0, c->stream, c->g, st, (const V4d*)A.tprim[s],
                       A.tpold, A.tprim[s + 1], A.tq[s + 1], A.tcspeed, (const double*)A.tdelta, A.tpartial,
                       (const Ctl*)A.tctl, s + 1, s + 1 == A.S ? 1 : 0, c->p.euler, 0.0, c->fs);
  }
  HIPCHK(hipGetLastError());
  return MF_OK;
}

// reverse of the taped iteration: A.Pb holds the adjoint of its output prim and
// receives the adjoint of its input prim
static int adj_reverse_iteration(mf_ctx* c, AdjState& A) {
  const int n = c->n;
  const dim3 gr(nblocks(c)), bl(BLOCK);
  int rc;
  if ((rc = adj_zero(c, A.poldb, 4 * (int64_t)n)) || (rc = adj_zero(c, A.deltab, n))) return rc;
  HalfSums hs{A.hxx, A.hyy, A.hxy};
  for (int s = A.S - 1; s >= 0; --s) {
    // update: Pb (adjoint of prim_{s+1}) -> Pstage (prim_s), poldb, frb, nrmb
    hipLaunchKernelGGL(k_update_rev, gr, bl, 0, c->stream, c->g, (const V4d*)A.tprim[s], (const V4d*)A.tpold,
                       (const V4d*)A.tprim[s + 1], (const V4d*)A.Pb, A.Pstage, A.poldb, A.frb, A.nrmb, s + 1,
                       c->p.euler, c->fs);
    if ((rc = adj_zero(c, A.qb, 4 * (int64_t)n)) || (rc = adj_zero(c, A.qmb, 8 * (int64_t)n)) ||
        (rc = adj_zero(c, A.dqb[A.I & 1], 8 * (int64_t)n)))
      return rc;
    State<double> st{A.tq[s], A.tdq[s * (A.I + 1) + A.I], A.tqm[s], (const double*)c->eps};
    AdjOut o{A.qb, A.dqb[A.I & 1], A.qmb, A.deltab, A.xyb, A.nrmb};
    // flux: frb -> q, dq, qm, delta, xy, nrm
    hipLaunchKernelGGL(k_flux_rev, gr, bl, 0, c->stream, c->g, hs, st, (const double*)A.tdelta,
                       (const V4d*)A.frb, o);
    // inner iterations, last first: dq^(j) adjoint -> dq^(j-1)
    for (int j = A.I; j >= 1; --j) {
      const int cur = j & 1, prv = (j - 1) & 1;
      if ((rc = adj_zero(c, A.dqb[prv], 8 * (int64_t)n))) return rc;
      hipLaunchKernelGGL(k_inner_rev, gr, bl, 0, c->stream, c->g, (const V4d*)A.tq[s],
                         (const V4d*)A.tdq[s * (A.I + 1) + j - 1], (const V4d*)A.tdq[s * (A.I + 1) + j],
                         (const V4d*)A.dqb[cur], A.dqb[prv], A.qb, A.xyb);
    }
    // LSQ: dq^(0), qm adjoints -> q, xy
    hipLaunchKernelGGL(k_lsq_rev, gr, bl, 0, c->stream, c->g, (const V4d*)A.tq[s],
                       (const V4d*)A.tdq[s * (A.I + 1)], (const V4d*)A.dqb[0], (const V4d*)A.qmb, A.qb, A.xyb);
    // q = q(prim_s): into the stage-input adjoint
    hipLaunchKernelGGL(k_qvar_rev, dim3(nblocks(c)), bl, 0, c->stream, n, (const V4d*)A.tprim[s],
                       (const V4d*)A.qb, A.Pstage);
    std::swap(A.Pb, A.Pstage);
  }
  // func_delta on the iteration input, then prim_old = prim (fpi_solver.F90:20-24)
  hipLaunchKernelGGL(k_delta_rev, gr, bl, 0, c->stream, c->g, (const V4d*)A.tprim[0], (const double*)A.deltab,
                     c->p.cfl, A.Pb, A.xyb);
  hipLaunchKernelGGL(k_axpy, dim3(1024), bl, 0, c->stream, (double*)A.Pb, (const double*)A.poldb, (int64_t)4 * n);
  HIPCHK(hipGetLastError());
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

static int adj_alloc(mf_ctx* c, AdjState& A, int snaps) {
  const int n = c->n;
  const int S = c->p.rks, I = c->p.inner_iterations;
  if (A.ready && A.S == S && A.I == I && (int)A.snap.size() >= snaps) return MF_OK;
  if (A.ready) return set_err(MF_ERR_STATE, "adjoint buffers already sized for a different schedule");
  int rc;
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
  if ((rc = dev_alloc(c, &A.tpold, n)) || (rc = dev_alloc(c, &A.tdelta, n)) ||
      (rc = dev_alloc(c, &A.tcspeed, n)) || (rc = dev_alloc(c, &A.tpartial, c->npartial)) ||
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
  A.snap.assign(snaps, nullptr);
  for (auto& p : A.snap) if ((rc = dev_alloc(c, &p, n))) return rc;
  A.snap_t.assign(snaps, -1);
  HIPCHK(hipEventCreate(&A.e0));
  HIPCHK(hipEventCreate(&A.e1));
  A.ready = true;
  return MF_OK;
}

static int adj_copy(mf_ctx* c, void* dst, const void* src) {
  HIPCHK(hipMemcpyAsync(dst, src, (size_t)c->n * sizeof(V4d), hipMemcpyDeviceToDevice, c->stream));
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
  return adj_reverse_iteration(c, A);
}

// Reverse iterations [t0, t1); snapshot `slot0` holds prim(t0); `free_slots` more free
static int adj_rev_range(mf_ctx* c, AdjState& A, int t0, int t1, int slot0, int free_slots) {
  int rc;
  if (t1 - t0 == 1) {
    if ((rc = adj_copy(c, A.tprim[0], A.snap[slot0]))) return rc;
    return adj_taped_reverse(c, A);
  }
  if (free_slots == 0) {
    for (int t = t1 - 1; t >= t0; --t) {
      if ((rc = adj_copy(c, c->prim[c->cur], A.snap[slot0])) || (rc = adj_advance(c, A, t0, t - t0)) ||
          (rc = adj_copy(c, A.tprim[0], c->prim[c->cur])) || (rc = adj_taped_reverse(c, A)))
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
  if (c->ssz != 8) return set_err(MF_ERR_STATE, "the adjoint needs an MF_FP64 context");
  if (!c->power0) return set_err(MF_ERR_STATE, "the adjoint supports power = 0 only");
  if (iters < 1 || snaps < 1) return set_err(MF_ERR_INVALID, "iters and snaps must be >= 1");
  if (c->nshape == 0) return set_err(MF_ERR_STATE, "J = Cd needs shape points");
  if (!c->adj) c->adj = new AdjState();
  AdjState& A = *c->adj;
  int rc;
  if ((rc = adj_alloc(c, A, snaps))) return rc;
  A.first_it = first_it;
  A.iters = iters;
  A.fwd_steps = A.taped = 0;
  A.peak_snaps = 1;
  for (auto& t : A.snap_t) t = -1;
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
  A.peak_snaps = (int)spine.size() + 1;
  std::vector<double> cl(c->p.shapes), cd(c->p.shapes), cm(c->p.shapes);
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
  hipEvent_t b0, b1;
  HIPCHK(hipEventCreate(&b0));
  HIPCHK(hipEventCreate(&b1));
  HIPCHK(hipEventRecord(b0, c->stream));
  // keep the final state; the recursion reuses the context's prim as its work buffer
  V4d* fin = A.Pstage;  // borrowed until the sweep starts
  (void)fin;
  std::vector<V4d> final_state(n);
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(final_state.data(), c->prim[c->cur], n * sizeof(V4d), hipMemcpyDeviceToHost));
  const int cur0 = c->cur;
  if ((rc = adj_zero(c, A.xyb, 2 * (int64_t)n)) || (rc = adj_zero(c, A.nrmb, 2 * (int64_t)n)) ||
      (rc = adj_zero(c, A.Pb, 4 * (int64_t)n)))
    return rc;
  // seed: J = sum of Cd over shapes on the final state (objective_function_b.F90:22-23)
  hipLaunchKernelGGL(k_forces_rev, dim3((c->nshape + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, c->stream,
                     (const V4d*)c->prim[c->cur], c->sh_m, A.sh_l, A.sh_r, c->sgeo, c->nshape, c->cp_scale,
                     std::cos(c->theta), std::sin(c->theta), c->g.xy, A.Pb, A.xyb, A.nrmb);
  HIPCHK(hipGetLastError());
  if ((rc = adj_rev_range(c, A, 0, A.iters, 0, (int)A.snap.size() - 1))) return rc;
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
  std::vector<D2> h(n);
  HIPCHK(hipMemcpy(h.data(), A.xyb, n * sizeof(D2), hipMemcpyDeviceToHost));
  for (int j = 0; j < n; ++j) {
    if (dJdx) dJdx[c->perm[j]] = h[j].x;
    if (dJdy) dJdy[c->perm[j]] = h[j].y;
  }
  c->cur = cur0;
  HIPCHK(hipMemcpy(c->prim[c->cur], final_state.data(), n * sizeof(V4d), hipMemcpyHostToDevice));
  if (q_variables(c)) return MF_ERR_HIP;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (stats) {
    stats->forward_steps = A.fwd_steps;
    stats->taped_steps = A.taped;
    stats->snapshots = A.peak_snaps;
    stats->reserved = 0;
    stats->primal_ms = A.primal_ms;
    stats->adjoint_ms = bms;
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
