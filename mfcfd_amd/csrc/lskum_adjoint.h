// Reverse (adjoint) mode of the LSKUM iteration, FP64, for gfx950.
//
// Hand-written reverse of every forward routine in lskum_device.h, in the order the
// forward evaluates them: each *_rev function takes the adjoints of a routine's
// outputs and adds the adjoints of its inputs.  Forward intermediates are recomputed
// from the routine's inputs (no tape inside a point).  Non-differentiable choices
// follow the primal's branch exactly, as the reference's Tapenade adjoint does
// (src_mpi_adjoint/sensitivity_files/limiters_b.F90:43-164: phi = 1 and phi capped
// at 1 carry no derivative; q_variables_b.F90:188-223: qm's max/min pass the adjoint
// to the selected point; state_update_b.F90:465-471: func_delta's min to its
// minimising neighbour).  d erf(s)/ds = 2/sqrt(pi) exp(-s^2), as the reference
// patched it (split_fluxes_b.F90:82-83).
#pragma once
#include "lskum_device.h"

namespace mf {

// The reverse routines contract multiply-adds into FMAs: the reverse is held to the
// tangent and to finite differences (1e-10 and 1e-6), not to rounding parity with a
// reference, and its recomputed forward values only steer the same branches as the
// primal's (limiter thresholds, stencil halves from the taped connectivity bits).
#pragma clang fp contract(fast)

constexpr double TWO_BY_SQRT_PI = 1.1283791670955126;
constexpr double INV_SQRT_PI = 0.5641895835477563;
constexpr double INV_GAMMA_M1 = 1.0 / Consts<double>::GAMMA_M1;

// ---------------------------------------------------------------------------------
// q-tilde -> primitive -> flux preliminaries, with every intermediate kept.  S is the
// arithmetic of the recomputation and of the reverse through the kinetic state: double
// for the FP64 adjoint; float for the mixed-precision adjoint, whose fluxes and limiter
// run in FP32 as the mixed primal's do (its sensitivities still accumulate in FP64).
template <class S>
struct KinFT {
  S qt[4];
  S beta, tmp, u1, u2, T1, rho, p, ut, un, b, sb, spb, S1, S2, ex1, ex2, B1, B2, e1, e2, prr, usq;
};
using KinF = KinFT<double>;

template <class S>
struct KinBT {
  S rho, ut, un, prr, usq, B1, B2, e1, e2;
};
using KinB = KinBT<double>;

template <class S>
__host__ __device__ inline void kinb_zero(KinBT<S>& b) {
  b.rho = b.ut = b.un = b.prr = b.usq = b.B1 = b.B2 = b.e1 = b.e2 = (S)0;
}

// The recomputation's elementary functions: the library's in FP64 (the primal's own
// operations); in FP32 the hardware's, as the mixed primal (kin_fast) uses them.
template <class S>
__host__ __device__ inline S rv_rcp(S x) {
  if constexpr (sizeof(S) == 4) return frcp(x);
  else return 1.0 / x;
}

// kin_of_qtilde's quantities with every intermediate kept.  The reverse only consumes
// their values (no branch depends on them), so both precisions use the identities
// b = rho / (2 p) = beta, p / rho = tmp = 1 / (2 beta) and sqrt(pi b) = sqrt(pi) sqrt(b),
// and one reciprocal for both B's: 2 divisions and 1 square root per end in FP64
// instead of the primal routine's 6 and 2 (the values agree to rounding).  In FP64,
// rho = exp(T1 - log(beta) / (gamma - 1)) is formed as exp(T1) beta^-2.5 with
// beta^-2.5 = (1 / beta)^3 sqrt(beta), from the reciprocal and the square root above:
// no logarithm (about 90 VALU instructions in FP64).
template <class S>
__host__ __device__ inline void kin_fwd(const S qt[4], S nx, S ny, KinFT<S>& k) {
  for (int r = 0; r < 4; ++r) k.qt[r] = qt[r];
  k.beta = -qt[3] * (S)0.5;
  if constexpr (sizeof(S) == 4) k.tmp = 0.5f * frcp(k.beta);
  else k.tmp = 0.5 / k.beta;
  k.u1 = qt[1] * k.tmp;
  k.u2 = qt[2] * k.tmp;
  k.T1 = qt[0] + k.beta * (k.u1 * k.u1 + k.u2 * k.u2);
  k.b = k.beta;  // b = rho / (2 p) = beta
  if constexpr (sizeof(S) == 4) {
    k.rho = fexp(k.T1 - flog(k.beta) * (1.0f / Consts<float>::GAMMA_M1));
    k.sb = __builtin_amdgcn_sqrtf(k.b);
    k.spb = k.sb * 1.7724538509f;  // sqrt(pi b)
  } else {
    k.sb = m_sqrt(k.b);
    k.spb = k.sb * 1.7724538509055160;  // sqrt(pi b)
    const double ib = 2.0 * k.tmp;      // 1 / beta
    k.rho = m_exp(k.T1) * ((ib * ib) * (ib * k.sb));
  }
  k.p = k.rho * k.tmp;
  const S tx = ny, ty = -nx;
  k.ut = k.u1 * tx + k.u2 * ty;
  k.un = k.u1 * nx + k.u2 * ny;
  k.S1 = k.ut * k.sb;
  k.S2 = k.un * k.sb;
  if constexpr (sizeof(S) == 4) {
    k.ex1 = fexp(-k.S1 * k.S1);
    k.ex2 = fexp(-k.S2 * k.S2);
    const float ispb = frcp(k.spb);
    k.B1 = 0.5f * k.ex1 * ispb;
    k.B2 = 0.5f * k.ex2 * ispb;
    k.e1 = ferf_from_exp(k.S1, k.ex1);
    k.e2 = ferf_from_exp(k.S2, k.ex2);
    k.prr = k.tmp;  // p / rho = 1 / (2 beta)
  } else {
    k.ex1 = m_exp(-k.S1 * k.S1);
    k.ex2 = m_exp(-k.S2 * k.S2);
    const double hispb = 0.5 / k.spb;
    k.B1 = k.ex1 * hispb;
    k.B2 = k.ex2 * hispb;
    k.e1 = m_erf(k.S1);
    k.e2 = m_erf(k.S2);
    k.prr = k.tmp;  // p / rho
  }
  k.usq = k.ut * k.ut + k.un * k.un;
}

template <class S>
__host__ __device__ inline Kin<S> kin_of(const KinFT<S>& f) {
  Kin<S> k;
  k.rho = f.rho;
  k.ut = f.ut;
  k.un = f.un;
  k.pr_by_rho = f.prr;
  k.u_sqr = f.usq;
  k.B1 = f.B1;
  k.B2 = f.B2;
  k.e1 = f.e1;
  k.e2 = f.e2;
  return k;
}

// reverse of flux_gx (split_fluxes.F90:8-96)
template <class S>
__host__ __device__ inline void flux_gx_rev(const KinFT<S>& k, bool plus, const S Gb[4], KinBT<S>& kb) {
  const S s = plus ? (S)1 : (S)-1;
  const S A1 = (S)0.5 * ((S)1 + s * k.e1);
  const S a7 = (S)7 * k.prr + k.usq, a6 = (S)6 * k.prr + k.usq;
  const S F0 = k.ut * A1 + s * k.B1;
  const S F1 = (k.prr + k.ut * k.ut) * A1 + s * k.ut * k.B1;
  const S F2 = k.ut * k.un * A1 + s * k.un * k.B1;
  const S F3 = (S)0.5 * k.ut * a7 * A1 + s * (S)0.5 * a6 * k.B1;
  kb.rho += Gb[0] * F0 + Gb[1] * F1 + Gb[2] * F2 + Gb[3] * F3;
  const S g0 = k.rho * Gb[0], g1 = k.rho * Gb[1], g2 = k.rho * Gb[2], g3 = k.rho * Gb[3];
  const S A1b = g0 * k.ut + g1 * (k.prr + k.ut * k.ut) + g2 * k.ut * k.un + g3 * (S)0.5 * k.ut * a7;
  kb.B1 += s * (g0 + g1 * k.ut + g2 * k.un + g3 * (S)0.5 * a6);
  kb.ut += g0 * A1 + g1 * ((S)2 * k.ut * A1 + s * k.B1) + g2 * k.un * A1 + g3 * (S)0.5 * a7 * A1;
  kb.un += g2 * (k.ut * A1 + s * k.B1);
  kb.prr += g1 * A1 + g3 * ((S)3.5 * k.ut * A1 + (S)3 * s * k.B1);
  kb.usq += g3 * (S)0.5 * (k.ut * A1 + s * k.B1);
  kb.e1 += (S)0.5 * s * A1b;
}

// reverse of flux_gy (split_fluxes.F90:99-190); outputs 1 and 2 swap roles
template <class S>
__host__ __device__ inline void flux_gy_rev(const KinFT<S>& k, bool plus, const S Gb[4], KinBT<S>& kb) {
  const S s = plus ? (S)1 : (S)-1;
  const S A2 = (S)0.5 * ((S)1 + s * k.e2);
  const S a7 = (S)7 * k.prr + k.usq, a6 = (S)6 * k.prr + k.usq;
  const S F0 = k.un * A2 + s * k.B2;
  const S F1 = k.ut * k.un * A2 + s * k.ut * k.B2;
  const S F2 = (k.prr + k.un * k.un) * A2 + s * k.un * k.B2;
  const S F3 = (S)0.5 * k.un * a7 * A2 + s * (S)0.5 * a6 * k.B2;
  kb.rho += Gb[0] * F0 + Gb[1] * F1 + Gb[2] * F2 + Gb[3] * F3;
  const S h0 = k.rho * Gb[0], hT = k.rho * Gb[1], hN = k.rho * Gb[2], h3 = k.rho * Gb[3];
  const S A2b = h0 * k.un + hN * (k.prr + k.un * k.un) + hT * k.ut * k.un + h3 * (S)0.5 * k.un * a7;
  kb.B2 += s * (h0 + hN * k.un + hT * k.ut + h3 * (S)0.5 * a6);
  kb.un += h0 * A2 + hN * ((S)2 * k.un * A2 + s * k.B2) + hT * k.ut * A2 + h3 * (S)0.5 * a7 * A2;
  kb.ut += hT * (k.un * A2 + s * k.B2);
  kb.prr += hN * A2 + h3 * ((S)3.5 * k.un * A2 + (S)3 * s * k.B2);
  kb.usq += h3 * (S)0.5 * (k.un * A2 + s * k.B2);
  kb.e2 += (S)0.5 * s * A2b;
}

// reverse of flux_quad (quadrant_fluxes.F90:12-248)
template <class S>
__host__ __device__ inline void flux_quad_rev(const KinFT<S>& k, bool a1p, bool a2p, const S Gb[4],
                                              KinBT<S>& kb) {
  const S s1 = a1p ? (S)1 : (S)-1, s2 = a2p ? (S)1 : (S)-1;
  const S A1 = (S)0.5 * ((S)1 + s1 * k.e1), A2 = (S)0.5 * ((S)1 + s2 * k.e2);
  const S a7 = (S)7 * k.prr + k.usq, a6 = (S)6 * k.prr + k.usq;
  const S t1b = k.ut * A1 + s1 * k.B1;
  const S P = (k.prr + k.ut * k.ut) * A1 + s1 * k.ut * k.B1;
  const S Q = k.un * A2 + s2 * k.B2;
  const S temp2 = (S)0.5 * k.ut * a7 * A1, temp3 = (S)0.5 * k.B1 * a6;
  const S R = temp2 + s1 * temp3;
  S rhob = Gb[0] * A2 * t1b + Gb[1] * A2 * P + Gb[2] * t1b * Q + Gb[3] * A2 * R;
  S A2b = Gb[0] * k.rho * t1b + Gb[1] * k.rho * P + Gb[3] * k.rho * R;
  S t1bb = Gb[0] * k.rho * A2 + Gb[2] * k.rho * Q;
  const S Pb = Gb[1] * k.rho * A2;
  const S Qb = Gb[2] * k.rho * t1b;
  const S Rb = Gb[3] * k.rho * A2;
  // temp4 = 0.5 rho un B2 t1b enters G3 with sign s2
  const S t4b = s2 * Gb[3];
  rhob += t4b * (S)0.5 * k.un * k.B2 * t1b;
  S unb = t4b * (S)0.5 * k.rho * k.B2 * t1b;
  S B2b = t4b * (S)0.5 * k.rho * k.un * t1b;
  t1bb += t4b * (S)0.5 * k.rho * k.un * k.B2;
  S utb = (S)0, prrb = (S)0, usqb = (S)0, A1b = (S)0, B1b = (S)0;
  // R = temp2 + s1 temp3
  utb += Rb * (S)0.5 * a7 * A1;
  prrb += Rb * (S)3.5 * k.ut * A1;
  usqb += Rb * (S)0.5 * k.ut * A1;
  A1b += Rb * (S)0.5 * k.ut * a7;
  const S t3b = s1 * Rb;
  B1b += t3b * (S)0.5 * a6;
  prrb += t3b * (S)3 * k.B1;
  usqb += t3b * (S)0.5 * k.B1;
  // Q = un A2 + s2 B2
  unb += Qb * A2;
  A2b += Qb * k.un;
  B2b += Qb * s2;
  // P = (prr + ut^2) A1 + s1 ut B1
  prrb += Pb * A1;
  utb += Pb * ((S)2 * k.ut * A1 + s1 * k.B1);
  A1b += Pb * (k.prr + k.ut * k.ut);
  B1b += Pb * s1 * k.ut;
  // t1b = ut A1 + s1 B1
  utb += t1bb * A1;
  A1b += t1bb * k.ut;
  B1b += t1bb * s1;
  kb.rho += rhob;
  kb.ut += utb;
  kb.un += unb;
  kb.prr += prrb;
  kb.usq += usqb;
  kb.B1 += B1b;
  kb.B2 += B2b;
  kb.e1 += (S)0.5 * s1 * A1b;
  kb.e2 += (S)0.5 * s2 * A2b;
}

// reverse of kin_fwd: adds d/d(qt) into qtb and d/d(nx, ny) into nxb, nyb (FP64 sums)
template <class S>
__host__ __device__ inline void kin_rev(const KinFT<S>& k, const KinBT<S>& kb, S nx, S ny, S qtb[4], double& nxb,
                                        double& nyb) {
  // the derivatives' divisions as products with three reciprocals: 1/rho and 1/sb
  // divided, 1/p = 2 beta / rho (p = rho tmp, tmp = 1/(2 beta)), 1/spb = (1/sb)/sqrt(pi)
  // and 1/beta = 2 tmp; the forward values above keep the primal's own operations
  const S irho = rv_rcp(k.rho), isb = rv_rcp(k.sb);
  const S ip = (S)2 * k.beta * irho, ispb = (S)INV_SQRT_PI * isb, ibeta = (S)2 * k.tmp;
  S utb = kb.ut, unb = kb.un, rhob = kb.rho;
  // usq = ut^2 + un^2
  utb += (S)2 * k.ut * kb.usq;
  unb += (S)2 * k.un * kb.usq;
  // prr = p / rho
  S pb = kb.prr * irho;
  rhob -= kb.prr * k.prr * irho;
  // e = erf(S), B = 0.5 exp(-S^2) / spb
  S S1b = kb.e1 * (S)TWO_BY_SQRT_PI * k.ex1 - (S)2 * k.S1 * k.B1 * kb.B1;
  S S2b = kb.e2 * (S)TWO_BY_SQRT_PI * k.ex2 - (S)2 * k.S2 * k.B2 * kb.B2;
  const S spbb = -(kb.B1 * k.B1 + kb.B2 * k.B2) * ispb;
  // S1 = ut sb, S2 = un sb
  utb += S1b * k.sb;
  unb += S2b * k.sb;
  const S sbb = S1b * k.ut + S2b * k.un;
  // sb = sqrt(b), spb = sqrt(pi b)
  const S bb = (S)0.5 * sbb * isb + (S)0.5 * (S)Consts<double>::PI * spbb * ispb;
  // b = 0.5 rho / p
  rhob += (S)0.5 * bb * ip;
  pb -= bb * k.b * ip;
  // ut = u1 ny - u2 nx, un = u1 nx + u2 ny
  S u1b = utb * ny + unb * nx;
  S u2b = -utb * nx + unb * ny;
  nxb += (double)(-utb * k.u2 + unb * k.u1);
  nyb += (double)(utb * k.u1 + unb * k.u2);
  // p = rho tmp
  rhob += pb * k.tmp;
  S tmpb = pb * k.rho;
  // rho = exp(T2), T2 = T1 - log(beta)/(gamma-1): d rho = rho (dT1 - dbeta / (beta (gamma-1)))
  const S T2b = rhob * k.rho;
  const S T1b = T2b;
  S betab = -T2b * ibeta * (S)INV_GAMMA_M1;
  // T1 = qt1 + beta (u1^2 + u2^2)
  qtb[0] += T1b;
  betab += T1b * (k.u1 * k.u1 + k.u2 * k.u2);
  u1b += T1b * (S)2 * k.beta * k.u1;
  u2b += T1b * (S)2 * k.beta * k.u2;
  // u1 = qt2 tmp, u2 = qt3 tmp
  qtb[1] += u1b * k.tmp;
  qtb[2] += u2b * k.tmp;
  tmpb += u1b * k.qt[1] + u2b * k.qt[2];
  // tmp = 0.5 / beta, beta = -0.5 qt4
  betab -= tmpb * k.tmp * ibeta;
  qtb[3] += (S)-0.5 * betab;
}

// reverse of venkat (limiters.F90:64-110), one component; the adjoints of the point
// record accumulate in FP64
template <class S>
__host__ __device__ inline void venkat_rev(S qt0, S q, S qmax, S qmin, S epsi, S phib, S& qt0b, double& qb,
                                           double& qmaxb, double& qminb) {
  const S dn = qt0 - q;
  if (m_abs(dn) <= Consts<S>::THRESH) return;
  const bool use_max = dn > (S)0;
  const S dp = (use_max ? qmax : qmin) - q;
  const S e2 = epsi * epsi;
  const S num = (dp * dp + e2) * dn + (S)2 * dn * dn * dp;
  const S D = dp * dp + (S)2 * dn * dn + dn * dp + e2;
  const S den = D * dn;
  S t, iden;
  if constexpr (sizeof(S) == 4) {
    iden = frcp(den);
    t = num * iden;
  } else {
    t = num / den;  // the primal's value, so the branch is the primal's
    iden = 1.0 / den;
  }
  if (!(t < (S)1)) return;
  const S numb = phib * iden;
  const S denb = -phib * t * iden;
  const S dnb = numb * ((dp * dp + e2) + (S)4 * dn * dp) + denb * (D + dn * ((S)4 * dn + dp));
  const S dpb = numb * ((S)2 * dp * dn + (S)2 * dn * dn) + denb * (dn * ((S)2 * dp + dn));
  qt0b += dnb;
  qb -= (double)dnb;
  if (use_max) qmaxb += (double)dpb;
  else qminb += (double)dpb;
  qb -= (double)dpb;
}

// Adjoints of one end's point record
struct EndB {
  double q[4], dx[4], dy[4], qmax[4], qmin[4];
};

__host__ __device__ inline void endb_zero(EndB& e) {
  for (int r = 0; r < 4; ++r) e.q[r] = e.dx[r] = e.dy[r] = e.qmax[r] = e.qmin[r] = 0.0;
}

// Adjoints of one end's point record with the reconstruction's increment kept as one
// number per component: d = delx dq_x + dely dq_y, so the dq adjoints are db*delx and
// db*dely (recomputed by whoever holds delx, dely)
struct EndD {
  double q[4], d[4], qmax[4], qmin[4];
};

__host__ __device__ inline void endd_zero(EndD& e) {
  for (int r = 0; r < 4; ++r) e.q[r] = e.d[r] = e.qmax[r] = e.qmin[r] = 0.0;
}

// reverse of reconstruct, keeping d's adjoint (see EndD); adds delx, dely adjoints.
// The reconstruction's arithmetic in S, its adjoints accumulated in FP64.
template <class S>
__host__ __device__ inline void reconstruct_rev_d(const V4<S>& q, const V4<S>& dx, const V4<S>& dy,
                                                  const V4<S>& qmax, const V4<S>& qmin, S eps, S delx, S dely,
                                                  const S qtb[4], EndD& eb, double& delxb, double& delyb) {
  for (int r = 0; r < 4; ++r) {
    const S d = delx * dx[r] + dely * dy[r];
    const S qt0 = q[r] - (S)0.5 * d;
    const S phi = venkat(qt0, q[r], qmax[r], qmin[r], eps);
    eb.q[r] += (double)qtb[r];
    const S phib = (S)-0.5 * d * qtb[r];
    S db = (S)-0.5 * phi * qtb[r];
    S qt0b = (S)0;
    venkat_rev<S>(qt0, q[r], qmax[r], qmin[r], eps, phib, qt0b, eb.q[r], eb.qmax[r], eb.qmin[r]);
    eb.q[r] += (double)qt0b;
    db += (S)-0.5 * qt0b;
    eb.d[r] += (double)db;
    delxb += (double)(db * dx[r]);
    delyb += (double)(db * dy[r]);
  }
}

// reverse of reconstruct (wall_fluxes.F90:79-86)
__host__ __device__ inline void reconstruct_rev(const V4d& q, const V4d& dx, const V4d& dy,
                                                const V4d& qmax, const V4d& qmin, double eps,
                                                double delx, double dely, const double qtb[4],
                                                EndB& eb, double& delxb, double& delyb) {
  EndD t;
  endd_zero(t);
  reconstruct_rev_d(q, dx, dy, qmax, qmin, eps, delx, dely, qtb, t, delxb, delyb);
  for (int r = 0; r < 4; ++r) {
    eb.q[r] += t.q[r];
    eb.dx[r] += t.d[r] * delx;
    eb.dy[r] += t.d[r] * dely;
    eb.qmax[r] += t.qmax[r];
    eb.qmin[r] += t.qmin[r];
  }
}

// One end's split fluxes for an edge (as end_fluxes<CLS>) from its KinF
template <int CLS, class S>
__host__ __device__ inline void end_fluxes_f(const KinFT<S>& f, bool xp, bool yp, S Gx[4], S Gy[4]) {
  end_fluxes<CLS, S>(kin_of(f), xp, yp, Gx, Gy);
}

template <int CLS, class S>
__host__ __device__ inline void end_fluxes_rev(const KinFT<S>& f, bool xp, bool yp, const S Gxb[4],
                                               const S Gyb[4], KinBT<S>& kb) {
  if (CLS == INTERIOR) {
    flux_gx_rev(f, xp, Gxb, kb);
    flux_gy_rev(f, yp, Gyb, kb);
  } else {
    const bool a2p = (CLS == OUTER);
    flux_quad_rev(f, xp, a2p, Gxb, kb);
    flux_gy_rev(f, a2p, Gyb, kb);
  }
}

// The half coefficients as products per half (0 xpos, 1 xneg, 2 ypos, 3 yneg): A = c1 obd,
// B = sxy obd, so an edge's x coefficient is dsw A - dnw B of its x half and its y
// coefficient dnw A - dsw B of its y half (the flux reverse's form of coef_x / coef_y;
// its adjoints have the same shape)
struct HalfAB {
  double A[4], B[4];
};

__host__ __device__ inline double coef_x(const HalfAB& h, bool first, double dsw, double dnw) {
  const double a = first ? h.A[0] : h.A[1];
  const double b = first ? h.B[0] : h.B[1];
  return dsw * a - dnw * b;
}
__host__ __device__ inline double coef_y(const HalfAB& h, bool first, double dsw, double dnw) {
  const double a = first ? h.A[2] : h.A[3];
  const double b = first ? h.B[2] : h.B[3];
  return dnw * a - dsw * b;
}

// reverse of coef_x / coef_y (HalfAB form) for half d with the coefficient's adjoint cb
__host__ __device__ inline void coef_rev_ab(const HalfAB& h, int d, bool ydir, double dsw, double dnw, double cb,
                                            HalfAB& hb, double& dswb, double& dnwb) {
  double A = 0, B = 0;
  for (int j = 0; j < 4; ++j)
    if (j == d) {
      A = h.A[j];
      B = h.B[j];
    }
  // c = a A - bq B
  const double a = ydir ? dnw : dsw, bq = ydir ? dsw : dnw;
  for (int j = 0; j < 4; ++j)
    if (j == d) {
      hb.A[j] += cb * a;
      hb.B[j] += -cb * bq;
    }
  const double ab = cb * A, bqb = -cb * B;
  if (ydir) {
    dnwb += ab;
    dswb += bqb;
  } else {
    dswb += ab;
    dnwb += bqb;
  }
}

// ---------------------------------------------------------------------------------
// state update

// reverse of prim_to_cons<double> (state_update.F90:144-163)
__host__ __device__ inline void prim_to_cons_rev(const V4d& p, double nx, double ny, const double Ub[4],
                                                 double pb[4], double& nxb, double& nyb) {
  const double rho = p.x, t1 = rho * p.y, t2 = rho * p.z;
  // U4 = 2.5 p + 0.5 (t1^2 + t2^2) / rho; U2 = t1 ny - t2 nx; U3 = t1 nx + t2 ny
  const double t1b = Ub[1] * ny + Ub[2] * nx + Ub[3] * t1 / rho;
  const double t2b = -Ub[1] * nx + Ub[2] * ny + Ub[3] * t2 / rho;
  nxb += -Ub[1] * t2 + Ub[2] * t1;
  nyb += Ub[1] * t1 + Ub[2] * t2;
  pb[0] += Ub[0] - Ub[3] * 0.5 * (t1 * t1 + t2 * t2) / (rho * rho) + t1b * p.y + t2b * p.z;
  pb[1] += t1b * rho;
  pb[2] += t2b * rho;
  pb[3] += 2.5 * Ub[3];
}

// reverse of cons_ubar<double> (state_update.F90:239-294)
__host__ __device__ inline void cons_ubar_rev(const V4d& p, double nx, double ny, const Freestream& fs,
                                              const double Ubb[4], double pb[4], double& nxb,
                                              double& nyb) {
  const double rho_inf = 1.0, pr_inf = 1.0 / 1.4, gm1 = 1.4 - 1.0, pi = Consts<double>::PI;
  const double u1i = fs.u1_inf, u2i = fs.u2_inf;
  const double tx = ny, ty = -nx;
  const double u1ir = u1i * tx + u2i * ty, u2ir = u1i * nx + u2i * ny;
  const double e_inf = pr_inf / (rho_inf * gm1) + 0.5 * (u1ir * u1ir + u2ir * u2ir);
  const double beta_i = (0.5 * rho_inf) / pr_inf;
  const double S2i = u2ir * sqrt(beta_i);
  const double exi = exp(-S2i * S2i);
  const double B2i = exi / (2.0 * sqrt(pi * beta_i));
  const double A2ni = 0.5 * (1.0 - erf(S2i));
  const double rho = p.x, u1 = p.y, u2 = p.z, pr = p.w;
  const double u1r = u1 * tx + u2 * ty, u2r = u1 * nx + u2 * ny;
  const double e = pr / (rho * gm1) + 0.5 * (u1r * u1r + u2r * u2r);
  const double beta = rho / (2.0 * pr);
  const double S2 = u2r * sqrt(beta);
  const double ex = exp(-S2 * S2);
  const double B2 = ex / (2.0 * sqrt(pi * beta));
  const double A2p = 0.5 * (1.0 + erf(S2));
  // freestream half
  const double A2nib = rho_inf * (Ubb[0] + Ubb[1] * u1ir + Ubb[2] * u2ir + Ubb[3] * e_inf);
  double u1irb = Ubb[1] * rho_inf * A2ni;
  double u2irb = Ubb[2] * rho_inf * A2ni - Ubb[3] * 0.5 * rho_inf * B2i;
  const double B2ib = -Ubb[2] * rho_inf - Ubb[3] * 0.5 * rho_inf * u2ir;
  const double e_infb = Ubb[3] * rho_inf * A2ni;
  u1irb += e_infb * u1ir;
  u2irb += e_infb * u2ir;
  const double S2ib = -A2nib * 0.5 * TWO_BY_SQRT_PI * exi + B2ib * B2i * (-2.0 * S2i);
  u2irb += S2ib * sqrt(beta_i);
  nyb += u1irb * u1i + u2irb * u2i;
  nxb += -u1irb * u2i + u2irb * u1i;
  // interior half
  double rhob = Ubb[0] * A2p + Ubb[1] * u1r * A2p + Ubb[2] * (u2r * A2p + B2) + Ubb[3] * (A2p * e + 0.5 * u2r * B2);
  const double A2pb = Ubb[0] * rho + Ubb[1] * rho * u1r + Ubb[2] * rho * u2r + Ubb[3] * rho * e;
  double u1rb = Ubb[1] * rho * A2p;
  double u2rb = Ubb[2] * rho * A2p + Ubb[3] * 0.5 * rho * B2;
  const double B2b = Ubb[2] * rho + Ubb[3] * 0.5 * rho * u2r;
  const double eb = Ubb[3] * rho * A2p;
  double prb = eb / (rho * gm1);
  rhob -= eb * pr / (rho * rho * gm1);
  u1rb += eb * u1r;
  u2rb += eb * u2r;
  double S2b = A2pb * 0.5 * TWO_BY_SQRT_PI * ex + B2b * B2 * (-2.0 * S2);
  double betab = -B2b * B2 / (2.0 * beta);
  u2rb += S2b * sqrt(beta);
  betab += S2b * u2r / (2.0 * sqrt(beta));
  rhob += betab / (2.0 * pr);
  prb -= betab * beta / pr;
  pb[0] += rhob;
  pb[1] += u1rb * tx + u2rb * nx;
  pb[2] += u1rb * ty + u2rb * ny;
  pb[3] += prb;
  nxb += -u1rb * u2 + u2rb * u1;
  nyb += u1rb * u1 + u2rb * u2;
}

// reverse of update_point<CLS, double> (state_update.F90:10-141): from the adjoint of
// the new prim `pn` (pnb), adds adjoints of the old prim, prim_old (stage 3),
// flux_res and the normal.  The post-update conserved state is rebuilt from pn.
template <int CLS>
__host__ __device__ inline void update_rev(const V4d& p, const V4d& pold, const V4d& pn, double nx,
                                           double ny, int rk, double euler, const Freestream& fs,
                                           const V4d& pnb, double pb[4], double poldb[4],
                                           double frb[4], double& nxb, double& nyb) {
  const double obt = 1.0 / 3.0, tbt = 2.0 / 3.0, half_euler = 0.5 * euler;
  const double W1 = pn.x, W2 = pn.y * W1, W3 = pn.z * W1;
  const double Un1 = W2 * ny - W3 * nx;
  const double Un2 = (CLS == WALL) ? 0.0 : W2 * nx + W3 * ny;
  // prim = (W1, W2 t, W3 t, 0.4 (W4 - 0.5 t (W2^2 + W3^2))), t = 1/W1
  const double t = 1.0 / W1;
  const double W4b = 0.4 * pnb.w;
  const double W2b = pnb.y * t - 0.4 * pnb.w * t * W2;
  const double W3b = pnb.z * t - 0.4 * pnb.w * t * W3;
  const double tb = pnb.y * W2 + pnb.z * W3 - 0.2 * pnb.w * (W2 * W2 + W3 * W3);
  const double W1b = pnb.x - tb * t * t;
  // W2 = Un1 ny + Un2 nx, W3 = Un2 ny - Un1 nx
  double Unb[4] = {W1b, W2b * ny - W3b * nx, W2b * nx + W3b * ny, W4b};
  nyb += W2b * Un1 + W3b * Un2;
  nxb += W2b * Un2 - W3b * Un1;
  if (CLS == WALL) Unb[2] = 0.0;  // no-penetration: U(3) = 0 after the update
  double Ub[4], Uob[4] = {0, 0, 0, 0};
  for (int r = 0; r < 4; ++r) {
    if (rk != 3) {
      Ub[r] = Unb[r];
      frb[r] += -half_euler * Unb[r];
    } else {
      Uob[r] = tbt * Unb[r];
      Ub[r] = obt * Unb[r];
      frb[r] += -0.5 * obt * Unb[r];
    }
  }
  if (CLS == OUTER) {
    cons_ubar_rev(p, nx, ny, fs, Ub, pb, nxb, nyb);
    if (rk == 3) cons_ubar_rev(pold, nx, ny, fs, Uob, poldb, nxb, nyb);
  } else {
    prim_to_cons_rev(p, nx, ny, Ub, pb, nxb, nyb);
    if (rk == 3) prim_to_cons_rev(pold, nx, ny, Uob, poldb, nxb, nyb);
  }
}

// reverse of q_of_prim<double> (q_variables.F90:8-37)
__host__ __device__ inline void q_of_prim_rev(const V4d& p, const double qb[4], double pb[4]) {
  const double rho = p.x, u1 = p.y, u2 = p.z, pr = p.w;
  const double beta = 0.5 * rho / pr;
  const double betab = qb[0] * (2.5 / beta - (u1 * u1 + u2 * u2)) + 2.0 * u1 * qb[1] + 2.0 * u2 * qb[2] - 2.0 * qb[3];
  pb[0] += qb[0] / rho + betab * 0.5 / pr;
  pb[3] += -betab * beta / pr;
  pb[1] += -2.0 * beta * u1 * qb[0] + 2.0 * beta * qb[1];
  pb[2] += -2.0 * beta * u2 * qb[0] + 2.0 * beta * qb[2];
}

// reverse of wave_speed (state_update.F90:223-225): mod_u + 3 sqrt(p/rho)
__host__ __device__ inline void wave_speed_rev(const V4d& p, double cb, double pb[4]) {
  const double mod_u = sqrt(p.y * p.y + p.z * p.z);
  if (mod_u > 0.0) {
    pb[1] += cb * p.y / mod_u;
    pb[2] += cb * p.z / mod_u;
  }
  const double z = p.w / p.x;
  const double zb = cb * 3.0 / (2.0 * sqrt(z));
  pb[3] += zb / p.x;
  pb[0] -= zb * z / p.x;
}

#pragma clang fp contract(off)

}  // namespace mf
