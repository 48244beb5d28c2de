// Host-side wrappers around the device header's forward and reverse functions, so
// the test suite can check each reverse function against finite differences of its
// forward on the CPU (TEST INFRASTRUCTURE; not part of the shipped library).
#include "../../mfcfd_amd/csrc/lskum_adjoint.h"

using namespace mf;

static void flux_by_kind(int kind, const Kin<double>& k, double G[4]) {
  switch (kind) {
    case 0: flux_gx(k, true, G); break;
    case 1: flux_gx(k, false, G); break;
    case 2: flux_gy(k, true, G); break;
    case 3: flux_gy(k, false, G); break;
    case 4: flux_quad(k, false, false, G); break;
    case 5: flux_quad(k, true, false, G); break;
    case 6: flux_quad(k, true, true, G); break;
    default: flux_quad(k, false, true, G); break;
  }
}

static void flux_rev_by_kind(int kind, const KinF& f, const double Gb[4], KinB& kb) {
  switch (kind) {
    case 0: flux_gx_rev(f, true, Gb, kb); break;
    case 1: flux_gx_rev(f, false, Gb, kb); break;
    case 2: flux_gy_rev(f, true, Gb, kb); break;
    case 3: flux_gy_rev(f, false, Gb, kb); break;
    case 4: flux_quad_rev(f, false, false, Gb, kb); break;
    case 5: flux_quad_rev(f, true, false, Gb, kb); break;
    case 6: flux_quad_rev(f, true, true, Gb, kb); break;
    default: flux_quad_rev(f, false, true, Gb, kb); break;
  }
}

extern "C" {

// in: qt[4], nx, ny  ->  G[4]
void ac_flux_fwd(int kind, const double* in, double* G) {
  const Kin<double> k = kin_of_qtilde<double>(in, in[4], in[5]);
  flux_by_kind(kind, k, G);
}

void ac_flux_rev(int kind, const double* in, const double* Gb, double* inb) {
  KinF f;
  kin_fwd(in, in[4], in[5], f);
  KinB kb;
  kinb_zero(kb);
  flux_rev_by_kind(kind, f, Gb, kb);
  double qtb[4] = {0, 0, 0, 0}, nxb = 0, nyb = 0;
  kin_rev(f, kb, in[4], in[5], qtb, nxb, nyb);
  for (int r = 0; r < 4; ++r) inb[r] = qtb[r];
  inb[4] = nxb;
  inb[5] = nyb;
}

// in: q[4] dx[4] dy[4] qmax[4] qmin[4] eps delx dely  ->  qt[4]
void ac_recon_fwd(const double* in, double* qt) {
  V4d q{in[0], in[1], in[2], in[3]}, dx{in[4], in[5], in[6], in[7]}, dy{in[8], in[9], in[10], in[11]};
  V4d mx{in[12], in[13], in[14], in[15]}, mn{in[16], in[17], in[18], in[19]};
  reconstruct<double>(q, dx, dy, mx, mn, in[20], in[21], in[22], qt);
}

void ac_recon_rev(const double* in, const double* qtb, double* inb) {
  V4d q{in[0], in[1], in[2], in[3]}, dx{in[4], in[5], in[6], in[7]}, dy{in[8], in[9], in[10], in[11]};
  V4d mx{in[12], in[13], in[14], in[15]}, mn{in[16], in[17], in[18], in[19]};
  EndB e;
  endb_zero(e);
  double delxb = 0, delyb = 0;
  reconstruct_rev(q, dx, dy, mx, mn, in[20], in[21], in[22], qtb, e, delxb, delyb);
  for (int r = 0; r < 4; ++r) {
    inb[r] = e.q[r];
    inb[4 + r] = e.dx[r];
    inb[8 + r] = e.dy[r];
    inb[12 + r] = e.qmax[r];
    inb[16 + r] = e.qmin[r];
  }
  inb[20] = 0.0;
  inb[21] = delxb;
  inb[22] = delyb;
}

// in: p[4] pold[4] fr[4] nx ny; fs: u1_inf, u2_inf  ->  pn[4]
void ac_update_fwd(int cls, int rk, double euler, const double* in, const double* fs, double* pn) {
  V4d p{in[0], in[1], in[2], in[3]}, po{in[4], in[5], in[6], in[7]};
  const double fr[4] = {in[8], in[9], in[10], in[11]};
  Freestream f{fs[0], fs[1]};
  double rs;
  V4d o;
  if (cls == WALL) o = update_point<WALL, double>(p, po, fr, in[12], in[13], rk, euler, f, &rs);
  else if (cls == OUTER) o = update_point<OUTER, double>(p, po, fr, in[12], in[13], rk, euler, f, &rs);
  else o = update_point<INTERIOR, double>(p, po, fr, in[12], in[13], rk, euler, f, &rs);
  pn[0] = o.x;
  pn[1] = o.y;
  pn[2] = o.z;
  pn[3] = o.w;
}

void ac_update_rev(int cls, int rk, double euler, const double* in, const double* fs, const double* pnb,
                   double* inb) {
  double pn4[4];
  ac_update_fwd(cls, rk, euler, in, fs, pn4);
  V4d p{in[0], in[1], in[2], in[3]}, po{in[4], in[5], in[6], in[7]}, pn{pn4[0], pn4[1], pn4[2], pn4[3]};
  V4d b{pnb[0], pnb[1], pnb[2], pnb[3]};
  Freestream f{fs[0], fs[1]};
  double pb[4] = {0, 0, 0, 0}, pob[4] = {0, 0, 0, 0}, frb[4] = {0, 0, 0, 0}, nxb = 0, nyb = 0;
  if (cls == WALL) update_rev<WALL>(p, po, pn, in[12], in[13], rk, euler, f, b, pb, pob, frb, nxb, nyb);
  else if (cls == OUTER) update_rev<OUTER>(p, po, pn, in[12], in[13], rk, euler, f, b, pb, pob, frb, nxb, nyb);
  else update_rev<INTERIOR>(p, po, pn, in[12], in[13], rk, euler, f, b, pb, pob, frb, nxb, nyb);
  for (int r = 0; r < 4; ++r) {
    inb[r] = pb[r];
    inb[4 + r] = pob[r];
    inb[8 + r] = frb[r];
  }
  inb[12] = nxb;
  inb[13] = nyb;
}

void ac_qprim_fwd(const double* p, double* q) {
  const V4d o = q_of_prim<double>(V4d{p[0], p[1], p[2], p[3]});
  q[0] = o.x;
  q[1] = o.y;
  q[2] = o.z;
  q[3] = o.w;
}

void ac_qprim_rev(const double* p, const double* qb, double* pb) {
  for (int r = 0; r < 4; ++r) pb[r] = 0.0;
  q_of_prim_rev(V4d{p[0], p[1], p[2], p[3]}, qb, pb);
}

// venkat_limiter reverse at one point for all four components: in qt[4] q[4] qmax[4]
// qmin[4] eps, phib[4] -> out qtb[4] qb[4] qmaxb[4] qminb[4] (16 values)
void ac_venkat_rev(const double* in, const double* phib, double* out) {
  for (int j = 0; j < 16; ++j) out[j] = 0.0;
  for (int r = 0; r < 4; ++r)
    venkat_rev(in[r], in[4 + r], in[8 + r], in[12 + r], in[16], phib[r], out[r], out[4 + r], out[8 + r],
               out[12 + r]);
}

double ac_wave_fwd(const double* p) { return wave_speed<double>(V4d{p[0], p[1], p[2], p[3]}); }

void ac_wave_rev(const double* p, double cb, double* pb) {
  for (int r = 0; r < 4; ++r) pb[r] = 0.0;
  wave_speed_rev(V4d{p[0], p[1], p[2], p[3]}, cb, pb);
}

}  // extern "C"
