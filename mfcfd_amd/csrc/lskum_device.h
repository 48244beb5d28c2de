// Device code of the LSKUM fixed-point iteration for gfx950 (CDNA4, wave64).
//
// LPP lanes per point (1 by default); a lane walks the point's neighbour slots lane,
// lane + LPP, ... in the reference's conn order, so with one lane each least-squares
// sum accumulates in the reference's order (q_variables.F90:71-102,
// wall_fluxes.F90:55-99 and friends) and with more the lanes' partial sums are combined
// by shuffles.  The staged walks read the same slots from a block's halo in LDS.
//
// Device layout (points renumbered on the host: wall | outer | interior, each class
// in Hilbert-curve order):
//   xy, nrm        double2 [N]             coordinates, unit normals
//   eps            S       [N]             Venkatakrishnan epsilon (VL_CONST*min_dist)**3
//   conn           uint32  [W][N]          slot-major: (neighbour << 4) | direction bits
//   lsq_geo        double4 [N]             (sum dx^2, sum dy^2, sum dx dy, 1/det) of the
//                                          full stencil, point-fixed geometry (mixed
//                                          path: float4 lsq_geo_f, the solve's products)
//   dg_c1/sxy/obd  double4 [N]             per half-stencil (xpos, xneg, ypos, yneg) the
//                                          2x2 solve coefficients of the flux derivative
//   prim, q        V4<S>  [N]
//   dq, qm         V4<S>  [2N]             dq: [2i] d/dx, [2i+1] d/dy; qm: [2i] max, [2i+1] min
//   halos (Halo)   per block of 64 (LSQ, inner) or 32 (flux) points: the halo's point
//                  ids and [W][N] slot words (uint8 halo positions in 64-point blocks;
//                  uint16 position | half-stencil bits in the flux kernel's)
//
// S is double for the FP64 path and float for the mixed path; least-squares sums and
// solves, flux-derivative accumulations, residual norms and force sums are double in
// both.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mf {

constexpr int WALL = 0, OUTER = 1, INTERIOR = 2;
constexpr unsigned XPOS = 1u, XNEG = 2u, YPOS = 4u, YNEG = 8u;
constexpr int BLOCK = 256;
// Staged stencil walks: blocks of consecutive points (STAGE_P for the LSQ and inner
// kernels, STAGE_PF for the flux kernel, whose neighbour records are larger); a
// lanes-per-point value with the LPP_STAGED bit set selects a kernel's staged form
// (LPP_PULL: the adjoint inner reverse's pull form, also staged)
constexpr int STAGE_P = 64;
constexpr int STAGE_PF = 32;
constexpr int LPP_STAGED = 16;
constexpr int LPP_PULL = 64;

// One staging layout (blocks of P points).  Block b's halo is ids[ptr[b] .. ptr[b+1]):
// its own points first, then every other neighbour of its points once.  loc [W][stride]
// holds each slot's position in its block's halo (low 12 bits) and the slot's
// half-stencil bits (high 4); rows are padded to whole blocks, so a block's row segment
// is 16-byte aligned.
struct Halo {
  const int32_t* ptr = nullptr;
  const int32_t* ids = nullptr;
  const uint16_t* loc = nullptr;  // [W][stride] position | half-stencil bits << 12
  const uint8_t* loc8 = nullptr;  // [W][stride] position, 255 past hcap (capped halos)
  // halo entries a block holds in LDS, a multiple of 8 (0: staging not available).
  // A block whose halo is larger reads its entries past hcap from global memory.
  int32_t hcap = 0;
  int32_t stride = 0;  // N rounded up to a multiple of P
};

template <class T>
struct alignas(4 * sizeof(T)) V4 {
  T x, y, z, w;
  __host__ __device__ T& operator[](int r) { return (&x)[r]; }
  __host__ __device__ const T& operator[](int r) const { return (&x)[r]; }
};
using V4d = V4<double>;

// componentwise difference
template <class T>
__host__ __device__ inline V4<T> operator-(const V4<T>& a, const V4<T>& b) {
  return V4<T>{a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w};
}

struct alignas(16) D2 {
  double x, y;
};

// Scalar math per precision (ocml); FP64 follows libm semantics to within an ulp.
__host__ __device__ inline double m_exp(double v) { return exp(v); }
__host__ __device__ inline float m_exp(float v) { return expf(v); }
__host__ __device__ inline double m_log(double v) { return log(v); }
__host__ __device__ inline float m_log(float v) { return logf(v); }
__host__ __device__ inline double m_erf(double v) { return erf(v); }
__host__ __device__ inline float m_erf(float v) { return erff(v); }
__host__ __device__ inline double m_sqrt(double v) { return sqrt(v); }
__host__ __device__ inline float m_sqrt(float v) { return sqrtf(v); }
__host__ __device__ inline double m_abs(double v) { return fabs(v); }
__host__ __device__ inline float m_abs(float v) { return fabsf(v); }

template <class S> struct Consts;
template <> struct Consts<double> {
  // parameter.F90:11 pi = 4 atan(1); limiters.F90:78 REAL*4 literal 10e-6 as REAL*8
  static constexpr double PI = 3.141592653589793;
  static constexpr double THRESH = (double)10e-6f;
  static constexpr double GAMMA_M1 = 1.4 - 1.0;
};
template <> struct Consts<float> {
  static constexpr float PI = 3.14159265f;
  static constexpr float THRESH = 10e-6f;
  static constexpr float GAMMA_M1 = 0.4f;
};

struct Freestream;

struct Geom {
  const D2* xy;
  const D2* nrm;
  const int32_t* nbhs;
  const uint32_t* conn;
  const V4d* lsq_geo;
  const V4d* dg_c1;
  const V4d* dg_sxy;
  const V4d* dg_obd;
  int32_t n, nw, no;  // points; wall count; outer count (interior = rest)
  int32_t first_order;  // 1: q-tilde = q at both ends (no reconstruction)
  // a batch of independent cases in one cloud: each point's case and the cases'
  // freestreams (null for a single case: the kernels' Freestream argument)
  const int32_t* case_id = nullptr;
  const Freestream* fs_tab = nullptr;
  // staged walks: hl[0] in blocks of STAGE_P points, hl[1] in blocks of STAGE_PF
  Halo hl[2];
  int32_t width = 0;  // W, slots per point of conn and of the halos' loc
  // mixed path: lsq_geo as the FP32 solve coefficients (sum dy^2 / det, sum dx dy / det,
  // sum dx^2 / det): d/dx = a Sx - b Sy, d/dy = c Sy - b Sx for the sums Sx, Sy
  const V4<float>* lsq_geo_f = nullptr;
  // mixed path: the half-stencil coefficients in FP32, [2i] c1 * obd and [2i+1] sxy * obd
  // of the four halves (HalfCoefF)
  const V4<float>* dg_f = nullptr;
  // 1 when every interior normal is exactly (0, 1) (point_normals.F90:91-96, the
  // default): an interior point's (t, n) frame is then the (x, y) frame
  int32_t axis_interior = 0;
};

// halt codes of Ctl
constexpr int32_t HALT_NAN = 1;      // a NaN residue (q_lskum.F90:71): the run stops, no error
constexpr int32_t HALT_BARRIER = 2;  // the resident sweeps' grid barrier gave up: an error

struct Ctl {
  int32_t it;      // iterations completed in this run (index into the histories)
  int32_t halt;    // HALT_NAN or HALT_BARRIER once set; every later kernel returns at once
  double res_old;  // fpi_solver.F90:65-67
  int32_t first_it;  // iteration number of the run's first iteration (residue rule t <= 2)
  int32_t reserved;
};

// Grid barrier of the resident inner sweeps.  Blocks arrive at one of SWEEP_GROUPS
// counters (block b at counter b mod groups, each on its own 128-byte line, so arrivals
// of different groups do not contend); each group's last arrival arrives at `top`, and
// the last of those advances `gen`, which the waiting blocks poll.  Counters return to
// zero at every episode, so the state serves any number of launches without a reset;
// `abort` is set when a wait gives up, and then the state needs a reset (zero fill).
constexpr int SWEEP_GROUPS = 64;
struct SweepBar {
  uint32_t group[SWEEP_GROUPS * 32];
  uint32_t top, gen, abort, pad[29];
};

// Blocks are dealt round-robin to the 8 XCDs; remap so each XCD (own L2) walks one
// contiguous range of the Hilbert-ordered points and its neighbour gathers stay local.
__device__ inline int xcd_block(int b, int nb) {
  const int q = nb >> 3, r = nb & 7, x = b & 7, local = b >> 3;
  return x * q + (x < r ? x : r) + local;
}

// Every stencil kernel is launched with BLOCK threads; the constant spares each wave the
// load of the dispatch packet's workgroup size before its first address is known.
__device__ inline int point_id() { return xcd_block(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x; }

// Lanes per point: a point's stencil loop can be split over LPP consecutive lanes of a
// wave (edge s goes to lane s mod LPP) and the partial sums combined with shuffles.
// More lanes per point give more, shorter waves, which fills the chip on small clouds
// and shortens the tail of a launch; LPP = 1 keeps the reference's summation order.
template <int LPP>
__device__ inline double group_sum(double v) {
#pragma unroll
  for (int off = LPP / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
template <int LPP, class T>
__device__ inline T group_max(T v) {
#pragma unroll
  for (int off = LPP / 2; off > 0; off >>= 1) {
    const T o = __shfl_xor(v, off, 64);
    v = (o > v) ? o : v;
  }
  return v;
}
template <int LPP, class T>
__device__ inline T group_min(T v) {
#pragma unroll
  for (int off = LPP / 2; off > 0; off >>= 1) {
    const T o = __shfl_xor(v, off, 64);
    v = (o < v) ? o : v;
  }
  return v;
}

// A lane's stencil walk (slots lane, lane + LPP, ...) taken U slots at a time: the U
// connectivity words are loaded together, then the U neighbour records (load(word)),
// then each record is used in slot order (use(record, word, valid)).  One slot at a
// time, each edge costs two dependent memory round trips (word, then record); batched,
// U edges share them.  Slots past nb load the point's own record (self) and come with
// valid = false: use() is called for every slot, so the compiler cannot sink a record's
// load into a branch (which would serialise it again), and must leave its sums unchanged
// (select, not add) when !valid, so a lane's sums keep the one-slot walk's order and
// values bit for bit.
template <int U, int LPP, class Load, class Use>
__device__ __forceinline__ void walk_stencil(const uint32_t* __restrict__ cp, int n, int nb,
                                             int lane, uint32_t self, Load&& load, Use&& use) {
  for (int s0 = lane; s0 < nb; s0 += U * LPP) {
    uint32_t e[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int s = s0 + u * LPP;
      e[u] = s < nb ? cp[(size_t)s * n] : self;
    }
    decltype(load(self)) rec[U];
#pragma unroll
    for (int u = 0; u < U; ++u) rec[u] = load(e[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) use(rec[u], e[u], s0 + u * LPP < nb);
  }
}

// The staged form of a stencil walk.  A block holds P consecutive points, LPP lanes
// each, and first copies its halo (fill(j, k): point k's records into the caller's LDS
// arrays at halo position j) and its points' 8-bit slot words into LDS; after one
// barrier every lane walks its slots (lane, lane + LPP, ...) with use(j, s) reading halo
// position j for slot s.  Each neighbour record then crosses from L2 once per block
// instead of once per edge (about 2.8 instead of 23.5 record loads per point at 64
// points per block), and the walk's loads are LDS reads.  A lane's slots and their order
// are those of the global walk, so its sums are bit-identical.
//
// LDS holds the first h.hcap halo entries.  The few blocks with larger halos (at 64
// points per block, those of wall and outer points, whose neighbours lie along a curve
// rather than around the block) read the neighbours past it from global memory: such a
// slot's word is 255 and it goes to use_global(k, s) with k the neighbour's point id from
// the point's connectivity.  The check is made per block, so the other blocks walk
// exactly as before.
struct NoHook {
  __device__ void operator()() const {}
};

// ready() runs after the barrier, before the walk (the block's records are in LDS)
template <int P, int LPP, class Fill, class Use, class UseGlobal, class Ready = NoHook>
__device__ __forceinline__ void stage_and_walk(const Geom& g, const Halo& h, int b, int nb, uint8_t* s_cl,
                                               Fill&& fill, Use&& use, UseGlobal&& use_global,
                                               Ready&& ready = Ready{}) {
  const int h0 = h.ptr[b], hn = h.ptr[b + 1] - h0;
  const int hl = hn < h.hcap ? hn : h.hcap;
  for (int j = threadIdx.x; j < hl; j += P * LPP) fill(j, h.ids[h0 + j]);
  // the block's slot words: per slot a P-byte row segment in 16-byte loads
  constexpr int V = P / 16;
  const uint4* src = (const uint4*)(h.loc8 + (size_t)b * P);
  for (int j = threadIdx.x; j < g.width * V; j += P * LPP) {
    const int s = j / V;
    ((uint4*)s_cl)[j] = src[(size_t)s * (h.stride / 16) + (j - s * V)];
  }
  __syncthreads();
  ready();
  const int li = threadIdx.x / LPP;
  if (hn <= h.hcap) {
    // the next slot's halo position is read while this edge is evaluated
    int s = threadIdx.x % LPP;
    int jn = s < nb ? s_cl[s * P + li] : 0;
    for (; s < nb; s += LPP) {
      const int j = jn;
      if (s + LPP < nb) jn = s_cl[(s + LPP) * P + li];
      use(j, s);
    }
  } else {
    const int i = b * P + li;
    for (int s = threadIdx.x % LPP; s < nb; s += LPP) {
      const int j = s_cl[s * P + li];
      if (j < h.hcap) use(j, s);
      else use_global((int)(g.conn[(size_t)s * g.n + i] >> 4), s);
    }
  }
}

template <class S>
__device__ inline S to_s(double v) { return (S)v; }

// ---------------------------------------------------------------------------------
// q_variables.F90:8-37 (one point)
template <class S>
__host__ __device__ inline V4<S> q_of_prim(const V4<S>& p) {
  S rho = p.x, u1 = p.y, u2 = p.z, pr = p.w;
  S beta = (S)0.5 * rho / pr;
  V4<S> q;
  q.x = m_log(rho) + (m_log(beta) * (S)2.5) - beta * (u1 * u1 + u2 * u2);
  S two_times_beta = (S)2.0 * beta;
  q.y = two_times_beta * u1;
  q.z = two_times_beta * u2;
  q.w = -two_times_beta;
  return q;
}

// state_update.F90:223-225: mod_u + 3 sqrt(p/rho) of a neighbour, the denominator of
// its local time step (stored per point so func_delta gathers one value)
template <class S>
__host__ __device__ inline double wave_speed(const V4<S>& p) {
  double rho = p.x, u1 = p.y, u2 = p.z, pr = p.w;
  double mod_u = sqrt(u1 * u1 + u2 * u2);
  return mod_u + 3.0 * sqrt(pr / rho);
}

// limiters.F90:64-110 venkat_limiter, one component
template <class S>
__host__ __device__ inline S venkat(S qt, S qk, S qmax, S qmin, S epsi) {
  S del_neg = qt - qk;
  if (m_abs(del_neg) <= Consts<S>::THRESH) return (S)1.0;
  S del_pos = (del_neg > (S)0) ? qmax - qk : qmin - qk;
  S num = (del_pos * del_pos) + (epsi * epsi);
  num = num * del_neg + (S)2.0 * del_neg * del_neg * del_pos;
  S den = del_pos * del_pos + (S)2.0 * del_neg * del_neg;
  den = den + del_neg * del_pos + epsi * epsi;
  den = den * del_neg;
  S t = num / den;
  return (t < (S)1.0) ? t : (S)1.0;
}

// venkat with del_pos's two candidates qmax - qk, qmin - qk (pmax, pmin) and epsi^2
// (e2) given, as the stencil kernels precompute them once per point: the same
// operations in the same order, so the same bits
template <class S>
__host__ __device__ inline S venkat_pre(S qt, S qk, S pmax, S pmin, S e2) {
  S del_neg = qt - qk;
  if (m_abs(del_neg) <= Consts<S>::THRESH) return (S)1.0;
  S del_pos = (del_neg > (S)0) ? pmax : pmin;
  S num = (del_pos * del_pos) + e2;
  num = num * del_neg + (S)2.0 * del_neg * del_neg * del_pos;
  S den = del_pos * del_pos + (S)2.0 * del_neg * del_neg;
  den = den + del_neg * del_pos + e2;
  den = den * del_neg;
  S t = num / den;
  return (t < (S)1.0) ? t : (S)1.0;
}

// State at one end of an edge after the limited reconstruction, in the point's
// (t, n) frame, with the shared pieces of every split flux (split_fluxes.F90:19-31,
// quadrant_fluxes.F90:219-234).
template <class S>
struct Kin {
  S rho, ut, un, pr_by_rho, u_sqr, B1, B2, e1, e2;
};

// q_variables.F90:230-256 qtilde_to_primitive, then the flux preliminaries
template <class S>
__host__ __device__ inline Kin<S> kin_of_qtilde(const S qt[4], S nx, S ny) {
  S beta = -qt[3] * (S)0.5;
  S tmp = (S)0.5 / beta;
  S u1 = qt[1] * tmp;
  S u2 = qt[2] * tmp;
  S temp1 = qt[0] + beta * (u1 * u1 + u2 * u2);
  S temp2 = temp1 - (m_log(beta) / Consts<S>::GAMMA_M1);
  S rho = m_exp(temp2);
  S pr = rho * tmp;
  Kin<S> k;
  S tx = ny, ty = -nx;
  k.ut = u1 * tx + u2 * ty;
  k.un = u1 * nx + u2 * ny;
  S b = (S)0.5 * rho / pr;
  S sb = m_sqrt(b);
  S spb = m_sqrt(Consts<S>::PI * b);
  S S1 = k.ut * sb;
  S S2 = k.un * sb;
  k.B1 = (S)0.5 * m_exp(-S1 * S1) / spb;
  k.B2 = (S)0.5 * m_exp(-S2 * S2) / spb;
  k.e1 = m_erf(S1);
  k.e2 = m_erf(S2);
  k.pr_by_rho = pr / rho;
  k.u_sqr = k.ut * k.ut + k.un * k.un;
  k.rho = rho;
  return k;
}

// ---------------------------------------------------------------------------------
// Single-precision math of the mixed path.  That path is held to |dCl|, |dCd| <= 1e-4
// of the FP64 path, not to ulp parity, so it uses the hardware transcendental
// instructions (exp2, log2, reciprocal and reciprocal square root, each about 1 ulp)
// instead of the correctly rounded library sequences, and erf from the exp(-s^2) the
// fluxes need anyway.  The FP64 path keeps the library functions and the reference's
// operation order.
__device__ inline float fexp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504f); }
__device__ inline float flog(float x) { return __builtin_amdgcn_logf(x) * 0.693147181f; }
__device__ inline float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline float frsq(float x) { return __builtin_amdgcn_rsqf(x); }

// erf(s) given ex = exp(-s^2): Abramowitz & Stegun 7.1.26, erf(x) = 1 - t (a1 + t (a2 +
// t (a3 + t (a4 + t a5)))) exp(-x^2), t = 1/(1 + p x), x >= 0, |error| <= 1.5e-7; odd in s.
// With s given as c x (c > 0), pass p / c.
__device__ inline float ferf_from_exp(float s, float ex, float p = 0.3275911f) {
  const float x = fabsf(s);
  const float t = frcp(1.0f + p * x);
  const float poly = t * (0.254829592f + t * (-0.284496736f + t * (1.421413741f + t * (-1.453152027f + t * 1.061405429f))));
  const float e = 1.0f - poly * ex;
  return s < 0.0f ? -e : e;
}

// The flux preliminaries (q-tilde -> primitive -> split-flux pieces) of the mixed
// path: kin_of_qtilde's quantities with the identities b = rho / (2 p) = beta,
// p / rho = 1 / (2 beta) = rsq(beta)^2 / 2 and ut^2 + un^2 = u1^2 + u2^2 used instead of
// recomputing them.  The exponentials run in base 2: S1, S2 are scaled by sqrt(log2 e),
// so exp(-s^2) = exp2(-S^2) and erf's t = 1 / (1 + p s) takes p / sqrt(log2 e).
template <bool AXIS = false>
__device__ inline Kin<float> kin_fast(const float qt[4], float nx, float ny) {
  constexpr float LOG2E = 1.44269504f, SQRT_LOG2E = 1.20112241f;
  const float beta = -qt[3] * 0.5f;
  const float rs = frsq(beta);
  const float tmp = (0.5f * rs) * rs;  // 1 / (2 beta)
  const float u1 = qt[1] * tmp, u2 = qt[2] * tmp;
  const float uu = u1 * u1 + u2 * u2;
  const float temp1 = qt[0] + beta * uu;
  // rho = exp(temp1 - log(beta) / (gamma - 1)), 1 / (gamma - 1) = 2.5
  const float rho = __builtin_amdgcn_exp2f(temp1 * LOG2E - 2.5f * __builtin_amdgcn_logf(beta));
  Kin<float> k;
  if constexpr (AXIS) {
    // normal (0, 1): ut = u1 ny - u2 nx = u1 and un = u2 exactly (up to the sign of a zero)
    k.ut = u1;
    k.un = u2;
  } else {
    k.ut = u1 * ny - u2 * nx;
    k.un = u1 * nx + u2 * ny;
  }
  const float sbl = (beta * rs) * SQRT_LOG2E;  // sqrt(beta log2 e)
  const float S1 = k.ut * sbl, S2 = k.un * sbl;
  const float ex1 = __builtin_amdgcn_exp2f(-S1 * S1), ex2 = __builtin_amdgcn_exp2f(-S2 * S2);
  const float half_inv_spb = 0.282094792f * rs;  // 0.5 / sqrt(pi beta)
  k.B1 = ex1 * half_inv_spb;
  k.B2 = ex2 * half_inv_spb;
  k.e1 = ferf_from_exp(S1, ex1, 0.3275911f / SQRT_LOG2E);
  k.e2 = ferf_from_exp(S2, ex2, 0.3275911f / SQRT_LOG2E);
  k.pr_by_rho = tmp;
  k.u_sqr = uu;
  k.rho = rho;
  return k;
}

// The stencil kernels' entry points: FP64 keeps the reference's routines, the mixed
// path takes the fast ones
__host__ __device__ inline Kin<double> kin_eval(const double qt[4], double nx, double ny) {
  return kin_of_qtilde(qt, nx, ny);
}
__device__ inline Kin<float> kin_eval(const float qt[4], float nx, float ny) { return kin_fast(qt, nx, ny); }

// x +- y with the sign chosen per lane: bitwise equal to the reference's separate
// "+" and "-" routines (x + (-y) == x - y in IEEE arithmetic)
template <class S>
__host__ __device__ inline S pm(S x, S y, bool plus) { return plus ? x + y : x - y; }

// split_fluxes.F90:8-96 flux_Gxp (plus) / flux_Gxn (minus)
template <class S>
__host__ __device__ inline void flux_gx(const Kin<S>& k, bool plus, S G[4]) {
  S A1 = (S)0.5 * pm((S)1.0, k.e1, plus);
  G[0] = k.rho * pm(k.ut * A1, k.B1, plus);
  S temp1 = k.pr_by_rho + k.ut * k.ut;
  S temp2 = pm(temp1 * A1, k.ut * k.B1, plus);
  G[1] = k.rho * temp2;
  temp1 = pm(k.ut * k.un * A1, k.un * k.B1, plus);
  G[2] = k.rho * temp1;
  temp1 = ((S)7.0 * k.pr_by_rho) + k.u_sqr;
  temp2 = (S)0.5 * k.ut * temp1 * A1;
  temp1 = ((S)6.0 * k.pr_by_rho) + k.u_sqr;
  G[3] = k.rho * pm(temp2, (S)0.5 * temp1 * k.B1, plus);
}

// split_fluxes.F90:99-190 flux_Gyp (plus) / flux_Gyn (minus)
template <class S>
__host__ __device__ inline void flux_gy(const Kin<S>& k, bool plus, S G[4]) {
  S A2 = (S)0.5 * pm((S)1.0, k.e2, plus);
  G[0] = k.rho * pm(k.un * A2, k.B2, plus);
  S temp1 = k.pr_by_rho + k.un * k.un;
  S temp2 = pm(temp1 * A2, k.un * k.B2, plus);
  G[2] = k.rho * temp2;
  temp1 = pm(k.ut * k.un * A2, k.ut * k.B2, plus);
  G[1] = k.rho * temp1;
  temp1 = ((S)7.0 * k.pr_by_rho) + k.u_sqr;
  temp2 = (S)0.5 * k.un * temp1 * A2;
  temp1 = ((S)6.0 * k.pr_by_rho) + k.u_sqr;
  G[3] = k.rho * pm(temp2, (S)0.5 * temp1 * k.B2, plus);
}

// quadrant_fluxes.F90:12-248: GxI (a1-, a2-), GxII (a1+, a2-), GxIII (a1+, a2+),
// GxIV (a1-, a2+)
template <class S>
__host__ __device__ inline void flux_quad(const Kin<S>& k, bool a1p, bool a2p, S G[4]) {
  S A1 = (S)0.5 * pm((S)1.0, k.e1, a1p);
  S A2 = (S)0.5 * pm((S)1.0, k.e2, a2p);
  S t1b = pm(k.ut * A1, k.B1, a1p);
  G[0] = k.rho * A2 * t1b;
  S temp1 = k.pr_by_rho + k.ut * k.ut;
  S temp2 = pm(temp1 * A1, k.ut * k.B1, a1p);
  G[1] = k.rho * A2 * temp2;
  temp2 = pm(k.un * A2, k.B2, a2p);
  G[2] = k.rho * t1b * temp2;
  temp1 = ((S)7.0 * k.pr_by_rho) + k.u_sqr;
  temp2 = (S)0.5 * k.ut * temp1 * A1;
  temp1 = ((S)6.0 * k.pr_by_rho) + k.u_sqr;
  S temp3 = (S)0.5 * k.B1 * temp1;
  S temp4 = (S)0.5 * k.rho * k.un * k.B2 * t1b;
  G[3] = pm(k.rho * A2 * pm(temp2, temp3, a1p), temp4, a2p);
}

template <class S>
struct State {
  const V4<S>* q;
  const V4<S>* dq;
  const V4<S>* qm;
  const S* eps;
};

// Limited mid-point reconstruction of one end (wall_fluxes.F90:79-86): qtilde =
// q - 0.5*phi*(delx*dq_x + dely*dq_y), phi from the end's own qm and epsilon.
template <class S>
__host__ __device__ inline void reconstruct(const V4<S>& q, const V4<S>& dx, const V4<S>& dy,
                                   const V4<S>& qmax, const V4<S>& qmin, S eps, S delx,
                                   S dely, S qt[4]) {
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const S d = delx * dx[r] + dely * dy[r];
    const S phi = venkat(q[r] - (S)0.5 * d, q[r], qmax[r], qmin[r], eps);
    qt[r] = q[r] - (S)0.5 * phi * d;
  }
}

// reconstruct for the stencil kernels (see kin_eval), from the end's pmax = qmax - q,
// pmin = qmin - q and e2 = eps^2 (precomputed per point; same bits as reconstruct)
__host__ __device__ inline void reconstruct_eval(const V4d& q, const V4d& dx, const V4d& dy, const V4d& pmax,
                                                 const V4d& pmin, double e2, double delx, double dely,
                                                 double qt[4]) {
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const double d = delx * dx[r] + dely * dy[r];
    const double phi = venkat_pre(q[r] - 0.5 * d, q[r], pmax[r], pmin[r], e2);
    qt[r] = q[r] - 0.5 * phi * d;
  }
}
// The mixed path's limited reconstruction.  The reference's limiter ratio has the
// factor del_neg in both numerator and denominator (limiters.F90:91-97); past the
// threshold del_neg != 0, so it is divided out: phi = min(1, num / den) with
// a = p^2 + e^2, num = a + 2 del_neg p, den = a + del_neg (2 del_neg + p), p = del_pos;
// den = (p + del_neg/2)^2 + 7/4 del_neg^2 + e^2 > 0.  del_neg = qtilde - q is taken as
// -d/2 exactly, so num = a - d p, den = a + del_neg (p - d) and qtilde = q + phi del_neg;
// a reciprocal replaces the division.
__device__ inline void reconstruct_eval(const V4<float>& q, const V4<float>& dx, const V4<float>& dy,
                                        const V4<float>& pmax, const V4<float>& pmin, float e2, float delx,
                                        float dely, float qt[4]) {
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const float d = delx * dx[r] + dely * dy[r];
    const float dn = -0.5f * d;
    const float p = dn > 0.0f ? pmax[r] : pmin[r];
    const float a = p * p + e2;
    const float num = a - d * p;
    const float den = a + dn * (p - d);
    const float t = fminf(num * frcp(den), 1.0f);
    const float phi = (fabsf(dn) <= Consts<float>::THRESH) ? 1.0f : t;
    qt[r] = q[r] + phi * dn;
  }
}

// The mixed path's interior split fluxes (split_fluxes.F90:8-190 regrouped): with
// rA = rho A and rB = +-rho B of a direction, Gx = (ut rA1 + rB1, (p/rho + ut^2) rA1 +
// ut rB1, un Gx(1), ut (7/2 p/rho + u^2/2) rA1 + (3 p/rho + u^2/2) rB1), and Gy the same
// with ut and un exchanged.  sx, sy = +1 / -1 for the + / - halves.
__device__ inline void interior_fluxes_fast(const Kin<float>& k, float sx, float sy, float Gx[4], float Gy[4]) {
  const float hu = 0.5f * k.u_sqr;
  const float c7 = 3.5f * k.pr_by_rho + hu, c6 = 3.0f * k.pr_by_rho + hu;
  const float rA1 = k.rho * (0.5f * sx * k.e1 + 0.5f);
  const float rB1 = (sx * k.rho) * k.B1;
  Gx[0] = k.ut * rA1 + rB1;
  Gx[1] = (k.ut * k.ut + k.pr_by_rho) * rA1 + k.ut * rB1;
  Gx[2] = k.un * Gx[0];
  Gx[3] = (k.ut * c7) * rA1 + c6 * rB1;
  const float rA2 = k.rho * (0.5f * sy * k.e2 + 0.5f);
  const float rB2 = (sy * k.rho) * k.B2;
  Gy[0] = k.un * rA2 + rB2;
  Gy[2] = (k.un * k.un + k.pr_by_rho) * rA2 + k.un * rB2;
  Gy[1] = k.ut * Gy[0];
  Gy[3] = (k.un * c7) * rA2 + c6 * rB2;
}

// Split fluxes of one end for one edge: the x flux of the neighbour's x half and
// the y flux of its y half.  `xp`/`yp` pick the + or - half.
template <int CLS, class S>
__host__ __device__ inline void end_fluxes(const Kin<S>& k, bool xp, bool yp, S Gx[4], S Gy[4]) {
  if constexpr (CLS == INTERIOR && sizeof(S) == 4) {
    interior_fluxes_fast(k, xp ? 1.0f : -1.0f, yp ? 1.0f : -1.0f, Gx, Gy);
  } else if (CLS == INTERIOR) {
    flux_gx(k, xp, Gx);
    flux_gy(k, yp, Gy);
  } else {
    // wall: xpos GxII, xneg GxI, every neighbour Gyn; outer: xpos GxIII, xneg GxIV,
    // every neighbour Gyp (wall_fluxes.F90, outer_fluxes.F90)
    const bool a2p = (CLS == OUTER);
    flux_quad(k, xp, a2p, Gx);
    flux_gy(k, a2p, Gy);
  }
}

// Per-point solve coefficients of the four half-stencils (xpos, xneg, ypos, yneg):
// the t-derivative of a half is (sum dG*ds*c1 - sum dG*dn*sxy)*obd
// (interior_fluxes.F90:180-183), the n-derivative (sum dG*dn*c1 - sum dG*ds*sxy)*obd
// (:378-381), with the point-fixed sums c1, sxy, obd precomputed on the host.  Both
// are linear in the edge terms, so each edge adds dG times its own coefficient
// straight into flux_res.
struct HalfCoef {
  double c1[4], sxy[4], obd[4];
};

// `first` selects half 0 (xpos) / 2 (ypos), else half 1 (xneg) / 3 (yneg); selects,
// not a runtime array index, so the coefficients stay in registers
__host__ __device__ inline double coef_x(const HalfCoef& h, bool first, double dsw, double dnw) {
  const double c1 = first ? h.c1[0] : h.c1[1];
  const double sxy = first ? h.sxy[0] : h.sxy[1];
  const double obd = first ? h.obd[0] : h.obd[1];
  return (dsw * c1 - dnw * sxy) * obd;
}
__host__ __device__ inline double coef_y(const HalfCoef& h, bool first, double dsw, double dnw) {
  const double c1 = first ? h.c1[2] : h.c1[3];
  const double sxy = first ? h.sxy[2] : h.sxy[3];
  const double obd = first ? h.obd[2] : h.obd[3];
  return (dnw * c1 - dsw * sxy) * obd;
}

// The mixed path's form of the same coefficients: a = c1 * obd and b = sxy * obd per
// half in FP32 (8 registers instead of 24), so an edge's coefficient is dsw * a - dnw * b
// (x) or dnw * a - dsw * b (y).  Their rounding (~1e-7 relative) is that of the FP32
// flux differences they multiply.
struct HalfCoefF {
  float a[4], b[4];
};
__device__ inline float coef_x(const HalfCoefF& h, bool first, float dsw, float dnw) {
  const float a = first ? h.a[0] : h.a[1];
  const float b = first ? h.b[0] : h.b[1];
  return dsw * a - dnw * b;
}
__device__ inline float coef_y(const HalfCoefF& h, bool first, float dsw, float dnw) {
  const float a = first ? h.a[2] : h.a[3];
  const float b = first ? h.b[2] : h.b[3];
  return dnw * a - dsw * b;
}

// Flux residual of point i (flux_residual.F90:14-73) for class CLS: the sum over the
// point's half-stencils of the least-squares derivatives of the kinetic split fluxes
// (wall_fluxes.F90:20-292, outer_fluxes.F90:18-300, interior_fluxes.F90:19-404).
// Returns flux_res(:, i) in fr; with LPP > 1, this lane's share (edges lane, lane +
// LPP, ...), to be summed over the point's lanes.
// Where flux_point reads a neighbour's records (k: a point id for the global walk, a
// halo position for the staged one) and how it walks its slots (walk(f) calls f(k, m)
// per slot of the lane, m the slot's half-stencil bits).
template <class S>
struct GlobalNbrs {
  using scalar = S;
  const Geom& g;
  const State<S>& st;
  int i, nb, lane, lpp;
  __device__ D2 xy(int k) const { return g.xy[k]; }
  __device__ V4<S> q(int k) const { return st.q[k]; }
  __device__ V4<S> dqx(int k) const { return st.dq[2 * k]; }
  __device__ V4<S> dqy(int k) const { return st.dq[2 * k + 1]; }
  __device__ V4<S> qmax(int k) const { return st.qm[2 * k]; }
  __device__ V4<S> qmin(int k) const { return st.qm[2 * k + 1]; }
  __device__ S eps(int k) const { return st.eps[k]; }
  // the limiter's inputs of a record: qmax - q, qmin - q, eps^2
  __device__ V4<S> pmax(int k) const { return st.qm[2 * k] - st.q[k]; }
  __device__ V4<S> pmin(int k) const { return st.qm[2 * k + 1] - st.q[k]; }
  __device__ S e2(int k) const {
    const S e = st.eps[k];
    return e * e;
  }
  // the point's own record index; slot s's neighbour index and half-stencil bits
  __device__ int own() const { return i; }
  __device__ void slot(int s, int& k, unsigned& m) const {
    const uint32_t pk = g.conn[(size_t)s * g.n + i];
    k = (int)(pk >> 4);
    m = pk & 15u;
  }
  template <class F>
  __device__ void walk(F&& f) const {
    const uint32_t* cp = g.conn + i;
    // the next slot's connectivity word is loaded while this edge is evaluated, so
    // each edge starts with one memory round trip (its neighbour's records), not two
    uint32_t pk_next = lane < nb ? cp[(size_t)lane * g.n] : 0u;
    for (int s = lane; s < nb; s += lpp) {
      const uint32_t pk = pk_next;
      if (s + lpp < nb) pk_next = cp[(size_t)(s + lpp) * g.n];
      f((int)(pk >> 4), pk & 15u);
    }
  }
};

// A block's staged halo in LDS (structure of arrays, one entry per halo position).
// Staged with PRE (the forward flux kernels), qmax, qmin and eps hold the limiter's
// inputs qmax - q, qmin - q and eps^2 instead.
template <class S>
struct HaloRecs {
  // per halo entry its five 4-vectors side by side: q, dq_x, dq_y, then qmax and qmin (a
  // PRE halo: qmax - q and qmin - q), so that an entry's fields sit at fixed offsets
  // from one address and the LDS reads take them as immediate offsets
  V4<S>* rec;  // [H][5]
  D2* xy;      // [H]
  S* eps;      // [H], epsilon (a PRE halo: epsilon^2)
};

template <class S, int P, int LPP, bool PRE = false>
struct StagedNbrs {
  using scalar = S;
  const Geom& g;
  const HaloRecs<S>& r;
  const uint16_t* s_cl;
  int li, nb, lane;
  __device__ D2 xy(int k) const { return r.xy[k]; }
  __device__ V4<S> q(int k) const { return r.rec[5 * k]; }
  __device__ V4<S> dqx(int k) const { return r.rec[5 * k + 1]; }
  __device__ V4<S> dqy(int k) const { return r.rec[5 * k + 2]; }
  __device__ V4<S> qmax(int k) const {
    static_assert(!PRE, "a PRE halo holds qmax - q");
    return r.rec[5 * k + 3];
  }
  __device__ V4<S> qmin(int k) const {
    static_assert(!PRE, "a PRE halo holds qmin - q");
    return r.rec[5 * k + 4];
  }
  __device__ S eps(int k) const {
    static_assert(!PRE, "a PRE halo holds eps^2");
    return r.eps[k];
  }
  __device__ V4<S> pmax(int k) const {
    if constexpr (PRE) return r.rec[5 * k + 3];
    else return r.rec[5 * k + 3] - r.rec[5 * k];
  }
  __device__ V4<S> pmin(int k) const {
    if constexpr (PRE) return r.rec[5 * k + 4];
    else return r.rec[5 * k + 4] - r.rec[5 * k];
  }
  __device__ S e2(int k) const {
    if constexpr (PRE) return r.eps[k];
    else return r.eps[k] * r.eps[k];
  }
  __device__ int own() const { return li; }  // a block's own points lead its halo
  __device__ void slot(int s, int& k, unsigned& m) const {
    const unsigned e = s_cl[s * P + li];
    k = (int)(e & 0xFFFu);
    m = e >> 12;
  }
  template <class F>
  __device__ void walk(F&& f) const {
    // the next slot's word is read while this edge is evaluated
    unsigned en = lane < nb ? s_cl[lane * P + li] : 0u;
    for (int s = lane; s < nb; s += LPP) {
      const unsigned e = en;
      if (s + LPP < nb) en = s_cl[(s + LPP) * P + li];
      f((int)(e & 0xFFFu), e >> 12);
    }
  }
};

// A STAGE_PF block's flux halo into LDS (records of State st and the coordinates, then
// the block's slot words), ended by a barrier; every thread of the block takes part.
// The LDS layout (HaloRecs, then the slot words) is stage_lds_flux's.
// Returns false, without staging, for a block whose halo exceeds the LDS capacity (the
// few blocks of wall and outer points): its points then walk the global records with
// the same lanes and slots (GlobalNbrs), which gives the same bits.  Every thread
// reaches the barrier either way.
template <class S, int LPP, bool PRE = false>
__device__ inline bool stage_flux_halo(const Geom& g, const State<S>& st, int b, unsigned char* smem,
                                       HaloRecs<S>& r, uint16_t*& s_cl) {
  const Halo& h = g.hl[1];
  const int H = h.hcap;
  r.rec = (V4<S>*)smem;
  r.xy = (D2*)(r.rec + 5 * H);
  r.eps = (S*)(r.xy + H);
  s_cl = (uint16_t*)(r.eps + H);
  const int h0 = h.ptr[b], hn = h.ptr[b + 1] - h0;
  if (hn > H) {
    __syncthreads();
    return false;
  }
  for (int j = threadIdx.x; j < hn; j += STAGE_PF * LPP) {
    const int k = h.ids[h0 + j];
    const V4<S> qk = st.q[k];
    V4<S>* e = r.rec + 5 * j;
    e[0] = qk;
    e[1] = st.dq[2 * k];
    e[2] = st.dq[2 * k + 1];
    if constexpr (PRE) {
      e[3] = st.qm[2 * k] - qk;
      e[4] = st.qm[2 * k + 1] - qk;
      const S ep = st.eps[k];
      r.eps[j] = ep * ep;
    } else {
      e[3] = st.qm[2 * k];
      e[4] = st.qm[2 * k + 1];
      r.eps[j] = st.eps[k];
    }
    r.xy[j] = g.xy[k];
  }
  constexpr int V = STAGE_PF * 2 / 16;
  const uint4* src = (const uint4*)(h.loc + (size_t)b * STAGE_PF);
  for (int j = threadIdx.x; j < g.width * V; j += STAGE_PF * LPP) {
    const int sl = j / V;
    ((uint4*)s_cl)[j] = src[(size_t)sl * (h.stride / 8) + (j - sl * V)];
  }
  __syncthreads();
  return true;
}

// AXIS (mixed path, interior points, Geom::axis_interior): the point's normal is (0, 1),
// so the edge offsets and the ends' velocities need no rotation into its frame (12 of an
// edge's ~350 VALU instructions).  The frame values equal the rotation's (x 1 and x 0
// are exact, up to the sign of a zero); the fluxes computed from them can differ in the
// last bit, as FMA contraction pairs their products by the shape of the code.
template <int CLS, class S, bool POWER0, int LPP, bool AXIS = false, class Nbrs>
__device__ inline void flux_point_src(const Geom& g, const State<S>& st, int i, double delta, double power,
                                      double fr[4], const Nbrs& nbrs) {
  static_assert(!AXIS || (CLS == INTERIOR && sizeof(S) == 4), "the axis frame is the mixed interior form");
  const D2 pi = g.xy[i];
  const D2 ni = g.nrm[i];
  const double nx = ni.x, ny = ni.y, tx = ny, ty = -nx;
  const S nxs = (S)nx, nys = (S)ny;
  const V4<S> qi = st.q[i];
  const V4<S> dxi = st.dq[2 * i], dyi = st.dq[2 * i + 1];
  // the own end's limiter inputs, once per point
  const V4<S> pmaxi = st.qm[2 * i] - qi, pmini = st.qm[2 * i + 1] - qi;
  const S epsi = st.eps[i];
  const S e2i = epsi * epsi;
  // FP64: the reference's coefficient arithmetic; the mixed path: FP32 edge geometry
  // and coefficients (HalfCoefF)
  constexpr bool F32 = sizeof(S) == 4;
  using R = std::conditional_t<F32, float, double>;
  std::conditional_t<F32, HalfCoefF, HalfCoef> hc;
  if constexpr (F32) {
    const V4<float> a = g.dg_f[2 * i], b = g.dg_f[2 * i + 1];
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      hc.a[d] = a[d];
      hc.b[d] = b[d];
    }
  } else {
    const V4d a = g.dg_c1[i], b = g.dg_sxy[i], c = g.dg_obd[i];
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      hc.c1[d] = a[d];
      hc.sxy[d] = b[d];
      hc.obd[d] = c[d];
    }
  }
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  // an end's kinetic state in the point's frame
  auto kin = [&](const S* qt) {
    if constexpr (F32) return kin_fast<AXIS>(qt, nxs, nys);
    else return kin_eval(qt, nxs, nys);
  };
  nbrs.walk([&](int k, unsigned m) {
    const D2 pkk = nbrs.xy(k);
    const double delx = pkk.x - pi.x, dely = pkk.y - pi.y;
    R dels, deln;
    if constexpr (F32) {
      const float fx = (float)delx, fy = (float)dely;
      if constexpr (AXIS) {
        dels = fx;
        deln = fy;
      } else {
        dels = fx * nys - fy * nxs;
        deln = fx * nxs + fy * nys;
      }
    } else {
      dels = delx * tx + dely * ty;
      deln = delx * nx + dely * ny;
    }
    R w = 1.0;
    if (!POWER0) w = (R)pow(sqrt((double)dels * dels + (double)deln * deln), power);
    const R dsw = dels * w, dnw = deln * w;
    const bool xp = (m & XPOS) != 0u;
    const bool tie_x = xp && (m & XNEG) != 0u;
    // interior: y half from the mask; wall: yneg (slot 3); outer: ypos (slot 2)
    const bool yp = (CLS == INTERIOR) ? (m & YPOS) != 0u : (CLS == OUTER);
    const bool tie_y = (CLS == INTERIOR) && yp && (m & YNEG) != 0u;
    const R cx = coef_x(hc, xp, dsw, dnw);
    const R cy = coef_y(hc, yp, dsw, dnw);

    const S dx = (S)delx, dy = (S)dely;
    S qti[4], qtk[4];
    if (g.first_order) {
      const V4<S> qk = nbrs.q(k);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        qti[r] = qi[r];
        qtk[r] = qk[r];
      }
    } else {
      if constexpr (F32) {
        // the own end's records read again per edge (the staged halo holds the point
        // at its own position), not held across the walk
        int o = nbrs.own();
        asm volatile("" : "+v"(o));
        reconstruct_eval(nbrs.q(o), nbrs.dqx(o), nbrs.dqy(o), nbrs.pmax(o), nbrs.pmin(o), nbrs.e2(o), dx, dy, qti);
      } else {
        reconstruct_eval(qi, dxi, dyi, pmaxi, pmini, e2i, dx, dy, qti);
      }
      const V4<S> qk = nbrs.q(k);
      const V4<S> dxk = nbrs.dqx(k), dyk = nbrs.dqy(k);
      reconstruct_eval(qk, dxk, dyk, nbrs.pmax(k), nbrs.pmin(k), nbrs.e2(k), dx, dy, qtk);
    }
    S gxi[4], gyi[4], gxk[4], gyk[4];
    if constexpr (sizeof(S) == 4) {
      // the mixed path's short sequences: both ends in straight-line code (a loop
      // over the ends moves its inputs and results between iterations)
      end_fluxes<CLS>(kin(qti), xp, yp, gxi, gyi);
      end_fluxes<CLS>(kin(qtk), xp, yp, gxk, gyk);
    } else {
    // FP64: the two ends one after the other, so one end's library-transcendental
    // temporaries live at a time
#pragma unroll 1
    for (int e = 0; e < 2; ++e) {
      S qt[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) qt[r] = e ? qtk[r] : qti[r];
      S Gx[4], Gy[4];
      end_fluxes<CLS>(kin(qt), xp, yp, Gx, Gy);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (e) {
          gxk[r] = Gx[r];
          gyk[r] = Gy[r];
        } else {
          gxi[r] = Gx[r];
          gyi[r] = Gy[r];
        }
      }
    }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      // the end fluxes' difference in the state precision (for FP64 the same as after
      // widening; for FP32 its rounding is at most that of the FP32 fluxes themselves),
      // accumulated in FP64
      const double dgx = (double)(gxk[r] - gxi[r]);
      const double dgy = (double)(gyk[r] - gyi[r]);
      acc[r] = acc[r] + dgx * cx + dgy * cy;
    }
    if (tie_x || tie_y) {
      // delta_s == 0 (or delta_n == 0 at an interior point): the neighbour also
      // belongs to the xneg (yneg) half, whose flux has the opposite split
      const double cx2 = tie_x ? coef_x(hc, false, dsw, dnw) : 0.0;
      const double cy2 = tie_y ? coef_y(hc, false, dsw, dnw) : 0.0;
      S Gx[4], Gy[4];
      end_fluxes<CLS>(kin(qti), false, false, Gx, Gy);
      S Hx[4], Hy[4];
      end_fluxes<CLS>(kin(qtk), false, false, Hx, Hy);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (tie_x) acc[r] = acc[r] + (double)(Hx[r] - Gx[r]) * cx2;
        if (tie_y) acc[r] = acc[r] + (double)(Hy[r] - Gy[r]) * cy2;
      }
    }
  });
  // flux_residual.F90:36 (wall: 2*delta), :50 and :68 (delta)
  const double scale = (CLS == WALL) ? 2.0 * delta : delta;
#pragma unroll
  for (int r = 0; r < 4; ++r) fr[r] = scale * acc[r];
}

// flux_point_src with the mixed path's interior points in the axis frame when
// Geom::axis_interior says their normals are (0, 1).  Every walk, staged or global, goes
// through it: the two frames' code contracts its products into FMAs differently, so a
// walk's bits depend on the frame form it runs, and the walks keep the same bits only if
// they run the same form.
template <int CLS, class S, bool POWER0, int LPP, class Nbrs>
__device__ inline void flux_point_frame(const Geom& g, const State<S>& st, int i, double delta, double power,
                                        double fr[4], const Nbrs& nbrs) {
  if constexpr (CLS == INTERIOR && sizeof(S) == 4) {
    if (g.axis_interior) {
      flux_point_src<CLS, S, POWER0, LPP, true>(g, st, i, delta, power, fr, nbrs);
      return;
    }
  }
  flux_point_src<CLS, S, POWER0, LPP>(g, st, i, delta, power, fr, nbrs);
}

// A point's flux residual from the global walk
template <int CLS, class S, bool POWER0, int LPP = 1>
__device__ inline void flux_point(const Geom& g, const State<S>& st, int i, double delta, double power,
                                  double fr[4], int lane = 0) {
  flux_point_frame<CLS, S, POWER0, LPP>(g, st, i, delta, power, fr,
                                        GlobalNbrs<S>{g, st, i, g.nbhs[i], lane, LPP});
}

// state_update.F90:144-163 primitive_to_conserved (rotated frame)
template <class S>
__host__ __device__ inline void prim_to_cons(const V4<S>& p, S nx, S ny, S U[4]) {
  S rho = p.x;
  U[0] = rho;
  S temp1 = rho * p.y;
  S temp2 = rho * p.z;
  U[3] = (S)2.5 * p.w + (S)0.5 * (temp1 * temp1 + temp2 * temp2) / rho;
  U[1] = temp1 * ny - temp2 * nx;
  U[2] = temp1 * nx + temp2 * ny;
}

struct Freestream {
  double u1_inf, u2_inf;  // q_inf(2:3), parameter.F90:29-32
};

// state_update.F90:239-294 conserved_vector_Ubar
template <class S>
__host__ __device__ inline void cons_ubar(const V4<S>& p, S nx, S ny, const Freestream& fs, S Ub[4]) {
  const S rho_inf = (S)1.0, pr_inf = (S)(1.0 / 1.4), gm1 = (S)(1.4 - 1.0);
  const S pi = Consts<S>::PI;
  S u1_inf = (S)fs.u1_inf, u2_inf = (S)fs.u2_inf;
  S tx = ny, ty = -nx;
  S u1_inf_rot = u1_inf * tx + u2_inf * ty;
  S u2_inf_rot = u1_inf * nx + u2_inf * ny;
  S temp1 = (u1_inf_rot * u1_inf_rot + u2_inf_rot * u2_inf_rot);
  S e_inf = pr_inf / (rho_inf * gm1) + (S)0.5 * (temp1);
  S beta = ((S)0.5 * rho_inf) / pr_inf;
  S S2 = u2_inf_rot * m_sqrt(beta);
  S B2_inf = m_exp(-S2 * S2) / ((S)2.0 * m_sqrt(pi * beta));
  S A2n_inf = (S)0.5 * ((S)1.0 - m_erf(S2));
  S rho = p.x, u1 = p.y, u2 = p.z, pr = p.w;
  S u1_rot = u1 * tx + u2 * ty;
  S u2_rot = u1 * nx + u2 * ny;
  temp1 = (u1_rot * u1_rot + u2_rot * u2_rot);
  S e = pr / (rho * gm1) + (S)0.5 * (temp1);
  beta = (rho) / ((S)2.0 * pr);
  S2 = u2_rot * m_sqrt(beta);
  S B2 = m_exp(-S2 * S2) / ((S)2.0 * m_sqrt(pi * beta));
  S A2p = (S)0.5 * ((S)1.0 + m_erf(S2));
  Ub[0] = (rho_inf * A2n_inf) + (rho * A2p);
  Ub[1] = (rho_inf * u1_inf_rot * A2n_inf) + (rho * u1_rot * A2p);
  temp1 = rho_inf * (u2_inf_rot * A2n_inf - B2_inf);
  S temp2 = rho * (u2_rot * A2p + B2);
  Ub[2] = temp1 + temp2;
  temp1 = (rho_inf * A2n_inf * e_inf - (S)0.5 * rho_inf * u2_inf_rot * B2_inf);
  temp2 = (rho * A2p * e + (S)0.5 * rho * u2_rot * B2);
  Ub[3] = temp1 + temp2;
}

// state_update.F90:10-141 for one point of class CLS.  Returns the updated prim and
// (U1_new - U1_old)^2 (0 for outer points, which the residual skips).
template <int CLS, class S>
__host__ __device__ inline V4<S> update_point(const V4<S>& p, const V4<S>& pold, const double fr[4],
                                     double nxd, double nyd, int rk, double euler,
                                     const Freestream& fs, double* res_sqr) {
  const S nx = (S)nxd, ny = (S)nyd;
  S U[4], Uo[4];
  if (CLS == OUTER) {
    cons_ubar(p, nx, ny, fs, U);
    if (rk == 3) cons_ubar(pold, nx, ny, fs, Uo);
  } else {
    prim_to_cons(p, nx, ny, U);
    if (rk == 3) prim_to_cons(pold, nx, ny, Uo);
  }
  const S temp = U[0];
  const S obt = (S)(1.0 / 3.0), tbt = (S)(2.0 / 3.0);
  const S half_euler = (S)(0.5 * euler);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const S f = (S)fr[r];
    if (rk != 3) U[r] = U[r] - (half_euler * f);
    else U[r] = tbt * Uo[r] + obt * (U[r] - (S)0.5 * f);
  }
  if (CLS == WALL) U[2] = (S)0.0;
  const S U2_rot = U[1], U3_rot = U[2];
  U[1] = U2_rot * ny + U3_rot * nx;
  U[2] = U3_rot * ny - U2_rot * nx;
  if (CLS == OUTER) *res_sqr = 0.0;
  else {
    const double d = (double)U[0] - (double)temp;  // exact difference of the S values
    *res_sqr = d * d;
  }
  V4<S> out;
  out.x = U[0];
  const S t = (S)1.0 / U[0];
  out.y = U[1] * t;
  out.z = U[2] * t;
  out.w = (S)0.4 * (U[3] - ((S)0.5 * t) * (U[1] * U[1] + U[2] * U[2]));
  return out;
}

// Wave64 + LDS sum of one double per thread; returns the block total in thread 0.
__device__ inline double block_sum(double v) {
  __shared__ double red[BLOCK / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
  }
  return t;
}

}  // namespace mf
