// MI355X (gfx950) LSKUM fixed-point iteration: kernels and the C ABI of include/mfcfd.h.
//
// Host side: builds the device layout from the reference's point arrays (normals,
// stencil split, Hilbert renumbering, point-fixed least-squares geometry), owns the
// device state of one cloud, and replays one fpi_solver iteration
// (fpi_solver.F90:13-83) as a captured HIP graph.  Device side: lskum_device.h (point
// and edge routines), lskum_kernels.h (stencil kernels and launchers, FP64 instances
// here, the mixed path's in mfcfd_f32.hip), mfcfd_adjoint.hip (the reverse sweep).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/mfcfd.h"
#include "lskum_kernels.h"

using namespace mf;

// =================================================================================
// Kernels
// =================================================================================

// This thread's share of p[begin, end) (elements begin + threadIdx.x + k * blockDim.x):
// eight independent accumulators, so eight loads are in flight at a time instead of one
// dependent load per add; the grouping is fixed, so the result is deterministic.
__device__ inline double strided_sum(const double* __restrict__ p, int begin, int end) {
  double v[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int st = (int)blockDim.x;
  int j = begin + (int)threadIdx.x;
  for (; j + 7 * st < end; j += 8 * st) {
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] += p[j + u * st];
  }
  for (; j < end; j += st) v[0] += p[j];
  return ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
}

// Sum of the residual partials and the fpi_solver.F90:57-70 residue logic; one block.
__global__ __launch_bounds__(BLOCK) void k_finalize(const double* __restrict__ partial, int np,
                                                    int n, int restart,
                                                    Ctl* __restrict__ ctl,
                                                    double* __restrict__ res_hist,
                                                    double* __restrict__ resnew_hist) {
  if (ctl->halt) return;
  const double tot = block_sum(strided_sum(partial, 0, np));
  if (threadIdx.x == 0) {
    const int it = ctl->it;
    const int t = ctl->first_it + it;
    const double res_new = sqrt(tot) / (double)n;
    double residue;
    if (t <= 2 && restart == 0) {
      ctl->res_old = res_new;
      residue = 0.0;
    } else {
      residue = log10(res_new / ctl->res_old);
    }
    res_hist[it] = residue;
    resnew_hist[it] = res_new;
    if (isnan(residue)) ctl->halt = HALT_NAN;
    ctl->it = it + 1;
  }
}

__global__ void k_begin_run(Ctl* ctl, int set_res_old, double res_old, int first_it) {
  ctl->it = 0;
  ctl->halt = 0;
  ctl->first_it = first_it;
  if (set_res_old) ctl->res_old = res_old;
}

// Residual norms of a batch, one block per case: sqrt(sum over the case's wall and
// interior points of (dU1)^2) / (its point count), and fpi_solver.F90:63-70's residue
// with a res_old per case.  rng per case: wall range, interior range, point count.
// Cases are independent: a NaN in one does not stop the others.
__global__ __launch_bounds__(BLOCK) void k_finalize_batch(const double* __restrict__ rs_pt,
                                                          const int32_t* __restrict__ rng, int ncases,
                                                          int restart,
                                                          const Ctl* __restrict__ ctl,
                                                          double* __restrict__ res_old,
                                                          double* __restrict__ res_hist,
                                                          double* __restrict__ resnew_hist) {
  const int c = blockIdx.x;
  const int* r = rng + 5 * c;
  const double tot = block_sum(strided_sum(rs_pt, r[0], r[1]) + strided_sum(rs_pt, r[2], r[3]));
  if (threadIdx.x == 0) {
    const int it = ctl->it;
    const double res_new = sqrt(tot) / (double)r[4];
    double residue;
    if (ctl->first_it + it <= 2 && restart == 0) {
      res_old[c] = res_new;
      residue = 0.0;
    } else {
      residue = log10(res_new / res_old[c]);
    }
    res_hist[(size_t)it * ncases + c] = residue;
    resnew_hist[(size_t)it * ncases + c] = res_new;
  }
}

__global__ void k_tick(Ctl* ctl) { ctl->it += 1; }

// Sum of partials into one scalar (unfused state_update's sum_res_sqr)
__global__ __launch_bounds__(BLOCK) void k_sum(const double* __restrict__ partial, int np,
                                               double* __restrict__ out) {
  const double t = block_sum(strided_sum(partial, 0, np));
  if (threadIdx.x == 0) *out = t;
}

// Plain vectorised copy: the known-traffic kernel for counter calibration
__global__ __launch_bounds__(BLOCK) void k_copy(const D2* __restrict__ a, D2* __restrict__ b,
                                                int64_t n) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * blockDim.x)
    b[j] = a[j];
}

// =================================================================================
// Host side
// =================================================================================

static thread_local std::string g_err;

static int set_err(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

// Error path of the allocation-free entry point (mfcfd_noalloc.hip): same message slot
int mf_noalloc_fail(int code, const char* msg) { return set_err(code, "%s", msg); }

#define HIPCHK(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return set_err(MF_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                     __FILE__, __LINE__);                                              \
  } while (0)

static const double kPi = 4.0 * std::atan(1.0);  // parameter.F90:11

// Polls of the resident sweeps' grid barrier before a wait gives up (about a second at
// a poll per microsecond; a wait normally ends within microseconds)
static constexpr uint32_t kDefaultBarrierPolls = 1u << 20;

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct mf_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  mf_params p{};
  int n = 0, W = 0, nw = 0, no = 0, ni = 0, nshape = 0;
  int64_t total_nbhs = 0;
  size_t ssz = 8;  // bytes per state scalar
  bool power0 = true;
  std::vector<int32_t> perm;   // internal -> original
  std::vector<int32_t> iperm;  // original -> internal
  std::vector<double> nx_h, ny_h;  // original order
  Geom g{};
  Freestream fs{};
  double theta = 0, cp_scale = 1;
  // device arrays
  std::vector<void*> allocs;
  void* eps = nullptr;
  const double* eps64 = nullptr;  // epsilon in FP64 (the adjoint's reverse kernels)
  void* prim[2] = {nullptr, nullptr};
  void* q[2] = {nullptr, nullptr};
  void* dq[2] = {nullptr, nullptr};
  void* qm = nullptr;
  void* prim_old = nullptr;
  int cur = 0, dq_cur = 0;
  V4d* flux_res = nullptr;
  double *delta = nullptr, *cspeed = nullptr, *partial = nullptr, *scalar = nullptr;
  int npartial = 0;
  Ctl* ctl = nullptr;
  double *res_hist = nullptr, *resnew_hist = nullptr, *cd_hist = nullptr;
  int hist_cap = 0;
  int32_t *sh_m = nullptr, *sh_id = nullptr;
  double *sgeo = nullptr, *cp = nullptr, *clcdcm = nullptr;
  // captured iterations: [phase] one iteration, [2 + phase] kGraphIters iterations
  hipGraphExec_t graph[4] = {nullptr, nullptr, nullptr, nullptr};
  // kept for the adjoint: internal ids of left/right (-1 for interior points), the
  // half-stencil sums, the shape points
  std::vector<int32_t> left_i, right_i, h_sh_m;
  std::vector<V4d> h_hxx, h_hyy, h_hxy;
  struct AdjState* adj = nullptr;
  struct TanState* tan = nullptr;  // mf_tangent's dual-number state
  // lanes per point of the stencil kernels (1, 2, 4 or 8; mf_autotune picks them)
  int lpp_lsq = 1, lpp_inner = 1, lpp_flux = 1;
  int lpp_rev[4] = {0, 0, 0, 0};  // adjoint: flux, lsq, inner, gather (0: timed)
  // shape ids (shapes per case x cases) and their per-id force constants
  int nsid = 0;
  int32_t *sh_ptr = nullptr, *sh_pos = nullptr;
  double* shape_par = nullptr;
  std::vector<double> h_shape_par;
  // a batch of cases (ncases > 0): per-point case ids, freestreams, residual pieces
  int ncases = 0;
  std::vector<int32_t> case_first;  // original-order offset of each case [ncases+1]
  int32_t* case_rng = nullptr;      // per case: wall range, interior range, point count
  double *rs_pt = nullptr, *res_old_case = nullptr;
  V4<float>* qsum = nullptr;  // mixed path: the LSQ first-pass sums per point [2N]
  Ctl* tune_ctl = nullptr;
  // mixed path, lpp_inner with MF_LPP_RESIDENT: the resident sweeps' grid barrier and
  // the polls before a wait at it gives up
  SweepBar* sbar = nullptr;
  uint32_t barrier_polls = kDefaultBarrierPolls;
};

template <class T>
static int dev_alloc(mf_ctx* c, T** ptr, size_t count) {
  void* p = nullptr;
  HIPCHK(hipMalloc(&p, std::max<size_t>(count * sizeof(T), 256)));
  c->allocs.push_back(p);
  *ptr = static_cast<T*>(p);
  return MF_OK;
}

template <class T>
static int upload(mf_ctx* c, T** ptr, const std::vector<T>& h) {
  int rc = dev_alloc(c, ptr, h.size());
  if (rc) return rc;
  HIPCHK(hipMemcpy(*ptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return MF_OK;
}

// Hilbert index of (x, y) in [0, 2^order)^2
static uint64_t hilbert_d(uint32_t order, uint64_t x, uint64_t y) {
  uint64_t d = 0;
  const uint64_t n = 1ull << order;
  for (uint64_t s = n >> 1; s > 0; s >>= 1) {
    const uint64_t rx = (x & s) ? 1 : 0, ry = (y & s) ? 1 : 0;
    d += s * s * ((3 * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) {
        x = n - 1 - x;
        y = n - 1 - y;
      }
      std::swap(x, y);
    }
  }
  return d;
}

static inline unsigned split_mask(int flag1, double dels, double deln) {
  // generate_connectivity.F90:39-256
  unsigned m = 0;
  if (dels <= 0.0) m |= XPOS;
  if (dels >= 0.0) m |= XNEG;
  if (flag1 == 1) {
    if (deln <= 0.0) m |= YPOS;
    if (deln >= 0.0) m |= YNEG;
  } else if (flag1 == 0) {
    m |= YNEG;
  } else {
    m |= YPOS;
  }
  return m;
}

static LaunchEnv env_of(const mf_ctx* c) {
  return LaunchEnv{c->stream, c->g, c->n, c->power0, c->p.power, c->p.cfl, c->p.euler, c->fs};
}

template <class S>
static int launch_q_variables(mf_ctx* c) {
  launch_q_variables_any<S>(env_of(c), (const V4<S>*)c->prim[c->cur], (V4<S>*)c->q[c->cur], c->cspeed);
  HIPCHK(hipGetLastError());
  return MF_OK;
}

static int q_variables(mf_ctx* c) {
  return c->ssz == 8 ? launch_q_variables<double>(c) : launch_q_variables<float>(c);
}

static int upload_prim_original(mf_ctx* c, const double* prim) {
  const int n = c->n;
  if (c->ssz == 8) {
    std::vector<V4d> h(n);
    for (int j = 0; j < n; ++j) {
      const int o = c->perm[j];
      h[j] = V4d{prim[4 * o], prim[4 * o + 1], prim[4 * o + 2], prim[4 * o + 3]};
    }
    HIPCHK(hipMemcpy(c->prim[c->cur], h.data(), n * sizeof(V4d), hipMemcpyHostToDevice));
  } else {
    std::vector<V4<float>> h(n);
    for (int j = 0; j < n; ++j) {
      const int o = c->perm[j];
      h[j] = V4<float>{(float)prim[4 * o], (float)prim[4 * o + 1], (float)prim[4 * o + 2],
                       (float)prim[4 * o + 3]};
    }
    HIPCHK(hipMemcpy(c->prim[c->cur], h.data(), n * sizeof(V4<float>), hipMemcpyHostToDevice));
  }
  return MF_OK;
}

// Downloads `count` V4<S> records per point from `src` and scatters them to the
// reference layout (point-major, `count*4` doubles per point, original order).
static int download_v4(mf_ctx* c, const void* src, int count, std::vector<double>& out) {
  const int n = c->n;
  out.assign((size_t)n * count * 4, 0.0);
  if (c->ssz == 8) {
    std::vector<V4d> h((size_t)n * count);
    HIPCHK(hipMemcpy(h.data(), src, h.size() * sizeof(V4d), hipMemcpyDeviceToHost));
    for (int j = 0; j < n; ++j)
      for (int a = 0; a < count; ++a)
        for (int r = 0; r < 4; ++r) out[((size_t)c->perm[j] * count + a) * 4 + r] = h[(size_t)count * j + a][r];
  } else {
    std::vector<V4<float>> h((size_t)n * count);
    HIPCHK(hipMemcpy(h.data(), src, h.size() * sizeof(V4<float>), hipMemcpyDeviceToHost));
    for (int j = 0; j < n; ++j)
      for (int a = 0; a < count; ++a)
        for (int r = 0; r < 4; ++r) out[((size_t)c->perm[j] * count + a) * 4 + r] = h[(size_t)count * j + a][r];
  }
  return MF_OK;
}

static int upload_v4(mf_ctx* c, void* dst, int count, const double* in) {
  const int n = c->n;
  if (c->ssz == 8) {
    std::vector<V4d> h((size_t)n * count);
    for (int j = 0; j < n; ++j)
      for (int a = 0; a < count; ++a)
        for (int r = 0; r < 4; ++r) h[(size_t)count * j + a][r] = in[((size_t)c->perm[j] * count + a) * 4 + r];
    HIPCHK(hipMemcpy(dst, h.data(), h.size() * sizeof(V4d), hipMemcpyHostToDevice));
  } else {
    std::vector<V4<float>> h((size_t)n * count);
    for (int j = 0; j < n; ++j)
      for (int a = 0; a < count; ++a)
        for (int r = 0; r < 4; ++r) h[(size_t)count * j + a][r] = (float)in[((size_t)c->perm[j] * count + a) * 4 + r];
    HIPCHK(hipMemcpy(dst, h.data(), h.size() * sizeof(V4<float>), hipMemcpyHostToDevice));
  }
  return MF_OK;
}

extern "C" {

const char* mf_last_error(void) { return g_err.c_str(); }

void mf_default_params(mf_params* p) {
  std::memset(p, 0, sizeof(*p));
  p->cfl = 0.0;
  p->vl_const = 150.0;
  p->power = 0.0;
  p->mach = 0.63;
  p->aoa = 2.0;
  p->euler = 1.0;
  p->rks = 4;
  p->inner_iterations = 0;
  p->shapes = 1;
  p->interior_points_normal_flag = 0;
  p->format = 1;
  p->restart = 0;
  p->precision = MF_FP64;
  p->reorder = 1;
}

struct AdjState;
static void adj_release(mf_ctx* c);
static void adj_forget_tuning(mf_ctx* c);
static void tan_release(mf_ctx* c);

int mf_destroy(mf_ctx* c) {
  if (!c) return MF_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  adj_release(c);
  tan_release(c);
  for (auto& gexec : c->graph)
    if (gexec) (void)hipGraphExecDestroy(gexec);
  for (void* a : c->allocs) (void)hipFree(a);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return MF_OK;
}

// Cases of a batch: each point's case (original order) and the cases' flow conditions
struct BatchSpec {
  int ncases = 0;
  std::vector<int32_t> case_of, first;  // first[c]: original index of case c's point 1
  std::vector<double> mach, aoa;
};

static int create_impl(const mf_geometry* geo, const mf_params* prm, mf_ctx* c, const BatchSpec* bs);

int mf_create(const mf_geometry* geo, const mf_params* prm, mf_ctx** out) {
  if (!geo || !prm || !out) return set_err(MF_ERR_INVALID, "mf_create: null argument");
  *out = nullptr;
  mf_ctx* c = new mf_ctx();
  int rc = create_impl(geo, prm, c, nullptr);
  if (rc != MF_OK) {
    std::string keep = g_err;
    mf_destroy(c);
    g_err = keep;
    return rc;
  }
  *out = c;
  return MF_OK;
}

}  // extern "C"

// Halo of the staged stencil walks in blocks of P consecutive (internal) points: each
// block's own points, then every other neighbour of them once (ascending), and per slot
// its position in that list with the slot's half-stencil bits above it.  Leaves
// h.hcap = 0 (staging unavailable) when a halo outgrows the 12-bit position.
// LDS holds the 99th percentile of the blocks' halo sizes rather than the largest (the
// largest belong to the few blocks of wall and outer points), so more blocks fit on a
// CU.  byte_words (the 64-point halos of the LSQ and inner kernels and their reverses):
// the slot words are 8-bit positions (loc8), the capacity at most 248, and 255 marks a
// position past it, whose neighbour the kernel reads through the point's own
// connectivity.  Otherwise (the flux kernels' 32-point halos) the 16-bit slot words
// (loc) carry the half-stencil bits, and a block whose halo exceeds the capacity walks
// the global records instead.
static int build_halo(mf_ctx* c, const std::vector<uint32_t>& conn, const std::vector<int32_t>& nbh, int P,
                      bool byte_words, Halo& h) {
  const int n = c->n, W = c->W;
  const int nblk = (n + P - 1) / P;
  const int stride = nblk * P;
  std::vector<int32_t> ptr(nblk + 1, 0), ids;
  std::vector<uint16_t> loc((size_t)W * stride, 0);
  ids.reserve((size_t)n * 4);
  std::vector<int32_t> other, sizes;
  sizes.reserve(nblk);
  h = Halo{};
  for (int b = 0; b < nblk; ++b) {
    const int p0 = b * P, p1 = std::min(n, p0 + P);
    other.clear();
    for (int i = p0; i < p1; ++i)
      for (int s = 0; s < nbh[i]; ++s) {
        const int k = (int)(conn[(size_t)s * n + i] >> 4);
        if (k < p0 || k >= p1) other.push_back(k);
      }
    std::sort(other.begin(), other.end());
    other.erase(std::unique(other.begin(), other.end()), other.end());
    const int own = p1 - p0, hn = own + (int)other.size();
    if (hn > 4096 - 8) return MF_OK;
    sizes.push_back(hn);
    for (int i = p0; i < p1; ++i) ids.push_back(i);
    ids.insert(ids.end(), other.begin(), other.end());
    for (int i = p0; i < p1; ++i)
      for (int s = 0; s < nbh[i]; ++s) {
        const uint32_t w = conn[(size_t)s * n + i];
        const int k = (int)(w >> 4);
        const int pos = (k >= p0 && k < p1)
                            ? k - p0
                            : own + (int)(std::lower_bound(other.begin(), other.end(), k) - other.begin());
        loc[(size_t)s * stride + i] = (uint16_t)(pos | ((w & 15u) << 12));
      }
    ptr[b + 1] = (int32_t)ids.size();
  }
  std::sort(sizes.begin(), sizes.end());
  int cap = sizes[std::min(nblk - 1, (int)((int64_t)nblk * 99 / 100))];
  if (byte_words) cap = std::min(cap, 248);
  const int hcap = (cap + 7) / 8 * 8;  // keeps every LDS array of the staged kernels 16-byte aligned
  int32_t *d_ptr, *d_ids;
  int rc;
  if ((rc = upload(c, &d_ptr, ptr)) || (rc = upload(c, &d_ids, ids))) return rc;
  if (byte_words) {
    std::vector<uint8_t> loc8(loc.size());
    for (size_t e = 0; e < loc.size(); ++e) {
      const int pos = loc[e] & 0xFFF;
      loc8[e] = (uint8_t)(pos < hcap ? pos : 255);
    }
    uint8_t* d_loc8;
    if ((rc = upload(c, &d_loc8, loc8))) return rc;
    h.loc8 = d_loc8;
  } else {
    uint16_t* d_loc;
    if ((rc = upload(c, &d_loc, loc))) return rc;
    h.loc = d_loc;
  }
  h.ptr = d_ptr;
  h.ids = d_ids;
  h.hcap = hcap;
  h.stride = stride;
  return MF_OK;
}

static int build_halos(mf_ctx* c, const std::vector<uint32_t>& conn, const std::vector<int32_t>& nbh) {
  c->g.width = c->W;
  int rc;
  if ((rc = build_halo(c, conn, nbh, STAGE_P, true, c->g.hl[0])) ||
      (rc = build_halo(c, conn, nbh, STAGE_PF, false, c->g.hl[1])))
    return rc;
  return MF_OK;
}

static int create_impl(const mf_geometry* geo, const mf_params* prm, mf_ctx* c, const BatchSpec* bs) {
  const int n = geo->n_points, W = geo->max_nbhs;
  if (n <= 0 || W <= 0) return set_err(MF_ERR_INVALID, "n_points and max_nbhs must be > 0");
  if ((int64_t)n >= (1ll << 27)) return set_err(MF_ERR_INVALID, "n_points must be < 2^27");
  if (!geo->x || !geo->y || !geo->min_dist || !geo->left || !geo->right || !geo->flag_1 ||
      !geo->flag_2 || !geo->nbhs || !geo->conn)
    return set_err(MF_ERR_INVALID, "mf_geometry has a null array");
  if (prm->precision != MF_FP64 && prm->precision != MF_MIXED)
    return set_err(MF_ERR_INVALID, "precision must be MF_FP64 or MF_MIXED");
  if (prm->rks != 1 && prm->rks != 4) return set_err(MF_ERR_INVALID, "rks must be 1 or 4");
  if (prm->inner_iterations < 0) return set_err(MF_ERR_INVALID, "inner_iterations < 0");
  if (prm->shapes < 1) return set_err(MF_ERR_INVALID, "shapes must be >= 1");
  if (prm->format == 2 && (!geo->nx || !geo->ny))
    return set_err(MF_ERR_INVALID, "format 2 needs file normals");
  if (prm->first_order != 0 && prm->first_order != 1)
    return set_err(MF_ERR_INVALID, "first_order must be 0 or 1");
  if (prm->block_size != 0 && prm->block_size != BLOCK)
    return set_err(MF_ERR_INVALID, "block_size must be 0 or %d", BLOCK);
  c->p = *prm;
  c->n = n;
  c->W = W;
  c->ssz = prm->precision == MF_FP64 ? 8 : 4;
  c->power0 = (prm->power == 0.0);

  // ---- validate the cloud (point_preprocessor.F90:357-385 reads the same fields)
  const int32_t* conn = geo->conn;
  auto nbr = [&](int i, int k) { return conn[(size_t)k * n + i] - 1; };  // 0-based
  for (int i = 0; i < n; ++i) {
    const int f = geo->flag_1[i];
    if (f < 0 || f > 2) return set_err(MF_ERR_INVALID, "point %d: flag_1 %d not in {0,1,2}", i + 1, f);
    if (geo->flag_2[i] > prm->shapes || geo->flag_2[i] < 0)
      return set_err(MF_ERR_INVALID, "point %d: flag_2 %d outside 0..shapes (shapes value wrong)", i + 1, geo->flag_2[i]);
    if (geo->flag_2[i] > 0 && f != 0)
      return set_err(MF_ERR_INVALID, "point %d: flag_2 > 0 on a non-wall point", i + 1);
    if (geo->nbhs[i] < 0 || geo->nbhs[i] > W)
      return set_err(MF_ERR_INVALID, "point %d: nbhs %d outside 0..%d", i + 1, geo->nbhs[i], W);
    for (int k = 0; k < geo->nbhs[i]; ++k) {
      const int j = nbr(i, k);
      if (j < 0 || j >= n) return set_err(MF_ERR_INVALID, "point %d: conn(%d) = %d out of range", i + 1, k + 1, j + 1);
    }
    if (f != 1) {
      const int l = geo->left[i] - 1, r = geo->right[i] - 1;
      if (l < 0 || l >= n || r < 0 || r >= n)
        return set_err(MF_ERR_INVALID, "boundary point %d: left/right out of range", i + 1);
    }
  }

  // ---- normals: point_normals.F90:8-105
  std::vector<double> nx(n, 0.0), ny(n, 0.0);
  for (int m = 0; m < n; ++m) {
    const int f = geo->flag_1[m];
    if (f == 0 || f == 2) {
      const int l = geo->left[m] - 1, r = geo->right[m] - 1;
      const double lx = geo->x[l], ly = geo->y[l], mx = geo->x[m], my = geo->y[m];
      const double rx = geo->x[r], ry = geo->y[r];
      const double nx1 = my - ly, nx2 = ry - my, ny1 = mx - lx, ny2 = rx - mx;
      double vx = 0.5 * (nx1 + nx2), vy = 0.5 * (ny1 + ny2);
      const double det = std::sqrt(vx * vx + vy * vy);
      nx[m] = -vx / det;
      ny[m] = vy / det;
    } else if (prm->format == 2) {
      nx[m] = geo->nx[m];
      ny[m] = geo->ny[m];
    } else if (prm->interior_points_normal_flag == 0) {
      nx[m] = 0.0;
      ny[m] = 1.0;
    } else {
      nx[m] = 1.0;
      ny[m] = 0.0;
    }
  }
  c->nx_h = nx;
  c->ny_h = ny;

  // ---- renumbering: class-major (wall, outer, interior), Hilbert order within a class
  std::vector<uint64_t> key(n, 0);
  if (prm->reorder) {
    double x0 = geo->x[0], x1 = geo->x[0], y0 = geo->y[0], y1 = geo->y[0];
    for (int i = 0; i < n; ++i) {
      x0 = std::min(x0, geo->x[i]); x1 = std::max(x1, geo->x[i]);
      y0 = std::min(y0, geo->y[i]); y1 = std::max(y1, geo->y[i]);
    }
    const double span = std::max(x1 - x0, y1 - y0) * (1.0 + 1e-9) + 1e-300;
    const uint32_t order = 21;
    const double scale = (double)((1u << order) - 1) / span;
    for (int i = 0; i < n; ++i)
      key[i] = hilbert_d(order, (uint64_t)((geo->x[i] - x0) * scale), (uint64_t)((geo->y[i] - y0) * scale));
  }
  auto cls = [&](int i) { return geo->flag_1[i] == 0 ? 0 : (geo->flag_1[i] == 2 ? 1 : 2); };
  // a batch: by case within each class, so every case is one contiguous range of each
  // class (its residual sums and its neighbour gathers stay together)
  auto case_of = [&](int i) { return bs ? bs->case_of[i] : 0; };
  c->perm.resize(n);
  std::iota(c->perm.begin(), c->perm.end(), 0);
  std::stable_sort(c->perm.begin(), c->perm.end(), [&](int a, int b) {
    const int ca = cls(a), cb = cls(b);
    if (ca != cb) return ca < cb;
    const int ka = case_of(a), kb = case_of(b);
    if (ka != kb) return ka < kb;
    return key[a] < key[b];
  });
  c->iperm.resize(n);
  for (int j = 0; j < n; ++j) c->iperm[c->perm[j]] = j;
  for (int i = 0; i < n; ++i) {
    if (geo->flag_1[i] == 0) ++c->nw;
    else if (geo->flag_1[i] == 2) ++c->no;
    else ++c->ni;
  }

  // ---- point-fixed geometry, accumulated in the reference's order (original points)
  std::vector<D2> xy(n), nrm(n);
  std::vector<int32_t> nbh_d(n);
  std::vector<uint32_t> conn_d((size_t)W * n, 0u);
  std::vector<V4d> lsq(n), c1(n), sxyv(n), obd(n);
  c->h_hxx.assign(n, V4d{0, 0, 0, 0});
  c->h_hyy.assign(n, V4d{0, 0, 0, 0});
  c->h_hxy.assign(n, V4d{0, 0, 0, 0});
  std::vector<double> eps_d(n);
  for (int j = 0; j < n; ++j) {
    const int i = c->perm[j];
    xy[j] = D2{geo->x[i], geo->y[i]};
    nrm[j] = D2{nx[i], ny[i]};
    nbh_d[j] = geo->nbhs[i];
    c->total_nbhs += geo->nbhs[i];
    // limiters.F90:88-89, (VL_CONST*min_dist)**3 as (e*e)*e
    double e = prm->vl_const * geo->min_dist[i];
    eps_d[j] = e * e * e;
    const double xi = geo->x[i], yi = geo->y[i];
    const double tx = ny[i], ty = -nx[i];
    // q_variables.F90:84-96 full-stencil sums
    double sxx = 0, syy = 0, sxy = 0;
    // wall/outer/interior_fluxes: per half-stencil sums in the rotated frame
    double hxx[4] = {0, 0, 0, 0}, hyy[4] = {0, 0, 0, 0}, hxy[4] = {0, 0, 0, 0};
    for (int k = 0; k < geo->nbhs[i]; ++k) {
      const int o = nbr(i, k);
      const double delx = geo->x[o] - xi, dely = geo->y[o] - yi;
      const double dist = std::sqrt(delx * delx + dely * dely);
      const double w = std::pow(dist, prm->power);
      sxx = sxx + delx * delx * w;
      syy = syy + dely * dely * w;
      sxy = sxy + delx * dely * w;
      const double dels = delx * tx + dely * ty;
      const double deln = delx * nx[i] + dely * ny[i];
      const unsigned m = split_mask(geo->flag_1[i], dels, deln);
      conn_d[(size_t)k * n + j] = ((uint32_t)c->iperm[o] << 4) | m;
      const double distr = std::sqrt(dels * dels + deln * deln);
      const double wr = std::pow(distr, prm->power);
      const double dsw = dels * wr, dnw = deln * wr;
      for (int d = 0; d < 4; ++d) {
        if (!(m & (1u << d))) continue;
        hxx[d] = hxx[d] + dels * dsw;
        hyy[d] = hyy[d] + deln * dnw;
        hxy[d] = hxy[d] + dels * dnw;
      }
    }
    const double det = sxx * syy - sxy * sxy;
    lsq[j] = V4d{sxx, syy, sxy, 1.0 / det};
    double cc[4], ob[4];
    for (int d = 0; d < 4; ++d) {
      const double dt = hxx[d] * hyy[d] - hxy[d] * hxy[d];
      ob[d] = 1.0 / dt;
      cc[d] = (d < 2) ? hyy[d] : hxx[d];
    }
    c1[j] = V4d{cc[0], cc[1], cc[2], cc[3]};
    c->h_hxx[j] = V4d{hxx[0], hxx[1], hxx[2], hxx[3]};
    c->h_hyy[j] = V4d{hyy[0], hyy[1], hyy[2], hyy[3]};
    c->h_hxy[j] = V4d{hxy[0], hxy[1], hxy[2], hxy[3]};
    sxyv[j] = V4d{hxy[0], hxy[1], hxy[2], hxy[3]};
    obd[j] = V4d{ob[0], ob[1], ob[2], ob[3]};
  }

  // ---- forces (compute_force_coeffs.F90:40-78): shape points grouped by shape id
  // (case * shapes + flag_2 - 1), in original order within an id
  const int ncase_eff = bs ? bs->ncases : 1;
  c->nsid = prm->shapes * ncase_eff;
  std::vector<int32_t> sh_m, sh_id, sh_pos, sh_ptr(c->nsid + 1, 0);
  std::vector<double> sgeo;
  {
    std::vector<int> order;
    for (int m = 0; m < n; ++m)
      if (geo->flag_2[m] > 0) order.push_back(m);
    std::vector<int32_t> ordinal(n, -1);
    for (size_t j = 0; j < order.size(); ++j) ordinal[order[j]] = (int32_t)j;
    auto sid = [&](int m) { return case_of(m) * prm->shapes + geo->flag_2[m] - 1; };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return sid(a) < sid(b); });
    for (int m : order) {
      const int l = geo->left[m] - 1, r = geo->right[m] - 1;
      const double lx = geo->x[l], ly = geo->y[l], mx = geo->x[m], my = geo->y[m];
      const double rx = geo->x[r], ry = geo->y[r];
      double ds1 = (mx - lx) * (mx - lx) + (my - ly) * (my - ly);
      ds1 = std::sqrt(ds1);
      double ds2 = (rx - mx) * (rx - mx) + (ry - my) * (ry - my);
      ds2 = std::sqrt(ds2);
      const double ds = 0.5 * (ds1 + ds2);
      sh_m.push_back(c->iperm[m]);
      sh_id.push_back(sid(m));
      sh_pos.push_back(ordinal[m]);
      sgeo.insert(sgeo.end(), {nx[m], ny[m], ds, mx - 0.25, my});
      ++sh_ptr[sid(m) + 1];
    }
    for (int k = 0; k < c->nsid; ++k) sh_ptr[k + 1] += sh_ptr[k];
  }
  c->nshape = (int)sh_m.size();
  c->h_sh_m = sh_m;
  c->left_i.assign(n, -1);
  c->right_i.assign(n, -1);
  for (int j = 0; j < n; ++j) {
    const int i = c->perm[j];
    if (geo->flag_1[i] != 1) {
      c->left_i[j] = c->iperm[geo->left[i] - 1];
      c->right_i[j] = c->iperm[geo->right[i] - 1];
    }
  }

  // ---- parameters: parameter.F90:16-34, compute_force_coeffs.F90:34, per case
  std::vector<Freestream> fs_case(ncase_eff);
  std::vector<double> th_case(ncase_eff), cps_case(ncase_eff);
  for (int k = 0; k < ncase_eff; ++k) {
    const double mach = bs ? bs->mach[k] : prm->mach, aoa = bs ? bs->aoa[k] : prm->aoa;
    th_case[k] = aoa * kPi / 180.0;
    fs_case[k].u1_inf = mach * std::cos(th_case[k]);
    fs_case[k].u2_inf = mach * std::sin(th_case[k]);
    cps_case[k] = 0.5 * 1.0 * mach * mach;
  }
  c->theta = th_case[0];
  c->fs = fs_case[0];
  c->cp_scale = cps_case[0];
  c->h_shape_par.resize(3 * (size_t)c->nsid);
  for (int sd = 0; sd < c->nsid; ++sd) {
    const int k = sd / prm->shapes;
    c->h_shape_par[3 * sd] = cps_case[k];
    c->h_shape_par[3 * sd + 1] = std::cos(th_case[k]);
    c->h_shape_par[3 * sd + 2] = std::sin(th_case[k]);
  }

  // ---- device
  HIPCHK(hipGetDevice(&c->device));
  HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  int rc;
  D2 *d_xy, *d_nrm;
  int32_t* d_nbhs;
  uint32_t* d_conn;
  V4d *d_lsq, *d_c1, *d_sxy, *d_obd;
  if ((rc = upload(c, &d_xy, xy)) || (rc = upload(c, &d_nrm, nrm)) ||
      (rc = upload(c, &d_nbhs, nbh_d)) || (rc = upload(c, &d_conn, conn_d)) ||
      (rc = upload(c, &d_lsq, lsq)) || (rc = upload(c, &d_c1, c1)) ||
      (rc = upload(c, &d_sxy, sxyv)) || (rc = upload(c, &d_obd, obd)))
    return rc;
  c->g = Geom{d_xy, d_nrm, d_nbhs, d_conn, d_lsq, d_c1, d_sxy, d_obd, n, c->nw, c->no, prm->first_order ? 1 : 0};
  c->g.axis_interior = 1;
  for (int j = c->nw + c->no; j < n; ++j)
    if (!(nrm[j].x == 0.0 && nrm[j].y == 1.0)) {
      c->g.axis_interior = 0;
      break;
    }
  if ((rc = build_halos(c, conn_d, nbh_d))) return rc;
  if (c->ssz == 4) {
    // the mixed path's solve coefficients (sum dy^2, sum dx dy, sum dx^2) / det in FP32
    std::vector<V4<float>> lf(n);
    for (int j = 0; j < n; ++j)
      lf[j] = V4<float>{(float)(lsq[j].y * lsq[j].w), (float)(lsq[j].z * lsq[j].w), (float)(lsq[j].x * lsq[j].w), 0.0f};
    V4<float>* d_lf;
    if ((rc = upload(c, &d_lf, lf))) return rc;
    c->g.lsq_geo_f = d_lf;
    // and the flux kernel's half-stencil coefficients (c1 * obd, sxy * obd) in FP32
    std::vector<V4<float>> df((size_t)2 * n);
    for (int j = 0; j < n; ++j)
      for (int d = 0; d < 4; ++d) {
        df[2 * j][d] = (float)(c1[j][d] * obd[j][d]);
        df[2 * j + 1][d] = (float)(sxyv[j][d] * obd[j][d]);
      }
    V4<float>* d_df;
    if ((rc = upload(c, &d_df, df))) return rc;
    c->g.dg_f = d_df;
  }
  {
    double* e;
    if ((rc = upload(c, &e, eps_d))) return rc;
    c->eps64 = e;
    c->eps = e;
  }
  if (c->ssz == 4) {
    std::vector<float> ef(eps_d.begin(), eps_d.end());
    float* e;
    if ((rc = upload(c, &e, ef))) return rc;
    c->eps = e;
  }
  const size_t v4b = 4 * c->ssz;
  for (int b = 0; b < 2; ++b) {
    char *pp, *qq, *dd;
    if ((rc = dev_alloc(c, &pp, (size_t)n * v4b)) || (rc = dev_alloc(c, &qq, (size_t)n * v4b)) ||
        (rc = dev_alloc(c, &dd, (size_t)2 * n * v4b)))
      return rc;
    c->prim[b] = pp;
    c->q[b] = qq;
    c->dq[b] = dd;
  }
  char *qmb, *pob;
  if ((rc = dev_alloc(c, &qmb, (size_t)2 * n * v4b)) || (rc = dev_alloc(c, &pob, (size_t)n * v4b))) return rc;
  if (c->ssz == 4 && (rc = dev_alloc(c, &c->qsum, (size_t)2 * n))) return rc;
  c->qm = qmb;
  c->prim_old = pob;
  // room for the partial sums of the largest lanes-per-point grid (8 lanes)
  // one residual partial per flux+update block: 8 lanes per point, or the staged walk
  c->npartial = (int)std::max(((int64_t)n * 8 + BLOCK - 1) / BLOCK, ((int64_t)n + STAGE_PF - 1) / STAGE_PF);
  if ((rc = dev_alloc(c, &c->flux_res, n)) || (rc = dev_alloc(c, &c->delta, n)) ||
      (rc = dev_alloc(c, &c->cspeed, n)) || (rc = dev_alloc(c, &c->partial, c->npartial)) ||
      (rc = dev_alloc(c, &c->scalar, 4)) || (rc = dev_alloc(c, &c->ctl, 1)))
    return rc;
  HIPCHK(hipMemset(c->flux_res, 0, (size_t)n * sizeof(V4d)));
  HIPCHK(hipMemset(c->delta, 0, (size_t)n * sizeof(double)));
  HIPCHK(hipMemset(c->prim_old, 0, (size_t)n * v4b));
  HIPCHK(hipMemset(c->dq[0], 0, (size_t)2 * n * v4b));
  HIPCHK(hipMemset(c->qm, 0, (size_t)2 * n * v4b));
  Ctl ctl0{0, 0, prm->restart ? prm->res_old : 0.0};
  HIPCHK(hipMemcpy(c->ctl, &ctl0, sizeof(Ctl), hipMemcpyHostToDevice));
  if (c->nshape > 0) {
    if ((rc = upload(c, &c->sh_m, sh_m)) || (rc = upload(c, &c->sh_id, sh_id)) ||
        (rc = upload(c, &c->sh_pos, sh_pos)) || (rc = upload(c, &c->sgeo, sgeo)) ||
        (rc = dev_alloc(c, &c->cp, c->nshape)))
      return rc;
  }
  if ((rc = upload(c, &c->sh_ptr, sh_ptr)) || (rc = upload(c, &c->shape_par, c->h_shape_par)) ||
      (rc = dev_alloc(c, &c->clcdcm, 3 * (size_t)c->nsid)))
    return rc;

  // ---- a batch: case of every point, the cases' freestreams, residual ranges
  if (bs) {
    c->ncases = bs->ncases;
    c->case_first = bs->first;
    std::vector<int32_t> cid(n), rng(5 * (size_t)bs->ncases, 0);
    for (int j = 0; j < n; ++j) cid[j] = bs->case_of[c->perm[j]];
    // points are sorted (class, case): each case is one range of the wall and of the
    // interior classes
    for (int k = 0; k < bs->ncases; ++k) {
      rng[5 * k] = rng[5 * k + 1] = 0;
      rng[5 * k + 2] = rng[5 * k + 3] = c->nw + c->no;
      rng[5 * k + 4] = bs->first[k + 1] - bs->first[k];
    }
    for (int j = c->nw - 1; j >= 0; --j) rng[5 * cid[j]] = j;
    for (int j = 0; j < c->nw; ++j) rng[5 * cid[j] + 1] = j + 1;
    for (int j = n - 1; j >= c->nw + c->no; --j) rng[5 * cid[j] + 2] = j;
    for (int j = c->nw + c->no; j < n; ++j) rng[5 * cid[j] + 3] = j + 1;
    int32_t* d_cid;
    Freestream* d_fs;
    if ((rc = upload(c, &d_cid, cid)) || (rc = upload(c, &d_fs, fs_case)) || (rc = upload(c, &c->case_rng, rng)) ||
        (rc = dev_alloc(c, &c->rs_pt, n)) || (rc = dev_alloc(c, &c->res_old_case, bs->ncases)))
      return rc;
    HIPCHK(hipMemset(c->rs_pt, 0, (size_t)n * sizeof(double)));
    std::vector<double> ro(bs->ncases, prm->restart ? prm->res_old : 0.0);
    HIPCHK(hipMemcpy(c->res_old_case, ro.data(), ro.size() * sizeof(double), hipMemcpyHostToDevice));
    c->g.case_id = d_cid;
    c->g.fs_tab = d_fs;
  }

  // ---- initial_conditions.F90:9-24: freestream (of the point's case)
  std::vector<double> prim0((size_t)4 * n);
  for (int i = 0; i < n; ++i) {
    const Freestream& f = fs_case[case_of(i)];
    prim0[4 * i] = 1.0;
    prim0[4 * i + 1] = f.u1_inf;
    prim0[4 * i + 2] = f.u2_inf;
    prim0[4 * i + 3] = 1.0 / 1.4;
  }
  if ((rc = upload_prim_original(c, prim0.data()))) return rc;
  if ((rc = q_variables(c))) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  return MF_OK;
}

// ---------------------------------------------------------------------------------
// kernel launchers
// ---------------------------------------------------------------------------------

static inline int nblocks(const mf_ctx* c) { return (c->n + BLOCK - 1) / BLOCK; }

template <class S>
static State<S> state_of(mf_ctx* c, int qbuf, int dqbuf) {
  return State<S>{(const V4<S>*)c->q[qbuf], (const V4<S>*)c->dq[dqbuf], (const V4<S>*)c->qm,
                  (const S*)c->eps};
}

static inline dim3 lane_blocks(const mf_ctx* c, int lpp) { return lane_grid(c->n, lpp); }

// eval_q_derivatives (+ func_delta) of q into dq, qm (and delta)
template <class S>
static void launch_lsq_on(mf_ctx* c, const V4<S>* q, V4<S>* dq, V4<S>* qm, const double* cspeed,
                          double* delta, bool with_delta, const Ctl* ctl, int lpp, V4<float>* qsum = nullptr) {
  launch_lsq_any<S>(env_of(c), q, dq, qm, cspeed, delta, with_delta, ctl, lpp, qsum);
}

// The mixed path's LSQ kernel also stores its first-pass sums (qsum) for the inner
// sweeps; an iteration without inner sweeps skips that store (32 bytes per point).
template <class S>
static void launch_lsq(mf_ctx* c, int qbuf, bool with_delta, const Ctl* ctl, bool sums = true) {
  launch_lsq_on<S>(c, (const V4<S>*)c->q[qbuf], (V4<S>*)c->dq[0], (V4<S>*)c->qm, c->cspeed, c->delta, with_delta,
                   ctl, c->lpp_lsq, sums ? c->qsum : nullptr);
}

// one inner iteration: dq_in -> dq_out
template <class S>
static void launch_inner_on(mf_ctx* c, const V4<S>* q, const V4<S>* a, V4<S>* b, const Ctl* ctl, int lpp,
                            const V4<float>* qsum = nullptr) {
  launch_inner_any<S>(env_of(c), q, a, b, ctl, lpp, qsum);
}

// the per-sweep shape of an inner-iteration shape (the resident sweeps' staged twin)
static int sweep_shape(int lpp) {
  return (lpp & MF_LPP_RESIDENT) ? (LPP_STAGED | (lpp & ~MF_LPP_RESIDENT)) : lpp;
}

template <class S>
static void launch_inner(mf_ctx* c, int qbuf, int din, const Ctl* ctl) {
  launch_inner_on<S>(c, (const V4<S>*)c->q[qbuf], (const V4<S>*)c->dq[din], (V4<S>*)c->dq[din ^ 1], ctl,
                     sweep_shape(c->lpp_inner), c->qsum);
}

// every inner sweep of a stage in one launch (lpp_inner with MF_LPP_RESIDENT): dq[0]
// in, the last sweep's dq in dq[sweeps & 1], as after `sweeps` launch_inner calls
static void launch_sweeps(mf_ctx* c, int sweeps, Ctl* ctl) {
  launch_sweeps_f32(env_of(c), (V4f*)c->dq[0], (V4f*)c->dq[1], sweeps, ctl, c->qsum, c->sbar, c->barrier_polls,
                    c->lpp_inner & ~MF_LPP_RESIDENT);
}

// fused flux residual + state update (rk) + q_variables of the new state
template <class S>
static void launch_flux_update_on(mf_ctx* c, const State<S>& st, const V4<S>* prim, V4<S>* pold, V4<S>* pout,
                                  V4<S>* qout, double* cspeed, const double* delta, double* partial,
                                  const Ctl* ctl, int rk, bool last, int lpp, double* rs_pt = nullptr) {
  launch_flux_update_any<S>(env_of(c), st, prim, pold, pout, qout, cspeed, delta, partial, ctl, rk, last, lpp,
                            rs_pt);
}

template <class S>
static void launch_flux_update(mf_ctx* c, int buf, int dqbuf, int rk, bool last) {
  launch_flux_update_on<S>(c, state_of<S>(c, buf, dqbuf), (const V4<S>*)c->prim[buf], (V4<S>*)c->prim_old,
                           (V4<S>*)c->prim[buf ^ 1], (V4<S>*)c->q[buf ^ 1], c->cspeed, c->delta, c->partial,
                           c->ctl, rk, last, c->lpp_flux, c->ncases > 0 ? c->rs_pt : nullptr);
}

template <class S>
static void launch_forces(mf_ctx* c, int buf, const Ctl* ctl, double* cd_hist) {
  const ForcesArgs f{c->sh_m, c->sh_pos, c->sgeo, c->sh_ptr, c->shape_par, c->nsid, c->cp, c->clcdcm};
  launch_forces_any<S>(env_of(c), (const V4<S>*)c->prim[buf], f, cd_hist, ctl);
}

static void launch_finalize(mf_ctx* c) {
  if (c->ncases > 0) {
    hipLaunchKernelGGL(k_finalize_batch, dim3(c->ncases), dim3(BLOCK), 0, c->stream, c->rs_pt, c->case_rng,
                       c->ncases, c->p.restart, c->ctl, c->res_old_case, c->res_hist,
                       c->resnew_hist);
    hipLaunchKernelGGL(k_tick, dim3(1), dim3(1), 0, c->stream, c->ctl);
    return;
  }
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(BLOCK), 0, c->stream, c->partial,
                     flux_blocks(c->n, c->lpp_flux), c->n, c->p.restart, c->ctl,
                     c->res_hist, c->resnew_hist);
}

// Kernel-kind ids used by the profiler
enum { KK_LSQ = 0, KK_INNER = 1, KK_FLUX = 2, KK_FIN = 3, KK_FORCES = 4 };

struct Timer {
  hipStream_t s;
  std::vector<hipEvent_t>* ev;  // pairs
  std::vector<int>* kind;
  void begin(int) {
    if (!ev) return;
    hipEvent_t a;
    (void)hipEventCreate(&a);
    (void)hipEventRecord(a, s);
    ev->push_back(a);
  }
  void end(int k) {
    if (!ev) return;
    hipEvent_t b;
    (void)hipEventCreate(&b);
    (void)hipEventRecord(b, s);
    ev->push_back(b);
    kind->push_back(k);
  }
};

// One fpi_solver iteration (fpi_solver.F90:20-70) starting from buffer `phase`;
// returns the phase after it.  Used for graph capture and for profiled eager runs.
template <class S>
static int record_iteration(mf_ctx* c, int phase, Timer* tm) {
  int buf = phase;
  // first order: the fluxes take q at both ends, so dq, qm and the inner iterations are
  // not needed; stage 1's LSQ kernel still runs for the time steps it fuses
  const bool first = c->p.first_order != 0;
  const int I = first ? 0 : c->p.inner_iterations;
  for (int rk = 1; rk <= c->p.rks; ++rk) {
    if (rk == 1 || !first) {
      if (tm) tm->begin(KK_LSQ);
      launch_lsq<S>(c, buf, rk == 1, c->ctl, I > 0);
      if (tm) tm->end(KK_LSQ);
    }
    int d = 0;
    if (is_f32<S> && I > 0 && (c->lpp_inner & MF_LPP_RESIDENT)) {
      if (tm) tm->begin(KK_INNER);
      launch_sweeps(c, I, c->ctl);
      if (tm) tm->end(KK_INNER);
      d = I & 1;
    } else {
      for (int it = 0; it < I; ++it) {
        if (tm) tm->begin(KK_INNER);
        launch_inner<S>(c, buf, d, c->ctl);
        if (tm) tm->end(KK_INNER);
        d ^= 1;
      }
    }
    if (tm) tm->begin(KK_FLUX);
    launch_flux_update<S>(c, buf, d, rk, rk == c->p.rks);
    if (tm) tm->end(KK_FLUX);
    buf ^= 1;
  }
  if (c->nshape > 0) {
    if (tm) tm->begin(KK_FORCES);
    launch_forces<S>(c, buf, c->ctl, c->cd_hist);
    if (tm) tm->end(KK_FORCES);
  }
  if (tm) tm->begin(KK_FIN);
  launch_finalize(c);
  if (tm) tm->end(KK_FIN);
  return buf;
}

static int record(mf_ctx* c, int phase, Timer* tm) {
  return c->ssz == 8 ? record_iteration<double>(c, phase, tm) : record_iteration<float>(c, phase, tm);
}

static int ensure_history(mf_ctx* c, int iters) {
  if (iters <= c->hist_cap && c->res_hist) return MF_OK;
  // histories are captured into the graphs: reallocate and drop the graphs, once the
  // launches already queued from them have finished
  HIPCHK(hipStreamSynchronize(c->stream));
  for (auto& gexec : c->graph)
    if (gexec) {
      HIPCHK(hipGraphExecDestroy(gexec));
      gexec = nullptr;
    }
  int cap = std::max(iters, 1024);
  const size_t per = (size_t)std::max(c->ncases, 1);
  int rc;
  if ((rc = dev_alloc(c, &c->res_hist, cap * per)) || (rc = dev_alloc(c, &c->resnew_hist, cap * per)) ||
      (rc = dev_alloc(c, &c->cd_hist, cap * (size_t)std::max(c->nsid, 1))))
    return rc;
  c->hist_cap = cap;
  return MF_OK;
}

// Iterations per graph launch on long runs: one launch replays this many captured
// iterations back to back (the iteration counter, residue histories and halt flag are
// device-side, so the replays need nothing from the host between iterations)
static constexpr int kGraphIters = 16;

// the graph of `n` iterations (1 or kGraphIters) from buffer phase `phase`
static int build_graph(mf_ctx* c, int phase, int n) {
  hipGraphExec_t& slot = c->graph[(n == 1 ? 0 : 2) + phase];
  if (slot) return MF_OK;
  hipGraph_t gr;
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int ph = phase;
  for (int j = 0; j < n; ++j) ph = record(c, ph, nullptr);
  HIPCHK(hipStreamEndCapture(c->stream, &gr));
  hipError_t e = hipGraphInstantiate(&slot, gr, nullptr, nullptr, 0);
  (void)hipGraphDestroy(gr);
  if (e != hipSuccess) return set_err(MF_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  return MF_OK;
}

// The graphs the next runs replay (one and kGraphIters iterations, from the current
// buffer phase and, when an iteration flips the phase, from the other), built when the
// launch shapes are chosen so that the first run after a tuning does not capture and
// instantiate them (about 2 ms for a 16-iteration graph at 2.56M points)
static int prebuild_graphs(mf_ctx* c) {
  int rc;
  if ((rc = ensure_history(c, kGraphIters))) return rc;
  for (int f = 0; f < ((c->p.rks & 1) ? 2 : 1); ++f)
    for (int n : {1, kGraphIters})
      if ((rc = build_graph(c, c->cur ^ f, n))) return rc;
  return MF_OK;
}

static int run_queue(mf_ctx* c, int32_t first_it, int32_t iters) {
  int rc;
  if ((rc = ensure_history(c, iters))) return rc;
  // refresh q and wave speeds from prim: the reference routines may have changed prim
  if ((rc = q_variables(c))) return rc;
  // first_it lives in the control block, so the captured graphs serve every first_it
  hipLaunchKernelGGL(k_begin_run, dim3(1), dim3(1), 0, c->stream, c->ctl, c->p.restart ? 1 : 0, c->p.res_old,
                     (int)first_it);
  HIPCHK(hipGetLastError());
  int phase = c->cur;
  int it = 0;
  for (; it + kGraphIters <= iters; it += kGraphIters) {
    if ((rc = build_graph(c, phase, kGraphIters))) return rc;
    HIPCHK(hipGraphLaunch(c->graph[2 + phase], c->stream));
    phase ^= (kGraphIters * c->p.rks) & 1;
  }
  for (; it < iters; ++it) {
    if ((rc = build_graph(c, phase, 1))) return rc;
    HIPCHK(hipGraphLaunch(c->graph[phase], c->stream));
    phase ^= (c->p.rks & 1);
  }
  c->dq_cur = c->p.first_order ? 0 : c->p.inner_iterations & 1;
  return MF_OK;
}

static int check_sweeps(mf_ctx* c);

extern "C" {

int mf_run_async(mf_ctx* c, int32_t first_it, int32_t iters) {
  if (!c || iters < 0) return set_err(MF_ERR_INVALID, "mf_run_async: bad argument");
  const int start = c->cur;
  int rc = run_queue(c, first_it, iters);
  if (rc) return rc;
  c->cur = start ^ ((iters * c->p.rks) & 1);
  return MF_OK;
}

int mf_sync(mf_ctx* c) {
  if (!c) return set_err(MF_ERR_INVALID, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  return check_sweeps(c);
}

int mf_run(mf_ctx* c, int32_t first_it, int32_t iters, double* residue_hist,
           double* res_new_hist, double* cd_hist, int32_t* iters_done) {
  if (!c || iters < 0) return set_err(MF_ERR_INVALID, "mf_run: bad argument");
  if (c->ncases > 0) return set_err(MF_ERR_STATE, "a batch context runs with mf_batch_run");
  const int start = c->cur;
  int rc = run_queue(c, first_it, iters);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  Ctl ctl;
  HIPCHK(hipMemcpy(&ctl, c->ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
  const int done = ctl.it;
  c->cur = start ^ ((done * c->p.rks) & 1);
  if (iters_done) *iters_done = done;
  if ((rc = check_sweeps(c))) return rc;
  if (residue_hist && done) HIPCHK(hipMemcpy(residue_hist, c->res_hist, done * sizeof(double), hipMemcpyDeviceToHost));
  if (res_new_hist && done) HIPCHK(hipMemcpy(res_new_hist, c->resnew_hist, done * sizeof(double), hipMemcpyDeviceToHost));
  if (cd_hist && done && c->nshape > 0) {
    std::vector<double> h((size_t)done * c->nsid);
    HIPCHK(hipMemcpy(h.data(), c->cd_hist, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int j = 0; j < done; ++j) cd_hist[j] = h[(size_t)j * c->nsid];
  }
  if (iters_done) *iters_done = done;
  return MF_OK;
}

int mf_profile_kernels(mf_ctx* c, int32_t iters, double* mean_ms, int32_t* launches) {
  if (!c || iters <= 0 || !mean_ms) return set_err(MF_ERR_INVALID, "mf_profile_kernels: bad argument");
  int rc;
  if ((rc = ensure_history(c, iters))) return rc;
  if ((rc = q_variables(c))) return rc;
  hipLaunchKernelGGL(k_begin_run, dim3(1), dim3(1), 0, c->stream, c->ctl, c->p.restart ? 1 : 0, c->p.res_old, 1);
  std::vector<hipEvent_t> ev;
  std::vector<int> kind;
  Timer tm{c->stream, &ev, &kind};
  int phase = c->cur;
  for (int it = 0; it < iters; ++it) phase = record(c, phase, &tm);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  double sum[MF_NUM_KERNEL_KINDS] = {0};
  int cnt[MF_NUM_KERNEL_KINDS] = {0};
  for (size_t j = 0; j < kind.size(); ++j) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ev[2 * j], ev[2 * j + 1]));
    sum[kind[j]] += ms;
    cnt[kind[j]] += 1;
  }
  for (auto e : ev) (void)hipEventDestroy(e);
  for (int k = 0; k < MF_NUM_KERNEL_KINDS; ++k) {
    mean_ms[k] = cnt[k] ? sum[k] / cnt[k] : 0.0;
    if (launches) launches[k] = cnt[k] / iters;
  }
  Ctl ctl;
  HIPCHK(hipMemcpy(&ctl, c->ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
  c->cur = c->cur ^ ((ctl.it * c->p.rks) & 1);
  c->dq_cur = c->p.first_order ? 0 : c->p.inner_iterations & 1;
  return check_sweeps(c);
}

}  // extern "C"

static void drop_graphs(mf_ctx* c) {
  (void)hipStreamSynchronize(c->stream);
  for (auto& gexec : c->graph)
    if (gexec) {
      (void)hipGraphExecDestroy(gexec);
      gexec = nullptr;
    }
}

static bool valid_lpp(int v) { return v == 1 || v == 2 || v == 4 || v == 8; }

// Largest dynamic LDS of one block of a staged kernel
static constexpr size_t kMaxStageLds = 64 * 1024;

// The resident inner sweeps with `lanes` lanes per point: mixed path, halos in LDS, and
// every 64-point block of the cloud resident at once by the occupancy calculator (single
// cases: a batch's residue logic does not stop at a halt)
static bool valid_resident(const mf_ctx* c, int lanes) {
  if (c->ssz != 4 || c->ncases > 0 || (lanes != 1 && lanes != 2 && lanes != 4) || c->g.hl[0].hcap <= 0) return false;
  const size_t lds = sweeps_lds(c->g.hl[0].hcap, c->g.width, lanes);
  if (lds > kMaxStageLds) return false;
  // a tenth of the calculator's capacity held back: the dispatcher does not fill every
  // slot of every CU
  const int64_t blocks = (c->n + STAGE_P - 1) / STAGE_P;
  return 10 * blocks <= 9 * (int64_t)sweeps_capacity_f32(lanes, c->power0, lds);
}

// A forward stencil kernel's shape: kind 0 LSQ, 1 inner iteration, 2 flux+update.  The
// staged walks (LPP_STAGED | lanes) exist when the context's halos fit in LDS: 1, 2 or 4
// lanes for the LSQ and inner kernels, 2, 4 or 8 for the flux kernel.
static bool valid_shape(const mf_ctx* c, int kind, int v) {
  if (valid_lpp(v)) return true;
  if (kind == 1 && (v & MF_LPP_RESIDENT)) return valid_resident(c, v & ~MF_LPP_RESIDENT);
  if (!(v & LPP_STAGED)) return false;
  const int l = v & ~LPP_STAGED;
  const Halo& h = c->g.hl[kind == 2 ? 1 : 0];
  if (h.hcap <= 0) return false;
  if (kind == 2 ? (l != 2 && l != 4 && l != 8) : (l != 1 && l != 2 && l != 4)) return false;
  if ((kind == 2 ? STAGE_PF : STAGE_P) * l > BLOCK) return false;
  const size_t lds = kind == 0   ? stage_lds_lsq(h.hcap, c->g.width, c->ssz, true)
                     : kind == 1 ? stage_lds_inner(h.hcap, c->g.width, c->ssz, c->ssz == 4)
                                 : stage_lds_flux(h.hcap, c->g.width, c->ssz);
  return lds <= kMaxStageLds;
}

// staged shapes of the adjoint's LSQ reverse (kind 1) and inner reverse (kind 2): 1, 2
// or 4 lanes, 64-point blocks, halo in LDS
static bool valid_rev_staged(const mf_ctx* c, int kind, int lpp) {
  const int l = lpp & ~LPP_STAGED;
  if (!(lpp & LPP_STAGED) || (l != 1 && l != 2 && l != 4) || c->g.hl[0].hcap <= 0) return false;
  if (STAGE_P * l > BLOCK) return false;
  const size_t lds = kind == 1 ? stage_lds_lsq_rev(c->g.hl[0].hcap, c->g.width)
                               : stage_lds_inner_rev(c->g.hl[0].hcap, c->g.width);
  return lds <= kMaxStageLds;
}

// the pull forms of the adjoint's LSQ and inner reverses: MF_LPP_PULL | 1, 2 or 4 lanes
// (their halo is built with the adjoint's buffers and sized to fit LDS)
static bool valid_pull(int lpp) {
  const int l = lpp & ~LPP_PULL;
  return (lpp & LPP_PULL) && (l == 1 || l == 2 || l == 4);
}

// mean time (us) of `reps` launches of fn on the context's stream, after 2 untimed
template <class F>
static int time_launches(mf_ctx* c, int reps, F&& fn, double* us) {
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  fn();
  fn();
  HIPCHK(hipEventRecord(a, c->stream));
  for (int r = 0; r < reps; ++r) fn();
  HIPCHK(hipEventRecord(b, c->stream));
  HIPCHK(hipEventSynchronize(b));
  HIPCHK(hipGetLastError());
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *us = 1e3 * (double)ms / reps;
  return MF_OK;
}

// The resident sweeps' barrier state, allocated and zeroed on first use
static int ensure_sweep_bar(mf_ctx* c) {
  if (c->sbar) return MF_OK;
  int rc;
  if ((rc = dev_alloc(c, &c->sbar, 1))) return rc;
  HIPCHK(hipMemset(c->sbar, 0, sizeof(SweepBar)));
  return MF_OK;
}

// After a run: a wait at the resident sweeps' barrier gave up (Ctl::halt ==
// HALT_BARRIER).  The barrier state is zeroed, the context goes back to the staged
// inner kernel with the same lanes (its graphs are rebuilt), and the run reports an error.
static int barrier_gave_up(mf_ctx* c) {
  drop_graphs(c);
  if (c->sbar) HIPCHK(hipMemset(c->sbar, 0, sizeof(SweepBar)));
  c->lpp_inner = sweep_shape(c->lpp_inner);
  return set_err(MF_ERR_HIP,
                 "resident inner sweeps: a wait at the grid barrier gave up (blocks not all resident); the run "
                 "stopped and the context now launches one sweep at a time");
}

// After queued work has finished: an error if a wait at the resident sweeps' barrier has
// given up since the last check (the barrier's abort flag stays set until this resets it)
static int check_sweeps(mf_ctx* c) {
  if (!c->sbar) return MF_OK;
  uint32_t abort_flag = 0;
  HIPCHK(hipMemcpy(&abort_flag, &c->sbar->abort, sizeof(uint32_t), hipMemcpyDeviceToHost));
  return abort_flag ? barrier_gave_up(c) : MF_OK;
}

// The timed launches write only scratch between iterations: dq, qm, and the next
// buffers of prim and q (stage 2: no prim_old, wave speed or residual writes).
static int prepare_timing(mf_ctx* c) {
  int rc;
  if ((rc = q_variables(c))) return rc;
  if (!c->tune_ctl && (rc = dev_alloc(c, &c->tune_ctl, 1))) return rc;
  const Ctl z{0, 0, 1.0};
  HIPCHK(hipMemcpy(c->tune_ctl, &z, sizeof(Ctl), hipMemcpyHostToDevice));
  return MF_OK;
}

// Mean time of one launch of kernel `kind` (0 LSQ, 1 inner sweep, 2 flux+update) with
// lanes-per-point shape lpp over reps back-to-back launches.  The inner sweep reads the
// dq the LSQ kernel last wrote to dq[0].
template <class S>
static int time_shape(mf_ctx* c, int kind, int lpp, int reps, double* us) {
  const int cur = c->cur;
  auto q = (const V4<S>*)c->q[cur];
  if (kind == 0)
    return time_launches(c, reps, [&] {
      launch_lsq_on<S>(c, q, (V4<S>*)c->dq[0], (V4<S>*)c->qm, c->cspeed, c->delta, false, nullptr, lpp,
                       c->p.inner_iterations > 0 ? c->qsum : nullptr);
    }, us);
  if (kind == 1 && (lpp & MF_LPP_RESIDENT)) {
    // one launch runs every inner sweep of a stage: the time per sweep
    const int I = std::max(c->p.inner_iterations, 1);
    int rc;
    if ((rc = ensure_sweep_bar(c))) return rc;
    const int keep = c->lpp_inner;
    c->lpp_inner = lpp;
    rc = time_launches(c, reps, [&] { launch_sweeps(c, I, c->tune_ctl); }, us);
    c->lpp_inner = keep;
    if (rc) return rc;
    *us /= I;
    Ctl tc;
    HIPCHK(hipMemcpy(&tc, c->tune_ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    if (tc.halt) {
      // a wait gave up: not a candidate
      *us = 1e300;
      const Ctl z{0, 0, 1.0};
      HIPCHK(hipMemset(c->sbar, 0, sizeof(SweepBar)));
      HIPCHK(hipMemcpy(c->tune_ctl, &z, sizeof(Ctl), hipMemcpyHostToDevice));
    }
    return MF_OK;
  }
  if (kind == 1)
    return time_launches(c, reps, [&] {
      launch_inner_on<S>(c, q, (const V4<S>*)c->dq[0], (V4<S>*)c->dq[1], nullptr, sweep_shape(lpp), c->qsum);
    }, us);
  const State<S> st = state_of<S>(c, cur, 0);
  return time_launches(c, reps, [&] {
    launch_flux_update_on<S>(c, st, (const V4<S>*)c->prim[cur], (V4<S>*)c->prim_old, (V4<S>*)c->prim[cur ^ 1],
                             (V4<S>*)c->q[cur ^ 1], c->cspeed, c->delta, c->partial, c->tune_ctl, 2, false, lpp);
  }, us);
}

// the LSQ kernel at the context's own shape, so that the inner sweeps read its dq
template <class S>
static void refresh_dq(mf_ctx* c, int lpp) {
  launch_lsq_on<S>(c, (const V4<S>*)c->q[c->cur], (V4<S>*)c->dq[0], (V4<S>*)c->qm, c->cspeed, c->delta, false,
                   nullptr, lpp, c->qsum);
}

template <class S>
static int autotune_impl(mf_ctx* c, int reps, mf_tuning* t) {
  int rc;
  if ((rc = prepare_timing(c))) return rc;
  const int cand[4] = {1, 2, 4, 8};
  // and every staged walk that fits
  std::vector<int> cand_st[3];
  for (int k = 0; k < 3; ++k) {
    cand_st[k].assign(cand, cand + 4);
    for (int l : {1, 2, 4, 8})
      if (valid_shape(c, k, LPP_STAGED | l)) cand_st[k].push_back(LPP_STAGED | l);
  }
  // and the resident inner sweeps
  if (c->p.inner_iterations > 0)
    for (int l : {1, 2, 4})
      if (valid_shape(c, 1, MF_LPP_RESIDENT | l)) cand_st[1].push_back(MF_LPP_RESIDENT | l);
  double best[3] = {1e300, 1e300, 1e300};
  int pick[3] = {1, 1, 1};
  for (int k = 0; k < 3; ++k) {
    if (k == 1) refresh_dq<S>(c, 1);
    for (int lpp : cand_st[k]) {
      double us;
      if ((rc = time_shape<S>(c, k, lpp, reps, &us))) return rc;
      if (us < best[k]) best[k] = us, pick[k] = lpp;
    }
  }
  c->lpp_lsq = pick[0];
  c->lpp_inner = pick[1];
  c->lpp_flux = pick[2];
  drop_graphs(c);
  if ((rc = prebuild_graphs(c))) return rc;
  if (t)
    *t = mf_tuning{pick[0], pick[1], pick[2], (int32_t)c->barrier_polls, best[0], best[1], best[2],
                   c->lpp_rev[0], c->lpp_rev[1], c->lpp_rev[2], c->lpp_rev[3]};
  return MF_OK;
}

template <class S>
static int time_kernels_impl(mf_ctx* c, int reps, double* us) {
  int rc;
  if ((rc = prepare_timing(c))) return rc;
  const int lpp[3] = {c->lpp_lsq, c->lpp_inner, c->lpp_flux};
  for (int k = 0; k < 3; ++k) {
    if (k == 1) refresh_dq<S>(c, lpp[0]);
    if ((rc = time_shape<S>(c, k, lpp[k], reps, &us[k]))) return rc;
  }
  return MF_OK;
}

extern "C" {

int mf_autotune(mf_ctx* c, int32_t reps, mf_tuning* out) {
  if (!c) return set_err(MF_ERR_INVALID, "null context");
  if (reps <= 0) reps = 10;
  return c->ssz == 8 ? autotune_impl<double>(c, reps, out) : autotune_impl<float>(c, reps, out);
}

int mf_time_kernels(mf_ctx* c, int32_t reps, double* us) {
  if (!c || !us) return set_err(MF_ERR_INVALID, "mf_time_kernels: null argument");
  if (c->ncases > 0) return set_err(MF_ERR_STATE, "mf_time_kernels times a single-case context");
  if (reps <= 0) reps = 20;
  return c->ssz == 8 ? time_kernels_impl<double>(c, reps, us) : time_kernels_impl<float>(c, reps, us);
}

int mf_get_tuning(mf_ctx* c, mf_tuning* out) {
  if (!c || !out) return set_err(MF_ERR_INVALID, "mf_get_tuning: null argument");
  *out = mf_tuning{c->lpp_lsq, c->lpp_inner, c->lpp_flux, (int32_t)c->barrier_polls, 0.0, 0.0, 0.0,
                   c->lpp_rev[0], c->lpp_rev[1], c->lpp_rev[2], c->lpp_rev[3]};
  return MF_OK;
}

int mf_set_tuning(mf_ctx* c, const mf_tuning* t) {
  if (!c || !t) return set_err(MF_ERR_INVALID, "mf_set_tuning: null argument");
  if (!valid_shape(c, 0, t->lpp_lsq) || !valid_shape(c, 1, t->lpp_inner) || !valid_shape(c, 2, t->lpp_flux))
    return set_err(MF_ERR_INVALID,
                   "lanes per point must be 1, 2, 4 or 8 (LSQ and inner kernels also MF_LPP_STAGED | 1, 2 or 4 "
                   "when the context's halos fit in LDS; the inner kernel MF_LPP_RESIDENT | 1, 2 or 4 on a "
                   "mixed context whose blocks all fit on the device)");
  if (t->barrier_polls < 0) return set_err(MF_ERR_INVALID, "barrier_polls must be >= 0");
  const int32_t rv[4] = {t->lpp_flux_rev, t->lpp_lsq_rev, t->lpp_inner_rev, t->lpp_gather};
  for (int k = 0; k < 4; ++k)
    if (rv[k] != 0 && !valid_lpp(rv[k]) && !(k == 0 && valid_shape(c, 2, rv[k])) &&
        !((k == 1 || k == 2) && valid_rev_staged(c, k, rv[k])) && !((k == 1 || k == 2) && valid_pull(rv[k])))
      return set_err(MF_ERR_INVALID,
                     "reverse lanes per point must be 0 (timed), 1, 2, 4 or 8 (the flux, LSQ and inner reverses "
                     "also their staged shapes, the LSQ and inner reverses MF_LPP_PULL | 1, 2 or 4)");
  int rc;
  if ((t->lpp_inner & MF_LPP_RESIDENT) && (rc = ensure_sweep_bar(c))) return rc;
  c->lpp_lsq = t->lpp_lsq;
  c->lpp_inner = t->lpp_inner;
  c->lpp_flux = t->lpp_flux;
  c->barrier_polls = t->barrier_polls > 0 ? (uint32_t)t->barrier_polls : kDefaultBarrierPolls;
  for (int k = 0; k < 4; ++k) c->lpp_rev[k] = rv[k];
  adj_forget_tuning(c);
  drop_graphs(c);
  return prebuild_graphs(c);
}

int mf_batch_run(mf_ctx* c, int32_t first_it, int32_t iters, double* residue, double* res_new, double* cd) {
  if (!c || iters < 0) return set_err(MF_ERR_INVALID, "mf_batch_run: bad argument");
  if (c->ncases == 0) return set_err(MF_ERR_STATE, "mf_batch_run needs a batch context (mf_create_batch)");
  const int start = c->cur;
  int rc = run_queue(c, first_it, iters);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  c->cur = start ^ ((iters * c->p.rks) & 1);
  const size_t nr = (size_t)iters * c->ncases;
  if (residue && nr) HIPCHK(hipMemcpy(residue, c->res_hist, nr * sizeof(double), hipMemcpyDeviceToHost));
  if (res_new && nr) HIPCHK(hipMemcpy(res_new, c->resnew_hist, nr * sizeof(double), hipMemcpyDeviceToHost));
  if (cd && iters && c->nshape > 0)
    HIPCHK(hipMemcpy(cd, c->cd_hist, (size_t)iters * c->nsid * sizeof(double), hipMemcpyDeviceToHost));
  return MF_OK;
}

int mf_batch_cases(mf_ctx* c, int32_t* ncases, int32_t* first) {
  if (!c) return set_err(MF_ERR_INVALID, "null context");
  if (ncases) *ncases = c->ncases;
  if (first)
    for (size_t k = 0; k < c->case_first.size(); ++k) first[k] = c->case_first[k];
  return MF_OK;
}

int mf_create_batch(const mf_case* cases, int32_t ncases, const mf_params* prm, mf_ctx** out) {
  if (!cases || !prm || !out || ncases < 1) return set_err(MF_ERR_INVALID, "mf_create_batch: bad argument");
  *out = nullptr;
  // one cloud of the cases' points, in case order, with their links offset
  BatchSpec bs;
  bs.ncases = ncases;
  bs.first.assign(ncases + 1, 0);
  int W = 0;
  int64_t total = 0;
  for (int k = 0; k < ncases; ++k) {
    const mf_geometry* g = cases[k].geometry;
    if (!g || g->n_points <= 0 || g->max_nbhs <= 0 || !g->x || !g->y || !g->min_dist || !g->left || !g->right ||
        !g->flag_1 || !g->flag_2 || !g->nbhs || !g->conn)
      return set_err(MF_ERR_INVALID, "case %d: incomplete geometry", k + 1);
    W = std::max(W, (int)g->max_nbhs);
    total += g->n_points;
    if (total >= (1ll << 27)) return set_err(MF_ERR_INVALID, "a batch must hold fewer than 2^27 points");
    bs.first[k + 1] = (int32_t)total;
    bs.mach.push_back(cases[k].mach);
    bs.aoa.push_back(cases[k].aoa);
  }
  const int n = (int)total;
  std::vector<double> x(n), y(n), fnx(n, 0.0), fny(n, 1.0), md(n);
  std::vector<int32_t> left(n), right(n), f1(n), f2(n), nb(n), conn((size_t)W * n, 0);
  bs.case_of.resize(n);
  for (int k = 0; k < ncases; ++k) {
    const mf_geometry* g = cases[k].geometry;
    const int o = bs.first[k], m = g->n_points;
    for (int i = 0; i < m; ++i) {
      const int j = o + i;
      x[j] = g->x[i];
      y[j] = g->y[i];
      if (g->nx && g->ny) {
        fnx[j] = g->nx[i];
        fny[j] = g->ny[i];
      }
      md[j] = g->min_dist[i];
      left[j] = g->left[i] > 0 ? g->left[i] + o : 0;
      right[j] = g->right[i] > 0 ? g->right[i] + o : 0;
      f1[j] = g->flag_1[i];
      f2[j] = g->flag_2[i];
      nb[j] = g->nbhs[i];
      if (nb[j] < 0 || nb[j] > g->max_nbhs) return set_err(MF_ERR_INVALID, "case %d point %d: bad nbhs", k + 1, i + 1);
      for (int s = 0; s < nb[j]; ++s) {
        const int32_t v = g->conn[(size_t)s * m + i];
        if (v < 1 || v > m) return set_err(MF_ERR_INVALID, "case %d point %d: conn(%d) = %d out of range", k + 1, i + 1, s + 1, v);
        conn[(size_t)s * n + j] = v + o;
      }
      bs.case_of[j] = k;
    }
  }
  const mf_geometry merged{n, W, x.data(), y.data(), fnx.data(), fny.data(), md.data(), left.data(), right.data(),
                           f1.data(), f2.data(), nb.data(), conn.data()};
  mf_ctx* c = new mf_ctx();
  int rc = create_impl(&merged, prm, c, &bs);
  if (rc != MF_OK) {
    std::string keep = g_err;
    mf_destroy(c);
    g_err = keep;
    return rc;
  }
  *out = c;
  return MF_OK;
}

int mf_set_prim(mf_ctx* c, const double* prim) {
  if (!c || !prim) return set_err(MF_ERR_INVALID, "mf_set_prim: null argument");
  int rc = upload_prim_original(c, prim);
  if (rc) return rc;
  return q_variables(c);
}

int mf_get_prim(mf_ctx* c, double* prim) {
  if (!c || !prim) return set_err(MF_ERR_INVALID, "mf_get_prim: null argument");
  HIPCHK(hipStreamSynchronize(c->stream));
  std::vector<double> out;
  int rc = download_v4(c, c->prim[c->cur], 1, out);
  if (rc) return rc;
  std::memcpy(prim, out.data(), out.size() * sizeof(double));
  return MF_OK;
}

int mf_get_field(mf_ctx* c, int32_t field, double* out) {
  if (!c || !out) return set_err(MF_ERR_INVALID, "mf_get_field: null argument");
  HIPCHK(hipStreamSynchronize(c->stream));
  const int n = c->n;
  std::vector<double> v;
  int rc = MF_OK;
  switch (field) {
    case MF_FIELD_Q: rc = download_v4(c, c->q[c->cur], 1, v); break;
    case MF_FIELD_PRIM_OLD: rc = download_v4(c, c->prim_old, 1, v); break;
    case MF_FIELD_DQ:
    case MF_FIELD_QM: {
      std::vector<double> t;
      rc = download_v4(c, field == MF_FIELD_DQ ? c->dq[c->dq_cur] : c->qm, 2, t);
      if (rc) return rc;
      // (point, a, r) -> reference (a fastest, then r): out[i*8 + r*2 + a]
      v.assign((size_t)8 * n, 0.0);
      for (int i = 0; i < n; ++i)
        for (int a = 0; a < 2; ++a)
          for (int r = 0; r < 4; ++r) v[(size_t)8 * i + 2 * r + a] = t[(size_t)8 * i + 4 * a + r];
      break;
    }
    case MF_FIELD_FLUX_RES: {
      std::vector<V4d> h(n);
      HIPCHK(hipMemcpy(h.data(), c->flux_res, n * sizeof(V4d), hipMemcpyDeviceToHost));
      v.assign((size_t)4 * n, 0.0);
      for (int j = 0; j < n; ++j)
        for (int r = 0; r < 4; ++r) v[(size_t)4 * c->perm[j] + r] = h[j][r];
      break;
    }
    case MF_FIELD_DELTA: {
      std::vector<double> h(n);
      HIPCHK(hipMemcpy(h.data(), c->delta, n * sizeof(double), hipMemcpyDeviceToHost));
      v.assign(n, 0.0);
      for (int j = 0; j < n; ++j) v[c->perm[j]] = h[j];
      break;
    }
    case MF_FIELD_NX: v = c->nx_h; break;
    case MF_FIELD_NY: v = c->ny_h; break;
    default: return set_err(MF_ERR_INVALID, "mf_get_field: unknown field %d", field);
  }
  if (rc) return rc;
  std::memcpy(out, v.data(), v.size() * sizeof(double));
  return MF_OK;
}

int mf_set_field(mf_ctx* c, int32_t field, const double* in) {
  if (!c || !in) return set_err(MF_ERR_INVALID, "mf_set_field: null argument");
  HIPCHK(hipStreamSynchronize(c->stream));
  const int n = c->n;
  switch (field) {
    case MF_FIELD_Q: return upload_v4(c, c->q[c->cur], 1, in);
    case MF_FIELD_PRIM_OLD: return upload_v4(c, c->prim_old, 1, in);
    case MF_FIELD_DQ:
    case MF_FIELD_QM: {
      std::vector<double> t((size_t)8 * n);
      for (int i = 0; i < n; ++i)
        for (int a = 0; a < 2; ++a)
          for (int r = 0; r < 4; ++r) t[(size_t)8 * i + 4 * a + r] = in[(size_t)8 * i + 2 * r + a];
      return upload_v4(c, field == MF_FIELD_DQ ? c->dq[c->dq_cur] : c->qm, 2, t.data());
    }
    case MF_FIELD_DELTA: {
      std::vector<double> h(n);
      for (int j = 0; j < n; ++j) h[j] = in[c->perm[j]];
      HIPCHK(hipMemcpy(c->delta, h.data(), n * sizeof(double), hipMemcpyHostToDevice));
      return MF_OK;
    }
    default: return set_err(MF_ERR_INVALID, "mf_set_field: field %d is not settable", field);
  }
}

int mf_eval_q_variables(mf_ctx* c) {
  if (!c) return set_err(MF_ERR_INVALID, "null context");
  return q_variables(c);
}

int mf_func_delta(mf_ctx* c) {
  if (!c) return set_err(MF_ERR_INVALID, "null context");
  if (c->ssz == 8) launch_func_delta_any<double>(env_of(c), (const V4d*)c->prim[c->cur], c->delta);
  else launch_func_delta_any<float>(env_of(c), (const V4<float>*)c->prim[c->cur], c->delta);
  HIPCHK(hipGetLastError());
  return MF_OK;
}

int mf_eval_q_derivatives(mf_ctx* c) {
  if (!c) return set_err(MF_ERR_INVALID, "null context");
  if (c->ssz == 8) launch_lsq<double>(c, c->cur, false, nullptr);
  else launch_lsq<float>(c, c->cur, false, nullptr);
  HIPCHK(hipGetLastError());
  c->dq_cur = 0;
  return MF_OK;
}

int mf_eval_q_inner_loop(mf_ctx* c) {
  if (!c) return set_err(MF_ERR_INVALID, "null context");
  if (c->ssz == 8) launch_inner<double>(c, c->cur, c->dq_cur, nullptr);
  else launch_inner<float>(c, c->cur, c->dq_cur, nullptr);
  HIPCHK(hipGetLastError());
  c->dq_cur ^= 1;
  return MF_OK;
}

int mf_cal_flux_residual(mf_ctx* c) {
  if (!c) return set_err(MF_ERR_INVALID, "null context");
  if (c->ssz == 8) launch_flux_any<double>(env_of(c), state_of<double>(c, c->cur, c->dq_cur), c->delta, c->flux_res);
  else launch_flux_any<float>(env_of(c), state_of<float>(c, c->cur, c->dq_cur), c->delta, c->flux_res);
  HIPCHK(hipGetLastError());
  return MF_OK;
}

int mf_state_update(mf_ctx* c, int32_t rk, double* sum_res_sqr) {
  if (!c) return set_err(MF_ERR_INVALID, "null context");
  if (rk < 1 || rk > 4) return set_err(MF_ERR_INVALID, "rk must be 1..4");
  if (c->ssz == 8)
    launch_update_any<double>(env_of(c), (V4d*)c->prim[c->cur], (const V4d*)c->prim_old, c->flux_res, c->partial, rk);
  else
    launch_update_any<float>(env_of(c), (V4<float>*)c->prim[c->cur], (const V4<float>*)c->prim_old, c->flux_res,
                             c->partial, rk);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(BLOCK), 0, c->stream, c->partial, nblocks(c), c->scalar);
  HIPCHK(hipGetLastError());
  if (sum_res_sqr) {
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(sum_res_sqr, c->scalar, sizeof(double), hipMemcpyDeviceToHost));
  }
  return MF_OK;
}

int32_t mf_shape_points(mf_ctx* c) { return c ? c->nshape : 0; }

int mf_compute_cl_cd_cm(mf_ctx* c, double* cl, double* cd, double* cm, double* cp) {
  if (!c) return set_err(MF_ERR_INVALID, "null context");
  if (c->nshape == 0) return set_err(MF_ERR_STATE, "cloud has no shape points");
  if (c->ssz == 8) launch_forces<double>(c, c->cur, nullptr, nullptr);
  else launch_forces<float>(c, c->cur, nullptr, nullptr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  std::vector<double> h(3 * (size_t)c->nsid);
  HIPCHK(hipMemcpy(h.data(), c->clcdcm, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int s = 0; s < c->nsid; ++s) {
    if (cl) cl[s] = h[3 * s];
    if (cd) cd[s] = h[3 * s + 1];
    if (cm) cm[s] = h[3 * s + 2];
  }
  if (cp) HIPCHK(hipMemcpy(cp, c->cp, c->nshape * sizeof(double), hipMemcpyDeviceToHost));
  return MF_OK;
}

int mf_copy_probe(int64_t n_bytes, int32_t reps, double* ms_per_copy) {
  if (n_bytes < 16 || reps < 1) return set_err(MF_ERR_INVALID, "mf_copy_probe: bad size");
  const int64_t n = n_bytes / 16;
  D2 *a = nullptr, *b = nullptr;
  HIPCHK(hipMalloc(&a, n * 16));
  HIPCHK(hipMalloc(&b, n * 16));
  HIPCHK(hipMemset(a, 0, n * 16));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  const int grid = 4096;
  hipLaunchKernelGGL(k_copy, dim3(grid), dim3(BLOCK), 0, 0, a, b, n);  // warm
  HIPCHK(hipEventRecord(e0, 0));
  for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(BLOCK), 0, 0, a, b, n);
  HIPCHK(hipEventRecord(e1, 0));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  if (ms_per_copy) *ms_per_copy = ms / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIPCHK(hipFree(a));
  HIPCHK(hipFree(b));
  return MF_OK;
}

int mf_layout(mf_ctx* c, mf_layout_info* info) {
  if (!c || !info) return set_err(MF_ERR_INVALID, "null argument");
  info->n_points = c->n;
  info->n_wall = c->nw;
  info->n_outer = c->no;
  info->n_interior = c->ni;
  info->n_shape = c->nshape;
  info->max_nbhs = c->W;
  info->total_nbhs = c->total_nbhs;
  info->state_bytes = (int64_t)c->ssz;
  return MF_OK;
}

}  // extern "C"

#include "mfcfd_adjoint.hip"
#include "mfcfd_tangent.hip"

static void adj_release(mf_ctx* c) {
  if (!c->adj) return;
  if (c->adj->tape_buf) (void)hipFree(c->adj->tape_buf);
  if (c->adj->e0) (void)hipEventDestroy(c->adj->e0);
  if (c->adj->e1) (void)hipEventDestroy(c->adj->e1);
  delete c->adj;
  c->adj = nullptr;
}
