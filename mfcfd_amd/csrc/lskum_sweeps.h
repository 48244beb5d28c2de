// Resident inner sweeps of the mixed path: all inner iterations of one stage
// (q_variables.F90:115-227, eval_q_inner_loop + eval_update_innerloop, Jacobi) in one
// launch.  One block per 64-point block of the staged walk, every block resident at
// once, and a grid barrier between sweeps.  A block stages its halo's ids (registers),
// coordinates and slot words (LDS) and its points' first-pass sums and solve factors
// (registers) once; each sweep then reloads only the halo's dq.  The per-sweep kernel
// (k_inner, staged) restages everything and pays a launch per sweep.
//
// The sums, their order and the solve are those of the staged k_inner with the same
// lanes per point (the same edge function, the same slot order per lane), so the two
// give bit-identical dq.
//
// The wait at the barrier is bounded: after spin_limit polls, or once another block has
// given up, a block sets the barrier's abort flag and Ctl::halt = HALT_BARRIER and
// returns, and every later kernel of the run returns at once.  The host checks halt
// after the run, reports an error and stops using this kernel.  The launcher is only
// used when the occupancy calculator says every block fits on the chip at once.
#pragma once

#include "lskum_kernels.h"

// halo ids a thread keeps: a 64-point block's halo in LDS has at most 248 entries
// (byte slot words)
template <int LPP>
constexpr int sweep_ids() {
  return (248 + STAGE_P * LPP - 1) / (STAGE_P * LPP);
}

// The dq the sweeps exchange between blocks moves by agent-scope coherent accesses
// (relaxed atomics, 8 bytes each): a store is written through to the device's coherence
// point and a load reads from it, so no cache of another XCD can hold a stale or an
// unwritten copy.  The ordering the barrier needs is then only that of completion: a
// block's stores complete (s_waitcnt) before it arrives, and the halo loads of the next
// sweep are issued after the wait has seen every arrival.  Agent-scope release and
// acquire fences would instead write back and invalidate the XCD's whole L2 once per
// block and barrier, which costs more than the sweep itself at 2,500 blocks.
struct F2 {
  float x, y;
};

__device__ inline V4<float> load_coherent(const V4<float>* p) {
  uint64_t* q = reinterpret_cast<uint64_t*>(const_cast<V4<float>*>(p));
  const F2 a = __builtin_bit_cast(F2, __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  const F2 b = __builtin_bit_cast(F2, __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  return V4<float>{a.x, a.y, b.x, b.y};
}

__device__ inline void store_coherent(V4<float>* p, const V4<float>& v) {
  uint64_t* q = reinterpret_cast<uint64_t*>(p);
  __hip_atomic_store(q, __builtin_bit_cast(uint64_t, F2{v.x, v.y}), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(q + 1, __builtin_bit_cast(uint64_t, F2{v.z, v.w}), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

// Arrive at the grid barrier and wait for every block of the launch.  The caller's
// stores have completed (s_waitcnt after them); the block barrier orders them before
// thread 0's arrival, and thread 0's observed generation before the block's next loads.
// Returns false in every thread of the block when the wait gave up.
__device__ inline bool sweep_barrier(SweepBar* bar, uint32_t& gen, uint32_t spin_limit, Ctl* ctl, int* s_flag) {
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t nblk = gridDim.x;
    const uint32_t ng = nblk < (uint32_t)SWEEP_GROUPS ? nblk : (uint32_t)SWEEP_GROUPS;
    const uint32_t grp = blockIdx.x % ng;
    const uint32_t members = nblk / ng + (grp < nblk % ng ? 1u : 0u);
    uint32_t* gc = &bar->group[grp * 32];
    if (__hip_atomic_fetch_add(gc, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == members - 1) {
      __hip_atomic_store(gc, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __builtin_amdgcn_s_waitcnt(0);
      if (__hip_atomic_fetch_add(&bar->top, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == ng - 1) {
        __hip_atomic_store(&bar->top, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_s_waitcnt(0);
        __hip_atomic_fetch_add(&bar->gen, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    int ok = 1;
    uint32_t polls = 0;
    while (__hip_atomic_load(&bar->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
      if (__hip_atomic_load(&bar->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u || ++polls > spin_limit) {
        __hip_atomic_store(&bar->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&ctl->halt, HALT_BARRIER, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    gen += 1;
    *s_flag = ok;
  }
  __syncthreads();
  return *s_flag != 0;
}

// `sweeps` inner sweeps: sweep s reads dq0 (s even) or dq1 and writes the other, so the
// last one leaves dq in dq0 when `sweeps` is even, as the per-sweep launches do.
template <bool POWER0, int LPP>
__global__ __launch_bounds__(STAGE_P * LPP) void k_sweeps(Geom g, V4<float>* dq0, V4<float>* dq1, int sweeps,
                                                          double power, Ctl* __restrict__ ctl,
                                                          const V4<float>* __restrict__ qsum,
                                                          SweepBar* __restrict__ bar, uint32_t spin_limit) {
  if (__hip_atomic_load(&ctl->halt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  __shared__ int s_flag;
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int P = STAGE_P, NT = STAGE_P * LPP, NID = sweep_ids<LPP>();
  const Halo& h = g.hl[0];
  const int H = h.hcap;
  V4<float>* s_dx = (V4<float>*)smem;
  V4<float>* s_dy = s_dx + H;
  D2* s_xy = (D2*)(s_dy + H);
  uint8_t* s_cl = (uint8_t*)(s_xy + H);
  // per point: first-pass sums x, y and solve factors; with one lane per point in
  // registers (the block is one wave, so registers do not limit residency, LDS does)
  constexpr bool TAIL_LDS = LPP > 1;
  V4<float>* s_tail = (V4<float>*)(s_cl + g.width * P);

  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int li = (int)threadIdx.x / LPP, lane = (int)threadIdx.x % LPP;
  const bool live = b * P + li < g.n;
  const int i = live ? b * P + li : g.n - 1;
  const int nb = live ? g.nbhs[i] : 0;
  uint32_t gen = 0;
  if (threadIdx.x == 0) gen = __hip_atomic_load(&bar->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);

  // for the whole launch: the halo ids (registers), coordinates and slot words, and the
  // block's first-pass sums and solve factors (LDS)
  V4<float> sa{}, sb{}, f{};
  if constexpr (TAIL_LDS) {
    for (int j = threadIdx.x; j < 3 * P; j += NT) {
      const int p = j / 3, k = b * P + p < g.n ? b * P + p : g.n - 1, c = j - 3 * p;
      s_tail[j] = c < 2 ? qsum[2 * k + c] : g.lsq_geo_f[k];
    }
  } else {
    sa = qsum[2 * i];
    sb = qsum[2 * i + 1];
    f = g.lsq_geo_f[i];
  }
  const int h0 = h.ptr[b], hn = h.ptr[b + 1] - h0;
  const int hl = hn < H ? hn : H;
  int id[NID];
#pragma unroll
  for (int m = 0; m < NID; ++m) {
    const int j = (int)threadIdx.x + m * NT;
    id[m] = j < hl ? h.ids[h0 + j] : i;  // past the halo: a valid point, loaded and not kept
    if (j < hl) s_xy[j] = g.xy[id[m]];
  }
  constexpr int V = P / 16;
  const uint4* wsrc = (const uint4*)(h.loc8 + (size_t)b * P);
  for (int j = threadIdx.x; j < g.width * V; j += NT) {
    const int s = j / V;
    ((uint4*)s_cl)[j] = wsrc[(size_t)s * (h.stride / 16) + (j - s * V)];
  }

  for (int sw = 0; sw < sweeps; ++sw) {
    const V4<float>* src = (sw & 1) ? dq1 : dq0;
    V4<float>* dst = (sw & 1) ? dq0 : dq1;
    if (sw > 0 && !sweep_barrier(bar, gen, spin_limit, ctl, &s_flag)) return;
    // every load issued before the first LDS write (coherent loads keep their order)
    V4<float> hdx[NID], hdy[NID];
#pragma unroll
    for (int m = 0; m < NID; ++m) {
      hdx[m] = load_coherent(src + 2 * id[m]);
      hdy[m] = load_coherent(src + 2 * id[m] + 1);
    }
#pragma unroll
    for (int m = 0; m < NID; ++m) {
      const int j = (int)threadIdx.x + m * NT;
      if (j < hl) {
        s_dx[j] = hdx[m];
        s_dy[j] = hdy[m];
      }
    }
    __syncthreads();
    // own points lead the halo
    const D2 pi = s_xy[li];
    const V4<float> dxi = s_dx[li], dyi = s_dy[li];
    double sxq[4] = {0.0, 0.0, 0.0, 0.0}, syq[4] = {0.0, 0.0, 0.0, 0.0};
    if (hn <= H) {
#pragma unroll 1
      for (int s = lane; s < nb; s += LPP) {
        const int j = s_cl[s * P + li];
        inner_edge_sums<float, POWER0>(pi, dxi, dyi, s_xy[j], s_dx[j], s_dy[j], power, true, sxq, syq);
      }
    } else {
#pragma unroll 1
      for (int s = lane; s < nb; s += LPP) {
        const int j = s_cl[s * P + li];
        if (j < H) {
          inner_edge_sums<float, POWER0>(pi, dxi, dyi, s_xy[j], s_dx[j], s_dy[j], power, true, sxq, syq);
        } else {
          const int k = (int)(g.conn[(size_t)s * g.n + i] >> 4);
          inner_edge_sums<float, POWER0>(pi, dxi, dyi, g.xy[k], load_coherent(src + 2 * k),
                                         load_coherent(src + 2 * k + 1), power, true, sxq, syq);
        }
      }
    }
    if (LPP > 1) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        sxq[r] = group_sum<LPP>(sxq[r]);
        syq[r] = group_sum<LPP>(syq[r]);
      }
    }
    if (live && lane == 0) {
      if constexpr (TAIL_LDS) {
        sa = s_tail[3 * li];
        sb = s_tail[3 * li + 1];
        f = s_tail[3 * li + 2];
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        sxq[r] = (double)sa[r] + sxq[r];
        syq[r] = (double)sb[r] + syq[r];
      }
      V4<float> ox, oy;
      lsq_apply_f(f, sxq, syq, ox, oy);
      store_coherent(dst + 2 * i, ox);
      store_coherent(dst + 2 * i + 1, oy);
      __builtin_amdgcn_s_waitcnt(0);  // complete before the block arrives at the barrier
    }
  }
}

// Blocks of k_sweeps<POWER0, lpp> that fit on the device at once (0: not launchable)
inline int sweeps_capacity(int lpp, bool power0, size_t lds) {
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  const void* fn = nullptr;
  switch (lpp) {
    case 1: fn = power0 ? (const void*)k_sweeps<true, 1> : (const void*)k_sweeps<false, 1>; break;
    case 2: fn = power0 ? (const void*)k_sweeps<true, 2> : (const void*)k_sweeps<false, 2>; break;
    case 4: fn = power0 ? (const void*)k_sweeps<true, 4> : (const void*)k_sweeps<false, 4>; break;
    default: return 0;
  }
  // LDS is allocated in chunks: count it rounded up to 2 KiB
  const size_t lds_alloc = (lds + 2047) / 2048 * 2048;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, STAGE_P * lpp, lds_alloc) != hipSuccess) return 0;
  return per * cus;
}

inline void launch_sweeps_env(const LaunchEnv& e, V4<float>* dq0, V4<float>* dq1, int sweeps, Ctl* ctl,
                              const V4<float>* qsum, SweepBar* bar, uint32_t spin_limit, int lpp) {
  const dim3 gr = stage_grid(e.n);
  const size_t lds = sweeps_lds(e.g.hl[0].hcap, e.g.width, lpp);
  with_lpp_staged(lpp, [&]<int L>() {
    const dim3 bl(STAGE_P * L);
    if (e.power0)
      hipLaunchKernelGGL((k_sweeps<true, L>), gr, bl, lds, e.stream, e.g, dq0, dq1, sweeps, e.power, ctl, qsum, bar,
                         spin_limit);
    else
      hipLaunchKernelGGL((k_sweeps<false, L>), gr, bl, lds, e.stream, e.g, dq0, dq1, sweeps, e.power, ctl, qsum,
                         bar, spin_limit);
  });
}
