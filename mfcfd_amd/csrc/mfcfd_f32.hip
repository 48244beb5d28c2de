// The mixed path (FP32 state, FP64 sums) of the LSKUM stencil kernels: this unit
// instantiates the launchers of lskum_kernels.h for S = float (the LSQ kernel's in
// mfcfd_lsq32.hip) and is compiled with FMA contraction (the FP64 path, in mfcfd.hip,
// is not).
#include "lskum_kernels.h"
#include "lskum_sweeps.h"

void launch_inner_f32(const LaunchEnv& e, const V4f* q, const V4f* a, V4f* b, const Ctl* ctl, int lpp,
                      const V4f* qsum) {
  launch_inner_env<float>(e, q, a, b, ctl, lpp, qsum);
}

void launch_sweeps_f32(const LaunchEnv& e, V4f* dq0, V4f* dq1, int sweeps, Ctl* ctl, const V4f* qsum, SweepBar* bar,
                       uint32_t spin_limit, int lpp) {
  launch_sweeps_env(e, dq0, dq1, sweeps, ctl, qsum, bar, spin_limit, lpp);
}

int sweeps_capacity_f32(int lpp, bool power0, size_t lds) { return sweeps_capacity(lpp, power0, lds); }

void launch_flux_update_f32(const LaunchEnv& e, const State<float>& st, const V4f* prim, V4f* pold, V4f* pout,
                            V4f* qout, double* cspeed, const double* delta, double* partial, const Ctl* ctl,
                            int rk, bool last, int lpp, double* rs_pt) {
  launch_flux_update_env<float>(e, st, prim, pold, pout, qout, cspeed, delta, partial, ctl, rk, last, lpp, rs_pt);
}

void launch_forces_f32(const LaunchEnv& e, const V4f* prim, const ForcesArgs& f, double* cd_hist, const Ctl* ctl) {
  launch_forces_env<float>(e, prim, f, cd_hist, ctl);
}

void launch_q_variables_f32(const LaunchEnv& e, const V4f* prim, V4f* q, double* cspeed) {
  launch_q_variables_env<float>(e, prim, q, cspeed);
}

void launch_func_delta_f32(const LaunchEnv& e, const V4f* prim, double* delta) {
  launch_func_delta_env<float>(e, prim, delta);
}

void launch_flux_f32(const LaunchEnv& e, const State<float>& st, const double* delta, V4d* flux_res) {
  launch_flux_env<float>(e, st, delta, flux_res);
}

void launch_update_f32(const LaunchEnv& e, V4f* prim, const V4f* pold, const V4d* flux_res, double* partial,
                       int rk) {
  launch_update_env<float>(e, prim, pold, flux_res, partial, rk);
}
