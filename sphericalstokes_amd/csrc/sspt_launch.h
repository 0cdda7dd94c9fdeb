// sspt_launch.h -- host-side launchers of the kernels (stream-ordered, no sync).
#pragma once
#include "sspt_internal.h"

namespace sspt {

void launch_equal3(ss_ctx *c, const FieldSet &s);
void launch_maxloc(ss_ctx *c, double *SM, const double *eta);
// Spread etaSM in place (ss_params.etasm_spread = m, 1 <= m <= SS_SPREAD_MAX):
// three separable passes r, th, phi of max_{|a|<=m} SM(clamp(i+a)) * 0.1^|a|,
// with tmp (a box of the same layout) as the intermediate.
constexpr int SS_SPREAD_MAX = 8;
void launch_etasm_spread(ss_ctx *c, double *SM, double *tmp, int m);
void launch_copy_bc(ss_ctx *c, double *A, int n1, int n2, int n3);
void launch_compute_P(ss_ctx *c, const FieldSet &s, double _dr, double _dth, double _dph, double dmp, double vpdt,
                      double lr);
void launch_compute_TII(ss_ctx *c, const FieldSet &s);
void launch_power_law(ss_ctx *c, const FieldSet &s, double tau_C, double n_exp, double relax, double emin,
                      double emax);
void launch_compute_V(ss_ctx *c, const FieldSet &s, double _dr, double _dth, double _dph, double g);
void launch_bc_R(ss_ctx *c, const FieldSet &s);
void launch_bc_th(ss_ctx *c, const FieldSet &s, int mode, double r, double sh);
void launch_bc_ph(ss_ctx *c, const FieldSet &s, int mode, double r, double sh);
// Reduces max|V - prev| and max|V| into c->red[0..5]; with write_err the
// difference is stored into prev (check_err! semantics).
void launch_check_err(ss_ctx *c, const FieldSet &s, const double *VRprev, const double *VTprev, const double *VPprev,
                      int write_err);
void launch_fill(ss_ctx *c, double *A, int n1, int n2, int n3, int o1, double v);
void launch_copy_field(ss_ctx *c, double *D, const double *S, int n1, int n2, int n3, int o1);
void launch_init_inclusion(ss_ctx *c, double *A, double r, double radius, double in);
void launch_init_velocity(ss_ctx *c, double *VP, double fact, double r);
void launch_to_box(ss_ctx *c, double *dst, const double *src, int n1, int n2, int n3, int o1, int layout);
void launch_from_box(ss_ctx *c, double *dst, const double *src, int n1, int n2, int n3, int o1, int layout);

// fused PT iteration (sspt_fused.hip); `out` receives the new state.  When
// given, the events bracket the main (r-marching) kernel on c->stream.
void launch_fused_iteration(ss_ctx *c, const FieldSet &in, const FieldSet &out, const ss_params &p, bool last,
                            bool check, hipEvent_t ev_main0 = nullptr, hipEvent_t ev_main1 = nullptr);

} // namespace sspt
