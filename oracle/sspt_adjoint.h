/*
 * sspt_adjoint.h -- CPU reference of the adjoint PT iteration.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP adjoint (libsspt.so).
 * Only tests/ may load it.  Arrays use the oracle's layout (sspt_oracle.h).
 */
#ifndef SSPT_ADJOINT_H
#define SSPT_ADJOINT_H
#include "sspt_oracle.h"
#ifdef __cplusplus
extern "C" {
#endif

/* adjoint state, one array per state field (shapes of the forward fields) */
enum { ORC_A_VR, ORC_A_VT, ORC_A_VP, ORC_A_P, ORC_A_RR, ORC_A_TT, ORC_A_PP, ORC_A_RT, ORC_A_RP, ORC_A_TP,
       ORC_A_ETA, ORC_A_EPL, ORC_A_N };
typedef struct orc_adj { double *a[ORC_A_N]; } orc_adj;

size_t orc_adj_len(int nr, int nth, int nph, int which);
/* lam <- T^T lam at the fixed point held in s (s->etaSM, s->DT_R, s->tII as
 * the last forward iteration left them); g_rho / g_ini (centre shape, may be
 * NULL) accumulate (dPhi/drho)^T lam and (dPhi/deta_ini)^T lam. */
void orc_adj_step(const orc_state *s, const orc_params *p, orc_adj *lam, double *g_rho, double *g_ini);
/* J and dJ/dVR (VR shape, may be NULL) */
double orc_adj_objective(const orc_state *s, double r, double radius, double *dJ);
/* Iterates lam = T^T lam + dJ/dx (lam holds the start value, usually 0) and
 * then adds the gradients to g_rho, g_ini.  Returns the iterations done,
 * negative if tol was not reached. */
int orc_adj_solve(const orc_state *s, const orc_params *p, double r, double radius, orc_adj *lam, int maxiter,
                  int ncheck, double tol, double *g_rho, double *g_ini);

#ifdef __cplusplus
}
#endif
#endif
