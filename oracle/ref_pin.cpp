// Test-only harness: exposes the reference's own header code to the oracle
// tests with a plain C interface. Compiled from the reference sources where
// they lie (see Makefile); nothing here is part of the product.
//
//   calcInternalFluxNew / calcInternalFluxSymmetricNew /
//   calcInternalCompactFluxNewGrad   src/charles/ideal_gas/IdealGasSolverFlux.hpp:9-215
//   calcRiemannFlux                  src/charles/ideal_gas/IdealGasSolverFlux.hpp:218-306
//   calcNscbcIgOutletRhs             src/charles/BcCommon.hpp:419-470
//   calcNscbcIgSubsonicInletRhs      src/charles/BcCommon.hpp:311-337
//   calcSvd33                        src/core/common/SvdNew.hpp:24
//   MiscUtils::fast_pow_posx         src/core/common/MiscUtils.hpp:273-276
//   RkWeights::erk_wgt3              src/charles/RkWeights.hpp:6-8
//   AlgebraicWM::solve_tau           src/charles/wm/AlgebraicWM.hpp:40-113
#include "FlowSolver.hpp"
#include "BcCommon.hpp"
#include "IdealGasSolverFlux.hpp"
#include "RkWeights.hpp"
#include "SvdNew.hpp"
#include "AlgebraicWM.hpp"
#include "EquilibriumWM.hpp"

#include <cstring>

static IdealGasState state_of(const double* v) {
  IdealGasState s;
  for (int i = 0; i < 3; ++i) s.u[i] = v[i];
  s.sp_vol = v[3];
  s.p = v[4];
  s.h = v[5];
  return s;
}

static AuxGasProp aux_of(const double* v) {
  AuxGasProp a;
  a.mu_total = v[0]; a.loc_total = v[1]; a.half_usq = v[2]; a.rho = v[3]; a.a_sgs = v[4];
  a.sos = v[5]; a.fgr = v[6]; a.symh_fax = v[7]; a.symp_fax = v[8];
  return a;
}

static void rhs_out(const IdealGasRhs& r, double* o) {
  o[0] = r.rho;
  for (int i = 0; i < 3; ++i) o[1 + i] = r.rhou[i];
  o[4] = r.rhoE;
}

extern "C" {

void ref_flux_riemann(double* flux, const double* n_bf, const double* cv0, const double* cv1, double gammaL,
                      double gammaR) {
  IdealGasState a = state_of(cv0), b = state_of(cv1);
  IdealGasRhs f;
  calcRiemannFlux(f, n_bf, a, b, gammaL, gammaR);
  rhs_out(f, flux);
}

// rhs[5] = d/dt (rho_bc, p_bc, u_bc[3]) of an NSCBC outlet face; bc = rho, p,
// u[3] of the face; cv = rho, p, u[3] of its cell; returns un_bc
double ref_nscbc_outlet_rhs(double* rhs, const double* unit_n, double delta, const double* bc, const double* cv,
                            double gamma, double L_ref, double sigma, double p_ref) {
  return calcNscbcIgOutletRhs(rhs, unit_n, delta, bc[0], bc + 2, bc[1], cv[0], cv + 2, cv[1], gamma, L_ref, sigma,
                              p_ref);
}

// the same for an NSCBC_OUTLET_MDOT face (u_avg is not used by the rates)
double ref_nscbc_mdot_rhs(double* rhs, const double* unit_n, double delta, const double* bc, const double* cv,
                          double gamma, double rhou_avg, double rhou_target, double T_relax) {
  return calcNscbcIgOutletMdotWeakRhs(rhs, unit_n, delta, bc[0], bc + 2, bc[1], cv[0], cv + 2, cv[1], gamma, rhou_avg,
                                      0.0, rhou_target, T_relax);
}

// aux_out = u_avg[3], inv_v_avg, h_stag, p_avg
void ref_flux_skew(double* flux, const double* nc, const double* cv0, const double* cv1, double* aux_out) {
  FaNC fa;
  for (int i = 0; i < 3; ++i) { fa.n[i] = nc[i]; fa.c[i] = nc[3 + i]; }
  IdealGasState a = state_of(cv0), b = state_of(cv1);
  IdealGasRhs f;
  double u_avg[3], inv_v_avg, h_stag, p_avg;
  calcInternalFluxNew(f, fa, a, b, u_avg, inv_v_avg, h_stag, p_avg);
  rhs_out(f, flux);
  for (int i = 0; i < 3; ++i) aux_out[i] = u_avg[i];
  aux_out[3] = inv_v_avg;
  aux_out[4] = h_stag;
  aux_out[5] = p_avg;
}

void ref_flux_symmetric(double* flux, const double* nc, const double* u_avg_in, double inv_v_avg, double h_stag,
                        double p_avg, const double* cv0, const double* cv1) {
  FaNC fa;
  for (int i = 0; i < 3; ++i) { fa.n[i] = nc[i]; fa.c[i] = nc[3 + i]; }
  IdealGasState a = state_of(cv0), b = state_of(cv1);
  double u_avg[3] = {u_avg_in[0], u_avg_in[1], u_avg_in[2]};
  IdealGasRhs f;
  calcInternalFluxSymmetricNew(f, fa, u_avg, inv_v_avg, h_stag, p_avg, a, b);
  rhs_out(f, flux);
}

// flux is in/out (the compact terms are subtracted from it)
void ref_flux_compact(double* flux, const double* fc, const double* cv0, const double* cv1, const double* aux0,
                      const double* aux1, const double* u_avg, double inv_v_avg, double h_stag, const double* g0,
                      const double* g1) {
  FaCompact fa;
  for (int i = 0; i < 3; ++i) fa.unit_n[i] = fc[i];
  fa.area_over_delta = fc[3];
  fa.area = fc[4];
  IdealGasState a = state_of(cv0), b = state_of(cv1);
  AuxGasProp A0 = aux_of(aux0), A1 = aux_of(aux1);
  double G0[3][3], G1[3][3];
  std::memcpy(G0, g0, sizeof(G0));
  std::memcpy(G1, g1, sizeof(G1));
  IdealGasRhs f;
  f.rho = flux[0];
  for (int i = 0; i < 3; ++i) f.rhou[i] = flux[1 + i];
  f.rhoE = flux[4];
  calcInternalCompactFluxNewGrad(f, fa, a, b, A0, A1, u_avg, inv_v_avg, h_stag, G0, G1);
  rhs_out(f, flux);
}

double ref_fast_pow_posx(double x, double p) { return MiscUtils::fast_pow_posx(x, p); }

int ref_svd33(const double* A, double* U, double* V, double* sigma, double* max_err) {
  return calcSvd33(U, V, sigma, A, *max_err);
}

void ref_erk_wgt3(double* w) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) w[3 * i + j] = RkWeights::erk_wgt3[i][j];
}

// rhs[0] of an NSCBC mass-flow inlet face; bc, cv = rho, p, u[3]
int ref_nscbc_inlet_rhs(double* rhs, const double* unit_n, double delta, const double* bc, const double* cv,
                        double gamma) {
  return calcNscbcIgSubsonicInletRhs(rhs, unit_n, delta, bc[0], bc + 2, bc[1], cv[0], cv + 2, cv[1], gamma);
}

// wall stress of the algebraic wall model
double ref_wm_solve_tau(double u1, double y1, double rho, double mu, double u_tau_guess) {
  return AlgebraicWM::solve_tau(u1, y1, rho, mu, u_tau_guess);
}

int ref_wm_eq_solve_qt(double* tau_w, double* q_w, double u1, double T1, double rho1, double y1, double mu1,
                       double Pr_lam, double cp, double* T_w, double relax, double y_plus0, double tol, int use_guess) {
  return EquilibriumWM::solve_qt(*tau_w, *q_w, u1, T1, rho1, y1, mu1, Pr_lam, cp, *T_w, relax, y_plus0, tol,
                                 use_guess != 0);
}

double ref_wm_q_wall(double T1, double Tw, double tau_w, double rho, double mu, double locp, double cp, double y1) {
  return AlgebraicWM::compute_q_wall_approx(T1, Tw, tau_w, rho, mu, locp, cp, y1);
}

}  // extern "C"
