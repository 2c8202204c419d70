// Device-side data of the ideal-gas step (structure-of-arrays, Morton cell order).
#pragma once
#include <cstdint>

namespace cgdev {

constexpr int kSlice = 64;   // cells per SELL slice == one wavefront
constexpr int kBlock = 256;  // threads per block (4 slices)

// Mesh / operator arrays, all device pointers.
struct Geom {
  int ncv, nfa, nef;
  const double *inv_vol, *vol, *sa_to_vol, *dcmag, *dx13, *dx23;
  // extended half-faces in SELL-64 layout: slot(k, lane) = ef_ptr[slice] + k*64 + lane,
  // k < ef_deg[cell]; ef_code = pair index | (cell is the pair's hi side) << 31
  const int *ef_ptr, *ef_deg, *ef_nbr, *ef_code;
  // pair geometry used by the flux (grouping applied)
  const double *pn0, *pn1, *pn2, *pc0, *pc1, *pc2;
  // compact geometry of the first nfa pairs (grouping applied) and raw normals
  const double *un0, *un1, *un2, *area, *aod;
  const double *nf0, *nf1, *nf2;
  // gradient stencil without the diagonal, SELL-64 layout
  const int *gr_ptr, *gr_deg, *gr_nbr;
  const double *gc0, *gc1, *gc2;
};

// Conserved state and the per-stage primitives derived from it.
struct State {
  double *rho, *u0, *u1, *u2, *rhoE;
  double *spv, *p, *h, *sos;
};

struct StepInfo {
  double dt, time;
  int64_t step;
  unsigned long long cfl_bits;  // max convective CFL at dt = 1 (bit pattern, >= 0)
  int nan;
  int pad;
};

struct Work {
  double* dudx;      // [9][ncv], component i*3+j = du_i/dx_j
  // lagged coefficients of AuxGasProp, refreshed once per step
  double *mu_total, *loc_total, *a_sgs_c, *fgr, *symh, *symp;
  // per-step fields
  double *T, *logp, *invRT;
  double *mu_lam, *mu_sgs, *loc_lam, *loc_sgs, *a_sgs, *vv2, *cfl;
  double* rhs0;      // [5][ncv]
  double* rhs1;      // [5][ncv]
  StepInfo* info;
  double* partial;   // reduction scratch [nblocks][8]
  double* diag;      // [8]
};

struct Params {
  double gamma, R_gas, gm1, Rgogm1, p_ref, inv_pref, inv_Tref, mu_ref, mu_power_law, invPr_lam;
  double vreman_coeff, a_sgs_coeff, a_sgs_coeff2, lambda_dil_coeff, invPr_t, invgm1;
  double coeff_p, coeff_h, gr_nrm;
  int sgs_model, diss_mode, limit_asgs, dt_mode;
  double dt_data0, dt_data1;
};

}  // namespace cgdev

// launchers (ig_kernels.hip)
namespace cgk {
int launch_flux(int stage, int mode, const cgdev::Geom& g, const cgdev::State& in, const cgdev::State& out,
                const cgdev::Work& w, const cgdev::Params& P, double* rhs_out, void* stream);
int launch_grad(const cgdev::Geom& g, const cgdev::State& st, const cgdev::Work& w, void* stream);
int launch_eos(const cgdev::Geom& g, const cgdev::State& st, const cgdev::Work& w, const cgdev::Params& P,
               void* stream);
int launch_props(const cgdev::Geom& g, const cgdev::State& st, const cgdev::Work& w, const cgdev::Params& P,
                 int fields, void* stream);
int launch_begin_step(const cgdev::Work& w, const cgdev::Params& P, void* stream);
int launch_cfl(const cgdev::Geom& g, const cgdev::State& st, const cgdev::Work& w, double dt, void* stream);
int launch_diag(const cgdev::Geom& g, const cgdev::State& st, const cgdev::Work& w, int nblocks, void* stream);
int launch_fill(double* p, double v, int64_t n, void* stream);
}  // namespace cgk
