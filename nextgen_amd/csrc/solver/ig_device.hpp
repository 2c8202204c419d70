// Device-side data of the ideal-gas step (Morton cell order).
//
// Per-cell data that neighbours gather is packed in records so that one
// neighbour costs one cache line per record instead of one line per field:
//   light[c][8]  u0 u1 u2 sp_vol p h rho sos        (IdealGasState + rho, sos; per stage)
//   aux[c][16]   mu_total loc_total a_sgs fgr symh_fax symp_fax - dudx[9]
//                (AuxGasProp's lagged coefficients, per step; the velocity
//                gradient, per stage)
// Data a cell only reads for itself (rhoE, geometry, RK registers) stays
// structure-of-arrays so a wavefront reads it coalesced.
#pragma once
#include <cstdint>

namespace cgdev {

constexpr int kSlice = 64;   // cells per SELL slice == one wavefront
constexpr int kBlock = 256;  // threads per block (4 slices)
constexpr int kXcd = 8;      // XCDs a launch is spread over (round-robin by block id)

// light record fields
enum { L_U0 = 0, L_U1, L_U2, L_SPV, L_P, L_H, L_RHO, L_SOS, L_N = 8 };
// aux record fields
enum { A_MU = 0, A_LOC, A_ASGS, A_FGR, A_SYMH, A_SYMP, A_PAD, A_D = 7, A_N = 16 };

// Mesh / operator arrays, all device pointers.
//
// Each cell's half-faces (its extended pairs, in pair order) are split in two
// SELL-64 regions: the compact ones (cf_*, ef_nc[cell] slots) and the purely
// extended ones (xf_*, xf_deg[cell] slots). Slot k of a cell is at
// ptr[slice] + k*64 + lane, so at every k the 64 lanes of a wavefront read one
// contiguous row. The pair geometry is stored per slot (both cells of a pair
// hold a copy), which keeps every geometry read coalesced.
struct Geom {
  int ncv, nfa, nef;
  const double *inv_vol, *vol, *sa_to_vol, *dcmag, *dx13, *dx23;
  // compact half-faces
  const int *cf_ptr, *ef_nc;
  const int* cf_nbr;                                   // neighbour | (cell is the pair's hi side) << 31
  const double *csn0, *csn1, *csn2, *csc0, *csc1, *csc2;  // pair n and c (grouping applied)
  const double *cu0, *cu1, *cu2, *carea, *caod;        // unit normal, area, area/delta (grouping applied)
  const double *cn0, *cn1, *cn2;                       // raw area vector (CFL)
  // purely extended half-faces
  const int *xf_ptr, *xf_deg;
  const int* xf_nbr;                                   // neighbour | hi-side flag << 31
  const double *xsn0, *xsn1, *xsn2, *xsc0, *xsc1, *xsc2;
  // per-block halo for LDS staging (null when staging is off): the sorted
  // cells block b's extended half-faces reach are halo_v[halo_ptr[b] .. halo_ptr[b+1]);
  // xf_loc[slot] = halo position of the slot's neighbour | hi-side flag << 15
  const int *halo_ptr, *halo_v;
  const int16_t* xf_loc;
  int halo_max;
  // gradient stencil without the diagonal, SELL-64 layout
  const int *gr_ptr, *gr_deg, *gr_nbr;
  const double *gc0, *gc1, *gc2;
};

// One buffer of the ping-pong state.
struct State {
  double* light;  // [ncv][8]
  double* rhoE;   // [ncv]
};

struct StepInfo {
  double dt, time;
  int64_t step;
  unsigned long long cfl_bits;  // max convective CFL at dt = 1 (bit pattern, >= 0)
  int nan;
  int pad;
};

struct Work {
  double* aux;       // [ncv][16]
  double* logp;      // [ncv] log(p/p_ref) of the end-of-step state
  // per-step fields (diagnostic copies)
  double *T, *mu_lam, *mu_sgs, *loc_lam, *loc_sgs, *a_sgs, *vv2, *cfl;
  double* rhs0;      // [5][ncv]
  double* rhs1;      // [5][ncv]
  double* racc;      // [5][ncv] compact-face part of the stage RHS
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
int launch_flux_compact(const cgdev::Geom& g, const cgdev::State& in, const cgdev::Work& w, void* stream);
int launch_flux_ext(int stage, int mode, const cgdev::Geom& g, const cgdev::State& in, const cgdev::State& out,
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
// gather a light / aux field into a SoA array
int launch_unpack(const double* rec, int stride, int field, int64_t n, double* out, void* stream);
}  // namespace cgk
