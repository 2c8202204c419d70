// Device-side data of the ideal-gas step (Morton cell order).
//
// The step runs in one of two working precisions R:
//   double  FP64 throughout (the reference arithmetic);
//   float   mixed: face geometry, neighbour records, fluxes and their
//           accumulation, gradients, lagged coefficients and RK registers in
//           FP32; the conserved state rho, u, rhoE, the RK3 state update, the
//           recovery of p, T from rhoE and the conservation sums in FP64.
//
// Per-cell data that neighbours gather is packed in records so that one
// neighbour costs one cache line per record instead of one line per field:
//   light[c][8]  u0 u1 u2 sp_vol p h rho sos        (IdealGasState + rho, sos; per stage)
//   aux[c][16]   mu_total loc_total a_sgs fgr symh_fax symp_fax - dudx[9]
//                (AuxGasProp's lagged coefficients, per step; the velocity
//                gradient, per stage)
// Data a cell only reads for itself (rhoE, geometry, RK registers) stays
// structure-of-arrays so a wavefront reads it coalesced.
//
// In FP64 the light record is also the state (rho and u live there, rhoE
// beside it). In mixed precision the state is the FP64 array cons[5][ncv]
// (rho, u0, u1, u2, rhoE) and the FP32 light record is derived from it, with
// p and h stored relative to the constant gauges p_g and h_g:
//   light.p = p - p_g,  light.h = h - h_g.
// Differences of p and h are unchanged by the gauge. Levels enter the flux in
// two places: p_avg n in the momentum flux, whose constant part p_g n sums to
// zero over a cell's half-faces (the operator's discrete GCL), so it is
// dropped; and h in the energy flux, which is carried as
//   F_rhoE' = F_rhoE - h_g F_rho
// through the accumulation and restored in FP64 in the RK update
// (rhs_rhoE = rhs_rhoE' + h_g rhs_rho). The p v terms of the compact closure
// use the absolute pressure (see flux_compact).
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

// Mesh / operator arrays, all device pointers; geometry in the working precision R.
//
// The compact pairs (the Voronoi faces, pair indices 0..nfa-1) are evaluated
// once per pair by their own kernel (thread per pair, per-pair SoA geometry),
// which writes each pair's flux to w.cflux. Each cell's half-faces (its
// extended pairs, in pair order) are split in two SELL-64 regions: the compact
// ones (cf_*, ef_nc[pos] slots: the pair index, to gather cflux) and the
// purely extended ones (xf_*, xf_deg[pos] slots). Slot k at lane position pos
// is at ptr[pos / 64] + k*64 + pos % 64, so at every k the 64 lanes of a
// wavefront read one contiguous row. Position pos holds cell
// (pos & ~63) | xperm[pos]: within each 64-cell slice the lanes take the cells
// in decreasing order of their row counts (SELL-C-sigma with C = sigma = 64),
// so the lanes still active in a slice's tail rows are contiguous and cache
// lines of finished lanes are not fetched; a permutation within the slice
// touches the same cache lines for every per-cell array. The gradient region
// has its own order (gperm). The extended pair geometry is stored per slot
// (both cells of a pair hold a copy), which keeps every geometry read coalesced.
template <class R>
struct GeomT {
  int ncv, nfa, nef;
  const double *inv_vol, *vol, *sa_to_vol, *dcmag, *dx13, *dx23;
  // compact pairs (thread per pair)
  const int *plo, *phi;                           // the pair's cells (lo < hi in pair orientation)
  const R *pn0, *pn1, *pn2, *pc0, *pc1, *pc2;     // pair n and c (grouping applied)
  const R *pu0, *pu1, *pu2, *parea, *paod;        // unit normal, area, area/delta (grouping applied)
  // lane order of the half-face regions: position -> cell offset in its slice
  const uint8_t* xperm;
  // compact half-faces
  const int *cf_ptr, *ef_nc;
  const int* cf_nbr;                              // neighbour | (cell is the pair's hi side) << 31
  const int* cf_pair;                             // pair index | (cell is the pair's hi side) << 31
  const R *cfn0, *cfn1, *cfn2;                    // the pair's raw area vector n_fa (CFL), per slot
  // purely extended half-faces
  const int *xf_ptr, *xf_deg;
  const int* xf_nbr;                              // neighbour | hi-side flag << 31
  const R *xsn0, *xsn1, *xsn2, *xsc0, *xsc1, *xsc2;
  // per-block halo for LDS staging (null when staging is off): the sorted
  // cells block b's extended half-faces reach are halo_v[halo_ptr[b] .. halo_ptr[b+1]);
  // xf_loc[slot] = halo position of the slot's neighbour | hi-side flag << 15
  const int *halo_ptr, *halo_v;
  const int16_t* xf_loc;
  int halo_max;
  // gradient stencil without the diagonal, SELL-64 layout in its own lane
  // order (gperm, as xperm)
  const uint8_t* gperm;
  const int *gr_ptr, *gr_deg, *gr_nbr;
  const R *gc0, *gc1, *gc2;
  // DSM_LOCAL: the test-filter weight of each compact slot's face (FP64,
  // |n_fa| |x_vv1 - x_vv0| / 6) and each cell's Voronoi radius r_vv
  const double* cfw;
  const double* r_vv;
  // boundary faces (null / 0 on a periodic mesh): cell c's faces are
  // bf_ptr[c] .. bf_ptr[c+1] in zone-major order; per face its zone, outward
  // area vector and area_over_delta_bf (FP64 in either precision)
  int nbf;
  const int *bf_ptr, *bf_zone;
  const int* bf_cell;  // each face's cell (device order)
  const double *bfn0, *bfn1, *bfn2, *bf_aod;
};

// One buffer of the ping-pong state.
template <class R>
struct StateT {
  R* light;      // [ncv][8]
  double* rhoE;  // [ncv]; in mixed precision it aliases cons + 4*ncv
  double* cons;  // mixed precision: [5][ncv] rho u0 u1 u2 rhoE (null in FP64)
  double* sc;    // passive scalars [nsc][ncv] (FP64; null without scalars)
};

struct StepInfo {
  double dt, time;
  int64_t step;
  unsigned long long cfl_bits;  // max convective CFL at dt = 1 (bit pattern, >= 0)
  int nan;
  int pad;
};

template <class R>
struct WorkT {
  R* aux;            // [ncv][16]
  double* logp;      // [ncv] log(p/p_ref) of the end-of-step state (FP64 sensor)
  // per-step fields (diagnostic copies)
  double *T, *mu_lam, *mu_sgs, *loc_lam, *loc_sgs, *a_sgs, *vv2, *cfl;
  R* rhs0;           // [5][ncv]
  R* rhs1;           // [5][ncv]
  R* cflux;          // [nfa][6] flux of each compact pair (lo -> hi), this stage
  StepInfo* info;
  double* partial;   // reduction scratch [nblocks][8]
  double* diag;      // [8]
  // DSM_LOCAL only (null otherwise): the filtered fields [ncv][16] (rho, u, the
  // model tensor smag rho S and rho u u, deviatoric-symmetric storage) and the
  // time-relaxed L:M and M:M [ncv] carried from step to step
  double* dsm;
  double *lijmij, *mijmij;
  // passive scalars: RK3 registers of the scalar RHS [nsc][ncv] (null without scalars)
  double *rhs_sc0, *rhs_sc1;
  // NSCBC outlet faces (null without them): per boundary face (device order)
  // its state rho, p, u[3] and the rates of the step's three stages [3][5]
  double *nscbc, *nscbc_rhs;
  // zones whose condition needs sums over the zone (null without them), per
  // zone [16][3], set by k_zone_sums before a stage: NSCBC_OUTLET_MDOT the sums
  // over its faces' states of |n_bf|, rho_bc u_bc . n_bf, u_bc . n_bf; CBC_TOTAL_PT
  // USE_AVG over its cells of rho |n_bf|, rho u . n_bf (at init and stage 1);
  // NSCBC_MT T_TOTAL of |n_bf| Ma^2 and |n_bf|, Ma = (rhoun_bc / rho) / c
  double* zsum;
  // SPONGE (null without one): per cell the relaxed state T, u[3] [ncv][4] and
  // the volume coefficient strength x_sp^2 (negative outside the sponge) [ncv]
  double *sponge, *sponge_coeff;
  // WM_ALG_ADIABATIC, WM_EQ_ADIABATIC and WM_ALG_ISOTHERMAL faces (null without
  // them): per boundary face its wall stress tau_wall (FP64, in the snapshots;
  // with a WM_EQ_ADIABATIC zone followed by each face's adiabatic wall
  // temperature T_wall, [2][nbf], both snapshotted as WmAdiabatic::storeData) and, with
  // an isothermal zone, its wall heat flux q_wall (FP64, not in the snapshots,
  // as in the reference: each step's stage 1 sets it before it is used); with a
  // WM_EQ_ISOTHERMAL zone wm_q is [2][nbf], its faces' stress in the second half
  // (WmIsothermal::storeData stores nothing for that model)
  double *wm_tau, *wm_q;
  // CBC_TOTAL_PT faces (null without them): per boundary face its inflow
  // normal velocity un_bc (FP64; not in the snapshots, as in the reference)
  double* tpt_un;
  // passive scalars on a mesh with boundary faces (null otherwise): each
  // boundary face's mass flux of the stage (FP64, the Riemann flux's F_rho on
  // CBC, NSCBC and sponge faces), written by the flux pass for the scalar pass
  double* mf_bf;
};

constexpr int kMaxBcZones = 16;
constexpr int kMaxScalars = 8;  // passive scalars per solve
enum {
  BC_NONE = 0, BC_SLIP = 1, BC_WALL_ADIABATIC = 2, BC_WALL_ISOTHERMAL = 3,
  BC_CBC_UPT = 4, BC_CBC_MPT = 5, BC_CBC_RUNP = 6, BC_CBC_RUP = 7, BC_NSCBC_OUTLET_P = 8, BC_SPONGE = 9,
  BC_WM_ALG_ADIABATIC = 10, BC_CBC_TOTAL_PT = 11, BC_NSCBC_MT = 12,
  BC_WM_ALG_ISOTHERMAL = 13, BC_NSCBC_OUTLET_MDOT = 14, BC_WM_EQ_ADIABATIC = 15,
  BC_WM_EQ_ISOTHERMAL = 16
};
// the zones whose faces carry a mass flux (the Riemann flux against a boundary state)
__host__ __device__ inline bool bc_riemann(int type) {
  return (type >= BC_CBC_UPT && type <= BC_SPONGE) || type == BC_CBC_TOTAL_PT || type == BC_NSCBC_MT ||
         type == BC_NSCBC_OUTLET_MDOT;
}
// the zones whose faces carry an evolved NSCBC state
__host__ __device__ inline bool bc_nscbc(int type) {
  return type == BC_NSCBC_OUTLET_P || type == BC_NSCBC_MT || type == BC_NSCBC_OUTLET_MDOT;
}

struct Params {
  double gamma, R_gas, gm1, Rgogm1, p_ref, inv_pref, inv_Tref, mu_ref, mu_power_law, invPr_lam;
  double vreman_coeff, a_sgs_coeff, a_sgs_coeff2, lambda_dil_coeff, invPr_t, invgm1, sigma_coeff;
  double coeff_p, coeff_h, gr_nrm;
  int sgs_model, diss_mode, limit_asgs, dt_mode;
  double dt_data0, dt_data1;
  double p_g, h_g;  // mixed-precision gauges (p_ref and h at T_ref)
  double cp, rho_ref, inv_rhoref;  // full Gibbs-remainder sensor
  int full_gr;                     // USE_FULL_GR
  double dsm_alpha;                // DELTA_R_DSM^2
  int nsc, res_time;               // passive scalars; index of RES_TIME or -1
  // boundary conditions of the mesh's boundary zones (by zone index)
  int bc_type[kMaxBcZones];
  double bc_T[kMaxBcZones];
  double bc_u[kMaxBcZones][3];
  double bc_p[kMaxBcZones], bc_rho[kMaxBcZones], bc_q[kMaxBcZones];
  double bc_area[kMaxBcZones];  // the zone's area (sum of its faces' |n_bf|), CBC_MPT
  double bc_sigma[kMaxBcZones], bc_L[kMaxBcZones];  // NSCBC outlet sigma and L (p_ref in bc_p)
  int sponge_zone;                 // the SPONGE zone, or -1
  double sponge_p, sponge_trelax;  // its pressure and relaxation time
  double bc_sc[kMaxBcZones][kMaxScalars];  // scalars entering through a zone (CBC values, 0 otherwise)
  double Pr_lam;                           // PR_LAM (the equilibrium wall model's laminar conductivity)
  int bc_mode[kMaxBcZones];  // CBC_TOTAL_PT 1: USE_AVG; NSCBC_MT 1: T_TOTAL (bc_T then the total temperature)
};

using Geom = GeomT<double>;
using State = StateT<double>;
using Work = WorkT<double>;

// One solve as the kernels see it. A launch covers any number of solves: the
// kernels take a device table of Inst records and a block map whose entry
// for each launched block is (instance << kBlkBits) | instance-local block.
// The ping-pong index `cur` (which st[] holds the current state) is shared
// by all instances of a launch, which advance in lockstep stage by stage.
template <class R>
struct Inst {
  GeomT<R> g;
  StateT<R> st[2];
  WorkT<R> w;
  Params P;
  int nblocks;   // blocks of kBlock cells
  int npblocks;  // blocks of kBlock compact pairs
};

constexpr int kBlkBits = 20;  // up to 2^20 blocks (268M cells) per instance, 2^11 instances

}  // namespace cgdev

// Launchers (ig_kernels.inl), instantiated for R = double (ig_kernels_f64.hip)
// and R = float (ig_kernels_f32.hip). `I` is the device instance table,
// `bmap` / `pbmap` the device block maps of the cell and compact-pair
// launches, `lds` the dynamic LDS bytes of the extended-face pass (0: the
// global-gather variant; every instance then needs no halo).
namespace cgk {
template <class R>
struct Launch {
  const cgdev::Inst<R>* I;
  const int* bmap;   // cell blocks
  int nblk;
  int ninst;
  size_t lds;
  const int* pbmap;  // compact-pair blocks
  int npblk;
  int tile;          // extended slots fetched per batch in the LDS pass (2, 4 or 8)
  int sigma;         // some instance uses SGS_MODEL SIGMA (its SGS pass runs first)
  int full_gr;       // some instance uses USE_FULL_GR (its sensor pass runs first)
  int dsm;           // some instance uses SGS_MODEL DSM_LOCAL (its two passes run first)
  int scalars;       // most passive scalars of an instance (0: no scalar pass after each stage's flux)
  size_t sc_lds;     // dynamic LDS bytes of the scalar pass's halo variant (0: global gathers)
  int boundary;      // some instance has boundary faces (the properties pass keeps mu_lam, loc_lam)
  int zsum;          // some instance has zones needing zone sums (k_zone_sums precedes each stage's flux pass)
  int wm_nbf;        // most boundary faces of an instance with wall-model zones (0: none; k_wm_solve precedes stage 1)
};
template <class R>
int launch_flux(int stage, int mode, const Launch<R>& L, int cur, double* rhs_out, void* stream);
template <class R>
int launch_flux_compact(const Launch<R>& L, int cur, void* stream);
template <class R>
int launch_flux_ext(int stage, int mode, const Launch<R>& L, int cur, double* rhs_out, void* stream);
template <class R>
int launch_grad(const Launch<R>& L, int cur, void* stream);
template <class R>
int launch_eos(const Launch<R>& L, int cur, void* stream);
template <class R>
// fields: 0 the step's pass; 1 the same and the per-cell diagnostic copies (T,
// mu_lam, ...); 2 the copies again for a read of the current state (DSM_LOCAL's
// relaxation state is then left as it is)
int launch_props(const Launch<R>& L, int cur, int fields, void* stream);
// per-cell convective CFL into the work field cfl, or (max_bits not null) only
// its maximum into *max_bits as the bit pattern of a non-negative double (the
// caller zeroes it first)
template <class R>
int launch_cfl(const Launch<R>& L, int cur, double dt, void* stream, unsigned long long* max_bits = nullptr);
// the 8 diagnostic sums of every instance; partial / out: device buffers for the
// block partials and the sums (one instance; null: each instance's own buffers)
template <class R>
int launch_diag(const Launch<R>& L, int cur, void* stream, double* partial = nullptr, double* out = nullptr);
// rk3Step of a stage from an FP64 stage RHS rhs[5][ncv] (device cell order), then
// the stage's scalar pass (the unfused stage of a solve with a source hook)
template <class R>
int launch_rk_apply(int stage, const Launch<R>& L, int cur, const double* rhs, void* stream);
template <class R>
int launch_begin_step(const Launch<R>& L, void* stream);
// Nscbc::initialHook / Sponge::initialHook: every NSCBC outlet face's state
// and (sponge true) the sponge's relaxed state from their cells
template <class R>
int launch_nscbc_init(const Launch<R>& L, int cur, bool sponge, void* stream);
// the zone sums of stage `stage` (0: those of the initial hooks) into each
// instance's zsum; the flux pass of a step launches it itself
template <class R>
int launch_zone_sums(const Launch<R>& L, int cur, int stage, void* stream);
// FP64 primitives p, T, sos, h, sp_vol of one instance's state into out[5][ncv]
template <class R>
int launch_prims(const Launch<R>& L, int inst, int ncv, int cur, double* out, void* stream);
// layout setup: the per-slot pair geometry (extended n, c; compact n_fa and
// filter weight) gathered from the per-pair FP64 arrays on the device
template <class R>
int launch_layout_geom(int64_t nx, const int* xpair, const double* nU, const double* cU, R* const x6[6], int64_t nc,
                       const int* cpair, const double* nfa, const double* fw, R* const cn[3], double* cw,
                       void* stream);
int launch_fill(double* p, double v, int64_t n, void* stream);
int launch_fill(float* p, float v, int64_t n, void* stream);
}  // namespace cgk
