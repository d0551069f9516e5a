// hg_kernels.hpp -- device data layout and kernel launch interface.
//
// Layout in HBM (one context): every field is an interior-only SoA array
// of n1*n2*n3 doubles, x fastest (the reference's order without ghosts).
// Ghost cells are never stored: the sweep kernels evaluate the boundary
// conditions of boundarycondition.f90 on the fly when a stencil reaches
// past the domain, from the interior cells they mirror or wrap to.
//
//   W[NW]   primitive fields  d v1 v2 v3 e eint [b1 b2 b3 phi]; d, e and B
//           are the same buffers as U[0], U[4], U[5..7]
//   U[NV]   conserved u(:,:,:,1..ufnmax)      (rungekutta.f90)
//   U0[NV]  uorg, the RK stage-1 copy
//   DU[3][NV] per-direction flux divergence  idxi_d * (F(c-1/2) - F(c+1/2))
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hormone_hip.h"

namespace hg {

enum { WD = 0, WV1, WV2, WV3, WE, WEINT, WB1, WB2, WB3, WPHI, NWMAX };
enum { NVMAX = 9 };
static constexpr int SEG = 60;  // cells a 64-lane sweep wavefront updates (2-cell halos)

// Wavefronts of one direction's sweep: ncol sequences of nsub lines laid end
// to end, each line with its 2 ghost cells at either end (n+4 positions);
// wave s of a sequence holds positions 60*s .. 60*s+63 and updates those of
// its lanes 2..61, so the last interior cell, at position nsub*(n+4)-3,
// needs (nsub*(n+4)-5)/60 + 1 waves.
__host__ __device__ inline long sweep_waves(long ncol, long nsub, int n) {
  return ncol * ((nsub * (long)(n + 4) - 5) / SEG + 1);
}

// stop reasons of the device-resident main loop
enum { STOP_NONE = 0, STOP_END = 1, STOP_OUTPUT = 2, STOP_MAXSTEPS = 3, STOP_ERROR = 4 };

// Device-resident scalars of one context: the hydro step inputs (dt, ch),
// the CFL reduction results and the main-loop clock of hormone.f90:124-179.
struct Scalars {
  double dt, ch;
  double dtimin, cfmax;
  double time, t_out, t_end, dt_out;
  long long tn, tnlim, tn_out, steps_left;
  int outstyle, endstyle, stop, pad0;
  unsigned long long err;  // first error key (NOERR = none), see errkey()
};

// geometry of one sweep direction, indices 0-based interior (ghosts -2,-1,n,n+1)
struct Axis {
  const double *x, *xi, *idx, *idxi, *dx, *dxi;  // base-shifted: valid for c = -2..n+1
  // mixed-precision path: per-cell limiter geometry evaluated once on the
  // host, (cf, cb, (x3-x2)/(x2-x1), 1/(x3-x2)), and the minmod weight
  // idx(i)*idx(i+1)/(idx(i)+idx(i+1))
  const double4 *limg;
  const double *mmw;
  int n;
  unsigned nlines;   // lines along this axis (ncell / n)
  unsigned nsub;     // lines laid end to end in one sweep sequence (lane_cell)
  long stride;       // element stride of this axis in the SoA arrays
  int sc_lo, sc_hi;  // scalar BC codes (bcNis, bcNos)
  int vc_lo, vc_hi;  // vector BC codes (bcNiv, bcNov)
  // codes 4, 5 (linear), 9 (Dirichlet) and 10 (flux) on this axis: ext is
  // set when any of the four codes is one of them.  G[side] holds, for a
  // face with such a code, the ghost values the boundary code does not
  // compute from the interior, [GF_*][layer 0 (next to the face), 1][line];
  // FX[side] the flux of a flux-boundary face (conserved-variable order,
  // NVMAX values).
  int ext;
  const double *G[2];
  const double *FX[2];
};

// fields of the ghost-value buffers
enum { GF_D = 0, GF_P, GF_V1, GF_V2, GF_V3, GF_B1, GF_B2, GF_B3, GF_PHI, NGF };

struct Dev {
  int n1, n2, n3;  // interior extents
  long ncell;
  int nv, nw;      // conserved / primitive field counts
  int mag, dim, limiter;
  int solve[3];
  int i9wv;        // index of the GLM variable or -1
  double gamma, courant;
  Axis ax[3];
  double *W[NWMAX];
  double *U[NVMAX], *U0[NVMAX];
  double *DU[3][NVMAX];
  // advected species (compswitch 2): mass fractions SPC[n*ncell + m], the
  // stage-1 partial densities SPC0 (spcorg) and the per-direction species
  // flux divergence DUS[(d*nspc + n)*ncell + m]
  int nspc;
  double *SPC, *SPC0, *DUS;
  double *FM[3];  // face mass flux per direction, (n+1) faces per line
  // one sweep direction fused with the RK update: that launch writes the
  // new state to the other buffer set Uo/Wo (W aliasing U as above)
  int fdir;        // the direction fused with the RK update, -1 if none
  double *Uo[NVMAX], *Wo[NWMAX];
  double *part;    // CFL partials: 2 doubles per block of 256 cells
  Scalars *sc;
  int nblk_cfl;
  long waves[3];   // sweep wavefronts per direction (0 if not solved)
  long woff[3];    // first wavefront of each direction in a sweep launch
  int lds_max;     // LDS bytes one block may use on this device
  int sweep_bw;    // sweep wavefronts per block (1, 2, 4, 8 or 16)
  // hyperbolic self-gravity (gravswitch 3): the potential is stored with a
  // one-cell ghost layer ((n1+2)(n2+2)(n3+2), Dirichlet values), twice:
  // gphi[gcur] holds the current potential, the other the stage-1 one; psi
  // (interior) likewise in gpsi and gpsi1 (k_grav_substep2, k_grav_stage1).
  int mixed;       // HG_PREC_FP32MIX
  double damp;     // phidamp factor exp(-alpha9wave*courant), evaluated on the host
  int grav;
  double *gphi[2], *gpsi, *gpsi1;
  int gcur;          // which gphi buffer holds the current potential
  double fourpig;    // 4*pi*G
};

// index of interior cell (i,j,k) in the potential's padded layout
__host__ __device__ inline long gidx(const Dev &D, int i, int j, int k) {
  return ((long)(k + 1) * (D.n2 + 2) + (j + 1)) * (long)(D.n1 + 2) + (i + 1);
}

// No error recorded: errors lower the key with atomicMin, so "none" must be
// the largest value.
constexpr unsigned long long NOERR = ~0ull;

// Error keys: the smallest key is the first error in the reference's serial
// loop order (stage, then numflux before the cell update, then direction,
// then k, j, i).  Cell indices are 0-based interior, ghosts down to -2.
__host__ __device__ inline unsigned long long errkey(int stage, int phase, int dir, int i,
                                                     int j, int k, const Dev &D) {
  unsigned long long cell =
      ((unsigned long long)(k + 2) * (D.n2 + 4) + (unsigned long long)(j + 2)) * (D.n1 + 4) +
      (unsigned long long)(i + 2);
  unsigned long long key = ((unsigned long long)stage * 2 + phase) * 4 + dir;
  return (key << 42) + cell + 1;
}

// A batch of independent boxes advanced by the same launches: flattened
// work-lists of (box, sweep wavefront) and (box, 256-cell block) items over
// a device array of the members' Dev records.
struct Batch {
  const Dev *boxes;
  int nbox;
  // (box, wavefront of the direction) items, direction by direction: those
  // of direction d are sitems[dbeg[d] .. dend[d]), each box's range padded
  // to a multiple of 16 items
  const int2 *sitems;
  long nsitems;
  long dbeg[3], dend[3];
  int lds_max;         // LDS bytes one block may use
  const int2 *citems;
  long ncitems;
  // the x-fused form (every member sweeps x with its RK update), with three
  // state sets per box: bpar[q] is the set p that holds box q's state at
  // the start of its step; stage 1 writes set p+1, stage 2 set p+2, stage 3
  // set p+1 again (mod 3), the later stages reading u0 from set p.  boxes
  // holds nine records per box, 9q + 3p + (stage - 1); rk: stages per step
  int fused, rk;
  int *bpar;
};

// one direction's sweep over the batch (k_bsweep_dir)
void launch_bsweep_dir(const Batch &B, int dir, int mag, int lim, int mixed, int stage,
                       unsigned grid, int block_waves, hipStream_t s);
// the x-fused form's x sweep with the RK update over the batch (k_bsweep_rk)
void launch_bsweep_rk(const Batch &B, int mag, int lim, int mixed, int stage, double faco,
                      double fact, double facn, bool last, unsigned grid, int block_waves,
                      hipStream_t s);
void launch_brk(const Batch &B, int mag, int stage, double faco, double fact, double facn,
                bool last, unsigned grid, hipStream_t s);
void launch_bstep_end_cfl(const Batch &B, hipStream_t s);  // per box: step_end, cfl_final

// RK coefficients (faco, fact, facn) of get_runge_coeff, rungekutta.f90:146-183
void runge_coeff(int stage, int rktype, double *faco, double *fact, double *facn);

void launch_pressure(const Dev &D, hipStream_t s);
void launch_clear_error(const Dev &D, hipStream_t s);
void launch_timestep_out(const Dev &D, double *out, hipStream_t s);  // dt, ch, status
void launch_set_step(const Dev &D, double dt, double ch, hipStream_t s);
void launch_set_clock(const Dev &D, double time, double t_out, double t_end, double dt_out,
                      int outstyle, int endstyle, long long tn, long long tnlim,
                      long long tn_out, long long steps, hipStream_t s);
// Optional kernel timing: a launch given a KEv carries the two events on the
// dispatch itself (hipExtLaunchKernel), so their elapsed time is the kernel's
// own duration, with no separate marker packets around it.
struct KEv {
  hipEvent_t start, stop;
};

void launch_sweep(const Dev &D, int stage, hipStream_t s, const KEv *ev = nullptr);
// RK coefficients and options of the fused last sweep
struct RKArgs {
  double faco, fact, facn, damp;
  int cfl, dampf, grav;
};
// the fused path of one RK stage: the sweeps of the other directions, then
// D.fdir's sweep with the RK update into D.Uo/D.Wo; returns the number
// of CFL partials the fused launch wrote (rk.cfl).  part 1 enqueues only the
// other directions' sweep (it needs neither dt nor the gravity potential),
// part 2 only the fused launch, 0 both.
int launch_fused_stage(const Dev &D, int stage, const RKArgs &rk, hipStream_t s,
                       const KEv *ev_sweep = nullptr, const KEv *ev_fused = nullptr,
                       int part = 0);
void launch_species(const Dev &D, hipStream_t s);
size_t sweep_tile_bytes(const Dev &D, int bw);
// cfl: also the next step's CFL partials; damp: phidamp
void launch_rk(const Dev &D, int stage, double faco, double fact, double facn, bool cfl,
               bool damp, double damp_factor, hipStream_t s, const KEv *ev = nullptr);
void launch_grav_substep(const Dev &D, int gk, double dtg, double cg, double damp_factor,
                         hipStream_t s, const KEv *ev = nullptr);
void launch_cfl(const Dev &D, hipStream_t s, const KEv *ev = nullptr);
void launch_cfl_final(const Dev &D, hipStream_t s);
void launch_step_end_cfl(const Dev &D, hipStream_t s);  // step_end, then cfl_final
void launch_phidamp(const Dev &D, double factor, hipStream_t s);
void launch_kat(int which, int n, const double *in, double *out, hipStream_t s);

}  // namespace hg
