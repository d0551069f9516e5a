// hg_kernels.hpp -- device data layout and kernel launch interface.
//
// Layout in HBM (one context): every field is an interior-only SoA array
// of n1*n2*n3 doubles, x fastest (the reference's order without ghosts).
// Ghost cells are never stored: the sweep kernels evaluate the boundary
// conditions of boundarycondition.f90 on the fly when a stencil reaches
// past the domain, from the interior cells they mirror or wrap to.
//
//   W[NW]   primitive fields  d v1 v2 v3 e eint [b1 b2 b3 phi]
//   U[NV]   conserved u(:,:,:,1..ufnmax)      (rungekutta.f90)
//   U0[NV]  uorg, the RK stage-1 copy
//   DU[3][NV] per-direction flux divergence  idxi_d * (F(c-1/2) - F(c+1/2))
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hg {

enum { WD = 0, WV1, WV2, WV3, WE, WEINT, WB1, WB2, WB3, WPHI, NWMAX };
enum { NVMAX = 9 };

// device-resident scalars shared by the kernels of one context
struct Scalars {
  double dt, ch;          // hydro step inputs
  double dtimin, cfmax;   // CFL reduction results
  unsigned long long err; // first error key (0 = none), see errkey()
  double pad[3];
};

// geometry of one sweep direction, indices 0-based interior (ghosts -2,-1,n,n+1)
struct Axis {
  const double *x, *xi, *idx, *idxi, *dx, *dxi;  // base-shifted: valid for c = -2..n+1
  int n;
  long stride;      // element stride of this axis in the SoA arrays
  int sc_lo, sc_hi; // scalar BC codes (bcNis, bcNos)
  int vc_lo, vc_hi; // vector BC codes (bcNiv, bcNov)
};

struct Dev {
  int n1, n2, n3;          // interior extents
  long ncell;
  int nv, nw;              // conserved / primitive field counts
  int mag, dim, limiter;
  int solve[3];
  int i9wv;                // index of the GLM variable or -1
  double gamma;
  Axis ax[3];
  double *W[NWMAX];
  double *U[NVMAX], *U0[NVMAX];
  double *DU[3][NVMAX];
  double *part;            // CFL partials: 2 doubles per block
  Scalars *sc;
  int nblk_cfl;
};

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

void launch_pressure(const Dev &D, hipStream_t s);
void launch_set_scalars(const Dev &D, double dt, double ch, hipStream_t s);
void launch_sweep(const Dev &D, int dir, int stage, hipStream_t s);
void launch_rk(const Dev &D, int stage, double faco, double fact, double facn,
               hipStream_t s);
void launch_cfl(const Dev &D, hipStream_t s);
void launch_phidamp(const Dev &D, double factor, hipStream_t s);
void launch_kat(int which, int n, const double *in, double *out, hipStream_t s);

}  // namespace hg
