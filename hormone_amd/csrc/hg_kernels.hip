// hg_kernels.hip -- FP64 kernels of the explicit hydro step on gfx950.
//
// One RK stage is: one "sweep" kernel per solved direction (MUSCL slopes,
// e_int guard, face states, HLLD/GLM flux and the per-direction flux
// difference), then one "rk" kernel (SSP-RK combination, primitive recovery,
// e_int check).  The timestep kernel reduces the CFL condition.
//
// Sweep mapping: a wavefront owns 64 consecutive positions of a sequence of
// lines laid end to end along the sweep direction, each line with its two
// ghost cells at either end (lane_cell); lane l holds one cell.  Neighbour
// values travel by cross-lane shuffles: slopes need l-1 and l+1, the face
// between l and l+1 needs the right state from l+1, and the flux difference
// of cell l needs the face flux of l-1.  Lanes 2..61 produce a result when
// they hold an interior cell.  Boundary ghosts are evaluated on the fly.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "hg_kernels.hpp"
#include "hg_physics.hpp"

namespace hg {

// Workgroups go to the 8 XCDs round robin (workgroup b to XCD b % 8), and
// each XCD has its own L2.  xcd_block renumbers them so that each XCD takes
// one contiguous range of the logical blocks (XCD x the x-th range; the
// first nb % 8 ranges are one block longer): data that neighbouring blocks
// share then sits in one L2 instead of being fetched once per XCD.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb) {
  const unsigned x = b & 7u, q = nb >> 3, r = nb & 7u;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + (b >> 3);
}



// A pointer read out of device memory (a batch's box record) is generic to
// the compiler, which then loads and stores through it with flat
// instructions; passing it through the global address space lets it emit
// global ones (with a scalar base).  The state, divergence and coordinate
// arrays are all device (global) memory.  A no-op for pointers the compiler
// already knows to be global (kernel arguments).
template <class T>
__device__ __forceinline__ T *gptr(T *p) {
  return (T *)(__attribute__((address_space(1))) T *)p;
}

// The wavefront's index within its block, visibly uniform: index arithmetic
// on it (the packed-line divisions) then runs on the scalar unit.
__device__ __forceinline__ int wave_in_block() {
  return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

struct Cell {
  double d, v[3], e, eint, b[3], phi;
};

__device__ __forceinline__ int ghost_src(int code, bool lo, int c, int n) {
  // boundarycondition.f90: 0 periodic, 1 reflective, 2/3 outgoing/free
  int s;
  if (code == 0) s = lo ? c + n : c - n;
  else if (code == 1) s = lo ? -1 - c : 2 * n - 1 - c;
  else s = lo ? 0 : n - 1;
  return s < 0 ? 0 : (s > n - 1 ? n - 1 : s);  // lanes past the ghosts: unused
}

// One cell of a sweep line, interior or ghost, from the values stored at its
// source cells: f[] holds W[d v1 v2 v3 e eint b1 b2 b3 phi], the scalars
// (d, e, eint, phi) read at the scalar-code source and v, B at the
// vector-code source.  Ghost cells (boundarycondition.f90) transform the
// normal components (codes 0..3), then rebuild eint and e (:221-243) from
// the ghost pressure: gm1*eint of the source cell, or, for scalar codes 4,
// 5, 9 and 10, the pressure ghost_ext left in f[WEINT].
template <int DIR, bool MAG>
__device__ __forceinline__ Cell finish_cell(const Dev &D, const Axis &A, int c,
                                            const double f[NWMAX]) {
  Cell q;
  const int n = A.n;
  const bool ghost = c < 0 || c >= n;
  const bool lo = c < 0;
  const int vc = lo ? A.vc_lo : A.vc_hi;
  const int sc = lo ? A.sc_lo : A.sc_hi;
  q.d = f[WD];
  q.phi = MAG ? f[WPHI] : 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double v = f[WV1 + a];
    double b = MAG ? f[WB1 + a] : 0.0;
    if (ghost && a == DIR && vc <= 3) {
      if (vc == 1) {
        v = -v;
        b = -b;
      } else if (vc == 2) {
        if (DIR == 0 && lo) v = mn2(0.0, v);
        else v = v * (lo ? 0.5 - sgn(0.5, v) : 0.5 + sgn(0.5, v));
      }
    }
    q.v[a] = v;
    q.b[a] = b;
  }
  if (!ghost) {
    q.eint = f[WEINT];
    q.e = f[WE];
  } else {
    const double gm1 = D.gamma - 1.0;
    const double p = sc >= 4 ? f[WEINT] : gm1 * f[WEINT];
    q.eint = p / gm1;
    q.e = q.eint + 0.5 * (q.d * (q.v[0] * q.v[0] + q.v[1] * q.v[1] + q.v[2] * q.v[2]) +
                          q.b[0] * q.b[0] + q.b[1] * q.b[1] + q.b[2] * q.b[2]);
  }
  return q;
}

// Source rows of cell c of a line (both the cell itself when interior).
__device__ __forceinline__ void cell_sources(const Axis &A, long base, int c, long &ms,
                                             long &mv) {
  const int n = A.n;
  ms = base + (long)c * A.stride;
  mv = ms;
  if (c < 0 || c >= n) {
    const bool lo = c < 0;
    ms = base + (long)ghost_src(lo ? A.sc_lo : A.sc_hi, lo, c, n) * A.stride;
    mv = base + (long)ghost_src(lo ? A.vc_lo : A.vc_hi, lo, c, n) * A.stride;
  }
}

// Ghost cell c of line `line` at a face whose scalar or vector code is 4, 5
// (linear extrapolation), 9 (Dirichlet) or 10 (flux): overwrites the raw
// values the code defines.  Scalars: d and the pressure (left in f[WEINT])
// and phi; vectors: v and B.
// * 4/5, boundarycondition.f90:302-326 (x1 outer), :963-985 (x3 inner),
//   :1222-1244 (x3 outer): q(ghost 1) and q(ghost 2) extrapolated from the
//   edge cell and its inner neighbour with the cell-width ratios, d and p
//   reset to the edge value when the outer ghost is not positive, phi (and
//   species) copied from the edge.  Vectors :419-468, :1074-1120,
//   :1335-1382; 5 also multiplies the normal velocity by the outgoing plug
//   of the edge velocity.  Below the x3 inner boundary the reference never
//   writes v2 and b2 of the outer ghost (:1081,1088,1104,1114 assign v3/b3
//   there), which keep the values of G.
// * 9: d0, p0, v0, b0 of the ghost cells (:104-116, :196-212, ...), here G;
//   phi is not written by the reference, G again.
// * 10: the reference writes no ghost value at all: G.
template <int DIR, bool MAG>
__device__ __forceinline__ void ghost_ext(const Dev &D, const Axis &A, long base, long line,
                                          int c, double f[NWMAX]) {
  const int n = A.n;
  const bool lo = c < 0;
  const int sc = lo ? A.sc_lo : A.sc_hi, vc = lo ? A.vc_lo : A.vc_hi;
  if (sc < 4 && vc < 4) return;
  const int layer = (c == -1 || c == n) ? 0 : 1;  // lanes past the ghosts: unused
  const long nl = A.nlines;
  const double *G = A.G[lo ? 0 : 1] + (long)layer * nl + line;  // field g at G[g * 2 * nl]
  const long me = base + (long)(lo ? 0 : n - 1) * A.stride;      // edge cell
  const long mn = base + (long)(lo ? 1 : n - 2) * A.stride;      // its inner neighbour
  const double r1 = gptr(A.dx)[lo ? 0 : n - 1], r1d = gptr(A.dx)[lo ? 1 : n - 2];
  const double r2 = gptr(A.dx)[lo ? -1 : n], r2d = gptr(A.dx)[lo ? 0 : n - 1];
  auto lin = [&](double qe, double qn, double &q1, double &q2) {
    if (lo) {
      q1 = qe - (qn - qe) * r1 / r1d;
      q2 = q1 - (qe - q1) * r2 / r2d;
    } else {
      q1 = qe + (qe - qn) * r1 / r1d;
      q2 = q1 + (q1 - qe) * r2 / r2d;
    }
  };
  if (sc == 4 || sc == 5) {
    const double gm1 = D.gamma - 1.0;
    double q1, q2;
    const double de = gptr(D.W[WD])[me];
    lin(de, gptr(D.W[WD])[mn], q1, q2);
    f[WD] = q2 <= 0.0 ? de : (layer ? q2 : q1);
    const double pe = gm1 * gptr(D.W[WEINT])[me];
    lin(pe, gm1 * gptr(D.W[WEINT])[mn], q1, q2);
    f[WEINT] = q2 <= 0.0 ? pe : (layer ? q2 : q1);
    if (MAG) f[WPHI] = gptr(D.W[WPHI])[me];
  } else if (sc >= 9) {
    f[WD] = G[GF_D * 2 * nl];
    f[WEINT] = G[GF_P * 2 * nl];
    if (MAG) f[WPHI] = G[GF_PHI * 2 * nl];
  }
  if (vc == 4 || vc == 5) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      double q1, q2;
      lin(gptr(D.W[WV1 + a])[me], gptr(D.W[WV1 + a])[mn], q1, q2);
      f[WV1 + a] = layer ? q2 : q1;
      if (MAG) {
        lin(gptr(D.W[WB1 + a])[me], gptr(D.W[WB1 + a])[mn], q1, q2);
        f[WB1 + a] = layer ? q2 : q1;
      }
    }
    if (DIR == 2 && lo && layer == 1) {
      f[WV2] = G[GF_V2 * 2 * nl];
      if (MAG) f[WB2] = G[GF_B2 * 2 * nl];
    }
    if (vc == 5) {
      const double vn = gptr(D.W[WV1 + DIR])[me];
      f[WV1 + DIR] = f[WV1 + DIR] * (lo ? 0.5 - sgn(0.5, vn) : 0.5 + sgn(0.5, vn));
    }
  } else if (vc >= 9) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      f[WV1 + a] = G[(GF_V1 + a) * 2 * nl];
      if (MAG) f[WB1 + a] = G[(GF_B1 + a) * 2 * nl];
    }
  }
}

// Raw values of cell c of a line (interior, or ghost by its face's code).
template <int DIR, bool MAG>
__device__ __forceinline__ void raw_cell(const Dev &D, const Axis &A, long base, long line,
                                         int c, double f[NWMAX]) {
  long ms, mv;
  cell_sources(A, base, c, ms, mv);
  f[WD] = gptr(D.W[WD])[ms];
  f[WE] = gptr(D.W[WE])[ms];
  f[WEINT] = gptr(D.W[WEINT])[ms];
#pragma unroll
  for (int a = 0; a < 3; ++a) f[WV1 + a] = gptr(D.W[WV1 + a])[mv];
#pragma unroll
  for (int a = 0; a < 3; ++a) f[WB1 + a] = MAG ? gptr(D.W[WB1 + a])[mv] : 0.0;
  f[WPHI] = MAG ? gptr(D.W[WPHI])[ms] : 0.0;
  if (A.ext && (c < 0 || c >= A.n)) ghost_ext<DIR, MAG>(D, A, base, line, c, f);
}

// One cell read straight from HBM by its own lane.
// EXT = false: the caller's boxes have no code 4, 5, 9 or 10 (batches).
template <int DIR, bool MAG, bool EXT = true>
__device__ __forceinline__ Cell load_cell(const Dev &D, const Axis &A, long base, long line,
                                          int c) {
  if (EXT && A.ext) {
    double f[NWMAX];
    raw_cell<DIR, MAG>(D, A, base, line, c, f);
    return finish_cell<DIR, MAG>(D, A, c, f);
  }
  long ms, mv;
  cell_sources(A, base, c, ms, mv);
  double f[NWMAX];
  f[WD] = gptr(D.W[WD])[ms];
  f[WE] = gptr(D.W[WE])[ms];
  f[WEINT] = gptr(D.W[WEINT])[ms];
#pragma unroll
  for (int a = 0; a < 3; ++a) f[WV1 + a] = gptr(D.W[WV1 + a])[mv];
  if (MAG) {
#pragma unroll
    for (int a = 0; a < 3; ++a) f[WB1 + a] = gptr(D.W[WB1 + a])[mv];
    f[WPHI] = gptr(D.W[WPHI])[ms];
  }
  return finish_cell<DIR, MAG>(D, A, c, f);
}

// ---- LDS-staged lines for the strided (y, z) sweeps ------------------------
// A block's bw wavefronts own bw consecutive waves of one direction, i.e.
// (mostly) neighbouring lines of consecutive x index over the same 64-cell
// segment, so the block's cells form a tile [64 cells][bw lines] whose rows
// are (mostly) contiguous in HBM.  Loading the tile row by row (one element
// per thread and field) replaces 64 lane-strided accesses per instruction --
// one cache line and, along z, one page per lane -- by bw/64-row coalesced
// ones; the flux divergences go back out the same way.  Each column finds
// its own line and segment, so a block may span the end of an x row or of a
// segment; columns past the last wave of the direction (padding of a
// partial block) work on a copy of the last line and store nothing.  LDS
// layout [field][cell row][line], rows padded to bw+1, then the index table.
__device__ __forceinline__ int tile_ld(int bw) { return bw + 1; }

// The tile's index table, two arrays of 64*bw entries after the fields:
// entry e = r*bw + col belongs to row r of column col, that is lane r of
// the block's wavefront col, and also to thread e, which loads and stores
// that row.  src[e] = the cell's source elements for the scalar and the
// vector fields (a ghost's boundary source), dst[e] = (the element its flux
// divergence goes to, or TILE_NONE, and its cell index c along the line).
// The owning lane computes them once from its own lane_cell; the loading
// and storing threads read them instead of repeating the packed-line
// divisions.  Element indices fit in 32 bits (every box holds fewer than
// 2^32 elements).
constexpr unsigned TILE_NONE = 0xffffffffu;
__device__ __forceinline__ uint2 *tile_src(double *lds, int nf, int bw) {
  return reinterpret_cast<uint2 *>(lds + nf * 64 * tile_ld(bw));
}
__device__ __forceinline__ uint2 *tile_dst(double *lds, int nf, int bw) {
  return tile_src(lds, nf, bw) + 64 * bw;
}

// The fused launch's lane path keeps its own cell's primitives in LDS
// between the load and the RK update: d, e and B are U[0], U[4] and
// U[5..7] (the same buffers) and d, v feed the gravity source, so the
// update reads none of them from memory again, and they hold no registers
// across the flux computation.  Layout [field][thread], fields d e v1 v2
// v3 (b1 b2 b3); each thread reads back only what it wrote.
__host__ __device__ inline int park_fields(bool mag) { return mag ? 8 : 5; }
template <bool MAG>
__device__ __forceinline__ void park_cell(double *lds, const Cell &q) {
  double *p = lds + threadIdx.x;
  const int bs = blockDim.x;
  p[0] = q.d;
  p[bs] = q.e;
#pragma unroll
  for (int a = 0; a < 3; ++a) p[(2 + a) * bs] = q.v[a];
  if (MAG) {
#pragma unroll
    for (int a = 0; a < 3; ++a) p[(5 + a) * bs] = q.b[a];
  }
}
template <bool MAG>
__device__ __forceinline__ Cell parked_cell(const double *lds) {
  const double *p = lds + threadIdx.x;
  const int bs = blockDim.x;
  Cell q;
  q.d = p[0];
  q.e = p[bs];
#pragma unroll
  for (int a = 0; a < 3; ++a) q.v[a] = p[(2 + a) * bs];
#pragma unroll
  for (int a = 0; a < 3; ++a) q.b[a] = MAG ? p[(5 + a) * bs] : 0.0;
  q.eint = 0.0;
  q.phi = 0.0;
  return q;
}

// ---- packed sweep lines ----------------------------------------------------
// The lines of a direction, each extended by its two ghost cells at either
// end (n+4 positions), lie end to end in sequences: all x lines form one
// sequence (row order j + k*n2), the y lines of column i one (k order), the
// z lines of column i another (j order).  Wave w of the direction takes
// column i = w % n1 (y, z; x has the one sequence) and positions
// 60*s .. 60*s+63 of it, s = w / n1: lane r holds position 60*s + r.
// Consecutive waves are neighbouring columns at the same positions, so the
// rows of a y/z tile stay contiguous in HBM.  Lanes 2..61 of the waves hold
// every position from 2 on exactly once, and the stencil c-2 .. c+2 of an
// interior cell c lies inside its own padded line, so a wavefront no longer
// stops at the end of a line: a 256-cell line costs 260/60 wavefronts
// instead of 5.
// lane_cell: the line of position r of wave w (numbered as A.nlines counts
// them: x j + k*n2, y and z i + (k or j)*n1), its base element and the cell
// c (-2 .. n+1) of the position; false past the sequence's end, where c is
// n+1 of the last line and the lane stores nothing.
template <int DIR>
__device__ __forceinline__ bool lane_cell(const Dev &D, const Axis &A, long w, int r,
                                          long &line, long &base, int &c) {
  const unsigned n4 = (unsigned)A.n + 4u;
  unsigned col = 0, s = (unsigned)w;
  if (DIR != 0) {
    col = (unsigned)w % (unsigned)D.n1;
    s = (unsigned)w / (unsigned)D.n1;
  }
  const unsigned pos = s * (unsigned)SEG + (unsigned)r;
  unsigned sub = pos / n4;
  c = (int)(pos - sub * n4) - 2;
  const bool in = sub < A.nsub;
  if (!in) {
    sub = A.nsub - 1;
    c = A.n + 1;
  }
  if (DIR == 0) {
    line = sub;
    base = (long)sub * D.n1;
  } else {
    line = (long)col + (long)sub * D.n1;
    base = DIR == 1 ? (long)col + (long)sub * D.n1 * D.n2 : line;
  }
  return in;
}

// cell-0 coordinates of a line (error reports)
template <int DIR>
__device__ __forceinline__ void line_origin(const Dev &D, long line, int &ci, int &cj, int &ck) {
  if (DIR == 0) {
    ci = 0; cj = (int)(line % D.n2); ck = (int)(line / D.n2);
  } else if (DIR == 1) {
    ci = (int)(line % D.n1); cj = 0; ck = (int)(line / D.n1);
  } else {
    ci = (int)(line % D.n1); cj = (int)(line / D.n1); ck = 0;
  }
}

// (i, j, k) of element m of a field; 32-bit arithmetic (hg_create keeps
// every box below 2^32 elements)
__device__ __forceinline__ void cell_ijk(const Dev &D, long m, int &i, int &j, int &k) {
  const unsigned u = (unsigned)m, n1 = (unsigned)D.n1, n2 = (unsigned)D.n2;
  const unsigned q = u / n1;
  const unsigned kk = q / n2;
  i = (int)(u - q * n1);
  j = (int)(q - kk * n2);
  k = (int)kk;
}

template <int DIR, bool MAG, bool EXT = true>
__device__ __forceinline__ Cell tile_cell(const Dev &D, const Axis &A, long wave0, long base,
                                          int c, bool valid, double *lds) {
  const int bw = blockDim.x >> 6, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ld = tile_ld(bw), fs = 64 * ld;
  const int nf = MAG ? NWMAX : WEINT + 1;
  uint2 *const src = tile_src(lds, nf, bw), *const dst = tile_dst(lds, nf, bw);
  {  // this lane's entry: lane r of column w
    long ms, mv;
    cell_sources(A, base, c, ms, mv);
    const bool st = valid && lane >= 2 && lane <= 61 && c >= 0 && c < A.n;
    const int e = lane * bw + w;
    src[e] = make_uint2((unsigned)ms, (unsigned)mv);
    dst[e] = make_uint2(st ? (unsigned)(base + (long)c * A.stride) : TILE_NONE, (unsigned)c);
  }
  __syncthreads();
  {  // cooperative load: thread t -> (row r, column col), table entry t
    // bw is a power of two (sweep blocks of 1, 2, 4, 8 or 16 wavefronts)
    const int t = threadIdx.x, r = t >> __builtin_ctz(bw), col = t & (bw - 1);
    const uint2 sv = src[t];
    const int cr = (int)dst[t].y;
    double *p = lds + r * ld + col;
    if (EXT && A.ext && (cr < 0 || cr >= A.n)) {
      // a ghost the boundary code computes (raw_cell needs its line)
      const long last = D.waves[DIR] - 1;
      const long wc = wave0 + col < last ? wave0 + col : last;
      long ln, bc;
      int cl;
      lane_cell<DIR>(D, A, wc, r, ln, bc, cl);
      double f[NWMAX];
      raw_cell<DIR, MAG>(D, A, bc, ln, cl, f);
#pragma unroll
      for (int k = 0; k < NWMAX; ++k)
        if (k < nf) p[k * fs] = f[k];
    } else {
      const long ms = sv.x, mv = sv.y;
      p[WD * fs] = gptr(D.W[WD])[ms];
      p[WE * fs] = gptr(D.W[WE])[ms];
      p[WEINT * fs] = gptr(D.W[WEINT])[ms];
#pragma unroll
      for (int a = 0; a < 3; ++a) p[(WV1 + a) * fs] = gptr(D.W[WV1 + a])[mv];
      if (MAG) {
#pragma unroll
        for (int a = 0; a < 3; ++a) p[(WB1 + a) * fs] = gptr(D.W[WB1 + a])[mv];
        p[WPHI * fs] = gptr(D.W[WPHI])[ms];
      }
    }
  }
  __syncthreads();
  double f[NWMAX];
  const double *p = lds + lane * ld + w;
#pragma unroll
  for (int k = 0; k < NWMAX; ++k) f[k] = k < nf ? p[k * fs] : 0.0;
  return finish_cell<DIR, MAG>(D, A, c, f);
}

// Thread t's row of the tile on the way out: true with its element index m
// when the row holds an interior cell that its column updates.
template <bool MAG>
__device__ __forceinline__ bool tile_row_cell(double *lds, int bw, int t, long &m) {
  const unsigned e = tile_dst(lds, MAG ? NWMAX : WEINT + 1, bw)[t].x;
  m = e;
  return e != TILE_NONE;
}

// Flux-divergence stores of a line: du[v] is the value of conserved
// variable v for this lane's cell c (lanes 2..61 with 0 <= c < n own one).
template <int DIR, bool MAG, bool TILE>
__device__ __forceinline__ void store_du(const Dev &D, const Axis &A, long base, int c,
                                         int lane, const double du[NVMAX], double *lds) {
  const int n = A.n;
  double *const *DU = D.DU[DIR];
  const int nv = MAG ? (D.i9wv >= 0 ? 9 : 8) : 5;
  if (!TILE) {
    if (lane >= 2 && lane <= 61 && c >= 0 && c < n) {
      const long m = base + (long)c * A.stride;
#pragma unroll
      for (int v = 0; v < NVMAX; ++v)
        if (v < nv) gptr(DU[v])[m] = du[v];
    }
    return;
  }
  const int bw = blockDim.x >> 6, w = threadIdx.x >> 6;
  const int ld = tile_ld(bw), fs = 64 * ld;
  __syncthreads();  // every wavefront has read its cells out of the tile
  {
    double *p = lds + lane * ld + w;
#pragma unroll
    for (int v = 0; v < NVMAX; ++v)
      if (v < nv) p[v * fs] = du[v];
  }
  __syncthreads();
  const int t = threadIdx.x, r = t >> __builtin_ctz(bw), col = t & (bw - 1);
  long m;
  if (tile_row_cell<MAG>(lds, bw, t, m)) {
    const double *p = lds + r * ld + col;
#pragma unroll
    for (int v = 0; v < NVMAX; ++v)
      if (v < nv) gptr(DU[v])[m] = p[v * fs];
  }
}

// FP64 reciprocal for the mixed-precision path: hardware estimate refined
// by two Newton steps (full FP64 accuracy, not correctly rounded).
__device__ __forceinline__ double rcp64(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// Neighbour exchange along the line: up1 gives lane l the value of lane l-1,
// dn1 the value of lane l+1 (the end lanes' results are never used).  DPP
// wavefront shifts (wave_shr:1 = 0x138, wave_shl:1 = 0x130): one VALU move
// per 32-bit half and no LDS round trip (measured 1-3% faster per sweep
// than ds_bpermute shuffles).
__device__ __forceinline__ int wshr1(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x138, 0xf, 0xf, true); }
__device__ __forceinline__ int wshl1(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x130, 0xf, 0xf, true); }
__device__ __forceinline__ double up1(double x) {
  return __hiloint2double(wshr1(__double2hiint(x)), wshr1(__double2loint(x)));
}
__device__ __forceinline__ double dn1(double x) {
  return __hiloint2double(wshl1(__double2hiint(x)), wshl1(__double2loint(x)));
}

// Limited slopes of a cell (interpolation.f90, loop c = -1 .. n) and the
// e_int guard (interpolation.f90:143-163).  qv, ql, qr: d e m1 m2 m3 b1 b2 b3
// eint of the cell and of its left and right neighbours.  SH: divisions by
// the grid denominators shared between the variables (qdiv).
template <bool MAG, int LIM, bool SH>
__device__ __forceinline__ void cell_slopes(const double qv[9], const double ql[9],
                                            const double qr[9], double xm, double x0, double xp,
                                            double xim, double xi0, double ida, double idb,
                                            double S[8]) {
  const double hl = x0 - xim, hr = xi0 - x0;
  // grid factors and denominators shared by every variable of the cell
  const double dxp = xp - x0;
  const double cf = dxp / (xi0 - x0), cb = (x0 - xm) / (x0 - xim);
  const auto Dm = mkden<SH>(LIM == LIM_MINMOD ? ida + idb : x0 - xm);
  const auto Dp = mkden<SH>(dxp);
  double fl[8], fr[8];
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    if (!MAG && s >= 5) {
      S[s] = 0.0;
      fl[s] = fr[s] = 0.0;
      continue;
    }
    const double du = (LIM == LIM_MINMOD)
                          ? minmod_d(ql[s], qv[s], qr[s], ida, idb, Dm)
                          : modified_mc_d(ql[s], qv[s], qr[s], dxp, cf, cb, Dm, Dp);
    S[s] = du;
    fl[s] = qv[s] - hl * du;
    fr[s] = qv[s] + hr * du;
  }
  double eintl = fl[1] - 0.5 * (fl[2] * fl[2] + fl[3] * fl[3] + fl[4] * fl[4]) / fl[0];
  double eintr = fr[1] - 0.5 * (fr[2] * fr[2] + fr[3] * fr[3] + fr[4] * fr[4]) / fr[0];
  if (MAG) {
    eintl = eintl - 0.5 * (fl[5] * fl[5] + fl[6] * fl[6] + fl[7] * fl[7]);
    eintr = eintr - 0.5 * (fr[5] * fr[5] + fr[6] * fr[6] + fr[7] * fr[7]);
  }
  const double el_ = ql[8], em_ = qv[8], er_ = qr[8];
  if (eintl >= mx2(el_, em_) || eintl <= mn2(el_, em_) || eintr >= mx2(em_, er_) ||
      eintr <= mn2(em_, er_)) {
#pragma unroll
    for (int s = 0; s < 8; ++s) S[s] = 0.0;
  }
}

// Face states, EOS, fast speeds and HLLD/GLM flux t[9] of the face between
// the lane's cell (left: L = d e m1 m2 m3 b1 b2 b3 phi, slopes S) and the
// right neighbour (R, RS) -- numflux.f90 with the cyclic rotation normal a =
// DIR, then b, c (numflux.f90:73-93,246-266,419-439).  Returns the face
// densities and internal energies for the positivity check.  SH: divisions
// by dl, dr and the HLLD denominators shared between numerators (qdiv).
template <int DIR, bool MAG, bool SH>
__device__ __forceinline__ void face_flux(const double L[9], const double S[8],
                                          const double R_[9], const double RS[8], double dx1,
                                          double dx2, double gam, double ch, double t[9],
                                          double fe[4]) {
  constexpr int a_ = DIR, b_ = (DIR + 1) % 3, c_ = (DIR + 2) % 3;
  const double dl = L[0] + dx1 * S[0];
  const double dr = R_[0] - dx2 * RS[0];
  const double el = L[1] + dx1 * S[1];
  const double er = R_[1] - dx2 * RS[1];
  const auto Dl = mkden<SH>(dl), Dr = mkden<SH>(dr);  // every division by dl, dr below
  const double v1l = qdiv(Dl, L[2 + a_] + dx1 * S[2 + a_]);
  const double v1r = qdiv(Dr, R_[2 + a_] - dx2 * RS[2 + a_]);
  const double v2l = qdiv(Dl, L[2 + b_] + dx1 * S[2 + b_]);
  const double v2r = qdiv(Dr, R_[2 + b_] - dx2 * RS[2 + b_]);
  const double v3l = qdiv(Dl, L[2 + c_] + dx1 * S[2 + c_]);
  const double v3r = qdiv(Dr, R_[2 + c_] - dx2 * RS[2 + c_]);
  double b1l = 0, b1r = 0, b2l = 0, b2r = 0, b3l = 0, b3r = 0, phil = 0, phir = 0;
  if (MAG) {
    b1l = L[5 + a_] + dx1 * S[5 + a_];
    b1r = R_[5 + a_] - dx2 * RS[5 + a_];
    b2l = L[5 + b_] + dx1 * S[5 + b_];
    b2r = R_[5 + b_] - dx2 * RS[5 + b_];
    b3l = L[5 + c_] + dx1 * S[5 + c_];
    b3r = R_[5 + c_] - dx2 * RS[5 + c_];
    // dphi is identically zero: phi + dx*0 == phi + 0
    phil = L[8] + 0.0;
    phir = R_[8] + 0.0;
  }
  const double eil = get_eint(MAG, el, dl, v1l, v2l, v3l, b1l, b2l, b3l);
  const double eir = get_eint(MAG, er, dr, v1r, v2r, v3r, b1r, b2r, b3r);
  fe[0] = dl;
  fe[1] = eil;
  fe[2] = dr;
  fe[3] = eir;
  // eos_p_cs (eos.f90:195-242); the erad term 4*0/(9d) adds +0 to a
  // positive cs2 and is omitted.
  double ptl = (gam - 1.0) * eil;
  double ptr = (gam - 1.0) * eir;
  const double csl = sqrt(qdiv(Dl, gam * ptl));
  const double csr = sqrt(qdiv(Dr, gam * ptr));
  if (MAG) {
    ptl = ptl + 0.5 * (b1l * b1l + b2l * b2l + b3l * b3l);
    ptr = ptr + 0.5 * (b1r * b1r + b2r * b2r + b3r * b3r);
  }
  const double cfl = get_cf_d(Dl, csl, b1l, b2l, b3l);
  const double cfr = get_cf_d(Dr, csr, b1r, b2r, b3r);
  hlld_d<SH>(t, cfl, cfr, v1l, v1r, v2l, v2r, v3l, v3r, dl, dr, el, er, ptl, ptr, b1l, b1r,
             b2l, b2r, b3l, b3r, phil, phir, ch);
}

// Mass flux of the face c+1/2 for the species pass (compswitch 2): FM[d]
// holds n+1 faces per line, face c+1/2 at line*(n+1) + c+1 (c = -1..n-1).
// Lanes 1..61 cover their segment's faces; the face two segments share is
// written twice with the same value.
template <int DIR>
__device__ __forceinline__ void store_mass_flux(const Dev &D, long line, int c, int lane,
                                                double f0) {
  const int n = D.ax[DIR].n;
  if (lane >= 1 && lane <= 61 && c >= -1 && c <= n - 1)
    D.FM[DIR][line * (n + 1) + (c + 1)] = f0;
}

// Flux boundary (scalar code 10): the reference skips the Riemann solve of
// the boundary face (numflux.f90:59-62, :232-235, :405-408) and takes the
// flux fluxboundary sets there (fluxboundary.f90:66-78; the flux array
// keeps its zero initial value at the faces it does not set).  Face c+1/2
// of the line is such a face:
__device__ __forceinline__ bool flux_face(const Axis &A, int c) {
  return (c == -1 && A.sc_lo == 10) || (c == A.n - 1 && A.sc_hi == 10);
}

// the flux given for the face (conserved-variable order) in the sweep's
// face-local order
template <int DIR>
__device__ __forceinline__ void boundary_flux(const Axis &A, int c, double t[9]) {
  const double *F = A.FX[c < 0 ? 0 : 1];
  constexpr int a_ = DIR, b_ = (DIR + 1) % 3, c_ = (DIR + 2) % 3;
  t[0] = F[0];
  t[1] = F[1 + a_];
  t[2] = F[1 + b_];
  t[3] = F[1 + c_];
  t[4] = F[4];
  t[5] = F[5 + a_];
  t[6] = F[5 + b_];
  t[7] = F[5 + c_];
  t[8] = F[8];
}

// RK update of the cells a fused sweep has just finished (defined with the RK
// kernels below)
template <int DIR, bool MAG, bool TILE>
__device__ __forceinline__ void fused_rk(const Dev &D, const Axis &A, long base, int c,
                                         int lane, const double du[NVMAX], double *lds,
                                         int stage, const RKArgs &rk, double &dti,
                                         double &cfm);

template <int DIR, bool MAG, int LIM, bool TILE = false, bool SPC = false, bool FUSED = false,
          bool EXT = true>
__device__ __forceinline__ void sweep_line(const Dev &D, long wave, int lane, int stage,
                                           double *lds = nullptr, const RKArgs *rk = nullptr,
                                           double *dti = nullptr, double *cfm = nullptr) {
  const Axis &A = D.ax[DIR];
  const int n = A.n;
  // a tiled block's padding wavefronts (past the direction's last wave) take
  // part in the block's loads and barriers on a copy of the last wave and
  // store and report nothing
  const bool valid = !TILE || wave < D.waves[DIR];
  const long wave0 = wave - wave_in_block();  // first wave of the block (tiled)
  if (!valid) wave = D.waves[DIR] - 1;
  long line, base;
  int c;
  lane_cell<DIR>(D, A, wave, lane, line, base, c);
  const Cell q = TILE ? tile_cell<DIR, MAG, EXT>(D, A, wave0, base, c, valid, lds)
                      : load_cell<DIR, MAG, EXT>(D, A, base, line, c);
  if (FUSED && !TILE) park_cell<MAG>(lds, q);
  // d e m1 m2 m3 b1 b2 b3 and eint (slopes) / phi (face) of the cell
  const double qv[9] = {q.d, q.e, q.d * q.v[0], q.d * q.v[1], q.d * q.v[2],
                        q.b[0], q.b[1], q.b[2], q.eint};

  // ---- slopes of cell c (interpolation.f90, loop c = -1 .. n) -----------
  double S[8];  // dd de dm1 dm2 dm3 db1 db2 db3 (dphi is always 0)
  const int cc = c < -2 ? -2 : (c > n + 1 ? n + 1 : c);
  if (LIM == LIM_FLAT) {
#pragma unroll
    for (int s = 0; s < 8; ++s) S[s] = 0.0;
  } else {
    double ql[9], qr[9];
#pragma unroll
    for (int s = 0; s < 9; ++s) {
      ql[s] = up1(qv[s]);
      qr[s] = dn1(qv[s]);
    }
    const int cl = cc > -1 ? cc : -1, cr = cc < n ? cc : n;  // keep indices in range
    const double xm = gptr(A.x)[cl - 1], x0 = gptr(A.x)[cl], xp = gptr(A.x)[cr + 1];
    const double xim = gptr(A.xi)[cl - 1], xi0 = gptr(A.xi)[cl];
    const double ida = gptr(A.idx)[cl], idb = gptr(A.idx)[cr + 1];
    cell_slopes<MAG, LIM, true>(qv, ql, qr, xm, x0, xp, xim, xi0, ida, idb, S);
  }

  // ---- face c+1/2 (numflux.f90, faces c = -1 .. n-1) ---------------------
  // right cell's primitives and slopes from lane+1
  double L[9], R_[9];
#pragma unroll
  for (int s = 0; s < 8; ++s) L[s] = qv[s];
  L[8] = q.phi;
#pragma unroll
  for (int s = 0; s < 9; ++s) R_[s] = (!MAG && s >= 5) ? 0.0 : dn1(L[s]);
  double RS[8];
#pragma unroll
  for (int s = 0; s < 8; ++s) RS[s] = (!MAG && s >= 5) ? 0.0 : dn1(S[s]);

  const int cf_ = cc < n - 1 ? (cc > -1 ? cc : -1) : n - 1;  // clamp face index
  const double dx1 = gptr(A.xi)[cf_] - gptr(A.x)[cf_], dx2 = gptr(A.x)[cf_ + 1] - gptr(A.xi)[cf_];
  const double gam = D.gamma, ch = gptr(D.sc)->ch;
  double t[9], fe[4];
  face_flux<DIR, MAG, true>(L, S, R_, RS, dx1, dx2, gam, ch, t, fe);
  const bool fxb = EXT && A.ext && flux_face(A, c);
  if (fxb) boundary_flux<DIR>(A, c, t);
  // faces between lanes l and l+1 have valid slopes on both sides for l = 1..61
  const bool face_valid = valid && (c >= -1 && c <= n - 1) && lane >= 1 && lane <= 61 && !fxb;
  if (face_valid && (fe[1] <= 0.0 || fe[0] <= 0.0 || fe[3] <= 0.0 || fe[2] <= 0.0)) {
    int ii, jj, kk;
    line_origin<DIR>(D, line, ii, jj, kk);
    if (DIR == 0) ii = c; else if (DIR == 1) jj = c; else kk = c;
    atomicMin(&gptr(D.sc)->err, errkey(stage, 0, DIR + 1, ii, jj, kk, D));
  }

  // ---- flux difference of cell c (rungekutta.f90:42-55, one direction) --
  double tm[9];
#pragma unroll
  for (int s = 0; s < 9; ++s) tm[s] = up1(t[s]);
  // equation order: icnt imo1 imo2 imo3 iene img1 img2 img3 i9wv
  constexpr int a_ = DIR, b_ = (DIR + 1) % 3, c_ = (DIR + 2) % 3;
  const double ix = gptr(A.idxi)[cc];
  double du[NVMAX];
  du[0] = ix * (tm[0] - t[0]);
  du[1 + a_] = ix * (tm[1] - t[1]);
  du[1 + b_] = ix * (tm[2] - t[2]);
  du[1 + c_] = ix * (tm[3] - t[3]);
  du[4] = ix * (tm[4] - t[4]);
  du[5 + a_] = ix * (tm[5] - t[5]);
  du[5 + b_] = ix * (tm[6] - t[6]);
  du[5 + c_] = ix * (tm[7] - t[7]);
  du[8] = ix * (tm[8] - t[8]);
  if (FUSED) {
    fused_rk<DIR, MAG, TILE>(D, A, base, c, lane, du, lds, stage, *rk, *dti, *cfm);
    return;
  }
  store_du<DIR, MAG, TILE>(D, A, base, c, lane, du, lds);
  if (SPC) store_mass_flux<DIR>(D, line, c, lane, t[0]);
}

// One launch covers every solved direction: wavefronts [0, waves[0]) sweep
// x, then y from woff[1], then z from woff[2] (the y and z ranges start on a
// multiple of 16 wavefronts, so a block never mixes directions).  Blocks are
// dispatched in order, so a compute unit mostly runs one direction's code at
// a time.  TILE: the y/z lines of a block go through the LDS tile (the
// launcher checks that a block's lines have consecutive x index); a
// separate instantiation keeps each kernel's code within the instruction
// cache.
template <bool MAG, int LIM, bool TILE, bool SPC = false, bool EXT = false>
__global__ __launch_bounds__(1024) void k_sweep(Dev D, int stage) {
  extern __shared__ double lds[];
  if (gptr(D.sc)->stop) return;
  const int lane = threadIdx.x & 63;
  const long wave = (long)blockIdx.x * (blockDim.x >> 6) + wave_in_block();
  if (wave < D.woff[1]) {
    if (wave < D.waves[0]) sweep_line<0, MAG, LIM, false, SPC, false, EXT>(D, wave, lane, stage);
    return;
  }
  // tiled: a block past the direction's last wave has nothing to do; a
  // partly filled one runs its padding wavefronts too (barriers)
  if (wave < D.woff[2]) {
    const long w1 = wave - D.woff[1], w0 = w1 - (threadIdx.x >> 6);
    if (TILE ? w0 >= D.waves[1] : w1 >= D.waves[1]) return;
    sweep_line<1, MAG, LIM, TILE, SPC, false, EXT>(D, w1, lane, stage, lds);
    return;
  }
  {
    const long w2 = wave - D.woff[2], w0 = w2 - (threadIdx.x >> 6);
    if (TILE ? w0 >= D.waves[2] : w2 >= D.waves[2]) return;
    sweep_line<2, MAG, LIM, TILE, SPC, false, EXT>(D, w2, lane, stage, lds);
  }
}

// ---- FP32-mixed sweep -------------------------------------------------------
// Same mapping and stencil as sweep_line.  The Riemann solver (fast speeds,
// star states, state selection) runs in FP32 in perturbation form
// (hlld_delta), with the hardware reciprocal and square root.  Everything
// whose FP64 form costs no more than FP32 on this hardware (additions and
// multiplications run at the same rate in both when not packed) stays FP64:
// the cell state, the limited slopes (with the limiter divisions replaced by
// per-cell geometry factors), the face states, the face internal energy
// (the subtraction of kinetic and magnetic energy from the total energy), the
// e_int guard, the base flux and the flux-divergence accumulation.  Keeping
// the limiter and guard decisions in FP64 matters: a decision that flips
// changes a face state by a whole slope, and such flips dominated the error
// of an all-FP32 reconstruction.

// Modified MC limiter (fluxlimiter.f90:58-75) on the one-sided differences
// dm = u2-u1, dp = u3-u2 with g = (cf, cb, (x3-x2)/(x2-x1), 1/(x3-x2)).
__device__ __forceinline__ double mc_geo(double dm, double dp, double4 g) {
  const double top = dm * g.z;
  const double bot = dp;
  const double sig = sgn(1.0, bot);
  return sig * mx2(0.0, mn3(sig * 0.5 * (bot + top), sig * g.x * bot, sig * g.y * top)) * g.w;
}

// Generalised minmod (fluxlimiter.f90:40-56); w = ida*idb/(ida+idb).
__device__ __forceinline__ double minmod_geo(double dm, double dp, double ida, double idb,
                                             double w) {
  const double x = 1.4 * dp * idb;
  const double y = 1.4 * dm * ida;
  const double z = (dp + dm) * w;
  const double s = sgn(1.0, x);
  return s * mx2(0.0, mn3(fabs(x), s * y, s * z));
}

template <int DIR, bool MAG, int LIM, bool TILE = false, bool SPC = false, bool EXT = true,
          bool FUSED = false>
__device__ __forceinline__ void sweep_line_mix(const Dev &D, long wave, int lane, int stage,
                                               double *lds = nullptr, const RKArgs *rk = nullptr,
                                               double *dti = nullptr, double *cfm = nullptr) {
  const Axis &A = D.ax[DIR];
  const int n = A.n;
  // a tiled block's padding wavefronts (past the direction's last wave) take
  // part in the block's loads and barriers on a copy of the last wave and
  // store and report nothing
  const bool valid = !TILE || wave < D.waves[DIR];
  const long wave0 = wave - wave_in_block();  // first wave of the block (tiled)
  if (!valid) wave = D.waves[DIR] - 1;
  long line, base;
  int c;
  lane_cell<DIR>(D, A, wave, lane, line, base, c);
  const Cell q = TILE ? tile_cell<DIR, MAG, EXT>(D, A, wave0, base, c, valid, lds)
                      : load_cell<DIR, MAG, EXT>(D, A, base, line, c);
  if (FUSED && !TILE) park_cell<MAG>(lds, q);
  const double m0 = q.d * q.v[0], m1 = q.d * q.v[1], m2 = q.d * q.v[2];
  const double qd[8] = {q.d, q.e, m0, m1, m2, q.b[0], q.b[1], q.b[2]};
  const int cc = c < -2 ? -2 : (c > n + 1 ? n + 1 : c);

  double S[8];
  if (LIM == LIM_FLAT) {
#pragma unroll
    for (int s = 0; s < 8; ++s) S[s] = 0.0;
  } else {
    const int cl = cc > -1 ? cc : -1, cr = cc < n ? cc : n;
    const double4 geo = gptr(A.limg)[cl];
    const double ida = gptr(A.idx)[cl], idb = gptr(A.idx)[cr + 1], mw = gptr(A.mmw)[cl];
    const double hl = gptr(A.x)[cl] - gptr(A.xi)[cl - 1], hr = gptr(A.xi)[cl] - gptr(A.x)[cl];
    double fl[8], fr[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      if (!MAG && s >= 5) {
        S[s] = 0.0;
        fl[s] = fr[s] = 0.0;
        continue;
      }
      const double u0 = up1(qd[s]), u2 = dn1(qd[s]);
      const double dm = qd[s] - u0, dp = u2 - qd[s];
      const double du = (LIM == LIM_MINMOD) ? minmod_geo(dm, dp, ida, idb, mw) : mc_geo(dm, dp, geo);
      S[s] = du;
      fl[s] = qd[s] - hl * du;
      fr[s] = qd[s] + hr * du;
    }
    // e_int guard (interpolation.f90:143-163)
    double eintl = fl[1] - 0.5 * (fl[2] * fl[2] + fl[3] * fl[3] + fl[4] * fl[4]) * rcp64(fl[0]);
    double eintr = fr[1] - 0.5 * (fr[2] * fr[2] + fr[3] * fr[3] + fr[4] * fr[4]) * rcp64(fr[0]);
    if (MAG) {
      eintl = eintl - 0.5 * (fl[5] * fl[5] + fl[6] * fl[6] + fl[7] * fl[7]);
      eintr = eintr - 0.5 * (fr[5] * fr[5] + fr[6] * fr[6] + fr[7] * fr[7]);
    }
    const double el_ = up1(q.eint), em_ = q.eint, er_ = dn1(q.eint);
    if (eintl >= mx2(el_, em_) || eintl <= mn2(el_, em_) || eintr >= mx2(em_, er_) ||
        eintr <= mn2(em_, er_)) {
#pragma unroll
      for (int s = 0; s < 8; ++s) S[s] = 0.0;
    }
  }

  // face c+1/2: right cell from lane+1
  double rq[8], RS[8];
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    rq[s] = (!MAG && s >= 5) ? 0.0 : dn1(qd[s]);
    RS[s] = (!MAG && s >= 5) ? 0.0 : dn1(S[s]);
  }
  const double rphi = MAG ? dn1(q.phi) : 0.0;
  const int cf_ = cc < n - 1 ? (cc > -1 ? cc : -1) : n - 1;
  const double dx1 = gptr(A.xi)[cf_] - gptr(A.x)[cf_], dx2 = gptr(A.x)[cf_ + 1] - gptr(A.xi)[cf_];
  constexpr int a_ = DIR, b_ = (DIR + 1) % 3, c_ = (DIR + 2) % 3;
  double L[8], R_[8];
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    L[s] = qd[s] + dx1 * S[s];
    R_[s] = rq[s] - dx2 * RS[s];
  }
  // face primitives in FP64: velocities, internal energy (the subtraction
  // of kinetic and magnetic energy from the total energy), total pressure
  const double idl = rcp64(L[0]), idr = rcp64(R_[0]);
  const double vl[3] = {L[2 + a_] * idl, L[2 + b_] * idl, L[2 + c_] * idl};
  const double vr[3] = {R_[2 + a_] * idr, R_[2 + b_] * idr, R_[2 + c_] * idr};
  double eil = L[1] - 0.5 * L[0] * (vl[0] * vl[0] + vl[1] * vl[1] + vl[2] * vl[2]);
  double eir = R_[1] - 0.5 * R_[0] * (vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]);
  double bbl = 0.0, bbr = 0.0;
  if (MAG) {
    bbl = 0.5 * (L[5] * L[5] + L[6] * L[6] + L[7] * L[7]);
    bbr = 0.5 * (R_[5] * R_[5] + R_[6] * R_[6] + R_[7] * R_[7]);
    eil = eil - bbl;
    eir = eir - bbr;
  }
  const bool fxb = EXT && A.ext && flux_face(A, c);
  const bool face_valid = valid && (c >= -1 && c <= n - 1) && lane >= 1 && lane <= 61 && !fxb;
  if (face_valid && (eil <= 0.0 || L[0] <= 0.0 || eir <= 0.0 || R_[0] <= 0.0)) {
    int ii, jj, kk;
    line_origin<DIR>(D, line, ii, jj, kk);
    if (DIR == 0) ii = c; else if (DIR == 1) jj = c; else kk = c;
    atomicMin(&gptr(D.sc)->err, errkey(stage, 0, DIR + 1, ii, jj, kk, D));
  }
  const double gm1 = D.gamma - 1.0;
  const double pl = gm1 * eil, pr = gm1 * eir;
  // fast speeds in FP32 (eos_p_cs, get_cf)
  const float gam = (float)D.gamma;
  const float dlf = (float)L[0], drf = (float)R_[0];
  const float csl = rsqrt_(gam * (float)pl * __builtin_amdgcn_rcpf(dlf));
  const float csr = rsqrt_(gam * (float)pr * __builtin_amdgcn_rcpf(drf));
  float cfl, cfr;
  if (MAG) {
    cfl = get_cf(dlf, csl, (float)L[5 + a_], (float)L[5 + b_], (float)L[5 + c_]);
    cfr = get_cf(drf, csr, (float)R_[5 + a_], (float)R_[5 + b_], (float)R_[5 + c_]);
  } else {
    cfl = csl;  // get_cf with B = 0 (exactly sqrt(0.5*(cs^2 + |cs^2|)))
    cfr = csr;
  }
  // GLM face values (hlldflux.f90:37-38) in FP64
  const double ch = gptr(D.sc)->ch;
  double b1 = 0.0, phi = 0.0;
  if (MAG) {
    b1 = 0.5 * ((L[5 + a_] + R_[5 + a_]) - (rphi - q.phi) * rcp64(ch));
    phi = 0.5 * ((q.phi + rphi) - (R_[5 + a_] - L[5 + a_]) * ch);
  }
  const double fl[8] = {L[0], vl[0], vl[1], vl[2], MAG ? L[5 + b_] : 0.0,
                        MAG ? L[5 + c_] : 0.0, L[1], pl + bbl};
  const double fr[8] = {R_[0], vr[0], vr[1], vr[2], MAG ? R_[5 + b_] : 0.0,
                        MAG ? R_[5 + c_] : 0.0, R_[1], pr + bbr};
  double t[9];
  hlld_delta(t, fl, fr, b1, cfl, cfr);
  t[5] = phi;
  t[8] = ch * ch * b1;
  if (fxb) boundary_flux<DIR>(A, c, t);

  // flux difference in FP64
  double tm[9];
#pragma unroll
  for (int s = 0; s < 9; ++s) tm[s] = (!MAG && s >= 5) ? 0.0 : up1(t[s]);
  // equation order: icnt imo1 imo2 imo3 iene img1 img2 img3 i9wv
  const double ix = gptr(A.idxi)[cc];
  double du[NVMAX];
  du[0] = ix * (tm[0] - t[0]);
  du[1 + a_] = ix * (tm[1] - t[1]);
  du[1 + b_] = ix * (tm[2] - t[2]);
  du[1 + c_] = ix * (tm[3] - t[3]);
  du[4] = ix * (tm[4] - t[4]);
  du[5 + a_] = ix * (tm[5] - t[5]);
  du[5 + b_] = ix * (tm[6] - t[6]);
  du[5 + c_] = ix * (tm[7] - t[7]);
  du[8] = ix * (tm[8] - t[8]);
  if (FUSED) {
    fused_rk<DIR, MAG, TILE>(D, A, base, c, lane, du, lds, stage, *rk, *dti, *cfm);
    return;
  }
  store_du<DIR, MAG, TILE>(D, A, base, c, lane, du, lds);
  if (SPC) store_mass_flux<DIR>(D, line, c, lane, t[0]);
}

template <bool MAG, int LIM, bool TILE, bool SPC = false, bool EXT = false>
__global__ __launch_bounds__(1024) void k_sweep_mix(Dev D, int stage) {
  extern __shared__ double lds[];
  if (gptr(D.sc)->stop) return;
  const int lane = threadIdx.x & 63;
  const long wave = (long)blockIdx.x * (blockDim.x >> 6) + wave_in_block();
  if (wave < D.woff[1]) {
    if (wave < D.waves[0]) sweep_line_mix<0, MAG, LIM, false, SPC, EXT>(D, wave, lane, stage);
    return;
  }
  if (wave < D.woff[2]) {
    const long w1 = wave - D.woff[1], w0 = w1 - (threadIdx.x >> 6);
    if (TILE ? w0 >= D.waves[1] : w1 >= D.waves[1]) return;
    sweep_line_mix<1, MAG, LIM, TILE, SPC, EXT>(D, w1, lane, stage, lds);
    return;
  }
  {
    const long w2 = wave - D.woff[2], w0 = w2 - (threadIdx.x >> 6);
    if (TILE ? w0 >= D.waves[2] : w2 >= D.waves[2]) return;
    sweep_line_mix<2, MAG, LIM, TILE, SPC, EXT>(D, w2, lane, stage, lds);
  }
}

// ---- one direction per launch ---------------------------------------------
// The same sweep as k_sweep / k_sweep_mix for one direction only (no
// species, no boundary codes 4/5/9/10: the fused forms that use it refuse
// those).  One direction's code is about a third of the all-direction
// kernel's, well inside the instruction cache.  The y and z blocks are
// renumbered XCD-aware: neighbouring blocks hold neighbouring tile columns,
// whose rows share 128-byte lines (a row of an 8-wavefront tile is 64
// bytes), so on one XCD the second half of a line is an L2 hit instead of a
// second fetch by another XCD.  Measured on the 256^3 MHD polytrope: y+z
// sweep 2244 -> 2052 us; the same renumbering of the streaming x blocks (the
// fused x sweep) and of the gravity sub-steps made them 3-6% slower.
template <int DIR, bool MAG, int LIM, bool TILE, bool MIX>
__global__ __launch_bounds__(1024) void k_sweep_dir(Dev D, int stage) {
  extern __shared__ double lds[];
  if (gptr(D.sc)->stop) return;
  const int lane = threadIdx.x & 63;
  const unsigned b = DIR == 0 ? blockIdx.x : xcd_block(blockIdx.x, gridDim.x);
  const long w = (long)b * (blockDim.x >> 6) + wave_in_block();
  // tiled: a partly filled last block runs its padding wavefronts too
  if (TILE ? w - (threadIdx.x >> 6) >= D.waves[DIR] : w >= D.waves[DIR]) return;
  if (MIX) sweep_line_mix<DIR, MAG, LIM, TILE, false, false>(D, w, lane, stage, lds);
  else sweep_line<DIR, MAG, LIM, TILE, false, false, false>(D, w, lane, stage, lds);
}

// ---- advected species (compswitch 2) -----------------------------------------
// One wavefront per 64-cell segment of a line, as the sweeps: the species'
// limited slopes (interpolation.f90:135-141; the e_int guard does not reset
// them), face values, upwinding with the face's mass flux (stored by the
// sweep) and renormalisation by the upwind sum (numflux.f90:200-212), and
// the flux divergence of each cell, idxi*(F(c-1/2) - F(c+1/2))
// (rungekutta.f90:66-77 with detg = 1).  FP64 in both precision modes.
template <int DIR, int LIM>
__device__ __forceinline__ void species_line(const Dev &D, long wave, int lane) {
  const Axis &A = D.ax[DIR];
  const int n = A.n;
  long line, base;
  int c;
  lane_cell<DIR>(D, A, wave, lane, line, base, c);
  const int cc = c < -2 ? -2 : (c > n + 1 ? n + 1 : c);
  long ms, mv;
  cell_sources(A, base, c, ms, mv);
  const int cl = cc > -1 ? cc : -1, cr = cc < n ? cc : n;  // as the hydro slopes
  const double xm = gptr(A.x)[cl - 1], x0 = gptr(A.x)[cl], xp = gptr(A.x)[cr + 1];
  const double xim = gptr(A.xi)[cl - 1], xi0 = gptr(A.xi)[cl];
  const double ida = gptr(A.idx)[cl], idb = gptr(A.idx)[cr + 1];
  const double dxp = xp - x0;
  const double cf = dxp / (xi0 - x0), cb = (x0 - xm) / (x0 - xim);
  const auto Dm = mkden<true>(LIM == LIM_MINMOD ? ida + idb : x0 - xm);
  const auto Dp = mkden<true>(dxp);
  const int cf_ = cc < n - 1 ? (cc > -1 ? cc : -1) : n - 1;  // face index, as the sweep
  const double dx1 = gptr(A.xi)[cf_] - gptr(A.x)[cf_], dx2 = gptr(A.x)[cf_ + 1] - gptr(A.xi)[cf_];
  const long N = D.ncell;
  double sl[HG_SPN_MAX], sr[HG_SPN_MAX];
  double suml = 0.0, sumr = 0.0;  // sum(spcl(1:spn)), sum(spcr(1:spn))
#pragma unroll
  for (int q = 0; q < HG_SPN_MAX; ++q) {
    if (q >= D.nspc) break;
    const double v = D.SPC[q * N + ms];
    double sp = 0.0;
    if (LIM != LIM_FLAT) {
      const double vl = up1(v), vr = dn1(v);
      sp = LIM == LIM_MINMOD ? minmod_d(vl, v, vr, ida, idb, Dm)
                             : modified_mc_d(vl, v, vr, dxp, cf, cb, Dm, Dp);
    }
    sl[q] = v + dx1 * sp;
    sr[q] = dn1(v) - dx2 * dn1(sp);
    suml = suml + sl[q];
    sumr = sumr + sr[q];
  }
  const bool face = c >= -1 && c <= n - 1;
  const double f0 = face ? D.FM[DIR][line * (n + 1) + (c + 1)] : 0.0;
  const double sg = sgn(0.5, f0);
  const double fix = 1.0 / (suml * (0.5 + sg) + sumr * (0.5 - sg));
  const double ix = gptr(A.idxi)[cc];
  const bool own = lane >= 2 && lane <= 61 && c >= 0 && c < n;
  const long m = base + (long)c * A.stride;
#pragma unroll
  for (int q = 0; q < HG_SPN_MAX; ++q) {
    if (q >= D.nspc) break;
    const double fl = fix * f0 * (sl[q] * (0.5 + sg) + sr[q] * (0.5 - sg));
    const double dv = ix * (up1(fl) - fl);
    if (own) D.DUS[((long)DIR * D.nspc + q) * N + m] = dv;
  }
}

template <int LIM>
__global__ __launch_bounds__(256) void k_species(Dev D) {
  if (gptr(D.sc)->stop) return;
  const int lane = threadIdx.x & 63;
  const long wave = (long)blockIdx.x * (blockDim.x >> 6) + wave_in_block();
  if (wave < D.woff[1]) {
    if (wave < D.waves[0]) species_line<0, LIM>(D, wave, lane);
  } else if (wave < D.woff[2]) {
    if (wave - D.woff[1] < D.waves[1]) species_line<1, LIM>(D, wave - D.woff[1], lane);
  } else if (wave - D.woff[2] < D.waves[2]) {
    species_line<2, LIM>(D, wave - D.woff[2], lane);
  }
}

// ---- CFL condition of one cell (timestep.f90:35-57, dti_cell :125-183) ----
template <bool MAG>
__device__ __forceinline__ void cfl_cell(const Dev &D, long m, double d, double eint,
                                         double v1, double v2, double v3, double b1, double b2,
                                         double b3, double &dti, double &cfm) {
  int i, j, k;
  cell_ijk(D, m, i, j, k);
  const double gam = D.gamma;
  const double p = (gam - 1.0) * eint;
  const double cs = sqrt(gam * p / d);
  double cf1, cf2, cf3;
  if (MAG) {
    cf1 = get_cf(d, cs, b1, b2, b3) + fabs(v1);
    cf2 = get_cf(d, cs, b2, b1, b3) + fabs(v2);
    cf3 = get_cf(d, cs, b3, b1, b2) + fabs(v3);
  } else {
    cf1 = cs + fabs(v1);
    cf2 = cs + fabs(v2);
    cf3 = cs + fabs(v3);
  }
  const double x1 = D.ax[0].dxi[i], x2 = D.ax[1].dxi[j], x3 = D.ax[2].dxi[k];
  const double dvol = x1 * x2 * x3;  // metric.f90:44-47
  const double sa1 = x2 * x3, sa2 = x1 * x3, sa3 = x1 * x2;
  const double o1 = D.solve[0] ? 1.0 : 0.0, o2 = D.solve[1] ? 1.0 : 0.0,
               o3 = D.solve[2] ? 1.0 : 0.0;
  dti = dvol / (cf1 * o1 * (sa1 + sa1) + cf2 * o2 * (sa2 + sa2) + cf3 * o3 * (sa3 + sa3));
  cfm = mx3(cf1, cf2, cf3);
  if (eint <= 0.0 || d <= 0.0) atomicMin(&gptr(D.sc)->err, errkey(0, 1, 0, i, j, k, D));
}

__device__ __forceinline__ double wave_min(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = mn2(x, __shfl_xor(x, o, 64));
  return x;
}
__device__ __forceinline__ double wave_max(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = mx2(x, __shfl_xor(x, o, 64));
  return x;
}

// block (up to 1024 threads) reduction of (min dti, max cf) into D.part[blk]
__device__ __forceinline__ void cfl_block_reduce(const Dev &D, double dti, double cfm,
                                                 long blk) {
  __shared__ double smin[16], smax[16];
  dti = wave_min(dti);
  cfm = wave_max(cfm);
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    smin[w] = dti;
    smax[w] = cfm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = smin[0], b = smax[0];
    for (int q = 1; q < nw; ++q) {
      a = mn2(a, smin[q]);
      b = mx2(b, smax[q]);
    }
    gptr(D.part)[2 * blk] = a;
    gptr(D.part)[2 * blk + 1] = b;
  }
}

// ---- RK update + primitive recovery ------------------------------------
// CFL: also evaluate the next step's CFL condition from the new primitives
// (the reference's timestep reads exactly these values at the start of the
// next step).  DAMP: apply phidamp (source.f90:239-260) to the new psi.
// Gravity source (source.f90:14-85 Cartesian, get_fieldforce :177-228):
// g_a = -(non-uniform central difference of phi) * rho with the stage-start
// rho and v; src(imo_a) = g_a, src(iene) = g.v.
// qc: the cell's current primitives when the caller holds them (the fused
// launch's lane path), else read from W.
__device__ __forceinline__ void grav_source(const Dev &D, long m, double src[NVMAX],
                                            const Cell *qc = nullptr) {
  int i, j, k;
  cell_ijk(D, m, i, j, k);
  const double *ph = D.gphi[D.gcur];
  const long g = gidx(D, i, j, k);
  const long st[3] = {1, (long)(D.n1 + 2), (long)(D.n1 + 2) * (D.n2 + 2)};
  const int q3[3] = {i, j, k};
  const double d = qc ? qc->d : gptr(D.W[WD])[m];
  double gr[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int q = q3[a];
    const double *DX = D.ax[a].dx, *IDX = D.ax[a].idx;
    gr[a] = -((DX[q] * IDX[q + 1] * ph[g + st[a]] - DX[q + 1] * IDX[q] * ph[g - st[a]]) /
                  (DX[q] + DX[q + 1]) +
              (DX[q + 1] - DX[q]) * IDX[q] * IDX[q + 1] * ph[g]) *
            d;
  }
#pragma unroll
  for (int v = 0; v < NVMAX; ++v) src[v] = 0.0;
  src[1] = gr[0];
  src[2] = gr[1];
  src[3] = gr[2];
  if (qc) src[4] = gr[0] * qc->v[0] + gr[1] * qc->v[1] + gr[2] * qc->v[2];
  else src[4] = gr[0] * gptr(D.W[WV1])[m] + gr[1] * gptr(D.W[WV2])[m] + gr[2] * gptr(D.W[WV3])[m];
}

// LD >= 0: direction LD's flux divergence comes from duL[] (the fused last
// sweep) instead of memory.  The flags are plain arguments so that one
// inlined copy serves every variant; k_rk passes constants.
// qc: the cell's current primitives (parked_cell) when the caller holds
// them: d, e and B are then not read from U again.
template <bool MAG, int LD = -1>
__device__ __forceinline__ void rk_cell_rt(const Dev &D, long m, int stage, double faco,
                                           double fact, double facn, double damp, bool CFL,
                                           bool DAMP, bool GRAV, double &dti, double &cfm,
                                           const double *duL = nullptr,
                                           const Cell *qc = nullptr) {
  // the fused form writes the new state to the other buffer set (the
  // launch's neighbours still read this one); k_rk updates in place
  double *const *Uo = LD >= 0 ? D.Uo : D.U;
  double *const *Wo = LD >= 0 ? D.Wo : D.W;
  const double ft = fact * gptr(D.sc)->dt;
  double src[NVMAX];
  if (GRAV) grav_source(D, m, src, qc);
  double un[NVMAX];
  double dold = 0.0;  // stage-start density (primitive d = u(icnt))
  // a bound known at compile time for hydro, so that the loop unrolls and
  // un[] stays in registers
  const int nv = MAG ? D.nv : 5;
#pragma unroll
  for (int v = 0; v < NVMAX; ++v) {
    un[v] = 0.0;
    if (v >= nv) continue;
    // inactive directions contribute idxi*(0-0) = +0 (allocations.f90:214-240)
    const double a = LD == 0 ? duL[v] : (D.solve[0] ? gptr(D.DU[0][v])[m] : 0.0);
    const double b = LD == 1 ? duL[v] : (D.solve[1] ? gptr(D.DU[1][v])[m] : 0.0);
    const double c = LD == 2 ? duL[v] : (D.solve[2] ? gptr(D.DU[2][v])[m] : 0.0);
    const double acc = ((a + b) + c) + (GRAV ? src[v] : 0.0);  // + src
    double u;
    if (qc && v == 0) u = qc->d;
    else if (qc && v == 4) u = qc->e;
    else if (qc && MAG && v >= 5 && v <= 7) u = qc->b[v - 5];
    else u = gptr(D.U[v])[m];
    double u0;
    if (stage == 1) {
      u0 = u;
      // (no U0: the fused form's three state sets keep the step's starting
      // values in their own set)
      if (D.U0[v]) gptr(D.U0[v])[m] = u;
    } else {
      u0 = gptr(D.U0[v])[m];
    }
    un[v] = faco * u0 + facn * u + ft * acc;
    gptr(Uo[v])[m] = un[v];
    if (v == 0) dold = u;
  }
  // species (rungekutta.f90:61-85): spcorg = spc*uorg(icnt) at stage 1;
  // spc = (faco*spcorg + facn*spc*d + fact*dt*(divergences)) / u(icnt)
  const long N = D.ncell;
#pragma unroll
  for (int q = 0; q < HG_SPN_MAX; ++q) {
    if (q >= D.nspc) break;
    const double x = D.SPC[q * N + m];
    double x0;
    if (stage == 1) {
      x0 = x * dold;
      D.SPC0[q * N + m] = x0;
    } else {
      x0 = D.SPC0[q * N + m];
    }
    const double a = D.solve[0] ? D.DUS[(0L * D.nspc + q) * N + m] : 0.0;
    const double b = D.solve[1] ? D.DUS[(1L * D.nspc + q) * N + m] : 0.0;
    const double cz = D.solve[2] ? D.DUS[(2L * D.nspc + q) * N + m] : 0.0;
    D.SPC[q * N + m] = (faco * x0 + facn * x * dold + ft * ((a + b) + cz)) / un[0];
  }
  // primitive (rungekutta.f90:193-232) + internalenergy (eos.f90:805-842)
  const double d = un[0];
  const double v1 = un[1] / d, v2 = un[2] / d, v3 = un[3] / d;
  const double e = un[4];
  // W[d], W[e] and W[B] are U[0], U[4] and U[5..7] (same buffers)
  gptr(Wo[WV1])[m] = v1;
  gptr(Wo[WV2])[m] = v2;
  gptr(Wo[WV3])[m] = v3;
  double b1 = 0, b2 = 0, b3 = 0;
  if (MAG) {
    b1 = un[5];
    b2 = un[6];
    b3 = un[7];
    if (D.i9wv >= 0) gptr(Wo[WPHI])[m] = DAMP ? un[8] * damp : un[8];
  }
  const double eint = get_eint(MAG, e, d, v1, v2, v3, b1, b2, b3);
  gptr(Wo[WEINT])[m] = eint;
  if (eint <= 0.0) {
    int i, j, k;
    cell_ijk(D, m, i, j, k);
    atomicMin(&gptr(D.sc)->err, errkey(stage, 1, 0, i, j, k, D));
  }
  if (CFL) cfl_cell<MAG>(D, m, d, eint, v1, v2, v3, b1, b2, b3, dti, cfm);
}

template <bool MAG, bool CFL, bool DAMP, bool GRAV>
__device__ __forceinline__ void rk_cell(const Dev &D, long m, int stage, double faco,
                                        double fact, double facn, double damp, double &dti,
                                        double &cfm) {
  rk_cell_rt<MAG>(D, m, stage, faco, fact, facn, damp, CFL, DAMP, GRAV, dti, cfm);
}

template <bool MAG, bool CFL, bool DAMP, bool GRAV>
__global__ __launch_bounds__(256) void k_rk(Dev D, int stage, double faco, double fact,
                                            double facn, double damp) {
  if (gptr(D.sc)->stop) return;
  const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double dti = __builtin_inf(), cfm = 0.0;
  if (m < D.ncell)
    rk_cell<MAG, CFL, DAMP, GRAV>(D, m, stage, faco, fact, facn, damp, dti, cfm);
  if (CFL) cfl_block_reduce(D, dti, cfm, blockIdx.x);
}

// ---- one sweep direction fused with the RK update ---------------------------
// The divergence of direction DIR comes straight from the sweep (registers,
// or the LDS tile when tiled); the other directions' from memory.  Each cell
// is updated by the thread that holds its divergence.  The launch reads the
// state from one buffer set and writes the new state to the other
// (Dev::Uo/Wo), so neighbouring segments still read the old values.
template <int DIR, bool MAG, bool TILE>
__device__ __forceinline__ void fused_rk(const Dev &D, const Axis &A, long base, int c,
                                         int lane, const double du[NVMAX], double *lds,
                                         int stage, const RKArgs &rk, double &dti,
                                         double &cfm) {
  const int n = A.n;
  const int nv = MAG ? (D.i9wv >= 0 ? 9 : 8) : 5;
  if (!TILE) {  // the lane's own cell, its primitives parked in LDS
    if (lane >= 2 && lane <= 61 && c >= 0 && c < n) {
      const Cell qc = parked_cell<MAG>(lds);
      rk_cell_rt<MAG, DIR>(D, base + (long)c * A.stride, stage, rk.faco, rk.fact, rk.facn,
                           rk.damp, rk.cfl, rk.dampf, rk.grav, dti, cfm, du, &qc);
    }
    return;
  }
  const int bw = blockDim.x >> 6, w = threadIdx.x >> 6;
  const int ld = tile_ld(bw), fs = 64 * ld;
  __syncthreads();  // every wavefront has read its cells out of the tile
  {
    double *p = lds + lane * ld + w;
#pragma unroll
    for (int v = 0; v < NVMAX; ++v)
      if (v < nv) p[v * fs] = du[v];
  }
  __syncthreads();
  const int t = threadIdx.x, r = t >> __builtin_ctz(bw), col = t & (bw - 1);
  long m;
  if (tile_row_cell<MAG>(lds, bw, t, m)) {
    const double *p = lds + r * ld + col;
    double d2[NVMAX];
#pragma unroll
    for (int v = 0; v < NVMAX; ++v) d2[v] = v < nv ? p[v * fs] : 0.0;
    rk_cell_rt<MAG, DIR>(D, m, stage, rk.faco, rk.fact, rk.facn, rk.damp, rk.cfl, rk.dampf,
                         rk.grav, dti, cfm, d2);
  }
}

// One wavefront per segment of a DIR line, as k_sweep; CFL partials (last
// stage) one per block.
template <int DIR, bool MAG, int LIM, bool TILE, bool MIX = false>
__global__ __launch_bounds__(1024) void k_sweep_rk(Dev D, int stage, RKArgs rk) {
  extern __shared__ double lds[];
  if (gptr(D.sc)->stop) {
    // stages enqueued after the run stopped: the host still makes the
    // output set the current one, so hand the state over unchanged
    const long nt = (long)gridDim.x * blockDim.x;
    for (long m = (long)blockIdx.x * blockDim.x + threadIdx.x; m < D.ncell; m += nt) {
      for (int v = 0; v < D.nv; ++v) gptr(D.Uo[v])[m] = gptr(D.U[v])[m];
      for (int w = 0; w < D.nw; ++w) {  // d, e, B alias U; a shared phi stays
        const bool aliased = w == WD || w == WE || (MAG && w >= WB1 && w <= WB3);
        if (!aliased && D.Wo[w] != D.W[w]) gptr(D.Wo[w])[m] = gptr(D.W[w])[m];
      }
    }
    return;
  }
  const int lane = threadIdx.x & 63;
  const long wave = (long)blockIdx.x * (blockDim.x >> 6) + wave_in_block();
  double dti = __builtin_inf(), cfm = 0.0;
  if (TILE || wave < D.waves[DIR]) {
    if (MIX)
      sweep_line_mix<DIR, MAG, LIM, TILE, false, false, true>(D, wave, lane, stage, lds, &rk,
                                                               &dti, &cfm);
    else
      sweep_line<DIR, MAG, LIM, TILE, false, true, false>(D, wave, lane, stage, lds, &rk, &dti,
                                                           &cfm);
  }
  if (rk.cfl) cfl_block_reduce(D, dti, cfm, blockIdx.x);
}

// ---- pressure(): eint from e, d, v, B after an upload ---------------------
template <bool MAG>
__global__ __launch_bounds__(256) void k_pressure(Dev D) {
  const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= D.ncell) return;
  double b1 = 0, b2 = 0, b3 = 0;
  if (MAG) {
    b1 = gptr(D.W[WB1])[m];
    b2 = gptr(D.W[WB2])[m];
    b3 = gptr(D.W[WB3])[m];
  }
  const double eint = get_eint(MAG, gptr(D.W[WE])[m], gptr(D.W[WD])[m], gptr(D.W[WV1])[m], gptr(D.W[WV2])[m],
                               gptr(D.W[WV3])[m], b1, b2, b3);
  gptr(D.W[WEINT])[m] = eint;
  if (eint <= 0.0) {
    int i, j, k;
    cell_ijk(D, m, i, j, k);
    atomicMin(&gptr(D.sc)->err, errkey(0, 1, 0, i, j, k, D));
  }
}

// ---- hyperbolic self-gravity sub-step (gravity_hyperbolic.f90:61-125) -----
// The potential is stored twice with its ghost layer (the Dirichlet boundary
// values, the same in both buffers; no kernel writes a ghost), psi twice
// without.  A sub-step reads phi from gphi[in] and psi from gpsi and writes
// the new values to gphi[1-in] and gpsi1; the host then swaps the buffers.
// lap_coeff for crdnt 0 (new; the reference has no Cartesian branch):
// c0 = -((idx1(i)+idx1(i+1))*dxi2*dxi3 + (idx2(j)+idx2(j+1))*dxi1*dxi3 +
// (idx3(k)+idx3(k+1))*dxi1*dxi2), c1(i) = dxi2*dxi3*idx1(i+1), c2, c3
// alike, normalised by the cell volume (get_lapphi_hgsrc, :230-280).

// Cell (i, j, k) and element m of a gravity launch: a block is blockDim.y
// rows of blockDim.x cells of one z plane, blocks numbered x fastest, then
// y, then z.  False past the box.  ROW: one row per block (blockDim.y = 1),
// so j and k are the same for the whole block and the y and z grid factors
// of the Laplacian are scalar loads, not vector ones.  CPT cells of the row
// per thread (i .. i+CPT-1; the launcher has CPT*blockDim.x divide n1).
template <bool ROW, int CPT = 1>
__device__ __forceinline__ bool grav_cell(const Dev &D, int &i, int &j, int &k, long &m) {
  const unsigned bc = blockDim.x * CPT;
  const unsigned gx = ((unsigned)D.n1 + bc - 1) / bc;
  const unsigned gy = ROW ? (unsigned)D.n2 : ((unsigned)D.n2 + blockDim.y - 1) / blockDim.y;
  const unsigned t = blockIdx.x / gx;
  i = (int)(((blockIdx.x - t * gx) * blockDim.x + threadIdx.x) * CPT);
  j = ROW ? (int)(t % gy) : (int)((t % gy) * blockDim.y + threadIdx.y);
  k = (int)(t / gy);
  m = ((long)k * D.n2 + j) * D.n1 + i;
  return i < D.n1 && j < D.n2;
}

// The Laplacian of the potential at cell (i, j, k) from its value there (v)
// and at its x, y, z neighbours (xm, xp, ...; the unsolved directions' are
// not read).  One operation order for every caller.
// grav_lap_f takes the grid factors themselves: x1 = dxi1(i), a1 = idx1(i),
// b1 = idx1(i+1), and the same for y (j) and z (k).  The unsolved
// directions' terms are computed and not taken (a select, not a branch, so
// that no load is held back behind one).
__device__ __forceinline__ double grav_lap_f(const Dev &D, double x1, double a1, double b1,
                                             double x2, double a2, double b2, double x3,
                                             double a3, double b3, double v, double xm,
                                             double xp, double ym, double yp, double zm,
                                             double zp) {
  const double c0 = -((a1 + b1) * x2 * x3 + (a2 + b2) * x1 * x3 + (a3 + b3) * x1 * x2);
  double lap = c0 * v;
  lap = lap + x2 * x3 * a1 * xm;
  lap = lap + x2 * x3 * b1 * xp;
  const double ly = (lap + x1 * x3 * a2 * ym) + x1 * x3 * b2 * yp;
  lap = D.n2 > 1 ? ly : lap;
  const double lz = (lap + x1 * x2 * a3 * zm) + x1 * x2 * b3 * zp;
  lap = D.n3 > 1 ? lz : lap;
  return lap / (x1 * x2 * x3);
}
__device__ __forceinline__ double grav_lapphi(const Dev &D, int i, int j, int k, double v,
                                              double xm, double xp, double ym, double yp,
                                              double zm, double zp) {
  const double *X1 = D.ax[0].dxi, *X2 = D.ax[1].dxi, *X3 = D.ax[2].dxi;
  const double *I1 = D.ax[0].idx, *I2 = D.ax[1].idx, *I3 = D.ax[2].idx;
  return grav_lap_f(D, X1[i], I1[i], I1[i + 1], X2[j], I2[j], I2[j + 1], X3[k], I3[k],
                    I3[k + 1], v, xm, xp, ym, yp, zm, zp);
}

// One stage of the wave update with its RK combination (hyperbolic_gravity_
// step, :73-116): phi = faco*phi0 + facn*phi + fact*dtg*c*psi, psi =
// faco*psi0 + facn*psi + fact*dtg*c*(lap(phi) - 4 pi G rho).
__device__ __forceinline__ double grav_phi(double faco, double facn, double fact, double dtg,
                                           double cg, double p0, double p, double q) {
  return faco * p0 + facn * p + fact * dtg * cg * q;
}
__device__ __forceinline__ double grav_psi(double faco, double facn, double fact, double dtg,
                                           double cg, double q0, double q, double lapphi,
                                           double hgsrc) {
  return faco * q0 + facn * q + fact * dtg * cg * (lapphi - hgsrc);
}

// grktype 1: the single stage, with the damping factor (:114).
template <bool ROW>
__global__ __launch_bounds__(256) void k_grav_stage1(Dev D, int in, double faco, double fact,
                                                     double facn, double dtg, double cg,
                                                     double damp) {
  int i, j, k;
  long m;
  if (!grav_cell<ROW>(D, i, j, k, m)) return;
  const double *ph = D.gphi[in];
  const long g = gidx(D, i, j, k);
  const long sy = D.n1 + 2, sz = (long)(D.n1 + 2) * (D.n2 + 2);
  const double lapphi = grav_lapphi(D, i, j, k, ph[g], ph[g - 1], ph[g + 1], ph[g - sy],
                                    ph[g + sy], ph[g - sz], ph[g + sz]);
  const double hgsrc = D.fourpig * gptr(D.W[WD])[m];
  const double p = ph[g], q = D.gpsi[m];
  D.gphi[1 - in][g] = grav_phi(faco, facn, fact, dtg, cg, p, p, q);
  D.gpsi1[m] = grav_psi(faco, facn, fact, dtg, cg, q, q, lapphi, hgsrc) * damp;
}

// grktype 2: both stages in one launch.  Stage 1's potential update has no
// stencil (phi1 = faco*phi0 + facn*phi0 + fact*dtg*c*psi0 with the stage-1
// coefficients), so phi1 at the cell and at its six neighbours follows from
// phi0 and psi0 there (at a ghost it is the boundary value, which stage 1
// leaves alone), psi1 at the cell from the Laplacian of phi0, and stage 2
// runs at once.  Each value is the same expression on the same operands as
// in two separate stage launches, so the result is the same bit for bit,
// while phi0, psi0 and rho are read once and the new phi and psi written
// once: 40 bytes per cell and sub-step against 96 for two launches (the
// neighbours' phi0 and psi0 come from the caches).  Stage 2 applies the
// damping factor.
// CPT cells i .. i+CPT-1 of a row per thread (CPT = 2: the compiler merges
// the two cells' overlapping and adjacent loads and their stores into
// 16-byte ones, so a cell takes about half the vector memory instructions).
// Element offsets are 64-bit and the neighbours' are added to them, so
// that the adjacency is visible; the new values are stored after every cell
// is computed (a store could alias a later load).
template <bool ROW, int CPT>
__global__ __launch_bounds__(256) void k_grav_substep2(Dev D, int in, double fo1, double ft1,
                                                       double fn1, double fo2, double ft2,
                                                       double fn2, double dtg, double cg,
                                                       double damp) {
  int i, j, k;
  long m;
  if (!grav_cell<ROW, CPT>(D, i, j, k, m)) return;
  const double *ph = D.gphi[in], *ps = D.gpsi;
  const long g = gidx(D, i, j, k);
  const long sy = D.n1 + 2, sz = (long)(D.n1 + 2) * (D.n2 + 2);
  const long my = D.n1, mz = (long)D.n1 * D.n2;
  // Every neighbour's psi0 is loaded unconditionally, from the cell itself
  // where the neighbour is a ghost (the value is then not used): a load under
  // a branch waits for its data before the branch joins, which made the six
  // of them six memory round trips one after another.
  const bool iym = j > 0, iyp = j < D.n2 - 1, izm = k > 0, izp = k < D.n3 - 1;
  const long oym = iym ? -my : 0, oyp = iyp ? my : 0, ozm = izm ? -mz : 0, ozp = izp ? mz : 0;
  const double *X1 = D.ax[0].dxi + i, *I1 = D.ax[0].idx + i;
  const double x2 = D.ax[1].dxi[j], a2 = D.ax[1].idx[j], b2 = D.ax[1].idx[j + 1];
  const double x3 = D.ax[2].dxi[k], a3 = D.ax[2].idx[k], b3 = D.ax[2].idx[k + 1];
  const double *rho = gptr(D.W[WD]);
  // every load first, then one scheduling barrier: the compiler would
  // otherwise spread them over several waits to save registers
  double P[CPT][7], Q[CPT][7], R[CPT], XF[CPT][3];
#pragma unroll
  for (int c = 0; c < CPT; ++c) {
    const long gc = g + c, mc = m + c;
    const bool ixm = c > 0 || i > 0, ixp = c < CPT - 1 || i + c < D.n1 - 1;
    XF[c][0] = X1[c];
    XF[c][1] = I1[c];
    XF[c][2] = I1[c + 1];
    P[c][0] = ph[gc];
    P[c][1] = ph[gc - 1];
    P[c][2] = ph[gc + 1];
    P[c][3] = ph[gc - sy];
    P[c][4] = ph[gc + sy];
    P[c][5] = ph[gc - sz];
    P[c][6] = ph[gc + sz];
    Q[c][0] = ps[mc];
    Q[c][1] = ps[ixm ? mc - 1 : mc];
    Q[c][2] = ps[ixp ? mc + 1 : mc];
    Q[c][3] = ps[mc + oym];
    Q[c][4] = ps[mc + oyp];
    Q[c][5] = ps[mc + ozm];
    Q[c][6] = ps[mc + ozp];
    R[c] = rho[mc];
  }
  __builtin_amdgcn_sched_barrier(0);
  double pn[CPT], qn[CPT];
#pragma unroll
  for (int c = 0; c < CPT; ++c) {
    const double x1 = XF[c][0], a1 = XF[c][1], b1 = XF[c][2];
    // stage 1
    const double p = P[c][0], q = Q[c][0];
    const double pxm = P[c][1], pxp = P[c][2], pym = P[c][3], pyp = P[c][4];
    const double pzm = P[c][5], pzp = P[c][6];
    const double hgsrc = D.fourpig * R[c];
    const double lap0 =
        grav_lap_f(D, x1, a1, b1, x2, a2, b2, x3, a3, b3, p, pxm, pxp, pym, pyp, pzm, pzp);
    const double q1 = grav_psi(fo1, fn1, ft1, dtg, cg, q, q, lap0, hgsrc);
    // phi1 at the cell and its neighbours (interior: the stage-1 update there)
    const bool ixm = c > 0 || i > 0, ixp = c < CPT - 1 || i + c < D.n1 - 1;
    auto phi1 = [&](double pv, bool inside, double qv) {
      return inside ? grav_phi(fo1, fn1, ft1, dtg, cg, pv, pv, qv) : pv;
    };
    const double p1 = grav_phi(fo1, fn1, ft1, dtg, cg, p, p, q);
    const double xm = phi1(pxm, ixm, Q[c][1]), xp = phi1(pxp, ixp, Q[c][2]);
    // the unsolved directions' values are not used (grav_lap_f)
    const double ym = phi1(pym, iym, Q[c][3]), yp = phi1(pyp, iyp, Q[c][4]);
    const double zm = phi1(pzm, izm, Q[c][5]), zp = phi1(pzp, izp, Q[c][6]);
    // stage 2 over grvphiorg = (phi0, psi0)
    const double lap1 =
        grav_lap_f(D, x1, a1, b1, x2, a2, b2, x3, a3, b3, p1, xm, xp, ym, yp, zm, zp);
    pn[c] = grav_phi(fo2, fn2, ft2, dtg, cg, p, p1, q1);
    qn[c] = grav_psi(fo2, fn2, ft2, dtg, cg, q, q1, lap1, hgsrc) * damp;
  }
  // each array's values together, so that adjacent stores merge
#pragma unroll
  for (int c = 0; c < CPT; ++c) D.gpsi1[m + c] = qn[c];
#pragma unroll
  for (int c = 0; c < CPT; ++c) D.gphi[1 - in][g + c] = pn[c];
}

// ---- CFL condition as its own pass (the reference's timestep) ------------
template <bool MAG>
__global__ __launch_bounds__(256) void k_cfl(Dev D) {
  const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double dti = __builtin_inf(), cfm = 0.0;
  if (m < D.ncell) {
    double b1 = 0, b2 = 0, b3 = 0;
    if (MAG) {
      b1 = gptr(D.W[WB1])[m];
      b2 = gptr(D.W[WB2])[m];
      b3 = gptr(D.W[WB3])[m];
    }
    cfl_cell<MAG>(D, m, gptr(D.W[WD])[m], gptr(D.W[WEINT])[m], gptr(D.W[WV1])[m], gptr(D.W[WV2])[m], gptr(D.W[WV3])[m],
                  b1, b2, b3, dti, cfm);
  }
  cfl_block_reduce(D, dti, cfm, blockIdx.x);
}

// Final CFL reduction and dt (timestep.f90:74-89): dt = courant*min(dti),
// clipped to t_out / t_end as the reference does; ch = max fast speed.
// One block of up to 1024 threads; each thread folds every blockDim-th
// partial pair, four loads in flight.
__device__ __forceinline__ void cfl_final(const Dev &D) {
  Scalars *sc = D.sc;
  if (sc->stop) return;
  __shared__ double smin[16], smax[16];
  double a = __builtin_inf(), b = 0.0;
  const double2 *P = reinterpret_cast<const double2 *>(D.part);
  const int nb = D.nblk_cfl, st = blockDim.x;
  int q = threadIdx.x;
  for (; q + 3 * st < nb; q += 4 * st) {
    const double2 p0 = P[q], p1 = P[q + st], p2 = P[q + 2 * st], p3 = P[q + 3 * st];
    a = mn2(mn2(mn2(mn2(a, p0.x), p1.x), p2.x), p3.x);
    b = mx2(mx2(mx2(mx2(b, p0.y), p1.y), p2.y), p3.y);
  }
  for (; q < nb; q += st) {
    a = mn2(a, P[q].x);
    b = mx2(b, P[q].y);
  }
  a = wave_min(a);
  b = wave_max(b);
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    smin[w] = a;
    smax[w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    a = smin[0];
    b = smax[0];
    for (int q = 1; q < nw; ++q) {
      a = mn2(a, smin[q]);
      b = mx2(b, smax[q]);
    }
    sc->dtimin = a;
    sc->cfmax = b;
    double dt = D.courant * a;
    const double time = sc->time, t_out = sc->t_out, t_end = sc->t_end;
    if (sc->outstyle == 1) {
      double t_ref = t_out;  // left unset by the reference unless endstyle == 1
      if (sc->endstyle == 1) t_ref = mn2(t_out, t_end);
      if (t_ref - time - dt > 0.0 && t_ref - time - 2.0 * dt < 0.0)
        dt = 0.5 * (t_ref - time);
      else
        dt = mn2(dt, t_out - time);
    }
    if (sc->endstyle == 1) dt = mn2(dt, t_end - time);
    sc->dt = dt;
    sc->ch = b;
  }
}

__global__ __launch_bounds__(1024) void k_cfl_final(Dev D) { cfl_final(D); }

// End of one main-loop iteration (hormone.f90:147-177): advance the clock,
// then the end and output checks in the reference's order.
__device__ __forceinline__ void step_end(const Dev &D) {
  Scalars *sc = D.sc;
  if (sc->stop) return;
  if (sc->err != NOERR) {
    sc->stop = STOP_ERROR;
    return;
  }
  sc->time = sc->time + sc->dt;
  sc->tn = sc->tn + 1;
  sc->steps_left = sc->steps_left - 1;
  if (sc->endstyle == 1) {
    if (sc->time >= sc->t_end) {
      sc->stop = STOP_END;
      return;
    }
  } else if (sc->endstyle == 2) {
    if (sc->tn >= sc->tnlim) {
      sc->stop = STOP_END;
      return;
    }
  }
  if (sc->outstyle == 1) {
    if (sc->time >= sc->t_out) {
      sc->t_out = sc->t_out + sc->dt_out;
      sc->stop = STOP_OUTPUT;
      return;
    }
  } else if (sc->outstyle == 2) {
    if (sc->tn != 0 && sc->tn % sc->tn_out == 0) {
      sc->stop = STOP_OUTPUT;
      return;
    }
  }
  if (sc->steps_left <= 0) sc->stop = STOP_MAXSTEPS;
}


// ---- batched boxes: persistent kernels over flattened work-lists ---------
// One launch per direction (as the single-box x-fused form, a third of the
// all-direction code each, within the instruction cache).  Its work items
// are (box, wavefront of the direction), every box's range padded to a
// multiple of 16 items that do nothing past the box's last wavefront.  A
// block takes its items in aligned chunks of one per wavefront, so a chunk
// never mixes boxes and its y/z lines can share an LDS tile (TILE: the
// launch carries the LDS); y and z chunks go to the XCDs in contiguous
// ranges (xcd_block), as the single-box y/z launches.
// Record of box q at RK stage `stage`: the box's own record, or in the
// x-fused form the one for the box's starting set and this stage.
//
// The records and bpar do not change during a sweep launch, but the compiler
// cannot see that past the launch's own stores: read through the constant
// address space, they become scalar loads that stay in SGPRs, as a
// single-box launch's kernel-argument record does.
__device__ __forceinline__ const Dev &batch_box(const Batch &B, int q, int stage) {
  typedef const __attribute__((address_space(4))) Dev CDev;
  typedef const __attribute__((address_space(4))) int CInt;
  int r = q;
  if (B.fused) r = 9 * q + 3 * *((CInt *)B.bpar + q) + (stage - 1);
  return *(const Dev *)((CDev *)B.boxes + r);
}

// Work item i of a list, the same for the whole wavefront: a scalar load
// (the lists do not change during a launch either).
__device__ __forceinline__ int2 batch_item(const int2 *list, long i) {
  typedef const __attribute__((address_space(4))) int CInt;
  const CInt *p = (CInt *)(list + i);
  return make_int2(p[0], p[1]);
}

// One chunk (bw work items, one per wavefront) of direction DIR's sweep.
template <int DIR, bool MAG, int LIM, bool MIX, bool TILE>
__device__ __forceinline__ void bsweep_chunk(const Batch &B, int stage, long g, double *lds) {
  const int lane = threadIdx.x & 63;
  const int bw = blockDim.x >> 6;
  const long wid = threadIdx.x >> 6;
  const long ibeg = B.dbeg[DIR], nit = B.dend[DIR] - ibeg;
  // the wavefront's index in the block, visibly uniform, so that the work
  // item and the box record are scalar loads
  const long it = g * bw + __builtin_amdgcn_readfirstlane((int)wid);
  if (it >= nit) return;  // only in a last, untiled chunk
  const int2 x = batch_item(B.sitems, ibeg + it);
  const Dev &D = batch_box(B, __builtin_amdgcn_readfirstlane(x.x), stage);
  if (gptr(D.sc)->stop) return;  // uniform over the chunk (one box)
  const long w = __builtin_amdgcn_readfirstlane(x.y);
  // tiled: padding waves of a partly filled chunk join its barriers
  if (TILE ? w - wid >= D.waves[DIR] : w >= D.waves[DIR]) return;
  if (MIX) sweep_line_mix<DIR, MAG, LIM, TILE, false, false>(D, w, lane, stage, lds);
  else sweep_line<DIR, MAG, LIM, TILE, false, false, false>(D, w, lane, stage, lds);
}

// PERSIST: fewer blocks than chunks, each block loops over its chunks;
// else one chunk per block (no loop state held across the sweep).
template <int DIR, bool MAG, int LIM, bool MIX, bool TILE, bool PERSIST>
__global__ __launch_bounds__(1024) void k_bsweep_dir(Batch B, int stage) {
  extern __shared__ double lds[];
  const unsigned b0 = DIR == 0 ? blockIdx.x : xcd_block(blockIdx.x, gridDim.x);
  if (!PERSIST) {
    bsweep_chunk<DIR, MAG, LIM, MIX, TILE>(B, stage, b0, lds);
    return;
  }
  const int bw = blockDim.x >> 6;
  const long nchunk = (B.dend[DIR] - B.dbeg[DIR] + bw - 1) / bw;
  for (long g = b0; g < nchunk; g += gridDim.x) {
    bsweep_chunk<DIR, MAG, LIM, MIX, TILE>(B, stage, g, lds);
    // the next chunk rewrites the tile's index table, which other rows may
    // still be reading for their stores (every return in bsweep_chunk is
    // uniform over a tiled block)
    if (TILE) __syncthreads();
  }
}

// The x-fused form: x sweep and RK update of every box (rk_cell_rt through
// fused_rk, as k_sweep_rk<0>), reading the record's state set and writing
// the other.  The CFL partials of the last stage go one per chunk of the
// box (its wavefronts w .. w+bw-1 give partial w/bw, as a single-box launch
// numbers its blocks); the phidamp factor is the box's own.
template <bool MAG, int LIM, bool MIX>
__device__ __forceinline__ void bsweep_rk_chunk(const Batch &B, int stage, long g, double faco,
                                                double fact, double facn, int last,
                                                double *lds) {
  const int lane = threadIdx.x & 63;
  const int bw = blockDim.x >> 6;
  const long wid = threadIdx.x >> 6;
  // whole chunks only: every box's range is a multiple of 16 items and bw
  // divides 16
  const int2 x =
      batch_item(B.sitems, B.dbeg[0] + g * bw + __builtin_amdgcn_readfirstlane((int)wid));
  const Dev &D = batch_box(B, __builtin_amdgcn_readfirstlane(x.x), stage);
  if (gptr(D.sc)->stop) return;  // uniform over the chunk (one box)
  const long w = __builtin_amdgcn_readfirstlane(x.y);
  if (w - wid >= D.waves[0]) return;  // a chunk of padding only
  const RKArgs rk{faco, fact, facn, D.damp, last, last && MAG && D.i9wv >= 0, 0};
  double dti = __builtin_inf(), cfm = 0.0;
  if (w < D.waves[0]) {
    if (MIX)
      sweep_line_mix<0, MAG, LIM, false, false, false, true>(D, w, lane, stage, lds, &rk, &dti,
                                                             &cfm);
    else
      sweep_line<0, MAG, LIM, false, false, true, false>(D, w, lane, stage, lds, &rk, &dti,
                                                         &cfm);
  }
  if (last) {
    cfl_block_reduce(D, dti, cfm, (w - wid) / bw);
    __syncthreads();  // the reduction's shared slots are reused next chunk
  }
}

// PERSIST as k_bsweep_dir.
template <bool MAG, int LIM, bool MIX, bool PERSIST>
__global__ __launch_bounds__(1024) void k_bsweep_rk(Batch B, int stage, double faco,
                                                    double fact, double facn, int last) {
  extern __shared__ double lds[];
  if (!PERSIST) {
    bsweep_rk_chunk<MAG, LIM, MIX>(B, stage, blockIdx.x, faco, fact, facn, last, lds);
    return;
  }
  const int bw = blockDim.x >> 6;
  const long nchunk = (B.dend[0] - B.dbeg[0] + bw - 1) / bw;
  for (long g = blockIdx.x; g < nchunk; g += gridDim.x)
    bsweep_rk_chunk<MAG, LIM, MIX>(B, stage, g, faco, fact, facn, last, lds);
}

template <bool MAG, bool LAST>
__global__ __launch_bounds__(256) void k_brk(Batch B, int stage, double faco, double fact,
                                             double facn) {
  for (long it = blockIdx.x; it < B.ncitems; it += gridDim.x) {
    const int2 x = batch_item(B.citems, it);
    const Dev &D = batch_box(B, __builtin_amdgcn_readfirstlane(x.x), stage);
    if (gptr(D.sc)->stop) continue;  // uniform over the block
    const long m = (long)x.y * 256 + threadIdx.x;
    double dti = __builtin_inf(), cfm = 0.0;
    if (m < D.ncell) {
      if (LAST && MAG && D.i9wv >= 0)
        rk_cell<MAG, true, true, false>(D, m, stage, faco, fact, facn, D.damp, dti, cfm);
      else
        rk_cell<MAG, LAST, false, false>(D, m, stage, faco, fact, facn, 1.0, dti, cfm);
    }
    if (LAST) {
      cfl_block_reduce(D, dti, cfm, x.y);
      __syncthreads();  // the reduction's shared slots are reused next item
    }
  }
}

// The end of one step and the dt of the next in one launch: step_end on one
// thread, then (if the run goes on) the CFL reduction of cfl_final.
__global__ __launch_bounds__(1024) void k_step_end_cfl(Dev D) {
  if (threadIdx.x == 0) step_end(D);
  __syncthreads();
  cfl_final(D);
}

// One block per box: close its step, then compute its next dt and ch.  In
// the x-fused form a box that ran this step (it was not stopped at its
// start) now holds its state in the set its last stage wrote.
__global__ __launch_bounds__(256) void k_bstep_end_cfl(Batch B) {
  const int q = blockIdx.x;
  const Dev &D = B.boxes[B.fused ? 9 * q : q];
  if (threadIdx.x == 0) {
    const bool ran = gptr(D.sc)->stop == STOP_NONE;
    step_end(D);
    // the last stage wrote set p+1 (one or three stages) or p+2 (two)
    if (B.fused && ran) B.bpar[q] = (B.bpar[q] + (B.rk == 2 ? 2 : 1)) % 3;
  }
  __syncthreads();
  cfl_final(D);
}

__global__ void k_clear_error(Scalars *sc) { sc->err = NOERR; }

// dt, ch and the error status of a finished CFL reduction into a caller's
// device array (the allocation-free timestep)
__global__ void k_timestep_out(const Scalars *sc, double *out) {
  out[0] = sc->dt;
  out[1] = sc->ch;
  out[2] = sc->err == NOERR ? 0.0 : (double)HG_ERR_NEG_EINT;
}

__global__ void k_set_step(Scalars *sc, double dt, double ch) {
  sc->dt = dt;
  sc->ch = ch;
  sc->stop = STOP_NONE;
}

__global__ void k_set_clock(Scalars *sc, double time, double t_out, double t_end,
                            double dt_out, int outstyle, int endstyle, long long tn,
                            long long tnlim, long long tn_out, long long steps) {
  sc->time = time;
  sc->t_out = t_out;
  sc->t_end = t_end;
  sc->dt_out = dt_out;
  sc->outstyle = outstyle;
  sc->endstyle = endstyle;
  sc->tn = tn;
  sc->tnlim = tnlim;
  sc->tn_out = tn_out;
  sc->steps_left = steps;
  sc->stop = STOP_NONE;
}

__global__ __launch_bounds__(256) void k_phidamp(Dev D, double factor) {
  const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= D.ncell) return;
  gptr(D.W[WPHI])[m] = gptr(D.W[WPHI])[m] * factor;
}

// ---- known-answer evaluation of the device arithmetic -----------------------
// SH: the forms the sweep runs, with divisions by a shared denominator (qdiv)
template <bool SH>
__device__ void kat_one(int which, long t, const double *in, double *out) {
  if (which == 0) {  // hlldflux: 25 inputs (hlldflux.f90:19-21 order), 9 outputs
    const double *a = in + 25 * t;
    double f[9];
    hlld_d<SH>(f, a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], a[10], a[11],
               a[12], a[13], a[14], a[15], a[16], a[17], a[18], a[19], a[20], a[21], a[22]);
    for (int s = 0; s < 9; ++s) out[9 * t + s] = f[s];
  } else if (which == 1) {  // get_cf: d cs b1 b2 b3
    const double *a = in + 5 * t;
    out[t] = get_cf_d(mkden<SH>(a[0]), a[1], a[2], a[3], a[4]);
  } else if (which == 2) {  // modified_mc: u0 u1 u2 xm x0 xp xim xi0
    const double *a = in + 8 * t;
    const double xm = a[3], x0 = a[4], xp = a[5], xim = a[6], xi0 = a[7];
    const double dxp = xp - x0;
    const double cf = dxp / (xi0 - x0), cb = (x0 - xm) / (x0 - xim);
    out[t] = modified_mc_d(a[0], a[1], a[2], dxp, cf, cb, mkden<SH>(x0 - xm), mkden<SH>(dxp));
  } else if (which == 3) {  // minmod: u0 u1 u2 dxa dxb
    const double *a = in + 5 * t;
    out[t] = minmod_d(a[0], a[1], a[2], a[3], a[4], mkden<SH>(a[3] + a[4]));
  } else if (which == 4) {  // division and square root: a b -> a/b, sqrt(a)
    const double *a = in + 2 * t;
    out[2 * t] = a[0] / a[1];
    out[2 * t + 1] = sqrt(a[0]);
  } else if (which == 5) {  // division by a shared denominator: b a0 a1 a2 a3 -> ai/b
    const double *a = in + 5 * t;
    const Den64 Db = mkden<true>(a[0]);
    for (int s = 0; s < 4; ++s) out[4 * t + s] = qdiv(Db, a[1 + s]);
  } else if (which == 6) {  // eos_p_cs as the face states use it: d eint gamma -> p cs
    const double *a = in + 3 * t;
    const double p = (a[2] - 1.0) * a[1];
    out[2 * t] = p;
    out[2 * t + 1] = sqrt(qdiv(mkden<SH>(a[0]), a[2] * p));
  } else if (which == 7) {  // get_eint: etot d v1 v2 v3 b1 b2 b3 mag_on -> eint
    const double *a = in + 9 * t;
    out[t] = get_eint(a[8] > 0.5, a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7]);
  }
}

// which 0..7 as hg_kat; + 16: the sweep's shared-denominator form
__global__ void k_kat(int which, int n, const double *in, double *out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  if (which >= 16) kat_one<true>(which - 16, t, in, out);
  else kat_one<false>(which, t, in, out);
}

// ---- launchers ----------------------------------------------------------------
static inline unsigned nblk(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

void runge_coeff(int stage, int rk, double *faco, double *fact, double *facn) {
  if (rk == 3) {
    if (stage == 1) { *faco = 1.0; *fact = 1.0; *facn = 0.0; }
    else if (stage == 2) { *faco = 0.75; *fact = 0.25; *facn = *fact; }
    else { *faco = 1.0 / 3.0; *fact = 2.0 / 3.0; *facn = *fact; }
  } else if (rk == 2) {
    if (stage == 1) { *faco = 1.0; *fact = 1.0; *facn = 0.0; }
    else { *faco = 0.5; *fact = 0.5; *facn = *fact; }
  } else {
    *faco = 1.0; *fact = 1.0; *facn = 0.0;
  }
}

// Launch k<<<g, b, 0, s>>>(args...); with ev the dispatch itself carries the
// start/stop events.
template <typename... KA, typename... A>
static inline void klaunch(void (*k)(KA...), dim3 g, dim3 b, hipStream_t s, const KEv *ev,
                           A... args) {
  if (ev)
    hipExtLaunchKernelGGL(k, g, b, 0, s, ev->start, ev->stop, 0, KA(args)...);
  else
    k<<<g, b, 0, s>>>(KA(args)...);
}

// Same, with dynamic LDS.
template <typename... KA, typename... A>
static inline void klaunch_lds(void (*k)(KA...), unsigned g, unsigned b, size_t lds,
                               hipStream_t s, const KEv *ev, A... args) {
  if (ev)
    hipExtLaunchKernelGGL(k, dim3(g), dim3(b), (uint32_t)lds, s, ev->start, ev->stop, 0,
                          KA(args)...);
  else
    k<<<g, b, lds, s>>>(KA(args)...);
}

// Bytes of LDS the y/z tile of a sweep block needs (the fields of 64 cells
// of bw lines, then the index table, 16 bytes per cell), or 0 when the
// block has fewer than 4 lines, no y/z direction is solved or the tile
// would not fit.
size_t sweep_tile_bytes(const Dev &D, int bw) {
  if (bw < 4 || !(D.solve[1] || D.solve[2])) return 0;
  const size_t nf = D.mag ? NWMAX : WEINT + 1;
  const size_t bytes = (nf * 64 * (size_t)(bw + 1) + 128 * (size_t)bw) * sizeof(double);
  return bytes <= (size_t)D.lds_max ? bytes : 0;
}

void launch_pressure(const Dev &D, hipStream_t s) {
  if (D.mag) k_pressure<true><<<nblk(D.ncell, 256), 256, 0, s>>>(D);
  else k_pressure<false><<<nblk(D.ncell, 256), 256, 0, s>>>(D);
}

void launch_clear_error(const Dev &D, hipStream_t s) { k_clear_error<<<1, 1, 0, s>>>(D.sc); }

void launch_timestep_out(const Dev &D, double *out, hipStream_t s) {
  k_timestep_out<<<1, 1, 0, s>>>(D.sc, out);
}

void launch_set_step(const Dev &D, double dt, double ch, hipStream_t s) {
  k_set_step<<<1, 1, 0, s>>>(D.sc, dt, ch);
}

void launch_set_clock(const Dev &D, double time, double t_out, double t_end, double dt_out,
                      int outstyle, int endstyle, long long tn, long long tnlim,
                      long long tn_out, long long steps, hipStream_t s) {
  k_set_clock<<<1, 1, 0, s>>>(D.sc, time, t_out, t_end, dt_out, outstyle, endstyle, tn,
                              tnlim, tn_out, steps);
}

// Sweep launch: D.sweep_bw wavefronts per block.  Consecutive wavefronts of
// a block own neighbouring lines (consecutive i for the y and z sweeps), so
// with 16 wavefronts a block uses whole 128-byte cache lines of the strided
// y/z loads.
template <bool MAG, int LIM, bool EXT>
static void sweep_variant_x(const Dev &D, int stage, unsigned g, int bs, size_t tile,
                            hipStream_t s, const KEv *ev) {
  if (D.mixed) {
    if (D.nspc) klaunch_lds(k_sweep_mix<MAG, LIM, false, true, EXT>, g, bs, 0, s, ev, D, stage);
    else if (tile) klaunch_lds(k_sweep_mix<MAG, LIM, true, false, EXT>, g, bs, tile, s, ev, D, stage);
    else klaunch_lds(k_sweep_mix<MAG, LIM, false, false, EXT>, g, bs, 0, s, ev, D, stage);
  } else {
    if (D.nspc) klaunch_lds(k_sweep<MAG, LIM, false, true, EXT>, g, bs, 0, s, ev, D, stage);
    else if (tile) klaunch_lds(k_sweep<MAG, LIM, true, false, EXT>, g, bs, tile, s, ev, D, stage);
    else klaunch_lds(k_sweep<MAG, LIM, false, false, EXT>, g, bs, 0, s, ev, D, stage);
  }
}

// boxes with a code 4, 5, 9 or 10 run their own instantiation, so that the
// others keep the sweep without those branches
template <bool MAG, int LIM>
static void sweep_variant(const Dev &D, int stage, unsigned g, int bs, size_t tile,
                          hipStream_t s, const KEv *ev) {
  if (D.ax[0].ext || D.ax[1].ext || D.ax[2].ext)
    sweep_variant_x<MAG, LIM, true>(D, stage, g, bs, tile, s, ev);
  else
    sweep_variant_x<MAG, LIM, false>(D, stage, g, bs, tile, s, ev);
}

void launch_sweep(const Dev &D, int stage, hipStream_t s, const KEv *ev) {
  const int bw = D.sweep_bw > 0 ? D.sweep_bw : 4;
  const unsigned g = nblk(D.woff[2] + D.waves[2], bw);
  const int bs = 64 * bw;
  // LDS tile for the y/z lines with 4+ lines per block (each column finds
  // its own line, so any x extent works)
  const size_t tile = sweep_tile_bytes(D, bw);
  if (D.mag) {
    if (D.limiter == LIM_FLAT) sweep_variant<true, LIM_FLAT>(D, stage, g, bs, tile, s, ev);
    else if (D.limiter == LIM_MINMOD) sweep_variant<true, LIM_MINMOD>(D, stage, g, bs, tile, s, ev);
    else sweep_variant<true, LIM_MC>(D, stage, g, bs, tile, s, ev);
  } else {
    if (D.limiter == LIM_FLAT) sweep_variant<false, LIM_FLAT>(D, stage, g, bs, tile, s, ev);
    else if (D.limiter == LIM_MINMOD) sweep_variant<false, LIM_MINMOD>(D, stage, g, bs, tile, s, ev);
    else sweep_variant<false, LIM_MC>(D, stage, g, bs, tile, s, ev);
  }
}

void launch_species(const Dev &D, hipStream_t s) {
  if (!D.nspc) return;
  const unsigned g = nblk(D.woff[2] + D.waves[2], 4);
  if (D.limiter == LIM_FLAT) klaunch(k_species<LIM_FLAT>, g, 256, s, nullptr, D);
  else if (D.limiter == LIM_MINMOD) klaunch(k_species<LIM_MINMOD>, g, 256, s, nullptr, D);
  else klaunch(k_species<LIM_MC>, g, 256, s, nullptr, D);
}

template <int DIR, bool MAG, int LIM>
static void sweep_rk_variant(const Dev &D, int stage, const RKArgs &rk, unsigned g, int bs,
                             size_t tile, hipStream_t s, const KEv *ev) {
  if constexpr (DIR > 0) {  // x lines are lane lines, never tiled
    if (tile) {
      if (D.mixed) klaunch_lds(k_sweep_rk<DIR, MAG, LIM, true, true>, g, bs, tile, s, ev, D, stage, rk);
      else klaunch_lds(k_sweep_rk<DIR, MAG, LIM, true>, g, bs, tile, s, ev, D, stage, rk);
      return;
    }
  }
  // the lane path parks each thread's primitives (park_cell)
  const size_t park = (size_t)park_fields(MAG) * sizeof(double) * (size_t)bs;
  if (D.mixed) klaunch_lds(k_sweep_rk<DIR, MAG, LIM, false, true>, g, bs, park, s, ev, D, stage, rk);
  else klaunch_lds(k_sweep_rk<DIR, MAG, LIM, false>, g, bs, park, s, ev, D, stage, rk);
}

template <int DIR>
static void sweep_rk_dir(const Dev &D, int stage, const RKArgs &rk, unsigned g, int bs,
                         size_t tile, hipStream_t s, const KEv *ev) {
  if (D.mag) {
    if (D.limiter == LIM_FLAT) sweep_rk_variant<DIR, true, LIM_FLAT>(D, stage, rk, g, bs, tile, s, ev);
    else if (D.limiter == LIM_MINMOD) sweep_rk_variant<DIR, true, LIM_MINMOD>(D, stage, rk, g, bs, tile, s, ev);
    else sweep_rk_variant<DIR, true, LIM_MC>(D, stage, rk, g, bs, tile, s, ev);
  } else {
    if (D.limiter == LIM_FLAT) sweep_rk_variant<DIR, false, LIM_FLAT>(D, stage, rk, g, bs, tile, s, ev);
    else if (D.limiter == LIM_MINMOD) sweep_rk_variant<DIR, false, LIM_MINMOD>(D, stage, rk, g, bs, tile, s, ev);
    else sweep_rk_variant<DIR, false, LIM_MC>(D, stage, rk, g, bs, tile, s, ev);
  }
}

template <int DIR, bool MAG, int LIM>
static void sweep_dir_launch(const Dev &D, int stage, int bw, size_t tile, hipStream_t s,
                             const KEv *ev) {
  const unsigned g = nblk(D.waves[DIR], bw);
  const int bs = 64 * bw;
  const bool t = DIR > 0 && tile;
  if (D.mixed) {
    if (t) klaunch_lds(k_sweep_dir<DIR, MAG, LIM, true, true>, g, bs, tile, s, ev, D, stage);
    else klaunch_lds(k_sweep_dir<DIR, MAG, LIM, false, true>, g, bs, 0, s, ev, D, stage);
  } else {
    if (t) klaunch_lds(k_sweep_dir<DIR, MAG, LIM, true, false>, g, bs, tile, s, ev, D, stage);
    else klaunch_lds(k_sweep_dir<DIR, MAG, LIM, false, false>, g, bs, 0, s, ev, D, stage);
  }
}

template <bool MAG, int LIM>
static void sweep_dirs_yz_v(const Dev &D, int stage, int bw, size_t tile, hipStream_t s,
                            const KEv *ev) {
  // one event pair around both launches: start on the first, stop on the last
  const bool y = D.waves[1] > 0, z = D.waves[2] > 0;
  KEv e1{}, e2{};
  const KEv *p1 = nullptr, *p2 = nullptr;
  if (ev) {
    e1 = KEv{ev->start, z ? nullptr : ev->stop};
    e2 = KEv{y ? nullptr : ev->start, ev->stop};
    p1 = &e1;
    p2 = &e2;
  }
  if (y) sweep_dir_launch<1, MAG, LIM>(D, stage, bw, tile, s, p1);
  if (z) sweep_dir_launch<2, MAG, LIM>(D, stage, bw, tile, s, p2);
}

// The y and z sweeps of a stage as two launches of k_sweep_dir.
static void sweep_dirs_yz(const Dev &D, int stage, int bw, size_t tile, hipStream_t s,
                          const KEv *ev) {
  if (D.mag) {
    if (D.limiter == LIM_FLAT) sweep_dirs_yz_v<true, LIM_FLAT>(D, stage, bw, tile, s, ev);
    else if (D.limiter == LIM_MINMOD) sweep_dirs_yz_v<true, LIM_MINMOD>(D, stage, bw, tile, s, ev);
    else sweep_dirs_yz_v<true, LIM_MC>(D, stage, bw, tile, s, ev);
  } else {
    if (D.limiter == LIM_FLAT) sweep_dirs_yz_v<false, LIM_FLAT>(D, stage, bw, tile, s, ev);
    else if (D.limiter == LIM_MINMOD) sweep_dirs_yz_v<false, LIM_MINMOD>(D, stage, bw, tile, s, ev);
    else sweep_dirs_yz_v<false, LIM_MC>(D, stage, bw, tile, s, ev);
  }
}

int launch_fused_stage(const Dev &D, int stage, const RKArgs &rk, hipStream_t s,
                       const KEv *ev, const KEv *ev_fused, int part) {
  const int L = D.fdir;
  const int bw = D.sweep_bw > 0 ? D.sweep_bw : 4;
  const int bs = 64 * bw;
  const size_t tile = sweep_tile_bytes(D, bw);
  // the other directions first: for L = 0, y and z in one launch each
  // (k_sweep_dir, one event pair around both); for L = 1, 2 the directions
  // before L in one all-direction launch over their wave ranges
  if (L == 0 && part != 2) sweep_dirs_yz(D, stage, bw, tile, s, ev);
  Dev Do = D;
  const long nwo = L == 0 ? 0 : D.woff[L];
  if (nwo > 0 && part != 2) {
    const unsigned g0 = nblk(nwo, bw);
    if (D.mag) {
      if (D.limiter == LIM_FLAT) sweep_variant<true, LIM_FLAT>(Do, stage, g0, bs, tile, s, ev);
      else if (D.limiter == LIM_MINMOD) sweep_variant<true, LIM_MINMOD>(Do, stage, g0, bs, tile, s, ev);
      else sweep_variant<true, LIM_MC>(Do, stage, g0, bs, tile, s, ev);
    } else {
      if (D.limiter == LIM_FLAT) sweep_variant<false, LIM_FLAT>(Do, stage, g0, bs, tile, s, ev);
      else if (D.limiter == LIM_MINMOD) sweep_variant<false, LIM_MINMOD>(Do, stage, g0, bs, tile, s, ev);
      else sweep_variant<false, LIM_MC>(Do, stage, g0, bs, tile, s, ev);
    }
  }
  const unsigned g = nblk(D.waves[L], bw);
  if (part == 1) return (int)g;
  if (L == 0) sweep_rk_dir<0>(D, stage, rk, g, bs, 0, s, ev_fused);
  else if (L == 1) sweep_rk_dir<1>(D, stage, rk, g, bs, tile, s, ev_fused);
  else sweep_rk_dir<2>(D, stage, rk, g, bs, tile, s, ev_fused);
  return (int)g;
}

template <bool MAG, bool GRAV>
static void rk_mag(const Dev &D, int stage, double faco, double fact, double facn, bool cfl,
                   bool damp, double df, hipStream_t s, const KEv *ev) {
  const unsigned g = nblk(D.ncell, 256);
  if (cfl && damp) klaunch(k_rk<MAG, true, true, GRAV>, g, 256, s, ev, D, stage, faco, fact, facn, df);
  else if (cfl) klaunch(k_rk<MAG, true, false, GRAV>, g, 256, s, ev, D, stage, faco, fact, facn, df);
  else if (damp) klaunch(k_rk<MAG, false, true, GRAV>, g, 256, s, ev, D, stage, faco, fact, facn, df);
  else klaunch(k_rk<MAG, false, false, GRAV>, g, 256, s, ev, D, stage, faco, fact, facn, df);
}

void launch_rk(const Dev &D, int stage, double faco, double fact, double facn, bool cfl,
               bool damp, double damp_factor, hipStream_t s, const KEv *ev) {
  if (D.grav) {
    if (D.mag) rk_mag<true, true>(D, stage, faco, fact, facn, cfl, damp, damp_factor, s, ev);
    else rk_mag<false, true>(D, stage, faco, fact, facn, cfl, false, damp_factor, s, ev);
  } else {
    if (D.mag) rk_mag<true, false>(D, stage, faco, fact, facn, cfl, damp, damp_factor, s, ev);
    else rk_mag<false, false>(D, stage, faco, fact, facn, cfl, false, damp_factor, s, ev);
  }
}

// One gravity sub-step of RK order gk (1 or 2) with the damping factor;
// the new phi and psi land in gphi[1-gcur] and gpsi1 (the caller swaps).
void launch_grav_substep(const Dev &D, int gk, double dtg, double cg, double damp_factor,
                         hipStream_t s, const KEv *ev) {
  // One row per block (ROW): two cells per thread when 2 x 256, 2 x 128 or
  // 2 x 64 cells divide a row, else one when 256, 128 or 64 do, or 256 cells
  // when the box has fewer than 4 rows; else 64 x 4 cells.
  unsigned bx = 256, by = 1;
  int cpt = 1;
  bool row = D.n2 < 4;
  for (unsigned w = 256; w >= 64 && !row; w /= 2)
    if (D.n1 % (int)(2 * w) == 0) {
      bx = w;
      cpt = 2;
      row = true;
    }
  for (unsigned w = 256; w >= 64 && !row; w /= 2)
    if (D.n1 % (int)w == 0) {
      bx = w;
      row = true;
    }
  if (!row) {
    bx = 64;
    by = 4;
  }
  double fo1, ft1, fn1;
  runge_coeff(1, gk, &fo1, &ft1, &fn1);
  if (gk == 1) {  // one cell per thread
    const unsigned g = nblk(D.n1, (int)bx) * nblk(D.n2, (int)by) * (unsigned)D.n3;
    const dim3 b(bx, by);
    if (row) klaunch(k_grav_stage1<true>, g, b, s, ev, D, D.gcur, fo1, ft1, fn1, dtg, cg, damp_factor);
    else klaunch(k_grav_stage1<false>, g, b, s, ev, D, D.gcur, fo1, ft1, fn1, dtg, cg, damp_factor);
    return;
  }
  const unsigned g = nblk(D.n1, (int)bx * cpt) * nblk(D.n2, (int)by) * (unsigned)D.n3;
  const dim3 b(bx, by);
  double fo2, ft2, fn2;
  runge_coeff(2, gk, &fo2, &ft2, &fn2);
  if (cpt == 2)
    klaunch(k_grav_substep2<true, 2>, g, b, s, ev, D, D.gcur, fo1, ft1, fn1, fo2, ft2, fn2, dtg,
            cg, damp_factor);
  else if (row)
    klaunch(k_grav_substep2<true, 1>, g, b, s, ev, D, D.gcur, fo1, ft1, fn1, fo2, ft2, fn2, dtg,
            cg, damp_factor);
  else
    klaunch(k_grav_substep2<false, 1>, g, b, s, ev, D, D.gcur, fo1, ft1, fn1, fo2, ft2, fn2, dtg,
            cg, damp_factor);
}

void launch_cfl(const Dev &D, hipStream_t s, const KEv *ev) {
  if (D.mag) klaunch(k_cfl<true>, (unsigned)D.nblk_cfl, 256, s, ev, D);
  else klaunch(k_cfl<false>, (unsigned)D.nblk_cfl, 256, s, ev, D);
}

// threads of the final CFL reduction: enough to keep many partial loads in
// flight on large grids, one wavefront on small ones
static inline int cfl_final_threads(const Dev &D) {
  return D.nblk_cfl > 16384 ? 1024 : (D.nblk_cfl > 2048 ? 512 : 256);
}

void launch_cfl_final(const Dev &D, hipStream_t s) {
  k_cfl_final<<<1, cfl_final_threads(D), 0, s>>>(D);
}


void launch_phidamp(const Dev &D, double factor, hipStream_t s) {
  k_phidamp<<<nblk(D.ncell, 256), 256, 0, s>>>(D, factor);
}

template <int DIR, bool MAG, int LIM>
static void bsweep_dir_mix(const Batch &B, int mixed, int stage, unsigned g, int bs,
                           hipStream_t s) {
  // the y/z tile when it fits (a separate instantiation each way: one
  // kernel holding both paths outgrows the instruction cache)
  const int bw = bs / 64;
  const size_t nf = MAG ? NWMAX : WEINT + 1;
  size_t tile = (nf * 64 * (size_t)(bw + 1) + 128 * (size_t)bw) * sizeof(double);
  if (DIR == 0 || bw < 4 || tile > (size_t)B.lds_max) tile = 0;
  const bool persist = (long)g < (B.dend[DIR] - B.dbeg[DIR] + bw - 1) / bw;
  if constexpr (DIR > 0) {
    if (tile) {
      if (persist) {
        if (mixed) k_bsweep_dir<DIR, MAG, LIM, true, true, true><<<g, bs, tile, s>>>(B, stage);
        else k_bsweep_dir<DIR, MAG, LIM, false, true, true><<<g, bs, tile, s>>>(B, stage);
      } else {
        if (mixed) k_bsweep_dir<DIR, MAG, LIM, true, true, false><<<g, bs, tile, s>>>(B, stage);
        else k_bsweep_dir<DIR, MAG, LIM, false, true, false><<<g, bs, tile, s>>>(B, stage);
      }
      return;
    }
  }
  if (persist) {
    if (mixed) k_bsweep_dir<DIR, MAG, LIM, true, false, true><<<g, bs, 0, s>>>(B, stage);
    else k_bsweep_dir<DIR, MAG, LIM, false, false, true><<<g, bs, 0, s>>>(B, stage);
  } else {
    if (mixed) k_bsweep_dir<DIR, MAG, LIM, true, false, false><<<g, bs, 0, s>>>(B, stage);
    else k_bsweep_dir<DIR, MAG, LIM, false, false, false><<<g, bs, 0, s>>>(B, stage);
  }
}

template <int DIR>
static void bsweep_dir_v(const Batch &B, int mag, int lim, int mixed, int stage, unsigned g,
                         int bs, hipStream_t s) {
  if (mag) {
    if (lim == LIM_FLAT) bsweep_dir_mix<DIR, true, LIM_FLAT>(B, mixed, stage, g, bs, s);
    else if (lim == LIM_MINMOD) bsweep_dir_mix<DIR, true, LIM_MINMOD>(B, mixed, stage, g, bs, s);
    else bsweep_dir_mix<DIR, true, LIM_MC>(B, mixed, stage, g, bs, s);
  } else {
    if (lim == LIM_FLAT) bsweep_dir_mix<DIR, false, LIM_FLAT>(B, mixed, stage, g, bs, s);
    else if (lim == LIM_MINMOD) bsweep_dir_mix<DIR, false, LIM_MINMOD>(B, mixed, stage, g, bs, s);
    else bsweep_dir_mix<DIR, false, LIM_MC>(B, mixed, stage, g, bs, s);
  }
}

void launch_bsweep_dir(const Batch &B, int dir, int mag, int lim, int mixed, int stage,
                       unsigned g, int block_waves, hipStream_t s) {
  const int bs = 64 * block_waves;
  if (dir == 0) bsweep_dir_v<0>(B, mag, lim, mixed, stage, g, bs, s);
  else if (dir == 1) bsweep_dir_v<1>(B, mag, lim, mixed, stage, g, bs, s);
  else bsweep_dir_v<2>(B, mag, lim, mixed, stage, g, bs, s);
}

template <bool MAG, int LIM>
static void bsweep_rk_mix(const Batch &B, int mixed, int stage, double faco, double fact,
                          double facn, bool last, unsigned g, int bs, hipStream_t s) {
  // the lane path parks each thread's primitives (park_cell)
  const size_t park = (size_t)park_fields(MAG) * sizeof(double) * (size_t)bs;
  const int l = last ? 1 : 0;
  const bool persist = (long)g < (B.dend[0] - B.dbeg[0] + bs / 64 - 1) / (bs / 64);
  if (persist) {
    if (mixed) k_bsweep_rk<MAG, LIM, true, true><<<g, bs, park, s>>>(B, stage, faco, fact, facn, l);
    else k_bsweep_rk<MAG, LIM, false, true><<<g, bs, park, s>>>(B, stage, faco, fact, facn, l);
  } else {
    if (mixed) k_bsweep_rk<MAG, LIM, true, false><<<g, bs, park, s>>>(B, stage, faco, fact, facn, l);
    else k_bsweep_rk<MAG, LIM, false, false><<<g, bs, park, s>>>(B, stage, faco, fact, facn, l);
  }
}

void launch_bsweep_rk(const Batch &B, int mag, int lim, int mixed, int stage, double faco,
                      double fact, double facn, bool last, unsigned g, int block_waves,
                      hipStream_t s) {
  const int bs = 64 * block_waves;
  if (mag) {
    if (lim == LIM_FLAT) bsweep_rk_mix<true, LIM_FLAT>(B, mixed, stage, faco, fact, facn, last, g, bs, s);
    else if (lim == LIM_MINMOD) bsweep_rk_mix<true, LIM_MINMOD>(B, mixed, stage, faco, fact, facn, last, g, bs, s);
    else bsweep_rk_mix<true, LIM_MC>(B, mixed, stage, faco, fact, facn, last, g, bs, s);
  } else {
    if (lim == LIM_FLAT) bsweep_rk_mix<false, LIM_FLAT>(B, mixed, stage, faco, fact, facn, last, g, bs, s);
    else if (lim == LIM_MINMOD) bsweep_rk_mix<false, LIM_MINMOD>(B, mixed, stage, faco, fact, facn, last, g, bs, s);
    else bsweep_rk_mix<false, LIM_MC>(B, mixed, stage, faco, fact, facn, last, g, bs, s);
  }
}

void launch_brk(const Batch &B, int mag, int stage, double faco, double fact, double facn,
                bool last, unsigned g, hipStream_t s) {
  if (mag) {
    if (last) k_brk<true, true><<<g, 256, 0, s>>>(B, stage, faco, fact, facn);
    else k_brk<true, false><<<g, 256, 0, s>>>(B, stage, faco, fact, facn);
  } else {
    if (last) k_brk<false, true><<<g, 256, 0, s>>>(B, stage, faco, fact, facn);
    else k_brk<false, false><<<g, 256, 0, s>>>(B, stage, faco, fact, facn);
  }
}

void launch_step_end_cfl(const Dev &D, hipStream_t s) {
  k_step_end_cfl<<<1, cfl_final_threads(D), 0, s>>>(D);
}

void launch_bstep_end_cfl(const Batch &B, hipStream_t s) {
  k_bstep_end_cfl<<<B.nbox, 256, 0, s>>>(B);
}

void launch_kat(int which, int n, const double *in, double *out, hipStream_t s) {
  k_kat<<<nblk(n, 128), 128, 0, s>>>(which, n, in, out);
}

}  // namespace hg
