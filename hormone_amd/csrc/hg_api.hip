// hg_api.hip -- the C ABI of include/hormone_hip.h over a device context.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/hormone_hip.h"
#include "hg_kernels.hpp"

using namespace hg;

// One state set: the conserved fields and the primitives (d, e and B of W
// are the U arrays).
struct StateSet {
  double *U[NVMAX], *W[NWMAX];
};

struct hg_ctx {
  hg_desc desc;
  Dev D;
  hipStream_t stream = nullptr;
  int device = 0;
  int is, js, ks;        // Fortran index of the first interior cell
  int nxg, nyg, nzg;     // padded extents
  std::vector<void *> allocs;
  double *coord = nullptr;  // device coordinate block
  Scalars *h_sc = nullptr;  // pinned host mirror of the device scalars
  bool eint_stale = true;
  double ch = 0.0;
  int err_i = 0, err_j = 0, err_k = 0, err_dir = 0;
  double *h_stage = nullptr;  // pinned staging for field transfers
  double dtg_unit = 0.0;      // tools.f90:154-184 (gravswitch 3)
  // profiling (the reference's profiler categories, profiler.f90:5-51)
  bool prof = false;
  unsigned prof_mask = 0;  // categories that carry events (bit HG_PROF_*)
  double prof_ms[HG_PROF_N] = {0};
  long long prof_n[HG_PROF_N] = {0};
  std::vector<hipEvent_t> ev;   // event pool
  std::vector<int> ev_kind;     // kind of each recorded pair
  int ev_used = 0;
  KEv kev{};
  bool fused = false;   // last sweep direction fused with the RK update
  long part_cap = 0;    // CFL partial slots allocated
  // the second state buffer set of the fused form (W aliasing U as in D)
  double *alt_U[NVMAX] = {}, *alt_W[NWMAX] = {};
  bool alt_ready = false;
  // the fused form's third state set: its conserved part is D.U0's arrays,
  // its own W fields these (d, e, B alias U as in the other sets)
  double *alt2_W[NWMAX] = {};
  bool alt2_ready = false;
  bool in_batch = false;
  // the batch it is in, and its three state sets as the batch's device
  // records number them (set 0 is the one D held when they were built)
  struct hg_batch *batch = nullptr;
  StateSet bsets[3] = {};
  // hg_run_gravity: the gravity sub-steps run on their own stream, beside
  // stage 1's flux sweep; gdone joins them back before stage 1's update
  hipStream_t gstream = nullptr;
  hipEvent_t gdone = nullptr;
};

#define HIPCHK(x)                                   \
  do {                                              \
    hipError_t e_ = (x);                            \
    if (e_ != hipSuccess) return HG_ERR_DEVICE;     \
  } while (0)

static const char kVersion[] = "hormone_hip 0.1 (gfx950)";

extern "C" const char *hg_version(void) { return kVersion; }

static int dev_alloc(hg_ctx *c, size_t bytes, void **p) {
  if (hipMalloc(p, bytes) != hipSuccess) return HG_ERR_NOMEM;
  c->allocs.push_back(*p);
  return HG_OK;
}

// The codes each face accepts (the cases of boundarycondition.f90's select
// blocks; any other code stops the reference with "Error from x? ...
// boundary condition"): 0-3, 9 and 10 everywhere; linear 4/5 at x1 outer
// (:302, :419, :441), x3 inner (:963, :1074, :1096) and x3 outer, where
// the scalar code takes 4 only (:1222; vectors :1335, :1357).  b indexes
// hg_desc.bc: 0..5 scalar codes (1is 1os 2is 2os 3is 3os), 6..11 vector.
static bool bc_accepted(int code, int b) {
  if ((code >= 0 && code <= 3) || code == 9 || code == 10) return true;
  const int face = b % 6;
  if (code == 4 || code == 5) return face == 1 || face == 4 || (face == 5 && (b >= 6 || code == 4));
  return false;
}

static int validate(const hg_desc *d) {
  if (!d) return HG_ERR_ARG;
  if (d->crdnt != 0 || d->eostype != 0 || (d->compswitch != 0 && d->compswitch != 2))
    return HG_ERR_UNSUPPORTED;
  if (d->compswitch == 2 && (d->spn < 1 || d->spn > HG_SPN_MAX)) return HG_ERR_ARG;
  if (d->ie < d->is || d->je < d->js || d->ke < d->ks) return HG_ERR_ARG;
  // the kernels index cells, sweep positions and the potential's padded
  // grid in 32 bits (4.3e9 elements: far beyond one device's memory at
  // ~50 FP64 fields per cell)
  if ((double)(d->ie - d->is + 5) * (d->je - d->js + 5) * (d->ke - d->ks + 5) >= 4294967295.0)
    return HG_ERR_UNSUPPORTED;
  if (d->rktype < 1 || d->rktype > 3) return HG_ERR_ARG;
  if (d->flux_limiter < 0 || d->flux_limiter > 2) return HG_ERR_ARG;
  for (int b = 0; b < 12; ++b)
    if (!bc_accepted(d->bc[b], b)) return HG_ERR_UNSUPPORTED;
  // species ghosts under Dirichlet (spc0) and flux (spcflx) boundaries are
  // not carried
  if (d->compswitch == 2)
    for (int b = 0; b < 12; ++b)
      if (d->bc[b] >= 9) return HG_ERR_UNSUPPORTED;
  // checksetup.f90:82-107: periodic must pair with periodic
  for (int q = 0; q < 6; ++q) {
    int a = d->bc[2 * q], b = d->bc[2 * q + 1];
    if (a * b == 0 && a + b != 0) return HG_ERR_ARG;
  }
  if (d->gravswitch != 0 && d->gravswitch != 3) return HG_ERR_UNSUPPORTED;
  if (d->gravswitch == 3 && (d->grktype < 1 || d->grktype > 2)) return HG_ERR_ARG;
  if (d->precision != HG_PREC_FP64 && d->precision != HG_PREC_FP32MIX) return HG_ERR_ARG;
  const double *const *arr[] = {&d->x1, &d->xi1, &d->dx1, &d->dxi1, &d->idx1, &d->idxi1,
                                &d->x2, &d->xi2, &d->dx2, &d->dxi2, &d->idx2, &d->idxi2,
                                &d->x3, &d->xi3, &d->dx3, &d->dxi3, &d->idx3, &d->idxi3};
  for (auto a : arr)
    if (!*a) return HG_ERR_ARG;
  return HG_OK;
}

extern "C" int hg_create(const hg_desc *desc, hg_ctx **out) {
  if (!out) return HG_ERR_ARG;
  *out = nullptr;
  int st = validate(desc);
  if (st) return st;
  hg_ctx *c = new (std::nothrow) hg_ctx();
  if (!c) return HG_ERR_NOMEM;
  c->desc = *desc;
  c->device = desc->device;
  auto fail = [&](int code) {
    hg_destroy(c);
    return code;
  };
  if (hipSetDevice(c->device) != hipSuccess) return fail(HG_ERR_DEVICE);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(HG_ERR_DEVICE);
  Dev &D = c->D;
  std::memset(&D, 0, sizeof(D));
  c->is = desc->is;
  c->js = desc->js;
  c->ks = desc->ks;
  D.n1 = desc->ie - desc->is + 1;
  D.n2 = desc->je - desc->js + 1;
  D.n3 = desc->ke - desc->ks + 1;
  c->nxg = D.n1 + 4;
  c->nyg = D.n2 + 4;
  c->nzg = D.n3 + 4;
  D.ncell = (long)D.n1 * D.n2 * D.n3;
  D.solve[0] = D.n1 > 1;
  D.solve[1] = D.n2 > 1;
  D.solve[2] = D.n3 > 1;
  D.dim = D.solve[0] + D.solve[1] + D.solve[2];
  D.mag = desc->mag_on ? 1 : 0;
  D.limiter = desc->flux_limiter;
  D.gamma = desc->gamma;
  D.courant = desc->courant;
  D.nv = D.mag ? 8 : 5;
  D.i9wv = -1;
  if (D.mag && D.dim > 1) {
    D.nv = 9;
    D.i9wv = 8;
  }
  D.nw = D.mag ? 10 : 6;
  D.mixed = desc->precision == HG_PREC_FP32MIX;
  D.damp = std::exp(-desc->alpha9wave * desc->courant);  // source.f90:253
  // coordinates: 6 arrays x 3 axes, each n+4 long
  const int nn[3] = {D.n1, D.n2, D.n3};
  size_t tot = 0;
  for (int a = 0; a < 3; ++a) tot += 6 * (size_t)(nn[a] + 4);
  if ((st = dev_alloc(c, tot * sizeof(double), (void **)&c->coord))) return fail(st);
  {
    std::vector<double> h(tot);
    const double *src[3][6] = {
        {desc->x1, desc->xi1, desc->dx1, desc->dxi1, desc->idx1, desc->idxi1},
        {desc->x2, desc->xi2, desc->dx2, desc->dxi2, desc->idx2, desc->idxi2},
        {desc->x3, desc->xi3, desc->dx3, desc->dxi3, desc->idx3, desc->idxi3}};
    size_t off = 0;
    for (int a = 0; a < 3; ++a) {
      Axis &A = D.ax[a];
      const double *dp[6];
      for (int q = 0; q < 6; ++q) {
        std::memcpy(&h[off], src[a][q], sizeof(double) * (nn[a] + 4));
        dp[q] = c->coord + off + 2;  // element 2 holds interior cell 0
        off += nn[a] + 4;
      }
      A.x = dp[0];
      A.xi = dp[1];
      A.dx = dp[2];
      A.dxi = dp[3];
      A.idx = dp[4];
      A.idxi = dp[5];
      A.n = nn[a];
      A.nlines = (unsigned)(D.ncell / nn[a]);
      // x lines form one sweep sequence, y and z lines one per column i
      A.nsub = a == 0 ? A.nlines : A.nlines / (unsigned)D.n1;
      A.stride = a == 0 ? 1 : (a == 1 ? (long)D.n1 : (long)D.n1 * D.n2);
      A.sc_lo = desc->bc[2 * a];
      A.sc_hi = desc->bc[2 * a + 1];
      A.vc_lo = desc->bc[6 + 2 * a];
      A.vc_hi = desc->bc[6 + 2 * a + 1];
      A.ext = A.sc_lo >= 4 || A.sc_hi >= 4 || A.vc_lo >= 4 || A.vc_hi >= 4;
    }
    if (hipMemcpyAsync(c->coord, h.data(), tot * sizeof(double), hipMemcpyHostToDevice,
                       c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
      return fail(HG_ERR_DEVICE);
    // mixed-precision limiter geometry, indices -2..n+1 per axis
    for (int a = 0; a < 3; ++a) {
      const int m = nn[a] + 4;
      std::vector<double> g4(4 * (size_t)m, 0.0), gw(m, 0.0);
      const double *x = src[a][0], *xi = src[a][1], *idx = src[a][4];
      for (int q = 1; q < m - 1; ++q) {  // array element q holds cell q-2
        const double xm = x[q - 1], x0 = x[q], xp = x[q + 1], xim = xi[q - 1], xi0 = xi[q];
        g4[4 * q + 0] = (xp - x0) / (xi0 - x0);
        g4[4 * q + 1] = (x0 - xm) / (x0 - xim);
        g4[4 * q + 2] = (xp - x0) / (x0 - xm);
        g4[4 * q + 3] = 1.0 / (xp - x0);
        gw[q] = idx[q] * idx[q + 1] / (idx[q] + idx[q + 1]);
      }
      double *dg = nullptr, *dw = nullptr;
      if ((st = dev_alloc(c, sizeof(double) * 4 * m, (void **)&dg))) return fail(st);
      if ((st = dev_alloc(c, sizeof(double) * m, (void **)&dw))) return fail(st);
      if (hipMemcpyAsync(dg, g4.data(), sizeof(double) * 4 * m, hipMemcpyHostToDevice,
                         c->stream) != hipSuccess ||
          hipMemcpyAsync(dw, gw.data(), sizeof(double) * m, hipMemcpyHostToDevice, c->stream) !=
              hipSuccess ||
          hipStreamSynchronize(c->stream) != hipSuccess)
        return fail(HG_ERR_DEVICE);
      D.ax[a].limg = reinterpret_cast<const double4 *>(dg) + 2;
      D.ax[a].mmw = dw + 2;
    }
  }
  const size_t fb = sizeof(double) * (size_t)D.ncell;
  // d, e and B are the same numbers in primitive and conserved form
  // (rungekutta.f90:193-232 copies them), so those W fields alias U and the
  // RK update writes them once
  auto aliased = [&](int w) { return w == WD || w == WE || (D.mag && w >= WB1 && w <= WB3); };
  for (int w = 0; w < D.nw; ++w)
    if (!aliased(w))
      if ((st = dev_alloc(c, fb, (void **)&D.W[w]))) return fail(st);
  for (int v = 0; v < D.nv; ++v) {
    if ((st = dev_alloc(c, fb, (void **)&D.U[v]))) return fail(st);
    if ((st = dev_alloc(c, fb, (void **)&D.U0[v]))) return fail(st);
    for (int a = 0; a < 3; ++a)
      if (D.solve[a])
        if ((st = dev_alloc(c, fb, (void **)&D.DU[a][v]))) return fail(st);
  }
  D.nspc = desc->compswitch == 2 ? desc->spn : 0;
  if (D.nspc) {
    const size_t sb = fb * (size_t)D.nspc;
    if ((st = dev_alloc(c, sb, (void **)&D.SPC))) return fail(st);
    if ((st = dev_alloc(c, sb, (void **)&D.SPC0))) return fail(st);
    if ((st = dev_alloc(c, sb * 3, (void **)&D.DUS))) return fail(st);
    if (hipMemsetAsync(D.SPC, 0, sb, c->stream) != hipSuccess) return fail(HG_ERR_DEVICE);
    for (int a = 0; a < 3; ++a)
      if (D.solve[a]) {
        const long nl = D.ncell / nn[a];
        if ((st = dev_alloc(c, sizeof(double) * (size_t)nl * (nn[a] + 1), (void **)&D.FM[a])))
          return fail(st);
      }
  }
  D.W[WD] = D.U[0];
  D.W[WE] = D.U[4];
  if (D.mag)
    for (int a = 0; a < 3; ++a) D.W[WB1 + a] = D.U[5 + a];
  for (int w = 0; w < D.nw; ++w)
    if (hipMemsetAsync(D.W[w], 0, fb, c->stream) != hipSuccess) return fail(HG_ERR_DEVICE);
  D.nblk_cfl = (int)((D.ncell + 255) / 256);
  D.sweep_bw = 4;
  for (int a = 0; a < 3; ++a)
    D.waves[a] = D.solve[a] ? sweep_waves(a == 0 ? 1 : D.n1, D.ax[a].nsub, nn[a]) : 0;
  // y and z wavefront ranges start on a multiple of 16 (the largest block),
  // so a sweep block never holds two directions
  D.woff[0] = 0;
  D.woff[1] = (D.waves[0] + 15) / 16 * 16;
  D.woff[2] = D.woff[1] + (D.waves[1] + 15) / 16 * 16;
  {
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) !=
        hipSuccess)
      lds = 0;
    D.lds_max = lds;
  }
  if ((st = dev_alloc(c, sizeof(double) * 2 * D.nblk_cfl, (void **)&D.part))) return fail(st);
  c->part_cap = D.nblk_cfl;
  D.fdir = -1;
  if (desc->gravswitch == 3) {
    D.grav = 1;
    D.fourpig = 4.0 * std::acos(-1.0) * desc->G;
    const size_t gb = sizeof(double) * (size_t)(D.n1 + 2) * (D.n2 + 2) * (D.n3 + 2);
    for (int q = 0; q < 2; ++q) {
      if ((st = dev_alloc(c, gb, (void **)&D.gphi[q]))) return fail(st);
      if (hipMemsetAsync(D.gphi[q], 0, gb, c->stream) != hipSuccess) return fail(HG_ERR_DEVICE);
    }
    if ((st = dev_alloc(c, fb, (void **)&D.gpsi))) return fail(st);
    if ((st = dev_alloc(c, fb, (void **)&D.gpsi1))) return fail(st);
    if (hipMemsetAsync(D.gpsi, 0, fb, c->stream) != hipSuccess) return fail(HG_ERR_DEVICE);
    D.gcur = 0;
    // dtg_unit = min over cells of dtgrav_cell with cgrav = 1 (tools.f90:154-184)
    const double o1 = D.solve[0] ? 1.0 : 0.0, o2 = D.solve[1] ? 1.0 : 0.0,
                 o3 = D.solve[2] ? 1.0 : 0.0;
    double m = INFINITY;
    for (int k = 0; k < D.n3; ++k)
      for (int j = 0; j < D.n2; ++j)
        for (int i = 0; i < D.n1; ++i) {
          const double x1 = desc->dxi1[i + 2], x2 = desc->dxi2[j + 2], x3 = desc->dxi3[k + 2];
          const double dvol = x1 * x2 * x3;
          const double sa1 = x2 * x3, sa2 = x1 * x3, sa3 = x1 * x2;
          const double v =
              dvol / (1.0 * (o1 * (sa1 + sa1) + o2 * (sa2 + sa2) + o3 * (sa3 + sa3)));
          m = v < m ? v : m;
        }
    c->dtg_unit = m;
  }
  // ghost values and boundary fluxes of the faces with codes 4, 5, 9, 10:
  // G starts at the reference's allocation values (allocations.f90:71-78:
  // d = p = 1, v = B = phi = 0), FX at zero
  for (int a = 0; a < 3; ++a) {
    Axis &A = D.ax[a];
    for (int side = 0; side < 2; ++side) {
      const int sc = side ? A.sc_hi : A.sc_lo, vc = side ? A.vc_hi : A.vc_lo;
      if (sc < 4 && vc < 4) continue;
      const size_t nl = A.nlines, ng = (size_t)NGF * 2 * nl;
      std::vector<double> g(ng, 0.0);
      std::fill(g.begin() + GF_D * 2 * nl, g.begin() + (GF_D + 1) * 2 * nl, 1.0);
      std::fill(g.begin() + GF_P * 2 * nl, g.begin() + (GF_P + 1) * 2 * nl, 1.0);
      double *gp = nullptr, *fx = nullptr;
      if ((st = dev_alloc(c, sizeof(double) * ng, (void **)&gp))) return fail(st);
      if ((st = dev_alloc(c, sizeof(double) * NVMAX, (void **)&fx))) return fail(st);
      const double zero[NVMAX] = {};
      if (hipMemcpyAsync(gp, g.data(), sizeof(double) * ng, hipMemcpyHostToDevice,
                         c->stream) != hipSuccess ||
          hipMemcpyAsync(fx, zero, sizeof(zero), hipMemcpyHostToDevice, c->stream) !=
              hipSuccess ||
          hipStreamSynchronize(c->stream) != hipSuccess)
        return fail(HG_ERR_DEVICE);
      A.G[side] = gp;
      A.FX[side] = fx;
    }
  }
  if ((st = dev_alloc(c, sizeof(Scalars), (void **)&D.sc))) return fail(st);
  // Every initialisation write goes to the context's own (non-blocking)
  // stream: work on the null stream is not ordered with it, and a zeroing
  // memset landing after the error-key clear would read as an error.
  if (hipMemsetAsync(D.sc, 0, sizeof(Scalars), c->stream) != hipSuccess)
    return fail(HG_ERR_DEVICE);
  launch_clear_error(D, c->stream);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(HG_ERR_DEVICE);
  if (hipHostMalloc((void **)&c->h_sc, sizeof(Scalars), hipHostMallocDefault) != hipSuccess)
    return fail(HG_ERR_NOMEM);
  if (hipHostMalloc((void **)&c->h_stage, fb, hipHostMallocDefault) != hipSuccess)
    return fail(HG_ERR_NOMEM);
  *out = c;
  return HG_OK;
}

extern "C" int hg_destroy(hg_ctx *c) {
  if (!c) return HG_OK;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->gstream) hipStreamSynchronize(c->gstream);
  for (hipEvent_t e : c->ev) hipEventDestroy(e);
  if (c->gdone) hipEventDestroy(c->gdone);
  if (c->gstream) hipStreamDestroy(c->gstream);
  for (void *p : c->allocs) hipFree(p);
  if (c->h_sc) hipHostFree(c->h_sc);
  if (c->h_stage) hipHostFree(c->h_stage);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return HG_OK;
}

extern "C" int hg_ufnmax(const hg_ctx *c) { return c ? c->D.nv : 0; }

static double *field_dev(hg_ctx *c, int field, bool upload) {
  Dev &D = c->D;
  switch (field) {
    case HG_F_D: return D.W[WD];
    case HG_F_V1: return D.W[WV1];
    case HG_F_V2: return D.W[WV2];
    case HG_F_V3: return D.W[WV3];
    case HG_F_E: return D.W[WE];
    case HG_F_B1: return D.mag ? D.W[WB1] : nullptr;
    case HG_F_B2: return D.mag ? D.W[WB2] : nullptr;
    case HG_F_B3: return D.mag ? D.W[WB3] : nullptr;
    case HG_F_PHI: return D.mag ? D.W[WPHI] : nullptr;
    case HG_F_EINT: return upload ? nullptr : D.W[WEINT];
    case HG_F_P: return upload ? nullptr : D.W[WEINT];
    case HG_F_GRVPSI: return D.grav ? D.gpsi : nullptr;
    default:
      if (field >= HG_F_U && field < HG_F_U + D.nv) return D.U[field - HG_F_U];
      if (field >= HG_F_SPC && field < HG_F_SPC + D.nspc)
        return D.SPC + (size_t)(field - HG_F_SPC) * D.ncell;
      return nullptr;
  }
}

// potential: interior plus one ghost layer, into both stage buffers
static int upload_grvphi(hg_ctx *c, const double *a) {
  const Dev &D = c->D;
  const size_t n = (size_t)(D.n1 + 2) * (D.n2 + 2) * (D.n3 + 2);
  std::vector<double> h(n);
  size_t q = 0;
  for (int k = -1; k <= D.n3; ++k)
    for (int j = -1; j <= D.n2; ++j)
      for (int i = -1; i <= D.n1; ++i)
        h[q++] = a[((size_t)(k + 2) * c->nyg + (j + 2)) * c->nxg + (i + 2)];
  for (int b = 0; b < 2; ++b)
    HIPCHK(hipMemcpyAsync(D.gphi[b], h.data(), n * sizeof(double), hipMemcpyHostToDevice,
                          c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->D.gcur = 0;
  return HG_OK;
}

static int batch_member_current(hg_ctx *c);

extern "C" int hg_upload(hg_ctx *c, int32_t field, const double *a) {
  if (!c || !a) return HG_ERR_ARG;
  if (int st = batch_member_current(c)) return st;
  if (field == HG_F_GRVPHI || field == HG_F_GRVPSI) {
    if (!c->D.grav) return HG_ERR_ARG;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (field == HG_F_GRVPHI) return upload_grvphi(c, a);
  }
  if (!c->D.mag && field >= HG_F_B1 && field <= HG_F_PHI) return HG_OK;  // B == 0
  double *dst = field_dev(c, field, true);
  if (!dst) return HG_ERR_ARG;
  HIPCHK(hipSetDevice(c->device));
  const Dev &D = c->D;
  double *h = c->h_stage;
  HIPCHK(hipStreamSynchronize(c->stream));  // staging buffer reuse
  long q = 0;
  for (int k = 0; k < D.n3; ++k)
    for (int j = 0; j < D.n2; ++j) {
      const double *row = a + ((size_t)(k + 2) * c->nyg + (j + 2)) * c->nxg + 2;
      std::memcpy(h + q, row, sizeof(double) * D.n1);
      q += D.n1;
    }
  HIPCHK(hipMemcpyAsync(dst, h, sizeof(double) * D.ncell, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (field < HG_F_EINT) c->eint_stale = true;
  return HG_OK;
}

static int check_error(hg_ctx *c) {
  HIPCHK(hipMemcpyAsync(c->h_sc, c->D.sc, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  unsigned long long key = c->h_sc->err;
  if (key == NOERR) return HG_OK;
  const Dev &D = c->D;
  unsigned long long cell = (key - 1) & ((1ull << 42) - 1);
  unsigned long long cls = (key - 1) >> 42;
  int dir = (int)(cls % 4);
  int phase = (int)((cls / 4) % 2);
  long n1 = D.n1 + 4, n2 = D.n2 + 4;
  c->err_i = (int)(cell % n1) - 2 + c->is;
  c->err_j = (int)((cell / n1) % n2) - 2 + c->js;
  c->err_k = (int)(cell / (n1 * n2)) - 2 + c->ks;
  c->err_dir = dir;
  // leave the device flag set: the state is invalid from here on
  return phase == 0 ? HG_ERR_FACE_EOS : HG_ERR_NEG_EINT;
}

static int ensure_eint(hg_ctx *c) {
  if (!c->eint_stale) return HG_OK;
  launch_pressure(c->D, c->stream);
  HIPCHK(hipGetLastError());
  c->eint_stale = false;
  return check_error(c);
}

extern "C" int hg_download(hg_ctx *c, int32_t field, double *a) {
  if (!c || !a) return HG_ERR_ARG;
  if (int st = batch_member_current(c)) return st;
  HIPCHK(hipSetDevice(c->device));
  const Dev &D = c->D;
  if (!D.mag && field >= HG_F_B1 && field <= HG_F_PHI) {
    for (int k = 0; k < D.n3; ++k)
      for (int j = 0; j < D.n2; ++j) {
        double *row = a + ((size_t)(k + 2) * c->nyg + (j + 2)) * c->nxg + 2;
        std::memset(row, 0, sizeof(double) * D.n1);
      }
    return HG_OK;
  }
  if (field == HG_F_GRVPHI) {
    if (!D.grav) return HG_ERR_ARG;
    HIPCHK(hipStreamSynchronize(c->stream));
    const size_t n = (size_t)(D.n1 + 2) * (D.n2 + 2) * (D.n3 + 2);
    std::vector<double> h(n);
    HIPCHK(hipMemcpyAsync(h.data(), D.gphi[D.gcur], n * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < D.n3; ++k)
      for (int j = 0; j < D.n2; ++j)
        for (int i = 0; i < D.n1; ++i)
          a[((size_t)(k + 2) * c->nyg + (j + 2)) * c->nxg + (i + 2)] = h[gidx(D, i, j, k)];
    return HG_OK;
  }
  double *src = field_dev(c, field, false);
  if (!src) return HG_ERR_ARG;
  int st = ensure_eint(c);
  if (st && st != HG_ERR_NEG_EINT) return st;
  double *h = c->h_stage;
  HIPCHK(hipMemcpyAsync(h, src, sizeof(double) * D.ncell, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  const double gm1 = c->desc.gamma - 1.0;
  long q = 0;
  for (int k = 0; k < D.n3; ++k)
    for (int j = 0; j < D.n2; ++j) {
      double *row = a + ((size_t)(k + 2) * c->nyg + (j + 2)) * c->nxg + 2;
      if (field == HG_F_P)
        for (int i = 0; i < D.n1; ++i) row[i] = gm1 * h[q + i];  // eos_p
      else
        std::memcpy(row, h + q, sizeof(double) * D.n1);
      q += D.n1;
    }
  return HG_OK;
}

// ---- profiling -----------------------------------------------------------------
// Events for the next timed launch of the given kind (nullptr when profiling
// is off); the launch records both on its own dispatch.
static const KEv *prof_ev(hg_ctx *c, int kind) {
  if (!c->prof || !(c->prof_mask & (1u << kind))) return nullptr;
  if ((size_t)(2 * c->ev_used + 2) > c->ev.size()) {
    for (int q = 0; q < 64; ++q) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      c->ev.push_back(e);
    }
  }
  if ((size_t)c->ev_used >= c->ev_kind.size()) c->ev_kind.resize(c->ev_used + 64);
  c->ev_kind[c->ev_used] = kind;
  c->kev = KEv{c->ev[2 * c->ev_used], c->ev[2 * c->ev_used + 1]};
  ++c->ev_used;
  return &c->kev;
}

// prof_ev's pair copied into `slot`: prof_ev hands out one KEv per context,
// which the next call overwrites
static const KEv *prof_ev_into(hg_ctx *c, int kind, KEv &slot) {
  const KEv *e = prof_ev(c, kind);
  if (!e) return nullptr;
  slot = *e;
  return &slot;
}

// after the stream has been synchronised: fold recorded pairs into totals
static void prof_collect(hg_ctx *c) {
  for (int q = 0; q < c->ev_used; ++q) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev[2 * q], c->ev[2 * q + 1]) == hipSuccess) {
      c->prof_ms[c->ev_kind[q]] += ms;
      c->prof_n[c->ev_kind[q]] += 1;
    } else {
      // a pair that was never recorded is dropped, and so is its status: HIP
      // keeps the last failed call's status for the next hipGetLastError,
      // where it would surface as a failed launch (HG_ERR_DEVICE) of a later
      // step
      (void)hipGetLastError();
    }
  }
  c->ev_used = 0;
}

extern "C" int hg_profile(hg_ctx *c, int32_t enable) {
  if (!c) return HG_ERR_ARG;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  prof_collect(c);
  c->prof = enable != 0;
  c->prof_mask = (enable & HG_PROF_ONLY) ? (unsigned)enable & ((1u << HG_PROF_N) - 1)
                                         : (1u << HG_PROF_N) - 1;
  for (int k = 0; k < HG_PROF_N; ++k) {
    c->prof_ms[k] = 0.0;
    c->prof_n[k] = 0;
  }
  return HG_OK;
}

extern "C" int hg_profile_read(const hg_ctx *c, int32_t kind, double *ms_total,
                               int64_t *launches, double *bytes_per_launch) {
  if (!c || kind < 0 || kind >= HG_PROF_N) return HG_ERR_ARG;
  const Dev &D = c->D;
  const double cell = (double)D.ncell * sizeof(double);
  // the W fields that are not U (d, e and B alias U)
  const int wr = D.nw - (D.mag ? 5 : 2);
  // directions the sweep launches cover: every solved one, or all but the
  // fused one
  const int nsw = D.dim - (c->fused ? 1 : 0);
  double bytes = 0.0;
  switch (kind) {
    case HG_PROF_SWEEP:  // read the primitives once and write one divergence per direction
      bytes = cell * (double)nsw * (D.nw + D.nv);
      break;
    case HG_PROF_RK: {
      // read U, U0 and the swept divergences, write U and the W fields that
      // are not U (stage 1 writes U0 and reads none: the same count); the
      // fused launch also reads its direction's primitives instead of a
      // divergence, and with its three state sets stage 1 neither reads
      // nor writes a U0 (averaged over the stages)
      const int rk = c->desc.rktype;
      bytes = cell * (c->fused ? D.nw + nsw * D.nv + 2.0 * D.nv +
                                     (double)D.nv * (rk - 1) / rk + wr
                               : (double)(D.dim * D.nv + 3 * D.nv + wr));
      break;
    }
    case HG_PROF_CFL:
      bytes = cell * (double)(D.nw - (D.mag ? 2 : 1));
      break;
    case HG_PROF_GRAV:  // per sub-step launch (grktype 2: both stages in one): phi,
                        // psi, rho read, phi and psi written
      bytes = cell * 5.0;
      break;
    default:
      bytes = 0.0;
  }
  if (ms_total) *ms_total = c->prof_ms[kind];
  if (launches) *launches = c->prof_n[kind];
  if (bytes_per_launch) *bytes_per_launch = bytes;
  return HG_OK;
}

// Back-to-back sweeps between two events: the sweep kernel's own duration
// without per-dispatch events (which add a completion wait to every launch).
// The sweep only reads the primitives and overwrites the flux divergences,
// which the next RK stage recomputes first, so the state is unchanged.  The
// stop flag is lifted for the replay and put back.
extern "C" int hg_time_sweep(hg_ctx *c, int32_t launches, double *ms_total) {
  if (!c || launches <= 0 || !ms_total) return HG_ERR_ARG;
  HIPCHK(hipSetDevice(c->device));
  int stop = 0;
  int *dstop = &c->D.sc->stop;
  HIPCHK(hipMemcpyAsync(&stop, dstop, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  const int zero = 0;
  HIPCHK(hipMemcpyAsync(dstop, &zero, sizeof(int), hipMemcpyHostToDevice, c->stream));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    hipEventDestroy(e0);
    return HG_ERR_DEVICE;
  }
  int st = HG_OK;
  if (hipEventRecord(e0, c->stream) != hipSuccess) st = HG_ERR_DEVICE;
  for (int q = 0; q < launches && !st; ++q) launch_sweep(c->D, 1, c->stream);
  if (!st && (hipGetLastError() != hipSuccess || hipEventRecord(e1, c->stream) != hipSuccess))
    st = HG_ERR_DEVICE;
  if (!st && hipMemcpyAsync(dstop, &stop, sizeof(int), hipMemcpyHostToDevice, c->stream) != hipSuccess)
    st = HG_ERR_DEVICE;
  if (!st && hipStreamSynchronize(c->stream) != hipSuccess) st = HG_ERR_DEVICE;
  float ms = 0.f;
  if (!st && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) st = HG_ERR_DEVICE;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  *ms_total = ms;
  return st;
}

// ---- the hot-path entry points ------------------------------------------------
static void grav_speeds(const hg_ctx *c, double dt, double cfmax, double *cgrav,
                        double *dtgrav);

extern "C" int hg_timestep(hg_ctx *c, double time, double t_out, double t_end,
                           int32_t outstyle, int32_t endstyle, double fixed_dt, double *dt_out,
                           double *ch_out, double *cgrav_out, double *dtgrav_out) {
  if (!c) return HG_ERR_ARG;
  HIPCHK(hipSetDevice(c->device));
  if (fixed_dt > 0.0) {
    if (dt_out) *dt_out = fixed_dt;
    if (ch_out) *ch_out = c->ch;
    return HG_OK;
  }
  int st = ensure_eint(c);
  if (st) return st;
  const Dev &D = c->D;
  launch_set_clock(D, time, t_out, t_end, 0.0, outstyle, endstyle, 0, 0, 1, 1, c->stream);
  launch_cfl(D, c->stream, prof_ev(c, HG_PROF_CFL));
  launch_cfl_final(D, c->stream);
  HIPCHK(hipGetLastError());
  if ((st = check_error(c))) return st;
  prof_collect(c);
  c->ch = c->h_sc->ch;
  if (dt_out) *dt_out = c->h_sc->dt;
  if (ch_out) *ch_out = c->h_sc->ch;
  double cgrav = 0.0, dtgrav = 0.0;
  if (c->D.grav) grav_speeds(c, c->h_sc->dt, c->h_sc->cfmax, &cgrav, &dtgrav);
  if (cgrav_out) *cgrav_out = cgrav;
  if (dtgrav_out) *dtgrav_out = dtgrav;
  return HG_OK;
}

// gravity_hyperbolic (gravity_hyperbolic.f90:19-50) with gsrc = rho
// (gravity.f90:114-150): tngrav = ceiling((time+dt-grvtime)/dtgrav) RK
// sub-steps of the wave equation, then grvtime and cgrav_old advance.
// The sub-steps of one gravity call, enqueued on the context's stream:
// tngrav, the sub-step length and every sub-step's psi damping factor are
// host arithmetic on (time, dt, grvtime, cgrav, cgrav_old, dtgrav), in the
// reference's order.  Returns tngrav, or 0 when time + dt does not lie
// beyond grvtime.
static int enqueue_gravity(hg_ctx *c, double time, double dt, double *grvtime, double cgrav,
                           double *cgrav_old, double *dtgrav, hipStream_t stream) {
  const int tngrav = (int)std::ceil((time + dt - *grvtime) / *dtgrav);
  if (tngrav < 1) return 0;
  *dtgrav = (time + dt - *grvtime) / (double)tngrav;
  const int gk = c->desc.grktype;
  for (int l = 1; l <= tngrav; ++l) {
    const double damp =
        std::exp(-(1.0 - *cgrav_old / cgrav + c->desc.alphagrv * c->desc.hgcfl));
    // one launch per sub-step (both stages of grktype 2 in one); the new
    // values are in the second buffers
    launch_grav_substep(c->D, gk, *dtgrav, cgrav, damp, stream, prof_ev(c, HG_PROF_GRAV));
    c->D.gcur = 1 - c->D.gcur;
    std::swap(c->D.gpsi, c->D.gpsi1);
    *grvtime = *grvtime + *dtgrav;
    *cgrav_old = cgrav;
  }
  return tngrav;
}

// cgrav and dtgrav of timestep.f90:92-97 from the step's dt and the
// largest fast speed cfmax.
static void grav_speeds(const hg_ctx *c, double dt, double cfmax, double *cgrav,
                        double *dtgrav) {
  double cg = c->desc.HGfac * cfmax;
  if (c->desc.maxtngrv > 0) {
    const double lim = c->desc.hgcfl * c->dtg_unit * (double)c->desc.maxtngrv / dt;
    cg = lim < cg ? lim : cg;
  }
  const double b = c->desc.hgcfl * c->dtg_unit / cg;
  *cgrav = cg;
  *dtgrav = b < dt ? b : dt;
}

extern "C" int hg_gravity(hg_ctx *c, double time, double dt, double *grvtime, double cgrav,
                          double *cgrav_old, double *dtgrav, int32_t *nsub) {
  if (!c || !grvtime || !cgrav_old || !dtgrav) return HG_ERR_ARG;
  if (!c->D.grav) return HG_ERR_STATE;
  HIPCHK(hipSetDevice(c->device));
  int st = ensure_eint(c);
  if (st) return st;
  const int tngrav = enqueue_gravity(c, time, dt, grvtime, cgrav, cgrav_old, dtgrav, c->stream);
  if (tngrav < 1) return HG_ERR_ARG;
  HIPCHK(hipGetLastError());
  if (nsub) *nsub = tngrav;
  HIPCHK(hipStreamSynchronize(c->stream));
  prof_collect(c);
  return HG_OK;
}

// The part of RK stage 1 that needs neither the step's dt nor the gravity
// potential of this step: the flux sweep (of every direction, or of the
// directions before the fused one) and the species pass.  It only reads the
// primitives and ch, which the previous step's end has already set.
static void enqueue_stage1_sweep(hg_ctx *c) {
  const Dev &D = c->D;
  if (c->fused) {
    RKArgs a{};
    launch_fused_stage(D, 1, a, c->stream, nullptr, nullptr, 1);
    return;
  }
  launch_sweep(D, 1, c->stream);
  launch_species(D, c->stream);
}

// The RK stages of one step.  Returns the number of CFL partials the last
// stage wrote (fuse_tail): one per 256 cells, or one per block of the fused
// last sweep.  sweep1_done: stage 1's sweep was enqueued beforehand
// (enqueue_stage1_sweep).
static int enqueue_stages(hg_ctx *c, bool fuse_tail, double damp, bool sweep1_done = false) {
  const Dev &D = c->D;
  const int rk = c->desc.rktype;
  int npart = D.nblk_cfl;
  // the fused form's three state sets (current, second, third)
  StateSet cur{}, alt{}, third{};
  if (c->fused) {
    for (int v = 0; v < D.nv; ++v) {
      cur.U[v] = c->D.U[v];
      alt.U[v] = c->alt_U[v];
      third.U[v] = c->D.U0[v];
    }
    for (int w = 0; w < D.nw; ++w) {
      cur.W[w] = c->D.W[w];
      alt.W[w] = c->alt_W[w];
      third.W[w] = c->alt2_W[w];
    }
  }
  for (int stage = 1; stage <= rk; ++stage) {
    double faco, fact, facn;
    runge_coeff(stage, rk, &faco, &fact, &facn);
    const bool pre = sweep1_done && stage == 1;
    if (c->fused) {
      const bool last = stage == rk;
      RKArgs a{faco, fact, facn, damp, fuse_tail && last, fuse_tail && last && damp > 0.0,
               D.grav};
      // a pair for the sweep of the directions before the fused one only
      // when there are such directions (every pair taken must be recorded)
      KEv k0{}, k1{};
      const long npre = D.fdir == 0 ? D.woff[2] - D.woff[1] + D.waves[2] : D.woff[D.fdir];
      const KEv *e0 = npre > 0 && !pre ? prof_ev_into(c, HG_PROF_SWEEP, k0) : nullptr;
      const KEv *e1 = prof_ev_into(c, HG_PROF_RK, k1);
      // three state sets: stage 1 reads the current set and writes the
      // second; each later stage writes whichever of the second and the
      // third it does not read, and takes the step's starting values u0
      // from the current set, which no stage of the step writes, so stage
      // 1 stores no U0 copy
      const StateSet &in = stage == 1 ? cur : (stage % 2 == 0 ? alt : third);
      const StateSet &out = stage % 2 == 1 ? alt : third;
      Dev Df = D;
      for (int v = 0; v < D.nv; ++v) {
        Df.U[v] = in.U[v];
        Df.Uo[v] = out.U[v];
        Df.U0[v] = stage == 1 ? nullptr : cur.U[v];
      }
      for (int w = 0; w < D.nw; ++w) {
        Df.W[w] = in.W[w];
        Df.Wo[w] = out.W[w];
      }
      const int g = launch_fused_stage(Df, stage, a, c->stream, e0, e1, pre ? 2 : 0);
      if (last && fuse_tail) npart = g;
      continue;
    }
    if (!pre) {
      launch_sweep(D, stage, c->stream, prof_ev(c, HG_PROF_SWEEP));
      launch_species(D, c->stream);
    }
    const bool last = stage == rk;
    launch_rk(D, stage, faco, fact, facn, fuse_tail && last, fuse_tail && last && damp > 0.0,
              damp, c->stream, prof_ev(c, HG_PROF_RK));
  }
  if (c->fused) {
    // the last stage's output becomes the current set; the step's starting
    // set is free again and takes the role its slot had
    const bool in_alt = rk % 2 == 1;
    const StateSet &ncur = in_alt ? alt : third;
    const StateSet &nalt = in_alt ? cur : alt;
    const StateSet &nthird = in_alt ? third : cur;
    for (int v = 0; v < D.nv; ++v) {
      c->D.U[v] = ncur.U[v];
      c->alt_U[v] = nalt.U[v];
      c->D.U0[v] = nthird.U[v];
    }
    for (int w = 0; w < D.nw; ++w) {
      c->D.W[w] = ncur.W[w];
      c->alt_W[w] = nalt.W[w];
      c->alt2_W[w] = nthird.W[w];
    }
  }
  return npart;
}

extern "C" int hg_hydro_step(hg_ctx *c, double dt, double ch) {
  if (!c) return HG_ERR_ARG;
  HIPCHK(hipSetDevice(c->device));
  int st = ensure_eint(c);
  if (st) return st;
  launch_set_step(c->D, dt, ch, c->stream);
  enqueue_stages(c, false, 0.0);
  HIPCHK(hipGetLastError());
  st = check_error(c);
  prof_collect(c);
  return st;
}

extern "C" int hg_phidamp(hg_ctx *c, double factor) {
  if (!c) return HG_ERR_ARG;
  if (!c->D.mag) return HG_OK;
  HIPCHK(hipSetDevice(c->device));
  launch_phidamp(c->D, factor, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return HG_OK;
}

extern "C" int hg_run(hg_ctx *c, double *time, int64_t *tn, double *t_out, double dt_out,
                      double t_end, int32_t outstyle, int32_t endstyle, int64_t tnlim,
                      int64_t tn_out, int64_t max_steps, double damp_factor,
                      int32_t *stop_reason) {
  if (!c || !time || !tn || !t_out || max_steps < 1) return HG_ERR_ARG;
  if (outstyle == 2 && tn_out < 1) return HG_ERR_ARG;
  if (c->D.grav) return HG_ERR_UNSUPPORTED;  // gravity runs use hg_run_gravity
  HIPCHK(hipSetDevice(c->device));
  int st = ensure_eint(c);
  if (st) return st;
  const Dev &D = c->D;
  const bool damp = D.mag && D.dim > 1 && damp_factor > 0.0;
  launch_set_clock(D, *time, *t_out, t_end, dt_out, outstyle, endstyle, *tn, tnlim, tn_out,
                   max_steps, c->stream);
  launch_cfl(D, c->stream, prof_ev(c, HG_PROF_CFL));
  // Steps are enqueued 16 at a time (never more than max_steps in all) and
  // the host reads the stop flag between groups; steps enqueued after a
  // stop return at once.  Each step ends with one launch that closes it
  // (step_end) and computes the next step's dt and ch (cfl_final).
  const int batch = 16;
  int reason = STOP_NONE;
  int64_t left = max_steps, tn_before = *tn;
  launch_cfl_final(D, c->stream);
  for (;;) {
    const int nb = left < batch ? (int)left : batch;
    if (nb <= 0) {  // the device stops at max_steps; never enqueue past it
      reason = STOP_MAXSTEPS;
      break;
    }
    const int ev_base = c->ev_used;
    for (int s = 0; s < nb; ++s) {
      Dev De = D;  // the step end reduces the partials the last stage wrote
      De.nblk_cfl = enqueue_stages(c, true, damp ? damp_factor : 0.0);
      launch_step_end_cfl(De, c->stream);
    }
    left -= nb;
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->h_sc, D.sc, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    // time only the launches of steps that ran (every step takes the same
    // number of event pairs)
    const int64_t ran = c->h_sc->tn - tn_before;
    tn_before = c->h_sc->tn;
    if (c->prof)
      c->ev_used = ev_base + (int)(std::min<int64_t>(ran, nb) * ((c->ev_used - ev_base) / nb));
    prof_collect(c);
    reason = c->h_sc->stop;
    if (reason != STOP_NONE) break;
  }
  *time = c->h_sc->time;
  *tn = c->h_sc->tn;
  *t_out = c->h_sc->t_out;
  c->ch = c->h_sc->ch;
  if (stop_reason) *stop_reason = reason;
  if (reason == STOP_ERROR) return check_error(c);
  return HG_OK;
}

// The main loop with hyperbolic self-gravity (hormone.f90:124-179 with
// gravswitch 3): per step the device's previous step end has computed dt,
// ch and the largest fast speed; the host reads them (one pinned copy and
// an event, not a stream synchronisation), evaluates cgrav, dtgrav, tngrav
// and the damping factors exactly as hg_timestep and hg_gravity do, and
// enqueues the gravity sub-steps, the RK stages and the step end.  Stage
// 1's flux sweep needs neither dt nor the new potential, so it is enqueued
// before the host waits: the GPU sweeps while the host reads the scalars.
// The gravity sub-steps go to a second stream, after the previous step's
// end: they stream the potential through HBM while the sweep keeps the
// VALUs busy, and stage 1's RK update waits for them.  They read only the
// density of the state the sweep reads too (the update writes the other
// buffer set) and write only the potential, which no sweep reads; the next
// step's sub-steps start after this step's end, so no stage still reads
// the potential they overwrite.
extern "C" int hg_run_gravity(hg_ctx *c, double *time, int64_t *tn, double *t_out,
                              double dt_out, double t_end, int32_t outstyle, int32_t endstyle,
                              int64_t tnlim, int64_t tn_out, int64_t max_steps,
                              double damp_factor, double *grvtime, double *cgrav_old,
                              int64_t *nsub_total, int32_t *stop_reason) {
  if (!c || !time || !tn || !t_out || !grvtime || !cgrav_old || max_steps < 1)
    return HG_ERR_ARG;
  if (outstyle == 2 && tn_out < 1) return HG_ERR_ARG;
  if (!c->D.grav) return HG_ERR_STATE;
  HIPCHK(hipSetDevice(c->device));
  int st = ensure_eint(c);
  if (st) return st;
  const Dev &D = c->D;
  const bool damp = D.mag && D.dim > 1 && damp_factor > 0.0;
  // stage 1's sweep runs while the host waits; it carries no profiling
  // events, so every timed launch is one of a step that ran
  const bool overlap = true;
  hipEvent_t ready;
  HIPCHK(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
  if (!c->gstream) {
    HIPCHK(hipStreamCreateWithFlags(&c->gstream, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&c->gdone, hipEventDisableTiming));
  }
  auto fail = [&](int code) {
    hipStreamSynchronize(c->gstream);
    hipStreamSynchronize(c->stream);
    hipEventDestroy(ready);
    return code;
  };
  auto post = [&]() -> bool {  // scalars to the host, then the next sweep
    if (hipMemcpyAsync(c->h_sc, D.sc, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream) !=
            hipSuccess ||
        hipEventRecord(ready, c->stream) != hipSuccess)
      return false;
    if (overlap) enqueue_stage1_sweep(c);
    return hipGetLastError() == hipSuccess;
  };
  launch_set_clock(D, *time, *t_out, t_end, dt_out, outstyle, endstyle, *tn, tnlim, tn_out,
                   max_steps, c->stream);
  launch_cfl(D, c->stream, prof_ev(c, HG_PROF_CFL));
  launch_cfl_final(D, c->stream);
  if (!post()) return fail(HG_ERR_DEVICE);
  int64_t nsum = 0;
  int reason = STOP_NONE;
  for (;;) {
    if (hipEventSynchronize(ready) != hipSuccess) return fail(HG_ERR_DEVICE);
    if (c->prof) prof_collect(c);
    const Scalars s = *c->h_sc;
    if (s.stop != STOP_NONE) {
      reason = s.stop;
      break;
    }
    double cgrav, dtgrav;
    grav_speeds(c, s.dt, s.cfmax, &cgrav, &dtgrav);
    // gravity.f90:41-60: the potential of step 0 is the initial one
    if (s.tn == 0) {
      *grvtime = 0.0;
    } else {
      hipStream_t gs = c->gstream;
      if (hipStreamWaitEvent(gs, ready, 0) != hipSuccess) return fail(HG_ERR_DEVICE);
      const int ns = enqueue_gravity(c, s.time, s.dt, grvtime, cgrav, cgrav_old, &dtgrav, gs);
      if (ns < 1) return fail(HG_ERR_STATE);
      if (hipEventRecord(c->gdone, gs) != hipSuccess ||
          hipStreamWaitEvent(c->stream, c->gdone, 0) != hipSuccess)
        return fail(HG_ERR_DEVICE);
      nsum += ns;
    }
    Dev De = D;  // the step end reduces the partials the last stage wrote
    De.nblk_cfl = enqueue_stages(c, true, damp ? damp_factor : 0.0, overlap);
    launch_step_end_cfl(De, c->stream);
    if (hipGetLastError() != hipSuccess || !post()) return fail(HG_ERR_DEVICE);
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(HG_ERR_DEVICE);
  hipEventDestroy(ready);
  prof_collect(c);
  *time = c->h_sc->time;
  *tn = c->h_sc->tn;
  *t_out = c->h_sc->t_out;
  c->ch = c->h_sc->ch;
  if (nsub_total) *nsub_total = nsum;
  if (stop_reason) *stop_reason = reason;
  if (reason == STOP_ERROR) return check_error(c);
  return HG_OK;
}

// Launch shape of the sweep kernel for this context: wavefronts per block.
extern "C" int hg_set_launch(hg_ctx *c, int32_t sweep_block_waves) {
  if (!c) return HG_ERR_ARG;
  const int w = sweep_block_waves;
  if (w != 1 && w != 2 && w != 4 && w != 8 && w != 16) return HG_ERR_ARG;
  HIPCHK(hipStreamSynchronize(c->stream));
  c->D.sweep_bw = w;
  return HG_OK;
}

extern "C" int hg_set_ghosts(hg_ctx *c, int32_t field, const double *padded) {
  if (!c || !padded) return HG_ERR_ARG;
  int gf;
  switch (field) {
    case HG_F_D: gf = GF_D; break;
    case HG_F_P: gf = GF_P; break;
    case HG_F_V1: case HG_F_V2: case HG_F_V3: gf = GF_V1 + (field - HG_F_V1); break;
    case HG_F_B1: case HG_F_B2: case HG_F_B3: gf = GF_B1 + (field - HG_F_B1); break;
    case HG_F_PHI: gf = GF_PHI; break;
    default: return HG_ERR_ARG;
  }
  HIPCHK(hipSetDevice(c->device));
  Dev &D = c->D;
  const long nx = c->nxg, ny = c->nyg;
  auto at = [&](long i, long j, long k) { return padded[(i + 2) + nx * ((j + 2) + ny * (k + 2))]; };
  std::vector<double> h;
  for (int a = 0; a < 3; ++a) {
    const Axis &A = D.ax[a];
    for (int side = 0; side < 2; ++side) {
      if (!A.G[side]) continue;
      const long nl = A.nlines;
      h.assign(2 * (size_t)nl, 0.0);
      for (int layer = 0; layer < 2; ++layer) {
        const long g = side ? A.n + layer : -1 - layer;  // ghost index along the axis
        for (long line = 0; line < nl; ++line) {
          double v;
          if (a == 0) v = at(g, line % D.n2, line / D.n2);
          else if (a == 1) v = at(line % D.n1, g, line / D.n1);
          else v = at(line % D.n1, line / D.n1, g);
          h[(size_t)layer * nl + line] = v;
        }
      }
      HIPCHK(hipMemcpyAsync(const_cast<double *>(A.G[side]) + (size_t)gf * 2 * nl, h.data(),
                            sizeof(double) * h.size(), hipMemcpyHostToDevice, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
    }
  }
  return HG_OK;
}

extern "C" int hg_set_boundary_flux(hg_ctx *c, int32_t face, const double *flux) {
  if (!c || !flux || face < 0 || face > 5) return HG_ERR_ARG;
  const Axis &A = c->D.ax[face / 2];
  const int side = face % 2;
  if ((side ? A.sc_hi : A.sc_lo) != 10) return HG_ERR_ARG;
  HIPCHK(hipSetDevice(c->device));
  double f[NVMAX] = {};
  for (int v = 0; v < c->D.nv; ++v) f[v] = flux[v];
  HIPCHK(hipMemcpyAsync(const_cast<double *>(A.FX[side]), f, sizeof(f), hipMemcpyHostToDevice,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return HG_OK;
}

// The fused form's buffers: the second and third state sets (d, e and B
// alias U in every set, as in hg_create; a phi the update does not write,
// 1D MHD, is shared) and one CFL partial per block of direction L's fused
// launch (one wavefront per block at most).
static int ensure_fused_buffers(hg_ctx *c, int L) {
  Dev &D = c->D;
  if (!c->alt_ready) {
    const size_t fb = sizeof(double) * (size_t)D.ncell;
    for (int v = 0; v < D.nv; ++v) {
      int st = dev_alloc(c, fb, (void **)&c->alt_U[v]);
      if (st) return st;
    }
    for (int w = 0; w < D.nw; ++w) {
      if (w == WD) c->alt_W[w] = c->alt_U[0];
      else if (w == WE) c->alt_W[w] = c->alt_U[4];
      else if (D.mag && w >= WB1 && w <= WB3) c->alt_W[w] = c->alt_U[5 + (w - WB1)];
      else if (w == WPHI && D.i9wv < 0) c->alt_W[w] = D.W[w];
      else if (int st = dev_alloc(c, fb, (void **)&c->alt_W[w])) return st;
    }
    c->alt_ready = true;
  }
  if (D.waves[L] > c->part_cap) {
    double *p = nullptr;
    int st = dev_alloc(c, sizeof(double) * 2 * (size_t)D.waves[L], (void **)&p);
    if (st) return st;
    D.part = p;
    c->part_cap = D.waves[L];
  }
  // the third state set: U is D.U0's arrays, W its own fields
  if (!c->alt2_ready) {
    const size_t fb = sizeof(double) * (size_t)D.ncell;
    for (int w = 0; w < D.nw; ++w) {
      if (w == WD) c->alt2_W[w] = D.U0[0];
      else if (w == WE) c->alt2_W[w] = D.U0[4];
      else if (D.mag && w >= WB1 && w <= WB3) c->alt2_W[w] = D.U0[5 + (w - WB1)];
      else if (w == WPHI && D.i9wv < 0) c->alt2_W[w] = D.W[w];
      else if (int st = dev_alloc(c, fb, (void **)&c->alt2_W[w])) return st;
    }
    c->alt2_ready = true;
  }
  return HG_OK;
}

// The context's three state sets (current, second, third) and back.
static void get_sets(const hg_ctx *c, StateSet s[3]) {
  for (int v = 0; v < c->D.nv; ++v) {
    s[0].U[v] = c->D.U[v];
    s[1].U[v] = c->alt_U[v];
    s[2].U[v] = c->D.U0[v];
  }
  for (int w = 0; w < c->D.nw; ++w) {
    s[0].W[w] = c->D.W[w];
    s[1].W[w] = c->alt_W[w];
    s[2].W[w] = c->alt2_W[w];
  }
}

static void put_sets(hg_ctx *c, const StateSet &cur, const StateSet &alt,
                     const StateSet &third) {
  for (int v = 0; v < c->D.nv; ++v) {
    c->D.U[v] = cur.U[v];
    c->alt_U[v] = alt.U[v];
    c->D.U0[v] = third.U[v];
  }
  for (int w = 0; w < c->D.nw; ++w) {
    c->D.W[w] = cur.W[w];
    c->alt_W[w] = alt.W[w];
    c->alt2_W[w] = third.W[w];
  }
}

// A batch member's own record onto the set p of its batch numbering (the
// set that holds its state), the others after it.
static void member_to_set(hg_ctx *c, int p) {
  put_sets(c, c->bsets[p], c->bsets[(p + 1) % 3], c->bsets[(p + 2) % 3]);
}

extern "C" int hg_set_fused(hg_ctx *c, int32_t on) {
  if (!c) return HG_ERR_ARG;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  Dev &D = c->D;
  if (!on) {
    c->fused = false;
    D.fdir = -1;
    return HG_OK;
  }
  // no species; on = 1: the last solved direction (y or z),
  // on = 2: x
  int L = -1;
  if (on == 2) {
    L = D.solve[0] ? 0 : -1;
  } else if (on == 1) {
    for (int a = 1; a < 3; ++a)
      if (D.solve[a]) L = a;
  } else {
    return HG_ERR_ARG;
  }
  if (D.nspc || L < 0) return HG_ERR_UNSUPPORTED;
  for (int a = 0; a < 3; ++a)
    if (D.ax[a].ext) return HG_ERR_UNSUPPORTED;  // fused kernels carry codes 0-3 only
  if (c->in_batch) return HG_ERR_UNSUPPORTED;
  int st = ensure_fused_buffers(c, L);
  if (st) return st;
  D.fdir = L;
  c->fused = true;
  return HG_OK;
}

extern "C" int hg_last_error(const hg_ctx *c, int32_t *i, int32_t *j, int32_t *k,
                             int32_t *dir) {
  if (!c) return HG_ERR_ARG;
  if (i) *i = c->err_i;
  if (j) *j = c->err_j;
  if (k) *k = c->err_k;
  if (dir) *dir = c->err_dir;
  return HG_OK;
}

// values per record of each hg_kat kind (kind + 16: the same counts)
static const int kat_nin[8] = {25, 5, 8, 5, 2, 5, 3, 9}, kat_nout[8] = {9, 1, 1, 1, 2, 4, 2, 1};

static int kat_kind(int32_t which) { return which >= 16 ? which - 16 : which; }

extern "C" int hg_kat(int32_t which, int32_t n, const double *in, double *out) {
  const int w = kat_kind(which);
  if (w < 0 || w > 7 || n < 0 || !in || !out) return HG_ERR_ARG;
  if (n == 0) return HG_OK;
  double *din = nullptr, *dout = nullptr;
  const size_t bi = sizeof(double) * (size_t)n * kat_nin[w];
  const size_t bo = sizeof(double) * (size_t)n * kat_nout[w];
  if (hipMalloc((void **)&din, bi) != hipSuccess) return HG_ERR_NOMEM;
  if (hipMalloc((void **)&dout, bo) != hipSuccess) {
    hipFree(din);
    return HG_ERR_NOMEM;
  }
  int st = HG_OK;
  if (hipMemcpy(din, in, bi, hipMemcpyHostToDevice) != hipSuccess) st = HG_ERR_DEVICE;
  if (!st) {
    launch_kat(which, n, din, dout, nullptr);
    if (hipGetLastError() != hipSuccess) st = HG_ERR_DEVICE;
  }
  if (!st && hipMemcpy(out, dout, bo, hipMemcpyDeviceToHost) != hipSuccess) st = HG_ERR_DEVICE;
  hipFree(din);
  hipFree(dout);
  return st;
}


// =============================================================================
// Batches: many independent boxes advanced by one set of launches per step.
// =============================================================================
struct hg_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<hg_ctx *> members;
  Dev *d_boxes = nullptr;
  int2 *d_sitems = nullptr, *d_citems = nullptr;
  size_t cap_boxes = 0, cap_s = 0, cap_c = 0;
  Batch B{};
  int mag = -1, lim = -1, mixed = -1, rktype = -1;
  bool dirty = true;
  bool use_graph = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // pinned host copies of the members' scalars after the last step, in
  // member order; sc_valid[q] is 0 until member q has been stepped
  Scalars *h_sc = nullptr;
  size_t cap_hsc = 0;
  std::vector<char> sc_valid;
  // launch shapes: sweep block size (wavefronts per block) and grid sizes of
  // the persistent sweep / rk launches (0 = one item per wavefront / block)
  int sweep_block_waves = 4;
  unsigned sweep_grid = 0, rk_grid = 0;
  // the x-fused form (hg_batch_set_fused): taken when every member sweeps
  // x; d_bpar[r] is the state set of rec_members[r], the member order of
  // the device records (the order of the last rebuild)
  bool want_fused = true, fused = false;
  int *d_bpar = nullptr;
  size_t cap_bpar = 0;
  std::vector<hg_ctx *> rec_members;
};

static void batch_drop_graph(hg_batch *b) {
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// After steps of the x-fused form a member's state may sit in its second
// buffer set (d_bpar[r] = 1): swap the context's sets back so that its own
// record (hg_ctx::D) holds the state again.  `only`: just that member (it
// leaves the batch); nullptr: every member of the device records.
static int batch_normalize(hg_batch *b, hg_ctx *only) {
  if (!b->fused || b->rec_members.empty()) return HG_OK;
  HIPCHK(hipStreamSynchronize(b->stream));
  std::vector<int> par(b->rec_members.size());
  HIPCHK(hipMemcpy(par.data(), b->d_bpar, par.size() * sizeof(int), hipMemcpyDeviceToHost));
  for (size_t r = 0; r < par.size(); ++r) {
    hg_ctx *c = b->rec_members[r];
    if (!c || (only && c != only)) continue;
    member_to_set(c, par[r]);
    b->rec_members[r] = nullptr;  // its records are stale until the rebuild
  }
  return HG_OK;
}

// A member's own record (hg_ctx::D) onto the set that holds its state now,
// for hg_download / hg_upload between batch steps; the device records stay
// as they are (they number the sets by bsets).
static int batch_member_current(hg_ctx *c) {
  hg_batch *b = c->batch;
  if (!b || !b->fused) return HG_OK;
  for (size_t r = 0; r < b->rec_members.size(); ++r)
    if (b->rec_members[r] == c) {
      HIPCHK(hipSetDevice(b->device));
      HIPCHK(hipStreamSynchronize(b->stream));
      int par = 0;
      HIPCHK(hipMemcpy(&par, b->d_bpar + r, sizeof(int), hipMemcpyDeviceToHost));
      member_to_set(c, par);
      return HG_OK;
    }
  return HG_OK;  // added since the last build: its state is in its own set
}

extern "C" int hg_batch_create(int32_t device, hg_batch **out) {
  if (!out) return HG_ERR_ARG;
  *out = nullptr;
  hg_batch *b = new (std::nothrow) hg_batch();
  if (!b) return HG_ERR_NOMEM;
  b->device = device;
  if (hipSetDevice(device) != hipSuccess ||
      hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return HG_ERR_DEVICE;
  }
  *out = b;
  return HG_OK;
}

extern "C" int hg_batch_destroy(hg_batch *b) {
  if (!b) return HG_OK;
  hipSetDevice(b->device);
  hipStreamSynchronize(b->stream);
  batch_drop_graph(b);
  if (b->d_boxes) hipFree(b->d_boxes);
  if (b->d_sitems) hipFree(b->d_sitems);
  if (b->d_citems) hipFree(b->d_citems);
  if (b->h_sc) hipHostFree(b->h_sc);
  const int st = batch_normalize(b, nullptr);  // members left in the batch
  for (hg_ctx *c : b->members) {  // ordinary contexts again
    c->in_batch = false;
    c->batch = nullptr;
  }
  if (b->d_bpar) hipFree(b->d_bpar);
  hipStreamDestroy(b->stream);
  delete b;
  return st;
}

// Room for `need` member records in the pinned host array, keeping the
// records already there.
static int batch_hsc_reserve(hg_batch *b, size_t need) {
  if (need <= b->cap_hsc) return HG_OK;
  const size_t cap = need * 2;
  Scalars *p = nullptr;
  if (hipHostMalloc((void **)&p, cap * sizeof(Scalars), hipHostMallocDefault) != hipSuccess)
    return HG_ERR_NOMEM;
  if (b->h_sc) {
    std::memcpy(p, b->h_sc, b->cap_hsc * sizeof(Scalars));
    hipHostFree(b->h_sc);
  }
  b->h_sc = p;
  b->cap_hsc = cap;
  return HG_OK;
}

extern "C" int hg_batch_add(hg_batch *b, hg_ctx *c, double time, double t_out, double dt_out,
                            double t_end, int32_t outstyle, int32_t endstyle, int64_t tnlim,
                            int64_t tn_out) {
  if (!b || !c) return HG_ERR_ARG;
  if (c->device != b->device || c->D.grav || c->D.nspc || c->D.ax[0].ext || c->D.ax[1].ext ||
      c->D.ax[2].ext || c->fused)
    return HG_ERR_UNSUPPORTED;
  for (hg_ctx *m : b->members)
    if (m == c) return HG_ERR_ARG;
  if (!b->members.empty() &&
      (c->D.mag != b->mag || c->D.limiter != b->lim || c->D.mixed != b->mixed ||
       c->desc.rktype != b->rktype))
    return HG_ERR_UNSUPPORTED;  // one batch runs one kernel variant
  HIPCHK(hipSetDevice(b->device));
  int st = batch_hsc_reserve(b, b->members.size() + 1);
  if (st) return st;
  st = ensure_eint(c);
  if (st) return st;
  HIPCHK(hipStreamSynchronize(c->stream));
  // the box's clock and its first CFL partials, on the batch stream
  launch_set_clock(c->D, time, t_out, t_end, dt_out, outstyle, endstyle, 0, tnlim,
                   tn_out < 1 ? 1 : tn_out, (long long)1 << 62, b->stream);
  launch_cfl(c->D, b->stream);
  launch_cfl_final(c->D, b->stream);  // its first dt; later ones close each step
  HIPCHK(hipGetLastError());
  b->mag = c->D.mag;
  b->lim = c->D.limiter;
  b->mixed = c->D.mixed;
  b->rktype = c->desc.rktype;
  b->members.push_back(c);
  b->sc_valid.push_back(0);
  c->in_batch = true;
  c->batch = b;
  b->dirty = true;
  return HG_OK;
}

extern "C" int hg_batch_remove(hg_batch *b, hg_ctx *c) {
  if (!b || !c) return HG_ERR_ARG;
  for (size_t q = 0; q < b->members.size(); ++q)
    if (b->members[q] == c) {
      HIPCHK(hipStreamSynchronize(b->stream));
      const int st = batch_normalize(b, c);
      if (st) return st;
      b->members.erase(b->members.begin() + q);
      b->sc_valid.erase(b->sc_valid.begin() + q);
      c->in_batch = false;
      c->batch = nullptr;
      // keep the cached member states aligned with the member list (the
      // array holds a record per member since hg_batch_add)
      for (size_t r = q; r < b->members.size(); ++r) b->h_sc[r] = b->h_sc[r + 1];
      b->dirty = true;
      return HG_OK;
    }
  return HG_ERR_ARG;
}

template <class T>
static int grow(T **p, size_t *cap, size_t need) {
  if (need <= *cap) return HG_OK;
  if (*p) hipFree(*p);
  *p = nullptr;
  size_t n = need + need / 2 + 16;
  if (hipMalloc((void **)p, n * sizeof(T)) != hipSuccess) return HG_ERR_NOMEM;
  *cap = n;
  return HG_OK;
}

static void batch_enqueue_step(hg_batch *b) {
  const Batch &B = b->B;
  // per direction: chunks of one item per wavefront, or the persistent grid
  unsigned gs[3];
  for (int d = 0; d < 3; ++d) {
    const long sw = (B.dend[d] - B.dbeg[d] + b->sweep_block_waves - 1) / b->sweep_block_waves;
    gs[d] = b->sweep_grid ? (unsigned)std::min<long>(b->sweep_grid, sw) : (unsigned)sw;
  }
  const unsigned gr = b->rk_grid ? (unsigned)std::min<long>(b->rk_grid, B.ncitems)
                                 : (unsigned)B.ncitems;
  for (int stage = 1; stage <= b->rktype; ++stage) {
    double faco, fact, facn;
    runge_coeff(stage, b->rktype, &faco, &fact, &facn);
    const bool last = stage == b->rktype;
    if (B.fused) {  // y, z, then x with the RK update
      for (int d = 1; d < 3; ++d)
        if (gs[d] > 0)
          launch_bsweep_dir(B, d, b->mag, b->lim, b->mixed, stage, gs[d], b->sweep_block_waves,
                            b->stream);
      launch_bsweep_rk(B, b->mag, b->lim, b->mixed, stage, faco, fact, facn, last, gs[0],
                       b->sweep_block_waves, b->stream);
      continue;
    }
    for (int d = 0; d < 3; ++d)
      if (gs[d] > 0)
        launch_bsweep_dir(B, d, b->mag, b->lim, b->mixed, stage, gs[d], b->sweep_block_waves,
                          b->stream);
    launch_brk(B, b->mag, stage, faco, fact, facn, last, gr, b->stream);
  }
  launch_bstep_end_cfl(B, b->stream);
}

static int batch_rebuild(hg_batch *b) {
  batch_drop_graph(b);
  int st = batch_normalize(b, nullptr);
  if (st) return st;
  const size_t nb = b->members.size();
  // the x-fused form when every member sweeps x
  bool fused = b->want_fused && nb > 0 && b->rktype <= 3;
  for (hg_ctx *c : b->members) fused = fused && c->D.waves[0] > 0;
  for (size_t q = 0; fused && q < nb; ++q)
    if ((st = ensure_fused_buffers(b->members[q], 0))) return st;
  std::vector<Dev> boxes(fused ? 9 * nb : nb);
  std::vector<int2> si, ci;
  for (size_t q = 0; q < nb; ++q) {
    hg_ctx *c = b->members[q];
    const Dev &D = c->D;
    if (!fused) {
      boxes[q] = D;
      for (long k = 0; k < D.nblk_cfl; ++k) ci.push_back(make_int2((int)q, (int)k));
      continue;
    }
    // record 9q + 3p + (s - 1): starting set p, stage s (set p+1, p+2, p+1
    // written; u0 from set p after stage 1, which stores no U0); the last
    // stage's CFL partials are one per chunk of x wavefronts
    get_sets(c, c->bsets);
    for (int p = 0; p < 3; ++p)
      for (int st3 = 1; st3 <= 3; ++st3) {
        const StateSet &in = c->bsets[st3 == 1 ? p : (st3 == 2 ? (p + 1) % 3 : (p + 2) % 3)];
        const StateSet &out = c->bsets[st3 == 2 ? (p + 2) % 3 : (p + 1) % 3];
        Dev r = D;
        for (int v = 0; v < D.nv; ++v) {
          r.U[v] = in.U[v];
          r.Uo[v] = out.U[v];
          r.U0[v] = st3 == 1 ? nullptr : c->bsets[p].U[v];
        }
        for (int w = 0; w < D.nw; ++w) {
          r.W[w] = in.W[w];
          r.Wo[w] = out.W[w];
        }
        r.fdir = 0;
        r.nblk_cfl = (int)((D.waves[0] + b->sweep_block_waves - 1) / b->sweep_block_waves);
        boxes[9 * q + 3 * p + (st3 - 1)] = r;
      }
  }
  // direction by direction, each box's wavefronts of the direction padded
  // to a multiple of 16, so every box's range starts on a multiple of 16
  // and aligned chunks never mix boxes
  for (int d = 0; d < 3; ++d) {
    b->B.dbeg[d] = (long)si.size();
    for (size_t q = 0; q < nb; ++q) {
      const long nwv = (b->members[q]->D.waves[d] + 15) / 16 * 16;
      for (long w = 0; w < nwv; ++w) si.push_back(make_int2((int)q, (int)w));
    }
    b->B.dend[d] = (long)si.size();
  }
  if ((st = grow(&b->d_boxes, &b->cap_boxes, boxes.size() ? boxes.size() : 1))) return st;
  if ((st = grow(&b->d_sitems, &b->cap_s, si.size() ? si.size() : 1))) return st;
  if ((st = grow(&b->d_citems, &b->cap_c, ci.size() ? ci.size() : 1))) return st;
  if ((st = grow(&b->d_bpar, &b->cap_bpar, nb ? nb : 1))) return st;
  if (nb) {
    // every member's state is in its own set now (batch_normalize)
    const std::vector<int> zero(nb, 0);
    HIPCHK(hipMemcpyAsync(b->d_bpar, zero.data(), nb * sizeof(int), hipMemcpyHostToDevice,
                          b->stream));
    HIPCHK(hipMemcpyAsync(b->d_boxes, boxes.data(), boxes.size() * sizeof(Dev),
                          hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(b->d_sitems, si.data(), si.size() * sizeof(int2),
                          hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipMemcpyAsync(b->d_citems, ci.data(), ci.size() * sizeof(int2),
                          hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
  }
  b->B.boxes = b->d_boxes;
  b->B.nbox = (int)nb;
  b->B.sitems = b->d_sitems;
  b->B.nsitems = (long)si.size();
  b->B.lds_max = nb ? b->members[0]->D.lds_max : 0;
  b->B.citems = b->d_citems;
  b->B.ncitems = (long)ci.size();
  b->B.fused = fused ? 1 : 0;
  b->B.rk = b->rktype;
  b->B.bpar = b->d_bpar;
  b->fused = fused;
  b->rec_members = b->members;
  if ((st = batch_hsc_reserve(b, nb ? nb : 1))) return st;
  if (b->use_graph && nb) {
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    batch_enqueue_step(b);
    HIPCHK(hipStreamEndCapture(b->stream, &b->graph));
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  }
  b->dirty = false;
  return HG_OK;
}

// Advance every member by nsteps of its own steps (members that reached
// their end or an output stop simply idle).  *nrunning receives the number
// of members still running afterwards.
extern "C" int hg_batch_step(hg_batch *b, int32_t nsteps, int32_t *nrunning) {
  if (!b || nsteps < 0) return HG_ERR_ARG;
  HIPCHK(hipSetDevice(b->device));
  if (b->members.empty()) {
    if (nrunning) *nrunning = 0;
    return HG_OK;
  }
  int st;
  if (b->dirty && (st = batch_rebuild(b))) return st;
  for (int q = 0; q < nsteps; ++q) {
    if (b->exec) HIPCHK(hipGraphLaunch(b->exec, b->stream));
    else batch_enqueue_step(b);
  }
  HIPCHK(hipGetLastError());
  for (size_t q = 0; q < b->members.size(); ++q)
    HIPCHK(hipMemcpyAsync(&b->h_sc[q], b->members[q]->D.sc, sizeof(Scalars),
                          hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  int run = 0;
  for (size_t q = 0; q < b->members.size(); ++q) {
    hg_ctx *c = b->members[q];
    b->sc_valid[q] = 1;
    *c->h_sc = b->h_sc[q];
    c->ch = b->h_sc[q].ch;
    if (b->h_sc[q].stop == STOP_NONE) ++run;
  }
  if (nrunning) *nrunning = run;
  return HG_OK;
}

// Clock of one member after the last hg_batch_step (stop: 0 running,
// 1 end, 2 output due, 4 error).
extern "C" int hg_batch_member_state(hg_batch *b, hg_ctx *c, double *time, int64_t *tn,
                                     int32_t *stop) {
  if (!b || !c) return HG_ERR_ARG;
  for (size_t q = 0; q < b->members.size(); ++q)
    if (b->members[q] == c) {
      if (!b->sc_valid[q]) return HG_ERR_STATE;  // not stepped since it was added
      const Scalars &s = b->h_sc[q];
      if (time) *time = s.time;
      if (tn) *tn = s.tn;
      if (stop) *stop = s.stop;
      if (s.stop == STOP_ERROR) return check_error(c);
      return HG_OK;
    }
  return HG_ERR_ARG;
}

// Launch shapes (set by the autotuner in hormone_amd/batch.py or by hand):
// sweep wavefronts per block (1, 2 or 4), persistent grid sizes (0 = one
// work item per wavefront / block), and whether steps replay a hipGraph.
extern "C" int hg_batch_config(hg_batch *b, int32_t sweep_block_waves, int32_t sweep_grid,
                               int32_t rk_grid, int32_t use_graph) {
  if (!b) return HG_ERR_ARG;
  const int w = sweep_block_waves;
  if (w != 1 && w != 2 && w != 4 && w != 8 && w != 16) return HG_ERR_ARG;
  if (sweep_grid < 0 || rk_grid < 0) return HG_ERR_ARG;
  HIPCHK(hipStreamSynchronize(b->stream));
  b->sweep_block_waves = sweep_block_waves;
  b->sweep_grid = (unsigned)sweep_grid;
  b->rk_grid = (unsigned)rk_grid;
  b->use_graph = use_graph != 0;
  b->dirty = true;
  return HG_OK;
}

// The x-fused form for the batch (on = 1, the default: taken when every
// member sweeps x; every member then holds a second copy of its state) or
// the separate RK update (on = 0).  Results are the same bits either way.
extern "C" int hg_batch_set_fused(hg_batch *b, int32_t on) {
  if (!b || (on != 0 && on != 1)) return HG_ERR_ARG;
  HIPCHK(hipSetDevice(b->device));
  HIPCHK(hipStreamSynchronize(b->stream));
  b->want_fused = on != 0;
  b->dirty = true;
  return HG_OK;
}

extern "C" int hg_batch_form(const hg_batch *b, int32_t *fused) {
  if (!b || !fused) return HG_ERR_ARG;
  *fused = b->fused ? 1 : 0;
  return HG_OK;
}

// ---- allocation-free timestep (hg_timestep_noalloc) -------------------------
// The device memory this entry point works in besides the caller's arrays:
// one Scalars record (the clock the dt clipping reads, the CFL results and
// the error key), then the CFL partials, two doubles per 256-cell block.
// A file-scope device array, one copy per GPU, held until the process ends.
static constexpr size_t NOALLOC_REC = 256;  // the Scalars record, padded
static constexpr size_t NOALLOC_BYTES =
    NOALLOC_REC + 2 * sizeof(double) * ((size_t)HG_NOALLOC_MAX_CELLS / 256);
static_assert(sizeof(Scalars) <= NOALLOC_REC, "scratch record too small");
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

// The device address of `scratch` on each GPU, found once per device with
// that device current (the host-side name is not a device address).
static constexpr int NOALLOC_MAX_DEV = 64;
static void *g_scratch_ptr[NOALLOC_MAX_DEV];
static std::mutex g_scratch_mu;

static int noalloc_scratch(int dev, unsigned char **ptr) {
  if (dev < 0 || dev >= NOALLOC_MAX_DEV) return HG_ERR_DEVICE;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return HG_ERR_DEVICE;
    g_scratch_ptr[dev] = p;
  }
  *ptr = static_cast<unsigned char *>(g_scratch_ptr[dev]);
  return HG_OK;
}

extern "C" int hg_timestep_noalloc(const hg_desc *desc, const double *const *prim,
                                   const double *const *dxi, double time, double t_out,
                                   double t_end, int32_t outstyle, int32_t endstyle,
                                   double *out, void *stream) {
  if (!desc || !prim || !dxi || !out) return HG_ERR_ARG;
  if (desc->ie < desc->is || desc->je < desc->js || desc->ke < desc->ks) return HG_ERR_ARG;
  const long n1 = desc->ie - desc->is + 1, n2 = desc->je - desc->js + 1,
             n3 = desc->ke - desc->ks + 1;
  if ((double)n1 * (double)n2 * (double)n3 > (double)HG_NOALLOC_MAX_CELLS)
    return HG_ERR_SCRATCH;
  const bool mag = desc->mag_on != 0;
  for (int q = 0; q < (mag ? 8 : 5); ++q)
    if (!prim[q]) return HG_ERR_ARG;
  for (int a = 0; a < 3; ++a)
    if (!dxi[a]) return HG_ERR_ARG;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  unsigned char *base = nullptr;
  int st = noalloc_scratch(dev, &base);
  if (st) return st;
  // a Dev record of kernel arguments over the caller's arrays and the
  // scratch (built on the host stack; nothing is allocated)
  Dev D;
  std::memset(&D, 0, sizeof(D));
  D.n1 = (int)n1;
  D.n2 = (int)n2;
  D.n3 = (int)n3;
  D.ncell = n1 * n2 * n3;
  D.solve[0] = n1 > 1;
  D.solve[1] = n2 > 1;
  D.solve[2] = n3 > 1;
  D.mag = mag;
  D.gamma = desc->gamma;
  D.courant = desc->courant;
  D.W[WD] = const_cast<double *>(prim[0]);
  for (int a = 0; a < 3; ++a) D.W[WV1 + a] = const_cast<double *>(prim[1 + a]);
  D.W[WEINT] = const_cast<double *>(prim[4]);
  if (mag)
    for (int a = 0; a < 3; ++a) D.W[WB1 + a] = const_cast<double *>(prim[5 + a]);
  for (int a = 0; a < 3; ++a) D.ax[a].dxi = dxi[a];
  D.sc = reinterpret_cast<Scalars *>(base);
  D.part = reinterpret_cast<double *>(base + NOALLOC_REC);
  D.nblk_cfl = (int)((D.ncell + 255) / 256);
  hipStream_t s = static_cast<hipStream_t>(stream);
  // every call starts from a cleared error key and its own clock
  launch_clear_error(D, s);
  launch_set_clock(D, time, t_out, t_end, 0.0, outstyle, endstyle, 0, 0, 1, 1, s);
  launch_cfl(D, s);
  launch_cfl_final(D, s);
  launch_timestep_out(D, out, s);
  HIPCHK(hipGetLastError());
  return HG_OK;
}
