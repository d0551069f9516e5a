/*
 * hormone_oracle.c -- CPU restatement of HORMONE's Cartesian explicit hydro
 * step (test infrastructure only; see hormone_oracle.h).
 *
 * Every routine below is a statement-by-statement restatement of the cited
 * reference routine, keeping the Fortran evaluation order (left to right,
 * x**2 == x*x) and the reference's MIN/MAX/SIGN semantics as compiled by
 * flang/gfortran:  max(a,b) = (a>b ? a : b), min(a,b) = (a<b ? a : b),
 * n-ary forms fold left, sign(a,b) = copysign(|a|,b) (signed zero honoured).
 * Build with -ffp-contract=off (no FMA contraction) and no fast-math.
 */
#include "hormone_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NV 9 /* max conserved variables handled (MHD with GLM) */

static inline double fmax2(double a, double b) { return a > b ? a : b; }
static inline double fmin2(double a, double b) { return a < b ? a : b; }
static inline double fmin3(double a, double b, double c) {
  return fmin2(fmin2(a, b), c);
}
static inline double fmax3(double a, double b, double c) {
  return fmax2(fmax2(a, b), c);
}
static inline double fsign(double a, double b) { return copysign(fabs(a), b); }

struct or_ctx {
  hg_desc prm;
  int is, ie, js, je, ks, ke;
  int nxg, nyg, nzg;
  size_t ncell;
  int solve[3], dim;
  int ufnmax;
  int icnt, imo1, imo2, imo3, iene, img1, img2, img3, i9wv; /* 0-based */
  /* coordinates, base-shifted so c->x[d][i] is valid for i = s-2..e+2 */
  double *xs[3], *xis[3], *dxs[3], *dxis[3], *idxs[3], *idxis[3];
  double *mem_coord;
  /* primitive and auxiliary fields (padded) */
  double *d, *v1, *v2, *v3, *e, *p, *eint, *b1, *b2, *b3, *phi, *ptot, *cs;
  double *u[NV], *uorg[NV], *src[NV];
  double *flux[3][NV];
  /* slopes per direction: dd de dm1 dm2 dm3 db1 db2 db3 dphi */
  double *sl[3][9];
  /* gravity */
  double *grvphi, *grvpsi, *grvphiorg[2], *lapphi, *hgsrc, *gsrc;
  double *lapc[5]; /* lap_coeff(-1:3,...) for crdnt 0 (no reference counterpart) */
  /* advected species (compswitch 2): spc, spcorg, dspc, spcflx
     (allocations.f90:294-300) */
  int spn;
  double *spc[HG_SPN_MAX], *spcorg[HG_SPN_MAX];
  double *dspc[3][HG_SPN_MAX], *spcflx[3][HG_SPN_MAX];
  double *dti;
  double ch;
  int err_i, err_j, err_k, err_dir;
  /* flux of each flux-boundary face (code 10), faces x1i x1o x2i x2o x3i
     x3o, conserved-variable order; zero until set */
  double fxb[6][NV];
};

static int g_threads = 0;
void or_set_threads(int n) { g_threads = n; }
int or_get_threads(void) {
#ifdef _OPENMP
  return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
  return 1;
#endif
}
#define NT (or_get_threads())

#define ID(c, i, j, k)                                                      \
  ((size_t)((i) - (c)->is + 2) +                                            \
   (size_t)(c)->nxg * ((size_t)((j) - (c)->js + 2) +                        \
                       (size_t)(c)->nyg * (size_t)((k) - (c)->ks + 2)))

/* ------------------------------------------------------------------------ */
/* create / destroy                                                          */
/* ------------------------------------------------------------------------ */

static double *alloc_field(size_t n) {
  double *a = (double *)calloc(n, sizeof(double));
  return a;
}

or_ctx *or_create(const hg_desc *desc, int *status) {
  int st = HG_OK;
  or_ctx *c = NULL;
  if (!desc) { st = HG_ERR_ARG; goto fail; }
  if (desc->crdnt != 0 || desc->eostype != 0 ||
      (desc->compswitch != 0 && desc->compswitch != 2)) {
    st = HG_ERR_UNSUPPORTED; goto fail;
  }
  if (desc->compswitch == 2 && (desc->spn < 1 || desc->spn > HG_SPN_MAX)) {
    st = HG_ERR_ARG; goto fail;
  }
  for (int b = 0; b < 12; ++b) {
    /* the cases of boundarycondition.f90's select blocks: 0-3, 9, 10 on
       every face; 4/5 at x1 outer (:302,:419,:441), x3 inner
       (:963,:1074,:1096), x3 outer (scalar 4 only :1222; vectors
       :1335,:1357) */
    int code = desc->bc[b], face = b % 6, ok = 0;
    if ((code >= 0 && code <= 3) || code == 9 || code == 10) ok = 1;
    else if (code == 4 || code == 5)
      ok = face == 1 || face == 4 || (face == 5 && (b >= 6 || code == 4));
    if (!ok || (desc->compswitch == 2 && code >= 9)) { st = HG_ERR_UNSUPPORTED; goto fail; }
  }
  if (desc->ie < desc->is || desc->je < desc->js || desc->ke < desc->ks) {
    st = HG_ERR_ARG; goto fail;
  }
  if (desc->rktype < 1 || desc->rktype > 3) { st = HG_ERR_ARG; goto fail; }
  c = (or_ctx *)calloc(1, sizeof(or_ctx));
  c->prm = *desc;
  c->is = desc->is; c->ie = desc->ie;
  c->js = desc->js; c->je = desc->je;
  c->ks = desc->ks; c->ke = desc->ke;
  c->nxg = c->ie - c->is + 5;
  c->nyg = c->je - c->js + 5;
  c->nzg = c->ke - c->ks + 5;
  c->ncell = (size_t)c->nxg * c->nyg * c->nzg;
  /* checksetup.f90:41-51 */
  c->solve[0] = c->ie > c->is;
  c->solve[1] = c->je > c->js;
  c->solve[2] = c->ke > c->ks;
  c->dim = c->solve[0] + c->solve[1] + c->solve[2];
  if (c->dim == 0) { st = HG_ERR_ARG; goto fail; }
  /* equation indices, checksetup.f90:62-75 (0-based here) */
  c->icnt = 0; c->imo1 = 1; c->imo2 = 2; c->imo3 = 3; c->iene = 4;
  c->ufnmax = 5;
  c->img1 = c->img2 = c->img3 = c->i9wv = -1;
  if (desc->mag_on) {
    c->img1 = 5; c->img2 = 6; c->img3 = 7; c->ufnmax = 8;
    if (c->dim > 1) { c->i9wv = 8; c->ufnmax = 9; }
  }
  {
    int n[3] = {c->nxg, c->nyg, c->nzg};
    int s[3] = {c->is, c->js, c->ks};
    const double *src[3][6] = {
        {desc->x1, desc->xi1, desc->dx1, desc->dxi1, desc->idx1, desc->idxi1},
        {desc->x2, desc->xi2, desc->dx2, desc->dxi2, desc->idx2, desc->idxi2},
        {desc->x3, desc->xi3, desc->dx3, desc->dxi3, desc->idx3, desc->idxi3}};
    size_t tot = 6 * (size_t)(n[0] + n[1] + n[2]);
    c->mem_coord = (double *)calloc(tot, sizeof(double));
    double *q = c->mem_coord;
    for (int d = 0; d < 3; ++d) {
      double **dst[6] = {&c->xs[d], &c->xis[d], &c->dxs[d],
                         &c->dxis[d], &c->idxs[d], &c->idxis[d]};
      for (int a = 0; a < 6; ++a) {
        if (!src[d][a]) { st = HG_ERR_ARG; goto fail; }
        memcpy(q, src[d][a], sizeof(double) * n[d]);
        *dst[a] = q - (s[d] - 2);
        q += n[d];
      }
    }
  }
  {
    size_t N = c->ncell;
    double **f[] = {&c->d, &c->v1, &c->v2, &c->v3, &c->e, &c->p, &c->eint,
                    &c->b1, &c->b2, &c->b3, &c->phi, &c->ptot, &c->cs,
                    &c->dti};
    for (size_t a = 0; a < sizeof(f) / sizeof(f[0]); ++a) *f[a] = alloc_field(N);
    /* allocations.f90:71-78: d = p = e = eint = ptot = 1 in every cell
       (ghost cells no boundary code writes keep them) */
    for (size_t m = 0; m < N; ++m)
      c->d[m] = c->p[m] = c->e[m] = c->eint[m] = c->ptot[m] = 1.0;
    for (int n = 0; n < c->ufnmax; ++n) {
      c->u[n] = alloc_field(N);
      c->uorg[n] = alloc_field(N);
      c->src[n] = alloc_field(N);
      for (int d = 0; d < 3; ++d) c->flux[d][n] = alloc_field(N);
    }
    for (int d = 0; d < 3; ++d)
      for (int a = 0; a < 9; ++a) c->sl[d][a] = alloc_field(N);
    c->spn = desc->compswitch == 2 ? desc->spn : 0;
    for (int n = 0; n < c->spn; ++n) {
      /* zero-filled: the flux arrays of an unsolved direction stay 0 and
         enter the species update as idxi*(0-0) */
      c->spc[n] = alloc_field(N);
      c->spcorg[n] = alloc_field(N);
      for (int d = 0; d < 3; ++d) {
        c->dspc[d][n] = alloc_field(N);
        c->spcflx[d][n] = alloc_field(N);
      }
    }
    if (desc->gravswitch == 3) {
      c->grvphi = alloc_field(N); c->grvpsi = alloc_field(N);
      c->grvphiorg[0] = alloc_field(N); c->grvphiorg[1] = alloc_field(N);
      c->lapphi = alloc_field(N); c->hgsrc = alloc_field(N);
      c->gsrc = alloc_field(N);
      for (int q = 0; q < 5; ++q) c->lapc[q] = alloc_field(N);
    }
  }
  if (desc->gravswitch == 3) {
    /* Cartesian lap_coeff: the 3D form of gravity_hyperbolic.f90:382-406
       with unit metric factors (the reference stops for crdnt 0, :413-417;
       this is new).  Normalisation lap_coeff(-1) = cell volume. */
    const double *X1 = c->dxis[0], *X2 = c->dxis[1], *X3 = c->dxis[2];
    const double *I1 = c->idxs[0], *I2 = c->idxs[1], *I3 = c->idxs[2];
    for (int k = c->ks - 1; k <= c->ke; ++k)
      for (int j = c->js - 1; j <= c->je; ++j)
        for (int i = c->is - 1; i <= c->ie; ++i) {
          size_t m = ID(c, i, j, k);
          c->lapc[1][m] = -((I1[i] + I1[i + 1]) * X2[j] * X3[k] +
                            (I2[j] + I2[j + 1]) * X1[i] * X3[k] +
                            (I3[k] + I3[k + 1]) * X1[i] * X2[j]);
          c->lapc[2][m] = X2[j] * X3[k] * I1[i + 1];
          c->lapc[3][m] = X1[i] * X3[k] * I2[j + 1];
          c->lapc[4][m] = X1[i] * X2[j] * I3[k + 1];
          c->lapc[0][m] = X1[i] * X2[j] * X3[k];
        }
  }
  if (status) *status = HG_OK;
  return c;
fail:
  if (c) or_destroy(c);
  if (status) *status = st;
  return NULL;
}

void or_destroy(or_ctx *c) {
  if (!c) return;
  double *f[] = {c->d, c->v1, c->v2, c->v3, c->e, c->p, c->eint, c->b1,
                 c->b2, c->b3, c->phi, c->ptot, c->cs, c->dti, c->grvphi,
                 c->grvpsi, c->grvphiorg[0], c->grvphiorg[1], c->lapphi,
                 c->hgsrc, c->gsrc, c->lapc[0], c->lapc[1], c->lapc[2], c->lapc[3],
                 c->lapc[4]};
  for (size_t a = 0; a < sizeof(f) / sizeof(f[0]); ++a) free(f[a]);
  for (int n = 0; n < NV; ++n) {
    free(c->u[n]); free(c->uorg[n]); free(c->src[n]);
    for (int d = 0; d < 3; ++d) free(c->flux[d][n]);
  }
  for (int d = 0; d < 3; ++d)
    for (int a = 0; a < 9; ++a) free(c->sl[d][a]);
  for (int n = 0; n < HG_SPN_MAX; ++n) {
    free(c->spc[n]); free(c->spcorg[n]);
    for (int d = 0; d < 3; ++d) { free(c->dspc[d][n]); free(c->spcflx[d][n]); }
  }
  free(c->mem_coord);
  free(c);
}

int or_ufnmax(const or_ctx *c) { return c->ufnmax; }

static double *field_ptr(or_ctx *c, int field) {
  switch (field) {
  case HG_F_D: return c->d;
  case HG_F_V1: return c->v1;
  case HG_F_V2: return c->v2;
  case HG_F_V3: return c->v3;
  case HG_F_E: return c->e;
  case HG_F_B1: return c->b1;
  case HG_F_B2: return c->b2;
  case HG_F_B3: return c->b3;
  case HG_F_PHI: return c->phi;
  case HG_F_EINT: return c->eint;
  case HG_F_P: return c->p;
  case HG_F_GRVPHI: return c->grvphi;
  case HG_F_GRVPSI: return c->grvpsi;
  default:
    if (field >= HG_F_U && field < HG_F_U + c->ufnmax) return c->u[field - HG_F_U];
    if (field >= HG_F_SPC && field < HG_F_SPC + c->spn) return c->spc[field - HG_F_SPC];
    return NULL;
  }
}

int or_upload(or_ctx *c, int field, const double *a) {
  double *f = field_ptr(c, field);
  if (!f || !a) return HG_ERR_ARG;
  if (field == HG_F_GRVPHI) {
    /* the potential's ghost layer carries the gravity boundary values */
    memcpy(f, a, sizeof(double) * c->ncell);
    return HG_OK;
  }
  for (int k = c->ks; k <= c->ke; ++k)
    for (int j = c->js; j <= c->je; ++j)
      for (int i = c->is; i <= c->ie; ++i) f[ID(c, i, j, k)] = a[ID(c, i, j, k)];
  return HG_OK;
}

/* ghost layers of the faces with codes 4, 5, 9, 10 (hg_set_ghosts) */
int or_set_ghosts(or_ctx *c, int field, const double *a) {
  double *f = NULL;
  switch (field) {
    case HG_F_D: f = c->d; break;
    case HG_F_P: f = c->p; break;
    case HG_F_V1: f = c->v1; break;
    case HG_F_V2: f = c->v2; break;
    case HG_F_V3: f = c->v3; break;
    case HG_F_B1: f = c->b1; break;
    case HG_F_B2: f = c->b2; break;
    case HG_F_B3: f = c->b3; break;
    case HG_F_PHI: f = c->phi; break;
    default: return HG_ERR_ARG;
  }
  if (!a) return HG_ERR_ARG;
  int s_[3] = {c->is, c->js, c->ks}, e_[3] = {c->ie, c->je, c->ke};
  for (int axis = 0; axis < 3; ++axis)
    for (int side = 0; side < 2; ++side) {
      if (c->prm.bc[2 * axis + side] < 4 && c->prm.bc[6 + 2 * axis + side] < 4) continue;
      int lo[3] = {c->is, c->js, c->ks}, hi[3] = {c->ie, c->je, c->ke};
      lo[axis] = side ? e_[axis] + 1 : s_[axis] - 2;
      hi[axis] = side ? e_[axis] + 2 : s_[axis] - 1;
      for (int k = lo[2]; k <= hi[2]; ++k)
        for (int j = lo[1]; j <= hi[1]; ++j)
          for (int i = lo[0]; i <= hi[0]; ++i) f[ID(c, i, j, k)] = a[ID(c, i, j, k)];
    }
  return HG_OK;
}

int or_set_boundary_flux(or_ctx *c, int face, const double *flux) {
  if (!flux || face < 0 || face > 5 || c->prm.bc[face] != 10) return HG_ERR_ARG;
  for (int v = 0; v < NV; ++v) c->fxb[face][v] = v < c->ufnmax ? flux[v] : 0.0;
  return HG_OK;
}

int or_download(or_ctx *c, int field, double *a) {
  double *f = field_ptr(c, field);
  if (!f || !a) return HG_ERR_ARG;
  for (int k = c->ks; k <= c->ke; ++k)
    for (int j = c->js; j <= c->je; ++j)
      for (int i = c->is; i <= c->ie; ++i) a[ID(c, i, j, k)] = f[ID(c, i, j, k)];
  return HG_OK;
}

int or_last_error(const or_ctx *c, int *i, int *j, int *k, int *dir) {
  if (i) *i = c->err_i;
  if (j) *j = c->err_j;
  if (k) *k = c->err_k;
  if (dir) *dir = c->err_dir;
  return HG_OK;
}

/* ------------------------------------------------------------------------ */
/* EOS helpers, eostype 0 (eos.f90)                                          */
/* ------------------------------------------------------------------------ */

/* get_eint, eos.f90:846-865 */
static inline double get_eint(int mag_on, double etot, double d, double v1,
                              double v2, double v3, double b1, double b2,
                              double b3) {
  double vsq = v1 * v1 + v2 * v2 + v3 * v3;
  double eint = etot - 0.5 * d * vsq;
  if (mag_on) {
    double bsq = b1 * b1 + b2 * b2 + b3 * b3;
    eint = eint - 0.5 * bsq;
  }
  return eint;
}

/* eos_p_cs, eos.f90:195-242 (eostype 0); returns ierr */
static inline int eos_p_cs(double gamma, double d, double eint, int has_erad,
                           double erad, double *p, double *cs) {
  if (eint <= 0.0 || d <= 0.0) return 1;
  double pp = (gamma - 1.0) * eint; /* eos_p, eos.f90:258-260 */
  double cs2 = gamma * pp / d;
  if (has_erad) cs2 = cs2 + 4.0 * erad / (9.0 * d);
  *p = pp;
  *cs = sqrt(cs2);
  return 0;
}

/* get_cf, eos.f90:351-365 */
static inline double get_cf(double d, double cs, double b1, double b2,
                            double b3) {
  double asq = cs * cs;
  double bb1 = b1 * b1 / d;
  double bb2 = b2 * b2 / d;
  double bb3 = b3 * b3 / d;
  double bbsq = bb1 + bb2 + bb3;
  return sqrt(0.5 * (asq + bbsq +
                     sqrt((asq + bbsq) * (asq + bbsq) - 4.0 * asq * bb1)));
}

/* ------------------------------------------------------------------------ */
/* limiters, fluxlimiter.f90:40-75                                           */
/* ------------------------------------------------------------------------ */

static inline double minmod(const double u[3], const double dx[2]) {
  double x = 1.4 * (u[2] - u[1]) * dx[1];
  double y = 1.4 * (u[1] - u[0]) * dx[0];
  double z = (u[2] - u[0]) * dx[0] * dx[1] / (dx[0] + dx[1]);
  double s = fsign(1.0, x);
  return s * fmax2(0.0, fmin3(fabs(x), s * y, s * z));
}

static inline double modified_mc(const double u[3], const double x[3],
                                 const double xi[2]) {
  double cf = (x[2] - x[1]) / (xi[1] - x[1]);
  double cb = (x[1] - x[0]) / (x[1] - xi[0]);
  double top = (u[1] - u[0]) * (x[2] - x[1]) / (x[1] - x[0]);
  double bot = u[2] - u[1];
  double sig = fsign(1.0, bot);
  return sig * fmax2(0.0, fmin3(sig * 0.5 * (bot + top), sig * cf * bot,
                                sig * cb * top)) /
         (x[2] - x[1]);
}

static inline double limiter(int which, const double u[3], const double dx[2],
                             const double x[3], const double xi[2]) {
  return which == HG_LIMITER_MINMOD ? minmod(u, dx) : modified_mc(u, x, xi);
}

/* ------------------------------------------------------------------------ */
/* HLLD + GLM flux, hlldflux.f90:19-179                                      */
/* ------------------------------------------------------------------------ */

static inline void fstar(double f[10], double d, double v1, double v2,
                         double v3, double b1, double b2, double b3, double e,
                         double p, double phi, double ch, double er) {
  f[0] = d * v1;
  f[1] = d * v1 * v1 + p - b1 * b1;
  f[2] = d * v1 * v2 - b1 * b2;
  f[3] = d * v1 * v3 - b1 * b3;
  f[4] = (e + p) * v1 - b1 * (v1 * b1 + v2 * b2 + v3 * b3);
  f[5] = phi;
  f[6] = b2 * v1 - b1 * v2;
  f[7] = b3 * v1 - b1 * v3;
  f[8] = ch * ch * b1;
  f[9] = er * v1;
}

static void hlldflux(double f[10], double cfl, double cfr, double v1l,
                     double v1r, double v2l, double v2r, double v3l,
                     double v3r, double dl, double dr, double el, double er,
                     double ptl, double ptr, double b1l, double b1r,
                     double b2l, double b2r, double b3l, double b3r,
                     double phil, double phir, double ch, double erl,
                     double err) {
  double b1 = 0.5 * ((b1l + b1r) - (phir - phil) / ch);
  double phi = 0.5 * ((phil + phir) - (b1r - b1l) * ch);

  double sl = fmin3(v1r - cfr, v1l - cfl, 0.0);
  double sr = fmax3(v1r + cfr, v1l + cfl, 0.0);

  double sm = ((sr - v1r) * dr * v1r - (sl - v1l) * dl * v1l - ptr + ptl) /
              ((sr - v1r) * dr - (sl - v1l) * dl);

  double dla = dl * (sl - v1l) / (sl - sm);
  double dra = dr * (sr - v1r) / (sr - sm);

  double erla = erl * (sl - v1l) / (sl - sm);
  double erra = err * (sr - v1r) / (sr - sm);

  double sla = sm - fabs(b1) / sqrt(dla);
  double sra = sm + fabs(b1) / sqrt(dra);

  double pta = ((sr - v1r) * dr * ptl - (sl - v1l) * dl * ptr +
                dl * dr * (sr - v1r) * (sl - v1l) * (v1r - v1l)) /
               ((sr - v1r) * dr - (sl - v1l) * dl);

  double v2la = v2l - b1 * b2l * (sm - v1l) / (dl * (sl - v1l) * (sl - sm) - b1 * b1);
  double v2ra = v2r - b1 * b2r * (sm - v1r) / (dr * (sr - v1r) * (sr - sm) - b1 * b1);
  double v3la = v3l - b1 * b3l * (sm - v1l) / (dl * (sl - v1l) * (sl - sm) - b1 * b1);
  double v3ra = v3r - b1 * b3r * (sm - v1r) / (dr * (sr - v1r) * (sr - sm) - b1 * b1);

  double b2la = b2l * (dl * (sl - v1l) * (sl - v1l) - b1 * b1) /
                (dl * (sl - v1l) * (sl - sm) - b1 * b1);
  double b2ra = b2r * (dr * (sr - v1r) * (sr - v1r) - b1 * b1) /
                (dr * (sr - v1r) * (sr - sm) - b1 * b1);
  double b3la = b3l * (dl * (sl - v1l) * (sl - v1l) - b1 * b1) /
                (dl * (sl - v1l) * (sl - sm) - b1 * b1);
  double b3ra = b3r * (dr * (sr - v1r) * (sr - v1r) - b1 * b1) /
                (dr * (sr - v1r) * (sr - sm) - b1 * b1);

  double ela = ((sl - v1l) * el - ptl * v1l + pta * sm +
                b1 * ((v1l * b1 + v2l * b2l + v3l * b3l) -
                      (sm * b1 + v2la * b2la + v3la * b3la))) /
               (sl - sm);
  double era = ((sr - v1r) * er - ptr * v1r + pta * sm +
                b1 * ((v1r * b1 + v2r * b2r + v3r * b3r) -
                      (sm * b1 + v2ra * b2ra + v3ra * b3ra))) /
               (sr - sm);

  double sb1 = fsign(1.0, b1);
  double v2aa = (sqrt(dla) * v2la + sqrt(dra) * v2ra + (b2ra - b2la) * sb1) /
                (sqrt(dla) + sqrt(dra));
  double v3aa = (sqrt(dla) * v3la + sqrt(dra) * v3ra + (b3ra - b3la) * sb1) /
                (sqrt(dla) + sqrt(dra));
  double b2aa = (sqrt(dla) * b2ra + sqrt(dra) * b2la +
                 sqrt(dla * dra) * (v2ra - v2la) * sb1) /
                (sqrt(dla) + sqrt(dra));
  double b3aa = (sqrt(dla) * b3ra + sqrt(dra) * b3la +
                 sqrt(dla * dra) * (v3ra - v3la) * sb1) /
                (sqrt(dla) + sqrt(dra));

  double eraa = era + sb1 * sqrt(dra) *
                          ((v2ra * b2ra + v3ra * b3ra) - (v2aa * b2aa + v3aa * b3aa));
  double elaa = ela - sb1 * sqrt(dla) *
                          ((v2la * b2la + v3la * b3la) - (v2aa * b2aa + v3aa * b3aa));

  double signsm = fsign(0.5, sm);
  double sortsla = 0.5 + fsign(0.5, sla);
  double sortsra = 0.5 + fsign(0.5, -sra);
  double sortslr = 0.5 - fsign(0.5, sla * sra);

  dla = (0.5 + signsm) * dla + (0.5 - signsm) * dra;
  erla = (0.5 + signsm) * erla + (0.5 - signsm) * erra;
  v2la = sortsla * v2la + sortslr * v2aa + sortsra * v2ra;
  v3la = sortsla * v3la + sortslr * v3aa + sortsra * v3ra;
  b2la = sortsla * b2la + sortslr * b2aa + sortsra * b2ra;
  b3la = sortsla * b3la + sortslr * b3aa + sortsra * b3ra;
  ela = sortsla * ela + (0.5 - fsign(0.5, sla * sm)) * elaa +
        (0.5 + fsign(0.5, -sm * sra)) * eraa + (0.5 - fsign(0.5, sra)) * era;

  fstar(f, dla, sm, v2la, v3la, b1, b2la, b3la, ela, pta, phi, ch, erla);
}

/* ------------------------------------------------------------------------ */
/* boundary conditions, boundarycondition.f90 (codes 0..3)                   */
/* ------------------------------------------------------------------------ */

/* ghost ptot / eint / e rebuild (boundarycondition.f90:221-243 etc.) */
static inline void ghost_energy(or_ctx *c, size_t g) {
  double gam = c->prm.gamma;
  c->ptot[g] = c->p[g] + 0.5 * (c->b1[g] * c->b1[g] + c->b2[g] * c->b2[g] +
                                c->b3[g] * c->b3[g]);
  c->eint[g] = c->p[g] / (gam - 1.0); /* eos_e, eos.f90:288-290 */
  c->e[g] = c->eint[g] +
            0.5 * (c->d[g] * (c->v1[g] * c->v1[g] + c->v2[g] * c->v2[g] +
                              c->v3[g] * c->v3[g]) +
                   c->b1[g] * c->b1[g] + c->b2[g] * c->b2[g] +
                   c->b3[g] * c->b3[g]);
}

/* Linear extrapolation of q into the two ghosts of one face line:
   boundarycondition.f90:302-326 (outer, edge e: q(e+1) = q(e) +
   (q(e)-q(e-1))*dx(e)/dx(e-1), q(e+2) = q(e+1) + (q(e+1)-q(e))*dx(e+1)/dx(e))
   and :963-985 (inner, edge s, mirrored).  E/N/G1/G2: edge, its inner
   neighbour, the ghosts next to and beyond the face; dx indexed by the
   axis index, eg = edge index. */
static inline void lin2(double *q, size_t E, size_t N, size_t G1, size_t G2, const double *dx,
                        int eg, int inner) {
  if (inner) {
    q[G1] = q[E] - (q[N] - q[E]) * dx[eg] / dx[eg + 1];
    q[G2] = q[G1] - (q[E] - q[G1]) * dx[eg - 1] / dx[eg];
  } else {
    q[G1] = q[E] + (q[E] - q[N]) * dx[eg] / dx[eg - 1];
    q[G2] = q[G1] + (q[G1] - q[E]) * dx[eg + 1] / dx[eg];
  }
}

/* codes 4/5 for one line of a face (scalars :302-326, :963-985,
   :1222-1244; vectors :419-468, :1074-1120, :1335-1382) */
static void linear_line(or_ctx *c, int axis, int inner, int sc, int vc, size_t E, size_t N,
                        size_t G1, size_t G2) {
  const double *dx = c->dxs[axis];
  int eg = inner ? (axis == 0 ? c->is : axis == 1 ? c->js : c->ks)
                 : (axis == 0 ? c->ie : axis == 1 ? c->je : c->ke);
  const int mag = c->prm.mag_on;
  if (sc == 4 || sc == 5) {
    lin2(c->d, E, N, G1, G2, dx, eg, inner);
    lin2(c->p, E, N, G1, G2, dx, eg, inner);
    if (c->d[G2] <= 0.0) c->d[G1] = c->d[G2] = c->d[E];
    if (c->p[G2] <= 0.0) c->p[G1] = c->p[G2] = c->p[E];
    if (mag) c->phi[G1] = c->phi[G2] = c->phi[E];
    for (int n = 0; n < c->spn; ++n) c->spc[n][G1] = c->spc[n][G2] = c->spc[n][E];
  }
  if (vc == 4 || vc == 5) {
    double *v[3] = {c->v1, c->v2, c->v3}, *b[3] = {c->b1, c->b2, c->b3};
    double vn = v[axis][E];
    double plug = inner ? 0.5 - fsign(0.5, vn) : 0.5 + fsign(0.5, vn);
    if (axis == 2 && inner) {
      /* :1078-1083 and :1085-1090 statement by statement: the second
         statement for v2 (b2) writes v3(ks-2) (b3(ks-2)) instead, from
         v2 (b2) differences; v2(ks-2) and b2(ks-2) keep their values */
      for (int w = 0; w < (mag ? 2 : 1); ++w) {
        double **q = w ? b : v;
        q[0][G1] = q[0][E] - (q[0][N] - q[0][E]) * dx[eg] / dx[eg + 1];
        q[0][G2] = q[0][G1] - (q[0][E] - q[0][G1]) * dx[eg - 1] / dx[eg];
        q[1][G1] = q[1][E] - (q[1][N] - q[1][E]) * dx[eg] / dx[eg + 1];
        q[2][G2] = q[2][G1] - (q[1][E] - q[1][G1]) * dx[eg - 1] / dx[eg];
        q[2][G1] = q[2][E] - (q[2][N] - q[2][E]) * dx[eg] / dx[eg + 1];
        q[2][G2] = q[2][G1] - (q[2][E] - q[2][G1]) * dx[eg - 1] / dx[eg];
      }
    } else {
      for (int a = 0; a < 3; ++a) {
        lin2(v[a], E, N, G1, G2, dx, eg, inner);
        if (mag) lin2(b[a], E, N, G1, G2, dx, eg, inner);
      }
    }
    if (vc == 5) {
      v[axis][G1] = v[axis][G1] * plug;
      v[axis][G2] = v[axis][G2] * plug;
    }
  }
}

/* The reference applies the scalar and the vector codes separately; for
   the supported codes (0..3) both copy from the same source cell except
   that reflective mirrors and periodic wraps.  We require the scalar and
   vector code of one face to be "compatible" (same source cell): periodic
   with periodic (enforced by checksetup.f90:82-107), 1 with 1, and 2/3
   with 2/3.  Mixed reflective/non-reflective pairs use separate copies. */
static void apply_face(or_ctx *c, int axis, int inner) {
  int sc = c->prm.bc[2 * axis + (inner ? 0 : 1)];
  int vc = c->prm.bc[6 + 2 * axis + (inner ? 0 : 1)];
  int s_[3] = {c->is, c->js, c->ks}, e_[3] = {c->ie, c->je, c->ke};
  int s = s_[axis], e = e_[axis];
  /* ranges of the two other axes: interior only */
  int lo[3] = {c->is, c->js, c->ks}, hi[3] = {c->ie, c->je, c->ke};
  int oa = (axis + 1) % 3, ob = (axis + 2) % 3;
#pragma omp parallel for collapse(2) num_threads(NT)
  for (int q = lo[ob]; q <= hi[ob]; ++q)
    for (int r = lo[oa]; r <= hi[oa]; ++r) {
      for (int gl = 0; gl < 2; ++gl) {
        /* ghost index along axis and its scalar / vector source */
        int g, ss, sv;
        if (inner) {
          g = s - 2 + gl; /* s-2, s-1 */
          ss = (sc == 0) ? e - 1 + gl : (sc == 1) ? s + 1 - gl : s;
          sv = (vc == 0) ? e - 1 + gl : (vc == 1) ? s + 1 - gl : s;
        } else {
          g = e + 1 + gl; /* e+1, e+2 */
          ss = (sc == 0) ? s + gl : (sc == 1) ? e - gl : e;
          sv = (vc == 0) ? s + gl : (vc == 1) ? e - gl : e;
        }
        int ijk_g[3], ijk_s[3], ijk_v[3];
        ijk_g[axis] = g; ijk_s[axis] = ss; ijk_v[axis] = sv;
        ijk_g[oa] = ijk_s[oa] = ijk_v[oa] = r;
        ijk_g[ob] = ijk_s[ob] = ijk_v[ob] = q;
        size_t G = ID(c, ijk_g[0], ijk_g[1], ijk_g[2]);
        size_t S = ID(c, ijk_s[0], ijk_s[1], ijk_s[2]);
        size_t V = ID(c, ijk_v[0], ijk_v[1], ijk_v[2]);
        /* scalars (codes 0-3; 4/5 below; 9/10 keep the ghost values
           or_set_ghosts put there) */
        if (sc <= 3) {
          c->d[G] = c->d[S];
          c->p[G] = c->p[S];
          if (c->prm.mag_on) c->phi[G] = c->phi[S];
          for (int n = 0; n < c->spn; ++n) c->spc[n][G] = c->spc[n][S];
        }
        if (vc > 3) continue;
        /* vectors */
        double *v[3] = {c->v1, c->v2, c->v3};
        double *b[3] = {c->b1, c->b2, c->b3};
        for (int a = 0; a < 3; ++a) {
          double vv = v[a][V];
          if (a == axis) {
            if (vc == 1) vv = -v[a][V];
            else if (vc == 2) {
              if (axis == 0 && inner) vv = fmin2(0.0, v[a][V]);
              else {
                double plug = inner ? 0.5 - fsign(0.5, v[a][V])
                                    : 0.5 + fsign(0.5, v[a][V]);
                vv = v[a][V] * plug;
              }
            }
          }
          v[a][G] = vv;
          if (c->prm.mag_on) b[a][G] = (a == axis && vc == 1) ? -b[a][V] : b[a][V];
        }
      }
      if ((sc == 4 || sc == 5) || (vc == 4 || vc == 5)) {
        int ie_[4] = {inner ? s : e, inner ? s + 1 : e - 1, inner ? s - 1 : e + 1,
                      inner ? s - 2 : e + 2};
        size_t m4[4];
        for (int t = 0; t < 4; ++t) {
          int ijk[3];
          ijk[axis] = ie_[t]; ijk[oa] = r; ijk[ob] = q;
          m4[t] = ID(c, ijk[0], ijk[1], ijk[2]);
        }
        linear_line(c, axis, inner, sc, vc, m4[0], m4[1], m4[2], m4[3]);
      }
      for (int gl = 0; gl < 2; ++gl) {
        int g = inner ? s - 2 + gl : e + 1 + gl;
        int ijk[3];
        ijk[axis] = g; ijk[oa] = r; ijk[ob] = q;
        ghost_energy(c, ID(c, ijk[0], ijk[1], ijk[2]));
      }
    }
}

/* boundarycondition, boundarycondition.f90:14-1440 (x always, y if solve_j,
   z if solve_k) */
static void boundarycondition(or_ctx *c) {
  apply_face(c, 0, 1);
  apply_face(c, 0, 0);
  if (c->solve[1]) { apply_face(c, 1, 1); apply_face(c, 1, 0); }
  if (c->solve[2]) { apply_face(c, 2, 1); apply_face(c, 2, 0); }
}

/* ------------------------------------------------------------------------ */
/* interpolation.f90:15-490                                                  */
/* ------------------------------------------------------------------------ */

static void interpolation(or_ctx *c) {
  if (c->prm.flux_limiter == HG_LIMITER_FLAT) return;
  const int mag = c->prm.mag_on, lim = c->prm.flux_limiter;
  for (int dir = 0; dir < 3; ++dir) {
    if (!c->solve[dir]) continue;
    int lo[3] = {c->is, c->js, c->ks}, hi[3] = {c->ie, c->je, c->ke};
    lo[dir] -= 1; hi[dir] += 1;
    size_t st = dir == 0 ? 1 : dir == 1 ? (size_t)c->nxg : (size_t)c->nxg * c->nyg;
    const double *X = c->xs[dir], *XI = c->xis[dir], *IDX = c->idxs[dir];
    double **S = c->sl[dir];
#pragma omp parallel for collapse(2) num_threads(NT)
    for (int k = lo[2]; k <= hi[2]; ++k)
      for (int j = lo[1]; j <= hi[1]; ++j)
        for (int i = lo[0]; i <= hi[0]; ++i) {
          int ii = dir == 0 ? i : dir == 1 ? j : k;
          /* interpolation.f90:59,62 (and y, z): no slopes in the ghost cell
             of a flux-boundary face */
          if ((ii == lo[dir] && c->prm.bc[2 * dir] == 10) ||
              (ii == hi[dir] && c->prm.bc[2 * dir + 1] == 10))
            continue;
          size_t m = ID(c, i, j, k), l = m - st, r = m + st;
          double dx[2] = {IDX[ii], IDX[ii + 1]};
          double x[3] = {X[ii - 1], X[ii], X[ii + 1]};
          double xi[2] = {XI[ii - 1], XI[ii]};
          double hl = X[ii] - XI[ii - 1], hr = XI[ii] - X[ii];
          double uu[3], du;
          double dlv, drv, elv, erv, m1l, m1r, m2l, m2r, m3l, m3r;
          double b1l, b1r, b2l, b2r, b3l, b3r;
#define SLOPE(EXPR_L, EXPR_M, EXPR_R, SLOT, OUTL, OUTR)                     \
  uu[0] = EXPR_L; uu[1] = EXPR_M; uu[2] = EXPR_R;                         \
  du = limiter(lim, uu, dx, x, xi);                                        \
  S[SLOT][m] = du;                                                         \
  OUTL = uu[1] - hl * du; OUTR = uu[1] + hr * du;
          SLOPE(c->d[l], c->d[m], c->d[r], 0, dlv, drv)
          SLOPE(c->e[l], c->e[m], c->e[r], 1, elv, erv)
          SLOPE(c->d[l] * c->v1[l], c->d[m] * c->v1[m], c->d[r] * c->v1[r], 2, m1l, m1r)
          SLOPE(c->d[l] * c->v2[l], c->d[m] * c->v2[m], c->d[r] * c->v2[r], 3, m2l, m2r)
          SLOPE(c->d[l] * c->v3[l], c->d[m] * c->v3[m], c->d[r] * c->v3[r], 4, m3l, m3r)
          if (mag) {
            SLOPE(c->b1[l], c->b1[m], c->b1[r], 5, b1l, b1r)
            SLOPE(c->b2[l], c->b2[m], c->b2[r], 6, b2l, b2r)
            SLOPE(c->b3[l], c->b3[m], c->b3[r], 7, b3l, b3r)
          } else {
            b1l = b1r = b2l = b2r = b3l = b3r = 0.0;
          }
#undef SLOPE
          /* species slopes (interpolation.f90:135-141): not reset by the
             e_int guard below */
          for (int n = 0; n < c->spn; ++n) {
            uu[0] = c->spc[n][l]; uu[1] = c->spc[n][m]; uu[2] = c->spc[n][r];
            c->dspc[dir][n][m] = limiter(lim, uu, dx, x, xi);
          }
          double eintl = elv - 0.5 * (m1l * m1l + m2l * m2l + m3l * m3l) / dlv;
          double eintr = erv - 0.5 * (m1r * m1r + m2r * m2r + m3r * m3r) / drv;
          if (mag) {
            eintl = eintl - 0.5 * (b1l * b1l + b2l * b2l + b3l * b3l);
            eintr = eintr - 0.5 * (b1r * b1r + b2r * b2r + b3r * b3r);
          }
          double el_ = c->eint[l], em_ = c->eint[m], er_ = c->eint[r];
          if (eintl >= fmax2(el_, em_) || eintl <= fmin2(el_, em_) ||
              eintr >= fmax2(em_, er_) || eintr <= fmin2(em_, er_)) {
            for (int a = 0; a < 9; ++a) S[a][m] = 0.0;
          }
        }
  }
}

/* ------------------------------------------------------------------------ */
/* numflux.f90:14-571                                                        */
/* ------------------------------------------------------------------------ */

static int numflux(or_ctx *c) {
  interpolation(c);
  const int mag = c->prm.mag_on;
  const double gam = c->prm.gamma, ch = c->ch;
  int err = 0;
  for (int dir = 0; dir < 3; ++dir) {
    if (!c->solve[dir]) continue;
    int lo[3] = {c->is, c->js, c->ks}, hi[3] = {c->ie, c->je, c->ke};
    lo[dir] -= 1;
    size_t st = dir == 0 ? 1 : dir == 1 ? (size_t)c->nxg : (size_t)c->nxg * c->nyg;
    const double *X = c->xs[dir], *XI = c->xis[dir];
    double **S = c->sl[dir];
    /* cyclic rotation a,b,c of the axes (numflux.f90:73-93,246-266,419-439) */
    int A = dir, B = (dir + 1) % 3, C = (dir + 2) % 3;
    double *vv[3] = {c->v1, c->v2, c->v3};
    double *bb[3] = {c->b1, c->b2, c->b3};
    int imo[3] = {c->imo1, c->imo2, c->imo3};
    int img[3] = {c->img1, c->img2, c->img3};
    double **F = c->flux[dir];
#pragma omp parallel for collapse(2) num_threads(NT) reduction(| : err)
    for (int k = lo[2]; k <= hi[2]; ++k)
      for (int j = lo[1]; j <= hi[1]; ++j)
        for (int i = lo[0]; i <= hi[0]; ++i) {
          int ii = dir == 0 ? i : dir == 1 ? j : k;
          /* numflux.f90:59-62 (and y, z): flux-boundary faces are not
             solved; their flux is set below */
          if ((ii == lo[dir] && c->prm.bc[2 * dir] == 10) ||
              (ii == hi[dir] && c->prm.bc[2 * dir + 1] == 10))
            continue;
          size_t L = ID(c, i, j, k), R = L + st;
          double dx1 = XI[ii] - X[ii], dx2 = X[ii + 1] - XI[ii];
          double dl = c->d[L] + dx1 * S[0][L];
          double dr = c->d[R] - dx2 * S[0][R];
          double el = c->e[L] + dx1 * S[1][L];
          double er = c->e[R] - dx2 * S[1][R];
          double v1l = (c->d[L] * vv[A][L] + dx1 * S[2 + A][L]) / dl;
          double v1r = (c->d[R] * vv[A][R] - dx2 * S[2 + A][R]) / dr;
          double v2l = (c->d[L] * vv[B][L] + dx1 * S[2 + B][L]) / dl;
          double v2r = (c->d[R] * vv[B][R] - dx2 * S[2 + B][R]) / dr;
          double v3l = (c->d[L] * vv[C][L] + dx1 * S[2 + C][L]) / dl;
          double v3r = (c->d[R] * vv[C][R] - dx2 * S[2 + C][R]) / dr;
          double b1l, b1r, b2l, b2r, b3l, b3r, phil, phir;
          if (mag) {
            b1l = bb[A][L] + dx1 * S[5 + A][L];
            b1r = bb[A][R] - dx2 * S[5 + A][R];
            b2l = bb[B][L] + dx1 * S[5 + B][L];
            b2r = bb[B][R] - dx2 * S[5 + B][R];
            b3l = bb[C][L] + dx1 * S[5 + C][L];
            b3r = bb[C][R] - dx2 * S[5 + C][R];
            phil = c->phi[L] + dx1 * S[8][L];
            phir = c->phi[R] - dx2 * S[8][R];
          } else {
            b1l = b1r = b2l = b2r = b3l = b3r = phil = phir = 0.0;
          }
          double eil = get_eint(mag, el, dl, v1l, v2l, v3l, b1l, b2l, b3l);
          double eir = get_eint(mag, er, dr, v1r, v2r, v3r, b1r, b2r, b3r);
          double ptl = 0, ptr = 0, csl = 0, csr = 0;
          int e1 = eos_p_cs(gam, dl, eil, 1, 0.0, &ptl, &csl);
          int e2 = eos_p_cs(gam, dr, eir, 1, 0.0, &ptr, &csr);
          if (e1 | e2) {
            err |= 1;
#pragma omp critical
            {
              c->err_i = i; c->err_j = j; c->err_k = k; c->err_dir = dir + 1;
            }
            continue;
          }
          ptl = ptl + 0.5 * (b1l * b1l + b2l * b2l + b3l * b3l);
          ptr = ptr + 0.5 * (b1r * b1r + b2r * b2r + b3r * b3r);
          double cfl = get_cf(dl, csl, b1l, b2l, b3l);
          double cfr = get_cf(dr, csr, b1r, b2r, b3r);
          double t[10];
          hlldflux(t, cfl, cfr, v1l, v1r, v2l, v2r, v3l, v3r, dl, dr, el, er,
                   ptl, ptr, b1l, b1r, b2l, b2r, b3l, b3r, phil, phir, ch,
                   0.0, 0.0);
          F[c->icnt][L] = t[0];
          F[imo[A]][L] = t[1];
          F[imo[B]][L] = t[2];
          F[imo[C]][L] = t[3];
          F[c->iene][L] = t[4];
          if (mag) {
            F[img[A]][L] = t[5];
            F[img[B]][L] = t[6];
            F[img[C]][L] = t[7];
            if (c->dim > 1) F[c->i9wv][L] = t[8];
          }
          if (c->spn) {
            /* species upwinded with the mass flux, renormalised by the
               upwind sum (numflux.f90:200-212) */
            double spcl[HG_SPN_MAX], spcr[HG_SPN_MAX], suml = 0.0, sumr = 0.0;
            for (int n = 0; n < c->spn; ++n) {
              spcl[n] = c->spc[n][L] + dx1 * c->dspc[dir][n][L];
              spcr[n] = c->spc[n][R] - dx2 * c->dspc[dir][n][R];
            }
            for (int n = 0; n < c->spn; ++n) { suml = suml + spcl[n]; sumr = sumr + spcr[n]; }
            double sgn = fsign(0.5, F[c->icnt][L]);
            double fix = 1.0 / (suml * (0.5 + sgn) + sumr * (0.5 - sgn));
            for (int n = 0; n < c->spn; ++n)
              c->spcflx[dir][n][L] = fix * F[c->icnt][L] *
                                     (spcl[n] * (0.5 + sgn) + spcr[n] * (0.5 - sgn));
          }
        }
  }
  /* fluxboundary (numflux.f90:566, fluxboundary.f90:66-78): the given
     flux on every face of a flux boundary */
  for (int dir = 0; dir < 3; ++dir) {
    if (!c->solve[dir]) continue;
    for (int side = 0; side < 2; ++side) {
      if (c->prm.bc[2 * dir + side] != 10) continue;
      int lo[3] = {c->is, c->js, c->ks}, hi[3] = {c->ie, c->je, c->ke};
      int f = side ? hi[dir] : lo[dir] - 1;
      lo[dir] = hi[dir] = f;
      for (int k = lo[2]; k <= hi[2]; ++k)
        for (int j = lo[1]; j <= hi[1]; ++j)
          for (int i = lo[0]; i <= hi[0]; ++i)
            for (int v = 0; v < c->ufnmax; ++v)
              c->flux[dir][v][ID(c, i, j, k)] = c->fxb[2 * dir + side][v];
    }
  }
  return err ? HG_ERR_FACE_EOS : HG_OK;
}

/* ------------------------------------------------------------------------ */
/* source.f90:14-85 (Cartesian), get_fieldforce source.f90:177-228           */
/* ------------------------------------------------------------------------ */

static void source(or_ctx *c) {
  const int grav = c->prm.gravswitch == 3;
#pragma omp parallel for collapse(2) num_threads(NT)
  for (int k = c->ks; k <= c->ke; ++k)
    for (int j = c->js; j <= c->je; ++j)
      for (int i = c->is; i <= c->ie; ++i) {
        size_t m = ID(c, i, j, k);
        for (int n = 0; n < c->ufnmax; ++n) c->src[n][m] = 0.0;
        if (!grav) continue;
        const double *ph = c->grvphi;
        double g[3];
        int ijk[3] = {i, j, k};
        size_t st[3] = {1, (size_t)c->nxg, (size_t)c->nxg * c->nyg};
        for (int a = 0; a < 3; ++a) {
          int q = ijk[a];
          const double *DX = c->dxs[a], *IDX = c->idxs[a];
          g[a] = -((DX[q] * IDX[q + 1] * ph[m + st[a]] -
                    DX[q + 1] * IDX[q] * ph[m - st[a]]) /
                       (DX[q] + DX[q + 1]) +
                   (DX[q + 1] - DX[q]) * IDX[q] * IDX[q + 1] * ph[m]) *
                 c->d[m];
        }
        c->src[c->imo1][m] = g[0];
        c->src[c->imo2][m] = g[1];
        c->src[c->imo3][m] = g[2];
        c->src[c->iene][m] = g[0] * c->v1[m] + g[1] * c->v2[m] + g[2] * c->v3[m];
      }
}

/* ------------------------------------------------------------------------ */
/* rungekutta.f90:16-98, get_runge_coeff :146-183, primitive :193-232,       */
/* pressure / internalenergy eos.f90:742-842                                 */
/* ------------------------------------------------------------------------ */

static void get_runge_coeff(int iter, int type, double *faco, double *fact,
                            double *facn) {
  if (type == 3) {
    if (iter == 1) { *faco = 1.0; *fact = 1.0; *facn = 0.0; }
    else if (iter == 2) { *faco = 0.75; *fact = 0.25; *facn = *fact; }
    else { *faco = 1.0 / 3.0; *fact = 2.0 / 3.0; *facn = *fact; }
  } else if (type == 2) {
    if (iter == 1) { *faco = 1.0; *fact = 1.0; *facn = 0.0; }
    else { *faco = 0.5; *fact = 0.5; *facn = *fact; }
  } else {
    *faco = 1.0; *fact = 1.0; *facn = 0.0;
  }
}

int or_pressure(or_ctx *c) {
  const int mag = c->prm.mag_on;
  const double gam = c->prm.gamma;
  int err = 0;
#pragma omp parallel for collapse(2) num_threads(NT) reduction(| : err)
  for (int k = c->ks; k <= c->ke; ++k)
    for (int j = c->js; j <= c->je; ++j)
      for (int i = c->is; i <= c->ie; ++i) {
        size_t m = ID(c, i, j, k);
        c->eint[m] = get_eint(mag, c->e[m], c->d[m], c->v1[m], c->v2[m],
                              c->v3[m], c->b1[m], c->b2[m], c->b3[m]);
        if (c->eint[m] <= 0.0) {
          err |= 1;
#pragma omp critical
          { c->err_i = i; c->err_j = j; c->err_k = k; c->err_dir = 0; }
        }
      }
  if (err) return HG_ERR_NEG_EINT;
#pragma omp parallel for collapse(2) num_threads(NT)
  for (int k = c->ks; k <= c->ke; ++k)
    for (int j = c->js; j <= c->je; ++j)
      for (int i = c->is; i <= c->ie; ++i) {
        size_t m = ID(c, i, j, k);
        c->p[m] = (gam - 1.0) * c->eint[m];
        c->ptot[m] = c->p[m];
        if (!mag) continue;
        double bsq = c->b1[m] * c->b1[m] + c->b2[m] * c->b2[m] + c->b3[m] * c->b3[m];
        c->ptot[m] = c->ptot[m] + 0.5 * bsq;
      }
  return HG_OK;
}

static int rungekutta(or_ctx *c, int rungen, double dt) {
  double faco, fact, facn;
  get_runge_coeff(rungen, c->prm.rktype, &faco, &fact, &facn);
  const double *I1 = c->idxis[0], *I2 = c->idxis[1], *I3 = c->idxis[2];
  size_t sx = 1, sy = (size_t)c->nxg, sz = (size_t)c->nxg * c->nyg;
  for (int n = 0; n < c->ufnmax; ++n) {
    double *U = c->u[n], *U0 = c->uorg[n], *S = c->src[n];
    double *F1 = c->flux[0][n], *F2 = c->flux[1][n], *F3 = c->flux[2][n];
#pragma omp parallel for collapse(2) num_threads(NT)
    for (int k = c->ks; k <= c->ke; ++k)
      for (int j = c->js; j <= c->je; ++j)
        for (int i = c->is; i <= c->ie; ++i) {
          size_t m = ID(c, i, j, k);
          if (rungen == 1) U0[m] = U[m];
          /* detg1 = detg2 = 1 and idetg = idxi for crdnt 0 (metric.f90:26-51) */
          U[m] = faco * U0[m] + facn * U[m] +
                 fact * dt *
                     (I1[i] * (1.0 * F1[m - sx] - 1.0 * F1[m]) +
                      I2[j] * (1.0 * F2[m - sy] - 1.0 * F2[m]) +
                      I3[k] * (F3[m - sz] - F3[m]) + S[m]);
        }
  }
  /* species (rungekutta.f90:61-85): mass fractions from the partial
     densities, with the stage-start primitive density d and the new u(icnt) */
  for (int n = 0; n < c->spn; ++n) {
    double *X = c->spc[n], *X0 = c->spcorg[n];
    double *F1 = c->spcflx[0][n], *F2 = c->spcflx[1][n], *F3 = c->spcflx[2][n];
    const double *D0 = c->uorg[c->icnt], *UN = c->u[c->icnt];
#pragma omp parallel for collapse(2) num_threads(NT)
    for (int k = c->ks; k <= c->ke; ++k)
      for (int j = c->js; j <= c->je; ++j)
        for (int i = c->is; i <= c->ie; ++i) {
          size_t m = ID(c, i, j, k);
          if (rungen == 1) X0[m] = X[m] * D0[m];
          X[m] = (faco * X0[m] + facn * X[m] * c->d[m] +
                  fact * dt *
                      (I1[i] * (1.0 * F1[m - sx] - 1.0 * F1[m]) +
                       I2[j] * (1.0 * F2[m - sy] - 1.0 * F2[m]) +
                       I3[k] * (F3[m - sz] - F3[m]))) /
                 UN[m];
        }
  }
  /* primitive */
  const int mag = c->prm.mag_on;
#pragma omp parallel for collapse(2) num_threads(NT)
  for (int k = c->ks; k <= c->ke; ++k)
    for (int j = c->js; j <= c->je; ++j)
      for (int i = c->is; i <= c->ie; ++i) {
        size_t m = ID(c, i, j, k);
        c->d[m] = c->u[c->icnt][m];
        c->v1[m] = c->u[c->imo1][m] / c->d[m];
        c->v2[m] = c->u[c->imo2][m] / c->d[m];
        c->v3[m] = c->u[c->imo3][m] / c->d[m];
        c->e[m] = c->u[c->iene][m];
        if (mag) {
          c->b1[m] = c->u[c->img1][m];
          c->b2[m] = c->u[c->img2][m];
          c->b3[m] = c->u[c->img3][m];
          if (c->dim > 1) c->phi[m] = c->u[c->i9wv][m];
        }
      }
  return or_pressure(c);
}

int or_hydro_step(or_ctx *c, double dt, double ch) {
  c->ch = ch;
  for (int rungen = 1; rungen <= c->prm.rktype; ++rungen) {
    boundarycondition(c);
    int st = numflux(c);
    if (st) return st;
    source(c);
    st = rungekutta(c, rungen, dt);
    if (st) return st;
  }
  return HG_OK;
}

int or_phidamp(or_ctx *c, double factor) {
#pragma omp parallel for collapse(2) num_threads(NT)
  for (int k = c->ks; k <= c->ke; ++k)
    for (int j = c->js; j <= c->je; ++j)
      for (int i = c->is; i <= c->ie; ++i) {
        size_t m = ID(c, i, j, k);
        c->phi[m] = c->phi[m] * factor;
      }
  return HG_OK;
}

/* ------------------------------------------------------------------------ */
/* timestep.f90:15-105, dti_cell :125-183                                    */
/* ------------------------------------------------------------------------ */

int or_timestep(or_ctx *c, double time, double t_out, double t_end,
                int outstyle, int endstyle, double fixed_dt, double *dt_out,
                double *ch_out, double *cgrav_out, double *dtgrav_out) {
  double dt;
  if (fixed_dt > 0.0) {
    dt = fixed_dt;
    if (dt_out) *dt_out = dt;
    if (ch_out) *ch_out = c->ch;
    return HG_OK;
  }
  const int mag = c->prm.mag_on;
  const double gam = c->prm.gamma;
  const double off1 = c->solve[0] ? 1.0 : 0.0, off2 = c->solve[1] ? 1.0 : 0.0,
               off3 = c->solve[2] ? 1.0 : 0.0;
  const double *DX1 = c->dxis[0], *DX2 = c->dxis[1], *DX3 = c->dxis[2];
  double cfmax = 0.0, dtimin = INFINITY;
#pragma omp parallel for collapse(2) num_threads(NT) reduction(max : cfmax) reduction(min : dtimin)
  for (int k = c->ks; k <= c->ke; ++k)
    for (int j = c->js; j <= c->je; ++j)
      for (int i = c->is; i <= c->ie; ++i) {
        size_t m = ID(c, i, j, k);
        double pp, cc;
        if (eos_p_cs(gam, c->d[m], c->eint[m], 0, 0.0, &pp, &cc) == 0) {
          c->p[m] = pp;
          c->cs[m] = cc;
        }
        double cs = c->cs[m], cf1, cf2, cf3;
        if (mag) {
          cf1 = get_cf(c->d[m], cs, c->b1[m], c->b2[m], c->b3[m]) + fabs(c->v1[m]);
          cf2 = get_cf(c->d[m], cs, c->b2[m], c->b1[m], c->b3[m]) + fabs(c->v2[m]);
          cf3 = get_cf(c->d[m], cs, c->b3[m], c->b1[m], c->b2[m]) + fabs(c->v3[m]);
        } else {
          cf1 = cs + fabs(c->v1[m]);
          cf2 = cs + fabs(c->v2[m]);
          cf3 = cs + fabs(c->v3[m]);
        }
        /* metric.f90:43-48: dvol, sa1..3 */
        double dvol = DX1[i] * DX2[j] * DX3[k];
        double sa1 = DX2[j] * DX3[k], sa2 = DX1[i] * DX3[k], sa3 = DX1[i] * DX2[j];
        double sa1m = DX2[j] * DX3[k], sa2m = DX1[i] * DX3[k], sa3m = DX1[i] * DX2[j];
        double dti = dvol / (cf1 * off1 * (sa1m + sa1) + cf2 * off2 * (sa2m + sa2) +
                             cf3 * off3 * (sa3m + sa3));
        c->dti[m] = dti;
        double cfm = fmax3(cf1, cf2, cf3);
        cfmax = fmax2(cfmax, cfm);
        dtimin = fmin2(dtimin, dti);
      }
  dt = c->prm.courant * dtimin;
  if (outstyle == 1) {
    double t_ref = t_out; /* used as is when endstyle != 1 */
    if (endstyle == 1) t_ref = fmin2(t_out, t_end);
    if (t_ref - time - dt > 0.0 && t_ref - time - 2.0 * dt < 0.0)
      dt = 0.5 * (t_ref - time);
    else
      dt = fmin2(dt, t_out - time);
  }
  if (endstyle == 1) dt = fmin2(dt, t_end - time);
  c->ch = cfmax;
  if (dt_out) *dt_out = dt;
  if (ch_out) *ch_out = cfmax;
  if (c->prm.gravswitch == 3) {
    /* dtg_unit: tools.f90:154-184 with dtgrav_cell (cgrav = 1) */
    double dtg_unit = INFINITY;
#pragma omp parallel for collapse(2) num_threads(NT) reduction(min : dtg_unit)
    for (int k = c->ks; k <= c->ke; ++k)
      for (int j = c->js; j <= c->je; ++j)
        for (int i = c->is; i <= c->ie; ++i) {
          double dvol = DX1[i] * DX2[j] * DX3[k];
          double sa1 = DX2[j] * DX3[k], sa2 = DX1[i] * DX3[k], sa3 = DX1[i] * DX2[j];
          double v = dvol / (1.0 * (off1 * (sa1 + sa1) + off2 * (sa2 + sa2) +
                                    off3 * (sa3 + sa3)));
          dtg_unit = fmin2(dtg_unit, v);
        }
    double cgrav = c->prm.HGfac * cfmax;
    if (c->prm.maxtngrv > 0)
      cgrav = fmin2(cgrav, c->prm.hgcfl * dtg_unit * (double)c->prm.maxtngrv / dt);
    double dtgrav = fmin2(dt, c->prm.hgcfl * dtg_unit / cgrav);
    if (cgrav_out) *cgrav_out = cgrav;
    if (dtgrav_out) *dtgrav_out = dtgrav;
  }
  return HG_OK;
}

/* ------------------------------------------------------------------------ */
/* hyperbolic self-gravity, gravity_hyperbolic.f90:19-125,230-280 and       */
/* gravity.f90:114-150 (gsrc = rho), for crdnt 0 with the potential's ghost */
/* layer held at the values uploaded with HG_F_GRVPHI (Dirichlet).          */
/* ------------------------------------------------------------------------ */

static void hg_stage(or_ctx *c, int grungen, double dtg, double cg) {
  double faco, fact, facn;
  get_runge_coeff(grungen, c->prm.grktype, &faco, &fact, &facn);
  const double fourpig = 4.0 * acos(-1.0) * c->prm.G;
  size_t st[3] = {1, (size_t)c->nxg, (size_t)c->nxg * c->nyg};
#pragma omp parallel for collapse(2) num_threads(NT)
  for (int k = c->ks; k <= c->ke; ++k)
    for (int j = c->js; j <= c->je; ++j)
      for (int i = c->is; i <= c->ie; ++i) {
        size_t m = ID(c, i, j, k);
        const double *ph = c->grvphi;
        double lap = c->lapc[1][m] * ph[m];
        lap = lap + c->lapc[2][m - st[0]] * ph[m - st[0]];
        lap = lap + c->lapc[2][m] * ph[m + st[0]];
        if (c->je > c->js) {
          lap = lap + c->lapc[3][m - st[1]] * ph[m - st[1]];
          lap = lap + c->lapc[3][m] * ph[m + st[1]];
        }
        if (c->ke > c->ks) {
          lap = lap + c->lapc[4][m - st[2]] * ph[m - st[2]];
          lap = lap + c->lapc[4][m] * ph[m + st[2]];
        }
        c->lapphi[m] = lap / c->lapc[0][m];
        c->hgsrc[m] = fourpig * c->gsrc[m];
      }
#pragma omp parallel for collapse(2) num_threads(NT)
  for (int k = c->ks; k <= c->ke; ++k)
    for (int j = c->js; j <= c->je; ++j)
      for (int i = c->is; i <= c->ie; ++i) {
        size_t m = ID(c, i, j, k);
        if (grungen == 1) {
          c->grvphiorg[0][m] = c->grvphi[m];
          c->grvphiorg[1][m] = c->grvpsi[m];
        }
        c->grvphi[m] = faco * c->grvphiorg[0][m] + facn * c->grvphi[m] +
                       fact * dtg * cg * c->grvpsi[m];
        c->grvpsi[m] = faco * c->grvphiorg[1][m] + facn * c->grvpsi[m] +
                       fact * dtg * cg * (c->lapphi[m] - c->hgsrc[m]);
      }
}

int or_gravity(or_ctx *c, double time, double dt, double *grvtime, double cgrav,
               double *cgrav_old, double *dtgrav, int *nsub) {
  if (c->prm.gravswitch != 3) return HG_ERR_ARG;
#pragma omp parallel for collapse(2) num_threads(NT)
  for (int k = c->ks; k <= c->ke; ++k)
    for (int j = c->js; j <= c->je; ++j)
      for (int i = c->is; i <= c->ie; ++i) {
        size_t m = ID(c, i, j, k);
        c->gsrc[m] = c->d[m];
      }
  int tngrav = (int)ceil((time + dt - *grvtime) / *dtgrav);
  *dtgrav = (time + dt - *grvtime) / (double)tngrav;
  for (int l = 1; l <= tngrav; ++l) {
    for (int g = 1; g <= c->prm.grktype; ++g) hg_stage(c, g, *dtgrav, cgrav);
    const double damp =
        exp(-(1.0 - *cgrav_old / cgrav + c->prm.alphagrv * c->prm.hgcfl));
#pragma omp parallel for collapse(2) num_threads(NT)
    for (int k = c->ks; k <= c->ke; ++k)
      for (int j = c->js; j <= c->je; ++j)
        for (int i = c->is; i <= c->ie; ++i) {
          size_t m = ID(c, i, j, k);
          c->grvpsi[m] = c->grvpsi[m] * damp;
        }
    *grvtime = *grvtime + *dtgrav;
    *cgrav_old = cgrav;
  }
  if (nsub) *nsub = tngrav;
  return HG_OK;
}

/* ------------------------------------------------------------------------ */
/* known-answer entry points                                                 */
/* ------------------------------------------------------------------------ */

void or_kat_hlldflux(const double *a, double *out10) {
  /* a[0..24]: cfl cfr v1l v1r v2l v2r v3l v3r dl dr el er ptl ptr b1l b1r
     b2l b2r b3l b3r phil phir ch erl err */
  /* argument order of hlldflux.f90:19-21 */
  hlldflux(out10, a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9],
           a[10], a[11], a[12], a[13], a[14], a[15], a[16], a[17], a[18],
           a[19], a[20], a[21], a[22], a[23], a[24]);
}

double or_kat_get_cf(double d, double cs, double b1, double b2, double b3) {
  return get_cf(d, cs, b1, b2, b3);
}

int or_kat_eos_p_cs(double d, double eint, double gamma, double erad,
                    int has_erad, double *p, double *cs) {
  return eos_p_cs(gamma, d, eint, has_erad, erad, p, cs);
}

double or_kat_modified_mc(const double *u3, const double *x3, const double *xi2) {
  return modified_mc(u3, x3, xi2);
}

double or_kat_minmod(const double *u3, const double *dx2) { return minmod(u3, dx2); }

double or_kat_get_eint(int mag_on, double etot, double d, double v1, double v2, double v3,
                       double b1, double b2, double b3) {
  return get_eint(mag_on, etot, d, v1, v2, v3, b1, b2, b3);
}
