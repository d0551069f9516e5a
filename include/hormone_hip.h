/*
 * hormone_hip.h -- C ABI of the MI355X (gfx950) explicit hydro step for
 * HORMONE's Cartesian grid.
 *
 * The reference has no plugin API: its hot path is four argument-less
 * Fortran subroutines working on module globals, called from the main loop
 * (src/main/hormone.f90:126,130,136,140):
 *     timestep   src/main/timestep.f90:15-105
 *     gravity    src/main/gravity.f90:18-64 -> gravity_hyperbolic.f90:19-50
 *     hydro_step src/main/hydro.f90:15-52
 *     phidamp    src/main/source.f90:239-260
 * Each entry point below replaces one of them (cited per function).  A
 * Fortran driver binds them with iso_c_binding (hormone_amd/fortran/
 * hormone_hip_mod.f90).  All pointers are host pointers; the library owns
 * the device copies, keeps the state resident between calls and retains no
 * host pointer after a call returns.  Every function returns an HG_* status.
 *
 * Array convention: a "padded" array is one 3D field laid out exactly like
 * the reference's allocations (src/main/allocations.f90:56-57,212): Fortran
 * order, x fastest, index ranges is-2:ie+2, js-2:je+2, ks-2:ke+2 (ghost width
 * 2 in every dimension, active or not): the address of a(is-2,js-2,ks-2),
 * which a Fortran caller passes by reference (hormone_hip_mod.f90 declares
 * it real(c_double) padded(*), so the array needs no TARGET attribute).
 */
#ifndef HORMONE_HIP_H
#define HORMONE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (mirror the reference's stop sites) ------------------ */
enum {
  HG_OK = 0,
  HG_ERR_ARG = 1,         /* bad argument / descriptor                       */
  HG_ERR_UNSUPPORTED = 2, /* crdnt/=0, eostype/=0, compswitch 1 or >2, a
                             boundary code a face does not take (see
                             hg_desc.bc), bc 9/10 with species, gravswitch
                             other than 0 and 3                              */
  HG_ERR_DEVICE = 3,      /* a HIP runtime call failed                       */
  HG_ERR_FACE_EOS = 4,    /* numflux.f90:579-593 error_flux: rho<=0 or
                             e_int<=0 at a reconstructed face state          */
  HG_ERR_NEG_EINT = 5,    /* eos.f90:833-839 internalenergy: e_int<=0 in a
                             cell after the RK update                        */
  HG_ERR_NOMEM = 6,       /* device allocation failed                        */
  HG_ERR_STATE = 7,       /* call order violated (e.g. step before upload)   */
  HG_ERR_SCRATCH = 8      /* input larger than an allocation-free entry
                             point's static scratch (hg_timestep_noalloc)   */
};

/* flux_limiter (settings: 'flat', 'minmod', 'modified_mc';
   src/main/fluxlimiter.f90:27-35, interpolation.f90:42-45) */
enum { HG_LIMITER_FLAT = 0, HG_LIMITER_MINMOD = 1, HG_LIMITER_MODIFIED_MC = 2 };

/* precision of the arithmetic on the path */
enum { HG_PREC_FP64 = 0, HG_PREC_FP32MIX = 1 };

/* boundary codes accepted (boundarycondition.f90:30-38 and the faces' own
   select blocks): 0 periodic, 1 reflective, 2 outgoing, 3 free everywhere;
   9 Dirichlet and 10 flux everywhere (not with species); 4 linear and 5
   linear + outgoing at x1 outer, x3 inner and x3 outer (scalar code 5 not
   at x3 outer).  See hg_desc.bc. */

/* fields for hg_upload / hg_download.  Primitive fields are the reference's
   physval arrays (modules.f90:106-138); HG_F_U + n (n = 0..ufnmax-1) is the
   conserved variable u(:,:,:,n+1) in the reference's equation order
   icnt,imo1,imo2,imo3,iene[,img1,img2,img3[,i9wv]] (checksetup.f90:62-75). */
enum { HG_SPN_MAX = 8 }; /* species advected per context (compswitch 2) */

enum {
  HG_F_D = 0, HG_F_V1 = 1, HG_F_V2 = 2, HG_F_V3 = 3, HG_F_E = 4,
  HG_F_B1 = 5, HG_F_B2 = 6, HG_F_B3 = 7, HG_F_PHI = 8,
  HG_F_EINT = 9,      /* download only: internal energy                   */
  HG_F_P = 10,        /* download only: gas pressure (gamma-1)*eint       */
  HG_F_GRVPHI = 11,   /* gravitational potential (gravswitch 3)           */
  HG_F_GRVPSI = 12,   /* its time derivative variable psi                 */
  HG_F_U = 16,        /* + n: conserved variable n                        */
  HG_F_SPC = 32       /* + n: mass fraction of species n+1, spc(n+1,:,:,:)
                         (compswitch 2)                                    */
};

/* Everything the hot path reads from the reference's settings, grid and
   physval modules (modules.f90:8-184), gathered into one struct. */
typedef struct hg_desc {
  /* interior index ranges (grid module: is,ie,js,je,ks,ke).  A dimension is
     solved when its range holds more than one cell (checksetup.f90:41-44). */
  int32_t is, ie, js, je, ks, ke;
  /* Cartesian coordinate arrays of the grid module (gridset.f90:43-158),
     each covering indices is-2..ie+2 (resp. j, k): the address of
     x1(is-2); read during hg_create only (hormone_hip_mod.f90's
     hg_create_cartesian sets them from arrays passed by reference). */
  const double *x1, *xi1, *dx1, *dxi1, *idx1, *idxi1;
  const double *x2, *xi2, *dx2, *dxi2, *idx2, *idxi2;
  const double *x3, *xi3, *dx3, *dxi3, *idx3, *idxi3;
  int32_t crdnt;        /* must be 0 (Cartesian)                            */
  int32_t eostype;      /* must be 0 (ideal gas)                            */
  int32_t compswitch;   /* 0 (uniform composition) or 2 (advected species)  */
  int32_t mag_on;       /* 0/1                                              */
  int32_t flux_limiter; /* HG_LIMITER_*                                     */
  int32_t rktype;       /* 1, 2 or 3                                        */
  int32_t bc[12];       /* bc1is bc1os bc2is bc2os bc3is bc3os
                           bc1iv bc1ov bc2iv bc2ov bc3iv bc3ov
                           0 periodic, 1 reflective, 2 outgoing, 3 free,
                           4 linear, 5 linear + outgoing (x1 outer, x3
                           inner, x3 outer; scalar 5 not at x3 outer),
                           9 Dirichlet (hg_set_ghosts), 10 flux
                           (hg_set_boundary_flux); 9/10 not with species */
  double gamma;         /* physval: gamma                                   */
  double courant;       /* settings: courant                                */
  double alpha9wave;    /* settings: alpha9wave                             */
  int32_t gravswitch;   /* 0 (off) or 3 (hyperbolic self-gravity)           */
  int32_t grktype;      /* RK order of the gravity wave update (1 or 2)     */
  double HGfac, hgcfl, alphagrv;
  double G;             /* gravitational constant (constants: G)            */
  int32_t maxtngrv;     /* settings: maxtngrv (<=0: no limit)               */
  int32_t precision;    /* HG_PREC_*                                        */
  int32_t device;       /* HIP device ordinal                               */
  int32_t spn;          /* settings: spn, species advected when compswitch=2
                           (1..HG_SPN_MAX)                                  */
} hg_desc;

typedef struct hg_ctx hg_ctx;

/* Library version; no reference counterpart. */
const char *hg_version(void);

/* Validate a descriptor and create a device-resident context (replaces the
   hot-path part of allocations.f90:14-335 + metric.f90:26-51 for crdnt=0).
   On error *out is NULL and the status says why. */
int hg_create(const hg_desc *desc, hg_ctx **out);

/* Release the context and its device memory. */
int hg_destroy(hg_ctx *ctx);

/* Number of conserved variables (checksetup.f90:62-75: ufnmax). */
int hg_ufnmax(const hg_ctx *ctx);

/* Copy the interior of one padded host field into the device state
   (HG_F_D..HG_F_PHI, HG_F_GRVPHI, HG_F_GRVPSI, HG_F_U+n).  For HG_F_GRVPHI
   the first ghost layer is copied too: it is the potential's boundary.  After primitive
   fields change, e_int and p are recomputed on the device from e, rho, v, B
   before the next step, exactly as pressure() does (eos.f90:742-842). */
int hg_upload(hg_ctx *ctx, int32_t field, const double *padded);

/* Copy the interior of one device field into a padded host array; ghost
   cells of the host array are left untouched.  Blocks until done. */
int hg_download(hg_ctx *ctx, int32_t field, double *padded);

/* Replaces timestep (timestep.f90:15-105) for fixed_dt<=0: per-cell CFL
   (dti_cell, timestep.f90:125-183), max-reduction of the fast speed and
   min-reduction of dti, dt=courant*min(dti), clipping to t_out/t_end
   (timestep.f90:74-85, outstyle/endstyle as in the reference), ch=cfmax,
   and for gravswitch 3: cgrav, dtgrav (timestep.f90:92-97).
   fixed_dt>0 returns dt=fixed_dt and leaves ch untouched, like the reference.
   Any output pointer may be NULL. */
int hg_timestep(hg_ctx *ctx, double time, double t_out, double t_end,
                int32_t outstyle, int32_t endstyle, double fixed_dt,
                double *dt, double *ch, double *cgrav, double *dtgrav);

/* Replaces hydro_step (hydro.f90:15-52) for the serial Cartesian case: for
   rungen = 1..rktype: boundary conditions (boundarycondition.f90),
   MUSCL slopes + e_int guard (interpolation.f90), face states + HLLD/GLM flux
   (numflux.f90, hlldflux.f90), gravity source (source.f90:14-85),
   SSP-RK update (rungekutta.f90:16-98) and primitive recovery
   (rungekutta.f90:193-232, eos.f90:742-842).  dt and ch are the values the
   reference holds in its grid/physval modules when it calls hydro_step.
   On HG_ERR_FACE_EOS / HG_ERR_NEG_EINT, hg_last_error reports where. */
int hg_hydro_step(hg_ctx *ctx, double dt, double ch);

/* Replaces gravity for gravswitch 3 and tn > 0 (gravity.f90:18-64 ->
   gravity_hyperbolic.f90:19-50): gsrc = rho, tngrav =
   ceiling((time+dt-grvtime)/dtgrav) sub-steps of the hyperbolic wave
   equation (RK order grktype, hyperbolic_gravity_step :61-125) with the psi
   damping exp(-(1-cgrav_old/cgrav+alphagrv*hgcfl)) evaluated on the host,
   then grvtime += dtgrav and cgrav_old = cgrav per sub-step.  Cartesian
   lap_coeff (the reference has none, gravity_hyperbolic.f90:413-417):
   c0 = -((idx1(i)+idx1(i+1))*dxi2*dxi3 + (idx2(j)+idx2(j+1))*dxi1*dxi3 +
   (idx3(k)+idx3(k+1))*dxi1*dxi2), c1(i) = dxi2*dxi3*idx1(i+1) (c2, c3
   alike), normalised by dxi1*dxi2*dxi3.  The potential's boundary is the
   ghost layer uploaded with HG_F_GRVPHI, held fixed (Dirichlet).
   dtgrav is updated in place; *nsub (may be NULL) receives tngrav. */
int hg_gravity(hg_ctx *ctx, double time, double dt, double *grvtime, double cgrav,
               double *cgrav_old, double *dtgrav, int32_t *nsub);

/* Replaces phidamp (source.f90:239-260): multiplies the primitive psi (not
   the conserved u(:,:,:,i9wv)) by factor = exp(-alpha9wave*courant), which
   the caller evaluates on the host. */
int hg_phidamp(hg_ctx *ctx, double factor);

/* Device-resident main loop; no single counterpart in the reference: it
   runs the hydro-only iteration of hormone.f90:124-179 -- timestep, hydro
   step, phidamp (when damp_factor > 0, mag_on and dim > 1), time += dt,
   tn += 1, then the end check (endstyle 1: time >= t_end; 2: tn >= tnlim)
   and the output check (outstyle 1: time >= t_out, then t_out += dt_out;
   2: mod(tn, tn_out) == 0) -- without returning to the host between steps.
   Returns when the run ends, an output is due (the caller writes it and calls
   again), max_steps steps have been taken, or an error occurs; *stop_reason
   is 1, 2, 3 or 4 respectively.  time, tn and t_out are updated in place.
   Gives the same state, bit for bit, as the per-step entry points.
   Contexts with gravswitch 3 use hg_run_gravity (HG_ERR_UNSUPPORTED here). */
int hg_run(hg_ctx *ctx, double *time, int64_t *tn, double *t_out, double dt_out,
           double t_end, int32_t outstyle, int32_t endstyle, int64_t tnlim,
           int64_t tn_out, int64_t max_steps, double damp_factor,
           int32_t *stop_reason);

/* hg_run for gravswitch 3: the iteration of hormone.f90:124-179 with the
   gravity call between timestep and hydro_step -- timestep (with cgrav and
   dtgrav, timestep.f90:92-97), at tn > 0 gravity_hyperbolic (tngrav
   sub-steps, as hg_gravity; at tn == 0 grvtime = 0, gravity.f90:41-60),
   hydro step, phidamp, clock and end/output checks as hg_run.  grvtime and
   cgrav_old are the gravmod module values, updated in place; *nsub_total
   (may be NULL) receives the sub-steps taken in this call.  The host reads
   one small record of device scalars per step (no stream synchronisation);
   the step's first flux sweep is already running while it does.  Same
   state, bit for bit, as hg_timestep + hg_gravity + hg_hydro_step +
   hg_phidamp. */
int hg_run_gravity(hg_ctx *ctx, double *time, int64_t *tn, double *t_out, double dt_out,
                   double t_end, int32_t outstyle, int32_t endstyle, int64_t tnlim,
                   int64_t tn_out, int64_t max_steps, double damp_factor, double *grvtime,
                   double *cgrav_old, int64_t *nsub_total, int32_t *stop_reason);

/* Kernel timing with HIP events on the context's stream, by category
   (cf. the reference's profiler, profiler.f90:5-51).  hg_profile(ctx, 1)
   resets and enables every category; hg_profile(ctx, HG_PROF_ONLY | mask)
   only those whose bit (1 << HG_PROF_*) is in mask (each timed launch waits
   for its completion signal, so timing fewer launches disturbs a run less);
   hg_profile(ctx, 0) disables.  hg_profile_read returns the summed
   milliseconds, the number of launches and the HBM bytes one launch of the
   kernel itself moves (its own reads and writes, intermediates included).
   The sweep category times the flux sweeps (k_sweep); the RK category the
   RK update, or the fused sweep + RK launch when a direction is fused. */
enum { HG_PROF_SWEEP = 0, HG_PROF_RK = 1, HG_PROF_CFL = 2, HG_PROF_GRAV = 3,
       HG_PROF_N = 4, HG_PROF_ONLY = 0x100 };
int hg_profile(hg_ctx *ctx, int32_t enable);
int hg_profile_read(const hg_ctx *ctx, int32_t kind, double *ms_total,
                    int64_t *launches, double *bytes_per_launch);

/* Ghost values of one field at the faces whose scalar or vector code is 4,
   5, 9 or 10: the values those codes do not compute from the interior.
   Replaces the d0/p0/v10../b10.. arrays a Dirichlet face copies into its
   ghost cells (boundarycondition.f90:104-116, 196-212 and the other faces;
   set by dirichletbound, dirichlet.f90:14-20) and, where the reference
   leaves a ghost cell unwritten (phi under 9, everything under 10, v2 and
   b2 of the outer ghost below the x3 inner boundary under 4/5,
   boundarycondition.f90:1081,1088,1104,1114), the value that cell holds.
   field: HG_F_D, HG_F_P, HG_F_V1..V3, HG_F_B1..B3, HG_F_PHI; padded: a
   padded array like hg_upload's, of which only the two ghost layers of
   those faces are read.  Before any call the values are the reference's
   allocation values (allocations.f90:71-78: d = p = 1, v = B = phi = 0).
   Call again between steps to change them (time-dependent Dirichlet
   data). */
int hg_set_ghosts(hg_ctx *ctx, int32_t field, const double *padded);

/* The flux through a flux-boundary face (scalar code 10), which the
   reference sets in fluxboundary (fluxboundary.f90:17-78) after skipping
   that face's Riemann solve (numflux.f90:59-62, 566).  face: 0..5 = x1
   inner, x1 outer, x2 inner, x2 outer, x3 inner, x3 outer; flux: ufnmax
   values in conserved-variable order, the same at every cell of the face.
   Until set it is zero, the value the reference's flux array keeps where
   fluxboundary does not write it. */
int hg_set_boundary_flux(hg_ctx *ctx, int32_t face, const double *flux);

/* Fuse one direction's sweep with the RK update: on = 1 the last solved
   direction (y or z), on = 2 the x direction, 0 off (the default).  One
   launch per RK stage does that direction's reconstruction, Riemann solves
   and flux difference and updates each cell, so that direction's flux
   divergence is never written to or read back from HBM; the other
   directions are swept first.  Results are bit-identical either way, in
   FP64 and FP32-mixed.  Not with species, boundary codes 4, 5, 9, 10 or
   while the context is in a batch (HG_ERR_UNSUPPORTED).  The first call allocates a second copy
   of the state (the launch reads one and writes the other).  No
   counterpart in the reference (a launch-shape choice). */
int hg_set_fused(hg_ctx *ctx, int32_t on);

/* The sweep kernel's duration without per-dispatch events: `launches`
   back-to-back sweeps of the current state between two events on the
   context's stream; *ms_total = their elapsed milliseconds.  The state is
   unchanged (the sweep overwrites only the flux divergences).  No
   counterpart in the reference; for measurement. */
int hg_time_sweep(hg_ctx *ctx, int32_t launches, double *ms_total);

/* ---- batches of independent boxes (no counterpart in the reference) ------
   Many contexts of possibly different sizes and dimensions, sharing one
   kernel variant (mag_on, flux_limiter, precision, rktype; no gravity, no species),
   advance together: each step is one launch per kernel over flattened
   work-lists of (box, sweep wavefront) and (box, cell block) items, replayed
   from a captured hipGraph.  Each box keeps its own dt, clock and stop
   condition (the same as hg_run's), so its result is bit-identical to
   running it alone.  While a context is in a batch, call nothing else on it
   but hg_download / hg_upload between steps; after hg_batch_remove (or
   hg_batch_destroy) it is an ordinary context again. */
typedef struct hg_batch hg_batch;
int hg_batch_create(int32_t device, hg_batch **out);
int hg_batch_destroy(hg_batch *batch);
/* Add a context with its main-loop clock (as hg_run's arguments). */
int hg_batch_add(hg_batch *batch, hg_ctx *ctx, double time, double t_out, double dt_out,
                 double t_end, int32_t outstyle, int32_t endstyle, int64_t tnlim,
                 int64_t tn_out);
int hg_batch_remove(hg_batch *batch, hg_ctx *ctx);
/* Advance every running member by nsteps of its own steps; *nrunning gets
   the number still running. */
int hg_batch_step(hg_batch *batch, int32_t nsteps, int32_t *nrunning);
/* A member's clock after the last hg_batch_step; stop: 0 running, 1 end,
   2 output due, 4 error (then the member's error status is returned). */
int hg_batch_member_state(hg_batch *batch, hg_ctx *ctx, double *time, int64_t *tn,
                          int32_t *stop);
/* Launch shapes: sweep wavefronts per block (1, 2, 4, 8, 16), persistent grid
   sizes (0 = one work item per wavefront / block) and hipGraph replay. */
int hg_batch_config(hg_batch *batch, int32_t sweep_block_waves, int32_t sweep_grid,
                    int32_t rk_grid, int32_t use_graph);
/* The batch's form: on = 1 (the default) sweeps x together with the RK
   update whenever every member sweeps x (the single-box x-fused form of
   hg_set_fused(ctx, 2); each member then holds a second copy of its state,
   the stages alternating between the two), on = 0 the separate RK update.
   The same bits either way.  hg_batch_form: *fused = the form the batch's
   launches were last built for. */
int hg_batch_set_fused(hg_batch *batch, int32_t on);
int hg_batch_form(const hg_batch *batch, int32_t *fused);

/* Launch shape of a context's sweep kernel: wavefronts per block (1, 2, 4,
   8 or 16; default 4).  Changes which wavefront computes which line, never
   the arithmetic.  No counterpart in the reference. */
int hg_set_launch(hg_ctx *ctx, int32_t sweep_block_waves);

/* Cell (Fortran indices) and direction (1..3, 0 for a cell error) of the
   last HG_ERR_FACE_EOS / HG_ERR_NEG_EINT. */
int hg_last_error(const hg_ctx *ctx, int32_t *i, int32_t *j, int32_t *k,
                  int32_t *dir);

/* hg_timestep without a context and without any device allocation.  No
   counterpart in the reference: it is for callers that cannot allocate
   device memory -- code under HIP graph capture, tight serving loops,
   hosts that keep the state in their own device arrays.  The result is
   hg_timestep's (timestep.f90:35-89 with the dt clipping of :74-85), bit for
   bit, on the same state.
   desc: extents (is..ke), mag_on, gamma and courant are read; nothing else.
   prim: device pointers d, v1, v2, v3, eint and, with mag_on, b1, b2, b3:
     interior fields, nx*ny*nz doubles each, x fastest.
   dxi: device pointers to the interior cell widths dxi1(is:ie), dxi2(js:je),
     dxi3(ks:ke).
   out: device pointer to 3 doubles, written when the stream reaches them:
     dt, ch (= the largest fast speed, cfmax) and 0, or HG_ERR_NEG_EINT when a
     cell holds rho <= 0 or e_int <= 0 (dt is then meaningless).  cgrav and
     dtgrav follow from ch on the host (HGfac*ch, timestep.f90:92-97).
   stream: a hipStream_t (NULL: the null stream).  The call only enqueues
     kernels; it allocates, frees and synchronises nothing, so it may be
     captured into a graph.
   Limit: nx*ny*nz <= HG_NOALLOC_MAX_CELLS, else HG_ERR_SCRATCH.  The working
   memory is one static device array per GPU (HG_NOALLOC_MAX_CELLS/256
   partial pairs plus a 256-byte record, 4 MiB) that every call on that GPU
   shares and that keeps the last call's contents: calls that overlap on one
   GPU (different streams or host threads at once) race. */
enum { HG_NOALLOC_MAX_CELLS = 1 << 26 };
int hg_timestep_noalloc(const hg_desc *desc, const double *const *prim,
                        const double *const *dxi, double time, double t_out,
                        double t_end, int32_t outstyle, int32_t endstyle,
                        double *out, void *stream);

/* Known-answer evaluation of the device arithmetic; no counterpart in the
   reference (its routines are module procedures without a C interface).
   which: 0 hlldflux (25 inputs in hlldflux.f90:19-21 argument order
   cfl..ch,erl,err -> 9 outputs tmpflux(1:9)), 1 get_cf (d cs b1 b2 b3 -> cf),
   2 modified_mc (u(1:3) x(1:3) xi(1:2) -> slope), 3 minmod (u(1:3) dx(1:2)
   -> slope), 4 IEEE a/b and sqrt(a) (a b -> a/b sqrt(a)), 5 the sweep's
   division by a shared denominator (b a0 a1 a2 a3 -> a0/b .. a3/b), 6
   eos_p_cs as the face states evaluate it (eos.f90:195-242 with erad = 0:
   d eint gamma -> p cs), 7 get_eint (eos.f90:846-865: etot d v1 v2 v3 b1
   b2 b3 mag_on -> eint).  which + 16: the same routine in the form the
   sweep kernels run it, with divisions by a shared denominator (kinds 0-3
   and 6).  n records, row-major.  For tests of bit-exactness against the
   reference. */
int hg_kat(int32_t which, int32_t n, const double *in, double *out);


#ifdef __cplusplus
}
#endif

#endif /* HORMONE_HIP_H */
