// hg_physics.hpp -- device arithmetic of HORMONE's hydro step for gfx950.
//
// Every function keeps the reference's Fortran evaluation order (left to
// right, x**2 == x*x) so that, compiled without FMA contraction
// (-ffp-contract=off) and with IEEE division and square root, the FP64 path
// rounds exactly like the reference.  MIN/MAX/SIGN follow the reference's
// compiled semantics: max(a,b) = a>b ? a : b, min(a,b) = a<b ? a : b, n-ary
// forms fold left, sign(a,b) = copysign(|a|,b).
//
// Functions are templated on the arithmetic type R so the same code serves
// the FP64 path (R = double) and the reconstruction/Riemann part of the
// FP32-mixed path (R = float).
#pragma once

#include <hip/hip_runtime.h>

namespace hg {

template <class R> __device__ __forceinline__ R mx2(R a, R b) { return a > b ? a : b; }
template <class R> __device__ __forceinline__ R mn2(R a, R b) { return a < b ? a : b; }
template <class R> __device__ __forceinline__ R mx3(R a, R b, R c) { return mx2(mx2(a, b), c); }
template <class R> __device__ __forceinline__ R mn3(R a, R b, R c) { return mn2(mn2(a, b), c); }
__device__ __forceinline__ double sgn(double a, double b) { return copysign(fabs(a), b); }
__device__ __forceinline__ float sgn(float a, float b) { return copysignf(fabsf(a), b); }
// FP64: IEEE (correctly rounded) square root and division, as the reference.
// FP32 (mixed-precision path only): the hardware v_sqrt_f32 / v_rcp_f32
// approximations (about 1 ulp), which is what makes that path cheaper.
__device__ __forceinline__ double rsqrt_(double a) { return sqrt(a); }
__device__ __forceinline__ float rsqrt_(float a) { return __builtin_amdgcn_sqrtf(a); }
__device__ __forceinline__ double fdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double abs_(double a) { return fabs(a); }
__device__ __forceinline__ float abs_(float a) { return fabsf(a); }

enum { LIM_FLAT = 0, LIM_MINMOD = 1, LIM_MC = 2 };

// fluxlimiter.f90:40-56 (generalised minmod); dx = idx(i:i+1)
template <class R>
__device__ __forceinline__ R minmod(R u0, R u1, R u2, R dxa, R dxb) {
  R x = R(1.4) * (u2 - u1) * dxb;
  R y = R(1.4) * (u1 - u0) * dxa;
  R z = fdiv((u2 - u0) * dxa * dxb, dxa + dxb);
  R s = sgn(R(1), x);
  return s * mx2(R(0), mn3(abs_(x), s * y, s * z));
}

// fluxlimiter.f90:58-75 (modified MC).  cf, cb and the centre spacings are
// pure grid functions; they are evaluated here from the same coordinate
// operands as in the reference so they round identically.
template <class R>
__device__ __forceinline__ R modified_mc(R u0, R u1, R u2, R xm, R x0, R xp,
                                         R xim, R xi0) {
  R cf = fdiv(xp - x0, xi0 - x0);
  R cb = fdiv(x0 - xm, x0 - xim);
  R top = fdiv((u1 - u0) * (xp - x0), x0 - xm);
  R bot = u2 - u1;
  R sig = sgn(R(1), bot);
  return fdiv(sig * mx2(R(0), mn3(sig * R(0.5) * (bot + top), sig * cf * bot, sig * cb * top)),
              xp - x0);
}

// get_eint, eos.f90:846-865
template <class R>
__device__ __forceinline__ R get_eint(bool mag, R etot, R d, R v1, R v2, R v3, R b1,
                                      R b2, R b3) {
  R vsq = v1 * v1 + v2 * v2 + v3 * v3;
  R eint = etot - R(0.5) * d * vsq;
  if (mag) eint = eint - R(0.5) * (b1 * b1 + b2 * b2 + b3 * b3);
  return eint;
}

// get_cf, eos.f90:351-365
template <class R>
__device__ __forceinline__ R get_cf(R d, R cs, R b1, R b2, R b3) {
  R asq = cs * cs;
  R bb1 = fdiv(b1 * b1, d);
  R bb2 = fdiv(b2 * b2, d);
  R bb3 = fdiv(b3 * b3, d);
  R bbsq = bb1 + bb2 + bb3;
  return rsqrt_(R(0.5) * (asq + bbsq + rsqrt_((asq + bbsq) * (asq + bbsq) - R(4) * asq * bb1)));
}

// hlldflux + fstar, hlldflux.f90:19-179, without the radiation component
// (erl = err = 0 on this path; its flux only feeds irad).  f[0..8] are
// tmpflux(1..9).
template <class R>
__device__ __forceinline__ void hlld(R f[9], R cfl, R cfr, R v1l, R v1r, R v2l, R v2r,
                                     R v3l, R v3r, R dl, R dr, R el, R er, R ptl,
                                     R ptr, R b1l, R b1r, R b2l, R b2r, R b3l, R b3r,
                                     R phil, R phir, R ch) {
  const R h = R(0.5);
  R b1 = h * ((b1l + b1r) - fdiv(phir - phil, ch));
  R phi = h * ((phil + phir) - (b1r - b1l) * ch);

  R sl = mn3(v1r - cfr, v1l - cfl, R(0));
  R sr = mx3(v1r + cfr, v1l + cfl, R(0));

  R slv = sl - v1l, srv = sr - v1r;
  R den = srv * dr - slv * dl;
  R sm = fdiv(srv * dr * v1r - slv * dl * v1l - ptr + ptl, den);

  R slm = sl - sm, srm = sr - sm;
  R dla = fdiv(dl * slv, slm);
  R dra = fdiv(dr * srv, srm);

  R sqla = rsqrt_(dla), sqra = rsqrt_(dra);
  R sla = sm - fdiv(abs_(b1), sqla);
  R sra = sm + fdiv(abs_(b1), sqra);

  R pta = fdiv(srv * dr * ptl - slv * dl * ptr + dl * dr * srv * slv * (v1r - v1l), den);

  R b1b1 = b1 * b1;
  R denl = dl * slv * slm - b1b1;
  R denr = dr * srv * srm - b1b1;
  R v2la = v2l - fdiv(b1 * b2l * (sm - v1l), denl);
  R v2ra = v2r - fdiv(b1 * b2r * (sm - v1r), denr);
  R v3la = v3l - fdiv(b1 * b3l * (sm - v1l), denl);
  R v3ra = v3r - fdiv(b1 * b3r * (sm - v1r), denr);

  R numl = dl * slv * slv - b1b1;
  R numr = dr * srv * srv - b1b1;
  R b2la = fdiv(b2l * numl, denl);
  R b2ra = fdiv(b2r * numr, denr);
  R b3la = fdiv(b3l * numl, denl);
  R b3ra = fdiv(b3r * numr, denr);

  R ela = fdiv(slv * el - ptl * v1l + pta * sm +
                   b1 * ((v1l * b1 + v2l * b2l + v3l * b3l) - (sm * b1 + v2la * b2la + v3la * b3la)),
               slm);
  R era = fdiv(srv * er - ptr * v1r + pta * sm +
                   b1 * ((v1r * b1 + v2r * b2r + v3r * b3r) - (sm * b1 + v2ra * b2ra + v3ra * b3ra)),
               srm);

  R sb1 = sgn(R(1), b1);
  R sqs = sqla + sqra;
  R v2aa = fdiv(sqla * v2la + sqra * v2ra + (b2ra - b2la) * sb1, sqs);
  R v3aa = fdiv(sqla * v3la + sqra * v3ra + (b3ra - b3la) * sb1, sqs);
  R sqp = rsqrt_(dla * dra);
  R b2aa = fdiv(sqla * b2ra + sqra * b2la + sqp * (v2ra - v2la) * sb1, sqs);
  R b3aa = fdiv(sqla * b3ra + sqra * b3la + sqp * (v3ra - v3la) * sb1, sqs);

  R vbaa = v2aa * b2aa + v3aa * b3aa;
  R eraa = era + sb1 * sqra * ((v2ra * b2ra + v3ra * b3ra) - vbaa);
  R elaa = ela - sb1 * sqla * ((v2la * b2la + v3la * b3la) - vbaa);

  R signsm = sgn(h, sm);
  R sortsla = h + sgn(h, sla);
  R sortsra = h + sgn(h, -sra);
  R sortslr = h - sgn(h, sla * sra);

  R d_ = (h + signsm) * dla + (h - signsm) * dra;
  R v2_ = sortsla * v2la + sortslr * v2aa + sortsra * v2ra;
  R v3_ = sortsla * v3la + sortslr * v3aa + sortsra * v3ra;
  R b2_ = sortsla * b2la + sortslr * b2aa + sortsra * b2ra;
  R b3_ = sortsla * b3la + sortslr * b3aa + sortsra * b3ra;
  R e_ = sortsla * ela + (h - sgn(h, sla * sm)) * elaa + (h + sgn(h, -sm * sra)) * eraa +
         (h - sgn(h, sra)) * era;

  // fstar, hlldflux.f90:156-179 with (d,v1,v2,v3,b1,b2,b3,e,p) =
  // (d_, sm, v2_, v3_, b1, b2_, b3_, e_, pta)
  f[0] = d_ * sm;
  f[1] = d_ * sm * sm + pta - b1 * b1;
  f[2] = d_ * sm * v2_ - b1 * b2_;
  f[3] = d_ * sm * v3_ - b1 * b3_;
  f[4] = (e_ + pta) * sm - b1 * (sm * b1 + v2_ * b2_ + v3_ * b3_);
  f[5] = phi;
  f[6] = b2_ * sm - b1 * v2_;
  f[7] = b3_ * sm - b1 * v3_;
  f[8] = ch * ch * b1;
}

}  // namespace hg
