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

// ---- division by a shared denominator ---------------------------------------
// FP64 (SH = true): q = a / b by the same instruction sequence as the
// compiler's IEEE division -- v_div_scale of b and of a, v_rcp_f64 of the
// scaled b refined by two Newton steps, v_div_fmas, v_div_fixup -- with the
// reciprocal of b computed once and shared by every numerator.
// v_div_scale leaves b unscaled except for operands of extreme exponent (b
// denormal or beyond 2^1022, a tiny, or a/b near the overflow or denormal
// range); then the shared reciprocal is not the one a / b would use, and
// qdiv recomputes that lane's quotient with a plain division.  Where the
// scaled b is NaN (a or b zero or NaN) v_div_fixup returns the special
// result (0, inf or NaN) without reading the reciprocal.  Every quotient is
// therefore bit-identical to a / b.
struct Den64 {
  double b, r;
};
template <class R> struct DenX {  // plain division (FP64 IEEE, FP32 approximate)
  R b;
};
template <bool SH, class R> __device__ __forceinline__ auto mkden(R b) {
  if constexpr (SH) {
    double r = __builtin_amdgcn_rcp(b);
    double e = fma(-b, r, 1.0);
    r = fma(r, e, r);
    e = fma(-b, r, 1.0);
    r = fma(r, e, r);
    return Den64{b, r};
  } else {
    return DenX<R>{b};
  }
}
__device__ __forceinline__ double qdiv(const Den64 &D, double a) {
  bool flag, vcc;
  const double sb = __builtin_amdgcn_div_scale(a, D.b, false, &flag);  // scaled b
  const double n = __builtin_amdgcn_div_scale(a, D.b, true, &vcc);    // scaled a
  const double m = n * D.r;
  const double t = fma(-D.b, m, n);
  double q = __builtin_amdgcn_div_fixup(__builtin_amdgcn_div_fmas(t, D.r, m, vcc), D.b, a);
  if (__builtin_islessgreater(sb, D.b)) {
    // rare: keep this a branch (the compiler would otherwise speculate the
    // full division on every lane and select)
    double x = a;
    asm volatile("" : "+v"(x));
    q = x / D.b;
  }
  return q;
}
template <class R> __device__ __forceinline__ R qdiv(const DenX<R> &D, R a) {
  return fdiv(a, D.b);
}

enum { LIM_FLAT = 0, LIM_MINMOD = 1, LIM_MC = 2 };

// fluxlimiter.f90:40-56 (generalised minmod); dx = idx(i:i+1)
// Dab: the denominator dxa + dxb (a grid function, shared by every variable
// of the cell).
template <class R, class DN>
__device__ __forceinline__ R minmod_d(R u0, R u1, R u2, R dxa, R dxb, const DN &Dab) {
  R x = R(1.4) * (u2 - u1) * dxb;
  R y = R(1.4) * (u1 - u0) * dxa;
  R z = qdiv(Dab, (u2 - u0) * dxa * dxb);
  R s = sgn(R(1), x);
  return s * mx2(R(0), mn3(abs_(x), s * y, s * z));
}
template <class R>
__device__ __forceinline__ R minmod(R u0, R u1, R u2, R dxa, R dxb) {
  return minmod_d(u0, u1, u2, dxa, dxb, mkden<false>(dxa + dxb));
}

// fluxlimiter.f90:58-75 (modified MC).  cf, cb and the centre spacings are
// pure grid functions; they are evaluated here from the same coordinate
// operands as in the reference so they round identically.
//
// modified_mc_d takes the per-cell grid factors: dxp = xp - x0,
// cf = (xp - x0)/(xi0 - x0), cb = (x0 - xm)/(x0 - xim) and the shared
// denominators Dm = x0 - xm and Dp = xp - x0.
template <class R, class DN>
__device__ __forceinline__ R modified_mc_d(R u0, R u1, R u2, R dxp, R cf, R cb, const DN &Dm,
                                           const DN &Dp) {
  R top = qdiv(Dm, (u1 - u0) * dxp);
  R bot = u2 - u1;
  R sig = sgn(R(1), bot);
  return qdiv(Dp, sig * mx2(R(0), mn3(sig * R(0.5) * (bot + top), sig * cf * bot, sig * cb * top)));
}
template <class R>
__device__ __forceinline__ R modified_mc(R u0, R u1, R u2, R xm, R x0, R xp,
                                         R xim, R xi0) {
  R cf = fdiv(xp - x0, xi0 - x0);
  R cb = fdiv(x0 - xm, x0 - xim);
  return modified_mc_d(u0, u1, u2, xp - x0, cf, cb, mkden<false>(x0 - xm), mkden<false>(xp - x0));
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
// get_cf_d takes the density as a shared denominator.
template <class R, class DN>
__device__ __forceinline__ R get_cf_d(const DN &Dd, R cs, R b1, R b2, R b3) {
  R asq = cs * cs;
  R bb1 = qdiv(Dd, b1 * b1);
  R bb2 = qdiv(Dd, b2 * b2);
  R bb3 = qdiv(Dd, b3 * b3);
  R bbsq = bb1 + bb2 + bb3;
  return rsqrt_(R(0.5) * (asq + bbsq + rsqrt_((asq + bbsq) * (asq + bbsq) - R(4) * asq * bb1)));
}
template <class R>
__device__ __forceinline__ R get_cf(R d, R cs, R b1, R b2, R b3) {
  return get_cf_d(mkden<false>(d), cs, b1, b2, b3);
}

// hlldflux + fstar, hlldflux.f90:19-179, without the radiation component
// (erl = err = 0 on this path; its flux only feeds irad).  f[0..8] are
// tmpflux(1..9).  SH: FP64 divisions by shared denominators (qdiv).
template <bool SH, class R>
__device__ __forceinline__ void hlld_d(R f[9], R cfl, R cfr, R v1l, R v1r, R v2l, R v2r,
                                       R v3l, R v3r, R dl, R dr, R el, R er, R ptl,
                                       R ptr, R b1l, R b1r, R b2l, R b2r, R b3l, R b3r,
                                       R phil, R phir, R ch) {
  const R h = R(0.5);
  R b1 = h * ((b1l + b1r) - fdiv(phir - phil, ch));
  R phi = h * ((phil + phir) - (b1r - b1l) * ch);

  R sl = mn3(v1r - cfr, v1l - cfl, R(0));
  R sr = mx3(v1r + cfr, v1l + cfl, R(0));

  R slv = sl - v1l, srv = sr - v1r;
  const auto Dden = mkden<SH>(srv * dr - slv * dl);
  R sm = qdiv(Dden, srv * dr * v1r - slv * dl * v1l - ptr + ptl);
  R pta = qdiv(Dden, srv * dr * ptl - slv * dl * ptr + dl * dr * srv * slv * (v1r - v1l));

  R slm = sl - sm, srm = sr - sm;
  const auto Dslm = mkden<SH>(slm), Dsrm = mkden<SH>(srm);
  R dla = qdiv(Dslm, dl * slv);
  R dra = qdiv(Dsrm, dr * srv);

  R sqla = rsqrt_(dla), sqra = rsqrt_(dra);
  R sla = sm - fdiv(abs_(b1), sqla);
  R sra = sm + fdiv(abs_(b1), sqra);

  R b1b1 = b1 * b1;
  const auto Dl = mkden<SH>(dl * slv * slm - b1b1);
  const auto Dr = mkden<SH>(dr * srv * srm - b1b1);
  R v2la = v2l - qdiv(Dl, b1 * b2l * (sm - v1l));
  R v2ra = v2r - qdiv(Dr, b1 * b2r * (sm - v1r));
  R v3la = v3l - qdiv(Dl, b1 * b3l * (sm - v1l));
  R v3ra = v3r - qdiv(Dr, b1 * b3r * (sm - v1r));

  R numl = dl * slv * slv - b1b1;
  R numr = dr * srv * srv - b1b1;
  R b2la = qdiv(Dl, b2l * numl);
  R b2ra = qdiv(Dr, b2r * numr);
  R b3la = qdiv(Dl, b3l * numl);
  R b3ra = qdiv(Dr, b3r * numr);

  R ela = qdiv(Dslm, slv * el - ptl * v1l + pta * sm +
                         b1 * ((v1l * b1 + v2l * b2l + v3l * b3l) -
                               (sm * b1 + v2la * b2la + v3la * b3la)));
  R era = qdiv(Dsrm, srv * er - ptr * v1r + pta * sm +
                         b1 * ((v1r * b1 + v2r * b2r + v3r * b3r) -
                               (sm * b1 + v2ra * b2ra + v3ra * b3ra)));

  R sb1 = sgn(R(1), b1);
  const auto Dsqs = mkden<SH>(sqla + sqra);
  R v2aa = qdiv(Dsqs, sqla * v2la + sqra * v2ra + (b2ra - b2la) * sb1);
  R v3aa = qdiv(Dsqs, sqla * v3la + sqra * v3ra + (b3ra - b3la) * sb1);
  R sqp = rsqrt_(dla * dra);
  R b2aa = qdiv(Dsqs, sqla * b2ra + sqra * b2la + sqp * (v2ra - v2la) * sb1);
  R b3aa = qdiv(Dsqs, sqla * b3ra + sqra * b3la + sqp * (v3ra - v3la) * sb1);

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
template <class R>
__device__ __forceinline__ void hlld(R f[9], R cfl, R cfr, R v1l, R v1r, R v2l, R v2r,
                                     R v3l, R v3r, R dl, R dr, R el, R er, R ptl,
                                     R ptr, R b1l, R b1r, R b2l, R b2r, R b3l, R b3r,
                                     R phil, R phir, R ch) {
  hlld_d<false>(f, cfl, cfr, v1l, v1r, v2l, v2r, v3l, v3r, dl, dr, el, er, ptl, ptr, b1l, b1r,
                b2l, b2r, b3l, b3r, phil, phir, ch);
}

// ---- FP32-mixed HLLD in perturbation form ----------------------------------
// The Riemann solver of hlld() above (hlldflux.f90:19-179, same star states,
// same state selection), rearranged so that FP32 only ever carries
// differences.  Every star state is written as the left face state plus a
// correction, and the returned flux is the FP64 flux function of the left
// state (fstar with the GLM normal field b1 and the left total pressure)
// plus an FP32 correction.  In exact arithmetic this equals hlld(); in
// rounding, the FP32 error scales with the jump across the face instead of
// with the flux itself, so smooth and near-static flows (a hydrostatic star,
// a weak shear layer) stay close to the FP64 result while the wave speeds,
// divisions and square roots run at FP32 cost.
//
// l, r: FP64 face primitives d, v1 (normal), v2, v3, b2, b3, e (total
// energy), pt (total pressure).  b1: GLM normal field (FP64).  cfl, cfr:
// fast speeds.  Writes f[0..4], f[6], f[7]; the GLM components f[5], f[8]
// are the caller's.
__device__ __forceinline__ void hlld_delta(double f[9], const double l[8], const double r[8],
                                           double b1, float cfl, float cfr) {
  const float h = 0.5f;
  const float dl = (float)l[0], dr = (float)r[0];
  const float v1l = (float)l[1], v1r = (float)r[1];
  const float v2l = (float)l[2], v2r = (float)r[2], v3l = (float)l[3], v3r = (float)r[3];
  const float b2l = (float)l[4], b2r = (float)r[4], b3l = (float)l[5], b3r = (float)r[5];
  const float b1f = (float)b1;
  // jumps across the face, differenced in FP64
  const float Dd = (float)(r[0] - l[0]), Dv1 = (float)(r[1] - l[1]);
  const float Dv2 = (float)(r[2] - l[2]), Dv3 = (float)(r[3] - l[3]);
  const float Db2 = (float)(r[4] - l[4]), Db3 = (float)(r[5] - l[5]);
  const float De = (float)(r[6] - l[6]), Dpt = (float)(r[7] - l[7]);
  const float ehl = (float)(l[6] + l[7]), ehr = (float)(r[6] + r[7]);

  const float sl = mn3(v1r - cfr, v1l - cfl, 0.0f);
  const float sr = mx3(v1r + cfr, v1l + cfl, 0.0f);
  const float slv = sl - v1l, srv = sr - v1r;
  const float den = srv * dr - slv * dl;
  const float dv = fdiv(srv * dr * Dv1 - Dpt, den);  // sm - v1l
  const float dvr = dv - Dv1;                         // sm - v1r
  const float sm = v1l + dv;
  const float slm = slv - dv, srm = srv - dvr;        // sl - sm, sr - sm
  const float dp = dl * slv * dv;                     // pta - ptl
  const float dpr = dr * srv * dvr;                   // pta - ptr
  const float ddla = fdiv(dl * dv, slm);              // dla - dl
  const float ddra = fdiv(dr * dvr, srm);             // dra - dr
  const float dla = dl + ddla, dra = dr + ddra;
  const float sqla = rsqrt_(dla), sqra = rsqrt_(dra);
  const float ab1 = abs_(b1f);
  const float sla = sm - fdiv(ab1, sqla), sra = sm + fdiv(ab1, sqra);

  const float b1b1 = b1f * b1f;
  const float idenl = fdiv(1.0f, dl * slv * slm - b1b1);
  const float idenr = fdiv(1.0f, dr * srv * srm - b1b1);
  // single-star transverse states relative to their own side
  const float dv2l = -b1f * b2l * dv * idenl, dv3l = -b1f * b3l * dv * idenl;
  const float dv2r = -b1f * b2r * dvr * idenr, dv3r = -b1f * b3r * dvr * idenr;
  const float gl = dl * slv * dv * idenl, gr = dr * srv * dvr * idenr;
  const float db2l = b2l * gl, db3l = b3l * gl, db2r = b2r * gr, db3r = b3r * gr;
  const float b2la = b2l + db2l, b3la = b3l + db3l, b2ra = b2r + db2r, b3ra = b3r + db3r;
  // (v.B) of the side state minus that of its star state
  const float xl = -b1f * dv - (v2l * db2l + dv2l * b2la) - (v3l * db3l + dv3l * b3la);
  const float xr = -b1f * dvr - (v2r * db2r + dv2r * b2ra) - (v3r * db3r + dv3r * b3ra);
  const float dela = fdiv(ehl * dv + dp * sm + b1f * xl, slm);   // ela - el
  const float dera = fdiv(ehr * dvr + dpr * sm + b1f * xr, srm);  // era - er

  // double-star (Alfven) states, relative to the left face state
  const float sb1 = sgn(1.0f, b1f);
  const float isqs = fdiv(1.0f, sqla + sqra);
  const float sqp = rsqrt_(dla * dra);
  const float dv2ra = Dv2 + dv2r, dv3ra = Dv3 + dv3r;  // v2ra - v2l
  const float db2ra = Db2 + db2r, db3ra = Db3 + db3r;  // b2ra - b2l
  const float dv2aa = (sqla * dv2l + sqra * dv2ra + (db2ra - db2l) * sb1) * isqs;
  const float dv3aa = (sqla * dv3l + sqra * dv3ra + (db3ra - db3l) * sb1) * isqs;
  const float db2aa = (sqla * db2ra + sqra * db2l + sqp * (dv2ra - dv2l) * sb1) * isqs;
  const float db3aa = (sqla * db3ra + sqra * db3l + sqp * (dv3ra - dv3l) * sb1) * isqs;
  const float b2aa = b2l + db2aa, b3aa = b3l + db3aa;
  // (v2 b2 + v3 b3) of a state minus that of the Alfven state
  const float pla = v2l * (db2l - db2aa) + dv2l * b2la - dv2aa * b2aa +
                    v3l * (db3l - db3aa) + dv3l * b3la - dv3aa * b3aa;
  const float pra = v2l * (db2ra - db2aa) + dv2ra * b2ra - dv2aa * b2aa +
                    v3l * (db3ra - db3aa) + dv3ra * b3ra - dv3aa * b3aa;
  const float delaa = dela - sb1 * sqla * pla;              // elaa - el
  const float deraa = (De + dera) + sb1 * sqra * pra;       // eraa - el

  // state selection (hlldflux.f90:140-154), as corrections to the left state
  const float signsm = sgn(h, sm);
  const float wla = h + sgn(h, sla), wra = h + sgn(h, -sra), waa = h - sgn(h, sla * sra);
  const float dd_ = (h + signsm) * ddla + (h - signsm) * (Dd + ddra);
  const float dv2_ = wla * dv2l + waa * dv2aa + wra * dv2ra;
  const float dv3_ = wla * dv3l + waa * dv3aa + wra * dv3ra;
  const float db2_ = wla * db2l + waa * db2aa + wra * db2ra;
  const float db3_ = wla * db3l + waa * db3aa + wra * db3ra;
  const float de_ = wla * dela + (h - sgn(h, sla * sm)) * delaa +
                    (h + sgn(h, -sm * sra)) * deraa + (h - sgn(h, sra)) * (De + dera);
  const float v2_ = v2l + dv2_, v3_ = v3l + dv3_, b2_ = b2l + db2_, b3_ = b3l + db3_;

  // fstar(selected) - fstar(left), hlldflux.f90:156-179
  const float dlv1l = (float)(l[0] * l[1]);
  const float df0 = dd_ * sm + dl * dv;
  const float df1 = df0 * sm + dlv1l * dv + dp;
  const float df2 = df0 * v2_ + dlv1l * dv2_ - b1f * db2_;
  const float df3 = df0 * v3_ + dlv1l * dv3_ - b1f * db3_;
  const float dvb = b1f * dv + (dv2_ * b2_ + v2l * db2_) + (dv3_ * b3_ + v3l * db3_);
  const float df4 = (de_ + dp) * sm + ehl * dv - b1f * dvb;
  const float df6 = db2_ * sm + b2l * dv - b1f * dv2_;
  const float df7 = db3_ * sm + b3l * dv - b1f * dv3_;

  // FP64 flux function of the left state
  const double u = l[1], du = l[0] * l[1];
  f[0] = du + (double)df0;
  f[1] = du * u + l[7] - b1 * b1 + (double)df1;
  f[2] = du * l[2] - b1 * l[4] + (double)df2;
  f[3] = du * l[3] - b1 * l[5] + (double)df3;
  f[4] = (l[6] + l[7]) * u - b1 * (u * b1 + l[2] * l[4] + l[3] * l[5]) + (double)df4;
  f[6] = l[4] * u - b1 * l[2] + (double)df6;
  f[7] = l[5] * u - b1 * l[3] + (double)df7;
}

}  // namespace hg
