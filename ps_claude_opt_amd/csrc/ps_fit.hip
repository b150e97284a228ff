// MI355X (gfx950) kernels of the period scan: one persistent workgroup per
// (trial frequency, initial pole) pair runs that pair's whole
// Levenberg-Marquardt fit (the reference spreads one LM iteration over
// ~13 + 4 * curves kernel launches, start_CUDA.cu:1017-1175).
//
// Per mrqcof evaluation (reference Start.cu:2262-2828):
//   curv          facet areas Area_f = Darea_f exp(sum P_l^m (c cos + s sin))
//   points        per-point record (rotation angle, pole-frame Earth and Sun
//                 directions, phase-function terms) and a counting sort of
//                 the points by view-direction cell
//   phase 1       over 16-point tiles of the sorted points, claimed by the
//                 waves one at a time; per 4-facet step lane l owns point
//                 (l & 15) and facet (4k + (l >> 4)); it evaluates the
//                 scattering law and the derivative sums on the VALU and
//                 feeds the facet weight w = Area * S as the A operand of
//                 v_mfma_f64_16x16x4f64 against the basis table (B operand,
//                 read through L2), which yields the shape Jacobian of 16
//                 points x 16 coefficients per block.  Steps in which no
//                 point of the tile sees any of the 4 facets are skipped;
//                 sorting by view direction makes such steps common.  Raw
//                 Jacobian rows go to the workgroup's scratch slab.
//   curve sums    brightness sum L of each relative curve (from the model
//                 brightness alone) and its factor c = sig lp / L
//   phase 2       one read of the rows: the augmented product
//                 [cJ|r]^T W [cJ|r] (r = obs - c y) is an FP64 MFMA SYRK
//                 split into tile chains, each wave owning lower 16x16
//                 block pairs of one chain in registers; per relative curve
//                 the chain also sums d = sum J, u = sum w y J, sum w y^2
//                 and sum w y r
//   correction    the relative renormalisation J' = c (J - y d / L)
//                 (Start.cu:2656-2715) applied algebraically: a rank-2 term
//                 per curve, as one MFMA chain per block pair on the
//                 combined matrix.  Alpha, beta and chi^2 come out at once.
// mrqmin: full-pivot Gauss-Jordan in LDS by all waves, accept / reject,
// Iter2, FinishPole (Start.cu:2838-3021, 300-392, 5129-5248, 5056-5125).
//
// Operation order is fixed (see the CPU oracle): facet partial sums by
// facet mod 4 combined (P0+P1)+(P2+P3); the MFMA gather is a fused
// multiply-add chain over facets in ascending order; tile brightness sums
// are four partials over points p mod 4 combined (G0+G1)+(G2+G3), curve sums
// add the tile sums in order; a chain's per-curve sums are per-lane chains
// over its tiles of the curve combined (S0+S1)+(S2+S3), the chains as the
// matrices; each element of the augmented normal matrix is
// the combination (C0+C1)+(C2+C3) (or C0+C1 above 64 columns) of fused
// multiply-add chains over the tiles T = g (mod chains), points in order,
// followed by one fused multiply-add chain of the renormalisation
// correction over the curves in order.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "ps_math.h"
#include "ps_kernels.h"

// This file is compiled once per workgroup size: PS_FIT_WAVES wavefronts per
// workgroup, everything in namespace PS_FIT_NS.  The variants give
// bit-identical results (the point chains of the normal equations do not
// depend on the wave count) and differ only in speed.
#ifndef PS_FIT_WAVES
#define PS_FIT_WAVES 4
#endif
#ifndef PS_FIT_NS
#define PS_FIT_NS psw4
#endif

typedef double v4d __attribute__((ext_vector_type(4)));

namespace PS_FIT_NS {

#ifdef PS_PHASE_TIMING
// Debug builds only: cycles spent per phase by thread 0 of every workgroup
// (including its waits at the phase's closing barrier).
__device__ unsigned long long g_phase_cycles[16];
#define PS_TSTART() unsigned long long ps_t0_ = clock64()
#define PS_TMARK(i)                                                  \
  do {                                                               \
    if (threadIdx.x == 0) {                                          \
      const unsigned long long t1_ = clock64();                      \
      atomicAdd(&g_phase_cycles[i], t1_ - ps_t0_);                   \
      ps_t0_ = t1_;                                                  \
    } else {                                                         \
      ps_t0_ = clock64();                                            \
    }                                                                \
  } while (0)
#else
#define PS_TSTART() \
  do {              \
  } while (0)
#define PS_TMARK(i) \
  do {              \
  } while (0)
#endif

namespace {

constexpr int NW = PS_FIT_WAVES;  // waves per workgroup
constexpr int WG = 64 * NW;
constexpr double kPi = 3.14159265358979323846;
constexpr double kTiny = 1e-8;  // constants.h:19

__device__ __forceinline__ v4d mfma(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// 1/x: hardware reciprocal refined by two Newton steps (agrees with the
// correctly rounded quotient; PS_IEEE_DIV selects the IEEE division).
__device__ __forceinline__ double recip(double x) {
#ifdef PS_IEEE_DIV
  return 1.0 / x;
#else
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
#endif
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lane and wave index, opaque to the optimiser: values derived from them
// are computed where each phase needs them instead of being hoisted out of
// the persistent pair loop (where they would stay live across every phase).
__device__ __forceinline__ int lane_id() {
  int l = threadIdx.x & 63;
  asm volatile("" : "+v"(l));
  return l;
}

__device__ __forceinline__ int tid_opaque() {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

__device__ __forceinline__ double xsum16_32(double v) {
  v = v + __shfl_xor(v, 16);
  v = v + __shfl_xor(v, 32);
  return v;
}

struct Inv {
  double omega, A, invD, kph, invD2, cl, cls, sb, cb, sl, clam;
};

__device__ __forceinline__ void store_inv(const Inv& v, double* d) {
  d[0] = v.omega;
  d[1] = v.A;
  d[2] = v.invD;
  d[3] = v.kph;
  d[4] = v.invD2;
  d[5] = v.cl;
  d[6] = v.cls;
  d[7] = v.sb;
  d[8] = v.cb;
  d[9] = v.sl;
  d[10] = v.clam;
}

__device__ __forceinline__ Inv load_inv(const double* d) {
  Inv v;
  v.omega = d[0];
  v.A = d[1];
  v.invD = d[2];
  v.kph = d[3];
  v.invD2 = d[4];
  v.cl = d[5];
  v.cls = d[6];
  v.sb = d[7];
  v.cb = d[8];
  v.sl = d[9];
  v.clam = d[10];
  return v;
}

// Pair invariants of one parameter vector a[] (a[i] = parameter i+1).
__device__ __forceinline__ Inv make_inv(const PsDevProblem& P, const double* a) {
  const int C = P.ncoef;
  Inv v;
  v.omega = a[C + 2];
  v.A = a[C + 3];
  v.invD = 1.0 / a[C + 4];
  v.kph = a[C + 5];
  v.invD2 = v.invD * v.invD;
  v.cl = psm::exp(a[C + 6]);
  v.cls = a[C + 7];
  psm::sincos(a[C], &v.sb, &v.cb);
  psm::sincos(a[C + 1], &v.sl, &v.clam);
  return v;
}

// Body-frame direction g = Rz(phi) P(beta') Rz(lambda) e and its
// derivatives with respect to beta', lambda and omega.
// An ecliptic direction e (Earth or Sun) in the pole frame: u = e turned
// by lambda about z, then v = u turned by beta about y (v_y = u_y).
struct Frame {
  double ux, uy, vx, vz;
};

__device__ __forceinline__ Frame pole_frame(const Inv& v, double ex, double ey, double ez) {
  Frame r;
  r.ux = fma(v.sl, ey, v.clam * ex);
  r.uy = fma(-v.sl, ex, v.clam * ey);
  r.vx = fma(-v.sb, ez, v.cb * r.ux);
  r.vz = fma(v.sb, r.ux, v.cb * ez);
  return r;
}

// Body-frame direction g at rotation angle (cos cf, sin sf).
__device__ __forceinline__ void body_dir(double cf, double sf, double vx, double uy, double vz, double* g) {
  g[0] = fma(sf, uy, cf * vx);
  g[1] = fma(-sf, vx, cf * uy);
  g[2] = vz;
}

// Derivatives of g with respect to beta, lambda and the rotation rate.
__device__ __forceinline__ void body_derivs(const Inv& v, double cf, double sf, double t, const Frame& r,
                                            const double* g, double* db, double* dl, double* dw) {
  db[0] = -(cf * r.vz);
  db[1] = sf * r.vz;
  db[2] = r.vx;
  const double w = v.cb * r.uy;
  dl[0] = fma(cf, w, -(sf * r.ux));
  dl[1] = fma(-sf, w, -(cf * r.ux));
  dl[2] = v.sb * r.uy;
  dw[0] = t * g[1];
  dw[1] = -(t * g[0]);
  dw[2] = 0.0;
}

// Fields of the per-point record (PS_REC doubles, one 128-byte line): the
// rotation angle, the pole-frame Earth and Sun directions, the epoch and the
// phase-function terms of one evaluation.
enum : int {
  RC_SF = 0, RC_CF, RC_VX, RC_UY, RC_VZ, RC_UX, RC_VX0, RC_UY0, RC_VZ0, RC_UX0,
  RC_TT, RC_SCALE, RC_FF, RC_D2, RC_ALPH
};
static_assert(RC_ALPH < PS_REC, "point record overflows its line");

// View-direction cell of a point: the cube-map cell (6 faces x 6 x 6,
// rows traversed alternately forwards and backwards) of g + g0.  Points of
// one cell see nearly the same facets.  Only the speed of the facet pass
// depends on it, so it is computed in single precision.
constexpr int kCellSide = 6;
constexpr int kCells = 6 * kCellSide * kCellSide;
static_assert(kCells <= 256, "cells are stored as bytes");

__device__ __forceinline__ int view_cell(const double* g, const double* g0) {
  const float h0 = (float)(g[0] + g0[0]), h1 = (float)(g[1] + g0[1]), h2 = (float)(g[2] + g0[2]);
  const float a0 = fabsf(h0), a1 = fabsf(h1), a2 = fabsf(h2);
  int ax;
  float hm, u, w;
  if (a0 >= a1 && a0 >= a2) {
    ax = 0, hm = h0, u = h1, w = h2;
  } else if (a1 >= a2) {
    ax = 1, hm = h1, u = h2, w = h0;
  } else {
    ax = 2, hm = h2, u = h0, w = h1;
  }
  const float r = 0.5f * kCellSide * __builtin_amdgcn_rcpf(fmaxf(fabsf(hm), 1e-30f));
  const float hi = kCellSide - 1;
  int iu = (int)fminf(fmaxf(fmaf(u, r, 0.5f * kCellSide), 0.0f), hi);
  const int iw = (int)fminf(fmaxf(fmaf(w, r, 0.5f * kCellSide), 0.0f), hi);
  if (iw & 1) iu = kCellSide - 1 - iu;
  return ((2 * ax + (hm > 0.0f ? 1 : 0)) * kCellSide + iw) * kCellSide + iu;
}

// LDS layout of one workgroup (dynamic shared memory).
struct Lds {
  double* work;    // [K*K] normal matrix / augmented Gauss-Jordan matrix
  double* nrm;     // [F + PS_FACET_PAD][4] facet normal x, y, z and the current facet area (padding: 0)
  const double* ytab;  // [F][16][nbs] shape basis (global table, read through L2)
  double* cgg;     // [ma]
  double* atry;    // [ma]
  double* rowbuf;  // [NW][16][16] per-point non-shape values of phase 1 (aliases work)
  double* red;     // [2*NW] reduction scratch
  double* inv;     // [16] pair invariants of the evaluated parameters
};

// Pointers into one workgroup's scratch slab.
struct Slab {
  double *jbuf, *ybuf, *tsy, *crv, *csum, *ttab, *part, *best, *prec;
  int* perm;
  unsigned char* pkey;
};

__device__ __forceinline__ Slab make_slab(const PsDevProblem& P, double* base) {
  const PsSlab o = ps_slab(P);
  Slab s;
  s.jbuf = base + o.jbuf;
  s.ybuf = base + o.ybuf;
  s.tsy = base + o.tsy;
  s.crv = base + o.crv;
  s.csum = base + o.csum;
  s.ttab = base + o.ttab;
  s.part = base + o.part;
  s.best = base + o.best;
  s.prec = base + o.prec;
  s.perm = reinterpret_cast<int*>(base + o.perm);
  s.pkey = reinterpret_cast<unsigned char*>(base + o.pkey);
  return s;
}

__device__ __forceinline__ Lds carve_lds(double* base, const PsDevProblem& P) {
  Lds L;
  const int K = P.mfit + 1;
  double* p = base;
  // the tile rows of phase 1 live in the normal-matrix area: during the
  // facet pass the matrix is dead (the Gauss-Jordan solution has been
  // consumed; the evaluation writes the new matrix after its last barrier)
  L.work = p;
  L.rowbuf = p;
  p += K * K > NW * 256 ? K * K : NW * 256;
  L.nrm = p;
  p += 4 * (P.nfac + PS_FACET_PAD);
  L.ytab = P.yl;
  L.cgg = p;
  p += P.ma;
  L.atry = p;
  p += P.ma;
  L.red = p;
  p += 2 * NW;
  L.inv = p;
  return L;
}

// Copies the facet normals into the workgroup's LDS once per launch.
__device__ __forceinline__ void stage_tables(const PsDevProblem& P, const Lds& L) {
  const int F = P.nfac;
  for (int f = threadIdx.x; f < F + PS_FACET_PAD; f += WG) {
    const bool in = f < F;
    L.nrm[4 * f + 0] = in ? P.nx[f] : 0.0;
    L.nrm[4 * f + 1] = in ? P.ny[f] : 0.0;
    L.nrm[4 * f + 2] = in ? P.nz[f] : 0.0;
    L.nrm[4 * f + 3] = 0.0;
  }
  __syncthreads();
}


// Facet areas of parameter vector a (Start.cu:421-528) into nrm[4f+3].
__device__ __forceinline__ void curv_block(const PsDevProblem& P, const double* a, double* nrm) {
  const int F = P.nfac, Lm = P.lmax, M = P.mmax;
  for (int f = tid_opaque(); f < F; f += WG) {
    double g = 0.0;
    int n = 0;
    for (int l = 0; l <= Lm; ++l) {
      const double fsum = a[n++];
      g = fma(P.pleg[l * F + f], fsum, g);
    }
    for (int m = 1; m <= M; ++m) {
      const double fcm = P.fc[m * F + f], fsm = P.fs[m * F + f];
      for (int l = m; l <= Lm; ++l) {
        double fsum = a[n] * fcm;
        fsum = fma(a[n + 1], fsm, fsum);
        n += 2;
        g = fma(P.pleg[(m * (Lm + 1) + l) * F + f], fsum, g);
      }
    }
    nrm[4 * f + 3] = P.darea[f] * psm::exp(g);
  }
}

// Rotation angle of point pt and its phase-function terms (Start.cu:2095-2195).
struct PointScalars {
  double sf, cf, tt, scale, ff, d2, alph;
};

__device__ __forceinline__ void rotation_angle(const PsDevProblem& P, const Inv& v, double tt, double* sf, double* cf) {
  double f = fma(v.omega, tt, P.phi0);
  f = fma(-(2.0 * kPi), round(f * (1.0 / (2.0 * kPi))), f);
  psm::sincos(f, sf, cf);
}

__device__ __forceinline__ PointScalars point_scalars(const PsDevProblem& P, const Inv& v, double tt, double alph) {
  PointScalars q;
  q.tt = tt;
  rotation_angle(P, v, tt, &q.sf, &q.cf);
  q.alph = alph;
  q.ff = psm::exp2(-1.44269504088896 * (q.alph * v.invD));
  q.scale = fma(v.kph, q.alph, fma(v.A, q.ff, 1.0));
  q.d2 = ((v.A * q.ff) * q.alph) * v.invD2;
  return q;
}

// Phase 1 for facet-pass tile t, the points perm[16t .. 16t+15] of the
// evaluation's view-direction order, executed by one wave.  A point's raw
// Jacobian row depends on that point alone, so the grouping changes no
// result; points that see the same facets share the facet steps with any
// visible facet, and the steps without one are skipped.  The rows go to the
// slab at each point's own index.
template <int NB, int NBS>
__device__ __forceinline__ unsigned phase1_tile(const PsDevProblem& P, const Lds& L, const double* prec, const int* perm,
                                                int t, double* jbuf, double* ybuf, double* rowbuf) {
  const int lane = lane_id(), ps = lane & 15, sl = lane >> 4;
  const int F = P.nfac, QS = P.qs, C = P.ncoef, N = P.ndata;
  const int pos = t * 16 + ps;
  const bool valid = pos < N;
  const int pt = valid ? perm[pos] : 0;
  (void)NB;
  // body-frame directions from the point's record (records are stored in
  // the sorted order; the directions are the only point values live in the
  // facet loop); padding lanes see nothing
  double g[3] = {0.0, 0.0, 0.0}, g0[3] = {0.0, 0.0, 0.0};
  const double cl = L.inv[5], cls = L.inv[6];
  if (valid) {
    const double* o = prec + (size_t)pos * PS_REC;
    const double sf = o[RC_SF], cf = o[RC_CF];
    body_dir(cf, sf, o[RC_VX], o[RC_UY], o[RC_VZ], g);
    body_dir(cf, sf, o[RC_VX0], o[RC_UY0], o[RC_VZ0], g0);
  }
  double t4 = 0.0, t5 = 0.0;
  double U0 = 0.0, U1 = 0.0, U2 = 0.0, V0 = 0.0, V1 = 0.0, V2 = 0.0;
  v4d gat[NBS];
#pragma unroll
  for (int b = 0; b < NBS; ++b) gat[b] = v4d{0.0, 0.0, 0.0, 0.0};
  // lane facet f = 4k + sl: normal/area at nrm[4f..4f+3], basis values of
  // shape columns 16b + ps at ytab[(16f + ps) * NBS + b].  Software
  // pipeline over two register sets (A: even steps, B: odd steps; the
  // triangulation's step count F / 4 = 2 nrows^2 is even): the normals of a
  // step are read from LDS one step ahead, its basis values from L2 as soon
  // as the step two before has issued its MFMAs.  Two sets keep the
  // pipeline free of register-rotation moves.
  const int NK = F / 4;
  const double* __restrict__ nr = L.nrm + 4 * sl;
  const double* __restrict__ ytab = L.ytab;  // uniform base; step k's rows at ytab + 64 NBS k
  const int ly = (16 * sl + ps) * NBS;       // this lane's offset in a step's rows
  unsigned nvis = 0;  // this lane's visible (point, facet) pairs
  auto step = [&](double n0, double n1, double n2, double ar, const double* y) {
    const double mu = psm::dot3(g[0], g[1], g[2], n0, n1, n2);
    const double mu0 = psm::dot3(g0[0], g0[1], g0[2], n0, n1, n2);
    const bool vis = (mu > kTiny) && (mu0 > kTiny);
    const unsigned long long vmask = __builtin_amdgcn_ballot_w64(vis);
    if (vmask != 0) {  // a whole step of invisible facets adds exact zeros
      double w = 0.0;
      if (vis) {
        ++nvis;
        // S = mu mu0 (cls rd + cl), rd = 1 / (mu + mu0); dS/dmu =
        // mu0 (c2 mu0 + cl) and dS/dmu0 = mu (c2 mu + cl), c2 = cls rd^2
        // (Start.cu:947-974, regrouped).  The brightness, sum ar S =
        // cls t5 + cl t4, is formed per point.
        const double rd = recip(mu + mu0);
        const double am = ar * (mu * mu0);
        w = am * fma(cls, rd, cl);
        const double c2 = cls * (rd * rd);
        const double ad = ar * (mu0 * fma(c2, mu0, cl));
        const double ad0 = ar * (mu * fma(c2, mu, cl));
        U0 = fma(ad, n0, U0);
        U1 = fma(ad, n1, U1);
        U2 = fma(ad, n2, U2);
        V0 = fma(ad0, n0, V0);
        V1 = fma(ad0, n1, V1);
        V2 = fma(ad0, n2, V2);
        t4 = t4 + am;
        t5 = fma(am, rd, t5);
      }
#pragma unroll
      for (int b = 0; b < NBS; ++b) gat[b] = mfma(w, y[b], gat[b]);
    }
  };
  // the prefetches of the last steps read up to three steps past the end:
  // the zero padding of both tables (PS_FACET_PAD facets)
  double a0 = nr[0], a1 = nr[1], a2 = nr[2], aar = nr[3];
  double yA[NBS], yB[NBS];
#pragma unroll
  for (int b = 0; b < NBS; ++b) {
    yA[b] = ytab[ly + b];
    yB[b] = ytab[64 * NBS + ly + b];
  }
  // Every load issued so far completes here, once per tile, so that the
  // loop's own waits count only the prefetches in flight (otherwise the
  // wait before the even step's MFMAs drains the odd step's prefetch too).
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
  for (int k = 0; k < NK; k += 2) {
    const double* nb = nr + 16 * (k + 1);
    const double b0 = nb[0], b1 = nb[1], b2 = nb[2], bar = nb[3];
    step(a0, a1, a2, aar, yA);
    const double* y2 = ytab + (size_t)(k + 2) * 64 * NBS;
#pragma unroll
    for (int b = 0; b < NBS; ++b) yA[b] = y2[ly + b];
    const double* na = nr + 16 * (k + 2);
    a0 = na[0];
    a1 = na[1];
    a2 = na[2];
    aar = na[3];
    if (k + 1 < NK) step(b0, b1, b2, bar, yB);  // odd step count: F = 4 (mod 8)
    const double* y3 = ytab + (size_t)(k + 3) * 64 * NBS;
#pragma unroll
    for (int b = 0; b < NBS; ++b) yB[b] = y3[ly + b];
  }
  t4 = xsum16_32(t4);
  t5 = xsum16_32(t5);
  const double br = fma(cls, t5, cl * t4);
  U0 = xsum16_32(U0);
  U1 = xsum16_32(U1);
  U2 = xsum16_32(U2);
  V0 = xsum16_32(V0);
  V1 = xsum16_32(V1);
  V2 = xsum16_32(V2);
  if (sl == 0) {
    double db[3], dl[3], dw[3], db0[3], dl0[3], dw0[3], gg[3];
    // reloaded through an index the optimiser cannot see through, so the
    // loads are not merged with the ones before the facet loop
    const int z = lane >> 6;
    const Inv v = load_inv(L.inv + z);
    double sf = 0.0, cf = 0.0, tt = 0.0, scale = 0.0, ff = 0.0, d2 = 0.0, alph = 0.0;
    Frame fe{0.0, 0.0, 0.0, 0.0}, fs{0.0, 0.0, 0.0, 0.0};
    if (valid) {
      const double* o = prec + (size_t)(pos + z) * PS_REC;
      sf = o[RC_SF];
      cf = o[RC_CF];
      fe = Frame{o[RC_UX], o[RC_UY], o[RC_VX], o[RC_VZ]};
      fs = Frame{o[RC_UX0], o[RC_UY0], o[RC_VX0], o[RC_VZ0]};
      tt = o[RC_TT];
      scale = o[RC_SCALE];
      ff = o[RC_FF];
      d2 = o[RC_D2];
      alph = o[RC_ALPH];
    }
    body_dir(cf, sf, fe.vx, fe.uy, fe.vz, gg);
    body_derivs(v, cf, sf, tt, fe, gg, db, dl, dw);
    body_dir(cf, sf, fs.vx, fs.uy, fs.vz, gg);
    body_derivs(v, cf, sf, tt, fs, gg, db0, dl0, dw0);
    auto tsum = [&](const double* dv, const double* dv0) {
      double r = dv[0] * U0;
      r = fma(dv[1], U1, r);
      r = fma(dv[2], U2, r);
      r = fma(dv0[0], V0, r);
      r = fma(dv0[1], V1, r);
      r = fma(dv0[2], V2, r);
      return r;
    };
    const double T1 = tsum(db, db0), T2 = tsum(dl, dl0), T3 = tsum(dw, dw0);
    const double sc = scale;
    double* rb = rowbuf + ps * 16;
    rb[0] = sc * T1;
    rb[1] = sc * T2;
    rb[2] = sc * T3;
    rb[3] = br * ff;
    rb[4] = br * d2;
    rb[5] = br * alph;
    rb[6] = (sc * t4) * cl;
    rb[7] = sc * t5;
    rb[8] = br * sc;
    rb[9] = sc;
  }
  wave_sync();
  // rows out: lane (ps, sl) writes column 16b + ps of the tile's points
  // 4r + sl.  Blocks below NBS - 1 hold shape columns only (NBS =
  // ceil((C - 1) / 16)); FULL: all 16 points of the tile exist (every tile
  // but the last)
  auto rows_out = [&](auto full_tag) __attribute__((always_inline)) {
    constexpr bool FULL = decltype(full_tag)::value;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int ptl = 4 * r + sl;
      const bool ok = FULL || t * 16 + ptl < N;
      const int prow = __shfl(pt, ptl);  // lane ptl holds the point of tile row ptl
      const double* rb = rowbuf + ptl * 16;
      double* jr = jbuf + (size_t)prow * QS;
      const double sc9 = rb[9];
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int q = 16 * b + ps;
        double val;
        if (b < NBS - 1) {
          val = sc9 * gat[b][r];
        } else {
          const int k = q - (C - 1);  // non-shape column index, 0..7
          const double ns = (k >= 0 && k <= 7) ? rb[k >= 0 && k <= 7 ? k : 0] : 0.0;
          val = (b < NBS && q <= C - 2) ? sc9 * gat[b < NBS ? b : 0][r] : ns;
        }
        if (FULL || ok) jr[q] = val;
      }
      if (ok && ps == 0) ybuf[prow] = rb[8];
    }
  };
  if (t * 16 + 16 <= N)
    rows_out(std::true_type{});
  else
    rows_out(std::false_type{});
  wave_sync();
  return nvis;
}

// ---------------------------------------------------------------------------
// Phase 2: the augmented normal matrix on FP64 MFMA.  Relative curves are
// renormalised algebraically (Start.cu:2656-2715 restated): the chains
// accumulate (W Z)^T Z for rows Z = [c J | r | ... | z | y] (r = obs - c y,
// c = sig lp / L; z and y are the last two columns of the last block)
// whose weighted side is [w c J | w r | ... | 1 | w y] on relative curves
// and 0 in the two sum columns otherwise.  Row z of the product then holds
// D = sum c J, row y holds U = sum w y c J,
// Y = sum w y^2 and R = sum w y r: at the end of each relative curve the
// chain writes those two rows out and clears them.  The correction
// -(a_k b_k^T + b_k a_k^T) per curve (a = U - (Y / 2L) D with a[mf] = R,
// b = D / L) is applied afterwards, so the rows are read once and no
// per-curve pass over them precedes the product.
// The waves form P2<NB>::GT point chains; chain g accumulates the tiles
// T = g (mod GT) (data tiles in curve order, then the convexity tile) and,
// for NB > 4, its two waves split the 16x16 block pairs (ib >= jb).  Every
// element of a chain is one fused multiply-add chain over its points in
// order; the chains' partial matrices go to the slab and are added as
// (C0 + C1) + (C2 + C3) (GT = 4) or C0 + C1 (GT = 2).
// ---------------------------------------------------------------------------
template <int NB>
struct P2 {
  static constexpr int NP = NB * (NB + 1) / 2;  // lower block pairs
  static constexpr int GT = NB <= 4 ? 4 : 2;    // point chains
  static constexpr int GB = NW >= GT ? NW / GT : 1;   // waves per chain
  static constexpr int CPW = NW >= GT ? 1 : GT / NW;  // chains per wave (run in turn)
  static constexpr int NPW = (NP + GB - 1) / GB;
};
static_assert(P2<4>::GT == 4 && P2<5>::GT == 2, "chain count must match ps_tile_chains");

// Phase 2 rows: lane (ps, sl) handles compact column 16b + ps of the points
// 4r + sl (r = 0..3) of a tile.  The column's kind is fixed per lane and
// block: a Jacobian column (q < mfit), the residual (q == mfit), padding,
// or in the last block the sum column z (q = 16 NB - 2) and the brightness
// column y (q = 16 NB - 1).
template <int NB>
struct P2Cols {
  unsigned jm = 0, rm = 0;
  bool z = false, y = false;  // the lane's last-block column is z / y
  __device__ __forceinline__ P2Cols(int ps, int mf) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int q = 16 * b + ps;
      if (q < mf) jm |= 1u << b;
      if (q == mf) rm |= 1u << b;
    }
    z = ps == 14;
    y = ps == 15;
  }
};

// The products of one row into the wave's block pairs: element (i, j)
// accumulates (weighted side)_i * (plain side)_j.
template <int NB, int NPW, int GB>
__device__ __forceinline__ void p2_products(const double* X, const double* XW, int gb, v4d* acc) {
  int idx = 0;
#pragma unroll
  for (int ib = 0; ib < NB; ++ib)
#pragma unroll
    for (int jb = 0; jb <= ib; ++jb) {
      if (idx % GB == gb) acc[idx / GB] = mfma(XW[ib], X[jb], acc[idx / GB]);
      ++idx;
    }
}

// One wave's share of phase 2: its chain's data tiles, the next row's loads
// issued before the current row's products, then (on the chain that owns
// it) the convexity tile, which closes the chain.  Writes the chain's
// partial matrix and, at the end of each relative curve, rows z and y of
// that curve.  Compact column q is parameter q + 2 - ia1, stored in the raw
// row at column q - ia1; with ia1 the lane of compact column 0 holds the
// size scale, whose derivative is y (zero after renormalisation).
template <int NB>
__device__ __forceinline__ void phase2_chain(const PsDevProblem& P, const Slab& S, double conw_w) {
  using SH = P2<NB>;
  const int tid = tid_opaque(), lane = tid & 63, ps = lane & 15, sl = lane >> 4, w = tid >> 6;
  const int gb = w % SH::GB;
  const int mf = P.mfit, C = P.ncoef, QS = P.qs, ia1 = P.ia1;
  const int CWS = ps_curve_sums_width(QS);
  const P2Cols<NB> K2(ps, mf);
  const bool p1 = ia1 != 0 && ps == 0;  // block 0 of this lane is the size scale
  const int c0 = ps - ia1 < 0 ? 0 : ps - ia1;  // row column of block 0 (any valid one for the size scale)
  const int c1 = ps - ia1 + 16;                // row column of block 1; block b: c1 + 16 (b - 1)
  // one chain gt of this wave
  auto run_chain = [&](const int gt) __attribute__((always_inline)) {
    v4d acc[SH::NPW];
#pragma unroll
    for (int i = 0; i < SH::NPW; ++i) acc[i] = v4d{0.0, 0.0, 0.0, 0.0};
    struct Row {
      double J[NB];
      double y, sig, obs, s2;
    };
    // with fl: rows z and y (16 NB - 2, 16 NB - 1: block row NB - 1, r = 3,
    // lanes sl = 2, 3) of this chain's matrix out for curve c, then cleared
    auto flush = [&](int c, bool fl) __attribute__((always_inline)) {
      double* o = S.csum + ((size_t)gt * P.ncurves + c) * CWS + (sl == 2 ? 0 : QS) + ps;
#pragma unroll
      for (int jb = 0; jb < NB; ++jb) {
        constexpr int base = (NB - 1) * NB / 2;
        const int idx = base + jb;
        if (idx % SH::GB == gb) {
          if (fl && sl >= 2) o[16 * jb] = acc[idx / SH::GB][3];
          acc[idx / SH::GB][3] = (fl && sl >= 2) ? 0.0 : acc[idx / SH::GB][3];
        }
      }
    };
    // ALLJ: blocks 0 .. NB-2 hold Jacobian columns only (16 (NB - 1) <= mfit,
    // true for every input the triangulation and basis produce); only the
    // last block then needs the per-lane column kind.
    auto tiles = [&](auto allj_tag) __attribute__((always_inline)) {
      constexpr bool ALLJ = decltype(allj_tag)::value;
      // first point, points in the tile, curve, curve length (negative:
      // absolute curve), one load; past the last tile: a one-point tile at
      // point 0
      auto info = [&](int T) __attribute__((always_inline)) {
        return T < P.ntiles ? *reinterpret_cast<const int4*>(P.tile_info + 4 * T) : make_int4(0, 1, 0, 1);
      };
      auto load = [&](int4 t, int r) __attribute__((always_inline)) {
        Row R;
        const int pin = 4 * r + sl;
        const int pt = t.x + (pin < t.y ? pin : 0);
        const double* jr = S.jbuf + (size_t)pt * QS;
        R.J[0] = jr[c0];
#pragma unroll
        for (int b = 1; b < NB; ++b) R.J[b] = jr[c1 + 16 * (b - 1)];
        R.y = S.ybuf[pt];
        R.sig = P.sig[pt];
        R.obs = P.bright[pt];
        R.s2 = P.s2w[pt];
        return R;
      };
      int4 ti = info(gt);
      Row nxt = load(ti, 0);
      for (int T = gt; T < P.ntiles; T += SH::GT) {
        const int4 tn = info(T + SH::GT);
        const int n_in = ti.y, c = ti.z;
        const bool rel = ti.w > 0;
        const double lpd = (double)(rel ? ti.w : -ti.w);
        const double rave = rel ? S.crv[c] : 0.0;
        // FULL: the tile has 16 points (every tile but a curve's last one);
        // REL: a relative curve
        auto rows = [&](auto full_tag, auto rel_tag) __attribute__((always_inline)) {
          constexpr bool FULL = decltype(full_tag)::value;
          constexpr bool REL = decltype(rel_tag)::value;
          Row rw[2];
          rw[0] = nxt;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            if (r < 3)
              rw[(r + 1) & 1] = load(ti, r + 1);
            else
              nxt = load(tn, 0);
            const Row& cur = rw[r & 1];
            const bool ok = FULL || 4 * r + sl < n_in;
            const double coef = REL ? (cur.sig * lpd) * rave : 1.0;
            const double res = cur.obs - coef * cur.y;
            const double s2v = ok ? cur.s2 : 0.0;
            double X[NB], XW[NB];
#pragma unroll
            for (int b = 0; b < NB; ++b) {
              double xr = cur.J[b];
              if (b == 0) xr = p1 ? (REL ? 0.0 : cur.y) : xr;
              const double xj = coef * xr;
              double xv, xw;
              if (ALLJ && b < NB - 1) {
                xv = ok ? xj : 0.0;
                xw = s2v * xv;
              } else {
                const bool isj = (K2.jm >> b) & 1u, isr = (K2.rm >> b) & 1u;
                const bool isz = b == NB - 1 && K2.z, isy = b == NB - 1 && K2.y;
                xv = isj ? xj : isr ? res : isy ? cur.y : 0.0;
                xv = ok ? xv : 0.0;
                xw = s2v * xv;
                if (isz) xw = (REL && ok) ? 1.0 : 0.0;
                if (isy) xw = REL ? s2v * cur.y : 0.0;
              }
              X[b] = xv;
              XW[b] = xw;
            }
            p2_products<NB, SH::NPW, SH::GB>(X, XW, gb, acc);
          }
        };
        if (rel) {
          if (n_in == 16)
            rows(std::true_type{}, std::true_type{});
          else
            rows(std::false_type{}, std::true_type{});
          // the chain's last tile of this curve (predicated, not branched
          // around: a branch here costs registers in the tile loop)
          flush(c, T + SH::GT >= P.ntiles || tn.z != c);
        } else {
          if (n_in == 16)
            rows(std::true_type{}, std::false_type{});
          else
            rows(std::false_type{}, std::false_type{});
        }
        ti = tn;
      }
    };
    if (16 * (NB - 1) <= mf)
      tiles(std::true_type{});
    else
      tiles(std::false_type{});
    // the convexity pseudo-curve (rows ndata .. ndata+2, Start.cu:2539-2612):
    // tile T = ntiles, the last tile of chain ntiles mod GT; only its first
    // row group holds rows (lanes sl < 3); raw rows, no sums
    if (P.ntiles % SH::GT == gt) {
      const bool ok = sl < 3;
      const int pt = P.ndata + (ok ? sl : 0);
      const double* jr = S.jbuf + (size_t)pt * QS;
      const double yv = S.ybuf[pt];
      const double res = 0.0 - yv;
      const double s2v = ok ? conw_w * 1.0 : 0.0;
      double X[NB], XW[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int par = 16 * b + ps + 2 - ia1;  // parameter of the lane's column (1-based)
        const double jv = b == 0 ? jr[c0] : jr[c1 + 16 * (b - 1)];
        double xv = (K2.jm >> b) & 1u ? (par == 1 ? yv : par <= C ? jv : 0.0) : (K2.rm >> b) & 1u ? res : 0.0;
        xv = ok ? xv : 0.0;
        X[b] = xv;
        XW[b] = s2v * xv;
      }
      p2_products<NB, SH::NPW, SH::GB>(X, XW, gb, acc);
    }
    double* pt = S.part + (size_t)gt * 256 * NB * NB;
    const int ld = 16 * NB;
    int idx = 0;
#pragma unroll
    for (int ib = 0; ib < NB; ++ib)
#pragma unroll
      for (int jb = 0; jb <= ib; ++jb) {
        if (idx % SH::GB == gb) {
#pragma unroll
          for (int r = 0; r < 4; ++r) pt[(size_t)(16 * ib + 4 * r + sl) * ld + 16 * jb + ps] = acc[idx / SH::GB][r];
        }
        ++idx;
      }
  };
  // the chains of this wave: one (NW >= GT), or GT / NW in turn
  if constexpr (SH::CPW == 1) {
    run_chain(w / SH::GB);
  } else {
#pragma unroll 1
    for (int cw = 0; cw < SH::CPW; ++cw) run_chain(w + cw * NW);
  }
}

// Brightness sums of the data tiles (16 consecutive points of one curve, in
// file order: four partials over points p mod 4, (G0 + G1) + (G2 + G3)),
// then per relative curve 1 / L, L = the curve's tile sums in order.
__device__ __forceinline__ void curve_sums(const PsDevProblem& P, const Slab& S) {
  const int tid = tid_opaque();
  for (int T = tid; T < P.ntiles; T += WG) {
    const int4 ti = *reinterpret_cast<const int4*>(P.tile_info + 4 * T);
    double g4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int p = 0; p < 16; ++p) {
      const double val = p < ti.y ? S.ybuf[ti.x + p] : 0.0;
      g4[p & 3] = g4[p & 3] + val;
    }
    S.tsy[T] = (g4[0] + g4[1]) + (g4[2] + g4[3]);
  }
  __syncthreads();
  for (int c = tid; c < P.ncurves; c += WG) {
    const int t0 = P.tile_off[c], t1 = P.tile_off[c + 1];
    if (P.tile_info[4 * t0 + 3] <= 0) continue;  // absolute curve
    double s = 0.0;
    for (int T = t0; T < t1; ++T) s = s + S.tsy[T];
    S.crv[c] = 1.0 / s;
  }
}

// The correction vectors of every curve (zero for absolute curves and the
// padding curve): the chains' rows z (D) and y (U) of the curve combined as
// the matrices are (a chain without a tile of the curve gives 0.0), then
// a = U - (Y / 2L) D with a[mf] = R, b = D / L with b[mf] = 0, where
// Y = U[16 NB - 1] and R = U[mf].
template <int NB>
__device__ __forceinline__ void correction_table(const PsDevProblem& P, const Slab& S) {
  using SH = P2<NB>;
  const int QS = P.qs, mf = P.mfit, NCP = ps_corr_curves(P.ncurves), CWS = ps_curve_sums_width(QS);
  auto comb = [](const double* x) { return SH::GT == 4 ? (x[0] + x[1]) + (x[2] + x[3]) : x[0] + x[1]; };
  for (int e = tid_opaque(); e < NCP * QS; e += WG) {
    const int c = e / QS, q = e - c * QS;
    double av = 0.0, bv = 0.0;
    if (c < P.ncurves && q <= mf && P.tile_info[4 * P.tile_off[c] + 3] > 0) {
      const int t0 = P.tile_off[c], nt = P.tile_off[c + 1] - t0;
      double dd[SH::GT], uu[SH::GT], yy[SH::GT];
#pragma unroll
      for (int g = 0; g < SH::GT; ++g) {
        const bool seen = nt >= SH::GT || ((g - t0) % SH::GT + SH::GT) % SH::GT < nt;
        const double* o = S.csum + ((size_t)g * P.ncurves + c) * CWS;
        dd[g] = seen ? o[q] : 0.0;
        uu[g] = seen ? o[QS + q] : 0.0;
        yy[g] = seen ? o[QS + QS - 1] : 0.0;
      }
      const double D = comb(dd), U = comb(uu), Y = comb(yy);
      const double rave = S.crv[c];
      if (q < mf) {
        av = fma(-(0.5 * (Y * rave)), D, U);
        bv = rave * D;
      } else {
        av = U;  // R = sum w y r
      }
    }
    S.ttab[(size_t)(2 * c) * QS + q] = av;
    S.ttab[(size_t)(2 * c + 1) * QS + q] = bv;
  }
}

// The chains' partial matrices combined, (C0 + C1) + (C2 + C3) or C0 + C1,
// then the correction as one MFMA chain per block pair: element (i, j)
// accumulates -a_k[i] b_k[j] and then -b_k[i] a_k[j] for every curve k in
// order.  Wave w takes the lower block pairs idx = w (mod NW).
template <int NB>
__device__ __forceinline__ void combine_correct(const PsDevProblem& P, const Slab& S, double* work) {
  using SH = P2<NB>;
  const int tid = tid_opaque(), lane = tid & 63, ps = lane & 15, sl = lane >> 4;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int K = P.mfit + 1, QS = P.qs, nk = 2 * ps_corr_curves(P.ncurves);
  const int ld = 16 * NB;
  const size_t pst = (size_t)256 * NB * NB;
  int idx = 0;
#pragma unroll
  for (int ib = 0; ib < NB; ++ib)
#pragma unroll
    for (int jb = 0; jb <= ib; ++jb) {
      if (idx % NW == w) {
        v4d acc;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const double* c0 = S.part + (size_t)(16 * ib + 4 * r + sl) * ld + 16 * jb + ps;
          acc[r] = SH::GT == 4 ? (c0[0] + c0[pst]) + (c0[2 * pst] + c0[3 * pst]) : c0[0] + c0[pst];
        }
        const double* ta = S.ttab + 16 * ib + ps;
        const double* tb = S.ttab + 16 * jb + ps;
        for (int k4 = 0; k4 < nk; k4 += 4) {
          const int k = k4 + sl;
          acc = mfma(-ta[(size_t)k * QS], tb[(size_t)(k ^ 1) * QS], acc);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = 16 * ib + 4 * r + sl, col = 16 * jb + ps;
          if (row < K && col < K) work[row * K + col] = acc[r];
        }
      }
      ++idx;
    }
}

// One mrqcof evaluation at a[] (LDS) into L.work (K*K: symmetric alpha,
// beta in row mfit, chi^2 at [mfit][mfit]); facet areas of a in L.nrm[4f+3];
// raw Jacobian rows of every point (and the convexity rows) in the slab.
template <int NB, int NBS>
__device__ __forceinline__ void evaluate(const PsDevProblem& P, const Lds& L, const double* a, double conw_r2, const Slab& S,
                         unsigned long long* vis_count, int claim) {
  using SH = P2<NB>;
  const int tid = tid_opaque(), lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave index, in a scalar register
  const int mf = P.mfit, K = mf + 1, QS = P.qs, C = P.ncoef, F = P.nfac, N = P.ndata;
  PS_TSTART();
  // view-direction cell counts, in the normal-matrix area (dead until the
  // end of the evaluation; the facet pass later reuses it for its rows)
  int* cell = reinterpret_cast<int*>(L.work);
  // the point passes up to the facet pass wait on memory: like the solve,
  // they run at wave priority 2 beside the other workgroups' facet passes
  __builtin_amdgcn_s_setprio(2);
  for (int e = tid; e < kCells; e += WG) cell[e] = 0;
  curv_block(P, a, L.nrm);
  int* next_tile = reinterpret_cast<int*>(L.red);  // phase-1 work counter (red is free until the solve)
  if (tid == 0) {
    store_inv(make_inv(P, a), L.inv);
    *next_tile = 0;
  }
  __syncthreads();
  // each point's view-direction cell, from its rotation angle and the
  // body-frame directions of Earth and Sun (Start.cu:2095-2195)
  {
    const Inv v = load_inv(L.inv);
    for (int pt = tid; pt < N; pt += WG) {
      const double tt = P.tim[pt];
      const double ex = P.ex[pt], ey = P.ey[pt], ez = P.ez[pt];
      const double sx = P.sx[pt], sy = P.sy[pt], sz = P.sz[pt];
      double sf, cf;
      rotation_angle(P, v, tt, &sf, &cf);
      const Frame fe = pole_frame(v, ex, ey, ez);
      const Frame fs = pole_frame(v, sx, sy, sz);
      double g[3], g0[3];
      body_dir(cf, sf, fe.vx, fe.uy, fe.vz, g);
      body_dir(cf, sf, fs.vx, fs.uy, fs.vz, g0);
      const int key = view_cell(g, g0);
      S.pkey[pt] = (unsigned char)key;
      atomicAdd(cell + key, 1);
    }
  }
  PS_TMARK(0);
  // convexity pseudo-curve (Start.cu:2539-2612), rows N..N+2:
  // y_j = sum Area n_j, J_j[c] = sum (Area n_j) Y_c
  for (int idx = tid; idx < 3 * QS; idx += WG) {
    const int j = idx / QS, q = idx - j * QS;
    double s = 0.0;
    if (q <= C - 2) {
      const double* yq = L.ytab + (size_t)(q & 15) * P.nbs + (q >> 4);
      const size_t ys_ = (size_t)16 * P.nbs;  // basis stride of one facet
#pragma unroll 8
      for (int f = 0; f < F; ++f) s = fma(L.nrm[4 * f + 3] * L.nrm[4 * f + j], yq[f * ys_], s);
    }
    S.jbuf[(size_t)(N + j) * QS + q] = s;
    if (q == QS - 1) {
      double ys = 0.0;
#pragma unroll 8
      for (int f = 0; f < F; ++f) ys = ys + L.nrm[4 * f + 3] * L.nrm[4 * f + j];
      S.ybuf[N + j] = ys;
    }
  }
  __syncthreads();  // the cells and their counts are complete
  // counting sort of the points by cell: exclusive prefix of the counts
  // (wave 0; lane l owns cells 4l .. 4l + 3, kCells <= 256), then each point
  // takes the next slot of its cell.  The order within a cell depends on
  // timing; no result depends on the order.
  if (w == 0) {
    int c4[4], sum = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = 4 * lane + j;
      c4[j] = e < kCells ? cell[e] : 0;
      sum += c4[j];
    }
    int incl = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(incl, off);
      if (lane >= off) incl += y;
    }
    int base = incl - sum;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = 4 * lane + j;
      if (e < kCells) cell[e] = base;
      base += c4[j];
    }
  }
  __syncthreads();
  // each point's record (Start.cu:2095-2195: rotation angle, pole-frame
  // Earth and Sun directions, phase-function terms), computed with the same
  // arithmetic as the cells above and stored at the point's place in the
  // sorted order
  {
    const Inv v = load_inv(L.inv);
    for (int pt = tid; pt < N; pt += WG) {
      // the point's inputs, loaded together ahead of the arithmetic
      const double tt = P.tim[pt], al = P.alph[pt];
      const double ex = P.ex[pt], ey = P.ey[pt], ez = P.ez[pt];
      const double sx = P.sx[pt], sy = P.sy[pt], sz = P.sz[pt];
      const int pos = atomicAdd(cell + S.pkey[pt], 1);
      const PointScalars q = point_scalars(P, v, tt, al);
      const Frame fe = pole_frame(v, ex, ey, ez);
      const Frame fs = pole_frame(v, sx, sy, sz);
      double* o = S.prec + (size_t)pos * PS_REC;
      o[RC_SF] = q.sf;
      o[RC_CF] = q.cf;
      o[RC_VX] = fe.vx;
      o[RC_UY] = fe.uy;
      o[RC_VZ] = fe.vz;
      o[RC_UX] = fe.ux;
      o[RC_VX0] = fs.vx;
      o[RC_UY0] = fs.uy;
      o[RC_VZ0] = fs.vz;
      o[RC_UX0] = fs.ux;
      o[RC_TT] = q.tt;
      o[RC_SCALE] = q.scale;
      o[RC_FF] = q.ff;
      o[RC_D2] = q.d2;
      o[RC_ALPH] = q.alph;
      S.perm[pos] = pt;
    }
  }
  __syncthreads();  // the facet-pass order and the records are complete
  __builtin_amdgcn_s_setprio(0);
  PS_TMARK(9);
  // phase 1 over the points in view-direction order, 16 per tile; each wave
  // takes the next `claim` unclaimed tiles (tiles differ in cost by their
  // visible facet steps; the assignment changes no result)
  double* rowbuf = L.rowbuf + w * 256;
  unsigned long long nvis = 0;
  const int n_tiles = (N + 15) >> 4;
  claim = claim < 1 ? 1 : claim;  // every claim takes at least one tile: the loop ends
  for (;;) {
    int t0 = 0;
    if (lane == 0) t0 = atomicAdd(next_tile, claim);
    t0 = __builtin_amdgcn_readfirstlane(t0);
    if (t0 >= n_tiles) break;
    const int t1 = t0 + claim < n_tiles ? t0 + claim : n_tiles;
    for (int t = t0; t < t1; ++t) nvis += phase1_tile<NB, NBS>(P, L, S.prec, S.perm, t, S.jbuf, S.ybuf, rowbuf);
  }
  PS_TMARK(1);
  __syncthreads();
  // per relative curve: brightness sum and renormalisation factor
  curve_sums(P, S);
  __syncthreads();
  PS_TMARK(2);
  phase2_chain<NB>(P, S, conw_r2);
  __syncthreads();
  PS_TMARK(4);
  correction_table<NB>(P, S);
  __syncthreads();
  PS_TMARK(5);
  combine_correct<NB>(P, S, L.work);
  __syncthreads();
  for (int e = tid; e < mf * mf; e += WG) {
    const int i = e / mf, j = e - i * mf;
    if (j > i) L.work[i * K + j] = L.work[j * K + i];
  }
  __syncthreads();
  if (vis_count) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) nvis += __shfl_xor(nvis, off);
    if (lane == 0) atomicAdd(vis_count, nvis);
  }
  PS_TMARK(6);
}

// Pivot candidates are ordered by |entry|, then by position (row * 128 +
// column): the larger position wins a tie, which is the NR rule's last
// occurrence in row-major order.  |entry| >= 0 is compared through its bit
// pattern (monotonic for non-negative doubles; NaN never becomes a
// candidate); "no candidate" is (0, -1), below every real candidate.
__device__ __forceinline__ bool cand_better(unsigned long long ok, int oi, unsigned long long k, int i) {
  return ok > k || (ok == k && oi > i);
}

// One step of the wave argmax: combine with the candidate that DPP control
// CTRL brings in (lanes outside ROWMASK keep their own).
template <int CTRL, int ROWMASK>
__device__ __forceinline__ void cand_step(unsigned long long& k, int& i) {
  const int lo = (int)(unsigned)k, hi = (int)(unsigned)(k >> 32);
  const int olo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROWMASK, 0xF, false);
  const int ohi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROWMASK, 0xF, false);
  const int oi = __builtin_amdgcn_update_dpp(i, i, CTRL, ROWMASK, 0xF, false);
  const unsigned long long ok = ((unsigned long long)(unsigned)ohi << 32) | (unsigned)olo;
  if (cand_better(ok, oi, k, i)) {
    k = ok;
    i = oi;
  }
}

// One step of a wave maximum of a double (lanes outside ROWMASK keep their
// own value).
// lane l's value of v, in scalar registers
__device__ __forceinline__ double readlane_d(double v, int l) {
  const unsigned long long b = __double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = __builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double(((unsigned long long)hi << 32) | lo);
}

// b, through a register the compiler cannot see into (keeps a per-lane flag
// from being merged into the per-row uniform tests that use it)
__device__ __forceinline__ bool opaque_flag(bool b) {
  int v = b;
  asm volatile("" : "+v"(v));
  return v != 0;
}

template <int CTRL, int ROWMASK>
__device__ __forceinline__ double max_step(double v) {
  const unsigned long long b = __double_as_longlong(v);
  const int lo = (int)(unsigned)b, hi = (int)(unsigned)(b >> 32);
  const int olo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROWMASK, 0xF, false);
  const int ohi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROWMASK, 0xF, false);
  return fmax(v, __longlong_as_double(((long long)(unsigned)ohi << 32) | (unsigned)olo));
}

template <int CTRL, int ROWMASK>
__device__ __forceinline__ int imax_step(int v) {
  const int o = __builtin_amdgcn_update_dpp(v, v, CTRL, ROWMASK, 0xF, false);
  return o > v ? o : v;
}

// The wave's best pivot candidate, as wave_argmax, in two plain maxima: the
// largest |entry| (t; -1 where a lane has no candidate), then the largest
// position among the lanes that hold it (the NR rule's last occurrence).
// Both in lane 63 after the same DPP pattern, read back to every lane; a
// third of wave_argmax's instructions, with no compare-and-select chains.
__device__ __forceinline__ void wave_argmax2(double t, int idx, unsigned long long& k, int& i) {
  double m = max_step<0xB1, 0xF>(t);
  m = max_step<0x4E, 0xF>(m);
  m = max_step<0x141, 0xF>(m);
  m = max_step<0x140, 0xF>(m);
  m = max_step<0x142, 0xA>(m);
  m = max_step<0x143, 0xC>(m);
  const unsigned long long mb = __double_as_longlong(m);
  const unsigned mlo = __builtin_amdgcn_readlane((int)(unsigned)mb, 63);
  const unsigned mhi = __builtin_amdgcn_readlane((int)(unsigned)(mb >> 32), 63);
  const double M = __longlong_as_double(((long long)mhi << 32) | mlo);
  int j = t == M ? idx : -1;
  j = imax_step<0xB1, 0xF>(j);
  j = imax_step<0x4E, 0xF>(j);
  j = imax_step<0x141, 0xF>(j);
  j = imax_step<0x140, 0xF>(j);
  j = imax_step<0x142, 0xA>(j);
  j = imax_step<0x143, 0xC>(j);
  i = M < 0.0 ? -1 : __builtin_amdgcn_readlane(j, 63);
  k = M < 0.0 ? 0ull : (unsigned long long)__double_as_longlong(M);
}

// The wave's best candidate, in lane 63: quad butterflies, the half-row and
// row mirrors (every lane of a 16-lane row then holds the row's best), then
// the row broadcasts 15 (rows 1, 3) and 31 (rows 2, 3).
__device__ __forceinline__ void wave_argmax(unsigned long long& k, int& i) {
  cand_step<0xB1, 0xF>(k, i);   // quad_perm [1, 0, 3, 2]
  cand_step<0x4E, 0xF>(k, i);   // quad_perm [2, 3, 0, 1]
  cand_step<0x141, 0xF>(k, i);  // row_half_mirror
  cand_step<0x140, 0xF>(k, i);  // row_mirror
  cand_step<0x142, 0xA>(k, i);  // row_bcast:15
  cand_step<0x143, 0xC>(k, i);  // row_bcast:31
  const unsigned lo = __builtin_amdgcn_readlane((int)(unsigned)k, 63);
  const unsigned hi = __builtin_amdgcn_readlane((int)(unsigned)(k >> 32), 63);
  i = __builtin_amdgcn_readlane(i, 63);
  k = ((unsigned long long)hi << 32) | lo;
}

// Damped full-pivot Gauss-Jordan (Start.cu:2838-3021) by the first GW
// waves of the workgroup on the augmented matrix [A | b] in LDS (rows
// 0..mf-1, stride K = mf + 1, b in column mf); further waves only keep the
// barriers (a pivot search in more waves costs more issue slots than it
// saves).  Lane l of every solving wave owns columns l and l + 64; wave w
// owns the rows ll = w (mod GW).  NR gaussj pivot rule: the
// largest |entry| among unpivoted rows and columns, the last occurrence in
// row-major order on ties; each wave searches its rows while eliminating
// them, reduces its lanes' candidates with DPP, and the waves' candidates
// meet in red[0 .. 2*NW).  Two barriers per pivot: every wave reads rows
// irow and icol before the second; after it, the waves that own those rows
// store the swapped rows (icol: the scaled pivot row, irow: old row icol)
// and every wave eliminates its rows, loading up to 16 of them per batch.
// a holds K * K doubles: row mf is the solve's own scratch (a zero slot and
// the dummy slots below).  Returns 0, or 2 for a zero pivot (or nothing
// selectable); the solution is left in column mf.  Every thread of the
// workgroup must call it.  TWO: columns beyond 64 exist (K > 64).
constexpr int GW = NW > 4 ? 4 : NW;  // waves that solve

template <bool TWO>
__device__ __forceinline__ int gaussj_block(double* a, int mf, int K, double* red) {
  const int tid = tid_opaque(), lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave index, in a scalar register
  const int w0 = w < GW ? w : mf;                          // first owned row (none beyond GW)
  const int c0 = lane, c1 = lane + 64;
  const bool h0 = c0 < K, h1 = TWO && c1 < K;
  // byte addresses of the elimination's loads and stores, row lr at
  // e0 + lr * rs0: a lane with no column of the matrix (c >= K) loads and
  // stores a dummy slot in row mf (columns 1..K-1), which the solve never
  // reads, instead of branching around each row's load and store.  Row mf,
  // column 0 holds 0.0: the pivot column's lane loads it as its base.
  char* const ab = reinterpret_cast<char*>(a);
  char* const zero = ab + 8 * mf * K;
  const unsigned rs0 = h0 ? 8u * K : 0u, rs1 = h1 ? 8u * K : 0u;
  char* const e0 = ab + 8 * (h0 ? c0 : mf * K + 1 + min(c0 - K, K - 2));
  char* const e1 = ab + 8 * (h1 ? c1 : mf * K + 1 + min(c1 - K, K - 2));
  if (tid == 0) *reinterpret_cast<double*>(zero) = 0.0;  // read after the first barrier
  unsigned long long u0 = 0, u1 = 0;  // pivoted rows / columns
  auto used = [&](int k) -> bool {
    if constexpr (TWO)
      return k < 64 ? (u0 >> k) & 1ull : (u1 >> (k - 64)) & 1ull;
    else
      return (u0 >> k) & 1ull;  // every index is below mf <= 63
  };
  double big = 0.0;  // the candidate's signed value: compared by magnitude
  int bidx = -1;     // row * 128 + column
  // one mask operation and one select per register (no nested selects)
  auto consider = [&](bool ok, double v, int idx) {
    const bool take = ok & (fabs(v) >= fabs(big));
    big = take ? v : big;
    bidx = take ? idx : bidx;
  };
  for (int ll = w0; ll < mf; ll += GW) {
    consider(c0 < mf, *reinterpret_cast<const double*>(e0 + __umul24(ll, rs0)), ll * 128 + c0);
    if constexpr (TWO) consider(c1 < mf, *reinterpret_cast<const double*>(e1 + __umul24(ll, rs1)), ll * 128 + c1);
  }
  for (int i = 0; i < mf; ++i) {
    unsigned long long key;
    int kidx;
    if constexpr (NW <= 4) {
      wave_argmax2(bidx >= 0 ? fabs(big) : -1.0, bidx, key, kidx);
    } else {
      // the 8-wave kernel (128 VGPRs) ran 1.1 % slower with the two maxima
      key = bidx >= 0 ? (unsigned long long)__double_as_longlong(fabs(big)) : 0ull;
      kidx = bidx;
      wave_argmax(key, kidx);
    }
    unsigned long long* ured = reinterpret_cast<unsigned long long*>(red);
    long long* ired = reinterpret_cast<long long*>(red + NW);
    if (lane == 0 && w < GW) {
      ured[w] = key;
      ired[w] = kidx;
    }
    __syncthreads();
    unsigned long long wk[GW];
    int wi[GW];
#pragma unroll
    for (int ww = 0; ww < GW; ++ww) {
      wk[ww] = ured[ww];
      wi[ww] = (int)ired[ww];
    }
    key = 0;
    kidx = -1;
#pragma unroll
    for (int ww = 0; ww < GW; ++ww)
      if (cand_better(wk[ww], wi[ww], key, kidx)) {
        key = wk[ww];
        kidx = wi[ww];
      }
    bidx = __builtin_amdgcn_readfirstlane(kidx);
    if (bidx < 0) return 2;
    const int irow = bidx >> 7, icol = bidx & 127;
    if (icol < 64)
      u0 |= 1ull << icol;
    else
      u1 |= 1ull << (icol - 64);
    // this lane's columns still open for the next pivot (opaque to the
    // compiler, so that each row tests them with one mask operation)
    const bool open0 = opaque_flag(c0 < mf && !used(c0));
    const bool open1 = TWO && opaque_flag(c1 < mf && !used(c1));
    // rows irow and icol before anyone rewrites them
    double pv0 = 0.0, pv1 = 0.0, cv0 = 0.0, cv1 = 0.0;
    if (h0) {
      pv0 = a[irow * K + c0];
      cv0 = a[icol * K + c0];
    }
    if (h1) {
      pv1 = a[irow * K + c1];
      cv1 = a[icol * K + c1];
    }
    const double piv = a[irow * K + icol];  // row irow, column icol
    __syncthreads();
    if (piv == 0.0) {
      // the rows are swapped before the zero pivot is found
      if (w == 0 && irow != icol) {
        if (h0) {
          a[irow * K + c0] = cv0;
          a[icol * K + c0] = pv0;
        }
        if (h1) {
          a[irow * K + c1] = cv1;
          a[icol * K + c1] = pv1;
        }
      }
      return 2;
    }
    const double pivinv = 1.0 / piv;
    const double pn0 = (c0 == icol ? 1.0 : pv0) * pivinv;
    const double pn1 = (c1 == icol ? 1.0 : pv1) * pivinv;
    // the row swap: row icol becomes the scaled pivot row, row irow old
    // row icol (each written by the wave that owns it)
    if (icol % GW == w) {
      if (h0) a[icol * K + c0] = pn0;
      if (h1) a[icol * K + c1] = pn1;
    }
    if (irow != icol && irow % GW == w) {
      if (h0) a[irow * K + c0] = cv0;
      if (h1) a[irow * K + c1] = cv1;
    }
    big = 0.0;
    bidx = -1;
    // the elimination's loads: the pivot column's lane reads the zero slot
    // (its entries become 0 - pn * dum)
    const bool ic0 = c0 == icol, ic1 = TWO && c1 == icol;
    const unsigned lrs0 = ic0 ? 0u : rs0, lrs1 = ic1 ? 0u : rs1;
    char* const le0 = ic0 ? zero : e0;
    char* const le1 = ic1 ? zero : e1;
    // eliminate this wave's rows
    for (int l0 = w0; l0 < mf; l0 += 16 * GW) {
      double b0[16], b1[16];
      // lane u of each 16 holds row l0 + u*GW's pivot-column entry (rows
      // past the last read the dummy row), broadcast row by row
      const int lr16 = min(l0 + (lane & 15) * GW, mf);
      const double dmv = a[lr16 * K + icol];
      // rows of the batch still unpivoted, as bits
      const unsigned cand16 = (unsigned)__builtin_amdgcn_ballot_w64(!used(min(lr16, mf - 1)));
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int lr = min(l0 + u * GW, mf);
        b0[u] = *reinterpret_cast<const double*>(le0 + __umul24(lr, lrs0));
        if constexpr (TWO) b1[u] = *reinterpret_cast<const double*>(le1 + __umul24(lr, lrs1));
      }
      // the batch's rows to eliminate (below mf, not the pivot row), as bits
      const int nrow = (mf - l0 + GW - 1) / GW;
      unsigned todo = nrow >= 16 ? 0xFFFFu : (1u << nrow) - 1u;
      if (icol >= l0 && icol < l0 + 16 * GW && (icol - l0) % GW == 0) todo &= ~(1u << ((icol - l0) / GW));
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        if (!((todo >> u) & 1u)) continue;
        const int ll = l0 + u * GW;
        const double dum = readlane_d(dmv, u);
        const bool cand = (cand16 >> u) & 1u;
        {
          const double base = b0[u];
          const double nv = fma(-pn0, dum, base);
          *reinterpret_cast<double*>(e0 + __umul24(ll, rs0)) = nv;
          consider(cand & open0, nv, ll * 128 + c0);
        }
        if constexpr (TWO) {
          const double base = b1[u];
          const double nv = fma(-pn1, dum, base);
          *reinterpret_cast<double*>(e1 + __umul24(ll, rs1)) = nv;
          consider(cand & open1, nv, ll * 128 + c1);
        }
      }
    }
  }
  return 0;
}

struct PairState {
  double alamda, chisq, ochisq, iter_diff, dev_old, dev_new, chck;
  unsigned long long nvis;  // visible (point, facet) pairs over all evaluations
  int niter;
};

}  // namespace

// ---------------------------------------------------------------------------
// Persistent LM fit kernel.
// ---------------------------------------------------------------------------
namespace {

// The whole LM fit of work item `item` on problem P (Start.cu:4346-5248):
// initialisation, up to n_iter_max iterations of mrqmin + Iter2, FinishPole.
template <int NB, int NBS>
__device__ __forceinline__ void fit_pair(const PsFitArgs& A, const PsDevProblem& P, const Lds& L, PairState& S,
                                         double* slab, int item) {
  const int tid = threadIdx.x;
  const int mf = A.P0.mfit, K = mf + 1, C = A.P0.ncoef, ma = A.P0.ma, F = A.P0.nfac;
  const Slab SL = make_slab(P, slab);
  double* best = SL.best;  // [K*K] accepted alpha | beta | chi^2
  const double conw_r2_rows = A.item_cw[2 * item], conw_r_dev = A.item_cw[2 * item + 1];
  const int n = A.item_n[item], m = A.item_m[item];  // 1-based
  // pair initialisation (Start.cu:4346-4418)
  if (tid == 0) {
    const double period = 1.0 / (P.f_start - (double)(n - 1) * P.f_step);
    const double d2r = kPi / 180.0;
    L.cgg[C] = d2r * 90 - d2r * P.beta_pole[m - 1];
    L.cgg[C + 1] = d2r * P.lambda_pole[m - 1];
    L.cgg[C + 2] = (24.0 * 2.0 * kPi) / period;
    if (A.item_a0) {
      const double* a0 = A.item_a0 + (size_t)item * ma;
      L.cgg[C + 3] = a0[C + 3];
      L.cgg[C + 4] = a0[C + 4];
      L.cgg[C + 5] = a0[C + 5];
      L.cgg[C + 6] = a0[C + 6];
    } else {
      L.cgg[C + 3] = P.par[0];
      L.cgg[C + 4] = P.par[1];
      L.cgg[C + 5] = P.par[2];
      L.cgg[C + 6] = P.lcl;
    }
    L.cgg[C + 7] = 1;
    S.alamda = -1;
    S.niter = 0;
    S.iter_diff = 1e40;
    S.dev_old = 1e30;
    S.dev_new = 0;
    S.chck = 0;
    S.chisq = 0;
    S.ochisq = 0;
    S.nvis = 0;
  }
  for (int i = tid; i < C; i += WG) L.cgg[i] = A.item_a0 ? A.item_a0[(size_t)item * ma + i] : P.cg_first[i];
  __syncthreads();
  for (;;) {
    // Iter1Begin (Start.cu:4421-4462)
    const int niter = S.niter;
    const int n_iter_max = A.n_iter_fixed > 0 ? A.n_iter_fixed : P.n_iter_max;
    const double iter_diff_max = A.n_iter_fixed > 0 ? 0.0 : P.iter_diff_max;
    const int n_iter_min = A.n_iter_fixed > 0 ? 0 : P.n_iter_min;
    const bool is_niter = ((niter < n_iter_max) && (S.iter_diff > iter_diff_max)) || (niter < n_iter_min);
    if (!is_niter) break;
    const bool is_alambda = S.alamda < 0;
    __syncthreads();
    if (is_alambda && tid == 0) S.alamda = P.alamda_start;
    // pass 0 (first iteration only): normal equations at cgg;
    // pass 1: damped solve (mrqmin step 1, Start.cu:2838-3021), then the
    // normal equations at the trial point atry.
#pragma unroll 1
    for (int pass = is_alambda ? 0 : 1; pass < 2; ++pass) {
      if (pass == 1) {
        PS_TSTART();
        // [alpha (diagonal * (1 + lambda)) | beta]
        const double ccc = 1 + S.alamda;
        for (int e = tid; e < mf * K; e += WG) {
          const int i = e / K, j = e - i * K;
          double vv;
          if (j < mf) {
            vv = best[i * K + j];
            if (i == j) vv *= ccc;
          } else {
            vv = best[mf * K + i];
          }
          L.work[e] = vv;
        }
        __syncthreads();
        // the solve is latency-bound: its waves issue first while the other
        // workgroups on the CU run their facet passes (only the arbitration
        // order changes)
        __builtin_amdgcn_s_setprio(2);
        const int err = gaussj_block<(NB > 4)>(L.work, mf, K, L.red);
        __builtin_amdgcn_s_setprio(0);
        __syncthreads();
        if (err == 0 || err == 2)
          for (int q = tid; q < mf; q += WG) L.atry[P.cmap[q]] = L.cgg[P.cmap[q]] + L.work[q * K + mf];
        __syncthreads();
        PS_TMARK(7);
      }
      evaluate<NB, NBS>(P, L, pass == 0 ? L.cgg : L.atry, conw_r2_rows, SL, &S.nvis, A.claim_tiles);
      if (pass == 0) {
        for (int e = tid; e < K * K; e += WG) best[e] = L.work[e];
        if (tid == 0) S.ochisq = L.work[mf * K + mf];
        for (int i = tid; i < ma; i += WG) L.atry[i] = L.cgg[i];
        __syncthreads();
      }
    }
    const double chisq_try = L.work[mf * K + mf];
    // mrqmin step 2 (Start.cu:300-392)
    const bool accept = chisq_try < S.ochisq;
    if (accept) {
      for (int e = tid; e < K * K; e += WG) best[e] = L.work[e];
      for (int i = tid; i < ma; i += WG) L.cgg[i] = L.atry[i];
    }
    __syncthreads();
    if (tid == 0) {
      if (accept) {
        S.alamda = S.alamda * P.alamda_incrr;
        S.chisq = chisq_try;
      } else {
        S.alamda = P.alamda_incr * S.alamda;
        S.chisq = S.ochisq;
      }
      S.niter = S.niter + 1;
    }
    __syncthreads();
    // Iter2 (Start.cu:5129-5248)
    PS_TSTART();
    if (S.niter == 1 || S.chisq < S.ochisq) {
      curv_block(P, L.cgg, L.nrm);
      __syncthreads();
      if (tid < 3) {
        double s = 0.0;
        for (int f = 0; f < F; ++f) s = fma(L.nrm[4 * f + 3], L.nrm[4 * f + tid], s);
        L.red[tid] = s;
      }
      __syncthreads();
      if (tid == 0) {
        S.ochisq = S.chisq;
        const double v0 = L.red[0], v1 = L.red[1], v2 = L.red[2];
        const double aa = fma(v2, v2, fma(v1, v1, v0 * v0));
        const double conwr2 = conw_r_dev * conw_r_dev;
        const double rchisq = fma(-aa, conwr2, S.chisq);
        S.dev_new = sqrt(rchisq / (double)P.ndata);
        S.chck = sqrt(aa);
        const double diff = S.dev_old - S.dev_new;
        if (diff > 1e-10) {
          S.iter_diff = diff;
          S.dev_old = S.dev_new;
        }
      }
    }
    __syncthreads();
    PS_TMARK(8);
  }
  // FinishPole (Start.cu:5056-5125)
  if (tid == 0) {
    double* o = A.out + (size_t)item * PS_OUT_STRIDE;
    double dev = 1e40, dark = 0, per = 0, la = 0, be = 0;
    if (!(S.dev_new >= 1e40)) {
      double tot = 0.0;
      for (int f = 0; f < F; ++f) tot = tot + L.nrm[4 * f + 3];
      tot = 1.0 / tot;
      double dd = (S.chck * 100.0) * tot;
      if (dd != dd) dd = 1.0;
      dev = S.dev_new;
      dark = dd;
      per = (2 * kPi) / L.cgg[C + 2];
      const double r2d = 180.0 / kPi;
      const double lt = r2d * L.cgg[C + 1];
      la = lt + (lt < 0 ? 360 : 0);
      be = 90 - r2d * L.cgg[C];
    }
    o[0] = dev;
    o[1] = dark;
    o[2] = per;
    o[3] = la;
    o[4] = be;
    o[5] = S.chisq;
    o[6] = (double)S.niter;
    o[7] = 1.0;
    o[8] = (double)S.nvis;
  }
  if (A.out_a)
    for (int i = tid; i < ma; i += WG) A.out_a[(size_t)item * ma + i] = L.cgg[i];
  __syncthreads();
}

}  // namespace

// ---------------------------------------------------------------------------
// Persistent LM fit kernel: workgroups take work items from an atomic queue
// until it is empty (BATCH: items of several problems, described in global
// memory; otherwise every item is on problem P0, passed as an argument).
// ---------------------------------------------------------------------------
// Register budget: the resident wavefronts per SIMD the allocation targets.
// Up to 64 Jacobian columns a 4-wave fit fits in 168 VGPRs (three
// workgroups per CU; the few spills are in per-iteration bookkeeping, not in
// the facet, SYRK or Gauss-Jordan loops) and an 8-wave fit in 128 (two
// workgroups per CU, four waves per SIMD: each wave holds half of its
// chain's block pairs); wider fits take up to 256 (two waves per SIMD).
// PS_FIT_WAVES_PER_EU overrides it.
template <int NB>
constexpr int fit_waves_per_eu() {
#ifdef PS_FIT_WAVES_PER_EU
  return PS_FIT_WAVES_PER_EU;
#else
  return NB <= 4 ? (NW == 4 ? 3 : NW == 8 ? 4 : 2) : 2;  // 1 and 2 waves per pair: LDS limits the workgroups per CU
#endif
}

template <int NB, int NBS, bool BATCH>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(fit_waves_per_eu<NB>()))) void ps_fit_kernel(
    PsFitArgs A) {
  extern __shared__ double lds_dyn[];
  __shared__ PairState S;
  __shared__ int item_s;
  // geometry and basis shared by every problem of the launch
  const Lds L = carve_lds(lds_dyn, A.P0);
  stage_tables(A.P0, L);
  double* const slab = A.scratch + (size_t)blockIdx.x * A.scratch_stride;
  for (;;) {
    if (threadIdx.x == 0) item_s = atomicAdd(A.counter, 1);
    __syncthreads();
    const int item = item_s;
    __syncthreads();
    if (item >= A.n_items) break;
    if constexpr (BATCH)
      fit_pair<NB, NBS>(A, A.probs[A.item_p[item]], L, S, slab, item);
    else
      fit_pair<NB, NBS>(A, A.P0, L, S, slab, item);
  }
}

// ---------------------------------------------------------------------------
// Single evaluation / Gauss-Jordan / math kernels for the parity tests.
// ---------------------------------------------------------------------------
template <int NB, int NBS>
__global__ __launch_bounds__(WG) void ps_eval_kernel(PsFitArgs A, const double* a, double* out_work,
                                                     double* out_area, double* out_j) {
  extern __shared__ double lds_dyn[];
  const PsDevProblem& P = A.P0;
  const Lds L = carve_lds(lds_dyn, P);
  stage_tables(P, L);
  const int tid = threadIdx.x, K = P.mfit + 1;
  const Slab SL = make_slab(P, A.scratch);
  for (int i = tid; i < P.ma; i += WG) L.cgg[i] = a[i];
  __syncthreads();
  evaluate<NB, NBS>(P, L, L.cgg, A.conw_r2_rows, SL, nullptr, 1);
  for (int e = tid; e < K * K; e += WG) out_work[e] = L.work[e];
  for (int f = tid; f < P.nfac; f += WG) out_area[f] = L.nrm[4 * f + 3];
  if (out_j != nullptr) {
    // raw rows in the reference's layout: column 0 = model brightness,
    // column c = d/d(parameter c); column 1, the size scale, is the model
    // brightness itself (every area carries the factor exp(a_1))
    const int W = P.ma + 1;
    for (int e = tid; e < (P.ndata + 3) * W; e += WG) {
      const int j = e / W, col = e - j * W;
      double val;
      if (col <= 1)
        val = SL.ybuf[j];
      else
        val = (col - 2 < P.qs) ? SL.jbuf[(size_t)j * P.qs + (col - 2)] : 0.0;
      out_j[e] = val;
    }
  }
}

// Gauss-Jordan of one damped system (the parity tests' entry).
__global__ __launch_bounds__(WG) void ps_gaussj_kernel(int mf, const double* alpha, const double* beta, double lambda,
                                                       double* x, int* err) {
  extern __shared__ double lds_dyn[];
  const int K = mf + 1;
  double* work = lds_dyn;
  const int tid = threadIdx.x;
  const double ccc = 1 + lambda;
  for (int e = tid; e < mf * K; e += WG) {
    const int i = e / K, j = e - i * K;
    double v;
    if (j < mf) {
      v = alpha[i * mf + j];
      if (i == j) v *= ccc;
    } else {
      v = beta[i];
    }
    work[e] = v;
  }
  __syncthreads();
  const int e = K > 64 ? gaussj_block<true>(work, mf, K, work + K * K) : gaussj_block<false>(work, mf, K, work + K * K);
  __syncthreads();
  for (int j = tid; j < mf; j += WG) x[j] = work[j * K + mf];
  if (tid == 0) *err = e;
}

__global__ void ps_math_kernel(int n, const double* x, double* out, double* out2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s, c;
  psm::sincos(x[i], &s, &c);
  out[3 * i + 0] = psm::exp(x[i]);
  out[3 * i + 1] = psm::exp2(x[i]);
  out[3 * i + 2] = s;
  out2[i] = c;
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
size_t ps_fit_lds_bytes(const PsDevProblem& P) {
  const int K = P.mfit + 1;
  const size_t d = (size_t)(K * K > NW * 256 ? K * K : NW * 256) + 4 * (P.nfac + PS_FACET_PAD) + 2 * P.ma +
                   2 * NW + 16;
  return d * sizeof(double) + 64;
}

// (NB, NBS) instantiations: NB = 16-column blocks of a Jacobian row, NBS =
// blocks holding shape columns (NBS in {NB-1, NB}).
#define PS_FOR_EACH_SHAPE(X) X(1, 1) X(2, 1) X(2, 2) X(3, 2) X(3, 3) X(4, 3) X(4, 4) X(5, 4) X(5, 5) X(6, 5) X(6, 6)

hipError_t ps_launch_fit(const PsFitArgs& A, int grid, hipStream_t s) {
  const size_t lds = ps_fit_lds_bytes(A.P0);
#define PS_L(NB_, NBS_)                                                                                       \
  if (A.P0.nb == NB_ && A.P0.nbs == NBS_) {                                                                   \
    if (A.batch)                                                                                              \
      hipLaunchKernelGGL(HIP_KERNEL_NAME(ps_fit_kernel<NB_, NBS_, true>), dim3(grid), dim3(WG), lds, s, A);   \
    else                                                                                                      \
      hipLaunchKernelGGL(HIP_KERNEL_NAME(ps_fit_kernel<NB_, NBS_, false>), dim3(grid), dim3(WG), lds, s, A);  \
    return hipGetLastError();                                                                                 \
  }
  PS_FOR_EACH_SHAPE(PS_L)
#undef PS_L
  return hipErrorInvalidValue;
}

hipError_t ps_launch_eval(const PsFitArgs& A, hipStream_t s, const double* a, double* work, double* area,
                          double* jrows) {
  const size_t lds = ps_fit_lds_bytes(A.P0);
#define PS_L(NB_, NBS_)                                                                                        \
  if (A.P0.nb == NB_ && A.P0.nbs == NBS_) {                                                                    \
    hipLaunchKernelGGL(HIP_KERNEL_NAME(ps_eval_kernel<NB_, NBS_>), dim3(1), dim3(WG), lds, s, A, a, work, area, \
                       jrows);                                                                                 \
    return hipGetLastError();                                                                                  \
  }
  PS_FOR_EACH_SHAPE(PS_L)
#undef PS_L
  return hipErrorInvalidValue;
}

hipError_t ps_phase_cycles(unsigned long long* out, bool reset) {
#ifdef PS_PHASE_TIMING
  hipError_t e = hipMemcpyFromSymbol(out, HIP_SYMBOL(g_phase_cycles), sizeof(unsigned long long) * 16);
  if (e == hipSuccess && reset) {
    unsigned long long z[16] = {0};
    e = hipMemcpyToSymbol(HIP_SYMBOL(g_phase_cycles), z, sizeof z);
  }
  return e;
#else
  (void)reset;
  for (int i = 0; i < 16; ++i) out[i] = 0;
  return hipSuccess;
#endif
}

hipError_t ps_launch_gaussj(int mf, const double* alpha, const double* beta, double lambda, double* x, int* err,
                            hipStream_t s) {
  const int K = mf + 1;
  const size_t lds = ((size_t)K * K + 2 * NW) * sizeof(double);
  hipLaunchKernelGGL(ps_gaussj_kernel, dim3(1), dim3(WG), lds, s, mf, alpha, beta, lambda, x, err);
  return hipGetLastError();
}

hipError_t ps_launch_math(int n, const double* x, double* out, double* out2, hipStream_t s) {
  hipLaunchKernelGGL(ps_math_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, x, out, out2);
  return hipGetLastError();
}

int ps_fit_blocks_per_cu(int nb, int nbs, size_t lds) {
#define PS_O(NB_, NBS_)                                                                                       \
  if (nb == NB_ && nbs == NBS_) {                                                                             \
    int n0 = 0, n1 = 0;                                                                                       \
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n0, reinterpret_cast<const void*>(                      \
                                                         &ps_fit_kernel<NB_, NBS_, false>), WG, lds) != hipSuccess || \
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&n1, reinterpret_cast<const void*>(                      \
                                                         &ps_fit_kernel<NB_, NBS_, true>), WG, lds) != hipSuccess)  \
      return 0;                                                                                               \
    return n0 < n1 ? n0 : n1;                                                                                 \
  }
  PS_FOR_EACH_SHAPE(PS_O)
#undef PS_O
  return 0;
}

hipError_t ps_set_fit_attributes(int nb, int nbs, size_t lds) {
#define PS_A(NB_, NBS_)                                                                                          \
  if (nb == NB_ && nbs == NBS_) {                                                                                \
    const void* k[3] = {reinterpret_cast<const void*>(&ps_fit_kernel<NB_, NBS_, false>),                         \
                        reinterpret_cast<const void*>(&ps_fit_kernel<NB_, NBS_, true>),                          \
                        reinterpret_cast<const void*>(&ps_eval_kernel<NB_, NBS_>)};                              \
    for (const void* f : k) {                                                                                    \
      hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);               \
      if (e != hipSuccess) return e;                                                                             \
    }                                                                                                            \
    return hipSuccess;                                                                                           \
  }
  PS_FOR_EACH_SHAPE(PS_A)
#undef PS_A
  return hipErrorInvalidValue;
}

}  // namespace PS_FIT_NS
