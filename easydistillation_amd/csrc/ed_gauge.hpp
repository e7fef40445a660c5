// ed_gauge.hpp -- link-field preprocessing on gfx950: unitary projection and
// 3-D stout smearing of the spatial links U[3][Lt][Lz][Ly][Lx][3][3]
// (complex128), one thread per link.
//
// Both run once per configuration before elemental generation
// (lattice/generator/elemental.py:107-117 project_SU3, :175-242,264-277 stout_smear)
// and are memory-bound SU(3) algebra in FP64.  Stout smearing follows
// Morningstar and Peardon, Phys. Rev. D 69 (2004) 054501: the staple sum C,
// Omega = rho C U^dag, Q = (i/2)(Omega^dag - Omega) made traceless, and
// exp(iQ) = f0 + f1 Q + f2 Q^2 by Cayley-Hamilton (their eqs. 29-35).
#pragma once
#include <hip/hip_runtime.h>

namespace ed {
namespace gauge {

struct M3 {
  double2 a[9];
};

__device__ __forceinline__ double2 cadd(double2 x, double2 y) { return make_double2(x.x + y.x, x.y + y.y); }
__device__ __forceinline__ double2 csub(double2 x, double2 y) { return make_double2(x.x - y.x, x.y - y.y); }
__device__ __forceinline__ double2 cmul(double2 x, double2 y) {
  return make_double2(x.x * y.x - x.y * y.y, x.x * y.y + x.y * y.x);
}
__device__ __forceinline__ double2 cmulc(double2 x, double2 y) {  // x * conj(y)
  return make_double2(x.x * y.x + x.y * y.y, x.y * y.x - x.x * y.y);
}
__device__ __forceinline__ double2 ccmul(double2 x, double2 y) {  // conj(x) * y
  return make_double2(x.x * y.x + x.y * y.y, x.x * y.y - x.y * y.x);
}
__device__ __forceinline__ double2 cscale(double2 x, double s) { return make_double2(x.x * s, x.y * s); }

__device__ __forceinline__ M3 load_m3(const double2* __restrict__ p) {
  M3 m;
#pragma unroll
  for (int i = 0; i < 9; ++i) m.a[i] = p[i];
  return m;
}
__device__ __forceinline__ void store_m3(double2* __restrict__ p, const M3& m) {
#pragma unroll
  for (int i = 0; i < 9; ++i) p[i] = m.a[i];
}
// A B
__device__ __forceinline__ M3 mm(const M3& A, const M3& B) {
  M3 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double2 s = cmul(A.a[3 * i], B.a[j]);
      s = cadd(s, cmul(A.a[3 * i + 1], B.a[3 + j]));
      s = cadd(s, cmul(A.a[3 * i + 2], B.a[6 + j]));
      r.a[3 * i + j] = s;
    }
  return r;
}
// A B^dag
__device__ __forceinline__ M3 mmd(const M3& A, const M3& B) {
  M3 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double2 s = cmulc(A.a[3 * i], B.a[3 * j]);
      s = cadd(s, cmulc(A.a[3 * i + 1], B.a[3 * j + 1]));
      s = cadd(s, cmulc(A.a[3 * i + 2], B.a[3 * j + 2]));
      r.a[3 * i + j] = s;
    }
  return r;
}
// A^dag B
__device__ __forceinline__ M3 mdm(const M3& A, const M3& B) {
  M3 r;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double2 s = ccmul(A.a[i], B.a[j]);
      s = cadd(s, ccmul(A.a[3 + i], B.a[3 + j]));
      s = cadd(s, ccmul(A.a[6 + i], B.a[6 + j]));
      r.a[3 * i + j] = s;
    }
  return r;
}
__device__ __forceinline__ void acc(M3& A, const M3& B) {
#pragma unroll
  for (int i = 0; i < 9; ++i) A.a[i] = cadd(A.a[i], B.a[i]);
}

// Periodic site offset helper for the spatial lattice of one timeslice.
struct Lat {
  int Lx, Ly, Lz, Lt;
  __device__ __forceinline__ size_t link(int mu, int t, int x, int y, int z) const {
    return ((((size_t)mu * Lt + t) * Lz + z) * Ly + y) * Lx + x;
  }
};

__device__ __forceinline__ void step(int dir, int sgn, int& x, int& y, int& z, const Lat& L) {
  if (dir == 0) x = (x + sgn + L.Lx) % L.Lx;
  else if (dir == 1) y = (y + sgn + L.Ly) % L.Ly;
  else z = (z + sgn + L.Lz) % L.Lz;
}

// exp(iQ) for traceless Hermitian Q by Cayley-Hamilton (Morningstar-Peardon
// eqs. 29-35).  c0 = det Q = tr(Q^3)/3, c1 = tr(Q^2)/2; negative c0 uses
// f_j(-c0) = (-1)^j conj(f_j(c0)).  Q = 0 (c1 = 0) gives the identity limit.
__device__ __forceinline__ M3 expi_traceless_hermitian(const M3& Q) {
  const M3 Q2 = mm(Q, Q);
  double c1 = 0.5 * (Q2.a[0].x + Q2.a[4].x + Q2.a[8].x);
  double c0 = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) c0 += Q.a[3 * i + k].x * Q2.a[3 * k + i].x - Q.a[3 * i + k].y * Q2.a[3 * k + i].y;
  c0 *= (1.0 / 3.0);
  double2 f0, f1, f2;
  if (!(c1 > 0.0)) {
    f0 = make_double2(1.0, 0.0);
    f1 = make_double2(0.0, 0.0);
    f2 = make_double2(0.0, 0.0);
  } else {
    const bool neg = c0 < 0.0;
    const double ac0 = fabs(c0);
    const double c0max = 2.0 * (c1 / 3.0) * sqrt(c1 / 3.0);
    double r = ac0 / c0max;
    r = r > 1.0 ? 1.0 : r;
    const double th = acos(r);
    const double u = sqrt(c1 / 3.0) * cos(th / 3.0);
    const double w = sqrt(c1) * sin(th / 3.0);
    const double u2 = u * u, w2 = w * w;
    double xi0;
    if (fabs(w) > 0.05) xi0 = sin(w) / w;
    else xi0 = 1.0 - w2 / 6.0 * (1.0 - w2 / 20.0 * (1.0 - w2 / 42.0 * (1.0 - w2 / 72.0)));
    const double cw = cos(w);
    const double2 em = make_double2(cos(u), -sin(u));           // e^{-iu}
    const double2 e2 = make_double2(cos(2.0 * u), sin(2.0 * u));  // e^{2iu}
    // h0 = (u^2 - w^2) e^{2iu} + e^{-iu} (8u^2 cos w + 2iu(3u^2 + w^2) xi0)
    const double2 h0 = cadd(cscale(e2, u2 - w2), cmul(em, make_double2(8.0 * u2 * cw, 2.0 * u * (3.0 * u2 + w2) * xi0)));
    // h1 = 2u e^{2iu} - e^{-iu} (2u cos w - i(3u^2 - w^2) xi0)
    const double2 h1 = csub(cscale(e2, 2.0 * u), cmul(em, make_double2(2.0 * u * cw, -(3.0 * u2 - w2) * xi0)));
    // h2 = e^{2iu} - e^{-iu} (cos w + 3iu xi0)
    const double2 h2 = csub(e2, cmul(em, make_double2(cw, 3.0 * u * xi0)));
    const double den = 1.0 / (9.0 * u2 - w2);
    f0 = cscale(h0, den);
    f1 = cscale(h1, den);
    f2 = cscale(h2, den);
    if (neg) {
      f0.y = -f0.y;
      f1.x = -f1.x;
      f2.y = -f2.y;
    }
  }
  M3 R;
#pragma unroll
  for (int i = 0; i < 9; ++i) R.a[i] = cadd(cmul(f1, Q.a[i]), cmul(f2, Q2.a[i]));
  R.a[0] = cadd(R.a[0], f0);
  R.a[4] = cadd(R.a[4], f0);
  R.a[8] = cadd(R.a[8], f0);
  return R;
}

// One stout step: Uout_mu(x) = exp(iQ_mu(x)) Uin_mu(x), mu, nu over the three
// spatial directions, staples within the link's own timeslice.
__global__ __launch_bounds__(256) void stout_step_kernel(const double2* __restrict__ Uin, double2* __restrict__ Uout,
                                                         Lat L, double rho) {
  const size_t V3 = (size_t)L.Lx * L.Ly * L.Lz;
  const size_t n = 3 * (size_t)L.Lt * V3;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = (int)(i % V3);
  const int t = (int)((i / V3) % L.Lt);
  const int mu = (int)(i / (V3 * L.Lt));
  const int x = s % L.Lx, y = (s / L.Lx) % L.Ly, z = s / (L.Lx * L.Ly);
  const M3 U = load_m3(Uin + 9 * i);
  M3 C;
#pragma unroll
  for (int k = 0; k < 9; ++k) C.a[k] = make_double2(0.0, 0.0);
  for (int nu = 0; nu < 3; ++nu) {
    if (nu == mu) continue;
    // U_nu(x) U_mu(x+nu) U_nu(x+mu)^dag
    int ax = x, ay = y, az = z;
    step(nu, +1, ax, ay, az, L);
    int bx = x, by = y, bz = z;
    step(mu, +1, bx, by, bz, L);
    const M3 a1 = load_m3(Uin + 9 * L.link(nu, t, x, y, z));
    const M3 a2 = load_m3(Uin + 9 * L.link(mu, t, ax, ay, az));
    const M3 a3 = load_m3(Uin + 9 * L.link(nu, t, bx, by, bz));
    acc(C, mmd(mm(a1, a2), a3));
    // U_nu(x-nu)^dag U_mu(x-nu) U_nu(x-nu+mu)
    int cx = x, cy = y, cz = z;
    step(nu, -1, cx, cy, cz, L);
    int dx = cx, dy = cy, dz = cz;
    step(mu, +1, dx, dy, dz, L);
    const M3 b1 = load_m3(Uin + 9 * L.link(nu, t, cx, cy, cz));
    const M3 b2 = load_m3(Uin + 9 * L.link(mu, t, cx, cy, cz));
    const M3 b3 = load_m3(Uin + 9 * L.link(nu, t, dx, dy, dz));
    acc(C, mm(mdm(b1, b2), b3));
  }
  // Omega = rho C U^dag;  Q = (i/2)(Omega^dag - Omega) - (1/3) tr(.)
  const M3 Om = mmd(C, U);
  M3 Q;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const double2 o = Om.a[3 * a + b], od = Om.a[3 * b + a];
      // (i/2)(conj(od) - o)
      const double re = od.x - o.x, im = -od.y - o.y;
      Q.a[3 * a + b] = make_double2(-0.5 * rho * im, 0.5 * rho * re);
    }
  const double2 tr = make_double2((Q.a[0].x + Q.a[4].x + Q.a[8].x) / 3.0, (Q.a[0].y + Q.a[4].y + Q.a[8].y) / 3.0);
  Q.a[0] = csub(Q.a[0], tr);
  Q.a[4] = csub(Q.a[4], tr);
  Q.a[8] = csub(Q.a[8], tr);
  const M3 E = expi_traceless_hermitian(Q);
  store_m3(Uout + 9 * i, mm(E, U));
}

// One polar iteration U <- (U + (U^-1)^dag) / 2 of project_SU3
// (elemental.py:107-117).  Also evaluates, on the input U, the reference's
// stopping test max|U - (U^-1)^dag| <= tol and max|U U^dag - 1| <= tol and
// sets *flag when a link fails it.
__global__ __launch_bounds__(256) void polar_step_kernel(const double2* __restrict__ Uc, double2* __restrict__ Un,
                                                         size_t nlinks, double tol, int* __restrict__ flag) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nlinks) return;
  const M3 M = load_m3(Uc + 9 * i);
  // cofactors: inv[a][b] = cof[b][a] / det
  M3 cof;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, q1 = (q + 1) % 3, q2 = (q + 2) % 3;
      cof.a[3 * r + q] = csub(cmul(M.a[3 * r1 + q1], M.a[3 * r2 + q2]), cmul(M.a[3 * r1 + q2], M.a[3 * r2 + q1]));
    }
  const double2 det =
      cadd(cadd(cmul(M.a[0], cof.a[0]), cmul(M.a[1], cof.a[1])), cmul(M.a[2], cof.a[2]));
  const double dn = 1.0 / (det.x * det.x + det.y * det.y);
  const double2 idet = make_double2(det.x * dn, -det.y * dn);
  // (U^-1)^dag [a][b] = conj(inv[b][a]) = conj(cof[a][b] / det)
  M3 Xd;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const double2 v = cmul(cof.a[k], idet);
    Xd.a[k] = make_double2(v.x, -v.y);
  }
  double e1 = 0.0, e2 = 0.0;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const double2 d = csub(M.a[k], Xd.a[k]);
    e1 = fmax(e1, sqrt(d.x * d.x + d.y * d.y));
  }
  const M3 G = mmd(M, M);
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const double2 d = make_double2(G.a[k].x - ((k % 4) == 0 ? 1.0 : 0.0), G.a[k].y);
    e2 = fmax(e2, sqrt(d.x * d.x + d.y * d.y));
  }
  if (e1 > tol || e2 > tol || !(e1 == e1) || !(e2 == e2)) *flag = 1;
  M3 R;
#pragma unroll
  for (int k = 0; k < 9; ++k) R.a[k] = cscale(cadd(M.a[k], Xd.a[k]), 0.5);
  store_m3(Un + 9 * i, R);
}

// 3-D gauge-covariant Laplacian on a block of vectors of timeslice t
// (lattice/generator/eigenvector.py:11-26):
//   (L F)(x) = 6 F(x) - sum_mu [U_mu(x) F(x + mu) + U_mu(x - mu)^dag F(x - mu)]
// F and out [nvec][V3][3] complex128; one thread per (vector, site), sites
// fastest so neighbouring threads read neighbouring 48-byte site vectors.
__global__ __launch_bounds__(256) void laplacian_kernel(const double2* __restrict__ F, double2* __restrict__ out,
                                                        const double2* __restrict__ U, Lat L, int t, int nvec) {
  const size_t V3 = (size_t)L.Lx * L.Ly * L.Lz;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V3 * (size_t)nvec) return;
  const size_t e = i / V3;
  const int s = (int)(i % V3);
  const int x = s % L.Lx, y = (s / L.Lx) % L.Ly, z = s / (L.Lx * L.Ly);
  const double2* Fe = F + e * V3 * 3;
  double2 acc[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) acc[a] = cscale(Fe[(size_t)s * 3 + a], 6.0);
#pragma unroll
  for (int mu = 0; mu < 3; ++mu) {
    int fx = x, fy = y, fz = z, bx = x, by = y, bz = z;
    step(mu, +1, fx, fy, fz, L);
    step(mu, -1, bx, by, bz, L);
    const size_t sf = ((size_t)fz * L.Ly + fy) * L.Lx + fx, sb = ((size_t)bz * L.Ly + by) * L.Lx + bx;
    const double2* uf = U + 9 * L.link(mu, t, x, y, z);
    const double2* ub = U + 9 * L.link(mu, t, bx, by, bz);
    double2 pf[3], pb[3];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      pf[b] = Fe[sf * 3 + b];
      pb[b] = Fe[sb * 3 + b];
    }
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        acc[a] = csub(acc[a], cmul(uf[a * 3 + b], pf[b]));
        acc[a] = csub(acc[a], ccmul(ub[b * 3 + a], pb[b]));  // (U^dag)_{ab} = conj(U_{ba})
      }
  }
  double2* o = out + e * V3 * 3 + (size_t)s * 3;
#pragma unroll
  for (int a = 0; a < 3; ++a) o[a] = acc[a];
}

}  // namespace gauge
}  // namespace ed
