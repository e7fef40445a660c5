// ed_adjoint.hpp -- gfx950 device code for the adjoint (reverse mode) of
// elemental generation, FP64 throughout.
//
// Forward (per timeslice, lattice/generator/elemental.py:279-338):
//   W_n = nabla_{d_k} ... nabla_{d_1} V,
//   E[d,p,e,f] = w[e,f] sum_{s in splits(d)} c_s sum_{x,c} Phi_p(x) conj(W_{L_s}[e,x,c]) W_{R_s}[f,x,c].
// With G = dL/dRe(E) + i dL/dIm(E) (so dL = Re sum conj(G) dE) and the same
// convention for every input gradient, the reverse pass is
//   Gamma_d[e,f,x] = sum_p w[e,f] G[d,p,e,f] conj(Phi_p(x))
//   gW_{L_s}[e,x,c] += c_s sum_f conj(Gamma_d[e,f,x]) W_{R_s}[f,x,c]
//   gW_{R_s}[f,x,c] += c_s sum_e Gamma_d[e,f,x]       W_{L_s}[e,x,c]
// then, children before parents (W_child = nabla_mu W_parent), with
// nabla_mu^dagger = -nabla_mu on the periodic lattice,
//   gW_parent += -nabla_mu gW_child
//   gU_mu(x)  += sum_e [ gW_child[e,x] (x) conj(W_parent[e,x+mu])
//                       - W_parent[e,x] (x) conj(gW_child[e,x+mu]) ]
// and dL/dV = gW_0.  The displaced vectors W are recomputed by the forward
// staging and hop kernels, not stored between the passes.
#pragma once
#include "ed_kernels.hpp"

namespace ed {
namespace adj {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // a * conj(b)
  return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// --------------------------------------------------------------------------
// Weighted cotangent in the two orders the contraction adjoint reads:
//   Acol[m][f][r] = w[f][r] G[m][f][r],   Arow[m][f][r] = w[r][f] G[m][r][f]
// for every m = (tb, d, p) matrix of Ne x Ne.
// grid (ceil(Ne/32), ceil(Ne/32), nmat), block (32, 8)
// --------------------------------------------------------------------------
__global__ __launch_bounds__(256) void adj_prep_kernel(const double2* __restrict__ G, const double* __restrict__ wgt,
                                                       double2* __restrict__ Arow, double2* __restrict__ Acol, int Ne) {
  __shared__ double2 tile[32][33];
  const size_t off = (size_t)blockIdx.z * Ne * Ne;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int r0 = blockIdx.x * 32, f0 = blockIdx.y * 32;
  for (int j = ty; j < 32; j += 8) {
    const int f = f0 + j, r = r0 + tx;
    if (f < Ne && r < Ne) {
      double2 v = G[off + (size_t)f * Ne + r];
      const double w = wgt ? wgt[(size_t)f * Ne + r] : 1.0;
      Acol[off + (size_t)f * Ne + r] = make_double2(v.x * w, v.y * w);
    }
    // tile of G rows r0.., columns f0..
    const int rr = r0 + j, ff = f0 + tx;
    double2 v = make_double2(0.0, 0.0);
    if (rr < Ne && ff < Ne) {
      v = G[off + (size_t)rr * Ne + ff];
      const double w = wgt ? wgt[(size_t)rr * Ne + ff] : 1.0;
      v = make_double2(v.x * w, v.y * w);
    }
    tile[j][tx] = v;
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int f = f0 + j, r = r0 + tx;
    if (f < Ne && r < Ne) Arow[off + (size_t)f * Ne + r] = tile[tx][j];
  }
}

struct AdjArgs {
  Geom g;
  int D, P;
  const SplitTerm* splits;  // merged split terms of every derivative
  const int2* dsplit;       // per derivative: (first, count)
  const double2* phase;     // [V3][P] Phi_p(x)
};

// --------------------------------------------------------------------------
// Contraction adjoint.  A wave owns one site; lane r owns row r, so every
// gW[n][x][c][r] it touches is written by this thread alone (no atomics):
// as a left row it gathers sum_f conj(Gamma[r,f]) W_R[f], as a right row
// sum_e Gamma[e,r] W_L[e].  NS bounds the split count of one derivative.
// grid (ceil(V3 / 8), ceil(Ne / 64), TB), block 512
// --------------------------------------------------------------------------
template <int NS>
__global__ __launch_bounds__(512) void adj_contract_kernel(const double2* __restrict__ W, double2* __restrict__ gW,
                                                           const double2* __restrict__ Arow,
                                                           const double2* __restrict__ Acol, AdjArgs a) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int s = blockIdx.x * 8 + wave;
  const int r = blockIdx.y * 64 + lane;
  const int tb = blockIdx.z;
  if (s >= a.g.V3) return;  // whole wave: no barriers below
  const int Ne = a.g.Ne;
  const bool live = r < Ne;
  const int rc = live ? r : 0;
  const size_t slab = (size_t)a.g.nW * a.g.V3P * 3 * a.g.NeP;
  const double2* Wt = W + (size_t)tb * slab;
  double2* gWt = gW + (size_t)tb * slab;
  const size_t ne2 = (size_t)Ne * Ne;
  const double2* ph = a.phase + (size_t)s * a.P;
  for (int d = 0; d < a.D; ++d) {
    const int2 fs = a.dsplit[d];
    int nL[NS], nR[NS];
    double cf[NS];
    double2 accL[NS][3], accR[NS][3];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      if (k < fs.y) {
        const SplitTerm st = a.splits[fs.x + k];
        nL[k] = st.nL;
        nR[k] = st.nR;
        cf[k] = (double)st.coeff;
      } else {
        nL[k] = nR[k] = 0;
        cf[k] = 0.0;
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) accL[k][c] = accR[k][c] = make_double2(0.0, 0.0);
    }
    const double2* Ar = Arow + ((size_t)tb * a.D + d) * a.P * ne2 + rc;
    const double2* Ac = Acol + ((size_t)tb * a.D + d) * a.P * ne2 + rc;
    for (int f = 0; f < Ne; ++f) {
      double2 gr = make_double2(0.0, 0.0), gc = make_double2(0.0, 0.0);
      for (int p = 0; p < a.P; ++p) {
        const double2 phi = ph[p];
        const size_t o = (size_t)p * ne2 + (size_t)f * Ne;
        const double2 x1 = Ar[o], x2 = Ac[o];
        const double2 t1 = cmulc(x1, phi), t2 = cmulc(x2, phi);
        gr.x += t1.x;
        gr.y += t1.y;
        gc.x += t2.x;
        gc.y += t2.y;
      }
      const double2 grc = make_double2(gr.x, -gr.y);
#pragma unroll
      for (int k = 0; k < NS; ++k) {
        if (k < fs.y) {
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const double2 wr = Wt[widx(a.g, nR[k], s, c, f)];
            const double2 wl = Wt[widx(a.g, nL[k], s, c, f)];
            const double2 u = cmul(grc, wr), v = cmul(gc, wl);
            accL[k][c].x += cf[k] * u.x;
            accL[k][c].y += cf[k] * u.y;
            accR[k][c].x += cf[k] * v.x;
            accR[k][c].y += cf[k] * v.y;
          }
        }
      }
    }
    if (live) {
#pragma unroll
      for (int k = 0; k < NS; ++k) {
        if (k < fs.y) {
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            double2* pl = gWt + widx(a.g, nL[k], s, c, r);
            double2 v = *pl;
            *pl = make_double2(v.x + accL[k][c].x, v.y + accL[k][c].y);
            double2* pr = gWt + widx(a.g, nR[k], s, c, r);
            v = *pr;
            *pr = make_double2(v.x + accR[k][c].x, v.y + accR[k][c].y);
          }
        }
      }
    }
  }
}

__device__ __forceinline__ void neighbours(const Geom& g, int s, int mu, int& sf, int& sb) {
  const int x = s % g.Lx, y = (s / g.Lx) % g.Ly, z = s / (g.Lx * g.Ly);
  if (mu == 0) {
    sf = s - x + (x + 1 == g.Lx ? 0 : x + 1);
    sb = s - x + (x == 0 ? g.Lx - 1 : x - 1);
  } else if (mu == 1) {
    const int r = s - y * g.Lx;
    sf = r + (y + 1 == g.Ly ? 0 : y + 1) * g.Lx;
    sb = r + (y == 0 ? g.Ly - 1 : y - 1) * g.Lx;
  } else {
    const int r = s - z * g.Lx * g.Ly;
    sf = r + (z + 1 == g.Lz ? 0 : z + 1) * g.Lx * g.Ly;
    sb = r + (z == 0 ? g.Lz - 1 : z - 1) * g.Lx * g.Ly;
  }
}

// --------------------------------------------------------------------------
// Hop adjoint for one derivative level: for every parent slot,
//   gW_parent(x) -= sum_mu [U_mu(x) gW_child(x+mu) - U_mu(x-mu)^dag gW_child(x-mu)]
// child = child0 + 3 (parent - parent0) + mu.  Wave per site, lanes over
// eigenvectors; each (parent, x, e) has one writer.
// grid (ceil(V3 / 4), nparent, TB), block 256
// --------------------------------------------------------------------------
__global__ __launch_bounds__(256) void adj_hop_kernel(double2* __restrict__ gW, const double2* __restrict__ U, Geom g,
                                                      int tb0, int parent0, int child0) {
  const int s = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (s >= g.V3) return;
  const int pl = blockIdx.y, parent = parent0 + pl;
  const int tb = blockIdx.z;
  double2* gWt = gW + (size_t)tb * g.nW * (size_t)g.V3P * 3 * g.NeP;
  for (int e = lane; e < g.Ne; e += 64) {
    double ar[3] = {0, 0, 0}, ai[3] = {0, 0, 0};
#pragma unroll 1
    for (int mu = 0; mu < 3; ++mu) {
      int sf, sb;
      neighbours(g, s, mu, sf, sb);
      const int child = child0 + 3 * pl + mu;
      const double2* Um = U + ((size_t)mu * g.Lt + (tb0 + tb)) * (size_t)g.V3 * 9;
      double2 cf[3], cb[3];
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        cf[b] = gWt[widx(g, child, sf, b, e)];
        cb[b] = gWt[widx(g, child, sb, b, e)];
      }
#pragma unroll
      for (int a2 = 0; a2 < 3; ++a2) {
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          const double2 u = Um[(size_t)s * 9 + a2 * 3 + b];
          ar[a2] += u.x * cf[b].x - u.y * cf[b].y;
          ai[a2] += u.x * cf[b].y + u.y * cf[b].x;
          const double2 v = Um[(size_t)sb * 9 + b * 3 + a2];  // (U^dag)_{a b} = conj(U_{b a})
          ar[a2] -= v.x * cb[b].x + v.y * cb[b].y;
          ai[a2] -= v.x * cb[b].y - v.y * cb[b].x;
        }
      }
    }
#pragma unroll
    for (int a2 = 0; a2 < 3; ++a2) {
      double2* p = gWt + widx(g, parent, s, a2, e);
      const double2 v = *p;
      *p = make_double2(v.x - ar[a2], v.y - ai[a2]);
    }
  }
}

// --------------------------------------------------------------------------
// Link gradient: for every (mu, x) the sum over hop pairs (parent p, child
// c = (p, mu)) and eigenvectors e of
//   gW_c[e,x,a] conj(W_p[e,x+mu,b]) - W_p[e,x,a] conj(gW_c[e,x+mu,b]).
// Parents are the slots below the last level; child_of[3 p + mu] gives c.
// Output gU[mu][tl][x][3][3] with tl = tout0 + tb and ntout timeslices.
// grid (ceil(V3 / 4), 3, TB), block 256: wave per (x, mu), lanes over e,
// a butterfly reduction over the wave at the end.
// --------------------------------------------------------------------------
__global__ __launch_bounds__(256) void adj_link_kernel(const double2* __restrict__ W, const double2* __restrict__ gW,
                                                       const int* __restrict__ child_of, int nparent,
                                                       double2* __restrict__ gU, Geom g, int tout0, int ntout) {
  const int s = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int mu = blockIdx.y, tb = blockIdx.z;
  if (s >= g.V3) return;
  int sf, sb;
  neighbours(g, s, mu, sf, sb);
  const size_t slab = (size_t)g.nW * g.V3P * 3 * g.NeP;
  const double2* Wt = W + (size_t)tb * slab;
  const double2* gWt = gW + (size_t)tb * slab;
  double accr[9], acci[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) accr[k] = acci[k] = 0.0;
  for (int p = 0; p < nparent; ++p) {
    const int c = child_of[3 * p + mu];
    for (int e = lane; e < g.Ne; e += 64) {
      double2 gc[3], wpf[3], wp[3], gcf[3];
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        gc[b] = gWt[widx(g, c, s, b, e)];
        gcf[b] = gWt[widx(g, c, sf, b, e)];
        wp[b] = Wt[widx(g, p, s, b, e)];
        wpf[b] = Wt[widx(g, p, sf, b, e)];
      }
#pragma unroll
      for (int a2 = 0; a2 < 3; ++a2)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          const double2 t1 = cmulc(gc[a2], wpf[b]);
          const double2 t2 = cmulc(wp[a2], gcf[b]);
          accr[a2 * 3 + b] += t1.x - t2.x;
          acci[a2 * 3 + b] += t1.y - t2.y;
        }
    }
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      accr[k] += __shfl_xor(accr[k], off, 64);
      acci[k] += __shfl_xor(acci[k], off, 64);
    }
  }
  if (lane == 0) {
    double2* o = gU + (((size_t)mu * ntout + (tout0 + tb)) * g.V3 + s) * 9;
#pragma unroll
    for (int k = 0; k < 9; ++k) o[k] = make_double2(accr[k], acci[k]);
  }
}

// --------------------------------------------------------------------------
// dL/dV: gV[tl][e][m] = gW_0[m][e] (m = 3 x + c) for tl = tout0 + tb,
// a 32 x 32 tile transpose through LDS.
// grid (ceil(3 V3 / 32), ceil(Ne / 32), TB), block (32, 8)
// --------------------------------------------------------------------------
__global__ __launch_bounds__(256) void adj_out_v_kernel(const double2* __restrict__ gW, double2* __restrict__ gV,
                                                        Geom g, int tout0) {
  __shared__ double2 tile[32][33];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int m0 = blockIdx.x * 32, e0 = blockIdx.y * 32, tb = blockIdx.z;
  const int M = 3 * g.V3;
  const double2* gWt = gW + (size_t)tb * g.nW * (size_t)g.V3P * 3 * g.NeP;  // slot 0
  for (int j = ty; j < 32; j += 8) {
    const int m = m0 + j, e = e0 + tx;
    tile[j][tx] = (m < M && e < g.Ne) ? gWt[(size_t)m * g.NeP + e] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  double2* o = gV + (size_t)(tout0 + tb) * g.Ne * (size_t)M;
  for (int j = ty; j < 32; j += 8) {
    const int e = e0 + j, m = m0 + tx;
    if (e < g.Ne && m < M) o[(size_t)e * M + m] = tile[tx][j];
  }
}

}  // namespace adj
}  // namespace ed
