// ed_adjoint.hpp -- gfx950 device code for the adjoint (reverse mode) of
// elemental generation, FP64 throughout.
//
// Forward (per timeslice, lattice/generator/elemental.py:279-338):
//   W_n = nabla_{d_k} ... nabla_{d_1} V,
//   E[d,p,e,f] = w[e,f] sum_{s in splits(d)} c_s sum_{x,c} Phi_p(x) conj(W_{L_s}[e,x,c]) W_{R_s}[f,x,c].
// With G = dL/dRe(E) + i dL/dIm(E) (so dL = Re sum conj(G) dE) and the same
// convention for every input gradient, the reverse pass is
//   Gamma_d[e,f,x] = sum_p w[e,f] G[d,p,e,f] conj(Phi_p(x))   (opposite momenta paired)
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
// Weighted cotangent, momenta grouped, in the two orders the contraction
// adjoint reads.  With A_p = w o G_p and Phi_p = c + i s, a group q is either
// a pair (p1, p2) of opposite momenta, Phi_p2 = conj(Phi_p1), or one momentum
// (p2 < 0), or a padding group (p1 < 0, zero); writing c, s for Re, Im Phi_p1(x),
//   sum_{p in q} A_p conj(Phi_p) = S_q c + D_q s,
//   S_q = A_p1 + A_p2,   D_q = -i (A_p1 - A_p2)     (A_p2 = 0 for one momentum),
// so the phase sum of the adjoint costs two real-scaled terms per group.
// Both orders hold [tb][d][q][S, D][f < Ne][r < NeP] (rows r contiguous,
// zero for r >= Ne):
//   Bcol[..][f][r] from A_p[f][r] = w[f][r] G_p[f][r],
//   Brow[..][f][r] from A_p[r][f] = w[r][f] G_p[r][f].
// grid (ceil(NeP/32), ceil(Ne/32), nz), z = blockIdx.z + zoff < nb D Q; block (32, 8)
// --------------------------------------------------------------------------
__global__ __launch_bounds__(256) void adj_prep_kernel(const double2* __restrict__ G, const double* __restrict__ wgt,
                                                       const int2* __restrict__ groups, double2* __restrict__ Brow,
                                                       double2* __restrict__ Bcol, int Ne, int NeP, int P, int Q,
                                                       int zoff) {
  __shared__ double2 tS[32][33], tD[32][33];
  const size_t ne2 = (size_t)Ne * Ne, bm = (size_t)Ne * NeP;
  const int z = blockIdx.z + zoff, q = z % Q, td = z / Q;  // td = tb D + d
  const int2 gq = groups[q];
  // a padding group (gq.x < 0) has S = D = 0
  const double2* G1 = G + ((size_t)td * P + max(gq.x, 0)) * ne2;
  const double2* G2 = G + ((size_t)td * P + (gq.y >= 0 ? gq.y : max(gq.x, 0))) * ne2;
  const double pair = gq.y >= 0 ? 1.0 : 0.0;
  const double one = gq.x >= 0 ? 1.0 : 0.0;
  double2* oS_c = Bcol + ((size_t)z * 2) * bm;
  double2* oD_c = oS_c + bm;
  double2* oS_r = Brow + ((size_t)z * 2) * bm;
  double2* oD_r = oS_r + bm;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int r0 = blockIdx.x * 32, f0 = blockIdx.y * 32;
  for (int j = ty; j < 32; j += 8) {
    // column order: element (f, r) = (f0 + j, r0 + tx)
    int f = f0 + j, r = r0 + tx;
    if (f < Ne && r < NeP) {
      double2 sm = make_double2(0.0, 0.0), dm = make_double2(0.0, 0.0);
      if (r < Ne) {
        const size_t o = (size_t)f * Ne + r;
        const double w = one * (wgt ? wgt[o] : 1.0);
        const double2 a1 = G1[o], a2 = G2[o];
        sm = make_double2(w * (a1.x + pair * a2.x), w * (a1.y + pair * a2.y));
        dm = make_double2(w * (a1.y - pair * a2.y), -w * (a1.x - pair * a2.x));
      }
      oS_c[(size_t)f * NeP + r] = sm;
      oD_c[(size_t)f * NeP + r] = dm;
    }
    // row order: tile of G rows r0 + j, columns f0 + tx, written transposed
    r = r0 + j;
    f = f0 + tx;
    double2 sm = make_double2(0.0, 0.0), dm = make_double2(0.0, 0.0);
    if (r < Ne && f < Ne) {
      const size_t o = (size_t)r * Ne + f;
      const double w = one * (wgt ? wgt[o] : 1.0);
      const double2 a1 = G1[o], a2 = G2[o];
      sm = make_double2(w * (a1.x + pair * a2.x), w * (a1.y + pair * a2.y));
      dm = make_double2(w * (a1.y - pair * a2.y), -w * (a1.x - pair * a2.x));
    }
    tS[j][tx] = sm;
    tD[j][tx] = dm;
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int f = f0 + j, r = r0 + tx;
    if (f < Ne && r < NeP) {
      oS_r[(size_t)f * NeP + r] = tS[tx][j];
      oD_r[(size_t)f * NeP + r] = tD[tx][j];
    }
  }
}

// --------------------------------------------------------------------------
// out[m][j][i] = in[m][i][j] for nmat matrices of rows x cols complex128,
// through 32 x 32 LDS tiles (the workspace [slot][site, colour][row] <->
// [slot][row][site, colour] for the contraction adjoint).
// grid (ceil(cols/32), ceil(rows/32), nmat), block (32, 8)
// --------------------------------------------------------------------------
__global__ __launch_bounds__(256) void adj_transpose_kernel(const double2* __restrict__ in, double2* __restrict__ out,
                                                            int rows, int cols) {
  __shared__ double2 tile[32][33];
  const size_t off = (size_t)blockIdx.z * rows * cols;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  for (int j = ty; j < 32; j += 8) {
    const int r = r0 + j, c = c0 + tx;
    tile[j][tx] = (r < rows && c < cols) ? in[off + (size_t)r * cols + c] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int c = c0 + j, r = r0 + tx;
    if (c < cols && r < rows) out[off + (size_t)c * rows + r] = tile[tx][j];
  }
}

struct AdjArgs {
  Geom g;
  int D, Q;
};

// --------------------------------------------------------------------------
// Contraction adjoint on the site-major workspace WT[slot][row][site][c]
// (the transpose of W).  Lane = site x (64 consecutive sites per wave), and
// the wave holds RB consecutive rows r0 .. r0+RB-1.  Per derivative d, first
// for rows as left operands (Brow, gradient of the left slots), then as right
// operands (Bcol, gradient of the right slots):
//   Gamma[i](x) = sum_q S_q[f][r0+i] c_q(x) + D_q[f][r0+i] s_q(x)
//   acc[i][k][c](x) += op(Gamma[i](x)) WT[src_k][f][x][c]        over f
// (op = conj with the split's right slot as source, identity with its left),
// then gWT[dst_k][r0+i][x][c] += coeff_k acc[i][k][c].  The cotangent values
// S, D are the same for the whole wave: GC groups at a time are read with
// scalar loads issued together (the group count is padded to a multiple of
// GC with zero groups).  The phases c_q(x), s_q(x) and the W values are per
// lane, coalesced over consecutive sites; each W load serves RB rows and is
// issued one f ahead.  With QR > 0 the group count is exactly QR and every
// group's phases stay in registers; QR = 0 takes any multiple of GC and
// reloads the phases of groups past the first chunk.
// Every gWT element a thread updates is its own (its site, the wave's rows):
// no atomics.  phase [Q][V3P] (zero on padded sites and groups).  WT, gWT:
// per timeslice [nW][NeP][V3P][3].  B: per timeslice [D][Q][2][Ne][NeP].
// grid (ceil(V3P / 256), NeP / RB, TB), block 256
// --------------------------------------------------------------------------
template <int NS, int RB, int GC, int QR>
__global__ __launch_bounds__(256) void adj_contract_kernel(const double2* __restrict__ WT, double2* __restrict__ gWT,
                                                           const double2* __restrict__ Brow,
                                                           const double2* __restrict__ Bcol,
                                                           const double2* __restrict__ phase,
                                                           const SplitTerm* __restrict__ splits,
                                                           const int2* __restrict__ dsplit, AdjArgs a) {
  const int V3P = a.g.V3P, NeP = a.g.NeP, Ne = a.g.Ne, Q = a.Q;  // Q: a multiple of GC
  const int x = blockIdx.x * 256 + threadIdx.x;
  const bool live = x < V3P;
  const int xc = live ? x : 0;
  const int r0 = blockIdx.y * RB;  // < NeP
  const int tb = blockIdx.z;
  const size_t rowsz = (size_t)V3P * 3;  // one row of one slot
  const size_t slotsz = (size_t)NeP * rowsz;
  const double2* Wb = WT + (size_t)tb * a.g.nW * slotsz + (size_t)xc * 3;
  double2* gWb = gWT + (size_t)tb * a.g.nW * slotsz + (size_t)xc * 3;
  const size_t bm = (size_t)Ne * NeP;
  constexpr int NPH = QR > 0 ? QR : GC;
  double2 ph[NPH];  // phases of every group (QR > 0) or of the first chunk
#pragma unroll
  for (int q = 0; q < NPH; ++q) ph[q] = phase[(size_t)q * V3P + xc];
  for (int d = 0; d < a.D; ++d) {
    const int2 fs = dsplit[d];
#pragma unroll 1
    for (int side = 0; side < 2; ++side) {
      const double2* B = (side == 0 ? Brow : Bcol) + (size_t)(tb * a.D + d) * Q * 2 * bm + r0;
      const double2* src[NS];
#pragma unroll
      for (int k = 0; k < NS; ++k) {
        const int n = k < fs.y ? (side == 0 ? splits[fs.x + k].nR : splits[fs.x + k].nL) : 0;
        src[k] = Wb + (size_t)n * slotsz;
      }
      double2 acc[RB][NS][3];
#pragma unroll
      for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int k = 0; k < NS; ++k)
#pragma unroll
          for (int c = 0; c < 3; ++c) acc[i][k][c] = make_double2(0.0, 0.0);
      double2 wn[NS][3];  // W of the next f
#pragma unroll
      for (int k = 0; k < NS; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) wn[k][c] = src[k][c];
      for (int f = 0; f < Ne; ++f) {
        double2 w[NS][3];
        const int fn = min(f + 1, Ne - 1);
#pragma unroll
        for (int k = 0; k < NS; ++k)
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            w[k][c] = wn[k][c];
            wn[k][c] = src[k][(size_t)fn * rowsz + c];
          }
        double2 gam[RB];
#pragma unroll
        for (int i = 0; i < RB; ++i) gam[i] = make_double2(0.0, 0.0);
        const double2* Bf = B + (size_t)f * NeP;
        for (int q0 = 0; q0 < Q; q0 += GC) {
          double2 sv[GC][RB], dv[GC][RB];
#pragma unroll
          for (int qq = 0; qq < GC; ++qq)
#pragma unroll
            for (int i = 0; i < RB; ++i) {
              sv[qq][i] = Bf[(size_t)(2 * (q0 + qq)) * bm + i];
              dv[qq][i] = Bf[(size_t)(2 * (q0 + qq) + 1) * bm + i];
            }
#pragma unroll
          for (int qq = 0; qq < GC; ++qq) {
            // phase of group q0 + qq: from registers (selected without
            // dynamic register indexing), or reloaded past them
            double2 cs = ph[qq];
#pragma unroll
            for (int cc = 1; cc < NPH / GC; ++cc)
              if (q0 == cc * GC) cs = ph[cc * GC + qq];
            if (QR == 0 && q0 > 0) cs = phase[(size_t)(q0 + qq) * V3P + xc];
#pragma unroll
            for (int i = 0; i < RB; ++i) {
              gam[i].x = fma(sv[qq][i].x, cs.x, gam[i].x);
              gam[i].y = fma(sv[qq][i].y, cs.x, gam[i].y);
              gam[i].x = fma(dv[qq][i].x, cs.y, gam[i].x);
              gam[i].y = fma(dv[qq][i].y, cs.y, gam[i].y);
            }
          }
        }
        if (side == 0) {
#pragma unroll
          for (int i = 0; i < RB; ++i) gam[i].y = -gam[i].y;
        }
#pragma unroll
        for (int k = 0; k < NS; ++k) {
          if (k < fs.y) {
#pragma unroll
            for (int i = 0; i < RB; ++i)
#pragma unroll
              for (int c = 0; c < 3; ++c) {
                acc[i][k][c].x = fma(gam[i].x, w[k][c].x, acc[i][k][c].x);
                acc[i][k][c].x = fma(-gam[i].y, w[k][c].y, acc[i][k][c].x);
                acc[i][k][c].y = fma(gam[i].x, w[k][c].y, acc[i][k][c].y);
                acc[i][k][c].y = fma(gam[i].y, w[k][c].x, acc[i][k][c].y);
              }
          }
        }
      }
      if (live) {
#pragma unroll
        for (int k = 0; k < NS; ++k) {
          if (k < fs.y) {
            const SplitTerm st = splits[fs.x + k];
            const double cf = (double)st.coeff;
            double2* dst = gWb + (size_t)(side == 0 ? st.nL : st.nR) * slotsz;
#pragma unroll
            for (int i = 0; i < RB; ++i) {
              double2* pw = dst + (size_t)(r0 + i) * rowsz;
#pragma unroll
              for (int c = 0; c < 3; ++c) {
                const double2 v = pw[c];
                pw[c] = make_double2(v.x + cf * acc[i][k][c].x, v.y + cf * acc[i][k][c].y);
              }
            }
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
