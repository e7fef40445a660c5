// Mixed-form contraction with both stages on FP16 matrix instructions.
//
// Same work decomposition, LDS layout and barrier protocol as
// contract_mixed_ws_body (four producer waves form the Q tiles of a 32-site
// chunk, four consumer waves contract them against the momentum phases), with
// two changes:
//
// Stage 1 reads the workspace in the hsplit form (hi/lo FP16 pairs of
// W 2^wsh) and runs on v_mfma_f32_16x16x32_f16.  Lane group g4 of k-group m
// carries the split-colour pairs j = 8 m + 2 g4 and j + 1 (pairs beyond the
// last are zero), as two complex FP16 pairs per operand:
//   A  = (Lh c, Ll c, Lh' c', Ll' c')          (c: the split coefficient)
//   B1 = (Rh, Rl, Rh', Rl')    B2 = (Rl, Rh, Rl', Rh')
// so A.B1 + A.B2 pairs every FP16 part of L with every part of R and the dot
// of the (re, im) pairs gives c Re(conj(L) R) up to the 22-bit split; the
// imaginary part uses A' = ((-Li, Lr) c, ...) with the same B1, B2.  Every
// operand register is a load or one packed multiply (the swap rides on the
// operand selection), products of FP16 values are exact and sums FP32, as in
// the FP32 stage 1.  FP16 matrix instructions leave the vector ALU to the
// other wave of the SIMD, which FP32 16x16x4 instructions do not.
//
// The zero momentum is summed by the producers, in FP32 over their eight
// sites of a chunk relative to each lane's chunk-0 mean, then in FP64, as in
// contract_mixed_ws_body.
// --------------------------------------------------------------------------
#pragma once

#include "ed_kernels.hpp"

namespace ed {

typedef uint32_t u4v __attribute__((ext_vector_type(4)));
typedef _Float16 h8v __attribute__((ext_vector_type(8)));

// c (l.x, l.y) and c (-l.y, l.x) of a packed FP16 pair: one packed multiply
// each (the swap folds into its operand selection).  Plain vector code, not
// inline assembly: these registers feed and follow matrix instructions, and
// only instructions the compiler sees get the wait states that a vector
// write before a matrix read, or after a matrix read of an accumulator,
// requires.
__device__ __forceinline__ uint32_t pk_scale(uint32_t l, uint32_t c2) {
  return __builtin_bit_cast(uint32_t, __builtin_bit_cast(h2v, l) * __builtin_bit_cast(h2v, c2));
}
__device__ __forceinline__ uint32_t pk_scale_swap(uint32_t l, uint32_t cm2) {
  const h2v v = __builtin_bit_cast(h2v, l);
  return __builtin_bit_cast(uint32_t, h2v{v.y, v.x} * __builtin_bit_cast(h2v, cm2));
}

// (v.y, v.x) of a 64-bit pair
__device__ __forceinline__ uint64_t swap_pair(uint64_t v) { return (v >> 32) | (v << 32); }

// (re, im) times i^k (k uniform; exact)
__device__ __forceinline__ void rot_i(int k, double& re, double& im) {
  if (k & 1) {
    const double t = re;
    re = -im;
    im = t;
  }
  if (k & 2) {
    re = -re;
    im = -im;
  }
}
typedef uint64_t u2l __attribute__((ext_vector_type(2)));

// LDS layout of the Q tiles.  A site holds 16 rows e of 16 columns f in 512
// floats, real and imaginary parts in separate 16-float slots: row e = 4 G + r
// puts its real part in slot q and its imaginary part in slot 4 + q of the
// 128-float block G, q = r ^ G ^ (site / 8) (mod 4).  The imaginary part is
// always 64 floats after the real part and sites are 512 floats apart, so
// every access of a lane is a fixed per-lane base plus immediate offsets (one
// ds_write2 / ds_read2 per value pair, no register moves or address
// arithmetic).  Banks (64 of 4 bytes): a producer store (one site, lanes:
// G = g4, column l16) and a consumer load (one row, lanes: site 8 g4 + k,
// column l16) both put lane group g4 in its own bank quarter.
constexpr int QS16 = 16 * 32;
__device__ __forceinline__ int qre(int buf, int site, int row, int col) {
  return (buf * SC2 + site) * QS16 + 128 * (row >> 2) + 16 * (((row & 3) ^ (row >> 2) ^ (site >> 3)) & 3) + col;
}
constexpr int QIM = 64;  // imaginary part: qre(...) + QIM
constexpr int QFLOATS = 2 * SC2 * QS16 + 2;
static_assert(NPROD * 16 * 16 * (int)sizeof(double2) <= (int)sizeof(float) * QFLOATS,
              "the producers' zero-momentum partial sums fit in the Q buffers");

// Producer loop for exactly NG k-groups (8 split-colour pairs each).  UNIT:
// every split coefficient is +1 (the unit plan), so A is the loaded L itself
// and only the imaginary operand i L costs a packed multiply per register.
template <int NG, bool UNIT>
__device__ __forceinline__ void mixed16_producer(const uint2* __restrict__ W, const ContractArgs& a, const DPlan& dp,
                                                 int tb, int E0, int F0, float* __restrict__ Qf,
                                                 float (*wmax)[NWAVE]) {
  typedef float4v acc_t;
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int l16 = lane & 15, g4 = lane >> 4;
  const Geom g = a.g;
  const int nchunk = g.V3P / SC2;
  const acc_t zero = {0.f, 0.f, 0.f, 0.f};
  const uint32_t site_b = 3u * (uint32_t)g.NeP * 8u;     // bytes per site
  const size_t slice_b = (size_t)g.nW * g.V3P * site_b;  // < 2^31 (checked on the host)
  const uint2* Wt = W + (size_t)tb * (slice_b / 8);
  const __amdgpu_buffer_rsrc_t wr = __builtin_amdgcn_make_buffer_rsrc((void*)Wt, (short)0, (int)slice_b, 0x00020000);
  const int J = 3 * dp.nsplit;
  const int oob = (int)slice_b;  // past the buffer's end: the load returns 0 without a memory request
  int offL[NG][2], offR[NG][2];
  uint32_t cre[NG][2], cim[NG][2];  // (c, c) and (-c, c) as packed FP16
#pragma unroll
  for (int m = 0; m < NG; ++m) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int j = 8 * m + 2 * g4 + h;
      // pairs beyond the last and padded rows read past the buffer's end
      // (the slice size, < 2^31; offset + site offset < 2^32): the loads
      // return 0 without touching memory, and the coefficient is 0 as well
      const int jj = min(j, J - 1);
      const SplitTerm st = a.splits[dp.first + jj / 3];
      offL[m][h] = j < J ? woff8(g, st.nL, jj % 3, E0 + l16, oob) : oob;
      offR[m][h] = j < J ? woff8(g, st.nR, jj % 3, F0 + l16, oob) : oob;
      const _Float16 cf = j < J ? (_Float16)st.coeff : (_Float16)0.f;
      cre[m][h] = __builtin_bit_cast(uint32_t, h2v{cf, cf});
      cim[m][h] = __builtin_bit_cast(uint32_t, h2v{-cf, cf});
    }
  }
  // (-1, 1) from a runtime value: as a compile-time constant the compiler
  // rewrites the swap-and-multiply into a multiply and a separate rotate
  const _Float16 m1h = (_Float16)runtime_m1(g);
  const h8v cunit = {m1h, -m1h, m1h, -m1h, m1h, -m1h, m1h, -m1h};
  uint2 lv[NG][2][SPP], rv[NG][2][SPP];
  auto load = [&](int m, int i, int c0) {
    const int xo = (c0 + SPP * wave + i) * (int)site_b;  // uniform
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      lv[m][h][i] = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(wr, offL[m][h], xo, 0));
      rv[m][h][i] = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(wr, offR[m][h], xo, 0));
    }
  };

  // zero-momentum sums of this wave's sites, rows 4 g4 + r, column l16:
  // FP32 over the wave's SPP sites of a chunk, then FP64
  double zr[4] = {0, 0, 0, 0}, zi[4] = {0, 0, 0, 0};

  // one chunk: site i's MFMAs are issued before site i - 1's results are
  // stored, so the stores, sums and maximum overlap the matrix pipe
  auto stage1 = [&](int chunk, int buf) {
    const int c0n = min(chunk + 1, nchunk - 1) * SC2;  // the last chunk re-reads itself
    float mx = 0.f;
    float sr[4], si[4];
    // Q of site i to LDS, into the tile maximum and the zero-momentum sums
    // (padded sites hold Q = 0 exactly)
    auto finish = [&](int i, const acc_t& qr, const acc_t& qi) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int o = qre(buf, SPP * wave + i, mfma<float>::row(g4, r), l16);
        Qf[o] = qr[r];
        Qf[o + QIM] = qi[r];
        if (i == 0) {
          sr[r] = qr[r];
          si[r] = qi[r];
        } else {
          sr[r] += qr[r];
          si[r] += qi[r];
        }
      }
      mx = fmaxf(fmaxf(mx, fabsf(qr[0])), fabsf(qi[0]));
      mx = fmaxf(fmaxf(mx, fabsf(qr[1])), fabsf(qi[1]));
      mx = fmaxf(fmaxf(mx, fabsf(qr[2])), fabsf(qi[2]));
      mx = fmaxf(fmaxf(mx, fabsf(qr[3])), fabsf(qi[3]));
    };
    acc_t pr = zero, pi = zero;
#pragma unroll
    for (int i = 0; i < SPP; ++i) {
      acc_t qr = zero, qi = zero;
#pragma unroll
      for (int m = 0; m < NG; ++m) {
        const uint2 l0 = lv[m][0][i], l1 = lv[m][1][i], r0 = rv[m][0][i], r1 = rv[m][1][i];
        half8 Ar, Ai;
        if constexpr (UNIT) {
          // i L = (-Im, Re) per pair: one swap-and-multiply over the whole
          // operand (a packed multiply per register)
          const h8v L = __builtin_bit_cast(h8v, u4v{l0.x, l0.y, l1.x, l1.y});
          Ar = __builtin_bit_cast(half8, L);
          Ai = __builtin_bit_cast(half8, __builtin_shufflevector(L, L, 1, 0, 3, 2, 5, 4, 7, 6) * cunit);
        } else {
          Ar = __builtin_bit_cast(half8, u4v{pk_scale(l0.x, cre[m][0]), pk_scale(l0.y, cre[m][0]),
                                             pk_scale(l1.x, cre[m][1]), pk_scale(l1.y, cre[m][1])});
          Ai = __builtin_bit_cast(half8, u4v{pk_scale_swap(l0.x, cim[m][0]), pk_scale_swap(l0.y, cim[m][0]),
                                             pk_scale_swap(l1.x, cim[m][1]), pk_scale_swap(l1.y, cim[m][1])});
        }
        const half8 B1 = __builtin_bit_cast(half8, u4v{r0.x, r0.y, r1.x, r1.y});
        const half8 B2 = __builtin_bit_cast(
            half8, u2l{swap_pair(__builtin_bit_cast(uint64_t, r0)), swap_pair(__builtin_bit_cast(uint64_t, r1))});
        qr = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ar, B1, qr, 0, 0, 0);
        qi = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ai, B1, qi, 0, 0, 0);
        qr = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ar, B2, qr, 0, 0, 0);
        qi = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ai, B2, qi, 0, 0, 0);
        load(m, i, c0n);
      }
      if (i > 0) finish(i - 1, pr, pi);
      pr = qr;
      pi = qi;
      // keep each site's operand preparation with its own matrix
      // instructions: hoisted to the top of the chunk, it would wait there
      // for loads issued late in the previous chunk
      __builtin_amdgcn_sched_barrier(0);
    }
    finish(SPP - 1, pr, pi);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      zr[r] += (double)sr[r];
      zi[r] += (double)si[r];
    }
    mx = wave_max_nonneg(mx);
    if (lane == 0) wmax[buf][wave] = mx;
  };

  __builtin_amdgcn_s_setprio(2);
#pragma unroll
  for (int i = 0; i < SPP; ++i) {
#pragma unroll
    for (int m = 0; m < NG; ++m) load(m, i, 0);
  }
  stage1(0, 0);
  __syncthreads();
  for (int chunk = 0; chunk + 1 < nchunk; ++chunk) {
    stage1(chunk + 1, (chunk + 1) & 1);
    __syncthreads();
  }
  __syncthreads();  // matches the consumers' last chunk
  // every consumer has read its last chunk: LDS buffer 0 takes the four
  // producers' zero-momentum partial sums, rows x columns, as doubles
  double2* zs = (double2*)(void*)Qf;
#pragma unroll
  for (int r = 0; r < 4; ++r) zs[(wave * 16 + mfma<float>::row(g4, r)) * 16 + l16] = make_double2(zr[r], zi[r]);
  __syncthreads();
}

template <int NKS>
__device__ __forceinline__ void contract_mixed16_body(const uint2* __restrict__ W, const half8* __restrict__ ph16,
                                                      double2* __restrict__ out, const ContractArgs& a, const int w,
                                                      float* __restrict__ Qf, float (*wmax)[NWAVE],
                                                      half8 (*PHs)[4][64]) {
  typedef float4v acc_t;
  typedef uint32_t u4 __attribute__((ext_vector_type(4)));
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int l16 = lane & 15, g4 = lane >> 4;

  int tb, di, pb, te, tf;
  decode_item(a, w, tb, di, pb, te, tf);
  const DPlan dp = a.dplan[di];
  const int E0 = te * 16, F0 = tf * 16;
  const Geom g = a.g;
  const int nchunk = g.V3P / SC2;
  const acc_t zero = {0.f, 0.f, 0.f, 0.f};

  if (wave < NPROD) {
    // producer: one branch-free instance per k-group count (NKS groups of
    // eight split-colour pairs at most)
    const int ng = (3 * dp.nsplit + 7) >> 3;
    constexpr int NG2 = NKS >= 2 ? 2 : 1;
    if (a.unit) {
      if (NKS >= 2 && ng == 2) mixed16_producer<NG2, true>(W, a, dp, tb, E0, F0, Qf, wmax);
      else mixed16_producer<1, true>(W, a, dp, tb, E0, F0, Qf, wmax);
    } else {
      if (NKS >= 2 && ng == 2) mixed16_producer<NG2, false>(W, a, dp, tb, E0, F0, Qf, wmax);
      else mixed16_producer<1, false>(W, a, dp, tb, E0, F0, Qf, wmax);
    }
    return;
  }

  // ================================ consumer ================================
  const int cw = wave - NPROD;
  const half8* PH = ph16 + (size_t)pb * nchunk * 4 * 16 * 4;
  const __amdgpu_buffer_rsrc_t pr_ = __builtin_amdgcn_make_buffer_rsrc((void*)PH, (short)0, nchunk * 4 * 16 * 4 * 16, 0x00020000);
  // The chunk's phase operand, 4 parts (Re hi, Re lo, Im hi, Im lo) of 1 KB,
  // crosses the load path once per workgroup: consumer cw loads part cw two
  // chunks ahead, stores it to LDS one chunk ahead (double buffer PHs), and
  // every consumer reads all four parts from LDS.  Lane (l16, g4) holds
  // momentum l16, sites 8 g4 .. 8 g4 + 7 of a part.
  static_assert(NBC == 4, "one phase part per consumer wave");
  const int phoff = (l16 * 4 + g4) * 16;
  auto load_part = [&](int chunk) {
    return __builtin_bit_cast(half8, __builtin_amdgcn_raw_buffer_load_b128(pr_, phoff, chunk * 4096 + cw * 1024, 0));
  };
  half8 prh, prl, pih, pil;
  double D1[NBC][4], D2[NBC][4], D3[NBC][4], D4[NBC][4];
  float cr[NBC], ci[NBC];
#pragma unroll
  for (int b = 0; b < NBC; ++b) {
#pragma unroll
    for (int r = 0; r < 4; ++r) D1[b][r] = D2[b][r] = D3[b][r] = D4[b][r] = 0;
  }
  // stage 2 of one chunk for row b; qb points at the real part of this
  // lane's first site (8 g4) of the chunk, row NBC cw + b, column l16
  const float m1 = runtime_m1(a.g);
  auto stage2 = [&](int b, const float* qb, float sq, int ex) {
    const float csr = -cr[b] * sq, csi = -ci[b] * sq;
    uint32_t qrh[4], qrl[4], qih[4], qil[4];
#pragma unroll
    for (int v2 = 0; v2 < 4; ++v2) {
      float sr[2], si[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const float* q = qb + (2 * v2 + h) * QS16;
        sr[h] = fmaf(q[0], sq, csr);
        si[h] = fmaf(q[QIM], sq, csi);
      }
      split16x2(sr[0], sr[1], m1, qrh[v2], qrl[v2]);
      split16x2(si[0], si[1], m1, qih[v2], qil[v2]);
    }
    const half8 Qrh = __builtin_bit_cast(half8, u4{qrh[0], qrh[1], qrh[2], qrh[3]});
    const half8 Qrl = __builtin_bit_cast(half8, u4{qrl[0], qrl[1], qrl[2], qrl[3]});
    const half8 Qih = __builtin_bit_cast(half8, u4{qih[0], qih[1], qih[2], qih[3]});
    const half8 Qil = __builtin_bit_cast(half8, u4{qil[0], qil[1], qil[2], qil[3]});
    acc_t A1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(prh, Qrh, zero, 0, 0, 0);
    acc_t A2 = __builtin_amdgcn_mfma_f32_16x16x32_f16(pih, Qih, zero, 0, 0, 0);
    acc_t A3 = __builtin_amdgcn_mfma_f32_16x16x32_f16(pih, Qrh, zero, 0, 0, 0);
    acc_t A4 = __builtin_amdgcn_mfma_f32_16x16x32_f16(prh, Qih, zero, 0, 0, 0);
    A1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(prh, Qrl, A1, 0, 0, 0);
    A2 = __builtin_amdgcn_mfma_f32_16x16x32_f16(pih, Qil, A2, 0, 0, 0);
    A3 = __builtin_amdgcn_mfma_f32_16x16x32_f16(pih, Qrl, A3, 0, 0, 0);
    A4 = __builtin_amdgcn_mfma_f32_16x16x32_f16(prh, Qil, A4, 0, 0, 0);
    A1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(prl, Qrh, A1, 0, 0, 0);
    A2 = __builtin_amdgcn_mfma_f32_16x16x32_f16(pil, Qih, A2, 0, 0, 0);
    A3 = __builtin_amdgcn_mfma_f32_16x16x32_f16(pil, Qrh, A3, 0, 0, 0);
    A4 = __builtin_amdgcn_mfma_f32_16x16x32_f16(prl, Qih, A4, 0, 0, 0);
    const double inv = ldexp(1.0, ex - 28);  // 1 / (2^14 sq)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      D1[b][r] = fma((double)A1[r], inv, D1[b][r]);
      D2[b][r] = fma((double)A2[r], inv, D2[b][r]);
      D3[b][r] = fma((double)A3[r], inv, D3[b][r]);
      D4[b][r] = fma((double)A4[r], inv, D4[b][r]);
    }
  };

  half8 nxt = load_part(0);
  PHs[0][cw][lane] = nxt;
  if (nchunk > 1) nxt = load_part(1);
  __syncthreads();  // chunk 0: Q and phase in LDS
  {
    const int n0 = min(SC2, g.V3);
#pragma unroll
    for (int b = 0; b < NBC; ++b) {
      float sr = 0, si = 0;
      for (int xl = 0; xl < n0; ++xl) {
        const int o = qre(0, xl, NBC * cw + b, l16);
        sr += Qf[o];
        si += Qf[o + QIM];
      }
      cr[b] = sr / (float)n0;
      ci[b] = si / (float)n0;
    }
  }
  for (int chunk = 0; chunk < nchunk; ++chunk) {
    const int buf = chunk & 1;
    prh = PHs[buf][0][lane];
    prl = PHs[buf][1][lane];
    pih = PHs[buf][2][lane];
    pil = PHs[buf][3][lane];
    float tmax = 0.f;
#pragma unroll
    for (int v = 0; v < NPROD; ++v) tmax = fmaxf(tmax, wmax[buf][v]);
    int ex = 0;
    if (tmax > 0.f) (void)frexpf(2.f * tmax, &ex);
    const float sq = ldexpf(1.f, 14 - ex);
#pragma unroll
    for (int b = 0; b < NBC; ++b) {
      stage2(b, Qf + qre(buf, 8 * g4, NBC * cw + b, l16), sq, ex);
      __builtin_amdgcn_sched_barrier(0);
    }
    // buffer (chunk + 1) & 1 was last read in the previous chunk
    if (chunk + 1 < nchunk) PHs[(chunk + 1) & 1][cw][lane] = nxt;
    if (chunk + 2 < nchunk) nxt = load_part(chunk + 2);
    __syncthreads();
  }
  __syncthreads();  // the producers' zero-momentum partial sums are in LDS
  const double2* zs = (const double2*)(const void*)Qf;

  // ---- epilogue: values in units of 2^(2 wsh) ----
  const double wsc = ldexp(1.0, -2 * g.wsh);
  const int Ne = g.Ne, P = a.P;
  const size_t plane = (size_t)Ne * Ne;
  const size_t tbase = (size_t)(a.t_out0 + tb) * a.D * P;
  const int f = F0 + l16;
#pragma unroll
  for (int b = 0; b < NBC; ++b) {
    const int e = E0 + NBC * cw + b;
    const bool ok = (e < Ne) && (f < Ne);
    const double wef = ((ok && a.wgt) ? a.wgt[(size_t)e * Ne + f] : 1.0) * wsc;
    const double wfe = ((ok && a.wgt) ? a.wgt[(size_t)f * Ne + e] : 1.0) * wsc;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int bi = pb * 16 + mfma<float>::row(g4, r);
      if (!ok || bi >= a.nbase) continue;
      const int ip = a.plus_idx[bi], im = a.minus_idx[bi];
      const double S = a.psum[bi], mr = (double)cr[b] * S, mi = (double)ci[b] * S;
      double epr = D1[b][r] - D2[b][r] + mr, epi = D3[b][r] + D4[b][r] + mi;
      double emr = D1[b][r] + D2[b][r] + mr, emi = D4[b][r] - D3[b][r] + mi;
      rot_i(dp.rot, epr, epi);
      rot_i(dp.rot, emr, emi);
      if (ip >= 0) out[(tbase + (size_t)dp.d * P + ip) * plane + (size_t)e * Ne + f] = make_double2(epr * wef, epi * wef);
      if (im >= 0) out[(tbase + (size_t)dp.d * P + im) * plane + (size_t)e * Ne + f] = make_double2(emr * wef, emi * wef);
      if (partner_writes(dp, te, tf)) {
        const double s = (double)dp.sign;
        if (im >= 0)
          out[(tbase + (size_t)dp.partner * P + im) * plane + (size_t)f * Ne + e] =
              make_double2(s * epr * wfe, -s * epi * wfe);
        if (ip >= 0)
          out[(tbase + (size_t)dp.partner * P + ip) * plane + (size_t)f * Ne + e] =
              make_double2(s * emr * wfe, -s * emi * wfe);
      }
    }
    // zero momentum: the producers' sums of Q over their sites
    const int row = NBC * cw + b;
    const double2 z0 = zs[(0 * 16 + row) * 16 + l16], z1 = zs[(1 * 16 + row) * 16 + l16];
    const double2 z2 = zs[(2 * 16 + row) * 16 + l16], z3 = zs[(3 * 16 + row) * 16 + l16];
    double zre = (z0.x + z1.x) + (z2.x + z3.x), zim = (z0.y + z1.y) + (z2.y + z3.y);
    rot_i(dp.rot, zre, zim);
    if (ok && g4 == 0 && pb == 0 && a.zero_idx >= 0) {
      out[(tbase + (size_t)dp.d * P + a.zero_idx) * plane + (size_t)e * Ne + f] = make_double2(zre * wef, zim * wef);
      if (partner_writes(dp, te, tf)) {
        const double s = (double)dp.sign;
        out[(tbase + (size_t)dp.partner * P + a.zero_idx) * plane + (size_t)f * Ne + e] =
            make_double2(s * zre * wfe, -s * zim * wfe);
      }
    }
  }
}

template <int NKS>
__global__ __launch_bounds__(64 * NWAVE, 1) void contract_mixed16_kernel(const uint2* __restrict__ W,
                                                                          const half8* __restrict__ ph16,
                                                                          double2* __restrict__ out,
                                                                          ContractArgs a) {
  __shared__ __align__(16) float Qf[QFLOATS];
  __shared__ float wmax[2][NWAVE];
  __shared__ half8 PHs[2][4][64];
  const int w = item_index(a, blockIdx.x, gridDim.x);
  if (w >= a.nwork) return;
  contract_mixed16_body<NKS>(W, ph16, out, a, w, Qf, wmax, PHs);
}

// Batched form: one workgroup per work item of the problem table (items
// numbered problem after problem, mapped to XCDs as contract_mixed_item_kernel
// does); W of each problem holds the hi/lo FP16 workspace.
template <int NKS>
__global__ __launch_bounds__(64 * NWAVE, 1) void contract_mixed16_item_kernel(const GroupProblem* __restrict__ probs,
                                                                              int nprob, int total, int order) {
  __shared__ __align__(16) float Qf[QFLOATS];
  __shared__ float wmax[2][NWAVE];
  __shared__ half8 PHs[2][4][64];
  const int item = __builtin_amdgcn_readfirstlane(group_item(blockIdx.x >> 3, (total + 7) >> 3, order));
  if (item >= total) return;
  const int i = __builtin_amdgcn_readfirstlane(group_find(probs, nprob, item));
  const ContractArgs a = probs[i].a;
  contract_mixed16_body<NKS>((const uint2*)probs[i].W, (const half8*)probs[i].ph0, probs[i].out, a,
                             item - probs[i].work0, Qf, wmax, PHs);
}

}  // namespace ed
