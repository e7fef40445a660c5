// ed_kernels.hpp -- gfx950 device code for elemental generation.
//
// Data layout in HBM (one timeslice batch of TB slices):
//   W  [TB][nW][V3P][3][NeP]  complex<T>, nW = D displaced vectors
//      W_n = nabla_{d_k} ... nabla_{d_1} V for derivative(n) = (d_1..d_k)
//      (lattice/generator/elemental.py:279-288 applies tuple elements in
//      order).  Site-major with the eigenvector index fastest: the 16
//      eigenvectors of a contraction tile at one (site, colour) are one
//      contiguous segment, and the hop kernel maps lanes to eigenvectors.
//      NeP = Ne rounded up to 16, V3P = V3 rounded up to 32; padded rows and
//      sites hold zeros so the contraction needs no tail masks.
//   U  [3][Lt][V3][3][3]      complex<double> spatial links (elemental.py:103)
//   Vd [Lt][Ne][V3][3]        complex<float>: the reference stages V in a
//                             complex64 buffer (elemental.py:55)
//   PH [NPB][V3P][16] x2      real and imaginary phase tables of the base
//                             momenta (one block of 16 per NPB)
//   out[nt][D][P][Ne][Ne]     complex<double>
//
// The contraction evaluates, for every derivative d and base momentum b,
//   Q_d[e,f,x]  = sum_{splits s} c_s sum_c conj(W_{L_s}[e,x,c]) W_{R_s}[f,x,c]   (stage 1)
//   E_d[+b]     = sum_x Phi_b(x) Q_d[e,f,x],  E_d[-b] = sum_x conj(Phi_b(x)) Q_d[e,f,x] (stage 2)
// which equals the reference's sum over splits and momenta of
// contract("zyx,ezyxc,fzyxc->ef", c_s Phi_p, left.conj(), right)
// (elemental.py:309-329): stage 2 is linear in Q, so the split sum is taken
// per site before the phase contraction, and the four real products
// Re/Im(Phi) x Re/Im(Q) serve +b and -b at once.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace ed {

template <typename T> struct cplx;
template <> struct cplx<double> { using type = double2; };
template <> struct cplx<float> { using type = float2; };
template <typename T> using c_t = typename cplx<T>::type;

template <typename T> __device__ __forceinline__ c_t<T> mk(T r, T i) { c_t<T> v; v.x = r; v.y = i; return v; }

// Workspace element of the mixed form's FP16 stage 1: the complex value w,
// scaled by 2^wsh, as FP16 hi/lo pairs, x = (hi re, hi im) and
// y = (lo re, lo im), hi = fp16(w 2^wsh), lo = fp16(w 2^wsh - hi): 22 bits of
// each component in 8 bytes, the size of a complex64.
struct hsplit {};
template <> struct cplx<hsplit> { using type = uint2; };
typedef _Float16 h2v __attribute__((ext_vector_type(2)));

// read / write one workspace element as a complex double (wsh: the hsplit
// scale exponent, unused by the other forms)
__device__ __forceinline__ double2 wget(const double2 v, int) { return v; }
__device__ __forceinline__ double2 wget(const float2 v, int) { return make_double2(v.x, v.y); }
__device__ __forceinline__ double2 wget(const uint2 v, int wsh) {
  // hi + lo needs at most 22 significant bits: the FP32 sum is exact
  const h2v h = __builtin_bit_cast(h2v, v.x), l = __builtin_bit_cast(h2v, v.y);
  return make_double2(ldexp((double)((float)h.x + (float)l.x), -wsh), ldexp((double)((float)h.y + (float)l.y), -wsh));
}
template <typename WT> __device__ __forceinline__ WT wput(double re, double im, int wsh);
template <> __device__ __forceinline__ double2 wput<double2>(double re, double im, int) { return make_double2(re, im); }
template <> __device__ __forceinline__ float2 wput<float2>(double re, double im, int) {
  return make_float2((float)re, (float)im);
}
template <> __device__ __forceinline__ uint2 wput<uint2>(double re, double im, int wsh) {
  // The residual is formed from the FP32 rounding of w 2^wsh (2^-24 of it,
  // below the pair's 2^-22), exactly inside v_fma_mix.  With contraction
  // allowed the compiler folds the two roundings double -> float -> half
  // into one double -> half conversion, which has no instruction and runs as
  // ~40 integer operations: keep them apart.
#pragma clang fp contract(off)
  const float xr = ldexpf((float)re, wsh), xi = ldexpf((float)im, wsh);
  const h2v h = {(_Float16)xr, (_Float16)xi};
  const uint32_t hu = __builtin_bit_cast(uint32_t, h);
  uint32_t l;
  asm("v_fma_mixlo_f16 %0, %1, -1.0, %2 op_sel_hi:[1,0,0]" : "=v"(l) : "v"(hu), "v"(xr));
  asm("v_fma_mixhi_f16 %0, %1, -1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "+v"(l) : "v"(hu), "v"(xi));
  return make_uint2(hu, l);
}

// --------------------------------------------------------------------------
// MFMA wrappers: 16x16x4, one operand value per lane.
//   A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k (both types).
//   D[i][j] sits in lane j + 16 g, element r, with
//     f32: i = 4 g + r      f64: i = 4 r + g
//   (measured on gfx950 with one-hot operands); row(g, r) returns i.
// --------------------------------------------------------------------------
template <typename T> struct mfma;
template <> struct mfma<double> {
  typedef double acc __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc mac(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int r) { return 4 * r + g; }
};
template <> struct mfma<float> {
  typedef float acc __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc mac(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int r) { return 4 * g + r; }
};

struct Geom {
  int Lx, Ly, Lz, Lt;
  int V3, V3P, Ne, NeP;
  int nW;
  int wsh;  // scale exponent of an hsplit workspace (0 otherwise)
  int rot;  // 1: the hop stores i times each child (a level-k slot holds i^k W)
};

// element (n, x, c, e) of one timeslice's W block
__device__ __forceinline__ size_t widx(const Geom& g, int n, int x, int c, int e) {
  return (((size_t)n * g.V3P + x) * 3 + c) * g.NeP + e;
}

// Byte offset of W[n][site 0][c][e] (8-byte elements) for a raw buffer load,
// or oob -- at or past the end of the buffer -- for a padded eigenvector row
// (e >= Ne): such a load returns 0, like the zero padding it stands for,
// without a memory request.
__device__ __forceinline__ int woff8(const Geom& g, int n, int c, int e, int oob) {
  return e < g.Ne ? (int)(widx(g, n, 0, c, e) * 8) : oob;
}

// --------------------------------------------------------------------------
// Stage V(t) ([Ne][V3*3], complex64 or, for exact eigenvectors, complex128)
// into slot 0 of W ([V3P*3][NeP]) in the compute precision: a 64 x 64 tile transpose through LDS so that both the
// read (along sites) and the write (along eigenvectors) are contiguous.
// grid (ceil(3 V3 / 64), ceil(Ne / 64), TB), block 256
// --------------------------------------------------------------------------
template <typename T, typename VT>
__global__ __launch_bounds__(256) void stage_v_kernel(const VT* __restrict__ V, c_t<T>* __restrict__ W,
                                                       Geom g, int tb0) {
  __shared__ VT tile[64][65];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int m0 = blockIdx.x * 64, e0 = blockIdx.y * 64, tb = blockIdx.z;
  const int M = 3 * g.V3;
  const VT* Vt = V + (size_t)(tb0 + tb) * g.Ne * (size_t)M;
#pragma unroll 4
  for (int k = 0; k < 16; ++k) {
    const int el = ty * 16 + k, e = e0 + el, m = m0 + tx;
    VT v;
    v.x = 0;
    v.y = 0;
    if (e < g.Ne && m < M) v = Vt[(size_t)e * M + m];
    tile[el][tx] = v;
  }
  __syncthreads();
  c_t<T>* Wt = W + (size_t)tb * g.nW * (size_t)g.V3P * 3 * g.NeP;  // slot 0
#pragma unroll 4
  for (int k = 0; k < 16; ++k) {
    const int ml = ty * 16 + k, m = m0 + ml, e = e0 + tx;
    if (m < M && e < g.Ne) {
      const VT v = tile[tx][ml];
      Wt[(size_t)m * g.NeP + e] = wput<c_t<T>>((double)v.x, (double)v.y, g.wsh);
    }
  }
}

// --------------------------------------------------------------------------
// Covariant hops (elemental.py:282-287, periodic) of one derivative level, all
// three directions of a parent in one pass:
//   child(parent, mu)(x) = U_mu(x) in(x + mu) - U_mu(x - mu)^dag in(x - mu)
// for mu = x, y, z.  The six neighbour reads of a parent value by the three
// children of one site share the caches, so the parent is fetched from HBM
// about once per level instead of twice per child.  The site index is made
// wave-uniform so the 54 link values become scalar loads.  FP64 arithmetic,
// one rounding at the store.
// grid (ceil(V3 / 4), nparent, TB), block 256 (4 sites)
// --------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void hop3_kernel(c_t<T>* __restrict__ W, const double2* __restrict__ U, Geom g,
                                                   int tb0, int parent0, int child0) {
  const int s = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  if (s >= g.V3) return;
  const int pl = blockIdx.y;
  const int parent = parent0 + pl;
  const int tb = blockIdx.z;
  const int x = s % g.Lx, yy = (s / g.Lx) % g.Ly, z = s / (g.Lx * g.Ly);
  int sf[3], sb[3];
  sf[0] = s - x + (x + 1 == g.Lx ? 0 : x + 1);
  sb[0] = s - x + (x == 0 ? g.Lx - 1 : x - 1);
  {
    const int r = s - yy * g.Lx;
    sf[1] = r + (yy + 1 == g.Ly ? 0 : yy + 1) * g.Lx;
    sb[1] = r + (yy == 0 ? g.Ly - 1 : yy - 1) * g.Lx;
  }
  {
    const int r = s - z * g.Lx * g.Ly;
    sf[2] = r + (z + 1 == g.Lz ? 0 : z + 1) * g.Lx * g.Ly;
    sb[2] = r + (z == 0 ? g.Lz - 1 : z - 1) * g.Lx * g.Ly;
  }
  const c_t<T>* Wt = W + (size_t)tb * g.nW * (size_t)g.V3P * 3 * g.NeP;
  c_t<T>* Wo = const_cast<c_t<T>*>(Wt);
#pragma unroll
  for (int mu = 0; mu < 3; ++mu) {
    // one direction's 18 link values live at a time (scalar loads)
    const double2* Um = U + ((size_t)mu * g.Lt + (tb0 + tb)) * (size_t)g.V3 * 9;
    const double2* uf = Um + (size_t)s * 9;
    const double2* ub = Um + (size_t)sb[mu] * 9;
    double2 lf[9], lb[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      lf[k] = uf[k];
      lb[k] = ub[k];
    }
    const int child = child0 + 3 * pl + mu;
    // an hsplit workspace holds w 2^wsh; the hop is linear, so it reads and
    // writes those scaled values as they are (scale exponent 0 below)
    // (U_mu(x) af - U_mu(x - mu)^dag ab)_a for one eigenvector
    auto hop = [&](const double2* af, const double2* ab, int a) {
      // one FMA per real product (written out: the compound expressions
      // compile to a multiply and an add per term)
      double re = 0, im = 0;
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const double2 u = lf[a * 3 + b];
        re = fma(u.x, af[b].x, re);
        re = fma(-u.y, af[b].y, re);
        im = fma(u.x, af[b].y, im);
        im = fma(u.y, af[b].x, im);
        const double2 v = lb[b * 3 + a];  // (U^dag)_{ab} = conj(U_{ba})
        re = fma(-v.x, ab[b].x, re);
        re = fma(-v.y, ab[b].y, re);
        im = fma(-v.x, ab[b].y, im);
        im = fma(v.y, ab[b].x, im);
      }
      return g.rot ? wput<c_t<T>>(-im, re, 0) : wput<c_t<T>>(re, im, 0);
    };
    if constexpr (sizeof(c_t<T>) == 8) {
      // 8-byte elements: each lane takes the eigenvector pair (2 l, 2 l + 1),
      // one 16-byte load per colour and neighbour (NeP is even; the padded
      // partner of an odd Ne is zero and stays zero)
      typedef uint32_t u4 __attribute__((ext_vector_type(4)));
      typedef uint32_t u2 __attribute__((ext_vector_type(2)));
      for (int e = 2 * lane; e < g.Ne; e += 128) {
        double2 af[2][3], ab[2][3];
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          const u4 pf = *(const u4*)&Wt[widx(g, parent, sf[mu], b, e)];
          const u4 pb = *(const u4*)&Wt[widx(g, parent, sb[mu], b, e)];
          af[0][b] = wget(__builtin_bit_cast(c_t<T>, u2{pf.x, pf.y}), 0);
          af[1][b] = wget(__builtin_bit_cast(c_t<T>, u2{pf.z, pf.w}), 0);
          ab[0][b] = wget(__builtin_bit_cast(c_t<T>, u2{pb.x, pb.y}), 0);
          ab[1][b] = wget(__builtin_bit_cast(c_t<T>, u2{pb.z, pb.w}), 0);
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const u2 h0 = __builtin_bit_cast(u2, hop(af[0], ab[0], a));
          const u2 h1 = __builtin_bit_cast(u2, hop(af[1], ab[1], a));
          *(u4*)&Wo[widx(g, child, s, a, e)] = u4{h0.x, h0.y, h1.x, h1.y};
        }
      }
    } else {
      for (int e = lane; e < g.Ne; e += 64) {
        double2 af[3], ab[3];
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          af[b] = wget(Wt[widx(g, parent, sf[mu], b, e)], 0);
          ab[b] = wget(Wt[widx(g, parent, sb[mu], b, e)], 0);
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) Wo[widx(g, child, s, a, e)] = hop(af, ab, a);
      }
    }
  }
}

// --------------------------------------------------------------------------
// LDS-staged hop (same outputs as hop3_kernel, bit for bit).  One workgroup:
// one parent slot, hl_ec<T>() eigenvectors (16 for 8-byte workspace elements,
// 8 for 16-byte ones: 128 bytes per site and colour), one 8 x 8 tile of
// (x, y) sites, marching over all z.  LDS holds
//   * a ring of three planes (z - 1, z, z + 1) of the parent: 10 x 10 sites
//     (the tile and a one-site halo) x 3 colours x hl_ec<T>() eigenvectors, and
//   * the links the plane's 64 outputs need: U_x and U_y on the tile, U_x on
//     the halo column x0 - 1 and U_y on the halo row y0 - 1 (backward links),
//     and U_z on the tile for the planes z and z - 1 (two slots),
// so every parent element crosses the load path once per workgroup plus the
// x-y halo (100 / 64 of the tile), and every link once per eigenvector
// chunk.  The next plane's elements and links are loaded into registers
// while the current plane is computed and go to LDS between two barriers.
// Lanes: (site of the iteration, eigenvector e), each computing the three
// colours of its site's child elements from one read of the six neighbours
// (the arithmetic of hop3_kernel, in its order); wave w of NW takes the
// plane's sites SPI w ... SPI (w + 1) - 1, then SPI (w + NW) ...
// Requires Lx, Ly multiples of 8, Lz >= 3, NeP a multiple of hl_ec<T>().
// grid (Lx/8 * Ly/8, nparent * ceil(Ne/hl_ec<T>()), TB), block 64 * NW
// --------------------------------------------------------------------------
constexpr int HL_T = 8, HL_H = HL_T + 2, HL_SITES = HL_H * HL_H;
constexpr int HL_PSI_PIECES = HL_SITES * 3 * 8;      // 16-byte pieces of one plane (2400)
constexpr int HL_LX = 0, HL_LY = 64 * 9, HL_LXH = 2 * 64 * 9, HL_LYH = HL_LXH + 8 * 9;
constexpr int HL_LZ = HL_LYH + 8 * 9;                 // two slots of 64 * 9
constexpr int HL_LINKS = HL_LZ + 2 * 64 * 9;          // double2 (2448, 39 KB)
constexpr int HL_LINK_LOAD = HL_LZ + 64 * 9;          // pieces loaded per plane (1872)
template <typename T> constexpr int hl_ec() { return 128 / (int)sizeof(c_t<T>); }

template <typename T, int NW>
__global__ __launch_bounds__(64 * NW, 1) void hop3_lds_kernel(c_t<T>* __restrict__ W, const double2* __restrict__ U,
                                                              Geom g, int tb0, int parent0, int child0, int nparent) {
  typedef uint32_t u4 __attribute__((ext_vector_type(4)));
  constexpr int NT = 64 * NW;       // threads
  constexpr int EC = hl_ec<T>();
  constexpr int SPI = 64 / EC;      // sites per wave iteration
  constexpr int ITER = 64 / (NW * SPI);
  constexpr int PR = (HL_PSI_PIECES + NT - 1) / NT, LR = (HL_LINK_LOAD + NT - 1) / NT;
  static_assert(64 % (NW * SPI) == 0, "the waves' iterations cover the 64 tile sites");
  __shared__ __align__(16) u4 psi[3][HL_PSI_PIECES];
  __shared__ __align__(16) double2 lk[HL_LINKS];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int ntx = g.Lx / HL_T;
  const int x0 = (blockIdx.x % ntx) * HL_T, y0 = (blockIdx.x / ntx) * HL_T;
  const int pl = blockIdx.y % nparent, ec = blockIdx.y / nparent;
  const int e0 = ec * EC;
  const int tb = blockIdx.z, t = tb0 + tb;
  const int parent = parent0 + pl;
  const c_t<T>* Wt = W + (size_t)tb * g.nW * (size_t)g.V3P * 3 * g.NeP;
  c_t<T>* Wo = W + (size_t)tb * g.nW * (size_t)g.V3P * 3 * g.NeP;
  auto wrap = [](int v, int L) { return v < 0 ? v + L : (v >= L ? v - L : v); };
  auto link_ptr = [&](int mu, int x, int y, int z) {
    return U + (((size_t)mu * g.Lt + t) * g.V3 + ((size_t)z * g.Ly + y) * g.Lx + x) * 9;
  };

  // register staging of one plane: its parent pieces and its links
  u4 sp[PR], sl[LR];
  auto fetch_psi = [&](int z) {
#pragma unroll
    for (int r = 0; r < PR; ++r) {
      // pieces past the plane's last re-read the last one (stored nowhere):
      // every load is unconditional, so the staging stays in registers
      const int k = min(tid + NT * r, HL_PSI_PIECES - 1);
      const int sh = k / 24, rem = k - sh * 24, c = rem >> 3, pc = rem & 7;
      const int hy = sh / HL_H, hx = sh - hy * HL_H;
      const int x = wrap(x0 + hx - 1, g.Lx), y = wrap(y0 + hy - 1, g.Ly);
      const int s = (z * g.Ly + y) * g.Lx + x;
      sp[r] = ((const u4*)&Wt[widx(g, parent, s, c, e0)])[pc];
    }
  };
  auto store_psi = [&](int slot) {
#pragma unroll
    for (int r = 0; r < PR; ++r) {
      const int k = tid + NT * r;
      if (k < HL_PSI_PIECES) psi[slot][k] = sp[r];
    }
  };
  // links of plane z: U_x, U_y on the tile, the two backward halo strips, U_z
  auto fetch_links = [&](int z) {
#pragma unroll
    for (int r = 0; r < LR; ++r) {
      // region of piece k: tile U_x (q < 64), tile U_y (< 128), halo column
      // U_x (< 136), halo row U_y (< 144), tile U_z (< 208); pieces past the
      // last re-read it (stored nowhere)
      const int k = min(tid + NT * r, HL_LINK_LOAD - 1);
      const int q = k / 9, pc = k - q * 9;
      const int qt = q < 64 ? q : (q < 128 ? q - 64 : q - 144);  // tile site of the tile regions
      const int mu = q < 64 ? 0 : (q < 128 ? 1 : (q < 136 ? 0 : (q < 144 ? 1 : 2)));
      const int x = q < 128 || q >= 144 ? x0 + (qt & 7) : (q < 136 ? wrap(x0 - 1, g.Lx) : x0 + (q - 136));
      const int y = q < 128 || q >= 144 ? y0 + (qt >> 3) : (q < 136 ? y0 + (q - 128) : wrap(y0 - 1, g.Ly));
      sl[r] = ((const u4*)link_ptr(mu, x, y, z))[pc];
    }
  };
  // to LDS; the plane's U_z goes to slot zslot
  auto store_links = [&](int zslot) {
#pragma unroll
    for (int r = 0; r < LR; ++r) {
      const int k = tid + NT * r;
      if (k < HL_LINK_LOAD) lk[k < HL_LZ ? k : k + zslot * 64 * 9] = __builtin_bit_cast(double2, sl[r]);
    }
  };

  // planes Lz - 1, 0, 1 in ring slots 2, 0, 1; links of plane 0, U_z of
  // plane Lz - 1 in slot 1
  fetch_psi(g.Lz - 1);
  store_psi(2);
  fetch_psi(0);
  store_psi(0);
  fetch_psi(1);
  store_psi(1);
  fetch_links(g.Lz - 1);
  store_links(1);
  __syncthreads();  // the slot-1 U_z is written; the other links of plane Lz - 1 are overwritten next
  fetch_links(0);
  store_links(0);
  __syncthreads();

  const int sub = lane / EC, e = lane - sub * EC;
  const bool active = e0 + e < g.Ne;
  for (int z = 0; z < g.Lz; ++z) {
    const bool more = z + 1 < g.Lz;
    if (more) {
      fetch_psi(z + 2 < g.Lz ? z + 2 : z + 2 - g.Lz);
      fetch_links(z + 1);
    }
    const int sz = z % 3, sf = (z + 1) % 3, sb = (z + 2) % 3;  // planes z, z + 1, z - 1
    const int zc = z & 1;                                       // U_z slot of plane z
    for (int it = 0; it < ITER; ++it) {
      const int q = (wave + NW * it) * SPI + sub;  // tile site
      const int qx = q & 7, qy = q >> 3, hx = qx + 1, hy = qy + 1;
      const int s = (z * g.Ly + y0 + qy) * g.Lx + x0 + qx;
#pragma unroll
      for (int mu = 0; mu < 3; ++mu) {
        // neighbour elements (halo coordinates, ring slot) and links
        int fh, bh, fsl, bsl, lf0, lb0;
        if (mu == 0) {
          fh = hy * HL_H + hx + 1, bh = hy * HL_H + hx - 1, fsl = bsl = sz;
          lf0 = HL_LX + q * 9, lb0 = qx > 0 ? HL_LX + (q - 1) * 9 : HL_LXH + qy * 9;
        } else if (mu == 1) {
          fh = (hy + 1) * HL_H + hx, bh = (hy - 1) * HL_H + hx, fsl = bsl = sz;
          lf0 = HL_LY + q * 9, lb0 = qy > 0 ? HL_LY + (q - 8) * 9 : HL_LYH + qx * 9;
        } else {
          fh = bh = hy * HL_H + hx, fsl = sf, bsl = sb;
          lf0 = HL_LZ + zc * 64 * 9 + q * 9, lb0 = HL_LZ + (zc ^ 1) * 64 * 9 + q * 9;
        }
        double2 af[3], ab[3];
#pragma unroll
        for (int b = 0; b < 3; ++b) {
          af[b] = wget(((const c_t<T>*)psi[fsl])[(fh * 3 + b) * EC + e], 0);
          ab[b] = wget(((const c_t<T>*)psi[bsl])[(bh * 3 + b) * EC + e], 0);
        }
        const int child = child0 + 3 * pl + mu;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          // row a of U_mu(x) and column a of U_mu(x - mu): the same LDS
          // address across the site's lanes (broadcast)
          double re = 0, im = 0;
#pragma unroll
          for (int b = 0; b < 3; ++b) {
            const double2 u = lk[lf0 + a * 3 + b];
            re = fma(u.x, af[b].x, re);
            re = fma(-u.y, af[b].y, re);
            im = fma(u.x, af[b].y, im);
            im = fma(u.y, af[b].x, im);
            const double2 v = lk[lb0 + b * 3 + a];  // (U^dag)_{ab} = conj(U_{ba})
            re = fma(-v.x, ab[b].x, re);
            re = fma(-v.y, ab[b].y, re);
            im = fma(-v.x, ab[b].y, im);
            im = fma(v.y, ab[b].x, im);
          }
          if (active) Wo[widx(g, child, s, a, e0 + e)] = g.rot ? wput<c_t<T>>(-im, re, 0) : wput<c_t<T>>(re, im, 0);
        }
      }
    }
    __syncthreads();  // plane z - 1 and this plane's links are no longer read
    if (more) {
      store_psi(sb);
      store_links(zc ^ 1);
    }
    __syncthreads();
  }
}

// --------------------------------------------------------------------------
// Straight-path transport for displacement elementals
// (lattice/generator/displacement_elemental.py:53-71): from the six paths of
// length n - 1 (slots src0 + d forward, src0 + 3 + d backward; src0 < 0: all
// six are V in slot 0) the paths of length n,
//   forward  P_d(x)  <- U_d(x) P_d(x + d)
//   backward P_-d(x) <- U_d(x - d)^dag P_-d(x - d),
// are written to dst0 + d / dst0 + 3 + d (dst0 < 0: not kept) and their mean
// over the six to slot mslot.  FP64 arithmetic, one rounding per store.
// grid (ceil(V3 / 4), ceil(Ne / 64), TB), block 256: wave per site, lane per
// eigenvector.
// --------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void disp_kernel(c_t<T>* __restrict__ W, const double2* __restrict__ U, Geom g,
                                                   int tb0, int src0, int dst0, int mslot) {
  const int s = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  const int e = blockIdx.y * 64 + (threadIdx.x & 63);
  if (s >= g.V3 || e >= g.Ne) return;
  const int tb = blockIdx.z;
  const int x = s % g.Lx, yy = (s / g.Lx) % g.Ly, z = s / (g.Lx * g.Ly);
  int sf[3], sb[3];
  sf[0] = s - x + (x + 1 == g.Lx ? 0 : x + 1);
  sb[0] = s - x + (x == 0 ? g.Lx - 1 : x - 1);
  {
    const int r = s - yy * g.Lx;
    sf[1] = r + (yy + 1 == g.Ly ? 0 : yy + 1) * g.Lx;
    sb[1] = r + (yy == 0 ? g.Ly - 1 : yy - 1) * g.Lx;
  }
  {
    const int r = s - z * g.Lx * g.Ly;
    sf[2] = r + (z + 1 == g.Lz ? 0 : z + 1) * g.Lx * g.Ly;
    sb[2] = r + (z == 0 ? g.Lz - 1 : z - 1) * g.Lx * g.Ly;
  }
  c_t<T>* Wt = W + (size_t)tb * g.nW * (size_t)g.V3P * 3 * g.NeP;
  // an hsplit workspace holds w 2^wsh; transport and mean are linear, so
  // they read and write those scaled values as they are (exponent 0 below)
  double mr[3] = {0, 0, 0}, mi[3] = {0, 0, 0};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double2* Ud = U + ((size_t)d * g.Lt + (tb0 + tb)) * (size_t)g.V3 * 9;
    const double2* uf = Ud + (size_t)s * 9;
    const double2* ub = Ud + (size_t)sb[d] * 9;
    const int slot_f = src0 < 0 ? 0 : src0 + d, slot_b = src0 < 0 ? 0 : src0 + 3 + d;
    double2 pf[3], pb[3];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      pf[b] = wget(Wt[widx(g, slot_f, sf[d], b, e)], 0);
      pb[b] = wget(Wt[widx(g, slot_b, sb[d], b, e)], 0);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      double fr = 0, fi = 0, br = 0, bi = 0;
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const double2 u = uf[a * 3 + b];
        fr += u.x * pf[b].x - u.y * pf[b].y;
        fi += u.x * pf[b].y + u.y * pf[b].x;
        const double2 v = ub[b * 3 + a];  // (U^dag)_{ab} = conj(U_{ba})
        br += v.x * pb[b].x + v.y * pb[b].y;
        bi += v.x * pb[b].y - v.y * pb[b].x;
      }
      if (dst0 >= 0) {
        Wt[widx(g, dst0 + d, s, a, e)] = wput<c_t<T>>(fr, fi, 0);
        Wt[widx(g, dst0 + 3 + d, s, a, e)] = wput<c_t<T>>(br, bi, 0);
      }
      mr[a] += fr + br;
      mi[a] += fi + bi;
    }
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) Wt[widx(g, mslot, s, a, e)] = wput<c_t<T>>(mr[a] / 6.0, mi[a] / 6.0, 0);
}

// --------------------------------------------------------------------------
// Contraction.
// --------------------------------------------------------------------------
struct SplitTerm {
  int nL, nR;
  float coeff;  // (-1)^{#right}, merged duplicates add up
  int pad;
};

struct DPlan {
  int d;            // output derivative index
  int partner;      // rev(d) written from this d's result, or -1
  int sign;         // (-1)^{|d|} for the partner
  int first, nsplit;
  int mirror;       // rev(d) = d: only tiles te <= tf run; tile (te, tf) also
                    // writes the transposed tile (tf, te) through the partner rule
  int item0;        // first work item of this d within a timeslice
  int rot;          // the FP16 stage-1 kernel multiplies its result by i^rot
};

// The diagonal tile of a self-partnered derivative computes every (e, f)
// itself and writes no mirror.
__device__ __forceinline__ bool partner_writes(const DPlan& dp, int te, int tf) {
  return dp.partner >= 0 && !(dp.mirror && te == tf);
}

struct ContractArgs {
  Geom g;
  int D, P;
  int nd;         // primary derivative count
  int npb;        // momentum blocks of 16 base momenta
  int nbase;      // base momenta used (<= 16 npb)
  int nte, ntf;   // e / f tiles of 16
  int nwork;      // TB * nd * npb * nte * ntf
  int zero_idx;   // output index of p = (0,0,0) handled on VALU, or -1
  int items_t;    // work items per timeslice: sum over d of npb x (tiles of d)
  int t_out0;     // batch offset inside out (in timeslices)
  const DPlan* dplan;
  const SplitTerm* splits;
  const int* plus_idx;   // [npb*16] output index of +b or -1
  const int* minus_idx;  // [npb*16] output index of -b or -1
  const double* wgt;     // [Ne][Ne] blending weights or nullptr
  const double* psum;    // [npb*16] sum_x Phi_b(x): V3 when b = 0 mod L, else 0
  int strip;             // e-tile rows per strip of the tile order (0: the kernel's default)
  int order;             // single-lattice item order: 0 contiguous ranges per XCD, 1 round robin
  int unit;              // FP16 stage-1 kernel: the unit plan (every coefficient +1, results times i^rot)
};

// Tile k of one (timeslice, derivative, momentum block): all nte x ntf tiles,
// or for a self-partnered derivative the nte (nte + 1) / 2 with te <= tf.
// Tiles run in strips of TS e-tile rows: within a strip f-tile by f-tile, and
// within an f-tile the strip's rows, so 32 consecutive items (one per CU of
// an XCD, started together) read about TS e slabs and 32 / TS f slabs of the
// workspace, which then hit in that XCD's L2 (TS = 4 for the mixed kernels,
// 2 for the FP64/FP32 kernel, measured, and the defaults; ContractArgs::strip
// overrides them).
constexpr int TSTRIP = 4;
constexpr int TSTRIP_FP = 2;
__host__ __device__ inline void decode_tile(int k, int nte, int ntf, bool mirror, int TS, int& te, int& tf) {
  int s0 = 0;
  if (mirror) {
    for (;;) {  // strip rows [s0, s0 + h), columns tf >= s0, te <= tf
      const int h = min(TS, nte - s0);
      const int cnt = h * (nte - s0) - h * (h - 1) / 2;
      if (k < cnt) break;
      k -= cnt;
      s0 += TS;
    }
    const int h = min(TS, nte - s0);
    tf = s0;
    for (;;) {
      const int c = min(h, tf - s0 + 1);
      if (k < c) break;
      k -= c;
      ++tf;
    }
    te = s0 + k;
  } else {
    for (;;) {
      const int cnt = min(TS, nte - s0) * ntf;
      if (k < cnt) break;
      k -= cnt;
      s0 += TS;
    }
    const int h = min(TS, nte - s0);
    tf = k / h;
    te = s0 + (k - tf * h);
  }
}

// Work item w -> (timeslice tb, derivative di, momentum block pb, tile
// (te, tf)).  Items of one timeslice run derivative by derivative
// (DPlan::item0), each over npb blocks of its tiles (decode_tile).
template <int TS = TSTRIP>
__device__ __forceinline__ void decode_item(const ContractArgs& a, int w, int& tb, int& di, int& pb, int& te,
                                            int& tf) {
  tb = w / a.items_t;
  int r = w - tb * a.items_t;
  di = 0;
  while (di + 1 < a.nd && a.dplan[di + 1].item0 <= r) ++di;
  r -= a.dplan[di].item0;
  const bool mirror = a.dplan[di].mirror;
  const int per = mirror ? a.nte * (a.nte + 1) / 2 : a.nte * a.ntf;
  pb = r / per;
  decode_tile(r - pb * per, a.nte, a.ntf, mirror, a.strip > 0 ? a.strip : TS, te, tf);
}

constexpr int SC = 16;      // sites per chunk (4 k-steps of the stage-2 MFMA)
constexpr int NWAVE = 8;    // waves per workgroup
constexpr int SPW = SC / NWAVE;   // stage-1 sites per wave and chunk
constexpr int NBW = 16 / NWAVE;   // stage-2 N-blocks (eigenvector rows e) per wave

// Workgroup id -> work item such that each XCD (workgroups are dispatched
// round-robin over 8 XCDs) receives a contiguous range of work items, so the
// tiles of one (t, d) share that XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
  const int per = nblocks >> 3;  // nblocks is a multiple of 8
  return (bid & 7) * per + (bid >> 3);
}

// Work item of workgroup bid in a single-lattice launch: contiguous ranges
// of items per XCD (the hardware hands workgroup bid to XCD bid mod 8), or
// the items round robin over the XCDs.
__device__ __forceinline__ int item_index(const ContractArgs& a, int bid, int nblocks) {
  return a.order ? bid : xcd_remap(bid, nblocks);
}

// One workgroup (8 waves, two per SIMD) owns a 16 x 16 (e, f) tile, one
// primary derivative, one block of 16 base momenta and one timeslice, and
// sweeps all sites in chunks of 16:
//   stage 1: wave w forms the Q tiles of sites SPW*w.. of the chunk on MFMA
//            (K = split x colour, NKS k-steps) and stores them to LDS;
//   stage 2: wave w contracts the 16 sites of the chunk for the NBW rows
//            e = NBW*w.. (16-pair N-blocks) against the 16 momenta (M),
//            K = sites.
// The next chunk's stage-1 operands and phases are loaded into registers
// while the current chunk is contracted; LDS is double-buffered so one
// barrier per chunk suffices.
template <typename T, int NKS>
__device__ __forceinline__ void contract_body(const c_t<T>* __restrict__ W, const T* __restrict__ phr,
                                              const T* __restrict__ phi, double2* __restrict__ out,
                                              const ContractArgs& a, const int w, c_t<T> (*Qs)[SC][16][16]) {
  typedef typename mfma<T>::acc acc_t;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int l16 = lane & 15, g4 = lane >> 4;

  int tb, di, pb, te, tf;
  decode_item<TSTRIP_FP>(a, w, tb, di, pb, te, tf);
  const DPlan dp = a.dplan[di];
  const int E0 = te * 16, F0 = tf * 16;
  const Geom g = a.g;
  const size_t site = (size_t)3 * g.NeP;  // elements per site
  const c_t<T>* Wt = W + (size_t)tb * g.nW * g.V3P * site;
  const int J = __builtin_amdgcn_readfirstlane(3 * dp.nsplit);
  const int nks = (J + 3) >> 2;  // <= NKS, uniform

  // per k-step operand bases (site 0) and coefficients; j >= J reads a valid
  // element with coefficient 0
  const c_t<T>* Lb[NKS];
  const c_t<T>* Rb[NKS];
  T cf[NKS];
#pragma unroll
  for (int ks = 0; ks < NKS; ++ks) {
    const int j = 4 * ks + g4;
    if (ks < nks && j < J) {
      const SplitTerm st = a.splits[dp.first + j / 3];
      const int c = j % 3;
      Lb[ks] = Wt + widx(g, st.nL, 0, c, E0 + l16);
      Rb[ks] = Wt + widx(g, st.nR, 0, c, F0 + l16);
      cf[ks] = (T)st.coeff;
    } else {
      Lb[ks] = Wt;
      Rb[ks] = Wt;
      cf[ks] = 0;
    }
  }

  const T* PR = phr + (size_t)pb * g.V3P * 16;
  const T* PI = phi + (size_t)pb * g.V3P * 16;

  // prefetch registers
  c_t<T> lv[NKS][SPW], rv[NKS][SPW];
  T ph_r[4], ph_i[4];
  auto load_ops = [&](int c0) {
    const int xb = c0 + SPW * wave;
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks) {
      if (ks < nks) {
#pragma unroll
        for (int i = 0; i < SPW; ++i) {
          lv[ks][i] = Lb[ks][(size_t)(xb + i) * site];
          rv[ks][i] = Rb[ks][(size_t)(xb + i) * site];
        }
      }
    }
  };
  auto load_phase = [&](int c0) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const size_t xo = (size_t)(c0 + 4 * k + g4) * 16 + l16;
      ph_r[k] = PR[xo];
      ph_i[k] = PI[xo];
    }
  };

  acc_t A1[NBW], A2[NBW], A3[NBW], A4[NBW];
  double D1[NBW][4], D2[NBW][4], D3[NBW][4], D4[NBW][4];  // FP64 cross-chunk sums (T = float)
  T zr[NBW], zi[NBW];
  double Zr[NBW], Zi[NBW];
#pragma unroll
  for (int b = 0; b < NBW; ++b) {
    A1[b] = A2[b] = A3[b] = A4[b] = acc_t{0, 0, 0, 0};
    zr[b] = zi[b] = 0;
    Zr[b] = Zi[b] = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) D1[b][r] = D2[b][r] = D3[b][r] = D4[b][r] = 0;
  }
  // FP32 form: Q_d[e,f,x] often carries a large site-independent part (for
  // d = 0, |V_e(x)|^2 >= 0).  Stage 2 works on Q - c with c the mean over the
  // first chunk, and adds c * sum_x Phi_b(x) back in FP64, so FP32 rounding
  // scales with the fluctuation of Q rather than with its mean.
  T cr[NBW], ci[NBW];
#pragma unroll
  for (int b = 0; b < NBW; ++b) cr[b] = ci[b] = 0;

  // stage 1 of the chunk at c0 from the prefetched operands into Qs[buf];
  // the operands of the next chunk are requested right after use
  auto stage1 = [&](int c0, int buf) {
    acc_t qr[SPW], qi[SPW];
    const acc_t zero = {0, 0, 0, 0};
    T lr[SPW], li[SPW];
#pragma unroll
    for (int i = 0; i < SPW; ++i) {
      lr[i] = cf[0] * lv[0][i].x;
      li[i] = cf[0] * lv[0][i].y;
    }
    // conj(L) R = (lr rr + li ri) + i (lr ri - li rr); dependent MFMAs are
    // separated by independent ones
#pragma unroll
    for (int i = 0; i < SPW; ++i) qr[i] = mfma<T>::mac(lr[i], rv[0][i].x, zero);
#pragma unroll
    for (int i = 0; i < SPW; ++i) qi[i] = mfma<T>::mac(lr[i], rv[0][i].y, zero);
#pragma unroll
    for (int i = 0; i < SPW; ++i) qr[i] = mfma<T>::mac(li[i], rv[0][i].y, qr[i]);
#pragma unroll
    for (int i = 0; i < SPW; ++i) qi[i] = mfma<T>::mac(-li[i], rv[0][i].x, qi[i]);
#pragma unroll
    for (int ks = 1; ks < NKS; ++ks) {
      if (ks < nks) {
#pragma unroll
        for (int i = 0; i < SPW; ++i) {
          lr[i] = cf[ks] * lv[ks][i].x;
          li[i] = cf[ks] * lv[ks][i].y;
        }
#pragma unroll
        for (int i = 0; i < SPW; ++i) qr[i] = mfma<T>::mac(lr[i], rv[ks][i].x, qr[i]);
#pragma unroll
        for (int i = 0; i < SPW; ++i) qi[i] = mfma<T>::mac(lr[i], rv[ks][i].y, qi[i]);
#pragma unroll
        for (int i = 0; i < SPW; ++i) qr[i] = mfma<T>::mac(li[i], rv[ks][i].y, qr[i]);
#pragma unroll
        for (int i = 0; i < SPW; ++i) qi[i] = mfma<T>::mac(-li[i], rv[ks][i].x, qi[i]);
      }
    }
    if (c0 + SC < g.V3P) load_ops(c0 + SC);
#pragma unroll
    for (int i = 0; i < SPW; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) Qs[buf][SPW * wave + i][mfma<T>::row(g4, r)][l16] = mk<T>(qr[i][r], qi[i][r]);
  };

  // Software pipeline: the barrier interval of chunk k holds stage 2 of
  // chunk k (reading Qs[k & 1]) and stage 1 of chunk k + 1 (writing
  // Qs[(k + 1) & 1]).
  load_ops(0);
  load_phase(0);
  stage1(0, 0);
  __syncthreads();
  if constexpr (sizeof(T) == 4) {
    const int n0 = min(SC, g.V3);
#pragma unroll
    for (int b = 0; b < NBW; ++b) {
      T sr = 0, si = 0;
      for (int xl = 0; xl < n0; ++xl) {
        const c_t<T> q = Qs[0][xl][NBW * wave + b][l16];
        sr += q.x;
        si += q.y;
      }
      cr[b] = sr / (T)n0;
      ci[b] = si / (T)n0;
    }
  }
  for (int c0 = 0; c0 < g.V3P; c0 += SC) {
    const int buf = (c0 / SC) & 1;
    if (c0 + SC < g.V3P) stage1(c0 + SC, buf ^ 1);

    // ---- stage 2: phase contraction over the 16 sites of the chunk ----
#pragma unroll
    for (int k = 0; k < SC / 4; ++k) {
      const T pr = ph_r[k], pi = ph_i[k];
      const bool real_site = (c0 + 4 * k + g4) < g.V3;
#pragma unroll
      for (int b = 0; b < NBW; ++b) {
        c_t<T> qq = Qs[buf][4 * k + g4][NBW * wave + b][l16];
        if constexpr (sizeof(T) == 4) {
          qq.x = real_site ? qq.x - cr[b] : (T)0;
          qq.y = real_site ? qq.y - ci[b] : (T)0;
        }
        A1[b] = mfma<T>::mac(pr, qq.x, A1[b]);
        A2[b] = mfma<T>::mac(pi, qq.y, A2[b]);
        A3[b] = mfma<T>::mac(pi, qq.x, A3[b]);
        A4[b] = mfma<T>::mac(pr, qq.y, A4[b]);
        zr[b] += qq.x;
        zi[b] += qq.y;
      }
    }
    if (c0 + SC < g.V3P) load_phase(c0 + SC);

    if constexpr (sizeof(T) == 4) {
      // FP32 partial sums of one chunk are moved into FP64 accumulators.
#pragma unroll
      for (int b = 0; b < NBW; ++b) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          D1[b][r] += (double)A1[b][r];
          D2[b][r] += (double)A2[b][r];
          D3[b][r] += (double)A3[b][r];
          D4[b][r] += (double)A4[b][r];
        }
        A1[b] = A2[b] = A3[b] = A4[b] = acc_t{0, 0, 0, 0};
        Zr[b] += (double)zr[b];
        Zi[b] += (double)zi[b];
        zr[b] = zi[b] = 0;
      }
    }
    __syncthreads();
  }
  if constexpr (sizeof(T) == 8) {
#pragma unroll
    for (int b = 0; b < NBW; ++b) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        D1[b][r] = (double)A1[b][r];
        D2[b][r] = (double)A2[b][r];
        D3[b][r] = (double)A3[b][r];
        D4[b][r] = (double)A4[b][r];
      }
      Zr[b] = (double)zr[b];
      Zi[b] = (double)zi[b];
    }
  }

  // ---- epilogue ----
  const int Ne = g.Ne, P = a.P;
  const size_t plane = (size_t)Ne * Ne;
  const size_t tbase = (size_t)(a.t_out0 + tb) * a.D * P;
  const int f = F0 + l16;
#pragma unroll
  for (int b = 0; b < NBW; ++b) {
    const int e = E0 + NBW * wave + b;
    const bool ok = (e < Ne) && (f < Ne);
    const double wef = (ok && a.wgt) ? a.wgt[(size_t)e * Ne + f] : 1.0;
    const double wfe = (ok && a.wgt) ? a.wgt[(size_t)f * Ne + e] : 1.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int bi = pb * 16 + mfma<T>::row(g4, r);
      if (!ok || bi >= a.nbase) continue;
      const int ip = a.plus_idx[bi], im = a.minus_idx[bi];
      const double S = a.psum[bi], mr = (double)cr[b] * S, mi = (double)ci[b] * S;
      const double epr = D1[b][r] - D2[b][r] + mr, epi = D3[b][r] + D4[b][r] + mi;  // E_d[+b]
      const double emr = D1[b][r] + D2[b][r] + mr, emi = D4[b][r] - D3[b][r] + mi;  // E_d[-b]
      if (ip >= 0) out[(tbase + (size_t)dp.d * P + ip) * plane + (size_t)e * Ne + f] = make_double2(epr * wef, epi * wef);
      if (im >= 0) out[(tbase + (size_t)dp.d * P + im) * plane + (size_t)e * Ne + f] = make_double2(emr * wef, emi * wef);
      if (partner_writes(dp, te, tf)) {
        // E_{rev d}[p'] = sign * E_d[-p']^dagger
        const double s = (double)dp.sign;
        if (im >= 0)
          out[(tbase + (size_t)dp.partner * P + im) * plane + (size_t)f * Ne + e] =
              make_double2(s * epr * wfe, -s * epi * wfe);
        if (ip >= 0)
          out[(tbase + (size_t)dp.partner * P + ip) * plane + (size_t)f * Ne + e] =
              make_double2(s * emr * wfe, -s * emi * wfe);
      }
    }
    // zero momentum: sum the four row groups (sites 4k + g4) of each lane column
    double zre = Zr[b], zim = Zi[b];
    zre += __shfl_xor(zre, 16);
    zim += __shfl_xor(zim, 16);
    zre += __shfl_xor(zre, 32);
    zim += __shfl_xor(zim, 32);
    zre += (double)cr[b] * g.V3;
    zim += (double)ci[b] * g.V3;
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

// --------------------------------------------------------------------------
// Mixed form with the phase contraction on FP16 MFMA.
//
// Stage 1 is the FP32 MFMA of contract_kernel<float> (4 sites per wave,
// 32-site chunks).  Stage 2 runs v_mfma_f32_16x16x32_f16 (K = 32 sites) on a
// three-term split: with power-of-two scales sq (Q) and 2^14 (Phi), each
// operand x is held as hi = fp16(x s), lo = fp16(x s - hi), and
//   Phi Q ~ (Phi_hi Q_hi + Phi_hi Q_lo + Phi_lo Q_hi) / (2^14 sq),
// i.e. a 22-bit representation of both operands with FP32 accumulation over
// the 32 sites of a chunk, then FP64 across chunks.  sq is chosen per chunk
// from the tile maximum of |Q| (gathered through LDS before the barrier), so
// no value can overflow FP16 and all values within 2^-17 of the maximum keep
// their full 22 bits.
// --------------------------------------------------------------------------
constexpr int SC2 = 32;            // sites per chunk (one K = 32 MFMA)
constexpr int SPW2 = SC2 / NWAVE;  // stage-1 sites per wave and chunk
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float float4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void split16(float x, _Float16& hi, _Float16& lo) {
  hi = (_Float16)x;
  lo = (_Float16)(x - (float)hi);
}

// wave maximum of a non-negative value through DPP row shifts and row
// broadcasts (no LDS traffic); every lane receives the result
template <int CTRL, int RMASK>
__device__ __forceinline__ float dpp_max_step(float x) {
  const int y = __builtin_amdgcn_update_dpp(__float_as_int(x), __float_as_int(x), CTRL, RMASK, 0xF, false);
  return fmaxf(x, __int_as_float(y));
}
__device__ __forceinline__ float wave_max(float v) {
  v = dpp_max_step<0x111, 0xF>(v);  // row_shr:1
  v = dpp_max_step<0x112, 0xF>(v);  // row_shr:2
  v = dpp_max_step<0x114, 0xF>(v);  // row_shr:4
  v = dpp_max_step<0x118, 0xF>(v);  // row_shr:8
  v = dpp_max_step<0x142, 0xA>(v);  // row_bcast:15
  v = dpp_max_step<0x143, 0xC>(v);  // row_bcast:31
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// Three-term split of two FP32 values: hi = fp16(x) (round to nearest even,
// one packed v_cvt_pk_f16_f32), lo = fp16(x - hi) with the residual formed
// exactly inside one v_fma_mix per value (fma(hi, m1, x) with m1 = -1 a
// runtime value, so the compiler keeps the fused form instead of folding it
// to a subtraction), reading hi's two FP16 halves by operand selection:
// 3 instructions per pair.  The two asm statements are empty (they emit no
// instruction): they make x and hi opaque, so the compiler neither fuses the
// conversion with the FMA that produced x (computing x a second time) nor
// converts x0 and x1 to FP16 again for the residual.  Every instruction is
// the compiler's own, so it inserts the wait states a vector write before a
// matrix read needs.
__device__ __forceinline__ void split16x2(float x0, float x1, float m1, uint32_t& hi, uint32_t& lo) {
  typedef _Float16 h2 __attribute__((ext_vector_type(2)));
  asm volatile("" : "+v"(x0), "+v"(x1));
  uint32_t hb = __builtin_bit_cast(uint32_t, h2{(_Float16)x0, (_Float16)x1});
  asm volatile("" : "+v"(hb));
  hi = hb;
  const h2 h = __builtin_bit_cast(h2, hb);
  lo = __builtin_bit_cast(uint32_t, h2{(_Float16)__builtin_fmaf((float)h.x, m1, x0), (_Float16)__builtin_fmaf((float)h.y, m1, x1)});
}

// -1.0f from a kernel argument (Ne > 0): a value the compiler cannot fold
__device__ __forceinline__ float runtime_m1(const Geom& g) { return g.Ne > 0 ? -1.f : 1.f; }

template <int NKS>
__device__ __forceinline__ void contract_mixed_body(const float2* __restrict__ W, const half8* __restrict__ ph16,
                                                    double2* __restrict__ out, const ContractArgs& a, const int w,
                                                    float2 (*Qs)[SC2][16][16], float (*wmax)[NWAVE]) {
  typedef float4v acc_t;
  typedef float v2f __attribute__((ext_vector_type(2)));
  typedef uint32_t u4 __attribute__((ext_vector_type(4)));
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int l16 = lane & 15, g4 = lane >> 4;

  int tb, di, pb, te, tf;
  decode_item(a, w, tb, di, pb, te, tf);
  const DPlan dp = a.dplan[di];
  const int E0 = te * 16, F0 = tf * 16;
  const Geom g = a.g;
  const uint32_t site_b = 3u * (uint32_t)g.NeP * 8u;  // bytes per site
  const size_t slice_b = (size_t)g.nW * g.V3P * site_b;  // < 2^31 (the host runs this kernel only then)
  const float2* Wt = W + (size_t)tb * (slice_b / 8);
  const __amdgpu_buffer_rsrc_t wr = __builtin_amdgcn_make_buffer_rsrc((void*)Wt, (short)0, (int)slice_b, 0x00020000);
  const int J = __builtin_amdgcn_readfirstlane(3 * dp.nsplit);
  const int nks = (J + 3) >> 2;  // uniform
  const int nchunk = g.V3P / SC2;

  // per k-step byte offsets of site 0 (within the timeslice block) and
  // coefficients; unused (split, colour) slots and padded rows read past the
  // buffer's end (0, no memory request)
  const int oob = (int)slice_b;
  int offL[NKS], offR[NKS];
  float cf[NKS];
#pragma unroll
  for (int ks = 0; ks < NKS; ++ks) {
    const int j = 4 * ks + g4;
    if (ks < nks && j < J) {
      const SplitTerm st = a.splits[dp.first + j / 3];
      const int c = j % 3;
      offL[ks] = woff8(g, st.nL, c, E0 + l16, oob);
      offR[ks] = woff8(g, st.nR, c, F0 + l16, oob);
      cf[ks] = st.coeff;
    } else {
      offL[ks] = offR[ks] = oob;
      cf[ks] = 0.f;
    }
  }
  // phase operands: [npb][nchunk][4 parts][16 momenta][32 sites], parts =
  // (Re hi, Re lo, Im hi, Im lo); lane (g4, l16) reads momentum l16, sites 8 g4..8 g4+7
  const half8* PH = ph16 + (size_t)pb * nchunk * 4 * 16 * 4;
  const __amdgpu_buffer_rsrc_t pr_ = __builtin_amdgcn_make_buffer_rsrc((void*)PH, (short)0, nchunk * 4 * 16 * 4 * 16, 0x00020000);
  const int phoff = (l16 * 4 + g4) * 16;

  v2f lv[NKS][SPW2], rv[NKS][SPW2];
  half8 prh, prl, pih, pil;
  auto load_ops = [&](int c0) {
    const int xo = (c0 + SPW2 * wave) * (int)site_b;  // uniform
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks) {
      if (ks < nks) {
#pragma unroll
        for (int i = 0; i < SPW2; ++i) {
          lv[ks][i] = __builtin_bit_cast(v2f, __builtin_amdgcn_raw_buffer_load_b64(wr, offL[ks], xo + i * (int)site_b, 0));
          rv[ks][i] = __builtin_bit_cast(v2f, __builtin_amdgcn_raw_buffer_load_b64(wr, offR[ks], xo + i * (int)site_b, 0));
        }
      }
    }
  };
  auto load_phase = [&](int chunk) {
    const int co = chunk * 4 * 16 * 4 * 16;  // uniform
    prh = __builtin_bit_cast(half8, __builtin_amdgcn_raw_buffer_load_b128(pr_, phoff, co + 0 * 1024, 0));
    prl = __builtin_bit_cast(half8, __builtin_amdgcn_raw_buffer_load_b128(pr_, phoff, co + 1 * 1024, 0));
    pih = __builtin_bit_cast(half8, __builtin_amdgcn_raw_buffer_load_b128(pr_, phoff, co + 2 * 1024, 0));
    pil = __builtin_bit_cast(half8, __builtin_amdgcn_raw_buffer_load_b128(pr_, phoff, co + 3 * 1024, 0));
  };

  const acc_t zero = {0.f, 0.f, 0.f, 0.f};
  double D1[NBW][4], D2[NBW][4], D3[NBW][4], D4[NBW][4];
  double Zr[NBW], Zi[NBW];
  float cr[NBW], ci[NBW];
#pragma unroll
  for (int b = 0; b < NBW; ++b) {
    Zr[b] = Zi[b] = 0;
    cr[b] = ci[b] = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) D1[b][r] = D2[b][r] = D3[b][r] = D4[b][r] = 0;
  }

  // stage 2 of one chunk for N-block b.  Padded sites hold Q = 0 and Phi = 0:
  // their (0 - c) sq terms meet a zero phase and only enter the p = 0 sum,
  // which the epilogue corrects with V3P in place of V3.
  const float m1 = runtime_m1(a.g);
  auto stage2 = [&](int b, int c0, float sq, int ex) {
    const float csr = -cr[b] * sq, csi = -ci[b] * sq;
    uint32_t qrh[4], qrl[4], qih[4], qil[4];
    float zsr = 0.f, zsi = 0.f;
#pragma unroll
    for (int v2 = 0; v2 < 4; ++v2) {
      float sr[2], si[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int xl = 8 * g4 + 2 * v2 + h;
        const float2 q = Qs[(c0 / SC2) & 1][xl][NBW * wave + b][l16];
        // scaled and mean-subtracted value s = (q - c) sq; padded sites give 0
        sr[h] = fmaf(q.x, sq, csr);
        si[h] = fmaf(q.y, sq, csi);
        zsr += sr[h];
        zsi += si[h];
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
    // FP32 chunk sums into FP64, unscaled exactly (powers of two)
    const double inv = ldexp(1.0, ex - 28);  // 1 / (2^14 sq)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      D1[b][r] = fma((double)A1[r], inv, D1[b][r]);
      D2[b][r] = fma((double)A2[r], inv, D2[b][r]);
      D3[b][r] = fma((double)A3[r], inv, D3[b][r]);
      D4[b][r] = fma((double)A4[r], inv, D4[b][r]);
    }
    const double invq = ldexp(1.0, ex - 14);  // 1 / sq
    Zr[b] = fma((double)zsr, invq, Zr[b]);
    Zi[b] = fma((double)zsi, invq, Zi[b]);
  };

  // stage 1 of chunk `chunk` from the prefetched operands into Qs[buf] and
  // wmax[buf]; the operands of chunk + 1 are requested right after use
  auto stage1 = [&](int chunk, int buf) {
    const int c0 = chunk * SC2;
    acc_t qr[SPW2], qi[SPW2];
    float lr[SPW2], li[SPW2];
#pragma unroll
    for (int i = 0; i < SPW2; ++i) {
      lr[i] = cf[0] * lv[0][i].x;
      li[i] = cf[0] * lv[0][i].y;
    }
    // conj(L) R = (lr rr + li ri) + i (lr ri - li rr); dependent MFMAs are
    // separated by independent ones
#pragma unroll
    for (int i = 0; i < SPW2; ++i) qr[i] = mfma<float>::mac(lr[i], rv[0][i].x, zero);
#pragma unroll
    for (int i = 0; i < SPW2; ++i) qi[i] = mfma<float>::mac(lr[i], rv[0][i].y, zero);
#pragma unroll
    for (int i = 0; i < SPW2; ++i) qr[i] = mfma<float>::mac(li[i], rv[0][i].y, qr[i]);
#pragma unroll
    for (int i = 0; i < SPW2; ++i) qi[i] = mfma<float>::mac(-li[i], rv[0][i].x, qi[i]);
#pragma unroll
    for (int ks = 1; ks < NKS; ++ks) {
      if (ks < nks) {
#pragma unroll
        for (int i = 0; i < SPW2; ++i) {
          lr[i] = cf[ks] * lv[ks][i].x;
          li[i] = cf[ks] * lv[ks][i].y;
        }
#pragma unroll
        for (int i = 0; i < SPW2; ++i) qr[i] = mfma<float>::mac(lr[i], rv[ks][i].x, qr[i]);
#pragma unroll
        for (int i = 0; i < SPW2; ++i) qi[i] = mfma<float>::mac(lr[i], rv[ks][i].y, qi[i]);
#pragma unroll
        for (int i = 0; i < SPW2; ++i) qr[i] = mfma<float>::mac(li[i], rv[ks][i].y, qr[i]);
#pragma unroll
        for (int i = 0; i < SPW2; ++i) qi[i] = mfma<float>::mac(-li[i], rv[ks][i].x, qi[i]);
      }
    }
    if (chunk + 1 < nchunk) load_ops(c0 + SC2);
    const bool full = c0 + SC2 <= g.V3;  // uniform: no padded site in this chunk
    float m = 0.f;
#pragma unroll
    for (int i = 0; i < SPW2; ++i) {
#pragma unroll
      for (int r = 0; r < 4; ++r) Qs[buf][SPW2 * wave + i][mfma<float>::row(g4, r)][l16] = make_float2(qr[i][r], qi[i][r]);
      if (full || (c0 + SPW2 * wave + i) < g.V3) {
        m = fmaxf(m, fmaxf(fmaxf(fabsf(qr[i][0]), fabsf(qi[i][0])), fmaxf(fabsf(qr[i][1]), fabsf(qi[i][1]))));
        m = fmaxf(m, fmaxf(fmaxf(fabsf(qr[i][2]), fabsf(qi[i][2])), fmaxf(fabsf(qr[i][3]), fabsf(qi[i][3]))));
      }
    }
    m = wave_max(m);
    if (lane == 0) wmax[buf][wave] = m;
  };

  // Software pipeline: the barrier interval of chunk k holds stage 2 of
  // chunk k (reading Qs[k & 1]) and stage 1 of chunk k + 1 (writing
  // Qs[(k + 1) & 1]), so the MFMA pipe has independent work while LDS reads
  // and conversions of the other stage are in flight.
  load_ops(0);
  load_phase(0);
  stage1(0, 0);
  __syncthreads();
  {
    const int n0 = min(SC2, g.V3);
#pragma unroll
    for (int b = 0; b < NBW; ++b) {
      float sr = 0, si = 0;
      for (int xl = 0; xl < n0; ++xl) {
        const float2 q = Qs[0][xl][NBW * wave + b][l16];
        sr += q.x;
        si += q.y;
      }
      cr[b] = sr / (float)n0;
      ci[b] = si / (float)n0;
    }
  }
  for (int chunk = 0; chunk < nchunk; ++chunk) {
    const int c0 = chunk * SC2, buf = chunk & 1;
    // chunk scale: |Q - c| <= 2 max|Q| < 2^ex, scaled values stay below 2^14
    float tmax = 0.f;
#pragma unroll
    for (int v = 0; v < NWAVE; ++v) tmax = fmaxf(tmax, wmax[buf][v]);
    int ex = 0;
    if (tmax > 0.f) (void)frexpf(2.f * tmax, &ex);
    const float sq = ldexpf(1.f, 14 - ex);

    if (chunk + 1 < nchunk) stage1(chunk + 1, buf ^ 1);

    // ---- stage 2 (FP16 MFMA, three-term split) of chunk ----
#pragma unroll
    for (int b = 0; b < NBW; ++b) stage2(b, c0, sq, ex);
    if (chunk + 1 < nchunk) load_phase(chunk + 1);
    __syncthreads();
  }

  // ---- epilogue (as contract_kernel) ----
  const int Ne = g.Ne, P = a.P;
  const size_t plane = (size_t)Ne * Ne;
  const size_t tbase = (size_t)(a.t_out0 + tb) * a.D * P;
  const int f = F0 + l16;
#pragma unroll
  for (int b = 0; b < NBW; ++b) {
    const int e = E0 + NBW * wave + b;
    const bool ok = (e < Ne) && (f < Ne);
    const double wef = (ok && a.wgt) ? a.wgt[(size_t)e * Ne + f] : 1.0;
    const double wfe = (ok && a.wgt) ? a.wgt[(size_t)f * Ne + e] : 1.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int bi = pb * 16 + mfma<float>::row(g4, r);
      if (!ok || bi >= a.nbase) continue;
      const int ip = a.plus_idx[bi], im = a.minus_idx[bi];
      const double S = a.psum[bi], mr = (double)cr[b] * S, mi = (double)ci[b] * S;
      const double epr = D1[b][r] - D2[b][r] + mr, epi = D3[b][r] + D4[b][r] + mi;
      const double emr = D1[b][r] + D2[b][r] + mr, emi = D4[b][r] - D3[b][r] + mi;
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
    double zre = Zr[b], zim = Zi[b];
    zre += __shfl_xor(zre, 16);
    zim += __shfl_xor(zim, 16);
    zre += __shfl_xor(zre, 32);
    zim += __shfl_xor(zim, 32);
    zre += (double)cr[b] * g.V3P;  // the V3P - V3 padded sites entered as (0 - c)
    zim += (double)ci[b] * g.V3P;
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

// --------------------------------------------------------------------------
// Mixed form, warp-specialized.  Waves 0..3 (producers) run stage 1 -- the
// FP32 MFMA split sum, 8 sites each per 32-site chunk -- and publish Q and
// its tile maximum in LDS; waves 4..7 (consumers) run stage 2 -- scaling,
// FP16 hi/lo split and FP16 MFMA phase contraction, 4 eigenvector rows each.
// Wave w and wave w + 4 share a SIMD (a SIMD issues no vector ALU work while
// it runs FP32 16x16x4 MFMAs, so the two roles interleave rather than
// overlap there); what bounds the kernel is the producers' workspace loads,
// the CU's load path.  The roles keep separate loops with the same number of
// barriers (nchunk + 2), so neither role's registers are live in the other.
// Numerics, LDS layout and epilogue are those of contract_mixed_body.
// --------------------------------------------------------------------------
constexpr int NPROD = 4;                   // producer waves
// LDS row-major Q tile per site, padded by 16 bytes: the consumer's four lane
// groups read sites 8 apart, 8 x 2064 bytes = 128 (mod 256), so they fall in
// alternate halves of the 64 banks
constexpr int QSITE = 16 * 16 + 2;
constexpr int SPP = SC2 / NPROD;           // stage-1 sites per producer wave and chunk
constexpr int NBC = 16 / (NWAVE - NPROD);  // stage-2 rows per consumer wave

// integer maximum over the wave of a non-negative float (bit patterns of
// non-negative floats order like the floats); every lane receives it
template <int CTRL, int RMASK>
__device__ __forceinline__ int dpp_imax_step(int x) {
  const int y = __builtin_amdgcn_update_dpp(x, x, CTRL, RMASK, 0xF, false);
  return max(x, y);
}
__device__ __forceinline__ float wave_max_nonneg(float v) {
  int x = __float_as_int(v);
  x = dpp_imax_step<0x111, 0xF>(x);
  x = dpp_imax_step<0x112, 0xF>(x);
  x = dpp_imax_step<0x114, 0xF>(x);
  x = dpp_imax_step<0x118, 0xF>(x);
  x = dpp_imax_step<0x142, 0xA>(x);
  x = dpp_imax_step<0x143, 0xC>(x);
  return __int_as_float(__builtin_amdgcn_readlane(x, 63));
}

// Producer loop of contract_mixed_ws_body for exactly NK stage-1 k-steps:
// no branch inside the chunk loop, so the loads of the next chunk (issued as
// soon as this chunk's copy of each operand is consumed; the last chunk
// re-reads itself) stay in flight across the MFMAs and the barrier.
// WIDE = false: 32-bit buffer offsets into one timeslice's workspace (below
// 2^31 bytes); WIDE = true: 64-bit lane addresses, any workspace size.
template <int NK, bool WIDE>
__device__ __forceinline__ void mixed_ws_producer(const float2* __restrict__ W, const ContractArgs& a, const DPlan& dp,
                                                  int tb, int E0, int F0, float2 (*Qs)[SC2][QSITE],
                                                  float (*wmax)[NWAVE]) {
  typedef float4v acc_t;
  typedef float v2f __attribute__((ext_vector_type(2)));
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int l16 = lane & 15, g4 = lane >> 4;
  const Geom g = a.g;
  const int nchunk = g.V3P / SC2;
  const acc_t zero = {0.f, 0.f, 0.f, 0.f};
  const uint32_t site_b = 3u * (uint32_t)g.NeP * 8u;     // bytes per site
  const size_t slice_b = (size_t)g.nW * g.V3P * site_b;  // < 2^31 unless WIDE (chosen on the host)
  const float2* Wt = W + (size_t)tb * (slice_b / 8);
  const __amdgpu_buffer_rsrc_t wr =
      __builtin_amdgcn_make_buffer_rsrc((void*)Wt, (short)0, WIDE ? 0 : (int)slice_b, 0x00020000);
  const int J = 3 * dp.nsplit;
  // 32-bit form: unused (split, colour) slots and padded rows read past the
  // buffer's end (0, no memory request)
  const int oob = (int)slice_b;
  int offL[NK], offR[NK];
  // wide form: every lane reads workspace values.  A padded row (e >= Ne)
  // reads row Ne - 1, which lies in the same tile: its products are dropped
  // at the store and leave the tile maximum unchanged.  An unused slot reads
  // the first slot's operands with coefficient 0 (exact zero products, so the
  // sums equal the 32-bit form's bit for bit).
  const float2* pL[NK];
  const float2* pR[NK];
  const int site_e = 3 * g.NeP;  // elements per site
  float cf[NK];
#pragma unroll
  for (int ks = 0; ks < NK; ++ks) {
    const int j = 4 * ks + g4;
    if (WIDE) {
      const bool use = j < J;
      const SplitTerm st = a.splits[dp.first + (use ? j / 3 : 0)];
      const int c = use ? j % 3 : 0;
      pL[ks] = Wt + widx(g, st.nL, 0, c, min(E0 + l16, g.Ne - 1));
      pR[ks] = Wt + widx(g, st.nR, 0, c, min(F0 + l16, g.Ne - 1));
      cf[ks] = use ? st.coeff : 0.f;
      offL[ks] = offR[ks] = 0;
    } else if (j < J) {
      const SplitTerm st = a.splits[dp.first + j / 3];
      const int c = j % 3;
      offL[ks] = woff8(g, st.nL, c, E0 + l16, oob);
      offR[ks] = woff8(g, st.nR, c, F0 + l16, oob);
      cf[ks] = st.coeff;
    } else {
      offL[ks] = offR[ks] = oob;
      cf[ks] = 0.f;
    }
  }
  v2f lv[NK][SPP], rv[NK][SPP];
  auto load = [&](int ks, bool left, int c0) {
    if constexpr (WIDE) {
      const size_t xo = (size_t)(c0 + SPP * wave) * site_e;  // uniform
#pragma unroll
      for (int i = 0; i < SPP; ++i) {
        if (left)
          lv[ks][i] = *(const v2f*)(pL[ks] + xo + i * site_e);
        else
          rv[ks][i] = *(const v2f*)(pR[ks] + xo + i * site_e);
      }
      return;
    }
    const int xo = (c0 + SPP * wave) * (int)site_b;  // uniform
#pragma unroll
    for (int i = 0; i < SPP; ++i) {
      if (left)
        lv[ks][i] = __builtin_bit_cast(v2f, __builtin_amdgcn_raw_buffer_load_b64(wr, offL[ks], xo + i * (int)site_b, 0));
      else
        rv[ks][i] = __builtin_bit_cast(v2f, __builtin_amdgcn_raw_buffer_load_b64(wr, offR[ks], xo + i * (int)site_b, 0));
    }
  };
  // zero-momentum sums of this wave's sites, rows 4 g4 + r, column l16, of
  // Q - z0 with z0 this lane's mean over its sites of chunk 0 (added back in
  // FP64 at the end), so the FP32 partial sums carry only fluctuations
  double zr[4] = {0, 0, 0, 0}, zi[4] = {0, 0, 0, 0};
  float z0r[4] = {0.f, 0.f, 0.f, 0.f}, z0i[4] = {0.f, 0.f, 0.f, 0.f};

  auto stage1 = [&](int chunk, int buf, bool first) {
    const int c0n = min(chunk + 1, nchunk - 1) * SC2;  // the last chunk re-reads itself
    acc_t qr[SPP], qi[SPP];
    float lr[SPP], li[SPP];
#pragma unroll
    for (int ks = 0; ks < NK; ++ks) {
#pragma unroll
      for (int i = 0; i < SPP; ++i) {
        lr[i] = cf[ks] * lv[ks][i].x;
        li[i] = cf[ks] * lv[ks][i].y;
      }
      load(ks, true, c0n);
      // conj(L) R = (lr rr + li ri) + i (lr ri - li rr)
#pragma unroll
      for (int i = 0; i < SPP; ++i) qr[i] = mfma<float>::mac(lr[i], rv[ks][i].x, ks ? qr[i] : zero);
#pragma unroll
      for (int i = 0; i < SPP; ++i) qi[i] = mfma<float>::mac(lr[i], rv[ks][i].y, ks ? qi[i] : zero);
#pragma unroll
      for (int i = 0; i < SPP; ++i) qr[i] = mfma<float>::mac(li[i], rv[ks][i].y, qr[i]);
#pragma unroll
      for (int i = 0; i < SPP; ++i) qi[i] = mfma<float>::mac(-li[i], rv[ks][i].x, qi[i]);
      load(ks, false, c0n);
    }
    // Q to LDS, the tile maximum and the zero-momentum sums (padded sites
    // hold Q = 0 exactly)
    if (first) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float ar = 0.f, ai = 0.f;
#pragma unroll
        for (int i = 0; i < SPP; ++i) {
          ar += qr[i][r];
          ai += qi[i][r];
        }
        z0r[r] = ar * (1.f / SPP);
        z0i[r] = ai * (1.f / SPP);
      }
    }
    float m = 0.f;
    float sr[4] = {0.f, 0.f, 0.f, 0.f}, si[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0; i < SPP; ++i) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        Qs[buf][SPP * wave + i][mfma<float>::row(g4, r) * 16 + l16] = make_float2(qr[i][r], qi[i][r]);
        sr[r] += qr[i][r] - z0r[r];
        si[r] += qi[i][r] - z0i[r];
      }
      m = fmaxf(fmaxf(m, fabsf(qr[i][0])), fabsf(qi[i][0]));
      m = fmaxf(fmaxf(m, fabsf(qr[i][1])), fabsf(qi[i][1]));
      m = fmaxf(fmaxf(m, fabsf(qr[i][2])), fabsf(qi[i][2]));
      m = fmaxf(fmaxf(m, fabsf(qr[i][3])), fabsf(qi[i][3]));
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      zr[r] += (double)sr[r];
      zi[r] += (double)si[r];
    }
    m = wave_max_nonneg(m);
    if (lane == 0) wmax[buf][wave] = m;
  };

  // issue priority over the consumer wave of the same SIMD: the matrix pipe
  // is kept fed by the split-sum MFMAs, the consumer fills VALU slots
  __builtin_amdgcn_s_setprio(2);
#pragma unroll
  for (int ks = 0; ks < NK; ++ks) {
    load(ks, true, 0);
    load(ks, false, 0);
  }
  stage1(0, 0, true);
  __syncthreads();
  for (int chunk = 0; chunk + 1 < nchunk; ++chunk) {
    stage1(chunk + 1, (chunk + 1) & 1, false);
    __syncthreads();
  }
  __syncthreads();  // matches the consumers' last chunk
  // every consumer has read its last chunk: LDS buffer 0 takes the four
  // producers' zero-momentum partial sums, rows x columns, as doubles
  double2* zs = (double2*)(void*)&Qs[0][0][0];
  const double nsum = (double)SPP * nchunk;  // sites this wave summed, padded ones (Q = 0) included
#pragma unroll
  for (int r = 0; r < 4; ++r)
    zs[(wave * 16 + mfma<float>::row(g4, r)) * 16 + l16] =
        make_double2(fma((double)z0r[r], nsum, zr[r]), fma((double)z0i[r], nsum, zi[r]));
  __syncthreads();
}

template <int NKS, bool WIDE = false>
__device__ __forceinline__ void contract_mixed_ws_body(const float2* __restrict__ W, const half8* __restrict__ ph16,
                                                       double2* __restrict__ out, const ContractArgs& a, const int w,
                                                       float2 (*Qs)[SC2][QSITE], float (*wmax)[NWAVE],
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
    // producer: one branch-free instance per stage-1 k-step count
    const int nks = (3 * dp.nsplit + 3) >> 2;
    if (NKS >= 3 && nks == 3) mixed_ws_producer<(NKS >= 3 ? 3 : 1), WIDE>(W, a, dp, tb, E0, F0, Qs, wmax);
    else if (NKS >= 2 && nks == 2) mixed_ws_producer<(NKS >= 2 ? 2 : 1), WIDE>(W, a, dp, tb, E0, F0, Qs, wmax);
    else mixed_ws_producer<1, WIDE>(W, a, dp, tb, E0, F0, Qs, wmax);
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
  // stage 2 of one chunk for row b (padded sites hold Q = 0 and Phi = 0, so
  // their (0 - c) sq terms meet a zero phase).  qb points at this lane's
  // first site (8 g4) of the chunk, row 4 cw, column l16: the site (+QSITE
  // float2) and row (+16) offsets are immediates.
  const float m1 = runtime_m1(a.g);
  auto stage2 = [&](int b, const float2* qb, float sq, int ex) {
    const float csr = -cr[b] * sq, csi = -ci[b] * sq;
    uint32_t qrh[4], qrl[4], qih[4], qil[4];
#pragma unroll
    for (int v2 = 0; v2 < 4; ++v2) {
      float sr[2], si[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const float2 q = qb[(2 * v2 + h) * QSITE + b * 16];
        sr[h] = fmaf(q.x, sq, csr);
        si[h] = fmaf(q.y, sq, csi);
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
        const float2 q = Qs[0][xl][(NBC * cw + b) * 16 + l16];
        sr += q.x;
        si += q.y;
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
    // one row at a time: no interleaving across rows keeps a single row's LDS
    // operands, split values and MFMA results live
    const float2* qb = &Qs[buf][8 * g4][NBC * cw * 16 + l16];
#pragma unroll
    for (int b = 0; b < NBC; ++b) {
      stage2(b, qb, sq, ex);
      __builtin_amdgcn_sched_barrier(0);
    }
    // buffer (chunk + 1) & 1 was last read in the previous chunk
    if (chunk + 1 < nchunk) PHs[(chunk + 1) & 1][cw][lane] = nxt;
    if (chunk + 2 < nchunk) nxt = load_part(chunk + 2);
    __syncthreads();
  }
  __syncthreads();  // the producers' zero-momentum partial sums are in LDS
  const double2* zs = (const double2*)(const void*)&Qs[0][0][0];

  // ---- epilogue (as contract_mixed_body, rows NBC cw + b) ----
  const int Ne = g.Ne, P = a.P;
  const size_t plane = (size_t)Ne * Ne;
  const size_t tbase = (size_t)(a.t_out0 + tb) * a.D * P;
  const int f = F0 + l16;
#pragma unroll
  for (int b = 0; b < NBC; ++b) {
    const int e = E0 + NBC * cw + b;
    const bool ok = (e < Ne) && (f < Ne);
    const double wef = (ok && a.wgt) ? a.wgt[(size_t)e * Ne + f] : 1.0;
    const double wfe = (ok && a.wgt) ? a.wgt[(size_t)f * Ne + e] : 1.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int bi = pb * 16 + mfma<float>::row(g4, r);
      if (!ok || bi >= a.nbase) continue;
      const int ip = a.plus_idx[bi], im = a.minus_idx[bi];
      const double S = a.psum[bi], mr = (double)cr[b] * S, mi = (double)ci[b] * S;
      const double epr = D1[b][r] - D2[b][r] + mr, epi = D3[b][r] + D4[b][r] + mi;
      const double emr = D1[b][r] + D2[b][r] + mr, emi = D4[b][r] - D3[b][r] + mi;
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
    // zero momentum: the producers' sums of Q over their sites (no mean
    // subtraction), each lane g4 takes a quarter of the terms
    const int row = NBC * cw + b;
    const double2 z0 = zs[(0 * 16 + row) * 16 + l16], z1 = zs[(1 * 16 + row) * 16 + l16];
    const double2 z2 = zs[(2 * 16 + row) * 16 + l16], z3 = zs[(3 * 16 + row) * 16 + l16];
    const double zre = (z0.x + z1.x) + (z2.x + z3.x), zim = (z0.y + z1.y) + (z2.y + z3.y);
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

// --------------------------------------------------------------------------
// Launch forms of the contraction.
//
// contract_kernel / contract_mixed_f16_kernel: one workgroup per work item
// (grid = nwork rounded up to 8), items spread over the XCDs in contiguous
// ranges (xcd_remap).
//
// contract_group_kernel / contract_mixed_group_kernel: one persistent launch
// over several problems (lattices of different size and Ne).  Work items of
// all problems are numbered problem after problem (work0 = first item of a
// problem); workgroup b on XCD b % 8 walks the items of its XCD:
//   order 0: XCD x owns the contiguous range [x n8, (x+1) n8), n8 = ceil(n/8)
//   order 1: item = 8 k + x (round robin over XCDs)
// with stride gridDim/8, so any grid that is a multiple of 8 covers all
// items.  Every workgroup reaches the end of its range and returns.
// --------------------------------------------------------------------------
struct GroupProblem {
  ContractArgs a;
  const void* W;    // c_t<T> workspace (float2 for the mixed kernel)
  const void* ph0;  // real phase table (T), or the FP16 split table
  const void* ph1;  // imaginary phase table (T), unused by the FP16 kernel
  double2* out;
  int work0;
  int pad[3];
};

__device__ __forceinline__ int group_item(int k, int per, int order) {
  const int x = blockIdx.x & 7;
  return order ? 8 * k + x : x * per + k;
}

__device__ __forceinline__ int group_find(const GroupProblem* __restrict__ probs, int nprob, int item) {
  int i = 0;
  while (i + 1 < nprob && probs[i + 1].work0 <= item) ++i;
  return i;
}

template <typename T, int NKS>
__global__ __launch_bounds__(64 * NWAVE, 1) void contract_kernel(const c_t<T>* __restrict__ W, const T* __restrict__ phr,
                                                                 const T* __restrict__ phi, double2* __restrict__ out,
                                                                 ContractArgs a) {
  __shared__ c_t<T> Qs[2][SC][16][16];
  const int w = item_index(a, blockIdx.x, gridDim.x);
  if (w >= a.nwork) return;
  contract_body<T, NKS>(W, phr, phi, out, a, w, Qs);
}

template <int NKS>
__global__ __launch_bounds__(64 * NWAVE, 1) void contract_mixed_f16_kernel(const float2* __restrict__ W,
                                                                            const half8* __restrict__ ph16,
                                                                            double2* __restrict__ out,
                                                                            ContractArgs a) {
  __shared__ float2 Qs[2][SC2][16][16];
  __shared__ float wmax[2][NWAVE];
  const int w = item_index(a, blockIdx.x, gridDim.x);
  if (w >= a.nwork) return;
  contract_mixed_body<NKS>(W, ph16, out, a, w, Qs, wmax);
}

template <int NKS, bool WIDE>
__global__ __launch_bounds__(64 * NWAVE, 1) void contract_mixed_ws_kernel(const float2* __restrict__ W,
                                                                           const half8* __restrict__ ph16,
                                                                           double2* __restrict__ out,
                                                                           ContractArgs a) {
  __shared__ float2 Qs[2][SC2][QSITE];
  __shared__ float wmax[2][NWAVE];
  __shared__ half8 PHs[2][4][64];
  const int w = item_index(a, blockIdx.x, gridDim.x);
  if (w >= a.nwork) return;
  contract_mixed_ws_body<NKS, WIDE>(W, ph16, out, a, w, Qs, wmax, PHs);
}

template <typename T, int NKS>
__global__ __launch_bounds__(64 * NWAVE, 1) void contract_group_kernel(const GroupProblem* __restrict__ probs, int nprob,
                                                                       int total, int order) {
  __shared__ c_t<T> Qs[2][SC][16][16];
  const int per = (total + 7) >> 3;
  const int nj = gridDim.x >> 3;
  for (int k = blockIdx.x >> 3; k < per; k += nj) {
    const int item = group_item(k, per, order);
    if (item >= total) break;
    const int i = group_find(probs, nprob, item);
    const ContractArgs a = probs[i].a;
    __syncthreads();  // the previous item's LDS reads are done
    contract_body<T, NKS>((const c_t<T>*)probs[i].W, (const T*)probs[i].ph0, (const T*)probs[i].ph1, probs[i].out, a,
                          item - probs[i].work0, Qs);
  }
}

template <int NKS, bool WIDE = false>
__global__ __launch_bounds__(64 * NWAVE, 1) void contract_mixed_group_kernel(const GroupProblem* __restrict__ probs,
                                                                             int nprob, int total, int order) {
  __shared__ float2 Qs[2][SC2][QSITE];
  __shared__ float wmax[2][NWAVE];
  __shared__ half8 PHs[2][4][64];
  const int per = (total + 7) >> 3;
  const int nj = gridDim.x >> 3;
  for (int k = blockIdx.x >> 3; k < per; k += nj) {
    const int item = group_item(k, per, order);
    if (item >= total) break;
    const int i = group_find(probs, nprob, item);
    const ContractArgs a = probs[i].a;
    __syncthreads();
    if constexpr (NKS <= 3)
      contract_mixed_ws_body<NKS, WIDE>((const float2*)probs[i].W, (const half8*)probs[i].ph0, probs[i].out, a,
                                        item - probs[i].work0, Qs, wmax, PHs);
    else
      contract_mixed_body<NKS>((const float2*)probs[i].W, (const half8*)probs[i].ph0, probs[i].out, a,
                               item - probs[i].work0, (float2(*)[SC2][16][16])(void*)Qs, wmax);
  }
}

// One workgroup per work item of the problem table (the batch's launch shape
// "one workgroup per item"): no persistent loop, so no state lives across
// items.  Grid = 8 ceil(total / 8); item mapping as contract_mixed_group_kernel.
template <int NKS, bool WIDE = false>
__global__ __launch_bounds__(64 * NWAVE, 1) void contract_mixed_item_kernel(const GroupProblem* __restrict__ probs,
                                                                            int nprob, int total, int order) {
  __shared__ float2 Qs[2][SC2][QSITE];
  __shared__ float wmax[2][NWAVE];
  __shared__ half8 PHs[2][4][64];
  const int item = __builtin_amdgcn_readfirstlane(group_item(blockIdx.x >> 3, (total + 7) >> 3, order));
  if (item >= total) return;
  const int i = __builtin_amdgcn_readfirstlane(group_find(probs, nprob, item));  // uniform: scalar loads
  const ContractArgs a = probs[i].a;
  if constexpr (NKS <= 3)
    contract_mixed_ws_body<NKS, WIDE>((const float2*)probs[i].W, (const half8*)probs[i].ph0, probs[i].out, a,
                                      item - probs[i].work0, Qs, wmax, PHs);
  else
    contract_mixed_body<NKS>((const float2*)probs[i].W, (const half8*)probs[i].ph0, probs[i].out, a,
                             item - probs[i].work0, (float2(*)[SC2][16][16])(void*)Qs, wmax);
}

}  // namespace ed
