// ed_api.hip -- C ABI of the MI355X elemental-generation engine.
// Declarations and conventions: include/elemental_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <exception>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/elemental_hip.h"
#include "ed_kernels.hpp"
#include "ed_mixed16.hpp"
#include "ed_gauge.hpp"
#include "ed_adjoint.hpp"

using ed::ContractArgs;
using ed::DPlan;
using ed::Geom;
using ed::SplitTerm;

namespace {

struct EdError {
  int code;
  std::string msg;
};

#define HIPCHK(expr)                                                                        \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess)                                                                   \
      throw EdError{ED_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)};         \
  } while (0)

[[noreturn]] void fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  throw EdError{code, buf};
}

// ---------------------------------------------------------------------------
// Fixed workspace of ed_compute_noalloc.  Each device that loads this library
// holds its own copy until the process ends.  In host code the name
// `scratch` is only a stand-in for the device symbol: its device address is
// resolved once per device (resolve_scratch, called from ed_create with the
// device current) and only that address is used afterwards.
// ---------------------------------------------------------------------------
static __device__ __align__(256) char scratch[ED_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};

// derivative(n) of lattice/insertion/derivative.py:23-33: length-k tuples
// occupy [(3^k-1)/2, (3^(k+1)-1)/2); digits most significant first.
std::vector<int> derivative_tuple(int n) {
  int k = 0, p = 1;
  while (n >= p) {
    n -= p;
    ++k;
    p *= 3;
  }
  std::vector<int> d(k);
  for (int i = k - 1; i >= 0; --i) {
    d[i] = n % 3;
    n /= 3;
  }
  return d;
}

int derivative_index(const std::vector<int>& d) {
  int k = (int)d.size(), base = 0, p = 1;
  for (int i = 0; i < k; ++i) {
    base += p;
    p *= 3;
  }
  int v = 0;
  for (int x : d) v = v * 3 + x;
  return base + v;
}

int level_start(int k) {  // (3^k - 1) / 2
  int s = 0, p = 1;
  for (int i = 0; i < k; ++i) {
    s += p;
    p *= 3;
  }
  return s;
}

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void ensure(size_t count) {
    if (count <= n && p) return;
    reset();
    if (count == 0) return;
    if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
      p = nullptr;
      fail(ED_ERR_ALLOC, "hipMalloc of %zu bytes failed", count * sizeof(T));
    }
    n = count;
  }
  ~DevBuf() { reset(); }
};

}  // namespace

struct ed_context {
  int device = 0;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  std::string err;

  int Lx = 0, Ly = 0, Lz = 0, Lt = 0, Ne = 0, V3 = 0, V3P = 0, NeP = 0;
  bool lattice_set = false, gauge_set = false, eig_set = false, mom_set = false;
  int num_nabla = 0, D = 1;
  // displacement mode (displacement_elemental.py): D = distance + 1 outputs,
  // workspace slots M_0 = V, M_1 .. M_distance, and two sets of six paths
  bool disp = false;
  int distance = 0;
  int slots() const {
    if (!disp) return D;
    return distance + 1 + (distance >= 1 ? 6 : 0) + (distance >= 2 ? 6 : 0);
  }
  std::vector<int32_t> moms;  // [P][3]

  DevBuf<double2> U64;
  DevBuf<float2> V;
  DevBuf<double> wgt;
  bool use_wgt = false;

  // plan
  bool plan_dirty = true;
  int hermitian = 1;
  int nd = 0, npb = 1, nbase = 0, zero_idx = -1;
  // per-plan MFMA counts for the flop accounting, [0] over derivatives that
  // run every (e, f) tile, [1] over palindromic ones (upper-triangle tiles)
  int64_t mfma_per_chunk_stage1[2] = {0, 0};  // sum of 16 * 4 * nks(d)
  int64_t nd_tri[2] = {0, 0};                 // derivative counts
  int64_t total_splits = 0;           // split terms of elemental.py:309-321 over all d
  int nks_max = 1;                    // largest stage-1 k-step count over primary d
  int items_t = 0;                    // contraction work items per timeslice
  int nsplit_max = 1;                 // largest merged split-term count over primary d
  int64_t h16_groups_sum[2] = {0, 0};         // sum of FP16 stage-1 k-groups
  float cmax = 1.f;                   // largest |split coefficient|
  DevBuf<DPlan> dplan;
  DevBuf<SplitTerm> splits;
  // unit plan of the FP16 stage-1 kernel (build_plan): every coefficient +1
  int h16_unit_opt = 1;               // ED_OPT_H16_UNIT
  int hop_lds = -1;                   // ED_OPT_HOP_LDS: the LDS-staged hop where the shape allows
  bool unit_ok = false;
  bool ws_rot = false;                // the unit plan needs slot n stored times i^|n|
  int nsplit_u_max = 1;               // largest split count of the unit plan
  int64_t h16u_groups_sum[2] = {0, 0};
  DevBuf<DPlan> dplan_u;
  DevBuf<SplitTerm> splits_u;
  DevBuf<int> pidx;  // plus [npb*16] then minus [npb*16]
  DevBuf<double> psum;  // [npb*16] sum over sites of Phi_b
  DevBuf<double> ph64;
  DevBuf<float> ph32;
  DevBuf<ed::half8> ph16;  // [npb][V3P/32][4][16][32] fp16 hi/lo of 2^14 Phi
  int mixed_stage2 = -1;   // 1: FP16 three-term split, 0: FP32 MFMA, -1: FP16 from V3 >= 4096 on
  // FP16 stage 2: 1 warp-specialized kernel with FP32 stage 1 (up to 3
  // k-steps), 2 warp-specialized kernel with FP16 stage 1 (up to 2 k-groups),
  // 0 uniform waves
  int mixed_ws = 1;
  int wide_ws = 0;     // 64-bit workspace addresses in the warp-specialized kernel below 2^31 bytes too
  int strip = 0;       // e-tile rows per strip of the contraction's tile order (0: kernel default)
  int item_order = 0;  // contraction items over the XCDs: 0 contiguous ranges, 1 round robin
  float vmax = 0.f;        // largest |re|, |im| of the stored eigenvectors
  double vv_dev = -1.0;    // FP64 max|V^dag V - 1| over timeslices of the stored eigenvectors
  int ortho_gate = 0;      // n > 0: ed_set_eigvecs refuses vv_dev > 10^-n
  int wsh = 0;             // scale exponent of the FP16 stage-1 workspace
  int W_elem = 0;          // bytes per complex element of the data now in W (0: none)
  uint64_t gen = 0;        // bumped whenever buffers or plan a captured launch refers to change

  // reverse pass
  int bns_max = 1;        // largest merged split count of one derivative
  int nparent_all = 0;    // slots that have children (levels below num_nabla)
  DevBuf<SplitTerm> bsplits;
  DevBuf<int2> bdsplit;
  int nq = 0;                 // momentum groups: opposite pairs and single momenta
  DevBuf<int2> bgroups;       // [nq] (p1, p2 or -1)
  DevBuf<double2> phase_grp;  // [nq][V3P] Phi of each group's first momentum (0 on padded sites)
  DevBuf<int> child_of;       // [nparent_all][3]
  DevBuf<double2> bW, bA, bG, bOutV, bOutU;
  // W transposed to [slot][row][site][colour] for the contraction adjoint; the
  // same buffer then receives the gradient of W back in W's layout
  DevBuf<double2> bWT;
  DevBuf<double2> bgWT;  // gradient of W in the transposed layout
  bool bW_clear = false;

  // exact (complex128) eigenvectors for the FP64 form and the reverse pass
  int exact_v = 0;
  DevBuf<double2> V64;
  bool v64_set = false;

  // workspace
  int tbatch_opt = 0;
  int64_t ws_mb = 65536;  // 8 timeslices of 7.2 GB (48^3, Nev 200) fit; 288 GB of HBM
  DevBuf<char> W;
  DevBuf<double2> stage;

  // profiling
  int profile = 0;
  double prof_ms[4] = {0, 0, 0, 0};
  int64_t prof_n[4] = {0, 0, 0, 0};
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::vector<hipEvent_t> event_pool;

  ~ed_context() {
    for (auto& p : pending) {
      (void)hipEventDestroy(p.second.first);
      (void)hipEventDestroy(p.second.second);
    }
    for (auto ev : event_pool) (void)hipEventDestroy(ev);
  }

  Geom geom(int nW) const {
    Geom g;
    g.Lx = Lx;
    g.Ly = Ly;
    g.Lz = Lz;
    g.Lt = Lt;
    g.V3 = V3;
    g.V3P = V3P;
    g.Ne = Ne;
    g.NeP = NeP;
    g.nW = nW;
    g.wsh = 0;
    g.rot = 0;
    return g;
  }

  hipEvent_t get_event() {
    if (!event_pool.empty()) {
      hipEvent_t e = event_pool.back();
      event_pool.pop_back();
      return e;
    }
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    return e;
  }
  void drain_events() {
    for (auto& p : pending) {
      float ms = 0;
      HIPCHK(hipEventSynchronize(p.second.second));
      HIPCHK(hipEventElapsedTime(&ms, p.second.first, p.second.second));
      prof_ms[p.first] += ms;
      prof_n[p.first] += 1;
      event_pool.push_back(p.second.first);
      event_pool.push_back(p.second.second);
    }
    pending.clear();
  }
  template <class F>
  void launch(int kind, F&& f) {
    if (!profile) {
      f();
      HIPCHK(hipGetLastError());
      return;
    }
    hipEvent_t a = get_event(), b = get_event();
    HIPCHK(hipEventRecord(a, stream));
    f();
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(b, stream));
    pending.push_back({kind, {a, b}});
  }
};

namespace {

int guard(ed_context* ctx, const std::function<void()>& body);

void require(bool c, int code, const char* msg) {
  if (!c) fail(code, "%s", msg);
}

// ---------------------------------------------------------------------------
// planning
// ---------------------------------------------------------------------------
// Split terms of elemental.py:309-321 for the tuple t: bit j of the mask sends
// t[j] to the right vector, the left list is applied reversed, coefficient
// (-1)^{#right}; identical (left, right) pairs are merged and cancelled ones
// dropped.
std::vector<SplitTerm> merged_splits(const std::vector<int>& t) {
  std::map<std::pair<int, int>, float> acc;
  std::vector<std::pair<int, int>> order;
  const int k = (int)t.size();
  for (int pick = 0; pick < (1 << k); ++pick) {
    std::vector<int> right, left;
    for (int j = 0; j < k; ++j) ((pick >> j) & 1 ? right : left).push_back(t[j]);
    std::reverse(left.begin(), left.end());
    const std::pair<int, int> key{derivative_index(left), derivative_index(right)};
    const float cf = (right.size() % 2) ? -1.f : 1.f;
    if (!acc.count(key)) order.push_back(key);
    acc[key] += cf;
  }
  std::vector<SplitTerm> out;
  for (auto& key : order) {
    if (acc[key] == 0.f) continue;
    SplitTerm st{};
    st.nL = key.first;
    st.nR = key.second;
    st.coeff = acc[key];
    out.push_back(st);
  }
  return out;
}

// Shape of the contraction adjoint for a largest split count ns: rows per
// wave (the accumulators are rows x splits x colours) and momentum groups
// per chunk of scalar loads (rows x groups x 2 complex values in 64 SGPRs).
#ifndef ED_ADJ_RB2
#define ED_ADJ_RB2 4
#endif
void adj_shape(int ns, int& rb, int& gc) {
  rb = ns <= 1 ? 4 : ns <= 2 ? ED_ADJ_RB2 : ns <= 4 ? 2 : 1;
  gc = 8 / rb > 4 ? 4 : 8 / rb;
}

void build_plan(ed_context* c) {
  if (!c->plan_dirty) return;
  require(c->lattice_set, ED_ERR_STATE, "ed_set_lattice must be called first");
  require(c->mom_set, ED_ERR_STATE, "ed_set_momenta must be called first");
  HIPCHK(hipSetDevice(c->device));
  const int D = c->D;

  // derivative plan: primary d and the partner rev(d) derived by Hermiticity
  std::vector<DPlan> dpl;
  std::vector<SplitTerm> sp;
  c->mfma_per_chunk_stage1[0] = c->mfma_per_chunk_stage1[1] = 0;
  c->nd_tri[0] = c->nd_tri[1] = 0;
  c->total_splits = 0;
  c->nks_max = 1;
  c->nsplit_max = 1;
  c->h16_groups_sum[0] = c->h16_groups_sum[1] = 0;
  c->cmax = 1.f;
  // displacement mode: output n is the single term conj(V) . M_n
  for (int d = 0; c->disp && d < D; ++d) {
    DPlan p{};
    p.d = d;
    p.partner = -1;
    p.sign = 1;
    p.first = (int)sp.size();
    p.nsplit = 1;
    SplitTerm st{};
    st.nL = 0;
    st.nR = d;
    st.coeff = 1.f;
    sp.push_back(st);
    c->total_splits += 1;
    c->mfma_per_chunk_stage1[0] += (int64_t)ed::SC * 4;
    c->h16_groups_sum[0] += 1;
    c->nd_tri[0] += 1;
    dpl.push_back(p);
  }
  for (int d = 0; !c->disp && d < D; ++d) {
    std::vector<int> t = derivative_tuple(d);
    c->total_splits += (int64_t)1 << t.size();
    std::vector<int> r(t.rbegin(), t.rend());
    const int rd = derivative_index(r);
    int partner = -1, mirror = 0;
    if (c->hermitian && rd != d) {
      if (rd < d) continue;  // written by its partner
      partner = rd;
    } else if (c->hermitian) {
      // palindromic d: E_d[p] = (-1)^{|d|} E_d[-p]^dagger, so tiles below the
      // (e, f) diagonal are the transposes of those above it
      partner = d;
      mirror = 1;
    }
    DPlan p{};
    p.d = d;
    p.partner = partner;
    p.mirror = mirror;
    p.sign = (t.size() % 2) ? -1 : 1;
    p.first = (int)sp.size();
    for (const SplitTerm& st : merged_splits(t)) sp.push_back(st);
    p.nsplit = (int)sp.size() - p.first;
    c->mfma_per_chunk_stage1[mirror] += (int64_t)ed::SC * 4 * ((3 * p.nsplit + 3) / 4);
    c->nd_tri[mirror] += 1;
    c->nks_max = std::max(c->nks_max, (3 * p.nsplit + 3) / 4);
    c->nsplit_max = std::max(c->nsplit_max, p.nsplit);
    c->h16_groups_sum[mirror] += (3 * p.nsplit + 7) / 8;
    for (int k = p.first; k < (int)sp.size(); ++k) c->cmax = std::max(c->cmax, std::fabs(sp[k].coeff));
    dpl.push_back(p);
  }
  c->nd = (int)dpl.size();

  // momentum plan: base momenta b (canonical sign: first non-zero component
  // positive); E[+b] and E[-b] come from one contraction.  The first
  // (0,0,0) goes to the FP64 VALU zero-momentum sum; repeated momenta get
  // their own base entry.
  const int P = (int)(c->moms.size() / 3);
  struct Base {
    int n[3];
    int plus = -1, minus = -1;
  };
  std::vector<Base> base;
  c->zero_idx = -1;
  for (int p = 0; p < P; ++p) {
    int n[3] = {c->moms[3 * p], c->moms[3 * p + 1], c->moms[3 * p + 2]};
    int sgn = 0;
    for (int i = 0; i < 3 && !sgn; ++i) sgn = n[i] > 0 ? 1 : (n[i] < 0 ? -1 : 0);
    if (sgn == 0 && c->zero_idx < 0) {
      c->zero_idx = p;
      continue;
    }
    int cn[3] = {sgn < 0 ? -n[0] : n[0], sgn < 0 ? -n[1] : n[1], sgn < 0 ? -n[2] : n[2]};
    bool placed = false;
    for (auto& b : base) {
      if (b.n[0] == cn[0] && b.n[1] == cn[1] && b.n[2] == cn[2]) {
        int& slot = (sgn < 0) ? b.minus : b.plus;
        if (slot < 0) {
          slot = p;
          placed = true;
          break;
        }
      }
    }
    if (!placed) {
      Base b;
      b.n[0] = cn[0];
      b.n[1] = cn[1];
      b.n[2] = cn[2];
      ((sgn < 0) ? b.minus : b.plus) = p;
      base.push_back(b);
    }
  }
  c->nbase = (int)base.size();
  c->npb = std::max(1, (c->nbase + 15) / 16);
  const int NB = c->npb * 16;
  std::vector<int> idx(2 * NB, -1);
  for (int i = 0; i < c->nbase; ++i) {
    idx[i] = base[i].plus;
    idx[NB + i] = base[i].minus;
  }
  // FP64 phase table [npb][V3P][16]; exp(2 pi i frac) with the integer
  // products reduced modulo L first (phase.py:11-13,41-46 up to rounding)
  std::vector<double> pr((size_t)NB * c->V3P, 0.0), pi((size_t)NB * c->V3P, 0.0);
  for (int i = 0; i < c->nbase; ++i) {
    const int blk = i / 16, col = i % 16;
    for (int z = 0; z < c->Lz; ++z)
      for (int y = 0; y < c->Ly; ++y)
        for (int x = 0; x < c->Lx; ++x) {
          auto md = [](long a, long m) { long r = a % m; return r < 0 ? r + m : r; };
          const double fr = (double)md((long)base[i].n[0] * x, c->Lx) / c->Lx +
                            (double)md((long)base[i].n[1] * y, c->Ly) / c->Ly +
                            (double)md((long)base[i].n[2] * z, c->Lz) / c->Lz;
          const double ang = 2.0 * M_PI * fr;
          const size_t s = ((size_t)z * c->Ly + y) * c->Lx + x;
          pr[((size_t)blk * c->V3P + s) * 16 + col] = std::cos(ang);
          pi[((size_t)blk * c->V3P + s) * 16 + col] = std::sin(ang);
        }
  }
  std::vector<float> pr32(pr.begin(), pr.end()), pi32(pi.begin(), pi.end());
  // FP16 hi/lo split of 2^14 Phi for the mixed form's FP16 MFMA phase contraction
  const int nchunk2 = c->V3P / ed::SC2;
  std::vector<_Float16> p16((size_t)c->npb * nchunk2 * 4 * 16 * 32);
  for (int blk = 0; blk < c->npb; ++blk)
    for (int ch = 0; ch < nchunk2; ++ch)
      for (int col = 0; col < 16; ++col)
        for (int xs = 0; xs < 32; ++xs) {
          const size_t src = ((size_t)blk * c->V3P + (size_t)ch * 32 + xs) * 16 + col;
          const double vr = pr[src] * 16384.0, vi = pi[src] * 16384.0;
          const _Float16 rh = (_Float16)vr, ih = (_Float16)vi;
          const _Float16 rl = (_Float16)(vr - (double)rh), il = (_Float16)(vi - (double)ih);
          const size_t base = (((size_t)blk * nchunk2 + ch) * 4) * 16 * 32 + (size_t)col * 32 + xs;
          p16[base + 0 * 512] = rh;
          p16[base + 1 * 512] = rl;
          p16[base + 2 * 512] = ih;
          p16[base + 3 * 512] = il;
        }
  // sum_x Phi_b(x) is V3 when every component of b is a multiple of L, else 0
  std::vector<double> ps(NB, 0.0);
  for (int i = 0; i < c->nbase; ++i)
    if (base[i].n[0] % c->Lx == 0 && base[i].n[1] % c->Ly == 0 && base[i].n[2] % c->Lz == 0) ps[i] = (double)c->V3;

  // work items of each derivative within a timeslice: npb blocks of its tiles
  {
    const int nt = c->NeP / 16;
    int items = 0;
    for (DPlan& p : dpl) {
      p.item0 = items;
      items += c->npb * (p.mirror ? nt * (nt + 1) / 2 : nt * nt);
    }
    c->items_t = items;
  }

  // Unit plan of the FP16 stage-1 kernel.  A merged split's coefficient is
  // (-1)^{|nR|} times a positive multiplicity (merged_splits).  With slot n
  // of the workspace stored times i^{|n|} (the hop multiplies each child by
  // i), conj(L'') R'' = i^{|nR| - |nL|} conj(L) R = (-1)^{|nR|} i^{-|d|}
  // conj(L) R since |nL| + |nR| = |d|: every split then enters as many
  // times as its multiplicity with coefficient +1, and the result is
  // multiplied by i^{|d|}.  Displacement terms have coefficient 1 and need no
  // rotation.
  {
    std::vector<DPlan> dplu = dpl;
    std::vector<SplitTerm> spu;
    bool ok = true;
    c->nsplit_u_max = 1;
    c->h16u_groups_sum[0] = c->h16u_groups_sum[1] = 0;
    for (DPlan& p : dplu) {
      const int first = (int)spu.size();
      for (int k = p.first; k < p.first + p.nsplit; ++k) {
        const SplitTerm& st = sp[k];
        const int sgn = c->disp ? 1 : (((int)derivative_tuple(st.nR).size() & 1) ? -1 : 1);
        const float mult = st.coeff * (float)sgn;
        const int m = (int)mult;
        if ((float)m != mult || m < 1 || m > 4) ok = false;
        for (int r = 0; ok && r < m; ++r) spu.push_back(SplitTerm{st.nL, st.nR, 1.f, 0});
      }
      p.first = first;
      p.nsplit = (int)spu.size() - first;
      p.rot = c->disp ? 0 : ((int)derivative_tuple(p.d).size() & 3);
      c->nsplit_u_max = std::max(c->nsplit_u_max, p.nsplit);
      c->h16u_groups_sum[p.mirror] += (3 * p.nsplit + 7) / 8;
    }
    c->unit_ok = ok && !dplu.empty();
    c->ws_rot = c->unit_ok && !c->disp;
    if (c->unit_ok) {
      c->dplan_u.ensure(dplu.size());
      c->splits_u.ensure(spu.size());
      HIPCHK(hipMemcpy(c->dplan_u.p, dplu.data(), dplu.size() * sizeof(DPlan), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(c->splits_u.p, spu.data(), spu.size() * sizeof(SplitTerm), hipMemcpyHostToDevice));
    }
  }
  c->dplan.ensure(dpl.size());
  c->splits.ensure(std::max<size_t>(1, sp.size()));
  c->pidx.ensure(idx.size());
  c->psum.ensure(ps.size());
  c->ph64.ensure(2 * pr.size());
  c->ph32.ensure(2 * pr.size());
  c->ph16.ensure(p16.size() / 8);
  HIPCHK(hipMemcpy(c->dplan.p, dpl.data(), dpl.size() * sizeof(DPlan), hipMemcpyHostToDevice));
  if (!sp.empty()) HIPCHK(hipMemcpy(c->splits.p, sp.data(), sp.size() * sizeof(SplitTerm), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->pidx.p, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->psum.p, ps.data(), ps.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->ph64.p, pr.data(), pr.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->ph64.p + pr.size(), pi.data(), pi.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->ph32.p, pr32.data(), pr32.size() * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->ph32.p + pr32.size(), pi32.data(), pi32.size() * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->ph16.p, p16.data(), p16.size() * sizeof(_Float16), hipMemcpyHostToDevice));

  // reverse pass: merged splits of every derivative (no Hermitian pairing),
  // the momentum groups and Phi of each group's first momentum as [nq][V3P],
  // and the child slot of every (parent slot, direction)
  {
    std::vector<SplitTerm> bsp;
    std::vector<int2> bds(D);
    c->bns_max = 1;
    for (int d = 0; d < D; ++d) {
      const std::vector<SplitTerm> terms =
          c->disp ? std::vector<SplitTerm>{SplitTerm{0, d, 1.f, 0}} : merged_splits(derivative_tuple(d));
      bds[d] = make_int2((int)bsp.size(), (int)terms.size());
      bsp.insert(bsp.end(), terms.begin(), terms.end());
      c->bns_max = std::max(c->bns_max, (int)terms.size());
    }
    // momentum groups of the adjoint's phase sum: each momentum pairs with an
    // unused opposite one in the list (Phi_-n = conj(Phi_n)), else stands alone
    std::vector<int2> grp;
    {
      std::vector<char> used(P, 0);
      for (int p = 0; p < P; ++p) {
        if (used[p]) continue;
        used[p] = 1;
        int partner = -1;
        const int* n = &c->moms[3 * p];
        if (n[0] || n[1] || n[2])
          for (int o = p + 1; o < P && partner < 0; ++o)
            if (!used[o] && c->moms[3 * o] == -n[0] && c->moms[3 * o + 1] == -n[1] && c->moms[3 * o + 2] == -n[2]) {
              partner = o;
              used[o] = 1;
            }
        grp.push_back(make_int2(p, partner));
      }
      // padded with zero groups to whole chunks of the adjoint's scalar loads,
      // and up to 4 or 8 groups, whose phases the adjoint holds in registers
      int rb, gc;
      adj_shape(c->bns_max, rb, gc);
      const size_t want = grp.size() <= 4 ? 4 : grp.size() <= 8 ? 8 : (grp.size() + gc - 1) / gc * gc;
      while (grp.size() < want) grp.push_back(make_int2(-1, -1));
    }
    c->nq = (int)grp.size();
    std::vector<double2> pa((size_t)c->V3P * c->nq, make_double2(0.0, 0.0));
    for (int z = 0; z < c->Lz; ++z)
      for (int y = 0; y < c->Ly; ++y)
        for (int x = 0; x < c->Lx; ++x) {
          const size_t st = ((size_t)z * c->Ly + y) * c->Lx + x;
          for (int q = 0; q < c->nq; ++q) {
            const int p = grp[q].x;
            if (p < 0) continue;
            auto md = [](long a, long m) { long r = a % m; return r < 0 ? r + m : r; };
            const double fr = (double)md((long)c->moms[3 * p] * x, c->Lx) / c->Lx +
                              (double)md((long)c->moms[3 * p + 1] * y, c->Ly) / c->Ly +
                              (double)md((long)c->moms[3 * p + 2] * z, c->Lz) / c->Lz;
            const double ang = 2.0 * M_PI * fr;
            pa[(size_t)q * c->V3P + st] = make_double2(std::cos(ang), std::sin(ang));
          }
        }
    c->nparent_all = c->disp ? 0 : level_start(c->num_nabla);
    std::vector<int> child(3 * std::max(1, c->nparent_all), 0);
    for (int pnt = 0; pnt < c->nparent_all; ++pnt)
      for (int mu = 0; mu < 3; ++mu) {
        std::vector<int> t = derivative_tuple(pnt);
        t.push_back(mu);
        child[3 * pnt + mu] = derivative_index(t);
      }
    c->bsplits.ensure(bsp.size());
    c->bdsplit.ensure(bds.size());
    c->bgroups.ensure(grp.size());
    c->phase_grp.ensure(pa.size());
    c->child_of.ensure(child.size());
    HIPCHK(hipMemcpy(c->bsplits.p, bsp.data(), bsp.size() * sizeof(SplitTerm), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->bdsplit.p, bds.data(), bds.size() * sizeof(int2), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->bgroups.p, grp.data(), grp.size() * sizeof(int2), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->phase_grp.p, pa.data(), pa.size() * sizeof(double2), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->child_of.p, child.data(), child.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  c->plan_dirty = false;
  ++c->gen;
}

__global__ void c128_to_c64_kernel(const double2* __restrict__ in, float2* __restrict__ out, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = make_float2((float)in[i].x, (float)in[i].y);
}
// largest |re|, |im| over n complex64 values into *m (bit pattern of a
// non-negative float, so unsigned order is float order); *m starts at 0
__global__ void absmax_c64_kernel(const float2* __restrict__ in, size_t n, unsigned int* __restrict__ m) {
  float v = 0.f;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    v = fmaxf(v, fmaxf(fabsf(in[i].x), fabsf(in[i].y)));
  v = ed::wave_max_nonneg(v);
  if ((threadIdx.x & 63) == 0) atomicMax(m, __float_as_uint(v));
}
// FP64 max |(V_t^dag V_t)[e][f] - delta_ef| over one 32 x 32 tile (e, f) of
// timeslice t0 + blockIdx.z; each row of V_t holds K = 3 V3 complex64 values.
// Products of complex64 values are exact in FP64 and the sums run in FP64.
// The tile maximum is maxed into *dev as the bit pattern of a non-negative
// double (unsigned order is double order); *dev starts at 0.
__global__ __launch_bounds__(256) void gram_dev_kernel(const float2* __restrict__ V, int Ne, size_t K, int t0,
                                                      unsigned long long* __restrict__ dev) {
  constexpr int KC = 32;
  __shared__ double2 A[32][KC + 1], B[32][KC + 1];
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const int e0 = blockIdx.x * 32, f0 = blockIdx.y * 32;
  const float2* Vt = V + (size_t)(t0 + blockIdx.z) * Ne * K;
  double re[2][2] = {{0.0, 0.0}, {0.0, 0.0}}, im[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
  for (size_t k0 = 0; k0 < K; k0 += KC) {
    for (int i = tid; i < 32 * KC; i += 256) {
      const int r = i / KC, k = i % KC;
      const size_t kk = k0 + k;
      float2 a = make_float2(0.f, 0.f), b = make_float2(0.f, 0.f);
      if (kk < K) {
        if (e0 + r < Ne) a = Vt[(size_t)(e0 + r) * K + kk];
        if (f0 + r < Ne) b = Vt[(size_t)(f0 + r) * K + kk];
      }
      A[r][k] = make_double2(a.x, a.y);
      B[r][k] = make_double2(b.x, b.y);
    }
    __syncthreads();
    for (int k = 0; k < KC; ++k) {
      const double2 a[2] = {A[ty][k], A[ty + 16][k]};
      const double2 b[2] = {B[tx][k], B[tx + 16][k]};
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {  // conj(a) b
          re[i][j] += a[i].x * b[j].x + a[i].y * b[j].y;
          im[i][j] += a[i].x * b[j].y - a[i].y * b[j].x;
        }
    }
    __syncthreads();
  }
  double d = 0.0;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      const int e = e0 + ty + 16 * i, f = f0 + tx + 16 * j;
      if (e < Ne && f < Ne) d = fmax(d, hypot(re[i][j] - (e == f ? 1.0 : 0.0), im[i][j]));
    }
  for (int off = 32; off > 0; off >>= 1) d = fmax(d, __shfl_xor(d, off));
  if ((tid & 63) == 0) atomicMax(dev, (unsigned long long)__double_as_longlong(d));
}

__global__ void c64_to_c128_kernel(const float2* __restrict__ in, double2* __restrict__ out, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = make_double2((double)in[i].x, (double)in[i].y);
}

template <typename T>
size_t per_t_ws_bytes(const ed_context* c) {
  return (size_t)c->slots() * c->NeP * 3 * c->V3P * sizeof(ed::c_t<T>);
}

// one timeslice's FP32 workspace is addressable with 32-bit buffer offsets
bool ws_fits32(const ed_context* c) { return per_t_ws_bytes<float>(c) < ((size_t)1 << 31); }

// the warp-specialized FP32 stage-1 kernel, which also has a 64-bit-address
// form for larger workspaces (where it replaces the uniform kind too)
bool use_ws_kernel(const ed_context* c) { return (c->mixed_ws != 0 || !ws_fits32(c)) && c->nks_max <= 3; }
bool ws_wide(const ed_context* c) { return !ws_fits32(c) || c->wide_ws; }

// the mixed form takes the FP16-MFMA phase contraction when chosen and its
// kernel can address the workspace (every FP16 stage-2 kernel below 2^31
// bytes per timeslice, the warp-specialized one at any size)
// The FP16 phase table of one momentum block is read through a buffer
// resource with 32-bit byte offsets: V3P x 128 bytes must stay below 2^31
// (V3 below 2^24 sites); larger lattices take the FP32 phase contraction.
bool ph16_fits32(const ed_context* c) { return (size_t)c->V3P * 128 < ((size_t)1 << 31); }

bool use_f16_stage2(const ed_context* c) {
  const bool chosen = c->mixed_stage2 == 1 || (c->mixed_stage2 < 0 && c->V3 >= 4096);
  return chosen && ph16_fits32(c) && (ws_fits32(c) || use_ws_kernel(c));
}

// The mixed form on lattices with V3 < 4096 and the default phase
// contraction runs the FP64 kernels: there the FP32 split sum of stage 1
// leaves up to 1.2e-6 normwise on (d, p) blocks that are ~1 % of their
// derivative's largest block (a 5x7x9 rough lattice, 57 momenta), above the
// mixed form's 1e-6 bound, while the FP64 kernels cost little at that size.
// An explicit ED_OPT_MIXED_STAGE2 keeps the mixed kernels.
bool mixed_runs_fp64(const ed_context* c) { return c->mixed_stage2 < 0 && c->V3 < 4096; }

// the FP16 stage-1 form of the mixed contraction: FP16 stage 2 chosen, the
// warp-specialized FP16 kernel selected, at most two k-groups of eight
// split-colour pairs (five split terms) per derivative, and V3 >= 4096 (on
// smaller rough lattices the 22-bit workspace pairs push blocks that are
// small against their terms past 1e-6; there the FP32 stage 1 runs)
int h16_unit_groups(const ed_context* c) { return (3 * c->nsplit_u_max + 7) / 8; }
// the unit plan runs when its k-groups fit the kernel
bool h16_unit(const ed_context* c) { return c->h16_unit_opt && c->unit_ok && h16_unit_groups(c) <= 2; }
int h16_groups(const ed_context* c) { return h16_unit(c) ? h16_unit_groups(c) : (3 * c->nsplit_max + 7) / 8; }
bool h16_fits(const ed_context* c) {
  return use_f16_stage2(c) && ws_fits32(c) && h16_groups(c) <= 2 && c->V3 >= 4096;
}
bool use_h16(const ed_context* c) { return c->mixed_ws == 2 && h16_fits(c); }

// the contraction kernel the mixed form runs (the order of the tests in
// run_batches): see ED_Q_MIXED_KERNEL
int mixed_kernel_id(const ed_context* c) {
  if (mixed_runs_fp64(c)) return 0;
  if (!use_f16_stage2(c)) return 1;
  if (use_h16(c)) return 5;
  if (use_ws_kernel(c)) return ws_wide(c) ? 4 : 3;
  return 2;
}

// Scale exponent of the FP16 stage-1 workspace: every stored component,
// times the largest split coefficient, stays below 2^15.  Unitary links keep
// a colour vector's norm, so a hop at most doubles it and a straight path
// keeps it; a component is at most sqrt(3) times the largest eigenvector
// component times that growth.
int h16_shift(const ed_context* c) {
  const double growth = std::sqrt(3.0) * (c->disp ? 1.0 : std::ldexp(1.0, c->num_nabla));
  const double bound = (double)c->vmax * growth * std::max(1.0f, c->cmax);
  if (!(bound > 0)) return 0;
  return std::max(-60, std::min(60, (int)std::floor(std::log2(32768.0 / bound))));
}

// Workspace for TB timeslices in element type T; padded rows and sites must
// hold zeros (they are never written), so a new buffer, or one last used
// with the other element type, is cleared.
template <typename T>
ed::c_t<T>* ensure_workspace(ed_context* c, int TB) {
  const size_t ws = per_t_ws_bytes<T>(c) * TB;
  if (c->W.n < ws || c->W_elem != (int)sizeof(ed::c_t<T>)) {
    c->W.ensure(ws);
    HIPCHK(hipMemsetAsync(c->W.p, 0, c->W.n, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->W_elem = (int)sizeof(ed::c_t<T>);
    ++c->gen;
  }
  return (ed::c_t<T>*)(void*)c->W.p;
}

template <typename T>
int auto_tbatch(const ed_context* c, int nt) {
  int TB = c->tbatch_opt;
  if (TB <= 0) {
    size_t budget = (size_t)std::max<int64_t>(1, c->ws_mb) << 20;
    // at most this context's current workspace (reused) plus three quarters
    // of the free device memory, so a shared or crowded GPU gets smaller
    // batches instead of a failed allocation
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) budget = std::min(budget, c->W.n + free_b / 4 * 3);
    const size_t per_t = per_t_ws_bytes<T>(c);
    TB = (int)std::max<size_t>(1, budget / per_t);
    // whole timeslices per XCD: the contraction hands each of the 8 XCDs a
    // contiguous eighth of the work items, and the items of one timeslice
    // differ in cost by derivative.  Large timeslices (>= 1 GiB of
    // workspace) run eight per launch, which measured fastest at 32^3.
    if (TB >= 8 && nt >= 8) TB = per_t >= ((size_t)1 << 30) ? 8 : (std::min(TB, nt) & ~7);
  }
  return std::min(TB, nt);
}

// V(t) -> slot 0, then the covariant hops level by level, for timeslices
// tb0 .. tb0+nb-1 into W; launch(kind, f) wraps each launch.
template <typename T, class L>
void enqueue_stencil(ed_context* c, ed::c_t<T>* W, int tb0, int nb, hipStream_t st, L&& launch) {
  Geom g = c->geom(c->slots());
  if (std::is_same<T, ed::hsplit>::value) {
    g.wsh = c->wsh;
    g.rot = h16_unit(c) && c->ws_rot ? 1 : 0;
  }
  launch(ED_KERNEL_STAGE, [&] {
    dim3 grid((3 * c->V3 + 63) / 64, (c->Ne + 63) / 64, nb);
    if (sizeof(T) == 8 && c->exact_v && c->v64_set)
      ed::stage_v_kernel<T, double2><<<grid, 256, 0, st>>>(c->V64.p, W, g, tb0);
    else
      ed::stage_v_kernel<T, float2><<<grid, 256, 0, st>>>(c->V.p, W, g, tb0);
  });
  if (c->disp) {
    // paths of length n from those of length n - 1; sets A and B alternate
    const int A = c->distance + 1, B = A + 6;
    for (int n = 1; n <= c->distance; ++n) {
      const int src = n == 1 ? -1 : ((n & 1) ? B : A);
      const int dst = n == c->distance ? -1 : ((n & 1) ? A : B);
      launch(ED_KERNEL_STENCIL, [&] {
        dim3 grid((c->V3 + 3) / 4, (c->Ne + 63) / 64, nb);
        ed::disp_kernel<T><<<grid, 256, 0, st>>>(W, c->U64.p, g, tb0, src, dst, n);
      });
    }
    return;
  }
  for (int k = 1; k <= c->num_nabla; ++k) {
    const int parent0 = level_start(k - 1), nparent = level_start(k) - parent0;
    const int child0 = level_start(k);
    launch(ED_KERNEL_STENCIL, [&] {
      constexpr int EC = ed::hl_ec<T>();
      // LDS-staged by default for 16-byte elements only: for 8-byte ones the
      // links re-read per eigenvector chunk cost more than the plane reuse saves
      const bool lds = c->hop_lds > 0 || (c->hop_lds < 0 && sizeof(ed::c_t<T>) == 16);
      if (lds && c->Lx % ed::HL_T == 0 && c->Ly % ed::HL_T == 0 && c->Lz >= 3 && c->NeP % EC == 0) {
        // eigenvector chunks that hold an eigenvector (a chunk of padding only
        // writes nothing: padded rows stay zero)
        dim3 grid((c->Lx / ed::HL_T) * (c->Ly / ed::HL_T), nparent * ((c->Ne + EC - 1) / EC), nb);
        constexpr int NW = EC == 16 ? 16 : 8;  // one site per wave and iteration
        ed::hop3_lds_kernel<T, NW><<<grid, 64 * NW, 0, st>>>(W, c->U64.p, g, tb0, parent0, child0, nparent);
      } else {
        dim3 grid((c->V3 + 3) / 4, nparent, nb);
        ed::hop3_kernel<T><<<grid, 256, 0, st>>>(W, c->U64.p, g, tb0, parent0, child0);
      }
    });
  }
}

ContractArgs contract_args(const ed_context* c, int nb, int t_out0) {
  ContractArgs a{};
  a.g = c->geom(c->slots());
  a.D = c->D;
  a.P = (int)(c->moms.size() / 3);
  a.nd = c->nd;
  a.npb = c->npb;
  a.nbase = c->nbase;
  a.nte = c->NeP / 16;
  a.ntf = c->NeP / 16;
  a.items_t = c->items_t;
  a.nwork = nb * a.items_t;
  a.zero_idx = c->zero_idx;
  a.t_out0 = t_out0;
  a.dplan = c->dplan.p;
  a.splits = c->splits.p;
  a.plus_idx = c->pidx.p;
  a.minus_idx = c->pidx.p + c->npb * 16;
  a.wgt = c->use_wgt ? c->wgt.p : nullptr;
  a.psum = c->psum.p;
  a.strip = c->strip;
  a.order = c->item_order;
  return a;
}

__global__ void zero_fill_kernel(uint4* __restrict__ p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = make_uint4(0u, 0u, 0u, 0u);
}

// Elementals of timeslices t0 .. t0+nt-1 into out_dev (device) or out_host.
// With fixed != nullptr (ed_compute_noalloc) the workspace is the caller's
// fixed buffer of fixed_bytes (device address): nothing is allocated, no
// event is recorded and the host never waits on the stream; the used part of
// the buffer is cleared by a fill kernel first, because padded rows and sites
// must read zero and the buffer keeps whatever the previous call left.
template <typename T>
void run_batches(ed_context* c, int t0, int nt, double2* out_dev, void* out_host, char* fixed = nullptr,
                 size_t fixed_bytes = 0) {
  const T* ph = nullptr;
  if (sizeof(T) == 8) ph = (const T*)(void*)c->ph64.p;
  else ph = (const T*)(void*)c->ph32.p;
  const size_t phn = (size_t)c->npb * 16 * c->V3P;
  const bool noalloc = fixed != nullptr;
  int TB;
  ed::c_t<T>* W;
  if (noalloc) {
    const size_t per_t = per_t_ws_bytes<T>(c);
    if (per_t > fixed_bytes)
      fail(ED_ERR_UNSUPPORTED,
           "one timeslice needs %zu bytes of workspace; the no-allocation entry point holds at most %zu "
           "(use ed_compute)", per_t, fixed_bytes);
    TB = std::min<int>(nt, (int)std::min<size_t>(fixed_bytes / per_t, (size_t)nt));
    if (c->tbatch_opt > 0) TB = std::min(TB, c->tbatch_opt);
    W = (ed::c_t<T>*)(void*)fixed;
    const size_t n16 = per_t * TB / sizeof(uint4);
    zero_fill_kernel<<<(unsigned)std::min<size_t>(4096, (n16 + 255) / 256), 256, 0, c->stream>>>((uint4*)(void*)fixed,
                                                                                                   n16);
    HIPCHK(hipGetLastError());
  } else {
    TB = auto_tbatch<T>(c, nt);
    W = ensure_workspace<T>(c, TB);
  }
  const size_t out_t = (size_t)c->D * (c->moms.size() / 3) * c->Ne * c->Ne;

  double2* out = out_dev;
  if (!out) {
    c->stage.ensure(out_t * TB);
    out = c->stage.p;
  }
  auto launch = [c, noalloc](int kind, auto&& f) {
    if (noalloc) {
      f();
      HIPCHK(hipGetLastError());
    } else {
      c->launch(kind, f);
    }
  };

  for (int tb0 = t0; tb0 < t0 + nt; tb0 += TB) {
    const int nb = std::min(TB, t0 + nt - tb0);
    ContractArgs a = contract_args(c, nb, out_dev ? (tb0 - t0) : 0);
    const int nblk = (a.nwork + 7) / 8 * 8;
    if (sizeof(T) == 4 && use_h16(c)) {
      // the same bytes hold the hi/lo FP16 pairs; padding stays zero in both forms
      uint2* Wh = (uint2*)(void*)W;
      c->wsh = h16_shift(c);
      enqueue_stencil<ed::hsplit>(c, Wh, tb0, nb, c->stream, launch);
      a.g.wsh = c->wsh;
      if (h16_unit(c)) {
        a.unit = 1;
        a.dplan = c->dplan_u.p;
        a.splits = c->splits_u.p;
      }
      launch(ED_KERNEL_CONTRACT, [&] {
        if (h16_groups(c) <= 1)
          ed::contract_mixed16_kernel<1><<<nblk, 64 * ed::NWAVE, 0, c->stream>>>(Wh, c->ph16.p, out, a);
        else
          ed::contract_mixed16_kernel<2><<<nblk, 64 * ed::NWAVE, 0, c->stream>>>(Wh, c->ph16.p, out, a);
      });
    } else {
    enqueue_stencil<T>(c, W, tb0, nb, c->stream, launch);
    if (sizeof(T) == 4 && use_f16_stage2(c)) {
      const float2* Wf = (const float2*)(const void*)W;
      launch(ED_KERNEL_CONTRACT, [&] {
        if (use_ws_kernel(c)) {
          const bool wide = ws_wide(c);
          auto ws = [&](auto nks) {
            constexpr int K = decltype(nks)::value;
            if (wide)
              ed::contract_mixed_ws_kernel<K, true><<<nblk, 64 * ed::NWAVE, 0, c->stream>>>(Wf, c->ph16.p, out, a);
            else
              ed::contract_mixed_ws_kernel<K, false><<<nblk, 64 * ed::NWAVE, 0, c->stream>>>(Wf, c->ph16.p, out, a);
          };
          if (c->nks_max <= 1)
            ws(std::integral_constant<int, 1>{});
          else if (c->nks_max <= 2)
            ws(std::integral_constant<int, 2>{});
          else
            ws(std::integral_constant<int, 3>{});
        } else if (c->nks_max <= 1)
          ed::contract_mixed_f16_kernel<1><<<nblk, 64 * ed::NWAVE, 0, c->stream>>>(Wf, c->ph16.p, out, a);
        else if (c->nks_max <= 2)
          ed::contract_mixed_f16_kernel<2><<<nblk, 64 * ed::NWAVE, 0, c->stream>>>(Wf, c->ph16.p, out, a);
        else if (c->nks_max <= 3)
          ed::contract_mixed_f16_kernel<3><<<nblk, 64 * ed::NWAVE, 0, c->stream>>>(Wf, c->ph16.p, out, a);
        else if (c->nks_max <= 6)
          ed::contract_mixed_f16_kernel<6><<<nblk, 64 * ed::NWAVE, 0, c->stream>>>(Wf, c->ph16.p, out, a);
        else
          fail(ED_ERR_UNSUPPORTED, "derivative with %d split terms", 4 * c->nks_max / 3);
      });
    } else
    launch(ED_KERNEL_CONTRACT, [&] {
      if (c->nks_max <= 1)
        ed::contract_kernel<T, 1><<<nblk, 64 * ed::NWAVE, 0, c->stream>>>(W, ph, ph + phn, out, a);
      else if (c->nks_max <= 2)
        ed::contract_kernel<T, 2><<<nblk, 64 * ed::NWAVE, 0, c->stream>>>(W, ph, ph + phn, out, a);
      else if (c->nks_max <= 3)
        ed::contract_kernel<T, 3><<<nblk, 64 * ed::NWAVE, 0, c->stream>>>(W, ph, ph + phn, out, a);
      else if (c->nks_max <= 6)
        ed::contract_kernel<T, 6><<<nblk, 64 * ed::NWAVE, 0, c->stream>>>(W, ph, ph + phn, out, a);
      else
        fail(ED_ERR_UNSUPPORTED, "derivative with %d split terms", 4 * c->nks_max / 3);
    });
    }
    if (!out_dev) {
      HIPCHK(hipMemcpyAsync((double2*)out_host + (size_t)(tb0 - t0) * out_t, out, (size_t)nb * out_t * sizeof(double2),
                            hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
    }
  }
  if (c->profile && !noalloc) {
    HIPCHK(hipStreamSynchronize(c->stream));
    c->drain_events();
  }
}

// Reverse pass for timeslices t0 .. t0+nt-1 (FP64).  G [nt][D][P][Ne][Ne];
// gV [nt][Ne][V3][3] and gU [3][nt][V3][3][3] (either may be null).
void run_backward(ed_context* c, int t0, int nt, const double2* G, bool g_dev, double2* gV, double2* gU, bool o_dev) {
  const Geom g = c->geom(c->D);
  const int D = c->D, Ne = c->Ne, P = (int)(c->moms.size() / 3);
  const size_t wslab = (size_t)D * c->V3P * 3 * c->NeP;
  const size_t gmat = (size_t)D * P * Ne * Ne;
  const size_t bmat = (size_t)D * c->nq * 2 * Ne * c->NeP;  // grouped cotangent of one order, [D][nq][S, D][Ne][NeP]
  const size_t vt = (size_t)Ne * 3 * c->V3, ut = (size_t)27 * c->V3;
  size_t per_t = (3 * wslab + 2 * bmat) * sizeof(double2);
  if (!g_dev) per_t += gmat * sizeof(double2);
  if (!o_dev) per_t += ((gV ? vt : 0) + (gU ? ut : 0)) * sizeof(double2);
  int TB = c->tbatch_opt;
  if (TB <= 0) {
    size_t budget = (size_t)std::max<int64_t>(1, c->ws_mb) << 20;
    // at most the reverse-pass buffers already held (reused) plus three
    // quarters of the free device memory, as for the forward batch
    const size_t held = (c->bW.n + c->bWT.n + c->bgWT.n + c->bA.n + c->bG.n + c->bOutV.n + c->bOutU.n) *
                        sizeof(double2);
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) budget = std::min(budget, held + free_b / 4 * 3);
    TB = (int)std::max<size_t>(1, budget / per_t);
  }
  TB = std::min(TB, nt);
  double2* old = c->bW.p;
  c->bW.ensure(wslab * TB);
  if (c->bW.p != old || !c->bW_clear) {
    // padded rows and sites stay zero: staging and hops write only real ones
    HIPCHK(hipMemsetAsync(c->bW.p, 0, c->bW.n * sizeof(double2), c->stream));
    c->bW_clear = true;
  }
  c->bWT.ensure(wslab * TB);
  c->bgWT.ensure(wslab * TB);
  c->bA.ensure(2 * bmat * TB);
  if (!g_dev) c->bG.ensure(gmat * TB);
  if (!o_dev && gV) c->bOutV.ensure(vt * TB);
  if (!o_dev && gU) c->bOutU.ensure(ut * TB);
  double2* Brow = c->bA.p;
  double2* Bcol = c->bA.p + bmat * TB;
  const double2* U = c->U64.p;

  for (int tb0 = t0; tb0 < t0 + nt; tb0 += TB) {
    const int nb = std::min(TB, t0 + nt - tb0);
    const int tl = tb0 - t0;
    // recompute the displaced vectors in FP64
    enqueue_stencil<double>(c, c->bW.p, tb0, nb, c->stream, [](int, auto&& f) {
      f();
      HIPCHK(hipGetLastError());
    });
    // weighted cotangent in row and column order
    const double2* Gb = G + (size_t)tl * gmat;
    if (!g_dev) {
      HIPCHK(hipMemcpyAsync(c->bG.p, Gb, (size_t)nb * gmat * sizeof(double2), hipMemcpyHostToDevice, c->stream));
      Gb = c->bG.p;
    }
    const int nz = nb * D * c->nq;
    for (int z0 = 0; z0 < nz; z0 += 65535) {
      const unsigned cnt = (unsigned)std::min(65535, nz - z0);
      dim3 grid((c->NeP + 31) / 32, (Ne + 31) / 32, cnt);
      ed::adj::adj_prep_kernel<<<grid, dim3(32, 8), 0, c->stream>>>(Gb, c->use_wgt ? c->wgt.p : nullptr, c->bgroups.p,
                                                                     Brow, Bcol, Ne, c->NeP, P, c->nq, z0);
      HIPCHK(hipGetLastError());
    }
    // contraction adjoint on the transposed workspace [slot][row][site][colour]
    const int wrows = 3 * c->V3P, wcols = c->NeP;
    {
      dim3 grid((wcols + 31) / 32, (wrows + 31) / 32, nb * D);
      ed::adj::adj_transpose_kernel<<<grid, dim3(32, 8), 0, c->stream>>>(c->bW.p, c->bWT.p, wrows, wcols);
      HIPCHK(hipGetLastError());
    }
    HIPCHK(hipMemsetAsync(c->bgWT.p, 0, (size_t)nb * wslab * sizeof(double2), c->stream));
    ed::adj::AdjArgs a{};
    a.g = g;
    a.D = D;
    a.Q = c->nq;
    c->launch(ED_KERNEL_ADJOINT, [&] {
      int rb, gc;
      adj_shape(c->bns_max, rb, gc);
      auto go = [&](auto ns, auto rbc) {
        constexpr int NS = decltype(ns)::value, RB = decltype(rbc)::value;
        constexpr int GC = 8 / RB > 4 ? 4 : 8 / RB;
        dim3 grid((c->V3P + 255) / 256, c->NeP / RB, nb);
        const double2* const Wp = c->bWT.p;
        double2* const gp = c->bgWT.p;
        if (c->nq == 4)
          ed::adj::adj_contract_kernel<NS, RB, GC, 4><<<grid, 256, 0, c->stream>>>(
              Wp, gp, Brow, Bcol, c->phase_grp.p, c->bsplits.p, c->bdsplit.p, a);
        else if (c->nq == 8)
          ed::adj::adj_contract_kernel<NS, RB, GC, 8><<<grid, 256, 0, c->stream>>>(
              Wp, gp, Brow, Bcol, c->phase_grp.p, c->bsplits.p, c->bdsplit.p, a);
        else
          ed::adj::adj_contract_kernel<NS, RB, GC, 0><<<grid, 256, 0, c->stream>>>(
              Wp, gp, Brow, Bcol, c->phase_grp.p, c->bsplits.p, c->bdsplit.p, a);
      };
      using I1 = std::integral_constant<int, 1>;
      using I2 = std::integral_constant<int, 2>;
      using I4 = std::integral_constant<int, 4>;
      using I8 = std::integral_constant<int, 8>;
      if (c->bns_max > 8) fail(ED_ERR_UNSUPPORTED, "derivative with %d split terms", c->bns_max);
      if (c->bns_max <= 1)
        go(I1{}, I4{});
      else if (c->bns_max <= 2 && rb == 4)
        go(I2{}, I4{});
      else if (c->bns_max <= 2 && rb == 2)
        go(I2{}, I2{});
      else if (c->bns_max <= 2)
        go(I2{}, I1{});
      else if (c->bns_max <= 4)
        go(I4{}, I2{});
      else
        go(I8{}, I1{});
    });
    {
      dim3 grid((wrows + 31) / 32, (wcols + 31) / 32, nb * D);
      ed::adj::adj_transpose_kernel<<<grid, dim3(32, 8), 0, c->stream>>>(c->bgWT.p, c->bWT.p, wcols, wrows);
      HIPCHK(hipGetLastError());
    }
    // hop adjoint, deepest level first
    for (int k = c->num_nabla; k >= 1; --k) {
      const int parent0 = level_start(k - 1), nparent = level_start(k) - parent0;
      dim3 grid((c->V3 + 3) / 4, nparent, nb);
      ed::adj::adj_hop_kernel<<<grid, 256, 0, c->stream>>>(c->bWT.p, U, g, tb0, parent0, level_start(k));
      HIPCHK(hipGetLastError());
    }
    if (gU) {
      double2* o = o_dev ? gU : c->bOutU.p;
      dim3 grid((c->V3 + 3) / 4, 3, nb);
      ed::adj::adj_link_kernel<<<grid, 256, 0, c->stream>>>(c->bW.p, c->bWT.p, c->child_of.p, c->nparent_all, o, g,
                                                             o_dev ? tl : 0, o_dev ? nt : nb);
      HIPCHK(hipGetLastError());
      if (!o_dev)
        for (int mu = 0; mu < 3; ++mu)
          HIPCHK(hipMemcpyAsync(gU + ((size_t)mu * nt + tl) * ut / 3, c->bOutU.p + (size_t)mu * nb * ut / 3,
                                (size_t)nb * ut / 3 * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
    }
    if (gV) {
      double2* o = o_dev ? gV : c->bOutV.p;
      dim3 grid((3 * c->V3 + 31) / 32, (Ne + 31) / 32, nb);
      ed::adj::adj_out_v_kernel<<<grid, dim3(32, 8), 0, c->stream>>>(c->bWT.p, o, g, o_dev ? tl : 0);
      HIPCHK(hipGetLastError());
      if (!o_dev)
        HIPCHK(hipMemcpyAsync(gV + (size_t)tl * vt, c->bOutV.p, (size_t)nb * vt * sizeof(double2),
                              hipMemcpyDeviceToHost, c->stream));
    }
    if (!o_dev || !g_dev) HIPCHK(hipStreamSynchronize(c->stream));
  }
  if (c->profile) {
    HIPCHK(hipStreamSynchronize(c->stream));
    c->drain_events();
  }
}

// Device address of this device's `scratch`, resolved once per device with
// the device current (the caller has set it).
void resolve_scratch(int device) {
  if (device >= kMaxDevices) fail(ED_ERR_UNSUPPORTED, "device %d: at most %d devices", device, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) fail(ED_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch workspace");
  g_scratch_ptr[device] = (char*)p;
}

int guard(ed_context* ctx, const std::function<void()>& body) {
  try {
    body();
    if (ctx) ctx->err.clear();
    return ED_OK;
  } catch (const EdError& e) {
    if (ctx) ctx->err = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    if (ctx) ctx->err = e.what();
    return ED_ERR_ALLOC;
  } catch (...) {
    if (ctx) ctx->err = "unknown error";
    return ED_ERR_ARG;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Batched form: several contexts (lattices of different size and Ne) computed
// by one stream-ordered sequence -- per lattice the staging and hop kernels
// over all its timeslices, then one persistent grouped contraction per kernel
// kind over every lattice -- captured once into a hipGraph and replayed with a
// single launch.
// ---------------------------------------------------------------------------
struct ed_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  struct Item {
    ed_context* ctx;
    int mode;
    double2* out;
  };
  std::vector<Item> items;
  int grid_per_cu = 2;  // persistent grid = grid_per_cu x CUs; 0: one workgroup per item
  int order = 0;        // item order of the grouped kernels
  int strip = 0;        // tile-order strip height for every member (0: each context's own)
  int h16 = -1;         // FP16 split sum for mixed members: -1 as each context's ED_OPT_MIXED_WS, 0 never, 1 where it fits
  int use_graph = 1;
  int ncu = 256;
  // kinds: 0 FP64 MFMA, 1 FP32 MFMA (mixed, FP32 stage 2), 2 mixed FP16 stage 2,
  // 3 mixed with the FP16 split sum as well (hi/lo FP16 workspace)
  static constexpr int NKIND = 4;
  DevBuf<ed::GroupProblem> probs[NKIND];
  std::vector<ed::GroupProblem> host_probs[NKIND];
  int nprob[NKIND] = {0, 0, 0, 0}, total[NKIND] = {0, 0, 0, 0}, nks[NKIND] = {1, 1, 1, 1};
  // kind 2 runs the 64-bit-address form (some lattice's timeslice workspace
  // reaches 2^31 bytes, or ED_OPT_WIDE_WS is set on one)
  bool wide = false;
  std::vector<uint64_t> fingerprint;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  // ordering against the member contexts' streams: one event per item
  // (recorded on its context's stream before a run) and one recorded on the
  // batch stream after it
  std::vector<hipEvent_t> ctx_ev;
  hipEvent_t done_ev = nullptr;

  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    fingerprint.clear();
  }
  ~ed_batch() {
    drop_graph();
    for (auto e : ev)
      if (e) (void)hipEventDestroy(e);
    for (auto e : ctx_ev) (void)hipEventDestroy(e);
    if (done_ev) (void)hipEventDestroy(done_ev);
  }
};

namespace {

int bguard(ed_batch* b, const std::function<void()>& body) {
  try {
    body();
    if (b) b->err.clear();
    return ED_OK;
  } catch (const EdError& e) {
    if (b) b->err = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    if (b) b->err = e.what();
    return ED_ERR_ALLOC;
  } catch (...) {
    if (b) b->err = "unknown error";
    return ED_ERR_ARG;
  }
}

int batch_kind(const ed_batch* b, const ed_context* c, int mode) {
  if (mode == ED_MODE_FP64 || mixed_runs_fp64(c)) return 0;
  if (!use_f16_stage2(c)) return 1;
  const bool h16 = b->h16 < 0 ? use_h16(c) : (b->h16 > 0 && h16_fits(c));
  return h16 ? 3 : 2;
}

// Plans, workspaces (every timeslice at once) and the grouped problem tables;
// returns the fingerprint of everything a captured graph refers to.
std::vector<uint64_t> batch_prepare(ed_batch* b) {
  std::vector<ed::GroupProblem> tab[ed_batch::NKIND];
  std::vector<uint64_t> fp;
  b->wide = false;
  for (auto& it : b->items) {
    ed_context* c = it.ctx;
    require(c->lattice_set && c->gauge_set && c->eig_set && c->mom_set, ED_ERR_STATE,
            "every context of a batch needs lattice, gauge, eigenvectors and momenta");
    HIPCHK(hipSetDevice(c->device));
    build_plan(c);
    const int kind = batch_kind(b, c, it.mode);
    void* W = kind == 0 ? (void*)ensure_workspace<double>(c, c->Lt) : (void*)ensure_workspace<float>(c, c->Lt);
    ed::GroupProblem gp{};
    gp.a = contract_args(c, c->Lt, 0);
    if (b->strip > 0) gp.a.strip = b->strip;
    if (kind == 3) {
      // the same bytes hold the hi/lo FP16 pairs (as in run_batches)
      c->wsh = h16_shift(c);
      gp.a.g.wsh = c->wsh;
      if (h16_unit(c)) {
        gp.a.unit = 1;
        gp.a.dplan = c->dplan_u.p;
        gp.a.splits = c->splits_u.p;
      }
    }
    gp.W = W;
    if (kind == 0) {
      const size_t phn = (size_t)c->npb * 16 * c->V3P;
      gp.ph0 = c->ph64.p;
      gp.ph1 = c->ph64.p + phn;
    } else if (kind == 1) {
      const size_t phn = (size_t)c->npb * 16 * c->V3P;
      gp.ph0 = c->ph32.p;
      gp.ph1 = c->ph32.p + phn;
    } else {
      gp.ph0 = c->ph16.p;
      gp.ph1 = nullptr;
    }
    gp.out = it.out;
    gp.work0 = b->total[kind];
    tab[kind].push_back(gp);
    b->nks[kind] = std::max(b->nks[kind], kind == 3 ? h16_groups(c) : c->nks_max);
    if (kind == 2 && ws_wide(c)) b->wide = true;
    fp.push_back((uint64_t)(uintptr_t)c);
    fp.push_back(c->gen);
    fp.push_back((uint64_t)(uintptr_t)W);
    fp.push_back((uint64_t)(uintptr_t)c->U64.p);
    fp.push_back((uint64_t)(uintptr_t)c->V.p);
    fp.push_back((uint64_t)(uintptr_t)c->V64.p);
    fp.push_back((uint64_t)kind);
    fp.push_back((uint64_t)(int64_t)c->wsh);
    int64_t tot = (int64_t)b->total[kind] + gp.a.nwork;
    if (tot >= ((int64_t)1 << 31)) fail(ED_ERR_UNSUPPORTED, "batch has too many work items");
    b->total[kind] = (int)tot;
  }
  // the grouped kernels' 64-bit-address form covers up to four split terms
  require(!(b->wide && b->nks[2] > 3), ED_ERR_UNSUPPORTED,
          "a lattice with 2^31 or more bytes of workspace per timeslice shares the mixed batch with "
          "derivatives of more than four split terms; batch them separately");
  for (int k = 0; k < ed_batch::NKIND; ++k) {
    b->nprob[k] = (int)tab[k].size();
    if (tab[k].empty()) continue;
    const size_t bytes = tab[k].size() * sizeof(ed::GroupProblem);
    ed::GroupProblem* old = b->probs[k].p;
    b->probs[k].ensure(tab[k].size());
    // upload only what changed: a replayed graph may still be reading the table
    if (b->probs[k].p != old || b->host_probs[k].size() != tab[k].size() ||
        std::memcmp(b->host_probs[k].data(), tab[k].data(), bytes) != 0) {
      HIPCHK(hipStreamSynchronize(b->stream));
      HIPCHK(hipMemcpy(b->probs[k].p, tab[k].data(), bytes, hipMemcpyHostToDevice));
      b->host_probs[k] = tab[k];
    }
    fp.push_back((uint64_t)(uintptr_t)b->probs[k].p);
  }
  fp.push_back((uint64_t)b->grid_per_cu);
  fp.push_back((uint64_t)b->order);
  fp.push_back((uint64_t)b->strip);
  fp.push_back((uint64_t)b->wide);
  fp.push_back((uint64_t)(int64_t)b->h16);
  return fp;
}

int batch_grid(const ed_batch* b, int total) {
  const int full = (total + 7) / 8 * 8;
  if (b->grid_per_cu <= 0) return full;
  const int pers = (b->grid_per_cu * b->ncu + 7) / 8 * 8;
  return std::min(full, pers);
}

void batch_enqueue(ed_batch* b) {
  auto direct = [](int, auto&& f) {
    f();
    HIPCHK(hipGetLastError());
  };
  for (auto& it : b->items) {
    ed_context* c = it.ctx;
    const int kind = batch_kind(b, c, it.mode);
    if (kind == 0)
      enqueue_stencil<double>(c, (double2*)(void*)c->W.p, 0, c->Lt, b->stream, direct);
    else if (kind == 3)
      enqueue_stencil<ed::hsplit>(c, (uint2*)(void*)c->W.p, 0, c->Lt, b->stream, direct);
    else
      enqueue_stencil<float>(c, (float2*)(void*)c->W.p, 0, c->Lt, b->stream, direct);
  }
  const dim3 blk(64 * ed::NWAVE);
  if (b->nprob[3]) {
    // FP16 split sum: one workgroup per work item (the persistent shapes are
    // not built for it), at most two k-groups (h16_fits)
    const int tot = b->total[3];
    if (b->nks[3] <= 1)
      ed::contract_mixed16_item_kernel<1><<<(tot + 7) / 8 * 8, blk, 0, b->stream>>>(b->probs[3].p, b->nprob[3], tot,
                                                                                    b->order);
    else
      ed::contract_mixed16_item_kernel<2><<<(tot + 7) / 8 * 8, blk, 0, b->stream>>>(b->probs[3].p, b->nprob[3], tot,
                                                                                    b->order);
    HIPCHK(hipGetLastError());
  }
  for (int k = 0; k < 3; ++k) {
    if (!b->nprob[k]) continue;
    const int grid = batch_grid(b, b->total[k]);
    const ed::GroupProblem* P = b->probs[k].p;
    const int np = b->nprob[k], tot = b->total[k], ord = b->order, nks = b->nks[k];
    if (nks > 6) fail(ED_ERR_UNSUPPORTED, "derivative with more split terms than the batched kernels take");
    const int ks = nks <= 1 ? 1 : nks <= 2 ? 2 : nks <= 3 ? 3 : 6;
    const bool wide = k == 2 && b->wide;  // ks <= 3 then (batch_prepare)
#define ED_GROUP(KS, WD)                                                                                   \
  do {                                                                                                     \
    if (k == 0) ed::contract_group_kernel<double, KS><<<grid, blk, 0, b->stream>>>(P, np, tot, ord);      \
    else if (k == 1) ed::contract_group_kernel<float, KS><<<grid, blk, 0, b->stream>>>(P, np, tot, ord);  \
    else if (b->grid_per_cu <= 0) ed::contract_mixed_item_kernel<KS, WD><<<grid, blk, 0, b->stream>>>(P, np, tot, ord); \
    else ed::contract_mixed_group_kernel<KS, WD><<<grid, blk, 0, b->stream>>>(P, np, tot, ord);           \
  } while (0)
    if (ks == 1 && wide) ED_GROUP(1, true);
    else if (ks == 1) ED_GROUP(1, false);
    else if (ks == 2 && wide) ED_GROUP(2, true);
    else if (ks == 2) ED_GROUP(2, false);
    else if (ks == 3 && wide) ED_GROUP(3, true);
    else if (ks == 3) ED_GROUP(3, false);
    else ED_GROUP(6, false);
#undef ED_GROUP
    HIPCHK(hipGetLastError());
  }
}

// The batch stream waits for everything enqueued so far on each member
// context's stream (ed_compute, ed_set_eigvecs from device memory,
// ed_stout_smear write the workspace, V and the links the batch reads).
void batch_wait_members(ed_batch* b) {
  if (b->ctx_ev.size() < b->items.size()) {
    for (size_t i = b->ctx_ev.size(); i < b->items.size(); ++i) {
      hipEvent_t e;
      HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      b->ctx_ev.push_back(e);
    }
  }
  for (size_t i = 0; i < b->items.size(); ++i) {
    HIPCHK(hipEventRecord(b->ctx_ev[i], b->items[i].ctx->stream));
    HIPCHK(hipStreamWaitEvent(b->stream, b->ctx_ev[i], 0));
  }
}

// Each member context's stream waits for the run just enqueued on the batch
// stream, so later calls on a context see the batch's outputs and do not
// overwrite a workspace the batch still reads.
void batch_release_members(ed_batch* b) {
  if (!b->done_ev) HIPCHK(hipEventCreateWithFlags(&b->done_ev, hipEventDisableTiming));
  HIPCHK(hipEventRecord(b->done_ev, b->stream));
  for (auto& it : b->items) HIPCHK(hipStreamWaitEvent(it.ctx->stream, b->done_ev, 0));
}

void batch_run(ed_batch* b) {
  require(!b->items.empty(), ED_ERR_STATE, "the batch is empty");
  HIPCHK(hipSetDevice(b->device));
  for (int k = 0; k < ed_batch::NKIND; ++k) {
    b->total[k] = 0;
    b->nks[k] = 1;
  }
  if (!b->use_graph) {
    b->drop_graph();
    batch_prepare(b);
    batch_wait_members(b);
    batch_enqueue(b);
    batch_release_members(b);
    return;
  }
  // a fresh plan may reallocate buffers; recapture when anything moved
  std::vector<uint64_t> fp = batch_prepare(b);
  if (!b->exec || fp != b->fingerprint) {
    b->drop_graph();
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    try {
      batch_enqueue(b);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(b->stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIPCHK(hipStreamEndCapture(b->stream, &b->graph));
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->fingerprint = fp;
  }
  batch_wait_members(b);
  HIPCHK(hipGraphLaunch(b->exec, b->stream));
  batch_release_members(b);
}

double batch_time(ed_batch* b, int reps) {
  if (!b->ev[0]) {
    HIPCHK(hipEventCreate(&b->ev[0]));
    HIPCHK(hipEventCreate(&b->ev[1]));
  }
  batch_run(b);  // warm-up (and capture)
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipEventRecord(b->ev[0], b->stream));
  for (int r = 0; r < reps; ++r) batch_run(b);
  HIPCHK(hipEventRecord(b->ev[1], b->stream));
  HIPCHK(hipEventSynchronize(b->ev[1]));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
  return (double)ms / reps;
}

}  // namespace

extern "C" {

const char* ed_version(void) { return "elemental_hip 0.1 (gfx950)"; }

int ed_create(int device, ed_context** ctx) {
  if (!ctx) return ED_ERR_ARG;
  *ctx = nullptr;
  ed_context* c = new (std::nothrow) ed_context();
  if (!c) return ED_ERR_ALLOC;
  int rc = guard(c, [&] {
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) fail(ED_ERR_ARG, "device %d out of range (%d devices)", device, n);
    c->device = device;
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking));
    c->stream = c->own;
    resolve_scratch(device);
  });
  if (rc != ED_OK) {
    delete c;
    return rc;
  }
  *ctx = c;
  return ED_OK;
}

int ed_destroy(ed_context* ctx) {
  if (!ctx) return ED_ERR_ARG;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->own) (void)hipStreamDestroy(ctx->own);
  delete ctx;
  return ED_OK;
}

const char* ed_last_error(const ed_context* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int ed_set_stream(ed_context* ctx, void* s) {
  if (!ctx) return ED_ERR_ARG;
  return guard(ctx, [&] { ctx->stream = s ? (hipStream_t)s : ctx->own; });
}

int ed_set_lattice(ed_context* ctx, int Lx, int Ly, int Lz, int Lt, int Ne) {
  if (!ctx) return ED_ERR_ARG;
  return guard(ctx, [&] {
    ++ctx->gen;
    if (Lx < 1 || Ly < 1 || Lz < 1 || Lt < 1 || Ne < 1) fail(ED_ERR_ARG, "lattice sizes and Ne must be >= 1");
    const long v3 = (long)Lx * Ly * Lz;
    if (v3 > (1L << 26)) fail(ED_ERR_ARG, "spatial volume %ld too large", v3);
    if (Ne > 3 * v3) fail(ED_ERR_ARG, "Ne = %d exceeds 3*V3 = %ld", Ne, 3 * v3);
    HIPCHK(hipSetDevice(ctx->device));
    ctx->Lx = Lx;
    ctx->Ly = Ly;
    ctx->Lz = Lz;
    ctx->Lt = Lt;
    ctx->Ne = Ne;
    ctx->V3 = (int)v3;
    ctx->V3P = (int)((v3 + ed::SC2 - 1) / ed::SC2 * ed::SC2);  // multiple of both chunk sizes
    ctx->NeP = (Ne + 15) / 16 * 16;
    ctx->lattice_set = true;
    ctx->gauge_set = ctx->eig_set = ctx->v64_set = false;
    ctx->use_wgt = false;
    ctx->plan_dirty = true;
    ctx->W.reset();
    ctx->W_elem = 0;
    ctx->stage.reset();
    ctx->bW.reset();
    ctx->bW_clear = false;
  });
}

int ed_set_gauge(ed_context* ctx, const void* U, int dtype, int on_device) {
  if (!ctx) return ED_ERR_ARG;
  return guard(ctx, [&] {
    ++ctx->gen;
    require(ctx->lattice_set, ED_ERR_STATE, "ed_set_lattice must be called first");
    if (!U) fail(ED_ERR_ARG, "U is null");
    if (dtype != ED_C128 && dtype != ED_C64) fail(ED_ERR_ARG, "dtype must be ED_C128 or ED_C64");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t n = (size_t)3 * ctx->Lt * ctx->V3 * 9;
    ctx->U64.ensure(n);
    if (dtype == ED_C128) {
      HIPCHK(hipMemcpyAsync(ctx->U64.p, U, n * sizeof(double2), on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                            ctx->stream));
    } else {
      DevBuf<float2> tmp;
      tmp.ensure(n);
      HIPCHK(hipMemcpyAsync(tmp.p, U, n * sizeof(float2), on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                            ctx->stream));
      c64_to_c128_kernel<<<1024, 256, 0, ctx->stream>>>(tmp.p, ctx->U64.p, n);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->gauge_set = true;
  });
}

int ed_set_eigvecs(ed_context* ctx, const void* V, int dtype, int on_device) {
  if (!ctx) return ED_ERR_ARG;
  return guard(ctx, [&] {
    ++ctx->gen;
    require(ctx->lattice_set, ED_ERR_STATE, "ed_set_lattice must be called first");
    if (!V) fail(ED_ERR_ARG, "V is null");
    if (dtype != ED_C128 && dtype != ED_C64) fail(ED_ERR_ARG, "dtype must be ED_C128 or ED_C64");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t n = (size_t)ctx->Lt * ctx->Ne * ctx->V3 * 3;
    ctx->V.ensure(n);
    if (dtype == ED_C64) {
      HIPCHK(hipMemcpyAsync(ctx->V.p, V, n * sizeof(float2), on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                            ctx->stream));
    } else {
      // round complex128 to complex64 (the reference's c8 staging buffer)
      const size_t chunk = (size_t)1 << 24;
      DevBuf<double2> tmp;
      tmp.ensure(std::min(n, chunk));
      for (size_t s = 0; s < n; s += chunk) {
        const size_t m = std::min(chunk, n - s);
        HIPCHK(hipMemcpyAsync(tmp.p, (const double2*)V + s, m * sizeof(double2),
                              on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx->stream));
        c128_to_c64_kernel<<<1024, 256, 0, ctx->stream>>>(tmp.p, ctx->V.p + s, m);
        HIPCHK(hipGetLastError());
      }
    }
    ctx->v64_set = false;
    if (ctx->exact_v) {
      ctx->V64.ensure(n);
      if (dtype == ED_C128) {
        HIPCHK(hipMemcpyAsync(ctx->V64.p, V, n * sizeof(double2),
                              on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx->stream));
      } else {
        c64_to_c128_kernel<<<1024, 256, 0, ctx->stream>>>(ctx->V.p, ctx->V64.p, n);
        HIPCHK(hipGetLastError());
      }
      ctx->v64_set = true;
    } else {
      ctx->V64.reset();
    }
    {
      // scale of the FP16 stage-1 workspace: from the largest component
      DevBuf<unsigned int> mx;
      mx.ensure(1);
      HIPCHK(hipMemsetAsync(mx.p, 0, sizeof(unsigned int), ctx->stream));
      absmax_c64_kernel<<<1024, 256, 0, ctx->stream>>>(ctx->V.p, n, mx.p);
      HIPCHK(hipGetLastError());
      unsigned int h = 0;
      HIPCHK(hipMemcpyAsync(&h, mx.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(hipStreamSynchronize(ctx->stream));
      float f;
      std::memcpy(&f, &h, sizeof(f));
      if (!std::isfinite(f)) fail(ED_ERR_ARG, "eigenvectors hold a non-finite value");
      ctx->vmax = f;
    }
    {
      // FP64 orthonormality of the stored eigenvectors, every timeslice
      DevBuf<unsigned long long> dv;
      dv.ensure(1);
      HIPCHK(hipMemsetAsync(dv.p, 0, sizeof(unsigned long long), ctx->stream));
      const size_t K = (size_t)3 * ctx->V3;
      const unsigned nb = (unsigned)((ctx->Ne + 31) / 32);
      for (int t0 = 0; t0 < ctx->Lt; t0 += 1024) {
        const unsigned nz = (unsigned)std::min(1024, ctx->Lt - t0);
        gram_dev_kernel<<<dim3(nb, nb, nz), 256, 0, ctx->stream>>>(ctx->V.p, ctx->Ne, K, t0, dv.p);
        HIPCHK(hipGetLastError());
      }
      unsigned long long h = 0;
      HIPCHK(hipMemcpyAsync(&h, dv.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(hipStreamSynchronize(ctx->stream));
      std::memcpy(&ctx->vv_dev, &h, sizeof(double));
      ctx->eig_set = false;
      if (ctx->ortho_gate > 0 && !(ctx->vv_dev <= std::pow(10.0, -ctx->ortho_gate)))
        fail(ED_ERR_ARG, "eigenvectors are not orthonormal: max|V^dag V - 1| = %.3e > 1e-%d", ctx->vv_dev,
             ctx->ortho_gate);
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->eig_set = true;
  });
}

int ed_set_momenta(ed_context* ctx, const int32_t* n, int P) {
  if (!ctx) return ED_ERR_ARG;
  return guard(ctx, [&] {
    ++ctx->gen;
    if (P < 1 || !n) fail(ED_ERR_ARG, "momentum list must be non-empty");
    if (P > 4096) fail(ED_ERR_ARG, "at most 4096 momenta");
    ctx->moms.assign(n, n + 3 * P);
    ctx->mom_set = true;
    ctx->plan_dirty = true;
  });
}

int ed_set_derivatives(ed_context* ctx, int num_nabla) {
  if (!ctx) return ED_ERR_ARG;
  return guard(ctx, [&] {
    ++ctx->gen;
    if (num_nabla < 0 || num_nabla > 3) fail(ED_ERR_ARG, "num_nabla must be in [0, 3]");
    ctx->disp = false;
    ctx->distance = 0;
    ctx->num_nabla = num_nabla;
    ctx->D = level_start(num_nabla + 1);
    ctx->plan_dirty = true;
    ctx->W.reset();
    ctx->W_elem = 0;
    ctx->stage.reset();
    ctx->bW.reset();
    ctx->bW_clear = false;
  });
}

int ed_set_displacement(ed_context* ctx, int distance) {
  if (!ctx) return ED_ERR_ARG;
  return guard(ctx, [&] {
    ++ctx->gen;
    if (distance < 0 || distance > 64) fail(ED_ERR_ARG, "distance must be in [0, 64]");
    ctx->disp = true;
    ctx->distance = distance;
    ctx->num_nabla = 0;
    ctx->D = distance + 1;
    ctx->plan_dirty = true;
    ctx->W.reset();
    ctx->W_elem = 0;
    ctx->stage.reset();
    ctx->bW.reset();
    ctx->bW_clear = false;
  });
}

int ed_set_weights(ed_context* ctx, const double* coeff) {
  if (!ctx) return ED_ERR_ARG;
  return guard(ctx, [&] {
    ++ctx->gen;
    require(ctx->lattice_set, ED_ERR_STATE, "ed_set_lattice must be called first");
    HIPCHK(hipSetDevice(ctx->device));
    if (!coeff) {
      ctx->use_wgt = false;
      return;
    }
    const size_t n = (size_t)ctx->Ne * ctx->Ne;
    ctx->wgt.ensure(n);
    HIPCHK(hipMemcpy(ctx->wgt.p, coeff, n * sizeof(double), hipMemcpyHostToDevice));
    ctx->use_wgt = true;
  });
}

int ed_set_option(ed_context* ctx, int option, int64_t value) {
  if (!ctx) return ED_ERR_ARG;
  return guard(ctx, [&] {
    ++ctx->gen;
    switch (option) {
      case ED_OPT_HERMITIAN:
        ctx->hermitian = value ? 1 : 0;
        ctx->plan_dirty = true;
        break;
      case ED_OPT_TBATCH:
        if (value < 0) fail(ED_ERR_ARG, "batch must be >= 0");
        ctx->tbatch_opt = (int)value;
        break;
      case ED_OPT_WS_MB:
        if (value < 1) fail(ED_ERR_ARG, "workspace budget must be >= 1 MiB");
        ctx->ws_mb = value;
        break;
      case ED_OPT_PROFILE:
        ctx->profile = value ? 1 : 0;
        break;
      case ED_OPT_MIXED_WS:
        if (value < 0 || value > 2)
          fail(ED_ERR_ARG, "mixed kernel kind must be 0 (uniform), 1 (FP32 stage 1) or 2 (FP16 stage 1)");
        ctx->mixed_ws = (int)value;
        break;
      case ED_OPT_EXACT_V:
        ctx->exact_v = value ? 1 : 0;
        break;
      case ED_OPT_MIXED_STAGE2:
        if (value < -1 || value > 1) fail(ED_ERR_ARG, "mixed stage-2 kind must be -1 (auto), 0 (FP32) or 1 (FP16 split)");
        ctx->mixed_stage2 = (int)value;
        break;
      case ED_OPT_WIDE_WS:
        ctx->wide_ws = value ? 1 : 0;
        break;
      case ED_OPT_ORTHO_GATE:
        if (value < 0 || value > 15) fail(ED_ERR_ARG, "orthonormality gate must be in [0, 15]");
        ctx->ortho_gate = (int)value;
        break;
      case ED_OPT_STRIP:
        if (value < 0 || value > 64) fail(ED_ERR_ARG, "strip height must be in [0, 64]");
        ctx->strip = (int)value;
        break;
      case ED_OPT_ITEM_ORDER:
        if (value < 0 || value > 1) fail(ED_ERR_ARG, "item order must be 0 or 1");
        ctx->item_order = (int)value;
        break;
      case ED_OPT_H16_UNIT:
        if (value < 0 || value > 1) fail(ED_ERR_ARG, "ED_OPT_H16_UNIT must be 0 or 1");
        ctx->h16_unit_opt = (int)value;
        break;
      case ED_OPT_HOP_LDS:
        if (value < -1 || value > 1) fail(ED_ERR_ARG, "ED_OPT_HOP_LDS must be -1, 0 or 1");
        ctx->hop_lds = (int)value;
        break;
      default:
        fail(ED_ERR_ARG, "unknown option %d", option);
    }
  });
}

int ed_compute(ed_context* ctx, int t0, int nt, int mode, void* out, int out_on_device) {
  if (!ctx) return ED_ERR_ARG;
  return guard(ctx, [&] {
    require(ctx->lattice_set && ctx->gauge_set && ctx->eig_set && ctx->mom_set, ED_ERR_STATE,
            "lattice, gauge, eigenvectors and momenta must be set before ed_compute");
    if (!out) fail(ED_ERR_ARG, "out is null");
    if (t0 < 0 || nt < 1 || t0 + nt > ctx->Lt) fail(ED_ERR_ARG, "timeslice range [%d, %d) outside [0, %d)", t0, t0 + nt, ctx->Lt);
    if (mode != ED_MODE_FP64 && mode != ED_MODE_MIXED) fail(ED_ERR_ARG, "unknown mode %d", mode);
    HIPCHK(hipSetDevice(ctx->device));
    build_plan(ctx);
    if (mode == ED_MODE_FP64 || mixed_runs_fp64(ctx))
      run_batches<double>(ctx, t0, nt, out_on_device ? (double2*)out : nullptr, out_on_device ? nullptr : out);
    else
      run_batches<float>(ctx, t0, nt, out_on_device ? (double2*)out : nullptr, out_on_device ? nullptr : out);
  });
}

int ed_compute_noalloc(ed_context* ctx, int t0, int nt, int mode, void* out_device) {
  if (!ctx) return ED_ERR_ARG;
  return guard(ctx, [&] {
    require(ctx->lattice_set && ctx->gauge_set && ctx->eig_set && ctx->mom_set, ED_ERR_STATE,
            "lattice, gauge, eigenvectors and momenta must be set before ed_compute_noalloc");
    if (!out_device) fail(ED_ERR_ARG, "out_device is null");
    if (t0 < 0 || nt < 1 || t0 + nt > ctx->Lt) fail(ED_ERR_ARG, "timeslice range [%d, %d) outside [0, %d)", t0, t0 + nt, ctx->Lt);
    if (mode != ED_MODE_FP64 && mode != ED_MODE_MIXED) fail(ED_ERR_ARG, "unknown mode %d", mode);
    // building the plan allocates its tables: it must already be built
    require(!ctx->plan_dirty, ED_ERR_STATE,
            "the plan is not built: call ed_compute or ed_query once after the last ed_set_* / ed_set_option call");
    char* fixed = g_scratch_ptr[ctx->device];
    require(fixed != nullptr, ED_ERR_STATE, "no scratch workspace on this device");
    HIPCHK(hipSetDevice(ctx->device));
    if (mode == ED_MODE_FP64 || mixed_runs_fp64(ctx))
      run_batches<double>(ctx, t0, nt, (double2*)out_device, nullptr, fixed, ED_NOALLOC_SCRATCH_BYTES);
    else
      run_batches<float>(ctx, t0, nt, (double2*)out_device, nullptr, fixed, ED_NOALLOC_SCRATCH_BYTES);
  });
}

int ed_backward(ed_context* ctx, int t0, int nt, const void* grad_E, int grad_on_device, void* grad_V, void* grad_U,
                int out_on_device) {
  if (!ctx) return ED_ERR_ARG;
  return guard(ctx, [&] {
    require(ctx->lattice_set && ctx->gauge_set && ctx->eig_set && ctx->mom_set, ED_ERR_STATE,
            "lattice, gauge, eigenvectors and momenta must be set before ed_backward");
    if (ctx->disp) fail(ED_ERR_UNSUPPORTED, "ed_backward covers derivative elementals, not displacement mode");
    if (!grad_E) fail(ED_ERR_ARG, "grad_E is null");
    if (!grad_V && !grad_U) fail(ED_ERR_ARG, "grad_V and grad_U are both null");
    if (t0 < 0 || nt < 1 || t0 + nt > ctx->Lt) fail(ED_ERR_ARG, "timeslice range [%d, %d) outside [0, %d)", t0, t0 + nt, ctx->Lt);
    HIPCHK(hipSetDevice(ctx->device));
    build_plan(ctx);
    run_backward(ctx, t0, nt, (const double2*)grad_E, grad_on_device != 0, (double2*)grad_V, (double2*)grad_U,
                 out_on_device != 0);
  });
}

int ed_query(ed_context* ctx, int what, int64_t* value) {
  if (!ctx) return ED_ERR_ARG;
  return guard(ctx, [&] {
    if (!value) fail(ED_ERR_ARG, "value is null");
    build_plan(ctx);
    // work items per (t, momentum block): every tile for most derivatives,
    // the upper triangle for palindromic ones
    const int64_t nt = ctx->NeP / 16, tiles[2] = {nt * nt, nt * (nt + 1) / 2};
    auto tw = [&](const int64_t (&v)[2]) { return tiles[0] * v[0] + tiles[1] * v[1]; };
    const int64_t chunks = ctx->V3P / ed::SC;
    const int64_t P = (int64_t)(ctx->moms.size() / 3);
    switch (what) {
      case ED_Q_CONTRACT_FLOPS_PER_T: {
        // per work item and chunk: stage 1 (16 sites x 4 MFMAs per k-step) and
        // stage 2 (4 waves x 4 k-steps x 4 N-blocks x 4 real products)
        const int64_t mf = ctx->npb * chunks * (tw(ctx->mfma_per_chunk_stage1) + tw(ctx->nd_tri) * 256);
        *value = mf * 2 * 16 * 16 * 4;
        break;
      }
      case ED_Q_MIXED_F32_FLOPS_PER_T:
      case ED_Q_MIXED_F16_FLOPS_PER_T: {
        // the mixed FP16 kernel: per (tile, d, 32-site chunk) stage 1 has 32 * 4 * nks(d)
        // FP32 MFMAs 16x16x4 and stage 2 has 16 N-blocks x 12 FP16 MFMAs 16x16x32
        const int64_t chunks2 = ctx->V3P / ed::SC2;
        if (use_h16(ctx)) {
          // FP16 stage 1: 32 sites x 4 MFMAs 16x16x32 per k-group; stage 2:
          // 16 N-blocks x 12 MFMAs
          *value = what == ED_Q_MIXED_F32_FLOPS_PER_T
                       ? 0
                       : ctx->npb * chunks2 *
                             (32 * 4 * tw(h16_unit(ctx) ? ctx->h16u_groups_sum : ctx->h16_groups_sum) +
                              tw(ctx->nd_tri) * 16 * 12) *
                             2 * 16 * 16 * 32;
          break;
        }
        if (what == ED_Q_MIXED_F32_FLOPS_PER_T)
          *value = ctx->npb * chunks2 * (2 * tw(ctx->mfma_per_chunk_stage1)) * 2 * 16 * 16 * 4;
        else
          *value = ctx->npb * chunks2 * tw(ctx->nd_tri) * 16 * 12 * 2 * 16 * 16 * 32;
        break;
      }
      case ED_Q_ALGO_FLOPS_PER_T:
        *value = 8 * (int64_t)ctx->Ne * ctx->Ne * 3 * ctx->V3 * ctx->total_splits * P;
        break;
      case ED_Q_PRIMARY_DERIVATIVES: *value = ctx->nd; break;
      case ED_Q_MOMENTUM_BLOCKS: *value = ctx->npb; break;
      case ED_Q_HOPS_PER_T: *value = ctx->disp ? 6 * ctx->distance : ctx->D - 1; break;
      case ED_Q_SPLITS: *value = ctx->total_splits; break;
      case ED_Q_MIXED_KERNEL: *value = mixed_kernel_id(ctx); break;
      default: fail(ED_ERR_ARG, "unknown query %d", what);
    }
  });
}

int ed_get_option(ed_context* ctx, int option, int64_t* value) {
  if (!ctx) return ED_ERR_ARG;
  return guard(ctx, [&] {
    if (!value) fail(ED_ERR_ARG, "value is null");
    switch (option) {
      case ED_OPT_HERMITIAN: *value = ctx->hermitian; break;
      case ED_OPT_TBATCH: *value = ctx->tbatch_opt; break;
      case ED_OPT_WS_MB: *value = ctx->ws_mb; break;
      case ED_OPT_PROFILE: *value = ctx->profile; break;
      case ED_OPT_MIXED_STAGE2: *value = ctx->mixed_stage2; break;
      case ED_OPT_EXACT_V: *value = ctx->exact_v; break;
      case ED_OPT_MIXED_WS: *value = ctx->mixed_ws; break;
      case ED_OPT_WIDE_WS: *value = ctx->wide_ws; break;
      case ED_OPT_ORTHO_GATE: *value = ctx->ortho_gate; break;
      case ED_OPT_STRIP: *value = ctx->strip; break;
      case ED_OPT_ITEM_ORDER: *value = ctx->item_order; break;
      case ED_OPT_H16_UNIT: *value = ctx->h16_unit_opt; break;
      case ED_OPT_HOP_LDS: *value = ctx->hop_lds; break;
      default: fail(ED_ERR_ARG, "unknown option %d", option);
    }
  });
}

int ed_tune(ed_context* ctx, int mode, int t0, int nt, int reps, double* best_ms) {
  if (!ctx) return ED_ERR_ARG;
  return guard(ctx, [&] {
    require(ctx->lattice_set && ctx->gauge_set && ctx->eig_set && ctx->mom_set, ED_ERR_STATE,
            "lattice, gauge, eigenvectors and momenta must be set before ed_tune");
    if (t0 < 0 || nt < 1 || t0 + nt > ctx->Lt) fail(ED_ERR_ARG, "timeslice range [%d, %d) outside [0, %d)", t0, t0 + nt, ctx->Lt);
    if (mode != ED_MODE_FP64 && mode != ED_MODE_MIXED) fail(ED_ERR_ARG, "unknown mode %d", mode);
    if (reps < 1) fail(ED_ERR_ARG, "reps must be >= 1");
    HIPCHK(hipSetDevice(ctx->device));
    build_plan(ctx);
    DevBuf<double2> out;
    out.ensure((size_t)nt * ctx->D * (ctx->moms.size() / 3) * ctx->Ne * ctx->Ne);
    hipEvent_t ev0, ev1;
    HIPCHK(hipEventCreate(&ev0));
    HIPCHK(hipEventCreate(&ev1));
    const int prof = ctx->profile;
    ctx->profile = 0;
    const bool fp = mode == ED_MODE_FP64 || mixed_runs_fp64(ctx);
    auto run = [&] {
      if (fp) run_batches<double>(ctx, t0, nt, out.p, nullptr);
      else run_batches<float>(ctx, t0, nt, out.p, nullptr);
    };
    // candidate mixed kernels: each distinct kernel the routing gives for
    // kinds 1 (warp-specialized), 0 (uniform waves) and 2 (FP16 stage 1)
    std::vector<int> kinds;
    std::vector<int> seen;
    const int kind0 = ctx->mixed_ws;
    for (int k : {1, 0, 2}) {
      if (fp && k != kind0) continue;
      ctx->mixed_ws = k;
      const int id = fp ? -1 : mixed_kernel_id(ctx);
      if (std::find(seen.begin(), seen.end(), id) != seen.end()) continue;
      seen.push_back(id);
      kinds.push_back(k);
    }
    double best = 1e300;
    int bk = kind0, bs = ctx->strip, bo = ctx->item_order;
    std::exception_ptr err;
    try {
      for (int k : kinds)
        for (int st : {1, 2, 4, 8})
          for (int o = 0; o < 2; ++o) {
            ctx->mixed_ws = k;
            ctx->strip = st;
            ctx->item_order = o;
            run();  // warm-up (workspace of this kind)
            HIPCHK(hipEventRecord(ev0, ctx->stream));
            for (int r = 0; r < reps; ++r) run();
            HIPCHK(hipEventRecord(ev1, ctx->stream));
            HIPCHK(hipEventSynchronize(ev1));
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
            if (ms / reps < best) {
              best = ms / reps;
              bk = k;
              bs = st;
              bo = o;
            }
          }
    } catch (...) {
      err = std::current_exception();
    }
    ctx->mixed_ws = bk;
    ctx->strip = bs;
    ctx->item_order = bo;
    ctx->profile = prof;
    ++ctx->gen;
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    if (err) std::rethrow_exception(err);
    if (best_ms) *best_ms = best;
  });
}

int ed_orthonormality(ed_context* ctx, double* max_dev) {
  if (!ctx) return ED_ERR_ARG;
  return guard(ctx, [&] {
    if (!max_dev) fail(ED_ERR_ARG, "max_dev is null");
    require(ctx->vv_dev >= 0.0, ED_ERR_STATE, "ed_set_eigvecs has not measured any eigenvectors yet");
    *max_dev = ctx->vv_dev;
  });
}

int ed_synchronize(ed_context* ctx) {
  if (!ctx) return ED_ERR_ARG;
  return guard(ctx, [&] {
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->drain_events();
  });
}

int ed_profile_read(ed_context* ctx, int kernel, double* total_ms, int64_t* launches) {
  if (!ctx) return ED_ERR_ARG;
  return guard(ctx, [&] {
    if (kernel < 0 || kernel > 3) fail(ED_ERR_ARG, "unknown kernel family %d", kernel);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->drain_events();
    if (total_ms) *total_ms = ctx->prof_ms[kernel];
    if (launches) *launches = ctx->prof_n[kernel];
  });
}

int ed_profile_reset(ed_context* ctx) {
  if (!ctx) return ED_ERR_ARG;
  return guard(ctx, [&] {
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->drain_events();
    for (int k = 0; k < 4; ++k) {
      ctx->prof_ms[k] = 0;
      ctx->prof_n[k] = 0;
    }
  });
}

int ed_project_su3(ed_context* ctx, int max_iter, int* iterations) {
  if (!ctx) return ED_ERR_ARG;
  return guard(ctx, [&] {
    ++ctx->gen;
    require(ctx->gauge_set, ED_ERR_STATE, "ed_set_gauge must be called first");
    if (max_iter < 0) fail(ED_ERR_ARG, "max_iter must be >= 0");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t nl = (size_t)3 * ctx->Lt * ctx->V3;
    DevBuf<double2> nxt;
    nxt.ensure(9 * nl);
    DevBuf<int> flag;
    flag.ensure(1);
    const unsigned blocks = (unsigned)((nl + 255) / 256);
    int it = 0;
    for (;; ++it) {
      HIPCHK(hipMemsetAsync(flag.p, 0, sizeof(int), ctx->stream));
      ed::gauge::polar_step_kernel<<<blocks, 256, 0, ctx->stream>>>(ctx->U64.p, nxt.p, nl, 1e-15, flag.p);
      HIPCHK(hipGetLastError());
      int h = 0;
      HIPCHK(hipMemcpyAsync(&h, flag.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(hipStreamSynchronize(ctx->stream));
      if (!h) break;  // every link passed the test: U is the result
      if (it == max_iter) fail(ED_ERR_UNSUPPORTED, "unitary projection did not converge in %d iterations", max_iter);
      std::swap(ctx->U64.p, nxt.p);
      std::swap(ctx->U64.n, nxt.n);
    }
    if (iterations) *iterations = it;
  });
}

int ed_stout_smear(ed_context* ctx, int nstep, double rho) {
  if (!ctx) return ED_ERR_ARG;
  return guard(ctx, [&] {
    ++ctx->gen;
    require(ctx->gauge_set, ED_ERR_STATE, "ed_set_gauge must be called first");
    if (nstep < 0) fail(ED_ERR_ARG, "nstep must be >= 0");
    if (!std::isfinite(rho)) fail(ED_ERR_ARG, "rho must be finite");
    HIPCHK(hipSetDevice(ctx->device));
    if (nstep == 0) return;
    const size_t nl = (size_t)3 * ctx->Lt * ctx->V3;
    DevBuf<double2> nxt;
    nxt.ensure(9 * nl);
    const ed::gauge::Lat L{ctx->Lx, ctx->Ly, ctx->Lz, ctx->Lt};
    const unsigned blocks = (unsigned)((nl + 255) / 256);
    for (int s = 0; s < nstep; ++s) {
      ed::gauge::stout_step_kernel<<<blocks, 256, 0, ctx->stream>>>(ctx->U64.p, nxt.p, L, rho);
      HIPCHK(hipGetLastError());
      std::swap(ctx->U64.p, nxt.p);
      std::swap(ctx->U64.n, nxt.n);
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
  });
}

int ed_laplacian(ed_context* ctx, int t, const void* F, void* out, int nvec, int on_device) {
  if (!ctx) return ED_ERR_ARG;
  return guard(ctx, [&] {
    require(ctx->gauge_set, ED_ERR_STATE, "ed_set_gauge must be called first");
    if (!F || !out) fail(ED_ERR_ARG, "F and out must be non-null");
    if (t < 0 || t >= ctx->Lt) fail(ED_ERR_ARG, "timeslice %d outside [0, %d)", t, ctx->Lt);
    if (nvec < 1) fail(ED_ERR_ARG, "nvec must be >= 1");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t n = (size_t)nvec * ctx->V3 * 3;
    const double2* src = (const double2*)F;
    double2* dst = (double2*)out;
    DevBuf<double2> hin, hout;
    if (!on_device) {
      hin.ensure(n);
      hout.ensure(n);
      HIPCHK(hipMemcpyAsync(hin.p, F, n * sizeof(double2), hipMemcpyHostToDevice, ctx->stream));
      src = hin.p;
      dst = hout.p;
    }
    const ed::gauge::Lat L{ctx->Lx, ctx->Ly, ctx->Lz, ctx->Lt};
    const size_t threads = (size_t)nvec * ctx->V3;
    ed::gauge::laplacian_kernel<<<(unsigned)((threads + 255) / 256), 256, 0, ctx->stream>>>(src, dst, ctx->U64.p, L, t,
                                                                                          nvec);
    HIPCHK(hipGetLastError());
    if (!on_device) {
      HIPCHK(hipMemcpyAsync(out, dst, n * sizeof(double2), hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(hipStreamSynchronize(ctx->stream));
    }
  });
}

int ed_get_gauge(ed_context* ctx, void* U, int on_device) {
  if (!ctx) return ED_ERR_ARG;
  return guard(ctx, [&] {
    require(ctx->gauge_set, ED_ERR_STATE, "ed_set_gauge must be called first");
    if (!U) fail(ED_ERR_ARG, "U is null");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t n = (size_t)3 * ctx->Lt * ctx->V3 * 9;
    HIPCHK(hipMemcpyAsync(U, ctx->U64.p, n * sizeof(double2), on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  });
}

int ed_batch_create(int device, ed_batch** b) {
  if (!b) return ED_ERR_ARG;
  *b = nullptr;
  ed_batch* x = new (std::nothrow) ed_batch();
  if (!x) return ED_ERR_ALLOC;
  int rc = bguard(x, [&] {
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) fail(ED_ERR_ARG, "device %d out of range (%d devices)", device, n);
    x->device = device;
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking));
    int cu = 0;
    HIPCHK(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device));
    x->ncu = std::max(8, cu);
  });
  if (rc != ED_OK) {
    delete x;
    return rc;
  }
  *b = x;
  return ED_OK;
}

int ed_batch_destroy(ed_batch* b) {
  if (!b) return ED_ERR_ARG;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  b->drop_graph();
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return ED_OK;
}

const char* ed_batch_last_error(const ed_batch* b) { return b ? b->err.c_str() : "null batch"; }

int ed_batch_add(ed_batch* b, ed_context* ctx, int mode, void* out_device) {
  if (!b) return ED_ERR_ARG;
  return bguard(b, [&] {
    if (!ctx || !out_device) fail(ED_ERR_ARG, "context and output must be non-null");
    if (mode != ED_MODE_FP64 && mode != ED_MODE_MIXED) fail(ED_ERR_ARG, "unknown mode %d", mode);
    if (ctx->device != b->device) fail(ED_ERR_ARG, "context is on device %d, batch on %d", ctx->device, b->device);
    for (auto& it : b->items)
      if (it.ctx == ctx) fail(ED_ERR_ARG, "a context can appear in a batch only once");
    b->items.push_back({ctx, mode, (double2*)out_device});
    b->drop_graph();
  });
}

int ed_batch_set_option(ed_batch* b, int option, int64_t value) {
  if (!b) return ED_ERR_ARG;
  return bguard(b, [&] {
    switch (option) {
      case ED_BOPT_GRID_PER_CU:
        if (value < 0 || value > 64) fail(ED_ERR_ARG, "workgroups per CU must be in [0, 64]");
        b->grid_per_cu = (int)value;
        break;
      case ED_BOPT_ORDER:
        if (value < 0 || value > 1) fail(ED_ERR_ARG, "item order must be 0 or 1");
        b->order = (int)value;
        break;
      case ED_BOPT_GRAPH:
        b->use_graph = value ? 1 : 0;
        break;
      case ED_BOPT_STRIP:
        if (value < 0 || value > 64) fail(ED_ERR_ARG, "strip height must be in [0, 64]");
        b->strip = (int)value;
        break;
      case ED_BOPT_H16:
        if (value < -1 || value > 1) fail(ED_ERR_ARG, "ED_BOPT_H16 must be -1, 0 or 1");
        b->h16 = (int)value;
        break;
      default:
        fail(ED_ERR_ARG, "unknown batch option %d", option);
    }
    b->drop_graph();
  });
}

int ed_batch_get_option(ed_batch* b, int option, int64_t* value) {
  if (!b) return ED_ERR_ARG;
  return bguard(b, [&] {
    if (!value) fail(ED_ERR_ARG, "value is null");
    switch (option) {
      case ED_BOPT_GRID_PER_CU: *value = b->grid_per_cu; break;
      case ED_BOPT_ORDER: *value = b->order; break;
      case ED_BOPT_GRAPH: *value = b->use_graph; break;
      case ED_BOPT_STRIP: *value = b->strip; break;
      case ED_BOPT_H16: *value = b->h16; break;
      default: fail(ED_ERR_ARG, "unknown batch option %d", option);
    }
  });
}

int ed_batch_run(ed_batch* b) {
  if (!b) return ED_ERR_ARG;
  return bguard(b, [&] { batch_run(b); });
}

int ed_batch_synchronize(ed_batch* b) {
  if (!b) return ED_ERR_ARG;
  return bguard(b, [&] {
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
  });
}

int ed_batch_time(ed_batch* b, int reps, double* ms_per_run) {
  if (!b) return ED_ERR_ARG;
  return bguard(b, [&] {
    if (reps < 1 || !ms_per_run) fail(ED_ERR_ARG, "reps must be >= 1 and ms_per_run non-null");
    *ms_per_run = batch_time(b, reps);
  });
}

int ed_batch_tune(ed_batch* b, int reps, double* best_ms) {
  if (!b) return ED_ERR_ARG;
  return bguard(b, [&] {
    if (reps < 1) fail(ED_ERR_ARG, "reps must be >= 1");
    const int grids[] = {1, 2, 4, 0};
    const int strips[] = {1, 2, 4, 8};
    double best = 1e300;
    int bg = b->grid_per_cu, bo = b->order, bs = b->strip, bh = b->h16;
    // FP16 split sum off, then on where it fits (only when some member can
    // take it; its kernel has the one-workgroup-per-item shape only)
    bool any_h16 = false;
    for (auto& it : b->items) {
      HIPCHK(hipSetDevice(it.ctx->device));
      build_plan(it.ctx);
      any_h16 = any_h16 || (it.mode != ED_MODE_FP64 && !mixed_runs_fp64(it.ctx) && h16_fits(it.ctx));
    }
    HIPCHK(hipSetDevice(b->device));
    for (int h : {0, 1}) {
      if (h && !any_h16) continue;
      for (int gsel : grids)
        for (int o = 0; o < 2; ++o)
          for (int st : strips) {
            b->h16 = h;
            b->grid_per_cu = gsel;
            b->order = o;
            b->strip = st;
            b->drop_graph();
            const double ms = batch_time(b, reps);
            if (ms < best) {
              best = ms;
              bg = gsel;
              bo = o;
              bs = st;
              bh = h;
            }
          }
    }
    b->grid_per_cu = bg;
    b->order = bo;
    b->strip = bs;
    b->h16 = bh;
    b->drop_graph();
    if (best_ms) *best_ms = best;
  });
}

}  // extern "C"
