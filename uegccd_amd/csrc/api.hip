// Host side of libuegccd: per-basis index maps, device buffers, the iteration
// driver and the C ABI declared in include/uegccd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <utility>
#include <new>
#include <tuple>
#include <vector>

#include "../../include/uegccd.h"
#include "uegccd_internal.h"

using namespace uegccd;

namespace {

// The tile table of the offline sweep (python -m uegccd_amd.tune writes
// uegccd_amd/tuned_tiles.json; the build turns it into this include): per
// class of systems, bounded by the largest pair-momentum block (pairs), the
// largest virtual list of a block and the precision, the tile choice that
// ran fastest.  Empty before the first sweep: every class takes the defaults.
struct TunedTiles {
  int noo_max, nvv_max, mixed;
  int ladder_group, quad_steps, hh_steps;
};
const TunedTiles kTunedTiles[] = {
#include "tuned_tiles.inc"
    {0, 0, -1, 2, 8, 8}  // end marker
};

// The tuned tiles of a system whose largest block has noo pairs and nvv
// virtual columns: the smallest class that covers it.
Tiles tuned_tiles(int noo, int nvv, bool mixed) {
  Tiles t;
  int best = -1;
  for (int k = 0; kTunedTiles[k].mixed >= 0; ++k) {
    const TunedTiles& e = kTunedTiles[k];
    if (e.mixed != (mixed ? 1 : 0) || noo > e.noo_max || nvv > e.nvv_max) continue;
    if (best < 0 || e.noo_max < kTunedTiles[best].noo_max ||
        (e.noo_max == kTunedTiles[best].noo_max && e.nvv_max < kTunedTiles[best].nvv_max))
      best = k;
  }
  if (best >= 0) {
    t.ladder_group = kTunedTiles[best].ladder_group;
    t.quad_steps = kTunedTiles[best].quad_steps;
    t.hh_steps = kTunedTiles[best].hh_steps;
  }
  return t;
}

constexpr double kPi = 3.14159265358979323846264338327950288;

#define HIPCHK(expr)                                                    \
  do {                                                                  \
    hipError_t e_ = (expr);                                             \
    if (e_ != hipSuccess) {                                             \
      std::fprintf(stderr, "uegccd: %s failed: %s\n", #expr,            \
                   hipGetErrorString(e_));                              \
      return UEGCCD_ERR_HIP;                                            \
    }                                                                   \
  } while (0)

// FindTol (source/HEG.f90:1073-1086).
void find_tol(double rs, double* tolmax, double* failratio, double* df) {
  *tolmax = 1.0e-8;
  *failratio = 100;
  *df = 2.0;
  if (rs > 5.1) *tolmax = 1.0e-6;
  if (rs > 5.1) *df = 5.0;
  if (rs > 9.9) *tolmax = 1.0e-5;
  if (rs > 80.0) *tolmax = 1.0e-3;
  if (rs < 0.95) *failratio = 1e6;
}

// Which ERI variant SetPointers would bind (source/Main.f90:274-318).
enum EriKind { ERI_CASED, ERI_NOCASE, ERI_CORE, ERI_GAP, ERI_BASE };
EriKind eri_kind(const uegccd_system& s) {
  int sum = 0;
  for (int k = 0; k < 15; ++k) sum += std::abs(s.irange[k]);
  if ((sum > 0 && s.gapii > 0.0 && s.coren > 0) || s.safe_eri) return ERI_CASED;
  if (s.gapii > 0.0 && s.coren > 0) return ERI_NOCASE;
  if (s.coren > 0) return ERI_CORE;
  if (s.gapii > 0.0) return ERI_GAP;
  return ERI_BASE;
}

// The range flag of each Coulomb role: the DummyFlag the reference passes at
// that call, a Max over the term's IRange keys (source/CCD.f90:216-217,
// 491-492, 555-556, 620-641, 666-672, 726-752, 782-788, 855-881, 910-916);
// only CasedERI reads it, every other variant computes flag 0.  irange[] is
// in Input-key order (source/IO.f90:88-102): Rings, XRings, Ladders,
// Mosaics, DriverDir, DriverEx, Energy, LinRings, QuadRings, DRings,
// ExRings, LinLadd, QuadLadd, DLadders, ExLadders.
void role_flags(const uegccd_system& s, EriKind kind, int* flag) {
  for (int r = 0; r < kRoles; ++r) flag[r] = 0;
  if (kind != ERI_CASED) return;
  const int* R = s.irange;
  enum { RING, XRING, LADDER, MOSAIC, DRVDIR, DRVEX, ENERGY, LINRING, QUADRING, DRING, EXRING,
         LINLADD, QUADLADD, DLADD, EXLADD };
  auto mx = [](int a, int b, int c) { return std::max(a, std::max(b, c)); };
  flag[RO_ENERGY] = R[ENERGY];
  flag[RO_DRVDIR] = R[DRVDIR];
  flag[RO_DRVEX] = R[DRVEX];
  flag[RO_MOSAIC] = R[MOSAIC];
  flag[RO_LAD_LD] = mx(R[LADDER], R[DLADD], R[LINLADD]);
  flag[RO_LAD_LX] = mx(R[LADDER], R[EXLADD], R[LINLADD]);
  flag[RO_LAD_QD] = mx(R[LADDER], R[DLADD], R[QUADLADD]);
  flag[RO_LAD_QX] = mx(R[LADDER], R[EXLADD], R[QUADLADD]);
  flag[RO_RING_LD] = mx(R[RING], R[DRING], R[LINRING]);
  flag[RO_RING_LX] = mx(R[RING], R[EXRING], R[LINRING]);
  flag[RO_RING_QD] = mx(R[RING], R[DRING], R[QUADRING]);
  flag[RO_RING_QX] = mx(R[RING], R[EXRING], R[QUADRING]);
  flag[RO_XR_LD] = mx(R[XRING], R[DRING], R[LINRING]);
  flag[RO_XR_LX] = mx(R[XRING], R[EXRING], R[LINRING]);
  flag[RO_XR_QD] = mx(R[XRING], R[DRING], R[QUADRING]);
  flag[RO_XR_QX] = mx(R[XRING], R[EXRING], R[QUADRING]);
}

// CoreN masks every integral with an index <= CoreN (source/HEG.f90:760-763,
// 819-820, 883-884) in CasedERI, NoCaseERI and CoreERI.
bool core_masked(const uegccd_system& s, EriKind kind) {
  return s.coren > 0 && (kind == ERI_CASED || kind == ERI_NOCASE || kind == ERI_CORE);
}

// Whether the roles read by k_slice (V slabs of the ring and crossed-ring
// linear products; v(s - k) of the mosaic and both quadratic directs; the Y
// tables of both quadratic exchanges) need more than one table per array.
bool slice_split_of(const int* rs) {
  return rs[RO_RING_LX] != rs[RO_XR_LX] || rs[RO_MOSAIC] != rs[RO_RING_QD] ||
         rs[RO_MOSAIC] != rs[RO_XR_QD] || rs[RO_RING_QX] != rs[RO_XR_QX];
}
// Whether the v(n_s - n_k) factors k_main reads (drivers, linear ring and
// crossed-ring directs, energy) come from more than one table.
bool main_split_of(const int* rs) {
  const int e = rs[RO_ENERGY];
  return rs[RO_DRVDIR] != e || rs[RO_DRVEX] != e || rs[RO_RING_LD] != e || rs[RO_XR_LD] != e;
}

// One entry of the Coulomb table of range flag `flag` at |dn|^2 = q
// (source/HEG.f90:727-757: flag 0 the full kernel, 1 its short-range Yukawa
// part, 2 the long-range rest, 3 zero), then the GapII screening of the
// variants that apply it (source/HEG.f90:770-772, 829-831, 938-940).
double coulomb(const uegccd_system& s, EriKind kind, int flag, int q) {
  const double L = s.L, mu2 = std::pow(s.screening, 2.0);
  double v;
  if (q == 0) {
    switch (flag) {
      case 0: v = -s.madelung; break;
      case 1: v = (1.0 / std::pow(L, 3.0)) * 4.0 * kPi / mu2; break;
      case 2: v = -s.madelung - (1.0 / std::pow(L, 3.0)) * 4.0 * kPi / mu2; break;
      default: v = 0.0; break;
    }
  } else {
    const double qs = (double)q * std::pow(2.0 * kPi / L, 2.0);
    switch (flag) {
      case 0: v = 1.0 / (L * kPi * (double)q); break;
      case 1: v = (1.0 / std::pow(L, 3.0)) * 4.0 * kPi / (mu2 + qs); break;
      case 2: v = (1.0 / std::pow(L, 3.0)) * 4.0 * kPi / (mu2 + qs) * mu2 / qs; break;
      default: v = 0.0; break;
    }
  }
  const bool gap = kind == ERI_CASED || kind == ERI_NOCASE || kind == ERI_GAP;
  if (gap && s.gapii > 0.0 && v > s.gapii) v = 0.0;
  return v;
}

struct DeviceArena {
  char* base = nullptr;
  size_t size = 0, used = 0;
  template <typename T>
  T* take(size_t count) {
    size_t bytes = (count * sizeof(T) + 255) & ~size_t(255);
    T* p = reinterpret_cast<T*>(base + used);
    used += bytes;
    return p;
  }
};

size_t aligned(size_t bytes) { return (bytes + 255) & ~size_t(255); }

}  // namespace

struct uegccd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  uegccd_options opt{};
  int o = 0, v = 0, nao = 0, n3 = 0, nK = 0, n2max = 0, maxiter = 1000;
  int64_t nvalid = 0;
  double tolmax = 1e-8, failratio = 100, df = 2.0;
  double boxL = 0.0;  // box length L
  // host copies of the maps
  std::vector<int> kb_poff, kb_pairs, kb_voff, pair_kb, pair_pos;
  int nva = 0;                     // entries of the blocks' virtual lists (built on the device)
  std::vector<uint64_t> vbits;     // bit av*nK + K: virtual a has a virtual partner in block K
  int box_w = 0, box_center = 0;   // momentum box of the device-built maps
  std::vector<int> kb_toff;        // K-blocked amplitude layout: block offsets
  std::vector<int> slab_coff;        // history layout: first compact index of each virtual slice
  std::vector<int4> nv4;
  std::vector<double> eig, vtab, voo, vso;  // Coulomb tables by slot: v(|dn|^2); v(n_k - n_j) over
                                            // occupied k, j; v(n_s - n_k)
  int ntab = 1;                    // table slots (one per range flag in use; slot 0 is flag 0)
  int rslot[kRoles] = {};          // slot of each Coulomb role
  int coren = 0;                   // core orbitals frozen by the CoreN masking (0: none)
  std::vector<int> core_elems;     // valid elements with a core index (held at zero)
  double screening = 0.0;          // HEGData%ScreeningDistance
  std::vector<int4> hh_items;   // (0, K, row0, 0) hole-hole units
  std::vector<int4> hh_items2;  // (0, K, row0, part) the same, split in two over the W columns of
                                // blocks of more than 16 pairs, parts 1 and 2 (a solve launched alone)
  std::vector<int4> lad_items;  // (0, K, row0, col0) particle-particle tiles of 16 columns
  std::vector<int4> lad_items2; // the same of 32 columns (two tiles per block, in batches)
  std::vector<int> qb_roff, qb_rows, qb_neg, qrow, qb_of;  // hole-particle sets P(q)
  std::vector<int4> qb_info;
  std::vector<int4> con_items;  // (0, kind, arg, arg) of k_contract
  // device
  DeviceArena arena;
  int* d_box = nullptr;    // scratch of the device-built maps (launch_plan)
  int* d_civ = nullptr;
  PlanArgs pa{};           // host copy of the planning arguments (batches collect them)
  SolveDev sd{};           // host copy of the device descriptor
  SolveDev* d_sd = nullptr;
  BatchDev batch{};        // batch of one
  State* d_state = nullptr;
  double* h_trace = nullptr;
  int last_niter = 0;
  // in-solve timing of one kernel (uegccd_time_kernel): -1 off
  int timed_kernel = -1;
  std::vector<hipEvent_t> tev;   // before/after pair of the timed pass of a chunk
  hipEvent_t span[2] = {nullptr, nullptr};  // brackets of a solve's iteration (created once)
  State* h_states = nullptr;      // pinned host copies of the polled states
  int h_states_cap = 0;
  double timed_us = 0.0;
  int64_t timed_launches = 0;
  bool owns_memory = false;  // arena from hipMalloc (else carved from a server pool)
  bool owns_stream = false;
  // the allocation-free solve: no events, no pinned buffers (polls into
  // pageable memory, host-clock timing)
  bool no_alloc = false;
  // captured chunks of passes of the context's own batch, by chunk length
  std::map<int, hipGraphExec_t> graphs;
  Tiles tiles;               // tile choice of the tuned kernels (tuned table or uegccd_set_tiles)
  int noo_max = 0, nvv_max = 0;  // largest pair-momentum block: pairs, virtual columns
  ElemCode ecode{};          // bit widths of the element map entries
  int vo2max = 0;            // largest |n_a - n_k|^2 of a virtual a and an occupied k
  // the structure factor's cells (transfer n_a - n_l, its P(q), and whether
  // its last writer in the reference's loop order is a masked core element),
  // in the reference's print order; built by the first structure-factor call
  struct SfCell {
    int x, y, z, q, core;
  };
  std::vector<SfCell> sf_cells;
  bool sf_ready = false;
};

namespace {

// The elements (i, j, a) that have a virtual momentum partner, in element
// order x = i + o j + o^2 (a - o), from the host's valid bits: f(x, core),
// core when i or j is a core orbital of the CoreN masking (those stay out of
// the valid set).  The device's element map cmap lists the others in the
// same order.
template <typename F>
void for_each_element(const uegccd_ctx* c, F f) {
  const int o = c->o, o2 = o * o, nK = c->nK, coren = c->coren;
  const uint64_t* bits = c->vbits.data();
  for (int av = 0; av < c->v; ++av) {
    const size_t row = (size_t)av * nK;
    for (int pr = 0; pr < o2; ++pr) {
      const size_t e = row + c->pair_kb[pr];
      if ((bits[e >> 6] >> (e & 63)) & 1ull) f(pr + o2 * av, pr % o < coren || pr / o < coren);
    }
  }
}
// The valid elements: f(x) for each, in element order (the order of cmap).
template <typename F>
void for_each_valid(const uegccd_ctx* c, F f) {
  for_each_element(c, [&f](int x, bool core) {
    if (!core) f(x);
  });
}

int build_maps(uegccd_ctx* c, const uegccd_system& s) {
  const int o = s.nocc, nao = s.nao, v = nao - o;
  // (NOcc is staged per slice in LDS; element indices are ints)
  if (o > kMaxOcc || (int64_t)o * o * v >= (int64_t(1) << 31)) return UEGCCD_ERR_UNSUPPORTED;
  c->o = o;
  c->v = v;
  c->nao = nao;
  c->boxL = s.L;
  c->n3 = o * o * v;
  c->nv4.resize(nao);
  int mx = 0;
  for (int p = 0; p < nao; ++p) {
    int4 q;
    q.x = s.nvec[p];
    q.y = s.nvec[nao + p];
    q.z = s.nvec[2 * nao + p];
    q.w = 0;
    c->nv4[p] = q;
    mx = std::max({mx, std::abs(q.x), std::abs(q.y), std::abs(q.z)});
  }
  int mo = 0;  // largest |component| of an occupied momentum
  for (int i = 0; i < o; ++i)
    mo = std::max({mo, std::abs(c->nv4[i].x), std::abs(c->nv4[i].y), std::abs(c->nv4[i].z)});
  // momentum -> orbital lookup over a box holding every n_i + n_j - n_p and
  // n_i + n_p - n_j (i, j occupied: half-width max|n| + 2 mo).  The box index is
  // affine in the momentum, so index(n_p + n_q - n_r) = center + lin[p] +
  // lin[q] - lin[r]: FindIndex is one add and one load.
  const int h = mx + 2 * mo + 1, w = 2 * h + 1;
  const int center = h + w * h + w * w * h;
  c->box_w = w;
  c->box_center = center;
  // (the temporaries of planning are kept per planning thread: a request
  // then touches no freshly mapped pages for them)
  thread_local std::vector<int> box, kid, civ, qid;
  box.assign((size_t)w * w * w, -1);
  std::vector<int> lin(nao);
  for (int p = 0; p < nao; ++p) {
    const int4 n = c->nv4[p];
    lin[p] = n.x + w * n.y + w * w * n.z;
    int& slot = box[center + lin[p]];
    if (slot >= 0) return UEGCCD_ERR_INVALID;  // duplicate plane wave
    slot = p;
  }
  // pair-momentum blocks, in order of first occurrence of K over x = i + o*j
  // (block ids looked up in a box over the sums of two occupied momenta, of
  // half-width 2 mo)
  const int wk = 4 * mo + 1, ck = 2 * mo * (1 + wk + wk * wk);
  kid.assign((size_t)wk * wk * wk, -1);
  thread_local std::vector<int> lk, kcnt, kfirst;
  lk.resize(o);
  for (int i = 0; i < o; ++i) lk[i] = c->nv4[i].x + wk * c->nv4[i].y + wk * wk * c->nv4[i].z;
  c->pair_kb.assign(o * o, 0);
  c->pair_pos.assign(o * o, 0);
  kcnt.clear();
  kfirst.clear();
  for (int x = 0; x < o * o; ++x) {
    int& K = kid[ck + lk[x % o] + lk[x / o]];
    if (K < 0) {
      K = (int)kcnt.size();
      kcnt.push_back(0);
      kfirst.push_back(x);
    }
    c->pair_kb[x] = K;
    c->pair_pos[x] = kcnt[K]++;
  }
  c->nK = (int)kcnt.size();
  c->kb_poff.assign(c->nK + 1, 0);
  for (int K = 0; K < c->nK; ++K) c->kb_poff[K + 1] = c->kb_poff[K] + kcnt[K];
  c->kb_pairs.assign(o * o, 0);
  for (int x = 0; x < o * o; ++x) c->kb_pairs[c->kb_poff[c->pair_kb[x]] + c->pair_pos[x]] = x;
  c->kb_voff.assign(c->nK + 1, 0);
  // The partner test of every loop in source/CCD.f90, FindIndex(i,j,a) > NOcc:
  // b = n_i + n_j - n_a depends on (i,j) only through K = n_i + n_j, so a
  // block's virtual list holds the a whose partner K - n_a is a virtual
  // orbital, and element (i,j,a) is valid exactly when a is on the list of
  // the block of (i,j).  The host keeps which (a, K) are valid (the bit set
  // vbits) and the counts the layout needs; the lists themselves and the
  // element map are built on the device at placement (launch_plan).
  const int nK = c->nK;
  civ.resize((size_t)nK * v);
  c->vbits.assign(((size_t)nK * v + 63) / 64, 0ull);
  std::vector<int> kbase(nK), nvk(nK, 0);
  for (int K = 0; K < nK; ++K) kbase[K] = center + lin[kfirst[K] % o] + lin[kfirst[K] / o];
  // With CoreN masking, the elements with a core index (i or j <= CoreN)
  // leave the valid set: every masked integral of the reference multiplies
  // such an amplitude or lands on such an element, so they stay exactly
  // zero and the others see the unmasked terms (the frozen core).
  const EriKind kind0 = eri_kind(s);
  const int coren = core_masked(s, kind0) ? std::min(s.coren, o) : 0;
  c->coren = coren;
  std::vector<int> nfree(nK, 0);  // pairs of each block with no core index
  for (int pr = 0; pr < o * o; ++pr)
    nfree[c->pair_kb[pr]] += (pr % o < coren || pr / o < coren) ? 0 : 1;
  // partners slice by slice, counted per block and (the valid elements of
  // each slice) per slice (two loops: each vectorises)
  c->slab_coff.assign((size_t)v + 1, 0);
  for (int av = 0; av < v; ++av) {
    int* crow = civ.data() + (size_t)av * nK;
    const int la = lin[o + av];
    for (int K = 0; K < nK; ++K) {
      const int b = box[kbase[K] - la];
      crow[K] = b >= o ? b : -1;
    }
    int nslab = 0;
    for (int K = 0; K < nK; ++K) {
      const int ok = crow[K] >= 0;
      nvk[K] += ok;
      nslab += ok * nfree[K];
    }
    c->slab_coff[av + 1] = c->slab_coff[av] + nslab;
  }
  {
    const size_t ne = (size_t)nK * v;
    uint64_t* const bits = c->vbits.data();
    for (size_t w0 = 0; w0 < ne; w0 += 64) {  // (branch-free: half the tests pass, at random)
      const size_t n = std::min<size_t>(64, ne - w0);
      uint64_t m = 0;
      for (size_t k = 0; k < n; ++k) m |= (uint64_t)(civ[w0 + k] >= 0) << k;
      bits[w0 >> 6] = m;
    }
  }
  for (int K = 0; K < nK; ++K) c->kb_voff[K + 1] = c->kb_voff[K] + nvk[K];
  c->nva = c->kb_voff[nK];
  c->noo_max = c->nvv_max = 0;
  for (int K = 0; K < c->nK; ++K) {
    c->noo_max = std::max(c->noo_max, c->kb_poff[K + 1] - c->kb_poff[K]);
    c->nvv_max = std::max(c->nvv_max, c->kb_voff[K + 1] - c->kb_voff[K]);
  }
  // K-blocked layout of the valid amplitudes: block K holds, for each entry ci
  // of its virtual list, its noo pairs contiguously (position toff + ci*noo + p).
  // The valid elements in element order (their compact index there is the
  // layout of the DIIS history) start each virtual slice at slab_coff.
  c->kb_toff.assign(c->nK + 1, 0);
  for (int K = 0; K < c->nK; ++K)
    c->kb_toff[K + 1] = c->kb_toff[K] + (c->kb_voff[K + 1] - c->kb_voff[K]) * (c->kb_poff[K + 1] - c->kb_poff[K]);
  c->nvalid = c->slab_coff[v];
  // element map entries: element, position in tk and partner in 64 bits
  c->ecode.xb = bits_for(c->n3);
  c->ecode.pb = bits_for(c->kb_toff[c->nK]);
  if (c->ecode.xb + c->ecode.pb + bits_for(nao) > 64) return UEGCCD_ERR_UNSUPPORTED;
  // the valid elements with a core index (held at zero), for the input checks
  c->core_elems.clear();
  if (coren)
    for_each_element(c, [c](int x, bool core) {
      if (core) c->core_elems.push_back(x);
    });
  // ladder work: per 16 pair rows of a block, one hole-hole unit (its W rows
  // and the hole-hole products of those rows) and the particle-particle
  // tiles of 16 a columns; the largest blocks' units first.  A solve launched
  // alone takes the split units (two, over the halves of the W columns, in a
  // block of more than 16 pairs: the largest units held its contraction
  // launch twice as long as a particle-particle tile); a batch, which fills
  // the chip, the whole ones.
  c->hh_items.clear();
  c->hh_items2.clear();
  c->lad_items.clear();
  c->lad_items2.clear();
  for (int K = 0; K < c->nK; ++K) {
    const int noo = c->kb_poff[K + 1] - c->kb_poff[K];
    const int nvv = c->kb_voff[K + 1] - c->kb_voff[K];
    if (nvv == 0) continue;
    for (int r0 = 0; r0 < noo; r0 += 16) {
      c->hh_items.push_back(make_int4(0, K, r0, 0));
      if (noo > 16) {
        c->hh_items2.push_back(make_int4(0, K, r0, 1));
        c->hh_items2.push_back(make_int4(0, K, r0, 2));
      } else {
        c->hh_items2.push_back(make_int4(0, K, r0, 0));
      }
      for (int c0 = 0; c0 < nvv; c0 += 16) c->lad_items.push_back(make_int4(0, K, r0, c0));
      for (int c0 = 0; c0 < nvv; c0 += 32) c->lad_items2.push_back(make_int4(0, K, r0, c0));
    }
  }
  auto by_work = [c](const int4& a, const int4& b) {
    auto work = [c](int K) {
      const int64_t noo = c->kb_poff[K + 1] - c->kb_poff[K], nvv = c->kb_voff[K + 1] - c->kb_voff[K];
      return noo * nvv;
    };
    return work(a.y) > work(b.y);
  };
  std::stable_sort(c->hh_items.begin(), c->hh_items.end(), by_work);
  std::stable_sort(c->hh_items2.begin(), c->hh_items2.end(), by_work);
  // Coulomb table, flag 0 of the ERI variants (source/HEG.f90:729-772), up to
  // the largest |n_p - n_q|^2 of the basis.  |n_p - n_q| <= |n_p| + |n_q|, so
  // with the plane waves in decreasing |n| the pairs of p can raise the
  // maximum only while (|n_p| + |n_q|)^2 exceeds it; the scan stops at the
  // first that cannot (d2 is an integer: a bound below best + 1 is enough,
  // which also covers the rounding of the square roots).
  int n2max = 0;
  {
    thread_local std::vector<int> byn, cnt;
    thread_local std::vector<double> rn;
    int n2top = 0;
    for (int p = 0; p < nao; ++p) {
      const int4 a = c->nv4[p];
      n2top = std::max(n2top, a.x * a.x + a.y * a.y + a.z * a.z);
    }
    cnt.assign(n2top + 2, 0);  // counting sort by |n|^2, largest first
    for (int p = 0; p < nao; ++p) {
      const int4 a = c->nv4[p];
      ++cnt[n2top - (a.x * a.x + a.y * a.y + a.z * a.z) + 1];
    }
    for (int k = 1; k <= n2top + 1; ++k) cnt[k] += cnt[k - 1];
    byn.assign(nao, 0);
    rn.assign(nao, 0.0);
    for (int p = 0; p < nao; ++p) {
      const int4 a = c->nv4[p];
      const int e = cnt[n2top - (a.x * a.x + a.y * a.y + a.z * a.z)]++;
      byn[e] = p;
      rn[e] = std::sqrt((double)(a.x * a.x + a.y * a.y + a.z * a.z));
    }
    for (int u = 0; u < nao; ++u) {
      if ((2.0 * rn[u]) * (2.0 * rn[u]) < n2max + 1.0 - 1e-6) break;
      const int4 a = c->nv4[byn[u]];
      for (int t = 0; t < nao; ++t) {
        if ((rn[u] + rn[t]) * (rn[u] + rn[t]) < n2max + 1.0 - 1e-6) break;
        const int4 b = c->nv4[byn[t]];
        const int dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
        n2max = std::max(n2max, dx * dx + dy * dy + dz * dz);
      }
    }
  }
  // hole-particle pair sets P(q) = {(i,p): n_p - n_i = q}, in order of first
  // occurrence over (p, i), for the quadratic ring products
  {
    // (a counting sort: transfer ids and counts in one pass, rows in another;
    // transfer ids looked up in a box of half-width max|n| + mo)
    const int hq = mx + mo, wq = 2 * hq + 1, cq = hq * (1 + wq + wq * wq);
    qid.assign((size_t)wq * wq * wq, -1);   // block of transfer q by box index
    thread_local std::vector<int> qlin, qcount, lq;
    lq.resize(nao);
    for (int p = 0; p < nao; ++p) lq[p] = c->nv4[p].x + wq * c->nv4[p].y + wq * wq * c->nv4[p].z;
    qlin.clear();
    qcount.clear();
    c->qb_of.assign((size_t)o * v, -1);
    for (int pv = 0; pv < v; ++pv)
      for (int i = 0; i < o; ++i) {
        const int ql = lq[o + pv] - lq[i];
        int& q = qid[cq + ql];
        if (q < 0) {
          q = (int)qlin.size();
          qlin.push_back(ql);
          qcount.push_back(0);
        }
        ++qcount[q];
        c->qb_of[(size_t)pv * o + i] = q;
      }
    const int nq = (int)qlin.size();
    c->qb_roff.assign(nq + 1, 0);
    for (int q = 0; q < nq; ++q) c->qb_roff[q + 1] = c->qb_roff[q] + qcount[q];
    c->qb_rows.assign((size_t)o * v, 0);
    c->qrow.assign((size_t)o * v, -1);
    {
      std::vector<int> fill(c->qb_roff.begin(), c->qb_roff.end() - 1);
      for (int pv = 0; pv < v; ++pv)
        for (int i = 0; i < o; ++i) {
          const int q = c->qb_of[(size_t)pv * o + i], r = fill[q]++;
          c->qrow[(size_t)pv * o + i] = r;
          c->qb_rows[r] = i + o * pv;
        }
    }
    // (the partner slices d_q(l) of each transfer, qb_d, are looked up on the
    // device at placement)
    c->qb_neg.assign(nq, -1);
    for (int q = 0; q < nq; ++q) c->qb_neg[q] = qid[cq - qlin[q]];
    c->qb_info.assign(nq, make_int4(0, 0, 0, 0));
    for (int q = 0; q < nq; ++q) {
      const int qn = c->qb_neg[q];
      c->qb_info[q] = make_int4(c->qb_roff[q], c->qb_roff[q + 1] - c->qb_roff[q],
                                qn >= 0 ? c->qb_roff[qn] : 0,
                                qn >= 0 ? c->qb_roff[qn + 1] - c->qb_roff[qn] : 0);
    }
    // wave items of the contraction pass: the Joo sums over slice ranges
    // (first: they read the most data per wave), 16x16 tiles of the
    // per-transfer ring products
    c->con_items.clear();
    {
      size_t n = s.do_mosaic ? kJooParts : 0;
      if (s.do_ring || s.do_xring)
        for (int q = 0; q < nq; ++q)
          n += (size_t)((c->qb_info[q].y + 15) / 16) * ((c->qb_info[q].w + 15) / 16) *
               ((s.do_ring ? 1 : 0) + (s.do_xring ? 1 : 0));
      c->con_items.reserve(n);
    }
    if (s.do_mosaic)
      for (int p = 0; p < kJooParts; ++p) c->con_items.push_back(make_int4(0, CK_JOO, p, 0));
    if (s.do_ring || s.do_xring)
      for (int q = 0; q < nq; ++q) {
        const int4 in = c->qb_info[q];
        const int nt = ((in.y + 15) / 16) * ((in.w + 15) / 16);
        for (int t = 0; t < nt; ++t) {
          if (s.do_ring) c->con_items.push_back(make_int4(0, CK_RING, q, t));
          if (s.do_xring) c->con_items.push_back(make_int4(0, CK_XRING, q, t));
        }
      }
  }
  c->n2max = n2max;
  // Coulomb tables: slot 0 holds flag 0, each other range flag a role uses
  // gets a slot of its own (usually there is none)
  const EriKind kind = eri_kind(s);
  int flag[kRoles];
  role_flags(s, kind, flag);
  int slot_of_flag[kMaxTabSlots] = {0, -1, -1, -1};
  int flags_of_slot[kMaxTabSlots] = {0, 0, 0, 0};
  c->ntab = 1;
  for (int r = 0; r < kRoles; ++r) {
    const int f = flag[r];
    if (f < 0 || f > 3) return UEGCCD_ERR_INVALID;
    if (slot_of_flag[f] < 0) {
      slot_of_flag[f] = c->ntab;
      flags_of_slot[c->ntab++] = f;
    }
    c->rslot[r] = slot_of_flag[f];
  }
  for (int t = 0; t < c->ntab; ++t)
    if ((flags_of_slot[t] == 1 || flags_of_slot[t] == 2) && !(s.screening > 0.0)) return UEGCCD_ERR_INVALID;
  const size_t nt = (size_t)n2max + 1;
  c->vtab.assign(c->ntab * nt, 0.0);
  for (int t = 0; t < c->ntab; ++t)
    for (int q = 0; q <= n2max; ++q) c->vtab[t * nt + q] = coulomb(s, kind, flags_of_slot[t], q);
  // A negative integral is the checked variants' -100 sentinel, on which the
  // reference skips the excitation ("Disallowed excitation not trapped!").
  // Only the roles whose transfer can vanish (an occupied or a virtual pair
  // with itself) ever meet v(0); refuse a negative one there.
  if (kind != ERI_BASE) {
    const int zero_roles[] = {RO_LAD_LD, RO_LAD_LX, RO_RING_LX, RO_XR_LX};
    const int on[] = {s.do_ladder, s.do_ladder, s.do_ring, s.do_xring};
    for (int k = 0; k < 4; ++k)
      if (on[k] && c->vtab[c->rslot[zero_roles[k]] * nt] < 0.0) return UEGCCD_ERR_UNSUPPORTED;
  }
  // the Coulomb elements every slice block needs, from the same tables:
  // V(k, j) = v(n_k - n_j) over occupied k, j (k fastest) and v(n_s - n_k)
  auto d2 = [&](int p, int q) {
    const int4 a = c->nv4[p], b = c->nv4[q];
    const int x = a.x - b.x, y = a.y - b.y, z = a.z - b.z;
    return x * x + y * y + z * z;
  };
  c->voo.assign((size_t)c->ntab * o * o, 0.0);
  c->vso.assign((size_t)c->ntab * v * o, 0.0);
  for (int t = 0; t < c->ntab; ++t) {
    const double* vt = c->vtab.data() + t * nt;
    double* voo = c->voo.data() + (size_t)t * o * o;
    double* vso = c->vso.data() + (size_t)t * v * o;
    for (int j = 0; j < o; ++j)
      for (int k = 0; k < o; ++k) voo[k + (size_t)o * j] = vt[d2(k, j)];
    for (int sv = 0; sv < v; ++sv)
      for (int k = 0; k < o; ++k) vso[(size_t)sv * o + k] = vt[d2(o + sv, k)];
  }
  c->vo2max = 0;
  for (int sv = 0; sv < v; ++sv)
    for (int k = 0; k < o; ++k) c->vo2max = std::max(c->vo2max, d2(o + sv, k));
  // LDS of the slice blocks (one table; two with split slice roles)
  const bool split = slice_split_of(c->rslot);
  if (slice_lds_bytes(o, split) + (int)sizeof(int4) * kMaxOcc +
      (int)sizeof(double) * ((split ? 2 : 1) * kMaxN2 + 16 + kMaxDiis * kMaxDiis + kMaxDiis) +
      slice_fin_lds(slice_threads(o)) > kMaxDynLds)
    return UEGCCD_ERR_UNSUPPORTED;
  c->eig.assign(s.eigen, s.eigen + nao);
  return UEGCCD_OK;
}

// Element-wise blocks of a context: one thread per valid element (at least
// one block, which writes the partials of an empty sum).
int elem_blocks_of(const uegccd_ctx* c) {
  return (int)std::max<int64_t>(1, (c->nvalid + kElemPerBlock - 1) / kElemPerBlock);
}

// Doubles of jring / jxring: (J2, J3) pairs per (row of P(q), l); the
// structure-factor sums reuse them for 3 doubles per hole-particle pair and
// per transfer.
static size_t jring_doubles(const uegccd_ctx* c) {
  const size_t ovo = (size_t)c->o * c->v * c->o, ov3 = 3 * (size_t)c->o * c->v;
  return 2 * ((std::max<size_t>(2 * ovo, ov3) + 1) / 2 + 1);
}

// Device bytes of a context: the arena ctx_create carves its buffers from.
// Blocks of a k_contract range, padded to a multiple of the 8 XCDs.
static int pad8(int n) { return (n + 7) & ~7; }
// Partials width k_main needs for a solve: 4 + its DIIS dots, 16 or 32.
static int main_nq_of(const SolveDev& S) {
  return (S.mode == UEGCCD_MODE_DIIS && S.ndiis > 12) ? 32 : 16;
}

// Scratch of the device-built maps: the momentum box and the partner table.
static size_t plan_scratch_ints(const uegccd_ctx* c) {
  return (size_t)c->box_w * c->box_w * c->box_w + (size_t)c->v * c->nK;
}

size_t arena_bytes(const uegccd_ctx* c, int ndiis) {
  const size_t n3 = c->n3;
  const size_t eb = elem_blocks_of(c);
  size_t bytes = 0;
  bytes += aligned(sizeof(int4) * c->nao) + aligned(sizeof(double) * c->nao);
  bytes += aligned(sizeof(double) * c->vtab.size());                    // vtab
  bytes += aligned(sizeof(double) * c->voo.size()) + aligned(sizeof(double) * c->vso.size());
  bytes += 3 * aligned(sizeof(int) * (c->nK + 1)) + aligned(sizeof(int) * c->kb_pairs.size());
  bytes += 2 * aligned(sizeof(int) * std::max<size_t>(1, c->nva));   // kb_va, kb_vb
  bytes += aligned(sizeof(int4) * std::max<size_t>(1, c->nva));      // kb_nva
  bytes += 2 * aligned(sizeof(int) * c->o * c->o);
  bytes += aligned(sizeof(double2) * n3) * 3;                         // tcur, gout, rout
  bytes += 2 * aligned(sizeof(double2) * std::max<size_t>(1, (size_t)c->nvalid * ndiis));  // histT, histE
  bytes += aligned(sizeof(int) * (c->v + 1));                          // slab_coff
  bytes += aligned(sizeof(uint2) * std::max<size_t>(1, c->nvalid));      // cmap
  bytes += aligned(sizeof(int) * plan_scratch_ints(c));                // box, civ (launch_plan)
  bytes += aligned(sizeof(double2) * std::max<size_t>(1, c->kb_toff.back()));  // tk
  bytes += aligned(sizeof(int) * (c->nK + 1));                         // kb_toff
  bytes += aligned(sizeof(double2) * n3);                              // lad
  bytes += 2 * aligned(sizeof(double2) * std::max(1, c->kb_toff.back()));  // hh, hh2
  bytes += 2 * aligned(sizeof(double) * jring_doubles(c));             // jring, jxring
  bytes += aligned(sizeof(double2) * n3);                              // tv
  bytes += 2 * aligned(sizeof(double) * c->o * c->v);                  // rsum, csum
  bytes += aligned(sizeof(double) * c->v) + aligned(sizeof(double) * c->v * c->o);
  bytes += aligned(sizeof(double) * c->o * kJooParts);
  bytes += aligned(sizeof(double) * eb * kPartialStride);
  bytes += 2 * aligned(sizeof(State)) + aligned(sizeof(SolveDev));
  bytes += aligned(sizeof(double) * 2 * c->maxiter);
  bytes += aligned(sizeof(int) * (8 + 2 * kSliceClasses)) +
           aligned(sizeof(int4) * std::max<size_t>(1, c->lad_items.size())) +
           aligned(sizeof(int4) * std::max<size_t>(1, c->lad_items2.size())) + aligned(sizeof(int) * 2) +
           aligned(sizeof(int4) * std::max<size_t>(1, c->hh_items.size())) +
           aligned(sizeof(int4) * std::max<size_t>(1, c->hh_items2.size()));
  bytes += aligned(sizeof(int) * c->qb_roff.size()) + aligned(sizeof(int) * std::max<size_t>(1, c->qb_rows.size()));
  bytes += aligned(sizeof(int) * std::max<size_t>(1, c->qb_neg.size())) +
           aligned(sizeof(int) * std::max<size_t>(1, c->qb_neg.size() * (size_t)c->o));  // qb_d
  bytes += aligned(sizeof(int4) * std::max<size_t>(1, c->con_items.size()));
  bytes += aligned(sizeof(int4) * std::max<size_t>(1, c->qb_info.size()));
  bytes += 2 * aligned(sizeof(int) * std::max<size_t>(1, c->qrow.size()));  // qrow, qb_of
  bytes += aligned(sizeof(double4) * n3);                              // quad
  bytes += 4096;
  return bytes;
}

// Input values the reference accepts (range flags 0-3; CoreN at least 0);
// which integrals they select is worked out in build_maps.
int check_supported(const uegccd_system& s) {
  if (s.coren < 0) return UEGCCD_ERR_INVALID;
  for (int k = 0; k < 15; ++k)
    if (s.irange[k] < 0 || s.irange[k] > 3) return UEGCCD_ERR_INVALID;
  return UEGCCD_OK;
}

// Host image of an arena's constant section: each host array is written at
// the offset its device copy will have, so one copy uploads them all.
struct ArenaImage {
  DeviceArena* a;
  std::vector<char>* img;
  bool copy = true;  // false: the image already holds the arrays; put() only places them
  template <typename T>
  T* reserve(size_t count) {
    T* d = a->take<T>(std::max<size_t>(1, count));
    const size_t end = (size_t)(reinterpret_cast<char*>(d) - a->base) +
                       sizeof(T) * std::max<size_t>(1, count);
    if (img->size() < end) img->resize(end, 0);
    return d;
  }
  template <typename T>
  void write(T* d, const T* h, size_t count) {
    if (count) std::memcpy(img->data() + (reinterpret_cast<char*>(d) - a->base), h, sizeof(T) * count);
  }
  template <typename T>
  const T* put(const std::vector<T>& h) {
    T* d = reserve<T>(h.size());
    if (copy) write(d, h.data(), h.size());
    return d;
  }
};

// Host (i,j,a) arrays of each spin -> interleaved device amplitudes; entries
// with no virtual momentum partner are zeroed.  Under CoreN masking the
// elements with a core index must be zero on input (DrvMBPT makes them so):
// the library keeps them at zero, which is the reference's iteration only
// from such amplitudes.
// The caller's dense amplitude arrays into tcur: both copied as they are into
// the gout work buffer (2 n3 doubles, free outside residual calls), then each
// valid element's pair placed into the cleared tcur on the device (the
// entries of invalid elements are ignored, as the reference's loops skip
// them).
int copy_in_amplitudes(uegccd_ctx* c, const double* tab, const double* tba) {
  for (int x : c->core_elems)
    if ((tab && tab[x] != 0.0) || (tba && tba[x] != 0.0)) return UEGCCD_ERR_INVALID;
  const size_t n3 = c->n3;
  double* const sab = c->sd.gout;
  double* const sba = c->sd.gout + n3;
  HIPCHK(hipMemsetAsync(c->sd.tcur, 0, sizeof(double2) * n3, c->stream));
  if (tab) HIPCHK(hipMemcpyAsync(sab, tab, sizeof(double) * n3, hipMemcpyHostToDevice, c->stream));
  if (tba) HIPCHK(hipMemcpyAsync(sba, tba, sizeof(double) * n3, hipMemcpyHostToDevice, c->stream));
  launch_place_amplitudes(c->batch, c->nvalid, tab ? sab : nullptr, tba ? sba : nullptr, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return UEGCCD_OK;
}

// tcur into the caller's dense arrays (the valid elements' amplitudes, zeros
// elsewhere: tcur holds zeros at invalid elements), unpacked on the device
// into the gout and rout work buffers and copied from there.
int copy_out_amplitudes(uegccd_ctx* c, double* taa, double* tab, double* tba) {
  const size_t n3 = c->n3;
  double* const sab = c->sd.gout;
  double* const sba = c->sd.gout + n3;
  double* const saa = c->sd.rout;
  launch_split_amplitudes(c->batch, (int64_t)n3, sab, sba, saa, c->stream);
  HIPCHK(hipGetLastError());
  if (tab) HIPCHK(hipMemcpyAsync(tab, sab, sizeof(double) * n3, hipMemcpyDeviceToHost, c->stream));
  if (tba) HIPCHK(hipMemcpyAsync(tba, sba, sizeof(double) * n3, hipMemcpyDeviceToHost, c->stream));
  if (taa) HIPCHK(hipMemcpyAsync(taa, saa, sizeof(double) * n3, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return UEGCCD_OK;
}

// Entries of a context's K-block virtual lists a contraction block holds in
// LDS at once: the largest list, up to kClsChunk (longer ones go in parts).
int cls_entries(const uegccd_ctx* c) { return std::max(1, std::min(c->nvv_max, kClsChunk)); }

// Whether the CCD pass kernels take a context's basis: the slice blocks' LDS
// Coulomb table must hold every virtual-occupied difference of momenta, and a
// contraction block's LDS the ladder tables, a part of a pair-momentum
// block's virtual list and its W rows.  (dRPA, MP2, the energy and the
// structure factor run on any basis a context holds.)
int ccd_check(const uegccd_ctx* c) {
  if (c->vo2max > kMaxN2 - 2) return UEGCCD_ERR_UNSUPPORTED;
  const int lds = contract_lds_bytes(ladder_tab_doubles(c->sd), cls_entries(c), c->noo_max);
  return lds <= contract_lds_limit() ? UEGCCD_OK : UEGCCD_ERR_UNSUPPORTED;
}

int read_state(uegccd_ctx* c, State* st) {
  HIPCHK(hipMemcpyAsync(st, c->d_state, sizeof(State), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return UEGCCD_OK;
}

}  // namespace

extern "C" {

void uegccd_default_options(uegccd_options* opt) {
  std::memset(opt, 0, sizeof(*opt));
  opt->mode = UEGCCD_MODE_REFERENCE;
  opt->ndiis = 8;
  opt->diis_start = 3;
  opt->max_iter = 1000;
  opt->tol_dt = 0.0;
  opt->tol_e = 1e-10;
  opt->tol_r = 1e-8;
  opt->damping = 0.0;
  opt->precision = UEGCCD_PREC_FP64;
  opt->mixed_switch = 1e-5;
  opt->device = 0;
  opt->iters_per_sync = 8;
  opt->graph = 1;
}

const char* uegccd_strerror(int st) {
  switch (st) {
    case UEGCCD_OK: return "ok";
    case UEGCCD_ERR_INVALID: return "invalid argument or inconsistent system";
    case UEGCCD_ERR_ALLOC: return "allocation failed";
    case UEGCCD_ERR_HIP: return "HIP runtime error or no usable device";
    case UEGCCD_ERR_UNSUPPORTED: return "Input option not supported by this build";
    case UEGCCD_ERR_NOT_CONVERGED: return "CCD equations did not converge";
    case UEGCCD_ERR_NONFINITE: return "non-finite amplitudes or residual";
    case UEGCCD_ERR_BUSY: return "no free slot or HBM budget";
    case UEGCCD_ERR_NOT_FOUND: return "unknown request";
    case UEGCCD_ERR_SCRATCH:
      return "system larger than the allocation-free solve's device scratch (UEGCCD_NOALLOC_SCRATCH_MB)";
    default: return "unknown status";
  }
}

}  // extern "C"

namespace {

// Checks a system and options and builds the host maps of a context that is
// not yet on a device.
int plan_ctx(const uegccd_system* sys, const uegccd_options* opt_in, uegccd_ctx** out) {
  if (!sys || !out || !sys->nvec || !sys->eigen) return UEGCCD_ERR_INVALID;
  *out = nullptr;
  if (sys->nocc < 1 || sys->nao <= sys->nocc || !(sys->L > 0.0)) return UEGCCD_ERR_INVALID;
  int rc = check_supported(*sys);
  if (rc) return rc;
  uegccd_options opt;
  if (opt_in) opt = *opt_in; else uegccd_default_options(&opt);
  if (opt.mode != UEGCCD_MODE_REFERENCE && opt.mode != UEGCCD_MODE_DIIS) return UEGCCD_ERR_INVALID;
  if (opt.mode == UEGCCD_MODE_DIIS && (opt.ndiis < 2 || opt.ndiis > kMaxDiis)) return UEGCCD_ERR_INVALID;
  if (opt.precision != UEGCCD_PREC_FP64 && opt.precision != UEGCCD_PREC_MIXED) return UEGCCD_ERR_INVALID;
  uegccd_ctx* c = new (std::nothrow) uegccd_ctx();
  if (!c) return UEGCCD_ERR_ALLOC;
  c->device = opt.device;
  c->opt = opt;
  rc = build_maps(c, *sys);
  if (rc) { delete c; return rc; }
  find_tol(sys->rs, &c->tolmax, &c->failratio, &c->df);
  if (opt.tol_dt > 0.0) c->tolmax = opt.tol_dt;
  // DIIS mode steps undamped unless told otherwise: the extrapolation supplies
  // the stability the damping gives the plain iteration (measured over r_s
  // 0.5-50, N 14-54: fewer iterations in nearly every case, none diverging)
  if (opt.damping > 0.0) c->df = opt.damping;
  else if (opt.mode == UEGCCD_MODE_DIIS) c->df = 1.0;
  c->maxiter = opt.max_iter > 0 ? opt.max_iter : 1000;
  c->tiles = tuned_tiles(c->noo_max, c->nvv_max, opt.precision == UEGCCD_PREC_MIXED);
  *out = c;
  return UEGCCD_OK;
}

int ndiis_of(const uegccd_options& opt) { return opt.mode == UEGCCD_MODE_DIIS ? opt.ndiis : 1; }

// An upper bound of the bytes of a context's constant section (each array of
// layout_ctx's image, rounded up to the arena's 256-byte alignment).
size_t const_section_bound(const uegccd_ctx* c) {
  size_t b = 0;
  auto add = [&b](size_t bytes) { b += (std::max<size_t>(bytes, 1) + 255) & ~size_t(255); };
  add(sizeof(int4) * c->nv4.size());
  add(sizeof(double) * c->eig.size());
  add(sizeof(double) * (c->vtab.size() + c->voo.size() + c->vso.size()));
  add(sizeof(int) * (c->kb_poff.size() + c->kb_pairs.size() + c->kb_voff.size() + c->kb_toff.size()));
  add(sizeof(int) * (c->slab_coff.size() + c->pair_kb.size() + c->pair_pos.size()));
  add(sizeof(int) * (c->qb_roff.size() + c->qb_rows.size() + c->qb_neg.size() + c->qrow.size() +
                     c->qb_of.size()));
  add(sizeof(int4) * (c->qb_info.size() + c->hh_items.size() + c->hh_items2.size() + c->lad_items.size() +
                      c->con_items.size()));
  add(sizeof(int4) * c->lad_items2.size() + sizeof(int) * 2);
  add(sizeof(SolveDev) + sizeof(int) * (8 + 2 * kSliceClasses));
  return b + 24 * 256;  // (one alignment pad per array)
}

// Lays a planned context out in `bytes` (= arena_bytes) of device memory at
// `base`: the maps, the device descriptor and the batch-of-one tables form a
// constant section written into `image` (uploaded by the caller in one copy),
// the work buffers follow it from byte *work_off on (cleared by the caller).
// Only the descriptor depends on `base`: with fill false, `image` already
// holds the section from an earlier layout (at any base) and only the
// descriptor and the prefixes are written again.
// The kernel arguments of a folded update: solve S's partials and state.
FoldArgs fold_args(const SolveDev& S) {
  FoldArgs f;
  f.partials = S.partials;
  f.nblk = S.nblk_elem;
  f.nd = S.mode == UEGCCD_MODE_DIIS ? S.ndiis : 0;
  f.st = S.st;
  return f;
}

void layout_ctx(uegccd_ctx* c, const uegccd_system* sys, char* base, size_t bytes,
                std::vector<char>& image, size_t* work_off, bool fill) {
  const uegccd_options& opt = c->opt;
  const int ndiis = ndiis_of(opt);
  const size_t n3 = c->n3;
  const int eb = elem_blocks_of(c);
  c->arena.base = base;
  c->arena.size = bytes;
  c->arena.used = 0;
  if (fill) {
    image.clear();
    image.reserve(const_section_bound(c));  // one allocation for the constant section
  }
  ArenaImage im{&c->arena, &image, fill};

  SolveDev& S = c->sd;
  S = SolveDev{};
  S.o = c->o; S.v = c->v; S.nao = c->nao; S.nK = c->nK; S.n3 = c->n3; S.n2max = c->n2max;
  S.ecode = c->ecode;
  S.ndiis = ndiis;
  S.mode = opt.mode;
  S.prec0 = opt.precision == UEGCCD_PREC_MIXED ? 0 : 1;
  S.maxiter = c->maxiter;
  S.diis_start = std::max(1, opt.diis_start);
  S.diis_err_is_dt = 1;
  S.do_ring = sys->do_ring; S.do_xring = sys->do_xring;
  S.do_ladder = sys->do_ladder; S.do_mosaic = sys->do_mosaic;
  S.df = c->df;
  S.tol_dt = c->tolmax;
  S.tol_e = opt.tol_e;
  S.tol_r = opt.tol_r;
  S.mixed_switch = opt.precision == UEGCCD_PREC_MIXED ? opt.mixed_switch : 0.0;
  S.ntab = c->ntab;
  for (int r = 0; r < kRoles; ++r) S.rslot[r] = c->rslot[r];
  S.slice_split = slice_split_of(c->rslot) ? 1 : 0;
  S.main_split = main_split_of(c->rslot) ? 1 : 0;
  {
    // the ladder tiles stage the distinct tables of their four roles
    const int roles[4] = {RO_LAD_LD, RO_LAD_LX, RO_LAD_QD, RO_LAD_QX};
    S.lad_ntab = 0;
    for (int k = 0; k < 4; ++k) {
      const int sl = c->rslot[roles[k]];
      int idx = -1;
      for (int m = 0; m < S.lad_ntab; ++m)
        if (S.lad_slot[m] == sl) idx = m;
      if (idx < 0) {
        idx = S.lad_ntab++;
        S.lad_slot[idx] = sl;
      }
      S.lad_tab[k] = idx;
    }
    for (int k = S.lad_ntab; k < 4; ++k) S.lad_slot[k] = 0;
  }
  S.nv4 = im.put(c->nv4);
  S.eig = im.put(c->eig);
  S.vtab = im.put(c->vtab);
  S.voo = im.put(c->voo);
  S.vso = im.put(c->vso);
  S.kb_poff = im.put(c->kb_poff);
  S.kb_pairs = im.put(c->kb_pairs);
  S.kb_voff = im.put(c->kb_voff);
  S.kb_toff = im.put(c->kb_toff);
  S.slab_coff = im.put(c->slab_coff);
  S.nvalid = (int)c->nvalid;
  S.pair_kb = im.put(c->pair_kb);
  S.pair_pos = im.put(c->pair_pos);
  S.nqb = (int)c->qb_neg.size();
  S.qb_roff = im.put(c->qb_roff);
  S.qb_rows = im.put(c->qb_rows);
  S.qb_neg = im.put(c->qb_neg);
  S.qb_info = im.put(c->qb_info);
  S.qrow = im.put(c->qrow);
  S.qb_of = im.put(c->qb_of);
  c->d_sd = im.reserve<SolveDev>(1);
  // the solve's ladder tiles and contraction items (any batch reads them here)
  S.nhh = sys->do_ladder ? (int)c->hh_items.size() : 0;
  S.hh_items = im.put(c->hh_items);
  S.nhh2 = sys->do_ladder ? (int)c->hh_items2.size() : 0;
  S.hh_items2 = im.put(c->hh_items2);
  S.nlad = sys->do_ladder ? (int)c->lad_items.size() : 0;
  S.lad_items = im.put(c->lad_items);
  S.nlad2 = sys->do_ladder ? (int)c->lad_items2.size() : 0;
  S.lad_items2 = im.put(c->lad_items2);
  S.ncon = (int)c->con_items.size();
  S.con_items = im.put(c->con_items);
  // batch of one: block prefixes of the element, slice, ladder and contraction launches
  BatchDev& b = c->batch;
  // (every pass needs its slice blocks: they build the pass's amplitudes
  // from the history, whichever terms are on)
  const int sb = c->v;
  const int cls = slice_threads(c->o) / 256 - 1;
  int h_pref[8 + 2 * kSliceClasses] = {0, eb, 0, S.nlad, 0, S.ncon, 0, S.nhh2};
  h_pref[8 + 2 * cls + 1] = sb;
  int* d_pref = im.reserve<int>(8 + 2 * kSliceClasses);
  im.write(d_pref, h_pref, 8 + 2 * kSliceClasses);
  b.n = 1;
  b.solves = c->d_sd;
  b.elem_prefix = d_pref;
  b.lad_prefix = d_pref + 2;
  b.con_prefix = d_pref + 4;
  b.hh_prefix = d_pref + 6;
  b.drpa_sum_blocks = kDrpaJooParts * c->o + c->v + (int)c->qb_neg.size();
  b.elem_blocks = eb;
  b.slice_blocks = sb;
  b.slice_split = S.slice_split;
  for (int k = 0; k < kSliceClasses; ++k) {
    b.slice_prefix[k] = d_pref + 8 + 2 * k;
    b.slice_cls_blocks[k] = k == cls ? sb : 0;
    b.slice_lds[k] = k == cls ? slice_lds_bytes(c->o, S.slice_split) : 0;
  }
  b.nhh = S.nhh2;
  b.nhh_pad = pad8(S.nhh2);
  b.hh_split = 1;
  b.nlad = S.nlad;
  b.nlad_pad = pad8(S.nlad);
  {
    int* d_l2 = im.reserve<int>(2);
    const int h_l2[2] = {0, S.nlad2};
    im.write(d_l2, h_l2, 2);
    b.lad_prefix2 = d_l2;
    b.nlad2 = S.nlad2;
  }
  b.tab_doubles = ladder_tab_doubles(S);
  b.cls_max = cls_entries(c);
  b.con_lds = contract_lds_bytes(b.tab_doubles, b.cls_max, c->noo_max);
  b.tiles = c->tiles;
  b.main_nq = main_nq_of(S);
  // the planning launches of the batch of one (their arguments are written
  // once the scratch is placed, below)
  b.plan = im.reserve<PlanArgs>(1);
  {
    int* d_ppref = im.reserve<int>(6);
    const int hpp[6] = {0, plan_box_blocks(c->nao), 0, plan_list_blocks(c->nK, (int)c->qb_neg.size(), c->o), 0,
                        c->v};
    im.write(d_ppref, hpp, 6);
    for (int k = 0; k < 3; ++k) {
      b.plan_prefix[k] = d_ppref + 2 * k;
      b.plan_blocks[k] = hpp[2 * k + 1];
    }
  }
  b.ncon = S.ncon;
  // work buffers (cleared at placement), the device-built maps first
  c->arena.used = aligned(c->arena.used);
  *work_off = c->arena.used;
  S.kb_va = c->arena.take<int>(std::max(1, c->nva));
  S.kb_vb = c->arena.take<int>(std::max(1, c->nva));
  S.kb_nva = c->arena.take<int4>(std::max(1, c->nva));
  S.cmap = c->arena.take<uint2>(std::max<int64_t>(1, c->nvalid));
  S.qb_d = c->arena.take<int>(std::max<size_t>(1, (size_t)S.nqb * c->o));
  c->d_box = c->arena.take<int>(plan_scratch_ints(c));
  c->d_civ = c->d_box + (size_t)c->box_w * c->box_w * c->box_w;
  S.tcur = c->arena.take<double2>(n3);
  S.tk = c->arena.take<double2>(std::max<int64_t>(1, c->kb_toff[c->nK]));
  S.histT = c->arena.take<double2>(std::max<size_t>(1, (size_t)c->nvalid * ndiis));
  S.histE = c->arena.take<double2>(std::max<size_t>(1, (size_t)c->nvalid * ndiis));
  S.gout = c->arena.take<double>(2 * n3);
  S.rout = c->arena.take<double>(2 * n3);
  S.lad = c->arena.take<double2>(n3);
  S.hh = c->arena.take<double2>(std::max(1, c->kb_toff[c->nK]));
  S.hh2 = c->arena.take<double2>(std::max(1, c->kb_toff[c->nK]));
  S.jring = c->arena.take<double2>(jring_doubles(c) / 2);
  S.jxring = c->arena.take<double2>(jring_doubles(c) / 2);
  S.tv = c->arena.take<double2>(n3);
  S.quad = c->arena.take<double4>(n3);
  S.rsum = c->arena.take<double>((size_t)c->o * c->v);
  S.csum = c->arena.take<double>((size_t)c->o * c->v);
  S.jvv = c->arena.take<double>(c->v);
  S.joo_part = c->arena.take<double>((size_t)c->v * c->o);
  S.joo = c->arena.take<double>((size_t)c->o * kJooParts);
  S.nblk_elem = eb;
  S.partials = c->arena.take<double>((size_t)eb * kPartialStride /* slot-major: [q][block] */);
  c->d_state = c->arena.take<State>(1);
  S.st = c->d_state;
  S.st2 = c->arena.take<State>(1);
  S.trace = c->arena.take<double>(2 * (size_t)c->maxiter);
  b.fold = fold_args(S);
  im.write(c->d_sd, &S, 1);
  c->pa.sd = c->d_sd;
  c->pa.w = c->box_w;
  c->pa.center = c->box_center;
  c->pa.coren = c->coren;
  c->pa.box = c->d_box;
  c->pa.civ = c->d_civ;
  im.write(const_cast<PlanArgs*>(b.plan), &c->pa, 1);
}

// The device-built maps of a placed context (its work buffers cleared).
void launch_plan_maps(const uegccd_ctx* c, hipStream_t s) { launch_plan(c->batch, s); }

// Puts a planned context on its device in memory of its own: one allocation,
// one copy of the constant section, one clear of the work buffers.  Deletes c
// on failure.
int place_ctx(uegccd_ctx* c, const uegccd_system* sys, uegccd_ctx** out) {
  *out = nullptr;
  const uegccd_options opt = c->opt;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) { delete c; return UEGCCD_ERR_HIP; }
  if (opt.device < 0 || opt.device >= ndev) { delete c; return UEGCCD_ERR_INVALID; }
  if (hipSetDevice(opt.device) != hipSuccess) { delete c; return UEGCCD_ERR_HIP; }
  const size_t bytes = arena_bytes(c, ndiis_of(opt));
  char* base = nullptr;
  if (hipMalloc(&base, bytes) != hipSuccess) { delete c; return UEGCCD_ERR_ALLOC; }
  c->owns_memory = true;
  std::vector<char> image;
  size_t work_off = 0;
  layout_ctx(c, sys, base, bytes, image, &work_off, true);
  if (c->arena.used > c->arena.size) { uegccd_ctx_destroy(c); return UEGCCD_ERR_ALLOC; }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    uegccd_ctx_destroy(c);
    return UEGCCD_ERR_HIP;
  }
  c->owns_stream = true;
  if (hipMemcpyAsync(base, image.data(), image.size(), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipMemsetAsync(base + work_off, 0, bytes - work_off, c->stream) != hipSuccess) {
    uegccd_ctx_destroy(c);
    return UEGCCD_ERR_HIP;
  }
  launch_plan_maps(c, c->stream);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) {
    uegccd_ctx_destroy(c);
    return UEGCCD_ERR_HIP;
  }
  *out = c;
  return UEGCCD_OK;
}

}  // namespace

extern "C" {

int uegccd_ctx_create(const uegccd_system* sys, const uegccd_options* opt_in, uegccd_ctx** out) {
  if (!out) return UEGCCD_ERR_INVALID;
  *out = nullptr;
  uegccd_ctx* c = nullptr;
  int rc = plan_ctx(sys, opt_in, &c);
  if (rc) return rc;
  return place_ctx(c, sys, out);
}

int uegccd_footprint(const uegccd_system* sys, const uegccd_options* opt, int64_t* bytes) {
  if (!bytes) return UEGCCD_ERR_INVALID;
  uegccd_ctx* c = nullptr;
  int rc = plan_ctx(sys, opt, &c);
  if (rc) return rc;
  *bytes = (int64_t)arena_bytes(c, ndiis_of(c->opt));
  delete c;
  return UEGCCD_OK;
}

void uegccd_ctx_destroy(uegccd_ctx* c) {
  if (!c) return;
  hipSetDevice(c->device);
  for (auto& g : c->graphs) hipGraphExecDestroy(g.second);
  for (hipEvent_t e : c->tev) hipEventDestroy(e);
  for (hipEvent_t e : c->span)
    if (e) hipEventDestroy(e);
  if (c->h_states) hipHostFree(c->h_states);
  if (c->stream && c->owns_stream) {
    hipStreamSynchronize(c->stream);
    hipStreamDestroy(c->stream);
  }
  if (c->arena.base && c->owns_memory) hipFree(c->arena.base);
  delete c;
}

int uegccd_ctx_info(const uegccd_ctx* c, int32_t* nocc, int32_t* nao, int64_t* nvalid,
                    int32_t* nblocks) {
  if (!c) return UEGCCD_ERR_INVALID;
  if (nocc) *nocc = c->o;
  if (nao) *nao = c->nao;
  if (nvalid) *nvalid = c->nvalid;
  if (nblocks) *nblocks = c->nK;
  return UEGCCD_OK;
}

int uegccd_residual(uegccd_ctx* c, const double* tab, const double* tba, double* gab,
                    double* gba, double* rab, double* rba) {
  if (!c || !tab || !tba) return UEGCCD_ERR_INVALID;
  if (int rc = ccd_check(c)) return rc;
  HIPCHK(hipSetDevice(c->device));
  int rc = copy_in_amplitudes(c, tab, tba);
  if (rc) return rc;
  launch_state_reset(c->batch, PH_ITER, 1, c->stream);
  launch_init(c->batch, c->stream);  // (also lays the amplitudes out by K block)
  launch_iteration(c->batch, MF_WRITE_GR, c->stream);
  HIPCHK(hipGetLastError());
  const size_t n3 = c->n3;
  std::vector<double> h(4 * n3);
  HIPCHK(hipMemcpyAsync(h.data(), c->sd.gout, sizeof(double) * 2 * n3, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(h.data() + 2 * n3, c->sd.rout, sizeof(double) * 2 * n3, hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  // the valid elements' values; zeros elsewhere
  for (double* out : {gab, gba, rab, rba})
    if (out) std::fill(out, out + n3, 0.0);
  for_each_valid(c, [&](int x) {
    if (gab) gab[x] = h[x];
    if (gba) gba[x] = h[n3 + x];
    if (rab) rab[x] = h[2 * n3 + x];
    if (rba) rba[x] = h[3 * n3 + x];
  });
  return UEGCCD_OK;
}

int uegccd_energy(uegccd_ctx* c, const double* tab, const double* tba, double* ecorr) {
  if (!c || !tab || !tba || !ecorr) return UEGCCD_ERR_INVALID;
  HIPCHK(hipSetDevice(c->device));
  int rc = copy_in_amplitudes(c, tab, tba);
  if (rc) return rc;
  launch_state_reset(c->batch, PH_ITER, 1, c->stream);
  launch_init(c->batch, c->stream);
  launch_finalize(c->batch, FIN_SUM, c->stream);
  HIPCHK(hipGetLastError());
  State st;
  rc = read_state(c, &st);
  if (rc) return rc;
  *ecorr = st.ecorr;
  return st.nonfinite ? UEGCCD_ERR_NONFINITE : UEGCCD_OK;
}

}  // extern "C"

namespace {

// The structure factor of the amplitudes in tcur (from_mp2: first written as
// DrvMBPT's guess) on the cells the reference prints: with qvec or vals NULL
// a size query, else nval values per cell (SfG, SfSum, SfV[, SfDSum,
// SfXSum]) with SfV from table slot vslot.  Shared by CCEnergyAndSF (unit
// 80) and DrvMBPT's binning (unit 81).
int structure_factor_cells(uegccd_ctx* c, const double* tab, const double* tba, bool from_mp2, int vslot,
                           int nval, int32_t* ncell, int32_t* qvec, double* vals, double* energy) {
  if (!c || !ncell) return UEGCCD_ERR_INVALID;
  const int o = c->o, nq = (int)c->qb_neg.size();
  if (!c->sf_ready) {
    // the cells printed: transfers n_a - n_i or n_a - n_j of a valid element
    // (those with a core index too: the reference bins them, with zero
    // amplitudes and, under CoreN masking, a zero integral).  SfV of a cell
    // is the integral of its last writer in the reference's loop order (i,
    // then j, then a; direct before exchange): v(q), or 0 when that writer
    // is a masked core element.
    std::vector<char> touched(nq, 0);
    std::vector<int64_t> last(nq, -1);
    std::vector<char> last_core(nq, 0);
    auto bin = [&](int x, bool core) {
      const int i = x % o, j = (x / o) % o, av = x / (o * o);
      const int64_t key = (((int64_t)i * o + j) * c->v + av) * 2;
      const int qd = c->qb_of[(size_t)av * o + i], qx = c->qb_of[(size_t)av * o + j];
      touched[qd] = touched[qx] = 1;
      if (key > last[qd]) { last[qd] = key; last_core[qd] = core; }
      if (key + 1 > last[qx]) { last[qx] = key + 1; last_core[qx] = core; }
    };
    for_each_valid(c, [&](int x) { bin(x, false); });
    for (int x : c->core_elems) bin(x, true);
    auto& cells = c->sf_cells;
    cells.clear();
    for (int q = 0; q < nq; ++q) {
      if (!touched[q]) continue;
      const int r = c->qb_rows[c->qb_roff[q]], l = r % o, dv = r / o;
      const int4 na = c->nv4[o + dv], nl = c->nv4[l];
      cells.push_back({na.x - nl.x, na.y - nl.y, na.z - nl.z, q, last_core[q]});
    }
    std::sort(cells.begin(), cells.end(), [](const uegccd_ctx::SfCell& a, const uegccd_ctx::SfCell& b) {
      return a.x != b.x ? a.x < b.x : a.y != b.y ? a.y < b.y : a.z < b.z;
    });
    c->sf_ready = true;
  }
  const auto& cells = c->sf_cells;
  const int n = (int)cells.size();
  if (!qvec || !vals) {  // size query
    *ncell = n;
    return UEGCCD_OK;
  }
  if (*ncell < n || (!from_mp2 && (!tab || !tba))) return UEGCCD_ERR_INVALID;
  HIPCHK(hipSetDevice(c->device));
  int rc = UEGCCD_OK;
  launch_state_reset(c->batch, PH_ITER, 1, c->stream);
  if (from_mp2) {
    launch_mp2(c->batch, c->stream);
  } else {
    rc = copy_in_amplitudes(c, tab, tba);
    if (rc) return rc;
    launch_init(c->batch, c->stream);
  }
  launch_finalize(c->batch, FIN_SUM, c->stream);
  launch_sf(c->batch, c->v, nq, c->stream);
  HIPCHK(hipGetLastError());
  std::vector<double> sums(3 * (size_t)std::max(1, nq));
  HIPCHK(hipMemcpyAsync(sums.data(), c->sd.jxring, sizeof(double) * 3 * nq, hipMemcpyDeviceToHost, c->stream));
  State st;
  rc = read_state(c, &st);
  if (rc) return rc;
  const double sc = std::pow(2.0 * M_PI / c->boxL, 2.0);  // dK2scaled, source/HEG.f90:1025
  const double* vso = c->vso.data() + (size_t)vslot * c->v * o;
  for (int k = 0; k < n; ++k) {
    const auto& e = cells[k];
    const int r = c->qb_rows[c->qb_roff[e.q]], l = r % o, dv = r / o;
    qvec[3 * k] = e.x;
    qvec[3 * k + 1] = e.y;
    qvec[3 * k + 2] = e.z;
    double* out = vals + (size_t)nval * k;
    out[0] = std::sqrt((double)(e.x * e.x + e.y * e.y + e.z * e.z) * sc);
    out[1] = sums[3 * (size_t)e.q];
    out[2] = e.core ? 0.0 : vso[(size_t)dv * o + l];  // v(n_a - n_l) = v(q)
    if (nval > 3) {
      out[3] = sums[3 * (size_t)e.q + 1];
      out[4] = sums[3 * (size_t)e.q + 2];
    }
  }
  *ncell = n;
  if (energy) *energy = st.ecorr;
  return st.nonfinite ? UEGCCD_ERR_NONFINITE : UEGCCD_OK;
}

}  // namespace

extern "C" {

int uegccd_structure_factor(uegccd_ctx* c, const double* tab, const double* tba, int32_t* ncell,
                            int32_t* qvec, double* vals, double* ecorr) {
  // SfV from the energy's table (IRangeEnergy, source/CCD.f90:290-291)
  if (!c) return UEGCCD_ERR_INVALID;
  return structure_factor_cells(c, tab, tba, false, c->rslot[RO_ENERGY], 5, ncell, qvec, vals, ecorr);
}

int uegccd_mp2_structure_factor(uegccd_ctx* c, int32_t* ncell, int32_t* qvec, double* vals, double* e2) {
  // DrvMBPT's ERI calls pass no flag (source/MP2.f90:61-62): SfV from flag 0
  return structure_factor_cells(c, nullptr, nullptr, true, 0, 3, ncell, qvec, vals, e2);
}

int uegccd_mp2(uegccd_ctx* c, double* taa, double* tab, double* tba, double* e2) {
  if (!c) return UEGCCD_ERR_INVALID;
  HIPCHK(hipSetDevice(c->device));
  launch_state_reset(c->batch, PH_ITER, 1, c->stream);
  launch_mp2(c->batch, c->stream);
  launch_finalize(c->batch, FIN_SUM, c->stream);
  HIPCHK(hipGetLastError());
  State st;
  int rc = read_state(c, &st);
  if (!rc) rc = copy_out_amplitudes(c, taa, tab, tba);
  if (rc) return rc;
  if (e2) *e2 = st.ecorr;
  return UEGCCD_OK;
}


int uegccd_trace(const uegccd_ctx* c, double* e, double* dt, int32_t n) {
  if (!c) return UEGCCD_ERR_INVALID;
  const int m = std::min<int>(n, std::min(c->last_niter, c->maxiter));
  if (m <= 0) return UEGCCD_OK;
  std::vector<double> h(2 * (size_t)m);
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpy(h.data(), c->sd.trace, sizeof(double) * 2 * m, hipMemcpyDeviceToHost));
  for (int k = 0; k < m; ++k) {
    if (e) e[k] = h[2 * k];
    if (dt) dt[k] = h[2 * k + 1];
  }
  return UEGCCD_OK;
}

int uegccd_solve(uegccd_ctx* c, double* taa, double* tab, double* tba, uegccd_result* res) {
  uegccd_ctx* cs[1] = {c};
  double* pab[1] = {tab};
  double* pba[1] = {tba};
  if (!c || !tab || !tba) return UEGCCD_ERR_INVALID;
  int rc = uegccd_solve_batch(cs, 1, pab, pba, res);
  if (taa && (rc == UEGCCD_OK || rc == UEGCCD_ERR_NOT_CONVERGED))
    for (int x = 0; x < c->n3; ++x) taa[x] = tab[x] + tba[x];
  return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Solves of one or more contexts, iterated together: every launch of a pass
// covers all systems' tiles.  Each system keeps its own state, convergence
// test and history; a finished system's blocks return at once.
namespace {

constexpr int kMaxChunk = 32;  // most passes launched between two host polls

struct BatchScratch {
  char* base = nullptr;
  size_t cap = 0;
  BatchDev b{};
  State* states = nullptr;  // [n] gathered states, one host copy per poll
};

// Pinned host staging for uploads that are queued on a stream while earlier
// work still runs: chunks are kept and handed out again after each reset
// (callers reset only once every copy from them has completed).
struct PinnedStage {
  std::vector<std::pair<char*, size_t>> chunks;
  size_t cur = 0, off = 0, used = 0;
  void reset() { cur = 0; off = 0; used = 0; }
  char* get(size_t bytes) {
    bytes = aligned(bytes);
    used += bytes;
    while (cur < chunks.size()) {
      if (off + bytes <= chunks[cur].second) {
        char* p = chunks[cur].first + off;
        off += bytes;
        return p;
      }
      ++cur;
      off = 0;
    }
    const size_t sz = std::max(bytes, size_t(32) << 20);
    char* p = nullptr;
    if (hipHostMalloc(reinterpret_cast<void**>(&p), sz, hipHostMallocDefault) != hipSuccess)
      return nullptr;
    chunks.push_back({p, sz});
    cur = chunks.size() - 1;
    off = bytes;
    return p;
  }
  void release() {
    for (auto& c : chunks) hipHostFree(c.first);
    chunks.clear();
    reset();
  }
};

// Copies a host image to the device: queued on `stream` through pinned
// staging when `stage` is given, else a blocking copy.
int upload_image(char* dst, const std::vector<char>& image, hipStream_t stream, PinnedStage* stage) {
  if (image.empty()) return UEGCCD_OK;
  if (stage) {
    char* p = stage->get(image.size());
    if (!p) return UEGCCD_ERR_ALLOC;
    std::memcpy(p, image.data(), image.size());
    HIPCHK(hipMemcpyAsync(dst, p, image.size(), hipMemcpyHostToDevice, stream));
  } else {
    HIPCHK(hipMemcpy(dst, image.data(), image.size(), hipMemcpyHostToDevice));
  }
  return UEGCCD_OK;
}

// Device tables of a batch: descriptors, block and item prefixes (each solve's
// ladder tiles and contraction items stay in its own memory) and a state
// gather buffer, in one (reused) allocation uploaded in one copy.
int make_batch(uegccd_ctx** cs, int n, BatchScratch* out, hipStream_t stream = nullptr,
               PinnedStage* stage = nullptr) {
  std::vector<int> ep(n + 1, 0), lp(n + 1, 0), lp2(n + 1, 0), cp(n + 1, 0), hp(n + 1, 0);
  std::vector<int> sp[kSliceClasses];
  int slds[kSliceClasses] = {};
  for (auto& v : sp) v.assign(n + 1, 0);
  // one k_slice instantiation serves the launch: the split one if any solve needs it
  int slice_split = 0, tab_doubles = 0, cls_max = 1, noo_max = 0;
  for (int k = 0; k < n; ++k) slice_split |= cs[k]->sd.slice_split;
  for (int k = 0; k < n; ++k) {
    uegccd_ctx* c = cs[k];
    ep[k + 1] = ep[k] + elem_blocks_of(c);
    const int cls = slice_threads(c->o) / 256 - 1;
    for (int q = 0; q < kSliceClasses; ++q) sp[q][k + 1] = sp[q][k] + (q == cls ? c->v : 0);
    slds[cls] = std::max(slds[cls], slice_lds_bytes(c->o, slice_split));
    lp[k + 1] = lp[k] + c->sd.nlad;
    lp2[k + 1] = lp2[k] + c->sd.nlad2;
    hp[k + 1] = hp[k] + c->sd.nhh;
    tab_doubles = std::max(tab_doubles, ladder_tab_doubles(c->sd));
    cls_max = std::max(cls_max, cls_entries(c));
    noo_max = std::max(noo_max, c->noo_max);
    cp[k + 1] = cp[k] + c->sd.ncon;
  }
  // the planning launches of the solves (used when a batch is planned at once)
  std::vector<PlanArgs> pas(n);
  std::vector<int> pp[3];
  for (auto& v : pp) v.assign(n + 1, 0);
  for (int k = 0; k < n; ++k) {
    const uegccd_ctx* c = cs[k];
    pas[k] = c->pa;
    pp[0][k + 1] = pp[0][k] + plan_box_blocks(c->nao);
    pp[1][k + 1] = pp[1][k] + plan_list_blocks(c->nK, (int)c->qb_neg.size(), c->o);
    pp[2][k + 1] = pp[2][k] + c->v;
  }
  size_t bytes = aligned(sizeof(SolveDev) * n) + (5 + kSliceClasses) * aligned(sizeof(int) * (n + 1)) +
                 aligned(sizeof(State) * n) + aligned(sizeof(PlanArgs) * n) + 3 * aligned(sizeof(int) * (n + 1));
  if (out->cap < bytes) {
    if (out->base) hipFree(out->base);
    out->base = nullptr;
    out->cap = 0;
    if (hipMalloc(&out->base, bytes) != hipSuccess) return UEGCCD_ERR_ALLOC;
    out->cap = bytes;
  }
  DeviceArena a;
  a.base = out->base;
  a.size = bytes;
  std::vector<char> image;
  ArenaImage im{&a, &image};
  BatchDev& b = out->b;
  b.n = n;
  b.solves = im.reserve<SolveDev>(n);
  for (int k = 0; k < n; ++k) im.write(b.solves + k, &cs[k]->sd, 1);
  b.elem_prefix = const_cast<int*>(im.put(ep));
  b.slice_blocks = 0;
  for (int q = 0; q < kSliceClasses; ++q) {
    b.slice_prefix[q] = const_cast<int*>(im.put(sp[q]));
    b.slice_cls_blocks[q] = sp[q][n];
    b.slice_lds[q] = slds[q];
    b.slice_blocks += sp[q][n];
  }
  b.lad_prefix = const_cast<int*>(im.put(lp));
  b.lad_prefix2 = const_cast<int*>(im.put(lp2));
  b.hh_prefix = const_cast<int*>(im.put(hp));
  b.nhh = hp[n];
  b.nhh_pad = pad8(hp[n]);
  b.hh_split = 0;
  b.fold = n == 1 ? fold_args(cs[0]->sd) : FoldArgs{};
  b.nlad_pad = pad8(lp[n]);
  b.tab_doubles = tab_doubles;
  b.cls_max = cls_max;
  b.con_lds = contract_lds_bytes(tab_doubles, cls_max, noo_max);
  // one instantiation per launch: the tiles of the solve with the most ladder tiles
  {
    int big = 0;
    for (int k = 1; k < n; ++k)
      if (cs[k]->sd.nlad > cs[big]->sd.nlad) big = k;
    b.tiles = cs[big]->tiles;
  }
  b.slice_split = slice_split;
  b.main_nq = 16;
  for (int k = 0; k < n; ++k) b.main_nq = std::max(b.main_nq, main_nq_of(cs[k]->sd));
  b.con_prefix = const_cast<int*>(im.put(cp));
  b.plan = im.put(pas);
  for (int q = 0; q < 3; ++q) {
    b.plan_prefix[q] = const_cast<int*>(im.put(pp[q]));
    b.plan_blocks[q] = pp[q][n];
  }
  b.ncon = cp[n];
  out->states = a.take<State>(n);
  b.nlad = lp[n];
  b.nlad2 = lp2[n];
  b.elem_blocks = ep[n];
  return upload_image(out->base, image, stream, stage);
}

// Passes to launch before the next host poll.  Each unfinished solve's
// convergence measure (max|R| against tol_r in DIIS mode, max|dT| against
// TolMax in the reference iteration) decays about geometrically; its rate
// over the last chunk predicts the passes still needed (+1 for the stopping
// pass, +1 for the final residual pass).  The batch waits for its slowest
// member.  Passes launched after a solve is done return at once, so a wrong
// guess costs either one more poll or a few empty passes, never a result.
int next_chunk(uegccd_ctx** cs, int n, const std::vector<State>& st, std::vector<double>& prev,
               int last, int first, int cap) {
  int want = 1;
  for (int k = 0; k < n; ++k) {
    const State& x = st[k];
    if (x.phase == PH_DONE) continue;
    if (x.phase == PH_FINAL) { want = std::max(want, 1); continue; }
    const bool diis = cs[k]->sd.mode == UEGCCD_MODE_DIIS;
    const double r = diis ? x.maxres_in : x.dt;
    const double tol = diis ? cs[k]->sd.tol_r : cs[k]->sd.tol_dt;
    const double r0 = prev[k];
    prev[k] = r;
    if (!(r > 0.0) || !(r0 > r) || !std::isfinite(r)) return first;  // no decay seen yet
    const double rate = std::pow(r / r0, 1.0 / std::max(1, last));
    int need = 2;
    if (r >= tol && rate < 1.0) need = (int)std::ceil(std::log(tol / r) / std::log(rate)) + 2;
    want = std::max(want, need);
  }
  return std::min(std::max(want, 1), cap);
}

// The chunk of `len` passes of the context's own batch as an instantiated HIP
// graph, captured on first use and kept (the kernels' arguments are the
// context's fixed device tables, so a replay is the same launches).  The
// passes hold no memset or copy nodes.
// Where a pass's reductions and state update run: folded into the next
// pass's k_slice, where every slice block reduces its solve's element-block
// partials for itself (about 1 us of a C2 pass, against 5 us for a launch of
// its own), for a single solve; in a finalize launch after every pass for a
// batch, where the slice blocks of its large solves would re-read hundreds
// of KB of partials each (a finalize launch per pass: batched stream 1,278
// -> 1,365 solves/s).

bool fold_state_update(int nsolves, const uegccd_ctx* c0) {
  // (the slice blocks reduce partials of DIIS histories up to 12 entries)
  return nsolves == 1 && !(c0->opt.mode == UEGCCD_MODE_DIIS && c0->opt.ndiis > 12);
}

int chunk_graph(uegccd_ctx* c, int len, hipGraphExec_t* out) {
  auto it = c->graphs.find(len);
  if (it != c->graphs.end()) {
    *out = it->second;
    return UEGCCD_OK;
  }
  prepare_launches();  // per-kernel attributes are set outside the capture
  hipStream_t s = c->stream;
  HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const bool fold = fold_state_update(1, c);
  for (int k = 0; k < len; ++k) {
    launch_iteration(c->batch, MF_ITERATE, s, nullptr, nullptr, fold && k > 0);
    if (!fold) launch_finalize(c->batch, FIN_ITER, s);
  }
  if (fold) launch_finalize(c->batch, FIN_ITER, s);
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(s, &g);
  if (ec != hipSuccess || !g) {
    std::fprintf(stderr, "uegccd: graph capture failed: %s\n", hipGetErrorString(ec));
    if (g) hipGraphDestroy(g);
    return UEGCCD_ERR_HIP;
  }
  hipGraphExec_t ex = nullptr;
  const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (ei != hipSuccess) {
    std::fprintf(stderr, "uegccd: graph instantiation failed: %s\n", hipGetErrorString(ei));
    return UEGCCD_ERR_HIP;
  }
  c->graphs[len] = ex;
  *out = ex;
  return UEGCCD_OK;
}

// Runs the iteration of a batch already holding its starting amplitudes in
// tcur: init (E, dT of the guess), then passes until every system is done.
int iterate(uegccd_ctx** cs, int n, const BatchDev& b, State* gather, std::vector<State>& st,
            float* ms_out) {
  uegccd_ctx* c0 = cs[0];
  hipStream_t s = c0->stream;
  const int prec64 = c0->opt.precision == UEGCCD_PREC_MIXED ? 0 : 1;
  for (int k = 0; k < n; ++k) {
    if ((cs[k]->opt.precision == UEGCCD_PREC_MIXED ? 0 : 1) != prec64) return UEGCCD_ERR_INVALID;
    if (int rc = ccd_check(cs[k])) return rc;
  }
  int maxiter = 0;
  for (int k = 0; k < n; ++k) maxiter = std::max(maxiter, cs[k]->maxiter);
  // events and the pinned state buffer are made once per context and reused
  // (the allocation-free solve makes neither: host clock, pageable polls)
  const bool no_alloc = c0->no_alloc;
  std::vector<State> pageable;
  State* hst = nullptr;
  if (no_alloc) {
    pageable.resize(n);
    hst = pageable.data();
  } else {
    if (!c0->span[0]) HIPCHK(hipEventCreate(&c0->span[0]));
    if (!c0->span[1]) HIPCHK(hipEventCreate(&c0->span[1]));
    if (c0->h_states_cap < n) {
      if (c0->h_states) hipHostFree(c0->h_states);
      c0->h_states = nullptr;
      c0->h_states_cap = 0;
      HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&c0->h_states), sizeof(State) * n, hipHostMallocDefault));
      c0->h_states_cap = n;
    }
    hst = c0->h_states;
  }
  const hipEvent_t e0 = c0->span[0], e1 = c0->span[1];
  const auto host_t0 = std::chrono::steady_clock::now();
  if (!no_alloc) HIPCHK(hipEventRecord(e0, s));
  launch_init(b, s);
  launch_finalize(b, FIN_INIT, s, prec64);
  HIPCHK(hipGetLastError());
  const int first = std::max(1, c0->opt.iters_per_sync);
  const int cap = std::max(first, kMaxChunk);
  int chunk = first;
  st.assign(n, State{});
  std::vector<double> prev_r(n, 0.0);
  int passes = 0;
  bool all_done = false;
  const bool timing = c0->timed_kernel >= 0 && !no_alloc;
  if (timing && c0->tev.size() < 2) {
    for (size_t k = c0->tev.size(); k < 2; ++k) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      c0->tev.push_back(e);
    }
  }
  std::vector<int> prev_niter(n, 1);  // as left by the init
  // The timed kernel carries its events in one pass per chunk, pass k0: an
  // event on a dispatch holds the next one back by a microsecond or two, so
  // the timed solves keep them few.  A timed first kernel of the pass takes
  // its `before` event from the pass before it, so it is timed in the second
  // pass of a chunk.
  const bool timed_first = timing && c0->timed_kernel == first_pass_kernel(b, MF_ITERATE);
  const int k0 = timed_first ? 1 : 0;
  // a single context's passes replay as captured graphs (not while a kernel
  // is timed with events, nor without device allocation: instantiating a
  // graph allocates)
  const bool persist = n == 1 && &b == &c0->batch && c0->opt.persistent && !timing && !no_alloc &&
                       !c0->sd.slice_split;
  const bool graphs = !persist && n == 1 && &b == &c0->batch && c0->opt.graph && !timing && !no_alloc;
  const int persist_lds = std::max(slice_lds_bytes(c0->o), b.con_lds);
  while (!all_done && passes <= maxiter + 2) {
    if (persist) {
      if (launch_persistent(b, persist_lds, chunk, s)) {
        std::fprintf(stderr, "uegccd: the persistent pass kernel could not be launched\n");
        return UEGCCD_ERR_HIP;
      }
    } else if (graphs) {
      hipGraphExec_t ex = nullptr;
      const int rc = chunk_graph(c0, chunk, &ex);
      if (rc) return rc;
      HIPCHK(hipGraphLaunch(ex, s));
    }
    // (a single solve's k_slice runs the reductions of the pass before it,
    // and the chunk's last pass has a finalize launch of its own; a batch
    // has a finalize launch after every pass)
    const bool fold = fold_state_update(n, c0);
    for (int k = 0; k < chunk && !graphs && !persist; ++k) {
      if (timing && k == k0) {
        const KernelTimer t{c0->timed_kernel, c0->tev[0], c0->tev[1]};
        launch_iteration(b, MF_ITERATE, s, &t, nullptr, fold && k > 0);
      } else {
        const hipEvent_t tail = (timing && k + 1 == k0 && k0 < chunk) ? c0->tev[0] : nullptr;
        launch_iteration(b, MF_ITERATE, s, nullptr, tail, fold && k > 0);
      }
      if (!fold) launch_finalize(b, FIN_ITER, s);
    }
    if (!graphs && !persist && fold) launch_finalize(b, FIN_ITER, s);
    passes += chunk;
    HIPCHK(hipGetLastError());
    if (gather) {
      launch_gather_states(b, gather, s);
      HIPCHK(hipMemcpyAsync(hst, gather, sizeof(State) * n, hipMemcpyDeviceToHost, s));
    } else {
      for (int k = 0; k < n; ++k)
        HIPCHK(hipMemcpyAsync(hst + k, cs[k]->d_state, sizeof(State), hipMemcpyDeviceToHost, s));
    }
    HIPCHK(hipStreamSynchronize(s));
    std::copy(hst, hst + n, st.begin());
    if (timing) {
      // the timed pass counts when some solve was still iterating in it (the
      // chunk's iterating passes come first: its NIter advance)
      int iterating = 0;
      for (int k = 0; k < n; ++k) {
        iterating = std::max(iterating, st[k].niter - prev_niter[k]);
        prev_niter[k] = st[k].niter;
      }
      if (k0 < chunk && iterating > k0) {
        float us = 0.f;
        HIPCHK(hipEventElapsedTime(&us, c0->tev[0], c0->tev[1]));
        c0->timed_us += 1000.0 * us;
        c0->timed_launches += 1;
      }
    }
    all_done = true;
    for (int k = 0; k < n; ++k) all_done = all_done && st[k].phase == PH_DONE;
    if (!all_done) chunk = next_chunk(cs, n, st, prev_r, chunk, first, cap);
  }
  float ms = 0.f;
  if (no_alloc) {
    HIPCHK(hipStreamSynchronize(s));
    ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - host_t0).count();
  } else {
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipEventSynchronize(e1));
    hipEventElapsedTime(&ms, e0, e1);
  }
  if (ms_out) *ms_out = ms;
  for (int k = 0; k < n; ++k) cs[k]->last_niter = st[k].niter;
  return UEGCCD_OK;
}

int fill_result(uegccd_ctx* c, const State& st, float ms, uegccd_result* res) {
  int status = UEGCCD_OK;
  if (st.nonfinite) status = UEGCCD_ERR_NONFINITE;
  else if (st.phase != PH_DONE || !st.converged) status = UEGCCD_ERR_NOT_CONVERGED;
  if (res) {
    std::memset(res, 0, sizeof(*res));
    res->ecorr = st.ecorr;
    res->dt = st.dt;
    res->maxres = st.maxres;
    res->niter = st.niter;
    res->failed = (st.maxres > c->failratio * c->tolmax) || status != UEGCCD_OK;
    res->iters_fp32 = st.iters_fp32;
    res->status = status;
    res->solve_ms = ms;
  }
  return status;
}


}  // namespace

// DrvdRPA (source/dRPA.f90:23-143) on the context's basis: the rings-only CCD
// (with mosaics when the system asks for them) on the spin-summed T2,
// iterated as the reference does (damped Jacobi; its DIIS is inert) until
// max|dT| < TolMax, then the residual check with FailRatio 10000.
extern "C" int uegccd_drpa(uegccd_ctx* c, double* t2, uegccd_result* res, double* esosex) {
  if (!c || !res) return UEGCCD_ERR_INVALID;
  if (c->opt.mode != UEGCCD_MODE_REFERENCE || c->opt.precision != UEGCCD_PREC_FP64)
    return UEGCCD_ERR_UNSUPPORTED;
  // the reduced ring form takes every Coulomb element from one table: range
  // flags other than 0 are not run here
  if (c->ntab != 1) return UEGCCD_ERR_UNSUPPORTED;
  HIPCHK(hipSetDevice(c->device));
  const auto t_start = std::chrono::steady_clock::now();
  const BatchDev& b = c->batch;
  hipStream_t s = c->stream;
  launch_state_reset(b, PH_ITER, 1, s);
  if (t2) {
    int rc = copy_in_amplitudes(c, t2, nullptr);  // spin-summed T2 in the first component
    if (rc) return rc;
  } else {
    launch_mp2(b, s, 1);
  }
  launch_init(b, s, 1);
  launch_finalize(b, FIN_INIT, s, 1);
  launch_drpa_start(b, (int64_t)c->n3, s);
  HIPCHK(hipGetLastError());
  State st{};
  int rc = read_state(c, &st);
  if (rc) return rc;
  // passes in chunks between state polls, sized from the decay of max|dT| as
  // in the CCD iteration; passes after the last one return at once
  const int first = std::max(1, c->opt.iters_per_sync), cap = std::max(first, kMaxChunk);
  std::vector<double> prev_dt(1, 0.0);
  std::vector<State> polled(1);
  for (int passes = 0, chunk = first; st.phase != PH_DONE && passes <= c->maxiter + 2;) {
    chunk = std::min(chunk, std::max(1, c->maxiter + 3 - passes));
    for (int k = 0; k < chunk; ++k) launch_drpa_pass(b, s);
    HIPCHK(hipGetLastError());
    passes += chunk;
    rc = read_state(c, &st);
    if (rc) return rc;
    polled[0] = st;
    if (st.phase != PH_DONE) chunk = next_chunk(&c, 1, polled, prev_dt, chunk, first, cap);
  }
  const float ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  c->last_niter = st.niter;  // the (E, dT) lines uegccd_trace returns
  const int status = fill_result(c, st, ms, res);
  res->failed = (st.maxres > 10000.0 * c->tolmax) || status != UEGCCD_OK;  // source/dRPA.f90:103-104
  // Ec(SOSEX) of the final amplitudes (CCEnergy with CSosex 1, source/dRPA.f90:200-208),
  // reduced on the device like every other energy
  if (esosex) {
    launch_init(b, s, 2);
    launch_finalize(b, FIN_SUM, s);
    HIPCHK(hipGetLastError());
    State se{};
    rc = read_state(c, &se);
    if (rc) return rc;
    *esosex = se.ecorr;
  }
  if (t2) {
    rc = copy_out_amplitudes(c, nullptr, t2, nullptr);
    if (rc) return rc;
  }
  return status;
}

extern "C" int uegccd_solve_batch(uegccd_ctx** cs, int32_t n, double** tab, double** tba,
                                  uegccd_result* res) {
  if (!cs || n <= 0 || !tab || !tba) return UEGCCD_ERR_INVALID;
  for (int k = 0; k < n; ++k)
    if (!cs[k] || !tab[k] || !tba[k] || cs[k]->device != cs[0]->device) return UEGCCD_ERR_INVALID;
  uegccd_ctx* c0 = cs[0];
  HIPCHK(hipSetDevice(c0->device));
  BatchScratch scratch;
  const BatchDev* b = &c0->batch;
  if (n > 1) {
    int rc = make_batch(cs, n, &scratch);
    if (rc) { if (scratch.base) hipFree(scratch.base); return rc; }
    b = &scratch.b;
  }
  int rc = UEGCCD_OK;
  for (int k = 0; k < n && !rc; ++k) rc = copy_in_amplitudes(cs[k], tab[k], tba[k]);
  std::vector<State> st;
  float ms = 0.f;
  if (!rc) rc = iterate(cs, n, *b, scratch.states, st, &ms);
  if (scratch.base) {
    hipStreamSynchronize(c0->stream);
    hipFree(scratch.base);
  }
  if (rc) return rc;
  int worst = UEGCCD_OK;
  for (int k = 0; k < n; ++k) {
    int r2 = copy_out_amplitudes(cs[k], nullptr, tab[k], tba[k]);
    if (r2) return r2;
    int status = fill_result(cs[k], st[k], ms, res ? &res[k] : nullptr);
    if (status != UEGCCD_OK && worst == UEGCCD_OK) worst = status;
  }
  return worst;
}

extern "C" int uegccd_solve_resident(uegccd_ctx* c, int32_t from_mp2, uegccd_result* res) {
  if (!c) return UEGCCD_ERR_INVALID;
  HIPCHK(hipSetDevice(c->device));
  if (from_mp2) {
    launch_state_reset(c->batch, PH_ITER, 1, c->stream);
    launch_mp2(c->batch, c->stream);
    HIPCHK(hipGetLastError());
  }
  std::vector<State> st;
  float ms = 0.f;
  uegccd_ctx* cs[1] = {c};
  int rc = iterate(cs, 1, c->batch, nullptr, st, &ms);
  if (rc) return rc;
  return fill_result(c, st[0], ms, res);
}

extern "C" int uegccd_connectivity(uegccd_ctx* c, int32_t* counts, int32_t nbins) {
  if (!c || !counts || nbins <= 0 || nbins > 4096) return UEGCCD_ERR_INVALID;
  HIPCHK(hipSetDevice(c->device));
  int* d = nullptr;
  if (hipMalloc(&d, sizeof(int) * (nbins + 1)) != hipSuccess) return UEGCCD_ERR_ALLOC;
  std::vector<int> h(nbins + 1, 0);
  int rc = UEGCCD_OK;
  if (hipMemsetAsync(d, 0, sizeof(int) * (nbins + 1), c->stream) != hipSuccess ||
      connectivity_on_device(c->batch, (int)c->nvalid, nbins, d, d + nbins, c->stream) ||
      hipMemcpyAsync(h.data(), d, sizeof(int) * (nbins + 1), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    rc = UEGCCD_ERR_HIP;
  hipFree(d);
  if (rc) return rc;
  if (h[nbins]) return UEGCCD_ERR_INVALID;  // a transfer beyond the histogram
  for (int k = 0; k < nbins; ++k) counts[k] = h[k];
  return UEGCCD_OK;
}

extern "C" int uegccd_get_amplitudes(uegccd_ctx* c, double* taa, double* tab, double* tba) {
  if (!c) return UEGCCD_ERR_INVALID;
  HIPCHK(hipSetDevice(c->device));
  return copy_out_amplitudes(c, taa, tab, tba);
}

extern "C" int uegccd_time_kernel(uegccd_ctx* c, int32_t which) {
  if (!c || which < -1 || which > 3 || which == 2) return UEGCCD_ERR_INVALID;
  c->timed_kernel = which;
  c->timed_us = 0.0;
  c->timed_launches = 0;
  return UEGCCD_OK;
}

extern "C" int uegccd_kernel_stats(uegccd_ctx* c, double* total_us, int64_t* launches) {
  if (!c || !total_us || !launches) return UEGCCD_ERR_INVALID;
  *total_us = c->timed_us;
  *launches = c->timed_launches;
  c->timed_us = 0.0;
  c->timed_launches = 0;
  return UEGCCD_OK;
}

extern "C" int uegccd_set_tiles(uegccd_ctx* c, const uegccd_tiles* t) {
  if (!c) return UEGCCD_ERR_INVALID;
  Tiles x = tuned_tiles(c->noo_max, c->nvv_max, c->opt.precision == UEGCCD_PREC_MIXED);
  if (t) {
    if ((t->ladder_group != 1 && t->ladder_group != 2 && t->ladder_group != 4) ||
        (t->quad_steps != 4 && t->quad_steps != 8) || (t->hh_steps != 4 && t->hh_steps != 8))
      return UEGCCD_ERR_INVALID;
    x.ladder_group = t->ladder_group;
    x.quad_steps = t->quad_steps;
    x.hh_steps = t->hh_steps;
  }
  c->tiles = x;
  c->batch.tiles = x;
  for (auto& g : c->graphs) hipGraphExecDestroy(g.second);  // captured with the old kernels
  c->graphs.clear();
  return UEGCCD_OK;
}

extern "C" int uegccd_get_tiles(const uegccd_ctx* c, uegccd_tiles* t, int32_t* noo_max, int32_t* nvv_max) {
  if (!c || !t) return UEGCCD_ERR_INVALID;
  t->ladder_group = c->tiles.ladder_group;
  t->quad_steps = c->tiles.quad_steps;
  t->hh_steps = c->tiles.hh_steps;
  if (noo_max) *noo_max = c->noo_max;
  if (nvv_max) *nvv_max = c->nvv_max;
  return UEGCCD_OK;
}

extern "C" int uegccd_kernel_time(uegccd_ctx* c, int32_t which, int32_t reps, double* avg_us) {
  if (!c || !avg_us || reps <= 0 || which < 0 || which > 3 || which == 2) return UEGCCD_ERR_INVALID;
  if (int rc = ccd_check(c)) return rc;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  // the ladders in the context's starting precision (FP32 for mixed precision)
  const int prec64 = c->opt.precision == UEGCCD_PREC_MIXED ? 0 : 1;
  launch_state_reset(c->batch, PH_ITER, prec64, s);
  launch_one(c->batch, which, s);  // warm
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, s));
  for (int k = 0; k < reps; ++k) launch_one(c->batch, which, s);
  HIPCHK(hipEventRecord(e1, s));
  HIPCHK(hipEventSynchronize(e1));
  HIPCHK(hipGetLastError());
  float ms = 0.f;
  hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  *avg_us = 1000.0 * ms / reps;
  launch_state_reset(c->batch, PH_DONE, 1, s);
  HIPCHK(hipStreamSynchronize(s));
  return UEGCCD_OK;
}

// ---------------------------------------------------------------------------
// Plane-wave basis, init_HEG + SetupBasis + change_rs (source/HEG.f90:16-43,
// 150-324, 494-533), CalcMadelung (source/HEG.f90:89-148) and print_HF's E_HF
// (source/HEG.f90:581-615).
namespace {

double madelung_ewald(double omega, double L) {
  const double kappa = 2.8 / std::pow(omega, 1.0 / 3.0);
  const double term2 = -kPi / (std::pow(kappa, 2.0) * std::pow(L, 3.0));
  const double term4 = -2.0 * kappa / std::sqrt(kPi);
  // the reference accumulates cubes of half-width 1..10 and keeps the last
  double recip = 0.0, real = 0.0;
  const int l4 = 10;
  for (int a = -l4; a <= l4; ++a)
    for (int b = -l4; b <= l4; ++b)
      for (int c = -l4; c <= l4; ++c) {
        const int n2 = a * a + b * b + c * c;
        if (n2 == 0) continue;
        const double k2 = (1.0 / std::pow(L, 2.0)) * (double)n2;
        recip += (1.0 / std::pow(L, 3.0)) * (1.0 / (kPi * k2)) *
                 std::exp(-std::pow(kPi, 2.0) * k2 / std::pow(kappa, 2.0));
        const double r = L * std::sqrt((double)n2);
        real += std::erfc(kappa * r) / r;
      }
  return real + recip + term2 + term4;
}

// The Ewald sum depends on (Omega, L) alone, and a twist-averaging or r_s-scan
// stream repeats few of those: each pair's value is kept for the process (a
// few thousand at most), and a repeated pair gets the same bits.
double madelung_cached(double omega, double L) {
  static std::mutex mu;
  static std::map<std::pair<uint64_t, uint64_t>, double> memo;
  uint64_t ko, kl;
  std::memcpy(&ko, &omega, sizeof ko);
  std::memcpy(&kl, &L, sizeof kl);
  const auto key = std::make_pair(ko, kl);
  {
    std::lock_guard<std::mutex> lk(mu);
    const auto it = memo.find(key);
    if (it != memo.end()) return it->second;
  }
  const double v = madelung_ewald(omega, L);
  std::lock_guard<std::mutex> lk(mu);
  if (memo.size() < 4096) memo.emplace(key, v);
  return v;
}

// Geometry at r_s = 1, then rescaled to r_s as change_rs does
// (source/HEG.f90:494-533), and the Madelung constant.
struct HegGeom {
  double L1, rescale, L, omega, mad;
};

HegGeom heg_geometry(const uegccd_heg_params* p) {
  HegGeom g;
  g.L1 = 1.0 * std::pow(4.0 / 3.0 * kPi * p->nel, 1.0 / 3.0);
  g.rescale = p->rs / 1.0;
  g.L = g.L1 * g.rescale;
  g.omega = std::pow(g.L1, 3) * std::pow(g.rescale, 3.0);
  g.mad = madelung_cached(g.omega, g.L) * p->madelungfactor;
  return g;
}

// L, Omega, Madelung, E_HF, ScreeningDistance, DenomFactor, TolMax, FailRatio.
void heg_scalars(const uegccd_heg_params* p, const HegGeom& g, double ehf, double* scalars) {
  const double dens = p->nel / std::pow(g.L, 3.0);
  const double kf = std::pow(3.0 * std::pow(kPi, 2.0) * dens, 1.0 / 3.0);
  double tolmax, failratio, df;
  find_tol(p->rs, &tolmax, &failratio, &df);
  scalars[0] = g.L;
  scalars[1] = g.omega;
  scalars[2] = g.mad;
  scalars[3] = ehf;
  scalars[4] = std::pow(4.0 / kPi * kf, 0.5);
  scalars[5] = df;
  scalars[6] = tolmax;
  scalars[7] = failratio;
}

int fcut_count(double f, int nel) {
  int m = (int)(std::pow(f, 1.5) * nel / 2.0);
  const double t = std::pow(f, 1.5) * nel / 2.0;
  if (std::fabs(m - t) > std::fabs(m + 1 - t)) m += 1;
  return m;
}

// The basis builder: out(nb) gives the n-vector (nb*3, column-major) and
// eigenvalue buffers for the nb orbitals found (either may be null).
template <typename Out>
int heg_build_core(const uegccd_heg_params* p, int32_t* nao_out, Out out, double* scalars) {
  if (!p || !nao_out || p->nel <= 0 || (p->nel % 2) != 0 || p->maxkpoint < 0 || !(p->rs > 0.0) ||
      (p->cutoff_twisted != 0 && p->cutoff_twisted != 1))
    return UEGCCD_ERR_INVALID;
  const int nel = p->nel, no = nel / 2;
  const int cell = (int)std::sqrt((double)p->maxkpoint) + 3;
  struct PW { int n[3]; double nt2; double nt[3]; };
  std::vector<PW> pw;
  for (int a = -cell; a <= cell; ++a)
    for (int b = -cell; b <= cell; ++b)
      for (int c = -cell; c <= cell; ++c) {
        PW w;
        w.n[0] = a; w.n[1] = b; w.n[2] = c;
        w.nt[0] = a + p->twist[0];
        w.nt[1] = b + p->twist[1];
        w.nt[2] = c + p->twist[2];
        w.nt2 = w.nt[0] * w.nt[0] + w.nt[1] * w.nt[1] + w.nt[2] * w.nt[2];
        if (p->cutoff_twisted ? w.nt2 > (double)p->maxkpoint : a * a + b * b + c * c > p->maxkpoint) continue;
        pw.push_back(w);
      }
  // ties keep enumeration order, as the reference's bubble sort on strict '<'
  std::stable_sort(pw.begin(), pw.end(), [](const PW& x, const PW& y) { return x.nt2 < y.nt2; });
  int nb = (int)pw.size();
  const int mf = fcut_count(p->fcut, nel);
  if (mf < nb) nb = mf;
  const int mc = fcut_count(p->corefcut, nel);
  if ((p->corefcut - 1.0) < 1e-6 && mc != no) return UEGCCD_ERR_INVALID;
  nb = nb - mc + no;
  if (nb <= no) return UEGCCD_ERR_INVALID;
  *nao_out = nb;
  int32_t* nvec;
  double* eigen;
  std::tie(nvec, eigen) = out(nb);
  if (!nvec && !eigen && !scalars) return UEGCCD_OK;
  const HegGeom g = heg_geometry(p);
  const double L1 = g.L1, rescale = g.rescale, L = g.L, omega = g.omega, mad = g.mad;
  std::vector<double> kv((size_t)nb * 3);
  for (int q = 0; q < nb; ++q) {
    const PW& w = pw[q < no ? q : q + mc - no];
    for (int d = 0; d < 3; ++d) {
      kv[3 * q + d] = (w.nt[d] * (2.0 * kPi / L1)) * (1.0 / rescale);
      if (nvec) nvec[(size_t)d * nb + q] = w.n[d];
    }
  }
  if (eigen) {
    for (int q = 0; q < nb; ++q) {
      const double* k1 = &kv[3 * q];
      double e = 0.5 * (k1[0] * k1[0] + k1[1] * k1[1] + k1[2] * k1[2]);
      for (int l = 0; l < no; ++l) {
        if (l != q) {
          const double* k2 = &kv[3 * l];
          const double d0 = k1[0] - k2[0], d1 = k1[1] - k2[1], d2v = k1[2] - k2[2];
          e = e - (1.0 / omega) * (4.0 * kPi / (d0 * d0 + d1 * d1 + d2v * d2v)) * p->exchangefactor;
        } else {
          e = e + mad - p->gap;
        }
      }
      eigen[q] = e;
    }
  }
  if (scalars) {
    double ehf = 0.0;
    for (int q = 0; q < no; ++q) {
      const double* k1 = &kv[3 * q];
      ehf = ehf + 0.5 * (k1[0] * k1[0] + k1[1] * k1[1] + k1[2] * k1[2]);
      for (int l = 0; l < no; ++l) {
        if (l == q) continue;
        const double* k2 = &kv[3 * l];
        const double d0 = k1[0] - k2[0], d1 = k1[1] - k2[1], d2v = k1[2] - k2[2];
        ehf = ehf - 0.5 * (1.0 / omega) * (4.0 * kPi / (d0 * d0 + d1 * d1 + d2v * d2v));
      }
    }
    ehf = ehf * 2.0 + mad * 0.5 * nel;
    heg_scalars(p, g, ehf, scalars);
  }
  return UEGCCD_OK;
}

}  // namespace

extern "C" int uegccd_heg_build(const uegccd_heg_params* p, int32_t* nao_out, int32_t* nvec,
                                double* eigen, double* scalars) {
  return heg_build_core(p, nao_out, [&](int) { return std::make_pair(nvec, eigen); }, scalars);
}

// print_ex_SF (source/HEG.f90:617-674), which print_HF writes to unit 82
// after every basis set-up (at r_s = 1, source/HEG.f90:319) and every change
// of r_s (source/HEG.f90:531): over ordered pairs of distinct occupied
// orbitals, the count at q = n_l1 - n_l2 and the exchange integral
// (1/Omega) 4 pi / |k_l1 - k_l2|^2 of the cell's last pair in loop order.
// A host routine: O(NOcc^2) bookkeeping beside the basis builder.
extern "C" int uegccd_exchange_structure_factor(const uegccd_heg_params* p, int32_t* ncell, int32_t* qvec,
                                                int32_t* count, double* exv) {
  if (!p || !ncell) return UEGCCD_ERR_INVALID;
  int32_t nao = 0;
  int rc = uegccd_heg_build(p, &nao, nullptr, nullptr, nullptr);
  if (rc) return rc;
  std::vector<int32_t> nvec((size_t)nao * 3);
  rc = uegccd_heg_build(p, &nao, nvec.data(), nullptr, nullptr);
  if (rc) return rc;
  const int no = p->nel / 2;
  const HegGeom g = heg_geometry(p);
  std::vector<double> kv((size_t)no * 3);
  for (int q = 0; q < no; ++q)
    for (int d = 0; d < 3; ++d) {
      const double nt = nvec[(size_t)d * nao + q] + p->twist[d];
      kv[3 * q + d] = (nt * (2.0 * kPi / g.L1)) * (1.0 / g.rescale);
    }
  std::map<std::tuple<int, int, int>, std::pair<int, double>> cells;  // (q_x, q_y, q_z) order
  for (int l1 = 0; l1 < no; ++l1)
    for (int l2 = 0; l2 < no; ++l2) {
      if (l1 == l2) continue;
      const auto key = std::make_tuple(nvec[l1] - nvec[l2], nvec[(size_t)nao + l1] - nvec[(size_t)nao + l2],
                                       nvec[2 * (size_t)nao + l1] - nvec[2 * (size_t)nao + l2]);
      const double d0 = kv[3 * l1] - kv[3 * l2], d1 = kv[3 * l1 + 1] - kv[3 * l2 + 1],
                   d2v = kv[3 * l1 + 2] - kv[3 * l2 + 2];
      auto& cell = cells[key];
      cell.first += 1;
      cell.second = (1.0 / g.omega) * (4.0 * kPi / (d0 * d0 + d1 * d1 + d2v * d2v));
    }
  const int n = (int)cells.size();
  if (!qvec || !count || !exv) {
    *ncell = n;
    return UEGCCD_OK;
  }
  if (*ncell < n) return UEGCCD_ERR_INVALID;
  int k = 0;
  for (const auto& e : cells) {
    qvec[3 * k] = std::get<0>(e.first);
    qvec[3 * k + 1] = std::get<1>(e.first);
    qvec[3 * k + 2] = std::get<2>(e.first);
    count[k] = e.second.first;
    exv[k] = e.second.second;
    ++k;
  }
  *ncell = n;
  return UEGCCD_OK;
}

// ---------------------------------------------------------------------------
// Request server: continuous batching of independent solves.  Each step runs
// a few passes of every admitted solve in shared launches; between steps,
// finished solves leave (their memory is freed) and queued ones join.  While
// the passes run, the host builds the next requests' bases and index maps.
namespace {

using Clock = std::chrono::steady_clock;

// Plan states: a planner thread (or the scheduling thread itself) claims a
// request 0 -> 1 and publishes its plan with 1 -> 2 (release); the scheduling
// thread reads the plan only after seeing 2 (acquire).
enum PlanState : int { PLAN_NONE = 0, PLAN_BUSY = 1, PLAN_DONE = 2 };

struct Request {
  uegccd_request req{};
  uegccd_request_result res{};     // read by poll (scheduling thread only)
  uegccd_ctx* ctx = nullptr;       // planned (host maps) or placed on the device
  bool placed = false;
  std::vector<int32_t> nvec;
  std::vector<double> eigen;
  uegccd_system sys{};
  std::atomic<int> plan_state{PLAN_NONE};
  int plan_rc = UEGCCD_OK;
  int32_t nao = 0;                 // results of planning, copied into res on admission
  double ehf = 0.0;
  int64_t footprint = 0;
  double plan_ms = 0.0;            // host time of plan_request
  std::vector<char> image;         // constant section laid out by the planner (descriptor at admission)
};

}  // namespace

namespace {

// Device memory of the server's contexts: 1 GiB slabs (or one slab of the
// budget when that is smaller), first-fit with coalescing free lists.  Slabs
// are kept until the server is destroyed.
struct DevicePool {
  struct Slab {
    char* base = nullptr;
    size_t size = 0;
    std::map<size_t, size_t> free;  // offset -> length
  };
  std::vector<Slab> slabs;
  size_t slab_bytes = size_t(1) << 30;

  char* alloc(size_t bytes) {
    bytes = aligned(bytes);
    for (Slab& sl : slabs)
      for (auto it = sl.free.begin(); it != sl.free.end(); ++it)
        if (it->second >= bytes) {
          const size_t off = it->first, len = it->second;
          sl.free.erase(it);
          if (len > bytes) sl.free[off + bytes] = len - bytes;
          return sl.base + off;
        }
    Slab sl;
    sl.size = std::max(slab_bytes, bytes);
    if (hipMalloc(&sl.base, sl.size) != hipSuccess) return nullptr;
    if (sl.size > bytes) sl.free[bytes] = sl.size - bytes;
    slabs.push_back(std::move(sl));
    return slabs.back().base;
  }
  void release(char* p, size_t bytes) {
    bytes = aligned(bytes);
    for (Slab& sl : slabs) {
      if (p < sl.base || p >= sl.base + sl.size) continue;
      auto it = sl.free.emplace((size_t)(p - sl.base), bytes).first;
      auto nx = std::next(it);
      if (nx != sl.free.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        sl.free.erase(nx);
      }
      if (it != sl.free.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
          pv->second += it->second;
          sl.free.erase(it);
        }
      }
      return;
    }
  }
  void clear() {
    for (Slab& sl : slabs) hipFree(sl.base);
    slabs.clear();
  }
};

// Host bytes staged for uploads per scheduling step before admission waits
// for the next step.
constexpr size_t kStagePerStep = size_t(64) << 20;

// One step's admissions travel to the device in one copy: the constant
// images of the admitted contexts packed into a pinned buffer, then the table
// of copy jobs; a scatter kernel puts each image in its arena and clears
// each work region (one API call and one launch, not two per request).
struct AdmitUpload {
  char* host = nullptr;  // pinned, kStagePerStep + the job table
  char* dev = nullptr;   // the same size on the device
  size_t cap = 0, used = 0;
  std::vector<CopyJob> jobs;
  std::vector<size_t> src_off;  // each job's source offset in the buffer (kNone: a clear)
  static constexpr size_t kNone = ~size_t(0);
  static size_t table_bytes() { return aligned(sizeof(CopyJob) * 2 * 4096); }
  int ensure() {
    if (host) return UEGCCD_OK;
    cap = kStagePerStep;
    if (hipHostMalloc(reinterpret_cast<void**>(&host), cap + table_bytes(), hipHostMallocDefault) != hipSuccess)
      return UEGCCD_ERR_ALLOC;
    if (hipMalloc(&dev, cap + table_bytes()) != hipSuccess) return UEGCCD_ERR_ALLOC;
    return UEGCCD_OK;
  }
  bool fits(size_t bytes) const { return used + aligned(bytes) <= cap && jobs.size() + 2 <= 2 * 4096; }
  // an image to dst (its size rounded up to 16 bytes: the arena's next
  // buffer starts 256-aligned past it) and a clear of `clear` bytes at zdst
  void add(char* dst, const std::vector<char>& image, char* zdst, size_t clear) {
    const size_t n = (image.size() + 15) & ~size_t(15);
    std::memcpy(host + used, image.data(), image.size());
    std::memset(host + used + image.size(), 0, n - image.size());
    if (n) {
      jobs.push_back({dst, nullptr, n});
      src_off.push_back(used);
    }
    used += aligned(n);
    if (clear) {
      jobs.push_back({zdst, nullptr, clear});
      src_off.push_back(kNone);
    }
  }
  // Sends what was added since the last send (so that the device works on
  // the first admissions while the host prepares the next); `last` ends the
  // step's upload.
  size_t sent = 0, jobs_sent = 0;
  int send(hipStream_t s, bool last) {
    if (jobs.size() > jobs_sent) {
      const size_t nj = jobs.size() - jobs_sent;
      for (size_t k = jobs_sent; k < jobs.size(); ++k)
        jobs[k].src = src_off[k] == kNone ? nullptr : dev + src_off[k];
      CopyJob* table = reinterpret_cast<CopyJob*>(host + cap) + jobs_sent;
      std::memcpy(table, jobs.data() + jobs_sent, sizeof(CopyJob) * nj);
      if (used > sent) HIPCHK(hipMemcpyAsync(dev + sent, host + sent, used - sent, hipMemcpyHostToDevice, s));
      CopyJob* dtable = reinterpret_cast<CopyJob*>(dev + cap) + jobs_sent;
      HIPCHK(hipMemcpyAsync(dtable, table, sizeof(CopyJob) * nj, hipMemcpyHostToDevice, s));
      launch_scatter(dtable, (int)nj, s);
      sent = used;
      jobs_sent = jobs.size();
    }
    if (last) {
      jobs.clear();
      src_off.clear();
      used = sent = jobs_sent = 0;
    }
    return UEGCCD_OK;
  }
  void release() {
    if (host) hipHostFree(host);
    if (dev) hipFree(dev);
    host = dev = nullptr;
  }
};

}  // namespace

struct uegccd_server {
  uegccd_server_options opt{};
  hipStream_t stream = nullptr;
  DevicePool pool;
  PinnedStage stage;
  Clock::time_point t0;
  std::vector<std::unique_ptr<Request>> reqs;  // indexed by request id
  // planner threads: build bases and maps of queued requests ahead of admission
  std::vector<std::thread> planners;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> to_plan;    // guarded by mu
  // uploaded constant images kept for the planners to lay the next ones out
  // in (memory already mapped: no page faults), at most lookahead of them (mu)
  std::vector<std::vector<char>> spare_images;
  int planned_waiting = 0;         // claimed for planning, not yet admitted (mu)
  int lookahead = 128;
  bool stopping = false;           // mu
  std::vector<int64_t> queue;      // waiting, in arrival order (head at qhead)
  size_t qhead = 0;
  std::vector<int64_t> active;     // iterated together
  size_t active_bytes = 0;
  BatchScratch batch;
  BatchScratch admitted;           // the solves admitted by one step, set up by batched launches
  AdmitUpload upload;              // their images and clears, one copy and one launch per step
  bool batch_stale = true;
  size_t nbatch = 0;               // the first nbatch active solves form the batch
  State* h_states = nullptr;       // pinned, one per batch member
};

namespace {

double server_now(const uegccd_server* s) {
  return std::chrono::duration<double, std::milli>(Clock::now() - s->t0).count();
}

void drop_ctx(uegccd_server* s, Request& r) {
  if (!r.ctx) return;
  if (r.placed) s->pool.release(r.ctx->arena.base, r.ctx->arena.size);
  uegccd_ctx_destroy(r.ctx);  // owns neither memory nor stream: host teardown only
}

void finish_request(uegccd_server* s, Request& r, int status) {
  r.res.state = UEGCCD_REQ_DONE;
  r.res.status = status;
  r.res.t_done_ms = server_now(s);
  drop_ctx(s, r);
  r.ctx = nullptr;
  r.placed = false;
  r.nvec.clear();
  r.nvec.shrink_to_fit();
  r.eigen.clear();
  r.eigen.shrink_to_fit();
}

// Basis, eigenvalues and host maps of a request (init_HEG + SetUpCCScr work).
int plan_request(Request& r) {
  const uegccd_request& q = r.req;
  int32_t nao = 0;
  double sc[8];
  int rc = heg_build_core(&q.heg, &nao, [&r](int nb) {
    r.nvec.assign((size_t)nb * 3, 0);
    r.eigen.assign(nb, 0.0);
    return std::make_pair(r.nvec.data(), r.eigen.data());
  }, sc);
  if (rc) return rc;
  uegccd_system& s = r.sys;
  s.nocc = q.heg.nel / 2;
  s.nao = nao;
  s.nvec = r.nvec.data();
  s.eigen = r.eigen.data();
  s.L = sc[0];
  s.madelung = sc[2];
  s.rs = q.heg.rs;
  s.gapii = q.gapii;
  s.coren = q.coren;
  s.safe_eri = q.safe_eri;
  for (int k = 0; k < 15; ++k) s.irange[k] = q.irange[k];
  s.do_ring = q.do_ring;
  s.do_xring = q.do_xring;
  s.do_ladder = q.do_ladder;
  s.do_mosaic = q.do_mosaic;
  s.screening = sc[4];
  r.nao = nao;
  r.ehf = sc[3];
  rc = plan_ctx(&s, &q.opt, &r.ctx);
  if (rc) return rc;
  r.footprint = (int64_t)arena_bytes(r.ctx, ndiis_of(r.ctx->opt));
  // the constant section at a placeholder base: admission writes the descriptor
  size_t work_off = 0;
  layout_ctx(r.ctx, &r.sys, nullptr, (size_t)r.footprint, r.image, &work_off, true);
  return ccd_check(r.ctx);
}

void planner_loop(uegccd_server* s) {
  for (;;) {
    Request* r = nullptr;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] {
        return s->stopping || (!s->to_plan.empty() && s->planned_waiting < s->lookahead);
      });
      if (s->stopping) return;
      r = s->to_plan.front();
      s->to_plan.pop_front();
      int expect = PLAN_NONE;
      if (!r->plan_state.compare_exchange_strong(expect, PLAN_BUSY)) continue;
      ++s->planned_waiting;
      if (!s->spare_images.empty()) {
        r->image.swap(s->spare_images.back());
        s->spare_images.pop_back();
      }
    }
    const auto t0 = Clock::now();
    r->plan_rc = plan_request(*r);
    r->plan_ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      r->plan_state.store(PLAN_DONE, std::memory_order_release);
    }
    s->cv_done.notify_all();
  }
}

// The plan of the head request: planned inline when nothing runs (or there
// are no planner threads), awaited when a planner holds it and nothing runs,
// else left for a later step.  Returns false when admission should stop.
bool head_planned(uegccd_server* s, Request& r) {
  int st = r.plan_state.load(std::memory_order_acquire);
  const bool idle = s->active.empty();
  if (st == PLAN_NONE && (idle || s->planners.empty())) {
    int expect = PLAN_NONE;
    if (r.plan_state.compare_exchange_strong(expect, PLAN_BUSY)) {
      {
        std::lock_guard<std::mutex> lk(s->mu);
        ++s->planned_waiting;
      }
      const auto t0 = Clock::now();
      r.plan_rc = plan_request(r);
      r.plan_ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
      r.plan_state.store(PLAN_DONE, std::memory_order_release);
      return true;
    }
    st = r.plan_state.load(std::memory_order_acquire);
  }
  if (st == PLAN_DONE) return true;
  if (st == PLAN_BUSY && idle) {
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return r.plan_state.load(std::memory_order_acquire) == PLAN_DONE; });
    return true;
  }
  return false;
}

// A planned request leaves the look-ahead window (admitted or finished).
void release_plan_slot(uegccd_server* s) {
  {
    std::lock_guard<std::mutex> lk(s->mu);
    --s->planned_waiting;
  }
  s->cv_work.notify_one();
}

// Admits queued requests in arrival order while slots and memory allow; the
// solves admitted together get their device-built maps, MP2 guess, E(2) and
// first dT on the server stream from one launch of each kernel.
int admit(uegccd_server* s) {
  const size_t budget = (size_t)s->opt.hbm_budget_bytes;
  std::vector<uegccd_ctx*> fresh;
  while (s->qhead < s->queue.size() && (int)s->active.size() < s->opt.max_active) {
    Request& r = *s->reqs[s->queue[s->qhead]];
    if (!head_planned(s, r)) break;
    r.res.nao = r.nao;
    r.res.nocc = r.req.heg.nel / 2;
    r.res.ehf = r.ehf;
    r.res.footprint_bytes = r.footprint;
    r.res.plan_ms = r.plan_ms;
    if (r.plan_rc) {
      finish_request(s, r, r.plan_rc);
      release_plan_slot(s);
      ++s->qhead;
      continue;
    }
    const size_t fp = (size_t)r.res.footprint_bytes;
    if (fp > budget) {
      finish_request(s, r, UEGCCD_ERR_BUSY);
      release_plan_slot(s);
      ++s->qhead;
      continue;
    }
    if (s->active_bytes + fp > budget) break;
    if (int rc = s->upload.ensure()) return rc;
    if (!s->upload.fits(r.image.size())) {
      if (!fresh.empty() || r.image.size() + 256 <= kStagePerStep) break;  // (the next step takes the rest)
      finish_request(s, r, UEGCCD_ERR_ALLOC);  // an image larger than a step's upload
      release_plan_slot(s);
      ++s->qhead;
      continue;
    }
    char* mem = s->pool.alloc(fp);
    if (!mem) {
      if (!s->active.empty()) break;  // retried once solves have left
      finish_request(s, r, UEGCCD_ERR_ALLOC);
      release_plan_slot(s);
      ++s->qhead;
      continue;
    }
    uegccd_ctx* c = r.ctx;
    c->device = s->opt.device;
    c->opt.device = s->opt.device;
    c->stream = s->stream;
    size_t work_off = 0;
    layout_ctx(c, &r.sys, mem, fp, r.image, &work_off, false);  // the planner built the image
    r.placed = true;
    s->upload.add(mem, r.image, mem + work_off, fp - work_off);
    {
      std::lock_guard<std::mutex> lk(s->mu);  // (the staged copy is independent of the image)
      if ((int)s->spare_images.size() < s->lookahead) s->spare_images.push_back(std::move(r.image));
    }
    std::vector<char>().swap(r.image);
    fresh.push_back(c);
    if (fresh.size() % 32 == 0) {
      if (int rc = s->upload.send(s->stream, false)) return rc;
    }
    r.res.state = UEGCCD_REQ_RUNNING;
    r.res.t_start_ms = server_now(s);
    s->active.push_back(s->queue[s->qhead]);
    s->active_bytes += fp;
    release_plan_slot(s);
    s->batch_stale = true;
    ++s->qhead;
  }
  if (!fresh.empty()) {
    // (their images and clears in one copy and one launch, then one launch of
    // each set-up kernel over the admitted solves, in their own batch
    // tables; each solve starts in its own precision)
    if (int rc = s->upload.send(s->stream, true)) return rc;
    int rc = make_batch(fresh.data(), (int)fresh.size(), &s->admitted, s->stream, &s->stage);
    if (rc) return rc;
    const BatchDev& b = s->admitted.b;
    launch_plan(b, s->stream);
    launch_mp2(b, s->stream);
    launch_init(b, s->stream);
    launch_finalize(b, FIN_INIT, s->stream, -1);
    HIPCHK(hipGetLastError());
  }
  if (s->qhead > 1024 && s->qhead * 2 > s->queue.size()) {
    s->queue.erase(s->queue.begin(), s->queue.begin() + s->qhead);
    s->qhead = 0;
  }
  return UEGCCD_OK;
}

}  // namespace

extern "C" {

int uegccd_server_create(const uegccd_server_options* opt, uegccd_server** out) {
  if (!out) return UEGCCD_ERR_INVALID;
  *out = nullptr;
  uegccd_server_options o{};
  if (opt) o = *opt;
  if (o.hbm_budget_bytes <= 0) o.hbm_budget_bytes = int64_t(64) << 30;
  if (o.max_active <= 0) o.max_active = 64;
  if (o.iters_per_step <= 0) o.iters_per_step = 4;
  if (o.max_active > 4096 || o.planner_threads > 64) return UEGCCD_ERR_INVALID;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return UEGCCD_ERR_HIP;
  if (o.device < 0 || o.device >= ndev) return UEGCCD_ERR_INVALID;
  HIPCHK(hipSetDevice(o.device));
  uegccd_server* s = new (std::nothrow) uegccd_server();
  if (!s) return UEGCCD_ERR_ALLOC;
  s->opt = o;
  s->t0 = Clock::now();
  s->pool.slab_bytes = std::min(size_t(1) << 30, std::max(size_t(64) << 20, (size_t)o.hbm_budget_bytes));
  s->lookahead = std::min(1024, 2 * o.max_active);
  int nplan = o.planner_threads;
  if (nplan == 0) nplan = std::max(1, std::min(8, (int)std::thread::hardware_concurrency() / 2));
  for (int k = 0; k < nplan; ++k) s->planners.emplace_back(planner_loop, s);
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&s->h_states), sizeof(State) * o.max_active,
                    hipHostMallocDefault) != hipSuccess) {
    uegccd_server_destroy(s);
    return UEGCCD_ERR_HIP;
  }
  *out = s;
  return UEGCCD_OK;
}

void uegccd_server_destroy(uegccd_server* s) {
  if (!s) return;
  hipSetDevice(s->opt.device);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopping = true;
  }
  s->cv_work.notify_all();
  for (std::thread& t : s->planners) t.join();
  if (s->stream) hipStreamSynchronize(s->stream);
  for (auto& r : s->reqs) {
    drop_ctx(s, *r);
    r->ctx = nullptr;
  }
  s->pool.clear();
  s->stage.release();
  s->upload.release();
  if (s->batch.base) hipFree(s->batch.base);
  if (s->admitted.base) hipFree(s->admitted.base);
  if (s->h_states) hipHostFree(s->h_states);
  if (s->stream) hipStreamDestroy(s->stream);
  delete s;
}

int uegccd_server_submit(uegccd_server* s, const uegccd_request* req, int64_t* id) {
  if (!s || !req) return UEGCCD_ERR_INVALID;
  std::unique_ptr<Request> r(new (std::nothrow) Request());
  if (!r) return UEGCCD_ERR_ALLOC;
  r->req = *req;
  r->res.state = UEGCCD_REQ_QUEUED;
  r->res.t_submit_ms = server_now(s);
  const int64_t k = (int64_t)s->reqs.size();
  Request* rp = r.get();
  s->reqs.push_back(std::move(r));
  s->queue.push_back(k);
  if (!s->planners.empty()) {
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->to_plan.push_back(rp);
    }
    s->cv_work.notify_one();
  }
  if (id) *id = k;
  return UEGCCD_OK;
}

int uegccd_server_step(uegccd_server* s, int32_t* n_active, int32_t* n_queued) {
  if (!s) return UEGCCD_ERR_INVALID;
  HIPCHK(hipSetDevice(s->opt.device));
  s->stage.reset();  // every copy staged by the last step has completed
  // 1. passes of every admitted solve (solves already done skip their tiles)
  if (!s->active.empty()) {
    if (s->batch_stale) {
      std::vector<uegccd_ctx*> cs;
      for (int64_t k : s->active) cs.push_back(s->reqs[k]->ctx);
      int rc = make_batch(cs.data(), (int)cs.size(), &s->batch, s->stream, &s->stage);
      if (rc) return rc;
      s->batch_stale = false;
      s->nbatch = cs.size();
    }
    const bool fold = false;  // (a batch: a finalize launch after every pass)
    for (int k = 0; k < s->opt.iters_per_step; ++k) {
      launch_iteration(s->batch.b, MF_ITERATE, s->stream, nullptr, nullptr, fold && k > 0);
      if (!fold) launch_finalize(s->batch.b, FIN_ITER, s->stream);
    }
    if (fold) launch_finalize(s->batch.b, FIN_ITER, s->stream);
    launch_gather_states(s->batch.b, s->batch.states, s->stream);
    HIPCHK(hipMemcpyAsync(s->h_states, s->batch.states, sizeof(State) * s->nbatch,
                          hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipGetLastError());
  } else {
    s->nbatch = 0;
  }
  // 2. admissions, prepared on the host while the passes run (a new solve is
  //    never done after its first dT, so it needs no state read this step)
  const size_t nb = s->nbatch;
  int rc = admit(s);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(s->stream));
  // 3. retire finished solves
  const size_t na = s->active.size();
  std::vector<int64_t> keep;
  keep.reserve(na);
  for (size_t k = 0; k < na; ++k) {
    Request& r = *s->reqs[s->active[k]];
    if (k >= nb || s->h_states[k].phase != PH_DONE) { keep.push_back(s->active[k]); continue; }
    const State& st = s->h_states[k];
    uegccd_result res;
    const int status = fill_result(r.ctx, st, 0.f, &res);
    r.res.niter = st.niter;
    r.res.iters_fp32 = st.iters_fp32;
    r.res.ecorr = st.ecorr;
    r.res.e2 = st.e_guess;
    r.res.maxres = st.maxres;
    s->active_bytes -= (size_t)r.res.footprint_bytes;
    finish_request(s, r, status);
    s->batch_stale = true;
  }
  s->active.swap(keep);
  if (n_active) *n_active = (int32_t)s->active.size();
  if (n_queued) *n_queued = (int32_t)(s->queue.size() - s->qhead);
  return UEGCCD_OK;
}

int uegccd_server_drain(uegccd_server* s) {
  if (!s) return UEGCCD_ERR_INVALID;
  int32_t na = 0, nq = 0;
  do {
    int rc = uegccd_server_step(s, &na, &nq);
    if (rc) return rc;
  } while (na > 0 || nq > 0);
  return UEGCCD_OK;
}

int uegccd_server_poll(const uegccd_server* s, int64_t id, uegccd_request_result* res) {
  if (!s || !res) return UEGCCD_ERR_INVALID;
  if (id < 0 || id >= (int64_t)s->reqs.size()) return UEGCCD_ERR_NOT_FOUND;
  *res = s->reqs[id]->res;
  return UEGCCD_OK;
}

double uegccd_server_clock_ms(const uegccd_server* s) { return s ? server_now(s) : 0.0; }

}  // extern "C"

// ---------------------------------------------------------------------------
// The same basis built on the device (enumeration, ordering, orbitals, E_HF);
// geometry, Madelung and the FindTol scalars stay on the host.
extern "C" int uegccd_heg_build_device(const uegccd_heg_params* p, int32_t device,
                                       int32_t* nao_out, int32_t* nvec, double* eigen,
                                       double* scalars) {
  if (!p || !nao_out || p->nel <= 0 || (p->nel % 2) != 0 || p->maxkpoint < 0 || !(p->rs > 0.0) ||
      (p->cutoff_twisted != 0 && p->cutoff_twisted != 1))
    return UEGCCD_ERR_INVALID;
  const int nel = p->nel, no = nel / 2;
  if (no > kMaxOcc) return UEGCCD_ERR_UNSUPPORTED;
  const int cell = (int)std::sqrt((double)p->maxkpoint) + 3;
  const int w = 2 * cell + 1, total = w * w * w;
  int npad = 1;
  while (npad < total) npad <<= 1;
  if (npad > kMaxBasisPad) return UEGCCD_ERR_UNSUPPORTED;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return UEGCCD_ERR_HIP;
  if (device < 0 || device >= ndev) return UEGCCD_ERR_INVALID;
  HIPCHK(hipSetDevice(device));
  // key[npad] | idx[npad] | count | nvec[3*npad] | eigen[npad] | ehf
  const size_t bytes = aligned(sizeof(double) * npad) + aligned(sizeof(int) * npad) + aligned(sizeof(int)) +
                       aligned(sizeof(int) * 3 * npad) + aligned(sizeof(double) * npad) +
                       aligned(sizeof(double));
  char* base = nullptr;
  if (hipMalloc(&base, bytes) != hipSuccess) return UEGCCD_ERR_ALLOC;
  DeviceArena a;
  a.base = base;
  a.size = bytes;
  double* d_key = a.take<double>(npad);
  int* d_idx = a.take<int>(npad);
  int* d_count = a.take<int>(1);
  int* d_nvec = a.take<int>(3 * (size_t)npad);
  double* d_eig = a.take<double>(npad);
  double* d_ehf = a.take<double>(1);
  hipStream_t s = nullptr;
  int rc = UEGCCD_OK;
  auto done = [&](int r) {
    if (s) {
      hipStreamSynchronize(s);
      hipStreamDestroy(s);
    }
    hipFree(base);
    return r;
  };
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return done(UEGCCD_ERR_HIP);
  int count = 0;
  if (basis_on_device(cell, p->maxkpoint, p->cutoff_twisted, p->twist, npad, d_key, d_idx, d_count, s) ||
      hipMemcpyAsync(&count, d_count, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return done(UEGCCD_ERR_HIP);
  int nb = count;
  const int mf = fcut_count(p->fcut, nel);
  if (mf < nb) nb = mf;
  const int mc = fcut_count(p->corefcut, nel);
  if ((p->corefcut - 1.0) < 1e-6 && mc != no) return done(UEGCCD_ERR_INVALID);
  nb = nb - mc + no;
  if (nb <= no) return done(UEGCCD_ERR_INVALID);
  *nao_out = nb;
  if (!nvec && !eigen && !scalars) return done(UEGCCD_OK);
  const HegGeom g = heg_geometry(p);
  double ehf = 0.0;
  if (basis_orbitals_on_device(cell, nb, no, mc, nel, p->twist, g.L1, g.rescale, g.omega, g.mad,
                               p->gap, p->exchangefactor, d_idx, d_nvec, d_eig, d_ehf, s))
    return done(UEGCCD_ERR_HIP);
  if ((nvec && hipMemcpyAsync(nvec, d_nvec, sizeof(int) * 3 * (size_t)nb, hipMemcpyDeviceToHost, s) != hipSuccess) ||
      (eigen && hipMemcpyAsync(eigen, d_eig, sizeof(double) * nb, hipMemcpyDeviceToHost, s) != hipSuccess) ||
      hipMemcpyAsync(&ehf, d_ehf, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return done(UEGCCD_ERR_HIP);
  if (scalars) heg_scalars(p, g, ehf, scalars);
  return done(rc);
}

// ---------------------------------------------------------------------------
// Allocation-free solve (uegccd_solve_noalloc).  The context that
// uegccd_ctx_create would allocate (maps, descriptor, amplitudes, DIIS
// history, intermediates, partials, state) is laid out in `scratch`, a
// file-scope device array the code-object loader places on every device and
// keeps until the process ends.  Host code never uses the array's name (a
// host-side stand-in): the per-device address comes from hipGetSymbolAddress
// on the first call on that device, with that device current.
namespace {

constexpr size_t kNoAllocScratchBytes = size_t(UEGCCD_NOALLOC_SCRATCH_MB) << 20;
constexpr int kMaxScratchDevices = 64;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxScratchDevices] = {};

// The current device's address of `scratch` (device `dev` must be current).
int scratch_of_device(int dev, char** out) {
  if (dev < 0 || dev >= kMaxScratchDevices) return UEGCCD_ERR_INVALID;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      std::fprintf(stderr, "uegccd: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
      return UEGCCD_ERR_HIP;
    }
    g_scratch_ptr[dev] = static_cast<char*>(p);
  }
  *out = g_scratch_ptr[dev];
  return UEGCCD_OK;
}

}  // namespace

extern "C" int uegccd_solve_noalloc(const uegccd_system* sys, const uegccd_options* opt_in,
                                    double* taa, double* tab, double* tba, uegccd_result* res) {
  if (!sys || !tab || !tba) return UEGCCD_ERR_INVALID;
  uegccd_ctx* c = nullptr;
  int rc = plan_ctx(sys, opt_in, &c);
  if (rc) return rc;
  std::unique_ptr<uegccd_ctx, void (*)(uegccd_ctx*)> guard(c, uegccd_ctx_destroy);
  const size_t bytes = arena_bytes(c, ndiis_of(c->opt));
  if (bytes > kNoAllocScratchBytes) return UEGCCD_ERR_SCRATCH;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return UEGCCD_ERR_HIP;
  if (c->device < 0 || c->device >= ndev) return UEGCCD_ERR_INVALID;
  HIPCHK(hipSetDevice(c->device));
  char* base = nullptr;
  rc = scratch_of_device(c->device, &base);
  if (rc) return rc;
  c->no_alloc = true;
  c->stream = nullptr;  // the null stream: no stream object is created
  std::vector<char> image;
  size_t work_off = 0;
  layout_ctx(c, sys, base, bytes, image, &work_off, true);
  if (c->arena.used > c->arena.size) return UEGCCD_ERR_SCRATCH;
  // the constant section, then the work buffers cleared by a fill kernel (the
  // last call's contents are still there)
  HIPCHK(hipMemcpyAsync(base, image.data(), image.size(), hipMemcpyHostToDevice, c->stream));
  launch_zero(base + work_off, bytes - work_off, c->stream);
  launch_plan_maps(c, c->stream);
  HIPCHK(hipGetLastError());
  rc = copy_in_amplitudes(c, tab, tba);
  if (rc) return rc;
  std::vector<State> st;
  float ms = 0.f;
  uegccd_ctx* cs[1] = {c};
  rc = iterate(cs, 1, c->batch, nullptr, st, &ms);
  if (rc) return rc;
  rc = copy_out_amplitudes(c, taa, tab, tba);
  if (rc) return rc;
  return fill_result(c, st[0], ms, res);
}
