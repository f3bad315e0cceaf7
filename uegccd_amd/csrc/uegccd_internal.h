// Internal data model of libuegccd: device-resident per-solve descriptors and
// the iteration state the kernels read and advance on the device.
//
// Layout in HBM (per solve, o = NOcc, v = NAO - NOcc, n3 = o*o*v):
//   amplitudes, residuals, G: dense FP64 T(i,j,a) with i fastest, then j, then
//   the virtual a (the reference's T(NOcc,NOcc,NOcc+1:NAO) column-major order),
//   spin blocks "abab" then "abba" back to back (2*n3 doubles).
//   The valid elements (momentum partner b(i,j,a) = k_i + k_j - k_a a virtual
//   orbital of the basis) in element order: cmap[c] = elem_map(x, position in tk, b).
//   Element-wise kernels run over them (invalid amplitudes stay zero).
//   Pair-momentum blocks K = n_i + n_j: their occupied pairs (x = i + o*j) and
//   their virtual columns a (with partner b), as CSR lists.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace uegccd {

constexpr int kMaxDiis = 16;
constexpr int kPartialStride = 4 + kMaxDiis;  // E, dT, max|R|, non-finite count, DIIS dots
constexpr int kElemThreads = 256;                 // threads of the element-wise blocks
constexpr int kElemPerBlock = kElemThreads;       // amplitudes per element block (one per thread)
constexpr int kFinThreads = 256;                  // threads of the per-solve reduction block
constexpr int kMaxN2 = 1024;                  // |dn|^2 range of the slice blocks' LDS Coulomb table
constexpr int kMaxOcc = 64;                   // NOcc staged per slice in LDS

// Map entry of a valid element, 64 bits (x low word, y high word): the
// element x in the low xb bits, its position in tk in the next pb bits, its
// partner b above them.  The widths are per solve (xb the bits of n3 - 1,
// pb those of the tk size - 1), so a basis fits while xb + pb + the bits of
// nao - 1 stay within 64.
struct ElemCode {
  int xb, pb;
};
__host__ __device__ inline uint2 elem_map(ElemCode c, int x, int pos, int b) {
  const uint64_t m = (uint64_t)(unsigned)x | ((uint64_t)(unsigned)pos << c.xb) |
                     ((uint64_t)(unsigned)b << (c.xb + c.pb));
  return make_uint2((unsigned)m, (unsigned)(m >> 32));
}
__host__ __device__ inline uint64_t elem_word(uint2 m) { return (uint64_t)m.x | ((uint64_t)m.y << 32); }
__host__ __device__ inline int elem_x(ElemCode c, uint2 m) { return (int)(m.x & ((1u << c.xb) - 1u)); }
__host__ __device__ inline int elem_pos(ElemCode c, uint2 m) {
  return (int)((elem_word(m) >> c.xb) & ((1ull << c.pb) - 1ull));
}
__host__ __device__ inline int elem_partner(ElemCode c, uint2 m) {
  return (int)(elem_word(m) >> (c.xb + c.pb));
}
// Bits that hold the values 0 .. n - 1 (at least one).
inline int bits_for(int64_t n) {
  int k = 1;
  while (k < 63 && (int64_t(1) << k) < n) ++k;
  return k;
}

// The places where a Coulomb factor enters the reformulated terms, each with
// its own range flag (source/CCD.f90: the DummyFlag of every ERI call, a Max
// over the term's IRange keys).  Each role reads the table slot rslot[role]
// (slot 0 is always flag 0, the one DrvMBPT uses).
enum Role : int {
  RO_ENERGY = 0,  // v(a-i), v(a-j) of CCEnergy (IRangeEnergy)
  RO_DRVDIR,      // v(i-a) of the drivers (IRangeDriverDirect)
  RO_DRVEX,       // v(i-b) of the drivers (IRangeDriverExchange)
  RO_MOSAIC,      // v(c-k), v(c-l) of the mosaics (IRangeMosaic)
  RO_LAD_LD,      // ladders, linear direct: pp v(c-a), hh v(i-k)
  RO_LAD_LX,      // ladders, linear exchange: pp v(c-b), hh v(i-l)
  RO_LAD_QD,      // ladders, quadratic direct: hh v(c-k)
  RO_LAD_QX,      // ladders, quadratic exchange: hh v(c-l)
  RO_RING_LD,     // rings, linear direct: v(i-a) row-sum factor
  RO_RING_LX,     // rings, linear exchange: v(k-j) of the slice products TV
  RO_RING_QD,     // rings, quadratic direct: v(q) of J1/J2
  RO_RING_QX,     // rings, quadratic exchange: Y(k,l)
  RO_XR_LD,       // crossed rings, linear direct: v(j-a) column-sum factor
  RO_XR_LX,       // crossed rings, linear exchange: v(j-k) of the slice products VT
  RO_XR_QD,       // crossed rings, quadratic direct: v(q') of J1x/J2x
  RO_XR_QX,       // crossed rings, quadratic exchange: Y of the crossed rings
  kRoles
};
constexpr int kMaxTabSlots = 4;  // flags 0..3

enum Phase : int { PH_ITER = 0, PH_FINAL = 1, PH_DONE = 2 };
enum FinKind : int { FIN_ITER = 0, FIN_INIT = 1, FIN_SUM = 2 };
enum MainFlags : int { MF_ITERATE = 1, MF_WRITE_GR = 2, MF_HH2 = 4 };  // MF_HH2: add the hh2 half
// Where the amplitudes of the next residual pass come from (read by k_slice,
// which builds each slab of them and writes it back to tcur and tk).
enum AmpSource : int {
  SRC_TCUR = 0,   // tcur as it stands (the starting amplitudes)
  SRC_PLAIN = 1,  // the last update histT[slot_cur]
  SRC_DIIS = 2    // the DIIS extrapolation sum_k c_k histT[k] (k_slice solves for c)
};

struct State {
  int niter;        // NIter as counted at source/CCD.f90:62,445
  int phase;        // Phase
  int slot;         // history slot the next residual pass writes
  int slot_cur;     // slot written by the last completed pass
  int nhist;        // filled DIIS history entries
  int next_src;     // AmpSource of the next pass's amplitudes
  int prec64;       // 1: FP64 ladders, 0: FP32 ladders
  int iters_fp32;
  int nonfinite;
  int converged;
  double ecorr, ecorr_prev, dt, maxres, maxres_in;
  double e_guess;   // energy of the starting amplitudes (E(2) for the MP2 guess)
  double bmat[kMaxDiis * kMaxDiis];
};
static_assert(sizeof(State) % sizeof(double) == 0, "State is copied as doubles");

// Per-solve descriptor.  Ptr chooses how its pointers are typed: plain on
// the host (SolveDev), global-address-space in the kernels (SolveG), so that
// loads through them compile to global rather than flat memory instructions.
template <template <typename> class Ptr>
struct SolveDevT {
  int o, v, nao, nK;
  int n3;           // o*o*v
  int n2max;        // largest |n_p - n_q|^2 in the basis
  int ndiis;        // history length (1 in reference mode)
  int mode;         // UEGCCD_MODE_*
  int prec0;        // ladder precision a solve starts in: 1 FP64, 0 FP32 (mixed precision)
  int maxiter, diis_start, diis_err_is_dt;
  int do_ring, do_xring, do_ladder, do_mosaic;
  double df;        // DenomFactor
  double tol_dt, tol_e, tol_r, mixed_switch;
  // Coulomb tables, one slot per range flag in use (slot 0: flag 0); role r
  // reads slot rslot[r].  slice_split / main_split: the roles of k_slice /
  // k_main do not all share one slot per array (some range flag differs).
  int ntab;
  int rslot[kRoles];
  int slice_split, main_split;
  int lad_ntab;             // distinct slots of the four ladder roles
  int lad_tab[4];           // LAD_LD, LAD_LX, LAD_QD, LAD_QX -> index among those slots
  int lad_slot[4];          // the distinct slots, in that index order
  // constant per basis
  Ptr<const int4> nv4;      // [nao] (n1,n2,n3,0)
  Ptr<const double> eig;    // [nao]
  Ptr<const double> vtab;   // [ntab][n2max+1] Coulomb table v(|dn|^2)
  Ptr<const double> voo;    // [ntab][o*o] v(n_k - n_j) over occupied k, j (k fastest)
  Ptr<const double> vso;    // [ntab][v*o] v(n_s - n_k), slice (virtual) s, occupied k
  Ptr<const int> kb_poff;   // [nK+1]
  Ptr<const int> kb_pairs;  // [o*o] pair ids x = i + o*j grouped by K
  Ptr<const int> kb_voff;   // [nK+1]
  Ptr<const int> kb_va;     // virtual columns a (orbital index) grouped by K
  Ptr<const int> kb_vb;     // their partners b
  Ptr<const int4> kb_nva;   // (n_a, a) of the same list
  Ptr<const int> kb_toff;   // [nK+1] offsets of the K blocks in tk
  // the valid elements in element order (compact index c; the history layout)
  int nvalid;
  ElemCode ecode;           // bit widths of the cmap entries
  Ptr<const int> slab_coff; // [v+1] first compact index of each virtual slice
  Ptr<const uint2> cmap;    // [nvalid] elem_map(ecode, element x, position in tk, partner b)
  Ptr<const int> pair_kb;   // [o*o] block of pair x
  Ptr<const int> pair_pos;  // [o*o] position of pair x inside its block
  // hole-particle pair sets P(q) = {(i,p): n_p - n_i = q}, one per transfer q
  int nqb;
  Ptr<const int> qb_roff;   // [nqb+1] row offsets
  Ptr<const int> qb_rows;   // [o*v] pairs i + o*(p - o), grouped by q
  Ptr<const int> qb_neg;    // [nqb] block of -q, or -1 when P(-q) is empty
  Ptr<const int> qb_d;      // [nqb*o] virtual index of n_l + q for each l, or -1
  Ptr<const int4> qb_info;  // [nqb] (first row, rows, first row of -q, rows of -q)
  Ptr<const int> qrow;      // [v*o] row of pair (i, p) in the grouped order, at (p - o)*o + i
  Ptr<const int> qb_of;     // [v*o] transfer block of pair (i, p), at (p - o)*o + i
  // work items of the contraction launch: (0, K, row0, 0) hole-hole units,
  // (0, K, row0, col0) particle-particle ladder tiles, (0, kind, arg, arg)
  // ring / crossed-ring / Joo items
  Ptr<const int4> hh_items;
  int nhh;
  Ptr<const int4> hh_items2;  // the units split over the W column halves (a solve alone)
  int nhh2;
  Ptr<const int4> lad_items;   // one 16-column tile each
  int nlad;
  Ptr<const int4> lad_items2;  // two 16-column tiles each (the batches' blocks)
  int nlad2;
  Ptr<const int4> con_items;
  int ncon;
  // mutable buffers
  Ptr<double2> tcur;        // [n3] (abab, abba) amplitudes of the current pass (zero where
                            // invalid); k_slice writes them from the history
  Ptr<double2> tk;          // [nvalid] the same amplitudes by K block: entry ci of block K's
                            // virtual list, pair p at kb_toff[K] + ci*noo + p
  Ptr<double2> histT;       // [ndiis][nvalid] updated amplitudes by compact index, slot-major
  Ptr<double2> histE;       // [ndiis][nvalid] DIIS error vectors, the same layout
  Ptr<double> gout;         // [2*n3] G of the last pass (when MF_WRITE_GR)
  Ptr<double> rout;         // [2*n3] R of the last pass (when MF_WRITE_GR)
  Ptr<double2> lad;         // [n3] particle-particle ladder, (S, P) parts
  Ptr<double2> hh;          // [nvalid] hole-hole ladder, (S, P) parts, laid out as tk
  Ptr<double2> hh2;         // the same over the second half of the W columns of blocks of
                            // more than 16 pairs (zeros at the other blocks' entries)
  Ptr<double2> jring;       // [o*v rows][o] quadratic ring intermediates (J2, J3), rows
                            // in the grouped P(q) order (qrow)
  Ptr<double2> jxring;      // [o*v rows][o] quadratic crossed-ring intermediates, same rows
  Ptr<double2> tv;          // [n3] slice products (TV_ab + VT_ab, TV_ba + VT_ba)
  Ptr<double4> quad;        // [n3] quadratic ring (ab, ba) and crossed-ring (ab, ba) sums
  Ptr<double> rsum;         // [v*o] slice row sums
  Ptr<double> csum;         // [v*o] slice column sums
  Ptr<double> jvv;          // [v] mosaic Jvv
  Ptr<double> joo_part;     // [v*o] per-slice parts of the mosaic Joo
  Ptr<double> joo;          // [kJooParts][o] mosaic Joo summed over each slice range
  Ptr<double> partials;     // [nblk_elem * kPartialStride]
  int nblk_elem;
  Ptr<State> st;            // the state between passes
  Ptr<State> st2;           // the state of the running pass: equal to st between passes;
                            // k_slice writes it when it runs the previous pass's reductions,
                            // k_contract and k_main read it, k_main copies it back to st
  Ptr<double> trace;        // [2*maxiter] (E, dT) by NIter
};
template <typename T> using HostPtr = T*;
using SolveDev = SolveDevT<HostPtr>;

// Doubles of the solve's distinct ladder tables (n2max + 1 each), rounded up
// to keep what follows them in LDS 16-byte aligned.
inline int ladder_tab_doubles(const SolveDev& S) { return (S.lad_ntab * (S.n2max + 1) + 1) & ~1; }
// Dynamic LDS of a k_contract block: the ladder tables, a K block's virtual
// list (at most cls_max entries) and the hole-hole unit's 16 W rows (at most
// noo_max pairs).
inline int contract_lds_bytes(int tab_doubles, int cls_max, int noo_max) {
  return tab_doubles * 8 + cls_max * 16 + 16 * noo_max * 16;
}
// Entries of a K block's virtual list held in LDS at once (a multiple of 16
// times every k-step group of the tile loops); cls_max never exceeds it.
constexpr int kClsChunk = 4096;

constexpr int kSliceClasses = 3;


// Tile choices of the tuned kernels: k-steps per group of loads in the ladder
// tiles (1, 2, 4) and per round of loads in the ring (4, 8) and hole-hole
// (4, 8) contraction tiles.
struct Tiles {
  int ladder_group = 2, quad_steps = 8, hh_steps = 8;
};

// Launch geometry of a batch of solves iterated together.
struct PlanArgs;

// A batch of one's pass-state arrays, handed to its slice launches as kernel
// arguments: the folded state update requests them without first reading
// the solve's descriptor.
struct FoldArgs {
  const double* partials;  // the element blocks' partials (slot-major)
  int nblk;                // element blocks
  int nd;                  // DIIS dots among the partials of an iterating pass
  const void* st;          // the state between passes (State)
};

struct BatchDev {
  int n;
  SolveDev* solves;     // device array [n]
  int* elem_prefix;     // [n+1] 256-thread blocks over each solve's valid elements
  // k_slice: one launch per block size (256, 512, 768 threads, class c =
  // slice_threads / 256 - 1) over the solves of that size; one block per
  // virtual slice, the other solves have no blocks in the class's prefix
  int* slice_prefix[kSliceClasses];  // [n+1] each
  // k_contract: hole-hole units, ladder tiles (one block each) and ring /
  // crossed-ring / Joo items (one wave each) of each solve; the unit and tile
  // ranges of the launch are padded to multiples of 8 blocks
  int* hh_prefix;       // [n+1]
  int nhh, nhh_pad;
  int hh_split;         // the units are each solve's hh_items2 (a batch of one), else hh_items
  FoldArgs fold;        // a batch of one's state-update arrays (others: unused)
  int* lad_prefix;      // [n+1]
  int nlad, nlad_pad;
  int* lad_prefix2;     // [n+1] the same over the two-tile items
  int nlad2;
  int tab_doubles;      // largest ladder_tab_doubles of the batch
  int cls_max;          // largest K-block virtual list of the batch
  int con_lds;          // dynamic LDS bytes of k_contract
  int slice_split;                   // some solve's slice roles use different tables
  int elem_blocks;
  int drpa_sum_blocks;               // k_drpa_sums of a batch of one: kDrpaJooParts*o + v + nqb
  int slice_blocks;                  // over all classes
  int slice_cls_blocks[kSliceClasses];
  int slice_lds[kSliceClasses];      // dynamic LDS bytes (largest NOcc of the class)
  int* con_prefix;      // [n+1]
  int ncon;
  Tiles tiles;          // the tuned kernels' tile choice for the launch
  int main_nq;          // k_main's partials width: 16 (DIIS lengths up to 12) or 32
  // the device-built maps of the solves (launch_plan): each solve's PlanArgs
  // and the block prefixes of the planning launches (box; block lists and
  // transfer partners; element map)
  const PlanArgs* plan;     // [n]
  int* plan_prefix[3];      // [n+1] each
  int plan_blocks[3];
};

// k_contract wave items: a 16x16 tile of the quadratic ring (crossed-ring)
// products of one transfer q, the mosaic Joo sums of a slice range.
enum ContractKind : int { CK_RING = 0, CK_JOO = 2, CK_XRING = 3 };
constexpr int kJooParts = 32;  // slice ranges the mosaic Joo sum is split over (one wave each)
constexpr int kDrpaJooParts = 8;  // slab ranges of the dRPA Joo sum (one wave each)
static_assert(kDrpaJooParts <= kJooParts, "the dRPA Joo parts fit the joo buffer");
constexpr int kItemsPerBlock = kElemThreads / 64;  // k_contract wave items per block

// Dynamic LDS of a k_slice block: T^ab, T^ba, V(k,j) slabs and four o-vectors;
// with split tables also the crossed-ring V(j,k) and two more o-vectors.
inline int slice_lds_bytes(int o, bool split = false) {
  return (3 * o * o + 4 * o) * 8 + (split ? (o * o + 2 * o) * 8 : 0);
}
// Threads of a k_slice block (one block per virtual slice, one slab entry
// and J output per thread): 256 per 256 of the o*o entries, at most 768
// (a larger slab is looped over).
inline int slice_threads(int o) {
  const int g = (o * o + 255) / 256;
  return 256 * (g < 1 ? 1 : (g > 3 ? 3 : g));
}
constexpr int kMaxDynLds = 160 * 1024;        // LDS per CU on gfx950
// Static LDS of a k_slice block of bt threads for the previous pass's
// reductions: the partial sums, the reduction scratch and a state copy.
constexpr int slice_fin_lds(int bt) {
  return (int)sizeof(double) * (kPartialStride + (bt / 64) * 32) + (int)sizeof(State);
}

// Kernel launchers (kernels.hip).
// drpa: the spin-summed amplitudes of the dRPA iteration (in the first
// component of tcur, the second zero) and its energy (1: Ec(dRPA), 2:
// Ec(SOSEX))
void launch_init(const BatchDev& b, hipStream_t s, int drpa = 0);
void launch_mp2(const BatchDev& b, hipStream_t s, int drpa = 0);
// One dRPA pass (source/dRPA.f90:80-96) of a batch-of-one: the row and
// transfer sums, the per-element update, the reductions.
void launch_drpa_start(const BatchDev& b, int64_t n3, hipStream_t s);
// A batch of one's amplitudes in and out of tcur through staged dense arrays.
void launch_place_amplitudes(const BatchDev& b, int64_t nvalid, const double* ab, const double* ba, hipStream_t s);
void launch_split_amplitudes(const BatchDev& b, int64_t n3, double* ab, double* ba, double* aa, hipStream_t s);
void launch_sf(const BatchDev& b, int v, int nq, hipStream_t s);
void launch_drpa_pass(const BatchDev& b, hipStream_t s);
// Events timing one kernel of a pass (which: as launch_one, or 4 for the
// reductions and DIIS solve): `after` is written when it completes, `before`
// when the kernel launched just before it completes.  When `which` is the
// first kernel of the pass, `before` belongs to the previous pass: pass that
// pass `tail_stop` (the event is written when its last kernel completes).
struct KernelTimer {
  int which;
  hipEvent_t before, after;
};
// A chunk of `passes` passes of a batch of one as one cooperative,
// grid-synchronised launch (dyn_lds: the larger of the slice slabs and the
// ladder tables).  Returns nonzero when it cannot run (split slice tables,
// launch refused).
int launch_persistent(const BatchDev& b, int dyn_lds, int passes, hipStream_t s);
// One-time per-kernel set-up (dynamic LDS limits), done before any stream
// capture.
void prepare_launches();
// Dynamic LDS every k_contract instance can take (the LDS of a CU less the
// largest static LDS among them; 0 when the runtime cannot say).
int contract_lds_limit();
// One residual pass: k_slice, k_contract, k_main.  fin_prev: the pass's
// k_slice first runs the reductions and state update of the pass before it
// (every block the same; one writes the state), so that only the last pass
// of a chunk needs a launch_finalize(FIN_ITER) of its own.
void launch_iteration(const BatchDev& b, int main_flags, hipStream_t s,
                      const KernelTimer* timer = nullptr, hipEvent_t tail_stop = nullptr, int fin_prev = 0);
int first_pass_kernel(const BatchDev& b, int main_flags);
// FIN_INIT starts a fresh state (every field; prec64 the starting precision,
// or each solve's own prec0 when negative).
void launch_finalize(const BatchDev& b, int kind, hipStream_t s, int prec64 = 1);
void launch_state_reset(const BatchDev& b, int phase, int prec64, hipStream_t s);
void launch_gather_states(const BatchDev& b, State* out, hipStream_t s);
// One kernel of the pass: 0 slice, 1 contract, 3 main (2 is no longer a
// kernel of its own: the ladder tiles run in k_contract).
void launch_one(const BatchDev& b, int which, hipStream_t s);

// Maps of a context built on the device at placement, after its work buffers
// are cleared (the reference finds every momentum partner with FindIndex,
// source/HEG.f90:357-404; here one momentum box answers all of them):
//   box  [w^3]     orbital p + 1 at box index center + n.x + w n.y + w^2 n.z
//                  of every plane wave of the basis, 0 elsewhere (zero on entry)
//   civ  [v * nK]  entry of virtual a in block K's virtual list, or -1 (scratch)
// and, into the descriptor's arrays: each pair-momentum block's virtual list
// (kb_va, kb_vb, kb_nva), the element map cmap of the valid elements in
// element order (slab_coff gives each slice's first compact index), and the
// partner slices qb_d of the transfer sets.
struct PlanArgs {
  const SolveDev* sd;  // the context's device descriptor
  int w, center, coren;
  int* box;
  int* civ;
};
// Blocks of the three planning launches for one solve.
inline int plan_box_blocks(int nao) { return (nao + 255) / 256; }
inline int plan_list_blocks(int nK, int nq, int o) { return nK + (nq * o + 255) / 256; }
// The maps of every solve of a batch, three launches.
void launch_plan(const BatchDev& b, hipStream_t s);

// Copies (src non-null) or clears (src null) `bytes` (a multiple of 16) at
// dst, for n such jobs in one launch: the admissions of a server step (each
// context's constant image and the clear of its work buffers).
struct CopyJob {
  char* dst;
  const char* src;
  size_t bytes;
};
void launch_scatter(const CopyJob* jobs, int n, hipStream_t s);

// Clears `bytes` (a multiple of 8) of device memory at p with a fill kernel
// (no memset node: usable where a caller captures the stream).
void launch_zero(void* p, size_t bytes, hipStream_t s);

// MP2 connectivity histogram of a batch-of-one solve (nbins <= 4096 bins,
// bin k counts |dn|^2 = k+1); hist and overflow zeroed by the caller.
int connectivity_on_device(const BatchDev& b, int nvalid, int nbins, int* hist, int* overflow,
                           hipStream_t s);

// Plane-wave basis on the device: keys of the cube [-cell, cell]^3 (npad a
// power of two >= (2 cell + 1)^3, at most kMaxBasisPad) sorted with their
// enumeration index, the count of plane waves inside the cutoff; then the nb
// orbitals (n-vectors column-major, eigenvalues) and E_HF.
constexpr int kMaxBasisPad = 8192;
int basis_on_device(int cell, int maxk, int twisted, const double* twist, int npad, double* key, int* idx,
                    int* count, hipStream_t s);
int basis_orbitals_on_device(int cell, int nb, int no, int mc, int nel, const double* twist,
                             double L1, double rescale, double omega, double mad, double gap,
                             double xf, const int* idx, int* nvec, double* eigen, double* ehf,
                             hipStream_t s);

}  // namespace uegccd
