// common.hpp -- shared host/device definitions of libftlm_hip.so (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <complex>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "ftlm_hip.h"

namespace ftlm_hip {

// ------------------------------------------------------------------ errors
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define FTLM_HIP_CHECK(expr)                                                                    \
  do {                                                                                          \
    hipError_t _e = (expr);                                                                     \
    if (_e != hipSuccess)                                                                       \
      throw ::ftlm_hip::Error(FTLM_EHIP, std::string(#expr " failed: ") + hipGetErrorString(_e) + \
                                             " (" __FILE__ ":" + std::to_string(__LINE__) + ")"); \
  } while (0)

#define FTLM_REQUIRE(cond, code, msg)                       \
  do {                                                      \
    if (!(cond)) throw ::ftlm_hip::Error((code), (msg));    \
  } while (0)

// Runs `body`, converting exceptions into status codes + last-error text.
template <class F>
int guarded(F&& body) {
  try {
    body();
    return FTLM_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("out of host memory");
    return FTLM_ENOMEM;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return FTLM_EHIP;
  }
}

// ------------------------------------------------------------------ device buffers
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      throw Error(e == hipErrorOutOfMemory ? FTLM_ENOMEM : FTLM_EHIP,
                  "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed: " + hipGetErrorString(e));
    }
    n = count;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void upload(const T* h, size_t count, hipStream_t s = nullptr) {
    if (count) FTLM_HIP_CHECK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T>& v, hipStream_t s = nullptr) { upload(v.data(), v.size(), s); }
};

// ------------------------------------------------------------------ combinatorics
constexpr int kMaxSites = 32;
// Pull-hop entry (uint32): bits 0..19 source string rank, 20..28 bond id, 29 own-spin JW parity,
// 30..31 zero.  entry >> kTargetBits is therefore a 10-bit index (parity * 512 + bond) into a
// table of pre-signed coefficients.  Bond kZeroBond has coefficient 0: slot-major tables pad
// every string's hop list with kPadEntry up to a multiple of kSlotPad slots, so an unrolled
// group of hops never needs a bounds test (a padded hop reads string 0 and adds 0 * x).
constexpr int kTargetBits = 20;
constexpr uint32_t kTargetMask = (1u << kTargetBits) - 1u;
constexpr int kBondBits = 9;
constexpr uint32_t kBondMask = (1u << kBondBits) - 1u;
constexpr int kParityBit = kTargetBits + kBondBits;  // 29
constexpr uint32_t kZeroBond = kBondMask;             // 511
constexpr uint32_t kPadEntry = kZeroBond << kTargetBits;
constexpr int kCoefSlots = 2 << kBondBits;            // 1024 pre-signed coefficients
#ifndef FTLM_SLOT_PAD
#define FTLM_SLOT_PAD 8
#endif
constexpr int kSlotPad = FTLM_SLOT_PAD;

// Binomial table C[a][b], a,b <= 32 (values < 2^63).
struct Binom {
  int64_t c[kMaxSites + 2][kMaxSites + 2];
};

Binom host_binomials();
int64_t binom(int a, int b);

// ------------------------------------------------------------------ operator (sector / k-block)
enum OpKind { kSector = 0, kKBlock = 1, kCallback = 2, kGram = 3, kKStack = 4 };

// Step tables of the LDS-staged H*V passes (elements of 128 bytes: 16 complex64 or 8 complex128
// vectors).  One pass handles one spin's hops along lines of D elements that share the other
// spin's string.  A tile is (line, block of kStepBlock consecutive targets); its 1024 threads own
// the block's sums in registers: thread tid serves target j * 256 + tid / 4 in round j
// (kStepRounds rounds) and bytes (tid % 4) * 32 .. +32 of the element, so wavefront w serves the
// group of kStepGroup targets j * 256 + w * 16 + [0, 16) in round j.  A block's hops are cut into
// steps; a step covers sources [xoff, xoff + nx) (nx <= kStepSub) and, for each wavefront, at most
// kStepRows rows of 16 entries: for round j, K_j rows (K_j = the longest hop list of the round's
// group inside the step's source range), row k holding hop k of each of the 16 targets, padded
// with zero-coefficient entries.  Each target's hops appear in increasing source order across
// steps and rows.  Entry (uint32): bits 0..15 the source's byte offset in the step's LDS copy,
// (source - xoff) * 128, bits 16..31 8 * c for its coefficient c (own-spin parity * 256 +
// bond; kStepZero is the zero coefficient): the byte offset in a table of 8-byte complex64
// slots, half of it in a table of 16-byte complex128 slots.
// Arrays: blk[b] .. blk[b + 1] = the block's steps; desc[step * 16 + w] = {xoff | nx << 20,
// first row | rows << 26, K_0..K_3 (bytes), K_4..K_7}; ent = rows of 16 entries.
constexpr int kStepBlock = 2048;
constexpr int kStepRounds = 8;
constexpr int kStepWaves = 16;
constexpr int kStepGroup = 16;
constexpr int kStepSub = 320;
constexpr int kStepRows = 32;
constexpr int kStepPosBits = 16;
constexpr uint32_t kStepZero = 255;
constexpr int kStepCoefSlots = 512;
constexpr int kStepMaxBonds = 255;  // bonds per spin the step tables encode (ids 0..254)
constexpr int kStepThreads = kStepWaves * 64;         // threads of a step workgroup
constexpr int kStepRound = kStepBlock / kStepRounds;  // targets per round (4 lanes each)
constexpr int kStepWgPerCu = 160 * 1024 / ((kStepSub * 128 + kStepWaves * kStepRows * 64) * 2 + 8192 + 64);
static_assert(kStepRound == kStepWaves * kStepGroup && kStepRound * 4 == kStepThreads, "step tile shape");
static_assert(kStepRows * 4 <= 128 && kStepWgPerCu >= 1, "step buffers");

// Device view of a full (N_up, N_dn) sector with its per-spin pull-hop tables (entry layout above:
// bits 0..19 source string rank, 20..28 bond id, 29 own-spin JW parity).
struct SectorView {
  int64_t dim;
  int Dup, Ddn;
  int n_sites;
  double U, V;
  const uint32_t* up_mask;   // [Dup]
  const uint32_t* dn_mask;   // [Ddn]
  const int32_t* off_up;     // [Dup + 1]
  const uint32_t* ent_up;    // [nnz_up]
  const int32_t* off_dn;     // [Ddn + 1]
  const uint32_t* ent_dn;    // [nnz_dn]
  const double2* coef_up;    // [nbond_up] summed directed-bond coefficient
  const double2* coef_dn;    // [nbond_dn]
  const uint32_t* xmask_up;  // [nbond_up] down-string sites inside the JW string of an up bond
  const uint32_t* xmask_dn;  // [nbond_dn] up-string sites inside the JW string of a down bond
  const int32_t* nuu;        // [Dup]  sum over NN pairs of u_i u_j
  const int32_t* ndd;        // [Ddn]  sum over NN pairs of d_i d_j
  const uint4* planes;       // [Dup]  bit planes of nb_u(s) = sum of u over pair partners of s
  // slot-major copies of the hop tables: entT[k * stride + i] is entry k of string i
  // (k < cnt[i], kPadEntry beyond; stride = D rounded up to 256), so slot k of consecutive
  // strings is one contiguous run; maxcnt_* = slots per table (max count rounded up to kSlotPad)
  const uint32_t* entT_up;   // [maxcnt_up][stride_up]
  const int32_t* cnt_up;     // [Dup]
  const uint32_t* entT_dn;   // [maxcnt_dn][stride_dn]
  const int32_t* cnt_dn;     // [Ddn]
  int nbond_up, nbond_dn;
  int entT_stride_up, entT_stride_dn;
  int maxcnt_up, maxcnt_dn;
  // step tables (kStepBlock layout above); null when not built
  const int32_t* st_blk_up;  // [ceil(Dup / kStepBlock) + 1]
  const uint4* st_desc_up;   // [steps][kStepWaves]
  const uint32_t* st_ent_up;
  const int32_t* st_blk_dn;  // [ceil(Ddn / kStepBlock) + 1]
  const uint4* st_desc_dn;
  const uint32_t* st_ent_dn;
  // row step pass over caller-described lines (null for the sector's own rows): line l's element
  // p at (line_base[l] + p) * RT, its up string line_umask[l]
  const int64_t* line_base;
  const uint32_t* line_umask;
};

constexpr int kMaxBondsLds = 512;  // LDS slots per spin for bond data (ids < kZeroBond are bonds)

// Device view of an orbit-representative momentum block (filled by orbit.hip).
struct KBlockView;

// Launch shape of an operator's H*V kernels for one (R, dtype): the vectors per launch (rb), the
// workgroups per XCD of each two-pass sweep (bpx), the column pass's stripe width log2 (wlog) and
// the grid of the single-pass kernels (grid).  0 (wlog: -1) keeps the library's default.
struct HvShape {
  int rb = 0, bpx = 0, wlog = -1, grid = 0;
};

// Staging of the host-pointer apply (ftlm_hip_apply_host), per operator: device x / y blocks,
// two pinned chunks, a stream and two events, made on the first host apply and kept with the
// operator, so a call allocates nothing.  Calls on one operator are serialised by `mu`.
struct HostStage {
  std::mutex mu;
  DevBuf<double2> dx, dy;
  double2* pin[2] = {nullptr, nullptr};
  size_t pin_n = 0;  // complex elements per pinned chunk
  hipStream_t st = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  HostStage() = default;
  HostStage(const HostStage&) = delete;
  HostStage& operator=(const HostStage&) = delete;
  ~HostStage() {
    if (st) (void)hipStreamSynchronize(st);
    for (int i = 0; i < 2; ++i) {
      if (pin[i]) (void)hipHostFree(pin[i]);
      if (ev[i]) (void)hipEventDestroy(ev[i]);
    }
    if (st) (void)hipStreamDestroy(st);
  }
};

struct OpBase {
  int kind = kSector;
  int device = 0;
  int64_t dim = 0;
  std::map<int, HvShape> shapes;  // set by the autotuner, key R * 4 + dtype
  HvShape shape(int R, int dtype) const {
    const auto it = shapes.find(R * 4 + dtype);
    return it == shapes.end() ? HvShape{} : it->second;
  }
  mutable std::mutex stage_mu;                    // guards the creation of host_stage
  mutable std::unique_ptr<HostStage> host_stage;  // ftlm_hip_apply_host staging
  virtual ~OpBase() = default;
};

}  // namespace ftlm_hip

// The opaque C handle.
struct ftlm_op : ftlm_hip::OpBase {};

namespace ftlm_hip {

struct Sector : ftlm_op {
  int n_sites = 0, n_up = 0, n_dn = 0;
  int Dup = 0, Ddn = 0;
  double U = 0.0, V = 0.0;
  int nbond[2] = {0, 0};
  int64_t nnz[2] = {0, 0};
  bool real_h = true;  // every summed bond coefficient has zero imaginary part
  // strings in increasing mask order instead of the reference's lexicographic order (internal
  // sectors of momentum stacks only: their index_of and external interfaces are not used)
  bool numeric_order = false;
  DevBuf<uint32_t> up_mask, dn_mask;
  DevBuf<int32_t> off[2];
  DevBuf<uint32_t> ent[2];
  DevBuf<double2> coef[2];
  DevBuf<uint32_t> xmask[2];
  DevBuf<int32_t> nuu, ndd;
  DevBuf<uint4> planes;
  DevBuf<uint32_t> entT[2];
  DevBuf<int32_t> cnt[2];
  int maxcnt[2] = {0, 0};
  int entT_stride[2] = {0, 0};
  DevBuf<int32_t> st_blk[2];
  DevBuf<uint4> st_desc[2];
  DevBuf<uint32_t> st_ent[2];
  // per-XCD tile counters of the two H*V passes, one slice of kCtrSlice ints per stream that
  // applies the operator ([0..8] rows pass, [16..24] column pass; the last workgroup of a pass
  // returns its counters to zero).  Launches on one stream are ordered, so a slice is never used
  // by two launches at once; applies from different streams may overlap.
  static constexpr int kCtrSlice = 32, kCtrStreams = 64;
  DevBuf<int32_t> tile_ctr;  // [kCtrStreams][kCtrSlice]
  mutable std::mutex ctr_mu;
  mutable std::vector<const void*> ctr_stream;  // stream owning each slice, in order of first use
  // The counter slice of stream st (assigned on its first apply); throws FTLM_EUNSUPPORTED once
  // kCtrStreams distinct streams hold slices.
  int32_t* tile_counters(hipStream_t st) const;
  // host copies: directed bond (from, to) and summed coefficient per spin, in bond-id order, and
  // the NN pair list as sorted (min, max) sites (term-split derivative tables are built from them)
  std::vector<std::pair<int, int>> bond_key[2];
  std::vector<std::complex<double>> bond_coef[2];
  std::vector<std::pair<int, int>> pair_list;
  SectorView view() const;
};

// Interleaved complex element types of the two vector dtypes.
template <typename T>
struct Cx;
template <>
struct Cx<double> {
  using v2 = double2;
};
template <>
struct Cx<float> {
  using v2 = float2;
};

// ------------------------------------------------------------------ launches
inline unsigned grid_for(int64_t n, int block) { return static_cast<unsigned>((n + block - 1) / block); }

void check_launch(const char* what);

// Block-vector H*V dispatch (hv.hip).  xs: optional per-vector input scale [R] on the device
// (y = H (xs .* x)); dotp: optional partial sums of <xs.*x | y> per block, layout
// [nblocks][R] double2; returns the number of partial blocks written.
// force_real: use only the real parts of the bond coefficients (the reference's
// ftlm_log_partition_real applies Re(H) to the real and imaginary parts separately).
int hv_apply(const ftlm_op* op, const void* x, void* y, int R, int dtype, const double* xs, double2* dotp,
             hipStream_t s, bool force_real = false);
int hv_partial_blocks(const ftlm_op* op);
// Sector indices of packed states (u << 16) | d, all in the sector (sector.hip).
void sector_rank_packs(const Sector* S, const uint32_t* d_packs, int64_t count, int32_t* d_out, hipStream_t st);
// Grid of a single-pass apply (small sector or momentum block) with the default shape, and the
// largest grid a shape may ask for.
int hv_default_grid(const ftlm_op* op, int* grid_max);
// Row pass (down bonds only: y = part(S_down) of the two-pass order) over nlines caller-described
// lines of D_dn elements each: line l's element p at x + (lbase[l] + p) * R, with up string
// lumask[l], for all R vectors (the step passes for 128-byte chunks, the L2-gather pass
// otherwise, which add in the same order).  Needs D_dn >= 128 and real bond coefficients (hv.hip).
bool sector_has_step_tables(const Sector* S);
void hv_rows_lines(const Sector* S, const void* x, void* y, int R, int dtype, const double* xs, const int64_t* lbase,
                   const uint32_t* lumask, int nlines, hipStream_t st);
// A sector whose strings are ranked in increasing mask order (sector.hip; see Sector::numeric_order).
std::unique_ptr<Sector> sector_create_numeric(int n_sites, int n_up, int n_dn, const ftlm_hop_t* hops, int n_hops,
                                              const ftlm_pair_t* pairs, int n_pairs, double U, double V);
// Default stripe-width log2 of a two-pass sector's column pass for a launch of rb vectors (-1 for
// a single-pass sector), and the largest one the tile allows.
int hv_default_wlog(const ftlm_op* op, int R, int dtype, int rb, int* wlog_max);
// Momentum stacks (orbit.hip): the momentum blocks of one orbit graph side by side, one vector
// slot per (block j, sample r) over every orbit of the graph, layout [orbit][j][rk].  Slot (j, r) of
// orbit g holds block j's row map_j[g], or 0 where g is not a row of block j.
constexpr int kMaxStackBlocks = 16;
struct KStackMaps {
  const int32_t* map[kMaxStackBlocks];  // orbit -> row of block j, or -1
};
// The orbit graph a momentum block uses (blocks created with share= return the same pointer);
// nullptr for other operators.
const void* kblock_graph(const ftlm_op* op);
// A stack over blocks (all of one orbit graph and device, at most kMaxStackBlocks); the blocks
// must outlive it.
std::unique_ptr<ftlm_op> kstack_build(const std::vector<const ftlm_op*>& blocks);
int kstack_blocks(const ftlm_op* stack);
KStackMaps kstack_maps(const ftlm_op* stack);
// Random-phase start of a stack (lanczos.hip): slot (j, r) of orbit g gets the value the per-block
// start gives block j's row map_j[g] for sample id sample_ids[r] (0 off the block).
void kstack_phase_start(const ftlm_op* stack, int rk, uint64_t seed, const uint32_t* sample_ids, int dtype, void* q,
                        hipStream_t st);
// True when a momentum-block operator is Hermitian (translation-invariant hop coefficients).
bool kblock_hermitian(const ftlm_op* op);

// Derivative operators of a momentum block as term-split data (orbit.hip): up to four hop tables
// over the block's own bond ids ([j][512] complex) and up to four diagonal arrays (each
// derivative block's own orbit diagonal); tab[p] / dg[p] name parameter p's table / array or -1.
struct KbTermSplit {
  int ntab = 0, ndg = 0;
  std::vector<int> tab, dg;
  const double* diag[4] = {nullptr, nullptr, nullptr, nullptr};
  DevBuf<double2> dc;
};
bool kblock_termsplit(const ftlm_op* op, ftlm_op* const* dops, int npar, KbTermSplit* T);
// Block partial sums [blocks][8] of the term-split bilinear forms Re <a| . |b> (R vectors);
// kblock_bilinear_blocks gives the number of blocks.
int kblock_bilinear_blocks(const ftlm_op* op, int R);
void kblock_bilinear(const ftlm_op* op, const KbTermSplit& T, const double2* a, const double2* b, int R,
                     double2* part, hipStream_t st);

}  // namespace ftlm_hip
