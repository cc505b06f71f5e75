// Host-side mirror of the reference's two-site pivot (PivotMatrix2 + PivotVector
// layout) and the flattening of its block structure into the staged
// block-triple tables that the device engine executes.
//
// Reference objects mirrored here:
//   Tripod (Multipede<3>)   tensors/Multipede.h:31-160   index[q] = {in,out,mid}, block[q][a][0]
//   Biped                   tensors/Biped.h:62-104       in[q], out[q], block[q], dict{in,out}
//   PivotMatrix2            pivot/DmrgPivotMatrix2.h:15-76 (L, R, W12, W34, qloc, qOp, qlhs/qrhs/factor_cgcs)
//   PivotMatrix1            pivot/DmrgPivotMatrix1.h:11-66 (one-site: L, R, W, qloc, qOp)
//   PivotMatrix0            pivot/DmrgPivotMatrix0.h:10-27 (zero-site: L, R)
//   PivotVector             pivot/DmrgPivotVector.h:8-117  data[s] = Biped of dense col-major blocks
#pragma once
#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <unordered_map>
#include <memory>
#include <vector>

#include "sym.hpp"

namespace vmps {

struct DenseBlock {
  int64_t rows = 0, cols = 0;
  int64_t off = -1;  // element offset into the owning packed array; -1 = absent
  bool present() const { return rows > 0 && cols > 0 && off >= 0; }
  int64_t size() const { return rows * cols; }
};

struct Key2 {
  Qn a, b;
  bool operator==(const Key2& o) const { return a == o.a && b == o.b; }
};
struct Key2Hash {
  size_t operator()(const Key2& k) const { return QnHash()(k.a) * 1000003u ^ QnHash()(k.b); }
};
struct Key3 {
  Qn a, b, c;
  bool operator==(const Key3& o) const { return a == o.a && b == o.b && c == o.c; }
};
struct Key3Hash {
  size_t operator()(const Key3& k) const {
    return (QnHash()(k.a) * 1000003u ^ QnHash()(k.b)) * 1000033u ^ QnHash()(k.c);
  }
};

// Environment tensor (Multipede<3,...> with block[q][a][0]).
// Completion of queued device work that writes a device-resident
// environment (an environment update returned before its kernels finished):
// wait() returns once that work is done and releases what it held.  Safe to
// call from several threads and more than once.
struct DevReady {
  virtual ~DevReady() = default;
  virtual void wait() = 0;
};

struct Tripod {
  std::vector<std::array<Qn, 3>> index;          // {in, out, mid}
  std::vector<std::vector<DenseBlock>> block;    // [q][a]
  std::vector<double> data;                      // packed FP64 blocks
  // device-resident environment: the blocks' offsets refer to `dev` (FP64
  // on device dev_device) and `data` stays empty.  dev_hold shares the
  // ownership of that allocation: it stays valid while any copy of the
  // Tripod (in a vmps_denv, a pivot or a contraction) exists.
  const double* dev = nullptr;
  std::shared_ptr<const double> dev_hold;
  // set while the blocks at `dev` may still be being written: every reader of
  // `dev` calls wait_device() first (declared after dev_hold, so that a last
  // owner waits before it frees the blocks)
  std::shared_ptr<DevReady> dev_ready;
  int64_t dev_size = 0;
  int dev_device = -1;
  void wait_device() const {
    if (dev_ready) dev_ready->wait();
  }
  std::unordered_map<Key3, size_t, Key3Hash> dict;
  size_t dim() const { return index.size(); }
  int64_t size() const { return dev ? dev_size : (int64_t)data.size(); }  // packed elements
  const Qn& in(size_t q) const { return index[q][0]; }
  const Qn& out(size_t q) const { return index[q][1]; }
  const Qn& mid(size_t q) const { return index[q][2]; }
};

// One sparse MPO block, nonzeros stored in column-major order as Eigen's
// InnerIterator visits them (outer = column, rows ascending).
struct SparseBlock {
  int rows = 0, cols = 0;
  std::vector<int> row, col;
  std::vector<double> val;
};

struct WBiped {
  std::vector<Qn> in, out;
  std::vector<SparseBlock> block;
  std::unordered_map<Key2, size_t, Key2Hash> dict;
};

// Layout of a PivotVector: data[s] is a Biped of dense blocks; all blocks
// packed into one array in (s, q) order.
struct VLayout {
  std::vector<std::vector<Qn>> in, out;                   // [s][q]
  std::vector<std::vector<DenseBlock>> block;             // [s][q]
  std::vector<std::unordered_map<Key2, size_t, Key2Hash>> dict;  // [s]
  int64_t size = 0;
  int64_t nblocks = 0;
};

// Two-site combined local basis (Qbasis::combine of loc12 and loc34,
// tensors/Qbasis.h:686-798, indexing via outer_num/leftOffset :378-462).
struct TwoSiteBasis {
  std::vector<Qn> qloc;  // quantum number of each combined state s
  // s index of (s1, s3, qmerge); -1 if qmerge not in qloc12[s1] (x) qloc34[s3]
  int64_t index(int s1, int s3, const Qn& qmerge) const;
  void build(const Sym& sym, const std::vector<Qn>& qloc12, const std::vector<Qn>& qloc34);

 private:
  struct Src {
    Qn q1, q2;
    int64_t dim1, dim2;
  };
  std::map<Qn, int64_t> outer_;                    // qmerge -> outer_num
  std::map<Qn, std::vector<Src>> history_;         // sorted by (q1,q2)
  std::vector<Qn> q12_, q34_;
  std::vector<int64_t> inner12_, inner34_;         // inner_num of each local state
  std::map<Qn, int64_t> plain12_, plain34_;        // plain size of each local sector
};

// One summand of a pivot: PivotMatrix2Terms (DmrgPivotMatrix2.h:15-26) or
// PivotMatrix1Terms (DmrgPivotMatrix1.h:12-22; W, qOp in the "12" slots).
struct PivotTerm {
  Tripod L, R;
  std::vector<Qn> qop12, qop34;
  std::vector<std::vector<std::vector<WBiped>>> W12, W34;  // [s1][s2][k]
  // qlhs / qrhs / factor_cgcs (DmrgPivotMatrix2.h:66-70, DmrgPivotMatrix1.h:58-62).
  // Two-site rows: {s2s4,q24,qL,qR,s1,s2,k12,qW12,s3,s4,k34,qW34}; one-site
  // rows use the first six entries {s2,q2,qL,qR,k,qW}, the rest are 0.
  std::vector<std::array<size_t, 2>> qlhs;
  std::vector<std::vector<std::array<size_t, 12>>> qrhs;
  std::vector<std::vector<double>> factor_cgcs;
  bool has_blockstructure = false;
};

// A pivot of the zero-, one- or two-site effective Hamiltonian
// (PivotMatrix0 DmrgPivotMatrix0.h:10-27, PivotMatrix1 DmrgPivotMatrix1.h:24-66,
// PivotMatrix2 DmrgPivotMatrix2.h:28-76): a sum over terms sharing one
// wavefunction layout.
struct Pivot {
  int nsite = 2;
  Sym sym;
  std::vector<Qn> qloc12, qloc34;  // one-site: qloc12 is the local basis
  std::vector<PivotTerm> terms;    // at least one
  size_t cur = 0;                  // term the builder calls fill
  VLayout V;

  // Excited-state projection (DmrgPivotMatrix2.h:72-75, :249-272): packed
  // vectors in the V layout.
  std::vector<std::vector<double>> A0proj;
  double Epenalty = 0.0;

  PivotTerm& term() { return terms[cur]; }
  const PivotTerm& term() const { return terms[cur]; }
  // new empty term with the given operator bases; becomes the current one
  size_t add_term(const std::vector<Qn>& qop12, const std::vector<Qn>& qop34);
  int add_V_block(int s, const Qn& qin, const Qn& qout, int64_t rows, int64_t cols);
  void invalidate_blockstructure() {
    for (auto& t : terms) t.has_blockstructure = false;
  }
};

// Restatement of precalc_blockStructure for one term:
//   two-site: DmrgIndexGymnastics.h:698-856 with the AAWWAA matcher (:198-262);
//   one-site: DmrgIndexGymnastics.h:485-590 with the LAWA matcher (:32-77);
//   zero-site: no table (PivotMatrix0's HxV matches sectors directly).
// The reference returns entries in unordered_map order; here the keys are
// emitted sorted, entries under one key keep the reference's loop order.
void precalc_blockStructure(Pivot& P, PivotTerm& T);
void precalc_blockStructure(Pivot& P);  // every term

// ---- staged block-triple plan --------------------------------------------
// Terms tau = (out block o, L block l, V block v, R block r, coefficient c)
// from the OxV loop (DmrgPivotMatrix2.h:182-230), regrouped as
//   T_{l,v} = L_l . V_v              (distinct (l,v))
//   U_{o,r} = sum_tau c_tau T_{l,v}  (distinct (o,r))
//   Vout_o  = sum_r U_{o,r} . R_r
struct GemmProb {     // C(M x N, ldc=M) = sum over chunks A(M x K) . B(K x N)
  int64_t c_off;
  int32_t M, N;
  int32_t chunk_begin, chunk_end;
};
struct GemmChunk {
  int64_t a_off, b_off;
  int32_t K, lda, ldb;
  int32_t pad;
};
struct CombineList {  // U_u = sum_j c_j T_{t_j}
  int64_t u_off;
  int64_t size;
  int32_t begin, end;
};
struct Stats {
  int64_t n_terms = 0, n_T = 0, n_U = 0, n_out = 0;
  int64_t n_L = 0, n_R = 0, n_V = 0;
  double flops_alg = 0, flops_ref = 0;
  double flops_A = 0, flops_C = 0;  // stage A (L.V) and stage C (U.R) shares of flops_alg
  double bytes_alg_fp64 = 0;  // 8 * (distinct L + distinct R + Vin + Vout elements)
  int64_t L_elems = 0, R_elems = 0, V_elems = 0, T_elems = 0, U_elems = 0;
  int64_t max_block_dim = 0;
};

// Flattened per-tile descriptors of the device kernels (one load per tile).
constexpr int kTile = 32;  // output tile edge of one wavefront (split-BF16 wave kernels)
// The FP64 / FP32 wave kernels use 16 x 16 tiles (one MFMA accumulator) when
// every block is at most this large, so the MFMA work of the few big blocks
// of a small-block pivot spreads over 4x more SIMDs; 32 x 32 tiles otherwise.
constexpr int kSmallTileMaxDim = 64;
constexpr int kMaxTermsPerChunk = 64;  // stage C loads a chunk's term list with one load per lane
struct TileA {             // T(M x N, ld M) tile = L(M x K, ld M) . V(K x N, ld K)
  int64_t a_off, b_off, c_off;
  int32_t M, N, K;
  int16_t mt, nt;          // tile coordinates (units of kTile)
};
struct TileC {             // Vout(M x N, ld M) tile = sum over chunks U . R
  int64_t c_off;
  int32_t M, N;
  int32_t chunk_begin, chunk_end;
  int16_t mt, nt;
  int32_t pad;
};
struct ChunkC {            // one (U_u, R_r) pair: U_u = sum_j c_j T_j (M x K), R_r (K x N, ld K)
  int64_t r_off;
  int32_t K;
  int32_t term_begin, term_end;
  int32_t pad;
};
struct TermC {             // c_j T_j with T_j an M x K block at t_off
  int64_t t_off;
  double c;
};

// Split-BF16 layouts: k-contiguous operands with the k extent padded to a
// multiple of 8 (so an MFMA lane's 8 consecutive k are one aligned 16-byte
// load and the padding supplies the zeros past K).
inline int64_t pad8(int64_t x) { return (x + 7) & ~int64_t(7); }
// T and U blocks start at multiples of kElemAlign elements (16 bytes in FP32,
// 32 in FP64), so the stage B kernel can move them 16 bytes per lane.
constexpr int64_t kElemAlign = 4;
inline int64_t align_elems(int64_t x) { return (x + kElemAlign - 1) / kElemAlign * kElemAlign; }
struct TileA16 {           // T(M x Np, row-major) tile = Lt(M x Kp, row-major) . V(Kp x N, col-major)
  int64_t a_off, b_off, c_off;
  int32_t M, N, Kp, Np;
  int16_t mt, nt;
  int32_t pad;
};
struct ChunkC16 {          // (U_u, R_r): R padded col-major (Kp x N), T terms row-major (M x Kp)
  int64_t r_off;
  int32_t Kp;
  int32_t term_begin, term_end;
  int32_t pad;
};
struct TermC16 {
  int64_t t_off;
  float c;
  int32_t pad;
};
// Explicit stage B: U_u = sum_j c_j T_j formed by its own kernel.  A task
// covers elements [e0, e0 + n) of one U block (same shape and layout as its
// T blocks), reading terms [term_begin, term_end) of the term table.
struct CombTask {
  int64_t u_off;   // first element of the U block
  int64_t e0;      // first element of the task within the block
  int32_t n;       // elements of the task
  int32_t term_begin, term_end;
  int32_t pad;
};
constexpr int kCombTaskElems = 4096;
struct VConv {             // V block: packed FP64 (K x N, ld K) -> padded split BF16 (Kp x N)
  int64_t src_off, dst_off;
  int32_t K, N, Kp, pad;
};
constexpr int kBigTile = 128;  // workgroup tile edge of the LDS-staged kernels
constexpr int kBigK = 64;      // k depth of one LDS stage of those kernels
// A block is handled by the LDS-staged big-tile kernels when both its
// dimensions are at least this (below, the wave-tile kernels waste less).
constexpr int kBigMin = 96;
struct Bf16Layout {
  int64_t L_size = 0, R_size = 0;            // sizes of the split L / R operands
  int64_t V_size = 0, T_size = 0;            // padded sizes of the per-call arrays
  int64_t U_size = 0;                        // padded U (explicit stage B only)
  std::vector<CombTask> comb_tasks;          // explicit stage B on the padded layouts
  std::vector<TermC16> comb_terms;
  std::vector<TileA16> tilesA_big;           // 128 x 128 tiles of the large stage A problems
  std::vector<TileC> tilesC_big, tilesC_small;  // stage C tiles: 128-unit and 32-unit
  // split-K work items of the big stage C tiles: tile index, iteration range
  // [it_begin, it_end) over the flattened (chunk, 32-deep k block) loop, and
  // the partial-sum slot (-1 = write the output directly)
  std::vector<int32_t> splitC_tile, splitC_begin, splitC_end;
  std::vector<int64_t> splitC_slot;
  std::vector<int32_t> redC_tile;     // big tiles reduced from partials
  std::vector<int64_t> redC_slot;     // first partial slot of each reduced tile
  std::vector<int32_t> redC_nsplit;
  int64_t partial_slots = 0;          // number of 128 x 128 FP32 partial tiles
  std::vector<TileA16> tilesA;
  std::vector<ChunkC16> chunksC;
  std::vector<TermC16> termsC;
  std::vector<VConv> vconv;
  // device conversion of the FP64 environment blocks into the split
  // operands: R blocks exactly as V (K x N -> padded Kp x N); L blocks
  // transposed (rconv's format with src_off = block offset + first row,
  // N = rows of the task, pad = the block's leading dimension M; element
  // (r, k) of the task reads src[src_off + k M + r], writes dst[dst_off + r Kp + k])
  std::vector<VConv> lconv, rconv;
};

struct BlockDim {
  int64_t off, rows, cols;
};

// Literal per-term list: out block o, X block l, Y block v, Z block r, coefficient.
struct Term {
  int64_t out_v, l_blk, in_v, r_blk;
  double c;
};

// A sum of scalar-weighted block triple products  O_o += c X_l Y_v Z_r: the
// common form of the two-site / one-site / zero-site HxV (X = L, Y = Vin,
// Z = R) and of the environment updates (X = A_bra^T, Y = L, Z = A_ket for
// contract_L; X = A_ket, Y = R, Z = A_bra^T for contract_R).  X and Z are
// packed column-major in product orientation (transposes already applied).
// one device-to-device copy of a gather: n elements from src to dst_off
struct Gather {
  const double* src;
  int64_t dst_off, n;
};

struct TripleSet {
  std::vector<double> Xdata, Zdata;
  // environments resident on the device: the X / Z blocks are gathered there
  // (Xdata / Zdata stay empty); Xsize / Zsize count the packed elements
  std::vector<Gather> Xgather, Zgather;
  std::vector<std::shared_ptr<DevReady>> env_waits;  // pending writes of those environments
  int64_t Xsize = 0, Zsize = 0;
  std::vector<BlockDim> Xb, Zb;
  std::vector<BlockDim> Yb, Ob;  // blocks of the input and output vectors
  int64_t ysize = 0, osize = 0;
  std::vector<double> Ow;        // coeff_dot weight of each output block (empty: all 1)
  std::vector<Term> terms;
};

struct HostPlan {
  // packed operands actually referenced by terms
  std::vector<double> Ldata, Rdata;  // X and Z blocks (empty when gathered on the device)
  std::vector<Gather> Lgather, Rgather;  // device gathers of device-resident environments
  std::vector<std::shared_ptr<DevReady>> env_waits;  // waited on before the gathers run
  int64_t L_size = 0, R_size = 0;    // packed L / R elements
  std::vector<BlockDim> Lblk, Rblk, Vblk;  // blocks of Ldata, Rdata and the input vector
  int64_t vsize = 0;                 // output vector size (== input size for HxV plans)
  int64_t ysize = 0, osize = 0;      // input / output vector sizes
  // device work lists (built from the tables below by build_tiles): tiles of
  // edge kTile (from which the split-BF16 lists derive) and tiles of edge
  // wave_edge for the FP64 / FP32 kernels (the same lists when wave_edge == kTile)
  std::vector<TileA> tilesA;
  std::vector<TileC> tilesC;
  int wave_edge = kTile;
  std::vector<TileA> tilesA_w;
  std::vector<TileC> tilesC_w;
  std::vector<std::pair<int32_t, int32_t>> crange;  // chunksC range of each stage C problem
  // split-BF16: problems with both dimensions >= big_min take the LDS-staged
  // kBigTile tiles (big_min <= 0: none do)
  int big_min = kBigMin;
  // FP64 / FP32: problems with both dimensions >= big_min64 take the
  // LDS-staged kBigTile tiles tilesA_b / tilesC_b and are absent from the wave
  // lists (0: none do, the default)
  int big_min64 = 0;
  std::vector<TileA> tilesA_b;
  std::vector<TileC> tilesC_b;
  // stage B: false = fused into stage C's operand load (U never stored);
  // true = U formed by its own kernel (comb_tasks / comb_terms) and stage C's
  // chunks read U with one unit-weight term each (termsC t_off = U offsets)
  bool explicit_u = false;
  std::vector<CombTask> comb_tasks;
  std::vector<TermC> comb_terms;
  std::vector<ChunkC> chunksC;
  std::vector<TermC> termsC;
  std::vector<GemmProb> probA;       // stage A: one chunk each (L . V)
  std::vector<GemmChunk> chunkA;
  std::vector<CombineList> comb;     // stage B
  std::vector<int32_t> comb_t;       // T index per entry
  std::vector<double> comb_c;        // coefficient per entry
  std::vector<int64_t> T_off;        // element offset of T_t
  int64_t T_size = 0, U_size = 0;
  std::vector<GemmProb> probC;       // stage C: one problem per out block
  std::vector<GemmChunk> chunkC;     // (U_u, R_r) chunks
  // weights for the coeff_dot-weighted inner product: (offset, length, weight)
  std::vector<int64_t> seg_off;
  std::vector<int32_t> seg_len;
  std::vector<double> seg_w;
  Stats stats;
  // block list of the vector (offsets/sizes) for NaN checks and checks of the layout
  std::vector<int64_t> vblock_off, vblock_size;
};

// (term, environment sector, mid index) of an environment block
struct EnvRef {
  size_t t, q, a;
  bool operator<(const EnvRef& o) const {
    if (t != o.t) return t < o.t;
    if (q != o.q) return q < o.q;
    return a < o.a;
  }
};

// Per-term list of a pivot, in the order the reference's OxV visits them
// (terms of the sum one after the other).
void expand_terms(const Pivot& P, std::vector<Term>& terms, std::vector<EnvRef>& lblocks,
                  std::vector<EnvRef>& rblocks, std::vector<std::pair<int, int>>& vblocks);

void build_host_plan(const Pivot& P, HostPlan& H);
void triples_from_pivot(const Pivot& P, TripleSet& S);
void build_plan(const TripleSet& S, HostPlan& H);
void build_plan(TripleSet&& S, HostPlan& H);  // moves S's environment data into H
void build_tiles(HostPlan& H);
// rebuild the FP64 / FP32 tile lists tilesA_w / tilesC_w (wave tiles of edge
// 16 or 32) and tilesA_b / tilesC_b (big tiles of the problems >= big_min64)
void set_wave_edge(HostPlan& H, int edge, int big_min64);
// Launch order of big-tile work items for MI355X, which hands workgroup b of
// a launch to XCD b % kXcdCount (each XCD has its own L2): items are grouped
// by `group` (supertiles: 4 x 4 tiles of one block, which share operand
// panels), the groups dealt largest first to the least loaded XCD, and the XCD
// queues interleaved `per` items at a time (the items of one workgroup).
// Returns item indices in launch order; -1 marks padding slots, filled with
// empty items, that keep later items on their XCD.
constexpr int kXcdCount = 8;
std::vector<int64_t> xcd_order(const std::vector<int64_t>& group, const std::vector<int64_t>& cost, int per = 1);
// wave tile lists at least this long are grouped per XCD (set_wave_edge)
constexpr size_t kXcdMinTilesA = 2048, kXcdMinTilesC = 512;
// key of the plan's block-size histogram: the share of stage A and stage C
// flops in classes of (rows, columns), quantised, so that pivots of similar
// shape share one tuned tile configuration
std::string tile_signature(const HostPlan& H);
// Split-BF16 layouts and work lists of the BF16x3 path (built on demand; the
// split operands themselves are formed on the device from lconv / rconv).
void build_bf16_layout(const HostPlan& H, Bf16Layout& B);
// Throws VMPS_ERR_INTERNAL if any descriptor would address outside its buffer.
void validate_host_plan(const HostPlan& H);

}  // namespace vmps
