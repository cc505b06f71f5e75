// extern "C" boundary (include/vmps_hxv.h).  Every entry point catches all
// exceptions and turns them into a status code plus a thread-local message.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>

#include "denv.hpp"
#include "engine.hpp"
#include "lanczos.hpp"
#include "errors.hpp"
#include "pivot.hpp"
#include "vmps_hxv.h"

struct vmps_pivot2 {
  vmps::Pivot P;
  int ns = 0;
};
struct vmps_plan {
  vmps::DevPlan D;
  int npivots = 1;  // > 1: a batch of independent pivots (vmps_plan_create_batch)
  bool batch = false;
  std::vector<int64_t> offsets;  // batch: pivot i owns [offsets[i], offsets[i+1])
  std::unique_ptr<vmps::BatchKrylov> bk;
};

namespace vmps {
namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }
}  // namespace vmps

using vmps::VmpsError;

#define VMPS_API_BEGIN try {
#define VMPS_API_END                                  \
  return VMPS_OK;                                     \
  }                                                   \
  catch (const VmpsError& e) {                        \
    vmps::set_last_error(e.what());                   \
    return e.status;                                  \
  }                                                   \
  catch (const std::bad_alloc&) {                     \
    vmps::set_last_error("host allocation failed");   \
    return VMPS_ERR_OOM;                              \
  }                                                   \
  catch (const std::exception& e) {                   \
    vmps::set_last_error(e.what());                   \
    return VMPS_ERR_INTERNAL;                         \
  }

namespace {
vmps::Qn read_qn(const int* q, int nq) {
  vmps::Qn r;
  for (int i = 0; i < nq; ++i) r[i] = q[i];
  return r;
}
void require(bool cond, int status, const char* msg) {
  if (!cond) throw VmpsError(status, msg);
}

std::vector<vmps::Qn> read_qns(const int* q, int n, int nq) {
  std::vector<vmps::Qn> v;
  for (int i = 0; i < n; ++i) v.push_back(read_qn(q + i * nq, nq));
  return v;
}

void set_sym(vmps::Sym& sym, int nq, const int* kinds) {
  require(nq >= 1 && nq <= vmps::kMaxNq, VMPS_ERR_ARG, "nq must be 1 or 2");
  require(kinds != nullptr, VMPS_ERR_ARG, "kinds is NULL");
  sym.nq = nq;
  for (int i = 0; i < nq; ++i) {
    require(kinds[i] == VMPS_SYM_U1 || kinds[i] == VMPS_SYM_SU2, VMPS_ERR_ARG, "unknown symmetry kind");
    sym.kind[i] = kinds[i];
  }
}

void require_su2_positive(const vmps::Sym& sym, const std::vector<vmps::Qn>& qs) {
  for (int i = 0; i < sym.nq; ++i)
    if (sym.kind[i] == vmps::kSU2)
      for (auto& q : qs) require(q[i] >= 1, VMPS_ERR_ARG, "SU(2) quantum numbers are q = 2S+1 >= 1");
}

void init_vector_layout(vmps_pivot2* p, int ns) {
  p->ns = ns;
  p->P.V.in.resize(ns);
  p->P.V.out.resize(ns);
  p->P.V.block.resize(ns);
  p->P.V.dict.resize(ns);
}

// one Tripod sector from per-block pointers (NULL or ignored for absent
// blocks, rows * cols == 0): Tripod::push_back, tensors/Multipede.h:77-95
void add_env_sector(vmps_pivot2* p, int which, const int* qin, const int* qout, const int* qmid, int adim,
                    const int64_t* rows, const int64_t* cols, const double* const* ptrs) {
  const int nq = p->P.sym.nq;
  vmps::Tripod& E = which == 0 ? p->P.term().L : p->P.term().R;
  require(!E.dev, VMPS_ERR_STATE, "this environment is already given on the device");
  vmps::Key3 key{read_qn(qin, nq), read_qn(qout, nq), read_qn(qmid, nq)};
  require(!E.dict.count(key), VMPS_ERR_LAYOUT, "duplicate environment sector");
  for (int a = 0; a < adim; ++a) {
    require(rows[a] >= 0 && cols[a] >= 0, VMPS_ERR_ARG, "negative block dimension");
    require(rows[a] * cols[a] == 0 || ptrs[a], VMPS_ERR_ARG, "NULL block data");
  }
  std::vector<vmps::DenseBlock> blocks(adim);
  for (int a = 0; a < adim; ++a) {
    vmps::DenseBlock b;
    b.rows = rows[a];
    b.cols = cols[a];
    if (b.rows * b.cols > 0) {
      b.off = (int64_t)E.data.size();
      E.data.insert(E.data.end(), ptrs[a], ptrs[a] + b.rows * b.cols);
    }
    blocks[a] = b;
  }
  E.index.push_back({key.a, key.b, key.c});
  E.block.push_back(std::move(blocks));
  E.dict[key] = E.index.size() - 1;
}
}  // namespace

namespace {
// rocBLAS handle per device for the two-site vector build, made on first use
// and kept until the process ends; calls on one device are serialised.
struct BlasSlot {
  rocblas_handle h = nullptr;
  std::mutex m;
};
std::mutex g_blas_mutex;
std::map<int, BlasSlot*> g_blas;

BlasSlot& blas_for_device(int dev) {
  std::lock_guard<std::mutex> lock(g_blas_mutex);
  BlasSlot*& s = g_blas[dev];
  if (!s) {
    auto* slot = new BlasSlot();
    if (rocblas_create_handle(&slot->h) != rocblas_status_success) {
      delete slot;
      throw VmpsError(VMPS_ERR_HIP, "rocblas_create_handle failed");
    }
    s = slot;
  }
  return *s;
}
}  // namespace

extern "C" {

int vmps_abi_version(void) { return VMPS_ABI_VERSION; }
const char* vmps_last_error(void) { return vmps::last_error(); }

vmps_status vmps_pivot2_create(int nq, const int* kinds, int nloc12, const int* qloc12, int nloc34, const int* qloc34,
                               int nop12, const int* qop12, int nop34, const int* qop34, vmps_pivot2** out) {
  VMPS_API_BEGIN
  require(out != nullptr, VMPS_ERR_ARG, "out is NULL");
  *out = nullptr;
  require(nloc12 > 0 && nloc34 > 0 && nop12 > 0 && nop34 > 0, VMPS_ERR_ARG, "empty local or operator basis");
  require(qloc12 && qloc34 && qop12 && qop34, VMPS_ERR_ARG, "NULL basis array");
  auto p = std::make_unique<vmps_pivot2>();
  set_sym(p->P.sym, nq, kinds);
  p->P.nsite = 2;
  p->P.qloc12 = read_qns(qloc12, nloc12, nq);
  p->P.qloc34 = read_qns(qloc34, nloc34, nq);
  auto o12 = read_qns(qop12, nop12, nq), o34 = read_qns(qop34, nop34, nq);
  for (auto* v : {&p->P.qloc12, &p->P.qloc34, &o12, &o34}) require_su2_positive(p->P.sym, *v);
  p->P.add_term(o12, o34);
  vmps::TwoSiteBasis tb;
  tb.build(p->P.sym, p->P.qloc12, p->P.qloc34);
  init_vector_layout(p.get(), (int)tb.qloc.size());
  *out = p.release();
  VMPS_API_END
}

vmps_status vmps_pivot1_create(int nq, const int* kinds, int nloc, const int* qloc, int nop, const int* qop,
                               vmps_pivot2** out) {
  VMPS_API_BEGIN
  require(out != nullptr, VMPS_ERR_ARG, "out is NULL");
  *out = nullptr;
  require(nloc > 0 && nop > 0, VMPS_ERR_ARG, "empty local or operator basis");
  require(qloc && qop, VMPS_ERR_ARG, "NULL basis array");
  auto p = std::make_unique<vmps_pivot2>();
  set_sym(p->P.sym, nq, kinds);
  p->P.nsite = 1;
  p->P.qloc12 = read_qns(qloc, nloc, nq);
  auto o = read_qns(qop, nop, nq);
  require_su2_positive(p->P.sym, p->P.qloc12);
  require_su2_positive(p->P.sym, o);
  p->P.add_term(o, {});
  init_vector_layout(p.get(), nloc);
  *out = p.release();
  VMPS_API_END
}

vmps_status vmps_pivot0_create(int nq, const int* kinds, vmps_pivot2** out) {
  VMPS_API_BEGIN
  require(out != nullptr, VMPS_ERR_ARG, "out is NULL");
  *out = nullptr;
  auto p = std::make_unique<vmps_pivot2>();
  set_sym(p->P.sym, nq, kinds);
  p->P.nsite = 0;
  p->P.add_term({}, {});
  init_vector_layout(p.get(), 1);
  *out = p.release();
  VMPS_API_END
}

vmps_status vmps_pivot_add_term(vmps_pivot2* p, int nop12, const int* qop12, int nop34, const int* qop34,
                                int* term) {
  VMPS_API_BEGIN
  require(p != nullptr, VMPS_ERR_ARG, "NULL pivot");
  const int nq = p->P.sym.nq;
  const int ns = p->P.nsite;
  require(ns < 1 || (nop12 > 0 && qop12), VMPS_ERR_ARG, "empty operator basis");
  require(ns < 2 || (nop34 > 0 && qop34), VMPS_ERR_ARG, "empty operator basis");
  auto o12 = ns >= 1 ? read_qns(qop12, nop12, nq) : std::vector<vmps::Qn>{};
  auto o34 = ns == 2 ? read_qns(qop34, nop34, nq) : std::vector<vmps::Qn>{};
  require_su2_positive(p->P.sym, o12);
  require_su2_positive(p->P.sym, o34);
  size_t t = p->P.add_term(o12, o34);
  if (term) *term = (int)t;
  VMPS_API_END
}

vmps_status vmps_pivot_select_term(vmps_pivot2* p, int term) {
  VMPS_API_BEGIN
  require(p != nullptr, VMPS_ERR_ARG, "NULL pivot");
  require(term >= 0 && (size_t)term < p->P.terms.size(), VMPS_ERR_ARG, "term index out of range");
  p->P.cur = (size_t)term;
  VMPS_API_END
}

vmps_status vmps_pivot_info(const vmps_pivot2* p, int* nsite, int* nterms) {
  VMPS_API_BEGIN
  require(p != nullptr, VMPS_ERR_ARG, "NULL pivot");
  if (nsite) *nsite = p->P.nsite;
  if (nterms) *nterms = (int)p->P.terms.size();
  VMPS_API_END
}

void vmps_pivot2_destroy(vmps_pivot2* p) { delete p; }

vmps_status vmps_pivot2_local_basis(const vmps_pivot2* p, int* ns, int* qloc) {
  VMPS_API_BEGIN
  require(p && ns, VMPS_ERR_ARG, "NULL argument");
  *ns = p->ns;
  if (qloc) {
    std::vector<vmps::Qn> q;
    if (p->P.nsite == 2) {
      vmps::TwoSiteBasis tb;
      tb.build(p->P.sym, p->P.qloc12, p->P.qloc34);
      q = tb.qloc;
    } else if (p->P.nsite == 1) {
      q = p->P.qloc12;
    } else {
      vmps::Qn vac{};
      for (int i = 0; i < p->P.sym.nq; ++i) vac[i] = p->P.sym.kind[i] == vmps::kSU2 ? 1 : 0;
      q.assign(1, vac);
    }
    for (int s = 0; s < p->ns; ++s)
      for (int i = 0; i < p->P.sym.nq; ++i) qloc[s * p->P.sym.nq + i] = q[s][i];
  }
  VMPS_API_END
}

vmps_status vmps_pivot2_add_env(vmps_pivot2* p, int which, const int* qin, const int* qout, const int* qmid, int adim,
                                const int64_t* rows, const int64_t* cols, const double* data) {
  VMPS_API_BEGIN
  require(p && qin && qout && qmid, VMPS_ERR_ARG, "NULL argument");
  require(which == 0 || which == 1, VMPS_ERR_ARG, "which must be 0 (L) or 1 (R)");
  require(adim >= 0, VMPS_ERR_ARG, "adim < 0");
  require(adim == 0 || (rows && cols), VMPS_ERR_ARG, "NULL rows/cols");
  std::vector<const double*> ptrs((size_t)adim, nullptr);
  int64_t src = 0;
  for (int a = 0; a < adim; ++a) {
    require(rows[a] >= 0 && cols[a] >= 0, VMPS_ERR_ARG, "negative block dimension");
    if (rows[a] * cols[a] > 0) {
      require(data != nullptr, VMPS_ERR_ARG, "NULL data");
      ptrs[(size_t)a] = data + src;
      src += rows[a] * cols[a];
    }
  }
  add_env_sector(p, which, qin, qout, qmid, adim, rows, cols, ptrs.data());
  VMPS_API_END
}

vmps_status vmps_pivot2_add_env_device(vmps_pivot2* p, int which, const vmps_denv* E) {
  VMPS_API_BEGIN
  require(p && E, VMPS_ERR_ARG, "NULL argument");
  require(which == 0 || which == 1, VMPS_ERR_ARG, "which must be 0 (L) or 1 (R)");
  require(E->nq == p->P.sym.nq, VMPS_ERR_ARG, "symmetry mismatch");
  int dev = -1;
  VMPS_HIP_CHECK(hipGetDevice(&dev));
  require(dev == E->device, VMPS_ERR_ARG, "the environment lives on another device");
  vmps::Tripod& T = which == 0 ? p->P.term().L : p->P.term().R;
  require(T.dim() == 0 && !T.dev, VMPS_ERR_STATE, "this environment already has sectors");
  T = E->T;
  VMPS_API_END
}

vmps_status vmps_pivot2_add_W(vmps_pivot2* p, int which, int s1, int s2, int k, const int* qin, const int* qout,
                              int rows, int cols, int64_t nnz, const int* row, const int* col, const double* val) {
  VMPS_API_BEGIN
  require(p && qin && qout, VMPS_ERR_ARG, "NULL argument");
  require(which == 12 || which == 34, VMPS_ERR_ARG, "which must be 12 or 34");
  require(nnz >= 0 && (nnz == 0 || (row && col && val)), VMPS_ERR_ARG, "bad nonzero arrays");
  require(p->P.nsite == 2 || (p->P.nsite == 1 && which == 12), VMPS_ERR_ARG,
          "W slot 34 exists only for two-site pivots, zero-site pivots have no W");
  auto& W = which == 12 ? p->P.term().W12 : p->P.term().W34;
  const size_t nloc = which == 12 ? p->P.qloc12.size() : p->P.qloc34.size();
  const size_t nop = which == 12 ? p->P.term().qop12.size() : p->P.term().qop34.size();
  require(s1 >= 0 && (size_t)s1 < nloc && s2 >= 0 && (size_t)s2 < nloc && k >= 0 && (size_t)k < nop, VMPS_ERR_ARG,
          "W index out of range");
  require(rows >= 0 && cols >= 0, VMPS_ERR_ARG, "negative W block dimension");
  const int nq = p->P.sym.nq;
  vmps::WBiped& B = W[s1][s2][k];
  vmps::Key2 key{read_qn(qin, nq), read_qn(qout, nq)};
  require(!B.dict.count(key), VMPS_ERR_LAYOUT, "duplicate W block");
  vmps::SparseBlock sb;
  sb.rows = rows;
  sb.cols = cols;
  std::vector<int64_t> order(nnz);
  for (int64_t i = 0; i < nnz; ++i) {
    require(row[i] >= 0 && row[i] < rows && col[i] >= 0 && col[i] < cols, VMPS_ERR_ARG, "W nonzero out of range");
    order[i] = i;
  }
  // Eigen ColMajor InnerIterator order: column outer, row inner
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
    if (col[a] != col[b]) return col[a] < col[b];
    return row[a] < row[b];
  });
  for (int64_t i = 0; i < nnz; ++i) {
    int64_t j = order[i];
    if (i > 0 && col[j] == sb.col.back() && row[j] == sb.row.back())
      throw VmpsError(VMPS_ERR_ARG, "duplicate W nonzero");
    sb.row.push_back(row[j]);
    sb.col.push_back(col[j]);
    sb.val.push_back(val[j]);
  }
  B.in.push_back(key.a);
  B.out.push_back(key.b);
  B.block.push_back(std::move(sb));
  B.dict[key] = B.in.size() - 1;
  VMPS_API_END
}

vmps_status vmps_pivot2_add_V_block(vmps_pivot2* p, int s, const int* qin, const int* qout, int64_t rows, int64_t cols,
                                    int64_t* offset) {
  VMPS_API_BEGIN
  require(p && qin && qout, VMPS_ERR_ARG, "NULL argument");
  const int nq = p->P.sym.nq;
  int q = p->P.add_V_block(s, read_qn(qin, nq), read_qn(qout, nq), rows, cols);
  if (offset) *offset = p->P.V.block[s][q].off;
  p->P.invalidate_blockstructure();
  VMPS_API_END
}

// Whole-environment / whole-slot forms of the builder calls above: one call
// per Tripod, W slot or vector instead of one per sector or block.  Each
// item goes through the single-item entry point, so the checks are the same
// and the first failing item's status is returned.
vmps_status vmps_pivot2_add_env_sectors(vmps_pivot2* p, int which, int nsec, const int* qin, const int* qout,
                                        const int* qmid, const int* adim, const int64_t* rows, const int64_t* cols,
                                        const double* const* blocks) {
  VMPS_API_BEGIN
  require(p && nsec >= 0 && (nsec == 0 || (qin && qout && qmid && adim)), VMPS_ERR_ARG, "bad argument");
  require(which == 0 || which == 1, VMPS_ERR_ARG, "which must be 0 (L) or 1 (R)");
  const int nq = p->P.sym.nq;
  {  // one reservation for all sectors (per-sector growth would copy the data again each time)
    int64_t total = 0, nslot = 0;
    for (int q = 0; q < nsec; ++q) nslot += std::max(0, adim[q]);
    if (rows && cols)
      for (int64_t i = 0; i < nslot; ++i) total += std::max<int64_t>(0, rows[i]) * std::max<int64_t>(0, cols[i]);
    vmps::Tripod& E = which == 0 ? p->P.term().L : p->P.term().R;
    E.data.reserve(E.data.size() + (size_t)total);
  }
  int64_t slot = 0;
  for (int q = 0; q < nsec; ++q) {
    require(adim[q] >= 0, VMPS_ERR_ARG, "adim < 0");
    require(adim[q] == 0 || (rows && cols && blocks), VMPS_ERR_ARG, "NULL rows/cols/blocks");
    add_env_sector(p, which, qin + q * nq, qout + q * nq, qmid + q * nq, adim[q], rows + slot, cols + slot,
                   blocks + slot);
    slot += adim[q];
  }
  VMPS_API_END
}

vmps_status vmps_pivot2_add_W_blocks(vmps_pivot2* p, int which, int nblk, const int* idx, const int* qin,
                                     const int* qout, const int* dims, const int64_t* nnz, const int* row,
                                     const int* col, const double* val) {
  VMPS_API_BEGIN
  require(p && nblk >= 0 && (nblk == 0 || (idx && qin && qout && dims && nnz)), VMPS_ERR_ARG, "bad argument");
  const int nq = p->P.sym.nq;
  int64_t pos = 0;
  for (int b = 0; b < nblk; ++b) {
    require(nnz[b] >= 0, VMPS_ERR_ARG, "nnz < 0");
    const vmps_status st =
        vmps_pivot2_add_W(p, which, idx[3 * b], idx[3 * b + 1], idx[3 * b + 2], qin + b * nq, qout + b * nq,
                          dims[2 * b], dims[2 * b + 1], nnz[b], row ? row + pos : nullptr,
                          col ? col + pos : nullptr, val ? val + pos : nullptr);
    if (st != VMPS_OK) return st;
    pos += nnz[b];
  }
  VMPS_API_END
}

vmps_status vmps_pivot2_add_V_blocks(vmps_pivot2* p, int nblk, const int* s, const int* qin, const int* qout,
                                     const int64_t* rows, const int64_t* cols, int64_t* offsets) {
  VMPS_API_BEGIN
  require(p && nblk >= 0 && (nblk == 0 || (s && qin && qout && rows && cols)), VMPS_ERR_ARG, "bad argument");
  const int nq = p->P.sym.nq;
  for (int b = 0; b < nblk; ++b) {
    const vmps_status st = vmps_pivot2_add_V_block(p, s[b], qin + b * nq, qout + b * nq, rows[b], cols[b],
                                                   offsets ? offsets + b : nullptr);
    if (st != VMPS_OK) return st;
  }
  VMPS_API_END
}

vmps_status vmps_pivot2_vsize(const vmps_pivot2* p, int64_t* size) {
  VMPS_API_BEGIN
  require(p && size, VMPS_ERR_ARG, "NULL argument");
  *size = p->P.V.size;
  VMPS_API_END
}

vmps_status vmps_pivot2_precalc(vmps_pivot2* p) {
  VMPS_API_BEGIN
  require(p != nullptr, VMPS_ERR_ARG, "NULL pivot");
  vmps::precalc_blockStructure(p->P);
  VMPS_API_END
}

vmps_status vmps_pivot2_set_blockstructure(vmps_pivot2* p, int64_t nlhs, const int64_t* qlhs, const int64_t* rhs_begin,
                                           const int64_t* qrhs, const double* factors) {
  VMPS_API_BEGIN
  require(p && nlhs >= 0, VMPS_ERR_ARG, "bad argument");
  require(nlhs == 0 || (qlhs && rhs_begin), VMPS_ERR_ARG, "NULL table");
  const int64_t nrhs = nlhs ? rhs_begin[nlhs] : 0;
  require(nrhs == 0 || (qrhs && factors), VMPS_ERR_ARG, "NULL table");
  auto& P = p->P.term();
  P.qlhs.assign(nlhs, {0, 0});
  P.qrhs.assign(nlhs, {});
  P.factor_cgcs.assign(nlhs, {});
  for (int64_t i = 0; i < nlhs; ++i) {
    require(qlhs[2 * i] >= 0 && qlhs[2 * i + 1] >= 0, VMPS_ERR_ARG, "negative index");
    P.qlhs[i] = {(size_t)qlhs[2 * i], (size_t)qlhs[2 * i + 1]};
    require(rhs_begin[i] <= rhs_begin[i + 1] && rhs_begin[i] >= 0, VMPS_ERR_ARG, "rhs_begin not monotone");
    for (int64_t j = rhs_begin[i]; j < rhs_begin[i + 1]; ++j) {
      std::array<size_t, 12> e;
      for (int c = 0; c < 12; ++c) {
        require(qrhs[12 * j + c] >= 0, VMPS_ERR_ARG, "negative index");
        e[c] = (size_t)qrhs[12 * j + c];
      }
      P.qrhs[i].push_back(e);
      P.factor_cgcs[i].push_back(factors[j]);
    }
  }
  P.has_blockstructure = true;
  VMPS_API_END
}

vmps_status vmps_pivot2_get_blockstructure(const vmps_pivot2* p, int64_t* nlhs, int64_t* nrhs, int64_t* qlhs,
                                           int64_t* rhs_begin, int64_t* qrhs, double* factors) {
  VMPS_API_BEGIN
  require(p && nlhs && nrhs, VMPS_ERR_ARG, "NULL argument");
  require(p->P.term().has_blockstructure, VMPS_ERR_STATE, "no block structure");
  const auto& P = p->P.term();
  *nlhs = (int64_t)P.qlhs.size();
  int64_t n = 0;
  for (auto& v : P.qrhs) n += (int64_t)v.size();
  *nrhs = n;
  if (qlhs && rhs_begin && qrhs && factors) {
    int64_t j = 0;
    for (size_t i = 0; i < P.qlhs.size(); ++i) {
      qlhs[2 * i] = (int64_t)P.qlhs[i][0];
      qlhs[2 * i + 1] = (int64_t)P.qlhs[i][1];
      rhs_begin[i] = j;
      for (size_t q = 0; q < P.qrhs[i].size(); ++q, ++j) {
        for (int c = 0; c < 12; ++c) qrhs[12 * j + c] = (int64_t)P.qrhs[i][q][c];
        factors[j] = P.factor_cgcs[i][q];
      }
    }
    rhs_begin[P.qlhs.size()] = j;
  }
  VMPS_API_END
}

vmps_status vmps_pivot2_set_projectors(vmps_pivot2* p, double Epenalty, int n, const double* A0) {
  VMPS_API_BEGIN
  require(p && n >= 0 && (n == 0 || A0), VMPS_ERR_ARG, "bad argument");
  p->P.Epenalty = Epenalty;
  p->P.A0proj.assign(n, {});
  for (int i = 0; i < n; ++i) p->P.A0proj[i].assign(A0 + (size_t)i * p->P.V.size, A0 + (size_t)(i + 1) * p->P.V.size);
  VMPS_API_END
}

vmps_status vmps_plan_create(const vmps_pivot2* p, vmps_precision prec, vmps_plan** out) {
  VMPS_API_BEGIN
  require(p && out, VMPS_ERR_ARG, "NULL argument");
  *out = nullptr;
  require(prec == VMPS_PREC_FP64 || prec == VMPS_PREC_FP32 || prec == VMPS_PREC_BF16X3, VMPS_ERR_ARG,
          "unknown precision");
  for (auto& v : p->P.A0proj)
    require((int64_t)v.size() == p->P.V.size, VMPS_ERR_LAYOUT, "projector size differs from the vector size");
  auto plan = std::make_unique<vmps_plan>();
  plan->D.create(p->P, prec);
  *out = plan.release();
  VMPS_API_END
}

vmps_status vmps_plan_create_batch(const vmps_pivot2* const* pivots, int n, vmps_precision prec, vmps_plan** out,
                                   int64_t* offsets) {
  VMPS_API_BEGIN
  require(pivots && out && n > 0, VMPS_ERR_ARG, "bad argument");
  *out = nullptr;
  require(prec == VMPS_PREC_FP64 || prec == VMPS_PREC_FP32 || prec == VMPS_PREC_BF16X3, VMPS_ERR_ARG,
          "unknown precision");
  // one triple set over the concatenated vectors: pivot i owns [off_i, off_i + vsize_i)
  vmps::TripleSet S;
  std::vector<int64_t> starts;
  for (int i = 0; i < n; ++i) {
    require(pivots[i] != nullptr, VMPS_ERR_ARG, "NULL pivot in batch");
    require(pivots[i]->P.A0proj.empty(), VMPS_ERR_ARG, "projectors are not supported in a batch");
    vmps::TripleSet Si;
    vmps::triples_from_pivot(pivots[i]->P, Si);
    require(Si.Xgather.empty() && Si.Zgather.empty(), VMPS_ERR_ARG,
            "a batch takes host environments only (vmps_pivot2_add_env*)");
    const int64_t nx = (int64_t)S.Xb.size(), nz = (int64_t)S.Zb.size(), ny = (int64_t)S.Yb.size();
    const int64_t xo = (int64_t)S.Xdata.size(), zo = (int64_t)S.Zdata.size(), vo = S.ysize;
    if (offsets) offsets[i] = vo;
    starts.push_back(vo);
    for (auto b : Si.Xb) S.Xb.push_back({b.off + xo, b.rows, b.cols});
    for (auto b : Si.Zb) S.Zb.push_back({b.off + zo, b.rows, b.cols});
    for (auto b : Si.Yb) S.Yb.push_back({b.off + vo, b.rows, b.cols});
    for (auto b : Si.Ob) S.Ob.push_back({b.off + vo, b.rows, b.cols});
    S.Xdata.insert(S.Xdata.end(), Si.Xdata.begin(), Si.Xdata.end());
    S.Zdata.insert(S.Zdata.end(), Si.Zdata.begin(), Si.Zdata.end());
    S.Ow.insert(S.Ow.end(), Si.Ow.begin(), Si.Ow.end());
    for (const auto& t : Si.terms) S.terms.push_back({t.out_v + ny, t.l_blk + nx, t.in_v + ny, t.r_blk + nz, t.c});
    S.ysize += Si.ysize;
    S.osize += Si.osize;
  }
  if (offsets) offsets[n] = S.ysize;
  const int64_t S_ysize = S.ysize;
  auto plan = std::make_unique<vmps_plan>();
  plan->D.create_triples(std::move(S), prec);
  plan->npivots = n;
  plan->batch = true;
  for (int i = 0; i < n; ++i) plan->offsets.push_back(starts[i]);
  plan->offsets.push_back(S_ysize);
  *out = plan.release();
  VMPS_API_END
}

void vmps_plan_destroy(vmps_plan* plan) { delete plan; }

static void copy_stats(const vmps::Stats& s, vmps_plan_stats* st) {
  st->n_terms = s.n_terms;
  st->n_T = s.n_T;
  st->n_U = s.n_U;
  st->n_out = s.n_out;
  st->n_L = s.n_L;
  st->n_R = s.n_R;
  st->n_V = s.n_V;
  st->flops_alg = s.flops_alg;
  st->flops_ref = s.flops_ref;
  st->bytes_alg_fp64 = s.bytes_alg_fp64;
  st->L_elems = s.L_elems;
  st->R_elems = s.R_elems;
  st->V_elems = s.V_elems;
  st->T_elems = s.T_elems;
  st->U_elems = s.U_elems;
  st->max_block_dim = s.max_block_dim;
  st->flops_A = s.flops_A;
  st->flops_C = s.flops_C;
}

vmps_status vmps_plan_stats_get(const vmps_plan* plan, vmps_plan_stats* st) {
  VMPS_API_BEGIN
  require(plan && st, VMPS_ERR_ARG, "NULL argument");
  copy_stats(plan->D.H.stats, st);
  VMPS_API_END
}

vmps_status vmps_pivot2_stats(const vmps_pivot2* p, vmps_plan_stats* st) {
  VMPS_API_BEGIN
  require(p && st, VMPS_ERR_ARG, "NULL argument");
  vmps::HostPlan H;
  vmps::build_host_plan(p->P, H);
  copy_stats(H.stats, st);
  VMPS_API_END
}

vmps_status vmps_plan_set_precision(vmps_plan* plan, vmps_precision prec) {
  VMPS_API_BEGIN
  require(plan != nullptr, VMPS_ERR_ARG, "NULL plan");
  require(prec == VMPS_PREC_FP64 || prec == VMPS_PREC_FP32 || prec == VMPS_PREC_BF16X3, VMPS_ERR_ARG,
          "unknown precision");
  plan->D.prec = prec;
  plan->D.ensure_precision_buffers();
  VMPS_HIP_CHECK(hipStreamSynchronize(plan->D.stream));
  VMPS_API_END
}

vmps_status vmps_hxv(vmps_plan* plan, const double* vin, double* vout) {
  VMPS_API_BEGIN
  require(plan && (plan->D.H.vsize == 0 || (vin && vout)), VMPS_ERR_ARG, "NULL argument");
  auto& D = plan->D;
  const size_t bytes = (size_t)D.H.vsize * sizeof(double);
  for (int64_t i = 0; i < D.H.vsize; ++i)
    if (!std::isfinite(vin[i])) throw VmpsError(VMPS_ERR_NONFINITE, "non-finite entry in Vin at " + std::to_string(i));
  if (bytes) VMPS_HIP_CHECK(hipMemcpyAsync(D.vin.p, vin, bytes, hipMemcpyHostToDevice, D.stream));
  D.hxv(D.vin.p, D.vout.p, D.stream);
  if (bytes) VMPS_HIP_CHECK(hipMemcpyAsync(vout, D.vout.p, bytes, hipMemcpyDeviceToHost, D.stream));
  VMPS_HIP_CHECK(hipStreamSynchronize(D.stream));
  VMPS_API_END
}

vmps_status vmps_hxv_device(vmps_plan* plan, const double* d_vin, double* d_vout, void* stream) {
  VMPS_API_BEGIN
  require(plan && (plan->D.H.vsize == 0 || (d_vin && d_vout)), VMPS_ERR_ARG, "NULL argument");
  require(d_vin != d_vout || plan->D.H.vsize == 0, VMPS_ERR_ARG, "Vin and Vout must not alias");
  plan->D.hxv(d_vin, d_vout, (hipStream_t)stream);
  VMPS_API_END
}

double vmps_coupling_6j(int q1, int q2, int q3, int q4, int q5, int q6) {
  return vmps::wigner6j_q(q1, q2, q3, q4, q5, q6);
}

double vmps_coupling_9j(int q1, int q2, int q3, int q4, int q5, int q6, int q7, int q8, int q9) {
  return vmps::wigner9j_q(q1, q2, q3, q4, q5, q6, q7, q8, q9);
}

vmps_status vmps_dot(vmps_plan* plan, const double* x, const double* y, double* result) {
  VMPS_API_BEGIN
  require(plan && result && (plan->D.H.vsize == 0 || (x && y)), VMPS_ERR_ARG, "NULL argument");
  auto& D = plan->D;
  const size_t bytes = (size_t)D.H.vsize * sizeof(double);
  if (bytes) {
    VMPS_HIP_CHECK(hipMemcpyAsync(D.vin.p, x, bytes, hipMemcpyHostToDevice, D.stream));
    VMPS_HIP_CHECK(hipMemcpyAsync(D.vout.p, y, bytes, hipMemcpyHostToDevice, D.stream));
  }
  *result = D.dot(D.vin.p, D.vout.p, D.stream);
  VMPS_API_END
}

vmps_status vmps_plan_time_hxv(vmps_plan* plan, int warmup, int iters, double* ms_per_hxv) {
  VMPS_API_BEGIN
  require(plan && ms_per_hxv && iters > 0 && warmup >= 0, VMPS_ERR_ARG, "bad argument");
  auto& D = plan->D;
  for (int i = 0; i < warmup; ++i) D.hxv(D.vin.p, D.vout.p, D.stream);
  VMPS_HIP_CHECK(hipEventRecord(D.ev0, D.stream));
  for (int i = 0; i < iters; ++i) D.hxv(D.vin.p, D.vout.p, D.stream);
  VMPS_HIP_CHECK(hipEventRecord(D.ev1, D.stream));
  VMPS_HIP_CHECK(hipEventSynchronize(D.ev1));
  float ms = 0.f;
  VMPS_HIP_CHECK(hipEventElapsedTime(&ms, D.ev0, D.ev1));
  *ms_per_hxv = (double)ms / iters;
  VMPS_API_END
}

vmps_status vmps_plan_get_tiles(const vmps_plan* plan, vmps_tile_config* cfg) {
  VMPS_API_BEGIN
  require(plan && cfg, VMPS_ERR_ARG, "NULL argument");
  const auto& H = plan->D.H;
  *cfg = vmps_tile_config{H.wave_edge, H.big_min, H.big_min64, H.explicit_u ? 1 : 0, 0.0, 0, 0};
  VMPS_API_END
}

vmps_status vmps_plan_set_tiles(vmps_plan* plan, const vmps_tile_config* cfg) {
  VMPS_API_BEGIN
  require(plan && cfg, VMPS_ERR_ARG, "NULL argument");
  require(cfg->wave_edge == 16 || cfg->wave_edge == 32, VMPS_ERR_ARG, "wave_edge must be 16 or 32");
  require(cfg->big_min >= 0 && cfg->big_min64 >= 0, VMPS_ERR_ARG, "big-tile thresholds must be >= 0");
  require(cfg->explicit_u == 0 || cfg->explicit_u == 1, VMPS_ERR_ARG, "explicit_u must be 0 or 1");
  plan->D.retile(cfg->wave_edge, cfg->big_min, cfg->big_min64, cfg->explicit_u != 0);
  VMPS_API_END
}

namespace {
struct TileChoice {
  int wave_edge, big_min, big_min64;
  bool explicit_u;
};
std::mutex g_tune_mutex;
std::map<std::string, TileChoice> g_tune_cache;  // (precision, block-size histogram) -> choice
}  // namespace

vmps_status vmps_plan_autotune(vmps_plan* plan, int iters, vmps_tile_config* out) {
  VMPS_API_BEGIN
  require(plan && iters > 0, VMPS_ERR_ARG, "bad argument");
  auto& D = plan->D;
  const std::string key = std::to_string((int)D.prec) + "/" + vmps::tile_signature(D.H);
  auto result = [&](double ms, int from_cache, int tried) {
    return vmps_tile_config{D.H.wave_edge, D.H.big_min, D.H.big_min64, D.H.explicit_u ? 1 : 0, ms, from_cache, tried};
  };
  {
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      D.retile(it->second.wave_edge, it->second.big_min, it->second.big_min64, it->second.explicit_u);
      if (out) *out = result(0.0, 1, 0);
      return VMPS_OK;
    }
  }
  // Coordinate search from the plan's current configuration over the
  // parameters its precision uses (the others stay as they are): each
  // parameter's other values in turn, keeping the best so far.  Four
  // candidates instead of the eight (FP64 / FP32) or six (split-BF16) of the
  // full grid; one warm-up call per candidate absorbs its first-call set-up.
  auto measure = [&](const TileChoice& c) {
    D.retile(c.wave_edge, c.big_min, c.big_min64, c.explicit_u);
    D.hxv(D.vin.p, D.vout.p, D.stream);
    VMPS_HIP_CHECK(hipEventRecord(D.ev0, D.stream));
    for (int i = 0; i < iters; ++i) D.hxv(D.vin.p, D.vout.p, D.stream);
    VMPS_HIP_CHECK(hipEventRecord(D.ev1, D.stream));
    VMPS_HIP_CHECK(hipEventSynchronize(D.ev1));
    float ms = 0.f;
    VMPS_HIP_CHECK(hipEventElapsedTime(&ms, D.ev0, D.ev1));
    return (double)ms / iters;
  };
  TileChoice pick{D.H.wave_edge, D.H.big_min, D.H.big_min64, D.H.explicit_u};
  // Pivots with blocks of 256 and more always take the FP64 / FP32 big tiles:
  // without them such a pivot's HxV is 2-3x slower (C60 at M = 4096: 124 ms
  // against 45 ms), so the search neither starts from nor times that choice.
  const bool large = D.H.stats.max_block_dim >= 2 * vmps::kBigTile && D.prec != VMPS_PREC_BF16X3;
  if (large && pick.big_min64 <= 0) pick.big_min64 = vmps::kBigTile;
  double best = measure(pick);
  int tried = 1;
  auto consider = [&](TileChoice c) {
    if (c.wave_edge == pick.wave_edge && c.big_min == pick.big_min && c.big_min64 == pick.big_min64 &&
        c.explicit_u == pick.explicit_u)
      return;
    const double t = measure(c);
    ++tried;
    if (t < best) {
      best = t;
      pick = c;
    }
  };
  if (D.prec == VMPS_PREC_BF16X3) {
    const TileChoice base = pick;
    for (int b : {vmps::kBigMin, 2 * vmps::kBigMin, 0}) {
      TileChoice c = base;
      c.big_min = b;
      consider(c);
    }
  } else {
    for (int e : {16, 32}) {
      TileChoice c = pick;
      c.wave_edge = e;
      consider(c);
    }
    for (int b64 : {0, vmps::kBigTile}) {
      if (large && b64 == 0) continue;
      TileChoice c = pick;
      c.big_min64 = b64;
      consider(c);
    }
  }
  for (bool xu : {false, true}) {
    TileChoice c = pick;
    c.explicit_u = xu;
    consider(c);
  }
  D.retile(pick.wave_edge, pick.big_min, pick.big_min64, pick.explicit_u);
  {
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    g_tune_cache[key] = pick;
  }
  if (out) *out = result(best, 0, tried);
  VMPS_API_END
}

vmps_status vmps_autotune_clear_cache(void) {
  VMPS_API_BEGIN
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  g_tune_cache.clear();
  VMPS_API_END
}

vmps_status vmps_vec_alloc(vmps_plan* plan, double** d_v) {
  VMPS_API_BEGIN
  require(plan && d_v, VMPS_ERR_ARG, "NULL argument");
  *d_v = nullptr;
  const size_t bytes = std::max<size_t>(1, (size_t)plan->D.H.vsize) * sizeof(double);
  VMPS_HIP_CHECK(hipMalloc((void**)d_v, bytes));
  VMPS_HIP_CHECK(hipMemsetAsync(*d_v, 0, bytes, plan->D.stream));
  VMPS_HIP_CHECK(hipStreamSynchronize(plan->D.stream));
  VMPS_API_END
}

vmps_status vmps_vec_free(vmps_plan* plan, double* d_v) {
  VMPS_API_BEGIN
  require(plan != nullptr, VMPS_ERR_ARG, "NULL plan");
  if (d_v) {
    VMPS_HIP_CHECK(hipStreamSynchronize(plan->D.stream));
    VMPS_HIP_CHECK(hipFree(d_v));
  }
  VMPS_API_END
}

vmps_status vmps_vec_upload(vmps_plan* plan, double* d_v, const double* h_v) {
  VMPS_API_BEGIN
  require(plan && (plan->D.H.vsize == 0 || (d_v && h_v)), VMPS_ERR_ARG, "NULL argument");
  const size_t bytes = (size_t)plan->D.H.vsize * sizeof(double);
  if (bytes) VMPS_HIP_CHECK(hipMemcpyAsync(d_v, h_v, bytes, hipMemcpyHostToDevice, plan->D.stream));
  VMPS_HIP_CHECK(hipStreamSynchronize(plan->D.stream));
  VMPS_API_END
}

vmps_status vmps_vec_download(vmps_plan* plan, const double* d_v, double* h_v) {
  VMPS_API_BEGIN
  require(plan && (plan->D.H.vsize == 0 || (d_v && h_v)), VMPS_ERR_ARG, "NULL argument");
  const size_t bytes = (size_t)plan->D.H.vsize * sizeof(double);
  if (bytes) VMPS_HIP_CHECK(hipMemcpyAsync(h_v, d_v, bytes, hipMemcpyDeviceToHost, plan->D.stream));
  VMPS_HIP_CHECK(hipStreamSynchronize(plan->D.stream));
  VMPS_API_END
}

vmps_status vmps_vec_dot(vmps_plan* plan, const double* d_x, const double* d_y, double* result) {
  VMPS_API_BEGIN
  require(plan && result && (plan->D.H.vsize == 0 || (d_x && d_y)), VMPS_ERR_ARG, "NULL argument");
  *result = plan->D.dot(d_x, d_y, plan->D.stream);
  VMPS_API_END
}

vmps_status vmps_vec_norm(vmps_plan* plan, const double* d_x, double* result) {
  VMPS_API_BEGIN
  require(plan && result && (plan->D.H.vsize == 0 || d_x), VMPS_ERR_ARG, "NULL argument");
  *result = std::sqrt(std::max(0.0, plan->D.dot(d_x, d_x, plan->D.stream)));
  VMPS_API_END
}

vmps_status vmps_vec_axpby(vmps_plan* plan, double a, const double* d_x, double b, double* d_y) {
  VMPS_API_BEGIN
  require(plan && (plan->D.H.vsize == 0 || (d_x && d_y)), VMPS_ERR_ARG, "NULL argument");
  if (plan->D.H.vsize) vmps::dev_axpby(a, d_x, b, d_y, plan->D.H.vsize, plan->D.stream);
  VMPS_HIP_CHECK(hipStreamSynchronize(plan->D.stream));
  VMPS_API_END
}

vmps_status vmps_vec_fill_random(vmps_plan* plan, double* d_v, uint64_t seed) {
  VMPS_API_BEGIN
  require(plan && (plan->D.H.vsize == 0 || d_v), VMPS_ERR_ARG, "NULL argument");
  auto& D = plan->D;
  const int64_t n = D.H.vsize;
  if (n) {
    vmps::dev_fill_uniform(d_v, n, seed, D.stream);
    const double nrm = std::sqrt(D.dot(d_v, d_v, D.stream));
    require(nrm > 0.0, VMPS_ERR_INTERNAL, "random vector has zero norm");
    vmps::dev_axpby(0.0, d_v, 1.0 / nrm, d_v, n, D.stream);
    VMPS_HIP_CHECK(hipStreamSynchronize(D.stream));
  }
  VMPS_API_END
}

vmps_status vmps_contract_aa2_device(vmps_plan* plan, int n1, const int64_t* rows1, const int64_t* cols1,
                                     const double* const* A1, int n2, const int64_t* rows2, const int64_t* cols2,
                                     const double* const* A2, int nterms, const int64_t* out_off, const int* i1,
                                     const int* i2, const double* factor, double* d_out) {
  VMPS_API_BEGIN
  require(plan && !plan->batch, VMPS_ERR_ARG, "NULL or batch plan");
  require(n1 >= 0 && n2 >= 0 && nterms >= 0, VMPS_ERR_ARG, "negative count");
  require(n1 == 0 || (rows1 && cols1 && A1), VMPS_ERR_ARG, "NULL A1 table");
  require(n2 == 0 || (rows2 && cols2 && A2), VMPS_ERR_ARG, "NULL A2 table");
  require(nterms == 0 || (out_off && i1 && i2 && factor), VMPS_ERR_ARG, "NULL term table");
  auto& D = plan->D;
  const int64_t vsize = D.H.vsize;
  require(vsize == 0 || d_out, VMPS_ERR_ARG, "NULL output vector");
  // every term's operands chain and its block lies inside the vector
  std::vector<char> used1((size_t)n1, 0), used2((size_t)n2, 0);
  for (int t = 0; t < nterms; ++t) {
    require(i1[t] >= 0 && i1[t] < n1 && i2[t] >= 0 && i2[t] < n2, VMPS_ERR_ARG, "term block index out of range");
    const int64_t r = rows1[i1[t]], k = cols1[i1[t]], k2 = rows2[i2[t]], c = cols2[i2[t]];
    require(r >= 0 && k >= 0 && k2 >= 0 && c >= 0, VMPS_ERR_ARG, "negative block dimension");
    require(k == k2, VMPS_ERR_LAYOUT, "A1 and A2 blocks of a term do not chain");
    require(out_off[t] >= 0 && out_off[t] + r * c <= vsize, VMPS_ERR_LAYOUT, "term block outside the vector");
    require(r <= INT32_MAX && c <= INT32_MAX && k <= INT32_MAX, VMPS_ERR_ARG, "block too large");
    if (r * c == 0 || k == 0) continue;
    require(A1[i1[t]] && A2[i2[t]], VMPS_ERR_ARG, "NULL block data");
    used1[(size_t)i1[t]] = used2[(size_t)i2[t]] = 1;
  }
  // the blocks the terms use, packed on the device
  std::vector<int64_t> off1((size_t)n1, 0), off2((size_t)n2, 0);
  int64_t tot1 = 0, tot2 = 0;
  for (int b = 0; b < n1; ++b)
    if (used1[(size_t)b]) off1[(size_t)b] = tot1, tot1 += rows1[b] * cols1[b];
  for (int b = 0; b < n2; ++b)
    if (used2[(size_t)b]) off2[(size_t)b] = tot2, tot2 += rows2[b] * cols2[b];
  vmps::DevBuf<double> d1, d2;
  d1.alloc((size_t)std::max<int64_t>(1, tot1));
  d2.alloc((size_t)std::max<int64_t>(1, tot2));
  const hipStream_t s = D.stream;
  for (int b = 0; b < n1; ++b)
    if (used1[(size_t)b])
      VMPS_HIP_CHECK(hipMemcpyAsync(d1.p + off1[(size_t)b], A1[b], sizeof(double) * rows1[b] * cols1[b],
                                    hipMemcpyHostToDevice, s));
  for (int b = 0; b < n2; ++b)
    if (used2[(size_t)b])
      VMPS_HIP_CHECK(hipMemcpyAsync(d2.p + off2[(size_t)b], A2[b], sizeof(double) * rows2[b] * cols2[b],
                                    hipMemcpyHostToDevice, s));
  if (vsize) VMPS_HIP_CHECK(hipMemsetAsync(d_out, 0, sizeof(double) * vsize, s));
  int dev = 0;
  VMPS_HIP_CHECK(hipGetDevice(&dev));
  BlasSlot& slot = blas_for_device(dev);
  {
    std::lock_guard<std::mutex> hold(slot.m);
    if (rocblas_set_stream(slot.h, s) != rocblas_status_success)
      throw VmpsError(VMPS_ERR_HIP, "rocblas_set_stream failed");
    const double one = 1.0;
    // out block (r x c, column-major) += f A1 (r x k) . A2 (k x c)
    for (int t = 0; t < nterms; ++t) {
      const int64_t r = rows1[i1[t]], k = cols1[i1[t]], c = cols2[i2[t]];
      if (r * c == 0 || k == 0) continue;
      const rocblas_status st =
          rocblas_dgemm(slot.h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)r, (rocblas_int)c,
                        (rocblas_int)k, &factor[t], d1.p + off1[(size_t)i1[t]], (rocblas_int)r,
                        d2.p + off2[(size_t)i2[t]], (rocblas_int)k, &one, d_out + out_off[t], (rocblas_int)r);
      if (st != rocblas_status_success)
        throw VmpsError(VMPS_ERR_HIP, std::string("rocblas_dgemm: ") + rocblas_status_to_string(st));
    }
  }
  VMPS_HIP_CHECK(hipStreamSynchronize(s));  // the packed blocks are freed on return
  VMPS_API_END
}

void vmps_lanczos_opts_default(vmps_lanczos_opts* o) {
  if (!o) return;
  o->max_iter = 100;
  o->min_iter = 2;
  o->tol_eigval = 1e-8;
  o->tol_state = 1e-5;
  o->precision = VMPS_PREC_FP64;
  o->refine_fp64 = 1;
  o->refine_tol_eigval = 1e-12;
  o->refine_tol_state = 1e-7;
  o->seed = 1;
  o->graph = 1;
  o->refine_method = VMPS_REFINE_CORRECTION;
}

static vmps_status lanczos_impl(vmps_plan* plan, double* d_v, bool random_start, const vmps_lanczos_opts* opts,
                                vmps_lanczos_result* res) {
  VMPS_API_BEGIN
  require(plan && opts && res, VMPS_ERR_ARG, "NULL argument");
  require(!plan->batch || plan->npivots == 1, VMPS_ERR_STATE,
          "a batch plan holds independent pivots; use vmps_lanczos_batch for one eigenpair per pivot");
  require(opts->max_iter >= 1 && opts->tol_eigval >= 0 && opts->tol_state >= 0, VMPS_ERR_ARG, "bad options");
  require(opts->precision >= VMPS_PREC_FP64 && opts->precision <= VMPS_PREC_BF16X3, VMPS_ERR_ARG, "bad precision");
  require(opts->refine_method == VMPS_REFINE_CORRECTION || opts->refine_method == VMPS_REFINE_RESTART, VMPS_ERR_ARG,
          "bad refine_method");
  auto& D = plan->D;
  const auto t0 = std::chrono::steady_clock::now();
  std::memset(res, 0, sizeof(*res));
  if (random_start) {
    vmps::dev_fill_uniform(d_v, D.H.vsize, opts->seed, D.stream);
  }
  vmps::LanczosOpts lo;
  lo.max_iter = opts->max_iter;
  lo.min_iter = opts->min_iter;
  lo.tol_eigval = opts->tol_eigval;
  lo.tol_state = opts->tol_state;
  lo.graph = opts->graph != 0;
  const auto prec = (vmps_precision)opts->precision;
  if (prec != VMPS_PREC_FP64 && opts->refine_fp64) {
    // The mixed phase cannot resolve the state below the rounding level of its
    // products (~2^-24 FP32, ~2^-16 split-BF16): stop it at a few times that
    // level (the energy error is quadratic in the state error) and leave the
    // rest to the FP64 refinement.
    const double state_floor = prec == VMPS_PREC_FP32 ? 1e-6 : 2e-4;
    lo.tol_state = std::max(lo.tol_state, state_floor);
    lo.tol_eigval = std::max(lo.tol_eigval, state_floor * state_floor);
  }
  // a correction refinement forms the mixed vector's Rayleigh quotient and
  // residual with the FP64 operator anyway: no extra mixed HxV for them
  const bool corrections = prec != VMPS_PREC_FP64 && opts->refine_fp64 && opts->refine_method == VMPS_REFINE_CORRECTION;
  lo.final_rq = !corrections;
  vmps::LanczosResult r = vmps::lanczos_solve(D, d_v, lo, prec);
  res->iterations = r.iterations;
  res->hxv_calls = r.hxv_calls;
  res->energy_mixed = r.energy;
  res->residual_mixed = r.residual;
  res->energy = r.energy;
  res->residual = r.residual;
  res->converged = r.converged;
  res->hxv_fp64 = prec == VMPS_PREC_FP64 ? r.hxv_calls : 0;
  if (prec != VMPS_PREC_FP64 && opts->refine_fp64) {
    bool done = false;
    if (opts->refine_method == VMPS_REFINE_CORRECTION) {
      vmps::RefineOpts ro;
      ro.tol_eigval = opts->refine_tol_eigval;
      ro.tol_state = opts->refine_tol_state;
      vmps::RefineResult f = vmps::refine_eigenpair(D, d_v, ro, prec);
      res->energy_mixed = f.energy0;
      res->residual_mixed = f.residual0;
      res->refine_steps = f.outer;
      res->refine_cg = f.cg_iterations;
      res->hxv_calls += f.cg_iterations + f.hxv_fp64;
      res->hxv_fp64 += f.hxv_fp64;
      res->energy = f.energy;
      res->residual = f.residual;
      res->converged = f.converged;
      done = f.converged;
    }
    if (!done) {  // RESTART, or corrections that did not reach the tolerances
      vmps::LanczosOpts ro = lo;
      ro.tol_eigval = opts->refine_tol_eigval;
      ro.tol_state = opts->refine_tol_state;
      ro.min_iter = 2;
      vmps::LanczosResult f = vmps::lanczos_solve(D, d_v, ro, VMPS_PREC_FP64);
      res->iterations_refine = f.iterations;
      res->hxv_calls += f.hxv_calls;
      res->hxv_fp64 += f.hxv_calls;
      res->energy = f.energy;
      res->residual = f.residual;
      res->converged = f.converged;
    }
  }
  res->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  VMPS_API_END
}

vmps_status vmps_lanczos_device(vmps_plan* plan, double* d_v, const vmps_lanczos_opts* opts,
                                vmps_lanczos_result* res) {
  if (!plan || !d_v) {
    vmps::set_last_error("NULL argument");
    return VMPS_ERR_ARG;
  }
  return lanczos_impl(plan, d_v, false, opts, res);
}

vmps_status vmps_lanczos(vmps_plan* plan, double* v, const vmps_lanczos_opts* opts, vmps_lanczos_result* res) {
  VMPS_API_BEGIN
  require(plan != nullptr, VMPS_ERR_ARG, "NULL plan");
  auto& D = plan->D;
  const size_t bytes = (size_t)D.H.vsize * sizeof(double);
  if (v && bytes) {
    for (int64_t i = 0; i < D.H.vsize; ++i)
      if (!std::isfinite(v[i])) throw VmpsError(VMPS_ERR_NONFINITE, "non-finite entry in the start vector");
    VMPS_HIP_CHECK(hipMemcpyAsync(D.vin.p, v, bytes, hipMemcpyHostToDevice, D.stream));
  }
  vmps_status st = lanczos_impl(plan, D.vin.p, v == nullptr, opts, res);
  if (st != VMPS_OK) return st;
  if (v && bytes) {
    VMPS_HIP_CHECK(hipMemcpyAsync(v, D.vin.p, bytes, hipMemcpyDeviceToHost, D.stream));
    VMPS_HIP_CHECK(hipStreamSynchronize(D.stream));
  }
  VMPS_API_END
}

vmps_status vmps_lanczos_batch(vmps_plan* plan, double* v, const vmps_lanczos_opts* opts, vmps_lanczos_result* res) {
  VMPS_API_BEGIN
  require(plan && opts && res, VMPS_ERR_ARG, "NULL argument");
  require(plan->batch, VMPS_ERR_STATE, "not a batch plan (vmps_plan_create_batch)");
  require(opts->max_iter >= 1 && opts->tol_eigval >= 0 && opts->tol_state >= 0, VMPS_ERR_ARG, "bad options");
  require(opts->precision >= VMPS_PREC_FP64 && opts->precision <= VMPS_PREC_BF16X3, VMPS_ERR_ARG, "bad precision");
  auto& D = plan->D;
  const int P = plan->npivots;
  const auto t0 = std::chrono::steady_clock::now();
  std::memset(res, 0, sizeof(*res) * (size_t)P);
  const size_t bytes = (size_t)D.H.vsize * sizeof(double);
  if (v && bytes) {
    for (int64_t i = 0; i < D.H.vsize; ++i)
      if (!std::isfinite(v[i])) throw VmpsError(VMPS_ERR_NONFINITE, "non-finite entry in the start vector");
    VMPS_HIP_CHECK(hipMemcpyAsync(D.vin.p, v, bytes, hipMemcpyHostToDevice, D.stream));
  } else if (bytes) {
    vmps::dev_fill_uniform(D.vin.p, D.H.vsize, opts->seed, D.stream);
  }
  if (!plan->bk) {
    plan->bk = std::make_unique<vmps::BatchKrylov>();
    plan->bk->setup(D, plan->offsets);
  }
  vmps::LanczosOpts lo;
  lo.max_iter = opts->max_iter;
  lo.min_iter = opts->min_iter;
  lo.tol_eigval = opts->tol_eigval;
  lo.tol_state = opts->tol_state;
  lo.graph = opts->graph != 0;
  const auto prec = (vmps_precision)opts->precision;
  const bool refine = prec != VMPS_PREC_FP64 && opts->refine_fp64;
  if (refine) {  // the mixed phase stops at its rounding level (as vmps_lanczos)
    const double state_floor = prec == VMPS_PREC_FP32 ? 1e-6 : 2e-4;
    lo.tol_state = std::max(lo.tol_state, state_floor);
    lo.tol_eigval = std::max(lo.tol_eigval, state_floor * state_floor);
  }
  std::vector<vmps::LanczosResult> r = vmps::lanczos_solve_batch(D, *plan->bk, D.vin.p, lo, prec);
  for (int p = 0; p < P; ++p) {
    res[p].iterations = r[p].iterations;
    res[p].hxv_calls = r[p].hxv_calls;
    res[p].energy = res[p].energy_mixed = r[p].energy;
    res[p].residual = res[p].residual_mixed = r[p].residual;
    res[p].converged = r[p].converged;
    res[p].hxv_fp64 = prec == VMPS_PREC_FP64 ? r[p].hxv_calls : 0;
  }
  if (refine) {
    vmps::LanczosOpts ro = lo;
    ro.tol_eigval = opts->refine_tol_eigval;
    ro.tol_state = opts->refine_tol_state;
    ro.min_iter = 2;
    std::vector<vmps::LanczosResult> f = vmps::lanczos_solve_batch(D, *plan->bk, D.vin.p, ro, VMPS_PREC_FP64);
    for (int p = 0; p < P; ++p) {
      res[p].iterations_refine = f[p].iterations;
      res[p].hxv_calls += f[p].hxv_calls;
      res[p].hxv_fp64 += f[p].hxv_calls;
      res[p].energy = f[p].energy;
      res[p].residual = f[p].residual;
      res[p].converged = f[p].converged;
    }
  }
  if (v && bytes) {
    VMPS_HIP_CHECK(hipMemcpyAsync(v, D.vin.p, bytes, hipMemcpyDeviceToHost, D.stream));
    VMPS_HIP_CHECK(hipStreamSynchronize(D.stream));
  }
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  for (int p = 0; p < P; ++p) res[p].ms = ms;
  VMPS_API_END
}

vmps_status vmps_plan_time_stages(vmps_plan* plan, int warmup, int iters, double* ms) {
  VMPS_API_BEGIN
  require(plan && ms && iters > 0 && warmup >= 0, VMPS_ERR_ARG, "bad argument");
  auto& D = plan->D;
  for (int i = 0; i < warmup; ++i) D.hxv(D.vin.p, D.vout.p, D.stream);
  for (int st = 0; st < 4; ++st) {
    VMPS_HIP_CHECK(hipEventRecord(D.ev0, D.stream));
    for (int i = 0; i < iters; ++i) {
      if (st < 3)
        D.launch_stage(st, D.vin.p, D.vout.p, D.stream);
      else
        D.hxv(D.vin.p, D.vout.p, D.stream);
    }
    VMPS_HIP_CHECK(hipEventRecord(D.ev1, D.stream));
    VMPS_HIP_CHECK(hipEventSynchronize(D.ev1));
    float t = 0.f;
    VMPS_HIP_CHECK(hipEventElapsedTime(&t, D.ev0, D.ev1));
    ms[st] = (double)t / iters;
  }
  VMPS_API_END
}

}  // extern "C"
