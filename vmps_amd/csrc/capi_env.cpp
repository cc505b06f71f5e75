// extern "C" entry points of the environment updates (include/vmps_hxv.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>

#include "denv.hpp"
#include "engine.hpp"
#include "env.hpp"
#include "errors.hpp"
#include "vmps_hxv.h"

struct vmps_env {
  vmps::EnvProblem P;
  vmps::TripleSet S;
  vmps::Stats stats;
  std::vector<double> result;  // output vector in S.Ob order
  bool done = false;
  bool result_on_device = false;  // the result went to a vmps_denv
};

using vmps::VmpsError;

#define VMPS_API_BEGIN try {
#define VMPS_API_END                                \
  return VMPS_OK;                                   \
  }                                                 \
  catch (const VmpsError& e) {                      \
    vmps::set_last_error(e.what());                 \
    return e.status;                                \
  }                                                 \
  catch (const std::bad_alloc&) {                   \
    vmps::set_last_error("host allocation failed"); \
    return VMPS_ERR_OOM;                            \
  }                                                 \
  catch (const std::exception& e) {                 \
    vmps::set_last_error(e.what());                 \
    return VMPS_ERR_INTERNAL;                       \
  }

namespace {
vmps::Qn qn(const int* q, int nq) {
  vmps::Qn r;
  for (int i = 0; i < nq; ++i) r[i] = q[i];
  return r;
}
void require(bool cond, int status, const char* msg) {
  if (!cond) throw VmpsError(status, msg);
}
// one sector of Lold / Rold from per-block pointers (absent blocks: rows * cols == 0)
void add_old_sector(vmps::Tripod& E, int nq, const int* qin, const int* qout, const int* qmid, int adim,
                    const int64_t* rows, const int64_t* cols, const double* const* ptrs) {
  require(!E.dev, VMPS_ERR_STATE, "the environment is already given on the device");
  vmps::Key3 key{qn(qin, nq), qn(qout, nq), qn(qmid, nq)};
  require(!E.dict.count(key), VMPS_ERR_LAYOUT, "duplicate environment sector");
  for (int a = 0; a < adim; ++a) {
    require(rows[a] >= 0 && cols[a] >= 0, VMPS_ERR_ARG, "negative block dimension");
    require(rows[a] * cols[a] == 0 || ptrs[a], VMPS_ERR_ARG, "NULL block data");
  }
  std::vector<vmps::DenseBlock> blocks(adim);
  for (int a = 0; a < adim; ++a) {
    blocks[a].rows = rows[a];
    blocks[a].cols = cols[a];
    if (rows[a] * cols[a] > 0) {
      blocks[a].off = (int64_t)E.data.size();
      E.data.insert(E.data.end(), ptrs[a], ptrs[a] + rows[a] * cols[a]);
    }
  }
  E.index.push_back({key.a, key.b, key.c});
  E.block.push_back(std::move(blocks));
  E.dict[key] = E.index.size() - 1;
}

// sectors given as nsec items (qin, qout, qmid: nq ints each; adim slots
// each; rows, cols, blocks: one per slot)
void add_old_sectors(vmps::Tripod& E, int nq, int nsec, const int* qin, const int* qout, const int* qmid,
                     const int* adim, const int64_t* rows, const int64_t* cols, const double* const* blocks) {
  require(nsec >= 0 && (nsec == 0 || (qin && qout && qmid && adim)), VMPS_ERR_ARG, "bad argument");
  {  // one reservation for all sectors
    int64_t total = 0, nslot = 0;
    for (int q = 0; q < nsec; ++q) nslot += std::max(0, adim[q]);
    if (rows && cols)
      for (int64_t i = 0; i < nslot; ++i) total += std::max<int64_t>(0, rows[i]) * std::max<int64_t>(0, cols[i]);
    E.data.reserve(E.data.size() + (size_t)total);
  }
  int64_t slot = 0;
  for (int q = 0; q < nsec; ++q) {
    require(adim[q] >= 0, VMPS_ERR_ARG, "adim < 0");
    require(adim[q] == 0 || (rows && cols && blocks), VMPS_ERR_ARG, "NULL rows/cols/blocks");
    add_old_sector(E, nq, qin + q * nq, qout + q * nq, qmid + q * nq, adim[q], rows + slot, cols + slot,
                   blocks + slot);
    slot += adim[q];
  }
}

// copies the packed blocks of E to a new allocation on the current device
// and empties E.data
void move_to_device(vmps_denv* d) {
  VMPS_HIP_CHECK(hipGetDevice(&d->device));
  const int64_t n = (int64_t)d->T.data.size();
  double* buf = nullptr;
  VMPS_HIP_CHECK(hipMalloc((void**)&buf, std::max<int64_t>(1, n) * sizeof(double)));
  d->T.dev_hold = vmps::own_device(buf, d->device);
  if (n) VMPS_HIP_CHECK(hipMemcpy(buf, d->T.data.data(), n * sizeof(double), hipMemcpyHostToDevice));
  d->T.data.clear();
  d->T.data.shrink_to_fit();
  d->T.dev = buf;
  d->T.dev_size = n;
}

namespace {
// An environment update whose kernels were queued and not waited for: the
// plan that runs them and the device blocks they read (the old environment)
// and write (the result) stay alive until wait() has seen them finish.
class PendingUpdate final : public vmps::DevReady {
 public:
  PendingUpdate(std::unique_ptr<vmps::DevPlan> plan, hipEvent_t done, int device,
                std::shared_ptr<const double> reads, std::shared_ptr<const double> writes)
      : plan_(std::move(plan)), done_(done), device_(device), reads_(std::move(reads)), writes_(std::move(writes)) {}
  PendingUpdate(const PendingUpdate&) = delete;
  PendingUpdate& operator=(const PendingUpdate&) = delete;
  ~PendingUpdate() override {
    finish(false);
    if (done_) (void)hipEventDestroy(done_);
  }
  void wait() override { finish(true); }

 private:
  void finish(bool check) {
    std::lock_guard<std::mutex> lock(m_);
    if (!plan_) return;
    int cur = -1;
    const bool other = hipGetDevice(&cur) == hipSuccess && cur != device_;
    if (other) (void)hipSetDevice(device_);
    const hipError_t st = hipEventSynchronize(done_);
    plan_.reset();  // its stream is idle now
    reads_.reset();
    writes_.reset();
    if (other) (void)hipSetDevice(cur);
    if (check && st != hipSuccess)
      throw VmpsError(VMPS_ERR_HIP, std::string("environment update: ") + hipGetErrorString(st));
  }
  std::mutex m_;
  std::unique_ptr<vmps::DevPlan> plan_;
  hipEvent_t done_ = nullptr;
  int device_ = -1;
  std::shared_ptr<const double> reads_, writes_;
};
}  // namespace

void contract(vmps_env* e, vmps_precision prec, vmps_denv* dres) {
  require(e != nullptr, VMPS_ERR_ARG, "NULL env");
  require(prec == VMPS_PREC_FP64 || prec == VMPS_PREC_FP32 || prec == VMPS_PREC_BF16X3, VMPS_ERR_ARG,
          "unknown precision");
  if (e->P.old.dev) {
    int dev = -1;
    VMPS_HIP_CHECK(hipGetDevice(&dev));
    require(dev == e->P.old.dev_device, VMPS_ERR_ARG, "the old environment lives on another device");
  }
  vmps::plan_env(e->P, e->S);
  e->done = false;
  e->result_on_device = dres != nullptr;
  e->result.clear();
  if (!dres) e->result.assign(e->S.osize, 0.0);
  const int64_t osize = e->S.osize;
  double* dout = nullptr;  // device result handed to dres
  if (!e->S.terms.empty() && osize > 0) {
    auto plan = std::make_unique<vmps::DevPlan>();
    vmps::DevPlan& D = *plan;
    D.create_triples(std::move(e->S), prec);  // the plan takes the packed operands; e->S keeps the layout
    e->stats = D.H.stats;
    const size_t ib = (size_t)D.H.ysize * sizeof(double), ob = (size_t)D.H.osize * sizeof(double);
    if (ib) {
      e->P.old.wait_device();  // an old environment still being written by its own update
      VMPS_HIP_CHECK(hipMemcpyAsync(D.vin.p, e->P.old.dev ? e->P.old.dev : e->P.old.data.data(), ib,
                                    e->P.old.dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, D.stream));
    }
    D.hxv(D.vin.p, D.vout.p, D.stream);
    // A device result from device (or no) input returns with its kernels
    // queued: the host goes on (the next pivot's plan) while they run, and
    // every reader of the result waits on their event first.  Host input
    // (a copy from pageable memory that the caller may free) and host
    // results wait here.
    const bool queued = dres && (!ib || e->P.old.dev);
    if (dres) {
      dout = D.vout.p;  // the plan gives up its output buffer to dres
      D.vout.p = nullptr;
      D.vout.n = 0;
      VMPS_HIP_CHECK(hipGetDevice(&dres->device));
      dres->T.dev_hold = vmps::own_device(dout, dres->device);
    } else {
      VMPS_HIP_CHECK(hipMemcpyAsync(e->result.data(), D.vout.p, ob, hipMemcpyDeviceToHost, D.stream));
    }
    if (queued) {
      hipEvent_t done = nullptr;
      VMPS_HIP_CHECK(hipEventCreateWithFlags(&done, hipEventDisableTiming));
      const hipError_t st = hipEventRecord(done, D.stream);
      if (st != hipSuccess) {
        (void)hipEventDestroy(done);
        VMPS_HIP_CHECK(st);
      }
      dres->T.dev_ready = std::make_shared<PendingUpdate>(std::move(plan), done, dres->device, e->P.old.dev_hold,
                                                          dres->T.dev_hold);
    } else {
      VMPS_HIP_CHECK(hipStreamSynchronize(D.stream));
    }
  } else {
    e->stats = vmps::Stats();
  }
  if (dres) {
    dres->nq = e->P.sym.nq;
    VMPS_HIP_CHECK(hipGetDevice(&dres->device));
    if (!dout) {  // nothing to contract: a zero environment of the result's layout
      VMPS_HIP_CHECK(hipMalloc((void**)&dout, std::max<int64_t>(1, osize) * sizeof(double)));
      dres->T.dev_hold = vmps::own_device(dout, dres->device);
      if (osize) VMPS_HIP_CHECK(hipMemset(dout, 0, osize * sizeof(double)));
    }
    vmps::Tripod& T = dres->T;
    T.dev = dout;
    T.dev_size = osize;
    T.dev_device = dres->device;
    for (size_t q = 0; q < e->P.new_index.size(); ++q) {
      std::vector<vmps::DenseBlock> blocks((size_t)e->P.new_adim[q]);
      for (int a = 0; a < e->P.new_adim[q]; ++a) {
        const int64_t id = e->P.new_block[q][a];
        if (id < 0) continue;
        const vmps::BlockDim& b = e->S.Ob[id];
        blocks[a].rows = b.rows;
        blocks[a].cols = b.cols;
        if (b.rows * b.cols > 0) blocks[a].off = b.off;
      }
      T.index.push_back(e->P.new_index[q]);
      T.block.push_back(std::move(blocks));
      T.dict[vmps::Key3{e->P.new_index[q][0], e->P.new_index[q][1], e->P.new_index[q][2]}] = q;
    }
  }
  e->done = true;
}

// sector layout of T, and (data != NULL) its blocks packed in slot order
void tripod_get(const vmps::Tripod& T, int nq, const double* src, int* qin, int* qout, int* qmid, int* adim,
                int64_t* rows, int64_t* cols, double* data) {
  int64_t slot = 0, pos = 0;
  for (size_t q = 0; q < T.dim(); ++q) {
    for (int i = 0; i < nq; ++i) {
      qin[q * nq + i] = T.index[q][0][i];
      qout[q * nq + i] = T.index[q][1][i];
      qmid[q * nq + i] = T.index[q][2][i];
    }
    adim[q] = (int)T.block[q].size();
    for (const vmps::DenseBlock& b : T.block[q]) {
      rows[slot] = b.present() ? b.rows : 0;
      cols[slot] = b.present() ? b.cols : 0;
      if (data && b.present()) std::memcpy(data + pos, src + b.off, sizeof(double) * b.size());
      if (b.present()) pos += b.size();
      ++slot;
    }
  }
}
}  // namespace

extern "C" {

vmps_status vmps_env_create(int which, int nq, const int* kinds, int nloc, const int* qloc, int nop, const int* qop,
                            vmps_env** out) {
  VMPS_API_BEGIN
  require(out != nullptr, VMPS_ERR_ARG, "out is NULL");
  *out = nullptr;
  require(which == 0 || which == 1, VMPS_ERR_ARG, "which must be 0 (contract_L) or 1 (contract_R)");
  require(nq >= 1 && nq <= vmps::kMaxNq && kinds, VMPS_ERR_ARG, "bad symmetry");
  require(nloc > 0 && nop > 0 && qloc && qop, VMPS_ERR_ARG, "empty local or operator basis");
  auto e = std::make_unique<vmps_env>();
  e->P.which = which;
  e->P.sym.nq = nq;
  for (int i = 0; i < nq; ++i) {
    require(kinds[i] == VMPS_SYM_U1 || kinds[i] == VMPS_SYM_SU2, VMPS_ERR_ARG, "unknown symmetry kind");
    e->P.sym.kind[i] = kinds[i];
  }
  for (int i = 0; i < nloc; ++i) e->P.qloc.push_back(qn(qloc + i * nq, nq));
  for (int i = 0; i < nop; ++i) e->P.qop.push_back(qn(qop + i * nq, nq));
  e->P.Abra.resize(nloc);
  e->P.Aket.resize(nloc);
  e->P.W.assign(nloc, std::vector<std::vector<vmps::WBiped>>(nloc, std::vector<vmps::WBiped>(nop)));
  *out = e.release();
  VMPS_API_END
}

void vmps_env_destroy(vmps_env* e) { delete e; }

vmps_status vmps_env_add_old(vmps_env* e, const int* qin, const int* qout, const int* qmid, int adim,
                             const int64_t* rows, const int64_t* cols, const double* data) {
  VMPS_API_BEGIN
  require(e && qin && qout && qmid && adim >= 0 && (adim == 0 || (rows && cols)), VMPS_ERR_ARG, "bad argument");
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
  add_old_sector(e->P.old, e->P.sym.nq, qin, qout, qmid, adim, rows, cols, ptrs.data());
  e->done = false;
  VMPS_API_END
}

vmps_status vmps_env_add_A(vmps_env* e, int bra, int s, const int* qin, const int* qout, int64_t rows, int64_t cols,
                           const double* data) {
  VMPS_API_BEGIN
  require(e && qin && qout && rows >= 0 && cols >= 0, VMPS_ERR_ARG, "bad argument");
  require(s >= 0 && s < (int)e->P.qloc.size(), VMPS_ERR_ARG, "s out of range");
  require(rows * cols == 0 || data, VMPS_ERR_ARG, "NULL data");
  const int nq = e->P.sym.nq;
  vmps::DBiped& A = bra ? e->P.Abra[s] : e->P.Aket[s];
  vmps::Key2 key{qn(qin, nq), qn(qout, nq)};
  require(!A.dict.count(key), VMPS_ERR_LAYOUT, "duplicate A block");
  vmps::DenseBlock b;
  b.rows = rows;
  b.cols = cols;
  b.off = (int64_t)A.data.size();
  A.data.insert(A.data.end(), data, data + rows * cols);
  A.in.push_back(key.a);
  A.out.push_back(key.b);
  A.block.push_back(b);
  A.dict[key] = A.in.size() - 1;
  e->done = false;
  VMPS_API_END
}

vmps_status vmps_env_add_W(vmps_env* e, int s1, int s2, int k, const int* qin, const int* qout, int rows, int cols,
                           int64_t nnz, const int* row, const int* col, const double* val) {
  VMPS_API_BEGIN
  require(e && qin && qout && nnz >= 0 && (nnz == 0 || (row && col && val)), VMPS_ERR_ARG, "bad argument");
  const int nloc = (int)e->P.qloc.size(), nop = (int)e->P.qop.size();
  require(s1 >= 0 && s1 < nloc && s2 >= 0 && s2 < nloc && k >= 0 && k < nop, VMPS_ERR_ARG, "W index out of range");
  const int nq = e->P.sym.nq;
  vmps::WBiped& B = e->P.W[s1][s2][k];
  vmps::Key2 key{qn(qin, nq), qn(qout, nq)};
  require(!B.dict.count(key), VMPS_ERR_LAYOUT, "duplicate W block");
  std::vector<int64_t> order(nnz);
  for (int64_t i = 0; i < nnz; ++i) {
    require(row[i] >= 0 && row[i] < rows && col[i] >= 0 && col[i] < cols, VMPS_ERR_ARG, "W nonzero out of range");
    order[i] = i;
  }
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
    return col[a] != col[b] ? col[a] < col[b] : row[a] < row[b];
  });
  vmps::SparseBlock sb;
  sb.rows = rows;
  sb.cols = cols;
  for (int64_t i = 0; i < nnz; ++i) {
    sb.row.push_back(row[order[i]]);
    sb.col.push_back(col[order[i]]);
    sb.val.push_back(val[order[i]]);
  }
  B.in.push_back(key.a);
  B.out.push_back(key.b);
  B.block.push_back(std::move(sb));
  B.dict[key] = B.in.size() - 1;
  e->done = false;
  VMPS_API_END
}

// Whole-Tripod / whole-list forms of the three calls above (one call instead
// of one per sector or block); each item goes through the single-item call.
vmps_status vmps_env_add_old_sectors(vmps_env* e, int nsec, const int* qin, const int* qout, const int* qmid,
                                     const int* adim, const int64_t* rows, const int64_t* cols,
                                     const double* const* blocks) {
  VMPS_API_BEGIN
  require(e != nullptr, VMPS_ERR_ARG, "NULL env");
  add_old_sectors(e->P.old, e->P.sym.nq, nsec, qin, qout, qmid, adim, rows, cols, blocks);
  e->done = false;
  VMPS_API_END
}

vmps_status vmps_env_add_A_blocks(vmps_env* e, int bra, int nblk, const int* s, const int* qin, const int* qout,
                                  const int64_t* rows, const int64_t* cols, const double* const* blocks) {
  VMPS_API_BEGIN
  require(e && nblk >= 0 && (nblk == 0 || (s && qin && qout && rows && cols && blocks)), VMPS_ERR_ARG,
          "bad argument");
  const int nq = e->P.sym.nq;
  {  // one reservation per local index
    std::vector<vmps::DBiped>& As = bra ? e->P.Abra : e->P.Aket;
    std::vector<size_t> add(As.size(), 0);
    for (int b = 0; b < nblk; ++b)
      if (s[b] >= 0 && s[b] < (int)As.size() && rows[b] > 0 && cols[b] > 0) add[s[b]] += (size_t)(rows[b] * cols[b]);
    for (size_t i = 0; i < As.size(); ++i) As[i].data.reserve(As[i].data.size() + add[i]);
  }
  for (int b = 0; b < nblk; ++b) {
    const vmps_status st = vmps_env_add_A(e, bra, s[b], qin + b * nq, qout + b * nq, rows[b], cols[b], blocks[b]);
    if (st != VMPS_OK) return st;
  }
  VMPS_API_END
}

vmps_status vmps_env_add_W_blocks(vmps_env* e, int nblk, const int* idx, const int* qin, const int* qout,
                                  const int* dims, const int64_t* nnz, const int* row, const int* col,
                                  const double* val) {
  VMPS_API_BEGIN
  require(e && nblk >= 0 && (nblk == 0 || (idx && qin && qout && dims && nnz)), VMPS_ERR_ARG, "bad argument");
  const int nq = e->P.sym.nq;
  int64_t pos = 0;
  for (int b = 0; b < nblk; ++b) {
    require(nnz[b] >= 0, VMPS_ERR_ARG, "nnz < 0");
    const vmps_status st = vmps_env_add_W(e, idx[3 * b], idx[3 * b + 1], idx[3 * b + 2], qin + b * nq, qout + b * nq,
                                          dims[2 * b], dims[2 * b + 1], nnz[b], row ? row + pos : nullptr,
                                          col ? col + pos : nullptr, val ? val + pos : nullptr);
    if (st != VMPS_OK) return st;
    pos += nnz[b];
  }
  VMPS_API_END
}

vmps_status vmps_env_contract(vmps_env* e, vmps_precision prec) {
  VMPS_API_BEGIN
  contract(e, prec, nullptr);
  VMPS_API_END
}

vmps_status vmps_env_contract_device(vmps_env* e, vmps_precision prec, vmps_denv** out) {
  VMPS_API_BEGIN
  require(out != nullptr, VMPS_ERR_ARG, "out is NULL");
  *out = nullptr;
  auto d = std::make_unique<vmps_denv>();
  contract(e, prec, d.get());
  *out = d.release();
  VMPS_API_END
}

vmps_status vmps_env_add_old_device(vmps_env* e, const vmps_denv* E) {
  VMPS_API_BEGIN
  require(e && E, VMPS_ERR_ARG, "NULL argument");
  require(E->nq == e->P.sym.nq, VMPS_ERR_ARG, "symmetry mismatch");
  require(e->P.old.dim() == 0 && !e->P.old.dev, VMPS_ERR_STATE, "the old environment already has sectors");
  e->P.old = E->T;
  e->done = false;
  VMPS_API_END
}

vmps_status vmps_denv_create(int nq, int nsec, const int* qin, const int* qout, const int* qmid, const int* adim,
                             const int64_t* rows, const int64_t* cols, const double* const* blocks,
                             vmps_denv** out) {
  VMPS_API_BEGIN
  require(out != nullptr, VMPS_ERR_ARG, "out is NULL");
  *out = nullptr;
  require(nq >= 1 && nq <= vmps::kMaxNq, VMPS_ERR_ARG, "bad nq");
  auto d = std::make_unique<vmps_denv>();
  d->nq = nq;
  add_old_sectors(d->T, nq, nsec, qin, qout, qmid, adim, rows, cols, blocks);
  move_to_device(d.get());
  d->T.dev_device = d->device;
  *out = d.release();
  VMPS_API_END
}

void vmps_denv_destroy(vmps_denv* d) { delete d; }

vmps_status vmps_denv_size(const vmps_denv* d, int64_t* nsec, int64_t* nslots, int64_t* size) {
  VMPS_API_BEGIN
  require(d && nsec && nslots && size, VMPS_ERR_ARG, "NULL argument");
  *nsec = (int64_t)d->T.dim();
  int64_t ns = 0, n = 0;
  for (const auto& blocks : d->T.block) {
    ns += (int64_t)blocks.size();
    for (const auto& b : blocks)
      if (b.present()) n += b.size();
  }
  *nslots = ns;
  *size = n;
  VMPS_API_END
}

vmps_status vmps_denv_get(const vmps_denv* d, int* qin, int* qout, int* qmid, int* adim, int64_t* rows,
                          int64_t* cols, double* data) {
  VMPS_API_BEGIN
  require(d && qin && qout && qmid && adim && rows && cols, VMPS_ERR_ARG, "NULL argument");
  std::vector<double> host;
  if (data && d->T.dev_size) {
    int dev = -1;
    VMPS_HIP_CHECK(hipGetDevice(&dev));
    require(dev == d->device, VMPS_ERR_ARG, "the environment lives on another device");
    host.resize((size_t)d->T.dev_size);
    d->T.wait_device();
    VMPS_HIP_CHECK(hipMemcpy(host.data(), d->T.dev, host.size() * sizeof(double), hipMemcpyDeviceToHost));
  }
  tripod_get(d->T, d->nq, host.data(), qin, qout, qmid, adim, rows, cols, data);
  VMPS_API_END
}

vmps_status vmps_env_result_size(const vmps_env* e, int64_t* nsec, int64_t* nslots, int64_t* size) {
  VMPS_API_BEGIN
  require(e && nsec && nslots && size, VMPS_ERR_ARG, "NULL argument");
  require(e->done, VMPS_ERR_STATE, "vmps_env_contract has not run");
  *nsec = (int64_t)e->P.new_index.size();
  int64_t ns = 0;
  for (int a : e->P.new_adim) ns += a;
  *nslots = ns;
  *size = e->S.osize;
  VMPS_API_END
}

vmps_status vmps_env_result_get(const vmps_env* e, int* qin, int* qout, int* qmid, int* adim, int64_t* rows,
                                int64_t* cols, double* data) {
  VMPS_API_BEGIN
  require(e && qin && qout && qmid && adim && rows && cols, VMPS_ERR_ARG, "NULL argument");
  require(e->done, VMPS_ERR_STATE, "vmps_env_contract has not run");
  require(!e->result_on_device, VMPS_ERR_STATE, "the result was handed to a device environment");
  const int nq = e->P.sym.nq;
  int64_t slot = 0, pos = 0;
  for (size_t q = 0; q < e->P.new_index.size(); ++q) {
    for (int i = 0; i < nq; ++i) {
      qin[q * nq + i] = e->P.new_index[q][0][i];
      qout[q * nq + i] = e->P.new_index[q][1][i];
      qmid[q * nq + i] = e->P.new_index[q][2][i];
    }
    adim[q] = e->P.new_adim[q];
    for (int a = 0; a < e->P.new_adim[q]; ++a, ++slot) {
      const int64_t id = e->P.new_block[q][a];
      if (id < 0) {
        rows[slot] = cols[slot] = 0;
        continue;
      }
      const vmps::BlockDim& b = e->S.Ob[id];
      rows[slot] = b.rows;
      cols[slot] = b.cols;
      if (data) std::memcpy(data + pos, e->result.data() + b.off, sizeof(double) * b.rows * b.cols);
      pos += b.rows * b.cols;
    }
  }
  VMPS_API_END
}

vmps_status vmps_env_stats(const vmps_env* e, vmps_plan_stats* st) {
  VMPS_API_BEGIN
  require(e && st, VMPS_ERR_ARG, "NULL argument");
  std::memset(st, 0, sizeof(*st));
  st->n_terms = e->stats.n_terms;
  st->n_T = e->stats.n_T;
  st->n_U = e->stats.n_U;
  st->n_out = e->stats.n_out;
  st->flops_alg = e->stats.flops_alg;
  st->flops_ref = e->stats.flops_ref;
  st->flops_A = e->stats.flops_A;
  st->flops_C = e->stats.flops_C;
  VMPS_API_END
}

}  // extern "C"
