// Device plan: the staged block-triple tables resident on one GPU, plus the
// kernels that execute one effective-Hamiltonian application on it.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "errors.hpp"
#include "pivot.hpp"

namespace vmps {

#define VMPS_HIP_CHECK(x)                                                                  \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess)                                                                  \
      throw ::vmps::VmpsError(e_ == hipErrorOutOfMemory ? VMPS_ERR_OOM : VMPS_ERR_HIP,     \
                              std::string(#x) + ": " + hipGetErrorString(e_));             \
  } while (0)

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    release();
    n = count;
    if (count) VMPS_HIP_CHECK(hipMalloc((void**)&p, count * sizeof(T)));
  }
  void upload(const std::vector<T>& h, hipStream_t s) {
    alloc(h.size());
    if (!h.empty()) VMPS_HIP_CHECK(hipMemcpyAsync(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

struct DevPlan;

// Reusable non-blocking HIP streams, one free list per device.  Creating and
// destroying a stream costs ~0.8 ms on MI355X (measured), more than a whole
// HxV of a small pivot; plans and environment updates are created once per
// DMRG step, so they borrow a stream here instead.  A stream is returned
// only after the caller has synchronised it.  Default priority, like every
// stream of this library.
hipStream_t acquire_stream();         // on the current device
void release_stream(hipStream_t s);  // s idle; back to the current device's list
void release_stream(hipStream_t s, int device);  // s idle, created on `device`

// Krylov workspace: basis Q ((cap+1) x n), work vector w, fixed HxV input v,
// Gram-Schmidt coefficients c (two passes of cap + 1), the iteration index
// in device memory, the three scalars of every iteration (out3, 4 per
// iteration), and the HIP graph of kIterBatch iterations.  cap is the
// solve's iteration limit plus kIterBatch, room for the iterations a batch
// runs past the limit or past convergence.
struct Krylov {
  static constexpr int kIterBatch = 4;
  DevBuf<double> Q, w, v, c, scal, out3;
  DevBuf<int> kdev;
  DevBuf<double> ew;      // coeff_dot weight of every vector element
  DevBuf<double> gpart;   // partial sums of the basis dot products, (kmax + 3) x ngrp
  DevBuf<double> npart;   // partial sums of ||w||^2
  DevBuf<double> om;      // partial re-orthogonalisation: rows k-1, k, k+1 of the omega estimate
  DevBuf<int> gate;       // gate[k]: iteration k ran the full Gram-Schmidt pass
  double* host3 = nullptr;  // pinned readback of out3: two slots of 4 * kIterBatch doubles
  hipEvent_t done[2] = {nullptr, nullptr};  // a slot's readback has landed
  int ngrp = 0;           // element chunks of the basis dot products
  int64_t chunk = 0;
  int cap = 0;
  int64_t nvec = -1;
  int device = 0;  // where Q lives
  // one captured iteration per precision (a mixed solve and its FP64
  // refinement each keep theirs), valid for the plan generation it was made at
  hipGraph_t graph[3] = {nullptr, nullptr, nullptr};
  hipGraphExec_t exec[3] = {nullptr, nullptr, nullptr};
  hipGraphExec_t exec2[3] = {nullptr, nullptr, nullptr};  // a second instance of each graph
  int graph_gen[3] = {-1, -1, -1};
  void ensure(DevPlan& D, int kmax);
  void enqueue_iteration(DevPlan& D, hipStream_t s, int kofs);  // iteration *kdev + kofs; *kdev is not advanced
  // FP64 solves on vectors of the fused Gram-Schmidt path re-orthogonalise
  // partially (local passes, the whole basis when the omega estimate asks)
  bool partial_reortho(const DevPlan& D) const;
  bool full_reortho = false;  // set for a solve that found the partial scheme unreliable
  // an executable instance (copy 0 or 1) of the graph of D.prec (kIterBatch iterations)
  hipGraphExec_t ensure_graph(DevPlan& D, int copy);
  void release_graph();
  Krylov() = default;
  Krylov(const Krylov&) = delete;
  Krylov& operator=(const Krylov&) = delete;
  ~Krylov();
};

struct DevPlan {
  int device = 0;
  int ncu = 0;  // compute units of the device
  int gen = 0;  // bumped whenever device work lists or operand buffers change (invalidates captured graphs)
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  vmps_precision prec = VMPS_PREC_FP64;
  HostPlan H;

  // FP64 operands and intermediates
  DevBuf<double> L, R, T, vin, vout;
  // FP32 operands and intermediates (mixed modes)
  DevBuf<float> L32, R32, V32, T32;
  // explicit stage B (H.explicit_u): U in FP64 / FP32 and its task tables
  DevBuf<double> Ub;
  DevBuf<float> U32;
  DevBuf<CombTask> combTasks;
  DevBuf<TermC> combTerms;

  // split-BF16 x3 operands, intermediates and work lists
  DevBuf<uint16_t> Lhi, Llo, Rhi, Rlo, Vhi, Vlo;
  DevBuf<float> Tp;
  DevBuf<TileA16> bfA, bfAbig;
  DevBuf<TileC> bfCsmall, bfCbig;
  DevBuf<int32_t> wTile, wBegin, wEnd, rTile, rNs;  // split-K work items / reductions
  DevBuf<int64_t> wSlot, rSlot;
  DevBuf<float> Cpart;                              // FP32 partial tiles
  DevBuf<ChunkC16> bfC;
  DevBuf<TermC16> bfT;
  DevBuf<VConv> vconv;
  DevBuf<float> Up;                     // explicit stage B: padded FP32 U
  DevBuf<CombTask> bfCombTasks;
  DevBuf<TermC16> bfCombTerms;

  // work lists
  DevBuf<TileA> tilesA;
  DevBuf<TileC> tilesC;
  DevBuf<TileA> tilesAb;  // LDS-staged 128 x 128 tiles (H.big_min64)
  DevBuf<TileC> tilesCb;
  DevBuf<ChunkC> chunksC;
  DevBuf<TermC> termsC;
  DevBuf<int64_t> seg_off;
  DevBuf<int32_t> seg_len;
  DevBuf<double> seg_w;
  DevBuf<double> partial;  // dot-product partial sums
  DevBuf<double> scalars;  // small device scalars (dot results, projector overlaps)

  Krylov krylov;

  // excited-state projectors
  int nproj = 0;
  double Epenalty = 0.0;
  DevBuf<double> A0;

  ~DevPlan();
  void create(const Pivot& P, vmps_precision prec);
  void create_triples(const TripleSet& S, vmps_precision prec);
  void create_triples(TripleSet&& S, vmps_precision prec);  // takes S's environment data
  void upload(vmps_precision prec);  // device copies of H
  void ensure_precision_buffers();
  void ensure_bf16_buffers();
  void upload_bf16_lists(const Bf16Layout& B);
  bool bf16_ready = false;
  // switch the tile configuration: FP64 / FP32 wave tile edge (16 or 32) and
  // big-tile threshold (H.big_min64), the split-BF16 big-tile threshold
  // (H.big_min) and fused / explicit stage B
  // (H.explicit_u); rebuilds and uploads the work lists only
  void retile(int wave_edge, int big_min, int big_min64, bool explicit_u);
  void upload_lists();  // FP64 / FP32 work lists and the explicit stage B buffers
  void launch_stage_bf16x3(int stage, const double* d_vin, double* d_vout, hipStream_t s);
  void hxv(const double* d_vin, double* d_vout, hipStream_t s);
  // one stage of hxv alone: 0 = L.V, 2 = (sum c T).R; stage 1 (the combine) is fused into 2
  void launch_stage(int stage, const double* d_vin, double* d_vout, hipStream_t s);
  // workgroups (4 waves each) of a kernel whose waves loop over `items` work
  // items: one per wave when there are few, else as many as are resident at
  // once (`resident` workgroups per compute unit)
  int persistent_blocks(int64_t items, int resident) const {
    const int64_t cap = (int64_t)std::max(resident, 1) * std::max(ncu, 1);
    return (int)std::max<int64_t>(1, std::min<int64_t>((items + 3) / 4, cap));
  }
  int occ_a16 = 0;  // resident workgroups per compute unit of the split-BF16 small-tile stage A kernel
  double dot(const double* d_x, const double* d_y, hipStream_t s);
  void dot_async(const double* d_x, const double* d_y, double* d_result, hipStream_t s);
};

}  // namespace vmps
