// qLogNEI value(+grad) for many independent models in one step: five launches
// for the whole batch, whatever its size, each covering every model.
//
// Each stage of the one-model step (qlognei.hip) becomes one launch over a
// table of per-model launch arguments (built once by bta_qlognei_batch_prepare
// with qlognei_step_args, copied into the caller's workspace):
//   covariance tiles      block -> (model, column tile, row tile)
//   V = K_x L^-T          persistent workers over one list of (model, tile) of
//                         all models, longest K range first
//   U = V L^-1 + MC       workers over the adjoint tile list, then one
//                         workgroup per (model, restart) for the MC stage
//   gradient tiles        as the covariance tiles
//   finish                block -> (model, 256 lanes of its 4 R d)
// A block finds its model by binary search over the prefix sums of the
// per-model block counts.  The bodies are the one-model step's, so a model's
// results do not depend on the batch it runs in.  Nothing in the step
// synchronises, copies or clears memory, so it can be captured in a HIP graph.
// Models may differ in n, d, n_b and b; the batch shares q, the kernel kind,
// the mixed precision and the presence of the gradient.  Stands in for
// evaluating the acquisition functions of many SingleTaskGPs (one
// optimize_acqf step each, generation/gen.py:423-485) model after model.
#include "bta_common.h"
#include "gemm.h"
#include "gemm_h2_body.h"
#include "kernels.h"
#include "qlognei_step.h"
#include "qstep_body.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace bta {

// The item whose block range [start[i], start[i + 1]) holds blk (start[0] = 0,
// start[nm] = total).
__device__ __forceinline__ int find_item(const int* __restrict__ start, int nm, int blk) {
  int lo = 0, hi = nm;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (start[mid] <= blk) lo = mid;
    else hi = mid;
  }
  return lo;
}

template <int KD, int KIND>
__global__ __launch_bounds__(256) void qstep_covar_list_kernel(const QsCovarArgs* __restrict__ tab,
                                                               const int* __restrict__ start,
                                                               int nm) {
  const int m = find_item(start, nm, blockIdx.x);
  const QsCovarArgs a = tab[m];
  const int local = (int)blockIdx.x - start[m];
  const int nct = (int)(a.n_pad / QS_COL_TILE);
  qstep_covar_body<KD, 2, KIND>(a, local % nct, local / nct);
}

template <int KD, int KIND>
__global__ __launch_bounds__(256) void qstep_grad_list_kernel(const QsGradArgs* __restrict__ tab,
                                                              const int* __restrict__ start,
                                                              int nm) {
  const int m = find_item(start, nm, blockIdx.x);
  const QsGradArgs a = tab[m];
  const int local = (int)blockIdx.x - start[m];
  const int nct = (int)(a.n_pad / QS_COL_TILE);
  qstep_grad_body<KD, float, KIND>(a, local % nct, local / nct);
}

__global__ void qstep_finish_list_kernel(const FinishArgs* __restrict__ tab,
                                         const int* __restrict__ start, int nm) {
  const int m = find_item(start, nm, blockIdx.x);
  const FinishArgs f = tab[m];
  const int64_t gid = (int64_t)((int)blockIdx.x - start[m]) * blockDim.x + threadIdx.x;
  qstep_finish_body(f.Dpart, f.nct, f.R, f.d, f.extra, f.ls, f.dX, gid);
}

__global__ __launch_bounds__(256, 2) void gemm_h2_list_kernel(const H2Args* __restrict__ tab,
                                                              const int4* __restrict__ tiles,
                                                              int total) {
  __shared__ __attribute__((aligned(1024))) unsigned char lds[FUSED_LDS];
  gemm_h2_list_body<64>(tab, tiles, total, blockIdx.x, gridDim.x, lds);
}

// Workgroups [0, W): adjoint GEMM workers; [W, W + mc blocks): the MC stage,
// one workgroup per (model, restart), or per (model, two restarts) for a model
// whose McArgs::nr is 2, as in its one-model step.
template <int Q>
__global__ __launch_bounds__(256, 2) void adjoint_mc_list_kernel(
    const H2Args* __restrict__ adj, const int4* __restrict__ tiles, int total, int W,
    const McArgs* __restrict__ mc, const int* __restrict__ mc_start, int nm) {
  __shared__ __attribute__((aligned(1024))) unsigned char lds[FUSED_LDS];
  if ((int)blockIdx.x < W) {
    gemm_h2_list_body<64>(adj, tiles, total, blockIdx.x, W, lds);
  } else {
    const int blk = (int)blockIdx.x - W;
    const int m = find_item(mc_start, nm, blk);
    const McArgs a = mc[m];
    if (a.nr == 2)
      mc_body<Q, float, 2>(a, blk - mc_start[m], (double*)lds);
    else
      mc_body<Q, float, 1>(a, blk - mc_start[m], (double*)lds);
  }
}

// ---------------------------------------------------------------------------
// Host side: the plan.
// ---------------------------------------------------------------------------
struct BatchTiles {
  std::vector<int4> fwd, adj;
};

// K chunks of 32 of tile (tm, tn) of a 64-column split GEMM with a triangular
// B (BTA_GEMM_B_LOWER: K range [0, 64 (tn + 1)); B_UPPER: [64 tn, K)).
static int tile_chunks(const H2Args& g, int tn) {
  int kb = 0, ke = g.K;
  if (g.flags & BTA_GEMM_B_LOWER) ke = std::min(ke, (tn + 1) * 64);
  if (g.flags & BTA_GEMM_B_UPPER) kb = std::max(kb, tn * 64);
  return ke > kb ? (ke - kb + 31) / 32 : 0;
}

// Appends every 128 x 64 tile of problem p.
static void add_tiles(const H2Args& g, int p, std::vector<int4>& out) {
  for (int tn = 0; tn < g.N / 64; ++tn)
    for (int tm = 0; tm < g.M / H2_BM; ++tm) out.push_back(make_int4(p, tm, tn, 0));
}

// Longest K range first (ties: problem, then column, then row tile), so that
// h2_tile_of pairs heavy with light tiles across the whole batch.
static void order_tiles(const std::vector<H2Args>& tab, std::vector<int4>& t) {
  std::stable_sort(t.begin(), t.end(), [&](const int4& a, const int4& b) {
    const int ca = tile_chunks(tab[a.x], a.z), cb = tile_chunks(tab[b.x], b.z);
    return ca > cb;
  });
}

// Worker count for a tile list: the schedule of h2_tile_of with worker w on
// CU w % ncu (two workgroups per CU), per-tile cost 0.55 per K chunk plus a
// fixed 6, the candidate with the smallest busiest CU.
static int list_workers(const std::vector<H2Args>& tab, const std::vector<int4>& t) {
  const int total = (int)t.size();
  if (total == 0) return 0;
  const int ncu = num_cus();
  int best = 1;
  double bt = 1e300;
  for (int c : {2 * ncu, 3 * ncu / 2, ncu, ncu / 2}) {
    const int W = std::max(1, std::min(c, total));
    std::vector<double> cu(ncu, 0.0);
    for (int w = 0; w < W; ++w)
      for (int j = 0; j * W < total; ++j) {
        const int k = h2_tile_of(j, w, W);
        if (k < total) cu[w % ncu] += 0.55 * tile_chunks(tab[t[k].x], t[k].z) + 6.0;
      }
    const double m = *std::max_element(cu.begin(), cu.end());
    if (m < bt) {
      bt = m;
      best = W;
    }
  }
  return best;
}

// Offsets (bytes) of the launch tables behind the models' workspaces.
struct BatchLayout {
  std::vector<size_t> item_ws;
  size_t cov, fwd, adj, mc, grad, fin, cov_start, mc_start, fin_start, fwd_tiles, adj_tiles,
      total;
};

static int batch_layout(const bta_qlognei_batch_item* items, int nm, const BatchTiles* tiles,
                        BatchLayout& L) {
  WsPlan p;
  L.item_ws.resize(nm);
  for (int i = 0; i < nm; ++i) {
    if (!items[i].model || !items[i].state || items[i].b < 1 || items[i].b > (1 << 24)) {
      set_error("qlognei_batch: item %d has no model / state or a batch size out of range", i);
      return BTA_ERR_INVALID;
    }
    L.item_ws[i] = p.take(qlognei_ws_bytes(items[i].model, items[i].state, items[i].b));
  }
  L.cov = p.take(nm * sizeof(QsCovarArgs));
  L.fwd = p.take(nm * sizeof(H2Args));
  L.adj = p.take(nm * sizeof(H2Args));
  L.mc = p.take(nm * sizeof(McArgs));
  L.grad = p.take(nm * sizeof(QsGradArgs));
  L.fin = p.take(nm * sizeof(FinishArgs));
  L.cov_start = p.take((nm + 1) * sizeof(int));
  L.mc_start = p.take((nm + 1) * sizeof(int));
  L.fin_start = p.take((nm + 1) * sizeof(int));
  // tile lists: exact when known, else their upper bound (every 128 x 64 tile)
  size_t nf = 0, na = 0;
  if (tiles) {
    nf = tiles->fwd.size();
    na = tiles->adj.size();
  } else {
    for (int i = 0; i < nm; ++i) {
      const int64_t R = items[i].b * items[i].state->q;
      const int64_t t = ((R + 127) / 128) * (items[i].model->n_pad / 64);
      nf += t;
      na += t;
    }
  }
  L.fwd_tiles = p.take(nf * sizeof(int4));
  L.adj_tiles = p.take(na * sizeof(int4));
  L.total = p.off;
  return BTA_OK;
}

size_t qlognei_batch_ws_bytes(const bta_qlognei_batch_item* items, int nm) {
  if (!items || nm < 1) return 0;
  BatchLayout L;
  if (batch_layout(items, nm, nullptr, L)) return 0;
  return L.total;
}

int qlognei_batch_prepare(const bta_qlognei_batch_item* items, int nm, void* ws,
                          size_t ws_bytes, bta_qlognei_batch_plan* plan, hipStream_t s) {
  if (!items || nm < 1 || nm > (1 << 16) || !plan || !ws) {
    set_error("qlognei_batch_prepare: invalid arguments");
    return BTA_ERR_INVALID;
  }
  BatchLayout L;
  int rc = batch_layout(items, nm, nullptr, L);
  if (rc) return rc;
  if (ws_bytes < L.total) {
    set_error("qlognei_batch_prepare: workspace %zu < %zu", ws_bytes, L.total);
    return BTA_ERR_WORKSPACE;
  }
  std::vector<StepArgs> A(nm);
  const int q = items[0].state->q, kind = items[0].model->kind;
  const bool grad = items[0].grad != nullptr;
  int kd = 1;
  for (int i = 0; i < nm; ++i) {
    const bta_qlognei_batch_item& it = items[i];
    rc = qlognei_step_args(it.model, it.state, it.X, it.b, it.value, it.grad, it.status,
                           (char*)ws + L.item_ws[i], ws_bytes - L.item_ws[i], A[i]);
    if (rc) return rc;
    if (!A[i].split || A[i].q != q || it.model->kind != kind || A[i].grad != grad ||
        q > 5) {
      set_error("qlognei_batch_prepare: item %d: every item must use BTA_PREC_MIXED, one q "
                "(1..5), one kernel kind, and give a gradient buffer or none",
                i);
      return BTA_ERR_UNSUPPORTED;
    }
    int bn = 0, wk = 0;  // the split GEMMs' shape and alignment checks
    rc = gemm_h2_plan(A[i].fwd, &bn, &wk);
    if (!rc && grad) rc = gemm_h2_plan(A[i].adj, &bn, &wk);
    if (rc) return rc;
    const McLayout lay(q, it.state->n_b, it.state->S);
    if ((size_t)lay.total * sizeof(double) > MC_LDS_MAX) {
      set_error("qlognei_batch_prepare: item %d: MC stage needs %zu bytes of LDS (> %zu)", i,
                (size_t)lay.total * sizeof(double), MC_LDS_MAX);
      return BTA_ERR_UNSUPPORTED;
    }
    kd = std::max(kd, (it.model->d + 3) / 4);
  }
  std::vector<QsCovarArgs> cov(nm);
  std::vector<H2Args> fwd(nm), adj(nm);
  std::vector<McArgs> mc(nm);
  std::vector<QsGradArgs> gr(nm);
  std::vector<FinishArgs> fin(nm);
  std::vector<int> cov_start(nm + 1, 0), mc_start(nm + 1, 0), fin_start(nm + 1, 0);
  BatchTiles T;
  for (int i = 0; i < nm; ++i) {
    cov[i] = A[i].c;
    fwd[i] = A[i].fwd;
    adj[i] = A[i].adj;
    mc[i] = A[i].mc;
    gr[i] = A[i].g;
    fin[i] = A[i].fin;
    const int64_t cb = (A[i].np / QS_COL_TILE) * (A[i].R_pad / 64);
    const int64_t fb = (4 * A[i].R * A[i].c.d + 255) / 256;
    if ((int64_t)cov_start[i] + cb > (1 << 30) || (int64_t)mc_start[i] + A[i].b > (1 << 30) ||
        (int64_t)fin_start[i] + fb > (1 << 30)) {
      set_error("qlognei_batch_prepare: more than 2^30 blocks in one stage");
      return BTA_ERR_UNSUPPORTED;
    }
    cov_start[i + 1] = cov_start[i] + (int)cb;
    mc_start[i + 1] = mc_start[i] + (int)(A[i].mc.nr == 2 ? (A[i].b + 1) / 2 : A[i].b);
    fin_start[i + 1] = fin_start[i] + (int)fb;
    add_tiles(fwd[i], i, T.fwd);
    if (grad) add_tiles(adj[i], i, T.adj);
  }
  order_tiles(fwd, T.fwd);
  order_tiles(adj, T.adj);
  // the exact layout (tile lists no longer than their bound)
  BatchLayout E;
  rc = batch_layout(items, nm, &T, E);
  if (rc) return rc;
  std::vector<unsigned char> host(E.total - E.cov, 0);
  auto put = [&](size_t off, const void* src, size_t bytes) {
    if (bytes) memcpy(host.data() + (off - E.cov), src, bytes);
  };
  put(E.cov, cov.data(), nm * sizeof(QsCovarArgs));
  put(E.fwd, fwd.data(), nm * sizeof(H2Args));
  put(E.adj, adj.data(), nm * sizeof(H2Args));
  put(E.mc, mc.data(), nm * sizeof(McArgs));
  put(E.grad, gr.data(), nm * sizeof(QsGradArgs));
  put(E.fin, fin.data(), nm * sizeof(FinishArgs));
  put(E.cov_start, cov_start.data(), (nm + 1) * sizeof(int));
  put(E.mc_start, mc_start.data(), (nm + 1) * sizeof(int));
  put(E.fin_start, fin_start.data(), (nm + 1) * sizeof(int));
  put(E.fwd_tiles, T.fwd.data(), T.fwd.size() * sizeof(int4));
  put(E.adj_tiles, T.adj.data(), T.adj.size() * sizeof(int4));
  char* base = (char*)ws;
  BTA_HIP(hipMemcpyAsync(base + E.cov, host.data(), host.size(), hipMemcpyHostToDevice, s));
  BTA_HIP(hipStreamSynchronize(s));  // the host copy of the tables goes out of scope

  bta_qlognei_batch_plan& P = *plan;
  P = bta_qlognei_batch_plan{};
  P.ws = ws;
  P.n_items = nm;
  P.q = q;
  P.kind = kind;
  P.kd = kd;
  P.grad = grad ? 1 : 0;
  P.covar_blocks = cov_start[nm];
  P.mc_blocks = mc_start[nm];
  P.fin_blocks = fin_start[nm];
  P.fwd_tiles = (int)T.fwd.size();
  P.adj_tiles = (int)T.adj.size();
  P.fwd_workers = list_workers(fwd, T.fwd);
  P.adj_workers = grad ? list_workers(adj, T.adj) : 0;
  P.off_cov = E.cov;
  P.off_fwd = E.fwd;
  P.off_adj = E.adj;
  P.off_mc = E.mc;
  P.off_grad = E.grad;
  P.off_fin = E.fin;
  P.off_cov_start = E.cov_start;
  P.off_mc_start = E.mc_start;
  P.off_fin_start = E.fin_start;
  P.off_fwd_tiles = E.fwd_tiles;
  P.off_adj_tiles = E.adj_tiles;
  return BTA_OK;
}

template <int KD>
static void covar_list(const bta_qlognei_batch_plan& P, const char* b, hipStream_t s) {
  const dim3 g((unsigned)P.covar_blocks);
  const QsCovarArgs* t = (const QsCovarArgs*)(b + P.off_cov);
  const int* st = (const int*)(b + P.off_cov_start);
  if (P.kind == BTA_KERNEL_MATERN52)
    hipLaunchKernelGGL((qstep_covar_list_kernel<KD, BTA_KERNEL_MATERN52>), g, dim3(256), 0, s, t,
                       st, P.n_items);
  else
    hipLaunchKernelGGL((qstep_covar_list_kernel<KD, BTA_KERNEL_RBF>), g, dim3(256), 0, s, t, st,
                       P.n_items);
}

template <int KD>
static void grad_list(const bta_qlognei_batch_plan& P, const char* b, hipStream_t s) {
  const dim3 g((unsigned)P.covar_blocks);
  const QsGradArgs* t = (const QsGradArgs*)(b + P.off_grad);
  const int* st = (const int*)(b + P.off_cov_start);
  if (P.kind == BTA_KERNEL_MATERN52)
    hipLaunchKernelGGL((qstep_grad_list_kernel<KD, BTA_KERNEL_MATERN52>), g, dim3(256), 0, s, t,
                       st, P.n_items);
  else
    hipLaunchKernelGGL((qstep_grad_list_kernel<KD, BTA_KERNEL_RBF>), g, dim3(256), 0, s, t, st,
                       P.n_items);
}

template <int Q>
static void adjoint_mc_list(const bta_qlognei_batch_plan& P, const char* b, hipStream_t s) {
  const int W = P.grad ? P.adj_workers : 0;
  hipLaunchKernelGGL((adjoint_mc_list_kernel<Q>), dim3((unsigned)(W + P.mc_blocks)), dim3(256),
                     0, s, (const H2Args*)(b + P.off_adj), (const int4*)(b + P.off_adj_tiles),
                     P.grad ? P.adj_tiles : 0, W, (const McArgs*)(b + P.off_mc),
                     (const int*)(b + P.off_mc_start), P.n_items);
}

int qlognei_batch_value_grad(const bta_qlognei_batch_plan* plan, hipStream_t s) {
  if (!plan || !plan->ws || plan->n_items < 1 || plan->q < 1 || plan->q > 5 || plan->kd < 1 ||
      plan->kd > 8 || plan->covar_blocks < 1 || plan->mc_blocks < 1 || plan->fwd_workers < 1 ||
      (plan->grad && plan->adj_workers < 1)) {
    set_error("qlognei_batch_value_grad: the plan was not made by bta_qlognei_batch_prepare");
    return BTA_ERR_INVALID;
  }
  const bta_qlognei_batch_plan& P = *plan;
  const char* b = (const char*)P.ws;
  switch (P.kd) {
    case 1: covar_list<1>(P, b, s); break;
    case 2: covar_list<2>(P, b, s); break;
    case 3: covar_list<3>(P, b, s); break;
    case 4: covar_list<4>(P, b, s); break;
    case 5: covar_list<5>(P, b, s); break;
    case 6: covar_list<6>(P, b, s); break;
    case 7: covar_list<7>(P, b, s); break;
    default: covar_list<8>(P, b, s); break;
  }
  BTA_LAUNCH_CHECK("qstep_covar_list_kernel");
  hipLaunchKernelGGL(gemm_h2_list_kernel, dim3((unsigned)P.fwd_workers), dim3(256), 0, s,
                     (const H2Args*)(b + P.off_fwd), (const int4*)(b + P.off_fwd_tiles),
                     P.fwd_tiles);
  BTA_LAUNCH_CHECK("gemm_h2_list_kernel");
  switch (P.q) {
    case 1: adjoint_mc_list<1>(P, b, s); break;
    case 2: adjoint_mc_list<2>(P, b, s); break;
    case 3: adjoint_mc_list<3>(P, b, s); break;
    case 4: adjoint_mc_list<4>(P, b, s); break;
    default: adjoint_mc_list<5>(P, b, s); break;
  }
  BTA_LAUNCH_CHECK("adjoint_mc_list_kernel");
  if (!P.grad) return BTA_OK;
  switch (P.kd) {
    case 1: grad_list<1>(P, b, s); break;
    case 2: grad_list<2>(P, b, s); break;
    case 3: grad_list<3>(P, b, s); break;
    case 4: grad_list<4>(P, b, s); break;
    case 5: grad_list<5>(P, b, s); break;
    case 6: grad_list<6>(P, b, s); break;
    case 7: grad_list<7>(P, b, s); break;
    default: grad_list<8>(P, b, s); break;
  }
  BTA_LAUNCH_CHECK("qstep_grad_list_kernel");
  hipLaunchKernelGGL(qstep_finish_list_kernel, dim3((unsigned)P.fin_blocks), dim3(256), 0, s,
                     (const FinishArgs*)(b + P.off_fin), (const int*)(b + P.off_fin_start),
                     P.n_items);
  BTA_LAUNCH_CHECK("qstep_finish_list_kernel");
  return BTA_OK;
}

}  // namespace bta

extern "C" size_t bta_qlognei_batch_workspace_size(const bta_qlognei_batch_item* items,
                                                   int32_t n_items) {
  return bta::qlognei_batch_ws_bytes(items, n_items);
}

extern "C" int bta_qlognei_batch_prepare(const bta_qlognei_batch_item* items, int32_t n_items,
                                         void* ws, size_t ws_bytes,
                                         bta_qlognei_batch_plan* plan, bta_stream_t stream) {
  return bta::qlognei_batch_prepare(items, n_items, ws, ws_bytes, plan, (hipStream_t)stream);
}

extern "C" int bta_qlognei_batch_value_grad(const bta_qlognei_batch_plan* plan,
                                            bta_stream_t stream) {
  return bta::qlognei_batch_value_grad(plan, (hipStream_t)stream);
}
