// faces_batch.hip -- the batched step: many independent periodic grids (each
// its own n, mx, my, mz) summed by one launch over a work list of
// element-segment items, replayed from a captured HIP graph.  No counterpart
// in the reference, which steps one grid per process (gpu/main.cc:170-186).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include "faces_batch.hpp"
#include "faces_common.hpp"

namespace faces {

// ---------------------------------------------------------------------------
// the batched kernel
// ---------------------------------------------------------------------------
// Static schedule, no atomics: the work list is cut into one contiguous
// chunk of equal estimated cost per block, the chunks of XCD x (blocks with
// blockIdx % 8 == x, round-robin dispatch) forming a contiguous eighth of the
// list's cost.  A block's items of one grid are consecutive element segments and
// rows: they form a run, summed with the next item's loads in flight, and the
// neighbouring items whose sectors an item shares were summed just before by
// the same block, so those lines are still in its L2.

// (n, S) routines: per n, S from 1 up to about 2K points per item.
#define FACES_BATCH_CASES(X)                                                                  \
  X(2, 1) X(2, 2) X(2, 4) X(2, 8) X(2, 16) X(2, 32) X(3, 1) X(3, 2) X(3, 4) X(3, 8) X(3, 16)  \
  X(4, 1) X(4, 2) X(4, 4) X(4, 8) X(4, 16) X(5, 1) X(5, 2) X(5, 4) X(5, 8) X(6, 1) X(6, 2)   \
  X(6, 4) X(6, 8) X(7, 1) X(7, 2) X(7, 4) X(7, 6) X(8, 1) X(8, 2) X(8, 4) X(9, 1) X(9, 2)   \
  X(9, 3) X(10, 1) X(10, 2) X(11, 1) X(11, 2) X(12, 1)

// One out-of-line routine per (n, S): inlining all of them into the kernel's
// switch makes the register allocator size every path for the union of their
// live ranges (256 VGPRs and spills); as calls, each keeps its own allocation.
// Arguments of a call arrive in vector registers; they are block-uniform, so
// they are moved back to scalar registers before use.
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ long uni(long x)
{
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)(x & 0xffffffffu));
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)((unsigned long)x >> 32));
  return (long)(((unsigned long)hi << 32) | lo);
}

// A block's items of one grid are taken in list order (rows, segments
// within a row): measured faster than visiting groups of rows segment by
// segment.
struct RunOrder {
  int a;
  __device__ int operator()(int t) const { return a + t; }
};

template <int N, int S>
__device__ __noinline__ void batch_run(const BatchGrid G0, int it, int count, double *lds)
{
  BatchGrid G;
  G.u = (double *)uni((long)G0.u);
  G.ys = uni(G0.ys);
  G.ps = uni(G0.ps);
  G.field_bytes = (unsigned)uni((int)G0.field_bytes);
  G.n = uni(G0.n);
  G.mx = uni(G0.mx);
  G.my = uni(G0.my);
  G.mz = uni(G0.mz);
  G.S = uni(G0.S);
  G.nseg = uni(G0.nseg);
  G.items = uni(G0.items);
  G.first = uni(G0.first);
  RunOrder o;
  o.a = uni(it);
  stream_run<N, S>(G, o, uni(count), lds);
}

// Four blocks per CU: the callees are compiled within the matching register
// budget (measured faster than letting them grow and run two blocks per CU).
__global__ __launch_bounds__(kStreamThreads, 4) void k_batch(BatchLaunch D)
{
  extern __shared__ double lds[];
  const int nb = gridDim.x, nbx = nb / kXcds;
  const int c = (blockIdx.x % kXcds) * nbx + blockIdx.x / kXcds;  // this block's chunk
  // last grid whose first item is at or before u / whose cost prefix is at or
  // before t (the grids are few: binary search over scalar loads)
  auto grid_of_item = [&](int u) {
    int lo = 0, hi = D.ngrids - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (D.grids[mid].first <= u) lo = mid; else hi = mid - 1;
    }
    return lo;
  };
  auto item_at = [&](double t) {
    int g = 0;
    if (D.grids) {
      int lo = 0, hi = D.ngrids - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (D.grids[mid].cost0 <= t) lo = mid; else hi = mid - 1;
      }
      g = lo;
    }
    const BatchGrid &G = D.grids ? D.grids[g] : D.g0;
    const double k = (t - G.cost0) / G.icost;
    return G.first + (k <= 0.0 ? 0 : k >= (double)G.items ? G.items : (int)k);
  };
  int u = (c == 0) ? 0 : item_at(D.total * c / nb);
  const int end = (c == nb - 1) ? D.nunits : item_at(D.total * (c + 1) / nb);
  int g = (D.grids && u < end) ? grid_of_item(u) : 0;
  while (u < end) {
    const BatchGrid G = D.grids ? D.grids[g] : D.g0;
    const int last = min(end, G.first + G.items);
    const int it = u - G.first;
    switch (G.n * 64 + G.S) {
#define FACES_BATCH_CASE(NN, SS) \
  case NN * 64 + SS: batch_run<NN, SS>(G, it, last - u, lds); break;
      FACES_BATCH_CASES(FACES_BATCH_CASE)
#undef FACES_BATCH_CASE
      default: break;  // excluded on the host
    }
    u = last;
    g++;
  }
}

// ---------------------------------------------------------------------------
// the batched kernel, one item per block
// ---------------------------------------------------------------------------
// Every block takes one item of the work list; XCD x (block b runs on XCD
// b % 8) takes its contiguous share of the list (shares of equal estimated
// cost), dispatched in order; blocks past the share exit at once.
// A grid's items are in memory order, so an item's seam, y and z partners
// (one, nseg and my*nseg items back) were summed moments before on the same
// XCD and their shared lines meet in its L2; with no loop, short-lived
// blocks keep many items' loads in flight per CU.

template <int N, int S>
__device__ __forceinline__ void batch_item(const BatchGrid G0, int it, double *lds)
{
  BatchGrid G;
  G.u = (double *)uni((long)G0.u);
  G.ys = uni(G0.ys);
  G.ps = uni(G0.ps);
  G.field_bytes = (unsigned)uni((int)G0.field_bytes);
  G.n = uni(G0.n);
  G.mx = uni(G0.mx);
  G.my = uni(G0.my);
  G.mz = uni(G0.mz);
  G.S = uni(G0.S);
  G.nseg = uni(G0.nseg);
  G.items = uni(G0.items);
  G.first = uni(G0.first);
  item_once<N, S>(G, uni(it), lds);
}

__global__ __launch_bounds__(kStreamThreads) __attribute__((amdgpu_waves_per_eu(8, 8)))
void k_batch_flow(BatchLaunch D)
{
  extern __shared__ double lds[];
  const int xcd = blockIdx.x & 7, k = blockIdx.x >> 3;
  const int u = D.xs[xcd] + k;
  if (u >= D.xs[xcd + 1]) return;
  // the grid holding item u: the number of grids whose first item is at or
  // before u, counted 64 grids per step across the wave's lanes (grids are
  // sorted by first item)
  int g = 0;
  if (D.grids) {
    const int lane = threadIdx.x & 63;
    int cnt = 0;
    for (int base = 0; base < D.ngrids; base += 256) {
      int f[4];
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int gi = base + r * 64 + lane;
        f[r] = gi < D.ngrids ? D.firsts[gi] : 0x7fffffff;
      }
#pragma unroll
      for (int r = 0; r < 4; r++) cnt += __popcll(__ballot(f[r] <= u));
    }
    g = uni(cnt - 1);
  }
  const BatchGrid G = D.grids ? D.grids[g] : D.g0;
  switch (G.n * 64 + G.S) {
#define FACES_BATCH_CASE(NN, SS) \
  case NN * 64 + SS: batch_item<NN, SS>(G, u - G.first, lds); break;
    FACES_BATCH_CASES(FACES_BATCH_CASE)
#undef FACES_BATCH_CASE
    default: break;  // excluded on the host
  }
}

size_t batch_pack_desc(const std::vector<BatchGrid> &g, void *buf)
{
  BatchGrid *grids = static_cast<BatchGrid *>(buf);
  std::copy(g.begin(), g.end(), grids);
  int *firsts = reinterpret_cast<int *>(grids + g.size());
  for (size_t i = 0; i < g.size(); i++) firsts[i] = g[i].first;
  return (sizeof(BatchGrid) + sizeof(int)) * g.size();
}

bool batch_case_ok(int n, int S)
{
  if (S < 1 || S > 63) return false;
  switch (n * 64 + S) {
#define FACES_BATCH_CASE(NN, SS) \
  case NN * 64 + SS: return true;
    FACES_BATCH_CASES(FACES_BATCH_CASE)
#undef FACES_BATCH_CASE
    default: return false;
  }
}

int batch_lds_bytes(int n, int S)
{
  const int T = kStreamThreads, K = n - 1;
  const int own = S * K * n * n, part = 2 * S * K * n + S * n + n * n;
  return ((own + T - 1) / T + (part + T - 1) / T) * T * (int)sizeof(double);
}

int batch_seg(int n, int mx, int cap)
{
  // the smallest routine that holds the whole row, else the largest one
  int best = 0;
  for (int s = 1; s <= std::min(cap, 63); s++) {
    if (!batch_case_ok(n, s) || batch_lds_bytes(n, s) > 64 * 1024) continue;
    best = s;
    if (s >= mx) break;
  }
  return best;
}

int batch_blocks(const BatchLaunch &D, int blocks_per_cu)
{
  if (blocks_per_cu == kBatchOneItem) {
    int most = 0;
    for (int x = 0; x < kXcds; x++) most = std::max(most, D.xs[x + 1] - D.xs[x]);
    return kXcds * most;
  }
  long nb = (long)device_cus() * blocks_per_cu;
  nb = std::min(nb, (long)D.nunits + kXcds - 1);
  nb = (nb + kXcds - 1) / kXcds * kXcds;
  return (int)std::max((long)kXcds, nb);
}

int launch_batch(const BatchLaunch &D, int nblocks, int lds_bytes, hipStream_t st, bool one_item)
{
  if (D.nunits == 0) return FACES_OK;
  if (nblocks % kXcds) return fail(FACES_ERR_ARG, "batched launch of %d blocks", nblocks);
  if (one_item) {
    if (D.xs[0] != 0 || D.xs[kXcds] != D.nunits) return fail(FACES_ERR_ARG, "XCD shares do not cover the list");
    for (int x = 0; x < kXcds; x++)
      if (D.xs[x + 1] < D.xs[x] || (long)kXcds * (D.xs[x + 1] - D.xs[x]) > nblocks)
        return fail(FACES_ERR_ARG, "one-item launch of %d blocks for an XCD share of %d items", nblocks,
                    D.xs[x + 1] - D.xs[x]);
    k_batch_flow<<<dim3(nblocks), dim3(kStreamThreads), lds_bytes, st>>>(D);
    FACES_HIP(hipGetLastError());
    return FACES_OK;
  }
  k_batch<<<dim3(nblocks), dim3(kStreamThreads), lds_bytes, st>>>(D);
  FACES_HIP(hipGetLastError());
  return FACES_OK;
}

// Estimated cost of one item of a grid, in doubles moved: the mean segment's
// points plus a fixed per-item share, the latency of a short block.  (256
// mixed grids of n 4, 6 and 8, one item per block: share 0 598.8 us, 320
// 589.9, 640 584.8, 1000 581.1, 1600 575.2, 2600 583.6.)
static double item_cost(const BatchGrid &x)
{
  return (double)x.mx * x.n * x.n * x.n / x.nseg + 1600.0;
}

// XCD shares of equal cost: xs[x] is the first item whose cost prefix
// reaches x/kXcds of the total.
static void split_xcds(const std::vector<BatchGrid> &g, BatchLaunch *D)
{
  size_t i = 0;
  D->xs[0] = 0;
  for (int x = 1; x < kXcds; x++) {
    const double t = D->total * x / kXcds;
    while (i + 1 < g.size() && g[i].cost0 + g[i].items * g[i].icost < t) i++;
    const BatchGrid &G = g[i];
    const double k = std::ceil((t - G.cost0) / G.icost);
    const int it = G.first + (k <= 0.0 ? 0 : k >= (double)G.items ? G.items : (int)k);
    D->xs[x] = std::max(D->xs[x - 1], std::min(it, D->nunits));
  }
  D->xs[kXcds] = D->nunits;
}

int launch_single(double *u, int n, int mx, int my, int mz, long plane_stride, int seg_cap,
                  int blocks_per_cu, hipStream_t st)
{
  BatchLaunch D;
  BatchGrid &x = D.g0;
  x.u = u;
  x.n = n;
  x.mx = mx;
  x.my = my;
  x.mz = mz;
  x.ys = (long)n * n * n * mx;
  x.ps = plane_stride > 0 ? plane_stride : x.ys * my;
  x.field_bytes = (unsigned)(8 * ((mz - 1) * x.ps + x.ys * my));
  x.S = batch_seg(n, mx, seg_cap);
  if (x.S == 0) return fail(FACES_ERR_ARG, "grid n=%d mx=%d has no item shape", n, mx);
  x.nseg = (mx + x.S - 1) / x.S;
  x.items = my * mz * x.nseg;
  x.first = 0;
  x.cost0 = 0.0;
  x.icost = 1.0;
  D.grids = nullptr;
  D.firsts = nullptr;
  D.ngrids = 1;
  D.nunits = x.items;
  D.total = (double)x.items;
  split_xcds(std::vector<BatchGrid>(1, x), &D);
  return launch_batch(D, batch_blocks(D, blocks_per_cu), batch_lds_bytes(n, x.S), st,
                      blocks_per_cu == kBatchOneItem);
}

int plan_grids(std::vector<BatchGrid> &g, int cap, int *lds_bytes, BatchLaunch *D)
{
  std::vector<int> by_cost(g.size());
  std::iota(by_cost.begin(), by_cost.end(), 0);
  for (auto &x : g) {
    x.S = batch_seg(x.n, x.mx, cap);
    if (x.S == 0) return fail(FACES_ERR_ARG, "grid n=%d mx=%d has no batched item shape", x.n, x.mx);
    x.nseg = (x.mx + x.S - 1) / x.S;
    x.items = x.my * x.mz * x.nseg;
    x.icost = item_cost(x);
  }
  std::stable_sort(by_cost.begin(), by_cost.end(), [&](int a, int b) { return g[a].icost > g[b].icost; });
  // Deal the grids, costliest items first, into one bin per XCD in snake
  // order (0..7, 7..0, ...) and lay the bins end to end: each XCD's share of
  // equal cost then holds the same mix of item sizes, and within it the
  // items of expensive grids come first, so it ends on cheap items.  (Sorted
  // list cut into eighths of equal item count: one XCD held 1.29 x the mean
  // bytes of 256 mixed grids.)
  std::vector<int> order;
  order.reserve(g.size());
  for (int bin = 0; bin < kXcds; bin++)
    for (size_t i = 0; i < by_cost.size(); i++) {
      const int k = (int)(i % (2 * kXcds));
      if ((k < kXcds ? k : 2 * kXcds - 1 - k) == bin) order.push_back(by_cost[i]);
    }
  std::vector<BatchGrid> sorted;
  int first = 0, lds = 0;
  double cost = 0.0;
  for (int i : order) {
    BatchGrid x = g[i];
    x.first = first;
    x.cost0 = cost;
    first += x.items;
    cost += x.items * x.icost;
    lds = std::max(lds, batch_lds_bytes(x.n, x.S));
    sorted.push_back(x);
  }
  g.swap(sorted);
  *lds_bytes = lds;
  D->nunits = first;
  D->total = cost;
  split_xcds(g, D);
  return FACES_OK;
}

}  // namespace faces

using namespace faces;

struct faces_batch {
  int ngrids;
  int boundary;
  std::vector<int> n, mx, my, mz;
  std::vector<double *> u;
  std::vector<faces_plan *> plans;  // one-launch-per-grid baseline and fallback paths
  std::vector<int> solo;            // grids without a batched item shape (n > 12): own plan
  // device state: the work list's descriptors and its launch (grids of
  // every element size in one list: with XCD shares of equal cost one launch
  // ran 256 mixed grids of n 4, 6 and 8 in 589.6 us, one launch per n in
  // 605.1 us)
  BatchGrid *d_grids = nullptr;
  BatchLaunch D = {};
  int lds_bytes = 0;
  // tuning: one block per item (256 mixed grids, n 4-8, 2-24 elements a
  // side: 626 us, traffic 1.02 x the fields; chunks of the list with 32
  // blocks per CU 685 us, traffic 1.36 x)
  int seg_cap = 64, bpc = kBatchOneItem;
  // captured graph of the batched launch
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool dirty = true;  // descriptors or tuning changed since the last upload
};

namespace {

// Grids the work-list kernel takes: n within its routines and a field below
// 64 MiB.  The others step on their own plan after the batched launch: large
// grids run faster through the plain step's streaming kernel (and its 32-bit
// offsets would not reach fields of 4 GB).
bool batchable(int n, int mx, int my, int mz)
{
  return n >= 2 && n <= 12 && 8 * faces_field_doubles(n, mx, my, mz, 0) < 64L * 1024 * 1024;
}

int upload(faces_batch *b)
{
  std::vector<BatchGrid> g;
  b->solo.clear();
  for (int i = 0; i < b->ngrids; i++) {
    if (!batchable(b->n[i], b->mx[i], b->my[i], b->mz[i])) {
      b->solo.push_back(i);
      continue;
    }
    g.emplace_back();
    BatchGrid &x = g.back();
    x.u = b->u[i];
    x.n = b->n[i];
    x.mx = b->mx[i];
    x.my = b->my[i];
    x.mz = b->mz[i];
    x.ys = (long)x.n * x.n * x.n * x.mx;
    x.ps = x.ys * x.my;
    x.field_bytes = (unsigned)(8 * x.ps * x.mz);
  }
  b->D = BatchLaunch{};
  b->lds_bytes = 0;
  if (!g.empty()) {
    int st = plan_grids(g, b->seg_cap, &b->lds_bytes, &b->D);
    if (st) return st;
    b->D.grids = b->d_grids;
    b->D.firsts = reinterpret_cast<const int *>(b->d_grids + g.size());
    b->D.ngrids = (int)g.size();
    // launches of the old descriptors may still run on any stream
    FACES_HIP(hipDeviceSynchronize());
    std::vector<char> host(batch_desc_bytes((int)g.size()));
    const size_t bytes = batch_pack_desc(g, host.data());
    FACES_HIP(hipMemcpy(b->d_grids, host.data(), bytes, hipMemcpyHostToDevice));
  }
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  b->dirty = false;
  return FACES_OK;
}

int launch_list(const faces_batch *b, hipStream_t st)
{
  return launch_batch(b->D, batch_blocks(b->D, b->bpc), b->lds_bytes, st, b->bpc == kBatchOneItem);
}

int ensure_graph(faces_batch *b)
{
  if (b->dirty) {
    int st = upload(b);
    if (st) return st;
  }
  if (b->exec || b->D.nunits == 0) return FACES_OK;
  FACES_HIP(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
  int st = launch_list(b, b->cap_stream);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
  if (st) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  b->graph = g;
  FACES_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return FACES_OK;
}

bool fields_bound(const faces_batch *b)
{
  for (double *p : b->u)
    if (!p) return false;
  return true;
}

}  // namespace

extern "C" {

int faces_batch_create(faces_batch **out, int ngrids, const int *n, const int *mx, const int *my,
                       const int *mz, int boundary)
{
  if (!out) return fail(FACES_ERR_ARG, "null output pointer");
  *out = nullptr;
  if (ngrids < 1) return fail(FACES_ERR_ARG, "need at least one grid");
  if (!n || !mx || !my || !mz) return fail(FACES_ERR_ARG, "null shape array");
  if (boundary != FACES_BOUNDARY_OPEN && boundary != FACES_BOUNDARY_PERIODIC)
    return fail(FACES_ERR_ARG, "unknown boundary %d", boundary);
  faces_batch *b = new (std::nothrow) faces_batch();
  if (!b) return fail(FACES_ERR_NOMEM, "batch allocation");
  b->ngrids = ngrids;
  b->boundary = boundary;
  long total_units = 0;
  for (int i = 0; i < ngrids; i++) {
    if (n[i] < 2 || n[i] > 64 || mx[i] < 1 || my[i] < 1 || mz[i] < 1) {
      delete b;
      return fail(FACES_ERR_ARG, "grid %d: shape n=%d m=%dx%dx%d invalid", i, n[i], mx[i], my[i], mz[i]);
    }
    total_units += (long)mx[i] * my[i] * mz[i];
    b->n.push_back(n[i]);
    b->mx.push_back(mx[i]);
    b->my.push_back(my[i]);
    b->mz.push_back(mz[i]);
  }
  if (total_units >= (1L << 30)) {
    delete b;
    return fail(FACES_ERR_ARG, "batch too large");
  }
  b->u.assign(ngrids, nullptr);
  int st = FACES_OK;
  for (int i = 0; i < ngrids && !st; i++) {
    faces_plan *p = nullptr;
    st = faces_plan_create(&p, 0, 1, 1, 1, mx[i], my[i], mz[i], n[i], boundary, 0);
    if (!st) b->plans.push_back(p);
  }
  hipError_t e = hipSuccess;
  if (!st) e = hipMalloc(&b->d_grids, batch_desc_bytes(ngrids));
  if (!st && e == hipSuccess) e = hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking);
  if (!st && e == hipSuccess) e = hipDeviceSynchronize();
  if (!st && e != hipSuccess) st = hip_fail(e, "batch device state");
  if (st) {
    faces_batch_destroy(b);
    return st;
  }
  *out = b;
  return FACES_OK;
}

int faces_batch_destroy(faces_batch *b)
{
  if (!b) return FACES_OK;
  int st = FACES_OK;
  for (faces_plan *p : b->plans) {
    const int s = faces_plan_destroy(p);
    if (s && !st) st = s;
  }
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
  if (b->d_grids) (void)hipFree(b->d_grids);
  delete b;
  return st;
}

int faces_batch_set_fields(faces_batch *b, double *const *u)
{
  if (!b || !u) return fail(FACES_ERR_ARG, "null argument");
  for (int i = 0; i < b->ngrids; i++)
    if (!u[i]) return fail(FACES_ERR_ARG, "null field pointer for grid %d", i);
  for (int i = 0; i < b->ngrids; i++) b->u[i] = u[i];
  b->dirty = true;
  return FACES_OK;
}

int faces_batch_set_tuning(faces_batch *b, int block, int seg_len, int blocks_per_cu)
{
  if (!b) return fail(FACES_ERR_ARG, "null batch");
  if (block != 0 && block != kStreamThreads)
    return fail(FACES_ERR_ARG, "batched kernel blocks have %d threads", kStreamThreads);
  if (seg_len < 0 || seg_len > 64) return fail(FACES_ERR_ARG, "segment length %d outside [0,64]", seg_len);
  if (blocks_per_cu < kBatchOneItem || blocks_per_cu > 256)
    return fail(FACES_ERR_ARG, "blocks per CU %d outside [-1,256]", blocks_per_cu);
  if (seg_len) b->seg_cap = seg_len;
  if (blocks_per_cu) b->bpc = blocks_per_cu;
  b->dirty = true;
  return FACES_OK;
}

int faces_batch_share(faces_batch *b, int compute, void *stream)
{
  if (!b) return fail(FACES_ERR_ARG, "null batch");
  if (!fields_bound(b)) return fail(FACES_ERR_STATE, "fields not bound (faces_batch_set_fields)");
  if (b->boundary != FACES_BOUNDARY_PERIODIC || !compute)
    return faces_batch_share_per_grid(b, compute, stream);  // open / surface-only: per grid
  int st = ensure_graph(b);
  if (st) return st;
  if (b->exec) FACES_HIP(hipGraphLaunch(b->exec, (hipStream_t)stream));
  for (int i : b->solo) {
    st = faces_share(b->plans[i], b->u[i], compute, stream);
    if (st) return st;
  }
  return FACES_OK;
}

int faces_batch_share_per_grid(faces_batch *b, int compute, void *stream)
{
  if (!b) return fail(FACES_ERR_ARG, "null batch");
  if (!fields_bound(b)) return fail(FACES_ERR_STATE, "fields not bound (faces_batch_set_fields)");
  for (int i = 0; i < b->ngrids; i++) {
    const int st = faces_share(b->plans[i], b->u[i], compute, stream);
    if (st) return st;
  }
  return FACES_OK;
}

int faces_batch_info(const faces_batch *b, double *bytes, long *items)
{
  if (!b || !bytes || !items) return fail(FACES_ERR_ARG, "null argument");
  double s = 0;
  long it = 0;
  for (int i = 0; i < b->ngrids; i++) {
    s += 8.0 * faces_field_doubles(b->n[i], b->mx[i], b->my[i], b->mz[i], 0);
    const int S = batchable(b->n[i], b->mx[i], b->my[i], b->mz[i]) ? batch_seg(b->n[i], b->mx[i], b->seg_cap) : 0;
    it += S ? (long)b->my[i] * b->mz[i] * ((b->mx[i] + S - 1) / S) : 0;
  }
  *bytes = s;
  *items = it;
  return FACES_OK;
}

int faces_batch_tune(faces_batch *b, int reps, void *stream, double *out)
{
  if (!b || !out) return fail(FACES_ERR_ARG, "null argument");
  if (reps < 1 || reps > 1000) return fail(FACES_ERR_ARG, "reps %d outside [1,1000]", reps);
  if (!fields_bound(b)) return fail(FACES_ERR_STATE, "fields not bound (faces_batch_set_fields)");
  hipStream_t st = (hipStream_t)stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    return fail(FACES_ERR_HIP, "hipEventCreate");
  }
  double best = 1e30;
  int best_seg = b->seg_cap, best_bpc = b->bpc;
  int rc = FACES_OK;
  for (int seg : {2, 4, 8, 16, 64}) {
    for (int bpc : {kBatchOneItem, 2, 4, 8, 16, 32, 48, 64, 96, 128}) {
      b->seg_cap = seg;
      b->bpc = bpc;
      b->dirty = true;
      rc = upload(b);
      if (rc) break;
      rc = launch_list(b, st);  // warm-up
      for (int r = 0; r < reps && !rc; r++) {
        if (r == 0 && hipEventRecord(e0, st) != hipSuccess) rc = fail(FACES_ERR_HIP, "event");
        if (!rc) rc = launch_list(b, st);
      }
      if (rc) break;
      float ms = 0;
      if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        rc = fail(FACES_ERR_HIP, "timing events");
        break;
      }
      const double us = 1e3 * ms / reps;
      if (us < best) {
        best = us;
        best_seg = seg;
        best_bpc = bpc;
      }
    }
    if (rc) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->seg_cap = best_seg;
  b->bpc = best_bpc;
  b->dirty = true;
  if (rc) return rc;
  out[0] = kStreamThreads;
  out[1] = best_seg;
  out[2] = best_bpc;
  out[3] = best;
  return FACES_OK;
}

}  // extern "C"
