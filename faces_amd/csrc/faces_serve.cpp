// faces_serve.cpp -- the served step: a request loop with continuous batching.
//
// Requests (a dense device field and its shape) arrive at any time.  A worker
// thread keeps up to `nslots` launches in flight, one per HIP stream: whenever
// a slot is free it admits queued requests in arrival order -- up to
// max_batch of them, while the bytes of admitted, unfinished requests stay
// within the footprint cap -- and sends them as one batched launch (one work
// list over all their element segments, faces_batch.hip) on that slot's
// stream.  Requests that arrive while other launches run are admitted as soon
// as a slot and footprint free up, so launches overlap on the device.  Every
// request's result depends only on its own field (items never span grids), so
// it is the same alone or in any mix.  Latency is measured from arrival to the
// moment the worker observes the launch's completion event.
//
// Requests on one field are stepped in submission order: a request whose
// field is already in a launch (or held back behind another request on it)
// waits in the queue while later requests on other fields go ahead.  With
// two or more slots, large requests never take the last free slot, so small
// requests always have one.  Plans for requests the batched kernel does not
// take are created by the submitting thread, never inside the worker.
//
// No counterpart in the reference, which steps one grid per process.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "faces_batch.hpp"
#include "faces_common.hpp"

using namespace faces;

namespace {

using Clock = std::chrono::steady_clock;

// Requests at or above this footprint get a launch of their own.
constexpr double kLargeBytes = 64.0 * 1024 * 1024;

// Default launch shape of a batched launch, with two launches in flight: one
// block per work item (Poisson 20k req/s: p50 32 us, p99 203 us), against
// chunks of the launch's work list with 1, 2 or 4 blocks per CU (p50 39, 35,
// 42 us; p99 220, 219, 228 us), measured with the two slots' streams sharing
// a hardware queue.
constexpr int kServeBlocksPerCu = kBatchOneItem;

double now_us(Clock::time_point t0) { return std::chrono::duration<double, std::micro>(Clock::now() - t0).count(); }

struct Request {
  uint64_t ticket;
  int n, mx, my, mz, boundary, compute;
  double *u;
  double bytes;
  double t_arrival = 0, t_done = 0;
  bool done = false;
  int status = FACES_OK;
  faces_plan *plan = nullptr;  // plain-step plan (large or not batchable), set at submission
};

// Requests the batched kernel takes: periodic, compute = 1, rows with a work
// item length, 32-bit offsets.  Large requests take the plain step's plan.
bool batchable(int n, int mx, int boundary, int compute, double bytes)
{
  return boundary == FACES_BOUNDARY_PERIODIC && compute && batch_seg(n, mx, 64) > 0 && bytes < (double)kMaxBufferBytes;
}

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t event = nullptr;
  BatchGrid *d_grids = nullptr;
  BatchGrid *h_grids = nullptr;  // pinned staging
  std::vector<Request *> reqs;
  bool busy = false;
  bool large = false;  // the launch holds one large request
};

}  // namespace

struct faces_server {
  double cap;
  int max_batch;
  int bpc = kServeBlocksPerCu;  // launch shape of batched launches (faces_server_set_tuning)
  int device = 0;
  std::vector<Slot> slots;
  Clock::time_point t0;

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request *> queue;
  std::unordered_map<uint64_t, Request *> reqs;
  uint64_t next_ticket = 1;
  double admitted_bytes = 0, peak_bytes = 0;
  long n_admitted = 0, n_completed = 0, n_launches = 0;
  int large_busy = 0;                         // slots running a large request
  std::unordered_set<const double *> fields;  // fields of admitted, unfinished requests
  bool stop = false;
  int worker_error = FACES_OK;
  std::thread worker;

  // plans for large requests and those the batched kernel does not take
  // (open boundary, compute = 0), keyed by shape; created by the submitting
  // thread under plans_mu (never together with mu), destroyed with the server
  std::mutex plans_mu;
  std::map<std::tuple<int, int, int, int, int>, faces_plan *> plans;

  void run();
  int launch(Slot &s, std::vector<Request *> &batch);
  void release(Request *r, double t);  // under mu
};

void faces_server::release(Request *r, double t)
{
  r->t_done = t;
  r->done = true;
  admitted_bytes -= r->bytes;
  fields.erase(r->u);
}

int faces_server::launch(Slot &s, std::vector<Request *> &batch)
{
  std::vector<BatchGrid> g;
  std::vector<Request *> other;
  // A launch of one large request takes the plain step's kernel choice (the
  // streaming kernel where the rows split evenly), measured faster on large
  // grids than the work-list kernel.
  const bool alone = batch.size() == 1 && batch[0]->bytes >= kLargeBytes;
  for (Request *r : batch) {
    if (!alone && batchable(r->n, r->mx, r->boundary, r->compute, r->bytes)) {
      BatchGrid x;
      x.u = r->u;
      x.n = r->n;
      x.mx = r->mx;
      x.my = r->my;
      x.mz = r->mz;
      x.ys = (long)x.n * x.n * x.n * x.mx;
      x.ps = x.ys * x.my;
      x.field_bytes = (unsigned)(8 * x.ps * x.mz);
      g.push_back(x);
    } else {
      other.push_back(r);
    }
  }
  if (!g.empty()) {
    int lds = 0;
    BatchLaunch D = {};
    int st = plan_grids(g, 64, &lds, &D);
    if (st) return st;
    const size_t bytes = batch_pack_desc(g, s.h_grids);
    FACES_HIP(hipMemcpyAsync(s.d_grids, s.h_grids, bytes, hipMemcpyHostToDevice, s.stream));
    D.grids = s.d_grids;
    D.firsts = reinterpret_cast<const int *>(s.d_grids + g.size());
    D.ngrids = (int)g.size();
    st = launch_batch(D, batch_blocks(D, bpc), lds, s.stream, bpc == kBatchOneItem);
    if (st) return st;
  }
  for (Request *r : other) {
    if (!r->plan) return fail(FACES_ERR_STATE, "request without a plan");
    int st = faces_share(r->plan, r->u, r->compute, s.stream);
    if (st) return st;
  }
  FACES_HIP(hipEventRecord(s.event, s.stream));
  s.reqs = batch;
  s.busy = true;
  s.large = alone;
  if (alone) large_busy++;
  n_launches++;
  return FACES_OK;
}

void faces_server::run()
{
  if (hipSetDevice(device) != hipSuccess) {
    std::lock_guard<std::mutex> lk(mu);
    worker_error = fail(FACES_ERR_HIP, "server worker: hipSetDevice");
    stop = true;
    cv_done.notify_all();
    return;
  }
  std::unique_lock<std::mutex> lk(mu);
  while (true) {
    // completions
    bool any_busy = false;
    for (Slot &s : slots) {
      if (!s.busy) continue;
      const hipError_t q = hipEventQuery(s.event);
      if (q == hipErrorNotReady) {
        any_busy = true;
        continue;
      }
      const double t = now_us(t0);
      for (Request *r : s.reqs) {
        if (q != hipSuccess) r->status = hip_fail(q, "server launch");
        release(r, t);
        n_completed++;
      }
      s.reqs.clear();
      s.busy = false;
      if (s.large) large_busy--;
      s.large = false;
      cv_done.notify_all();
    }
    // admissions into free slots, in arrival order
    for (Slot &s : slots) {
      if (s.busy || queue.empty()) continue;
      // A launch holds either one large request or small ones only, so small
      // requests never wait for a large one that shares their launch; with
      // two or more slots, large requests leave the last free slot to small
      // ones.  A request on a field that an unfinished request holds (or an
      // earlier request passed over in this scan) stays queued; later
      // requests on other fields go ahead of it.  A request that does not
      // fit the footprint cap stops the scan: nothing overtakes it for
      // footprint.
      const bool keep_small = slots.size() > 1 && large_busy >= (int)slots.size() - 1;
      std::vector<Request *> batch;
      std::unordered_set<const double *> held;
      for (auto it = queue.begin(); it != queue.end() && (int)batch.size() < max_batch;) {
        Request *r = *it;
        const bool large = r->bytes >= kLargeBytes;
        if (fields.count(r->u) || held.count(r->u) || (large && (keep_small || !batch.empty()))) {
          held.insert(r->u);
          ++it;
          continue;
        }
        if (admitted_bytes + r->bytes > cap) break;
        it = queue.erase(it);
        admitted_bytes += r->bytes;
        peak_bytes = std::max(peak_bytes, admitted_bytes);
        fields.insert(r->u);
        n_admitted++;
        batch.push_back(r);
        if (large) break;
      }
      if (batch.empty()) break;  // the queue waits for footprint, a field or a slot
      const int st = launch(s, batch);
      if (st) {
        // part of the batch may already be enqueued: let it finish before
        // the caller learns of the failure (and may free its fields)
        (void)hipStreamSynchronize(s.stream);
        const double t = now_us(t0);
        for (Request *r : batch) {
          r->status = st;
          release(r, t);
        }
        cv_done.notify_all();
      } else {
        any_busy = true;
      }
    }
    if (stop && queue.empty() && !any_busy) break;
    if (any_busy) {
      // launches in flight: poll their events without holding the lock long
      lk.unlock();
      std::this_thread::yield();
      lk.lock();
    } else {
      cv_work.wait_for(lk, std::chrono::milliseconds(5));
    }
  }
}

extern "C" {

int faces_server_create(faces_server **out, double footprint_cap, int max_batch, int nstreams)
{
  if (!out) return fail(FACES_ERR_ARG, "null output pointer");
  *out = nullptr;
  if (!(footprint_cap > 0)) return fail(FACES_ERR_ARG, "footprint cap must be positive");
  if (max_batch < 1 || max_batch > 65536) return fail(FACES_ERR_ARG, "max_batch %d outside [1,65536]", max_batch);
  if (nstreams < 1 || nstreams > 16) return fail(FACES_ERR_ARG, "nstreams %d outside [1,16]", nstreams);
  faces_server *s = new (std::nothrow) faces_server();
  if (!s) return fail(FACES_ERR_NOMEM, "server allocation");
  s->cap = footprint_cap;
  s->max_batch = max_batch;
  s->t0 = Clock::now();
  hipError_t e = hipGetDevice(&s->device);
  s->slots.resize(nstreams);
  for (Slot &sl : s->slots) {
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.event, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMalloc(&sl.d_grids, batch_desc_bytes(max_batch));
    if (e == hipSuccess) e = hipHostMalloc(&sl.h_grids, batch_desc_bytes(max_batch), hipHostMallocDefault);
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    const int st = hip_fail(e, "server device state");
    faces_server_destroy(s);
    return st;
  }
  s->worker = std::thread([s] { s->run(); });
  *out = s;
  return FACES_OK;
}

int faces_server_destroy(faces_server *s)
{
  if (!s) return FACES_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  for (Slot &sl : s->slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    if (sl.event) (void)hipEventDestroy(sl.event);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
    if (sl.d_grids) (void)hipFree(sl.d_grids);
    if (sl.h_grids) (void)hipHostFree(sl.h_grids);
  }
  for (auto &kv : s->plans) faces_plan_destroy(kv.second);
  for (auto &kv : s->reqs) delete kv.second;
  delete s;
  return FACES_OK;
}

static int submit_at(faces_server *s, int n, int mx, int my, int mz, int boundary, int compute,
                     double *u, double t_arrival, uint64_t *ticket)
{
  if (!s || !u || !ticket) return fail(FACES_ERR_ARG, "null argument");
  if (n < 2 || n > 64 || mx < 1 || my < 1 || mz < 1) return fail(FACES_ERR_ARG, "bad shape");
  if (boundary != FACES_BOUNDARY_OPEN && boundary != FACES_BOUNDARY_PERIODIC)
    return fail(FACES_ERR_ARG, "unknown boundary %d", boundary);
  const double bytes = 8.0 * faces_field_doubles(n, mx, my, mz, 0);
  if (bytes > s->cap) return fail(FACES_ERR_ADMISSION, "request of %.0f bytes exceeds the cap %.0f", bytes, s->cap);
  // The plain-step plan of a large or non-batchable request is looked up, or
  // created on the server's device, here, by the submitting thread.
  faces_plan *plan = nullptr;
  if (bytes >= kLargeBytes || !batchable(n, mx, boundary, compute, bytes)) {
    std::lock_guard<std::mutex> lk(s->plans_mu);
    faces_plan *&p = s->plans[std::make_tuple(n, mx, my, mz, boundary)];
    if (!p) {
      int dev0 = 0;
      if (hipGetDevice(&dev0) != hipSuccess) return fail(FACES_ERR_HIP, "hipGetDevice");
      if (dev0 != s->device && hipSetDevice(s->device) != hipSuccess) return fail(FACES_ERR_HIP, "hipSetDevice");
      const int st = faces_plan_create(&p, 0, 1, 1, 1, mx, my, mz, n, boundary, 0);
      if (dev0 != s->device) (void)hipSetDevice(dev0);
      if (st) {
        p = nullptr;
        return st;
      }
    }
    plan = p;
  }
  Request *r = new (std::nothrow) Request();
  if (!r) return fail(FACES_ERR_NOMEM, "request allocation");
  r->n = n;
  r->mx = mx;
  r->my = my;
  r->mz = mz;
  r->boundary = boundary;
  r->compute = compute ? 1 : 0;
  r->u = u;
  r->bytes = bytes;
  r->plan = plan;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) {
      delete r;
      return fail(FACES_ERR_STATE, "server stopped");
    }
    r->ticket = s->next_ticket++;
    r->t_arrival = t_arrival >= 0 ? t_arrival : now_us(s->t0);
    s->reqs[r->ticket] = r;
    s->queue.push_back(r);
    *ticket = r->ticket;
  }
  s->cv_work.notify_one();
  return FACES_OK;
}

int faces_server_submit(faces_server *s, int n, int mx, int my, int mz, int boundary, int compute,
                        double *u, uint64_t *ticket)
{
  return submit_at(s, n, mx, my, mz, boundary, compute, u, -1.0, ticket);
}

int faces_server_wait(faces_server *s, uint64_t ticket, double timeout_s, double *latency_us)
{
  if (!s) return fail(FACES_ERR_ARG, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  const auto deadline = Clock::now() + std::chrono::duration<double>(timeout_s > 0 ? timeout_s : 1e9);
  // The ticket is looked up again after every wake-up: another thread may
  // have waited for it (and released it) meanwhile.
  auto it = s->reqs.find(ticket);
  while (it != s->reqs.end() && !it->second->done) {
    if (s->cv_done.wait_until(lk, deadline) == std::cv_status::timeout) {
      it = s->reqs.find(ticket);
      if (it != s->reqs.end() && !it->second->done)
        return fail(FACES_ERR_TIMEOUT, "request %llu not done", (unsigned long long)ticket);
      break;
    }
    it = s->reqs.find(ticket);
  }
  if (it == s->reqs.end())
    return fail(FACES_ERR_ARG, "unknown ticket %llu (never issued, or already waited for)", (unsigned long long)ticket);
  Request *r = it->second;
  const int st = r->status;
  if (latency_us) *latency_us = r->t_done - r->t_arrival;
  s->reqs.erase(it);
  delete r;
  if (st) return fail(st, "request failed: %s", faces_strerror(st));
  return FACES_OK;
}

int faces_server_run_trace(faces_server *s, int nreq, const double *arrival_us, const int *n,
                           const int *mx, const int *my, const int *mz, int boundary, int compute,
                           double *const *u, double *latency_us, double *stats)
{
  if (!s || nreq < 1 || !arrival_us || !n || !mx || !my || !mz || !u || !latency_us || !stats)
    return fail(FACES_ERR_ARG, "null argument");
  for (int i = 1; i < nreq; i++)
    if (arrival_us[i] < arrival_us[i - 1]) return fail(FACES_ERR_ARG, "arrivals must be sorted");
  std::vector<uint64_t> tickets(nreq, 0);
  long launches0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    launches0 = s->n_launches;
  }
  const double start = now_us(s->t0);
  int st = FACES_OK;
  for (int i = 0; i < nreq && !st; i++) {
    const double due = start + arrival_us[i];
    while (now_us(s->t0) < due) {
      if (due - now_us(s->t0) > 200.0) std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
    st = submit_at(s, n[i], mx[i], my[i], mz[i], boundary, compute, u[i], due, &tickets[i]);
  }
  double last = start;
  for (int i = 0; i < nreq; i++) {
    if (!tickets[i]) continue;
    double lat = 0;
    const int w = faces_server_wait(s, tickets[i], 600.0, &lat);
    if (w && !st) st = w;
    latency_us[i] = lat;
    last = std::max(last, start + arrival_us[i] + lat);
  }
  if (st) return st;
  std::vector<double> sorted(latency_us, latency_us + nreq);
  std::sort(sorted.begin(), sorted.end());
  auto pct = [&](double p) {
    const double x = p * (nreq - 1);
    const int i0 = (int)x;
    const int i1 = std::min(nreq - 1, i0 + 1);
    return sorted[i0] + (x - i0) * (sorted[i1] - sorted[i0]);
  };
  const double wall = (last - start) * 1e-6;
  stats[0] = wall;
  stats[1] = wall > 0 ? nreq / wall : 0;
  stats[2] = pct(0.50);
  stats[3] = pct(0.99);
  stats[4] = sorted[nreq - 1];
  {
    std::lock_guard<std::mutex> lk(s->mu);
    stats[5] = (double)(s->n_launches - launches0);
  }
  return FACES_OK;
}

int faces_server_set_tuning(faces_server *s, int blocks_per_cu)
{
  if (!s) return fail(FACES_ERR_ARG, "null server");
  if (blocks_per_cu < kBatchOneItem || blocks_per_cu == 0 || blocks_per_cu > 256)
    return fail(FACES_ERR_ARG, "blocks per CU %d outside {-1} and [1,256]", blocks_per_cu);
  std::lock_guard<std::mutex> lk(s->mu);
  s->bpc = blocks_per_cu;
  return FACES_OK;
}

int faces_server_stats(const faces_server *s, double *stats)
{
  if (!s || !stats) return fail(FACES_ERR_ARG, "null argument");
  faces_server *m = const_cast<faces_server *>(s);
  std::lock_guard<std::mutex> lk(m->mu);
  stats[0] = (double)m->n_admitted;
  stats[1] = (double)m->n_completed;
  stats[2] = (double)m->n_launches;
  stats[3] = m->peak_bytes;
  return FACES_OK;
}

}  // extern "C"
