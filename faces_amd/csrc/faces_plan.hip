// faces_plan.hip -- the plain exchange + DSS step behind the C ABI
// (replaces class Faces, gpu/v-26-messages/Faces.{h,cc}).
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "faces_common.hpp"
#include "faces_kernels.hpp"
#include "faces_stream.hpp"
#include "faces_flow.hpp"
#include "faces_batch.hpp"

namespace faces {

static thread_local std::string g_last_error;

int fail(int status, const char *fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return status;
}

int hip_fail(hipError_t e, const char *what)
{
  const int st = (e == hipErrorOutOfMemory) ? FACES_ERR_NOMEM : FACES_ERR_HIP;
  return fail(st, "%s: %s (%s)", what, hipGetErrorName(e), hipGetErrorString(e));
}

// Per-device attributes, cached (plans and servers may query them from
// several threads).
int device_cus()
{
  static std::atomic<int> cached[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  int v = cached[dev].load(std::memory_order_relaxed);
  if (v == 0) {
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
      v = 256;
    cached[dev].store(v, std::memory_order_relaxed);
  }
  return v;
}

// LDS bytes per CU (160 KiB on gfx950).
int device_lds_per_cu()
{
  static std::atomic<int> cached[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 160 * 1024;
  int v = cached[dev].load(std::memory_order_relaxed);
  if (v == 0) {
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess ||
        v <= 0)
      v = 160 * 1024;
    cached[dev].store(v, std::memory_order_relaxed);
  }
  return v;
}

// A non-blocking stream per device (created on first use, kept until the
// process ends) for the library's set-up work, so that set-up waits only for
// itself.
hipError_t setup_stream(hipStream_t *out)
{
  static std::mutex mu;
  static hipStream_t streams[64] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(mu);
  if (!streams[dev]) {
    e = hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking);
    if (e != hipSuccess) {
      streams[dev] = nullptr;
      return e;
    }
  }
  *out = streams[dev];
  return hipSuccess;
}

}  // namespace faces

using namespace faces;

struct faces_plan {
  int rank, lx, ly, lz, kx, ky, kz;
  int mx, my, mz, n;
  int boundary;
  long ps;
  int nbr[26];   // neighbour rank per slot, -1 = none
  int mask;      // face-neighbour bits x-,x+,y-,y+,z-,z+
  int path;      // FACES_PATH_*
  int block, seg_len, xcd_remap;
  int variant;   // FACES_VARIANT_*
  int lds_block; // threads per block of the LDS-staged kernel
  int stream_bpc; // streaming kernel: persistent blocks per CU (> 0), or one unit per block (<= 0)
  int stream_walk;   // streaming kernel traversal: 0 items, 1 x walk, 2 y walk, 3 row groups
  int stream_rows;   // walk 3: rows per group; walk 2: rows per column piece (the largest
                     // divisor of my not above it); 0 = default per walk
  int items_bpc;     // work-list item kernel: blocks per CU, a chunk of the list each
  int stream_planes; // 1: per-plane descriptors even below 4 GB
  int flow;          // periodic steps with compute = 1: 0 the streaming kernel; 1 the one-item
                     // kernel (faces_flow.hpp), 8-byte lanes; 2 the same with 16-byte OWN
                     // loads; 3 the owned runs in 16-byte pairs
  int surface_blocks; // staged step on the one-item kernel: surface sums by its items (3), as
                      // blocks before (1) or after (2) its items, or a launch of their own (0)
  // message arena: send slots then receive slots
  double *arena;
  long arena_doubles;
  long slot_cnt[26];
  Slots send, recv;
};

namespace {

constexpr int kMaxN = 64;

int neighbor_of(const faces_plan *p, int dx, int dy, int dz)
{
  // gpu/v-26-messages/Faces.cc:92-103, with the periodic rank grid wrapping.
  int kx = p->kx + dx, ky = p->ky + dy, kz = p->kz + dz;
  if (p->boundary == FACES_BOUNDARY_PERIODIC) {
    kx = (kx + p->lx) % p->lx;
    ky = (ky + p->ly) % p->ly;
    kz = (kz + p->lz) % p->lz;
  } else {
    if (kx < 0 || kx >= p->lx || ky < 0 || ky >= p->ly || kz < 0 || kz >= p->lz) return -1;
  }
  return (kz * p->ly + ky) * p->lx + kx;
}

void slot_dir(int k, int &dx, int &dy, int &dz)
{
  dx = dy = dz = 0;
  if (k < 6) {
    const int s = (k & 1) ? 1 : -1;
    if (k < 2) dx = s; else if (k < 4) dy = s; else dz = s;
  } else if (k < 18) {
    const int f = (k - 6) / 4, e = (k - 6) % 4;
    const int a = (e & 1) ? 1 : -1, b = (e >> 1) ? 1 : -1;
    if (f == 0) { dy = a; dz = b; }        // iedge_[0..3]  (0,+-1,+-1)
    else if (f == 1) { dx = a; dz = b; }   // iedge_[4..7]  (+-1,0,+-1)
    else { dx = a; dy = b; }               // iedge_[8..11] (+-1,+-1,0)
  } else {
    const int c = k - 18;
    dx = (c & 1) ? 1 : -1;
    dy = (c & 2) ? 1 : -1;
    dz = (c & 4) ? 1 : -1;
  }
}

Geom make_geom(const faces_plan *p, double *u)
{
  Geom g;
  g.u = u;
  g.n = p->n;
  g.mx = p->mx;
  g.my = p->my;
  g.mz = p->mz;
  g.ys = (long)p->n * p->n * p->n * p->mx;
  g.ps = p->ps;
  return g;
}

Bound bound_wrap()
{
  Bound b;
  b.x = b.y = b.z = Axis{1, 0, 0};
  return b;
}

Bound bound_recv(int mask)
{
  Bound b;
  b.x = Axis{0, mask & 1, (mask >> 1) & 1};
  b.y = Axis{0, (mask >> 2) & 1, (mask >> 3) & 1};
  b.z = Axis{0, (mask >> 4) & 1, (mask >> 5) & 1};
  return b;
}

SurfDesc surf_desc(const faces_plan *p, const Bound &b)
{
  SurfDesc s;
  const int K = p->n - 1;
  s.Gx = p->mx * K + (b.x.wrap ? 0 : 1);
  s.Gy = p->my * K + (b.y.wrap ? 0 : 1);
  s.Gz = p->mz * K + (b.z.wrap ? 0 : 1);
  s.nsx = b.x.wrap ? 1 : 2;
  s.nsy = b.y.wrap ? 1 : 2;
  s.nsz = b.z.wrap ? 1 : 2;
  s.cz = (long)s.nsz * s.Gx * s.Gy;
  s.cy = (long)s.nsy * s.Gx * (s.Gz - s.nsz);
  s.cx = (long)s.nsx * (s.Gy - s.nsy) * (s.Gz - s.nsz);
  return s;
}

int grid_for(long work, int block)
{
  const long cap = (long)device_cus() * 32;
  long g = (work + block - 1) / block;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return (int)g;
}

// Elements per item of the LDS kernel and the LDS bytes that takes.
int elem_seg(const faces_plan *p)
{
  const int want = p->seg_len > 0 ? p->seg_len : 4;
  return want > p->mx ? p->mx : want;
}

long elem_lds_bytes(const faces_plan *p, int S)
{
  const long n = p->n, K = n - 1;
  return (long)sizeof(double) * (S * (K * n * n + 2 * K * n + n) + n * n);
}

bool elem_lds_ok(const faces_plan *p)
{
  return p->n >= 2 && p->n <= 12 && elem_lds_bytes(p, elem_seg(p)) <= 64 * 1024;
}

int launch_elem_lds(const faces_plan *p, double *u, hipStream_t st)
{
  Geom G = make_geom(p, u);
  ElemItems I;
  I.my = p->my;
  I.mz = p->mz;
  I.seg = elem_seg(p);
  I.nseg = (p->mx + I.seg - 1) / I.seg;
  I.nitems = p->my * p->mz * I.nseg;
  I.xcd_remap = p->xcd_remap;
  const int nb = I.xcd_remap ? 8 * ((I.nitems + 7) / 8) : I.nitems;
  const size_t lds = (size_t)elem_lds_bytes(p, I.seg);
  const dim3 grid(nb), blk(p->lds_block);
#define FACES_LAUNCH_ELEM(NN) k_elem_lds<NN, 8><<<grid, blk, lds, st>>>(G, I);
  switch (p->n) {
    case 2: FACES_LAUNCH_ELEM(2); break;
    case 3: FACES_LAUNCH_ELEM(3); break;
    case 4: FACES_LAUNCH_ELEM(4); break;
    case 5: FACES_LAUNCH_ELEM(5); break;
    case 6: FACES_LAUNCH_ELEM(6); break;
    case 7: FACES_LAUNCH_ELEM(7); break;
    case 8: FACES_LAUNCH_ELEM(8); break;
    case 9: FACES_LAUNCH_ELEM(9); break;
    case 10: FACES_LAUNCH_ELEM(10); break;
    case 11: FACES_LAUNCH_ELEM(11); break;
    case 12: FACES_LAUNCH_ELEM(12); break;
    default: return fail(FACES_ERR_ARG, "LDS kernel has no n=%d specialisation", p->n);
  }
#undef FACES_LAUNCH_ELEM
  FACES_HIP(hipGetLastError());
  return FACES_OK;
}

// Footprint slots per thread of the streaming kernel for (n, S); mirrors
// StreamShape<N,S>::J.
int stream_slots(int n, int S)
{
  const int K = n - 1, T = kStreamThreads;
  const int own = S * K * n * n, part = 2 * S * K * n + S * n + n * n;
  return (own + T - 1) / T + (part + T - 1) / T;
}

constexpr int kStreamMaxSlots = 12;

// Segment lengths the streaming kernel is compiled for, per n: 1-4 for all,
// and longer ones where n is small (a 4-element item of n=4 is 256 points).
// S = 5 fills the 256-lane slots of n = 5-8 items closer to whole than S = 4
// (n=5: 750 of 768 footprint slots against 605; n=7: 1974 of 2048 against
// 1589 of 1792), S = 10 those of n = 5 (1475 of 1536).
bool stream_case(int n, int S)
{
  switch (S) {
    case 1: case 2: case 3: case 4: return true;
    case 5: return n >= 5 && n <= 8;
    case 10: return n == 5;
    case 6: return n <= 7;
    case 8: return n <= 6;
    case 16: return n <= 4;
    case 32: return n <= 3;
    case 64: return n == 2;
    default: return false;
  }
}

// Elements per item: the largest compiled S not above the requested length
// (default per n, measured on large grids) that divides mx and keeps the
// per-thread tables in registers.
int stream_seg(const faces_plan *p)
{
  if (p->n < 2 || p->n > 12) return 0;
  static const int kDefault[13] = {0, 0, 32, 16, 16, 10, 5, 4, 4, 3, 2, 1, 1};
  const int want = p->seg_len > 0 ? p->seg_len : kDefault[p->n];
  for (int s = std::min(want, 64); s >= 1; s--) {
    if (!stream_case(p->n, s) || p->mx % s) continue;
    if (stream_slots(p->n, s) <= kStreamMaxSlots) return s;
  }
  return 0;
}

// The largest compiled item length up to `want` elements that keeps the
// per-thread tables in registers, fits `cols` columns and (divides) divides
// them; 0 if none.
static int item_len(int n, int want, int cols, bool divides)
{
  for (int s = std::min(std::min(want, 64), cols); s >= 1; s--) {
    if (!stream_case(n, s) || (divides && cols % s)) continue;
    if (stream_slots(n, s) <= kStreamMaxSlots) return s;
  }
  return 0;
}

// Elements per item of the one-item kernel, with its own defaults per n
// (60^3 n=7: 5 elements 249-253 us, 4 elements 254-256 us, 6 elements
// 251-252 us; n=3, 32 elements against 16: 160^3 421 us against 483, 128^3
// 210 against 252, 96x100x200 203 against 232, 64 elements 451 and 222 us;
// n=2, 64 elements against 32: 128^3 50.5 us against 71.5, 256x128x128 94.6
// against 137.2).  Where the longest item dividing mx is shorter than the
// longest that fits by 3 or more elements (61 columns, n=7: 1 against 5),
// items of the longer length cover the first mx - mx % S columns and a
// second launch of items of flow_rest elements the rest: each element owns
// the nodes on its low faces whatever item it is in, so items of different
// lengths sum disjoint nodes.  (n=7, 60x60 rows and planes: 61 columns 243
// us split, 320 us on the work-list kernel; 62 columns 250 us split, 376 us
// with 2-element items; 63 columns 257 us split, 249 us with 3-element
// items.)
int flow_seg(const faces_plan *p)
{
  if (p->n < 2 || p->n > 12) return 0;
  static const int kDefault[13] = {0, 0, 64, 32, 16, 10, 5, 5, 4, 3, 2, 1, 1};
  const int want = p->seg_len > 0 ? p->seg_len : kDefault[p->n];
  const int sd = item_len(p->n, want, p->mx, true), sw = item_len(p->n, want, p->mx, false);
  return sw >= sd + 3 ? sw : sd;
}

// Elements per item of the second launch over the last mx % flow_seg
// columns; 0 if there are none.
int flow_rest(const faces_plan *p)
{
  const int S = flow_seg(p);
  if (S == 0 || p->mx % S == 0) return 0;
  return item_len(p->n, S, p->mx % S, true);
}

// Byte offsets of the streaming kernel are 32-bit: within the field, or,
// for fields of 4 GB and more, within one jz plane.
bool stream_big(const faces_plan *p)
{
  return p->stream_planes || 8 * faces_field_doubles(p->n, p->mx, p->my, p->mz, p->ps) >= kMaxBufferBytes;
}

bool stream_ok(const faces_plan *p)
{
  const long plane = 8L * p->n * p->n * p->n * p->mx * p->my;
  return stream_seg(p) > 0 && plane < kMaxBufferBytes;
}

// Items of S elements over the element columns [jxb, jxb + cols) (the
// streaming kernel: all columns).
template <int N, int S>
int launch_stream_ns(const faces_plan *p, const Geom &G, int mode, int tree, const SurfWork *W,
                     hipStream_t st, int jxb, int cols)
{
  if (p->flow && !stream_big(p)) {
    // one item per block, items in memory order (faces_flow.hpp)
    FlowItems F;
    F.my = p->my;
    F.mz = p->mz;
    F.nseg = cols / S;
    F.jxb = jxb;
    F.nitems = p->my * p->mz * F.nseg;
    F.field_bytes = (unsigned)(8 * faces_field_doubles(p->n, p->mx, p->my, p->mz, p->ps));
    F.xcd_remap = p->xcd_remap;
    F.nbr_mask = p->mask;
    unsigned nb = (unsigned)F.nitems;
    if (F.xcd_remap == 2) {  // the XCD with the largest share of a plane sets the grid
      const int per = p->my * F.nseg;
      int cmax = 0;
      for (int x = 0; x < 8; x++) cmax = std::max(cmax, (x + 1) * per / 8 - x * per / 8);
      nb = 8u * (unsigned)(cmax * p->mz);
    } else if (F.xcd_remap == 1) {
      nb = 8u * (unsigned)((F.nitems + 7) / 8);
    }
    // load and store widths: flow 3 the owned runs in 16-byte pairs, flow 2
    // 16-byte loads on 16-byte boundaries (a 16-byte aligned field; else 8
    // bytes), flow 1 8 bytes per lane; the tree in z, y, x order (the
    // 6-message schedule's) is compiled with the 16-byte pairs only
    int lw = p->flow == 4                                                     ? 3
             : p->flow == 3                                                   ? 2
             : (p->flow == 2 && reinterpret_cast<uintptr_t>(G.u) % 16 == 0) ? 1
                                                                             : 0;
    // the 6-message tree: the owned runs in pairs, the partner regions' too
    // with flow 4; mode 3: the owned runs in pairs
    if ((tree && lw < 2) || mode == 3) lw = 2;
    return launch_flow(N, S, lw, mode, tree, G, F, nb, W, st);
  }
  if (tree || W)
    return fail(FACES_ERR_ARG, "the z, y, x tree and the surface blocks are the one-item kernel's only");
  if (jxb != 0 || cols != p->mx) return fail(FACES_ERR_ARG, "the streaming kernel takes whole element rows");
  StreamItems I = {};
  I.my = p->my;
  I.mz = p->mz;
  I.nseg = p->mx / S;
  I.nitems = p->my * p->mz * I.nseg;
  const bool big = stream_big(p);
  I.field_bytes = big ? (unsigned)(8L * p->n * p->n * p->n * p->mx * p->my)
                      : (unsigned)(8 * faces_field_doubles(p->n, p->mx, p->my, p->mz, p->ps));
  I.xcd_remap = p->xcd_remap;
  I.walk = p->stream_walk;
  I.skip_surface = mode;
  I.nbr_mask = p->mask;
  I.rows = 1;
  if (I.walk == 2 || I.walk == 3) {
    auto divisor_upto = [&](int want) {
      for (int r = std::max(1, std::min(want, p->my)); r > 1; r--)
        if (p->my % r == 0) return r;
      return 1;
    };
    const bool small = 8.0 * faces_field_doubles(p->n, p->mx, p->my, p->mz, p->ps) < 256.0 * 1024 * 1024;
    I.rows = divisor_upto(p->stream_rows > 0   ? p->stream_rows
                          : (I.walk == 3)      ? 4
                          : (p->stream_bpc < 0) ? 10
                          : small              ? 8
                                               : 20);
    // Default column pieces of walk 2: up to 20 rows for persistent blocks
    // (measured best on 60^3 n=7), 8 on fields under 256 MiB (40^3 n=7: 69.9
    // us against 74.4 for 10 rows), up to 10 for one piece per block (60^3
    // n=8 385 us against 396, 40^3 n=12 395 against 409), shorter while they
    // would give fewer than 4 pieces per block (48^3 n=7: 12 rows 135 us, 16
    // rows 140; 40^3 n=10, 2 one-piece blocks per CU: 10 rows 231 us, 20 rows
    // 240), 2 on fields under 256 MiB (32^3 n=8, six fields in turn: 8 rows
    // 63.8 us, 4 rows 66.3).
    if (I.walk == 2 && p->stream_rows == 0 && p->stream_bpc != 0) {
      const long want = (p->stream_bpc > 0 ? (small ? 2L : 4L) * p->stream_bpc : 4L * -p->stream_bpc) * device_cus();
      while (I.rows > 1 && (long)p->mz * I.nseg * (p->my / I.rows) < want) I.rows = divisor_upto(I.rows - 1);
    }
  }
  const long units = (I.walk == 1)   ? (long)p->my * p->mz
                     : (I.walk == 2) ? (long)p->mz * I.nseg * (p->my / I.rows)
                     : (I.walk == 3) ? (long)p->mz * (p->my / I.rows)
                                     : I.nitems;
  // stream_bpc > 0: that many persistent blocks per CU; <= 0: one unit per
  // block, and for -c at most c blocks resident per CU, each block reserving
  // just over 1/(c+1) of the CU's LDS (the kernel uses only its footprint)
  long nb = p->stream_bpc > 0 ? (long)device_cus() * p->stream_bpc : units;
  if (nb > units) nb = units;
  if (I.xcd_remap) nb = (nb + 7) / 8 * 8;
  const size_t lds = stream_lds_doubles<N, S>(I.walk, I.rows) * sizeof(double);
  if (lds > 64 * 1024) return fail(FACES_ERR_ARG, "streaming kernel LDS %zu B above 64 KiB", lds);
  size_t lds_launch = lds;
  if (p->stream_bpc < 0) {
    const size_t share = (size_t)device_lds_per_cu() / (size_t)(1 - p->stream_bpc) + 1024;
    lds_launch = std::max(lds, std::min(share, (size_t)64 * 1024));
  }
  if (mode == 2) {
    if (big)
      k_stream<N, S, true, true><<<dim3((unsigned)nb), dim3(kStreamThreads), lds_launch, st>>>(G, I);
    else
      k_stream<N, S, false, true><<<dim3((unsigned)nb), dim3(kStreamThreads), lds_launch, st>>>(G, I);
  } else if (big) {
    k_stream<N, S, true><<<dim3((unsigned)nb), dim3(kStreamThreads), lds_launch, st>>>(G, I);
  } else {
    k_stream<N, S, false><<<dim3((unsigned)nb), dim3(kStreamThreads), lds_launch, st>>>(G, I);
  }
  FACES_HIP(hipGetLastError());
  return FACES_OK;
}

template <int N>
int launch_stream_n(const faces_plan *p, const Geom &G, int S, int mode, int tree, const SurfWork *W,
                    hipStream_t st, int jxb, int cols)
{
  switch (S) {
    case 1: return launch_stream_ns<N, 1>(p, G, mode, tree, W, st, jxb, cols);
    case 2: return launch_stream_ns<N, 2>(p, G, mode, tree, W, st, jxb, cols);
    case 3: return launch_stream_ns<N, 3>(p, G, mode, tree, W, st, jxb, cols);
    case 4: return launch_stream_ns<N, 4>(p, G, mode, tree, W, st, jxb, cols);
    default: break;
  }
  if constexpr (N >= 5 && N <= 8) {
    if (S == 5) return launch_stream_ns<N, 5>(p, G, mode, tree, W, st, jxb, cols);
  }
  if constexpr (N == 5) {
    if (S == 10) return launch_stream_ns<N, 10>(p, G, mode, tree, W, st, jxb, cols);
  }
  if constexpr (N <= 7) {
    if (S == 6) return launch_stream_ns<N, 6>(p, G, mode, tree, W, st, jxb, cols);
  }
  if constexpr (N <= 6) {
    if (S == 8) return launch_stream_ns<N, 8>(p, G, mode, tree, W, st, jxb, cols);
  }
  if constexpr (N <= 4) {
    if (S == 16) return launch_stream_ns<N, 16>(p, G, mode, tree, W, st, jxb, cols);
  }
  if constexpr (N <= 3) {
    if (S == 32) return launch_stream_ns<N, 32>(p, G, mode, tree, W, st, jxb, cols);
  }
  if constexpr (N == 2) {
    if (S == 64) return launch_stream_ns<N, 64>(p, G, mode, tree, W, st, jxb, cols);
  }
  return fail(FACES_ERR_ARG, "streaming kernel has no n=%d S=%d specialisation", N, S);
}

template <int N>
int launch_stream_cols(const faces_plan *p, const Geom &G, int S, int mode, int tree, const SurfWork *W,
                       hipStream_t st)
{
  const bool flow = p->flow && !stream_big(p);
  const int cols = flow ? p->mx - p->mx % S : p->mx;
  int rc = launch_stream_n<N>(p, G, S, mode, tree, W, st, 0, cols);
  if (rc || cols == p->mx) return rc;
  // the last mx % S columns: items of their own length, no surface blocks
  SurfWork w = {};
  if (W) {
    w = *W;
    w.nblk = 0;
  }
  return launch_stream_n<N>(p, G, flow_rest(p), mode, tree, W ? &w : nullptr, st, cols, p->mx - cols);
}

int launch_stream(const faces_plan *p, double *u, int mode, hipStream_t st, int tree = 0,
                  const SurfWork *W = nullptr)
{
  const Geom G = make_geom(p, u);
  const int S = (p->flow && !stream_big(p)) ? flow_seg(p) : stream_seg(p);
  if (S == 0) return fail(FACES_ERR_ARG, "no streaming item length for n=%d mx=%d", p->n, p->mx);
  if (p->flow && !stream_big(p) && p->mx % S && flow_rest(p) == 0)
    return fail(FACES_ERR_ARG, "no item length for the last %d columns", p->mx % S);
  switch (p->n) {
    case 2: return launch_stream_cols<2>(p, G, S, mode, tree, W, st);
    case 3: return launch_stream_cols<3>(p, G, S, mode, tree, W, st);
    case 4: return launch_stream_cols<4>(p, G, S, mode, tree, W, st);
    case 5: return launch_stream_cols<5>(p, G, S, mode, tree, W, st);
    case 6: return launch_stream_cols<6>(p, G, S, mode, tree, W, st);
    case 7: return launch_stream_cols<7>(p, G, S, mode, tree, W, st);
    case 8: return launch_stream_cols<8>(p, G, S, mode, tree, W, st);
    case 9: return launch_stream_cols<9>(p, G, S, mode, tree, W, st);
    case 10: return launch_stream_cols<10>(p, G, S, mode, tree, W, st);
    case 11: return launch_stream_cols<11>(p, G, S, mode, tree, W, st);
    case 12: return launch_stream_cols<12>(p, G, S, mode, tree, W, st);
    default: return fail(FACES_ERR_ARG, "streaming kernel has no n=%d specialisation", p->n);
  }
}

// The work-list item kernel (faces_batch.hip) takes ragged element rows: it
// serves grids whose mx leaves the streaming kernel only 1- or 2-element
// items.
bool items_ok(const faces_plan *p)
{
  const long bytes = 8 * faces_field_doubles(p->n, p->mx, p->my, p->mz, p->ps);
  return p->n >= 2 && p->n <= 12 && bytes < kMaxBufferBytes && batch_seg(p->n, p->mx, 64) > 0;
}

// The sum kernel faces_share launches on a periodic field with compute = 1.
int active_kernel(const faces_plan *p)
{
  switch (p->variant) {
    case FACES_VARIANT_DIRECT: return FACES_VARIANT_DIRECT;
    case FACES_VARIANT_LDS: return elem_lds_ok(p) ? FACES_VARIANT_LDS : FACES_VARIANT_DIRECT;
    case FACES_VARIANT_STREAM: return stream_ok(p) ? FACES_VARIANT_STREAM : FACES_VARIANT_DIRECT;
    case FACES_VARIANT_ITEMS: return items_ok(p) ? FACES_VARIANT_ITEMS : FACES_VARIANT_DIRECT;
    default:
      // the work-list kernel only where its items (ragged rows allowed) are
      // more than twice as long as the streaming kernel's (the one-item
      // kernel's, where periodic steps take it)
      if (stream_ok(p) &&
          !(items_ok(p) && batch_seg(p->n, p->mx, 64) > 2 * (p->flow && !stream_big(p) ? flow_seg(p) : stream_seg(p))))
        return FACES_VARIANT_STREAM;
      if (items_ok(p)) return FACES_VARIANT_ITEMS;
      return stream_ok(p) ? FACES_VARIANT_STREAM : FACES_VARIANT_DIRECT;
  }
}

int launch_box(const faces_plan *p, double *u, bool skip_surface, hipStream_t st)
{
  int kern = active_kernel(p);
  // interior only (open ranks): the streaming kernel skips the surface
  // nodes; the other kernels fall back to the direct one
  if (skip_surface && kern != FACES_VARIANT_STREAM)
    kern = (p->variant == FACES_VARIANT_AUTO && stream_ok(p)) ? FACES_VARIANT_STREAM : FACES_VARIANT_DIRECT;
  if (kern == FACES_VARIANT_STREAM) return launch_stream(p, u, skip_surface ? 1 : 0, st);
  if (kern == FACES_VARIANT_ITEMS)
    return launch_single(u, p->n, p->mx, p->my, p->mz, p->ps, p->seg_len > 0 ? p->seg_len : 64,
                         p->items_bpc, st);
  if (!skip_surface && p->variant == FACES_VARIANT_LDS && elem_lds_ok(p))
    return launch_elem_lds(p, u, st);
  Geom G = make_geom(p, u);
  RowItems R;
  R.my = p->my;
  R.mz = p->mz;
  R.seg_len = (p->seg_len <= 0 || p->seg_len > p->mx) ? p->mx : p->seg_len;
  R.segs = (p->mx + R.seg_len - 1) / R.seg_len;
  R.nitems = p->my * p->mz * R.segs;
  R.xcd_remap = p->xcd_remap;
  const int nb = R.xcd_remap ? 8 * ((R.nitems + 7) / 8) : R.nitems;
  const dim3 grid(nb), blk(p->block);
#define FACES_LAUNCH_BOX(NN)                                              \
  if (skip_surface) k_box<NN, true><<<grid, blk, 0, st>>>(G, R);          \
  else k_box<NN, false><<<grid, blk, 0, st>>>(G, R);
  FACES_DISPATCH_N(p->n, FACES_LAUNCH_BOX)
#undef FACES_LAUNCH_BOX
  FACES_HIP(hipGetLastError());
  return FACES_OK;
}

// Whether the interior sums (launch_box with the surface skipped) run on the
// one-item kernel, whose items leave every rank-surface location alone (mode
// 1 marks those copies and stores none of them): its launch can take the
// surface sums as extra blocks.
bool interior_takes_surface(const faces_plan *p)
{
  int kern = active_kernel(p);
  if (kern != FACES_VARIANT_STREAM)
    kern = (p->variant == FACES_VARIANT_AUTO && stream_ok(p)) ? FACES_VARIANT_STREAM : FACES_VARIANT_DIRECT;
  return p->surface_blocks && kern == FACES_VARIANT_STREAM && p->flow && !stream_big(p) && flow_seg(p) > 0;
}

// The interior sums and, as extra blocks of the same launch, the surface
// sums of the staged step (interior_takes_surface(p) must hold).
int launch_inner_surface(const faces_plan *p, double *u, const Bound &b, hipStream_t st)
{
  SurfWork W = {};
  W.B = b;
  W.recv = p->arena;
  W.sl = p->recv;
  W.S = surf_desc(p, b);
  const long total = W.S.cz + W.S.cy + W.S.cx;
  if (p->surface_blocks == 3) return launch_stream(p, u, 3, st, 0, &W);  // summed by the items
  W.nblk = total > 0 ? (unsigned)grid_for(total, kStreamThreads) : 0u;
  W.lead = p->surface_blocks == 1;
  return launch_stream(p, u, 1, st, 0, &W);
}

int launch_surface(const faces_plan *p, double *u, const Bound &b, hipStream_t st)
{
  Geom G = make_geom(p, u);
  SurfDesc S = surf_desc(p, b);
  const long total = S.cz + S.cy + S.cx;
  if (total == 0) return FACES_OK;
  const int blk = 256;
  const int grid = grid_for(total, blk);
  const double *recv = p->arena;
  Slots sl = p->recv;
#define FACES_LAUNCH_SURF(NN) k_surface<NN><<<grid, blk, 0, st>>>(G, b, recv, sl, S);
  FACES_DISPATCH_N(p->n, FACES_LAUNCH_SURF)
#undef FACES_LAUNCH_SURF
  FACES_HIP(hipGetLastError());
  return FACES_OK;
}

int launch_pack(const faces_plan *p, const double *u, bool direct, hipStream_t st)
{
  Geom G = make_geom(p, const_cast<double *>(u));
  PackDesc P;
  P.src_off[0] = 0;
  for (int k = 0; k < 26; k++) P.src_off[k + 1] = P.src_off[k] + p->slot_cnt[k];
  for (int k = 0; k < 26; k++) {
    if (!direct) {
      P.dst[k] = p->send.off[k];
    } else {
      const int o = opposite_slot(k);
      P.dst[k] = (p->nbr[o] >= 0) ? p->recv.off[o] : -1;
    }
  }
  long nb = 0;
  for (int r = 0; r < 6; r++) {
    P.bstart[r] = (int)nb;
    if (P.dst[r] < 0) continue;  // a skipped slot takes no blocks
    nb += (p->slot_cnt[r] + kPackFaceEntries - 1) / kPackFaceEntries;
  }
  P.bstart[6] = (int)nb;
  nb += (P.src_off[26] - P.src_off[6] + 255) / 256;
  P.bstart[7] = (int)nb;
  if (nb > 0x7fffffffL / 2) return fail(FACES_ERR_ARG, "pack launch of %ld blocks", nb);
  const dim3 grid((unsigned)nb), blk(256);
#define FACES_LAUNCH_PACK(NN) k_pack<NN><<<grid, blk, 0, st>>>(G, P, p->arena);
  FACES_DISPATCH_N(p->n, FACES_LAUNCH_PACK)
#undef FACES_LAUNCH_PACK
  FACES_HIP(hipGetLastError());
  return FACES_OK;
}

int launch_exchange(const faces_plan *p, hipStream_t st)
{
  XchgDesc X;
  int ns = 0;
  X.idx_off[0] = 0;
  for (int k = 0; k < 26; k++) {
    if (p->nbr[k] < 0) continue;
    X.src[ns] = p->send.off[opposite_slot(k)];
    X.dst[ns] = p->recv.off[k];
    X.idx_off[ns + 1] = X.idx_off[ns] + p->slot_cnt[k];
    ns++;
  }
  X.nslots = ns;
  if (ns == 0) return FACES_OK;
  const int blk = 256;
  const int grid = grid_for(X.idx_off[ns], blk);
  k_exchange<<<grid, blk, 0, st>>>(X, p->arena);
  FACES_HIP(hipGetLastError());
  return FACES_OK;
}

int check_plan(const faces_plan *p, const void *u)
{
  if (!p) return fail(FACES_ERR_ARG, "null plan");
  if (!u) return fail(FACES_ERR_ARG, "null field pointer");
  return FACES_OK;
}

}  // namespace

extern "C" {

const char *faces_strerror(int status)
{
  switch (status) {
    case FACES_OK: return "success";
    case FACES_ERR_ARG: return "invalid argument";
    case FACES_ERR_HIP: return "HIP runtime error";
    case FACES_ERR_NOMEM: return "out of memory";
    case FACES_ERR_STATE: return "invalid state";
    case FACES_ERR_ADMISSION: return "request exceeds footprint cap";
    case FACES_ERR_NODEV: return "no HIP device";
    case FACES_ERR_TIMEOUT: return "timed out";
    default: return "unknown status";
  }
}

const char *faces_last_error(void) { return g_last_error.c_str(); }

const char *faces_version(void) { return "faces_amd 0.1 (gfx950)"; }

long faces_field_doubles(int n, int mx, int my, int mz, long plane_stride)
{
  if (n < 1 || mx < 1 || my < 1 || mz < 1) return -1;
  const long dense = (long)n * n * n * mx * my;
  const long ps = plane_stride > 0 ? plane_stride : dense;
  return ps * (long)(mz - 1) + dense;
}

}  // extern "C"

namespace {

// Everything of a plan but its device arena: argument checks, rank
// coordinates, neighbour table, message sizes, launch defaults, and the
// arena layout (send slots, then receive slots, 512-byte aligned).  Host
// only: allocates nothing.
int plan_init(faces_plan *p, int rank, int lx, int ly, int lz, int mx, int my, int mz, int n,
              int boundary, long plane_stride)
{
  if (n < 2 || n > kMaxN) return fail(FACES_ERR_ARG, "element size n=%d outside [2,%d]", n, kMaxN);
  if (mx < 1 || my < 1 || mz < 1)
    return fail(FACES_ERR_ARG, "element counts must be positive (%d,%d,%d)", mx, my, mz);
  if (lx < 1 || ly < 1 || lz < 1)
    return fail(FACES_ERR_ARG, "rank grid must be positive (%d,%d,%d)", lx, ly, lz);
  if ((long)lx * ly * lz > (1L << 30)) return fail(FACES_ERR_ARG, "rank grid too large");
  if (rank < 0 || rank >= lx * ly * lz)
    return fail(FACES_ERR_ARG, "rank %d outside rank grid of %d", rank, lx * ly * lz);
  if (boundary != FACES_BOUNDARY_OPEN && boundary != FACES_BOUNDARY_PERIODIC)
    return fail(FACES_ERR_ARG, "unknown boundary %d", boundary);
  const long dense = (long)n * n * n * mx * my;
  if (plane_stride != 0 && plane_stride < dense)
    return fail(FACES_ERR_ARG, "plane stride %ld below dense plane %ld", plane_stride, dense);
  // Row-local node indices are int; global node counts per axis must fit too.
  if ((long)mx * (n - 1) + 1 > (1L << 30) || (long)my * (n - 1) + 1 > (1L << 30) ||
      (long)mz * (n - 1) + 1 > (1L << 30) || (long)mx * (n - 1) * (n - 1) * (n - 1) > (1L << 30))
    return fail(FACES_ERR_ARG, "grid too large for one plan");
  p->rank = rank;
  p->lx = lx; p->ly = ly; p->lz = lz;
  // gpu/v-26-messages/Faces.cc:33-35
  p->kz = rank / (lx * ly);
  p->ky = (rank - p->kz * lx * ly) / lx;
  p->kx = rank - (p->kz * ly + p->ky) * lx;
  p->mx = mx; p->my = my; p->mz = mz; p->n = n;
  p->boundary = boundary;
  p->ps = plane_stride > 0 ? plane_stride : dense;
  p->mask = 0;
  for (int k = 0; k < 26; k++) {
    int dx, dy, dz;
    slot_dir(k, dx, dy, dz);
    p->nbr[k] = neighbor_of(p, dx, dy, dz);
  }
  for (int k = 0; k < 6; k++)
    if (p->nbr[k] >= 0) p->mask |= 1 << k;
  p->path = FACES_PATH_AUTO;
  p->block = 256;
  p->seg_len = 0;
  // XCD order: each XCD takes an eighth of every plane's items (the one-item
  // kernel; 64^3 n=8 410 us against 423 for contiguous eighths of the list,
  // 80^3 n=6 356 against 397); the other kernels read it as 1, contiguous
  // eighths of their work list (n=2, 64-element items: 128^3 50.5 us
  // against 53.0, 256x128x128 94.6 against 120.7).
  p->xcd_remap = 2;
  p->variant = FACES_VARIANT_AUTO;
  p->lds_block = 256;
  // Blocks: 2 persistent blocks per CU for n = 6-7 (60^3 n=7 276 us against
  // 281-284 us for one unit per block); one unit (column piece or row) per
  // block, at most 2 resident per CU, for n >= 8 (64^3 n=8 448-450 us
  // against 480, 40^3 n=12 395 against 460, 48^3 n=9 287 against 305), at
  // most 3 for n <= 5 (128^3 n=4 478 against 518, 100^3 n=5 503 against 528).
  // Fields under 256 MiB keep 2 persistent blocks for n >= 8 (32^3 n=8, six
  // fields in turn: 66 us against 71).
  const bool small = 8.0 * faces_field_doubles(n, mx, my, mz, plane_stride) < 256.0 * 1024 * 1024;
  p->stream_bpc = (n >= 8) ? (small ? 2 : -2) : (n >= 6) ? 2 : -3;
  // column pieces walked in y for n >= 7 (60^3 n=7 275 us against 282-288 us
  // for rows walked in x, 64^3 n=8 476 against 511, 40^3 n=12 457 against
  // 472) where my has a divisor from 4 to 20 to cut the columns by (95x94x93
  // n=12, my = 2*47: 2-row pieces 6.41 ms against 5.92 ms for rows in x);
  // rows walked in x for n <= 6 (128^3 n=4 515 against 532)
  int piece = 1;
  for (int r = std::min(20, my); r > 1 && piece == 1; r--)
    if (my % r == 0) piece = r;
  p->stream_walk = (n >= 7 && piece >= 4) ? 2 : 1;
  p->stream_rows = 0;
  // work-list kernel on one grid (rows without a suitable divisor): 16
  // blocks per CU for n >= 7 (61x60x60 n=7 320 us against 382 us for 4), 32
  // for n <= 6 (79x80x80 n=6 475 against 565)
  p->items_bpc = (n >= 7) ? 16 : 32;
  p->stream_planes = 0;
  // The one-item kernel for periodic steps and open ranks: 60^3 n=7 230-236
  // us against 274 us for the streaming kernel's column walk, 64^3 n=8 410
  // against 449, 128^3 n=4 413-426 against 477.  Load width per n (us; flow
  // 2: 16-byte loads on 16-byte boundaries, 3: the owned runs in 16-byte
  // pairs, 4: the partner regions' runs in pairs too): 60^3 n=7 229.3 /
  // 228.3 / 226.2; 80^3 n=6 356 / 366 / 349.6; 48^3 n=9 247 / 247 / 242;
  // 40^3 n=11 280 / 275 / 267; 128^3 n=4 412 / 426 / 411; 100^3 n=5 413 /
  // 413 / 412; 64^3 n=8 409 / 415 / 408; but 160^3 n=3 494 / 482 / 527,
  // 44^3 n=10 266.5 / 281 / 272, 40^3 n=12 358 / 416 / 399, 256^3 n=2 728 /
  // 745 / 763.  n=2 takes 8-byte lanes (flow 1; 64-element items, flows 1
  // / 3: 128^3 50.5 / 54.3 us, 256x128x128 94.6 / 103.3).  n=3 with its 32-element items (flows 1 / 2 /
  // 3 / 4): 160^3 478.7 / 442.0 / 424.5 / 421.2, 128^3 245.7 / 222.2 / 215.0
  // / 209.7.
  static const int kFlowDefault[13] = {0, 0, 1, 4, 4, 4, 4, 4, 4, 4, 2, 4, 2};
  p->flow = (n >= 2 && n <= 12) ? kFlowDefault[n] : 2;
  // staged step, 60^3 n=7, interior rank of 3x3x3: surface blocks after the
  // items 259.7 us, before them 261, a launch of their own 267.5, summed by
  // the items (their scattered loads of received copies wait in the surface
  // items) 265.8; corner rank: 257.8, -, 265.4, 252.2
  p->surface_blocks = 2;

  // Message sizes, gpu/v-26-messages/Faces.cc:44-47, 64-66.
  const long nn = (long)n * n;
  for (int s = 0; s < 2; s++) {
    p->slot_cnt[0 + s] = nn * my * mz;
    p->slot_cnt[2 + s] = nn * mx * mz;
    p->slot_cnt[4 + s] = nn * mx * my;
  }
  for (int e = 0; e < 4; e++) {
    p->slot_cnt[6 + e] = (long)n * mx;
    p->slot_cnt[10 + e] = (long)n * my;
    p->slot_cnt[14 + e] = (long)n * mz;
  }
  for (int c = 0; c < 8; c++) p->slot_cnt[18 + c] = 1;
  // Slots start on 64-double (512 B) boundaries.
  long off = 0;
  for (int side = 0; side < 2; side++) {
    Slots &sl = side ? p->recv : p->send;
    for (int k = 0; k < 26; k++) {
      sl.off[k] = off;
      off += (p->slot_cnt[k] + 63) / 64 * 64;
    }
    sl.off[26] = off;
  }
  p->arena_doubles = off;
  p->arena = nullptr;
  return FACES_OK;
}

}  // namespace

extern "C" {

int faces_plan_create(faces_plan **out, int rank, int lx, int ly, int lz, int mx, int my, int mz,
                      int n, int boundary, long plane_stride)
{
  if (!out) return fail(FACES_ERR_ARG, "null output pointer");
  *out = nullptr;
  faces_plan *p = new (std::nothrow) faces_plan();
  if (!p) return fail(FACES_ERR_NOMEM, "plan allocation");
  int st = plan_init(p, rank, lx, ly, lz, mx, my, mz, n, boundary, plane_stride);
  if (st) {
    delete p;
    return st;
  }
  const long off = p->arena_doubles;
  hipError_t e = hipMalloc(&p->arena, sizeof(double) * (size_t)off);
  if (e != hipSuccess) {
    delete p;
    return hip_fail(e, "hipMalloc(message arena)");
  }
  // Receives from a missing neighbour leave zeros (gpu/DArray.h:81).  The
  // clear runs on the library's own setup stream and only that stream is
  // waited for, so creating a plan never waits for other work on the device.
  hipStream_t zs = nullptr;
  e = setup_stream(&zs);
  if (e == hipSuccess) e = hipMemsetAsync(p->arena, 0, sizeof(double) * (size_t)off, zs);
  if (e == hipSuccess) e = hipStreamSynchronize(zs);
  if (e != hipSuccess) {
    (void)hipFree(p->arena);
    delete p;
    return hip_fail(e, "hipMemset(message arena)");
  }
  *out = p;
  return FACES_OK;
}

int faces_plan_destroy(faces_plan *p)
{
  if (!p) return FACES_OK;
  hipError_t e = hipSuccess;
  if (p->arena) e = hipFree(p->arena);
  delete p;
  if (e != hipSuccess) return hip_fail(e, "hipFree(message arena)");
  return FACES_OK;
}

int faces_plan_set_path(faces_plan *p, int path)
{
  if (!p) return fail(FACES_ERR_ARG, "null plan");
  if (path < FACES_PATH_AUTO || path > FACES_PATH_SIX)
    return fail(FACES_ERR_ARG, "unknown path %d", path);
  p->path = path;
  return FACES_OK;
}

int faces_plan_set_tuning(faces_plan *p, int block, int seg_len, int xcd_remap)
{
  if (!p) return fail(FACES_ERR_ARG, "null plan");
  if (block < 64 || block > 1024 || block % 64)
    return fail(FACES_ERR_ARG, "block %d not a multiple of 64 in [64,1024]", block);
  if (seg_len < 0) return fail(FACES_ERR_ARG, "negative segment length");
  p->block = block;
  p->seg_len = seg_len;
  p->xcd_remap = xcd_remap ? 1 : 0;
  return FACES_OK;
}

int faces_plan_set_variant(faces_plan *p, int variant, int block)
{
  if (!p) return fail(FACES_ERR_ARG, "null plan");
  if (variant < FACES_VARIANT_AUTO || variant > FACES_VARIANT_ITEMS)
    return fail(FACES_ERR_ARG, "unknown kernel variant %d", variant);
  if (variant == FACES_VARIANT_ITEMS) {
    if (block < 0 || block > 256) return fail(FACES_ERR_ARG, "blocks per CU %d outside [0,256]", block);
    if (!items_ok(p))
      return fail(FACES_ERR_ARG, "item kernel does not fit this grid (n=%d, mx=%d)", p->n, p->mx);
    p->variant = variant;
    if (block) p->items_bpc = block;
    return FACES_OK;
  }
  if (variant == FACES_VARIANT_STREAM) {
    if (block < -1 || block > 32)
      return fail(FACES_ERR_ARG, "blocks per CU %d outside [-1,32] (-1: one item per block)", block);
    if (!stream_ok(p))
      return fail(FACES_ERR_ARG, "streaming kernel does not fit this grid (n=%d, mx=%d)", p->n, p->mx);
    p->variant = variant;
    if (block) p->stream_bpc = block < 0 ? 0 : block;
    return FACES_OK;
  }
  if (block != 0 && (block < 64 || block > 512 || block % 64))
    return fail(FACES_ERR_ARG, "LDS kernel block %d not a multiple of 64 in [64,512]", block);
  if (variant == FACES_VARIANT_LDS && !elem_lds_ok(p))
    return fail(FACES_ERR_ARG, "LDS row kernel does not fit this grid (n=%d, mx=%d)", p->n, p->mx);
  p->variant = variant;
  if (block) p->lds_block = block;
  return FACES_OK;
}

int faces_plan_set_option(faces_plan *p, int option, int value)
{
  if (!p) return fail(FACES_ERR_ARG, "null plan");
  switch (option) {
    case FACES_OPT_VARIANT: return faces_plan_set_variant(p, value, 0);
    case FACES_OPT_DIRECT_BLOCK: return faces_plan_set_tuning(p, value, p->seg_len, p->xcd_remap);
    case FACES_OPT_SEG:
      if (value < 0) return fail(FACES_ERR_ARG, "negative segment length");
      p->seg_len = value;
      return FACES_OK;
    case FACES_OPT_XCD:
      if (value < 0 || value > 2) return fail(FACES_ERR_ARG, "XCD order %d outside [0,2]", value);
      p->xcd_remap = value;
      return FACES_OK;
    case FACES_OPT_LDS_BLOCK:
      if (value < 64 || value > 512 || value % 64)
        return fail(FACES_ERR_ARG, "LDS kernel block %d not a multiple of 64 in [64,512]", value);
      p->lds_block = value;
      return FACES_OK;
    case FACES_OPT_STREAM_BPC:
      if (value < -4 || value == -1 || value > 32)
        return fail(FACES_ERR_ARG, "blocks per CU %d outside [-4,-2] and [0,32]", value);
      p->stream_bpc = value;
      return FACES_OK;
    case FACES_OPT_STREAM_WALK:
      if (value < 0 || value > 3) return fail(FACES_ERR_ARG, "walk %d outside [0,3]", value);
      p->stream_walk = value;
      return FACES_OK;
    case FACES_OPT_STREAM_ROWS:
      if (value < 0 || value > 64) return fail(FACES_ERR_ARG, "rows per group %d outside [0,64]", value);
      p->stream_rows = value;
      return FACES_OK;
    case FACES_OPT_STREAM_PLANES:
      p->stream_planes = value ? 1 : 0;
      return FACES_OK;
    case FACES_OPT_FLOW:
      if (value < 0 || value > 4) return fail(FACES_ERR_ARG, "one-item kernel mode %d outside [0,4]", value);
      p->flow = value;
      return FACES_OK;
    case FACES_OPT_SURFACE_BLOCKS:
      if (value < 0 || value > 3) return fail(FACES_ERR_ARG, "surface sums placement %d outside [0,3]", value);
      p->surface_blocks = value;
      return FACES_OK;
    default: return fail(FACES_ERR_ARG, "unknown option %d", option);
  }
}

int faces_plan_get_option(const faces_plan *p, int option, int *value)
{
  if (!p || !value) return fail(FACES_ERR_ARG, "null argument");
  switch (option) {
    case FACES_OPT_VARIANT: *value = p->variant; break;
    case FACES_OPT_DIRECT_BLOCK: *value = p->block; break;
    case FACES_OPT_SEG: *value = p->seg_len; break;
    case FACES_OPT_XCD: *value = p->xcd_remap; break;
    case FACES_OPT_LDS_BLOCK: *value = p->lds_block; break;
    case FACES_OPT_STREAM_BPC: *value = p->stream_bpc; break;
    case FACES_OPT_STREAM_WALK: *value = p->stream_walk; break;
    case FACES_OPT_STREAM_ROWS: *value = p->stream_rows; break;
    case FACES_OPT_STREAM_PLANES: *value = stream_big(p) ? 1 : 0; break;
    case FACES_OPT_ACTIVE_KERNEL: *value = active_kernel(p); break;
    case FACES_OPT_FLOW: *value = p->flow; break;
    case FACES_OPT_SURFACE_BLOCKS: *value = p->surface_blocks; break;
    default: return fail(FACES_ERR_ARG, "unknown option %d", option);
  }
  return FACES_OK;
}

int faces_plan_tune(faces_plan *p, double *u, int reps, void *stream, double *out)
{
  int st = check_plan(p, u);
  if (st) return st;
  if (reps < 1 || reps > 1000) return fail(FACES_ERR_ARG, "reps %d outside [1,1000]", reps);
  struct Cand {
    int variant, walk, rows, seg, bpc, flow, xcd;
  };
  std::vector<Cand> cands;
  const int seg0 = p->seg_len;
  if (stream_ok(p) && !stream_big(p)) {
    // the one-item kernel: the three longest item lengths, every load width,
    // both XCD orders
    std::vector<int> segs;
    for (int sl = 64; sl >= 1 && segs.size() < 3; sl--) {
      p->seg_len = sl;
      if (flow_seg(p) == sl) segs.push_back(sl);
    }
    for (int fl : {4, 3, 2, 1})
      for (int sl : segs)
        for (int x : {2, 1}) cands.push_back({FACES_VARIANT_STREAM, p->stream_walk, p->stream_rows, sl, p->stream_bpc, fl, x});
  }
  if (stream_ok(p)) {
    // the three longest compiled segment lengths that divide mx
    std::vector<int> segs;
    for (int sl = 32; sl >= 1 && segs.size() < 3; sl--) {
      p->seg_len = sl;
      if (stream_seg(p) == sl) segs.push_back(sl);
    }
    // (walk, rows): rows in x; column pieces in y of the default length and
    // of 10 rows; groups of 2 and 4 rows
    const int wr[5][2] = {{1, 1}, {2, 0}, {2, 10}, {3, 2}, {3, 4}};
    for (const auto &w : wr)
      for (int sl : segs)
        for (int bpc : {1, 2, 3, 4, -2, -3}) cands.push_back({FACES_VARIANT_STREAM, w[0], w[1], sl, bpc, 0, 1});
  }
  p->seg_len = seg0;
  if (items_ok(p))
    for (int bpc : {2, 4, 8, 16, 32, 64})
      cands.push_back({FACES_VARIANT_ITEMS, p->stream_walk, p->stream_rows, 0, bpc, 0, 1});
  if (cands.empty()) return fail(FACES_ERR_STATE, "no tunable kernel for this grid");

  const int v0 = p->variant, w0 = p->stream_walk, r0 = p->stream_rows, b0 = p->stream_bpc, i0 = p->items_bpc;
  const int f0 = p->flow, x0 = p->xcd_remap;
  auto apply = [&](const Cand &c) {
    p->variant = c.variant;
    p->stream_walk = c.walk;
    p->stream_rows = c.rows;
    p->seg_len = c.seg;
    p->flow = c.flow;
    p->xcd_remap = c.xcd;
    if (c.variant == FACES_VARIANT_ITEMS) p->items_bpc = c.bpc;
    else p->stream_bpc = c.bpc;
  };
  hipStream_t hs = (hipStream_t)stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    return fail(FACES_ERR_HIP, "hipEventCreate");
  }
  double best = 1e300;
  Cand bestc = cands[0];
  for (const Cand &c : cands) {
    apply(c);
    st = faces_share(p, u, 1, stream);  // warm-up
    if (!st && hipEventRecord(e0, hs) != hipSuccess) st = fail(FACES_ERR_HIP, "hipEventRecord");
    for (int r = 0; r < reps && !st; r++) st = faces_share(p, u, 1, stream);
    float ms = 0.f;
    if (!st && (hipEventRecord(e1, hs) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
      st = fail(FACES_ERR_HIP, "timing events");
    if (st) break;
    const double us = 1e3 * ms / reps;
    if (us < best) {
      best = us;
      bestc = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st) {  // restore the settings the plan had
    p->variant = v0;
    p->stream_walk = w0;
    p->stream_rows = r0;
    p->seg_len = seg0;
    p->stream_bpc = b0;
    p->items_bpc = i0;
    p->flow = f0;
    p->xcd_remap = x0;
    return st;
  }
  apply(bestc);
  if (out) {
    out[0] = bestc.variant;
    out[1] = bestc.flow ? -bestc.flow : bestc.walk;
    out[2] = bestc.variant == FACES_VARIANT_ITEMS ? batch_seg(p->n, p->mx, 64)
             : bestc.flow                         ? flow_seg(p)
                                                  : stream_seg(p);
    out[3] = bestc.flow ? bestc.xcd : bestc.bpc;
    out[4] = best;
  }
  return FACES_OK;
}

int faces_plan_neighbor(const faces_plan *p, int slot, int *rank)
{
  if (!p || !rank) return fail(FACES_ERR_ARG, "null argument");
  if (slot < 0 || slot >= 26) return fail(FACES_ERR_ARG, "slot %d outside [0,26)", slot);
  *rank = p->nbr[slot];
  return FACES_OK;
}

int faces_plan_message(faces_plan *p, int slot, int recv, double **dptr, long *count)
{
  if (!p || !dptr || !count) return fail(FACES_ERR_ARG, "null argument");
  if (slot < 0 || slot >= 26) return fail(FACES_ERR_ARG, "slot %d outside [0,26)", slot);
  const Slots &sl = recv ? p->recv : p->send;
  *dptr = p->arena + sl.off[slot];
  *count = p->slot_cnt[slot];
  return FACES_OK;
}

int faces_pack(faces_plan *p, const double *u, void *stream)
{
  int st = check_plan(p, u);
  if (st) return st;
  return launch_pack(p, u, false, (hipStream_t)stream);
}

int faces_exchange(faces_plan *p, void *stream)
{
  if (!p) return fail(FACES_ERR_ARG, "null plan");
  return launch_exchange(p, (hipStream_t)stream);
}

int faces_inner(faces_plan *p, double *u, void *stream)
{
  int st = check_plan(p, u);
  if (st) return st;
  return launch_box(p, u, true, (hipStream_t)stream);
}

int faces_unpack(faces_plan *p, double *u, void *stream)
{
  int st = check_plan(p, u);
  if (st) return st;
  return launch_surface(p, u, bound_recv(p->mask), (hipStream_t)stream);
}

int faces_inner_unpack(faces_plan *p, double *u, void *stream)
{
  int st = check_plan(p, u);
  if (st) return st;
  const hipStream_t s = (hipStream_t)stream;
  if (interior_takes_surface(p)) return launch_inner_surface(p, u, bound_recv(p->mask), s);
  st = launch_box(p, u, true, s);
  if (st) return st;
  return launch_surface(p, u, bound_recv(p->mask), s);
}

int faces_share(faces_plan *p, double *u, int compute, void *stream)
{
  int st = check_plan(p, u);
  if (st) return st;
  hipStream_t s = (hipStream_t)stream;
  int path = p->path;
  // An open rank takes the fused path where one streaming launch sums it
  // (its neighbours stood in for by self halos), else the staged one.
  const bool open_stream = p->boundary == FACES_BOUNDARY_OPEN && compute &&
                           active_kernel(p) == FACES_VARIANT_STREAM;
  if (path == FACES_PATH_AUTO)
    path = (p->boundary == FACES_BOUNDARY_PERIODIC || p->mask == 0 || p->mask == 63 || open_stream)
               ? FACES_PATH_FUSED
               : FACES_PATH_STAGED_DIRECT;
  if (path == FACES_PATH_SIX) {
    // z, y, x phases; self halos on the sides where a neighbour exists
    const int mask = (p->boundary == FACES_BOUNDARY_PERIODIC) ? 63 : p->mask;
    if (compute && p->flow && !stream_big(p) && flow_seg(p) > 0) {
      // One launch: a node's three phases are its sum tree with the levels
      // in z, y, x order, which the one-item kernel evaluates on its copies
      // (an open rank with all six face neighbours sums the periodic wrap
      // partners, as in the fused 26-message step).
      return launch_stream(p, u, mask == 63 ? 0 : 2, s, 1);
    }
    Geom G = make_geom(p, u);
    const int m[3] = {p->mx, p->my, p->mz};
    for (int axis = 2; axis >= 0; axis--) {
      const long total = (long)p->n * p->n * m[0] * m[1] * m[2];
      k_six<<<dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s>>>(
          G, axis, (mask >> (2 * axis)) & 1, (mask >> (2 * axis + 1)) & 1, compute ? 1 : 0);
      FACES_HIP(hipGetLastError());
    }
    return FACES_OK;
  }
  if (path == FACES_PATH_FUSED) {
    // An open rank whose six face neighbours all exist, stood in for by self
    // halos, sums exactly the periodic wrap partners (same tree, same bits).
    if (p->boundary == FACES_BOUNDARY_PERIODIC || p->mask == 63) {
      // One launch: every node with its wrap partners, in place.
      if (compute) return launch_box(p, u, false, s);
      return launch_surface(p, u, bound_wrap(), s);
    }
    // Other open ranks: one streaming launch in which each rank-surface node
    // sums the copies on its side, and those across the surface where the
    // rank has a neighbour on that side; without the streaming kernel, a
    // rank without neighbours sums interior nodes, then the rank surface.
    if (open_stream) return launch_stream(p, u, 2, s);
    if (p->mask != 0)
      return fail(FACES_ERR_ARG, "fused path of a rank with neighbours needs the streaming kernel "
                                 "and compute = 1: use the staged path");
    if (compute) {
      st = launch_box(p, u, true, s);
      if (st) return st;
    }
    return launch_surface(p, u, bound_recv(0), s);
  }
  // Staged: pack -> (exchange) -> inner -> unpack.  Inner and unpack touch
  // disjoint node sets, and pack reads only rank-surface copies, which inner
  // never writes (gpu/v-26-messages/Faces.cc:297-380 runs inner concurrently
  // with pack for the same reason).  On the one-item kernel, inner and unpack
  // are one launch: its surface blocks run beside its item blocks.
  st = launch_pack(p, u, path == FACES_PATH_STAGED_DIRECT, s);
  if (st) return st;
  if (path == FACES_PATH_STAGED) {
    st = launch_exchange(p, s);
    if (st) return st;
  }
  if (compute && interior_takes_surface(p)) return launch_inner_surface(p, u, bound_recv(p->mask), s);
  if (compute) {
    st = launch_box(p, u, true, s);
    if (st) return st;
  }
  return launch_surface(p, u, bound_recv(p->mask), s);
}

}  // extern "C"

namespace faces {

// The staged step -- pack straight into the receive slot of the opposite
// direction, interior sums, surface sums -- on a message arena the caller
// provides, with the plan kept on the host stack: nothing is allocated.
// Only receive slots are laid out (512-byte aligned); every slot the surface
// sums read is written by the same call's pack (a direction's slot is read
// only when its neighbour exists), so the arena needs no clearing.
int share_staged_on_arena(double *u, int rank, int lx, int ly, int lz, int mx, int my, int mz, int n,
                          int boundary, long plane_stride, int compute, double *arena, size_t arena_bytes,
                          const char *who, hipStream_t s)
{
  if (!u) return fail(FACES_ERR_ARG, "null field pointer");
  faces_plan p;
  int st = plan_init(&p, rank, lx, ly, lz, mx, my, mz, n, boundary, plane_stride);
  if (st) return st;
  long off = 0;
  for (int k = 0; k < 26; k++) {
    p.recv.off[k] = off;
    off += (p.slot_cnt[k] + 63) / 64 * 64;
  }
  p.recv.off[26] = off;
  p.send = p.recv;  // unused by the direct pack
  p.arena_doubles = off;
  if ((size_t)off * sizeof(double) > arena_bytes)
    return fail(FACES_ERR_ARG, "messages of %ld bytes exceed the %zu-byte arena of %s",
                off * (long)sizeof(double), arena_bytes, who);
  p.arena = arena;
  st = launch_pack(&p, u, true, s);
  if (st) return st;
  if (compute && interior_takes_surface(&p)) return launch_inner_surface(&p, u, bound_recv(p.mask), s);
  if (compute) {
    st = launch_box(&p, u, true, s);
    if (st) return st;
  }
  return launch_surface(&p, u, bound_recv(p.mask), s);
}

}  // namespace faces

extern "C" {

// In-place read-modify-write stream, one double2 per lane, one launch over
// the buffer: the bandwidth ceiling of an in-place pass on this device.
// `scale` is 1.0 at run time, so values stay as they are.
static __global__ __launch_bounds__(256) void k_rmw_stream(double *__restrict__ u, long n, double scale)
{
  const long i = blockIdx.x * 256L + threadIdx.x;
  const long n2 = n / 2;
  if (i < n2) {
    double2 v = reinterpret_cast<double2 *>(u)[i];
    v.x *= scale;
    v.y *= scale;
    reinterpret_cast<double2 *>(u)[i] = v;
  } else if (i == n2 && (n & 1)) {
    u[n - 1] *= scale;
  }
}

int faces_stream_ceiling(double *u, long n, int reps, void *stream, double *gbs)
{
  if (!u || !gbs) return fail(FACES_ERR_ARG, "null argument");
  if (reinterpret_cast<uintptr_t>(u) % 16) return fail(FACES_ERR_ARG, "buffer not 16-byte aligned");
  if (n < 2) return fail(FACES_ERR_ARG, "buffer of %ld doubles", n);
  if (reps < 1 || reps > 1000) return fail(FACES_ERR_ARG, "reps %d outside [1,1000]", reps);
  const long blocks = (n / 2 + 1 + 255) / 256;
  if (blocks > 0x7fffffffL) return fail(FACES_ERR_ARG, "buffer too large");
  hipStream_t st = (hipStream_t)stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    return fail(FACES_ERR_HIP, "hipEventCreate");
  }
  int rc = FACES_OK;
  float ms = 0.f;
  for (int r = 0; r < 2 + reps && !rc; r++) {
    if (r == 2 && hipEventRecord(e0, st) != hipSuccess) rc = fail(FACES_ERR_HIP, "hipEventRecord");
    if (!rc) {
      k_rmw_stream<<<dim3((unsigned)blocks), dim3(256), 0, st>>>(u, n, 1.0);
      if (hipGetLastError() != hipSuccess) rc = fail(FACES_ERR_HIP, "stream kernel launch");
    }
  }
  if (!rc && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
              hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
    rc = fail(FACES_ERR_HIP, "timing events");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  *gbs = ms > 0.f ? 16.0 * (double)n * reps / (ms * 1e-3) / 1e9 : 0.0;
  return FACES_OK;
}

int faces_fill(double *u, int n, int mx, int my, int mz, long plane_stride, int kind, uint64_t seed,
               void *stream)
{
  if (!u) return fail(FACES_ERR_ARG, "null field pointer");
  if (n < 1 || mx < 1 || my < 1 || mz < 1) return fail(FACES_ERR_ARG, "bad shape");
  if (kind < 0 || kind > 2) return fail(FACES_ERR_ARG, "unknown fill kind %d", kind);
  Geom G;
  G.u = u;
  G.n = n;
  G.mx = mx;
  G.my = my;
  G.mz = mz;
  G.ys = (long)n * n * n * mx;
  const long dense = G.ys * my;
  if (plane_stride != 0 && plane_stride < dense) return fail(FACES_ERR_ARG, "bad plane stride");
  G.ps = plane_stride > 0 ? plane_stride : dense;
  const long total = dense * mz;
  const int blk = 256;
  const int grid = grid_for(total, blk);
  k_fill<<<grid, blk, 0, (hipStream_t)stream>>>(G, kind, seed);
  FACES_HIP(hipGetLastError());
  return FACES_OK;
}

}  // extern "C"
