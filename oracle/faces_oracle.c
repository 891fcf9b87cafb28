/*
 * faces_oracle.c -- CPU oracle for the Faces exchange + DSS ("share") step.
 *
 * THIS FILE IS TEST INFRASTRUCTURE.  It is the reference result the GPU
 * library (libfaces_amd.so) is compared against, and the timed CPU baseline of the
 * benchmark.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  No product code path calls into it.
 *
 * It restates, in plain C + OpenMP, the reference algorithm of trey-ornl/faces
 * (paths relative to the reference root):
 *
 *   field layout     u(ix,iy,iz,jx,jy,jz), ix fastest          gpu/Array.h:24-26
 *                    optional padded jz-plane stride            gpu/DArray.h:95-99
 *   reference init   jz*1e4+jy*100+jx+iz*0.01+iy*1e-4+ix*1e-6   gpu/main.cc:21-37
 *   sweep DSS        x-pass, y-pass, z-pass of pairwise sums    gpu/Mugs.cc:136-203
 *                    compute=false predicate                   gpu/Mugs.cc:146-150,
 *                                                               167-171,188-192
 *   26 messages      pack maps                                  gpu/v-26-messages/Faces.cc:243-293
 *                    neighbour directions and sizes             gpu/v-26-messages/Faces.cc:37-75
 *                    unpack-and-sum with received copies        gpu/v-26-messages/Faces.cc:554-669
 *   multi-rank Mugs  pack after local sums, add received        gpu/Mugs.cc:205-412
 *   compare          |u-v|/(|u|+|v|) < 1e-15                    gpu/main.cc:39-105
 *
 * Boundary semantics.  The reference's neighbor() returns MPI_PROC_NULL outside
 * the rank grid (gpu/v-26-messages/Faces.cc:92-103), so a 1x1x1 run exchanges
 * nothing ("open").  The periodic form adds, in each pass, the wrap pair that
 * joins element 0's low face with element m-1's high face.  Every pass pairs
 * disjoint copies, so after x, y, z passes each shared node holds the tree
 * ((c000+c100)+(c010+c110)) + ((c001+c101)+(c011+c111)) of its copies, grouped
 * x first, then y, then z.  That tree is the canonical summation order the GPU
 * kernels evaluate; floating-point addition is commutative, so only the
 * grouping fixes the bits.
 *
 * Self-halo emulation.  For a rank with neighbours, the reference receives the
 * neighbour's packed boundary values.  On one device the neighbour is stood in
 * for by this rank's own opposite boundary: recv[d] <- send[opposite(d)]
 * (faces s <-> 1-s, edges e <-> 3-e, corners c <-> 7-c).  nbr_mask bit k says a
 * neighbour exists across face k = x-,x+,y-,y+,z-,z+; an edge or corner
 * neighbour exists iff all its face directions do (gpu/Mugs.cc:72-83).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* layout                                                                    */
/* ------------------------------------------------------------------------ */

typedef struct {
  int n, mx, my, mz;
  long ps; /* jz-plane stride in doubles (>= n^3*mx*my) */
} grid_t;

static grid_t mkgrid(int n, int mx, int my, int mz, long ps)
{
  grid_t g;
  g.n = n; g.mx = mx; g.my = my; g.mz = mz;
  const long dense = (long)n * n * n * mx * my;
  g.ps = (ps > 0) ? ps : dense;
  return g;
}

/* gpu/DArray.h:36-38 with the padded outer stride of gpu/DArray.h:95-99 */
static inline long IX(const grid_t *g, int ix, int iy, int iz, int jx, int jy, int jz)
{
  const long n = g->n;
  return ix + n * (iy + n * (iz + n * (jx + (long)g->mx * jy))) + (long)jz * g->ps;
}

/* Dense linear index, used by the seeded generators (padding slots skipped). */
static inline long DIX(const grid_t *g, int ix, int iy, int iz, int jx, int jy, int jz)
{
  const long n = g->n;
  return ix + n * (iy + n * (iz + n * (jx + (long)g->mx * (jy + (long)g->my * jz))));
}

long orc_field_doubles(int n, int mx, int my, int mz, long ps)
{
  grid_t g = mkgrid(n, mx, my, mz, ps);
  return g.ps * (long)(mz - 1) + (long)n * n * n * mx * my;
}

/* ------------------------------------------------------------------------ */
/* seeded generators (not in the reference, which fills by position only)   */
/* ------------------------------------------------------------------------ */

/* splitmix64 finaliser of (seed xor idx*golden) -- Steele, Lea & Flood 2014 */
uint64_t orc_hash(uint64_t seed, uint64_t idx)
{
  uint64_t z = seed ^ (idx * 0x9E3779B97F4A7C15ull);
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* kind 0: integer-valued in [0,1023]; kind 1: uniform in [-1,1) on a 2^-52
 * grid; kind 2: the reference's positional init (gpu/main.cc:21-37). */
void orc_fill(double *u, int n, int mx, int my, int mz, long ps, int kind, uint64_t seed)
{
  grid_t g = mkgrid(n, mx, my, mz, ps);
#pragma omp parallel for collapse(3) schedule(static)
  for (int jz = 0; jz < mz; jz++)
    for (int jy = 0; jy < my; jy++)
      for (int jx = 0; jx < mx; jx++)
        for (int iz = 0; iz < n; iz++)
          for (int iy = 0; iy < n; iy++)
            for (int ix = 0; ix < n; ix++) {
              double v;
              if (kind == 2) {
                v = jz * 10000 + jy * 100 + jx + iz * 0.01 + iy * 0.0001 + ix * 0.000001;
              } else {
                const uint64_t h = orc_hash(seed, (uint64_t)DIX(&g, ix, iy, iz, jx, jy, jz));
                if (kind == 0) v = (double)(h >> 54);
                else v = 2.0 * ((double)(h >> 11) * 0x1.0p-53) - 1.0;
              }
              u[IX(&g, ix, iy, iz, jx, jy, jz)] = v;
            }
}

/* ------------------------------------------------------------------------ */
/* sweep DSS: gpu/Mugs.cc:136-203, plus the periodic wrap pair of each pass   */
/* ------------------------------------------------------------------------ */

static inline void pair_sum(double *u, long ia, long ib)
{
  /* gpu/Mugs.cc:153-154 */
  u[ia] += u[ib];
  u[ib] = u[ia];
}

void orc_share_sweep(double *u, int n, int mx, int my, int mz, long ps, int periodic, int compute)
{
  grid_t g = mkgrid(n, mx, my, mz, ps);
  const int nm1 = n - 1, mxm1 = mx - 1, mym1 = my - 1, mzm1 = mz - 1;
  /* In the periodic form the first element's low face pairs with the last
   * element's high face; jx == 0 stands for that pair (partner mx-1). */
  const int jx0 = periodic ? 0 : 1, jy0 = periodic ? 0 : 1, jz0 = periodic ? 0 : 1;
#pragma omp parallel
  {
    /* x-pass: gpu/Mugs.cc:140-160 */
#pragma omp for collapse(3) schedule(static)
    for (int jz = 0; jz < mz; jz++)
      for (int jy = 0; jy < my; jy++)
        for (int jx = jx0; jx < mx; jx++) {
          const int jxl = (jx == 0) ? mxm1 : jx - 1;
          for (int iz = 0; iz < n; iz++)
            for (int iy = 0; iy < n; iy++) {
              /* A wrap pair lies on the rank surface, which compute=false
               * still sums (it is the exchanged part of the step). */
              if (compute || jx == 0 || (jz == 0 && iz == 0) || (jz == mzm1 && iz == nm1) ||
                  (jy == 0 && iy == 0) || (jy == mym1 && iy == nm1))
                pair_sum(u, IX(&g, 0, iy, iz, jx, jy, jz), IX(&g, nm1, iy, iz, jxl, jy, jz));
            }
        }
    /* y-pass: gpu/Mugs.cc:161-181 */
#pragma omp for collapse(3) schedule(static)
    for (int jz = 0; jz < mz; jz++)
      for (int jy = jy0; jy < my; jy++)
        for (int jx = 0; jx < mx; jx++) {
          const int jyl = (jy == 0) ? mym1 : jy - 1;
          for (int iz = 0; iz < n; iz++)
            for (int ix = 0; ix < n; ix++) {
              if (compute || jy == 0 || (jz == 0 && iz == 0) || (jz == mzm1 && iz == nm1) ||
                  (jx == 0 && ix == 0) || (jx == mxm1 && ix == nm1))
                pair_sum(u, IX(&g, ix, 0, iz, jx, jy, jz), IX(&g, ix, nm1, iz, jx, jyl, jz));
            }
        }
    /* z-pass: gpu/Mugs.cc:182-202 */
#pragma omp for collapse(3) schedule(static)
    for (int jz = jz0; jz < mz; jz++)
      for (int jy = 0; jy < my; jy++)
        for (int jx = 0; jx < mx; jx++) {
          const int jzl = (jz == 0) ? mzm1 : jz - 1;
          for (int iy = 0; iy < n; iy++)
            for (int ix = 0; ix < n; ix++) {
              if (compute || jz == 0 || (jy == 0 && iy == 0) || (jy == mym1 && iy == nm1) ||
                  (jx == 0 && ix == 0) || (jx == mxm1 && ix == nm1))
                pair_sum(u, IX(&g, ix, iy, 0, jx, jy, jz), IX(&g, ix, iy, nm1, jx, jy, jzl));
            }
        }
  }
}

/* ------------------------------------------------------------------------ */
/* the 6-message schedule: gpu/v-6-messages-basic/Faces.cc:65-172             */
/* ------------------------------------------------------------------------ */
/*
 * Three phases, z then y then x.  A phase sums every face pair across an
 * element boundary along its axis (Faces.cc:112-119 for z): the high face of
 * element j-1 with the low face of element j, both copies set to the sum.  At
 * the rank boundary the low face of element 0 adds the face received from
 * the low neighbour, and the high face of element m-1 the one from the high
 * neighbour (Faces.cc:114,119); with self halos (nbr_mask bit 2a = low,
 * 2a+1 = high neighbour exists along axis a) those are this field's own
 * opposite faces as packed before the phase.  Faces carry the edges and
 * corners summed by the earlier phases forward, so a node's tree is
 * ((z pairs) then y pairs) then x pairs.  compute = 0 sums only pairs on the
 * rank surface (the reference's 6-message variants ignore the flag,
 * Faces.cc:65; the 26-message predicate, gpu/Mugs.cc:146-150, is kept here).
 */
static void six_phase(double *u, const grid_t *g, int axis, int lo, int hi, int compute)
{
  const int n = g->n, nm1 = n - 1;
  const int m[3] = {g->mx, g->my, g->mz};
  const int ma = m[axis];
  const int b1 = (axis == 0) ? 1 : 0, b2 = (axis == 2) ? 1 : 2; /* the other two axes */
#pragma omp parallel for collapse(3) schedule(static)
  for (int q2 = 0; q2 < m[b2]; q2++)
    for (int q1 = 0; q1 < m[b1]; q1++)
      for (int j = 0; j < ma; j++) {
        for (int i2 = 0; i2 < n; i2++)
          for (int i1 = 0; i1 < n; i1++) {
            int jl[3], jh[3], il[3], ih[3];
            jl[b1] = jh[b1] = q1; jl[b2] = jh[b2] = q2;
            il[b1] = ih[b1] = i1; il[b2] = ih[b2] = i2;
            jl[axis] = j; il[axis] = 0;                          /* low face of element j */
            jh[axis] = (j == 0) ? ma - 1 : j - 1; ih[axis] = nm1; /* high face below it */
            const long ia = IX(g, il[0], il[1], il[2], jl[0], jl[1], jl[2]);
            const long ib = IX(g, ih[0], ih[1], ih[2], jh[0], jh[1], jh[2]);
            if (j > 0) {
              const int surf = (q1 == 0 && i1 == 0) || (q1 == m[b1] - 1 && i1 == nm1) ||
                               (q2 == 0 && i2 == 0) || (q2 == m[b2] - 1 && i2 == nm1);
              if (compute || surf) pair_sum(u, ia, ib);
            } else {
              const double a = u[ia], b = u[ib];
              if (lo) u[ia] = a + b;
              if (hi) u[ib] = b + a;
            }
          }
      }
}

int orc_share_six(double *u, int n, int mx, int my, int mz, long ps, int nbr_mask, int compute)
{
  if (n < 2 || mx < 1 || my < 1 || mz < 1 || nbr_mask < 0 || nbr_mask > 63) return -1;
  grid_t g = mkgrid(n, mx, my, mz, ps);
  for (int axis = 2; axis >= 0; axis--)
    six_phase(u, &g, axis, (nbr_mask >> (2 * axis)) & 1, (nbr_mask >> (2 * axis + 1)) & 1, compute);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* the 26 messages (gpu/v-26-messages/Faces.cc:37-75, 243-293)               */
/* ------------------------------------------------------------------------ */
/*
 * Slot numbering follows the reference's neighbour tables:
 *   faces   0..5   iface_: x-, x+, y-, y+, z-, z+       (s = side)
 *   edges   6..9   iedge_[0..3]  x-edges, e = sy + 2*sz  (n*mx doubles)
 *          10..13  iedge_[4..7]  y-edges, e = sx + 2*sz  (n*my)
 *          14..17  iedge_[8..11] z-edges, e = sx + 2*sy  (n*mz)
 *   corners 18..25 icorner_, c = sx + 2*sy + 4*sz        (1)
 * Face message element order: x-face (iy,iz,jy,jz), y-face (ix,iz,jx,jz),
 * z-face (ix,iy,jx,jy), first index fastest, as xfs/yfs/zfs in Faces.cc.
 */
typedef struct {
  double *buf[26];
  long cnt[26];
} msgs_t;

static void msgs_alloc(msgs_t *m, const grid_t *g)
{
  const long nn = (long)g->n * g->n;
  for (int s = 0; s < 2; s++) {
    m->cnt[0 + s] = nn * g->my * g->mz;
    m->cnt[2 + s] = nn * g->mx * g->mz;
    m->cnt[4 + s] = nn * g->mx * g->my;
  }
  for (int e = 0; e < 4; e++) {
    m->cnt[6 + e] = (long)g->n * g->mx;
    m->cnt[10 + e] = (long)g->n * g->my;
    m->cnt[14 + e] = (long)g->n * g->mz;
  }
  for (int c = 0; c < 8; c++) m->cnt[18 + c] = 1;
  for (int k = 0; k < 26; k++) m->buf[k] = (double *)calloc((size_t)m->cnt[k], sizeof(double));
}

static void msgs_free(msgs_t *m)
{
  for (int k = 0; k < 26; k++) free(m->buf[k]);
}

int orc_opposite_slot(int k)
{
  if (k < 6) return (k & ~1) | (1 - (k & 1));
  if (k < 18) { const int f = (k - 6) / 4, e = (k - 6) % 4; return 6 + 4 * f + (3 - e); }
  return 18 + (7 - (k - 18));
}

/* Does the neighbour for slot k exist?  Face bits: 0 x-, 1 x+, 2 y-, 3 y+, 4 z-, 5 z+ */
static int slot_exists(int k, int mask)
{
#define HAS(b) ((mask >> (b)) & 1)
  if (k < 6) return HAS(k);
  if (k < 10) { const int e = k - 6, sy = e & 1, sz = e >> 1; return HAS(2 + sy) && HAS(4 + sz); }
  if (k < 14) { const int e = k - 10, sx = e & 1, sz = e >> 1; return HAS(0 + sx) && HAS(4 + sz); }
  if (k < 18) { const int e = k - 14, sx = e & 1, sy = e >> 1; return HAS(0 + sx) && HAS(2 + sy); }
  { const int c = k - 18, sx = c & 1, sy = (c >> 1) & 1, sz = c >> 2; return HAS(sx) && HAS(2 + sy) && HAS(4 + sz); }
#undef HAS
}

/* Pack: Faces.cc:243-293 (same contents as Mugs.cc:209-286). */
static void pack(const double *u, const grid_t *g, msgs_t *m)
{
  const int n = g->n, nm1 = n - 1, mx = g->mx, my = g->my, mz = g->mz;
  for (int s = 0; s < 2; s++) {
    const int ie = s ? nm1 : 0;
    for (int jz = 0; jz < mz; jz++)
      for (int jy = 0; jy < my; jy++)
        for (int iz = 0; iz < n; iz++)
          for (int iy = 0; iy < n; iy++)
            m->buf[0 + s][iy + n * (iz + (long)n * (jy + (long)my * jz))] =
                u[IX(g, ie, iy, iz, s ? mx - 1 : 0, jy, jz)];
    for (int jz = 0; jz < mz; jz++)
      for (int jx = 0; jx < mx; jx++)
        for (int iz = 0; iz < n; iz++)
          for (int ix = 0; ix < n; ix++)
            m->buf[2 + s][ix + n * (iz + (long)n * (jx + (long)mx * jz))] =
                u[IX(g, ix, ie, iz, jx, s ? my - 1 : 0, jz)];
    for (int jy = 0; jy < my; jy++)
      for (int jx = 0; jx < mx; jx++)
        for (int iy = 0; iy < n; iy++)
          for (int ix = 0; ix < n; ix++)
            m->buf[4 + s][ix + n * (iy + (long)n * (jx + (long)mx * jy))] =
                u[IX(g, ix, iy, ie, jx, jy, s ? mz - 1 : 0)];
  }
  for (int e = 0; e < 4; e++) {
    const int a = e & 1, b = e >> 1;
    for (int jx = 0; jx < mx; jx++)
      for (int ix = 0; ix < n; ix++)
        m->buf[6 + e][ix + (long)n * jx] =
            u[IX(g, ix, a ? nm1 : 0, b ? nm1 : 0, jx, a ? my - 1 : 0, b ? mz - 1 : 0)];
    for (int jy = 0; jy < my; jy++)
      for (int iy = 0; iy < n; iy++)
        m->buf[10 + e][iy + (long)n * jy] =
            u[IX(g, a ? nm1 : 0, iy, b ? nm1 : 0, a ? mx - 1 : 0, jy, b ? mz - 1 : 0)];
    for (int jz = 0; jz < mz; jz++)
      for (int iz = 0; iz < n; iz++)
        m->buf[14 + e][iz + (long)n * jz] =
            u[IX(g, a ? nm1 : 0, b ? nm1 : 0, iz, a ? mx - 1 : 0, b ? my - 1 : 0, jz)];
  }
  for (int c = 0; c < 8; c++) {
    const int sx = c & 1, sy = (c >> 1) & 1, sz = c >> 2;
    m->buf[18 + c][0] = u[IX(g, sx ? nm1 : 0, sy ? nm1 : 0, sz ? nm1 : 0,
                             sx ? mx - 1 : 0, sy ? my - 1 : 0, sz ? mz - 1 : 0)];
  }
}

/* recv[d] <- send[opposite(d)] for every existing neighbour; others stay 0,
 * as a receive from MPI_PROC_NULL leaves the zeroed buffer (gpu/DArray.h:81). */
static void exchange_self(const msgs_t *send, msgs_t *recv, int mask)
{
  for (int k = 0; k < 26; k++) {
    if (slot_exists(k, mask))
      memcpy(recv->buf[k], send->buf[orc_opposite_slot(k)], sizeof(double) * (size_t)recv->cnt[k]);
    else
      memset(recv->buf[k], 0, sizeof(double) * (size_t)recv->cnt[k]);
  }
}

/* ------------------------------------------------------------------------ */
/* staged form: pack raw values -> exchange -> local tree with received      */
/* copies (gpu/v-26-messages/Faces.cc:173-295, 386-417, 554-669)             */
/* ------------------------------------------------------------------------ */
/*
 * Per axis a unique rank-local node g in [0, m*(n-1)] has up to two copies:
 * interior interface nodes have two local copies (element j at index 0 and
 * element j-1 at index n-1); a rank-surface node has one local copy plus,
 * when that neighbour exists, one remote copy (side s).
 */
typedef struct {
  int cnt;
  int j[2], i[2];
  int remote[2]; /* 1 = copy lives on the neighbour at side s[] */
  int s[2];
} axis_copies_t;

static void axis_copies(int gidx, int m, int n, int has_lo, int has_hi, axis_copies_t *c)
{
  const int G = m * (n - 1) + 1;
  c->cnt = 0;
  if (gidx == 0) {
    c->j[0] = 0; c->i[0] = 0; c->remote[0] = 0; c->cnt = 1;
    if (has_lo) { c->remote[1] = 1; c->s[1] = 0; c->j[1] = 0; c->i[1] = 0; c->cnt = 2; }
  } else if (gidx == G - 1) {
    c->j[0] = m - 1; c->i[0] = n - 1; c->remote[0] = 0; c->cnt = 1;
    if (has_hi) { c->remote[1] = 1; c->s[1] = 1; c->j[1] = m - 1; c->i[1] = n - 1; c->cnt = 2; }
  } else if (gidx % (n - 1) == 0) {
    const int j = gidx / (n - 1);
    c->j[0] = j; c->i[0] = 0; c->remote[0] = 0;
    c->j[1] = j - 1; c->i[1] = n - 1; c->remote[1] = 0;
    c->cnt = 2;
  } else {
    c->j[0] = gidx / (n - 1); c->i[0] = gidx % (n - 1); c->remote[0] = 0; c->cnt = 1;
  }
}

/* Value of copy (ax,ay,az): local u, or the matching entry of the received
 * message (index maps as in gpu/v-26-messages/Faces.cc:554-669). */
static double copy_value(const double *u, const grid_t *g, const msgs_t *r,
                         const axis_copies_t *X, int ax, const axis_copies_t *Y, int ay,
                         const axis_copies_t *Z, int az)
{
  const int n = g->n, mx = g->mx, my = g->my;
  const int rx = X->remote[ax], ry = Y->remote[ay], rz = Z->remote[az];
  const int ix = X->i[ax], jx = X->j[ax], iy = Y->i[ay], jy = Y->j[ay], iz = Z->i[az], jz = Z->j[az];
  const int sx = X->s[ax], sy = Y->s[ay], sz = Z->s[az];
  if (!rx && !ry && !rz) return u[IX(g, ix, iy, iz, jx, jy, jz)];
  if (rx && !ry && !rz) return r->buf[0 + sx][iy + n * (iz + (long)n * (jy + (long)my * jz))];
  if (!rx && ry && !rz) return r->buf[2 + sy][ix + n * (iz + (long)n * (jx + (long)mx * jz))];
  if (!rx && !ry && rz) return r->buf[4 + sz][ix + n * (iy + (long)n * (jx + (long)mx * jy))];
  if (!rx && ry && rz) return r->buf[6 + sy + 2 * sz][ix + (long)n * jx];
  if (rx && !ry && rz) return r->buf[10 + sx + 2 * sz][iy + (long)n * jy];
  if (rx && ry && !rz) return r->buf[14 + sx + 2 * sy][iz + (long)n * jz];
  return r->buf[18 + sx + 2 * sy + 4 * sz][0];
}

int orc_share_staged(double *u, int n, int mx, int my, int mz, long ps, int nbr_mask, int compute)
{
  if (n < 2 || mx < 1 || my < 1 || mz < 1) return -1;
  grid_t g = mkgrid(n, mx, my, mz, ps);
  msgs_t send, recv;
  msgs_alloc(&send, &g);
  msgs_alloc(&recv, &g);
  pack(u, &g, &send);
  exchange_self(&send, &recv, nbr_mask);
  const int Gx = mx * (n - 1) + 1, Gy = my * (n - 1) + 1, Gz = mz * (n - 1) + 1;
  const int hx0 = nbr_mask & 1, hx1 = (nbr_mask >> 1) & 1, hy0 = (nbr_mask >> 2) & 1,
            hy1 = (nbr_mask >> 3) & 1, hz0 = (nbr_mask >> 4) & 1, hz1 = (nbr_mask >> 5) & 1;
  /* Every unique node writes only its own local copies, and reads only its
   * own copies, so the loop is free of cross-iteration dependences. */
#pragma omp parallel for collapse(2) schedule(static)
  for (int gz = 0; gz < Gz; gz++)
    for (int gy = 0; gy < Gy; gy++) {
      axis_copies_t Z, Y, X;
      axis_copies(gz, mz, n, hz0, hz1, &Z);
      axis_copies(gy, my, n, hy0, hy1, &Y);
      for (int gx = 0; gx < Gx; gx++) {
        axis_copies(gx, mx, n, hx0, hx1, &X);
        const int surf = gx == 0 || gx == Gx - 1 || gy == 0 || gy == Gy - 1 || gz == 0 || gz == Gz - 1;
        if (!compute && !surf) continue;
        double q[2] = {0, 0};
        for (int az = 0; az < Z.cnt; az++) {
          double p[2] = {0, 0};
          for (int ay = 0; ay < Y.cnt; ay++) {
            double v = copy_value(u, &g, &recv, &X, 0, &Y, ay, &Z, az);
            if (X.cnt == 2) v = v + copy_value(u, &g, &recv, &X, 1, &Y, ay, &Z, az);
            p[ay] = v;
          }
          q[az] = (Y.cnt == 2) ? p[0] + p[1] : p[0];
        }
        const double r = (Z.cnt == 2) ? q[0] + q[1] : q[0];
        for (int az = 0; az < Z.cnt; az++)
          for (int ay = 0; ay < Y.cnt; ay++)
            for (int ax = 0; ax < X.cnt; ax++)
              if (!X.remote[ax] && !Y.remote[ay] && !Z.remote[az])
                u[IX(&g, X.i[ax], Y.i[ay], Z.i[az], X.j[ax], Y.j[ay], Z.j[az])] = r;
      }
    }
  msgs_free(&send);
  msgs_free(&recv);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* multi-rank Mugs order (gpu/Mugs.cc:86-416) with self-halo neighbours       */
/* ------------------------------------------------------------------------ */
/*
 * Local passes over rank-internal pairs (with the compute predicate), then
 * pack of the locally summed boundary (Mugs.cc:209-286), exchange, and the
 * in-place additions of the received faces, edges and corners
 * (Mugs.cc:337-412).  Same sums as the tree; a different association, so it
 * equals the tree bit for bit only where every partial sum is exact.
 */
int orc_share_mugs26(double *u, int n, int mx, int my, int mz, long ps, int nbr_mask, int compute)
{
  if (n < 2 || mx < 1 || my < 1 || mz < 1) return -1;
  grid_t g = mkgrid(n, mx, my, mz, ps);
  const int nm1 = n - 1, mxm1 = mx - 1, mym1 = my - 1, mzm1 = mz - 1;
  orc_share_sweep(u, n, mx, my, mz, ps, 0, compute);
  if (nbr_mask == 0) return 0; /* Mugs.cc:205: no exchange without neighbours */
  msgs_t send, recv;
  msgs_alloc(&send, &g);
  msgs_alloc(&recv, &g);
  pack(u, &g, &send);
  exchange_self(&send, &recv, nbr_mask);
  /* Mugs.cc:339-373 faces */
  for (int jz = 0; jz < mz; jz++)
    for (int jy = 0; jy < my; jy++)
      for (int iz = 0; iz < n; iz++)
        for (int iy = 0; iy < n; iy++) {
          const long k = iy + n * (iz + (long)n * (jy + (long)my * jz));
          u[IX(&g, 0, iy, iz, 0, jy, jz)] += recv.buf[0][k];
          u[IX(&g, nm1, iy, iz, mxm1, jy, jz)] += recv.buf[1][k];
        }
  for (int jz = 0; jz < mz; jz++)
    for (int jx = 0; jx < mx; jx++)
      for (int iz = 0; iz < n; iz++)
        for (int ix = 0; ix < n; ix++) {
          const long k = ix + n * (iz + (long)n * (jx + (long)mx * jz));
          u[IX(&g, ix, 0, iz, jx, 0, jz)] += recv.buf[2][k];
          u[IX(&g, ix, nm1, iz, jx, mym1, jz)] += recv.buf[3][k];
        }
  for (int jy = 0; jy < my; jy++)
    for (int jx = 0; jx < mx; jx++)
      for (int iy = 0; iy < n; iy++)
        for (int ix = 0; ix < n; ix++) {
          const long k = ix + n * (iy + (long)n * (jx + (long)mx * jy));
          u[IX(&g, ix, iy, 0, jx, jy, 0)] += recv.buf[4][k];
          u[IX(&g, ix, iy, nm1, jx, jy, mzm1)] += recv.buf[5][k];
        }
  /* Mugs.cc:375-402 edges */
  for (int jx = 0; jx < mx; jx++)
    for (int ix = 0; ix < n; ix++) {
      const long k = ix + (long)n * jx;
      u[IX(&g, ix, 0, 0, jx, 0, 0)] += recv.buf[6][k];
      u[IX(&g, ix, nm1, 0, jx, mym1, 0)] += recv.buf[7][k];
      u[IX(&g, ix, 0, nm1, jx, 0, mzm1)] += recv.buf[8][k];
      u[IX(&g, ix, nm1, nm1, jx, mym1, mzm1)] += recv.buf[9][k];
    }
  for (int jy = 0; jy < my; jy++)
    for (int iy = 0; iy < n; iy++) {
      const long k = iy + (long)n * jy;
      u[IX(&g, 0, iy, 0, 0, jy, 0)] += recv.buf[10][k];
      u[IX(&g, nm1, iy, 0, mxm1, jy, 0)] += recv.buf[11][k];
      u[IX(&g, 0, iy, nm1, 0, jy, mzm1)] += recv.buf[12][k];
      u[IX(&g, nm1, iy, nm1, mxm1, jy, mzm1)] += recv.buf[13][k];
    }
  for (int jz = 0; jz < mz; jz++)
    for (int iz = 0; iz < n; iz++) {
      const long k = iz + (long)n * jz;
      u[IX(&g, 0, 0, iz, 0, 0, jz)] += recv.buf[14][k];
      u[IX(&g, nm1, 0, iz, mxm1, 0, jz)] += recv.buf[15][k];
      u[IX(&g, 0, nm1, iz, 0, mym1, jz)] += recv.buf[16][k];
      u[IX(&g, nm1, nm1, iz, mxm1, mym1, jz)] += recv.buf[17][k];
    }
  /* Mugs.cc:405-412 corners */
  for (int c = 0; c < 8; c++) {
    const int sx = c & 1, sy = (c >> 1) & 1, sz = c >> 2;
    u[IX(&g, sx ? nm1 : 0, sy ? nm1 : 0, sz ? nm1 : 0, sx ? mxm1 : 0, sy ? mym1 : 0, sz ? mzm1 : 0)] +=
        recv.buf[18 + c][0];
  }
  msgs_free(&send);
  msgs_free(&recv);
  return 0;
}

/* Exposes the packed send messages so tests can check the index maps. */
int orc_pack(const double *u, int n, int mx, int my, int mz, long ps, double *out, long *offsets)
{
  grid_t g = mkgrid(n, mx, my, mz, ps);
  msgs_t m;
  msgs_alloc(&m, &g);
  pack(u, &g, &m);
  long off = 0;
  for (int k = 0; k < 26; k++) {
    offsets[k] = off;
    if (out) memcpy(out + off, m.buf[k], sizeof(double) * (size_t)m.cnt[k]);
    off += m.cnt[k];
  }
  offsets[26] = off;
  msgs_free(&m);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* compare: gpu/main.cc:39-105, with the max taken race-free                 */
/* ------------------------------------------------------------------------ */

static inline int64_t ordered_bits(double x)
{
  int64_t b;
  memcpy(&b, &x, sizeof b);
  return (b < 0) ? (int64_t)(0x8000000000000000ull - (uint64_t)b) : b;
}

/* out[0] = max relative difference |u-v|/(|u|+|v|) (0 where |u|+|v| <= 1e-15)
 * out[1] = linear index of that maximum (dense-with-padding index)
 * out[2] = number of points whose bits differ
 * out[3] = max distance in units in the last place
 * returns 1 if the reference criterion (max rel < 1e-15) passes. */
int orc_compare(const double *u, const double *v, int n, int mx, int my, int mz, long ps, double *out)
{
  grid_t g = mkgrid(n, mx, my, mz, ps);
  double mv = 0, mu = 0;
  long ml = 0, nd = 0;
#pragma omp parallel
  {
    double lmv = 0, lmu = 0;
    long lml = 0, lnd = 0;
#pragma omp for collapse(3) schedule(static)
    for (int jz = 0; jz < mz; jz++)
      for (int jy = 0; jy < my; jy++)
        for (int jx = 0; jx < mx; jx++)
          for (int iz = 0; iz < n; iz++)
            for (int iy = 0; iy < n; iy++)
              for (int ix = 0; ix < n; ix++) {
                const long i = IX(&g, ix, iy, iz, jx, jy, jz);
                const double ue = u[i], ve = v[i];
                const double denom = fabs(ue) + fabs(ve);
                const double d = (denom > 1e-15) ? fabs(ue - ve) / denom : 0;
                if (d > lmv) { lmv = d; lml = i; }
                int64_t a = ordered_bits(ue), b = ordered_bits(ve);
                if (a != b) {
                  lnd++;
                  const uint64_t dd = (a > b) ? (uint64_t)a - (uint64_t)b : (uint64_t)b - (uint64_t)a;
                  const double du = (double)dd;
                  if (du > lmu) lmu = du;
                }
              }
#pragma omp critical
    {
      if (lmv > mv || (lmv == mv && lml < ml && lmv > 0)) { mv = lmv; ml = lml; }
      if (lmu > mu) mu = lmu;
      nd += lnd;
    }
  }
  out[0] = mv;
  out[1] = (double)ml;
  out[2] = (double)nd;
  out[3] = mu;
  return mv < 1e-15;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Threads of the following parallel regions (the CPU baseline sets one per
 * core of the host's CPU share). */
void orc_set_num_threads(int t)
{
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
#else
  (void)t;
#endif
}

/* OpenMP binding as the runtime sees it: out[0] = omp_get_proc_bind()
 * (0 false, 1 true, 2 master, 3 close, 4 spread), out[1] = places. */
void orc_omp_binding(int *out)
{
#ifdef _OPENMP
  out[0] = (int)omp_get_proc_bind();
  out[1] = omp_get_num_places();
#else
  out[0] = 0;
  out[1] = 0;
#endif
}
