/*
 * faces_amd.h -- C ABI of the MI355X (gfx950) Faces exchange + DSS library.
 *
 * The library replaces the step API of trey-ornl/faces, class Faces
 * (gpu/v-26-messages/Faces.h:6-39, Faces.cc:10-678), with plain C entry points:
 * device pointers, sizes and an opaque HIP stream handle (void *; NULL is the
 * null stream).  No entry point exits the process: each returns FACES_OK or a
 * negative status, and faces_strerror / faces_last_error give the text.
 *
 * Field layout (gpu/DArray.h:36-38): u(ix,iy,iz,jx,jy,jz) at
 *   ix + n*(iy + n*(iz + n*(jx + mx*jy))) + jz*plane_stride
 * with plane_stride = n^3*mx*my for a dense field (pass 0), or the
 * reference's page-padded stride (gpu/DArray.h:95-99).
 *
 * Semantics: every copy of a shared node becomes the sum of all its copies,
 * grouped x, then y, then z (the order of gpu/Mugs.cc:136-203).
 *   FACES_BOUNDARY_OPEN      neighbours outside the rank grid do not exist, as
 *                            in the reference's neighbor() (Faces.cc:92-103);
 *                            existing neighbours are stood in for by this
 *                            rank's own opposite boundary (self halo).
 *   FACES_BOUNDARY_PERIODIC  the rank grid wraps; on one device every
 *                            neighbour is stood in for by the self halo, which
 *                            makes the field periodic.
 */
#ifndef FACES_AMD_H
#define FACES_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
#define FACES_OK 0
#define FACES_ERR_ARG (-1)       /* invalid argument */
#define FACES_ERR_HIP (-2)       /* HIP runtime error */
#define FACES_ERR_NOMEM (-3)     /* device or host allocation failed */
#define FACES_ERR_STATE (-4)     /* call not valid in the object's state */
#define FACES_ERR_ADMISSION (-5) /* request exceeds the server's footprint cap */
#define FACES_ERR_NODEV (-6)     /* no HIP device visible */
#define FACES_ERR_TIMEOUT (-7)   /* wait ended before the work completed */

#define FACES_BOUNDARY_OPEN 0
#define FACES_BOUNDARY_PERIODIC 1

/* Execution paths of faces_share. */
#define FACES_PATH_AUTO 0          /* fused when possible, else staged */
#define FACES_PATH_FUSED 1         /* gather-sum-scatter in place, one launch; for
                                      an open rank with neighbours (self halos)
                                      it needs the streaming kernel and
                                      compute = 1 */
#define FACES_PATH_STAGED 2        /* pack -> exchange copy -> inner + unpack */
#define FACES_PATH_STAGED_DIRECT 3 /* pack straight into receive slots */
#define FACES_PATH_SIX 4           /* the 6-message schedule: z, y, x face phases
                                      (gpu/v-6-messages-basic/Faces.cc:65-172);
                                      its summation tree is z-first, so random
                                      fields differ from the other paths by
                                      rounding.  compute = 1 takes one launch
                                      (the one-item kernel, tree z, y, x) where
                                      FACES_OPT_FLOW is not 0 and the field is
                                      below 4 GB; else three phase launches */

/* Field generators (faces_fill). */
#define FACES_FILL_INT 0  /* integer-valued, uniform in [0,1023] */
#define FACES_FILL_RAND 1 /* uniform in [-1,1) */
#define FACES_FILL_REF 2  /* reference init, gpu/main.cc:10-19 */

/* Text of a status code. */
const char *faces_strerror(int status);
/* Detail of the last failure on the calling thread ("" if none). */
const char *faces_last_error(void);
/* Library version string. */
const char *faces_version(void);

/* ---- plain step: replaces class Faces ------------------------------------ */
typedef struct faces_plan faces_plan;

/* Faces::Faces(rank,lx,ly,lz,mx,my,mz,n) -- gpu/v-26-messages/Faces.cc:10-82.
 * Builds the neighbour table and message sizes, allocates the 26 send and 26
 * receive slots once, and picks the launch shapes.  plane_stride 0 = dense. */
int faces_plan_create(faces_plan **plan, int rank, int lx, int ly, int lz, int mx, int my, int mz,
                      int n, int boundary, long plane_stride);

/* Faces::~Faces -- gpu/v-26-messages/Faces.cc:84-90. */
int faces_plan_destroy(faces_plan *plan);

/* Faces::share(u, compute) -- gpu/v-26-messages/Faces.cc:111-678.
 * Stream-ordered: returns once the work is enqueued on `stream` (the
 * reference returns after synchronising, Faces.cc:675-677; call
 * hipStreamSynchronize for that).  compute = 0 skips the rank-interior sums
 * and sums only nodes on the rank surface (Faces.cc:299, gpu/main.cc:157-161).
 * Allocates nothing; may be captured in a HIP graph. */
int faces_share(faces_plan *plan, double *u, int compute, void *stream);

/* The staged step (faces_share on FACES_PATH_STAGED_DIRECT: pack into the
 * receive slots, interior sums, surface sums) of rank `rank` of an
 * lx*ly*lz rank grid, without a plan and without allocating: its receive
 * messages live in a device array of FACES_NOALLOC_SCRATCH_BYTES that the
 * library holds on each GPU it runs on until the process ends.  No
 * counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The rank's 26 receive messages
 * (Faces.cc:44-47,64-66, each padded to 512 bytes) must fit the array
 * (16 MiB holds a 60^3 n=7 rank's 8.5 MB), else FACES_ERR_ARG.  Every call
 * on one GPU shares that GPU's array and keeps the last call's contents:
 * calls that overlap on one GPU (other streams, other threads) race.
 * Stream-ordered like faces_share. */
#define FACES_NOALLOC_SCRATCH_BYTES (16L * 1024 * 1024)
int faces_share_noalloc(double *u, int rank, int lx, int ly, int lz, int mx, int my, int mz, int n,
                        int boundary, long plane_stride, int compute, void *stream);

/* Selects the execution path (FACES_PATH_*). */
int faces_plan_set_path(faces_plan *plan, int path);

/* Launch shape of the main sum kernel: threads per block (64..1024, multiple
 * of 64), elements per row segment (0 = whole row), XCD-contiguous work order
 * (0/1).  No counterpart in the reference (gpu/gpu.h:73-78 fixes MAXTHREADS
 * at build time). */
int faces_plan_set_tuning(faces_plan *plan, int block, int seg_len, int xcd_remap);

/* Kernel of the in-place sum on periodic fields: FACES_VARIANT_DIRECT (one
 * lane per node, copies read straight from HBM), FACES_VARIANT_LDS (element
 * segments staged through LDS), FACES_VARIANT_STREAM (streaming
 * kernel with item-invariant offset tables), FACES_VARIANT_ITEMS (the
 * batched step's work-list kernel on this one grid: element segments of up
 * to 32 elements, the row's last one shorter, so any mx works); AUTO picks
 * ITEMS when its items are more than twice as long as STREAM's (mx without
 * a suitable divisor), else STREAM.
 * block: threads per block of the LDS kernel (64..512, multiple of 64),
 * persistent blocks per CU of the STREAM kernel (1..32, or -1 for one work
 * item per block; FACES_OPT_STREAM_BPC also takes the capped forms) or
 * blocks per CU of the ITEMS kernel (1..256, each block one contiguous chunk
 * of the work list); 0 keeps the current value. */
#define FACES_VARIANT_AUTO 0
#define FACES_VARIANT_DIRECT 1
#define FACES_VARIANT_LDS 2
#define FACES_VARIANT_STREAM 3
#define FACES_VARIANT_ITEMS 4
int faces_plan_set_variant(faces_plan *plan, int variant, int block);

/* Launch options of the in-place sum (no counterpart in the reference,
 * whose launch shape is fixed at build time, gpu/gpu.h:73-78):
 *   FACES_OPT_VARIANT        FACES_VARIANT_*
 *   FACES_OPT_DIRECT_BLOCK   threads per block of the DIRECT kernel
 *   FACES_OPT_SEG            elements per work item (0 = default)
 *   FACES_OPT_XCD            XCD work order: 2 (default) = each XCD takes an
 *                            eighth of every plane's items (the one-item
 *                            kernel; the others read it as 1), 1 =
 *                            contiguous eighths of the work list, 0 = one
 *                            global order
 *   FACES_OPT_LDS_BLOCK      threads per block of the LDS kernel
 *   FACES_OPT_STREAM_BPC     STREAM kernel: persistent blocks per CU (1-32);
 *                            0 = one work unit per block; -2, -3, -4 = one
 *                            work unit per block, at most 2, 3, 4 blocks
 *                            resident per CU (each block reserves that share
 *                            of the CU's LDS)
 *   FACES_OPT_STREAM_WALK    STREAM kernel traversal: 0 = independent items,
 *                            1 = a block walks element rows in x carrying the
 *                            seam column (default for n <= 6), 2 = a block
 *                            walks pieces of columns of rows in y carrying the
 *                            y-partner rows (default for n >= 7), 3 = a block
 *                            walks a group of rows segment by segment,
 *                            carrying both
 *   FACES_OPT_ACTIVE_KERNEL  (get only) the variant faces_share launches
 *   FACES_OPT_STREAM_ROWS    STREAM walk 3: rows per group; walk 2: rows per
 *                            piece of a column (the largest divisor of my not
 *                            above the value is used); 0 (default) = 4 rows
 *                            per group, and column pieces of up to 20 rows,
 *                            shorter while there are fewer than 4 per block
 *   FACES_OPT_STREAM_PLANES  STREAM kernel addressing: 1 = through one
 *                            descriptor per jz plane (what fields of 4 GB and
 *                            more use; set, it applies to any field); get:
 *                            whether the plan addresses per plane
 *   FACES_OPT_FLOW           periodic steps with compute = 1, open ranks and
 *                            the staged step's interior sums (fields below
 *                            4 GB): the one-item kernel, one work item per
 *                            block in memory order, with load widths 2 = OWN
 *                            runs loaded 16 bytes per lane (8 when the field
 *                            is not 16-byte aligned; default for n = 10,
 *                            12); 3 = loading and storing exactly the
 *                            locations each item owns, in 16-byte pairs that
 *                            may start on any double; 4 = the same for the
 *                            partner regions too (default for n = 3-9 and
 *                            11; the staged step's surface
 *                            sums done by its items, FACES_OPT_SURFACE_BLOCKS
 *                            3, take 3); 1 = 8-byte lanes (default for n =
 *                            2); 0 = the STREAM
 *                            walks above
 *   FACES_OPT_SURFACE_BLOCKS the staged step's surface sums where its interior
 *                            sums take the one-item kernel: 2 (default) =
 *                            extra blocks of that launch placed after its
 *                            items, 1 = before them, 3 = summed by that
 *                            launch's items from their local copies and the
 *                            received messages, 0 = a launch of their own
 *                            after it */
#define FACES_OPT_VARIANT 0
#define FACES_OPT_DIRECT_BLOCK 1
#define FACES_OPT_SEG 2
#define FACES_OPT_XCD 3
#define FACES_OPT_LDS_BLOCK 4
#define FACES_OPT_STREAM_BPC 5
#define FACES_OPT_STREAM_WALK 6
#define FACES_OPT_ACTIVE_KERNEL 7
#define FACES_OPT_STREAM_ROWS 8
#define FACES_OPT_STREAM_PLANES 9
#define FACES_OPT_FLOW 10
#define FACES_OPT_SURFACE_BLOCKS 11
int faces_plan_set_option(faces_plan *plan, int option, int value);
int faces_plan_get_option(const faces_plan *plan, int option, int *value);

/* Autotuner of the sum kernel (no counterpart in the reference, whose launch
 * shape is fixed at build time, gpu/gpu.h:73-78): times the streaming kernel
 * over walks, segment lengths and blocks per CU, the one-item kernel over
 * item lengths, load widths and XCD orders, and the work-list kernel over
 * blocks per CU, on this plan's grid, and keeps the fastest.  Every trial
 * shares `u` 1 + reps times (compute = 1): the field changes as after that
 * many shares, so tune on a scratch copy.  Synchronises `stream`.  out (may
 * be NULL), 5 doubles: variant, walk (the one-item kernel: -FACES_OPT_FLOW),
 * elements per item, blocks per CU (as FACES_OPT_STREAM_BPC: negative = one
 * work unit per block, at most that many resident per CU; the one-item
 * kernel: its FACES_OPT_XCD order), µs per share of the choice. */
int faces_plan_tune(faces_plan *plan, double *u, int reps, void *stream, double *out);

/* Neighbour rank of message slot 0..25, or -1 for none (MPI_PROC_NULL) --
 * Faces::neighbor, gpu/v-26-messages/Faces.cc:92-103. */
int faces_plan_neighbor(const faces_plan *plan, int slot, int *rank);

/* Device pointer and length (doubles) of message slot `slot` of the send
 * (recv = 0) or receive (recv = 1) side -- the buffers the reference posts to
 * MPI_Isend / MPI_Irecv (Faces.cc:144-171, 386-413). */
int faces_plan_message(faces_plan *plan, int slot, int recv, double **dptr, long *count);

/* The staged step in pieces, for an external transport (e.g. one process per
 * GPU).  pack: Faces.cc:173-295; exchange: self-halo copy recv[d] <-
 * send[opposite(d)] for every existing neighbour (stand-in for Faces.cc:
 * 144-171, 386-417); inner: Faces.cc:297-380; unpack: Faces.cc:419-671. */
int faces_pack(faces_plan *plan, const double *u, void *stream);
int faces_exchange(faces_plan *plan, void *stream);
int faces_inner(faces_plan *plan, double *u, void *stream);
int faces_unpack(faces_plan *plan, double *u, void *stream);
/* inner and unpack together once the messages have arrived (Faces.cc:297-380
 * and 419-671, whose kernels the reference runs in one step): one launch of
 * the one-item kernel where the plan's field takes it (FACES_OPT_SURFACE_BLOCKS
 * picks how), else faces_inner then faces_unpack. */
int faces_inner_unpack(faces_plan *plan, double *u, void *stream);

/* ---- field helpers ------------------------------------------------------- */

/* Fills a field with one of the FACES_FILL_* generators (gpu/main.cc:10-19 for
 * FACES_FILL_REF; splitmix64 of seed ^ index*golden for the others). */
int faces_fill(double *u, int n, int mx, int my, int mz, long plane_stride, int kind, uint64_t seed,
               void *stream);

/* Doubles spanned by a field (including plane padding). */
long faces_field_doubles(int n, int mx, int my, int mz, long plane_stride);

/* Measurement helper, no counterpart in the reference: the in-place
 * read-plus-write bandwidth (GB/s, 16 B per double per pass) of a plain
 * streaming kernel over n doubles at u (16-byte aligned), timed over reps
 * passes after two untimed ones -- the ceiling an in-place step can reach on
 * this device.  Values are unchanged (each is multiplied by 1.0; a
 * signalling NaN becomes quiet).  Synchronises `stream`. */
int faces_stream_ceiling(double *u, long n, int reps, void *stream, double *gbs);

/* ---- batched step: many independent grids, one launch -------------------- */
/* No counterpart in the reference, which steps one grid per process
 * (gpu/main.cc:170-186). */
typedef struct faces_batch faces_batch;

/* Describes `ngrids` grids (dense fields).  boundary as above; every grid is
 * one rank 1x1x1. */
int faces_batch_create(faces_batch **batch, int ngrids, const int *n, const int *mx, const int *my,
                       const int *mz, int boundary);
int faces_batch_destroy(faces_batch *batch);

/* Binds the device field of every grid (host array of `ngrids` device
 * pointers).  Rebuilds the captured graph on the next share. */
int faces_batch_set_fields(faces_batch *batch, double *const *u);

/* One step over all grids: one persistent work-list launch, replayed from a
 * captured HIP graph. */
int faces_batch_share(faces_batch *batch, int compute, void *stream);

/* Same step, one launch per grid (the baseline the batched form is measured
 * against). */
int faces_batch_share_per_grid(faces_batch *batch, int compute, void *stream);

/* Launch shape: threads per block (0 or 256), longest work item in elements
 * (1-64), blocks per CU (1-256: each block takes one contiguous chunk of the
 * work list of equal estimated cost; -1: one block per work item, each XCD
 * taking a contiguous share of the list of equal estimated cost, its items
 * dispatched in order); 0 keeps the current value.  The grids are dealt
 * into one bin per XCD by item cost, so every share holds the same mix of
 * item sizes. */
int faces_batch_set_tuning(faces_batch *batch, int block, int seg_len, int blocks_per_cu);

/* Times every candidate launch shape on the bound fields (reps replays each,
 * fields are modified) and keeps the fastest.  out[0..3] = block, seg_len,
 * blocks_per_cu, microseconds per step.  Not graph-capturable. */
int faces_batch_tune(faces_batch *batch, int reps, void *stream, double *out);

/* Total bytes of the bound fields and number of work items. */
int faces_batch_info(const faces_batch *batch, double *bytes, long *items);

/* ---- served step: continuous batching of requests ------------------------ */
/* No counterpart in the reference, which steps one grid per process
 * (gpu/main.cc:170-186). */
typedef struct faces_server faces_server;

/* footprint_cap: bytes of request fields admitted at once; max_batch: most
 * requests per launch; nstreams: launches that may overlap on the device. */
int faces_server_create(faces_server **srv, double footprint_cap, int max_batch, int nstreams);
int faces_server_destroy(faces_server *srv);

/* Enqueues a request on the dense field u (device, resident until done).
 * Returns FACES_ERR_ADMISSION if the field alone exceeds the cap. */
int faces_server_submit(faces_server *srv, int n, int mx, int my, int mz, int boundary, int compute,
                        double *u, uint64_t *ticket);

/* Waits for a request and releases its ticket: a ticket is waited for once
 * (again, or never issued: FACES_ERR_ARG); FACES_ERR_TIMEOUT after timeout_s
 * seconds (<= 0: no limit) leaves it pending.  latency_us = completion -
 * submission.  Submit and wait may be called from any threads. */
int faces_server_wait(faces_server *srv, uint64_t ticket, double timeout_s, double *latency_us);

/* Replays an arrival trace: request i is submitted at arrival_us[i] after the
 * start; fills latency_us[i] (arrival -> completion).  stats[0..5] = wall
 * seconds, requests per second, p50 us, p99 us, max us, launches. */
int faces_server_run_trace(faces_server *srv, int nreq, const double *arrival_us, const int *n,
                           const int *mx, const int *my, const int *mz, int boundary, int compute,
                           double *const *u, double *latency_us, double *stats);

/* stats[0..3] = admitted, completed, launches, peak admitted bytes. */
int faces_server_stats(const faces_server *srv, double *stats);

/* Launch shape of the server's batched launches: blocks per CU (1-256, a
 * chunk of the launch's work list each), or -1 (default) for one block per
 * work item.  Applies to launches made after the call. */
int faces_server_set_tuning(faces_server *srv, int blocks_per_cu);

#ifdef __cplusplus
}
#endif

#endif /* FACES_AMD_H */
