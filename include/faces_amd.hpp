/* faces_amd.hpp -- C++ host interface over the C ABI (faces_amd.h), shaped
 * like the reference's host classes so that a caller written against them
 * (gpu/main.cc) reads the same:
 *   faces_amd::Field   the device field, as DArray<double,6> (gpu/DArray.h)
 *   faces_amd::Faces   the plan and its share, as class Faces
 *                      (gpu/v-26-messages/Faces.h:6-39)
 * Header-only; link libfaces_amd.so.  Failures throw faces_amd::Error (the
 * reference prints and exits, gpu/gpu.h:61-69; a C++ caller chooses here). */
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "faces_amd.h"

namespace faces_amd {

/* A failed call: status() is the FACES_ERR_* code. */
class Error : public std::runtime_error {
 public:
  Error(int status, const std::string &what) : std::runtime_error(what), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

inline void check(int status, const char *where)
{
  if (status == FACES_OK) return;
  throw Error(status, std::string(where) + ": " + faces_strerror(status) + " (" + faces_last_error() + ")");
}

inline void check_hip(hipError_t e, const char *where)
{
  if (e != hipSuccess) throw Error(FACES_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

/* mx*my*mz elements of n^3 FP64 points in device memory, index
 * ix + n(iy + n(iz + n(jx + mx*jy))) + jz*plane_stride (gpu/DArray.h:36-38;
 * plane_stride 0 = dense). */
class Field {
 public:
  Field(int n, int mx, int my, int mz, long plane_stride = 0)
      : n_(n), mx_(mx), my_(my), mz_(mz),
        ps_(plane_stride > 0 ? plane_stride : (long)n * n * n * mx * my),
        size_(faces_field_doubles(n, mx, my, mz, plane_stride))
  {
    if (size_ <= 0) throw Error(FACES_ERR_ARG, "faces_amd::Field: bad shape");
    check_hip(hipMalloc(&values_, sizeof(double) * (size_t)size_), "hipMalloc(field)");
  }
  ~Field()
  {
    if (values_) (void)hipFree(values_);
  }
  Field(const Field &) = delete;
  Field &operator=(const Field &) = delete;
  Field(Field &&o) noexcept
      : n_(o.n_), mx_(o.mx_), my_(o.my_), mz_(o.mz_), ps_(o.ps_), size_(o.size_), values_(o.values_)
  {
    o.values_ = nullptr;
  }

  long index(int ix, int iy, int iz, int jx, int jy, int jz) const
  {
    return ix + (long)n_ * (iy + (long)n_ * (iz + (long)n_ * (jx + (long)mx_ * jy))) + jz * ps_;
  }
  double *data() { return values_; }
  const double *data() const { return values_; }
  long size() const { return size_; }  // doubles spanned, padding included
  int n() const { return n_; }
  int mx() const { return mx_; }
  int my() const { return my_; }
  int mz() const { return mz_; }
  long plane_stride() const { return ps_; }

  /* The whole field to the host (gpu/DArray.h:45-50). */
  void copy(std::vector<double> &host) const
  {
    host.resize((size_t)size_);
    check_hip(hipMemcpy(host.data(), values_, sizeof(double) * (size_t)size_, hipMemcpyDeviceToHost),
              "hipMemcpy(field to host)");
  }
  /* One of the FACES_FILL_* generators (FACES_FILL_REF: gpu/main.cc:10-19). */
  void fill(int kind, uint64_t seed = 42, hipStream_t stream = nullptr)
  {
    check(faces_fill(values_, n_, mx_, my_, mz_, ps_ == (long)n_ * n_ * n_ * mx_ * my_ ? 0 : ps_, kind, seed,
                     stream),
          "faces_fill");
  }

 private:
  int n_, mx_, my_, mz_;
  long ps_, size_;
  double *values_ = nullptr;
};

/* The exchange + DSS step of one rank (gpu/v-26-messages/Faces.h:6-39). */
class Faces {
 public:
  using Double6D = Field;

  /* Faces.cc:10-82.  boundary: FACES_BOUNDARY_PERIODIC (every message to
   * self's wrap partner) or FACES_BOUNDARY_OPEN (the reference's: no
   * neighbour outside the rank grid). */
  Faces(int rank, int lx, int ly, int lz, int mx, int my, int mz, int n,
        int boundary = FACES_BOUNDARY_PERIODIC, long plane_stride = 0)
  {
    check(faces_plan_create(&plan_, rank, lx, ly, lz, mx, my, mz, n, boundary, plane_stride),
          "faces_plan_create");
  }
  ~Faces()
  {
    if (plan_) (void)faces_plan_destroy(plan_);
  }
  Faces(const Faces &) = delete;
  Faces &operator=(const Faces &) = delete;

  /* Faces::share (Faces.cc:111-678): returns once the step is done, as the
   * reference does (Faces.cc:675-677). */
  void share(Double6D &u, bool compute = true)
  {
    check(faces_share(plan_, u.data(), compute ? 1 : 0, nullptr), "faces_share");
    check_hip(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
  }
  /* The same step enqueued on `stream` (capturable in a HIP graph). */
  void share_async(double *u, bool compute, hipStream_t stream)
  {
    check(faces_share(plan_, u, compute ? 1 : 0, stream), "faces_share");
  }
  /* Faces::neighbor (Faces.cc:92-103): -1 for none. */
  int neighbor(int slot) const
  {
    int r = -1;
    check(faces_plan_neighbor(plan_, slot, &r), "faces_plan_neighbor");
    return r;
  }
  faces_plan *plan() { return plan_; }

 private:
  faces_plan *plan_ = nullptr;
};

}  // namespace faces_amd
