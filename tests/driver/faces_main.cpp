// faces_main -- the reference's stdin driver (gpu/main.cc:108-199) on top of
// the faces_amd C ABI, with the oracle as its correctness check.  Test
// infrastructure: it links the oracle, so it lives beside the tests.
//
// stdin:  lx ly lz mx my mz n nk nj ni     (README.md:61; gpu/main.cc:136-137)
//   nk  plans built in turn ("face inits", gpu/main.cc:175)
//   nj  timed field initialisations per plan, after one untimed (:176-183)
//   ni  shares per initialisation; negative = |ni| shares with compute = 0
//       (:157-161)
// options (argv):
//   --boundary periodic|open   periodic (default): every message addressed to
//                              self; open: ranks outside the lx*ly*lz grid do
//                              not exist (MPI_PROC_NULL), as in the reference
//   --init int|rand|ref        field generator (default int: integer-valued
//                              [0,1023]; ref: the formula of gpu/main.cc:18)
//   --seed S                   generator seed (default 42)
//   --rank R                   the rank of the lx*ly*lz grid this GPU holds
//                              (default 0); --all-ranks checks every rank in turn
//   --reference                shorthand for --boundary open --init ref
//
// stdout reproduces the reference's lines: the echo of the input
// (gpu/main.cc:138-142), "k: " per plan (:174), "N tasks passed correctness
// check" (:104) and "time T avg Tmin min Tmax max" (:194, seconds of the timed
// shares).  Extra figures (GB/s, bit-exactness) go to stderr.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/faces_amd.hpp"

extern "C" {
long orc_field_doubles(int n, int mx, int my, int mz, long ps);
void orc_fill(double *u, int n, int mx, int my, int mz, long ps, int kind, uint64_t seed);
void orc_share_sweep(double *u, int n, int mx, int my, int mz, long ps, int periodic, int compute);
int orc_share_staged(double *u, int n, int mx, int my, int mz, long ps, int nbr_mask, int compute);
int orc_compare(const double *u, const double *v, int n, int mx, int my, int mz, long ps, double *out);
}

namespace {

int usage(const char *msg)
{
  std::fprintf(stderr, "faces_main: %s\n", msg);
  std::fprintf(stderr,
               "usage: faces_main [--boundary periodic|open] [--init int|rand|ref] [--seed S]\n"
               "                  [--rank R | --all-ranks] [--reference]  < 'lx ly lz mx my mz n nk nj ni'\n");
  return 2;
}

// Self-halo bits of a rank in an open lx*ly*lz grid: bit 2*axis for a low
// neighbour, 2*axis+1 for a high one (Faces.cc:33-35, 92-103).
int neighbour_mask(int rank, int lx, int ly, int lz)
{
  const int k[3] = {rank % lx, (rank / lx) % ly, rank / (lx * ly)};
  const int l[3] = {lx, ly, lz};
  int m = 0;
  for (int a = 0; a < 3; a++) {
    if (k[a] > 0) m |= 1 << (2 * a);
    if (k[a] < l[a] - 1) m |= 1 << (2 * a + 1);
  }
  return m;
}

}  // namespace

int main(int argc, char **argv)
{
  int boundary = FACES_BOUNDARY_PERIODIC, kind = FACES_FILL_INT, rank = 0;
  bool all_ranks = false;
  uint64_t seed = 42;
  for (int i = 1; i < argc; i++) {
    const std::string a = argv[i];
    auto next = [&](void) -> const char * { return i + 1 < argc ? argv[++i] : nullptr; };
    if (a == "--boundary") {
      const char *v = next();
      if (!v) return usage("--boundary needs a value");
      if (!std::strcmp(v, "periodic")) boundary = FACES_BOUNDARY_PERIODIC;
      else if (!std::strcmp(v, "open")) boundary = FACES_BOUNDARY_OPEN;
      else return usage("unknown boundary");
    } else if (a == "--init") {
      const char *v = next();
      if (!v) return usage("--init needs a value");
      if (!std::strcmp(v, "int")) kind = FACES_FILL_INT;
      else if (!std::strcmp(v, "rand")) kind = FACES_FILL_RAND;
      else if (!std::strcmp(v, "ref")) kind = FACES_FILL_REF;
      else return usage("unknown init");
    } else if (a == "--seed") {
      const char *v = next();
      if (!v) return usage("--seed needs a value");
      seed = std::strtoull(v, nullptr, 0);
    } else if (a == "--rank") {
      const char *v = next();
      if (!v) return usage("--rank needs a value");
      rank = std::atoi(v);
    } else if (a == "--all-ranks") {
      all_ranks = true;
    } else if (a == "--reference") {
      boundary = FACES_BOUNDARY_OPEN;
      kind = FACES_FILL_REF;
    } else {
      return usage(("unknown option " + a).c_str());
    }
  }

  int lx, ly, lz, mx, my, mz, n, nk, nj, ni;
  if (!(std::cin >> lx >> ly >> lz >> mx >> my >> mz >> n >> nk >> nj >> ni))
    return usage("expected 10 integers on stdin");
  if (lx < 1 || ly < 1 || lz < 1 || mx < 1 || my < 1 || mz < 1 || n < 2 || nk < 0 || nj < 0)
    return usage("invalid sizes");
  const int nranks = lx * ly * lz;
  if (rank < 0 || rank >= nranks) return usage("--rank outside the rank grid");

  // gpu/main.cc:138-142
  std::cout << lx << ' ' << ly << ' ' << lz << " tasks\n";
  std::cout << mx << ' ' << my << ' ' << mz << " local elements of size " << n << '\n';
  std::cout << nk << " face inits x " << nj << " element inits x ";
  if (ni < 0) std::cout << -ni << " shares without compute kernels" << std::endl;
  else std::cout << ni << " shares" << std::endl;
  const bool compute = ni >= 0;
  const int nshare = ni < 0 ? -ni : ni;

  const int r0 = all_ranks ? 0 : rank, r1 = all_ranks ? nranks : rank + 1;
  int npass = 0, nexact = 0;
  double tsum = 0, tmin = 1e300, tmax = 0;
  try {
    faces_amd::Faces::Double6D u(n, mx, my, mz);  // gpu/main.cc:169
    const long nd = u.size();
    std::vector<double> host, want(nd);
    for (int r = r0; r < r1; r++) {
      // the oracle's result for this rank: nshare shares of the initial field
      orc_fill(want.data(), n, mx, my, mz, 0, kind, seed);
      const int mask = boundary == FACES_BOUNDARY_OPEN ? neighbour_mask(r, lx, ly, lz) : 0;
      for (int i = 0; i < nshare; i++) {
        if (boundary == FACES_BOUNDARY_PERIODIC) orc_share_sweep(want.data(), n, mx, my, mz, 0, 1, compute);
        else orc_share_staged(want.data(), n, mx, my, mz, 0, mask, compute);
      }
      // gpu/main.cc:172-184: nk plans, each nj+1 initialisations (the first
      // untimed) of nshare shares
      double t = 0;
      for (int k = 0; k < nk; k++) {
        if (r == r0) std::cout << (k + 1) << ": " << std::flush;
        faces_amd::Faces faces(r, lx, ly, lz, mx, my, mz, n, boundary);
        for (int j = 0; j <= nj; j++) {
          u.fill(kind, seed);
          faces_amd::check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
          const auto t0 = std::chrono::steady_clock::now();
          for (int i = 0; i < nshare; i++) faces.share(u, compute);
          faces_amd::check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
          const auto t1 = std::chrono::steady_clock::now();
          if (j > 0) t += std::chrono::duration<double>(t1 - t0).count();
        }
      }
      if (nk == 0) {  // nothing stepped: the field is the initial one
        orc_fill(want.data(), n, mx, my, mz, 0, kind, seed);
        u.fill(kind, seed);
        faces_amd::check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
      }
      u.copy(host);
      double cmp[4];
      orc_compare(host.data(), want.data(), n, mx, my, mz, 0, cmp);
      // gpu/main.cc:65-80: pass when the largest |u-v|/(|u|+|v|) is below 1e-15
      if (cmp[0] < 1e-15) npass++;
      else std::cout << r << " FAIL " << cmp[0] << " (index " << (long)cmp[1] << ")" << std::endl;
      if (cmp[2] == 0) nexact++;
      tsum += t;
      tmin = std::min(tmin, t);
      tmax = std::max(tmax, t);
    }
  } catch (const faces_amd::Error &e) {
    std::fprintf(stderr, "faces_main: %s\n", e.what());
    return 1;
  }
  const int nchecked = r1 - r0;
  std::cout << npass << " tasks passed correctness check" << std::endl;
  std::cout << "time " << tsum / nchecked << " avg " << tmin << " min " << tmax << " max" << std::endl;
  const double steps = (double)nk * nj * nshare;
  if (steps > 0 && tsum > 0) {
    const double bytes = 16.0 * n * n * n * (double)mx * my * mz;
    std::fprintf(stderr, "faces_main: %d of %d ranks bit-exact; %.3f GB/s over %.0f timed shares per rank\n",
                 nexact, nchecked, bytes * steps / (tsum / nchecked) / 1e9, steps);
  } else {
    std::fprintf(stderr, "faces_main: %d of %d ranks bit-exact\n", nexact, nchecked);
  }
  return npass == nchecked ? 0 : 3;
}
