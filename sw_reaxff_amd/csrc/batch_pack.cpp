// Host-side packing of independent systems into one batch (swr_batch_pack).
// Ghost images follow LAMMPS Comm::borders for an orthogonal periodic box:
// per dimension, two swaps over the atoms present when the dimension starts.
#include <cmath>
#include <cstring>
#include <vector>

#include "swr_internal.h"

namespace {

// Slab membership of an atom in dimension d: an atom of the low slab
// (x <= lo + cut) is imaged at +L, one of the high slab (x >= hi - cut) at
// -L; an image keeps its other coordinates, so an atom ends up with
// prod_d (1 + low_d + high_d) copies in all.
struct Slabs {
  double c0[3], c1[3], L[3];
  Slabs(const double *lo, const double *hi, double cut) {
    for (int d = 0; d < 3; d++) {
      L[d] = hi[d] - lo[d];
      c0[d] = lo[d] + cut;
      c1[d] = hi[d] - cut;
    }
  }
};

struct Scratch {
  std::vector<double> x;
  std::vector<int> owner;
  std::vector<signed char> img;
};

// Ghost images of one periodic system into S (local atoms first, then ghosts
// in borders order); S is sized to n + ng beforehand.
void borders(int n, const double *x, const Slabs &B, Scratch &S) {
  std::memcpy(S.x.data(), x, sizeof(double) * 3 * (size_t)n);
  for (int i = 0; i < n; i++) S.owner[i] = i;
  std::memset(S.img.data(), 0, 3 * (size_t)n);
  size_t top = n;
  for (int d = 0; d < 3; d++) {
    const size_t nlast = top;
    for (int side = 0; side < 2; side++) {
      for (size_t a = 0; a < nlast; a++) {
        const double v = S.x[3 * a + d];
        if (side == 0 ? v <= B.c0[d] : v >= B.c1[d]) {
          for (int e = 0; e < 3; e++) {
            S.x[3 * top + e] = S.x[3 * a + e];
            S.img[3 * top + e] = S.img[3 * a + e];
          }
          if (side == 0) { S.x[3 * top + d] += B.L[d]; S.img[3 * top + d] += 1; }
          else { S.x[3 * top + d] -= B.L[d]; S.img[3 * top + d] -= 1; }
          S.owner[top] = S.owner[a];
          top++;
        }
      }
    }
  }
}

}  // namespace

extern "C" int swr_batch_pack(int nsys, const int *nat, const double *x, const int *type, const int *tag,
                              const double *lo, const double *hi, const int *periodic, double cutghost,
                              long long cap, double *x_out, int *type_out, int *tag_out, int *owner_out,
                              signed char *img_out, int *row_off, int *ghost_off, double *shift,
                              long long *N_out) {
  if (nsys <= 0 || !nat || !lo || !hi || !periodic || !N_out || !(cutghost >= 0)) {
    swr_set_error_msg("swr_batch_pack: bad arguments");
    return SWR_ERR_ARG;
  }
  std::vector<long long> first(nsys + 1, 0), nghost(nsys, 0);
  for (int k = 0; k < nsys; k++) {
    if (nat[k] < 0) { swr_set_error_msg("swr_batch_pack: negative atom count"); return SWR_ERR_ARG; }
    first[k + 1] = first[k] + nat[k];
  }
  if (first[nsys] > 0 && (!x || !type || !tag)) { swr_set_error_msg("swr_batch_pack: missing atoms"); return SWR_ERR_ARG; }
  // pass 1: ghost counts and extents (atoms, their images and the box)
  std::vector<double> elo(3 * (size_t)nsys), ehi(3 * (size_t)nsys);
  double size = 0;
  long long n = first[nsys], ng = 0;
  for (int k = 0; k < nsys; k++) {
    const double *xk = x ? x + 3 * first[k] : nullptr;
    const double *lk = lo + 3 * k, *hk = hi + 3 * k;
    double mn[3], mx[3];
    for (int d = 0; d < 3; d++) { mn[d] = lk[d]; mx[d] = hk[d]; }
    if (periodic[k]) {
      for (int d = 0; d < 3; d++)
        if (!(hk[d] > lk[d])) { swr_set_error_msg("swr_batch_pack: empty periodic box"); return SWR_ERR_ARG; }
      // one image per side and dimension covers the ghost shell only when
      // the box is at least cutghost long (LAMMPS then needs several swaps)
      for (int d = 0; d < 3; d++)
        if (hk[d] - lk[d] < cutghost) {
          swr_set_error_msg("swr_batch_pack: periodic box shorter than the ghost cutoff");
          return SWR_ERR_ARG;
        }
      const Slabs B(lk, hk, cutghost);
      for (int i = 0; i < nat[k]; i++) {
        long long copies = 1;
        for (int d = 0; d < 3; d++) {
          const double v = xk[3 * i + d];
          const bool low = v <= B.c0[d], high = v >= B.c1[d];
          copies *= 1 + (int)low + (int)high;
          mn[d] = v < mn[d] ? v : mn[d];
          mx[d] = v > mx[d] ? v : mx[d];
          if (low) { const double g = v + B.L[d]; mx[d] = g > mx[d] ? g : mx[d]; }
          if (high) { const double g = v - B.L[d]; mn[d] = g < mn[d] ? g : mn[d]; }
        }
        nghost[k] += copies - 1;
      }
    } else {
      for (int i = 0; i < nat[k]; i++)
        for (int d = 0; d < 3; d++) {
          const double v = xk[3 * i + d];
          mn[d] = v < mn[d] ? v : mn[d];
          mx[d] = v > mx[d] ? v : mx[d];
        }
    }
    for (int d = 0; d < 3; d++) {
      elo[3 * k + d] = mn[d];
      ehi[3 * k + d] = mx[d];
      size = std::fmax(size, mx[d] - mn[d]);
    }
    ng += nghost[k];
  }
  *N_out = n + ng;
  if (!x_out) return SWR_OK;
  if (cap < n + ng || !type_out || !tag_out || !owner_out || !img_out || !row_off || !ghost_off || !shift) {
    swr_set_error_msg("swr_batch_pack: output arrays missing or too small");
    return SWR_ERR_ARG;
  }
  if (n + ng > 0x7fffffffLL) { swr_set_error_msg("swr_batch_pack: batch too large"); return SWR_ERR_ARG; }
  size = size + cutghost + 1.0;
  int m = 1;
  while ((long long)m * m * m < nsys) m++;
  row_off[0] = 0;
  ghost_off[0] = 0;
  for (int k = 0; k < nsys; k++) {
    row_off[k + 1] = row_off[k] + nat[k];
    ghost_off[k + 1] = ghost_off[k] + (int)nghost[k];
    const double cell[3] = {(double)(k % m), (double)((k / m) % m), (double)(k / (m * m))};
    for (int d = 0; d < 3; d++) shift[3 * k + d] = cell[d] * size - elo[3 * k + d];
  }
  // pass 2: images (borders order) and the translated, renumbered output
  Scratch S;
  for (int k = 0; k < nsys; k++) {
    const int nk = nat[k];
    const size_t tot = (size_t)nk + (size_t)nghost[k];
    const double *xk = x + 3 * first[k];
    const double *sk = shift + 3 * k;
    const int a = row_off[k];
    const long long g0 = n + ghost_off[k];
    if (periodic[k]) {
      S.x.resize(3 * tot);
      S.owner.resize(tot);
      S.img.resize(3 * tot);
      borders(nk, xk, Slabs(lo + 3 * k, hi + 3 * k, cutghost), S);
    }
    for (size_t i = 0; i < tot; i++) {
      const long long r = (long long)i < nk ? a + (long long)i : g0 + ((long long)i - nk);
      const int own = periodic[k] ? S.owner[i] : (int)i;
      const double *xi = periodic[k] ? &S.x[3 * i] : xk + 3 * i;
      for (int d = 0; d < 3; d++) {
        x_out[3 * r + d] = xi[d] + sk[d];
        img_out[3 * r + d] = periodic[k] ? S.img[3 * i + d] : 0;
      }
      type_out[r] = type[first[k] + own];
      tag_out[r] = tag[first[k] + own] + a;
      owner_out[r] = own + a;
    }
  }
  return SWR_OK;
}
