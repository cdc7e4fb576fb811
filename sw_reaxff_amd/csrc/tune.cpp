// tune.cpp — the autotuner of launch shapes.
//
// Several kernels are compiled in more than one launch shape, whose speed
// depends on the workload (row lengths, system sizes, how many systems a
// batch holds) more than on anything known at compile time:
//
//   spmv_fp64   the FP64 CG's SpMV (k_cg_spmv): T threads per panel of H
//               rows, G lanes per row and U groups of 4 entries in flight
//               per lane (T512G8U4, T512G8U2, T256G8U4, T256G4U4);
//   spmv_fp32   the mixed-precision inner solve's FP32 SpMV (k_pcg_spmv);
//   spmv_resid  the mixed-precision outer FP64 residual (k_resid);
//   far         the far pass (k_far_ws): gatherer and writer waves per
//               workgroup (G4W2, G4W4, G6W2, G8W4);
//   panels      whether the SpMV panels keep their LDS column lists (on)
//               or gather every entry from global memory (off); not timed,
//               set with swr_set_shape (e.g. to compare the two paths).
//
// swr_autotune times every shape of every family on the system's current
// state with HIP events and keeps the fastest; later launches use it.  A
// shape changes only the order of floating-point sums inside a row, never a
// list or a cutoff decision.  With SWR_TUNE_CACHE naming a file, choices are
// appended there keyed by device, family and size class (log2 of the local
// atom count), and swr_autotune_load applies the cached ones to a system.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "swr_internal.h"

namespace {

const char *const kSpmvNames[] = {"T512G8U4", "T512G8U2", "T256G8U4", "T256G4U4"};
constexpr int kSpmvN = 4;
constexpr int kSpmvDefault = 0;   // T512G8U4
const char *const kFarNames[] = {"G4W2", "G4W4", "G6W2", "G8W4"};
constexpr int kFarN = SWR_FAR_NSHAPES;

const char *const kPanelNames[] = {"on", "off"};

struct Family {
  const char *name;
  const char *const *shapes;
  int nshapes;
  int dflt;
  bool timed;   // swr_autotune times it (otherwise set with swr_set_shape only)
};
const Family kFamilies[] = {{"spmv_fp64", kSpmvNames, kSpmvN, kSpmvDefault, true},
                            {"spmv_fp32", kSpmvNames, kSpmvN, kSpmvDefault, true},
                            {"spmv_resid", kSpmvNames, kSpmvN, kSpmvDefault, true},
                            {"far", kFarNames, kFarN, SWR_FAR_G4W2, true},
                            {"panels", kPanelNames, 2, 0, false}};
constexpr int kNFam = 5;

int *slot(swr_system *S, int fam) {
  switch (fam) {
    case 0: return &S->spmv_shape[0];
    case 1: return &S->spmv_shape[1];
    case 2: return &S->spmv_shape[2];
    case 3: return &S->far_shape;
    default: return &S->panel_shape;
  }
}

int family_index(const char *name) {
  for (int f = 0; f < kNFam; f++)
    if (!std::strcmp(kFamilies[f].name, name)) return f;
  return -1;
}

int shape_index(int fam, const char *name) {
  for (int k = 0; k < kFamilies[fam].nshapes; k++)
    if (!std::strcmp(kFamilies[fam].shapes[k], name)) return k;
  return -1;
}

int current(swr_system *S, int fam) {
  const int v = *slot(S, fam);
  return v >= 0 ? v : kFamilies[fam].dflt;
}

std::string device_name(swr_system *S) {
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, S->device) != hipSuccess) return "unknown";
  return std::string(p.gcnArchName) + ":" + std::to_string(p.multiProcessorCount) + "CU";
}

int size_class(swr_system *S) { return S->n > 0 ? (int)std::floor(std::log2((double)S->n)) : 0; }

// one timed family: every shape, the fastest kept; timings appended to rep
int tune_family(swr_system *S, int fam, int reps, std::ostringstream &rep) {
  const Family &F = kFamilies[fam];
  double best = 1e300;
  int best_k = current(S, fam);
  rep << "\"" << F.name << "\": {\"us\": {";
  for (int k = 0; k < F.nshapes; k++) {
    float ms = 0.f;
    int rc;
    if (fam == 0) rc = swr_qeq_time_spmv(S, k, reps, &ms);
    else if (fam == 1) rc = swr_qeq_time_mixed(S, 1, k, reps, &ms);
    else if (fam == 2) rc = swr_qeq_time_mixed(S, 0, k, reps, &ms);
    else rc = swr_lists_time_far(S, k, reps, &ms);
    if (rc) return rc;
    rep << (k ? ", " : "") << "\"" << F.shapes[k] << "\": " << (int)std::lround(ms * 1e3 * 10) / 10.0;
    if (ms < best) {
      best = ms;
      best_k = k;
    }
  }
  *slot(S, fam) = best_k;
  rep << "}, \"chosen\": \"" << F.shapes[best_k] << "\"}";
  return SWR_OK;
}

}  // namespace

extern "C" {

// Times the launch shapes of every tunable family the system's state allows
// (the FP64 SpMV once H exists, the mixed-precision SpMVs after a mixed
// solve, the far pass once the lists exist) and keeps the fastest of each.
// Run between steps; the far family re-runs the pair stages
// (swr_pair_compute) so the lists are those of the chosen shape afterwards.
int swr_autotune(swr_system *S, int reps) {
  if (!S) return SWR_ERR_ARG;
  if (reps < 1) reps = 5;
  int rc = swr_sync(S);
  if (rc) return rc;
  std::ostringstream rep;
  rep << "{\"device\": \"" << device_name(S) << "\", \"atoms\": " << S->n << ", \"reps\": " << reps;
  bool far_done = false;
  for (int fam = 0; fam < kNFam; fam++) {
    if (!kFamilies[fam].timed) continue;
    const bool ok = fam == 0 ? S->H_val != nullptr
                  : fam <= 2 ? (S->mixed && S->mixw != nullptr && S->H_val32 != nullptr)
                             : (S->far != nullptr && S->bonds != nullptr);
    if (!ok) continue;
    rep << ", ";
    if ((rc = tune_family(S, fam, reps, rep))) return rc;
    if (fam == 3) far_done = true;
  }
  rep << "}";
  S->tune_report = rep.str();
  if (far_done && (rc = swr_pair_compute(S))) return rc;
  if ((rc = swr_sync(S))) return rc;
  // optional cache: one line per family
  if (const char *path = std::getenv("SWR_TUNE_CACHE")) {
    std::ofstream f(path, std::ios::app);
    if (f)
      for (int fam = 0; fam < kNFam; fam++)
        if (*slot(S, fam) >= 0)
          f << device_name(S) << " " << kFamilies[fam].name << " " << size_class(S) << " "
            << kFamilies[fam].shapes[*slot(S, fam)] << "\n";
  }
  return SWR_OK;
}

// Applies the choices SWR_TUNE_CACHE holds for this device and the system's
// size class (the last line of each family wins); returns the number applied.
int swr_autotune_load(swr_system *S) {
  if (!S) return -SWR_ERR_ARG;
  const char *path = std::getenv("SWR_TUNE_CACHE");
  if (!path) return 0;
  std::ifstream f(path);
  if (!f) return 0;
  const std::string dev = device_name(S);
  const int sc = size_class(S);
  std::map<int, int> got;
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream is(line);
    std::string d, fam, shape;
    int c = -1;
    if (!(is >> d >> fam >> c >> shape) || d != dev || c != sc) continue;
    const int fi = family_index(fam.c_str());
    if (fi < 0) continue;
    const int k = shape_index(fi, shape.c_str());
    if (k >= 0) got[fi] = k;
  }
  for (auto &kv : got) *slot(S, kv.first) = kv.second;
  return (int)got.size();
}

// Forces a family's shape (names as in swr_tune_report), e.g. to reproduce
// a tuned run or to check that every shape gives the same results.
int swr_set_shape(swr_system *S, const char *family, const char *shape) {
  if (!S || !family || !shape) return SWR_ERR_ARG;
  const int fi = family_index(family);
  if (fi < 0) {
    swr_set_error_msg(std::string("unknown tuning family ") + family);
    return SWR_ERR_ARG;
  }
  const int k = shape_index(fi, shape);
  if (k < 0) {
    swr_set_error_msg(std::string("unknown shape ") + shape + " for " + family);
    return SWR_ERR_ARG;
  }
  if (int rc = swr_sync(S)) return rc;
  *slot(S, fi) = k;   // (panels: from the next H build, swr_qeq_init_matvec)
  return SWR_OK;
}

// The shape a family launches with now (its name into out, len bytes).
int swr_get_shape(swr_system *S, const char *family, char *out, int len) {
  if (!S || !family || !out || len <= 0) return SWR_ERR_ARG;
  const int fi = family_index(family);
  if (fi < 0) return SWR_ERR_ARG;
  std::snprintf(out, (size_t)len, "%s", kFamilies[fi].shapes[current(S, fi)]);
  return SWR_OK;
}

// JSON text of the last swr_autotune (timings per shape, the choices);
// returns its length (copies at most len - 1 bytes when out is given).
int64_t swr_tune_report(swr_system *S, char *out, int64_t len) {
  if (!S) return -1;
  const std::string &r = S->tune_report;
  if (out && len > 0) {
    const size_t k = std::min((size_t)(len - 1), r.size());
    std::memcpy(out, r.data(), k);
    out[k] = '\0';
  }
  return (int64_t)r.size();
}

}  // extern "C"
