// ssg_plan.hpp -- the plan object behind the C ABI (internal to the library).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "springsteel_hip.h"
#include "ssg_kernels.hpp"
#include "ssg_setup.hpp"

struct ssg_plan {
  ssg::Setup S;
  ssg::DevTables T{};
  int device = 0;
  std::vector<void*> allocs;
  double* cb = nullptr;    // [v][zb][iDim] forward Chebyshev modes (RZ)
  double* cm = nullptr;    // [v][zb][jDim] forward Chebyshev modes of every ring point (RLZ)
  int precision = SSG_FP64;
  float* X32 = nullptr;    // FP32 form: the synthesis input in float (cm, rc, y stay FP64)
  double* rc = nullptr;    // [v*nzb+zb][b_jDim] forward ring coefficients (RL, RLZ)
  double* y = nullptr;     // banded-solve scratch: [maxMin][nlines] (RZ), [maxMin][lanes][2] (RL, RLZ)
  double* X = nullptr;     // [nvb][Ktot][P|Q][3][ntb] phase-rotated SI values, ring synthesis input (RL, RLZ)
  double* cbd = nullptr;   // [3][v][zb][iDim] (RZ)
  // ring synthesis by chirp-z transforms: on when every ring's transform fits
  // kFftMaxM; the block table of its launches; RLZ output Y[v][5 ntb][jDim]
  // before the inverse Chebyshev stage (double, or float for the FP32 form)
  bool fft_syn = false;
  int* fft_blk = nullptr;
  int fft_nblocks = 0;
  // ring analysis by chirp-z transforms (both forms: the forward path is FP64)
  bool fft_ana = false;
  int* ana_blk = nullptr;
  int ana_nblocks = 0;
  void* ysyn = nullptr;
  void* h_phys = nullptr;    // device staging for host-memory calls (physical: double, or float for the FP32 form)
  double* h_spec = nullptr;
  double* h_spec2 = nullptr;                 // second spectral staging buffer (host-memory tile calls)
  double* ahat = nullptr;     // R3X background coefficients (spectral layout), set by ssg_plan_set_ahat
  double* sbt = nullptr;      // btilde = spectral - pq ahat, the solve's input while ahat is set
  long* filt_off = nullptr;   // applyFilter!: start (device layout) of each line the filter changes
  double* filt_w = nullptr;   // and its weight (0.0: store zeros)
  long filt_n = 0;            // lines in the table; 0: no filter set
  void* reg = nullptr;       // regular-grid evaluation scratch (grown on demand, freed with the plan)
  size_t reg_bytes = 0;
  bool timing = false;
  struct Rec { const char* name; hipEvent_t a, b; };
  std::vector<Rec> recs;               // stage launches recorded since the last query
  std::vector<hipEvent_t> pool;        // idle events
  std::vector<double> last_ms;         // per-stage average of the last query
  std::vector<const char*> last_names;
};

namespace ssg {
// records the message of a failing call for ssg_last_error and returns code
int set_error(int code, const std::string& msg);
}  // namespace ssg
