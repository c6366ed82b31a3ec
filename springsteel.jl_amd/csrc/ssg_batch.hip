// ssg_batch.hip -- the batching form: many RL / RLZ grids of one arithmetic
// form transformed by one sequence of batched launches captured in a hipGraph.
//
// Every stage kernel runs once per launch shape for all grids of the batch: a
// job is one grid's share of a stage, the jobs of one launch share the kernel
// instance, block size and dynamic LDS size, and each block finds its job from
// the prefix of block counts (batch_job).  The stage order is that of the
// single-grid path, so each grid's results are bitwise those of its own plan.
// The graph of a direction is captured once (and again after tuning) and then
// replayed with one hipGraphLaunch.
//
// The tuner times the candidate launch shapes of the three tunable stages
// (synthesis: column/quad tiles per wave and quad groups per block; analysis:
// row tiles per wave; SI: cell segments per lane) on each grid's own arrays
// and keeps the fastest per grid.  The physical arrays of an FP32 batch are
// float; its forward path, solve and SI run the FP64 kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <initializer_list>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "ssg_plan.hpp"

struct ssg_batch {
  std::vector<ssg_plan*> plans;
  std::vector<void*> phys;                    // element type R of the form: double, or float (FP32 form)
  std::vector<double*> spec;
  int precision = SSG_FP64;
  std::vector<void*> allocs;                  // job and start arrays of the current graphs
  std::vector<std::function<void(hipStream_t)>> launches[2];   // 0: spectral, 1: grid
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  int nlaunch[2] = {0, 0};
};

namespace {

#define BCHK(expr)                                                                               \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess) return ssg::set_error(SSG_EHIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

template <typename J>
struct Group {
  std::vector<J> jobs;
  std::vector<int> start{0};
  size_t lds = 0;
  J* d_jobs = nullptr;
  int* d_start = nullptr;
  int nblocks() const { return start.back(); }
  void add(const J& j, int blocks, size_t l) {
    jobs.push_back(j);
    start.push_back(start.back() + (blocks + 7) / 8 * 8);   // XCD-aligned job ranges
    lds = std::max(lds, l);
  }
};

template <typename J>
int upload_group(ssg_batch* b, Group<J>& g) {
  void* pj = nullptr;
  void* ps = nullptr;
  BCHK(hipMalloc(&pj, g.jobs.size() * sizeof(J)));
  b->allocs.push_back(pj);
  BCHK(hipMalloc(&ps, g.start.size() * sizeof(int)));
  b->allocs.push_back(ps);
  BCHK(hipMemcpy(pj, g.jobs.data(), g.jobs.size() * sizeof(J), hipMemcpyHostToDevice));
  BCHK(hipMemcpy(ps, g.start.data(), g.start.size() * sizeof(int), hipMemcpyHostToDevice));
  g.d_jobs = static_cast<J*>(pj);
  g.d_start = static_cast<int*>(ps);
  return SSG_OK;
}

void release_graphs(ssg_batch* b) {
  for (int d = 0; d < 2; d++) {
    if (b->exec[d]) (void)hipGraphExecDestroy(b->exec[d]);
    if (b->graph[d]) (void)hipGraphDestroy(b->graph[d]);
    b->exec[d] = nullptr;
    b->graph[d] = nullptr;
  }
  for (void* a : b->allocs) (void)hipFree(a);
  b->allocs.clear();
  b->launches[0].clear();
  b->launches[1].clear();
}

// the stage launches of one direction for the form whose physical arrays and
// synthesis are of type R (the forward path, the solve and SI run in FP64 in
// both forms): job arrays are built and uploaded here, the launches kept as
// closures for the capture
template <typename R>
int prepare(ssg_batch* b, int dir) {
  auto& L = b->launches[dir];
  int n = 0;
  if (dir == 0) {
    // stages whose LDS size differs between grids are grouped by it: a launch's LDS
    // size is its largest job's, and it sets the blocks per CU of every job
    std::map<size_t, Group<ssg::ChebJob<R>>> cheb;
    std::map<std::pair<int, int>, Group<ssg::AnaJob<double, double>>> ana_d;   // input FP64 modes (RLZ) / physical
    std::map<std::pair<int, int>, Group<ssg::AnaJob<double, R>>> ana_p;        // input physical of type R (RL)
    std::map<size_t, Group<ssg::SbJob>> sb;
    Group<ssg::AnaFftJob<double>> anaf_d;                                      // chirp-z analysis
    Group<ssg::AnaFftJob<float>> anaf_f;
    for (size_t i = 0; i < b->plans.size(); i++) {
      ssg_plan* p = b->plans[i];
      const ssg::DevTables& T = p->T;
      const bool rlz = T.geometry == SSG_RLZ;
      const R* phys = static_cast<const R*>(b->phys[i]);
      if (rlz) {
        const int kzp = (T.kDim + 3) / 4 * 4, bkp = (T.nzb + 15) / 16 * 16;
        const int blocks = ssg::cheb_fwd_pts_blocks(T);
        const size_t lds = ssg::cheb_fwd_pts_lds(T, sizeof(double));
        cheb[lds].add(ssg::ChebJob<R>{T, phys, p->cm, kzp, bkp, blocks / T.nvars}, blocks, lds);
      }
      const ssg::AnaLaunch A = ssg::plan_ana(T, sizeof(double));
      if (p->fft_ana) {
        if (rlz || sizeof(R) == sizeof(double)) {
          const double* src = rlz ? p->cm : reinterpret_cast<const double*>(phys);
          anaf_d.add(ssg::AnaFftJob<double>{T, p->ana_blk, src, p->rc}, p->ana_nblocks, 0);
        } else {
          anaf_f.add(ssg::AnaFftJob<float>{T, p->ana_blk, reinterpret_cast<const float*>(phys), p->rc}, p->ana_nblocks,
                     0);
        }
      } else if (rlz || sizeof(R) == sizeof(double)) {
        const double* src = rlz ? p->cm : reinterpret_cast<const double*>(phys);
        ana_d[{A.nct, A.mr}].add(ssg::AnaJob<double, double>{T, A, src, p->rc}, A.nblocks, A.lds);
      } else {
        ana_p[{A.nct, A.mr}].add(ssg::AnaJob<double, R>{T, A, phys, p->rc}, A.nblocks, A.lds);
      }
      sb[ssg::sb_lds(T)].add(ssg::SbJob{T, p->rc, b->spec[i]}, ssg::sb_blocks(T), ssg::sb_lds(T));
    }
    for (auto& kv : cheb) {
      Group<ssg::ChebJob<R>>& g = kv.second;
      if (int rc = upload_group(b, g)) return rc;
      L.push_back([j = g.d_jobs, s = g.d_start, nj = int(g.jobs.size()), nb = g.nblocks(), lds = g.lds](
                      hipStream_t st) { ssg::launch_cheb_fwd_pts_batch<R>(j, s, nj, nb, lds, st); });
      n++;
    }
    if (!anaf_d.jobs.empty()) {
      if (int rc = upload_group(b, anaf_d)) return rc;
      L.push_back([j = anaf_d.d_jobs, s = anaf_d.d_start, nj = int(anaf_d.jobs.size()), nb = anaf_d.nblocks()](
                      hipStream_t st) { ssg::launch_ana_fft_batch<double>(j, s, nj, nb, st); });
      n++;
    }
    if (!anaf_f.jobs.empty()) {
      if (int rc = upload_group(b, anaf_f)) return rc;
      L.push_back([j = anaf_f.d_jobs, s = anaf_f.d_start, nj = int(anaf_f.jobs.size()), nb = anaf_f.nblocks()](
                      hipStream_t st) { ssg::launch_ana_fft_batch<float>(j, s, nj, nb, st); });
      n++;
    }
    for (auto& kv : ana_d) {
      if (int rc = upload_group(b, kv.second)) return rc;
      L.push_back([j = kv.second.d_jobs, s = kv.second.d_start, nj = int(kv.second.jobs.size()),
                   nb = kv.second.nblocks(), nct = kv.first.first, mr = kv.first.second, lds = kv.second.lds](
                      hipStream_t st) { ssg::launch_ring_ana_batch<double, double>(j, s, nj, nb, nct, mr, lds, st); });
      n++;
    }
    if constexpr (sizeof(R) != sizeof(double)) {
      for (auto& kv : ana_p) {
        if (int rc = upload_group(b, kv.second)) return rc;
        L.push_back([j = kv.second.d_jobs, s = kv.second.d_start, nj = int(kv.second.jobs.size()),
                     nb = kv.second.nblocks(), nct = kv.first.first, mr = kv.first.second, lds = kv.second.lds](
                        hipStream_t st) { ssg::launch_ring_ana_batch<double, R>(j, s, nj, nb, nct, mr, lds, st); });
        n++;
      }
    }
    for (auto& kv : sb) {
      Group<ssg::SbJob>& g = kv.second;
      if (int rc = upload_group(b, g)) return rc;
      L.push_back([j = g.d_jobs, s = g.d_start, nj = int(g.jobs.size()), nb = g.nblocks(), lds = g.lds](
                      hipStream_t st) { ssg::launch_sb_ring_batch(j, s, nj, nb, lds, st); });
      n++;
    }
  } else {
    Group<ssg::SaJob> sa;
    Group<ssg::SiJob<R>> si;
    std::map<std::pair<int, int>, Group<ssg::SaSiJob<R>>> sasi;  // fused SA + SI, by (threads, form)
    std::map<std::pair<int, int>, Group<ssg::SynJob<R>>> syn;   // (shape, threads)
    Group<ssg::FftJob<R, R>> fft;                                // chirp-z synthesis
    std::map<std::pair<int, size_t>, Group<ssg::ChebSynJob<R>>> csyn;   // RLZ inverse Chebyshev stage, by (class, LDS)
    for (size_t i = 0; i < b->plans.size(); i++) {
      ssg_plan* p = b->plans[i];
      const ssg::DevTables& T = p->T;
      R* X = nullptr;
      if constexpr (sizeof(R) == sizeof(float)) X = p->X32; else X = p->X;
      if (ssg::sasi_supported(T)) {
        sasi[{ssg::sasi_threads(T), ssg::sasi_form(T)}].add(ssg::SaSiJob<R>{T, b->spec[i], X}, ssg::sasi_blocks(T),
                                                 ssg::sasi_lds(T));
      } else {
        sa.add(ssg::SaJob{T, b->spec[i], p->y}, ssg::sa_ring_blocks(T), 0);
        const int nseg = ssg::si_segments(T);
        si.add(ssg::SiJob<R>{T, nseg, p->y, X}, ssg::si_ring_blocks(T, nseg), 0);
      }
      R* phys = static_cast<R*>(b->phys[i]);
      if (p->fft_syn) {
        R* out = T.geometry == SSG_RLZ ? static_cast<R*>(p->ysyn) : phys;
        fft.add(ssg::FftJob<R, R>{T, p->fft_blk, X, out}, p->fft_nblocks, 0);
        if (T.geometry == SSG_RLZ)
          csyn[{ssg::syn_cheb_class(T), ssg::syn_cheb_lds(T, sizeof(R))}].add(
              ssg::ChebSynJob<R>{T, static_cast<const R*>(p->ysyn), phys}, ssg::syn_cheb_blocks(T),
              ssg::syn_cheb_lds(T, sizeof(R)));
      } else {
        const ssg::SynLaunch P = ssg::plan_syn(T, sizeof(R));
        const int shape = P.ct == 1 ? 1 : (P.qt == 1 ? 2 : 3);
        syn[{shape, P.nthr}].add(ssg::SynJob<R>{T, P, X, phys}, P.nblocks, P.lds);
      }
    }
    for (auto& kv : sasi) {
      if (int rc = upload_group(b, kv.second)) return rc;
      L.push_back([j = kv.second.d_jobs, s = kv.second.d_start, nj = int(kv.second.jobs.size()),
                   nb = kv.second.nblocks(), nthr = kv.first.first, mu = kv.first.second,
                   lds = kv.second.lds](hipStream_t st) {
        ssg::launch_sasi_batch<R>(j, s, nj, nb, nthr, lds, mu, st);
      });
      n++;
    }
    if (!sa.jobs.empty()) {
      if (int rc = upload_group(b, sa)) return rc;
      L.push_back([j = sa.d_jobs, s = sa.d_start, nj = int(sa.jobs.size()), nb = sa.nblocks()](hipStream_t st) {
        ssg::launch_sa_ring_batch(j, s, nj, nb, st);
      });
      if (int rc = upload_group(b, si)) return rc;
      L.push_back([j = si.d_jobs, s = si.d_start, nj = int(si.jobs.size()), nb = si.nblocks()](hipStream_t st) {
        ssg::launch_si_ring_batch<R>(j, s, nj, nb, st);
      });
      n += 2;
    }
    for (auto& kv : syn) {
      if (int rc = upload_group(b, kv.second)) return rc;
      L.push_back([j = kv.second.d_jobs, s = kv.second.d_start, nj = int(kv.second.jobs.size()),
                   nb = kv.second.nblocks(), ct = kv.first.first, nthr = kv.first.second, lds = kv.second.lds](
                      hipStream_t st) { ssg::launch_ring_syn_batch<R>(j, s, nj, nb, ct, nthr, lds, st); });
      n++;
    }
    if (!fft.jobs.empty()) {
      if (int rc = upload_group(b, fft)) return rc;
      L.push_back([j = fft.d_jobs, s = fft.d_start, nj = int(fft.jobs.size()), nb = fft.nblocks()](hipStream_t st) {
        ssg::launch_syn_fft_batch<R, R>(j, s, nj, nb, st);
      });
      n++;
    }
    for (auto& kv : csyn) {
      Group<ssg::ChebSynJob<R>>& g = kv.second;
      if (int rc = upload_group(b, g)) return rc;
      L.push_back([j = g.d_jobs, s = g.d_start, nj = int(g.jobs.size()), nb = g.nblocks(), lds = g.lds,
                   cls = kv.first.first](hipStream_t st) { ssg::launch_syn_cheb_batch<R>(j, s, nj, nb, lds, cls, st); });
      n++;
    }
  }
  b->nlaunch[dir] = n;
  return SSG_OK;
}

int build(ssg_batch* b) {
  release_graphs(b);
  for (int dir = 0; dir < 2; dir++) {
    const int rc = b->precision == SSG_FP32_REFINE ? prepare<float>(b, dir) : prepare<double>(b, dir);
    if (rc) return rc;
  }
  hipStream_t cs = nullptr;
  BCHK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  int rc = SSG_OK;
  for (int dir = 0; dir < 2 && rc == SSG_OK; dir++) {
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) { rc = ssg::set_error(SSG_EHIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e)); break; }
    for (auto& f : b->launches[dir]) f(cs);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(cs, &g);
    b->graph[dir] = g;
    if (e != hipSuccess) { rc = ssg::set_error(SSG_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e)); break; }
    e = hipGraphInstantiate(&b->exec[dir], g, nullptr, nullptr, 0);
    if (e != hipSuccess) rc = ssg::set_error(SSG_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  (void)hipStreamDestroy(cs);
  return rc;
}

template <typename F>
float time_ms(hipStream_t st, int reps, F&& launch) {
  hipEvent_t a = nullptr, z = nullptr;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&z) != hipSuccess) return 1e30f;
  launch();                                         // warm-up
  (void)hipEventRecord(a, st);
  for (int r = 0; r < reps; r++) launch();
  (void)hipEventRecord(z, st);
  (void)hipEventSynchronize(z);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, a, z);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(z);
  return ms / reps;
}

// R: the type of physical and of the synthesis (double, or float for the FP32 form)
template <typename R>
int tune_plan(ssg_plan* p, R* phys, double* spec, hipStream_t st) {
  ssg::DevTables& T = p->T;
  constexpr bool f32 = sizeof(R) == sizeof(float);
  R* X = nullptr;
  if constexpr (f32) X = p->X32; else X = p->X;
  // fill the intermediates with this grid's data
  if constexpr (f32) {
    if (int rc = ssg_spectral_transform_f32(p, phys, spec, SSG_MEM_DEVICE, st)) return rc;
    if (int rc = ssg_grid_transform_f32(p, spec, phys, SSG_MEM_DEVICE, st)) return rc;
  } else {
    if (int rc = ssg_spectral_transform(p, phys, spec, SSG_MEM_DEVICE, st)) return rc;
    if (int rc = ssg_grid_transform(p, spec, phys, SSG_MEM_DEVICE, st)) return rc;
  }
  // synthesis of the direct form: (column tiles per wave, quad groups per block)
  if (!p->fft_syn) {
    float best = 1e30f;
    ssg::Tuning keep = T.tune;
    for (int ct = 1; ct <= (f32 ? 3 : 2); ct++)
      for (int nwq = 4; nwq >= 1; nwq /= 2) {
        T.tune.syn_ct = ct;
        T.tune.syn_nwq = nwq;
        const ssg::SynLaunch P = ssg::plan_syn(T, sizeof(R));
        if (!P.ok || P.nwq != nwq) continue;
        const float ms = time_ms(st, 3, [&] { ssg::launch_ring_syn<R>(T, X, phys, st); });
        if (hipGetLastError() != hipSuccess) continue;
        if (ms < best) { best = ms; keep.syn_ct = ct; keep.syn_nwq = nwq; }
      }
    T.tune.syn_ct = keep.syn_ct;
    T.tune.syn_nwq = keep.syn_nwq;
  }
  // analysis of the direct form: row tiles per wave
  if (!p->fft_ana) {
    float best = 1e30f;
    int keep = 0;
    for (int mr = 2; mr >= 1; mr--) {
      T.tune.ana_mr = mr;
      const ssg::AnaLaunch A = ssg::plan_ana(T, sizeof(double));
      if (A.mr != mr) continue;
      float ms;
      if (T.geometry == SSG_RLZ) ms = time_ms(st, 3, [&] { ssg::launch_ring_ana<double, double>(T, p->cm, p->rc, st); });
      else ms = time_ms(st, 3, [&] { ssg::launch_ring_ana<double, R>(T, phys, p->rc, st); });
      if (ms < best) { best = ms; keep = mr; }
    }
    T.tune.ana_mr = keep;
  }
  // SA + SI: cell segments per SI lane
  {
    float best = 1e30f;
    int keep_seg = 0;
    for (int seg = 1; seg <= 8; seg *= 2) {
      T.tune.si_seg = seg;
      const float ms = time_ms(st, 3, [&] { ssg::launch_sa_si_ring<R>(T, spec, p->y, X, st); });
      if (ms < best) { best = ms; keep_seg = seg; }
    }
    T.tune.si_seg = keep_seg;
  }
  BCHK(hipGetLastError());
  return SSG_OK;
}

}  // namespace

extern "C" {

int ssg_batch_create(ssg_plan* const* plans, int nplans, void* const* physical, double* const* spectral,
                     ssg_batch** out) {
  if (!plans || !physical || !spectral || !out || nplans < 1)
    return ssg::set_error(SSG_EINVAL, "ArgumentError: null argument or empty batch");
  *out = nullptr;
  ssg_batch* b = new ssg_batch();
  for (int i = 0; i < nplans; i++) {
    ssg_plan* p = plans[i];
    if (!p || !physical[i] || !spectral[i]) { delete b; return ssg::set_error(SSG_EINVAL, "ArgumentError: null plan or array in the batch"); }
    if (p->S.geometry == SSG_RZ) { delete b; return ssg::set_error(SSG_ENOTSUP, "ENOTSUP: the batching form takes RL and RLZ grids"); }
    if (p->S.any_r3x) {
      delete b;
      return ssg::set_error(SSG_ENOTSUP, "ENOTSUP: grids with R3X sides run the single-grid calls (background coefficients)");
    }
    if (i > 0 && p->precision != plans[0]->precision) {
      delete b; return ssg::set_error(SSG_EINVAL, "ArgumentError: the plans of a batch must share one arithmetic form");
    }
    b->plans.push_back(p);
    b->phys.push_back(physical[i]);
    b->spec.push_back(spectral[i]);
  }
  b->precision = plans[0]->precision;
  if (int rc = build(b)) { release_graphs(b); delete b; return rc; }
  *out = b;
  return SSG_OK;
}

int ssg_batch_spectral_transform(ssg_batch* b, void* stream) {
  if (!b) return ssg::set_error(SSG_EINVAL, "ArgumentError: null batch");
  BCHK(hipGraphLaunch(b->exec[0], static_cast<hipStream_t>(stream)));
  return SSG_OK;
}

int ssg_batch_grid_transform(ssg_batch* b, void* stream) {
  if (!b) return ssg::set_error(SSG_EINVAL, "ArgumentError: null batch");
  BCHK(hipGraphLaunch(b->exec[1], static_cast<hipStream_t>(stream)));
  return SSG_OK;
}

int ssg_batch_tune(ssg_batch* b, void* stream) {
  if (!b) return ssg::set_error(SSG_EINVAL, "ArgumentError: null batch");
  hipStream_t st = static_cast<hipStream_t>(stream);
  for (size_t i = 0; i < b->plans.size(); i++) {
    ssg_plan* p = b->plans[i];
    const int rc = b->precision == SSG_FP32_REFINE
                       ? tune_plan<float>(p, static_cast<float*>(b->phys[i]), b->spec[i], st)
                       : tune_plan<double>(p, static_cast<double*>(b->phys[i]), b->spec[i], st);
    if (rc) return rc;
  }
  BCHK(hipStreamSynchronize(st));
  return build(b);
}

int ssg_batch_launches(const ssg_batch* b, int* spectral_launches, int* grid_launches) {
  if (!b) return ssg::set_error(SSG_EINVAL, "ArgumentError: null batch");
  if (spectral_launches) *spectral_launches = b->nlaunch[0];
  if (grid_launches) *grid_launches = b->nlaunch[1];
  return SSG_OK;
}

int ssg_plan_tune(ssg_plan* p, void* physical, double* spectral, void* stream) {
  if (!p || !physical || !spectral) return ssg::set_error(SSG_EINVAL, "ArgumentError: null plan or array");
  if (p->S.geometry == SSG_RZ) return SSG_OK;        // nothing to tune on the Cartesian grid
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int rc = p->precision == SSG_FP32_REFINE ? tune_plan<float>(p, static_cast<float*>(physical), spectral, st)
                                                 : tune_plan<double>(p, static_cast<double*>(physical), spectral, st);
  if (rc) return rc;
  BCHK(hipStreamSynchronize(st));
  return SSG_OK;
}

int ssg_plan_get_tuning(const ssg_plan* p, int* out5) {
  if (!p || !out5) return ssg::set_error(SSG_EINVAL, "ArgumentError: null plan or output");
  out5[0] = p->T.tune.syn_ct;
  out5[1] = p->T.tune.syn_nwq;
  out5[2] = p->T.tune.ana_mr;
  out5[3] = p->T.tune.si_seg;
  out5[4] = p->T.tune.sa_mode;
  return SSG_OK;
}

int ssg_plan_set_tuning(ssg_plan* p, const int* in5) {
  if (!p || !in5) return ssg::set_error(SSG_EINVAL, "ArgumentError: null plan or input");
  auto one_of = [](int v, std::initializer_list<int> ok) {
    for (int o : ok) if (v == o) return true;
    return false;
  };
  if (!one_of(in5[0], {0, 1, 2, 3}) || !one_of(in5[1], {0, 1, 2, 4}) || !one_of(in5[2], {0, 1, 2}) ||
      !one_of(in5[3], {0, 1, 2, 4, 8}) || !one_of(in5[4], {0, 2}))
    return ssg::set_error(SSG_EINVAL, "ArgumentError: tuning value out of range (the spline solve [4] is the FP64 "
                                      "substitution, 0 or 2)");
  p->T.tune.syn_ct = in5[0];
  p->T.tune.syn_nwq = in5[1];
  p->T.tune.ana_mr = in5[2];
  p->T.tune.si_seg = in5[3];
  p->T.tune.sa_mode = in5[4];
  return SSG_OK;
}

int ssg_batch_destroy(ssg_batch* b) {
  if (!b) return SSG_OK;
  (void)hipDeviceSynchronize();
  release_graphs(b);
  delete b;
  return SSG_OK;
}

}  // extern "C"
