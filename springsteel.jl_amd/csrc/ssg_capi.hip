// ssg_capi.hip -- the C-ABI of include/springsteel_hip.h: plan life cycle,
// device residency of the tables, and the stage sequence of each transform.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "springsteel_hip.h"
#include "ssg_kernels.hpp"
#include "ssg_setup.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return fail(SSG_EHIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

}  // namespace

#include "ssg_plan.hpp"

namespace ssg {
int set_error(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace ssg

namespace {

template <typename T>
int dalloc(ssg_plan* p, T** out, size_t count) {
  *out = nullptr;
  if (count == 0) count = 1;
  void* ptr = nullptr;
  hipError_t e = hipMalloc(&ptr, count * sizeof(T));
  if (e != hipSuccess) return fail(SSG_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  p->allocs.push_back(ptr);
  *out = static_cast<T*>(ptr);
  return SSG_OK;
}

std::vector<float> to_float(const std::vector<double>& v) { return std::vector<float>(v.begin(), v.end()); }

template <typename T>
int upload(ssg_plan* p, const T** out, const std::vector<T>& v) {
  T* d = nullptr;
  int rc = dalloc(p, &d, v.size());
  if (rc) return rc;
  if (!v.empty()) HIPCHK(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  *out = d;
  return SSG_OK;
}

void free_plan(ssg_plan* p) {
  (void)hipDeviceSynchronize();
  for (void* a : p->allocs) (void)hipFree(a);
  for (auto& r : p->recs) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
  for (hipEvent_t e : p->pool) (void)hipEventDestroy(e);
  delete p;
}

hipEvent_t take_event(ssg_plan* p) {
  if (!p->pool.empty()) { hipEvent_t e = p->pool.back(); p->pool.pop_back(); return e; }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// stage timing: an event pair around one kernel launch on the call's stream;
// nothing waits on them until ssg_plan_stage_times
struct StageMark {
  ssg_plan* p;
  hipStream_t st;
  hipEvent_t a = nullptr, b = nullptr;
  const char* name;
  StageMark(ssg_plan* p_, hipStream_t st_, const char* n) : p(p_), st(st_), name(n) {
    if (!p->timing) return;
    a = take_event(p);
    b = take_event(p);
    if (a) (void)hipEventRecord(a, st);
  }
  ~StageMark() {
    if (!a || !b) return;
    (void)hipEventRecord(b, st);
    p->recs.push_back({name, a, b});
  }
};

int run_spectral(ssg_plan* p, const double* phys, double* spec, hipStream_t st) {
  const ssg::DevTables& T = p->T;
  const int g = T.geometry;
  const bool f32 = p->precision == SSG_FP32_REFINE;
  if (g == SSG_RL && f32) {
    { StageMark m(p, st, "ring_ana"); ssg::launch_ring_ana<float, double>(T, phys, p->rc32, st); }
    { StageMark m(p, st, "sb"); ssg::launch_sb_ring<float>(T, p->rc32, spec, st); }
  } else if (g == SSG_RLZ && f32) {
    { StageMark m(p, st, "cheb_fwd"); ssg::launch_cheb_fwd_pts<float>(T, phys, p->cm32, st); }
    { StageMark m(p, st, "ring_ana"); ssg::launch_ring_ana<float, float>(T, p->cm32, p->rc32, st); }
    { StageMark m(p, st, "sb"); ssg::launch_sb_ring<float>(T, p->rc32, spec, st); }
  } else if (g == SSG_RL) {
    { StageMark m(p, st, "ring_ana"); ssg::launch_ring_ana<double, double>(T, phys, p->rc, st); }
    { StageMark m(p, st, "sb"); ssg::launch_sb_ring<double>(T, p->rc, spec, st); }
  } else if (g == SSG_RZ) {
    { StageMark m(p, st, "cheb_fwd"); ssg::launch_cheb_fwd(T, phys, T.iDim, p->cb, st); }
    { StageMark m(p, st, "sb"); ssg::launch_sb(T, p->cb, spec, st); }
  } else {
    { StageMark m(p, st, "cheb_fwd"); ssg::launch_cheb_fwd_pts<double>(T, phys, p->cm, st); }
    { StageMark m(p, st, "ring_ana"); ssg::launch_ring_ana<double, double>(T, p->cm, p->rc, st); }
    { StageMark m(p, st, "sb"); ssg::launch_sb_ring<double>(T, p->rc, spec, st); }
  }
  HIPCHK(hipGetLastError());
  return SSG_OK;
}

int run_grid(ssg_plan* p, const double* spec, double* phys, hipStream_t st) {
  const ssg::DevTables& T = p->T;
  const int g = T.geometry;
  const bool f32 = p->precision == SSG_FP32_REFINE;
  if (g != SSG_RZ && f32) {
    { StageMark m(p, st, "sa_si"); ssg::launch_sa_si_ring<float>(T, spec, p->y, p->X32, st); }
    { StageMark m(p, st, g == SSG_RL ? "ring_syn" : "ring_syn_cheb"); ssg::launch_ring_syn<float>(T, p->X32, phys, st); }
  } else if (g == SSG_RL) {
    { StageMark m(p, st, "sa_si"); ssg::launch_sa_si_ring<double>(T, spec, p->y, p->X, st); }
    { StageMark m(p, st, "ring_syn"); ssg::launch_ring_syn<double>(T, p->X, phys, st); }
  } else if (g == SSG_RZ) {
    { StageMark m(p, st, "sa_si"); ssg::launch_sa_si(T, spec, p->y, p->cbd, st); }
    { StageMark m(p, st, "cheb_inv"); ssg::launch_cheb_inv(T, p->cbd, T.iDim, phys, st); }
  } else {
    { StageMark m(p, st, "sa_si"); ssg::launch_sa_si_ring<double>(T, spec, p->y, p->X, st); }
    { StageMark m(p, st, "ring_syn_cheb"); ssg::launch_ring_syn<double>(T, p->X, phys, st); }
  }
  HIPCHK(hipGetLastError());
  return SSG_OK;
}

// average duration per launch of every stage recorded since the last query;
// with nothing recorded the previous result stays
int collect_times(ssg_plan* p) {
  if (p->recs.empty()) return SSG_OK;
  p->last_ms.clear();
  p->last_names.clear();
  std::vector<int> count;
  for (auto& r : p->recs) {
    HIPCHK(hipEventSynchronize(r.b));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
    size_t i = 0;
    while (i < p->last_names.size() && std::strcmp(p->last_names[i], r.name) != 0) i++;
    if (i == p->last_names.size()) { p->last_names.push_back(r.name); p->last_ms.push_back(0.0); count.push_back(0); }
    p->last_ms[i] += ms;
    count[i]++;
  }
  for (size_t i = 0; i < p->last_ms.size(); i++) p->last_ms[i] /= count[i];
  for (auto& r : p->recs) { p->pool.push_back(r.a); p->pool.push_back(r.b); }
  p->recs.clear();
  return SSG_OK;
}

}  // namespace

extern "C" {

const char* ssg_last_error(void) { return g_err.c_str(); }

int ssg_plan_create(const ssg_desc* desc, ssg_plan** out) {
  if (!desc || !out) return fail(SSG_EINVAL, "ArgumentError: null descriptor or output");
  *out = nullptr;
  ssg_plan* p = new ssg_plan();
  std::string err;
  int rc = ssg::build_setup(*desc, p->S, err);
  if (rc != SSG_OK) { delete p; return fail(rc, err); }
  if (desc->precision != SSG_FP64 && desc->precision != SSG_FP32_REFINE) {
    delete p; return fail(SSG_EINVAL, "ArgumentError: precision must be SSG_FP64 or SSG_FP32_REFINE");
  }
  if (desc->precision == SSG_FP32_REFINE && desc->geometry == SSG_RZ) {
    delete p; return fail(SSG_ENOTSUP, "ENOTSUP: the FP32 form is built for the ring grids (RL, RLZ) only");
  }
  p->precision = desc->precision;
  hipError_t he = hipGetDevice(&p->device);
  if (he != hipSuccess) { delete p; return fail(SSG_EHIP, std::string("hipGetDevice: ") + hipGetErrorString(he)); }
  const ssg::Setup& S = p->S;
  ssg::DevTables& T = p->T;
  T.geometry = S.geometry; T.nvars = S.nvars; T.nc = S.nc; T.mubar = S.mubar; T.iDim = S.iDim; T.b_iDim = S.b_iDim;
  T.K = S.K; T.kDim = S.kDim; T.bk = S.b_kDim; T.nzb = S.nzb; T.jDim = S.jDim; T.b_jDim = S.b_jDim;
  T.maxMin = S.maxMin; T.npts = S.npts; T.nspec = S.nspec;
  T.nl = (S.geometry == SSG_RZ) ? 1 : 1 + 2 * S.K;
  T.nlines = long(S.nvars) * S.nzb * T.nl;
  T.ntr = S.nvars * S.nzb;
  T.ntb = (S.geometry == SSG_RLZ) ? S.nzb : S.nvars;
  T.nvb = T.ntr / T.ntb;
  T.Ktot = (S.geometry == SSG_RZ) ? 0 : S.phaseOff[S.iDim] + S.iDim;
  // tables
  std::vector<double> siw(S.siw[0]);
  siw.insert(siw.end(), S.siw[1].begin(), S.siw[1].end());
  siw.insert(siw.end(), S.siw[2].begin(), S.siw[2].end());
  std::vector<double> Lall(size_t(S.nvars) * S.maxMin * 4, 0.0), Linv(size_t(S.nvars) * S.maxMin, 0.0);
  std::vector<int> gint(size_t(S.nvars) * 8, 0);
  std::vector<double> gdbl(size_t(S.nvars) * 4, 0.0);
  for (int v = 0; v < S.nvars; v++) {
    const ssg::SplineVar& sv = S.spl[v];
    std::copy(sv.L.begin(), sv.L.end(), Lall.begin() + size_t(v) * S.maxMin * 4);
    std::copy(sv.Linv.begin(), sv.Linv.end(), Linv.begin() + size_t(v) * S.maxMin);
    const int gi[8] = {sv.g.Mdim, sv.g.Min, sv.g.rankL, sv.g.rankR, sv.g.lk, sv.g.rk, 0, 0};
    std::copy(gi, gi + 8, gint.begin() + size_t(v) * 8);
    const double gd[4] = {sv.g.aL, sv.g.bL, sv.g.aR, sv.g.bR};
    std::copy(gd, gd + 4, gdbl.begin() + size_t(v) * 4);
  }
  if ((rc = upload(p, &T.sbw, S.sbw)) || (rc = upload(p, &T.siw, siw)) || (rc = upload(p, &T.L, Lall)) ||
      (rc = upload(p, &T.Linv, Linv)) || (rc = upload(p, &T.gint, gint)) || (rc = upload(p, &T.gdbl, gdbl))) {
    free_plan(p); return rc;
  }
  if (S.geometry != SSG_RZ) {
    if ((rc = upload(p, &T.ringN, S.ringN)) || (rc = upload(p, &T.ringRi, S.ringRi)) ||
        (rc = upload(p, &T.ringOff, S.ringOff)) || (rc = upload(p, &T.coefOff, S.coefOff)) ||
        (rc = upload(p, &T.phaseOff, S.phaseOff)) || (rc = upload(p, &T.twOff, S.twOff)) ||
        (rc = upload(p, &T.kmax, S.kmax)) || (rc = upload(p, &T.phaseC, S.phaseC)) ||
        (rc = upload(p, &T.phaseS, S.phaseS)) || (rc = upload(p, &T.twid, S.twid))) {
      free_plan(p); return rc;
    }
  }
  if (S.geometry != SSG_RL) {
    std::vector<double> fwd, inv;
    for (int v = 0; v < S.nvars; v++) {
      fwd.insert(fwd.end(), S.cheb[v].fwd.begin(), S.cheb[v].fwd.end());
      inv.insert(inv.end(), S.cheb[v].inv0.begin(), S.cheb[v].inv0.end());
      inv.insert(inv.end(), S.cheb[v].inv1.begin(), S.cheb[v].inv1.end());
      inv.insert(inv.end(), S.cheb[v].inv2.begin(), S.cheb[v].inv2.end());
    }
    if ((rc = upload(p, &T.fwd, fwd)) || (rc = upload(p, &T.inv, inv))) { free_plan(p); return rc; }
    // RLZ: the synthesis epilogue copies its variable's three matrices into
    // LDS as they are laid out here, transposed and zero padded
    std::vector<double> invt;
    if (S.geometry == SSG_RLZ) {
      const int bk = S.b_kDim, kz = S.kDim, bkp = ssg::syn_bkp(bk), kzp = ssg::syn_kzp(kz);
      invt.assign(size_t(S.nvars) * 3 * bkp * kzp, 0.0);
      for (int v = 0; v < S.nvars; v++)
        for (int m = 0; m < 3; m++) {
          const std::vector<double>& M = m == 0 ? S.cheb[v].inv0 : m == 1 ? S.cheb[v].inv1 : S.cheb[v].inv2;
          for (int zb = 0; zb < bk; zb++)
            for (int z = 0; z < kz; z++)
              invt[((size_t(v) * 3 + m) * bkp + zb) * kzp + z] = M[size_t(z) * bk + zb];
        }
      if ((rc = upload(p, &T.invt, invt))) { free_plan(p); return rc; }
    }
    if (p->precision == SSG_FP32_REFINE &&
        ((rc = upload(p, &T.f32.fwd, to_float(fwd))) ||
         (!invt.empty() && (rc = upload(p, &T.f32.invt, to_float(invt)))))) {
      free_plan(p); return rc;
    }
  }
  if (p->precision == SSG_FP32_REFINE) {
    // float copies of the tables the FP32 kernels read (the banded factor is
    // kept in both: FP32 substitutions, FP64 residual)
    if ((rc = upload(p, &T.f32.sbw, to_float(S.sbw))) || (rc = upload(p, &T.f32.siw, to_float(siw))) ||
        (rc = upload(p, &T.f32.L, to_float(Lall))) || (rc = upload(p, &T.f32.Linv, to_float(Linv))) ||
        (rc = upload(p, &T.f32.phaseC, to_float(S.phaseC))) || (rc = upload(p, &T.f32.phaseS, to_float(S.phaseS))) ||
        (rc = upload(p, &T.f32.twid, to_float(S.twid)))) {
      free_plan(p); return rc;
    }
  }
  // intermediates
  const size_t ntr = size_t(S.nvars) * S.nzb;
  // y: one double per line (RZ) or two per wavenumber lane (RL / RLZ)
  const size_t ylines = std::max(size_t(T.nlines), 2 * ntr * size_t(S.K + 1));
  if ((rc = dalloc(p, &p->y, size_t(S.maxMin) * ylines))) { free_plan(p); return rc; }
  if (S.geometry == SSG_RZ) {
    if ((rc = dalloc(p, &p->cb, size_t(S.nvars) * S.b_kDim * S.iDim)) ||
        (rc = dalloc(p, &p->cbd, size_t(3) * S.nvars * S.b_kDim * S.iDim))) { free_plan(p); return rc; }
  } else {
    if (!ssg::ring_ana_supported(T)) {
      free_plan(p);
      return fail(SSG_ENOTSUP, "ENOTSUP: more than 48 Chebyshev modes or 64 Chebyshev points (RLZ) or 48 variables (RL) per ring transform");
    }
    if (p->precision == SSG_FP32_REFINE) {
      if ((rc = dalloc(p, &p->rc32, ntr * S.b_jDim)) || (rc = dalloc(p, &p->X32, 6 * ntr * size_t(T.Ktot)))) {
        free_plan(p); return rc;
      }
      if (S.geometry == SSG_RLZ && (rc = dalloc(p, &p->cm32, ntr * size_t(S.jDim)))) { free_plan(p); return rc; }
    } else {
      if ((rc = dalloc(p, &p->rc, ntr * S.b_jDim)) || (rc = dalloc(p, &p->X, 6 * ntr * size_t(T.Ktot)))) {
        free_plan(p); return rc;
      }
      if (S.geometry == SSG_RLZ && (rc = dalloc(p, &p->cm, ntr * size_t(S.jDim)))) { free_plan(p); return rc; }
    }
  }
  *out = p;
  return SSG_OK;
}

int ssg_plan_destroy(ssg_plan* plan) {
  if (!plan) return SSG_OK;
  free_plan(plan);
  return SSG_OK;
}

int ssg_plan_get_dims(const ssg_plan* p, ssg_dims* out) {
  if (!p || !out) return fail(SSG_EINVAL, "ArgumentError: null plan or output");
  const ssg::Setup& S = p->S;
  out->npts = S.npts; out->nspec = S.nspec; out->nvars = S.nvars; out->nslots = S.nslots;
  out->iDim = S.iDim; out->b_iDim = S.b_iDim; out->jDim = S.jDim; out->kDim = S.kDim; out->b_kDim = S.b_kDim;
  out->kDim_wn = S.K;
  return SSG_OK;
}

static int check_call(ssg_plan* p, const void* a, const void* b, int mem) {
  if (!p) return fail(SSG_EINVAL, "ArgumentError: null plan");
  if (!a || !b) return fail(SSG_EINVAL, "ArgumentError: null array pointer");
  if (mem != SSG_MEM_HOST && mem != SSG_MEM_DEVICE) return fail(SSG_EINVAL, "ArgumentError: mem must be SSG_MEM_HOST or SSG_MEM_DEVICE");
  HIPCHK(hipSetDevice(p->device));
  return SSG_OK;
}

static int ensure_staging(ssg_plan* p) {
  if (p->h_phys) return SSG_OK;
  const ssg::Setup& S = p->S;
  int rc;
  if ((rc = dalloc(p, &p->h_phys, size_t(S.npts) * S.nvars * S.nslots))) return rc;
  if ((rc = dalloc(p, &p->h_spec, size_t(S.nspec) * S.nvars))) return rc;
  return SSG_OK;
}

int ssg_spectral_transform(ssg_plan* p, const double* physical, double* spectral, int mem, void* stream) {
  int rc = check_call(p, physical, spectral, mem);
  if (rc) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const ssg::Setup& S = p->S;
  if (mem == SSG_MEM_DEVICE) return run_spectral(p, physical, spectral, st);
  if ((rc = ensure_staging(p))) return rc;
  const size_t nin = size_t(S.npts) * S.nvars;   // slot 1 only
  HIPCHK(hipMemcpyAsync(p->h_phys, physical, nin * sizeof(double), hipMemcpyHostToDevice, st));
  if ((rc = run_spectral(p, p->h_phys, p->h_spec, st))) return rc;
  HIPCHK(hipMemcpyAsync(spectral, p->h_spec, size_t(S.nspec) * S.nvars * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return SSG_OK;
}

int ssg_grid_transform(ssg_plan* p, const double* spectral, double* physical, int mem, void* stream) {
  int rc = check_call(p, spectral, physical, mem);
  if (rc) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const ssg::Setup& S = p->S;
  if (mem == SSG_MEM_DEVICE) return run_grid(p, spectral, physical, st);
  if ((rc = ensure_staging(p))) return rc;
  HIPCHK(hipMemcpyAsync(p->h_spec, spectral, size_t(S.nspec) * S.nvars * sizeof(double), hipMemcpyHostToDevice, st));
  if ((rc = run_grid(p, p->h_spec, p->h_phys, st))) return rc;
  HIPCHK(hipMemcpyAsync(physical, p->h_phys, size_t(S.npts) * S.nvars * S.nslots * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return SSG_OK;
}

int ssg_plan_set_timing(ssg_plan* p, int enable) {
  if (!p) return fail(SSG_EINVAL, "ArgumentError: null plan");
  p->timing = enable != 0;
  int rc = collect_times(p);   // retire whatever was recorded before
  p->last_ms.clear();
  p->last_names.clear();
  return rc;
}

int ssg_plan_stage_times(ssg_plan* p, const char** names, double* ms, int n) {
  if (!p) return fail(SSG_EINVAL, "ArgumentError: null plan");
  int rc = collect_times(p);
  if (rc) return rc;
  const int m = int(p->last_ms.size());
  for (int i = 0; i < m && i < n; i++) {
    if (names) names[i] = p->last_names[i];
    if (ms) ms[i] = p->last_ms[i];
  }
  return m;
}

}  // extern "C"
