// ssg_capi.hip -- the C-ABI of include/springsteel_hip.h: plan life cycle,
// device residency of the tables, and the stage sequence of each transform.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
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

// bytes of the RLZ chirp-z synthesis output Y[v][5 ntb][jDim]
size_t ysyn_bytes(const ssg_plan* p) {
  const size_t es = p->precision == SSG_FP32_REFINE ? sizeof(float) : sizeof(double);
  return es * size_t(p->S.nvars) * 5 * size_t(p->S.nzb) * size_t(p->S.jDim);
}

std::vector<float> to_float(const std::vector<double>& v) { return std::vector<float>(v.begin(), v.end()); }

// (a[0], b[0], a[1], b[1], ...)
std::vector<double> interleave(const std::vector<double>& a, const std::vector<double>& b) {
  std::vector<double> out(2 * a.size());
  for (size_t e = 0; e < a.size(); e++) { out[2 * e] = a[e]; out[2 * e + 1] = b[e]; }
  return out;
}

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
  if (p->reg) (void)hipFree(p->reg);
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
  StageMark(ssg_plan* p_, hipStream_t st_, const char* n, bool on = true) : p(p_), st(st_), name(n) {
    if (!p->timing || !on) return;
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

// physical: double, or float for the FP32 form (whose forward path runs the
// FP64 kernels on it)
int run_spectral(ssg_plan* p, const void* physical, double* spec, hipStream_t st) {
  const ssg::DevTables& T = p->T;
  const int g = T.geometry;
  const bool f32 = p->precision == SSG_FP32_REFINE;
  const double* phys = static_cast<const double*>(physical);
  const float* physf = static_cast<const float*>(physical);
  // ring analysis: chirp-z transforms, or the direct MFMA transform for plans with a
  // ring beyond the largest chirp-z transform
  auto ana = [&](auto* in) {
    StageMark m(p, st, "ring_ana");
    if (p->fft_ana) ssg::launch_ana_fft(T, p->ana_blk, p->ana_nblocks, in, p->rc, st);
    else ssg::launch_ring_ana<double>(T, in, p->rc, st);
  };
  if (g == SSG_RL && f32) {
    ana(physf);
    { StageMark m(p, st, "sb"); ssg::launch_sb_ring(T, p->rc, spec, st); }
  } else if (g == SSG_RLZ && f32) {
    { StageMark m(p, st, "cheb_fwd"); ssg::launch_cheb_fwd_pts<float>(T, physf, p->cm, st); }
    ana(static_cast<const double*>(p->cm));
    { StageMark m(p, st, "sb"); ssg::launch_sb_ring(T, p->rc, spec, st); }
  } else if (g == SSG_RL) {
    ana(phys);
    { StageMark m(p, st, "sb"); ssg::launch_sb_ring(T, p->rc, spec, st); }
  } else if (g == SSG_RZ) {
    { StageMark m(p, st, "cheb_fwd"); ssg::launch_cheb_fwd(T, phys, T.iDim, p->cb, st); }
    { StageMark m(p, st, "sb"); ssg::launch_sb(T, p->cb, spec, st); }
  } else {
    { StageMark m(p, st, "cheb_fwd"); ssg::launch_cheb_fwd_pts<double>(T, phys, p->cm, st); }
    ana(static_cast<const double*>(p->cm));
    { StageMark m(p, st, "sb"); ssg::launch_sb_ring(T, p->rc, spec, st); }
  }
  HIPCHK(hipGetLastError());
  return SSG_OK;
}

// The intermediates of gridTransform: the plan's own buffers, or pieces of the
// library's static scratch array (ssg_grid_transform_noalloc).
struct GridScratch {
  double* y; double* X; float* X32; double* cbd; double* sbt;
  void* ysyn;   // RLZ chirp-z synthesis output (null: the direct synthesis)
  bool timed;   // stage events (ssg_plan_set_timing); off where nothing may be created
};
GridScratch plan_scratch(ssg_plan* p) { return GridScratch{p->y, p->X, p->X32, p->cbd, p->sbt, p->ysyn, true}; }

// Ring synthesis from X: chirp-z transforms (then, for RLZ, the inverse
// Chebyshev stage from Y), or the direct MFMA synthesis for plans with a ring
// beyond the largest chirp-z transform.
template <typename R>
void run_synthesis(ssg_plan* p, hipStream_t st, const GridScratch& w, const R* X, R* phys) {
  const ssg::DevTables& T = p->T;
  const bool rl = T.geometry == SSG_RL;
  if (p->fft_syn && (rl || w.ysyn)) {
    if (rl) {
      StageMark m(p, st, "ring_syn", w.timed);
      ssg::launch_syn_fft<R, R>(T, p->fft_blk, p->fft_nblocks, X, phys, st);
    } else {
      R* Y = static_cast<R*>(w.ysyn);
      { StageMark m(p, st, "ring_syn_fft", w.timed);
        ssg::launch_syn_fft<R, R>(T, p->fft_blk, p->fft_nblocks, X, Y, st); }
      { StageMark m(p, st, "syn_cheb", w.timed); ssg::launch_syn_cheb<R>(T, Y, phys, st); }
    }
  } else {
    StageMark m(p, st, rl ? "ring_syn" : "ring_syn_cheb", w.timed);
    ssg::launch_ring_syn<R>(T, X, phys, st);
  }
}

// physical: double, or float for the FP32 form
int run_grid(ssg_plan* p, const double* spec, void* physical, hipStream_t st, const GridScratch& w) {
  double* phys = static_cast<double*>(physical);
  const ssg::DevTables& T = p->T;
  const int g = T.geometry;
  const bool f32 = p->precision == SSG_FP32_REFINE;
  if (T.ahat) {
    // R3X background coefficients: the solve runs on btilde = b - pq ahat
    StageMark m(p, st, "ahat_btilde", w.timed);
    ssg::launch_ahat_btilde(T, spec, w.sbt, st);
    spec = w.sbt;
  }
  if (g != SSG_RZ && f32) {
    { StageMark m(p, st, "sa_si", w.timed); ssg::launch_sa_si_ring<float>(T, spec, w.y, w.X32, st); }
    run_synthesis<float>(p, st, w, w.X32, static_cast<float*>(physical));
  } else if (g == SSG_RZ) {
    { StageMark m(p, st, "sa_si", w.timed); ssg::launch_sa_si(T, spec, w.y, w.cbd, st); }
    { StageMark m(p, st, "cheb_inv", w.timed); ssg::launch_cheb_inv(T, w.cbd, T.iDim, phys, st); }
  } else {
    { StageMark m(p, st, "sa_si", w.timed); ssg::launch_sa_si_ring<double>(T, spec, w.y, w.X, st); }
    run_synthesis<double>(p, st, w, w.X, phys);
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
  T.sph = S.spherical ? 1 : 0;
  // tables
  std::vector<double> siw(S.siw[0]);
  siw.insert(siw.end(), S.siw[1].begin(), S.siw[1].end());
  siw.insert(siw.end(), S.siw[2].begin(), S.siw[2].end());
  std::vector<double> Lall(size_t(S.nvars) * S.maxMin * 4, 0.0), Linv(size_t(S.nvars) * S.maxMin, 0.0);
  std::vector<int> gint(size_t(S.nvars) * 8, 0);
  std::vector<double> gdbl(size_t(S.nvars) * 4, 0.0);
  std::vector<double> pqb;
  for (int v = 0; v < S.nvars; v++) {
    const ssg::SplineVar& sv = S.spl[v];
    std::copy(sv.L.begin(), sv.L.end(), Lall.begin() + size_t(v) * S.maxMin * 4);
    std::copy(sv.Linv.begin(), sv.Linv.end(), Linv.begin() + size_t(v) * S.maxMin);
    pqb.insert(pqb.end(), sv.pq.begin(), sv.pq.end());
    const int gi[8] = {sv.g.Mdim, sv.g.Min, sv.g.rankL, sv.g.rankR, sv.g.lk, sv.g.rk, sv.r3x ? 1 : 0, 0};
    std::copy(gi, gi + 8, gint.begin() + size_t(v) * 8);
    const double gd[4] = {sv.g.aL, sv.g.bL, sv.g.aR, sv.g.bR};
    std::copy(gd, gd + 4, gdbl.begin() + size_t(v) * 4);
  }
  std::vector<int> mins(S.nvars), ranks(S.nvars);
  for (int v = 0; v < S.nvars; v++) { mins[v] = S.spl[v].g.Min; ranks[v] = S.spl[v].g.rankL; }
  const ssg::SasiTables stab = ssg::sasi_tables(Lall, Linv, mins, ranks, S.maxMin);
  if ((rc = upload(p, &T.sbw, S.sbw)) || (rc = upload(p, &T.siw, siw)) || (rc = upload(p, &T.L, Lall)) ||
      (rc = upload(p, &T.Linv, Linv)) || (rc = upload(p, &T.gint, gint)) || (rc = upload(p, &T.gdbl, gdbl)) ||
      (rc = upload(p, &T.pqb, pqb)) || (rc = upload(p, &T.sasiA, stab.maps)) ||
      (rc = upload(p, &T.sasiR, stab.rows))) {
    free_plan(p); return rc;
  }
  if (S.geometry != SSG_RZ) {
    if ((rc = upload(p, &T.ringN, S.ringN)) || (rc = upload(p, &T.ringRi, S.ringRi)) ||
        (rc = upload(p, &T.ringOff, S.ringOff)) || (rc = upload(p, &T.coefOff, S.coefOff)) ||
        (rc = upload(p, &T.phaseOff, S.phaseOff)) || (rc = upload(p, &T.twOff, S.twOff)) ||
        (rc = upload(p, &T.kmax, S.kmax)) || (rc = upload(p, &T.phaseC, S.phaseC)) ||
        (rc = upload(p, &T.phaseS, S.phaseS)) || (rc = upload(p, &T.twid, S.twid)) ||
        (rc = upload(p, &T.phaseCS, interleave(S.phaseC, S.phaseS)))) {
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
    if (p->precision == SSG_FP32_REFINE && !invt.empty() && (rc = upload(p, &T.f32.invt, to_float(invt)))) {
      free_plan(p); return rc;
    }
  }
  if (p->precision == SSG_FP32_REFINE && S.geometry != SSG_RZ) {
    // float copy of the ring twiddles, which the FP32 synthesis reads (every
    // other stage of the FP32 form runs the FP64 kernels on the FP64 tables)
    if ((rc = upload(p, &T.f32.twid, to_float(S.twid)))) { free_plan(p); return rc; }
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
    if ((rc = dalloc(p, &p->rc, ntr * S.b_jDim))) { free_plan(p); return rc; }
    // chirp-z ring transforms: every ring within the largest transform.  The analysis
    // runs in FP64 in both forms; the synthesis in the form's arithmetic (the FP32
    // form's transforms are FP32, on float copies of the tables)
    const bool fits = std::all_of(S.fftL.begin(), S.fftL.end(), [](int L) { return L > 0; });
    p->fft_ana = fits;
    p->fft_syn = fits;
    if (fits) {
      const ssg::FftSynBlocks A = ssg::fft_blocks(S.fftL, S.iDim, ssg::ana_fft_pairs(S.nvars, S.nzb, S.geometry == SSG_RLZ));
      p->ana_nblocks = A.nblocks;
      const int* ablk = nullptr;
      if ((rc = upload(p, &T.fftL, S.fftL)) || (rc = upload(p, &T.fftHOff, S.fftHOff)) ||
          (rc = upload(p, &T.fftChirp, S.fftChirp)) || (rc = upload(p, &T.fftH, S.fftH)) ||
          (rc = upload(p, &T.fftTw, S.fftTw)) || (rc = upload(p, &ablk, A.blk))) { free_plan(p); return rc; }
      p->ana_blk = const_cast<int*>(ablk);
      if (p->precision == SSG_FP32_REFINE &&
          ((rc = upload(p, &T.f32.fftChirp, to_float(S.fftChirp))) || (rc = upload(p, &T.f32.fftH, to_float(S.fftH))) ||
           (rc = upload(p, &T.f32.fftTw, to_float(S.fftTw))))) { free_plan(p); return rc; }
    }
    if (p->fft_syn) {
      const ssg::FftSynBlocks B = ssg::fft_blocks(S.fftL, S.iDim, ssg::syn_fft_pairs(S.nvars, S.nzb));
      p->fft_nblocks = B.nblocks;
      const int* blk = nullptr;
      if ((rc = upload(p, &blk, B.blk))) { free_plan(p); return rc; }
      p->fft_blk = const_cast<int*>(blk);
      if (S.geometry == SSG_RLZ) {
        unsigned char* y = nullptr;
        if ((rc = dalloc(p, &y, ysyn_bytes(p)))) { free_plan(p); return rc; }
        p->ysyn = y;
      }
    }
    if (S.geometry == SSG_RLZ && (rc = dalloc(p, &p->cm, ntr * size_t(S.jDim)))) { free_plan(p); return rc; }
    if (p->precision == SSG_FP32_REFINE) rc = dalloc(p, &p->X32, 6 * ntr * size_t(T.Ktot));
    else rc = dalloc(p, &p->X, 6 * ntr * size_t(T.Ktot));
    if (rc) { free_plan(p); return rc; }
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

// the physical element type of a call must be the plan's
static int check_form(const ssg_plan* p, bool f32_call) {
  const bool f32 = p->precision == SSG_FP32_REFINE;
  if (f32 == f32_call) return SSG_OK;
  return fail(SSG_EINVAL, f32 ? "ArgumentError: an SSG_FP32_REFINE plan stores physical in Float32: use "
                                "ssg_spectral_transform_f32 / ssg_grid_transform_f32"
                              : "ArgumentError: an SSG_FP64 plan stores physical in Float64: use "
                                "ssg_spectral_transform / ssg_grid_transform");
}

static size_t phys_esize(const ssg_plan* p) {
  return p->precision == SSG_FP32_REFINE ? sizeof(float) : sizeof(double);
}

static int ensure_staging(ssg_plan* p) {
  if (p->h_phys) return SSG_OK;
  const ssg::Setup& S = p->S;
  int rc;
  unsigned char* ph = nullptr;
  if ((rc = dalloc(p, &ph, size_t(S.npts) * S.nvars * S.nslots * phys_esize(p)))) return rc;
  p->h_phys = ph;
  if ((rc = dalloc(p, &p->h_spec, size_t(S.nspec) * S.nvars))) return rc;
  return SSG_OK;
}

static int spectral_call(ssg_plan* p, const void* physical, double* spectral, int mem, void* stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  const ssg::Setup& S = p->S;
  if (mem == SSG_MEM_DEVICE) return run_spectral(p, physical, spectral, st);
  int rc;
  if ((rc = ensure_staging(p))) return rc;
  const size_t nin = size_t(S.npts) * S.nvars;   // slot 1 only
  HIPCHK(hipMemcpyAsync(p->h_phys, physical, nin * phys_esize(p), hipMemcpyHostToDevice, st));
  if ((rc = run_spectral(p, p->h_phys, p->h_spec, st))) return rc;
  HIPCHK(hipMemcpyAsync(spectral, p->h_spec, size_t(S.nspec) * S.nvars * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return SSG_OK;
}

static int grid_call(ssg_plan* p, const double* spectral, void* physical, int mem, void* stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  const ssg::Setup& S = p->S;
  if (mem == SSG_MEM_DEVICE) return run_grid(p, spectral, physical, st, plan_scratch(p));
  int rc;
  if ((rc = ensure_staging(p))) return rc;
  HIPCHK(hipMemcpyAsync(p->h_spec, spectral, size_t(S.nspec) * S.nvars * sizeof(double), hipMemcpyHostToDevice, st));
  if ((rc = run_grid(p, p->h_spec, p->h_phys, st, plan_scratch(p)))) return rc;
  HIPCHK(hipMemcpyAsync(physical, p->h_phys, size_t(S.npts) * S.nvars * S.nslots * phys_esize(p),
                        hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return SSG_OK;
}

int ssg_spectral_transform(ssg_plan* p, const double* physical, double* spectral, int mem, void* stream) {
  int rc = check_call(p, physical, spectral, mem);
  if (rc || (rc = check_form(p, false))) return rc;
  return spectral_call(p, physical, spectral, mem, stream);
}

int ssg_grid_transform(ssg_plan* p, const double* spectral, double* physical, int mem, void* stream) {
  int rc = check_call(p, spectral, physical, mem);
  if (rc || (rc = check_form(p, false))) return rc;
  return grid_call(p, spectral, physical, mem, stream);
}

int ssg_spectral_transform_f32(ssg_plan* p, const float* physical, double* spectral, int mem, void* stream) {
  int rc = check_call(p, physical, spectral, mem);
  if (rc || (rc = check_form(p, true))) return rc;
  return spectral_call(p, physical, spectral, mem, stream);
}

int ssg_grid_transform_f32(ssg_plan* p, const double* spectral, float* physical, int mem, void* stream) {
  int rc = check_call(p, spectral, physical, mem);
  if (rc || (rc = check_form(p, true))) return rc;
  return grid_call(p, spectral, physical, mem, stream);
}

// ----------------------------------------------------------------------------
// gridTransform without allocation: the intermediates in a static device array
// ----------------------------------------------------------------------------
namespace {

// The library's scratch for ssg_grid_transform_noalloc.  Every GPU the library
// runs on holds its own copy for the life of the process.  In host code the
// name `scratch` is only a stand-in for the symbol: its device address comes
// from hipGetSymbolAddress, once per device, and only that address is used.
constexpr size_t kScratchBytes = size_t(SSG_NOALLOC_SCRATCH_BYTES);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// device address of this device's scratch array (the device is current)
int scratch_address(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(SSG_EINVAL, "ArgumentError: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return fail(SSG_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  *out = static_cast<unsigned char*>(g_scratch_ptr[dev]);
  return SSG_OK;
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// Carves gridTransform's intermediates of plan p out of base (base may be
// null to size them); returns the bytes used.  The pieces are those
// ssg_plan_create allocates for the plan.
size_t carve_grid_scratch(const ssg_plan* p, unsigned char* base, GridScratch& w) {
  const ssg::DevTables& T = p->T;
  const ssg::Setup& S = p->S;
  const size_t ntr = size_t(S.nvars) * S.nzb;
  const size_t ylines = std::max(size_t(T.nlines), 2 * ntr * size_t(S.K + 1));
  size_t off = 0;
  auto take = [&](size_t bytes) { unsigned char* q = base ? base + off : nullptr; off += align256(bytes); return q; };
  w = GridScratch{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, false};
  w.y = reinterpret_cast<double*>(take(sizeof(double) * size_t(S.maxMin) * ylines));
  if (S.geometry == SSG_RZ) {
    w.cbd = reinterpret_cast<double*>(take(sizeof(double) * 3 * size_t(S.nvars) * S.b_kDim * S.iDim));
  } else if (p->precision == SSG_FP32_REFINE) {
    w.X32 = reinterpret_cast<float*>(take(sizeof(float) * 6 * ntr * size_t(T.Ktot)));
  } else {
    w.X = reinterpret_cast<double*>(take(sizeof(double) * 6 * ntr * size_t(T.Ktot)));
  }
  if (T.ahat) w.sbt = reinterpret_cast<double*>(take(sizeof(double) * size_t(S.nspec) * S.nvars));
  if (p->fft_syn && S.geometry == SSG_RLZ) w.ysyn = take(ysyn_bytes(p));
  return off;
}

// zero fill of n16 16-byte words (a kernel, not a memset node, so that a
// captured graph replays it correctly)
__global__ void k_zero16(uint4* __restrict__ dst, size_t n16) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n16; i += size_t(gridDim.x) * blockDim.x)
    dst[i] = make_uint4(0u, 0u, 0u, 0u);
}

}  // namespace

extern "C" int ssg_grid_transform_noalloc(ssg_plan* p, const double* spectral, double* physical, void* stream) {
  int rc = check_call(p, spectral, physical, SSG_MEM_DEVICE);
  if (rc || (rc = check_form(p, false))) return rc;
  GridScratch w;
  const size_t need = carve_grid_scratch(p, nullptr, w);
  if (need > kScratchBytes)
    return fail(SSG_ENOTSUP, "ENOTSUP: ssg_grid_transform_noalloc needs " + std::to_string(need) +
                             " bytes of scratch for this plan, more than the library's static " +
                             std::to_string(kScratchBytes) + " (use ssg_grid_transform)");
  unsigned char* base = nullptr;
  if ((rc = scratch_address(p->device, &base))) return rc;
  carve_grid_scratch(p, base, w);
  hipStream_t st = static_cast<hipStream_t>(stream);
  // the array keeps the previous call's contents: start every call from zeros
  const size_t n16 = need / 16;
  const unsigned nb = unsigned(std::min<size_t>((n16 + 255) / 256, 2048));
  hipLaunchKernelGGL(k_zero16, dim3(nb), dim3(256), 0, st, reinterpret_cast<uint4*>(base), n16);
  return run_grid(p, spectral, physical, st, w);
}

// ----------------------------------------------------------------------------
// regularGridTransform (src/transforms_cylindrical.jl:929-1020 RL, 1168-1287
// RLZ; src/transforms_cartesian.jl:946-1001 RZ)
// ----------------------------------------------------------------------------
namespace {

// Rows of the Chebyshev series evaluation at z (_cheb_eval_pts!, _cheb_dz_pts!,
// _cheb_dzz_pts!, src/transforms_cylindrical.jl:1030-1094): value m = 0, d/dz
// m = 1, d2/dz2 m = 2 of sum_k r[k] a[k], a = CAtransform!(b).  Evaluated in
// FP64 with the reference's expressions (t = acos(xi), sin / cos of the FP64
// product k * t), so that its rounding is the reference's: at a clamped angle
// (the z end points, t ~ 1.4e-5 from 0 or pi) the d2/dz2 formula divides a
// difference of two O(t) terms by sin(t)^3, and the rounding of k * t alone
// moves its value by up to 1e-2 relative; a more accurate evaluation would
// return a different number there than the reference does.
void cheb_point_row(const ssg::Setup& S, int m, double z, std::vector<double>& r) {
  const int N = S.kDim;
  r.assign(N, 0.0);
  const double scale = -0.5 * (S.d.kMax - S.d.kMin), offset = 0.5 * (S.d.kMin + S.d.kMax);
  const double eps = m == 0 ? 1e-14 : 1e-10;
  double xi = (z - offset) / scale;
  xi = xi < -1.0 + eps ? -1.0 + eps : (xi > 1.0 - eps ? 1.0 - eps : xi);
  const double t = std::acos(xi), st = std::sin(t), ct = std::cos(t);
  for (int k = 0; k < N; k++) {
    const double c = (k == 0 || k == N - 1) ? 1.0 : 2.0;   // DCT-I weights of the series
    const double kt = double(k) * t;
    if (m == 0) r[k] = c * std::cos(kt);
    else {
      const double dfdt = -c * double(k) * std::sin(kt);
      if (m == 1) r[k] = dfdt / (-scale * st);
      else {
        const double d2 = -c * double(k * k) * std::cos(kt);
        r[k] = (d2 * st - dfdt * ct) / (scale * scale * (st * st * st));
      }
    }
  }
}

int grow_reg(ssg_plan* p, size_t bytes) {
  if (p->reg_bytes >= bytes) return SSG_OK;
  HIPCHK(hipDeviceSynchronize());
  if (p->reg) { (void)hipFree(p->reg); p->reg = nullptr; p->reg_bytes = 0; }
  hipError_t e = hipMalloc(&p->reg, bytes);
  if (e != hipSuccess) { p->reg = nullptr; return fail(SSG_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e)); }
  p->reg_bytes = bytes;
  return SSG_OK;
}

}  // namespace

int ssg_regular_transform(ssg_plan* p, const double* spectral, const double* ip, int ni, const double* jp, int nj,
                          const double* kp, int nk, double* out, int mem, void* stream) {
  int rc = check_call(p, spectral, out, mem);
  if (rc) return rc;
  if (p->T.ahat) return fail(SSG_ENOTSUP, "ENOTSUP: regularGridTransform with R3X background coefficients set");
  const ssg::Setup& S = p->S;
  const ssg::DevTables& T = p->T;
  const int geo = S.geometry;
  const bool ring = geo != SSG_RZ, cheb = geo != SSG_RL;
  if (!ip || ni < 1) return fail(SSG_EINVAL, "ArgumentError: at least one i point");
  if (ring && (!jp || nj < 1)) return fail(SSG_EINVAL, "ArgumentError: at least one lambda point");
  if (cheb && (!kp || nk < 1)) return fail(SSG_EINVAL, "ArgumentError: at least one z point");
  for (int i = 0; i < ni; i++)
    if (!(ip[i] >= S.d.iMin && ip[i] <= S.d.iMax)) return fail(SSG_EDOMAIN, "DomainError: x outside spline domain");
  const int NJ = ring ? nj : 1, NK = cheb ? nk : 1;
  const long npr = long(ni) * NJ * NK;
  const int nv = S.nvars, bk = S.b_kDim, nks = 2 * S.K + 1;
  const long nline = ring ? long(T.ntr) * nks : T.nlines;
  hipStream_t st = static_cast<hipStream_t>(stream);

  // host tables: spline taps at the radii, Chebyshev point rows
  const double DX = (S.d.iMax - S.d.iMin) / S.nc, DXr = 1.0 / DX;
  std::vector<int> pxm(ni);
  std::vector<double> pw(size_t(3) * ni * 4, 0.0);
  for (int i = 0; i < ni; i++) {
    const int xm = int(std::ceil((ip[i] - S.d.iMin - 2.0 * DX) * DXr));
    pxm[i] = xm;
    for (int t = 0; t < 4; t++) {
      const int m = xm + t;
      if (m < -1 || m > S.nc + 1) continue;
      for (int d = 0; d < 3; d++) pw[(size_t(d) * ni + i) * 4 + t] = ssg::bspline(S.d.iMin, DX, DXr, m, ip[i], d);
    }
  }
  std::vector<double> E;      // [m][v][NK][bk]
  if (cheb) {
    E.assign(size_t(3) * nv * NK * bk, 0.0);
    std::vector<double> r;
    for (int m = 0; m < 3; m++)
      for (int z = 0; z < NK; z++) {
        cheb_point_row(S, m, kp[z], r);
        for (int v = 0; v < nv; v++) {
          const std::vector<double>& A = S.cheb[v].ca;   // [k][zb]
          for (int zb = 0; zb < bk; zb++) {
            long double acc = 0.0L;
            for (int k = 0; k < S.kDim; k++) acc += r[k] * (long double)A[size_t(k) * bk + zb];
            E[((size_t(m) * nv + v) * NK + z) * bk + zb] = double(acc);
          }
        }
      }
  }

  // device scratch: tables, S, the Fourier results (RLZ) and host staging
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t bS = al(sizeof(double) * 3 * size_t(ni) * nline);
  const size_t bF = geo == SSG_RLZ ? al(sizeof(double) * 5 * size_t(T.ntr) * ni * NJ) : 0;
  // Fourier rows [dl][ks][NJ] (built on the device from the angles) and the angles
  const size_t bT = ring ? al(sizeof(double) * 3 * size_t(nks) * NJ) : 0, bJ = ring ? al(sizeof(double) * NJ) : 0;
  const size_t bE = al(sizeof(double) * E.size());
  const size_t bW = al(sizeof(double) * pw.size()), bX = al(sizeof(int) * pxm.size());
  const size_t bIn = mem == SSG_MEM_HOST ? al(sizeof(double) * size_t(S.nspec) * nv) : 0;
  const size_t bOut = mem == SSG_MEM_HOST ? al(sizeof(double) * size_t(npr) * nv * S.nslots) : 0;
  if ((rc = grow_reg(p, bS + bF + bT + bJ + bE + bW + bX + bIn + bOut))) return rc;
  char* base = static_cast<char*>(p->reg);
  double* dS = reinterpret_cast<double*>(base); base += bS;
  double* dF = reinterpret_cast<double*>(base); base += bF;
  double* dT = reinterpret_cast<double*>(base); base += bT;
  double* dJ = reinterpret_cast<double*>(base); base += bJ;
  double* dE = reinterpret_cast<double*>(base); base += bE;
  double* dW = reinterpret_cast<double*>(base); base += bW;
  int* dX = reinterpret_cast<int*>(base); base += bX;
  const double* spec = spectral;
  double* dout = out;
  if (mem == SSG_MEM_HOST) {
    double* din = reinterpret_cast<double*>(base); base += bIn;
    dout = reinterpret_cast<double*>(base);
    HIPCHK(hipMemcpyAsync(din, spectral, sizeof(double) * size_t(S.nspec) * nv, hipMemcpyHostToDevice, st));
    spec = din;
  }
  if (ring) {
    HIPCHK(hipMemcpyAsync(dJ, jp, sizeof(double) * NJ, hipMemcpyHostToDevice, st));
    ssg::launch_trig_table(dJ, NJ, S.K, dT, st);
  }
  if (!E.empty()) HIPCHK(hipMemcpyAsync(dE, E.data(), sizeof(double) * E.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dW, pw.data(), sizeof(double) * pw.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dX, pxm.data(), sizeof(int) * pxm.size(), hipMemcpyHostToDevice, st));

  {
    StageMark m(p, st, "reg_sa");
    if (ring) ssg::launch_sa_ring_only(T, spec, p->y, st);
    else ssg::launch_sa_lines(T, spec, p->y, st);
  }
  { StageMark m(p, st, "reg_si"); ssg::launch_si_pts(T, ring, p->y, ni, dX, dW, dS, st); }
  const long sd = long(ni) * nline;                  // order stride of S
  if (geo == SSG_RL) {
    StageMark m(p, st, "reg_fourier");
    // (d, dl) -> slot: f, dr f, dr2 f, dl f, dl2 f
    const int fam[5][3] = {{0, 0, 0}, {1, 0, 1}, {2, 0, 2}, {0, 1, 3}, {0, 2, 4}};
    // one product per family over every (radius, variable) row: row m = ip * nv + v
    // of S is contiguous (nline = nv * nks), the trig rows are shared by all of them
    for (auto& f : fam) {
      ssg::GemmDesc g{};
      g.A = dS + f[0] * sd; g.ai = nks; g.ak = 1;
      g.B = dT + size_t(f[1]) * nks * NJ; g.bk = NJ; g.bj = 1;
      g.C = dout + long(f[2]) * nv * npr; g.msplit = nv; g.ci = npr; g.ci2 = NJ; g.cj = 1;
      g.M = ni * nv; g.N = NJ; g.K = nks; g.batch = 1;
      ssg::launch_gemm_f64(g, st);
    }
  } else if (geo == SSG_RLZ) {
    const long nij = long(ni) * NJ;
    {
      StageMark m(p, st, "reg_fourier");
      const int fam[5][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 0}, {2, 0}};   // (d, dl) of family f
      // row m = ip * ntr + t of S is contiguous (nline = ntr * nks)
      for (int f = 0; f < 5; f++) {
        ssg::GemmDesc g{};
        g.A = dS + fam[f][0] * sd; g.ai = nks; g.ak = 1;
        g.B = dT + size_t(fam[f][1]) * nks * NJ; g.bk = NJ; g.bj = 1;
        g.C = dF + long(f) * T.ntr * nij; g.msplit = T.ntr; g.ci = nij; g.ci2 = NJ; g.cj = 1;
        g.M = int(long(ni) * T.ntr); g.N = NJ; g.K = nks; g.batch = 1;
        ssg::launch_gemm_f64(g, st);
      }
    }
    StageMark m(p, st, "reg_cheb");
    // (family, Chebyshev row m, slot): f -> 1, 6, 7; dl f -> 4; dl2 f -> 5; dr f -> 2; dr2 f -> 3
    const int sl[7][3] = {{0, 0, 0}, {0, 1, 5}, {0, 2, 6}, {1, 0, 3}, {2, 0, 4}, {3, 0, 1}, {4, 0, 2}};
    for (auto& s : sl) {
      ssg::GemmDesc g{};
      g.A = dF + long(s[0]) * T.ntr * nij; g.ai = 1; g.ak = nij; g.sA = long(bk) * nij;
      g.B = dE + size_t(s[1]) * nv * NK * bk; g.bk = 1; g.bj = bk; g.sB = long(NK) * bk;
      g.C = dout + long(s[2]) * nv * npr; g.ci = NK; g.cj = 1; g.sC = npr;
      g.M = int(nij); g.N = NK; g.K = bk; g.batch = nv;
      ssg::launch_gemm_f64(g, st);
    }
  } else {
    StageMark m(p, st, "reg_cheb");
    // (dr, dk) -> slot (_rz_slot): f, dx f, dx2 f, dz f, dz2 f
    const int sl[5][3] = {{0, 0, 0}, {1, 0, 1}, {2, 0, 2}, {0, 1, 3}, {0, 2, 4}};
    for (auto& s : sl) {
      ssg::GemmDesc g{};
      g.A = dS + s[0] * sd; g.ai = nline; g.ak = 1; g.sA = bk;
      g.B = dE + size_t(s[1]) * nv * NK * bk; g.bk = 1; g.bj = bk; g.sB = long(NK) * bk;
      g.C = dout + long(s[2]) * nv * npr; g.ci = NK; g.cj = 1; g.sC = npr;
      g.M = ni; g.N = NK; g.K = bk; g.batch = nv;
      ssg::launch_gemm_f64(g, st);
    }
  }
  HIPCHK(hipGetLastError());
  if (mem == SSG_MEM_HOST) {
    HIPCHK(hipMemcpyAsync(out, dout, sizeof(double) * size_t(npr) * nv * S.nslots, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  return SSG_OK;
}

int ssg_evaluate_unstructured(ssg_plan* p, const double* spectral, const double* pts, long npts, double* out,
                              double fill, int oob_error, int mem, void* stream) {
  int rc = check_call(p, spectral, out, mem);
  if (rc) return rc;
  if (npts < 0 || (npts > 0 && !pts)) return fail(SSG_EINVAL, "ArgumentError: points");
  if (npts >= (1L << 31)) return fail(SSG_EINVAL, "ArgumentError: at most 2^31 - 1 points per call");
  if (p->T.ahat) return fail(SSG_ENOTSUP, "ENOTSUP: evaluate_unstructured with R3X background coefficients set");
  if (npts == 0) return SSG_OK;
  const ssg::Setup& S = p->S;
  const ssg::DevTables& T = p->T;
  const int geo = S.geometry, nv = S.nvars;
  const bool ring = geo != SSG_RZ, cheb = geo != SSG_RL;
  const int nks = ring ? 2 * S.K + 1 : 1;
  const long ntr = ring ? long(T.ntr) : T.nlines;
  if (cheb && S.b_kDim > 48) return fail(SSG_ENOTSUP, "unstructured evaluation supports at most 48 Chebyshev modes");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const double* px = pts;
  const double* pl = ring ? pts + npts : nullptr;
  const double* pz = geo == SSG_RZ ? pts + npts : (geo == SSG_RLZ ? pts + 2 * npts : nullptr);

  // bounds (_check_bounds_nd, src/interpolation.jl:1202-1232): the spline and
  // Chebyshev axes are bounded, the angle is periodic; NaN is outside
  std::vector<long> oob;
  std::vector<int> inb;
  inb.reserve(size_t(npts));
  for (long i = 0; i < npts; i++) {
    bool ok = px[i] >= S.d.iMin && px[i] <= S.d.iMax;
    if (pz) ok = ok && pz[i] >= S.d.kMin && pz[i] <= S.d.kMax;
    if (ok) inb.push_back(int(i));
    else oob.push_back(i);
  }
  if (oob_error && !oob.empty())
    return fail(SSG_EDOMAIN, "DomainError: " + std::to_string(oob.size()) + " points are outside the source domain");

  // first taps (SItransform, src/CubicBSpline.jl:1554-1567), a counting sort by
  // first tap, and groups of at most 64 points with one first tap (the kernel
  // evaluates the tap weights)
  const double DX = (S.d.iMax - S.d.iMin) / S.nc, DXr = 1.0 / DX;
  const long nin = long(inb.size());
  std::vector<int> xm(static_cast<size_t>(nin));
  const int nb = S.nc + 4;                              // first taps -2 .. nc + 1
  std::vector<long> cnt(size_t(nb) + 1, 0);
  for (long q = 0; q < nin; q++) {
    int m = int(std::ceil((px[inb[q]] - S.d.iMin - 2.0 * DX) * DXr));
    m = m < -2 ? -2 : (m > S.nc + 1 ? S.nc + 1 : m);
    xm[q] = m;
    cnt[size_t(m + 2) + 1]++;
  }
  for (int b = 0; b < nb; b++) cnt[size_t(b) + 1] += cnt[size_t(b)];
  std::vector<int> perm(static_cast<size_t>(nin)), sxm(static_cast<size_t>(nin));
  for (long q = 0; q < nin; q++) {
    const long at = cnt[size_t(xm[q] + 2)]++;
    perm[size_t(at)] = inb[q];
    sxm[size_t(at)] = xm[q];
  }
  std::vector<double> xs(static_cast<size_t>(nin)), lam, zs;
  if (ring) lam.resize(size_t(nin));
  if (cheb) zs.resize(size_t(nin));
  std::vector<int> gstart, gcount, gxm;
  for (long n = 0; n < nin; n++) {
    const int i = perm[size_t(n)], m0 = sxm[size_t(n)];
    xs[size_t(n)] = px[i];
    if (ring) lam[size_t(n)] = pl[i];
    if (cheb) zs[size_t(n)] = pz[i];
    if (gstart.empty() || gxm.back() != m0 || gcount.back() == 64) {
      gstart.push_back(int(n));
      gcount.push_back(0);
      gxm.push_back(m0);
    }
    gcount.back()++;
  }
  const int ngroups = int(gstart.size());
  std::vector<double> ca;
  if (cheb) {
    ca.reserve(size_t(nv) * S.kDim * S.b_kDim);
    for (int v = 0; v < nv; v++) ca.insert(ca.end(), S.cheb[v].ca.begin(), S.cheb[v].ca.end());
  }

  // device scratch: folded coefficients, one chunk of point products, tables, staging
  const long CH = 1L << 20;                             // points per chunk of groups
  auto al = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t bA = al(sizeof(double) * size_t(S.b_iDim) * nks * ntr);
  const size_t bC = cheb ? al(sizeof(double) * size_t(std::min(nin, CH + 64)) * ntr) : 0;
  const size_t bW = al(sizeof(double) * xs.size()), bL = al(sizeof(double) * lam.size());
  const size_t bZ = al(sizeof(double) * zs.size()), bP = al(sizeof(int) * perm.size());
  const size_t bG = al(sizeof(int) * 3 * size_t(ngroups)), bO = al(sizeof(long) * oob.size());
  const size_t bK = al(sizeof(double) * ca.size());
  const size_t bIn = mem == SSG_MEM_HOST ? al(sizeof(double) * size_t(S.nspec) * nv) : 0;
  const size_t bOut = mem == SSG_MEM_HOST ? al(sizeof(double) * size_t(npts) * nv) : 0;
  if ((rc = grow_reg(p, bA + bC + bW + bL + bZ + bP + bG + bO + bK + bIn + bOut))) return rc;
  char* base = static_cast<char*>(p->reg);
  double* dA = reinterpret_cast<double*>(base); base += bA;
  double* dC = reinterpret_cast<double*>(base); base += bC;
  double* dW = reinterpret_cast<double*>(base); base += bW;
  double* dL = reinterpret_cast<double*>(base); base += bL;
  double* dZ = reinterpret_cast<double*>(base); base += bZ;
  int* dP = reinterpret_cast<int*>(base); base += bP;
  int* dG = reinterpret_cast<int*>(base); base += bG;
  long* dO = reinterpret_cast<long*>(base); base += bO;
  double* dK = reinterpret_cast<double*>(base); base += bK;
  const double* spec = spectral;
  double* dout = out;
  if (mem == SSG_MEM_HOST) {
    double* din = reinterpret_cast<double*>(base); base += bIn;
    dout = reinterpret_cast<double*>(base);
    HIPCHK(hipMemcpyAsync(din, spectral, sizeof(double) * size_t(S.nspec) * nv, hipMemcpyHostToDevice, st));
    spec = din;
  }
  std::vector<int> groups;
  groups.reserve(size_t(3) * ngroups);
  groups.insert(groups.end(), gstart.begin(), gstart.end());
  groups.insert(groups.end(), gcount.begin(), gcount.end());
  groups.insert(groups.end(), gxm.begin(), gxm.end());
  auto up = [&](void* d, const void* h, size_t bytes) -> int {
    if (bytes) HIPCHK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st));
    return SSG_OK;
  };
  if ((rc = up(dW, xs.data(), sizeof(double) * xs.size())) || (rc = up(dL, lam.data(), sizeof(double) * lam.size())) ||
      (rc = up(dZ, zs.data(), sizeof(double) * zs.size())) || (rc = up(dP, perm.data(), sizeof(int) * perm.size())) ||
      (rc = up(dG, groups.data(), sizeof(int) * groups.size())) ||
      (rc = up(dO, oob.data(), sizeof(long) * oob.size())) || (rc = up(dK, ca.data(), sizeof(double) * ca.size())))
    return rc;

  if (nin > 0) {
    {
      StageMark m(p, st, "unstr_sa");
      if (ring) ssg::launch_sa_ring_only(T, spec, p->y, st);
      else ssg::launch_sa_lines(T, spec, p->y, st);
      ssg::launch_fold_lines(T, ring, p->y, dA, st);
    }
    for (int g0 = 0; g0 < ngroups;) {
      int g1 = g0;
      while (g1 < ngroups && long(gstart[g1]) + gcount[g1] - gstart[g0] <= CH) g1++;
      if (g1 == g0) g1++;
      const long cbase = gstart[g0], cn = long(gstart[g1 - 1]) + gcount[g1 - 1] - cbase;
      ssg::UnstrDesc d{};
      d.A2 = dA; d.nks = nks; d.ntr = int(ntr); d.b_iDim = S.b_iDim;
      d.gstart = dG + g0; d.gcount = dG + ngroups + g0; d.gxm = dG + 2 * ngroups + g0;
      d.x = dW; d.xmin = S.d.iMin; d.DX = DX; d.DXr = DXr; d.nc = S.nc;
      d.lam = ring ? dL : nullptr;
      d.C = cheb ? dC : nullptr; d.cbase = cbase;
      d.out = dout; d.perm = dP; d.nall = npts;
      {
        StageMark m(p, st, "unstr_points");
        ssg::launch_unstr_gemm(d, g1 - g0, st);
      }
      if (cheb) {
        StageMark m(p, st, "unstr_cheb");
        ssg::UnstrCheb c{};
        c.C = dC; c.npt = cn; c.ntr = int(ntr); c.nzb = S.b_kDim; c.kDim = S.kDim; c.nv = nv;
        c.ca = dK; c.z = dZ; c.cbase = cbase; c.zmin = S.d.kMin; c.zmax = S.d.kMax;
        c.out = dout; c.perm = dP; c.nall = npts;
        ssg::launch_unstr_cheb(c, st);
      }
      g0 = g1;
    }
  }
  ssg::launch_unstr_fill(dO, long(oob.size()), nv, npts, fill, dout, st);
  HIPCHK(hipGetLastError());
  // the host tables are freed on return: copies from pageable host memory have
  // consumed their source when hipMemcpyAsync returns
  if (mem == SSG_MEM_HOST) {
    HIPCHK(hipMemcpyAsync(out, dout, sizeof(double) * size_t(npts) * nv, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  return SSG_OK;
}

int ssg_plan_set_ahat(ssg_plan* p, const double* ahat, int mem) {
  if (!p) return fail(SSG_EINVAL, "ArgumentError: null plan");
  if (mem != SSG_MEM_HOST && mem != SSG_MEM_DEVICE) return fail(SSG_EINVAL, "ArgumentError: mem must be SSG_MEM_HOST or SSG_MEM_DEVICE");
  HIPCHK(hipSetDevice(p->device));
  HIPCHK(hipDeviceSynchronize());                   // no call of this plan is using the previous values
  if (!ahat || !p->S.any_r3x) {
    // cleared, or no variable has an R3X side: SAtransform! ignores ahat
    p->T.ahat = nullptr;
    return SSG_OK;
  }
  const size_t n = size_t(p->S.nspec) * p->S.nvars;
  int rc;
  if (!p->ahat && (rc = dalloc(p, &p->ahat, n))) return rc;
  if (!p->sbt && (rc = dalloc(p, &p->sbt, n))) return rc;
  HIPCHK(hipMemcpy(p->ahat, ahat, sizeof(double) * n,
                   mem == SSG_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice));
  p->T.ahat = p->ahat;
  return SSG_OK;
}

// applyFilter! (src/filtering.jl:262-428).  The weight of each spectral line is
// the variable's weight of the line's wavenumber (RL / SL), of its Chebyshev
// mode (RZ), or their product (RLZ / SLZ); the reference skips a block whose
// weight is isapprox 1.0, assigns 0.0 to one whose weight is 0.0 (isapprox with
// atol = 0 holds for 0.0 alone) and scales the others.  The table keeps the
// lines that change.
int ssg_plan_set_filter(ssg_plan* p, const double* fw, const double* cw) {
  if (!p) return fail(SSG_EINVAL, "ArgumentError: null plan");
  HIPCHK(hipSetDevice(p->device));
  HIPCHK(hipDeviceSynchronize());                   // no call of this plan is using the previous table
  p->filt_n = 0;
  const ssg::Setup& S = p->S;
  const int g = S.geometry;                         // SL / SLZ run as RL / RLZ
  const double* fwu = g == SSG_RZ ? nullptr : fw;  // RZ filters Chebyshev modes only
  const double* cwu = g == SSG_RL ? nullptr : cw;  // RL / SL filter wavenumbers only
  if (!fwu && !cwu) return SSG_OK;
  const int K = S.K, nl = g == SSG_RZ ? 1 : 1 + 2 * K;
  const long nlv = S.nspec / S.b_iDim;              // lines per variable
  const double rtol = 1.4901161193847656e-8;       // sqrt(eps(Float64)), isapprox's default
  std::vector<long> off;
  std::vector<double> w;
  for (int v = 0; v < S.nvars; v++) {
    for (long l = 0; l < nlv; l++) {
      const int zb = int(l / nl), pl = int(l % nl);
      const int k = (pl + 1) / 2;
      const double c = cwu ? cwu[long(v) * S.b_kDim + zb] : 1.0;
      const double f = fwu ? fwu[long(v) * (K + 1) + k] : 1.0;
      const double wl = g == SSG_RL ? f : g == SSG_RZ ? c : c * f;
      if (wl == 1.0 || (std::isfinite(wl) && std::fabs(wl - 1.0) <= rtol * std::max(std::fabs(wl), 1.0))) continue;
      off.push_back(long(v) * S.nspec + l * S.b_iDim);
      w.push_back(wl);
    }
  }
  if (off.empty()) return SSG_OK;
  const size_t cap = size_t(S.nvars) * size_t(nlv);
  int rc;
  if (!p->filt_off && (rc = dalloc(p, &p->filt_off, cap))) return rc;
  if (!p->filt_w && (rc = dalloc(p, &p->filt_w, cap))) return rc;
  HIPCHK(hipMemcpy(p->filt_off, off.data(), off.size() * sizeof(long), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p->filt_w, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice));
  p->filt_n = long(off.size());
  return SSG_OK;
}

int ssg_apply_filter(ssg_plan* p, double* spectral, int mem, void* stream) {
  int rc = check_call(p, spectral, spectral, mem);
  if (rc) return rc;
  if (p->filt_n == 0) return SSG_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const ssg::Setup& S = p->S;
  if (mem == SSG_MEM_DEVICE) {
    { StageMark m(p, st, "filter"); ssg::launch_apply_filter(spectral, p->filt_off, p->filt_w, p->filt_n, S.b_iDim, st); }
    HIPCHK(hipGetLastError());
    return SSG_OK;
  }
  if ((rc = ensure_staging(p))) return rc;
  const size_t n = size_t(S.nspec) * S.nvars;
  HIPCHK(hipMemcpyAsync(p->h_spec, spectral, n * sizeof(double), hipMemcpyHostToDevice, st));
  { StageMark m(p, st, "filter"); ssg::launch_apply_filter(p->h_spec, p->filt_off, p->filt_w, p->filt_n, S.b_iDim, st); }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(spectral, p->h_spec, n * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return SSG_OK;
}

// ----------------------------------------------------------------------------
// The tile pipeline (src/tiling.jl) on the ring grids
// ----------------------------------------------------------------------------
static int check_tile_plan(ssg_plan* p) {
  const ssg::DevTables& T = p->T;
  if (T.geometry == SSG_RZ) return fail(SSG_ENOTSUP, "the tile transforms take the ring grids (RL, RLZ, SL, SLZ)");
  if (p->precision != SSG_FP64) return fail(SSG_EINVAL, "ArgumentError: the tile transforms take SSG_FP64 plans");
  if (!ssg::sasi_supported(T)) return fail(SSG_ENOTSUP, "the tile transforms need b_iDim <= 256 (the fused solve)");
  return SSG_OK;
}

static int ensure_spec2(ssg_plan* p) {
  if (p->h_spec2) return SSG_OK;
  return dalloc(p, &p->h_spec2, size_t(p->S.nspec) * p->S.nvars);
}

int ssg_spline_transform(ssg_plan* p, const double* shared, double* spectral, int mem, void* stream) {
  int rc = check_call(p, shared, spectral, mem);
  if (rc || (rc = check_tile_plan(p))) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const ssg::DevTables& T = p->T;
  const size_t n = size_t(p->S.nspec) * p->S.nvars;
  const double* in = shared;
  double* out = spectral;
  if (mem == SSG_MEM_HOST) {
    if ((rc = ensure_staging(p)) || (rc = ensure_spec2(p))) return rc;
    HIPCHK(hipMemcpyAsync(p->h_spec, shared, n * sizeof(double), hipMemcpyHostToDevice, st));
    in = p->h_spec;
    out = p->h_spec2;
  }
  if (T.ahat) {
    // R3X: the solve runs on btilde = b - pq ahat and the background is added back
    StageMark m(p, st, "ahat_btilde");
    ssg::launch_ahat_btilde(T, in, p->sbt, st);
    in = p->sbt;
  }
  { StageMark m(p, st, "spline"); ssg::launch_sasi_to_a(T, in, out, st); }
  HIPCHK(hipGetLastError());
  if (mem == SSG_MEM_HOST) {
    HIPCHK(hipMemcpyAsync(spectral, out, n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  return SSG_OK;
}

int ssg_tile_transform(ssg_plan* p, const double* spectral, double* physical, int mem, void* stream) {
  int rc = check_call(p, spectral, physical, mem);
  if (rc || (rc = check_tile_plan(p))) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const ssg::DevTables& T = p->T;
  const ssg::Setup& S = p->S;
  const double* a = spectral;
  double* phys = physical;
  if (mem == SSG_MEM_HOST) {
    if ((rc = ensure_staging(p))) return rc;
    HIPCHK(hipMemcpyAsync(p->h_spec, spectral, size_t(S.nspec) * S.nvars * sizeof(double), hipMemcpyHostToDevice, st));
    a = p->h_spec;
    phys = static_cast<double*>(p->h_phys);
  }
  { StageMark m(p, st, "si"); ssg::launch_sasi_from_a(T, a, p->X, st); }
  run_synthesis<double>(p, st, plan_scratch(p), p->X, phys);
  HIPCHK(hipGetLastError());
  if (mem == SSG_MEM_HOST) {
    HIPCHK(hipMemcpyAsync(physical, phys, size_t(S.npts) * S.nvars * S.nslots * sizeof(double),
                          hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  return SSG_OK;
}

// a tile's lines into the patch's (add: sumSpectralTile!, else setSpectralTile!)
static int tile_embed_call(ssg_plan* patch, double* ps, ssg_plan* tile, const double* ts, int siL, int mem,
                           void* stream, bool add) {
  int rc = check_call(patch, ps, ts, mem);
  if (rc) return rc;
  if (!tile) return fail(SSG_EINVAL, "ArgumentError: null tile plan");
  const ssg::DevTables& P = patch->T;
  const ssg::DevTables& T = tile->T;
  const ssg::Setup& SP = patch->S;
  const ssg::Setup& ST = tile->S;
  if (P.geometry == SSG_RZ || T.geometry != P.geometry)
    return fail(SSG_ENOTSUP, "the tile sums take a ring-grid patch and a tile of its geometry");
  if (T.nvars != P.nvars || T.nzb != P.nzb)
    return fail(SSG_EINVAL, "ArgumentError: tile and patch differ in variables or Chebyshev modes");
  if (siL < 1 || siL - 1 + T.b_iDim > P.b_iDim || T.nl > P.nl)
    return fail(SSG_EBOUNDS, "BoundsError: the tile's spectral rows do not fit in the patch");
  hipStream_t st = static_cast<hipStream_t>(stream);
  double* pd = ps;
  const double* td = ts;
  const size_t np = size_t(SP.nspec) * SP.nvars, nt = size_t(ST.nspec) * ST.nvars;
  if (mem == SSG_MEM_HOST) {
    if ((rc = ensure_staging(patch)) || (rc = ensure_staging(tile))) return rc;
    // a grid summed into itself stages its two arrays apart
    double* tstage = tile->h_spec;
    if (tile == patch) {
      if ((rc = ensure_spec2(patch))) return rc;
      tstage = patch->h_spec2;
    }
    if (add) HIPCHK(hipMemcpyAsync(patch->h_spec, ps, np * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(tstage, ts, nt * sizeof(double), hipMemcpyHostToDevice, st));
    pd = patch->h_spec;
    td = tstage;
  }
  ssg::launch_tile_embed(pd, td, P.nvars, P.nzb, P.b_iDim, P.nl, SP.nspec, T.b_iDim, T.nl, ST.nspec, siL - 1, add, st);
  HIPCHK(hipGetLastError());
  if (mem == SSG_MEM_HOST) {
    HIPCHK(hipMemcpyAsync(ps, pd, np * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  return SSG_OK;
}

int ssg_sum_spectral_tile(ssg_plan* patch, double* patch_spectral, ssg_plan* tile, const double* tile_spectral,
                          int spectral_index_l, int mem, void* stream) {
  return tile_embed_call(patch, patch_spectral, tile, tile_spectral, spectral_index_l, mem, stream, true);
}

int ssg_set_spectral_tile(ssg_plan* patch, double* patch_spectral, ssg_plan* tile, const double* tile_spectral,
                          int spectral_index_l, int mem, void* stream) {
  return tile_embed_call(patch, patch_spectral, tile, tile_spectral, spectral_index_l, mem, stream, false);
}

int ssg_sum_shared_spectral(double* shared, long shared_rows, ssg_plan* tile, const double* tile_spectral,
                            int spectral_index_l, int mem, void* stream) {
  int rc = check_call(tile, shared, tile_spectral, mem);
  if (rc) return rc;
  const ssg::Setup& ST = tile->S;
  const long w = long(tile->T.b_iDim) - 3;           // rows spectralIndexL .. spectralIndexR - 3
  if (w <= 0) return SSG_OK;
  if (spectral_index_l < 1 || spectral_index_l - 1 + w > shared_rows)
    return fail(SSG_EBOUNDS, "BoundsError: the tile's rows do not fit in the shared array");
  if (mem == SSG_MEM_HOST) {
    // host arrays: a copy of w rows per variable column, nothing for the device to do
    for (int v = 0; v < ST.nvars; v++)
      std::memcpy(shared + long(v) * shared_rows + (spectral_index_l - 1), tile_spectral + long(v) * ST.nspec,
                  size_t(w) * sizeof(double));
    return SSG_OK;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  HIPCHK(hipMemcpy2DAsync(shared + (spectral_index_l - 1), size_t(shared_rows) * sizeof(double), tile_spectral,
                          size_t(ST.nspec) * sizeof(double), size_t(w) * sizeof(double), size_t(ST.nvars),
                          hipMemcpyDeviceToDevice, st));
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
