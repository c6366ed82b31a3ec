// ssg_ring_fft.hip -- ring synthesis (HC2R with azimuthal derivatives) and ring
// analysis (R2HC with the forward phase filter) by chirp-z transforms, and the
// RLZ inverse Chebyshev stage that follows the synthesis.  The transforms run in
// FP64, or in FP32 for the FP32 form's synthesis (float tables and output).
//
// Every output column of a ring is (src/Fourier.jl:463-468, 495-503)
//   y(phi_l) = dc + 2 sum_{k=1}^{kmax} [P_k cos(k phi_l) + Q_k sin(k phi_l)],  phi_l = 2 pi l / N,
// with (P, Q) the phase-rotated SI values in X (values, d/dr, d2/dr2) or the
// azimuthal derivative families built from the values: (k Q, -k P) for
// d/dlambda and (-k^2 P, -k^2 Q) for d2/dlambda2.  As a two-sided sum,
// y(l) = sum_{k=-K}^{K} c_k W^(k l) with W = exp(2 pi i / N), c_0 = dc,
// c_k = P_k - i Q_k and c_-k = conj(c_k); two real columns A and B share one
// complex sum with coefficients a = c^A + i c^B, whose real and imaginary parts
// are y^A and y^B.
//
// Bluestein: with w(n) = exp(i pi n^2 / N), W^(k l) = w(k) w(l) conj(w(l - k)), so
//   y(l) = w(l) [g (*) h](l + K),  g_u = a_{u-K} w(u - K),  h(n) = conj(w(n)),
// a linear convolution over l - k in [-K, N + K - 1], which a circular one of
// length M = 2^L >= N + 2K computes exactly: c = IFFT(FFT(g) . H) with H the
// ring's transform of h, scaled by 1/M (plan tables, long double on the host).
// The inverse runs as conj(FFT(conj(.))), so one forward FFT serves both.
//
// FFT: M = 16 T points in T = M / 16 threads, 16 per thread in registers,
// element t + e T in register e of thread t.  Stockham passes of radix 16 (and
// one of 2, 4 or 8 when L is not a multiple of 4): a pass reads the elements
// t + e T, multiplies by the twiddles of its stage, runs the DFT in registers
// and writes its outputs to the pass's Stockham positions in LDS, from which
// the next pass reads; the last pass leaves its outputs in natural order in
// registers, which is where the pointwise product with H and the inverse's
// first pass read them, and where the output stage stores them from.  LDS rows
// are padded by one element per 16, so the stride-16 writes of a first pass
// fall on distinct banks.  Exchanges move real and imaginary parts in turn, so a
// transform needs M reals of LDS.
#include "ssg_kernels.hpp"
#include "ssg_mfma.hpp"
#include "ssg_setup.hpp"

namespace ssg {

// dynamic LDS of a chirp-z block: M + M / 16 doubles for each of its 4096 / M
// transforms (the exchanges move real and imaginary parts in turn) and two
// doubles per wave for the pair norms
// (reals of esize bytes: double in the FP64 form, float in the FP32 form's synthesis)
__host__ __device__ constexpr size_t fft_red(size_t esize) { return size_t(kFftMaxM) * 17 / 16 * esize; }
__host__ __device__ constexpr size_t fft_lds(size_t esize) { return fft_red(esize) + 8 * sizeof(double); }

// waves per SIMD of the FP32 synthesis transforms (the batched kernel, with its job
// lookup, fits five without scratch)
#define SSG_FFT32_WAVES 6, 6
#define SSG_FFT32_BATCH_WAVES 5, 5

// arithmetic of a synthesis transform by its output type: FP64 for double output,
// FP32 for the FP32 form's float output
template <typename Yt> struct FftC { typedef double2 C; };
template <> struct FftC<float> { typedef float2 C; };

// exponent of the pair scale sb = 2^e; in FP32 arithmetic |e| is held to 100, so
// that sb and 1 / sb stay normal floats (columns further apart than that keep
// only part of the balance)
template <typename F> __device__ __forceinline__ int fft_scale_exp(int e) {
  if constexpr (sizeof(F) == sizeof(float)) return e < -100 ? -100 : e > 100 ? 100 : e;
  else return e;
}

namespace {

// complex arithmetic of the transforms: double2 (FP64 form) or float2 (FP32 form)
template <typename C> struct Cx;
template <> struct Cx<double2> {
  typedef double F;
  static __device__ __forceinline__ double2 mk(double a, double b) { return make_double2(a, b); }
};
template <> struct Cx<float2> {
  typedef float F;
  static __device__ __forceinline__ float2 mk(float a, float b) { return make_float2(a, b); }
};
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return Cx<C>::mk(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return Cx<C>::mk(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  return Cx<C>::mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename C> __device__ __forceinline__ C mul_mi(C a) { return Cx<C>::mk(a.y, -a.x); }   // -i a

// forward DFT of 2 and 4 points in place (sign -1)
template <typename C> __device__ __forceinline__ void dft2(C& a, C& b) {
  const C s = cadd(a, b), d = csub(a, b);
  a = s; b = d;
}
template <typename C> __device__ __forceinline__ void dft4(C& x0, C& x1, C& x2, C& x3) {
  const C s02 = cadd(x0, x2), d02 = csub(x0, x2), s13 = cadd(x1, x3), d13 = mul_mi(csub(x1, x3));
  x0 = cadd(s02, s13);
  x2 = csub(s02, s13);
  x1 = cadd(d02, d13);
  x3 = csub(d02, d13);
}

constexpr double kR2 = 0.70710678118654752440;   // sqrt(1/2)
constexpr double kC1 = 0.92387953251128675613;   // cos(pi / 8)
constexpr double kS1 = 0.38268343236508977173;   // sin(pi / 8)

// x w16^2 = x (1 - i) / sqrt 2 and x w16^6 = x (-1 - i) / sqrt 2
template <typename C> __device__ __forceinline__ C rot2(C x) {
  typedef typename Cx<C>::F F;
  return Cx<C>::mk(F(kR2) * (x.x + x.y), F(kR2) * (x.y - x.x));
}
template <typename C> __device__ __forceinline__ C rot6(C x) {
  typedef typename Cx<C>::F F;
  return Cx<C>::mk(F(kR2) * (x.y - x.x), -F(kR2) * (x.x + x.y));
}

// DFT of R points in x[0..R): R = 8 and 16 as (R/4) x 4 (n = 4 n1 + n2,
// k = k1 + (R/4) k2), leaving X_k in x[dft_pos(R, k)]
__host__ __device__ constexpr int dft_pos(int R, int k) {
  return R == 16 ? 4 * (k & 3) + (k >> 2) : R == 8 ? 4 * (k & 1) + (k >> 1) : k;
}
template <int R, typename C> struct Dft;
template <typename C> struct Dft<2, C> { static __device__ __forceinline__ void run(C* x) { dft2(x[0], x[1]); } };
template <typename C> struct Dft<4, C> {
  static __device__ __forceinline__ void run(C* x) { dft4(x[0], x[1], x[2], x[3]); }
};
template <typename C> struct Dft<8, C> {
  static __device__ __forceinline__ void run(C* x) {
#pragma unroll
    for (int n2 = 0; n2 < 4; n2++) dft2(x[n2], x[4 + n2]);
    // x[n2 + 4] *= w8^n2
    x[5] = rot2(x[5]);
    x[6] = mul_mi(x[6]);
    x[7] = rot6(x[7]);
    dft4(x[0], x[1], x[2], x[3]);
    dft4(x[4], x[5], x[6], x[7]);
  }
};
// the twiddles x[n2 + 4 k1] *= w16^(n2 k1) and the second-stage 4-point DFTs of a
// 16-point DFT
template <typename C> __device__ __forceinline__ void dft16_tail(C* x) {
  typedef typename Cx<C>::F F;
  const C w1 = Cx<C>::mk(F(kC1), F(-kS1)), w3 = Cx<C>::mk(F(kS1), F(-kC1)), w9 = Cx<C>::mk(F(-kC1), F(kS1));
  x[5] = cmul(x[5], w1);
  x[9] = rot2(x[9]);
  x[13] = cmul(x[13], w3);
  x[6] = rot2(x[6]);
  x[10] = mul_mi(x[10]);
  x[14] = rot6(x[14]);
  x[7] = cmul(x[7], w3);
  x[11] = rot6(x[11]);
  x[15] = cmul(x[15], w9);
#pragma unroll
  for (int k1 = 0; k1 < 4; k1++) dft4(x[4 * k1], x[4 * k1 + 1], x[4 * k1 + 2], x[4 * k1 + 3]);
}
template <typename C> struct Dft<16, C> {
  static __device__ __forceinline__ void run(C* x) {
#pragma unroll
    for (int n2 = 0; n2 < 4; n2++) dft4(x[n2], x[4 + n2], x[8 + n2], x[12 + n2]);
    dft16_tail(x);
  }
};

// dft<16> of an input whose entries 6..15 are zero: the first-stage 4-point DFTs
// of (x0, x4, 0, 0), (x1, x5, 0, 0), (x2, 0, 0, 0), (x3, 0, 0, 0) in closed form
template <typename C> __device__ __forceinline__ void dft16_head6(C* x) {
#pragma unroll
  for (int n2 = 0; n2 < 2; n2++) {
    const C a = x[n2], b = x[4 + n2];
    x[n2] = cadd(a, b);
    x[4 + n2] = cadd(a, mul_mi(b));
    x[8 + n2] = csub(a, b);
    x[12 + n2] = csub(a, mul_mi(b));
  }
#pragma unroll
  for (int n2 = 2; n2 < 4; n2++) x[4 + n2] = x[8 + n2] = x[12 + n2] = x[n2];
  dft16_tail(x);
}

// Shape of pass P of a 2^L-point transform: radix R, product LS of the earlier
// passes' radices, B = 16 / R butterflies per thread.
template <int L, int P>
struct PassOf {
  static constexpr int NP = L / 4, RL = 1 << (L % 4), NPASS = NP + (RL > 1 ? 1 : 0);
  static constexpr int R = P < NP ? 16 : RL;
  static constexpr int LS = 1 << (4 * P);
  static constexpr int B = 16 / R;
};

// Ordering of a transform's exchanges through LDS: a block barrier when the
// transform spans several waves, otherwise the wave's own LDS operations are
// performed in order and only the compiler must not move them across.
template <int TPP>
__device__ __forceinline__ void xsync() {
  if constexpr (TPP > 64) {
    __syncthreads();
  } else {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// LDS slot of element i in a transform's exchange array: one pad per 16, so that
// the stride-16 writes of a first pass fall on distinct banks; the slots of a
// thread's 16 writes (and of its 16 reads) are then one address plus constants.
__device__ __forceinline__ int swz(int i) { return i + (i >> 4); }

// A value the compiler must treat as new: address arithmetic derived from it is
// formed where it is used instead of being merged with the same arithmetic of an
// earlier pass and held in registers (or scratch) across the transforms.
__device__ __forceinline__ int fresh(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

// Stockham passes of a 2^L-point transform.  Butterfly s = t + q T (q < B) takes
// elements s + j M / R (register q + j B), multiplies element j by
// w_{LS R}^(j (s mod LS)) and writes X_j to (s / LS) LS R + (s mod LS) + j LS; on
// the last pass that position is t + (q + j B) T, register q + j B, so the
// result stays in registers in natural order.  The other passes (radix 16, one
// butterfly per thread, whose 16 outputs lie in one half of the array) exchange
// through the M reals of buf: the real parts first, then the imaginary parts.
template <int L, int P, bool HEAD6 = false, typename C>
__device__ __forceinline__ void fft_passes(C (&v)[16], typename Cx<C>::F* buf, int t, const C* __restrict__ tw) {
  typedef PassOf<L, P> PP;
  typedef typename Cx<C>::F F;
  if constexpr (P < PP::NPASS) {
    constexpr int T = (1 << L) / 16, R = PP::R, B = PP::B, LS = PP::LS;
    constexpr bool LAST = P == PP::NPASS - 1;
#pragma unroll
    for (int q = 0; q < B; q++) {
      const int s = t + q * T;
      const int k = fresh(s & (LS - 1));
      C x[R];
#pragma unroll
      for (int j = 0; j < R; j++) x[j] = v[q + j * B];
      if constexpr (LS > 1) {
        // w^j (w = w_{LS R}^k) from the table entries of w, w^2, w^4, w^8: at most
        // three products each, formed as they are applied
        constexpr int step = kFftMaxM / (LS * R);
        const int n1 = k * step;
        const C w1 = tw[n1];
        const C w2 = R > 2 ? tw[2 * n1] : w1;
        const C w4 = R > 4 ? tw[4 * n1] : w1;
        const C w8 = R > 8 ? tw[8 * n1] : w1;
        x[1] = cmul(x[1], w1);
        if constexpr (R > 2) {
          const C w3 = cmul(w2, w1);
          x[2] = cmul(x[2], w2);
          x[3] = cmul(x[3], w3);
          if constexpr (R > 4) {
            x[4] = cmul(x[4], w4);
            x[5] = cmul(x[5], cmul(w4, w1));
            x[6] = cmul(x[6], cmul(w4, w2));
            x[7] = cmul(x[7], cmul(w4, w3));
            if constexpr (R > 8) {
              x[8] = cmul(x[8], w8);
              x[9] = cmul(x[9], cmul(w8, w1));
              x[10] = cmul(x[10], cmul(w8, w2));
              x[11] = cmul(x[11], cmul(w8, w3));
              x[12] = cmul(x[12], cmul(w8, w4));
              const C w5 = cmul(w4, w1), w6 = cmul(w4, w2), w7 = cmul(w4, w3);
              x[13] = cmul(x[13], cmul(w8, w5));
              x[14] = cmul(x[14], cmul(w8, w6));
              x[15] = cmul(x[15], cmul(w8, w7));
            }
          }
        }
      }
      if constexpr (HEAD6 && P == 0) dft16_head6(x);
      else Dft<R, C>::run(x);
#pragma unroll
      for (int j = 0; j < R; j++) v[q + j * B] = x[dft_pos(R, j)];
    }
    if constexpr (!LAST) {
      static_assert(B == 1, "exchanges follow radix-16 passes");
      // write slots swz(base + j LS) = swz(base) + j (LS + LS / 16) (LS = 1 or a
      // multiple of 16) and, for T >= 16, read slots swz(t + e T) = swz(t) +
      // e (T + T / 16): one address each plus constant offsets
      F* const bw = buf + fresh(swz((t / LS) * (LS * 16) + (t & (LS - 1))));
      constexpr int WS = LS + LS / 16;
      auto rd = [&](int e) -> F& {
        if constexpr (T >= 16) return buf[swz(t) + e * (T + T / 16)];
        else return buf[swz(t + e * T)];
      };
#pragma unroll
      for (int j = 0; j < 16; j++) bw[j * WS] = v[j].x;
      xsync<T>();
#pragma unroll
      for (int e = 0; e < 16; e++) v[e].x = rd(e);
      xsync<T>();
#pragma unroll
      for (int j = 0; j < 16; j++) bw[j * WS] = v[j].y;
      xsync<T>();
#pragma unroll
      for (int e = 0; e < 16; e++) v[e].y = rd(e);
      xsync<T>();
      fft_passes<L, P + 1, false, C>(v, buf, t, tw);
    }
  }
}

// a synthesis column: its X runs, cap, family and output
struct SynCol {
  const void* p; const void* q;   // P and Q runs of its order (wavenumber k at offset k)
  int km;                         // wavenumbers above km are zero
  int fam;                        // 0 values, 1 d/dlambda, 2 d2/dlambda2, 3 d/dr, 4 d2/dr2; -1 none
};

// (P, Q) of a column at wavenumber k from the raw run values p, q: the family's
// derivation, zero above the cap, at k = 0 the dc (P) alone
__device__ __forceinline__ void col_coef(const SynCol& c, int k, double p, double q, double& P, double& Q) {
  const double kd = double(k);
  const bool f1 = c.fam == 1, f2 = c.fam == 2;
  P = f1 ? kd * q : f2 ? -kd * kd * p : p;
  Q = f1 ? -kd * p : f2 ? -kd * kd * q : q;
  const bool zero = c.fam < 0 || k > c.km;
  P = zero || (k == 0 && (f1 || f2)) ? 0.0 : P;
  Q = zero || k == 0 ? 0.0 : Q;
}

// Sums of x and y over the TPP consecutive threads of a transform (TPP a power of
// two; red: two doubles per wave of LDS past the transforms' space).  The order of
// the additions is fixed, so every call of a pair gives the same bits.
template <int TPP>
__device__ __forceinline__ void group_sum2(double& x, double& y, double* red) {
  constexpr int W = TPP < 64 ? TPP : 64;
#pragma unroll
  for (int o = W / 2; o > 0; o >>= 1) {
    x += __shfl_xor(x, o);
    y += __shfl_xor(y, o);
  }
  if constexpr (TPP > 64) {
    const int w = int(threadIdx.x) >> 6;
    if ((threadIdx.x & 63) == 0) { red[w] = x; red[4 + w] = y; }
    __syncthreads();
    const int g0 = (int(threadIdx.x) / TPP) * (TPP / 64);
    x = 0.0; y = 0.0;
#pragma unroll
    for (int i = 0; i < TPP / 64; i++) { x += red[g0 + i]; y += red[4 + g0 + i]; }
  }
}

// physical slot (0-based) of a family's values: 1, 4, 5, 2, 3
__device__ __forceinline__ int syn_slot(int fam) { return fam == 0 ? 0 : fam == 1 ? 3 : fam == 2 ? 4 : fam == 3 ? 1 : 2; }

}  // namespace

// One block: 4096 / M pair transforms of one ring (block table entry b: (ring, first pair)).
template <int L, typename S, typename Yt>
__device__ __forceinline__ void syn_fft_body(const DevTables& T, const int* __restrict__ blk, int b,
                                             const S* __restrict__ X, Yt* __restrict__ out) {
  typedef typename FftC<Yt>::C C;                     // double2, or float2 for float output
  typedef typename Cx<C>::F F;
  constexpr int M = 1 << L, TPP = M / 16;
  extern __shared__ __align__(16) unsigned char smraw[];
  const int r = blk[2 * b];
  if (r < 0) return;                                  // padding block (uniform)
  const int g = int(threadIdx.x) / TPP, t = int(threadIdx.x) - g * TPP;
  F* buf = reinterpret_cast<F*>(smraw) + g * (M + M / 16);
  C* gst = reinterpret_cast<C*>(buf);                 // staging of g (2K + 1 <= M / 2 entries)
  const RTab<F> RT = rtab<F>(T);
  const C* tw = reinterpret_cast<const C*>(RT.fftTw);
  // pairs stay within one variable (its results then depend on its own data only):
  // cpv = 5 nzb columns per variable, local column lc = f nzb + zb
  const int nzb = T.nzb, cpv = 5 * nzb, ppv = (cpv + 1) / 2, ntb = T.ntb;
  const int pair = blk[2 * b + 1] + g;
  const bool live = pair < T.nvars * ppv;             // groups past the last pair run on zeros
  const int var = live ? pair / ppv : 0;
  const int ca = live ? 2 * (pair - var * ppv) : 0;
  const int N = T.ringN[r], K = T.ringRi[r], kl = K + 1;
  const bool rlz = T.geometry == 3;
  // X blocks: RLZ one per variable (transform = mode), RL one for all (transform = variable)
  const int vb = rlz ? var : 0;
  const S* Xr = X + (long(vb) * T.Ktot + T.phaseOff[r] + r) * 6 * ntb;
  SynCol col[2];
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int c = ca + h;
    SynCol& C = col[h];
    C.fam = (live && c < cpv) ? c / nzb : -1;
    const int zb = c - (c / nzb) * nzb;
    const int tl = rlz ? zb : var;
    const int d = C.fam < 3 ? 0 : C.fam - 2;
    C.p = Xr + long(d * ntb + tl) * kl;
    C.q = Xr + long((3 + d) * ntb + tl) * kl;
    C.km = C.fam < 0 ? -1 : T.kmax[var * T.iDim + r];
  }
  const C* chirp = reinterpret_cast<const C*>(RT.fftChirp) + T.ringOff[r];
  const C* H = reinterpret_cast<const C*>(RT.fftH) + T.fftHOff[r];

  // g_u = a_{u-K} w(u - K), u <= 2K, staged in LDS: thread t forms the entries
  // u = K +- k of k = t + i TPP (at most three per thread, since M >= 6K); every
  // load is issued before any is used (indices clamped into the runs).  Rounding
  // in the transform is relative to the pair's norm, so column B enters scaled by
  // sb = 2^e (exact) with e such that both columns carry about the same norm, and a
  // column that is exactly zero comes out as exact zeros: each column's result is
  // then as accurate as its own transform's.
  double sb = 1.0;
  bool za = false, zb = false;
  {
    const S* pa = static_cast<const S*>(col[0].p);
    const S* qa = static_cast<const S*>(col[0].q);
    const S* pb = static_cast<const S*>(col[1].p);
    const S* qb = static_cast<const S*>(col[1].q);
    double rp[3][4];
    C cw[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const int kc = min(t + i * TPP, K);
      rp[i][0] = double(pa[kc]); rp[i][1] = double(qa[kc]);
      rp[i][2] = double(pb[kc]); rp[i][3] = double(qb[kc]);
      cw[i] = chirp[kc];
    }
    double na = 0.0, nb = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const int k = t + i * TPP;
      col_coef(col[0], k, rp[i][0], rp[i][1], rp[i][0], rp[i][1]);
      col_coef(col[1], k, rp[i][2], rp[i][3], rp[i][2], rp[i][3]);
      if (k <= K) {
        na += rp[i][0] * rp[i][0] + rp[i][1] * rp[i][1];
        nb += rp[i][2] * rp[i][2] + rp[i][3] * rp[i][3];
      }
    }
    group_sum2<TPP>(na, nb, reinterpret_cast<double*>(smraw + fft_red(sizeof(F))));
    za = na == 0.0;
    zb = nb == 0.0;
    if (!za && !zb && isfinite(na) && isfinite(nb)) sb = ldexp(1.0, fft_scale_exp<F>((ilogb(na) - ilogb(nb)) / 2));
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const int k = t + i * TPP;
      const double PA = rp[i][0], QA = rp[i][1], PB = sb * rp[i][2], QB = sb * rp[i][3];
      if (k <= K) {
        gst[K + k] = cmul(Cx<C>::mk(F(PA + QB), F(PB - QA)), cw[i]);
        gst[K - k] = cmul(Cx<C>::mk(F(PA - QB), F(QA + PB)), cw[i]);   // k = 0: the same value
      }
    }
  }
  C v[16];
  xsync<TPP>();
#pragma unroll
  for (int e = 0; e < 16; e++) {
    const int u = t + e * TPP;
    v[e] = u <= 2 * K ? gst[u] : Cx<C>::mk(F(0), F(0));
  }
  xsync<TPP>();
  fft_passes<L, 0, true, C>(v, buf, t, tw);           // inputs u <= 2K < 6 M / 16
  // pointwise product with H, conjugated: the inverse as conj(FFT(conj(.))); four
  // elements at a time, the next four's H requested before the current four are
  // multiplied (eight loads in flight at most, within the register budget)
  {
    C hc[4], hn[4];
#pragma unroll
    for (int e = 0; e < 4; e++) hc[e] = H[t + e * TPP];
#pragma unroll
    for (int e0 = 0; e0 < 16; e0 += 4) {
      if (e0 < 12) {
#pragma unroll
        for (int e = 0; e < 4; e++) hn[e] = H[t + (e0 + 4 + e) * TPP];
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const C z = cmul(v[e0 + e], hc[e]);
        v[e0 + e] = Cx<C>::mk(z.x, -z.y);
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int e = 0; e < 4; e++) hc[e] = hn[e];
    }
  }
  fft_passes<L, 0, false, C>(v, buf, t, tw);

  // y(l) = w(l) conj(v) at l = t + e T - K
  Yt* oa = nullptr;
  Yt* ob = nullptr;
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int c = ca + h;
    Yt* o = nullptr;
    if (col[h].fam >= 0)
      o = rlz ? out + (long(var) * cpv + c) * T.jDim + T.ringOff[r]
              : out + (long(syn_slot(col[h].fam)) * T.nvars + var) * T.npts + T.ringOff[r];
    (h == 0 ? oa : ob) = o;
  }
  if (!oa) return;
  const F isb = F(1.0 / sb);                          // exact: sb is a power of two
#pragma unroll
  for (int e = 0; e < 16; e++) {
    const int l = t + e * TPP - K;
    if (l >= 0 && l < N) {
      const C y = cmul(chirp[l], Cx<C>::mk(v[e].x, -v[e].y));
      __builtin_nontemporal_store(za ? Yt(0) : Yt(y.x), oa + l);
      if (ob) __builtin_nontemporal_store(zb ? Yt(0) : Yt(y.y * isb), ob + l);
    }
    if ((e & 3) == 3) __builtin_amdgcn_sched_barrier(0);
  }
}

// every launch class in one kernel: a block's class is its ring's
template <typename S, typename Yt>
__device__ __forceinline__ void syn_fft_block(const DevTables& T, const int* __restrict__ blk, int b,
                                              const S* __restrict__ X, Yt* __restrict__ out) {
  const int r = blk[2 * b];
  if (r < 0) return;                                  // padding block (uniform)
  switch (T.fftL[r]) {
    case 12: syn_fft_body<12, S, Yt>(T, blk, b, X, out); break;
    case 11: syn_fft_body<11, S, Yt>(T, blk, b, X, out); break;
    case 10: syn_fft_body<10, S, Yt>(T, blk, b, X, out); break;
    case 9: syn_fft_body<9, S, Yt>(T, blk, b, X, out); break;
    case 8: syn_fft_body<8, S, Yt>(T, blk, b, X, out); break;
    case 7: syn_fft_body<7, S, Yt>(T, blk, b, X, out); break;
    case 6: syn_fft_body<6, S, Yt>(T, blk, b, X, out); break;
    case 5: syn_fft_body<5, S, Yt>(T, blk, b, X, out); break;
    default: syn_fft_body<4, S, Yt>(T, blk, b, X, out); break;
  }
}

// FP64 transforms (double output): 16 complex doubles per thread, three waves per SIMD
template <typename S>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_syn_fft(DevTables T, const int* __restrict__ blk, const S* __restrict__ X, double* __restrict__ out) {
  syn_fft_block<S, double>(T, blk, int(blockIdx.x), X, out);
}

template <typename S>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_syn_fft_batch(const FftJob<S, double>* __restrict__ jobs, const int* __restrict__ start, int njobs) {
  int lin;
  const FftJob<S, double>& j = jobs[batch_job(start, njobs, lin)];
  syn_fft_block<S, double>(j.T, j.blk, lin, j.X, j.out);
}

// FP32 transforms (the FP32 form's float output): half the registers and LDS
template <typename S>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SSG_FFT32_WAVES)))
k_syn_fft32(DevTables T, const int* __restrict__ blk, const S* __restrict__ X, float* __restrict__ out) {
  syn_fft_block<S, float>(T, blk, int(blockIdx.x), X, out);
}

template <typename S>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SSG_FFT32_BATCH_WAVES)))
k_syn_fft32_batch(const FftJob<S, float>* __restrict__ jobs, const int* __restrict__ start, int njobs) {
  int lin;
  const FftJob<S, float>& j = jobs[batch_job(start, njobs, lin)];
  syn_fft_block<S, float>(j.T, j.blk, lin, j.X, j.out);
}

// ----------------------------------------------------------------------------
// Ring analysis by chirp-z transforms: FBtransform! (R2HC / N, src/Fourier.jl:
// 388-395) and apply_phasefilter_forward! (src/phase_filter.jl:105-143) of every
// ring column.  With w(n) = exp(i pi n^2 / N), W^(-k l) = conj(w(k)) conj(w(l))
// w(k - l), so the transform of z = x^A + i sb x^B (two real columns of one
// variable, B scaled as in the synthesis) is
//   Z_k = conj(w(k)) [g (*) h](k),  g_l = z_l conj(w(l)) (l < N),  h(n) = w(n),
// a convolution over k - l in [-(N - 1 + K), K]; in the circular one of length
// M its filter spectrum is conj(H), H being the synthesis table of the ring.
// X^A_k = (Z_k + conj(Z_-k)) / 2 and X^B_k = (Z_k - conj(Z_-k)) / (2 i sb) give
// re_k = Re X_k = sum x cos, im_k = Im X_k = -sum x sin; the column's ring in rc
// (line order) gets (re, im) / N rotated by -k ymin, and zeros above its cap.
// Columns: RLZ the Chebyshev modes of a variable (cm[var][zb][point], paired
// within the variable), RL the variables of physical slot 1 (one per transform).
// ----------------------------------------------------------------------------
template <int L, typename S>
__device__ __forceinline__ void ana_fft_body(const DevTables& T, const int* __restrict__ blk, int b,
                                             const S* __restrict__ src, double* __restrict__ rc) {
  constexpr int M = 1 << L, TPP = M / 16;
  extern __shared__ __align__(16) unsigned char smraw[];
  const int r = blk[2 * b];
  if (r < 0) return;                                  // padding block (uniform)
  const int g = int(threadIdx.x) / TPP, t = int(threadIdx.x) - g * TPP;
  double* buf = reinterpret_cast<double*>(smraw) + g * (M + M / 16);
  double2* zst = reinterpret_cast<double2*>(buf);     // Z_k, k = -K .. K, after the transforms
  const double2* tw = reinterpret_cast<const double2*>(T.fftTw);
  const bool rlz = T.geometry == 3;
  const int nzb = T.nzb;
  const int ppv = rlz ? (nzb + 1) / 2 : 1;            // transforms per variable
  const int pair = blk[2 * b + 1] + g;
  const bool live = pair < T.nvars * ppv;
  const int var = live ? pair / ppv : 0;
  const int za0 = live && rlz ? 2 * (pair - var * ppv) : 0;   // first column's mode
  const bool hasA = live, hasB = live && rlz && za0 + 1 < nzb;
  const int N = T.ringN[r], K = T.ringRi[r];
  // the columns' ring points: RLZ cm[var][zb][point], RL physical slot 1 of var
  const S* xa = rlz ? src + (long(var) * nzb + za0) * T.jDim + T.ringOff[r] : src + long(var) * T.npts + T.ringOff[r];
  const S* xb = xa + (hasB ? T.jDim : 0);
  const double2* chirp = reinterpret_cast<const double2*>(T.fftChirp) + T.ringOff[r];
  const double2* H = reinterpret_cast<const double2*>(T.fftH) + T.fftHOff[r];

  // g_l = (x^A_l + i sb x^B_l) conj(w(l)) for l = t + e T < N, in the registers the
  // first pass reads; sb balances the two columns' norms (exact power of two)
  double2 v[16];
  double na = 0.0, nb = 0.0;
#pragma unroll
  for (int e = 0; e < 16; e++) {
    const int l = t + e * TPP;
    const int lc = min(l, N - 1);
    const double a = double(xa[lc]), bb = double(xb[lc]);
    const bool in = l < N;
    v[e] = make_double2(in && hasA ? a : 0.0, in && hasB ? bb : 0.0);
    na += v[e].x * v[e].x;
    nb += v[e].y * v[e].y;
  }
  group_sum2<TPP>(na, nb, reinterpret_cast<double*>(smraw + fft_red(sizeof(double))));
  const bool za = na == 0.0, zb = nb == 0.0;
  double sb = 1.0;
  if (!za && !zb && isfinite(na) && isfinite(nb)) sb = ldexp(1.0, (ilogb(na) - ilogb(nb)) / 2);
#pragma unroll
  for (int e0 = 0; e0 < 16; e0 += 4) {
#pragma unroll
    for (int e = e0; e < e0 + 4; e++) {
      const int l = t + e * TPP;
      const double2 w = chirp[min(l, N - 1)];
      v[e] = cmul(make_double2(v[e].x, sb * v[e].y), make_double2(w.x, -w.y));
    }
    __builtin_amdgcn_sched_barrier(0);
  }
  fft_passes<L, 0>(v, buf, t, tw);
  // times conj(H), conjugated: the inverse as conj(FFT(conj(.))); the next four
  // elements' H requested before the current four are multiplied
  {
    double2 hc[4], hn[4];
#pragma unroll
    for (int e = 0; e < 4; e++) hc[e] = H[t + e * TPP];
#pragma unroll
    for (int e0 = 0; e0 < 16; e0 += 4) {
      if (e0 < 12) {
#pragma unroll
        for (int e = 0; e < 4; e++) hn[e] = H[t + (e0 + 4 + e) * TPP];
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int e = 0; e < 4; e++) v[e0 + e] = cmul(make_double2(v[e0 + e].x, -v[e0 + e].y), hc[e]);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int e = 0; e < 4; e++) hc[e] = hn[e];
    }
  }
  fft_passes<L, 0>(v, buf, t, tw);
  // c = conj(v); Z_k = conj(w(|k|)) c(k mod M) for |k| <= K, into LDS at K + k
#pragma unroll
  for (int e = 0; e < 16; e++) {
    const int u = t + e * TPP;
    const int k = u <= K ? u : u - M;                 // u in [M - K, M): k = u - M
    if (k >= -K && k <= K) {
      const double2 w = chirp[k < 0 ? -k : k];
      const double2 c = make_double2(v[e].x, -v[e].y);
      zst[K + k] = make_double2(w.x * c.x + w.y * c.y, w.x * c.y - w.y * c.x);   // conj(w) c
    }
  }
  xsync<TPP>();
  // columns in rc's ring layout (line order): DC, then (re, im) / N rotated by
  // -k ymin at 2k - 1 and 2k
  const double invN = 1.0 / double(N);
  const double isb = 1.0 / sb;                        // exact: sb is a power of two
  const double* pc = T.phaseC + T.phaseOff[r];
  const double* psn = T.phaseS + T.phaseOff[r];
  const int km = T.kmax[var * T.iDim + r];
  const int ta = rlz ? var * nzb + za0 : var;
  double* oa = rc + long(ta) * T.b_jDim + T.coefOff[r];
  double* ob = oa + T.b_jDim;
  for (int k = t; k <= K; k += TPP) {
    const double2 zp = zst[K + k], zm = zst[K - k];
    // X^A = (Z_k + conj(Z_-k)) / 2, X^B = (Z_k - conj(Z_-k)) / (2 i sb)
    const double ar = 0.5 * (zp.x + zm.x), ai = 0.5 * (zp.y - zm.y);
    const double br = 0.5 * (zp.y + zm.y) * isb, bi = -0.5 * (zp.x - zm.x) * isb;
    const bool cap = k > km;
    const double rea = za || cap ? 0.0 : ar * invN, ima = za || cap ? 0.0 : ai * invN;
    const double reb = zb || cap ? 0.0 : br * invN, imb = zb || cap ? 0.0 : bi * invN;
    if (k == 0) {
      if (hasA) oa[0] = rea;
      if (hasB) ob[0] = reb;
    } else {
      const double cc = pc[k - 1], ss = psn[k - 1];
      double* const pa = oa + (2 * k - 1);
      double* const pb = ob + (2 * k - 1);
      if (hasA) { pa[0] = cc * rea + ss * ima; pa[1] = -ss * rea + cc * ima; }
      if (hasB) { pb[0] = cc * reb + ss * imb; pb[1] = -ss * reb + cc * imb; }
    }
  }
}

template <typename S>
__device__ __forceinline__ void ana_fft_block(const DevTables& T, const int* __restrict__ blk, int b,
                                              const S* __restrict__ src, double* __restrict__ rc) {
  const int r = blk[2 * b];
  if (r < 0) return;                                  // padding block (uniform)
  switch (T.fftL[r]) {
    case 12: ana_fft_body<12, S>(T, blk, b, src, rc); break;
    case 11: ana_fft_body<11, S>(T, blk, b, src, rc); break;
    case 10: ana_fft_body<10, S>(T, blk, b, src, rc); break;
    case 9: ana_fft_body<9, S>(T, blk, b, src, rc); break;
    case 8: ana_fft_body<8, S>(T, blk, b, src, rc); break;
    case 7: ana_fft_body<7, S>(T, blk, b, src, rc); break;
    case 6: ana_fft_body<6, S>(T, blk, b, src, rc); break;
    case 5: ana_fft_body<5, S>(T, blk, b, src, rc); break;
    default: ana_fft_body<4, S>(T, blk, b, src, rc); break;
  }
}

template <typename S>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_ana_fft(DevTables T, const int* __restrict__ blk, const S* __restrict__ src, double* __restrict__ rc) {
  ana_fft_block<S>(T, blk, int(blockIdx.x), src, rc);
}

template <typename S>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3)))
k_ana_fft_batch(const AnaFftJob<S>* __restrict__ jobs, const int* __restrict__ start, int njobs) {
  int lin;
  const AnaFftJob<S>& j = jobs[batch_job(start, njobs, lin)];
  ana_fft_block<S>(j.T, j.blk, lin, j.src, j.rc);
}

template <typename S>
void launch_ana_fft(const DevTables& T, const int* blk, int nblocks, const S* src, double* rc, hipStream_t st) {
  if (nblocks > 0)
    hipLaunchKernelGGL((k_ana_fft<S>), dim3(nblocks), dim3(256), fft_lds(sizeof(double)), st, T, blk, src, rc);
}
template void launch_ana_fft<double>(const DevTables&, const int*, int, const double*, double*, hipStream_t);
template void launch_ana_fft<float>(const DevTables&, const int*, int, const float*, double*, hipStream_t);

template <typename S>
void launch_ana_fft_batch(const AnaFftJob<S>* jobs, const int* start, int njobs, int nblocks, hipStream_t st) {
  hipLaunchKernelGGL((k_ana_fft_batch<S>), dim3(nblocks), dim3(256), fft_lds(sizeof(double)), st, jobs, start, njobs);
}
template void launch_ana_fft_batch<double>(const AnaFftJob<double>*, const int*, int, int, hipStream_t);
template void launch_ana_fft_batch<float>(const AnaFftJob<float>*, const int*, int, int, hipStream_t);

FftSynBlocks fft_blocks(const std::vector<int>& fftL, int iDim, int npair) {
  FftSynBlocks B;
  for (int L = kFftMaxLog; L >= 4; L--) {
    const int G = kFftMaxM >> L;                      // pairs per block
    const int nb = (npair + G - 1) / G;               // blocks per ring
    // rings of this class, largest first (for ring load balance), dealt to the
    // eight XCD lanes by least load; a ring's blocks stay in one lane (one
    // XCD's L2 holds the ring's tables)
    std::vector<std::vector<int>> lane(8);
    for (int r = iDim - 1; r >= 0; r--) {
      if (fftL[r] != L) continue;
      int x = 0;
      for (int y = 1; y < 8; y++) if (lane[y].size() < lane[x].size()) x = y;
      for (int i = 0; i < nb; i++) { lane[x].push_back(r); lane[x].push_back(i * G); }
    }
    size_t len = 0;
    for (auto& l : lane) len = std::max(len, l.size() / 2);
    for (size_t i = 0; i < len; i++)
      for (int x = 0; x < 8; x++) {
        if (i < lane[x].size() / 2) { B.blk.push_back(lane[x][2 * i]); B.blk.push_back(lane[x][2 * i + 1]); }
        else { B.blk.push_back(-1); B.blk.push_back(0); }
      }
  }
  B.nblocks = int(B.blk.size() / 2);
  return B;
}

template <typename S, typename Yt>
void launch_syn_fft(const DevTables& T, const int* blk, int nblocks, const S* X, Yt* out, hipStream_t st) {
  if (nblocks <= 0) return;
  if constexpr (sizeof(Yt) == sizeof(float))
    hipLaunchKernelGGL((k_syn_fft32<S>), dim3(nblocks), dim3(256), fft_lds(sizeof(float)), st, T, blk, X, out);
  else
    hipLaunchKernelGGL((k_syn_fft<S>), dim3(nblocks), dim3(256), fft_lds(sizeof(double)), st, T, blk, X, out);
}

template <typename S, typename Yt>
void launch_syn_fft_batch(const FftJob<S, Yt>* jobs, const int* start, int njobs, int nblocks, hipStream_t st) {
  if constexpr (sizeof(Yt) == sizeof(float))
    hipLaunchKernelGGL((k_syn_fft32_batch<S>), dim3(nblocks), dim3(256), fft_lds(sizeof(float)), st, jobs, start,
                       njobs);
  else
    hipLaunchKernelGGL((k_syn_fft_batch<S>), dim3(nblocks), dim3(256), fft_lds(sizeof(double)), st, jobs, start,
                       njobs);
}
template void launch_syn_fft_batch<double, double>(const FftJob<double, double>*, const int*, int, int, hipStream_t);
template void launch_syn_fft_batch<float, float>(const FftJob<float, float>*, const int*, int, int, hipStream_t);

template void launch_syn_fft<double, double>(const DevTables&, const int*, int, const double*, double*, hipStream_t);
template void launch_syn_fft<float, float>(const DevTables&, const int*, int, const float*, float*, hipStream_t);

// RLZ inverse Chebyshev stage (src/Chebyshev.jl:473-535, 636-722 composed into the
// plan's (kDim x b_kDim) matrices; src/transforms_cylindrical.jl:780-844): every
// ring point's b_kDim modes of each family are contracted into kDim heights on
// MFMA: family 0 -> slots 1, 6, 7 (values, d/dz, d2/dz2), 1 -> 4, 2 -> 5, 3 -> 2,
// 4 -> 3 (values).  A block stages its variable's three matrices in LDS once and
// takes kChebPts consecutive ring points, 16 per wave at a time; the A operand
// (one family's modes of 16 points) of the next family is requested while the
// current one is contracted.  NKS = b_kDim padded to 4, over 4.
static constexpr int kChebPts = 256;
// LDS row stride of the matrices (doubles): padded to 8 or 16 mod 32 against bank
// conflicts between the k rows of an MFMA B operand; the pad of 8 keeps 64 heights
// within 80 KB, two blocks per CU
__host__ __device__ inline int syn_cheb_ldm(int kzp) { return kzp % 32 == 16 ? kzp : kzp + 8; }

// order of one wave's LDS accesses (its output tile)
__device__ __forceinline__ void cheb_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename R, int NKS>
__device__ __forceinline__ void syn_cheb_body(const DevTables& T, int lin, const R* __restrict__ Y,
                                              R* __restrict__ phys) {
  typedef typename Mfma<R>::v4 v4;
  extern __shared__ __align__(16) unsigned char smraw[];
  R* ml = reinterpret_cast<R*>(smraw);                 // [3 bkp][ldm]: ml[m bkp + zb][z] = M_m[z][zb]
  const int bk = T.nzb, kDim = T.kDim, bkp = 4 * NKS, kzp = syn_kzp(kDim), nzt = kzp >> 4, ldm = syn_cheb_ldm(kzp);
  const int nch = (T.jDim + kChebPts - 1) / kChebPts;
  const int v = lin / nch;
  if (v >= T.nvars) return;                           // block-uniform
  const int pbase = (lin - v * nch) * kChebPts;
  const R* Mg = rtab<R>(T).invt + long(v) * 3 * bkp * kzp;
  for (int e = int(threadIdx.x); e < 3 * bkp * kzp; e += 256) {
    const int row = e / kzp;
    ml[row * ldm + (e - row * kzp)] = Mg[e];
  }
  __syncthreads();
  const int lane = int(threadIdx.x) & 63, w = int(threadIdx.x) >> 6, kq = lane >> 4;
  const long jD = T.jDim;
  const R* Yv = Y + long(v) * 5 * bk * jD;
  R* physv = phys + long(v) * T.npts;
  const R* mlane = ml + kq * ldm + (lane & 15);
  // float outputs go out through a per-wave tile of 16 points x kDim past the matrices
  R* stile = ml + 3 * bkp * ldm + w * 16 * kzp;
  for (int tile = 0; tile < kChebPts / 64; tile++) {
    const int p0 = pbase + tile * 64 + w * 16;
    if (p0 >= jD) break;                              // wave-uniform
    const long pa = min(long(p0 + (lane & 15)), jD - 1);
    auto load_a = [&](int f, R (&dst)[NKS]) {
#pragma unroll
      for (int ks = 0; ks < NKS; ks++) {
        const int zb = 4 * ks + kq;
        const R x = Yv[(long(f) * bk + min(zb, bk - 1)) * jD + pa];
        dst[ks] = zb < bk ? x : R(0);
      }
    };
    R a[NKS], an[NKS];
    load_a(0, a);
#pragma unroll
    for (int f = 0; f < 5; f++) {
      if (f < 4) load_a(f + 1, an);
      const int nm = f == 0 ? 3 : 1;
      for (int m = 0; m < nm; m++) {
        const int slot = f == 0 ? (m == 0 ? 0 : m == 1 ? 5 : 6) : (f == 1 ? 3 : f == 2 ? 4 : f == 3 ? 1 : 2);
        R* outs = physv + long(slot) * T.nvars * T.npts;
        const R* mm = mlane + m * bkp * ldm;
        for (int zt = 0; zt < nzt; zt++) {
          v4 o = v4{0, 0, 0, 0};
#pragma unroll
          for (int ks = 0; ks < NKS; ks++) o = Mfma<R>::mma(a[ks], mm[4 * ks * ldm + zt * 16], o);
          const int z = zt * 16 + (lane & 15);
          if (z >= kDim) continue;
#pragma unroll
          for (int rr = 0; rr < 4; rr++) {
            if constexpr (sizeof(R) == sizeof(float)) {
              stile[Mfma<R>::row(lane, rr) * kDim + z] = o[rr];
            } else {
              const long p = p0 + Mfma<R>::row(lane, rr);
              if (p < jD) outs[p * kDim + z] = o[rr];
            }
          }
        }
        if constexpr (sizeof(R) == sizeof(float)) {
          // the wave's 16 points are 16 kDim consecutive outputs: 16-byte stores
          cheb_wave_sync();
          const int nv = int(min(long(16), jD - p0)) * kDim;
          R* dst = outs + long(p0) * kDim;
          const int nv4 = (reinterpret_cast<size_t>(dst) & 15) == 0 ? nv & ~3 : 0;   // 16-byte aligned part
          for (int i = lane; 4 * i < nv4; i += 64)
            *reinterpret_cast<float4*>(dst + 4 * i) = *reinterpret_cast<const float4*>(stile + 4 * i);
          for (int i = nv4 + lane; i < nv; i += 64) dst[i] = stile[i];
          cheb_wave_sync();
        }
      }
      if (f < 4) {
#pragma unroll
        for (int ks = 0; ks < NKS; ks++) a[ks] = an[ks];
      }
    }
  }
}

#define SSG_CHEB_CASES(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12)

// the b_kDim of a plan in one of three kernels by NKS: 1-4, 5-8, 9-12 (class C
// takes NKS 4C - 3 .. 4C), so that a kernel's registers are those of its
// largest NKS (116, 164 and 224 VGPRs in FP64: four, three and two waves per SIMD)
template <typename R, int C>
__device__ __forceinline__ void syn_cheb_block(const DevTables& T, int lin, const R* __restrict__ Y,
                                               R* __restrict__ phys) {
  switch (syn_bkp(T.nzb) / 4) {
#define SSG_CASE(n) \
  case n:           \
    if constexpr (n > 4 * C - 4 && n <= 4 * C) syn_cheb_body<R, n>(T, lin, Y, phys); \
    break;
    SSG_CHEB_CASES(SSG_CASE)
#undef SSG_CASE
    default: break;
  }
}

template <typename R, int C>
__global__ void __launch_bounds__(256) k_syn_cheb(DevTables T, const R* __restrict__ Y, R* __restrict__ phys) {
  syn_cheb_block<R, C>(T, int(blockIdx.x), Y, phys);
}

template <typename R, int C>
__global__ void __launch_bounds__(256) k_syn_cheb_batch(const ChebSynJob<R>* __restrict__ jobs,
                                                        const int* __restrict__ start, int njobs) {
  int lin;
  const ChebSynJob<R>& j = jobs[batch_job(start, njobs, lin)];
  syn_cheb_block<R, C>(j.T, lin, j.Y, j.phys);
}

int syn_cheb_class(const DevTables& T) { return (syn_bkp(T.nzb) / 4 + 3) / 4; }
int syn_cheb_blocks(const DevTables& T) { return (T.jDim + kChebPts - 1) / kChebPts * T.nvars; }
size_t syn_cheb_lds(const DevTables& T, size_t esize) {
  const size_t mats = esize * 3 * size_t(syn_bkp(T.nzb)) * size_t(syn_cheb_ldm(syn_kzp(T.kDim)));
  return esize == sizeof(float) ? mats + esize * 4 * 16 * size_t(syn_kzp(T.kDim)) : mats;
}

template <typename R>
void launch_syn_cheb(const DevTables& T, const R* Y, R* phys, hipStream_t st) {
  const dim3 g(syn_cheb_blocks(T));
  const size_t lds = syn_cheb_lds(T, sizeof(R));
  switch (syn_cheb_class(T)) {
    case 1: hipLaunchKernelGGL((k_syn_cheb<R, 1>), g, dim3(256), lds, st, T, Y, phys); break;
    case 2: hipLaunchKernelGGL((k_syn_cheb<R, 2>), g, dim3(256), lds, st, T, Y, phys); break;
    default: hipLaunchKernelGGL((k_syn_cheb<R, 3>), g, dim3(256), lds, st, T, Y, phys); break;
  }
}
template void launch_syn_cheb<double>(const DevTables&, const double*, double*, hipStream_t);
template void launch_syn_cheb<float>(const DevTables&, const float*, float*, hipStream_t);

template <typename R>
void launch_syn_cheb_batch(const ChebSynJob<R>* jobs, const int* start, int njobs, int nblocks, size_t lds, int cls,
                           hipStream_t st) {
  switch (cls) {
    case 1: hipLaunchKernelGGL((k_syn_cheb_batch<R, 1>), dim3(nblocks), dim3(256), lds, st, jobs, start, njobs); break;
    case 2: hipLaunchKernelGGL((k_syn_cheb_batch<R, 2>), dim3(nblocks), dim3(256), lds, st, jobs, start, njobs); break;
    default: hipLaunchKernelGGL((k_syn_cheb_batch<R, 3>), dim3(nblocks), dim3(256), lds, st, jobs, start, njobs); break;
  }
}
template void launch_syn_cheb_batch<double>(const ChebSynJob<double>*, const int*, int, int, size_t, int, hipStream_t);
template void launch_syn_cheb_batch<float>(const ChebSynJob<float>*, const int*, int, int, size_t, int, hipStream_t);

}  // namespace ssg
