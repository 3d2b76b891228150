// Uplink FFT stage of the detection path: the reference's UplinkFFT and
// PilotFFT operators (src/operators/fft/fft_op.cu:143-172; cuFFT C2C forward
// between a load and a store kernel, fft_callback.cuh:20-108), fused here
// into one kernel per launch. For OFDM symbol s of antenna b:
//   x[n]  = (re, im) / (max + 1) of the received integer pairs (int16/int8),
//   X[k]  = sum_n x[n] exp(-2 pi i k n / ca)           (unnormalised),
// then either
//   uplink: y[s][b][i] = X[(i + start + ca/2) % ca], i < num  (the [B][N]
//           layout the equalize reads), or
//   pilot:  the CSI element of user ue_start + s*sg + i % sg in group i / sg,
//           conj(X[col(i)]) * pilot[i], into [G][B][U] (A = H^H, the layout
//           the ZF kernels read); users >= U are skipped. Pilot symbol s of a
//           launch carries the s-th set of sg users.
// The downlink IFFT (DownlinkIFFT, fft_op.cu:168-174, ifft_callback.cuh:21-59)
// runs the same kernel the other way: complex rows [S][B][ca] (the precode
// output), the unnormalised inverse DFT, each part scaled by (max + 1) / ca,
// rounded to nearest even and saturated into integer (re, im) pairs.
//
// One workgroup of 256 threads per (antenna, symbol): samples are loaded
// coalesced and scaled into LDS in bit-reversed order, ca/2 twiddles are
// computed once into LDS (sincospif of exact arguments for the first octant,
// the rest by symmetry), and log2(ca)
// radix-2 decimation-in-time stages run in place, four per barrier (radix
// 2^4, sixteen elements per thread in registers); the store reads the
// shifted bins straight from LDS. Sizes: powers of
// two from 16 to 4096 (51 KB of LDS at 4096); 2048 and 4096 points run an
// instantiation with the size fixed at compile time: the LDS addresses are
// immediate offsets, the first pass runs in registers on the samples as
// loaded, u - w x is formed as 2u - (u + w x), and the last pass stores its
// bins to the outputs directly.
#include "eq_core.h"
#include "internal.h"

namespace mega {
namespace detail {
namespace {

constexpr int kFftThreads = 256;
constexpr uint32_t kFftMax = 4096;

// LDS slot of element i of the samples and of the twiddle table: one pad
// element every 32, so power-of-two strides across lanes (the samples' in the
// early passes, the twiddles' in the later ones: up to 128 entries apart)
// spread over the banks.
__host__ __device__ constexpr uint32_t pidx(uint32_t i) { return i + (i >> 5); }

// u + w v and u - w v (complex)
__device__ __forceinline__ float2 bfly_add(float2 u, float2 w, float2 v) {
  return make_float2(u.x + (w.x * v.x - w.y * v.y), u.y + (w.x * v.y + w.y * v.x));
}
__device__ __forceinline__ float2 bfly_sub(float2 u, float2 w, float2 v) {
  return make_float2(u.x - (w.x * v.x - w.y * v.y), u.y - (w.x * v.y + w.y * v.x));
}

// Stages st .. st+R-1 of the in-place decimation-in-time FFT over sx (input
// in bit-reversed order), 2^R elements per thread.
template <int R>
__device__ __forceinline__ void fft_pass(float2* sx, const float2* tw, uint32_t N, uint32_t L,
                                         uint32_t st) {
  constexpr int M = 1 << R;
  const uint32_t half = 1u << st;
  for (uint32_t j = threadIdx.x; j < (N >> R); j += kFftThreads) {
    const uint32_t k = j & (half - 1);
    const uint32_t i0 = ((j >> st) << (st + R)) + k;
    float2 v[M];
#pragma unroll
    for (int m = 0; m < M; ++m) v[m] = sx[pidx(i0 + m * half)];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int sp = 1 << q;
#pragma unroll
      for (int m = 0; m < M; ++m) {
        if (m & sp) continue;
        const float2 w = tw[pidx((k + (uint32_t)(m & (sp - 1)) * half) << (L - 1 - st - q))];
        const float2 u = v[m], x = v[m + sp];
        v[m] = bfly_add(u, w, x);
        v[m + sp] = bfly_sub(u, w, x);
      }
    }
#pragma unroll
    for (int m = 0; m < M; ++m) sx[pidx(i0 + m * half)] = v[m];
  }
}

// The same pass with the transform size 2^LC and the first stage ST known
// at compile time (the 2048- and 4096-point symbols): thread j's elements
// and twiddles sit at a per-thread base plus compile-time offsets, so the
// LDS accesses take immediate offsets instead of per-element address
// arithmetic. pass_offsets_hold checks the offset identities for every
// thread and element of each instantiation at compile time.
template <int R, int LC, int ST>
constexpr bool pass_offsets_hold() {
  constexpr uint32_t N = 1u << LC, half = 1u << ST, J = N >> R;
  for (uint32_t j = 0; j < J; ++j) {
    const uint32_t k = j & (half - 1), i0 = ((j >> ST) << (ST + R)) + k;
    for (uint32_t m = 0; m < (1u << R); ++m)
      if (pidx(i0 + m * half) != pidx(i0) + m * half + ((m * half) >> 5)) return false;
    for (int q = 0; q < R; ++q) {
      const uint32_t sh = LC - 1 - ST - q;
      for (uint32_t c = 0; c < (1u << q); ++c) {
        const uint32_t C = (c * half) << sh;
        if (pidx((k << sh) + C) != pidx(k << sh) + C + (C >> 5)) return false;
      }
    }
  }
  return true;
}

// The butterflies: u + w x takes four FMAs and u - w x = 2u - (u + w x) two
// more.
__device__ __forceinline__ void bfly_fma(float2& u, float2& x, float2 w) {
  const float2 a = make_float2(__builtin_fmaf(w.x, x.x, __builtin_fmaf(-w.y, x.y, u.x)),
                               __builtin_fmaf(w.x, x.y, __builtin_fmaf(w.y, x.x, u.y)));
  x = make_float2(__builtin_fmaf(2.f, u.x, -a.x), __builtin_fmaf(2.f, u.y, -a.y));
  u = a;
}

// With LAST (the pass that ends at stage LC - 1: there j = k and the
// elements are bins k + m half in natural order) the results go to
// emit(bin, value) instead of back to LDS.
template <int R, int LC, int ST, bool LAST = false, class Emit = int>
__device__ __forceinline__ void fft_pass_ct(float2* sx, const float2* tw, Emit&& emit = 0) {
  static_assert(pass_offsets_hold<R, LC, ST>(), "LDS offset identities");
  constexpr int M = 1 << R;
  constexpr uint32_t half = 1u << ST, J = (1u << LC) >> R;
  static_assert(J <= (uint32_t)kFftThreads, "one element block per thread");
  const uint32_t j = threadIdx.x;
  if (J < (uint32_t)kFftThreads && j >= J) return;
  const uint32_t k = j & (half - 1);
  float2* const xb = sx + pidx(((j >> ST) << (ST + R)) + k);
  float2 v[M];
#pragma unroll
  for (int m = 0; m < M; ++m) v[m] = xb[m * half + ((m * half) >> 5)];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int sp = 1 << q;
    const uint32_t sh = LC - 1 - ST - q;
    const float2* const tb = tw + pidx(k << sh);
#pragma unroll
    for (int m = 0; m < M; ++m) {
      if (m & sp) continue;
      const uint32_t c = (uint32_t)(m & (sp - 1));  // W_{2^(ST+q+1)}^(k + c half)
      const uint32_t C = (c * half) << sh;
      bfly_fma(v[m], v[m + sp], tb[C + (C >> 5)]);
    }
  }
  if constexpr (LAST) {
    static_assert(ST + R == LC, "the last pass ends the transform");
#pragma unroll
    for (int m = 0; m < M; ++m) emit(k + m * half, v[m]);
  } else {
#pragma unroll
    for (int m = 0; m < M; ++m) xb[m * half + ((m * half) >> 5)] = v[m];
  }
}

// The first pass (stages 0-3) of the compile-time sizes runs in registers
// on the samples as loaded: thread j < N/16 takes elements 16 j + m of the
// bit-reversed sequence, the samples rev_{LC-4}(j) + 2^(LC-4) rev4(m), and
// its twiddles W_{2^(q+1)}^c are the constants exp(-+i pi e / 8), e =
// c 2^(3-q): 1 and the quarter turn (e = 0, 4) as a copy and a swap, the
// others as FMAs. The result goes to LDS at 16 j + m (pad included).
constexpr float kC1 = 0.923879532511286756f;  // cos(pi/8)
constexpr float kS1 = 0.382683432365089772f;  // sin(pi/8)
constexpr float kR2 = 0.707106781186547524f;  // cos(pi/4)
constexpr float kCos16[8] = {1.f, kC1, kR2, kS1, 0.f, -kS1, -kR2, -kC1};
constexpr float kSin16[8] = {0.f, kS1, kR2, kC1, 1.f, kC1, kR2, kS1};

constexpr uint32_t rev_bits(uint32_t x, int bits) {
  uint32_t r = 0;
  for (int i = 0; i < bits; ++i) r |= ((x >> i) & 1u) << (bits - 1 - i);
  return r;
}
// The gather and the LDS slots of the first pass, checked for every thread
// and element: element 16 j + m of the bit-reversed sequence is sample
// rev_{LC-4}(j) + 2^(LC-4) rev4(m), and its slot is pidx(16 j) + m.
template <int LC>
constexpr bool first_pass_holds() {
  if constexpr (LC >= 8) {  // the sizes with a first pass in registers
    for (uint32_t j = 0; j < ((1u << LC) >> 4); ++j)
      for (uint32_t m = 0; m < 16; ++m) {
        if (rev_bits(16 * j + m, LC) != rev_bits(j, LC - 4) + (rev_bits(m, 4) << (LC - 4)))
          return false;
        if (pidx(16 * j + m) != pidx(16 * j) + m) return false;
      }
  }
  return true;
}

template <bool INV>
__device__ __forceinline__ void fft_first_pass(float2 (&v)[16]) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int sp = 1 << q;
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      if (m & sp) continue;
      const int e = (m & (sp - 1)) << (3 - q);
      const float2 u = v[m], x = v[m + sp];
      if (e == 0 || e == 4) {
        const float2 t = e == 0 ? x : INV ? make_float2(-x.y, x.x) : make_float2(x.y, -x.x);
        v[m] = make_float2(u.x + t.x, u.y + t.y);
        v[m + sp] = make_float2(u.x - t.x, u.y - t.y);
      } else {
        bfly_fma(v[m], v[m + sp], make_float2(kCos16[e], INV ? kSin16[e] : -kSin16[e]));
      }
    }
  }
}

template <typename Q>
__device__ __forceinline__ float2 load_iq(const Q* p, uint64_t n, float inv) {
  return make_float2((float)p[2 * n] * inv, (float)p[2 * n + 1] * inv);
}

enum FftMode { kUplink = 0, kPilot = 1, kIfft = 2 };

template <typename Q>
__device__ __forceinline__ Q quantize(float v) {
  constexpr float hi = sizeof(Q) == 2 ? 32767.f : 127.f;
  constexpr float lo = sizeof(Q) == 2 ? -32768.f : -128.f;
  if (v > hi) return (Q)hi;
  if (v < lo) return (Q)lo;
  return (Q)(int)__builtin_rintf(v);
}

// LC: log2 of the transform size when it is fixed at compile time (11, 12),
// else 0 and the size comes from the arguments.
template <typename Q, int MODE, int LC>
__global__ __launch_bounds__(kFftThreads) void fft_kernel(const FftArgs a) {
  __shared__ float2 sx[kFftMax + kFftMax / 32];
  __shared__ float2 tw[kFftMax / 2 + kFftMax / 64];
  const uint32_t b = blockIdx.x, s = blockIdx.y;
  const uint32_t N = LC ? 1u << LC : a.ca, L = LC ? LC : a.log2ca, H = N / 2;
  const uint64_t row = (uint64_t)s * a.in_sym + (uint64_t)b * N;
  constexpr uint32_t J0 = LC ? (1u << LC) >> 4 : 0;  // threads of the first pass
  static_assert(LC == 0 || (LC >= 8 && J0 <= (uint32_t)kFftThreads && first_pass_holds<LC>()),
                "first pass");
  float2 v0[16];
  if constexpr (LC != 0) {
    if (J0 == (uint32_t)kFftThreads || threadIdx.x < J0) {
      const uint32_t rj = __builtin_bitreverse32(threadIdx.x) >> (32 - (LC - 4));
      const float inv = 1.0f / a.qscale;
#pragma unroll
      for (uint32_t m = 0; m < 16; ++m) {
        const uint32_t n = rj + ((__builtin_bitreverse32(m) >> 28) << (LC - 4));
        if constexpr (MODE == kIfft)
          v0[m] = static_cast<const float2*>(a.in)[row + n];
        else
          v0[m] = load_iq(static_cast<const Q*>(a.in) + 2 * row, n, inv);
      }
    }
  } else if constexpr (MODE == kIfft) {
    const float2* in = static_cast<const float2*>(a.in) + row;
    for (uint32_t n = threadIdx.x; n < N; n += kFftThreads)
      sx[pidx(__builtin_bitreverse32(n) >> (32 - L))] = in[n];
  } else {
    const Q* in = static_cast<const Q*>(a.in) + 2 * row;
    const float inv = 1.0f / a.qscale;
    for (uint32_t n = threadIdx.x; n < N; n += kFftThreads)
      sx[pidx(__builtin_bitreverse32(n) >> (32 - L))] = load_iq(in, n, inv);
  }
  // twiddles exp(-+2 pi i k / N) from the first octant: with c, s = cos,
  // sin of 2 pi k / N (k <= N/8), W^k = (c, -+s), W^(N/4-k) = (s, -+c),
  // W^(N/4+k) = (-s, -+c), W^(N/2-k) = (-c, -+s); coinciding indices fall
  // to the same thread
  const float sgn = MODE == kIfft ? 1.0f : -1.0f;
  for (uint32_t k = threadIdx.x; k <= N / 8; k += kFftThreads) {
    float sn, cs;
    sincospif(2.0f * (float)k / (float)N, &sn, &cs);
    tw[pidx(k)] = make_float2(cs, sgn * sn);
    tw[pidx(N / 4 - k)] = make_float2(sn, sgn * cs);
    tw[pidx(N / 4 + k)] = make_float2(-sn, sgn * cs);
    if (k > 0) tw[pidx(H - k)] = make_float2(-cs, sgn * sn);
  }
  if constexpr (LC != 0) {
    if (J0 == (uint32_t)kFftThreads || threadIdx.x < J0) {
      fft_first_pass<MODE == kIfft>(v0);
      float2* const xb = sx + pidx(16 * threadIdx.x);
#pragma unroll
      for (int m = 0; m < 16; ++m) xb[m] = v0[m];  // pidx(16 j + m) = pidx(16 j) + m
    }
  }
  __syncthreads();
  // R stages per barrier (radix 2^R, R = 4, the last pass 1 to 3): a thread
  // takes the 2^R elements x[i + m h] (h = 2^st, i = its block's start + k,
  // k < h) and runs stages st .. st+R-1 in registers; stage st+q pairs
  // m and m + 2^q with W_{2^(st+q+1)}^(k + (m mod 2^q) h).
  // bin X[bb] of the transform to its output: y[i] for i = (bb - shift) mod
  // N below num (uplink), the CSI element of that subcarrier (pilot), or the
  // quantized sample bb (IFFT)
  const uint32_t shift = a.start + H;  // bins (i + shift) mod N, N a power of two
  auto emit = [&](uint32_t bb, float2 e) {
    if constexpr (MODE == kIfft) {
      Q* o = static_cast<Q*>(a.qout) + 2 * ((uint64_t)s * a.out_sym + (uint64_t)b * N);
      const float scale = a.qscale / (float)N;
      o[2 * bb] = quantize<Q>(e.x * scale);
      o[2 * bb + 1] = quantize<Q>(e.y * scale);
    } else {
      const uint32_t i = (bb - shift) & (N - 1);
      if (i >= a.num) return;
      if constexpr (MODE == kUplink) {
        a.out[(uint64_t)s * a.out_sym + (uint64_t)b * a.num + i] = e;
      } else {
        const uint32_t gi = div_by(i, a.sg, a.sg_inv);  // group of subcarrier i
        const uint32_t u = a.ue_start + s * a.sg + (i - gi * a.sg);
        if (u >= a.U) return;
        const float2 p = a.pilot[i];
        const uint64_t o = (uint64_t)gi * a.B * a.U + (uint64_t)b * a.U + u;
        a.out[(uint64_t)s * a.out_sym + o] =
            make_float2(e.x * p.x + e.y * p.y, e.x * p.y - e.y * p.x);
      }
    }
  };
  if constexpr (LC != 0) {
    // middle passes through LDS, the last one straight to the outputs
    constexpr int NP = (LC + 3) / 4;  // passes: the first (in registers) .. the last
    static_for<1, NP - 1>([&](auto pc) {
      fft_pass_ct<4, LC, 4 * decltype(pc)::value>(sx, tw);
      __syncthreads();
    });
    constexpr int RL = LC - 4 * (NP - 1);
    fft_pass_ct<RL, LC, 4 * (NP - 1), true>(sx, tw, emit);
    return;
  }
  uint32_t st = LC != 0 ? L : 0;
  for (; st + 4 <= L; st += 4) {
    fft_pass<4>(sx, tw, N, L, st);
    __syncthreads();
  }
  if (st + 3 == L) {
    fft_pass<3>(sx, tw, N, L, st);
    __syncthreads();
  } else if (st + 2 == L) {
    fft_pass<2>(sx, tw, N, L, st);
    __syncthreads();
  } else if (st + 1 == L) {
    fft_pass<1>(sx, tw, N, L, st);
    __syncthreads();
  }
  if constexpr (MODE == kIfft) {
    Q* o = static_cast<Q*>(a.qout) + 2 * ((uint64_t)s * a.out_sym + (uint64_t)b * N);
    const float scale = a.qscale / (float)N;
    for (uint32_t n = threadIdx.x; n < N; n += kFftThreads) {
      const float2 v = sx[pidx(n)];
      o[2 * n] = quantize<Q>(v.x * scale);
      o[2 * n + 1] = quantize<Q>(v.y * scale);
    }
  } else if constexpr (MODE == kUplink) {
    float2* y = a.out + (uint64_t)s * a.out_sym + (uint64_t)b * a.num;
    for (uint32_t i = threadIdx.x; i < a.num; i += kFftThreads) y[i] = sx[pidx((i + shift) & (N - 1))];
  } else {
    for (uint32_t i = threadIdx.x; i < a.num; i += kFftThreads) {
      const uint32_t u = a.ue_start + s * a.sg + i % a.sg;
      if (u >= a.U) continue;
      const float2 e = sx[pidx((i + shift) & (N - 1))], p = a.pilot[i];
      const uint64_t o = (uint64_t)(i / a.sg) * a.B * a.U + (uint64_t)b * a.U + u;
      a.out[(uint64_t)s * a.out_sym + o] =
          make_float2(e.x * p.x + e.y * p.y, e.x * p.y - e.y * p.x);
    }
  }
}

template <int MODE, int LC>
hipError_t launch_lc(const FftArgs& a, hipStream_t st) {
  const dim3 grid(a.B, a.S), block(kFftThreads);
  if (a.qbits == 16)
    hipLaunchKernelGGL((fft_kernel<int16_t, MODE, LC>), grid, block, 0, st, a);
  else
    hipLaunchKernelGGL((fft_kernel<int8_t, MODE, LC>), grid, block, 0, st, a);
  return hipGetLastError();
}

template <int MODE>
hipError_t launch_q(const FftArgs& a, hipStream_t st) {
  if (a.log2ca == 12) return launch_lc<MODE, 12>(a, st);
  if (a.log2ca == 11) return launch_lc<MODE, 11>(a, st);
  return launch_lc<MODE, 0>(a, st);
}

}  // namespace

hipError_t launch_fft(const FftArgs& a0, int mode, hipStream_t st, const char** why) {
  FftArgs a = a0;
  if (a.B == 0 || a.S == 0 || (mode != kIfft && a.num == 0)) return hipSuccess;
  if (a.ca < 16 || a.ca > kFftMax || (a.ca & (a.ca - 1)) != 0) {
    *why = "ofdm_ca must be a power of two from 16 to 4096";
    return hipErrorInvalidValue;
  }
  if (mode == kIfft) a.num = a.ca;  // every bin
  if (a.num > a.ca || a.start >= a.ca) {
    *why = "ofdm_num and ofdm_start must be below ofdm_ca";
    return hipErrorInvalidValue;
  }
  if (a.qbits != 16 && a.qbits != 8) {
    *why = "samples must be 16- or 8-bit integer pairs";
    return hipErrorInvalidValue;
  }
  if (a.B > 0x7fffffffu || a.S > 65535) {
    *why = "grid too large";
    return hipErrorInvalidValue;
  }
  if (mode == kPilot && (a.sg == 0 || a.U == 0)) return hipSuccess;
  a.log2ca = 31 - __builtin_clz(a.ca);
  a.sg_inv = div_inv(a.sg);
  a.qscale = a.qbits == 16 ? 32768.0f : 128.0f;
  if (mode == kIfft) return launch_q<kIfft>(a, st);
  return mode == kPilot ? launch_q<kPilot>(a, st) : launch_q<kUplink>(a, st);
}

}  // namespace detail
}  // namespace mega
