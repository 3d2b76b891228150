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
// instantiation with the size fixed at compile time (the same arithmetic,
// with the LDS addresses folded into immediate offsets).
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

// The butterflies: in the first pass (k = 0) the twiddles of c = 0 are 1 and
// those of a quarter turn are (0, -+1) (the table's own values there), so
// w x is x or a swap; otherwise u + w x takes four FMAs and u - w x =
// 2u - (u + w x) two more.
template <int R, int LC, int ST, bool INV>
__device__ __forceinline__ void fft_pass_ct(float2* sx, const float2* tw) {
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
      const float2 u = v[m], x = v[m + sp];
      if (ST == 0 && (c == 0 || 2 * c == (uint32_t)sp)) {
        // k = 0: w = 1, or the quarter turn w = (0, -+1) of exp(-+2 pi i / 4)
        const float2 t = c == 0 ? x : INV ? make_float2(-x.y, x.x) : make_float2(x.y, -x.x);
        v[m] = make_float2(u.x + t.x, u.y + t.y);
        v[m + sp] = make_float2(u.x - t.x, u.y - t.y);
      } else {
        const uint32_t C = (c * half) << sh;
        const float2 w = tb[C + (C >> 5)];
        const float2 a = make_float2(__builtin_fmaf(w.x, x.x, __builtin_fmaf(-w.y, x.y, u.x)),
                                     __builtin_fmaf(w.x, x.y, __builtin_fmaf(w.y, x.x, u.y)));
        v[m] = a;
        v[m + sp] = make_float2(__builtin_fmaf(2.f, u.x, -a.x), __builtin_fmaf(2.f, u.y, -a.y));
      }
    }
  }
#pragma unroll
  for (int m = 0; m < M; ++m) xb[m * half + ((m * half) >> 5)] = v[m];
}

// Bit-reversed LDS slot of sample tid + 256 i for N = 2^LC: the reversed
// index is (rev8(tid) << (LC - 8)) | rev(i), and the pad of the sum is the
// pad of the first part (checked at compile time).
template <int LC>
constexpr bool load_offsets_hold() {
  for (uint32_t t = 0; t < 256; ++t)
    for (uint32_t i = 0; i < (1u << (LC - 8)); ++i) {
      const uint32_t base = t << (LC - 8);
      if (pidx(base + i) != pidx(base) + i) return false;
    }
  return true;
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
  if constexpr (LC != 0) {
    static_assert(LC >= 8 && load_offsets_hold<LC>(), "load offsets");
    float2* const xb = sx + pidx((__builtin_bitreverse32(threadIdx.x) >> 24) << (LC - 8));
    const float inv = 1.0f / a.qscale;
#pragma unroll
    for (uint32_t i = 0; i < (1u << (LC - 8)); ++i) {
      const uint32_t n = threadIdx.x + kFftThreads * i;
      const uint32_t ri = __builtin_bitreverse32(i) >> (32 - (LC - 8));
      if constexpr (MODE == kIfft)
        xb[ri] = static_cast<const float2*>(a.in)[row + n];
      else
        xb[ri] = load_iq(static_cast<const Q*>(a.in) + 2 * row, n, inv);
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
  __syncthreads();
  // R stages per barrier (radix 2^R, R = 4, the last pass 1 to 3): a thread
  // takes the 2^R elements x[i + m h] (h = 2^st, i = its block's start + k,
  // k < h) and runs stages st .. st+R-1 in registers; stage st+q pairs
  // m and m + 2^q with W_{2^(st+q+1)}^(k + (m mod 2^q) h).
  if constexpr (LC != 0) {
    static_for<0, LC / 4>([&](auto pc) {
      fft_pass_ct<4, LC, 4 * decltype(pc)::value, MODE == kIfft>(sx, tw);
      __syncthreads();
    });
    if constexpr (LC % 4 != 0) {
      fft_pass_ct<LC % 4, LC, 4 * (LC / 4), MODE == kIfft>(sx, tw);
      __syncthreads();
    }
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
  const uint32_t shift = a.start + H;  // bins (i + shift) mod N, N a power of two
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
  a.qscale = a.qbits == 16 ? 32768.0f : 128.0f;
  if (mode == kIfft) return launch_q<kIfft>(a, st);
  return mode == kPilot ? launch_q<kPilot>(a, st) : launch_q<kUplink>(a, st);
}

}  // namespace detail
}  // namespace mega
