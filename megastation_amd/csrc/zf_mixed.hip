// Mixed-precision zero-forcing weights (the precision form of the detection
// path): per subcarrier group, one workgroup (one wave for up to 16 users,
// else four) computes
//   1. G = A A^H on the BF16 matrix cores with operands split in three
//      (prec_core.h; FP32-accurate products), FP32 accumulation, mirrored
//      to an exactly Hermitian matrix;
//   2. X0 = G^-1 by the FP32 sweep operator (zf_core.h), and the 1-norm
//      condition number kappa = |G|_1 |X0|_1 (exact: X0 is explicit);
//   3. one step of iterative refinement of the inverse,
//        R = I - G X0 (products and sums in FP64),  X = X0 + X0 R (FP32);
//   4. W = X A on the BF16 matrix cores, operands split in three.
// Both products carry FP32 accuracy because the condition number multiplies
// their errors into x_hat: with two-piece operands (about 16 bits) the Gram
// matrix the inverse is refined against, and W, would err by kappa * 2^-17,
// about 4e-7 * kappa_1 relative to max |x_hat| at 32 x 64; with three pieces
// the error is that of FP32 arithmetic.
// Groups with kappa above kCondMax, or whose FP32 sweep meets a pivot at or
// below the relative tolerance, take the FP64 route instead: Gram, sweep and
// W = G^-1 A all in FP64 (vector ALU), rounded to FP32 at the end; a pivot
// at or below U * 2^-52 * G_kk there marks the group singular (status k+1,
// W = 0).
//
// Outputs: W in the CSI layout (FP32, optional; may be the input buffer),
// W as split BF16 words for the mixed equalize kernel (optional; per group a
// plane of hi words then a plane of lo words, each U rows of split_row(B)
// packed (re, im) BF16 pairs, zero past B; prec_core.h), the status and the
// route (1 = FP64) per group.
#include "prec_core.h"
#include "zf_core.h"

namespace mega {
namespace detail {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr float kCondMax = 1e3f;  // kappa_1(G) above this takes the FP64 route

namespace {

// Threads per group: one wave for up to 16 users, four waves otherwise.
template <int UT>
__host__ __device__ constexpr int mix_threads() { return UT == 1 ? 64 : 256; }

template <int UT>
struct ZfMixShared {
  static constexpr int UP = 16 * UT;
  static constexpr int LDG = UP + 1;
  ZfShared<UT> sw;     // G, then -G^-1 (sweep); later X (refined inverse)
  float2 G0[UP * LDG];  // copy of G for the residual; later R; later split X
  float red[2];
  float part[mix_threads<UT>()];  // column-norm shares
};


// max over columns c < U of sum_i |M[i][c]| (|.| from the hardware square
// root; the norm only feeds the condition estimate), result in *out after the
// caller's next barrier. Every thread sums its share of a column into part,
// then wave 0 adds the shares and takes the maximum. All threads call it.
template <int UP, int LDG, int NTH>
__device__ __forceinline__ void col_norm1(const float2* M, uint32_t U, float* part, float* out) {
  constexpr int RS = NTH / UP, EPT = UP / RS;
  const int tid = threadIdx.x, c = tid % UP, r0 = tid / UP;
  float s = 0.f;
#pragma unroll
  for (int j = 0; j < EPT; ++j) {
    const int i = r0 + RS * j;
    const float2 v = M[i * LDG + c];
    s += i < (int)U ? __builtin_amdgcn_sqrtf(v.x * v.x + v.y * v.y) : 0.f;
  }
  part[tid] = s;
  __syncthreads();
  if (tid < 64) {
    const int cc = tid < UP ? tid : UP - 1;
    float t = 0.f;
#pragma unroll
    for (int r = 0; r < RS; ++r) t += part[r * UP + cc];
    if (tid >= (int)U) t = 0.f;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) t = fmaxf(t, __shfl_xor(t, off, 64));
    if (tid == 0) *out = t;
  }
}

// FP64 sweep on M (UP x LDG, Hermitian, padded with identity) -> -M^-1, one
// barrier per pivot; entries in registers, pivot row/column through prow/pcol
// (2 x UP each). Returns 0 or k+1. Ends with a barrier.
template <int UP, int LDG, int NTH>
__device__ int sweep_f64(double2* M, double2* prow, double2* pcol, double* diag,
                         uint32_t U) {
  constexpr int EPT = UP * UP / NTH;
  constexpr int RS = NTH / UP;
  static_assert((UP * UP) % NTH == 0 && NTH % UP == 0, "sweep layout");
  const int tid = threadIdx.x;
  const int c = tid % UP, r0 = tid / UP;
  for (int i = tid; i < UP; i += NTH) {
    if (i >= (int)U) M[i * LDG + i] = make_double2(1.0, 0.0);
    diag[i] = M[i * LDG + i].x;
  }
  __syncthreads();
  double2 a[EPT];
#pragma unroll
  for (int j = 0; j < EPT; ++j) a[j] = M[(r0 + RS * j) * LDG + c];
  auto publish = [&](int k, int buf) {
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
      const int i = r0 + RS * j;
      if (i == k) prow[buf * UP + c] = a[j];
      if (c == k) pcol[buf * UP + i] = a[j];
    }
  };
  publish(0, 0);
  __syncthreads();
  const double tol = 0x1p-52 * (double)U;
  int fail = 0;
  for (uint32_t k = 0; k < U; ++k) {
    const int buf = k & 1;
    const double piv = prow[buf * UP + k].x;
    if (!(piv > tol * diag[k]) || !isfinite(piv)) {
      fail = (int)k + 1;
      break;
    }
    const double ip = 1.0 / piv;
    const double2 akc = prow[buf * UP + c];
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
      const int i = r0 + RS * j;
      const double2 aik = pcol[buf * UP + i];
      const double2 base = (i == (int)k || c == (int)k) ? make_double2(0.0, 0.0) : a[j];
      const double2 x = (i == (int)k) ? make_double2(-1.0, 0.0) : aik;
      const double2 y = (c == (int)k) ? make_double2(-1.0, 0.0) : akc;
      const double pr = x.x * y.x - x.y * y.y, pi = x.x * y.y + x.y * y.x;
      a[j] = make_double2(base.x - pr * ip, base.y - pi * ip);
    }
    if (k + 1 < U) publish((int)k + 1, buf ^ 1);
    __syncthreads();
  }
  if (!fail) {
#pragma unroll
    for (int j = 0; j < EPT; ++j) M[(r0 + RS * j) * LDG + c] = a[j];
  }
  __syncthreads();
  return fail;
}

// W[u][b] to the FP32 CSI layout (w, optional) and to the split planes
// (ws, optional; U rows of Bp words per plane).
__device__ __forceinline__ void store_w(float2* w, uint32_t* ws, uint32_t U, uint32_t Bp,
                                        uint32_t u, uint32_t b, float2 v) {
  if (w) w[u + (uint64_t)U * b] = v;
  if (ws) {
    const SplitC s = split_c(v);
    ws[(uint64_t)u * Bp + b] = s.hi;
    ws[(uint64_t)U * Bp + (uint64_t)u * Bp + b] = s.lo;
  }
}

__device__ __forceinline__ void store_pad(uint32_t* ws, uint32_t U, uint32_t Bp, uint32_t u,
                                          uint32_t b) {
  if (ws) {
    ws[(uint64_t)u * Bp + b] = 0u;
    ws[(uint64_t)U * Bp + (uint64_t)u * Bp + b] = 0u;
  }
}

// ---- Gram on split BF16 products ------------------------------------------
// Antenna chunks (stage_kc) go through LDS once per group as split words:
// every thread loads its share of kStageBatch chunks at once (coalesced, the
// whole batch in flight), then per chunk splits it once in three and writes
// the hi, lo and rest planes [user][antenna] (rows of 4 more words, 16-byte
// aligned),
// double-buffered with one barrier per chunk. Output tiles are 8 users (rows, re/im
// interleaved) x 16 users (columns); only tiles reaching the lower triangle
// are computed (the caller mirrors), spread over the four waves.
// Chunk of antennas staged at a time: 64 for 17 to 32 users, else 32 (a
// small footprint for up to 16 users, many groups per CU; LDS room for 64
// users); rows padded by 4 words.
template <int UT>
__host__ __device__ constexpr int stage_kc() { return UT == 2 ? 64 : 32; }
constexpr int kPieces = 3;  // hi, lo, rest
constexpr int kStageBatch = 4;

// Staging buffers: one for one-wave groups (the smallest LDS footprint), two
// (double-buffered) otherwise.
template <int UT>
__host__ __device__ constexpr int stage_bufs() { return UT == 1 ? 1 : 2; }

template <int UT>
__host__ __device__ constexpr size_t gram_stage_bytes() {
  return (size_t)stage_bufs<UT>() * kPieces * 16 * UT * (stage_kc<UT>() + 4) * sizeof(uint32_t);
}

template <int UT>
struct LowerTiles {
  static constexpr int RT = 2 * UT, CT = UT;
  static constexpr bool lower(int rt, int ct) { return 8 * rt + 7 >= 16 * ct; }
  static constexpr int count() {
    int n = 0;
    for (int rt = 0; rt < RT; ++rt)
      for (int ct = 0; ct < CT; ++ct) n += lower(rt, ct) ? 1 : 0;
    return n;
  }
  static constexpr int N = count();
  // (rt, ct) of lower tile t, packed as rt * CT + ct
  static constexpr int at(int t) {
    for (int rt = 0; rt < RT; ++rt)
      for (int ct = 0; ct < CT; ++ct)
        if (lower(rt, ct) && t-- == 0) return rt * CT + ct;
    return 0;
  }
};

template <int UT>
__device__ __forceinline__ void gram_split(const float2* __restrict__ A, uint32_t U, uint32_t B,
                                           uint32_t* stage, float2* sG) {
  constexpr int UP = 16 * UT, LDG = UP + 1, NTH = mix_threads<UT>(), NW = NTH / 64;
  constexpr int kStageKC = stage_kc<UT>(), kStageKCP = kStageKC + 4;
  constexpr int PER = UP * kStageKC / NTH;  // staged elements per thread
  using T = LowerTiles<UT>;
  constexpr int TW = (T::N + NW - 1) / NW;  // tiles per wave (at most)
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int i16 = lane & 15, q4 = lane >> 4;
  const ARow h(i16 & 1);
  const uint32_t nch = (B + kStageKC - 1) / kStageKC;
  auto plane = [&](int buf, int piece) { return stage + (kPieces * buf + piece) * UP * kStageKCP; };

  // all chunks of a batch of up to kStageBatch are loaded at once, so the
  // group's A is in flight together (one HBM round trip per batch)
  constexpr int NB = UT == 2 ? kStageBatch : kStageBatch / 2;
  float2 pf[NB][PER];
  auto fetch = [&](uint32_t ch, float2 (&d)[PER]) {
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const uint32_t e = tid + NTH * j;
      const uint32_t u = e % UP, b = ch * kStageKC + e / UP;
      d[j] = A[min(u, U - 1) + (uint64_t)U * min(b, B - 1)];
    }
  };
  auto put = [&](uint32_t ch, const float2 (&d)[PER], int buf) {
    uint32_t* ph = plane(buf, 0);
    uint32_t* pl = plane(buf, 1);
    uint32_t* pr = plane(buf, 2);
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const uint32_t e = tid + NTH * j;
      const uint32_t u = e % UP, bl = e / UP;
      const bool ok = u < U && ch * kStageKC + bl < B;
      const SplitC3 sp = split_c3(ok ? d[j] : make_float2(0.f, 0.f));
      ph[u * kStageKCP + bl] = sp.hi;
      pl[u * kStageKCP + bl] = sp.lo;
      pr[u * kStageKCP + bl] = sp.rest;
    }
  };

  v4f acc[TW], acc2[TW], acc3[TW];
  int trow[TW], tcol[TW];  // first row / column of the wave's tiles (-1: none)
  const int wv = __builtin_amdgcn_readfirstlane(wave);
#pragma unroll
  for (int t = 0; t < TW; ++t) {
    acc[t] = acc2[t] = acc3[t] = v4f{0.f, 0.f, 0.f, 0.f};
    const int idx = wv + NW * t;
    const int rc = idx < T::N ? T::at(idx) : -1;
    trow[t] = rc < 0 ? -1 : 8 * (rc / T::CT);
    tcol[t] = rc < 0 ? -1 : 16 * (rc % T::CT);
  }
  // Per 16-antenna step, the operands of all the wave's tiles are read
  // first, then the six products of every tile go to three accumulators by
  // size (hi*hi; lo*hi + hi*lo; lo*lo + rest*hi + hi*rest), summed at the
  // end, smallest first.
  auto products = [&](int buf) {
    const uint32_t* ph = plane(buf, 0);
    const uint32_t* pl = plane(buf, 1);
    const uint32_t* pr = plane(buf, 2);
#pragma unroll
    for (int st = 0; st < kStageKC / 16; ++st) {
      const int k0 = 16 * st + 4 * q4;
      v4u ah[TW], al[TW], ar[TW], bh[TW], bl[TW], br[TW];
#pragma unroll
      for (int t = 0; t < TW; ++t) {
        const uint32_t u = max(trow[t], 0) + (i16 >> 1), v = max(tcol[t], 0) + i16;
        const v4u uh = *reinterpret_cast<const v4u*>(ph + u * kStageKCP + k0);
        const v4u ul = *reinterpret_cast<const v4u*>(pl + u * kStageKCP + k0);
        const v4u ur = *reinterpret_cast<const v4u*>(pr + u * kStageKCP + k0);
        const v4u vh = *reinterpret_cast<const v4u*>(ph + v * kStageKCP + k0);
        const v4u vl = *reinterpret_cast<const v4u*>(pl + v * kStageKCP + k0);
        const v4u vr = *reinterpret_cast<const v4u*>(pr + v * kStageKCP + k0);
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          ah[t][a] = a_word(uh[a], h);
          al[t][a] = a_word(ul[a], h);
          ar[t][a] = a_word(ur[a], h);
          bh[t][a] = vh[a] ^ 0x80000000u;  // conj(A[v][b])
          bl[t][a] = vl[a] ^ 0x80000000u;
          br[t][a] = vr[a] ^ 0x80000000u;
        }
      }
#pragma unroll
      for (int t = 0; t < TW; ++t) {
        if (trow[t] < 0) continue;  // uniform per wave
        acc[t] = mfma_bf16(ah[t], bh[t], acc[t]);
        acc2[t] = mfma_bf16(al[t], bh[t], acc2[t]);
        acc3[t] = mfma_bf16(al[t], bl[t], acc3[t]);
        acc2[t] = mfma_bf16(ah[t], bl[t], acc2[t]);
        acc3[t] = mfma_bf16(ar[t], bh[t], acc3[t]);
        acc3[t] = mfma_bf16(ah[t], br[t], acc3[t]);
      }
    }
  };
  // chunks ch0 .. ch0+NB-1 per batch; chunk c of the batch uses buffer
  // c % stage_bufs (with one buffer, a barrier keeps each chunk's products
  // ahead of the next chunk's writes)
  constexpr int NBUF = stage_bufs<UT>();
  for (uint32_t ch0 = 0; ch0 < nch; ch0 += NB) {
#pragma unroll
    for (int c = 0; c < NB; ++c) fetch(min(ch0 + c, nch - 1), pf[c]);
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      if (ch0 + c >= nch) break;  // uniform
      if (NBUF == 1 && c > 0) __syncthreads();
      put(ch0 + c, pf[c], c % NBUF);
      __syncthreads();
      products(c % NBUF);
    }
    __syncthreads();  // the last buffers are read before the next batch's writes
  }
  static_for<0, TW>([&](auto tc) {
    constexpr int t = decltype(tc)::value;
    if (trow[t] < 0) return;
    acc[t] += acc2[t] + acc3[t];  // hi*hi + ((lo*hi + hi*lo) + small terms)
    const uint32_t col = tcol[t] + i16;
#pragma unroll
    for (int r2 = 0; r2 < 2; ++r2) {
      const uint32_t row = trow[t] + 2 * q4 + r2;
      const bool ok = row < U && col < U;
      sG[row * LDG + col] =
          ok ? make_float2(acc[t][2 * r2], acc[t][2 * r2 + 1]) : make_float2(0.f, 0.f);
    }
  });
}

template <int UT>
__global__ __launch_bounds__(mix_threads<UT>(), UT == 1 ? 16 : UT == 2 ? 2 : 1) void zf_mixed(const float2* __restrict__ in,
                                                        float2* w_out, uint32_t* ws_out,
                                                        int32_t* info, int32_t* route,
                                                        uint32_t U, uint32_t B) {
  constexpr int UP = 16 * UT;
  constexpr int LDG = UP + 1;
  constexpr int NTH = mix_threads<UT>();
  constexpr int NW = NTH / 64;
  __shared__ ZfMixShared<UT> sh;
  // Gram staging (gram_split), later the FP64 route: UP x LDG, prow/pcol, diag
  extern __shared__ double2 sM64[];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int i16 = lane & 15, q4 = lane >> 4;
  const ARow h(i16 & 1);
  const uint32_t g = blockIdx.x;
  const float2* A = in + (uint64_t)g * U * B;
  float2* w = w_out ? w_out + (uint64_t)g * U * B : nullptr;
  const uint32_t Bp = split_row(B);
  uint32_t* ws = ws_out ? ws_out + (uint64_t)g * 2 * U * Bp : nullptr;
  float2* sG = sh.sw.G;

  // ---- 1. Gram on split BF16 products --------------------------------------
  gram_split<UT>(A, U, B, reinterpret_cast<uint32_t*>(sM64), sG);
  __syncthreads();
  // exact Hermitian symmetry (lower triangle mirrored) and the copy for R
  for (int e = tid; e < UP * UP; e += NTH) {
    const int i = e / UP, c = e % UP;
    float2 val = sG[i * LDG + c];
    if (i < c) {
      const float2 t = sG[c * LDG + i];
      val = make_float2(t.x, -t.y);
    }
    if (i == c) val.y = 0.f;
    sG[i * LDG + c] = val;
    sh.G0[i * LDG + c] = val;
  }
  __syncthreads();
  col_norm1<UP, LDG, NTH>(sh.G0, U, sh.part, &sh.red[0]);

  // ---- 2. FP32 sweep and condition number ------------------------------------
  int fail = zf_sweep<UT, NTH>(sh.sw, U);  // starts with the barrier it needs
  bool fp64 = fail != 0;
  if (!fp64) {
    col_norm1<UP, LDG, NTH>(sG, U, sh.part, &sh.red[1]);
    __syncthreads();
    const float kappa = sh.red[0] * sh.red[1];
    fp64 = !(kappa <= kCondMax);
  }

  if (!fp64) {
    // ---- 3. refinement: R = I - G X0 (FP64), X = X0 + X0 R -----------------
    // Both products on the matrix cores, one 16 x 16 output tile per wave and
    // step (v_mfma_f64_16x16x4_f64 for R: FP32 operands are exact in FP64,
    // products and sums in FP64; v_mfma_f32_16x16x4_f32 for X). Operand lane
    // l: A[row l%16][k l/16], B[k l/16][col l%16]; result lane l, register r:
    // column l%16 of row 4*(l/16)+r (FP32) or row l/16+4*r (FP64, its result
    // rows are interleaved). sG holds -X0; entries past U are zero in
    // G0 and off the diagonal in sG, so k runs over UP.
    {
      constexpr int TT = UT * UT;  // output tiles
      constexpr int KS = UP / 4;   // k steps
      // R = I + G0 (-X0), rounded to FP32, masked to the U x U block
      float2 rr[(TT + NW - 1) / NW][4];
#pragma unroll
      for (int q = 0; q < (TT + NW - 1) / NW; ++q) {
        const int tile = wave + NW * q;
        if (tile >= TT) break;  // uniform per wave
        const int ti = tile / UT, tj = tile % UT;
        v4d cre = {0, 0, 0, 0}, cim = {0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          const float2 a = sh.G0[(16 * ti + i16) * LDG + 4 * ks + q4];
          const float2 b = sG[(4 * ks + q4) * LDG + 16 * tj + i16];
          const double ar = a.x, ai = a.y, br = b.x, bi = b.y;
          cre = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, cre, 0, 0, 0);
          cre = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai, bi, cre, 0, 0, 0);
          cim = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bi, cim, 0, 0, 0);
          cim = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, br, cim, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = 16 * ti + q4 + 4 * r, c = 16 * tj + i16;  // FP64 result rows
          const bool in = i < (int)U && c < (int)U;
          rr[q][r] = in ? make_float2((float)((i == c ? 1.0 : 0.0) + cre[r]), (float)cim[r])
                        : make_float2(0.f, 0.f);
        }
      }
      __syncthreads();  // every read of G0 is done
#pragma unroll
      for (int q = 0; q < (TT + NW - 1) / NW; ++q) {
        const int tile = wave + NW * q;
        if (tile >= TT) break;
        const int ti = tile / UT, tj = tile % UT;
#pragma unroll
        for (int r = 0; r < 4; ++r)
          sh.G0[(16 * ti + q4 + 4 * r) * LDG + 16 * tj + i16] = rr[q][r];
      }
      __syncthreads();
      // X = X0 + X0 R = -((-X0) + (-X0) R), split into BF16 words for the apply
      float2 xn[(TT + NW - 1) / NW][4];
#pragma unroll
      for (int q = 0; q < (TT + NW - 1) / NW; ++q) {
        const int tile = wave + NW * q;
        if (tile >= TT) break;
        const int ti = tile / UT, tj = tile % UT;
        v4f cre = {0.f, 0.f, 0.f, 0.f}, cim = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          const float2 a = sG[(16 * ti + i16) * LDG + 4 * ks + q4];
          const float2 b = sh.G0[(4 * ks + q4) * LDG + 16 * tj + i16];
          cre = mfma4(a.x, b.x, cre);
          cre = mfma4(-a.y, b.y, cre);
          cim = mfma4(a.x, b.y, cim);
          cim = mfma4(a.y, b.x, cim);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float2 x0 = sG[(16 * ti + 4 * q4 + r) * LDG + 16 * tj + i16];
          xn[q][r] = make_float2(-(x0.x + cre[r]), -(x0.y + cim[r]));
        }
      }
      __syncthreads();  // every read of G0 (R) is done
      // split X in three for the apply: (hi, lo) words in G0 (reused as uint2
      // storage), the rest words in the staging area (free since the Gram)
      uint2* sXw = reinterpret_cast<uint2*>(sh.G0);
      uint32_t* sXr = reinterpret_cast<uint32_t*>(sM64);
#pragma unroll
      for (int q = 0; q < (TT + NW - 1) / NW; ++q) {
        const int tile = wave + NW * q;
        if (tile >= TT) break;
        const int ti = tile / UT, tj = tile % UT;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const SplitC3 sp = split_c3(xn[q][r]);
          sXw[(16 * ti + 4 * q4 + r) * LDG + 16 * tj + i16] = make_uint2(sp.hi, sp.lo);
          sXr[(16 * ti + 4 * q4 + r) * LDG + 16 * tj + i16] = sp.rest;
        }
      }
      __syncthreads();
    }
    const uint2* sX = reinterpret_cast<const uint2*>(sh.G0);
    const uint32_t* sXr = reinterpret_cast<const uint32_t*>(sM64);

    // ---- 4. W = X A on products of three-piece splits -------------------------
    // A tiles of 16 antennas x all users; a wave's tiles are loaded in
    // batches of NAB, all in flight together, then multiplied and stored.
    const uint32_t ntiles = (B + 15) / 16;
    constexpr int NAB = 1;
    const bool vec = U == UP;  // whole rows of four users: 16-byte loads
    float2 zb[NAB][UT][4];
    auto load_a = [&](uint32_t nt, float2 (&d)[UT][4]) {
      const uint32_t bc = min(16 * min(nt, ntiles - 1) + i16, B - 1);
#pragma unroll
      for (int ks = 0; ks < UT; ++ks) {
        if (vec) {
          const float4* p4 = reinterpret_cast<const float4*>(A + 16u * ks + 4 * q4 + (uint64_t)U * bc);
          const float4 x0 = p4[0], x1 = p4[1];
          d[ks][0] = make_float2(x0.x, x0.y);
          d[ks][1] = make_float2(x0.z, x0.w);
          d[ks][2] = make_float2(x1.x, x1.y);
          d[ks][3] = make_float2(x1.z, x1.w);
        } else {
#pragma unroll
          for (int a = 0; a < 4; ++a)
            d[ks][a] = A[min(16u * ks + 4 * q4 + a, U - 1) + (uint64_t)U * bc];
        }
      }
    };
    for (uint32_t n0 = wave; n0 < ntiles; n0 += NAB * NW) {
#pragma unroll
      for (int r = 0; r < NAB; ++r) load_a(n0 + r * NW, zb[r]);
#pragma unroll
      for (int r = 0; r < NAB; ++r) {
        const uint32_t nt = n0 + r * NW;
        if (nt >= ntiles) break;  // uniform per wave
        const uint32_t b = 16 * nt + i16;
        v4u bh[UT], bl[UT], br[UT];
#pragma unroll
        for (int ks = 0; ks < UT; ++ks)
#pragma unroll
          for (int a = 0; a < 4; ++a) {
            const bool ok = 16u * ks + 4 * q4 + a < U;
            const SplitC3 sp = split_c3(ok ? zb[r][ks][a] : make_float2(0.f, 0.f));
            bh[ks][a] = sp.hi;
            bl[ks][a] = sp.lo;
            br[ks][a] = sp.rest;
          }
#pragma unroll
        for (int mt = 0; mt < 2 * UT; ++mt) {
          const uint32_t u = 8 * mt + (i16 >> 1);
          v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int ks = 0; ks < UT; ++ks) {
            v4u ah, al, ar;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
              const uint2 xw = sX[u * LDG + 16 * ks + 4 * q4 + a];
              ah[a] = a_word(xw.x, h);
              al[a] = a_word(xw.y, h);
              ar[a] = a_word(sXr[u * LDG + 16 * ks + 4 * q4 + a], h);
            }
            acc = mfma_x6(ah, al, ar, bh[ks], bl[ks], br[ks], acc);
          }
          const uint32_t u0 = 8 * mt + 2 * q4;
          if (b < B) {
            if (u0 < U) store_w(w, ws, U, Bp, u0, b, make_float2(acc[0], acc[1]));
            if (u0 + 1 < U) store_w(w, ws, U, Bp, u0 + 1, b, make_float2(acc[2], acc[3]));
          } else if (b < Bp) {
            if (u0 < U) store_pad(ws, U, Bp, u0, b);
            if (u0 + 1 < U) store_pad(ws, U, Bp, u0 + 1, b);
          }
        }
      }
    }
  } else {
    // ---- FP64 route ------------------------------------------------------------
    double2* M = sM64;
    double2* prow = sM64 + UP * LDG;
    double2* pcol = prow + 2 * UP;
    double* diag = reinterpret_cast<double*>(pcol + 2 * UP);
    for (int e = tid; e < UP * UP; e += NTH) {
      const uint32_t i = e / UP, c = e % UP;
      double sr = 0.0, si = 0.0;
      if (i < U && c < U && i >= c) {
        for (uint32_t b = 0; b < B; ++b) {
          const float2 x = A[i + (uint64_t)U * b], y = A[c + (uint64_t)U * b];
          sr += (double)x.x * y.x + (double)x.y * y.y;
          si += (double)x.y * y.x - (double)x.x * y.y;
        }
      }
      M[i * LDG + c] = make_double2(sr, si);
    }
    __syncthreads();
    for (int e = tid; e < UP * UP; e += NTH) {
      const int i = e / UP, c = e % UP;
      if (i < c) {
        const double2 t = M[c * LDG + i];
        M[i * LDG + c] = make_double2(t.x, -t.y);
      }
    }
    __syncthreads();
    fail = sweep_f64<UP, LDG, NTH>(M, prow, pcol, diag, U);
    // W = G^-1 A in rounds of NTH weights, each read before any is written
    // (w may alias A).
    for (uint32_t base = 0; base < U * B; base += NTH) {
      const uint32_t e = base + tid;
      double sr = 0.0, si = 0.0;
      if (!fail && e < U * B) {
        const uint32_t u = e % U, b = e / U;
        for (uint32_t v = 0; v < U; ++v) {
          const double2 x = M[u * LDG + v];  // -G^-1
          const float2 z = A[v + (uint64_t)U * b];
          sr -= x.x * z.x - x.y * z.y;
          si -= x.x * z.y + x.y * z.x;
        }
      }
      __syncthreads();
      if (e < U * B) store_w(w, ws, U, Bp, e % U, e / U, make_float2((float)sr, (float)si));
      __syncthreads();
    }
    for (uint32_t e = tid; e < U * (Bp - B); e += NTH) store_pad(ws, U, Bp, e % U, B + e / U);
  }
  if (tid == 0) {
    if (info) info[g] = fail;
    if (route) route[g] = fp64 ? 1 : 0;
  }
}

template <int UT>
hipError_t launch_ut(uint32_t U, uint32_t B, uint32_t batch, const float2* in,
                     float2* w_out, uint32_t* wp, int32_t* info, int32_t* route,
                     hipStream_t st) {
  constexpr int UP = 16 * UT;
  const size_t f64 = (size_t)UP * (UP + 1) * sizeof(double2) + 4 * UP * sizeof(double2) +
                     UP * sizeof(double);
  const size_t lds = f64 > gram_stage_bytes<UT>() ? f64 : gram_stage_bytes<UT>();
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&zf_mixed<UT>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((zf_mixed<UT>), dim3(batch), dim3(mix_threads<UT>()), lds, st, in, w_out, wp,
                     info, route, U, B);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_zf_mixed(uint32_t U, uint32_t B, uint32_t batch, const float2* in,
                           float2* w_out, void* wp, int32_t* info, int32_t* route,
                           hipStream_t st, const char** why) {
  if (batch == 0) return hipSuccess;
  if (batch > 0x7fffffffu) {
    *why = "batch too large";
    return hipErrorInvalidValue;
  }
  uint32_t* w2 = static_cast<uint32_t*>(wp);
  if (U <= 16) return launch_ut<1>(U, B, batch, in, w_out, w2, info, route, st);
  if (U <= 32) return launch_ut<2>(U, B, batch, in, w_out, w2, info, route, st);
  if (U <= 64) return launch_ut<4>(U, B, batch, in, w_out, w2, info, route, st);
  *why = "more than 64 users";
  return hipErrorInvalidValue;
}

}  // namespace detail
}  // namespace mega
