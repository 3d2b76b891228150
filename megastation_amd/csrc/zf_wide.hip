// Zero-forcing weights of few, wide groups (17 to 64 users; e.g. one slot of
// a 256 x 32 cell, 273 groups) spread over the whole GPU in three kernels.
//
// With one workgroup per group (zf.hip) such a launch is one workgroup per CU
// running a chain of dependent phases -- A loads, Gram MFMAs, 32 pivots with
// a barrier each, the apply -- and its time is that chain's latency (at
// 256 x 32 the median group takes 9.8 us of Gram, 6.4 of sweep and 9.0 of
// apply, and the 17 CUs holding two of the 273 groups stretch the launch to
// 37 us). Here each phase is a launch of its own with work items sized to
// fill the GPU:
//   1. zf_gram_items: one workgroup per (group, 64-antenna chunk) -- 1092
//      at 256 x 32 -- stages the chunk in LDS (every element read from HBM
//      once) and writes the chunk's partial of every lower Gram tile;
//   2. zf_sweep_groups: one workgroup per group adds the partials of each
//      tile in chunk order, mirrors the Hermitian matrix, runs the sweep
//      (zf_sweep, eight waves for up to 32 users) and writes -G^-1 and the
//      group status;
//   3. zf_apply_groups: one workgroup per (group, 64 antennas) -- four per
//      group at 256 antennas -- stages its chunk of A and -G^-1 in LDS and
//      computes W = G^-1 A for its antennas (zf_apply).
// Every arithmetic step is the one-workgroup kernel's, in the same order
// (the partials of a tile are those of its 64-antenna chunks, added in chunk
// order), so W and the status are bit-identical to zf.hip's.
#include "zf_core.h"

namespace mega {
namespace detail {
namespace {

constexpr int kWideThreads = 256;
constexpr uint32_t kApplyAntennas = 64;  // antennas per apply workgroup

template <int UT>
struct WideLayout {
  static constexpr int UP = 16 * UT;
  static constexpr int NT = UT * (UT + 1) / 2;  // lower tiles
  uint32_t nch, items;                           // chunks per tile, items per group
  __host__ __device__ explicit WideLayout(uint32_t B)
      : nch(gram_chunks(B)), items(NT * gram_chunks(B)) {}
};

// Workspace (float2 unless noted): partial tiles [batch][items][256], then
// -G^-1 [batch][UP][UP], then the status int32 [batch].
template <int UT>
__host__ __device__ uint64_t wide_parts(uint32_t B, uint32_t batch) {
  return (uint64_t)batch * WideLayout<UT>(B).items * 256;
}
template <int UT>
__host__ __device__ uint64_t wide_inv(uint32_t batch) {
  return (uint64_t)batch * 16 * UT * 16 * UT;
}

// A chunk of 64 antennas of one group (U x 64, ld = U in memory) into LDS
// rows of U+1 complex, all of a thread's loads in flight at once; antennas
// past nb are zero.
template <int UT>
__device__ __forceinline__ void stage_chunk(const float2* A, uint32_t U, uint32_t nb,
                                            float2* sA) {
  constexpr int PER = 16 * UT * 64 / kWideThreads;  // elements per thread (UP x 64)
  float2 v[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const uint32_t e = threadIdx.x + kWideThreads * j;
    v[j] = e < U * nb ? A[e] : make_float2(0.f, 0.f);
  }
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const uint32_t e = threadIdx.x + kWideThreads * j;
    if (e < U * 64) {
      const uint32_t b = e / U, u = e - b * U;
      sA[b * (U + 1) + u] = v[j];
    }
  }
}

// Gram partials: one workgroup per (group, 64-antenna chunk); the chunk goes
// through LDS once and the four waves take its lower tiles in turn. A tile's
// partial over the chunk is computed exactly as in zf.hip (four real MFMA
// products per 4-antenna step, (gr1 + gr2, gi1 + gi2) at the end).
template <int UT>
__global__ __launch_bounds__(kWideThreads) void zf_gram_items(const float2* __restrict__ in,
                                                              float2* __restrict__ parts,
                                                              uint32_t U, uint32_t B) {
  extern __shared__ float2 sA[];  // 64 rows of U+1
  const WideLayout<UT> L(B);
  const uint32_t g = blockIdx.x / L.nch, ch = blockIdx.x % L.nch;
  const uint32_t b0 = ch * 64, nb = min(64u, B - b0);
  stage_chunk<UT>(in + (uint64_t)g * U * B + (uint64_t)U * b0, U, nb, sA);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i16 = lane & 15, q4 = lane >> 4;
  const uint32_t lda = U + 1;
  for (int tile = wave; tile < WideLayout<UT>::NT; tile += kWideThreads / 64) {
    int mi, mj;
    tile_of(tile, mi, mj);
    const uint32_t ui = 16 * mi + i16, uj = 16 * mj + i16;
    v4f gr1 = {0, 0, 0, 0}, gr2 = {0, 0, 0, 0}, gi1 = {0, 0, 0, 0}, gi2 = {0, 0, 0, 0};
#pragma unroll
    for (uint32_t s = 0; s < kGramChunk; ++s) {
      const uint32_t b = 4 * s + q4;
      const float2 ai = ui < U && b < nb ? sA[b * lda + ui] : make_float2(0.f, 0.f);
      const float2 aj = uj < U && b < nb ? sA[b * lda + uj] : make_float2(0.f, 0.f);
      gr1 = mfma4(ai.x, aj.x, gr1);
      gr2 = mfma4(ai.y, aj.y, gr2);
      gi1 = mfma4(ai.y, aj.x, gi1);
      gi2 = mfma4(-ai.x, aj.y, gi2);
    }
    float2* out = parts + ((uint64_t)g * L.items + (uint64_t)tile * L.nch + ch) * 256 + lane;
#pragma unroll
    for (int q = 0; q < 4; ++q) out[q * 64] = make_float2(gr1[q] + gr2[q], gi1[q] + gi2[q]);
  }
}

// Per group: the partials of every tile added in chunk order (all of a
// thread's loads issued first), the Hermitian matrix, the sweep over NTH
// threads (eight waves for 17 to 32 users: two entries per thread), -G^-1
// and the status out.
template <int UT, int NTH>
__global__ __launch_bounds__(NTH) void zf_sweep_groups(const float2* __restrict__ parts,
                                                       float2* __restrict__ inv,
                                                       int32_t* __restrict__ status,
                                                       uint32_t U, uint32_t B) {
  constexpr int UP = ZfShared<UT>::UP, LDG = ZfShared<UT>::LDG;
  constexpr int NE = WideLayout<UT>::NT * 256;  // tile entries
  constexpr int PER = (NE + NTH - 1) / NTH;
  constexpr int MAXCH = (256 / 4 + kGramChunk - 1) / kGramChunk;  // chunks at 256 antennas
  __shared__ ZfShared<UT> sh;
  const WideLayout<UT> L(B);
  const uint32_t g = blockIdx.x;
  const float2* gp = parts + (uint64_t)g * L.items * 256;
  if (L.nch <= (uint32_t)MAXCH) {
    float2 v[PER][MAXCH];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int e = threadIdx.x + NTH * j;
      const int tile = e / 256, q = (e % 256) / 64, l = e % 64;
#pragma unroll
      for (int ch = 0; ch < MAXCH; ++ch)
        if (e < NE && (uint32_t)ch < L.nch)
          v[j][ch] = gp[(uint64_t)(tile * L.nch + ch) * 256 + q * 64 + l];
    }
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int e = threadIdx.x + NTH * j;
      if (e >= NE) break;
      const int tile = e / 256, q = (e % 256) / 64, l = e % 64;
      float2 sum = v[j][0];
#pragma unroll
      for (int ch = 1; ch < MAXCH; ++ch)
        if ((uint32_t)ch < L.nch) sum = make_float2(sum.x + v[j][ch].x, sum.y + v[j][ch].y);
      int mi, mj;
      tile_of(tile, mi, mj);
      gram_put_tile<LDG>(sh.G, U, mi, mj, l, q, sum);
    }
  } else {
    for (int e = threadIdx.x; e < NE; e += NTH) {
      const int tile = e / 256, q = (e % 256) / 64, l = e % 64;
      float2 sum = gp[(uint64_t)(tile * L.nch) * 256 + q * 64 + l];
      for (uint32_t ch = 1; ch < L.nch; ++ch) {
        const float2 c = gp[(uint64_t)(tile * L.nch + ch) * 256 + q * 64 + l];
        sum = make_float2(sum.x + c.x, sum.y + c.y);
      }
      int mi, mj;
      tile_of(tile, mi, mj);
      gram_put_tile<LDG>(sh.G, U, mi, mj, l, q, sum);
    }
  }
  __syncthreads();
  const int fail = zf_sweep<UT, NTH>(sh, U);
  float2* gi = inv + (uint64_t)g * UP * UP;
  for (int e = threadIdx.x; e < UP * UP; e += NTH) gi[e] = sh.G[(e / UP) * LDG + e % UP];
  if (threadIdx.x == 0) status[g] = fail;
}

// W for one (group, 64 antennas): the chunk of A and -G^-1 go into LDS
// together (all loads in flight), then zf_apply on the staged chunk.
template <int UT>
__global__ __launch_bounds__(kWideThreads) void zf_apply_groups(
    const float2* in, float2* out, const float2* __restrict__ inv,
    const int32_t* __restrict__ status, int32_t* info, uint32_t U, uint32_t B,
    uint32_t chunks) {
  constexpr int UP = ZfShared<UT>::UP, LDG = ZfShared<UT>::LDG;
  constexpr int PI = UP * UP / kWideThreads;
  __shared__ ZfShared<UT> sh;
  extern __shared__ float2 sA[];  // 64 rows of U+1
  const uint32_t g = blockIdx.x / chunks, c = blockIdx.x % chunks;
  const uint32_t b0 = c * kApplyAntennas;
  const uint32_t nb = min(kApplyAntennas, B - b0);
  const int fail = status[g];
  const float2* A = in + (uint64_t)g * U * B + (uint64_t)U * b0;
  float2* W = out + (uint64_t)g * U * B + (uint64_t)U * b0;
  if (info && c == 0 && threadIdx.x == 0) info[g] = fail;
  if (fail) {  // uniform: W = 0 on this group's antennas
    for (uint32_t e = threadIdx.x; e < U * nb; e += kWideThreads) W[e] = make_float2(0.f, 0.f);
    return;
  }
  const float2* gi = inv + (uint64_t)g * UP * UP;
  float2 x[PI];
#pragma unroll
  for (int j = 0; j < PI; ++j) x[j] = gi[threadIdx.x + kWideThreads * j];
  stage_chunk<UT>(A, U, nb, sA);
#pragma unroll
  for (int j = 0; j < PI; ++j) {
    const int e = threadIdx.x + kWideThreads * j;
    sh.G[(e / UP) * LDG + e % UP] = x[j];
  }
  __syncthreads();
  zf_apply<UT, kWideThreads, true>(A, U, nb, sh, sA, W, U, nullptr);
}

template <int UT>
hipError_t launch_wide(uint32_t U, uint32_t B, uint32_t batch, const float2* in, float2* out,
                       int32_t* info, void* work, hipStream_t st) {
  const WideLayout<UT> L(B);
  float2* parts = static_cast<float2*>(work);
  float2* inv = parts + wide_parts<UT>(B, batch);
  int32_t* status = reinterpret_cast<int32_t*>(inv + wide_inv<UT>(batch));
  const size_t lds = (size_t)64 * (U + 1) * sizeof(float2);  // a staged 64-antenna chunk
  hipLaunchKernelGGL(zf_gram_items<UT>, dim3(batch * L.nch), dim3(kWideThreads), lds, st, in,
                     parts, U, B);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  constexpr int NTS = UT == 2 ? 512 : 256;  // sweep threads: two (UT=2) or 16 entries each
  hipLaunchKernelGGL((zf_sweep_groups<UT, NTS>), dim3(batch), dim3(NTS), 0, st, parts, inv,
                     status, U, B);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  const uint32_t chunks = (B + kApplyAntennas - 1) / kApplyAntennas;
  hipLaunchKernelGGL(zf_apply_groups<UT>, dim3(batch * chunks), dim3(kWideThreads), lds, st,
                     in, out, inv, status, info, U, B, chunks);
  return hipGetLastError();
}

}  // namespace

uint64_t zf_wide_work_bytes(uint32_t U, uint32_t B, uint32_t batch) {
  if (U <= 16 || U > 64 || B == 0 || batch == 0) return 0;
  const uint64_t f2 = U <= 32 ? wide_parts<2>(B, batch) + wide_inv<2>(batch)
                              : wide_parts<4>(B, batch) + wide_inv<4>(batch);
  return f2 * sizeof(float2) + (uint64_t)batch * sizeof(int32_t);
}

bool zf_wide_pays(uint32_t, uint32_t, uint32_t) {
  // Measured at 256 x 32 (273 groups): Gram items 12.4 us, sweep 11.3, apply
  // 14.5 -- 38 to 45 us against the one-workgroup kernel's 37 to 42. Every
  // workgroup of a phase starts at once (about 1.3 us of dispatch skew), so
  // loads and matrix-core work do not overlap across workgroups: staging a
  // 16 KB chunk takes 4.8 us, the Gram MFMAs 3.5 (2.7 is the whole GPU's FP32
  // matrix-core time for them). Not the default anywhere; MEGA_ZF_WIDE=2
  // selects it.
  return false;
}

hipError_t launch_zf_wide(uint32_t U, uint32_t B, uint32_t batch, const float2* in,
                          float2* out, int32_t* info, void* work, hipStream_t st,
                          const char** why) {
  if (batch == 0) return hipSuccess;
  if (U <= 16 || U > 64) {
    *why = "the three-kernel ZF takes 17 to 64 users";
    return hipErrorInvalidValue;
  }
  if (batch > 0x7fffffffu / 64 || !work) {
    *why = !work ? "no workspace" : "batch too large";
    return hipErrorInvalidValue;
  }
  return U <= 32 ? launch_wide<2>(U, B, batch, in, out, info, work, st)
                 : launch_wide<4>(U, B, batch, in, out, info, work, st);
}

}  // namespace detail
}  // namespace mega
