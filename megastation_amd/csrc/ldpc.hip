// LDPC consumer of the detection path's LLRs: layered normalized min-sum
// decoding of 5G NR code blocks (base graphs of 3GPP TS 38.212, ldpc_tables.h)
// behind the reference's LDPCDecoder contract (src/operators/ldpc/decoder.h:
// 73-86): per code block (Kb + mb) * Zc LLRs, binary16 or float, the first
// 2 * Zc (the punctured systematic bits) zero, LLR > 0 meaning bit 0; out the
// hard decisions of the Kb * Zc information bits, 32 per word, LSB first.
//
// One workgroup per code block, one thread per check node z < Zc of the
// current block row:
//   * the a-posteriori LLRs L of all (Kb + mb) * Zc variables live in LDS as
//     FP32; the circulant shift of every block, reduced modulo Zc, is put in
//     LDS once;
//   * the block rows are unrolled at compile time from the base graph's
//     structure (which columns a row touches), so the compressed check
//     messages of every row -- alpha*min1, alpha*min2, the index of the
//     minimum and the sign bits -- stay in registers across iterations;
//   * rows are processed in order (layered schedule) with one barrier
//     between rows: within a row the Zc checks touch disjoint variables.
// Per check and row, exactly the operations of the CPU oracle in the same
// order (round-to-nearest FP32 subtract, multiply and add, no contraction),
// so the decoded bits and the final LLRs equal the oracle's.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "eq_core.h"
#include "internal.h"
#include "ldpc_tables.h"

namespace mega {
namespace detail {
namespace {

template <int BG>
struct Bg;
template <>
struct Bg<1> {
  static constexpr int rows = ldpc::kBg1Rows, cols = ldpc::kBg1Cols, edges = ldpc::kBg1Edges;
  static constexpr int kb = 22;
  static constexpr int start(int r) { return ldpc::kBg1RowStart[r]; }
  static constexpr int col(int e) { return ldpc::kBg1Col[e]; }
  static constexpr const int16_t* shift(int s) { return ldpc::kBg1Shift[s]; }
};
template <>
struct Bg<2> {
  static constexpr int rows = ldpc::kBg2Rows, cols = ldpc::kBg2Cols, edges = ldpc::kBg2Edges;
  static constexpr int kb = 10;
  static constexpr int start(int r) { return ldpc::kBg2RowStart[r]; }
  static constexpr int col(int e) { return ldpc::kBg2Col[e]; }
  static constexpr const int16_t* shift(int s) { return ldpc::kBg2Shift[s]; }
};

constexpr int kMaxEdges = 320;  // >= the edge count of either base graph

struct LdpcParams {
  const void* llr;
  uint64_t llr_stride;  // elements between code blocks
  uint32_t* out;
  uint64_t out_stride;  // words between code blocks
  uint32_t zc, mb, iters, half;
  float alpha;
  int16_t shift[kMaxEdges];  // the lifting-size set's shift of every block
};
static_assert(ldpc::kBg1Edges <= kMaxEdges && ldpc::kBg2Edges <= kMaxEdges, "edge table");

template <int BG>
__global__ __launch_bounds__(384) void ldpc_decode(const LdpcParams p) {
  using G = Bg<BG>;
  extern __shared__ float L[];  // (kb + mb) * zc LLRs, then edges shifts, then bit words
  const uint32_t zc = p.zc, mb = p.mb;
  const uint32_t nv = (G::kb + mb) * zc;
  int16_t* sh = reinterpret_cast<int16_t*>(L + nv);
  uint32_t* bits = reinterpret_cast<uint32_t*>(L + nv + (G::edges + 1) / 2);
  const uint32_t nbw = (G::kb * zc + 31) / 32;
  const uint32_t cw = blockIdx.x;
  const uint32_t z = threadIdx.x;
  const bool act = z < zc;

  for (uint32_t v = threadIdx.x; v < nv; v += blockDim.x) {
    const uint64_t i = cw * p.llr_stride + v;
    L[v] = p.half ? (float)static_cast<const _Float16*>(p.llr)[i]
                  : static_cast<const float*>(p.llr)[i];
  }
  for (uint32_t e = threadIdx.x; e < (uint32_t)G::edges; e += blockDim.x)
    sh[e] = (int16_t)(p.shift[e] % (int)zc);
  for (uint32_t w = threadIdx.x; w < nbw; w += blockDim.x) bits[w] = 0u;
  __syncthreads();

  // compressed messages of this thread's check in every row
  float m1[G::rows], m2[G::rows];
  uint32_t pk[G::rows];  // sign bits 0..23, index of the minimum in bits 24..31
#pragma unroll
  for (int r = 0; r < G::rows; ++r) {
    m1[r] = 0.f;
    m2[r] = 0.f;
    pk[r] = 0u;
  }

  for (uint32_t it = 0; it < p.iters; ++it) {
    static_for<0, G::rows>([&](auto rc) {
      constexpr int r = decltype(rc)::value;
      constexpr int e0 = G::start(r), deg = G::start(r + 1) - G::start(r);
      if ((uint32_t)r >= mb) return;  // uniform
      if (act) {
        float q[deg];
        uint32_t vi[deg];
        float min1 = __builtin_inff(), min2 = __builtin_inff();
        uint32_t im = 0, sb = 0u, sp = 0u;
        const uint32_t amin = pk[r] >> 24;
#pragma unroll
        for (int e = 0; e < deg; ++e) {
          uint32_t idx = z + (uint32_t)sh[e0 + e];
          idx = idx >= zc ? idx - zc : idx;
          const uint32_t v = (uint32_t)G::col(e0 + e) * zc + idx;
          const float mag = (uint32_t)e == amin ? m2[r] : m1[r];
          const float rold = (pk[r] >> e) & 1u ? -mag : mag;
          const float qe = __fsub_rn(L[v], rold);
          const float a = fabsf(qe);
          if (a < min1) {
            min2 = min1;
            min1 = a;
            im = (uint32_t)e;
          } else if (a < min2) {
            min2 = a;
          }
          const uint32_t s = qe < 0.f ? 1u : 0u;
          sb |= s << e;
          sp ^= s;
          q[e] = qe;
          vi[e] = v;
        }
        const float n1 = __fmul_rn(p.alpha, min1), n2 = __fmul_rn(p.alpha, min2);
        uint32_t nsg = 0u;
#pragma unroll
        for (int e = 0; e < deg; ++e) {
          const float mag = (uint32_t)e == im ? n2 : n1;
          const uint32_t s = sp ^ ((sb >> e) & 1u);
          nsg |= s << e;
          L[vi[e]] = __fadd_rn(q[e], s ? -mag : mag);
        }
        m1[r] = n1;
        m2[r] = n2;
        pk[r] = nsg | (im << 24);
      }
      __syncthreads();
    });
  }

  // hard decisions of the information bits, packed through LDS
  for (uint32_t k = threadIdx.x; k < (uint32_t)G::kb * zc; k += blockDim.x)
    if (L[k] < 0.f) atomicOr(&bits[k / 32], 1u << (k % 32));
  __syncthreads();
  for (uint32_t w = threadIdx.x; w < nbw; w += blockDim.x)
    p.out[cw * p.out_stride + w] = bits[w];
}

}  // namespace

// Lifting-size set of Zc (Table 5.3.2-1), or -1.
int ldpc_set_index(uint32_t zc) {
  static const uint32_t a[8] = {2, 3, 5, 7, 9, 11, 13, 15};
  for (int i = 0; i < 8; ++i)
    for (uint32_t z = a[i]; z <= 384; z *= 2)
      if (z == zc) return i;
  return -1;
}

hipError_t launch_ldpc_decode(const LdpcArgs& a, hipStream_t st, const char** why) {
  if (a.num_cb == 0) return hipSuccess;
  const int set = ldpc_set_index(a.zc);
  if (set < 0) {
    *why = "lifting size is not one of TS 38.212 Table 5.3.2-1";
    return hipErrorInvalidValue;
  }
  LdpcParams p;
  p.llr = a.llr;
  p.llr_stride = a.llr_stride;
  p.out = a.out;
  p.out_stride = a.out_stride;
  p.zc = a.zc;
  p.mb = a.mb;
  p.iters = a.iters;
  p.half = a.llr_bits == 16;
  p.alpha = a.alpha;
  const uint32_t kb = a.bg == 1 ? 22u : 10u;
  const uint32_t edges = a.bg == 1 ? ldpc::kBg1Edges : ldpc::kBg2Edges;
  const int16_t* row = a.bg == 1 ? ldpc::kBg1Shift[set] : ldpc::kBg2Shift[set];
  for (uint32_t e = 0; e < (uint32_t)kMaxEdges; ++e) p.shift[e] = e < edges ? row[e] : 0;
  hipError_t e = hipSuccess;
  const size_t lds = (size_t)(kb + a.mb) * a.zc * sizeof(float) +
                     (size_t)(edges + 1) / 2 * sizeof(float) +
                     (size_t)(kb * a.zc + 31) / 32 * sizeof(uint32_t);
  const dim3 grid(a.num_cb), block((a.zc + 63) / 64 * 64);
  const void* fn = a.bg == 1 ? reinterpret_cast<const void*>(&ldpc_decode<1>)
                             : reinterpret_cast<const void*>(&ldpc_decode<2>);
  if (lds > 65536) {
    e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  if (a.bg == 1)
    hipLaunchKernelGGL(ldpc_decode<1>, grid, block, lds, st, p);
  else
    hipLaunchKernelGGL(ldpc_decode<2>, grid, block, lds, st, p);
  return hipGetLastError();
}

}  // namespace detail
}  // namespace mega
