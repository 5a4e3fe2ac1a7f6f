// store_probe.hip -- the write ceiling of the order-2 output shape on this GPU: a kernel that only stores, in the
// layout the evaluation kernels write (per wavefront 64 consecutive matrices: their F block of 9, J block of 81 and H
// block of 729 values, each one contiguous range), with the same 16-byte vector stores and nothing else (no loads, no
// arithmetic beyond addresses).  Its rate bounds any kernel that writes these three streams, so the evaluation
// kernels are measured against it as well as against the HBM peak.  One wavefront per workgroup: measured on MI355X
// (n = 2^22, FP64 matrix layout) 4.69-4.70 ms, against 4.72 ms with two and 4.73-4.78 ms with four per workgroup,
// and 4.69 ms also with the evaluation kernel's 25 KB of LDS per workgroup (six per CU).
// The _in form first reads the wave's input rows (nin doubles per matrix, coalesced through LDS, as the evaluation
// kernels do) and stores values that depend on them: the evaluation's whole traffic, 72 B read per 6552 B written in
// the matrix layout.  Those reads cost more than their bytes: interleaved with the write stream they took the probe
// from 4.78 to 5.04 ms at 2^22 (the DRAM turns from writes to reads and back), while a wait of the same length
// without reads cost nothing -- so the _in form is the ceiling an evaluation kernel is held to.  No counterpart in
// the reference.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "../../include/matfun_hip.h"

namespace {

constexpr int PROBE_WAVES = 1, PROBE_WAVE = 64;

// cnt values from p (16-byte aligned) by one wavefront: 16-byte stores, the last cnt mod (16 / sizeof(T)) values
// one by one
template <class T>
__device__ void fill_block(T* __restrict__ p, int64_t cnt, int lane, T v) {
  using V = typename std::conditional<sizeof(T) == 8, double2, float4>::type;
  constexpr int PER = 16 / (int)sizeof(T);
  V w;
  if constexpr (sizeof(T) == 8) w = V{v, v};
  else w = V{v, v, v, v};
  const int64_t nv = cnt / PER;
  V* q = reinterpret_cast<V*>(p);
  for (int64_t i = lane; i < nv; i += PROBE_WAVE) q[i] = w;
  for (int64_t i = nv * PER + lane; i < cnt; i += PROBE_WAVE) p[i] = v;
}

// IN: the input rows x (nin doubles per matrix) read first; the stored value lane + 1 + 0 x (exact for finite x)
template <class T, bool IN>
__global__ void __launch_bounds__(PROBE_WAVE * PROBE_WAVES)
    store_probe_kernel(int64_t n, int nf, const double* __restrict__ x, int nin, T* __restrict__ F,
                       T* __restrict__ J, T* __restrict__ H) {
  const int w = threadIdx.x / PROBE_WAVE, lane = threadIdx.x - w * PROBE_WAVE;
  const int64_t m0 = ((int64_t)blockIdx.x * PROBE_WAVES + w) * PROBE_WAVE;
  if (m0 >= n) return;
  const int64_t nm = n - m0 < PROBE_WAVE ? n - m0 : PROBE_WAVE;
  T v = (T)(lane + 1);
  if constexpr (IN) {
    __shared__ double xin[PROBE_WAVES][9 * PROBE_WAVE];
    for (int t = lane; t < nin * nm; t += PROBE_WAVE) xin[w][t] = x[nin * m0 + t];
    __syncthreads();
    v = (T)((double)(lane + 1) + 0.0 * xin[w][lane < nm ? nin * lane : 0]);
  }
  fill_block(F + nf * m0, nf * nm, lane, v);
  if (J) fill_block(J + (int64_t)nf * nf * m0, (int64_t)nf * nf * nm, lane, v);
  if (H) fill_block(H + (int64_t)nf * nf * nf * m0, (int64_t)nf * nf * nf * nm, lane, v);
}

template <class T>
int probe(int64_t n, int nf, const double* x, int nin, T* F, T* J, T* H, void* stream) {
  if (n < 0 || (nf != 9 && nf != 6) || (n > 0 && !F) || (H && !J)) return -1;
  if (x && nin != 9 && nin != 6) return -1;
  if (((uintptr_t)F | (uintptr_t)J | (uintptr_t)H) & 15) return -1;         // (16-byte stores)
  if (n == 0) return 0;
  const int64_t per = (int64_t)PROBE_WAVE * PROBE_WAVES, g = (n + per - 1) / per;
  if (g > INT32_MAX) return -2;
  if (x)
    hipLaunchKernelGGL((store_probe_kernel<T, true>), dim3((unsigned)g), dim3(PROBE_WAVE * PROBE_WAVES), 0,
                       (hipStream_t)stream, n, nf, x, nin, F, J, H);
  else
    hipLaunchKernelGGL((store_probe_kernel<T, false>), dim3((unsigned)g), dim3(PROBE_WAVE * PROBE_WAVES), 0,
                       (hipStream_t)stream, n, nf, x, 0, F, J, H);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace

extern "C" int mf_hip_store_probe(int64_t n, int nf, double* F, double* J, double* H, void* stream) {
  return probe<double>(n, nf, nullptr, 0, F, J, H, stream);
}

extern "C" int mf_hip_store_probe_f32(int64_t n, int nf, float* F, float* J, float* H, void* stream) {
  return probe<float>(n, nf, nullptr, 0, F, J, H, stream);
}

extern "C" int mf_hip_store_probe_in(int64_t n, const double* x, int nin, int nf, double* F, double* J, double* H,
                                     void* stream) {
  if (n > 0 && !x) return -1;
  return probe<double>(n, nf, x, nin, F, J, H, stream);
}

extern "C" int mf_hip_store_probe_in_f32(int64_t n, const double* x, int nin, int nf, float* F, float* J, float* H,
                                         void* stream) {
  if (n > 0 && !x) return -1;
  return probe<float>(n, nf, x, nin, F, J, H, stream);
}
