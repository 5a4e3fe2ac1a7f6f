// smf_kernels.hip -- symmetric 3x3 matrices on gfx950, one wavefront (64 matrices) per workgroup.
//   phase A, one matrix per lane: coalesced input through LDS, eigen-decomposition, F, the 21 distinct J values,
//            the 10 second divided differences; F / J staged in LDS and written as one contiguous region per wave
//   phase B, six lanes per matrix (one per Voigt row of H), ten matrices per pass: H rows from V and D2 staged in
//            LDS, gathered into the API layout through a lookup table and written contiguously (a pass covers
//            10 x 5.8 KB of consecutive output)
// Every store instruction of the wave writes 64 consecutive elements: the output (96 B .. 5.8 KB per matrix) bounds
// the kernel by HBM write bandwidth, which needs whole cache lines per instruction.
// OT = double: the FP64 form; OT = float: the mixed form (FP64 eigen-decomposition and divided differences, FP32
// derivative assembly and FP32 outputs).  PA: powers with one exponent per matrix (alv[i], read by its lane; a
// matrix whose exponent is not finite gets NaN results).
#include "smf_device.hpp"
#include "smf_launch.hpp"

namespace mf {

constexpr int WAVE = 64;
constexpr int SA = 27;             // phase-A tile per matrix: F (6) + Js (21)
constexpr int SB = 19;             // V (9) + D2 (10)
constexpr int HPP = 10;            // matrices per phase-B pass (6 lanes each)
constexpr int SH = 127;            // staged H values per matrix (6 rows x 21, +1: odd stride)

// out[t] = get(j, e) for t = j N + e < count, by the wave: with PAIRS two consecutive elements per lane and one
// vector store (16 B FP64 / 8 B FP32; out must be aligned to it), else one element per lane
template <int N, bool PAIRS, class OT, class Get>
MF_DEV void copy_out(OT* __restrict__ out, int count, int lane, Get get) {
  if (PAIRS) {
    typedef OT v2 __attribute__((ext_vector_type(2)));
    int t = 2 * lane, j = t / N, e = t - j * N;
    for (; t + 1 < count; t += 2 * WAVE) {
      const int e1 = e + 1 == N ? 0 : e + 1, j1 = e + 1 == N ? j + 1 : j;
      const v2 val = {get(j, e), get(j1, e1)};
      *reinterpret_cast<v2*>(out + t) = val;
      e += 2 * WAVE;
      while (e >= N) { e -= N; j++; }
    }
    if (t < count) out[t] = get(j, e);         // (odd count: the last element)
  } else {
    int j = lane / N, e = lane - j * N;
    for (int t = lane; t < count; t += WAVE) {
      out[t] = get(j, e);
      e += WAVE;
      while (e >= N) { e -= N; j++; }
    }
  }
}

template <int KIND, bool MAT, int ORDER, class OT, bool PAIRS, bool PA>
__global__ void __launch_bounds__(WAVE) smf_kernel(int64_t n, const double* __restrict__ x, OT* __restrict__ F,
                                                   OT* __restrict__ J, OT* __restrict__ H, double al,
                                                   const double* __restrict__ alv, SmfCoef c2) {
  constexpr int NIN = MAT ? 9 : 6, NF = MAT ? 9 : 6, NJ = NF * NF, NH = NJ * NF;
  constexpr int R1 = (WAVE * SA > HPP * SH ? WAVE * SA : HPP * SH);             // region 1 (OT): phase-A / H tiles
  constexpr int R1B = (R1 * (int)sizeof(OT) > WAVE * NIN * 8 ? R1 * (int)sizeof(OT) : WAVE * NIN * 8);
  __shared__ double raw1[(R1B + 7) / 8];                                         // (also the FP64 input block)
  __shared__ OT tb[ORDER >= 2 ? WAVE * SB : 1];
  __shared__ unsigned short lutJ[ORDER >= 1 ? NJ : 1], lutH[ORDER >= 2 ? NH : 1];
  double* xin = raw1;
  OT* t1 = reinterpret_cast<OT*>(raw1);
  const int lane = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * WAVE;
  const int nm = (int)(n - base < WAVE ? n - base : WAVE);

  // ---- lookup tables of the output layouts: distinct-value position | weight code << 8 (weight 2^-code, or 2)
  if (ORDER >= 1)
    for (int e = lane; e < NJ; e += WAVE) {
      double w;
      const int s = j_src(MAT, e, w);
      lutJ[e] = (unsigned short)(s | (w == 2.0 ? 3 << 8 : 0));
    }
  if (ORDER >= 2)
    for (int e = lane; e < NH; e += WAVE) {
      double w;
      const int s = h_src(MAT, e, w);
      lutH[e] = (unsigned short)(s | ((w == 1.0 ? 0 : (w == 0.5 ? 1 : 2)) << 8));
    }

  // ---- input, coalesced through LDS (lanes past the end take the identity)
  for (int t = lane; t < NIN * nm; t += WAVE) xin[t] = x[base * NIN + t];
  __syncthreads();
  double v[6];
  if (MAT) {
    double a[9];
#pragma unroll
    for (int k = 0; k < 9; k++) a[k] = lane < nm ? xin[9 * lane + k] : (k % 4 == 0 ? 1.0 : 0.0);
    matrix_to_voigt(a, v);
  } else {
#pragma unroll
    for (int k = 0; k < 6; k++) v[k] = lane < nm ? xin[6 * lane + k] : (k < 3 ? 1.0 : 0.0);
  }
  __syncthreads();
  bool alok = true;                            // (PA: a non-finite exponent computes with 1 and stores NaN)
  if (PA) {
    al = lane < nm ? alv[base + lane] : 1.0;
    alok = fabs(al) <= 1.7976931348623157e308;
    al = alok ? al : 1.0;
  }
  const double nanf = alok ? 1.0 : dnan();

  // ---- phase A
  {
    SmfLane<KIND, OT> s;
    double Fv[6];
    s.eigen(v, al, Fv);
#pragma unroll
    for (int k = 0; k < 6; k++) t1[lane * SA + k] = (OT)(Fv[k] * nanf);
    if (ORDER >= 1) {
      OT Js[21];
      s.jacobian(al, Js);
#pragma unroll
      for (int k = 0; k < 21; k++) t1[lane * SA + 6 + k] = Js[k] * (OT)nanf;
    }
    if (ORDER >= 2) {
      s.template second_dd<PA>(al, c2.c);
#pragma unroll
      for (int k = 0; k < 9; k++) tb[lane * SB + k] = (OT)s.V[k / 3][k % 3];
#pragma unroll
      for (int k = 0; k < 10; k++) tb[lane * SB + 9 + k] = (OT)(s.D2[k] * nanf);
    }
  }
  __syncthreads();
  for (int t = lane; t < NF * nm; t += WAVE) {
    const int j = t / NF, e = t - j * NF;
    F[base * NF + t] = t1[j * SA + f_src(MAT, e)];
  }
  if (ORDER < 1) return;
  copy_out<NJ, PAIRS>(J + base * NJ, NJ * nm, lane, [&](int j, int e) {
    const int u = lutJ[e];
    const OT val = t1[j * SA + 6 + (u & 255)];
    return (u >> 8) ? (OT)2 * val : val;
  });
  if (ORDER < 2) return;

  // ---- phase B: lane = (matrix jj of the pass, Voigt row i)
  const int jj = lane / 6, ir = lane - 6 * jj;
  for (int p0 = 0; p0 < nm; p0 += HPP) {
    const int cnt = nm - p0 < HPP ? nm - p0 : HPP;
    __syncthreads();                           // (previous tile fully read)
    if (jj < cnt) {
      const OT* sb = tb + (p0 + jj) * SB;
      const SmfHess<OT> hs(sb, sb + 9);
      OT h[21];
      hs.h_row(ir, h);
#pragma unroll
      for (int k = 0; k < 21; k++) t1[jj * SH + ir * 21 + k] = h[k];
    }
    __syncthreads();
    copy_out<NH, PAIRS>(H + (base + p0) * NH, NH * cnt, lane, [&](int j, int e) {
      const int u = lutH[e];
      const OT val = t1[j * SH + (u & 255)];
      return (u >> 8) == 0 ? val : ((u >> 8) == 1 ? (OT)0.5 * val : (OT)0.25 * val);
    });
  }
}

template <int KIND, bool MAT, class OT, bool PAIRS, bool PA>
static void launch_order(hipStream_t st, unsigned blocks, int64_t n, const double* x, OT* F, OT* J, OT* H, double al,
                         const double* alv, const SmfCoef& c2) {
  if (H)
    hipLaunchKernelGGL((smf_kernel<KIND, MAT, 2, OT, PAIRS, PA>), dim3(blocks), dim3(WAVE), 0, st, n, x, F, J, H, al,
                       alv, c2);
  else if (J)
    hipLaunchKernelGGL((smf_kernel<KIND, MAT, 1, OT, PAIRS, PA>), dim3(blocks), dim3(WAVE), 0, st, n, x, F, J, H, al,
                       alv, c2);
  else
    hipLaunchKernelGGL((smf_kernel<KIND, MAT, 0, OT, false, PA>), dim3(blocks), dim3(WAVE), 0, st, n, x, F, J, H, al,
                       alv, c2);
}

// paired stores need J and H aligned to two elements (the per-wave regions then stay aligned: 64 and 10 matrices of
// 81 / 729 (36 / 216) elements are even counts)
template <int KIND, bool MAT, class OT, bool PA>
static void launch_kind(hipStream_t st, unsigned blocks, int64_t n, const double* x, OT* F, OT* J, OT* H, double al,
                        const double* alv, const SmfCoef& c2) {
  const uintptr_t al2 = 2 * sizeof(OT) - 1;
  if ((!J || ((uintptr_t)J & al2) == 0) && (!H || ((uintptr_t)H & al2) == 0))
    launch_order<KIND, MAT, OT, true, PA>(st, blocks, n, x, F, J, H, al, alv, c2);
  else
    launch_order<KIND, MAT, OT, false, PA>(st, blocks, n, x, F, J, H, al, alv, c2);
}

template <class OT>
static void launch_any(hipStream_t st, int kind, bool mat, int64_t n, const double* x, OT* F, OT* J, OT* H, double al,
                       const SmfCoef& c2) {
  const unsigned blocks = (unsigned)((n + WAVE - 1) / WAVE);
#define MF_CASE(K)                                                                        \
  if (mat) launch_kind<K, true, OT, false>(st, blocks, n, x, F, J, H, al, nullptr, c2);   \
  else launch_kind<K, false, OT, false>(st, blocks, n, x, F, J, H, al, nullptr, c2);
  switch (kind) {
    case EXP: MF_CASE(EXP) break;
    case LOG: MF_CASE(LOG) break;
    case SQRT: MF_CASE(SQRT) break;
    default: MF_CASE(POW) break;
  }
#undef MF_CASE
}

template <class OT>
static void launch_pow(hipStream_t st, bool mat, int64_t n, const double* alv, const double* x, OT* F, OT* J, OT* H) {
  const unsigned blocks = (unsigned)((n + WAVE - 1) / WAVE);
  const SmfCoef unused{};
  if (mat) launch_kind<POW, true, OT, true>(st, blocks, n, x, F, J, H, 1.0, alv, unused);
  else launch_kind<POW, false, OT, true>(st, blocks, n, x, F, J, H, 1.0, alv, unused);
}

void smf_launch(hipStream_t st, int kind, bool mat, int64_t n, const double* x, double* F, double* J, double* H,
                double al, const SmfCoef& c2) {
  launch_any<double>(st, kind, mat, n, x, F, J, H, al, c2);
}

void smf_launch_mixed(hipStream_t st, int kind, bool mat, int64_t n, const double* x, float* F, float* J, float* H,
                      double al, const SmfCoef& c2) {
  launch_any<float>(st, kind, mat, n, x, F, J, H, al, c2);
}

void smf_launch_pow(hipStream_t st, bool mat, int64_t n, const double* alv, const double* x, double* F, double* J,
                    double* H) {
  launch_pow<double>(st, mat, n, alv, x, F, J, H);
}

void smf_launch_pow_mixed(hipStream_t st, bool mat, int64_t n, const double* alv, const double* x, float* F, float* J,
                          float* H) {
  launch_pow<float>(st, mat, n, alv, x, F, J, H);
}

}  // namespace mf
