// gmf_tile.hpp -- one tile (up to 64 general real 3x3 matrices, one wavefront) of the general kernels:
//   phase 1, one matrix per lane: eigenvalues, divided differences, real basis, contracted tensors (gmf_device.hpp);
//            r_m (27 values) and D3 (27) staged in LDS per matrix, D2 (9) and F (9) in a second region
//   phase 2, lane groups, assembling J and H into an LDS tile in the API layout that the wave then writes as one
//            contiguous block (16-byte stores for consecutive rows, copy_block): J with 9 lanes per matrix (9 values
//            each), 7 matrices per pass; H with 27 lanes per matrix (27 values each), 2 matrices per pass, stored one
//            matrix (5.8 KB FP64) at a time (FP32: both)
// OT = double: the FP64 form; OT = float: the mixed form (FP64 per-lane math, FP32 staged tensors, J / H assembly
// and outputs).  PA: powers with one exponent per matrix (alv[i], read by its lane; bt then holds 1 / j; a matrix
// whose exponent is not finite gets NaN results).  Used by the homogeneous kernels (gmf_kernels.hip, RunRows) and the
// heterogeneous batch kernel (batch.hip, ListRows).
#pragma once
#include "gmf_device.hpp"
#include "tile_rows.hpp"

namespace mf {

// the reciprocals of the series-length bound, in constant memory (uniform reads: scalar loads)
static __constant__ GmfRecip g_gmf_recip = gmf_recip_table();

// The group series of the NW x 64 matrices of a workgroup (NW tiles), dealt out by length: a wavefront runs to the
// longest series among its lanes, and the lengths spread widely (log, order 1: median 37 terms, 1% above 147), so the
// lanes post their series (inputs and term count) to LDS, rank them by term count, and lane t of the workgroup
// computes the series of rank t; the sums return through LDS to the lanes that posted them.  Same arithmetic as the
// lane's own series (gmf_series): the results do not depend on where a series runs.
// LDS pool: GmfPool::BYTES, shared by the workgroup's tiles and free while they are in phase 1
template <int ORDER, int NW>
struct GmfPool {
  static constexpr int NJ = WAVE * NW, ND = ORDER == 0 ? 3 : (ORDER == 1 ? 9 : 19);
  static constexpr int IN = NJ * ((int)sizeof(GmfSeriesIn) + 4 + 2);      // inputs, term counts, job of each rank
  static constexpr int OUT = NJ * ND * (int)sizeof(cpx);                  // sums, [entry][job]
  static constexpr int BYTES = IN > OUT ? IN : OUT;
};
template <int KIND, int ORDER, bool PA, int NW>
struct GmfSeriesShared {
  const double* bt;
  unsigned char* pool;            // LDS, 16-byte aligned
  int tid;                        // threadIdx.x
  static constexpr bool K_ONLY = false;
  // every lane of the workgroup calls it once per tile (five workgroup barriers)
  __device__ void operator()(const GmfSeriesIn& s, cpx* acc) const {
    using P = GmfPool<ORDER, NW>;
    GmfSeriesIn* jin = reinterpret_cast<GmfSeriesIn*>(pool);
    int* key = reinterpret_cast<int*>(pool + P::NJ * sizeof(GmfSeriesIn));
    short* job_of = reinterpret_cast<short*>(key + P::NJ);
    __syncthreads();                                   // (every tile has read its input block: the pool is free)
    jin[tid] = s;
    key[tid] = s.K;
    __syncthreads();
    int r = 0;                                         // rank: longer series first, ties by position
    for (int t = 0; t < P::NJ; t += 4) {
      const int4 k4 = *reinterpret_cast<const int4*>(key + t);
      r += (k4.x > s.K || (k4.x == s.K && t < tid)) ? 1 : 0;
      r += (k4.y > s.K || (k4.y == s.K && t + 1 < tid)) ? 1 : 0;
      r += (k4.z > s.K || (k4.z == s.K && t + 2 < tid)) ? 1 : 0;
      r += (k4.w > s.K || (k4.w == s.K && t + 3 < tid)) ? 1 : 0;
    }
    job_of[r] = (short)tid;
    __syncthreads();
    const int job = job_of[tid];
    const GmfSeriesIn mine = jin[job];
    __syncthreads();                                   // (inputs read: the pool takes the sums)
    cpx a[P::ND];
    gmf_series<KIND, ORDER, PA>(mine, bt, a);
    cpx* res = reinterpret_cast<cpx*>(pool);
#pragma unroll
    for (int c = 0; c < P::ND; c++) res[c * P::NJ + job] = a[c];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < P::ND; c++) acc[c] = res[c * P::NJ + tid];
    __syncthreads();                                   // (sums read: the tiles may stage into the pool)
  }
};

// LDS layout (elements of OT):
//   main [64][MS]   r1 (9), r2 (9); order 2: the 10 distinct D3 (MS = 28; order 1: 19)
//   x    [XS]       phase 1: F of the 64 matrices [0, 576), then D2 (6 distinct) and sc1 per matrix [XAUX, +448)
//                   (FP32 only: AUXR);
//                   J passes: the JMPP x 81 output tile [0, XAUX); H passes: the 729 values of one matrix (FP32:
//                   of both matrices of the pass, HT)
// (the FP64 input block of 64 x 9 doubles overlays the start, before phase 1)
template <int ORDER, class OT>
struct GmfLds {
  static constexpr int MS = ORDER >= 2 ? 28 : 19;
  static constexpr int JG = 9;                              // J: lanes per matrix
  static constexpr int JMPP = WAVE / JG;                    // J: matrices per pass (7)
  static constexpr int XAUX = JMPP * 81;
  // FP64: D2 and sc1 reach the J lanes from their matrix's lane by shuffles (no aux region: order 2 stays under
  // 20 KB, eight wavefronts per CU; order 1 at 14 KB, eleven); FP32: staged in LDS after the J tile
  static constexpr bool AUXR = sizeof(OT) == 8;
  static constexpr int HT = sizeof(OT) == 4 ? 1458 : 729;  // H passes: both matrices (FP32), else one at a time
  static constexpr int XJ = XAUX + (AUXR ? 0 : 448);
  static constexpr int XS0 = ORDER == 0 || XJ < 576 ? 576 : XJ;   // (F of the 64 matrices: 576)
  static constexpr int XS = ORDER >= 2 && HT > XS0 ? HT : XS0;
  static constexpr int MAIN = ORDER == 0 ? 0 : 64 * MS;
  static constexpr int TOT = MAIN + XS;
  static constexpr int BYTES = (TOT * (int)sizeof(OT) > WAVE * 9 * 8) ? TOT * (int)sizeof(OT) : WAVE * 9 * 8;
};

// one tile of nm <= 64 matrices (nm = 0: none, the wave still takes part in the shared series); lds:
// GmfLds<ORDER, OT>::BYTES bytes, 16-byte aligned; lane: the lane in the tile's wavefront.  NW = 1: each lane computes
// its own series; NW > 1: the tile is one of NW in a workgroup of NW wavefronts (every one calling gmf_tile once) that
// share out the series (GmfSeriesShared; pool: its LDS, tid = threadIdx.x).  Ends with the tile's outputs written and
// a wave barrier (the LDS may be reused).
// the nodes and group series sums of each matrix computed by earlier passes (the staged evaluation, gmf_sorted.hip),
// at index k: nodes nre[k] and nim[6 k ..] (gmf_stage_nodes), the real parts of the sums sre[k ND ..], their
// imaginary parts sim[k ND ..] for a matrix of series class 2 (else +0).  k = row - off for the tile matrix's input
// row, or, with pos0 >= 0, pos0 + its place in the tile (a batch class's matrices staged in tile order)
struct GmfStaged {
  const GmfNodesRe* nre;
  const double* nim;
  const double* sre;
  const double* sim;
  int64_t off, pos0;
};

// STAGED: each lane finishes its matrix from staged nodes and sums (gmf_finish) instead of running gmf_lane.
// pre: a transform of the inputs first (pre_inputs in tile_rows.hpp)
template <int KIND, int ORDER, class OT, bool PA, class R, int NW = 1, bool STAGED = false, class Pre = NoPre>
MF_DEV void gmf_tile(const R& rows, int nm, const double* __restrict__ A, OT* __restrict__ F, OT* __restrict__ J,
                     OT* __restrict__ H, double al, const double* __restrict__ alv, const double* __restrict__ bt,
                     unsigned char* lds, int lane, unsigned char* pool = nullptr, int tid = 0,
                     GmfStaged staged = GmfStaged{nullptr, nullptr, nullptr, nullptr, 0, -1}, const Pre& pre = Pre()) {
  using L = GmfLds<ORDER, OT>;
  constexpr int VEC = 16 / (int)sizeof(OT);             // (elements per 16-byte store)
  double* xin = reinterpret_cast<double*>(lds);
  OT* mainr = reinterpret_cast<OT*>(lds);
  OT* xr = mainr + L::MAIN;

  load_in<9>(xin, A, rows, nm, lane);
  wave_sync();
  double a[9];
#pragma unroll
  for (int k = 0; k < 9; k++) a[k] = lane < nm ? xin[9 * lane + k] : (k % 4 == 0 ? 1.0 : 0.0);
  pre_inputs(pre, a, xin, rows, nm, lane);
  bool alok = true;                                  // (PA: a non-finite exponent computes with 1 and stores NaN)
  if (PA) {
    al = lane < nm ? alv[rows.m(lane < nm ? lane : 0)] : 1.0;
    alok = fabs(al) <= 1.7976931348623157e308;
    al = alok ? al : 1.0;
  }
  wave_sync();

  // ---- phase 1
  GmfOut o;
  if constexpr (STAGED) {
    constexpr int ND = ORDER == 0 ? 3 : (ORDER == 1 ? 9 : 19);
    const int ln = lane < nm ? lane : 0;
    const int64_t row = staged.pos0 >= 0 ? staged.pos0 + ln : rows.m(ln) - staged.off;
    const GmfNodesRe nr = staged.nre[row];
    const GmfNodes st = gmf_unstage_nodes(nr, staged.nim + 6 * row);
    cpx acc[ND];
#pragma unroll
    for (int c = 0; c < ND; c++) acc[c] = mk(staged.sre[row * ND + c], 0.0);
    if (((nr.flags >> 9) & 3u) == 2u) {
#pragma unroll
      for (int c = 0; c < ND; c++) acc[c].im = staged.sim[row * ND + c];
    }
    gmf_finish<KIND, ORDER, PA>(a, al, st, acc, o);
  } else if constexpr (NW > 1)
    gmf_lane<KIND, ORDER, PA>(a, al, bt, o, g_gmf_recip.r, GmfSeriesShared<KIND, ORDER, PA, NW>{bt, pool, tid});
  else
    gmf_lane<KIND, ORDER, PA>(a, al, bt, o, g_gmf_recip.r, GmfSeriesLocal<KIND, ORDER, PA>{bt});
  if (PA) {                                          // (J and H carry the factors sc1, sc2)
    const double nanf = alok ? 1.0 : dnan();
#pragma unroll
    for (int k = 0; k < 9; k++) o.F[k] *= nanf;
    o.sc1 *= nanf;
    o.sc2 *= nanf;
  }
  const OT sc2 = ORDER >= 2 ? (OT)o.sc2 : (OT)0;     // (order 2: H's factor, read by the H lanes through a shuffle)
  if (ORDER >= 1) {
    OT* my = mainr + lane * L::MS;
#pragma unroll
    for (int k = 0; k < 9; k++) { my[k] = (OT)o.r[1][k]; my[9 + k] = (OT)o.r[2][k]; }
    if (ORDER >= 2) {
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = i; j < 3; j++)
#pragma unroll
          for (int k = j; k < 3; k++) my[18 + sym33s(i, j, k)] = (OT)o.D3[9 * i + 3 * j + k];
    }
  }
#pragma unroll
  for (int k = 0; k < 9; k++) xr[lane * 9 + k] = (OT)o.F[k];
  wave_sync();
  if constexpr (R::RUN) {
    copy_block<VEC>(F + rows.m(0) * 9, xr, 9 * nm, lane);
  } else if constexpr (R::SCATTER) {
    copy_rows_scattered<9, R_IN>(F, rows, xr, 9 * nm, lane);
  } else {                                           // (rows listed: runs of consecutive rows as blocks)
    for_row_runs<R_IN>(rows, nm, lane,
                       [&](int j0, int j1) { copy_block<VEC>(F + rows.m(j0) * 9, xr + 9 * j0, 9 * (j1 - j0), lane); });
  }
  if (ORDER >= 1) {
    OT d2r[7];                                       // D2 (6 distinct, by sym3) and sc1 of this lane's matrix
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = i; j < 3; j++) d2r[sym3(i, j)] = (OT)o.D2[3 * i + j];
    d2r[6] = (OT)o.sc1;
    if constexpr (!L::AUXR) {
      wave_sync();
#pragma unroll
      for (int k = 0; k < 7; k++) xr[L::XAUX + lane * 7 + k] = d2r[k];
      wave_sync();
    }

    // ---- phase 2, J: lane (gj, Lj) = matrix gj of the pass, Lj = 3 c + d: J_abcd for a, b = 0..2, each value at
    //      27 a + 9 b + 3 c + d of its matrix
    constexpr int JG = L::JG, JMPP = L::JMPP;
    const int gj = lane / JG, Lj = lane - gj * JG;
    for (int p0 = 0; p0 < nm; p0 += JMPP) {
      const int cnt = nm - p0 < JMPP ? nm - p0 : JMPP;
      OT dj[7];
      if constexpr (L::AUXR) {                       // (from the matrix's lane; every lane takes part)
#pragma unroll
        for (int k = 0; k < 7; k++) dj[k] = __shfl(d2r[k], (p0 + (gj < JMPP ? gj : 0)) & (WAVE - 1));
      }
      if (gj < cnt) {
        const int j = p0 + gj;
        const OT* s = mainr + j * L::MS;
        if constexpr (!L::AUXR) {
#pragma unroll
          for (int k = 0; k < 7; k++) dj[k] = xr[L::XAUX + j * 7 + k];
        }
        OT out[3];
#pragma unroll
        for (int b = 0; b < 3; b++) {
          gmf_j_lane(s, s + 9, dj, dj[6], b, Lj / 3, Lj % 3, out);
#pragma unroll
          for (int k = 0; k < 3; k++) xr[gj * 81 + 27 * k + 9 * b + Lj] = out[k];
        }
      }
      wave_sync();
      if constexpr (R::RUN) {
        copy_block<VEC>(J + rows.jr(p0) * 81, xr, 81 * cnt, lane);
      } else if constexpr (R::SCATTER) {
        copy_rows_scattered<81, R_J>(J, shift_rows(rows, p0), xr, 81 * cnt, lane);
      } else {
        const R sub = shift_rows(rows, p0);
        for_row_runs<R_J>(sub, cnt, lane, [&](int j0, int j1) {
          copy_block<VEC>(J + sub.jr(j0) * 81, xr + 81 * j0, 81 * (j1 - j0), lane);
        });
      }
      wave_sync();
    }
    if (ORDER >= 2) {
      constexpr int G = 27, MPP = WAVE / G;            // H: lane (g, Lg), Lg = 9 d + 3 e + f
      const int g = lane / G, Lg = lane - g * G;
      for (int p0 = 0; p0 < nm; p0 += MPP) {        // H: H_abcdef, (a,b,c) = 0..26 at 27 (9a + 3b + c) + (9d + 3e + f)
        const int cnt = nm - p0 < MPP ? nm - p0 : MPP;
        const OT s2 = __shfl(sc2, (p0 + g) & (WAVE - 1));   // (sc2 of the lane group's matrix, from its lane)
        const OT* s = mainr + (p0 + g) * L::MS;
        OT* ys = xr + g * 243;                         // the 27 lanes' X1 (gmf_h_y), in the H tile's space
        const int d = Lg / 9, e = (Lg / 3) % 3, f = Lg % 3;
        OT y[9], out[27];
        if (g < cnt) {
          gmf_h_y(s, s + 9, s + 18, d, e, f, y);
#pragma unroll
          for (int k = 0; k < 9; k++) ys[9 * Lg + k] = y[k];
        }
        wave_sync();
        if (g < cnt) gmf_h_out(s, s + 9, s2, d, e, f, y, ys, out);
        wave_sync();                                   // (X1 read: the H tile takes the outputs)
        if constexpr (L::HT == 1458 && R::RUN) {   // both matrices staged, stored as one block (consecutive rows)
          if (g < cnt) {
#pragma unroll
            for (int k = 0; k < 27; k++) xr[g * 729 + 27 * k + Lg] = out[k];
          }
          wave_sync();
          copy_block<VEC>(H + rows.hr(p0) * 729, xr, 729 * cnt, lane);
          wave_sync();
        } else {
          for (int gg = 0; gg < cnt; gg++) {       // one matrix at a time: its 729 values staged, then stored
            if (g == gg) {
#pragma unroll
              for (int k = 0; k < 27; k++) xr[27 * k + Lg] = out[k];
            }
            wave_sync();
            copy_block<VEC>(H + rows.hr(p0 + gg) * 729, xr, 729, lane);
            wave_sync();
          }
        }
      }
    }
  }
  wave_sync();                                   // (the LDS is free for the next tile)
}

}  // namespace mf
