// The FP64 finaliser of one accumulator row, shared by the finaliser kernel
// (panoptic.hip) and the AP kernel that runs it beside the AP classes
// (instance.hip).  Needs kFinalThreads threads and 16 * C bytes of dynamic LDS.
#pragma once

#include "panoptic_kernels.h"

namespace pe {

constexpr int kFinalThreads = 1024;

// One workgroup per accumulator row.  Confusion row and column sums come from
// one coalesced sweep into LDS (64-bit LDS atomics); per-class ratios follow
// the reference CPU formulas exactly; group means reduce in a fixed order
// (butterfly within each wave, then waves in index order).
__device__ inline void finalize_row(const FinalArgs& a, i64 k, uint4* lds4) {
  u64* rowsum = reinterpret_cast<u64*>(lds4);
  u64* colsum = rowsum + a.C;
  constexpr int kWaves = kFinalThreads / kWave;
  __shared__ double red[14][kWaves];
  __shared__ double total[14];
  const int C = a.C, tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int C1 = C + 1;
  const int CC = C * C1;  // the launcher caps C at 7680 (LDS), so C * (C + 1) fits an int
  const i64* tp = a.tp + k * C;
  const i64* fp = a.fp + k * C;
  const i64* fn = a.fn + k * C;
  const double* sums = a.sums + k * C;
  const i64* conf = a.confusion + k * CC;
  double* per_class = a.per_class + k * 4 * C;
  for (int c = tid; c < 2 * C; c += kFinalThreads) rowsum[c] = 0;
  __syncthreads();
  // Twenty-four loads in flight per lane (one step covers C up to 156): the
  // sweep is latency-bound otherwise.  Indices past the table are clamped
  // and their values dropped, so no load waits behind a branch.
  constexpr int U = 24;
  for (int i0 = tid; i0 < CC; i0 += kFinalThreads * U) {
    i64 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = conf[min(i0 + u * kFinalThreads, CC - 1)];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (i0 + u * kFinalThreads >= CC) v[u] = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = i0 + u * kFinalThreads;
      if (i >= CC) break;
      if (a.out_confusion) a.out_confusion[k * CC + i] = v[u];
      if (!v[u]) continue;
      const int r = i / C1, col = i - r * C1;
      atomicAdd(rowsum + r, static_cast<u64>(v[u]));
      if (col < C) atomicAdd(colsum + col, static_cast<u64>(v[u]));
    }
  }
  __syncthreads();
  double part[14];
#pragma unroll
  for (int j = 0; j < 14; ++j) part[j] = 0.0;
  for (int c = tid; c < C; c += kFinalThreads) {
    // Every load of the class first (one round trip), then the ratios.
    const i64 t = tp[c], f = fp[c], n = fn[c];
    const double s = sums[c];
    const i64 diag = conf[c * C1 + c];
    const bool thing = a.things[c] != 0;
    if (a.out_tp) {
      a.out_tp[k * C + c] = t;
      a.out_fp[k * C + c] = f;
      a.out_fn[k * C + c] = n;
      a.out_sums[k * C + c] = s;
    }
    const double denom = static_cast<double>(t) + 0.5 * static_cast<double>(f + n);
    const double dd = denom < 1.0 ? 1.0 : denom;
    const double pq = s / dd;
    const double sq = s / static_cast<double>(t < 1 ? 1 : t);
    const double rq = static_cast<double>(t) / dd;
    per_class[c] = pq;
    per_class[C + c] = sq;
    per_class[2 * C + c] = rq;
    const i64 uni = static_cast<i64>(rowsum[c]) + static_cast<i64>(colsum[c]) - diag;
    const double iou = uni > 0 ? static_cast<double>(diag) / static_cast<double>(uni) : __builtin_nan("");
    per_class[3 * C + c] = iou;
    if (uni > 0) {
      part[12] += iou;
      part[13] += 1.0;
    }
    if (denom > 0) {
      const int group = thing ? 1 : 2;
      part[0] += pq; part[1] += sq; part[2] += rq; part[3] += 1.0;
      part[group * 4] += pq; part[group * 4 + 1] += sq; part[group * 4 + 2] += rq; part[group * 4 + 3] += 1.0;
    }
  }
  // Only waves that own classes reduce (the others' partials are +0.0, and
  // adding +0.0 changes no sum).
  const int active_waves = min(kWaves, (C + kWave - 1) / kWave);
  if (wave < active_waves) {
#pragma unroll
    for (int j = 0; j < 14; ++j) part[j] = wave_sum(part[j]);
    if (lane == 0)
#pragma unroll
      for (int j = 0; j < 14; ++j) red[j][wave] = part[j];
  }
  __syncthreads();
  if (tid < 14) {
    double s = 0.0;
    for (int w = 0; w < active_waves; ++w) s += red[tid][w];
    total[tid] = s;
  }
  __syncthreads();
  if (tid < 3) {
    const double n = total[tid * 4 + 3];
    a.counts[k * 3 + tid] = static_cast<i64>(n);
    for (int j = 0; j < 3; ++j) a.summaries[k * 9 + tid * 3 + j] = n > 0 ? total[tid * 4 + j] / n : 0.0;
  }
  if (tid == 0) a.miou[k] = total[13] > 0 ? total[12] / total[13] : 0.0;
}

}  // namespace pe
