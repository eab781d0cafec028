// Record post-processing on the GPU: the per-trace waveform L2 misfit of the FWI drivers,
// optionally with the per-trace max normalisation of the synthetic traces, and its adjoint
// source, in one kernel over the [S][nt][R] records.
//
//   n[s,t,r] = a[s,t,r] / max_t |a[s,.,r]|     (normalize_a; fwi/acoustic_fwi.py:149-150)
//   rsd      = n - b,  e[s,r] = sum_t rsd^2 dt,  trace loss = sqrt(e)   (fwi/misfit/L2.py:26-29)
//   d loss / d n[t] = w rsd[t],  w = dt / sqrt(e)
//   d loss / d a[t] = w rsd[t] / m  -  [t == t*] sign(a[t*]) w (sum_t' rsd[t'] a[t']) / m^2
// with m = max_t |a| reached first at t* (the max's gradient goes to that sample, as torch's
// max over a dimension routes it); d loss / d b = -w rsd.  A trace with zero residual has
// e = 0 and, as in the reference, a NaN gradient.
// Least-squares mode (ADFWI_L2_SQUARED): trace loss = e / 2 and w = dt, the objective
// 1/2 sum (n - b)^2 dt, whose gradient is defined on every trace (no counterpart in the
// reference's misfit set).
//
// One workgroup per (shot, 64 receivers): each lane takes one receiver, so every time sample
// is a coalesced 256-byte row of the records, and the TW wavefronts of the workgroup take
// consecutive chunks of the time axis (enough wavefronts to keep the memory system busy; one
// lane per whole trace left most of the chip idle).  Three passes over the trace: the maximum,
// the residual energy and sum rsd * a, then the gradients; the wavefronts combine their
// partial maxima and sums through LDS in a fixed order (the first maximum in time wins, the
// sums are added wavefront 0 first), so results do not depend on timing.
#include <hip/hip_runtime.h>

#include "adfwi_amd.h"
#include "common.hpp"

namespace {

constexpr int TW = 8;   // wavefronts per workgroup: chunks of the time axis

__global__ __launch_bounds__(TW * 64) void trace_l2(const float *__restrict__ a, const float *__restrict__ b,
                                                    int nt, int R, float dt, int normalize, int squared,
                                                    float *__restrict__ loss,
                                                    float *__restrict__ ga, float *__restrict__ gb) {
  __shared__ float part_m[TW][64], part_e[TW][64], part_q[TW][64];
  __shared__ int part_t[TW][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ngr = (R + 63) / 64;
  const int s = blockIdx.x / ngr, r = (blockIdx.x - s * ngr) * 64 + lane;
  const bool act = r < R;   // lanes past the last receiver only take part in the reductions
  const int chunk = (nt + TW - 1) / TW;
  const int t0 = min(nt, wv * chunk), t1 = min(nt, t0 + chunk);
  const long long base = (long long)s * nt * R + r;
  float m = 1.0f;
  int tstar = 0;
  if (normalize) {
    float lm = -1.0f;
    int lt = 0;
    if (act) {
#pragma unroll 4
      for (int t = t0; t < t1; ++t) {
        const float v = fabsf(a[base + (long long)t * R]);
        if (v > lm) { lm = v; lt = t; }
      }
    }
    part_m[wv][lane] = lm;
    part_t[wv][lane] = lt;
    __syncthreads();
    m = -1.0f;
    for (int k = 0; k < TW; ++k) {   // earlier chunks first: the first maximum in time wins
      const float v = part_m[k][lane];
      if (v > m) { m = v; tstar = part_t[k][lane]; }
    }
  }
  float e = 0.0f, q = 0.0f;
  if (act) {
#pragma unroll 4
    for (int t = t0; t < t1; ++t) {
      const long long i = base + (long long)t * R;
      const float av = a[i];
      const float n = normalize ? av / m : av;
      const float rsd = n - b[i];
      e += (rsd * rsd) * dt;
      q += rsd * av;
    }
  }
  part_e[wv][lane] = e;
  part_q[wv][lane] = q;
  __syncthreads();
  e = 0.0f;
  q = 0.0f;
  for (int k = 0; k < TW; ++k) {
    e += part_e[k][lane];
    q += part_q[k][lane];
  }
  const float root = sqrtf(e);
  if (!act) return;
  if (wv == 0) loss[(long long)s * R + r] = squared ? 0.5f * e : root;
  if (!ga && !gb) return;
  const float w = squared ? dt : dt / root;
#pragma unroll 4
  for (int t = t0; t < t1; ++t) {
    const long long i = base + (long long)t * R;
    const float av = a[i];
    const float n = normalize ? av / m : av;
    const float g = w * (n - b[i]);
    if (ga) {
      float v = normalize ? g / m : g;
      if (normalize && t == tstar) {
        const float sg = av > 0.0f ? 1.0f : (av < 0.0f ? -1.0f : 0.0f);
        v = v - sg * ((w * q) / (m * m));
      }
      ga[i] = v;
    }
    if (gb) gb[i] = -g;
  }
}

}  // namespace

extern "C" int adfwi_trace_l2(const float *a, const float *b, int32_t nshots, int32_t nt, int32_t nrec, float dt,
                              int32_t flags, float *trace_loss, float *grad_a, float *grad_b,
                              adfwi_stream_t stream) {
  if (int rc = adfwi::fault_injected("adfwi_trace_l2")) return rc;
  ADFWI_CHECK_ARG(a && b && trace_loss, "null record or trace-loss pointer");
  ADFWI_CHECK_ARG(nshots >= 1 && nt >= 1 && nrec >= 1, "record shape %d x %d x %d invalid", nshots, nt, nrec);
  ADFWI_CHECK_ARG(dt > 0.0f, "dt must be positive (got %g)", (double)dt);
  ADFWI_CHECK_ARG((flags & ~(ADFWI_L2_NORMALIZE | ADFWI_L2_SQUARED)) == 0, "unknown flags 0x%x", flags);
  const long long nblocks = (long long)nshots * ((nrec + 63) / 64);
  ADFWI_CHECK_ARG(nblocks < (1LL << 31), "record shape %d x %d x %d too large for one launch", nshots, nt, nrec);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(trace_l2, dim3((unsigned)nblocks), dim3(TW * 64), 0, st, a, b, nt, nrec, dt,
                     (flags & ADFWI_L2_NORMALIZE) ? 1 : 0, (flags & ADFWI_L2_SQUARED) ? 1 : 0, trace_loss,
                     grad_a, grad_b);
  return adfwi::check_launch("trace_l2");
}
