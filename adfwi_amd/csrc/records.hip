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
//
// One thread per trace (s, r): consecutive threads take consecutive receivers, so every
// time sample is a coalesced row of the records.  Three passes over the trace: the maximum,
// the residual energy and sum rsd * a, then the gradients.
#include <hip/hip_runtime.h>

#include "adfwi_amd.h"
#include "common.hpp"

namespace {

__global__ __launch_bounds__(256) void trace_l2(const float *__restrict__ a, const float *__restrict__ b, int S,
                                                 int nt, int R, float dt, int normalize, float *__restrict__ loss,
                                                 float *__restrict__ ga, float *__restrict__ gb) {
  const long long tr = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tr >= (long long)S * R) return;
  const int s = (int)(tr / R), r = (int)(tr - (long long)s * R);
  const long long base = (long long)s * nt * R + r;
  float m = 1.0f;
  int tstar = 0;
  if (normalize) {
    m = -1.0f;
    for (int t = 0; t < nt; ++t) {
      const float v = fabsf(a[base + (long long)t * R]);
      if (v > m) { m = v; tstar = t; }
    }
  }
  float e = 0.0f, q = 0.0f;
  for (int t = 0; t < nt; ++t) {
    const long long i = base + (long long)t * R;
    const float av = a[i];
    const float n = normalize ? av / m : av;
    const float rsd = n - b[i];
    e += (rsd * rsd) * dt;
    q += rsd * av;
  }
  const float root = sqrtf(e);
  loss[tr] = root;
  if (!ga && !gb) return;
  const float w = dt / root;
  for (int t = 0; t < nt; ++t) {
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
                              int32_t normalize_a, float *trace_loss, float *grad_a, float *grad_b,
                              adfwi_stream_t stream) {
  if (int rc = adfwi::fault_injected("adfwi_trace_l2")) return rc;
  ADFWI_CHECK_ARG(a && b && trace_loss, "null record or trace-loss pointer");
  ADFWI_CHECK_ARG(nshots >= 1 && nt >= 1 && nrec >= 1, "record shape %d x %d x %d invalid", nshots, nt, nrec);
  ADFWI_CHECK_ARG(dt > 0.0f, "dt must be positive (got %g)", (double)dt);
  const long long ntr = (long long)nshots * nrec;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(trace_l2, dim3((unsigned)((ntr + 255) / 256)), dim3(256), 0, st, a, b, nshots, nt, nrec, dt,
                     normalize_a ? 1 : 0, trace_loss, grad_a, grad_b);
  return adfwi::check_launch("trace_l2");
}
