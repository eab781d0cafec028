// Acoustic p-u-w staggered-grid propagator for gfx950: fused forward step,
// fused discrete-adjoint step, and the C ABI around them.
//
// One forward step (reference: ADFWI/propagator/acoustic_kernels.py:113-174):
//   1. p <- (1-K1) p - A1 (c1 (u_j - u_{j-1} + w_i - w_{i-1}) + c2 (u_{j+1} - u_{j-2} + w_{i+1} - w_{i-2}))
//      on rows [fs+1, Z-3] x cols [2, X-3]
//   2. p[zs, xs] += dt * wavelet[t]
//   3. free surface: p[fs-1, :] = -p[fs+1, :]
//   4. u <- (1-K2) u - A2 (c1 (p_{j+1} - p_j) + c2 (p_{j+2} - p_{j-1}))  on [fs, Z-2] x [1, X-3]
//   5. w <- (1-K3) w - A2 (c1 (p_{i+1} - p_i) + c2 (p_{i+2} - p_{i-1}))  on [fs, Z-3] x [1, X-2]
//   6. free surface: w[fs-1, :] = w[fs, :]
//   7. records of p, u, w; 8. energy maps.
// Every expression keeps the reference's fp32 association (the library is
// compiled with -ffp-contract=off), so a step is bit-identical to the
// reference's CPU op sequence given identical coefficient fields.
//
// A workgroup owns a TZ x TX output tile for a group of shots.  Per shot it
// stages u/w halos and p on the extended tile [z0-1, z0+TZ+1] x [x0-1, x0+TX+1]
// in LDS, recomputes the new p there (source and free-surface row included),
// then updates u and w on the tile.  State is ping-ponged between two buffer
// sets, so halo reads never race with neighbouring tiles' writes.
#include "common.hpp"

namespace {

constexpr int TX = 64;   // tile width: one wavefront per row -> 256-byte coalesced rows
constexpr int TZ = 8;    // tile height
constexpr int NT = 256;  // threads per workgroup (4 wavefronts), each thread owns 2 tile points
constexpr int PTS = TZ * TX / NT;
constexpr int EH = TZ + 3, EW = TX + 3;  // extended tile rows z0-1.., cols x0-1..
constexpr int UH = TZ + 3, UW = TX + 6;  // u halo rows z0-1.., cols x0-3..
constexpr int WH = TZ + 6, WW = TX + 3;  // w halo rows z0-3.., cols x0-1..

constexpr float C1 = 9.0f / 8.0f;
constexpr float C2 = static_cast<float>(-1.0 / 24.0);

struct AcDev {
  int S, Z, X, pitch, fs, fsurf, nt, nrec;
  int zbase, ntx, spb;
  long long plane;  // Z * pitch
  long long field;  // S * Z * pitch
  const float *coef;
  const float *src;
  const int *src_z, *src_x, *rcv_z, *rcv_x, *bin_start, *bin_list;
};

__device__ __forceinline__ bool win_p(const AcDev &d, int i, int j) {
  return i >= d.fs + 1 && i <= d.Z - 3 && j >= 2 && j <= d.X - 3;
}
__device__ __forceinline__ bool win_u(const AcDev &d, int i, int j) {
  return i >= d.fs && i <= d.Z - 2 && j >= 1 && j <= d.X - 3;
}
__device__ __forceinline__ bool win_w(const AcDev &d, int i, int j) {
  return i >= d.fs && i <= d.Z - 3 && j >= 1 && j <= d.X - 2;
}
// valid cells of each field: p (Z, X), u (Z, X-1), w (Z-1, X)
__device__ __forceinline__ bool ok_p(const AcDev &d, int i, int j) {
  return i >= 0 && i < d.Z && j >= 0 && j < d.X;
}
__device__ __forceinline__ bool ok_u(const AcDev &d, int i, int j) {
  return i >= 0 && i < d.Z && j >= 0 && j < d.X - 1;
}
__device__ __forceinline__ bool ok_w(const AcDev &d, int i, int j) {
  return i >= 0 && i < d.Z - 1 && j >= 0 && j < d.X;
}

__device__ __forceinline__ void tile_origin(const AcDev &d, int tile, int &z0, int &x0) {
  const int tz = tile / d.ntx;
  z0 = d.zbase + tz * TZ;
  x0 = (tile - tz * d.ntx) * TX;
}

// ---------------------------------------------------------------------------
// forward step
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void ac_fwd_step(
    AcDev d, const float *__restrict__ in, float *__restrict__ out, int t,
    float *__restrict__ rec_p, float *__restrict__ rec_u, float *__restrict__ rec_w,
    float *__restrict__ energy, int acc_u, int acc_w,
    float *__restrict__ img_d, float *__restrict__ img_p) {
  __shared__ float s_u[UH][UW];
  __shared__ float s_w[WH][WW];
  __shared__ float s_p[EH][EW];
  __shared__ float s_k1[EH][EW];
  __shared__ float s_a1[EH][EW];
  __shared__ float s_tu[TZ][TX];
  __shared__ float s_tw[TZ][TX];

  const int tile = blockIdx.x;
  int z0, x0;
  tile_origin(d, tile, z0, x0);
  const int s_begin = blockIdx.y * d.spb;
  const int s_end = min(d.S, s_begin + d.spb);
  const int tid = threadIdx.x;
  const int lx = tid & (TX - 1), lz = tid / TX;
  const int pitch = d.pitch;
  const float *k1g = d.coef;
  const float *a1g = k1g + d.plane;
  const float *k2g = a1g + d.plane;
  const float *a2g = k2g + d.plane;
  const float *k3g = a2g + d.plane;

  // shot-independent coefficients: ext-tile 1-K1 and A1 in LDS, tile K2/A2/K3 in registers
  for (int k = tid; k < EH * EW; k += NT) {
    const int r = k / EW, c = k - r * EW;
    const int i = z0 - 1 + r, j = x0 - 1 + c;
    const bool ok = ok_p(d, i, j);
    s_k1[r][c] = ok ? k1g[(long long)i * pitch + j] : 1.0f;
    s_a1[r][c] = ok ? a1g[(long long)i * pitch + j] : 0.0f;
  }
  float ck2[PTS], ca2[PTS], ck3[PTS];
  float acc_pp[PTS], acc_uu[PTS], acc_ww[PTS];
#pragma unroll
  for (int q = 0; q < PTS; ++q) {
    const int i = z0 + lz + q * (NT / TX), j = x0 + lx;
    const bool ok = ok_p(d, i, j);
    const long long o = (long long)i * pitch + j;
    ck2[q] = ok ? k2g[o] : 1.0f;
    ca2[q] = ok ? a2g[o] : 0.0f;
    ck3[q] = ok ? k3g[o] : 1.0f;
    acc_pp[q] = acc_uu[q] = acc_ww[q] = 0.0f;
  }

  for (int s = s_begin; s < s_end; ++s) {
    const long long so = (long long)s * d.plane;
    const float *P = in + so;
    const float *U = in + d.field + so;
    const float *W = in + 2 * d.field + so;
    __syncthreads();  // previous shot finished with LDS
    for (int k = tid; k < UH * UW; k += NT) {
      const int r = k / UW, c = k - r * UW;
      const int i = z0 - 1 + r, j = x0 - 3 + c;
      s_u[r][c] = ok_u(d, i, j) ? U[(long long)i * pitch + j] : 0.0f;
    }
    for (int k = tid; k < WH * WW; k += NT) {
      const int r = k / WW, c = k - r * WW;
      const int i = z0 - 3 + r, j = x0 - 1 + c;
      s_w[r][c] = ok_w(d, i, j) ? W[(long long)i * pitch + j] : 0.0f;
    }
    for (int k = tid; k < EH * EW; k += NT) {
      const int r = k / EW, c = k - r * EW;
      const int i = z0 - 1 + r, j = x0 - 1 + c;
      s_p[r][c] = ok_p(d, i, j) ? P[(long long)i * pitch + j] : 0.0f;
    }
    __syncthreads();

    // new pressure on the extended tile (stage 1 + source)
    const int zs = d.src_z[s], xs = d.src_x[s];
    const float sv = d.src[(long long)s * d.nt + t];
    for (int k = tid; k < EH * EW; k += NT) {
      const int r = k / EW, c = k - r * EW;
      const int i = z0 - 1 + r, j = x0 - 1 + c;
      float pv = s_p[r][c];
      float dd = 0.0f;
      if (win_p(d, i, j)) {
        // u(i, j-2..j+1) -> s_u[r][c .. c+3]; w(i-2..i+1, j) -> s_w[r .. r+3][c]
        dd = C1 * (((s_u[r][c + 2] - s_u[r][c + 1]) + s_w[r + 2][c]) - s_w[r + 1][c]) +
             C2 * (((s_u[r][c + 3] - s_u[r][c]) + s_w[r + 3][c]) - s_w[r][c]);
        pv = s_k1[r][c] * pv - s_a1[r][c] * dd;
      }
      if (i == zs && j == xs) pv = pv + sv;
      s_p[r][c] = pv;
      if (img_d && r >= 1 && r <= TZ && c >= 1 && c <= TX && ok_p(d, i, j))
        img_d[so + (long long)i * pitch + j] = dd;
    }
    __syncthreads();
    // stage 3: free-surface mirror (only the first tile row block holds rows fs-1..fs+1)
    if (d.fsurf) {
      const int rf = d.fs - 1 - (z0 - 1);
      if (rf >= 0 && rf + 2 < EH) {
        for (int c = tid; c < EW; c += NT) s_p[rf][c] = -s_p[rf + 2][c];
      }
    }
    __syncthreads();

    // stages 4-6 on the tile
#pragma unroll
    for (int q = 0; q < PTS; ++q) {
      const int tr = lz + q * (NT / TX);
      const int r = tr + 1, c = lx + 1;
      const int i = z0 + tr, j = x0 + lx;
      float un = 0.0f, wn = 0.0f;
      if (ok_p(d, i, j)) {
        const float pn = s_p[r][c];
        un = s_u[r][c + 2];
        if (win_u(d, i, j))
          un = ck2[q] * un - ca2[q] * (C1 * (s_p[r][c + 1] - s_p[r][c]) + C2 * (s_p[r][c + 2] - s_p[r][c - 1]));
        wn = s_w[r + 2][c];
        if (win_w(d, i, j))
          wn = ck3[q] * wn - ca2[q] * (C1 * (s_p[r + 1][c] - s_p[r][c]) + C2 * (s_p[r + 2][c] - s_p[r - 1][c]));
        if (d.fsurf && i == d.fs - 1) {
          // w[fs-1] <- new w[fs]: recompute w at row fs from the same LDS data
          float w2 = s_w[r + 3][c];
          if (win_w(d, d.fs, j)) {
            const long long o2 = (long long)d.fs * pitch + j;
            w2 = k3g[o2] * w2 - a2g[o2] * (C1 * (s_p[r + 2][c] - s_p[r + 1][c]) + C2 * (s_p[r + 3][c] - s_p[r][c]));
          }
          wn = w2;
        }
        const long long o = so + (long long)i * pitch + j;
        out[o] = pn;
        if (ok_u(d, i, j)) out[d.field + o] = un; else un = 0.0f;
        if (ok_w(d, i, j)) out[2 * d.field + o] = wn; else wn = 0.0f;
        if (img_p) img_p[o] = pn;
        acc_pp[q] += pn * pn;
        if (acc_u) acc_uu[q] += un * un;
        if (acc_w) acc_ww[q] += wn * wn;
      }
      s_tu[tr][lx] = un;
      s_tw[tr][lx] = wn;
    }
    if (rec_p) {
      __syncthreads();
      const int b0 = d.bin_start[tile], b1 = d.bin_start[tile + 1];
      for (int k = b0 + tid; k < b1; k += NT) {
        const int rr = d.bin_list[k];
        const int zi = d.rcv_z[rr] - z0, xi = d.rcv_x[rr] - x0;
        const long long o = ((long long)s * d.nt + t) * d.nrec + rr;
        rec_p[o] = s_p[zi + 1][xi + 1];
        rec_u[o] = s_tu[zi][xi];
        rec_w[o] = s_tw[zi][xi];
      }
    }
  }

  if (energy) {
    const long long go = (long long)blockIdx.y * 3 * d.plane;
#pragma unroll
    for (int q = 0; q < PTS; ++q) {
      const int i = z0 + lz + q * (NT / TX), j = x0 + lx;
      if (!ok_p(d, i, j)) continue;
      const long long o = go + (long long)i * pitch + j;
      energy[o] += acc_pp[q];
      if (acc_u) energy[d.plane + o] += acc_uu[q];
      if (acc_w) energy[2 * d.plane + o] += acc_ww[q];
    }
  }
}

// ---------------------------------------------------------------------------
// adjoint step: transpose of stages 6..1 of step t, plus zero-lag imaging
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void ac_adj_step(
    AcDev d, const float *__restrict__ lin, float *__restrict__ lout, int t,
    const float *__restrict__ img_d, const float *__restrict__ img_p,
    const float *__restrict__ g_p, const float *__restrict__ g_u, const float *__restrict__ g_w,
    float *__restrict__ grad_a1, float *__restrict__ grad_a2, float *__restrict__ grad_src) {
  __shared__ float s_fu[UH][UW];   // A2 * masked lambda_u (u halo)
  __shared__ float s_fw[WH][WW];   // A2 * masked lambda_w' (w halo)
  __shared__ float s_l[EH][EW];    // lambda_p3 -> lambda_p2 -> A1 * mu_p
  __shared__ float s_a2u[UH][UW];  // A2 on the u halo, zero outside the u window
  __shared__ float s_a2w[WH][WW];  // A2 on the w halo, zero outside the w window
  __shared__ float s_a1[EH][EW];   // A1 on the ext tile, zero outside the p window
  __shared__ float s_pp[EH][EW];   // pressure image (for dL/dA2)
  __shared__ float s_tp[TZ][TX], s_tu[TZ][TX], s_tw[TZ][TX];

  const int tile = blockIdx.x;
  int z0, x0;
  tile_origin(d, tile, z0, x0);
  const int s_begin = blockIdx.y * d.spb;
  const int s_end = min(d.S, s_begin + d.spb);
  const int tid = threadIdx.x;
  const int lx = tid & (TX - 1), lz = tid / TX;
  const int pitch = d.pitch;
  const float *k1g = d.coef;
  const float *a1g = k1g + d.plane;
  const float *k2g = a1g + d.plane;
  const float *a2g = k2g + d.plane;
  const float *k3g = a2g + d.plane;
  const bool want_a2 = grad_a2 != nullptr && img_p != nullptr;

  for (int k = tid; k < UH * UW; k += NT) {
    const int r = k / UW, c = k - r * UW;
    const int i = z0 - 1 + r, j = x0 - 3 + c;
    s_a2u[r][c] = (ok_u(d, i, j) && win_u(d, i, j)) ? a2g[(long long)i * pitch + j] : 0.0f;
  }
  for (int k = tid; k < WH * WW; k += NT) {
    const int r = k / WW, c = k - r * WW;
    const int i = z0 - 3 + r, j = x0 - 1 + c;
    s_a2w[r][c] = (ok_w(d, i, j) && win_w(d, i, j)) ? a2g[(long long)i * pitch + j] : 0.0f;
  }
  for (int k = tid; k < EH * EW; k += NT) {
    const int r = k / EW, c = k - r * EW;
    const int i = z0 - 1 + r, j = x0 - 1 + c;
    s_a1[r][c] = (ok_p(d, i, j) && win_p(d, i, j)) ? a1g[(long long)i * pitch + j] : 0.0f;
  }
  float ck1[PTS], ck2[PTS], ck3[PTS], acc1[PTS], acc2[PTS];
#pragma unroll
  for (int q = 0; q < PTS; ++q) {
    const int i = z0 + lz + q * (NT / TX), j = x0 + lx;
    const bool ok = ok_p(d, i, j);
    const long long o = (long long)i * pitch + j;
    ck1[q] = (ok && win_p(d, i, j)) ? k1g[o] : 1.0f;
    ck2[q] = (ok && win_u(d, i, j)) ? k2g[o] : 1.0f;
    ck3[q] = (ok && win_w(d, i, j)) ? k3g[o] : 1.0f;
    acc1[q] = acc2[q] = 0.0f;
  }
  const int fsr = d.fsurf ? d.fs : -100;  // rows fs-1 / fs get the free-surface transpose

  for (int s = s_begin; s < s_end; ++s) {
    const long long so = (long long)s * d.plane;
    const float *LP = lin + so;
    const float *LU = lin + d.field + so;
    const float *LW = lin + 2 * d.field + so;
    __syncthreads();
    for (int k = tid; k < UH * UW; k += NT) {
      const int r = k / UW, c = k - r * UW;
      const int i = z0 - 1 + r, j = x0 - 3 + c;
      const float a = s_a2u[r][c];
      s_fu[r][c] = (a != 0.0f) ? a * LU[(long long)i * pitch + j] : 0.0f;
    }
    for (int k = tid; k < WH * WW; k += NT) {
      const int r = k / WW, c = k - r * WW;
      const int i = z0 - 3 + r, j = x0 - 1 + c;
      const float a = s_a2w[r][c];
      float v = 0.0f;
      if (a != 0.0f) {
        // stage-6 transpose: lambda_w'[fs] = lambda_w[fs] + lambda_w[fs-1]; row fs-1 is not in the window
        v = LW[(long long)i * pitch + j];
        if (i == fsr) v = v + LW[(long long)(i - 1) * pitch + j];
        v = a * v;
      }
      s_fw[r][c] = v;
    }
    for (int k = tid; k < EH * EW; k += NT) {
      const int r = k / EW, c = k - r * EW;
      const int i = z0 - 1 + r, j = x0 - 1 + c;
      s_l[r][c] = ok_p(d, i, j) ? LP[(long long)i * pitch + j] : 0.0f;
      if (want_a2) s_pp[r][c] = ok_p(d, i, j) ? img_p[so + (long long)i * pitch + j] : 0.0f;
    }
    __syncthreads();
    // stages 5/4 transposed into lambda_p3 on the ext tile
    for (int k = tid; k < EH * EW; k += NT) {
      const int r = k / EW, c = k - r * EW;
      const float dxt = C1 * (s_fu[r][c + 1] - s_fu[r][c + 2]) + C2 * (s_fu[r][c] - s_fu[r][c + 3]);
      const float dzt = C1 * (s_fw[r + 1][c] - s_fw[r + 2][c]) + C2 * (s_fw[r][c] - s_fw[r + 3][c]);
      s_l[r][c] = s_l[r][c] - (dxt + dzt);
    }
    __syncthreads();
    // stage 3 transposed: lambda_p2[fs+1] -= lambda_p3[fs-1]; lambda_p2[fs-1] = 0
    if (d.fsurf) {
      const int rf = d.fs - 1 - (z0 - 1);
      if (rf >= 0 && rf + 2 < EH) {
        for (int c = tid; c < EW; c += NT) {
          const float m = s_l[rf][c];
          s_l[rf + 2][c] = s_l[rf + 2][c] - m;
          s_l[rf][c] = 0.0f;
        }
      }
    }
    __syncthreads();
    // stage 2 transposed: d loss / d src[s, t] = lambda_p2 at the source
    const int zs = d.src_z[s], xs = d.src_x[s];
    if (grad_src && tid == 0 && zs >= z0 && zs < z0 + TZ && xs >= x0 && xs < x0 + TX)
      grad_src[(long long)s * d.nt + t] = s_l[zs - z0 + 1][xs - x0 + 1];
    float lp1[PTS];
#pragma unroll
    for (int q = 0; q < PTS; ++q) lp1[q] = s_l[lz + q * (NT / TX) + 1][lx + 1];
    __syncthreads();
    for (int k = tid; k < EH * EW; k += NT) {
      const int r = k / EW, c = k - r * EW;
      s_l[r][c] = s_a1[r][c] * s_l[r][c];
    }
    __syncthreads();
    // stage 1 transposed + outputs on the tile
#pragma unroll
    for (int q = 0; q < PTS; ++q) {
      const int tr = lz + q * (NT / TX);
      const int r = tr + 1, c = lx + 1;
      const int i = z0 + tr, j = x0 + lx;
      float op = 0.0f, ou = 0.0f, ow = 0.0f;
      if (ok_p(d, i, j)) {
        const long long o = (long long)i * pitch + j;
        const bool wp = win_p(d, i, j);
        op = wp ? ck1[q] * lp1[q] : lp1[q];
        if (ok_u(d, i, j)) {
          const float lu = LU[o];
          ou = ck2[q] * lu - (C1 * (s_l[r][c] - s_l[r][c + 1]) + C2 * (s_l[r][c - 1] - s_l[r][c + 2]));
          if (want_a2 && win_u(d, i, j))
            acc2[q] -= lu * (C1 * (s_pp[r][c + 1] - s_pp[r][c]) + C2 * (s_pp[r][c + 2] - s_pp[r][c - 1]));
        }
        if (ok_w(d, i, j)) {
          float lw = LW[o];
          if (i == fsr) lw = lw + LW[o - pitch];
          if (i == fsr - 1) lw = 0.0f;
          ow = ck3[q] * lw - (C1 * (s_l[r][c] - s_l[r + 1][c]) + C2 * (s_l[r - 1][c] - s_l[r + 2][c]));
          if (want_a2 && win_w(d, i, j))
            acc2[q] -= lw * (C1 * (s_pp[r + 1][c] - s_pp[r][c]) + C2 * (s_pp[r + 2][c] - s_pp[r - 1][c]));
        }
        if (grad_a1 && wp) acc1[q] -= lp1[q] * img_d[so + o];
      }
      s_tp[tr][lx] = op;
      s_tu[tr][lx] = ou;
      s_tw[tr][lx] = ow;
    }
    // receiver adjoint sources of step t-1 join the state after step t-1
    if (t >= 1 && (g_p || g_u || g_w)) {
      __syncthreads();
      const int b0 = d.bin_start[tile], b1 = d.bin_start[tile + 1];
      for (int k = b0 + tid; k < b1; k += NT) {
        const int rr = d.bin_list[k];
        const int zi = d.rcv_z[rr] - z0, xi = d.rcv_x[rr] - x0;
        const long long o = ((long long)s * d.nt + (t - 1)) * d.nrec + rr;
        if (g_p) atomicAdd(&s_tp[zi][xi], g_p[o]);
        if (g_u) atomicAdd(&s_tu[zi][xi], g_u[o]);
        if (g_w) atomicAdd(&s_tw[zi][xi], g_w[o]);
      }
      __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < PTS; ++q) {
      const int tr = lz + q * (NT / TX);
      const int i = z0 + tr, j = x0 + lx;
      if (!ok_p(d, i, j)) continue;
      const long long o = so + (long long)i * pitch + j;
      lout[o] = s_tp[tr][lx];
      if (ok_u(d, i, j)) lout[d.field + o] = s_tu[tr][lx];
      if (ok_w(d, i, j)) lout[2 * d.field + o] = s_tw[tr][lx];
    }
  }

  const long long go = (long long)blockIdx.y * d.plane;
#pragma unroll
  for (int q = 0; q < PTS; ++q) {
    const int i = z0 + lz + q * (NT / TX), j = x0 + lx;
    if (!ok_p(d, i, j)) continue;
    const long long o = go + (long long)i * pitch + j;
    if (grad_a1) grad_a1[o] += acc1[q];
    if (want_a2) grad_a2[o] += acc2[q];
  }
}

// lam[field][s][zr][xr] += g[s][t][r] for the three record components
__global__ void ac_seed(AcDev d, float *__restrict__ lam, const float *__restrict__ g_p,
                        const float *__restrict__ g_u, const float *__restrict__ g_w, int t) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long long)d.S * d.nrec) return;
  const int s = (int)(k / d.nrec), r = (int)(k - (long long)s * d.nrec);
  const long long o = (long long)s * d.plane + (long long)d.rcv_z[r] * d.pitch + d.rcv_x[r];
  const long long gi = ((long long)s * d.nt + t) * d.nrec + r;
  if (g_p) atomicAdd(&lam[o], g_p[gi]);
  if (g_u) atomicAdd(&lam[d.field + o], g_u[gi]);
  if (g_w) atomicAdd(&lam[2 * d.field + o], g_w[gi]);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
int layout_of(const adfwi_acoustic_desc *d, adfwi_layout *L) {
  if (!d) { adfwi::set_error("null descriptor"); return ADFWI_E_ARG; }
  if (d->abi_version != ADFWI_ABI_VERSION || d->struct_size != (int32_t)sizeof(adfwi_acoustic_desc)) {
    adfwi::set_error("acoustic descriptor version %d size %d, library expects %d / %d", d->abi_version,
                     d->struct_size, ADFWI_ABI_VERSION, (int)sizeof(adfwi_acoustic_desc));
    return ADFWI_E_VERSION;
  }
  ADFWI_CHECK_ARG(d->nshots >= 1, "nshots must be >= 1 (got %d)", d->nshots);
  ADFWI_CHECK_ARG(d->nz >= 8 && d->nx >= 8, "padded grid %d x %d too small (need >= 8 x 8)", d->nz, d->nx);
  ADFWI_CHECK_ARG(d->pitch >= d->nx, "pitch %d < nx %d", d->pitch, d->nx);
  ADFWI_CHECK_ARG(d->nt >= 1 && d->nrec >= 0, "nt %d / nrec %d invalid", d->nt, d->nrec);
  if (d->free_surface) {
    ADFWI_CHECK_ARG(d->fs >= 1 && d->fs + 3 <= d->nz - 3, "free-surface row %d outside grid of %d rows",
                    d->fs, d->nz);
  } else {
    ADFWI_CHECK_ARG(d->fs == 1, "fs must be 1 without a free surface (got %d)", d->fs);
  }
  L->tile_z = TZ;
  L->tile_x = TX;
  L->zbase = d->free_surface ? d->fs - 1 : 0;
  L->ntiles_z = (d->nz - L->zbase + TZ - 1) / TZ;
  L->ntiles_x = (d->nx + TX - 1) / TX;
  L->ntiles = L->ntiles_z * L->ntiles_x;
  int spb = d->shots_per_block;
  if (spb <= 0) {
    // enough workgroups to fill 256 CUs several times over, as few shot groups as that allows
    const long long want = 4096;
    long long groups = (want + L->ntiles - 1) / L->ntiles;
    if (groups > d->nshots) groups = d->nshots;
    if (groups < 1) groups = 1;
    spb = (int)((d->nshots + groups - 1) / groups);
  }
  if (spb > d->nshots) spb = d->nshots;
  L->shots_per_block = spb;
  L->ngroups = (d->nshots + spb - 1) / spb;
  L->nfields = 3;
  L->reserved = 0;
  L->field_floats = (int64_t)d->nshots * d->nz * d->pitch;
  L->state_floats = 3 * L->field_floats;
  L->accum_floats = (int64_t)L->ngroups * d->nz * d->pitch;
  return ADFWI_OK;
}

int check_device_desc(const adfwi_acoustic_desc *d, bool need_records) {
  ADFWI_CHECK_ARG(d->coef && d->src && d->src_z && d->src_x, "null coefficient/source pointer");
  if (need_records && d->nrec > 0)
    ADFWI_CHECK_ARG(d->rcv_z && d->rcv_x && d->rcv_bin_start && d->rcv_bin_list,
                    "null receiver table pointer");
  return ADFWI_OK;
}

AcDev make_dev(const adfwi_acoustic_desc *d, const adfwi_layout &L) {
  AcDev v;
  v.S = d->nshots; v.Z = d->nz; v.X = d->nx; v.pitch = d->pitch;
  v.fs = d->fs; v.fsurf = d->free_surface ? 1 : 0; v.nt = d->nt; v.nrec = d->nrec;
  v.zbase = L.zbase; v.ntx = L.ntiles_x; v.spb = L.shots_per_block;
  v.plane = (long long)d->nz * d->pitch;
  v.field = (long long)d->nshots * v.plane;
  v.coef = d->coef; v.src = d->src; v.src_z = d->src_z; v.src_x = d->src_x;
  v.rcv_z = d->rcv_z; v.rcv_x = d->rcv_x; v.bin_start = d->rcv_bin_start; v.bin_list = d->rcv_bin_list;
  return v;
}

}  // namespace

extern "C" {

int adfwi_acoustic_get_layout(const adfwi_acoustic_desc *d, adfwi_layout *out) {
  ADFWI_CHECK_ARG(out != nullptr, "null layout pointer");
  return layout_of(d, out);
}

int adfwi_acoustic_bin_receivers(const adfwi_acoustic_desc *d, const int32_t *rz, const int32_t *rx,
                                 int32_t *start, int32_t *list) {
  adfwi_layout L;
  int rc = layout_of(d, &L);
  if (rc) return rc;
  ADFWI_CHECK_ARG(start && (d->nrec == 0 || (rz && rx && list)), "null receiver binning array");
  for (int k = 0; k <= L.ntiles; ++k) start[k] = 0;
  for (int r = 0; r < d->nrec; ++r) {
    const int z = rz[r], x = rx[r];
    ADFWI_CHECK_ARG(z >= 0 && z < d->nz - 1 && x >= 0 && x < d->nx - 1,
                    "receiver %d at padded (%d, %d) outside the grid %d x %d", r, z, x, d->nz, d->nx);
    if (z < L.zbase) continue;  // rows above the free surface never change: records stay zero
    const int tile = ((z - L.zbase) / TZ) * L.ntiles_x + x / TX;
    start[tile + 1]++;
  }
  for (int k = 0; k < L.ntiles; ++k) start[k + 1] += start[k];
  // stable fill
  int *fill = new int[L.ntiles];
  for (int k = 0; k < L.ntiles; ++k) fill[k] = start[k];
  for (int r = 0; r < d->nrec; ++r) {
    const int z = rz[r], x = rx[r];
    if (z < L.zbase) continue;
    const int tile = ((z - L.zbase) / TZ) * L.ntiles_x + x / TX;
    list[fill[tile]++] = r;
  }
  delete[] fill;
  return ADFWI_OK;
}

int adfwi_acoustic_forward(const adfwi_acoustic_desc *d, adfwi_acoustic_fwd_args *a, adfwi_stream_t stream) {
  if (int rc = adfwi::fault_injected("adfwi_acoustic_forward")) return rc;
  adfwi_layout L;
  int rc = layout_of(d, &L);
  if (rc) return rc;
  ADFWI_CHECK_ARG(a != nullptr, "null forward args");
  const bool recs = a->rec_p || a->rec_u || a->rec_w;
  ADFWI_CHECK_ARG(!recs || (a->rec_p && a->rec_u && a->rec_w), "rec_p/rec_u/rec_w must be all set or all NULL");
  if ((rc = check_device_desc(d, recs))) return rc;
  ADFWI_CHECK_ARG(a->state_in && a->work[0] && a->work[1], "null state/work buffer");
  ADFWI_CHECK_ARG(a->work[0] != a->work[1], "work[0] and work[1] must be distinct buffers");
  ADFWI_CHECK_ARG(a->t0 >= 0 && a->nsteps >= 0 && a->t0 + a->nsteps <= d->nt,
                  "steps [%d, %d) outside [0, %d)", a->t0, a->t0 + a->nsteps, d->nt);
  ADFWI_CHECK_ARG(!a->ckpt || a->ckpt_every >= 1, "ckpt_every must be >= 1 with a checkpoint buffer");
  AcDev v = make_dev(d, L);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const dim3 grid(L.ntiles, L.ngroups), block(NT);
  const float *cur = a->state_in;
  for (int k = 0; k < a->nsteps; ++k) {
    const int t = a->t0 + k;
    float *nxt;
    if (a->ckpt && (t + 1) % a->ckpt_every == 0 && t + 1 < d->nt) {
      nxt = a->ckpt + (long long)((t + 1) / a->ckpt_every - 1) * L.state_floats;
    } else {
      nxt = (cur == a->work[0]) ? a->work[1] : a->work[0];
    }
    float *imd = a->image ? a->image + (long long)k * L.field_floats : nullptr;
    float *imp = a->image_p ? a->image_p + (long long)k * L.field_floats : nullptr;
    const int acc_u = (a->energy && t >= a->energy_u_from) ? 1 : 0;
    const int acc_w = (a->energy && t == d->nt - 1) ? 1 : 0;
    hipLaunchKernelGGL(ac_fwd_step, grid, block, 0, st, v, cur, nxt, t, a->rec_p, a->rec_u, a->rec_w,
                       a->energy, acc_u, acc_w, imd, imp);
    if ((rc = adfwi::check_launch("ac_fwd_step"))) return rc;
    cur = nxt;
  }
  a->state_out = cur;
  return ADFWI_OK;
}

int adfwi_acoustic_adjoint(const adfwi_acoustic_desc *d, adfwi_acoustic_adj_args *a, adfwi_stream_t stream) {
  if (int rc = adfwi::fault_injected("adfwi_acoustic_adjoint")) return rc;
  adfwi_layout L;
  int rc = layout_of(d, &L);
  if (rc) return rc;
  ADFWI_CHECK_ARG(a != nullptr, "null adjoint args");
  const bool recs = a->grad_rec_p || a->grad_rec_u || a->grad_rec_w;
  if ((rc = check_device_desc(d, recs))) return rc;
  ADFWI_CHECK_ARG(a->lam_in && a->work[0] && a->work[1], "null adjoint state/work buffer");
  ADFWI_CHECK_ARG(a->t_hi < d->nt && a->nsteps >= 0 && a->t_hi - a->nsteps + 1 >= 0,
                  "adjoint steps %d..%d outside [0, %d)", a->t_hi - a->nsteps + 1, a->t_hi, d->nt);
  ADFWI_CHECK_ARG(!a->grad_a1 || a->image, "grad_a1 needs the stencil image");
  ADFWI_CHECK_ARG(!a->grad_a2 || a->image_p, "grad_a2 needs the pressure image");
  ADFWI_CHECK_ARG(a->image_t0 <= a->t_hi - a->nsteps + 1 || !(a->image || a->image_p),
                  "image_t0 %d after the first adjoint step %d", a->image_t0, a->t_hi - a->nsteps + 1);
  AcDev v = make_dev(d, L);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const dim3 grid(L.ntiles, L.ngroups), block(NT);
  const float *cur = a->lam_in;
  for (int k = 0; k < a->nsteps; ++k) {
    const int t = a->t_hi - k;
    float *nxt = (cur == a->work[0]) ? a->work[1] : a->work[0];
    const long long slot = (long long)(t - a->image_t0) * L.field_floats;
    const float *imd = a->image ? a->image + slot : nullptr;
    const float *imp = a->image_p ? a->image_p + slot : nullptr;
    hipLaunchKernelGGL(ac_adj_step, grid, block, 0, st, v, cur, nxt, t, imd, imp, a->grad_rec_p,
                       a->grad_rec_u, a->grad_rec_w, a->grad_a1, a->grad_a2, a->grad_src);
    if ((rc = adfwi::check_launch("ac_adj_step"))) return rc;
    cur = nxt;
  }
  a->lam_out = cur;
  return ADFWI_OK;
}

int adfwi_acoustic_adjoint_seed(const adfwi_acoustic_desc *d, float *lam, const float *grad_rec_p,
                                const float *grad_rec_u, const float *grad_rec_w, int32_t t,
                                adfwi_stream_t stream) {
  adfwi_layout L;
  int rc = layout_of(d, &L);
  if (rc) return rc;
  ADFWI_CHECK_ARG(lam != nullptr, "null adjoint state");
  ADFWI_CHECK_ARG(t >= 0 && t < d->nt, "seed step %d outside [0, %d)", t, d->nt);
  if (d->nrec == 0 || !(grad_rec_p || grad_rec_u || grad_rec_w)) return ADFWI_OK;
  ADFWI_CHECK_ARG(d->rcv_z && d->rcv_x, "null receiver table pointer");
  AcDev v = make_dev(d, L);
  const long long n = (long long)d->nshots * d->nrec;
  hipLaunchKernelGGL(ac_seed, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), v, lam, grad_rec_p, grad_rec_u, grad_rec_w, t);
  return adfwi::check_launch("ac_seed");
}

}  // extern "C"
