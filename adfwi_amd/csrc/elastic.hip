// Elastic (isotropic / VTI / general 2-D anisotropic) velocity-stress propagator for gfx950:
// fused forward step with split-field PML, receiver gather, and the C ABI.
//
// One forward step (reference: ADFWI/propagator/elastic_kernels.py:339-418, 4th order, PML),
// on the window rows [2, Z-3] x cols [2, X-3] of the padded grid:
//   1. stress splits  s_x <- (pxn s_x + dt/dx (C.Dx v)) * (1/pxd),  s_z <- (pzn s_z + dt/dz (C.Dz v)) * (1/pzd)
//      for txx, tzz, txz with the D operators of the reference (Dxb/Dzb for the normal
//      stresses, Dxf/Dzf for the shear stress) applied to the velocity totals;
//   2. moment-tensor source: -MT/2 * w(t) into both splits of each stress component;
//   3. stress totals = split sums; free surface (rows 0..2): tzz[2]=0, tzz[1]=-tzz[3],
//      txz[1]=-txz[2], txz[0]=-txz[3];
//   4. velocity splits v_x <- (pxn v_x + dt b Dx(s) / dx) / pxd, v_z likewise with pzn/pzd;
//   5. velocity totals; free surface (rows 0..1, cols [2, X-3]): vz[1]=vz[2],
//      vx[1]=vz[1,j+1]-vz[1,j]+vz[2,j+1]-vz[2,j]+vx[3,j], vz[0]=vz[1];
//   6. records of the five totals.
// Every expression keeps the reference's fp32 association (compiled with
// -ffp-contract=off, IEEE division), so a step is bit-identical to the reference's CPU
// op sequence given identical coefficients.  C15/C35 terms are skipped for isotropic and
// VTI media (exact: they multiply zeros).
//
// State: the ten split fields [10][S][Z][pitch] in the order txx_x, txx_z, tzz_x, tzz_z,
// txz_x, txz_z, vx_x, vx_z, vz_x, vz_z.  The splits are identically zero outside the window
// (never updated there); totals and free-surface rows are derived when needed, so the
// reference's five total arrays need no storage.
//
// Work decomposition: one wavefront owns 64 columns x0-4 .. x0+59 of a TZ-row strip and
// writes the 55 columns x0 .. x0+54: stress is valid on lanes 2..60 (it reads velocities
// 2 columns away, plus one more column for the free-surface row), velocity on lanes 4..58.
// Each lane marches down its column: at iteration k it loads the velocity splits of row
// k+2, updates the stress of row k and the velocity of row k-2, keeping 5-row windows of
// velocity and stress totals in registers.  The four wavefronts of a workgroup advance
// four different shots of the same strip (their gradient images share one LDS reduction).
#include "common.hpp"

namespace {

constexpr int WAVE = 64;
constexpr int EHL = 4;              // lanes left of the first output column
constexpr int ESX = 55;             // output columns per strip
constexpr int EWPB = 4;             // wavefronts per workgroup
constexpr int ENT = WAVE * EWPB;
constexpr int ETZ_MAX = 64;

typedef __amdgpu_buffer_rsrc_t rsrc_t;
constexpr int RSRC_DW3 = 0x00020000;  // raw 32-bit data, gfx9 family
__device__ __forceinline__ rsrc_t rsrc(const void *p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, RSRC_DW3);
}
// the b32 builtins move raw 32-bit words: bit-cast, never value-convert
__device__ __forceinline__ float bld(rsrc_t r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ void bst(float v, rsrc_t r, unsigned voff, unsigned soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
}
__device__ __forceinline__ float shfl(float v, int lane) { return __shfl(v, lane, WAVE); }

enum { TXX_X = 0, TXX_Z, TZZ_X, TZZ_Z, TXZ_X, TXZ_Z, VX_X, VX_Z, VZ_X, VZ_Z, NFIELD };
enum { TXX = 0, TZZ, TXZ, VX, VZ, NFIELD5 };   // state of the damping (non-PML) scheme
enum { C11 = 0, C13, C33, C55, BX, BZ, C15, C35, NCOEF };

struct ElDev {
  int S, Z, X, pitch, fsurf, nt, nrec, spw;     // spw: shots per wavefront (runtime loop)
  int tz, nsx, nstrips;
  float dt, dt_dx, dt_dz, dx, dz, c0, c1, c2;
  int nn;                                      // stencil half-order: 2, or 3..5 (three-term scheme)
  long long plane, field;
  const float *coef, *pml_x, *pml_z, *src, *src_m, *damp;
  const float *src_fz;   // [S] vertical-force weight of each shot's wavelet, or NULL
  const int *src_z, *src_x, *rcv_z, *rcv_x;
};

// Receivers handled inside the 4th-order PML step kernels: map [Z + Z*pitch] holds the receivers
// of each row, then the receiver index of each point (-1: none; every receiver at its own point).
struct ElRec {          // forward: records of txx, tzz, txz, vx, vz [S][nt][R]
  const int *map;
  float *r[5];
};
struct ElInj {          // adjoint: record adjoints [S][nt][R] (NULL: zero)
  const int *map;
  const float *g[5];
};
static_assert(ETZ_MAX + 8 <= 128, "a strip's rows (with the adjoint's halo) fit one receiver row mask");

__device__ __forceinline__ bool strip_of(const ElDev &d, int &z0, int &x0) {
  const int n = (int)gridDim.x, b = (int)blockIdx.x;
  const int xcd = b & 7, per = n >> 3, rem = n & 7;
  const int strip = xcd * per + min(xcd, rem) + (b >> 3);   // XCD-contiguous strips
  if (strip >= d.nstrips) return false;
  const int sz = strip / d.nsx;
  z0 = 2 + sz * d.tz;                       // strips tile the window rows [2, Z-3]
  x0 = 2 + (strip - sz * d.nsx) * ESX;      // and the window columns [2, X-3]
  return true;
}

// ---------------------------------------------------------------------------
// forward step
// ---------------------------------------------------------------------------
template <bool ANISO, bool REC>
__global__ __launch_bounds__(ENT) void el_fwd_step(ElDev d, const float *__restrict__ in, float *__restrict__ out,
                                                   int t, ElRec rc) {
  int z0, x0;
  if (!strip_of(d, z0, x0)) return;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
  const int lane = threadIdx.x & (WAVE - 1);
  const int c = x0 - EHL + lane;
  const unsigned vc = 4u * (unsigned)min(max(c, 0), d.X - 1);
  const unsigned vs = 4u * (unsigned)max(c, 0);
  const bool out_col = lane >= EHL && lane < EHL + ESX && c <= d.X - 3;   // c >= 2 by the strip origin
  const bool cwin = c >= 2 && c <= d.X - 3;
  const int zmax = d.Z - 1, zw1 = d.Z - 3;
  const unsigned rowb = 4u * (unsigned)d.pitch, planeb = 4u * (unsigned)d.plane, fieldb = 4u * (unsigned)d.field;
  auto rb = [&](int k) -> unsigned { return (unsigned)min(max(k, 0), zmax) * rowb; };
  const rsrc_t rIn = rsrc(in, (unsigned)NFIELD * fieldb);
  const rsrc_t rOut = rsrc(out, (unsigned)NFIELD * fieldb);
  const rsrc_t rC = rsrc(d.coef, (unsigned)NCOEF * planeb);
  // per-row PML factors, read as part of each row's loads (uniform-address vector loads queue
  // behind the row data; a separate read would make the wave wait for every load in flight)
  const unsigned zB = 4u * (unsigned)d.Z;
  const rsrc_t rPz = rsrc(d.pml_z, 3u * zB);
  auto zo = [&](int k) -> unsigned { return 4u * (unsigned)min(max(k, 0), zmax); };
  const adfwi::RowMask rrows = adfwi::receiver_rows(REC ? rc.map : nullptr, d.Z, z0, lane);
  const float c0 = d.c0, c1 = d.c1;
  // per-column PML factors (x); rows come from pml_z
  const float pxn = d.pml_x[min(max(c, 0), d.X - 1)];
  const float pxd = d.pml_x[d.pitch + min(max(c, 0), d.X - 1)];
  const float pxi = d.pml_x[2 * d.pitch + min(max(c, 0), d.X - 1)];
  const int end_row = min(z0 + d.tz, zw1 + 1);   // one past the last output row
  const bool jj = c >= 2 && c <= d.X - 3;          // free-surface velocity columns J
  const bool jj1 = c + 1 >= 2 && c + 1 <= d.X - 3;

  const int s_first = (blockIdx.y * EWPB + wv) * d.spw;
  for (int si = 0; si < d.spw; ++si) {
    const int s = s_first + si;
    if (s >= d.S) break;
    const unsigned so = (unsigned)s * planeb;
    auto F = [&](int f) -> unsigned { return (unsigned)f * fieldb + so; };
    const int zs = d.src_z[s], xs = d.src_x[s];
    const float wsrc = d.src[(long long)s * d.nt + t];
    const float mxx = d.src_m[3 * s], mzz = d.src_m[3 * s + 1], mxz = d.src_m[3 * s + 2];
    const float fz = d.src_fz ? d.src_fz[s] : 0.0f;
    // windows: velocity totals rows k-2..k+2; stress totals txz/tzz rows k-4..k, txx rows k-2..k
    float vx0 = 0.f, vx1 = 0.f, vx2 = 0.f, vx3 = 0.f, vx4 = 0.f;
    float vz0 = 0.f, vz1 = 0.f, vz2 = 0.f, vz3 = 0.f, vz4 = 0.f;
    float xz0 = 0.f, xz1 = 0.f, xz2 = 0.f, xz3 = 0.f, xz4 = 0.f;
    float zz0 = 0.f, zz1 = 0.f, zz2 = 0.f, zz3 = 0.f, zz4 = 0.f;
    float xx0 = 0.f, xx1 = 0.f, xx2 = 0.f;
    // prologue: velocity rows z0-4 .. z0-1 (k = z0-2 needs rows z0-4 .. z0)
    for (int rr = z0 - 4; rr <= z0 - 1; ++rr) {
      const unsigned o = rb(rr);
      const float a = bld(rIn, vc, F(VX_X) + o) + bld(rIn, vc, F(VX_Z) + o);
      const float b = bld(rIn, vc, F(VZ_X) + o) + bld(rIn, vc, F(VZ_Z) + o);
      const bool ok = rr >= 0;
      vx1 = vx2; vx2 = vx3; vx3 = vx4; vx4 = ok ? a : 0.f;
      vz1 = vz2; vz2 = vz3; vz3 = vz4; vz4 = ok ? b : 0.f;
    }
    // everything iteration k reads from memory, loaded two iterations ahead: velocity splits of
    // row k+2, stress splits and moduli of row k, buoyancies of row k-2.  The old velocity splits
    // of row k-2 come from a register window of the splits loaded four iterations before.
    struct Row {
      float v[4], s[6], cm[ANISO ? 6 : 4], b[2];
      float pz[4];   // pml_z: 1 - dt/2 d and 1/(1 + dt/2 d) of row k; 1 - dt/2 d and 1 + dt/2 d of row k-2
    };
    float us[5][4] = {};                         // velocity splits of rows k-2 .. k+2
    auto load = [&](int k, Row &R) {
      const unsigned o2 = rb(k + 2), o0 = rb(k), om = rb(k - 2);
#pragma unroll
      for (int q = 0; q < 4; ++q) R.v[q] = bld(rIn, vc, F(VX_X + q) + o2);
#pragma unroll
      for (int q = 0; q < 6; ++q) R.s[q] = bld(rIn, vc, F(TXX_X + q) + o0);
      R.cm[0] = bld(rC, vc, C11 * planeb + o0); R.cm[1] = bld(rC, vc, C13 * planeb + o0);
      R.cm[2] = bld(rC, vc, C33 * planeb + o0); R.cm[3] = bld(rC, vc, C55 * planeb + o0);
      if constexpr (ANISO) {
        R.cm[4] = bld(rC, vc, C15 * planeb + o0); R.cm[5] = bld(rC, vc, C35 * planeb + o0);
      }
      if (k - 2 >= z0 && k - 2 < end_row) {   // only rows this strip updates
        R.b[0] = bld(rC, vc, BX * planeb + om); R.b[1] = bld(rC, vc, BZ * planeb + om);
      }
      R.pz[0] = bld(rPz, 0u, zo(k)); R.pz[1] = bld(rPz, 0u, 2u * zB + zo(k));
      R.pz[2] = bld(rPz, 0u, zo(k - 2)); R.pz[3] = bld(rPz, 0u, zB + zo(k - 2));
    };
    // three row buffers in rotation: the loads of row k+2 go out before row k is computed, into
    // the buffer row k-1 has just freed, so two rows of work cover each load and no register
    // holding a load in flight is ever copied (a copy waits for every load issued before it)
    auto step = [&](const int kk, const Row &cur) {
      const int k = z0 - 2 + kk;
      // 1. velocity totals of row k+2 into the window (zero outside the grid)
      {
        const bool ok = k + 2 <= zmax;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
#pragma unroll
          for (int q = 0; q < 4; ++q) us[e][q] = us[e + 1][q];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) us[4][q] = cur.v[q];
        const float a = cur.v[0] + cur.v[1], b = cur.v[2] + cur.v[3];
        vx0 = vx1; vx1 = vx2; vx2 = vx3; vx3 = vx4; vx4 = ok ? a : 0.f;
        vz0 = vz1; vz1 = vz2; vz2 = vz3; vz3 = vz4; vz4 = ok ? b : 0.f;
      }
      // free surface of the velocities of the previous step: rows 0, 1 from rows 2, 3
      if (d.fsurf && k == 2) {
        // window rows 0..4 = vx0..vx4
        const float vz2n = shfl(vz2, lane + 1);             // vz row 2, column j+1
        const float vz1m_n = jj1 ? vz2n : 0.0f;             // modified vz row 1 at column j+1
        const float vz1m = jj ? vz2 : vz1;
        if (jj) {
          vx1 = (((vz1m_n - vz1m) + vz2n) - vz2) + vx3;
          vz0 = vz1m;
        }
        vz1 = vz1m;
      }
      // 2. stress of row k
      const unsigned ok_ = rb(k);
      float sxx_x = cur.s[0], sxx_z = cur.s[1], szz_x = cur.s[2], szz_z = cur.s[3];
      float sxz_x = cur.s[4], sxz_z = cur.s[5];
      // x derivatives of the velocity totals of row k (all lanes take part in the shuffles)
      const float vxm1 = shfl(vx2, lane - 1), vxm2 = shfl(vx2, lane - 2), vxp1 = shfl(vx2, lane + 1);
      const float vzm1 = shfl(vz2, lane - 1), vzp1 = shfl(vz2, lane + 1), vzp2 = shfl(vz2, lane + 2);
      float vzm2 = 0.0f, vxp2 = 0.0f;
      if constexpr (ANISO) {
        vzm2 = shfl(vz2, lane - 2);
        vxp2 = shfl(vx2, lane + 2);
      }
      if (k >= 2 && k <= zw1 && cwin) {
        const float dxb_vx = c0 * (vx2 - vxm1) + c1 * (vxp1 - vxm2);
        const float dzb_vz = c0 * (vz2 - vz1) + c1 * (vz3 - vz0);
        const float dxf_vz = c0 * (vzp1 - vz2) + c1 * (vzp2 - vzm1);
        const float dzf_vx = c0 * (vx3 - vx2) + c1 * (vx4 - vx1);
        const float C11v = cur.cm[0], C13v = cur.cm[1], C33v = cur.cm[2], C55v = cur.cm[3];
        const float pzn = cur.pz[0], pzi = cur.pz[1];
        float sx_xx, sz_xx, sx_zz, sz_zz, sx_xz, sz_xz;
        if constexpr (ANISO) {
          const float dxb_vz = c0 * (vz2 - vzm1) + c1 * (vzp1 - vzm2);
          const float dzb_vx = c0 * (vx2 - vx1) + c1 * (vx3 - vx0);
          const float dxf_vx = c0 * (vxp1 - vx2) + c1 * (vxp2 - vxm1);
          const float dzf_vz = c0 * (vz3 - vz2) + c1 * (vz4 - vz1);
          const float C15v = cur.cm[4], C35v = cur.cm[5];
          sx_xx = C11v * dxb_vx + C15v * dxb_vz;
          sz_xx = C15v * dzb_vx + C13v * dzb_vz;
          sx_zz = C13v * dxb_vx + C35v * dxb_vz;
          sz_zz = C35v * dzb_vx + C33v * dzb_vz;
          sx_xz = C15v * dxf_vx + C55v * dxf_vz;
          sz_xz = C55v * dzf_vx + C35v * dzf_vz;
        } else {
          sx_xx = C11v * dxb_vx;
          sz_xx = C13v * dzb_vz;
          sx_zz = C13v * dxb_vx;
          sz_zz = C33v * dzb_vz;
          sx_xz = C55v * dxf_vz;
          sz_xz = C55v * dzf_vx;
        }
        sxx_x = (pxn * sxx_x + d.dt_dx * sx_xx) * pxi;
        sxx_z = (pzn * sxx_z + d.dt_dz * sz_xx) * pzi;
        szz_x = (pxn * szz_x + d.dt_dx * sx_zz) * pxi;
        szz_z = (pzn * szz_z + d.dt_dz * sz_zz) * pzi;
        sxz_x = (pxn * sxz_x + d.dt_dx * sx_xz) * pxi;
        sxz_z = (pzn * sxz_z + d.dt_dz * sz_xz) * pzi;
      }
      if (k == zs && c == xs) {
        sxx_x = sxx_x + mxx * wsrc; sxx_z = sxx_z + mxx * wsrc;
        szz_x = szz_x + mzz * wsrc; szz_z = szz_z + mzz * wsrc;
        sxz_x = sxz_x + mxz * wsrc; sxz_z = sxz_z + mxz * wsrc;
      }
      const bool own_k = out_col && k >= z0 && k < end_row;
      if (own_k) {
        bst(sxx_x, rOut, vs, F(TXX_X) + ok_); bst(sxx_z, rOut, vs, F(TXX_Z) + ok_);
        bst(szz_x, rOut, vs, F(TZZ_X) + ok_); bst(szz_z, rOut, vs, F(TZZ_Z) + ok_);
        bst(sxz_x, rOut, vs, F(TXZ_X) + ok_); bst(sxz_z, rOut, vs, F(TXZ_Z) + ok_);
      }
      if (REC && rrows.has(k)) {
        // stress records of row k: the split sums of the state after the step (the free-surface
        // tzz of row 2 is zero)
        const int mk = own_k ? rc.map[d.Z + (long long)k * d.pitch + c] : -1;
        if (mk >= 0) {
          const long long ro = ((long long)s * d.nt + t) * d.nrec + mk;
          rc.r[TXX][ro] = sxx_x + sxx_z;
          rc.r[TZZ][ro] = (d.fsurf && k == 2) ? 0.0f : szz_x + szz_z;
          rc.r[TXZ][ro] = sxz_x + sxz_z;
        }
      }
      xx0 = xx1; xx1 = xx2; xx2 = sxx_x + sxx_z;
      xz0 = xz1; xz1 = xz2; xz2 = xz3; xz3 = xz4; xz4 = sxz_x + sxz_z;
      zz0 = zz1; zz1 = zz2; zz2 = zz3; zz3 = zz4; zz4 = szz_x + szz_z;
      // 3. free surface of the stresses once row 3 exists (window rows -1..3 = *0..*4)
      if (d.fsurf && k == 3) {
        zz3 = 0.0f;         // tzz[2] = 0
        zz2 = -zz4;         // tzz[1] = -tzz[3]
        xz2 = -xz3;         // txz[1] = -txz[2]
        xz1 = -xz4;         // txz[0] = -txz[3]
      }
      // 4. velocity of row r = k-2 (stress rows r-2..r+2 = *0..*4 for txz/tzz, txx row r = xx0)
      const int r = k - 2;
      const float txp1 = shfl(xx0, lane + 1), txp2 = shfl(xx0, lane + 2), txm1 = shfl(xx0, lane - 1);
      const float tzm1 = shfl(xz2, lane - 1), tzm2 = shfl(xz2, lane - 2), tzp1 = shfl(xz2, lane + 1);
      const bool own_r = out_col && r >= z0 && r < end_row;
      if (own_r) {
        const unsigned orr = rb(r);
        const float dxf_txx = c0 * (txp1 - xx0) + c1 * (txp2 - txm1);
        const float dzb_txz = c0 * (xz2 - xz1) + c1 * (xz3 - xz0);
        const float dxb_txz = c0 * (xz2 - tzm1) + c1 * (tzp1 - tzm2);
        const float dzf_tzz = c0 * (zz3 - zz2) + c1 * (zz4 - zz1);
        const float bxv = cur.b[0], bzv = cur.b[1];
        const float pzn = cur.pz[2], pzd = cur.pz[3];
        const float nx_x = (pxn * us[0][0] + ((d.dt * bxv) * dxf_txx) / d.dx) / pxd;
        const float nx_z = (pzn * us[0][1] + ((d.dt * bxv) * dzb_txz) / d.dz) / pzd;
        float nz_x = (pxn * us[0][2] + ((d.dt * bzv) * dxb_txz) / d.dx) / pxd;
        float nz_z = (pzn * us[0][3] + ((d.dt * bzv) * dzf_tzz) / d.dz) / pzd;
        if (d.src_fz && r == zs && c == xs) {   // vertical force: half of f w into each vz split
          const float hf = 0.5f * (fz * wsrc);
          nz_x = nz_x + hf;
          nz_z = nz_z + hf;
        }
        bst(nx_x, rOut, vs, F(VX_X) + orr); bst(nx_z, rOut, vs, F(VX_Z) + orr);
        bst(nz_x, rOut, vs, F(VZ_X) + orr); bst(nz_z, rOut, vs, F(VZ_Z) + orr);
        if (REC && rrows.has(r)) {   // velocity records of row r
          const int mr = rc.map[d.Z + (long long)r * d.pitch + c];
          if (mr >= 0) {
            const long long ro = ((long long)s * d.nt + t) * d.nrec + mr;
            rc.r[VX][ro] = nx_x + nx_z;
            rc.r[VZ][ro] = nz_x + nz_z;
          }
        }
      }
    };
    Row R0, R1, R2;
    load(z0 - 2, R0);
    load(z0 - 1, R1);
    // iterations rounded up to a multiple of three: the extra rows lie below the strip, where a
    // step stores nothing (branch-free unrolled steps keep the compiler's load waits exact)
    const int nit = (d.tz + 4 + 2) / 3 * 3;
#pragma unroll 1
    for (int kk = 0; kk < nit; kk += 3) {
      load(z0 + kk, R2);
      step(kk, R0);
      load(z0 + kk + 1, R0);
      step(kk + 1, R1);
      load(z0 + kk + 2, R1);
      step(kk + 2, R2);
    }
  }
}

// five records of step t from the state after step t (totals; the free-surface tzz row NN is zero)
__global__ void el_gather(ElDev d, const float *__restrict__ st, float *__restrict__ r_txx,
                          float *__restrict__ r_tzz, float *__restrict__ r_txz, float *__restrict__ r_vx,
                          float *__restrict__ r_vz, int t) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long long)d.S * d.nrec) return;
  const int s = (int)(k / d.nrec), r = (int)(k - (long long)s * d.nrec);
  const int z = d.rcv_z[r];
  const long long o = (long long)s * d.plane + (long long)z * d.pitch + d.rcv_x[r];
  const long long f = d.field;
  const long long ro = ((long long)s * d.nt + t) * d.nrec + r;
  r_txx[ro] = st[TXX_X * f + o] + st[TXX_Z * f + o];
  r_tzz[ro] = (d.fsurf && z == d.nn) ? 0.0f : st[TZZ_X * f + o] + st[TZZ_Z * f + o];
  r_txz[ro] = st[TXZ_X * f + o] + st[TXZ_Z * f + o];
  r_vx[ro] = st[VX_X * f + o] + st[VX_Z * f + o];
  r_vz[ro] = st[VZ_X * f + o] + st[VZ_Z * f + o];
}

// ---------------------------------------------------------------------------
// adjoint step
// ---------------------------------------------------------------------------
// The transpose of one forward step, for the adjoint L of the split fields after step t
// (record adjoints of step t added to L: by el_inject beforehand, or, with INJ, here as each
// row of L is read, in the same order of operations):
//   a. velocity update^T: e = L(v') * dt b / (d pd) feeds the stress-total adjoints
//      Lt_xx = Dxf^T e_vxx, Lt_xz = Dzb^T e_vxz + Dxb^T e_vzx, Lt_zz = Dzf^T e_vzz;
//      gradient of bx / bz: L(v') * dt / (d pd) * (D T) with T the forward stress totals;
//   b. stress free surface^T (tzz[2] = 0, tzz[1] = -tzz[3], txz[1] = -txz[2], txz[0] = -txz[3]):
//      Lt_zz[2] = 0, Lt_zz[3] -= Lt_zz[1], Lt_xz[2] -= Lt_xz[1], Lt_xz[3] -= Lt_xz[0];
//   c. L(s') = L(s after step) + Lt for both splits; the source gradient is sum_c m_c (L(s_cx') + L(s_cz'));
//   d. stress update^T: L(s) = pn pi L(s'), a = dt/d * pi * L(s'); the velocity-total adjoints
//      Lv_x = Dxb^T(C11 a_xxx + C13 a_zzx) + Dzf^T(C55 a_xzz) [+ Dzb^T(C15 a_xxz + C35 a_zzz) + Dxf^T(C15 a_xzx)],
//      Lv_z = Dzb^T(C13 a_xxz + C33 a_zzz) + Dxf^T(C55 a_xzx) [+ Dxb^T(C15 a_xxx + C35 a_zzx) + Dzf^T(C35 a_xzz)],
//      and the C gradients a * (D V) with V the forward velocity totals before the step;
//   e. velocity free surface^T (vz[1] = vz[2], vx[1] = 2 vz[2,j+1] - 2 vz[2,j] + vx[3,j], vz[0] = vz[2]
//      on the columns J = [2, X-3]): Lv_z[2,j] += Lv_z[1,j] + Lv_z[0,j] - 2 Lv_x[1,j] + 2 Lv_x[1,j-1],
//      Lv_x[3,j] += Lv_x[1,j];
//   f. L(v) = pn / pd L(v') + Lv for both splits.
// Inside the window only; the splits outside it never change, so their adjoints are not needed.
// Lanes and rows march exactly as in the forward step.  The four wavefronts of a workgroup
// (four shots of one strip) march their rows in lockstep: every iteration each writes its
// gradient terms of the rows it finished to an LDS row buffer, and after one barrier the
// wavefront that owns a coefficient field adds the four terms, in wavefront order, to the
// workgroup's plane of that field (its old value was loaded at the start of the iteration).
// Nothing is held per strip, so strips can be tall: each strip re-reads four halo rows of
// every field, 4/tz of its loads.
constexpr int ETZA = 8;   // strip rows of the adjoints that keep a per-strip LDS gradient slab

template <bool ANISO, bool INJ>
__global__ __launch_bounds__(ENT) void el_adj_step(ElDev d, const float *__restrict__ lin, float *__restrict__ lout,
                                                   const float *__restrict__ fv, const float *__restrict__ fsn, int t,
                                                   float *__restrict__ gacc, float *__restrict__ gsrc, ElInj inj) {
  constexpr int NGF = ANISO ? 8 : 6;
  // gradient terms of the rows finished in one iteration, per wavefront, double-buffered by
  // iteration parity (one barrier per iteration keeps a buffer from being rewritten early)
  __shared__ float rowbuf[2][EWPB][NGF][WAVE];
  // PML factors 1 - dt/2 d, 1 + dt/2 d, 1/(1 + dt/2 d) of the rows the strip reads (z0-4 ..
  // z0+tz+3), staged once: read from LDS they cost no wait on the vector loads in flight
  constexpr int PZR = ETZ_MAX + 8;
  __shared__ float pzs[3][PZR];
  int z0, x0;
  if (!strip_of(d, z0, x0)) return;   // uniform over the workgroup (one strip)
  static_assert(3 * PZR <= ENT, "one PML factor per thread");
  if (threadIdx.x < 3 * PZR) {
    // (rows outside the grid are never read; a test rather than a clamp keeps the compiler's
    // row limit, which every uniform row clamp uses, in a scalar register)
    const int q = threadIdx.x / PZR, j = threadIdx.x - q * PZR, row = z0 - 4 + j;
    pzs[q][j] = (unsigned)row < (unsigned)d.Z ? d.pml_z[q * d.Z + row] : 0.0f;
  }
  __syncthreads();
  auto pzf = [&](int q, int row) -> float { return pzs[q][row - (z0 - 4)]; };
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
  const int lane = threadIdx.x & (WAVE - 1);
  const int c = x0 - EHL + lane;
  const unsigned vc = 4u * (unsigned)min(max(c, 0), d.X - 1);
  const unsigned vs = 4u * (unsigned)max(c, 0);
  const bool out_col = lane >= EHL && lane < EHL + ESX && c <= d.X - 3;
  const bool cwin = c >= 2 && c <= d.X - 3;
  const int zmax = d.Z - 1, zw1 = d.Z - 3;
  const unsigned rowb = 4u * (unsigned)d.pitch, planeb = 4u * (unsigned)d.plane, fieldb = 4u * (unsigned)d.field;
  auto rb = [&](int k) -> unsigned { return (unsigned)min(max(k, 0), zmax) * rowb; };
  const rsrc_t rLi = rsrc(lin, (unsigned)NFIELD * fieldb);
  const rsrc_t rLo = rsrc(lout, (unsigned)NFIELD * fieldb);
  const rsrc_t rFv = rsrc(fv, (unsigned)NFIELD * fieldb);
  const rsrc_t rFs = rsrc(fsn, (unsigned)NFIELD * fieldb);
  const rsrc_t rC = rsrc(d.coef, (unsigned)NCOEF * planeb);
  // rows z0-4 .. z0+tz+3 are read (tz <= 64)
  const adfwi::RowMask rrows = adfwi::receiver_rows(INJ ? inj.map : nullptr, d.Z, z0 - 4, lane);
  const float c0 = d.c0, c1 = d.c1;
  const int cc = min(max(c, 0), d.X - 1);
  const float pxn = d.pml_x[cc], pxd = d.pml_x[d.pitch + cc], pxi = d.pml_x[2 * d.pitch + cc];
  const float sx_keep = pxn * pxi, sx_a = d.dt_dx * pxi;    // stress split: old-value factor, source factor
  const float vx_keep = pxn / pxd, vx_w = d.dt / (d.dx * pxd);
  const int end_row = min(z0 + d.tz, zw1 + 1);
  const bool jj = c >= 2 && c <= d.X - 3;
  const bool jj1 = c + 1 >= 2 && c + 1 <= d.X - 3;
  const bool jm1 = c - 1 >= 2 && c - 1 <= d.X - 3;
  const int grp = blockIdx.y * EWPB + wv;
  // this wavefront adds fields f = wv and wv + 4 (C fields of row k, b fields 4, 5 of row k-2)
  const long long planef = (long long)d.Z * d.pitch;
  // the same planes for the old values' loads, which every lane issues, so that a load lands in
  // its own register and nothing waits for it before the end of the iteration (a load in a
  // branch was copied at the branch's end, which waited for it)
  const rsrc_t rG = rsrc(gacc + (long long)blockIdx.y * NCOEF * planef, (unsigned)NCOEF * planeb);
  float gf[NGF];

  // every wavefront runs the same number of iterations (the barrier is in the row loop); one
  // whose shot is past the last computes nothing and contributes zeros
  for (int si = 0; si < d.spw; ++si) {
    const int s_raw = grp * d.spw + si;
    const bool active = s_raw < d.S;
    const int s = active ? s_raw : d.S - 1;   // in-range indices for the (skipped) shot set-up
    const unsigned so = (unsigned)s * planeb;
    auto F = [&](int f) -> unsigned { return (unsigned)f * fieldb + so; };
    const int zs = d.src_z[s], xs = d.src_x[s];
    const float mxx = d.src_m[3 * s], mzz = d.src_m[3 * s + 1], mxz = d.src_m[3 * s + 2];
    float gmt = 0.0f;   // wavelet gradient of the moment-tensor injection, when a force term follows
    // forward velocity totals (rows k-2..k+2) and stress totals (txz/tzz rows k-4..k, txx k-2..k)
    float vx0 = 0.f, vx1 = 0.f, vx2 = 0.f, vx3 = 0.f, vx4 = 0.f;
    float vz0 = 0.f, vz1 = 0.f, vz2 = 0.f, vz3 = 0.f, vz4 = 0.f;
    float xz0 = 0.f, xz1 = 0.f, xz2 = 0.f, xz3 = 0.f, xz4 = 0.f;
    float zz0 = 0.f, zz1 = 0.f, zz2 = 0.f, zz3 = 0.f, zz4 = 0.f;
    float xx0 = 0.f, xx1 = 0.f, xx2 = 0.f;
    // e windows: e1 (vx_x), e3 (vz_x) rows k..k+2; e2 (vx_z), e4 (vz_z) rows k-2..k+2
    float e1a = 0.f, e1b = 0.f, e1c = 0.f, e3a = 0.f, e3b = 0.f, e3c = 0.f;
    float e20 = 0.f, e21 = 0.f, e22 = 0.f, e23 = 0.f, e24 = 0.f;
    float e40 = 0.f, e41 = 0.f, e42 = 0.f, e43 = 0.f, e44 = 0.f;
    // stress-product windows: P1, Q3 (, P3, Q1) rows k-2..k; P4, Q2 (, P2, Q4) rows k-4..k
    float p1a = 0.f, p1b = 0.f, p1c = 0.f, q3a = 0.f, q3b = 0.f, q3c = 0.f;
    float p40 = 0.f, p41 = 0.f, p42 = 0.f, p43 = 0.f, p44 = 0.f;
    float q20 = 0.f, q21 = 0.f, q22 = 0.f, q23 = 0.f, q24 = 0.f;
    float p3a = 0.f, p3b = 0.f, p3c = 0.f, q1a = 0.f, q1b = 0.f, q1c = 0.f;
    float p20 = 0.f, p21 = 0.f, p22 = 0.f, p23 = 0.f, p24 = 0.f;
    float q40 = 0.f, q41 = 0.f, q42 = 0.f, q43 = 0.f, q44 = 0.f;
    // free-surface rows kept for the transposes
    float ltzz1 = 0.f, ltxz0 = 0.f, ltxz1 = 0.f, lvz0 = 0.f, lvz1 = 0.f, lvx1 = 0.f;

    // the record adjoint of one field at receiver index m in step t, added to both of its
    // splits as the adjoint state's own values
    auto add2 = [&](const float *g, int m, float &a, float &b) {
      if (g) {
        const float v = g[((long long)s * d.nt + t) * d.nrec + m];
        a = a + v;
        b = b + v;
      }
    };
    // receiver index at this lane's point of row rr (-1: none or outside the window)
    auto map_at = [&](int rr) -> int {
      return (INJ && rrows.has(rr) && rr >= 2 && rr <= zw1 && cwin) ? inj.map[d.Z + (long long)rr * d.pitch + c]
                                                                     : -1;
    };
    // forward velocity totals and e of row rr (zero outside the grid / window)
    auto load_row = [&](int rr_, float &fvx, float &fvz, float &ea, float &eb, float &ec, float &ed) {
      const int rr = __builtin_amdgcn_readfirstlane(rr_);   // uniform (a rolled loop's counter)
      const unsigned o = rb(rr);
      const bool in_grid = rr >= 0 && rr <= zmax;
      const float a = bld(rFv, vc, F(VX_X) + o) + bld(rFv, vc, F(VX_Z) + o);
      const float b = bld(rFv, vc, F(VZ_X) + o) + bld(rFv, vc, F(VZ_Z) + o);
      fvx = in_grid ? a : 0.f;
      fvz = in_grid ? b : 0.f;
      ea = eb = ec = ed = 0.f;
      if (rr >= 2 && rr <= zw1 && cwin) {
        const float bxv = bld(rC, vc, BX * planeb + o), bzv = bld(rC, vc, BZ * planeb + o);
        const float vz_w = d.dt / (d.dz * pzf(1, rr));
        float l0 = bld(rLi, vc, F(VX_X) + o), l1 = bld(rLi, vc, F(VX_Z) + o);
        float l2 = bld(rLi, vc, F(VZ_X) + o), l3 = bld(rLi, vc, F(VZ_Z) + o);
        if constexpr (INJ) {
          const int m = map_at(rr);
          if (m >= 0) {
            add2(inj.g[VX], m, l0, l1);
            add2(inj.g[VZ], m, l2, l3);
          }
        }
        ea = l0 * (vx_w * bxv);
        eb = l1 * (vz_w * bxv);
        ec = l2 * (vx_w * bzv);
        ed = l3 * (vz_w * bzv);
      }
    };
    // rolled: unrolled, the record-adjoint variant needs more than the 168 VGPRs of 3 waves per SIMD
#pragma unroll 1
    for (int q = -2; q <= 1 && active; ++q) {
      float fvx, fvz, ea, eb, ec, ed;
      load_row(z0 - 2 + q, fvx, fvz, ea, eb, ec, ed);
      vx1 = vx2; vx2 = vx3; vx3 = vx4; vx4 = fvx;
      vz1 = vz2; vz2 = vz3; vz3 = vz4; vz4 = fvz;
      e1a = e1b; e1b = e1c; e1c = ea;
      e21 = e22; e22 = e23; e23 = e24; e24 = eb;
      e3a = e3b; e3b = e3c; e3c = ec;
      e41 = e42; e42 = e43; e43 = e44; e44 = ed;
    }
    // everything iteration k reads from memory, loaded one iteration ahead: row k+2 of the
    // forward velocity splits, the velocity adjoints and the buoyancies; row k of the stress
    // adjoints, the moduli and the forward stress splits; row k-2 of the velocity adjoints
    struct Row {
      float fv[4], lv[4], b[2], ls[6], cm[ANISO ? 6 : 4], fs[6];
    };
    // velocity adjoints of rows k-2 .. k+2 as loaded (the row k-2 ones are the old values the
    // velocity stage needs, so they are not loaded a second time)
    float lw[5][4] = {};
    auto load = [&](int k, Row &R) {
      if (!active) return;
      const unsigned o2 = rb(k + 2), o0 = rb(k);
#pragma unroll
      for (int q = 0; q < 4; ++q) R.fv[q] = bld(rFv, vc, F(VX_X + q) + o2);
#pragma unroll
      for (int q = 0; q < 4; ++q) R.lv[q] = bld(rLi, vc, F(VX_X + q) + o2);
      R.b[0] = bld(rC, vc, BX * planeb + o2); R.b[1] = bld(rC, vc, BZ * planeb + o2);
#pragma unroll
      for (int q = 0; q < 6; ++q) R.ls[q] = bld(rLi, vc, F(TXX_X + q) + o0);
      R.cm[0] = bld(rC, vc, C11 * planeb + o0); R.cm[1] = bld(rC, vc, C13 * planeb + o0);
      R.cm[2] = bld(rC, vc, C33 * planeb + o0); R.cm[3] = bld(rC, vc, C55 * planeb + o0);
      if constexpr (ANISO) {
        R.cm[4] = bld(rC, vc, C15 * planeb + o0); R.cm[5] = bld(rC, vc, C35 * planeb + o0);
      }
#pragma unroll
      for (int q = 0; q < 6; ++q) R.fs[q] = bld(rFs, vc, F(TXX_X + q) + o0);
    };
    Row cur;
    load(z0 - 2, cur);
#pragma unroll 1
    for (int kk = 0; kk < d.tz + 4; ++kk) {
      const int k = z0 - 2 + kk;
      // old plane values of this wavefront's fields at the rows finished this iteration
      float gold[2] = {0.0f, 0.0f};
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int f = wv + j * EWPB;
        const int row = (f == 4 || f == 5) ? k - 2 : k;
        // lanes that do not own the point take an offset past the planes: such a buffer load
        // returns zero without touching memory, so the load needs no branch
        const bool own = f < NGF && out_col && row >= z0 && row < end_row;   // (as at the store)
        gold[j] = bld(rG, own ? 4u * (unsigned)cc : 0x80000000u, (unsigned)min(f, NGF - 1) * planeb + rb(row));
      }
#pragma unroll
      for (int f = 0; f < NGF; ++f) gf[f] = 0.0f;
      Row nxt;
      load(k + 1, nxt);      // next row's loads go out before this row's stores
      if (active) {
      if constexpr (INJ) {
        // record adjoints of step t: velocity fields of row k+2, stress fields of row k
        if (rrows.has(k + 2)) {
          const int m = map_at(k + 2);
          if (m >= 0) {
            add2(inj.g[VX], m, cur.lv[0], cur.lv[1]);
            add2(inj.g[VZ], m, cur.lv[2], cur.lv[3]);
          }
        }
        if (rrows.has(k)) {
          const int m = map_at(k);
          if (m >= 0) {
            add2(inj.g[TXX], m, cur.ls[0], cur.ls[1]);
            if (!(d.fsurf && k == 2)) add2(inj.g[TZZ], m, cur.ls[2], cur.ls[3]);   // free-surface tzz: none
            add2(inj.g[TXZ], m, cur.ls[4], cur.ls[5]);
          }
        }
      }
      {
        const int rr = k + 2;
        const bool in_grid = rr <= zmax;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
#pragma unroll
          for (int q = 0; q < 4; ++q) lw[e][q] = lw[e + 1][q];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) lw[4][q] = cur.lv[q];
        const float fvx = in_grid ? cur.fv[0] + cur.fv[1] : 0.f;
        const float fvz = in_grid ? cur.fv[2] + cur.fv[3] : 0.f;
        float ea = 0.f, eb = 0.f, ec = 0.f, ed = 0.f;
        if (rr >= 2 && rr <= zw1 && cwin) {
          const float vz_w = d.dt / (d.dz * pzf(1, rr));
          ea = cur.lv[0] * (vx_w * cur.b[0]);
          eb = cur.lv[1] * (vz_w * cur.b[0]);
          ec = cur.lv[2] * (vx_w * cur.b[1]);
          ed = cur.lv[3] * (vz_w * cur.b[1]);
        }
        vx0 = vx1; vx1 = vx2; vx2 = vx3; vx3 = vx4; vx4 = fvx;
        vz0 = vz1; vz1 = vz2; vz2 = vz3; vz3 = vz4; vz4 = fvz;
        e1a = e1b; e1b = e1c; e1c = ea;
        e20 = e21; e21 = e22; e22 = e23; e23 = e24; e24 = eb;
        e3a = e3b; e3b = e3c; e3c = ec;
        e40 = e41; e41 = e42; e42 = e43; e43 = e44; e44 = ed;
      }
      if (d.fsurf && k == 2) {   // forward velocity free surface (rows 0, 1 from rows 2, 3)
        const float vz2n = shfl(vz2, lane + 1);
        const float vz1m_n = jj1 ? vz2n : 0.0f;
        const float vz1m = jj ? vz2 : vz1;
        if (jj) {
          vx1 = (((vz1m_n - vz1m) + vz2n) - vz2) + vx3;
          vz0 = vz1m;
        }
        vz1 = vz1m;
      }
      // a. stress-total adjoints of row k
      const float e1m1 = shfl(e1a, lane - 1), e1m2 = shfl(e1a, lane - 2), e1p1 = shfl(e1a, lane + 1);
      const float e3p1 = shfl(e3a, lane + 1), e3m1 = shfl(e3a, lane - 1), e3p2 = shfl(e3a, lane + 2);
      float lt_xx = c0 * (e1m1 - e1a) + c1 * (e1m2 - e1p1);
      float lt_xz = (c0 * (e22 - e23) + c1 * (e21 - e24)) + (c0 * (e3a - e3p1) + c1 * (e3m1 - e3p2));
      float lt_zz = c0 * (e41 - e42) + c1 * (e40 - e43);
      // b. stress free surface transposed
      if (d.fsurf) {
        if (k == 0) ltxz0 = lt_xz;
        if (k == 1) { ltzz1 = lt_zz; ltxz1 = lt_xz; }
        if (k == 2) { lt_zz = 0.0f; lt_xz = lt_xz - ltxz1; }
        if (k == 3) { lt_zz = lt_zz - ltzz1; lt_xz = lt_xz - ltxz0; }
      }
      // forward velocity derivatives of row k (for the C gradients)
      const float vxm1 = shfl(vx2, lane - 1), vxm2 = shfl(vx2, lane - 2), vxp1 = shfl(vx2, lane + 1);
      const float vzm1 = shfl(vz2, lane - 1), vzp1 = shfl(vz2, lane + 1), vzp2 = shfl(vz2, lane + 2);
      float vzm2 = 0.0f, vxp2 = 0.0f;
      if constexpr (ANISO) {
        vzm2 = shfl(vz2, lane - 2);
        vxp2 = shfl(vx2, lane + 2);
      }
      // c, d. stress splits of row k
      float P1 = 0.f, Q2 = 0.f, Q3 = 0.f, P4 = 0.f, P2 = 0.f, P3 = 0.f, Q1 = 0.f, Q4 = 0.f;
      if (k >= 2 && k <= zw1 && cwin) {
        const unsigned o = rb(k);
        const float lxx_x = cur.ls[0] + lt_xx, lxx_z = cur.ls[1] + lt_xx;
        const float lzz_x = cur.ls[2] + lt_zz, lzz_z = cur.ls[3] + lt_zz;
        const float lxz_x = cur.ls[4] + lt_xz, lxz_z = cur.ls[5] + lt_xz;
        const bool own_k = out_col && k >= z0 && k < end_row;
        if (own_k && k == zs && c == xs && gsrc) {
          const float gm = (mxx * (lxx_x + lxx_z) + mzz * (lzz_x + lzz_z)) + mxz * (lxz_x + lxz_z);
          if (d.src_fz) gmt = gm;        // the force term of row zs follows two rows later
          else gsrc[(long long)s * d.nt + t] = gm;
        }
        const float pzn = pzf(0, k), pzi = pzf(2, k);
        const float sz_keep = pzn * pzi, sz_a = d.dt_dz * pzi;
        if (own_k) {
          bst(sx_keep * lxx_x, rLo, vs, F(TXX_X) + o); bst(sz_keep * lxx_z, rLo, vs, F(TXX_Z) + o);
          bst(sx_keep * lzz_x, rLo, vs, F(TZZ_X) + o); bst(sz_keep * lzz_z, rLo, vs, F(TZZ_Z) + o);
          bst(sx_keep * lxz_x, rLo, vs, F(TXZ_X) + o); bst(sz_keep * lxz_z, rLo, vs, F(TXZ_Z) + o);
        }
        const float axx_x = sx_a * lxx_x, axx_z = sz_a * lxx_z;
        const float azz_x = sx_a * lzz_x, azz_z = sz_a * lzz_z;
        const float axz_x = sx_a * lxz_x, axz_z = sz_a * lxz_z;
        const float C11v = cur.cm[0], C13v = cur.cm[1], C33v = cur.cm[2], C55v = cur.cm[3];
        P1 = C11v * axx_x + C13v * azz_x;
        Q2 = C13v * axx_z + C33v * azz_z;
        Q3 = C55v * axz_x;
        P4 = C55v * axz_z;
        const float dxb_vx = c0 * (vx2 - vxm1) + c1 * (vxp1 - vxm2);
        const float dzb_vz = c0 * (vz2 - vz1) + c1 * (vz3 - vz0);
        const float dxf_vz = c0 * (vzp1 - vz2) + c1 * (vzp2 - vzm1);
        const float dzf_vx = c0 * (vx3 - vx2) + c1 * (vx4 - vx1);
        if constexpr (ANISO) {
          const float C15v = cur.cm[4], C35v = cur.cm[5];
          P2 = C15v * axx_z + C35v * azz_z;
          P3 = C15v * axz_x;
          Q1 = C15v * axx_x + C35v * azz_x;
          Q4 = C35v * axz_z;
          if (own_k) {
            const float dxb_vz = c0 * (vz2 - vzm1) + c1 * (vzp1 - vzm2);
            const float dzb_vx = c0 * (vx2 - vx1) + c1 * (vx3 - vx0);
            const float dxf_vx = c0 * (vxp1 - vx2) + c1 * (vxp2 - vxm1);
            const float dzf_vz = c0 * (vz3 - vz2) + c1 * (vz4 - vz1);
            gf[6] = axx_x * dxb_vz + axx_z * dzb_vx + axz_x * dxf_vx;
            gf[7] = azz_x * dxb_vz + azz_z * dzb_vx + axz_z * dzf_vz;
          }
        }
        if (own_k) {
          gf[0] = axx_x * dxb_vx;
          gf[1] = axx_z * dzb_vz + azz_x * dxb_vx;
          gf[2] = azz_z * dzb_vz;
          gf[3] = axz_x * dxf_vz + axz_z * dzf_vx;
        }
      }
      p1a = p1b; p1b = p1c; p1c = P1;
      q3a = q3b; q3b = q3c; q3c = Q3;
      p40 = p41; p41 = p42; p42 = p43; p43 = p44; p44 = P4;
      q20 = q21; q21 = q22; q22 = q23; q23 = q24; q24 = Q2;
      if constexpr (ANISO) {
        p3a = p3b; p3b = p3c; p3c = P3;
        q1a = q1b; q1b = q1c; q1c = Q1;
        p20 = p21; p21 = p22; p22 = p23; p23 = p24; p24 = P2;
        q40 = q41; q41 = q42; q42 = q43; q43 = q44; q44 = Q4;
      }
      // forward stress totals of row k (after the step), free surface once row 3 exists
      {
        const bool in_grid = k >= 0 && k <= zmax;
        const float a = cur.fs[0] + cur.fs[1];
        const float b = cur.fs[2] + cur.fs[3];
        const float e = cur.fs[4] + cur.fs[5];
        xx0 = xx1; xx1 = xx2; xx2 = in_grid ? a : 0.f;
        zz0 = zz1; zz1 = zz2; zz2 = zz3; zz3 = zz4; zz4 = in_grid ? b : 0.f;
        xz0 = xz1; xz1 = xz2; xz2 = xz3; xz3 = xz4; xz4 = in_grid ? e : 0.f;
      }
      if (d.fsurf && k == 3) {
        zz3 = 0.0f; zz2 = -zz4; xz2 = -xz3; xz1 = -xz4;
      }
      // e. velocity-total adjoints of row r = k-2
      const int r = k - 2;
      const float p1m1 = shfl(p1a, lane - 1), p1p1 = shfl(p1a, lane + 1), p1p2 = shfl(p1a, lane + 2);
      const float q3m1 = shfl(q3a, lane - 1), q3m2 = shfl(q3a, lane - 2), q3p1 = shfl(q3a, lane + 1);
      float lvx = (c0 * (p1a - p1p1) + c1 * (p1m1 - p1p2)) + (c0 * (p41 - p42) + c1 * (p40 - p43));
      float lvz = (c0 * (q22 - q23) + c1 * (q21 - q24)) + (c0 * (q3m1 - q3a) + c1 * (q3m2 - q3p1));
      if constexpr (ANISO) {
        const float p3m1 = shfl(p3a, lane - 1), p3m2 = shfl(p3a, lane - 2), p3p1 = shfl(p3a, lane + 1);
        const float q1m1 = shfl(q1a, lane - 1), q1p1 = shfl(q1a, lane + 1), q1p2 = shfl(q1a, lane + 2);
        lvx = lvx + ((c0 * (p22 - p23) + c1 * (p21 - p24)) + (c0 * (p3m1 - p3a) + c1 * (p3m2 - p3p1)));
        lvz = lvz + ((c0 * (q1a - q1p1) + c1 * (q1m1 - q1p2)) + (c0 * (q41 - q42) + c1 * (q40 - q43)));
      }
      if (d.fsurf) {
        if (r == 0) lvz0 = lvz;
        if (r == 1) { lvz1 = lvz; lvx1 = lvx; }
        if (r == 2) {
          const float lvx1m = shfl(lvx1, lane - 1);
          if (jj) lvz = lvz + (((lvz1 + lvz0) - 2.0f * lvx1) + (jm1 ? 2.0f * lvx1m : 0.0f));
        }
        if (r == 3 && jj) lvx = lvx + lvx1;
      }
      // forward stress derivatives of row r (for the b gradients)
      const float txp1 = shfl(xx0, lane + 1), txp2 = shfl(xx0, lane + 2), txm1 = shfl(xx0, lane - 1);
      const float tzm1 = shfl(xz2, lane - 1), tzm2 = shfl(xz2, lane - 2), tzp1 = shfl(xz2, lane + 1);
      const bool own_r = out_col && r >= z0 && r < end_row;
      if (own_r) {
        const unsigned o = rb(r);
        const float pzn = pzf(0, r), pzd = pzf(1, r);
        const float vz_keep = pzn / pzd, vz_w = d.dt / (d.dz * pzd);
        const float l1 = lw[0][0], l2 = lw[0][1], l3 = lw[0][2], l4 = lw[0][3];
        if (d.src_fz && gsrc && r == zs && c == xs)   // vertical force: f w / 2 entered both vz splits
          gsrc[(long long)s * d.nt + t] = gmt + 0.5f * (d.src_fz[s] * (l3 + l4));
        bst(vx_keep * l1 + lvx, rLo, vs, F(VX_X) + o); bst(vz_keep * l2 + lvx, rLo, vs, F(VX_Z) + o);
        bst(vx_keep * l3 + lvz, rLo, vs, F(VZ_X) + o); bst(vz_keep * l4 + lvz, rLo, vs, F(VZ_Z) + o);
        const float dxf_txx = c0 * (txp1 - xx0) + c1 * (txp2 - txm1);
        const float dzb_txz = c0 * (xz2 - xz1) + c1 * (xz3 - xz0);
        const float dxb_txz = c0 * (xz2 - tzm1) + c1 * (tzp1 - tzm2);
        const float dzf_tzz = c0 * (zz3 - zz2) + c1 * (zz4 - zz1);
        gf[4] = l1 * (vx_w * dxf_txx) + l2 * (vz_w * dzb_txz);
        gf[5] = l3 * (vx_w * dxb_txz) + l4 * (vz_w * dzf_tzz);
      }
      cur = nxt;
      }  // active
      // the four wavefronts' terms of the finished rows, summed in wavefront order
      const int par = kk & 1;
#pragma unroll
      for (int f = 0; f < NGF; ++f) rowbuf[par][wv][f][lane] = gf[f];
      __syncthreads();
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int f = wv + j * EWPB;
        const int row = (f == 4 || f == 5) ? k - 2 : k;
        float v = 0.0f;
        if (f < NGF)
          v = (rowbuf[par][0][f][lane] + rowbuf[par][1][f][lane]) + (rowbuf[par][2][f][lane] + rowbuf[par][3][f][lane]);
        // stored by every lane, past the planes where the lane does not own the point (such a
        // buffer store is dropped): the old value is then consumed on every path, and no load
        // into its register is left in flight when the next iteration reuses it
        const bool own = f < NGF && out_col && row >= z0 && row < end_row;
        bst(gold[j] + v, rG, own ? 4u * (unsigned)cc : 0x80000000u, (unsigned)min(f, NGF - 1) * planeb + rb(row));
      }
    }
  }
}

// adds the record adjoints of step t to both splits of each recorded field (free-surface tzz
// on row NN is identically zero there, so it receives nothing)
__global__ void el_inject(ElDev d, float *__restrict__ lam, const float *__restrict__ g_txx,
                          const float *__restrict__ g_tzz, const float *__restrict__ g_txz,
                          const float *__restrict__ g_vx, const float *__restrict__ g_vz, int t) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long long)d.S * d.nrec) return;
  const int s = (int)(k / d.nrec), r = (int)(k - (long long)s * d.nrec);
  const int z = d.rcv_z[r];
  const long long o = (long long)s * d.plane + (long long)z * d.pitch + d.rcv_x[r];
  const long long f = d.field;
  const long long ro = ((long long)s * d.nt + t) * d.nrec + r;
  const float *gs[5] = {g_txx, g_tzz, g_txz, g_vx, g_vz};
  for (int q = 0; q < 5; ++q) {
    if (!gs[q] || (q == 1 && d.fsurf && z == d.nn)) continue;
    const float v = gs[q][ro];
    atomicAdd(lam + (2 * q) * f + o, v);
    atomicAdd(lam + (2 * q + 1) * f + o, v);
  }
}

// ---------------------------------------------------------------------------
// damping scheme (reference: step_forward_ABL_4order, elastic_kernels.py:585-777)
// ---------------------------------------------------------------------------
// State [5][S][Z][pitch] = txx, tzz, txz, vx, vz.  One step, on the window rows [2, Z-3] x cols
// [2, X-3]:
//   1. s += dt ((C.Dx v)/dx + (C.Dz v)/dz) for txx, tzz, txz (Dxb/Dzb for the normal stresses,
//      Dxf/Dzf for the shear stress), then the source -MT/3 w(t);
//   2. free surface, kept in the state: tzz[2] = 0, tzz[1] = -tzz[3], txz[1] = -txz[2],
//      txz[0] = -txz[3];
//   3. vx += (dt bx) ((Dxf txx)/dx + (Dzb txz)/dz), vz += (dt bz) ((Dxb txz)/dx + (Dzf tzz)/dz);
//   4. free surface on the columns J = [2, X-3], from the new rows 2 and 3: vz[1] = vz[2],
//      vx[1] = vz[1,j+1] - vz[1,j] + vz[2,j+1] - vz[2,j] + vx[3,j], vz[0] = vz[1];
//   5. vx, vz *= damp everywhere (points that never change stay zero).
// The strip march is that of el_fwd_step: velocity rows k-2..k+2 and the new stress rows in
// registers; the first strip also writes the free-surface rows 0 and 1.
template <bool ANISO>
__global__ __launch_bounds__(ENT) void el_abl_fwd(ElDev d, const float *__restrict__ in, float *__restrict__ out,
                                                  int t) {
  int z0, x0;
  if (!strip_of(d, z0, x0)) return;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
  const int lane = threadIdx.x & (WAVE - 1);
  const int c = x0 - EHL + lane;
  const unsigned vc = 4u * (unsigned)min(max(c, 0), d.X - 1);
  const unsigned vs = 4u * (unsigned)max(c, 0);
  const bool out_col = lane >= EHL && lane < EHL + ESX && c <= d.X - 3;
  const bool cwin = c >= 2 && c <= d.X - 3;
  const int zmax = d.Z - 1, zw1 = d.Z - 3;
  const unsigned rowb = 4u * (unsigned)d.pitch, planeb = 4u * (unsigned)d.plane, fieldb = 4u * (unsigned)d.field;
  auto rb = [&](int k) -> unsigned { return (unsigned)min(max(k, 0), zmax) * rowb; };
  const rsrc_t rIn = rsrc(in, (unsigned)NFIELD5 * fieldb);
  const rsrc_t rOut = rsrc(out, (unsigned)NFIELD5 * fieldb);
  const rsrc_t rC = rsrc(d.coef, (unsigned)NCOEF * planeb);
  const rsrc_t rD = rsrc(d.damp, planeb);
  const float c0 = d.c0, c1 = d.c1, dt = d.dt, dx = d.dx, dz = d.dz;
  const int end_row = min(z0 + d.tz, zw1 + 1);
  const bool jj = c >= 2 && c <= d.X - 3;
  const bool jj1 = c + 1 >= 2 && c + 1 <= d.X - 3;
  const bool first = d.fsurf && z0 == 2;         // the strip that writes the free-surface rows

  const int s_first = (blockIdx.y * EWPB + wv) * d.spw;
  for (int si = 0; si < d.spw; ++si) {
    const int s = s_first + si;
    if (s >= d.S) break;
    const unsigned so = (unsigned)s * planeb;
    auto F = [&](int f) -> unsigned { return (unsigned)f * fieldb + so; };
    const int zs = d.src_z[s], xs = d.src_x[s];
    const float wsrc = d.src[(long long)s * d.nt + t];
    const float mxx = d.src_m[3 * s], mzz = d.src_m[3 * s + 1], mxz = d.src_m[3 * s + 2];
    float vx0 = 0.f, vx1 = 0.f, vx2 = 0.f, vx3 = 0.f, vx4 = 0.f;
    float vz0 = 0.f, vz1 = 0.f, vz2 = 0.f, vz3 = 0.f, vz4 = 0.f;
    float xz0 = 0.f, xz1 = 0.f, xz2 = 0.f, xz3 = 0.f, xz4 = 0.f;
    float zz0 = 0.f, zz1 = 0.f, zz2 = 0.f, zz3 = 0.f, zz4 = 0.f;
    float xx0 = 0.f, xx1 = 0.f, xx2 = 0.f;
    float nvz2 = 0.f;                            // new vz of row 2 before damping (free surface)
    for (int rr = z0 - 4; rr <= z0 - 1; ++rr) {
      const unsigned o = rb(rr);
      const bool ok = rr >= 0;
      const float a = bld(rIn, vc, F(VX) + o), b = bld(rIn, vc, F(VZ) + o);
      vx1 = vx2; vx2 = vx3; vx3 = vx4; vx4 = ok ? a : 0.f;
      vz1 = vz2; vz2 = vz3; vz3 = vz4; vz4 = ok ? b : 0.f;
    }
    struct Row {
      float v[2], s[3], cm[ANISO ? 6 : 4], b[2], dm;
    };
    auto load = [&](int k, Row &R) {
      const unsigned o2 = rb(k + 2), o0 = rb(k), om = rb(k - 2);
      R.v[0] = bld(rIn, vc, F(VX) + o2); R.v[1] = bld(rIn, vc, F(VZ) + o2);
      R.s[0] = bld(rIn, vc, F(TXX) + o0); R.s[1] = bld(rIn, vc, F(TZZ) + o0); R.s[2] = bld(rIn, vc, F(TXZ) + o0);
      R.cm[0] = bld(rC, vc, C11 * planeb + o0); R.cm[1] = bld(rC, vc, C13 * planeb + o0);
      R.cm[2] = bld(rC, vc, C33 * planeb + o0); R.cm[3] = bld(rC, vc, C55 * planeb + o0);
      if constexpr (ANISO) {
        R.cm[4] = bld(rC, vc, C15 * planeb + o0); R.cm[5] = bld(rC, vc, C35 * planeb + o0);
      }
      R.b[0] = bld(rC, vc, BX * planeb + om); R.b[1] = bld(rC, vc, BZ * planeb + om);
      R.dm = bld(rD, vc, om);
    };
    Row cur;
    load(z0 - 2, cur);
#pragma unroll 1
    for (int kk = 0; kk < d.tz + 4; ++kk) {
      const int k = z0 - 2 + kk;
      Row nxt;
      load(k + 1, nxt);
      {
        const bool ok = k + 2 <= zmax;
        vx0 = vx1; vx1 = vx2; vx2 = vx3; vx3 = vx4; vx4 = ok ? cur.v[0] : 0.f;
        vz0 = vz1; vz1 = vz2; vz2 = vz3; vz3 = vz4; vz4 = ok ? cur.v[1] : 0.f;
      }
      // 1. stress of row k
      const unsigned ok_ = rb(k);
      float sxx = cur.s[0], szz = cur.s[1], sxz = cur.s[2];
      const float vxm1 = shfl(vx2, lane - 1), vxm2 = shfl(vx2, lane - 2), vxp1 = shfl(vx2, lane + 1);
      const float vzm1 = shfl(vz2, lane - 1), vzp1 = shfl(vz2, lane + 1), vzp2 = shfl(vz2, lane + 2);
      float vzm2 = 0.0f, vxp2 = 0.0f;
      if constexpr (ANISO) {
        vzm2 = shfl(vz2, lane - 2);
        vxp2 = shfl(vx2, lane + 2);
      }
      if (k >= 2 && k <= zw1 && cwin) {
        const float dxb_vx = c0 * (vx2 - vxm1) + c1 * (vxp1 - vxm2);
        const float dzb_vz = c0 * (vz2 - vz1) + c1 * (vz3 - vz0);
        const float dxf_vz = c0 * (vzp1 - vz2) + c1 * (vzp2 - vzm1);
        const float dzf_vx = c0 * (vx3 - vx2) + c1 * (vx4 - vx1);
        const float C11v = cur.cm[0], C13v = cur.cm[1], C33v = cur.cm[2], C55v = cur.cm[3];
        if constexpr (ANISO) {
          const float dxb_vz = c0 * (vz2 - vzm1) + c1 * (vzp1 - vzm2);
          const float dzb_vx = c0 * (vx2 - vx1) + c1 * (vx3 - vx0);
          const float dxf_vx = c0 * (vxp1 - vx2) + c1 * (vxp2 - vxm1);
          const float dzf_vz = c0 * (vz3 - vz2) + c1 * (vz4 - vz1);
          const float C15v = cur.cm[4], C35v = cur.cm[5];
          sxx = sxx + dt * ((C11v * dxb_vx + C15v * dxb_vz) / dx + (C15v * dzb_vx + C13v * dzb_vz) / dz);
          szz = szz + dt * ((C13v * dxb_vx + C35v * dxb_vz) / dx + (C35v * dzb_vx + C33v * dzb_vz) / dz);
          sxz = sxz + dt * ((C15v * dxf_vx + C55v * dxf_vz) / dx + (C55v * dzf_vx + C35v * dzf_vz) / dz);
        } else {
          sxx = sxx + dt * ((C11v * dxb_vx) / dx + (C13v * dzb_vz) / dz);
          szz = szz + dt * ((C13v * dxb_vx) / dx + (C33v * dzb_vz) / dz);
          sxz = sxz + dt * ((C55v * dxf_vz) / dx + (C55v * dzf_vx) / dz);
        }
      }
      if (k == zs && c == xs) {
        sxx = sxx + mxx * wsrc;
        szz = szz + mzz * wsrc;
        sxz = sxz + mxz * wsrc;
      }
      if (d.fsurf && k == 2) szz = 0.0f;        // 2. tzz[2] = 0 (the value before it is not used)
      const bool own_k = out_col && k >= z0 && k < end_row;
      if (own_k) {
        bst(sxx, rOut, vs, F(TXX) + ok_); bst(szz, rOut, vs, F(TZZ) + ok_); bst(sxz, rOut, vs, F(TXZ) + ok_);
      }
      xx0 = xx1; xx1 = xx2; xx2 = sxx;
      xz0 = xz1; xz1 = xz2; xz2 = xz3; xz3 = xz4; xz4 = sxz;
      zz0 = zz1; zz1 = zz2; zz2 = zz3; zz3 = zz4; zz4 = szz;
      if (d.fsurf && k == 3) {                   // rows -1..3 = *0..*4: the rows 0 and 1
        zz2 = -zz4;
        xz2 = -xz3;
        xz1 = -xz4;
        if (out_col) {
          bst(zz2, rOut, vs, F(TZZ) + rb(1));
          bst(xz2, rOut, vs, F(TXZ) + rb(1));
          bst(xz1, rOut, vs, F(TXZ) + rb(0));
        }
      }
      // 3. velocity of row r = k-2 (old values: the window rows vx0, vz0)
      const int r = k - 2;
      const float txp1 = shfl(xx0, lane + 1), txp2 = shfl(xx0, lane + 2), txm1 = shfl(xx0, lane - 1);
      const float tzm1 = shfl(xz2, lane - 1), tzm2 = shfl(xz2, lane - 2), tzp1 = shfl(xz2, lane + 1);
      const bool own_r = out_col && r >= z0 && r < end_row;
      float nvx = 0.0f, nvz = 0.0f;
      if (r >= 2 && r <= zw1 && cwin) {
        const float dxf_txx = c0 * (txp1 - xx0) + c1 * (txp2 - txm1);
        const float dzb_txz = c0 * (xz2 - xz1) + c1 * (xz3 - xz0);
        const float dxb_txz = c0 * (xz2 - tzm1) + c1 * (tzp1 - tzm2);
        const float dzf_tzz = c0 * (zz3 - zz2) + c1 * (zz4 - zz1);
        nvx = vx0 + (dt * cur.b[0]) * (dxf_txx / dx + dzb_txz / dz);
        nvz = vz0 + (dt * cur.b[1]) * (dxb_txz / dx + dzf_tzz / dz);
      }
      if (d.src_fz && r == zs && c == xs) nvz = nvz + d.src_fz[s] * wsrc;   // vertical force
      if (own_r) {
        const unsigned orr = rb(r);
        bst(nvx * cur.dm, rOut, vs, F(VX) + orr);
        bst(nvz * cur.dm, rOut, vs, F(VZ) + orr);
      }
      // 4. free-surface velocity rows from the new rows 2 and 3 (every lane computes its column)
      if (first && r == 2) nvz2 = nvz;
      if (first && r == 3) {
        const float vz2n = shfl(nvz2, lane + 1);
        const float vz1m_n = jj1 ? vz2n : 0.0f;
        const float vz1m = jj ? nvz2 : 0.0f;
        if (out_col) {
          const float d1 = bld(rD, vc, rb(1)), d0 = bld(rD, vc, rb(0));
          const float nvx1 = (((vz1m_n - vz1m) + vz2n) - nvz2) + nvx;
          bst(nvx1 * d1, rOut, vs, F(VX) + rb(1));
          bst(vz1m * d1, rOut, vs, F(VZ) + rb(1));
          bst(vz1m * d0, rOut, vs, F(VZ) + rb(0));
        }
      }
      cur = nxt;
    }
  }
}

// five records of step t from the state after step t (damping scheme: stored totals)
__global__ void el_gather5(ElDev d, const float *__restrict__ st, float *__restrict__ r_txx,
                           float *__restrict__ r_tzz, float *__restrict__ r_txz, float *__restrict__ r_vx,
                           float *__restrict__ r_vz, int t) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long long)d.S * d.nrec) return;
  const int s = (int)(k / d.nrec), r = (int)(k - (long long)s * d.nrec);
  const long long o = (long long)s * d.plane + (long long)d.rcv_z[r] * d.pitch + d.rcv_x[r];
  const long long f = d.field;
  const long long ro = ((long long)s * d.nt + t) * d.nrec + r;
  r_txx[ro] = st[TXX * f + o];
  r_tzz[ro] = st[TZZ * f + o];
  r_txz[ro] = st[TXZ * f + o];
  r_vx[ro] = st[VX * f + o];
  r_vz[ro] = st[VZ * f + o];
}

// record adjoints of step t into the five-field adjoint state (tzz on the free-surface row NN
// is identically zero there, so it receives nothing)
__global__ void el_inject5(ElDev d, float *__restrict__ lam, const float *__restrict__ g_txx,
                           const float *__restrict__ g_tzz, const float *__restrict__ g_txz,
                           const float *__restrict__ g_vx, const float *__restrict__ g_vz, int t) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long long)d.S * d.nrec) return;
  const int s = (int)(k / d.nrec), r = (int)(k - (long long)s * d.nrec);
  const int z = d.rcv_z[r];
  const long long o = (long long)s * d.plane + (long long)z * d.pitch + d.rcv_x[r];
  const long long ro = ((long long)s * d.nt + t) * d.nrec + r;
  const float *gs[5] = {g_txx, g_tzz, g_txz, g_vx, g_vz};
  for (int q = 0; q < 5; ++q) {
    if (!gs[q] || (q == TZZ && d.fsurf && z == d.nn)) continue;
    atomicAdd(lam + q * d.field + o, gs[q][ro]);
  }
}

// Adjoint of one damping-scheme step (the forward's stages in reverse), for the adjoint L of the
// state after step t (record adjoints of step t already added by el_inject5):
//   5'. L2(v) = damp * L(v) everywhere;
//   4'. free surface of the velocities transposed into L1(v): on J, Lvz[2] += Lvz2[1] + Lvz2[0]
//       - 2 Lvx2[1,j] + 2 Lvx2[1,j-1], Lvx[3] += Lvx2[1]; the overwritten rows 0, 1 get zero;
//   3'. e = L1(v) dt b / d feeds the stress adjoints Lt (as el_adj_step), b gradients;
//   2'. stress free surface transposed; the stored tzz[2] = 0 is a constant, so L1(tzz)[2] = 0;
//   1'. L(s before) = L1(s); the velocity adjoints L(v before) = L1(v) + D^T(C dt/d L1(s));
//       C gradients; the source gradient sum_c m_c L1(s_c)(source).
// Rows and lanes march as in el_adj_step; the first strip also writes the velocity adjoints of
// the free-surface rows 0 and 1 (those rows are state here).
template <bool ANISO>
__global__ __launch_bounds__(ENT) void el_abl_adj(ElDev d, const float *__restrict__ lin, float *__restrict__ lout,
                                                  const float *__restrict__ fv, const float *__restrict__ fsn, int t,
                                                  float *__restrict__ gacc, float *__restrict__ gsrc) {
  constexpr int NGF = ANISO ? 8 : 6;
  __shared__ float slab[EWPB * NGF * ETZA * WAVE];
  int z0, x0;
  if (!strip_of(d, z0, x0)) return;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
  const int lane = threadIdx.x & (WAVE - 1);
  const int c = x0 - EHL + lane;
  const unsigned vc = 4u * (unsigned)min(max(c, 0), d.X - 1);
  const unsigned vs = 4u * (unsigned)max(c, 0);
  const bool out_col = lane >= EHL && lane < EHL + ESX && c <= d.X - 3;
  const bool cwin = c >= 2 && c <= d.X - 3;
  const int zmax = d.Z - 1, zw1 = d.Z - 3;
  const unsigned rowb = 4u * (unsigned)d.pitch, planeb = 4u * (unsigned)d.plane, fieldb = 4u * (unsigned)d.field;
  auto rb = [&](int k) -> unsigned { return (unsigned)min(max(k, 0), zmax) * rowb; };
  const rsrc_t rLi = rsrc(lin, (unsigned)NFIELD5 * fieldb);
  const rsrc_t rLo = rsrc(lout, (unsigned)NFIELD5 * fieldb);
  const rsrc_t rFv = rsrc(fv, (unsigned)NFIELD5 * fieldb);
  const rsrc_t rFs = rsrc(fsn, (unsigned)NFIELD5 * fieldb);
  const rsrc_t rC = rsrc(d.coef, (unsigned)NCOEF * planeb);
  const rsrc_t rD = rsrc(d.damp, planeb);
  const float c0 = d.c0, c1 = d.c1, dt = d.dt, dx = d.dx, dz = d.dz;
  const float dtdx = d.dt_dx, dtdz = d.dt_dz;
  const int end_row = min(z0 + d.tz, zw1 + 1);
  const bool jj = c >= 2 && c <= d.X - 3;
  const bool jm1 = c - 1 >= 2 && c - 1 <= d.X - 3;
  const bool first = d.fsurf && z0 == 2;
  const int grp = blockIdx.y * EWPB + wv;
  float *sl = slab + wv * (NGF * ETZA * WAVE) + lane;
  for (int q = 0; q < NGF * ETZA; ++q) sl[q * WAVE] = 0.0f;
  auto acc = [&](int f, int row, float v) { sl[(f * ETZA + (row - z0)) * WAVE] += v; };

  for (int si = 0; si < d.spw; ++si) {
    const int s = grp * d.spw + si;
    if (s >= d.S) break;
    const unsigned so = (unsigned)s * planeb;
    auto F = [&](int f) -> unsigned { return (unsigned)f * fieldb + so; };
    const int zs = d.src_z[s], xs = d.src_x[s];
    const float mxx = d.src_m[3 * s], mzz = d.src_m[3 * s + 1], mxz = d.src_m[3 * s + 2];
    float gmt = 0.0f;   // wavelet gradient of the moment-tensor injection, when a force term follows
    float vx0 = 0.f, vx1 = 0.f, vx2 = 0.f, vx3 = 0.f, vx4 = 0.f;     // forward velocities before the step
    float vz0 = 0.f, vz1 = 0.f, vz2 = 0.f, vz3 = 0.f, vz4 = 0.f;
    float xz0 = 0.f, xz1 = 0.f, xz2 = 0.f, xz3 = 0.f, xz4 = 0.f;     // forward stresses after the step
    float zz0 = 0.f, zz1 = 0.f, zz2 = 0.f, zz3 = 0.f, zz4 = 0.f;
    float xx0 = 0.f, xx1 = 0.f, xx2 = 0.f;
    float ux0 = 0.f, ux1 = 0.f, ux2 = 0.f, ux3 = 0.f, ux4 = 0.f;     // L1(vx) rows k-2..k+2
    float uz0 = 0.f, uz1 = 0.f, uz2 = 0.f, uz3 = 0.f, uz4 = 0.f;     // L1(vz)
    float e1a = 0.f, e1b = 0.f, e1c = 0.f, e3a = 0.f, e3b = 0.f, e3c = 0.f;
    float e20 = 0.f, e21 = 0.f, e22 = 0.f, e23 = 0.f, e24 = 0.f;
    float e40 = 0.f, e41 = 0.f, e42 = 0.f, e43 = 0.f, e44 = 0.f;
    float p1a = 0.f, p1b = 0.f, p1c = 0.f, q3a = 0.f, q3b = 0.f, q3c = 0.f;
    float p40 = 0.f, p41 = 0.f, p42 = 0.f, p43 = 0.f, p44 = 0.f;
    float q20 = 0.f, q21 = 0.f, q22 = 0.f, q23 = 0.f, q24 = 0.f;
    float p3a = 0.f, p3b = 0.f, p3c = 0.f, q1a = 0.f, q1b = 0.f, q1c = 0.f;
    float p20 = 0.f, p21 = 0.f, p22 = 0.f, p23 = 0.f, p24 = 0.f;
    float q40 = 0.f, q41 = 0.f, q42 = 0.f, q43 = 0.f, q44 = 0.f;
    float ltzz1 = 0.f, ltxz0 = 0.f, ltxz1 = 0.f;     // stress adjoints of rows 0, 1 (free surface)
    float wz0 = 0.f, wz1 = 0.f, wx1 = 0.f;           // damped velocity adjoints of rows 0, 1

    // row rr of: forward velocities, L1(v), and e (zero outside the grid / window)
    struct VRow {
      float fv[2], lv[2], b[2], dm;
    };
    auto vload = [&](int rr, VRow &R) {
      const unsigned o = rb(rr);
      R.fv[0] = bld(rFv, vc, F(VX) + o); R.fv[1] = bld(rFv, vc, F(VZ) + o);
      R.lv[0] = bld(rLi, vc, F(VX) + o); R.lv[1] = bld(rLi, vc, F(VZ) + o);
      R.b[0] = bld(rC, vc, BX * planeb + o); R.b[1] = bld(rC, vc, BZ * planeb + o);
      R.dm = bld(rD, vc, o);
    };
    auto vpush = [&](int rr, const VRow &R) {
      const bool in_grid = rr >= 0 && rr <= zmax;
      float lx = in_grid ? R.lv[0] * R.dm : 0.f, lz = in_grid ? R.lv[1] * R.dm : 0.f;   // 5'
      if (d.fsurf) {                                                                     // 4'
        if (rr == 0) { wz0 = lz; if (jj) lz = 0.f; }
        if (rr == 1) { wz1 = lz; wx1 = lx; if (jj) { lz = 0.f; lx = 0.f; } }
        if (rr == 2) {
          const float wx1m = shfl(wx1, lane - 1);
          if (jj) lz = lz + (((wz1 + wz0) - 2.0f * wx1) + (jm1 ? 2.0f * wx1m : 0.0f));
        }
        if (rr == 3 && jj) lx = lx + wx1;
      }
      float ea = 0.f, eb = 0.f, ec = 0.f, ed = 0.f;
      if (rr >= 2 && rr <= zw1 && cwin) {
        ea = lx * ((dt * R.b[0]) / dx);
        eb = lx * ((dt * R.b[0]) / dz);
        ec = lz * ((dt * R.b[1]) / dx);
        ed = lz * ((dt * R.b[1]) / dz);
      }
      vx0 = vx1; vx1 = vx2; vx2 = vx3; vx3 = vx4; vx4 = in_grid ? R.fv[0] : 0.f;
      vz0 = vz1; vz1 = vz2; vz2 = vz3; vz3 = vz4; vz4 = in_grid ? R.fv[1] : 0.f;
      ux0 = ux1; ux1 = ux2; ux2 = ux3; ux3 = ux4; ux4 = lx;
      uz0 = uz1; uz1 = uz2; uz2 = uz3; uz3 = uz4; uz4 = lz;
      e1a = e1b; e1b = e1c; e1c = ea;
      e20 = e21; e21 = e22; e22 = e23; e23 = e24; e24 = eb;
      e3a = e3b; e3b = e3c; e3c = ec;
      e40 = e41; e41 = e42; e42 = e43; e43 = e44; e44 = ed;
    };
    for (int rr = z0 - 4; rr <= z0 - 1; ++rr) {
      VRow R;
      vload(rr, R);
      vpush(rr, R);
    }
    struct SRow {
      float ls[3], cm[ANISO ? 6 : 4], fs[3];
    };
    auto sload = [&](int k, SRow &R) {
      const unsigned o = rb(k);
      R.ls[0] = bld(rLi, vc, F(TXX) + o); R.ls[1] = bld(rLi, vc, F(TZZ) + o); R.ls[2] = bld(rLi, vc, F(TXZ) + o);
      R.cm[0] = bld(rC, vc, C11 * planeb + o); R.cm[1] = bld(rC, vc, C13 * planeb + o);
      R.cm[2] = bld(rC, vc, C33 * planeb + o); R.cm[3] = bld(rC, vc, C55 * planeb + o);
      if constexpr (ANISO) {
        R.cm[4] = bld(rC, vc, C15 * planeb + o); R.cm[5] = bld(rC, vc, C35 * planeb + o);
      }
      R.fs[0] = bld(rFs, vc, F(TXX) + o); R.fs[1] = bld(rFs, vc, F(TZZ) + o); R.fs[2] = bld(rFs, vc, F(TXZ) + o);
    };
    VRow vcur;
    SRow scur;
    vload(z0, vcur);
    sload(z0 - 2, scur);
#pragma unroll 1
    for (int kk = 0; kk < d.tz + 4; ++kk) {
      const int k = z0 - 2 + kk;
      VRow vnxt;
      SRow snxt;
      vload(k + 3, vnxt);
      sload(k + 1, snxt);
      vpush(k + 2, vcur);
      // 3'. stress adjoints of row k from e
      const float e1m1 = shfl(e1a, lane - 1), e1m2 = shfl(e1a, lane - 2), e1p1 = shfl(e1a, lane + 1);
      const float e3p1 = shfl(e3a, lane + 1), e3m1 = shfl(e3a, lane - 1), e3p2 = shfl(e3a, lane + 2);
      float lt_xx = c0 * (e1m1 - e1a) + c1 * (e1m2 - e1p1);
      float lt_xz = (c0 * (e22 - e23) + c1 * (e21 - e24)) + (c0 * (e3a - e3p1) + c1 * (e3m1 - e3p2));
      float lt_zz = c0 * (e41 - e42) + c1 * (e40 - e43);
      // 2'. stress free surface transposed
      if (d.fsurf) {
        if (k == 0) ltxz0 = lt_xz;
        if (k == 1) { ltzz1 = lt_zz; ltxz1 = lt_xz; }
        if (k == 2) lt_xz = lt_xz - ltxz1;
        if (k == 3) { lt_zz = lt_zz - ltzz1; lt_xz = lt_xz - ltxz0; }
      }
      const float vxm1 = shfl(vx2, lane - 1), vxm2 = shfl(vx2, lane - 2), vxp1 = shfl(vx2, lane + 1);
      const float vzm1 = shfl(vz2, lane - 1), vzp1 = shfl(vz2, lane + 1), vzp2 = shfl(vz2, lane + 2);
      float vzm2 = 0.0f, vxp2 = 0.0f;
      if constexpr (ANISO) {
        vzm2 = shfl(vz2, lane - 2);
        vxp2 = shfl(vx2, lane + 2);
      }
      // 1'. stress row k
      float P1 = 0.f, Q2 = 0.f, Q3 = 0.f, P4 = 0.f, P2 = 0.f, P3 = 0.f, Q1 = 0.f, Q4 = 0.f;
      if (k >= 2 && k <= zw1 && cwin) {
        const unsigned o = rb(k);
        const float lxx = scur.ls[0] + lt_xx;
        const float lzz = (d.fsurf && k == 2) ? 0.0f : scur.ls[1] + lt_zz;
        const float lxz = scur.ls[2] + lt_xz;
        const bool own_k = out_col && k >= z0 && k < end_row;
        if (own_k && k == zs && c == xs && gsrc) {
          const float gm = (mxx * lxx + mzz * lzz) + mxz * lxz;
          if (d.src_fz) gmt = gm;        // the force term of row zs follows two rows later
          else gsrc[(long long)s * d.nt + t] = gm;
        }
        if (own_k) {
          bst(lxx, rLo, vs, F(TXX) + o); bst(lzz, rLo, vs, F(TZZ) + o); bst(lxz, rLo, vs, F(TXZ) + o);
        }
        const float axx_x = dtdx * lxx, axx_z = dtdz * lxx;
        const float azz_x = dtdx * lzz, azz_z = dtdz * lzz;
        const float axz_x = dtdx * lxz, axz_z = dtdz * lxz;
        const float C11v = scur.cm[0], C13v = scur.cm[1], C33v = scur.cm[2], C55v = scur.cm[3];
        P1 = C11v * axx_x + C13v * azz_x;
        Q2 = C13v * axx_z + C33v * azz_z;
        Q3 = C55v * axz_x;
        P4 = C55v * axz_z;
        const float dxb_vx = c0 * (vx2 - vxm1) + c1 * (vxp1 - vxm2);
        const float dzb_vz = c0 * (vz2 - vz1) + c1 * (vz3 - vz0);
        const float dxf_vz = c0 * (vzp1 - vz2) + c1 * (vzp2 - vzm1);
        const float dzf_vx = c0 * (vx3 - vx2) + c1 * (vx4 - vx1);
        if constexpr (ANISO) {
          const float C15v = scur.cm[4], C35v = scur.cm[5];
          P2 = C15v * axx_z + C35v * azz_z;
          P3 = C15v * axz_x;
          Q1 = C15v * axx_x + C35v * azz_x;
          Q4 = C35v * axz_z;
          if (own_k) {
            const float dxb_vz = c0 * (vz2 - vzm1) + c1 * (vzp1 - vzm2);
            const float dzb_vx = c0 * (vx2 - vx1) + c1 * (vx3 - vx0);
            const float dxf_vx = c0 * (vxp1 - vx2) + c1 * (vxp2 - vxm1);
            const float dzf_vz = c0 * (vz3 - vz2) + c1 * (vz4 - vz1);
            acc(6, k, axx_x * dxb_vz + axx_z * dzb_vx + axz_x * dxf_vx);
            acc(7, k, azz_x * dxb_vz + azz_z * dzb_vx + axz_z * dzf_vz);
          }
        }
        if (own_k) {
          acc(0, k, axx_x * dxb_vx);
          acc(1, k, axx_z * dzb_vz + azz_x * dxb_vx);
          acc(2, k, azz_z * dzb_vz);
          acc(3, k, axz_x * dxf_vz + axz_z * dzf_vx);
        }
      }
      p1a = p1b; p1b = p1c; p1c = P1;
      q3a = q3b; q3b = q3c; q3c = Q3;
      p40 = p41; p41 = p42; p42 = p43; p43 = p44; p44 = P4;
      q20 = q21; q21 = q22; q22 = q23; q23 = q24; q24 = Q2;
      if constexpr (ANISO) {
        p3a = p3b; p3b = p3c; p3c = P3;
        q1a = q1b; q1b = q1c; q1c = Q1;
        p20 = p21; p21 = p22; p22 = p23; p23 = p24; p24 = P2;
        q40 = q41; q41 = q42; q42 = q43; q43 = q44; q44 = Q4;
      }
      // forward stresses of row k after the step (stored with the free surface applied)
      {
        const bool in_grid = k >= 0 && k <= zmax;
        xx0 = xx1; xx1 = xx2; xx2 = in_grid ? scur.fs[0] : 0.f;
        zz0 = zz1; zz1 = zz2; zz2 = zz3; zz3 = zz4; zz4 = in_grid ? scur.fs[1] : 0.f;
        xz0 = xz1; xz1 = xz2; xz2 = xz3; xz3 = xz4; xz4 = in_grid ? scur.fs[2] : 0.f;
      }
      // velocity adjoints of row r = k-2: L1(v) + D^T(products)
      const int r = k - 2;
      const float p1m1 = shfl(p1a, lane - 1), p1p1 = shfl(p1a, lane + 1), p1p2 = shfl(p1a, lane + 2);
      const float q3m1 = shfl(q3a, lane - 1), q3m2 = shfl(q3a, lane - 2), q3p1 = shfl(q3a, lane + 1);
      float lvx = (c0 * (p1a - p1p1) + c1 * (p1m1 - p1p2)) + (c0 * (p41 - p42) + c1 * (p40 - p43));
      float lvz = (c0 * (q22 - q23) + c1 * (q21 - q24)) + (c0 * (q3m1 - q3a) + c1 * (q3m2 - q3p1));
      if constexpr (ANISO) {
        const float p3m1 = shfl(p3a, lane - 1), p3m2 = shfl(p3a, lane - 2), p3p1 = shfl(p3a, lane + 1);
        const float q1m1 = shfl(q1a, lane - 1), q1p1 = shfl(q1a, lane + 1), q1p2 = shfl(q1a, lane + 2);
        lvx = lvx + ((c0 * (p22 - p23) + c1 * (p21 - p24)) + (c0 * (p3m1 - p3a) + c1 * (p3m2 - p3p1)));
        lvz = lvz + ((c0 * (q1a - q1p1) + c1 * (q1m1 - q1p2)) + (c0 * (q41 - q42) + c1 * (q40 - q43)));
      }
      const float txp1 = shfl(xx0, lane + 1), txp2 = shfl(xx0, lane + 2), txm1 = shfl(xx0, lane - 1);
      const float tzm1 = shfl(xz2, lane - 1), tzm2 = shfl(xz2, lane - 2), tzp1 = shfl(xz2, lane + 1);
      const bool own_r = out_col && ((r >= z0 && r < end_row) || (first && r >= 0 && r < 2));
      if (own_r) {
        const unsigned o = rb(r);
        if (d.src_fz && gsrc && r == zs && c == xs)   // vertical force entered vz before the taper
          gsrc[(long long)s * d.nt + t] = gmt + d.src_fz[s] * uz0;
        bst(ux0 + lvx, rLo, vs, F(VX) + o);
        bst(uz0 + lvz, rLo, vs, F(VZ) + o);
        if (r >= z0) {
          const float dxf_txx = c0 * (txp1 - xx0) + c1 * (txp2 - txm1);
          const float dzb_txz = c0 * (xz2 - xz1) + c1 * (xz3 - xz0);
          const float dxb_txz = c0 * (xz2 - tzm1) + c1 * (tzp1 - tzm2);
          const float dzf_tzz = c0 * (zz3 - zz2) + c1 * (zz4 - zz1);
          acc(4, r, ux0 * (dt * (dxf_txx / dx + dzb_txz / dz)));
          acc(5, r, uz0 * (dt * (dxb_txz / dx + dzf_tzz / dz)));
        }
      }
      vcur = vnxt;
      scur = snxt;
    }
  }
  __syncthreads();
  if (!out_col) return;
  const long long planef = (long long)d.Z * d.pitch;
  float *g = gacc + (long long)blockIdx.y * NCOEF * planef + c;
  const float *sb = slab + lane;
  for (int f = wv; f < NGF; f += EWPB)
    for (int row = z0; row < end_row; ++row) {
      const int o = (f * ETZA + (row - z0)) * WAVE;
      const float v = (sb[o] + sb[NGF * ETZA * WAVE + o]) + (sb[2 * NGF * ETZA * WAVE + o] + sb[3 * NGF * ETZA * WAVE + o]);
      float *p = g + (long long)f * planef + (long long)row * d.pitch;
      *p = *p + v;
    }
}

// ---------------------------------------------------------------------------
// three-term (radius-3) PML scheme: fd_order 6, 8, 10
// ---------------------------------------------------------------------------
// Reference: step_forward_PML_6order (ADFWI/propagator/elastic_kernels.py:426-584) with the
// operators Dxfm_6 .. Dzbm_6 (:85-108).  With NN = fd_order / 2 the step is the 4th-order one
// with three-term differences c0 (a[i+1]-a[i]) + c1 (a[i+2]-a[i-1]) + c2 (a[i+3]-a[i-2]) (the
// first three coefficients of DiffCoef(NN)), on the window rows [NN, Z-NN-1] x cols
// [NN, X-NN-1], and the free-surface rows at h = NN+1: stress tzz[NN] = 0, tzz[NN-1] = -tzz[NN+1],
// txz[NN-1] = -txz[NN], txz[NN-2] = -txz[NN+1]; velocity (cols J = [NN, X-NN-1])
// vz[NN-1] = vz[NN], vx[NN-1] = vz[NN-1,j+1] - vz[NN-1,j] + vz[NN,j+1] - vz[NN,j] + vx[NN+1,j],
// vz[NN-2] = vz[NN-1].
//
// Lanes: a wavefront owns columns x0-6 .. x0+57 and writes x0 .. x0+50: stress is valid on
// lanes 3..59 (velocities 3 columns away, one more for the free-surface vx row), velocity on
// lanes 6..56.  Rows: at iteration k the velocity totals of row k+3 enter a 7-row window, the
// stress of row k is updated and the velocity of row k-3.  Strips are at least 8 rows, so that
// only the first strip reaches the free-surface rows.
constexpr int EHL3 = 6;
constexpr int ESX3 = 51;

__device__ __forceinline__ bool strip_of3(const ElDev &d, int &z0, int &x0) {
  const int n = (int)gridDim.x, b = (int)blockIdx.x;
  const int xcd = b & 7, per = n >> 3, rem = n & 7;
  const int strip = xcd * per + min(xcd, rem) + (b >> 3);
  if (strip >= d.nstrips) return false;
  const int sz = strip / d.nsx;
  z0 = d.nn + sz * d.tz;
  x0 = d.nn + (strip - sz * d.nsx) * ESX3;
  return true;
}

template <int N>
__device__ __forceinline__ void push(float (&w)[N], float v) {
#pragma unroll
  for (int i = 0; i < N - 1; ++i) w[i] = w[i + 1];
  w[N - 1] = v;
}

struct Fd3 {
  float c0, c1, c2;
  // x differences of a row value (every lane must take part in the shuffles)
  __device__ __forceinline__ float dxf(float v, int l) const {
    return c0 * (shfl(v, l + 1) - v) + c1 * (shfl(v, l + 2) - shfl(v, l - 1)) + c2 * (shfl(v, l + 3) - shfl(v, l - 2));
  }
  __device__ __forceinline__ float dxb(float v, int l) const {
    return c0 * (v - shfl(v, l - 1)) + c1 * (shfl(v, l + 1) - shfl(v, l - 2)) + c2 * (shfl(v, l + 2) - shfl(v, l - 3));
  }
  // z differences about window entry i
  template <int N>
  __device__ __forceinline__ float dzf(const float (&w)[N], int i) const {
    return c0 * (w[i + 1] - w[i]) + c1 * (w[i + 2] - w[i - 1]) + c2 * (w[i + 3] - w[i - 2]);
  }
  template <int N>
  __device__ __forceinline__ float dzb(const float (&w)[N], int i) const {
    return c0 * (w[i] - w[i - 1]) + c1 * (w[i + 1] - w[i - 2]) + c2 * (w[i + 2] - w[i - 3]);
  }
  // transposes
  __device__ __forceinline__ float dxfT(float e, int l) const {
    return c0 * (shfl(e, l - 1) - e) + c1 * (shfl(e, l - 2) - shfl(e, l + 1)) + c2 * (shfl(e, l - 3) - shfl(e, l + 2));
  }
  __device__ __forceinline__ float dxbT(float e, int l) const {
    return c0 * (e - shfl(e, l + 1)) + c1 * (shfl(e, l - 1) - shfl(e, l + 2)) + c2 * (shfl(e, l - 2) - shfl(e, l + 3));
  }
  template <int N>
  __device__ __forceinline__ float dzbT(const float (&w)[N], int i) const {
    return c0 * (w[i] - w[i + 1]) + c1 * (w[i - 1] - w[i + 2]) + c2 * (w[i - 2] - w[i + 3]);
  }
  template <int N>
  __device__ __forceinline__ float dzfT(const float (&w)[N], int i) const {
    return c0 * (w[i - 1] - w[i]) + c1 * (w[i - 2] - w[i + 1]) + c2 * (w[i - 3] - w[i + 2]);
  }
};

template <bool ANISO>
__global__ __launch_bounds__(ENT) void el3_fwd_step(ElDev d, const float *__restrict__ in, float *__restrict__ out,
                                                    int t) {
  int z0, x0;
  // PML factors 1 - dt/2 d, 1 + dt/2 d, 1/(1 + dt/2 d) of the rows the strip updates (z0-6 ..
  // z0+tz+2), staged once: read from LDS they cost no wait on the vector loads in flight
  constexpr int PZR = ETZ_MAX + 12;
  static_assert(3 * PZR <= ENT, "one PML factor per thread");
  __shared__ float pzs[3][PZR];
  if (!strip_of3(d, z0, x0)) return;   // uniform over the workgroup (one strip)
  if (threadIdx.x < 3 * PZR) {
    const int q = threadIdx.x / PZR, j = threadIdx.x - q * PZR, row = z0 - 6 + j;
    pzs[q][j] = (unsigned)row < (unsigned)d.Z ? d.pml_z[q * d.Z + row] : 0.0f;
  }
  __syncthreads();
  auto pzf = [&](int q, int row) -> float { return pzs[q][row - (z0 - 6)]; };
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
  const int lane = threadIdx.x & (WAVE - 1);
  const int nn = d.nn;
  const int c = x0 - EHL3 + lane;
  const unsigned vc = 4u * (unsigned)min(max(c, 0), d.X - 1);
  const unsigned vs = 4u * (unsigned)max(c, 0);
  const int zw1 = d.Z - nn - 1, xw1 = d.X - nn - 1;   // last window row / column
  const bool out_col = lane >= EHL3 && lane < EHL3 + ESX3 && c <= xw1;   // c >= nn by the strip origin
  const bool cwin = c >= nn && c <= xw1;
  const bool jj1 = c + 1 >= nn && c + 1 <= xw1;
  const int zmax = d.Z - 1;
  const unsigned rowb = 4u * (unsigned)d.pitch, planeb = 4u * (unsigned)d.plane, fieldb = 4u * (unsigned)d.field;
  auto rb = [&](int k) -> unsigned { return (unsigned)min(max(k, 0), zmax) * rowb; };
  const rsrc_t rIn = rsrc(in, (unsigned)NFIELD * fieldb);
  const rsrc_t rOut = rsrc(out, (unsigned)NFIELD * fieldb);
  const rsrc_t rC = rsrc(d.coef, (unsigned)NCOEF * planeb);
  const Fd3 D{d.c0, d.c1, d.c2};
  const int cc = min(max(c, 0), d.X - 1);
  const float pxn = d.pml_x[cc], pxd = d.pml_x[d.pitch + cc], pxi = d.pml_x[2 * d.pitch + cc];
  const int end_row = min(z0 + d.tz, zw1 + 1);
  const int nk = end_row - z0 + 6;

  const int s_first = (blockIdx.y * EWPB + wv) * d.spw;
  for (int si = 0; si < d.spw; ++si) {
    const int s = s_first + si;
    if (s >= d.S) break;
    const unsigned so = (unsigned)s * planeb;
    auto F = [&](int f) -> unsigned { return (unsigned)f * fieldb + so; };
    const int zs = d.src_z[s], xs = d.src_x[s];
    const float wsrc = d.src[(long long)s * d.nt + t];
    const float mxx = d.src_m[3 * s], mzz = d.src_m[3 * s + 1], mxz = d.src_m[3 * s + 2];
    // velocity totals rows k-3..k+3; stress totals txz/tzz rows k-6..k, txx rows k-3..k
    float vx[7] = {}, vz[7] = {}, xz[7] = {}, zz[7] = {}, xx[4] = {};
    float us[7][4] = {};                         // velocity splits of rows k-3 .. k+3 as loaded
    for (int rr = z0 - 6; rr <= z0 - 1; ++rr) {
      const unsigned o = rb(rr);
      const float a = bld(rIn, vc, F(VX_X) + o) + bld(rIn, vc, F(VX_Z) + o);
      const float b = bld(rIn, vc, F(VZ_X) + o) + bld(rIn, vc, F(VZ_Z) + o);
      const bool ok = rr >= 0;
      push(vx, ok ? a : 0.f);
      push(vz, ok ? b : 0.f);
    }
    // loaded one iteration ahead: velocity splits of row k+3, stress splits and moduli of row k
    struct Row {
      float v[4], s[6], cm[ANISO ? 6 : 4];
    };
    auto load = [&](int k, Row &R) {
      const unsigned o3 = rb(k + 3), o0 = rb(k);
#pragma unroll
      for (int q = 0; q < 4; ++q) R.v[q] = bld(rIn, vc, F(VX_X + q) + o3);
#pragma unroll
      for (int q = 0; q < 6; ++q) R.s[q] = bld(rIn, vc, F(TXX_X + q) + o0);
      R.cm[0] = bld(rC, vc, C11 * planeb + o0); R.cm[1] = bld(rC, vc, C13 * planeb + o0);
      R.cm[2] = bld(rC, vc, C33 * planeb + o0); R.cm[3] = bld(rC, vc, C55 * planeb + o0);
      if constexpr (ANISO) {
        R.cm[4] = bld(rC, vc, C15 * planeb + o0); R.cm[5] = bld(rC, vc, C35 * planeb + o0);
      }
    };
    Row cur;
    load(z0 - 3, cur);
#pragma unroll 1
    for (int kk = 0; kk < nk; ++kk) {
      const int k = z0 - 3 + kk;
      // buoyancies of row k-3, consumed at the end of the iteration (its old velocity splits
      // come from the register window us)
      float bb[2] = {};
      if (k - 3 >= z0 && k - 3 < end_row) {
        const unsigned om = rb(k - 3);
        bb[0] = bld(rC, vc, BX * planeb + om); bb[1] = bld(rC, vc, BZ * planeb + om);
      }
      Row nxt;
      load(k + 1, nxt);
      {
        const bool ok = k + 3 <= zmax;
#pragma unroll
        for (int e = 0; e < 6; ++e) {
#pragma unroll
          for (int q = 0; q < 4; ++q) us[e][q] = us[e + 1][q];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) us[6][q] = cur.v[q];
        push(vx, ok ? cur.v[0] + cur.v[1] : 0.f);
        push(vz, ok ? cur.v[2] + cur.v[3] : 0.f);
      }
      // free surface of the previous step's velocities: window entries 1..4 = rows NN-2..NN+1
      if (d.fsurf && k == nn) {
        const float vzn = shfl(vz[3], lane + 1);            // vz row NN, column j+1
        const float vz2m_n = jj1 ? vzn : 0.0f;              // new vz row NN-1 at column j+1
        const float vz2m = cwin ? vz[3] : vz[2];
        if (cwin) {
          vx[2] = (((vz2m_n - vz2m) + vzn) - vz[3]) + vx[4];
          vz[1] = vz2m;
        }
        vz[2] = vz2m;
      }
      // stress of row k (velocity row k = entry 3)
      float sxx_x = cur.s[0], sxx_z = cur.s[1], szz_x = cur.s[2], szz_z = cur.s[3];
      float sxz_x = cur.s[4], sxz_z = cur.s[5];
      const float dxb_vx = D.dxb(vx[3], lane), dxf_vz = D.dxf(vz[3], lane);
      float dxb_vz = 0.f, dxf_vx = 0.f;
      if constexpr (ANISO) {
        dxb_vz = D.dxb(vz[3], lane);
        dxf_vx = D.dxf(vx[3], lane);
      }
      if (k >= nn && k <= zw1 && cwin) {
        const float dzb_vz = D.dzb(vz, 3), dzf_vx = D.dzf(vx, 3);
        const float C11v = cur.cm[0], C13v = cur.cm[1], C33v = cur.cm[2], C55v = cur.cm[3];
        const float pzn = pzf(0, k), pzi = pzf(2, k);
        float sx_xx, sz_xx, sx_zz, sz_zz, sx_xz, sz_xz;
        if constexpr (ANISO) {
          const float dzb_vx = D.dzb(vx, 3), dzf_vz = D.dzf(vz, 3);
          const float C15v = cur.cm[4], C35v = cur.cm[5];
          sx_xx = C11v * dxb_vx + C15v * dxb_vz;
          sz_xx = C15v * dzb_vx + C13v * dzb_vz;
          sx_zz = C13v * dxb_vx + C35v * dxb_vz;
          sz_zz = C35v * dzb_vx + C33v * dzb_vz;
          sx_xz = C15v * dxf_vx + C55v * dxf_vz;
          sz_xz = C55v * dzf_vx + C35v * dzf_vz;
        } else {
          sx_xx = C11v * dxb_vx;
          sz_xx = C13v * dzb_vz;
          sx_zz = C13v * dxb_vx;
          sz_zz = C33v * dzb_vz;
          sx_xz = C55v * dxf_vz;
          sz_xz = C55v * dzf_vx;
        }
        sxx_x = (pxn * sxx_x + d.dt_dx * sx_xx) * pxi;
        sxx_z = (pzn * sxx_z + d.dt_dz * sz_xx) * pzi;
        szz_x = (pxn * szz_x + d.dt_dx * sx_zz) * pxi;
        szz_z = (pzn * szz_z + d.dt_dz * sz_zz) * pzi;
        sxz_x = (pxn * sxz_x + d.dt_dx * sx_xz) * pxi;
        sxz_z = (pzn * sxz_z + d.dt_dz * sz_xz) * pzi;
      }
      if (k == zs && c == xs) {
        sxx_x = sxx_x + mxx * wsrc; sxx_z = sxx_z + mxx * wsrc;
        szz_x = szz_x + mzz * wsrc; szz_z = szz_z + mzz * wsrc;
        sxz_x = sxz_x + mxz * wsrc; sxz_z = sxz_z + mxz * wsrc;
      }
      if (out_col && k >= z0 && k < end_row) {
        const unsigned o = rb(k);
        bst(sxx_x, rOut, vs, F(TXX_X) + o); bst(sxx_z, rOut, vs, F(TXX_Z) + o);
        bst(szz_x, rOut, vs, F(TZZ_X) + o); bst(szz_z, rOut, vs, F(TZZ_Z) + o);
        bst(sxz_x, rOut, vs, F(TXZ_X) + o); bst(sxz_z, rOut, vs, F(TXZ_Z) + o);
      }
      push(xx, sxx_x + sxx_z);
      push(xz, sxz_x + sxz_z);
      push(zz, szz_x + szz_z);
      // stress free surface once row NN+1 exists (entries 3..6 = rows NN-2..NN+1)
      if (d.fsurf && k == nn + 1) {
        zz[5] = 0.0f;       // tzz[NN] = 0
        zz[4] = -zz[6];     // tzz[NN-1] = -tzz[NN+1]
        xz[4] = -xz[5];     // txz[NN-1] = -txz[NN]
        xz[3] = -xz[6];     // txz[NN-2] = -txz[NN+1]
      }
      // velocity of row r = k-3 (txz/tzz rows r-3..r+3 = entries 0..6, txx row r = xx[0])
      const int r = k - 3;
      const float dxf_txx = D.dxf(xx[0], lane), dxb_txz = D.dxb(xz[3], lane);
      if (out_col && r >= z0 && r < end_row) {
        const unsigned o = rb(r);
        const float dzb_txz = D.dzb(xz, 3), dzf_tzz = D.dzf(zz, 3);
        const float bxv = bb[0], bzv = bb[1];
        const float pzn = pzf(0, r), pzd = pzf(1, r);
        const float nx_x = (pxn * us[0][0] + ((d.dt * bxv) * dxf_txx) / d.dx) / pxd;
        const float nx_z = (pzn * us[0][1] + ((d.dt * bxv) * dzb_txz) / d.dz) / pzd;
        float nz_x = (pxn * us[0][2] + ((d.dt * bzv) * dxb_txz) / d.dx) / pxd;
        float nz_z = (pzn * us[0][3] + ((d.dt * bzv) * dzf_tzz) / d.dz) / pzd;
        if (d.src_fz && r == zs && c == xs) {   // vertical force: half of f w into each vz split
          const float hf = 0.5f * (d.src_fz[s] * wsrc);
          nz_x = nz_x + hf;
          nz_z = nz_z + hf;
        }
        bst(nx_x, rOut, vs, F(VX_X) + o); bst(nx_z, rOut, vs, F(VX_Z) + o);
        bst(nz_x, rOut, vs, F(VZ_X) + o); bst(nz_z, rOut, vs, F(VZ_Z) + o);
      }
      cur = nxt;
    }
  }
}

// Adjoint of el3_fwd_step: the transposes listed for el_adj_step with the three-term
// operators and the free-surface rows moved to NN-2 .. NN+1:
//   stress free surface^T: Lt_zz[NN] = 0, Lt_zz[NN+1] -= Lt_zz[NN-1], Lt_xz[NN] -= Lt_xz[NN-1],
//   Lt_xz[NN+1] -= Lt_xz[NN-2];
//   velocity free surface^T (cols J): Lv_z[NN,j] += Lv_z[NN-1,j] + Lv_z[NN-2,j] - 2 Lv_x[NN-1,j]
//   + 2 Lv_x[NN-1,j-1], Lv_x[NN+1,j] += Lv_x[NN-1,j].
// Iteration k: forward velocity totals and e of row k+3 enter their windows, the stress
// adjoints of row k are formed, then the velocity adjoints of row k-3.  Strips of ETZA rows.
template <bool ANISO>
__global__ __launch_bounds__(ENT) void el3_adj_step(ElDev d, const float *__restrict__ lin, float *__restrict__ lout,
                                                    const float *__restrict__ fv, const float *__restrict__ fsn, int t,
                                                    float *__restrict__ gacc, float *__restrict__ gsrc) {
  constexpr int NGF = ANISO ? 8 : 6;
  __shared__ float slab[EWPB * NGF * ETZA * WAVE];
  // PML factors of the rows the strip reads (z0-6 .. z0+ETZA+5), staged once: read from LDS
  // they cost no wait on the vector loads in flight
  constexpr int PZR = ETZA + 12;
  __shared__ float pzs[3][PZR];
  int z0, x0;
  if (!strip_of3(d, z0, x0)) return;   // uniform over the workgroup (one strip)
  if (threadIdx.x < 3 * PZR) {
    const int q = threadIdx.x / PZR, j = threadIdx.x - q * PZR, row = z0 - 6 + j;
    pzs[q][j] = (unsigned)row < (unsigned)d.Z ? d.pml_z[q * d.Z + row] : 0.0f;
  }
  __syncthreads();
  auto pzf = [&](int q, int row) -> float { return pzs[q][row - (z0 - 6)]; };
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
  const int lane = threadIdx.x & (WAVE - 1);
  const int nn = d.nn;
  const int c = x0 - EHL3 + lane;
  const unsigned vc = 4u * (unsigned)min(max(c, 0), d.X - 1);
  const unsigned vs = 4u * (unsigned)max(c, 0);
  const int zw1 = d.Z - nn - 1, xw1 = d.X - nn - 1;
  const bool out_col = lane >= EHL3 && lane < EHL3 + ESX3 && c <= xw1;
  const bool cwin = c >= nn && c <= xw1;
  const bool jj1 = c + 1 >= nn && c + 1 <= xw1;
  const bool jm1 = c - 1 >= nn && c - 1 <= xw1;
  const int zmax = d.Z - 1;
  const unsigned rowb = 4u * (unsigned)d.pitch, planeb = 4u * (unsigned)d.plane, fieldb = 4u * (unsigned)d.field;
  auto rb = [&](int k) -> unsigned { return (unsigned)min(max(k, 0), zmax) * rowb; };
  const rsrc_t rLi = rsrc(lin, (unsigned)NFIELD * fieldb);
  const rsrc_t rLo = rsrc(lout, (unsigned)NFIELD * fieldb);
  const rsrc_t rFv = rsrc(fv, (unsigned)NFIELD * fieldb);
  const rsrc_t rFs = rsrc(fsn, (unsigned)NFIELD * fieldb);
  const rsrc_t rC = rsrc(d.coef, (unsigned)NCOEF * planeb);
  const Fd3 D{d.c0, d.c1, d.c2};
  const int cc = min(max(c, 0), d.X - 1);
  const float pxn = d.pml_x[cc], pxd = d.pml_x[d.pitch + cc], pxi = d.pml_x[2 * d.pitch + cc];
  const float sx_keep = pxn * pxi, sx_a = d.dt_dx * pxi;
  const float vx_keep = pxn / pxd, vx_w = d.dt / (d.dx * pxd);
  const int end_row = min(z0 + d.tz, zw1 + 1);
  const int nk = end_row - z0 + 6;
  const int grp = blockIdx.y * EWPB + wv;
  float *sl = slab + wv * (NGF * ETZA * WAVE) + lane;
  for (int q = 0; q < NGF * ETZA; ++q) sl[q * WAVE] = 0.0f;
  auto acc = [&](int f, int row, float v) { sl[(f * ETZA + (row - z0)) * WAVE] += v; };

  for (int si = 0; si < d.spw; ++si) {
    const int s = grp * d.spw + si;
    if (s >= d.S) break;
    const unsigned so = (unsigned)s * planeb;
    auto F = [&](int f) -> unsigned { return (unsigned)f * fieldb + so; };
    const int zs = d.src_z[s], xs = d.src_x[s];
    const float mxx = d.src_m[3 * s], mzz = d.src_m[3 * s + 1], mxz = d.src_m[3 * s + 2];
    float gmt = 0.0f;   // wavelet gradient of the moment-tensor injection, when a force term follows
    // forward velocity totals rows k-3..k+3; forward stress totals txz/tzz rows k-6..k, txx k-3..k
    float vx[7] = {}, vz[7] = {}, xz[7] = {}, zz[7] = {}, xx[4] = {};
    // e windows: e1 (vx_x), e3 (vz_x) rows k..k+3; e2 (vx_z), e4 (vz_z) rows k-3..k+3
    float e1[4] = {}, e3[4] = {}, e2[7] = {}, e4[7] = {};
    // stress-product windows: P1, Q3 (, P3, Q1) rows k-3..k; P4, Q2 (, P2, Q4) rows k-6..k
    float p1[4] = {}, q3[4] = {}, p4[7] = {}, q2[7] = {};
    float p3[ANISO ? 4 : 1] = {}, q1[ANISO ? 4 : 1] = {}, p2[ANISO ? 7 : 1] = {}, q4[ANISO ? 7 : 1] = {};
    // free-surface rows kept for the transposes
    float ltzz1 = 0.f, ltxz0 = 0.f, ltxz1 = 0.f, lvz0 = 0.f, lvz1 = 0.f, lvx1 = 0.f;

    auto push_row = [&](int rr, float fvx, float fvz, float lvxx, float lvxz, float lvzx, float lvzz, float bxv,
                        float bzv) {
      const bool in_grid = rr >= 0 && rr <= zmax;
      push(vx, in_grid ? fvx : 0.f);
      push(vz, in_grid ? fvz : 0.f);
      float ea = 0.f, eb = 0.f, ec = 0.f, ed = 0.f;
      if (rr >= nn && rr <= zw1 && cwin) {
        const float vz_w = d.dt / (d.dz * pzf(1, rr));
        ea = lvxx * (vx_w * bxv);
        eb = lvxz * (vz_w * bxv);
        ec = lvzx * (vx_w * bzv);
        ed = lvzz * (vz_w * bzv);
      }
      push(e1, ea);
      push(e2, eb);
      push(e3, ec);
      push(e4, ed);
    };
    for (int rr = z0 - 6; rr <= z0 - 1; ++rr) {
      const unsigned o = rb(rr);
      push_row(rr, bld(rFv, vc, F(VX_X) + o) + bld(rFv, vc, F(VX_Z) + o),
               bld(rFv, vc, F(VZ_X) + o) + bld(rFv, vc, F(VZ_Z) + o), bld(rLi, vc, F(VX_X) + o),
               bld(rLi, vc, F(VX_Z) + o), bld(rLi, vc, F(VZ_X) + o), bld(rLi, vc, F(VZ_Z) + o),
               bld(rC, vc, BX * planeb + o), bld(rC, vc, BZ * planeb + o));
    }
    // loaded one iteration ahead: row k+3 of the forward velocity splits, the velocity
    // adjoints and the buoyancies; row k of the stress adjoints and the moduli.  The forward
    // stress splits of row k and the velocity adjoints of row k-3, consumed late in the
    // iteration, are loaded at its start (fewer live registers: three waves per SIMD)
    struct Row {
      float fv[4], lv[4], b[2], ls[6], cm[ANISO ? 6 : 4];
    };
    auto load = [&](int k, Row &R) {
      const unsigned o3 = rb(k + 3), o0 = rb(k);
#pragma unroll
      for (int q = 0; q < 4; ++q) R.fv[q] = bld(rFv, vc, F(VX_X + q) + o3);
#pragma unroll
      for (int q = 0; q < 4; ++q) R.lv[q] = bld(rLi, vc, F(VX_X + q) + o3);
      R.b[0] = bld(rC, vc, BX * planeb + o3); R.b[1] = bld(rC, vc, BZ * planeb + o3);
#pragma unroll
      for (int q = 0; q < 6; ++q) R.ls[q] = bld(rLi, vc, F(TXX_X + q) + o0);
      R.cm[0] = bld(rC, vc, C11 * planeb + o0); R.cm[1] = bld(rC, vc, C13 * planeb + o0);
      R.cm[2] = bld(rC, vc, C33 * planeb + o0); R.cm[3] = bld(rC, vc, C55 * planeb + o0);
      if constexpr (ANISO) {
        R.cm[4] = bld(rC, vc, C15 * planeb + o0); R.cm[5] = bld(rC, vc, C35 * planeb + o0);
      }
    };
    Row cur;
    load(z0 - 3, cur);
#pragma unroll 1
    for (int kk = 0; kk < nk; ++kk) {
      const int k = z0 - 3 + kk;
      float fs[6], lr[4] = {};
      {
        const unsigned o0 = rb(k), om = rb(k - 3);
#pragma unroll
        for (int q = 0; q < 6; ++q) fs[q] = bld(rFs, vc, F(TXX_X + q) + o0);
        if (k - 3 >= z0 && k - 3 < end_row) {
#pragma unroll
          for (int q = 0; q < 4; ++q) lr[q] = bld(rLi, vc, F(VX_X + q) + om);
        }
      }
      Row nxt;
      load(k + 1, nxt);
      push_row(k + 3, cur.fv[0] + cur.fv[1], cur.fv[2] + cur.fv[3], cur.lv[0], cur.lv[1], cur.lv[2], cur.lv[3],
               cur.b[0], cur.b[1]);
      if (d.fsurf && k == nn) {   // forward velocity free surface (entries 1..4 = rows NN-2..NN+1)
        const float vzn = shfl(vz[3], lane + 1);
        const float vz2m_n = jj1 ? vzn : 0.0f;
        const float vz2m = cwin ? vz[3] : vz[2];
        if (cwin) {
          vx[2] = (((vz2m_n - vz2m) + vzn) - vz[3]) + vx[4];
          vz[1] = vz2m;
        }
        vz[2] = vz2m;
      }
      // a. stress-total adjoints of row k (e1/e3 row k = entry 0; e2/e4 row k = entry 3)
      float lt_xx = D.dxfT(e1[0], lane);
      float lt_xz = D.dzbT(e2, 3) + D.dxbT(e3[0], lane);
      float lt_zz = D.dzfT(e4, 3);
      // b. stress free surface transposed
      if (d.fsurf) {
        if (k == nn - 2) ltxz0 = lt_xz;
        if (k == nn - 1) { ltzz1 = lt_zz; ltxz1 = lt_xz; }
        if (k == nn) { lt_zz = 0.0f; lt_xz = lt_xz - ltxz1; }
        if (k == nn + 1) { lt_zz = lt_zz - ltzz1; lt_xz = lt_xz - ltxz0; }
      }
      // forward velocity derivatives of row k (for the C gradients)
      const float dxb_vx = D.dxb(vx[3], lane), dxf_vz = D.dxf(vz[3], lane);
      float dxb_vz = 0.f, dxf_vx = 0.f;
      if constexpr (ANISO) {
        dxb_vz = D.dxb(vz[3], lane);
        dxf_vx = D.dxf(vx[3], lane);
      }
      // c, d. stress splits of row k
      float P1 = 0.f, Q2 = 0.f, Q3 = 0.f, P4 = 0.f, P2 = 0.f, P3 = 0.f, Q1 = 0.f, Q4 = 0.f;
      if (k >= nn && k <= zw1 && cwin) {
        const unsigned o = rb(k);
        const float lxx_x = cur.ls[0] + lt_xx, lxx_z = cur.ls[1] + lt_xx;
        const float lzz_x = cur.ls[2] + lt_zz, lzz_z = cur.ls[3] + lt_zz;
        const float lxz_x = cur.ls[4] + lt_xz, lxz_z = cur.ls[5] + lt_xz;
        const bool own_k = out_col && k >= z0 && k < end_row;
        if (own_k && k == zs && c == xs && gsrc) {
          const float gm = (mxx * (lxx_x + lxx_z) + mzz * (lzz_x + lzz_z)) + mxz * (lxz_x + lxz_z);
          if (d.src_fz) gmt = gm;        // the force term of row zs follows three rows later
          else gsrc[(long long)s * d.nt + t] = gm;
        }
        const float pzn = pzf(0, k), pzi = pzf(2, k);
        const float sz_keep = pzn * pzi, sz_a = d.dt_dz * pzi;
        if (own_k) {
          bst(sx_keep * lxx_x, rLo, vs, F(TXX_X) + o); bst(sz_keep * lxx_z, rLo, vs, F(TXX_Z) + o);
          bst(sx_keep * lzz_x, rLo, vs, F(TZZ_X) + o); bst(sz_keep * lzz_z, rLo, vs, F(TZZ_Z) + o);
          bst(sx_keep * lxz_x, rLo, vs, F(TXZ_X) + o); bst(sz_keep * lxz_z, rLo, vs, F(TXZ_Z) + o);
        }
        const float axx_x = sx_a * lxx_x, axx_z = sz_a * lxx_z;
        const float azz_x = sx_a * lzz_x, azz_z = sz_a * lzz_z;
        const float axz_x = sx_a * lxz_x, axz_z = sz_a * lxz_z;
        const float C11v = cur.cm[0], C13v = cur.cm[1], C33v = cur.cm[2], C55v = cur.cm[3];
        P1 = C11v * axx_x + C13v * azz_x;
        Q2 = C13v * axx_z + C33v * azz_z;
        Q3 = C55v * axz_x;
        P4 = C55v * axz_z;
        if constexpr (ANISO) {
          const float C15v = cur.cm[4], C35v = cur.cm[5];
          P2 = C15v * axx_z + C35v * azz_z;
          P3 = C15v * axz_x;
          Q1 = C15v * axx_x + C35v * azz_x;
          Q4 = C35v * axz_z;
        }
        if (own_k) {
          const float dzb_vz = D.dzb(vz, 3), dzf_vx = D.dzf(vx, 3);
          if constexpr (ANISO) {
            const float dzb_vx = D.dzb(vx, 3), dzf_vz = D.dzf(vz, 3);
            acc(6, k, axx_x * dxb_vz + axx_z * dzb_vx + axz_x * dxf_vx);
            acc(7, k, azz_x * dxb_vz + azz_z * dzb_vx + axz_z * dzf_vz);
          }
          acc(0, k, axx_x * dxb_vx);
          acc(1, k, axx_z * dzb_vz + azz_x * dxb_vx);
          acc(2, k, azz_z * dzb_vz);
          acc(3, k, axz_x * dxf_vz + axz_z * dzf_vx);
        }
      }
      push(p1, P1);
      push(q3, Q3);
      push(p4, P4);
      push(q2, Q2);
      if constexpr (ANISO) {
        push(p3, P3);
        push(q1, Q1);
        push(p2, P2);
        push(q4, Q4);
      }
      // forward stress totals of row k (after the step), free surface once row NN+1 exists
      {
        const bool in_grid = k >= 0 && k <= zmax;
        push(xx, in_grid ? fs[0] + fs[1] : 0.f);
        push(zz, in_grid ? fs[2] + fs[3] : 0.f);
        push(xz, in_grid ? fs[4] + fs[5] : 0.f);
      }
      if (d.fsurf && k == nn + 1) {
        zz[5] = 0.0f; zz[4] = -zz[6]; xz[4] = -xz[5]; xz[3] = -xz[6];
      }
      // e. velocity-total adjoints of row r = k-3 (P1/Q3 row r = entry 0; P4/Q2 row r = entry 3)
      const int r = k - 3;
      float lvx = D.dxbT(p1[0], lane) + D.dzfT(p4, 3);
      float lvz = D.dzbT(q2, 3) + D.dxfT(q3[0], lane);
      if constexpr (ANISO) {
        lvx = lvx + (D.dzbT(p2, 3) + D.dxfT(p3[0], lane));
        lvz = lvz + (D.dxbT(q1[0], lane) + D.dzfT(q4, 3));
      }
      if (d.fsurf) {
        if (r == nn - 2) lvz0 = lvz;
        if (r == nn - 1) { lvz1 = lvz; lvx1 = lvx; }
        if (r == nn) {
          const float lvx1m = shfl(lvx1, lane - 1);
          if (cwin) lvz = lvz + (((lvz1 + lvz0) - 2.0f * lvx1) + (jm1 ? 2.0f * lvx1m : 0.0f));
        }
        if (r == nn + 1 && cwin) lvx = lvx + lvx1;
      }
      // forward stress derivatives of row r (for the b gradients)
      const float dxf_txx = D.dxf(xx[0], lane), dxb_txz = D.dxb(xz[3], lane);
      if (out_col && r >= z0 && r < end_row) {
        const unsigned o = rb(r);
        const float pzn = pzf(0, r), pzd = pzf(1, r);
        const float vz_keep = pzn / pzd, vz_w = d.dt / (d.dz * pzd);
        const float l1 = lr[0], l2 = lr[1], l3 = lr[2], l4 = lr[3];
        if (d.src_fz && gsrc && r == zs && c == xs)
          gsrc[(long long)s * d.nt + t] = gmt + 0.5f * (d.src_fz[s] * (l3 + l4));
        bst(vx_keep * l1 + lvx, rLo, vs, F(VX_X) + o); bst(vz_keep * l2 + lvx, rLo, vs, F(VX_Z) + o);
        bst(vx_keep * l3 + lvz, rLo, vs, F(VZ_X) + o); bst(vz_keep * l4 + lvz, rLo, vs, F(VZ_Z) + o);
        const float dzb_txz = D.dzb(xz, 3), dzf_tzz = D.dzf(zz, 3);
        acc(4, r, l1 * (vx_w * dxf_txx) + l2 * (vz_w * dzb_txz));
        acc(5, r, l3 * (vx_w * dxb_txz) + l4 * (vz_w * dzf_tzz));
      }
      cur = nxt;
    }
  }
  __syncthreads();
  if (!out_col) return;
  const long long planef = (long long)d.Z * d.pitch;
  float *g = gacc + (long long)blockIdx.y * NCOEF * planef + c;
  const float *sb = slab + lane;
  for (int f = wv; f < NGF; f += EWPB)
    for (int row = z0; row < end_row; ++row) {
      const int o = (f * ETZA + (row - z0)) * WAVE;
      const float v = (sb[o] + sb[NGF * ETZA * WAVE + o]) + (sb[2 * NGF * ETZA * WAVE + o] + sb[3 * NGF * ETZA * WAVE + o]);
      float *p = g + (long long)f * planef + (long long)row * d.pitch;
      *p = *p + v;
    }
}

// Three-term damping scheme (reference: step_forward_ABL_6order, elastic_kernels.py:781-912):
// the stages of el_abl_fwd with the three-term differences, the window [NN, Z-NN-1] x
// [NN, X-NN-1] and the free-surface rows kept in the state: tzz[NN] = 0, tzz[NN-1] = -tzz[NN+1],
// txz[NN-1] = -txz[NN], txz[NN-2] = -txz[NN+1]; on the columns J, from the new undamped rows NN
// and NN+1, vz[NN-1] = vz[NN], vx[NN-1] = vz[NN-1,j+1] - vz[NN-1,j] + vz[NN,j+1] - vz[NN,j] +
// vx[NN+1,j], vz[NN-2] = vz[NN-1]; then vx, vz *= damp.  Lanes and rows march as in
// el3_fwd_step; the first strip also writes the free-surface rows NN-2 and NN-1.
template <bool ANISO>
__global__ __launch_bounds__(ENT) void el3_abl_fwd(ElDev d, const float *__restrict__ in, float *__restrict__ out,
                                                   int t) {
  int z0, x0;
  if (!strip_of3(d, z0, x0)) return;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
  const int lane = threadIdx.x & (WAVE - 1);
  const int nn = d.nn;
  const int c = x0 - EHL3 + lane;
  const unsigned vc = 4u * (unsigned)min(max(c, 0), d.X - 1);
  const unsigned vs = 4u * (unsigned)max(c, 0);
  const int zw1 = d.Z - nn - 1, xw1 = d.X - nn - 1;
  const bool out_col = lane >= EHL3 && lane < EHL3 + ESX3 && c <= xw1;
  const bool cwin = c >= nn && c <= xw1;
  const bool jj1 = c + 1 >= nn && c + 1 <= xw1;
  const int zmax = d.Z - 1;
  const unsigned rowb = 4u * (unsigned)d.pitch, planeb = 4u * (unsigned)d.plane, fieldb = 4u * (unsigned)d.field;
  auto rb = [&](int k) -> unsigned { return (unsigned)min(max(k, 0), zmax) * rowb; };
  const rsrc_t rIn = rsrc(in, (unsigned)NFIELD5 * fieldb);
  const rsrc_t rOut = rsrc(out, (unsigned)NFIELD5 * fieldb);
  const rsrc_t rC = rsrc(d.coef, (unsigned)NCOEF * planeb);
  const rsrc_t rD = rsrc(d.damp, planeb);
  const Fd3 D{d.c0, d.c1, d.c2};
  const float dt = d.dt, dx = d.dx, dz = d.dz;
  const int end_row = min(z0 + d.tz, zw1 + 1);
  const int nk = end_row - z0 + 6;
  const bool first = d.fsurf && z0 == nn;         // the strip that writes the free-surface rows

  const int s_first = (blockIdx.y * EWPB + wv) * d.spw;
  for (int si = 0; si < d.spw; ++si) {
    const int s = s_first + si;
    if (s >= d.S) break;
    const unsigned so = (unsigned)s * planeb;
    auto F = [&](int f) -> unsigned { return (unsigned)f * fieldb + so; };
    const int zs = d.src_z[s], xs = d.src_x[s];
    const float wsrc = d.src[(long long)s * d.nt + t];
    const float mxx = d.src_m[3 * s], mzz = d.src_m[3 * s + 1], mxz = d.src_m[3 * s + 2];
    // velocities (state before the step) rows k-3..k+3; new stresses txz/tzz rows k-6..k, txx k-3..k
    float vx[7] = {}, vz[7] = {}, xz[7] = {}, zz[7] = {}, xx[4] = {};
    float nvzh = 0.f;                            // new vz of row NN before damping (free surface)
    for (int rr = z0 - 6; rr <= z0 - 1; ++rr) {
      const unsigned o = rb(rr);
      const bool ok = rr >= 0;
      const float a = bld(rIn, vc, F(VX) + o), b = bld(rIn, vc, F(VZ) + o);
      push(vx, ok ? a : 0.f);
      push(vz, ok ? b : 0.f);
    }
    // loaded one iteration ahead: velocities of row k+3, stresses and moduli of row k
    struct Row {
      float v[2], s[3], cm[ANISO ? 6 : 4];
    };
    auto load = [&](int k, Row &R) {
      const unsigned o3 = rb(k + 3), o0 = rb(k);
      R.v[0] = bld(rIn, vc, F(VX) + o3); R.v[1] = bld(rIn, vc, F(VZ) + o3);
      R.s[0] = bld(rIn, vc, F(TXX) + o0); R.s[1] = bld(rIn, vc, F(TZZ) + o0); R.s[2] = bld(rIn, vc, F(TXZ) + o0);
      R.cm[0] = bld(rC, vc, C11 * planeb + o0); R.cm[1] = bld(rC, vc, C13 * planeb + o0);
      R.cm[2] = bld(rC, vc, C33 * planeb + o0); R.cm[3] = bld(rC, vc, C55 * planeb + o0);
      if constexpr (ANISO) {
        R.cm[4] = bld(rC, vc, C15 * planeb + o0); R.cm[5] = bld(rC, vc, C35 * planeb + o0);
      }
    };
    Row cur;
    load(z0 - 3, cur);
#pragma unroll 1
    for (int kk = 0; kk < nk; ++kk) {
      const int k = z0 - 3 + kk;
      // buoyancies and taper of row k-3, consumed at the end of the iteration
      float bxr = 0.f, bzr = 0.f, dmr = 0.f;
      {
        const unsigned om = rb(k - 3);
        bxr = bld(rC, vc, BX * planeb + om); bzr = bld(rC, vc, BZ * planeb + om);
        dmr = bld(rD, vc, om);
      }
      Row nxt;
      load(k + 1, nxt);
      {
        const bool ok = k + 3 <= zmax;
        push(vx, ok ? cur.v[0] : 0.f);
        push(vz, ok ? cur.v[1] : 0.f);
      }
      // 1. stress of row k (velocity row k = entry 3)
      float sxx = cur.s[0], szz = cur.s[1], sxz = cur.s[2];
      const float dxb_vx = D.dxb(vx[3], lane), dxf_vz = D.dxf(vz[3], lane);
      float dxb_vz = 0.f, dxf_vx = 0.f;
      if constexpr (ANISO) {
        dxb_vz = D.dxb(vz[3], lane);
        dxf_vx = D.dxf(vx[3], lane);
      }
      if (k >= nn && k <= zw1 && cwin) {
        const float dzb_vz = D.dzb(vz, 3), dzf_vx = D.dzf(vx, 3);
        const float C11v = cur.cm[0], C13v = cur.cm[1], C33v = cur.cm[2], C55v = cur.cm[3];
        if constexpr (ANISO) {
          const float dzb_vx = D.dzb(vx, 3), dzf_vz = D.dzf(vz, 3);
          const float C15v = cur.cm[4], C35v = cur.cm[5];
          sxx = sxx + dt * ((C11v * dxb_vx + C15v * dxb_vz) / dx + (C15v * dzb_vx + C13v * dzb_vz) / dz);
          szz = szz + dt * ((C13v * dxb_vx + C35v * dxb_vz) / dx + (C35v * dzb_vx + C33v * dzb_vz) / dz);
          sxz = sxz + dt * ((C15v * dxf_vx + C55v * dxf_vz) / dx + (C55v * dzf_vx + C35v * dzf_vz) / dz);
        } else {
          sxx = sxx + dt * ((C11v * dxb_vx) / dx + (C13v * dzb_vz) / dz);
          szz = szz + dt * ((C13v * dxb_vx) / dx + (C33v * dzb_vz) / dz);
          sxz = sxz + dt * ((C55v * dxf_vz) / dx + (C55v * dzf_vx) / dz);
        }
      }
      if (k == zs && c == xs) {
        sxx = sxx + mxx * wsrc;
        szz = szz + mzz * wsrc;
        sxz = sxz + mxz * wsrc;
      }
      if (d.fsurf && k == nn) szz = 0.0f;        // 2. tzz[NN] = 0
      if (out_col && k >= z0 && k < end_row) {
        const unsigned o = rb(k);
        bst(sxx, rOut, vs, F(TXX) + o); bst(szz, rOut, vs, F(TZZ) + o); bst(sxz, rOut, vs, F(TXZ) + o);
      }
      push(xx, sxx);
      push(xz, sxz);
      push(zz, szz);
      if (d.fsurf && k == nn + 1) {              // entries 3..6 = rows NN-2..NN+1
        zz[4] = -zz[6];
        xz[4] = -xz[5];
        xz[3] = -xz[6];
        if (out_col) {
          bst(zz[4], rOut, vs, F(TZZ) + rb(nn - 1));
          bst(xz[4], rOut, vs, F(TXZ) + rb(nn - 1));
          bst(xz[3], rOut, vs, F(TXZ) + rb(nn - 2));
        }
      }
      // 3. velocity of row r = k-3 (old values: window entry 0)
      const int r = k - 3;
      const float dxf_txx = D.dxf(xx[0], lane), dxb_txz = D.dxb(xz[3], lane);
      float nvx = 0.0f, nvz = 0.0f;
      if (r >= nn && r <= zw1 && cwin) {
        const float dzb_txz = D.dzb(xz, 3), dzf_tzz = D.dzf(zz, 3);
        nvx = vx[0] + (dt * bxr) * (dxf_txx / dx + dzb_txz / dz);
        nvz = vz[0] + (dt * bzr) * (dxb_txz / dx + dzf_tzz / dz);
      }
      if (d.src_fz && r == zs && c == xs) nvz = nvz + d.src_fz[s] * wsrc;   // vertical force
      if (out_col && r >= z0 && r < end_row) {
        const unsigned o = rb(r);
        bst(nvx * dmr, rOut, vs, F(VX) + o);
        bst(nvz * dmr, rOut, vs, F(VZ) + o);
      }
      // 4. free-surface velocity rows from the new rows NN and NN+1
      if (first && r == nn) nvzh = nvz;
      if (first && r == nn + 1) {
        const float vzn = shfl(nvzh, lane + 1);
        const float vz1m_n = jj1 ? vzn : 0.0f;
        const float vz1m = cwin ? nvzh : 0.0f;
        if (out_col) {
          const float d1 = bld(rD, vc, rb(nn - 1)), d0 = bld(rD, vc, rb(nn - 2));
          const float nvx1 = (((vz1m_n - vz1m) + vzn) - nvzh) + nvx;
          bst(nvx1 * d1, rOut, vs, F(VX) + rb(nn - 1));
          bst(vz1m * d1, rOut, vs, F(VZ) + rb(nn - 1));
          bst(vz1m * d0, rOut, vs, F(VZ) + rb(nn - 2));
        }
      }
      cur = nxt;
    }
  }
}

// Adjoint of el3_abl_fwd: the stages of el_abl_adj with the three-term transposes and the
// free-surface rows NN-2 .. NN+1:
//   5'. L2(v) = damp * L(v);
//   4'. on J: Lvz[NN] += L2vz[NN-1] + L2vz[NN-2] - 2 L2vx[NN-1,j] + 2 L2vx[NN-1,j-1],
//       Lvx[NN+1] += L2vx[NN-1]; the overwritten rows NN-2, NN-1 get zero;
//   3'. e = L1(v) dt b / d feeds the stress adjoints; b gradients;
//   2'. stress free surface transposed; the stored tzz[NN] = 0 is a constant, so L1(tzz)[NN] = 0;
//   1'. L(s before) = L1(s); L(v before) = L1(v) + D^T(C dt/d L1(s)); C and source gradients.
// The first strip also writes the velocity adjoints of the free-surface rows NN-2 and NN-1
// (state in this scheme).  Strips of ETZA rows.
template <bool ANISO>
__global__ __launch_bounds__(ENT) void el3_abl_adj(ElDev d, const float *__restrict__ lin, float *__restrict__ lout,
                                                   const float *__restrict__ fv, const float *__restrict__ fsn, int t,
                                                   float *__restrict__ gacc, float *__restrict__ gsrc) {
  constexpr int NGF = ANISO ? 8 : 6;
  __shared__ float slab[EWPB * NGF * ETZA * WAVE];
  int z0, x0;
  if (!strip_of3(d, z0, x0)) return;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
  const int lane = threadIdx.x & (WAVE - 1);
  const int nn = d.nn;
  const int c = x0 - EHL3 + lane;
  const unsigned vc = 4u * (unsigned)min(max(c, 0), d.X - 1);
  const unsigned vs = 4u * (unsigned)max(c, 0);
  const int zw1 = d.Z - nn - 1, xw1 = d.X - nn - 1;
  const bool out_col = lane >= EHL3 && lane < EHL3 + ESX3 && c <= xw1;
  const bool cwin = c >= nn && c <= xw1;
  const bool jm1 = c - 1 >= nn && c - 1 <= xw1;
  const int zmax = d.Z - 1;
  const unsigned rowb = 4u * (unsigned)d.pitch, planeb = 4u * (unsigned)d.plane, fieldb = 4u * (unsigned)d.field;
  auto rb = [&](int k) -> unsigned { return (unsigned)min(max(k, 0), zmax) * rowb; };
  const rsrc_t rLi = rsrc(lin, (unsigned)NFIELD5 * fieldb);
  const rsrc_t rLo = rsrc(lout, (unsigned)NFIELD5 * fieldb);
  const rsrc_t rFv = rsrc(fv, (unsigned)NFIELD5 * fieldb);
  const rsrc_t rFs = rsrc(fsn, (unsigned)NFIELD5 * fieldb);
  const rsrc_t rC = rsrc(d.coef, (unsigned)NCOEF * planeb);
  const rsrc_t rD = rsrc(d.damp, planeb);
  const Fd3 D{d.c0, d.c1, d.c2};
  const float dt = d.dt, dx = d.dx, dz = d.dz;
  const float dtdx = d.dt_dx, dtdz = d.dt_dz;
  const int end_row = min(z0 + d.tz, zw1 + 1);
  const int nk = end_row - z0 + 6;
  const bool first = d.fsurf && z0 == nn;
  const int grp = blockIdx.y * EWPB + wv;
  float *sl = slab + wv * (NGF * ETZA * WAVE) + lane;
  for (int q = 0; q < NGF * ETZA; ++q) sl[q * WAVE] = 0.0f;
  auto acc = [&](int f, int row, float v) { sl[(f * ETZA + (row - z0)) * WAVE] += v; };

  for (int si = 0; si < d.spw; ++si) {
    const int s = grp * d.spw + si;
    if (s >= d.S) break;
    const unsigned so = (unsigned)s * planeb;
    auto F = [&](int f) -> unsigned { return (unsigned)f * fieldb + so; };
    const int zs = d.src_z[s], xs = d.src_x[s];
    const float mxx = d.src_m[3 * s], mzz = d.src_m[3 * s + 1], mxz = d.src_m[3 * s + 2];
    float gmt = 0.0f;   // wavelet gradient of the moment-tensor injection, when a force term follows
    // forward velocities before the step rows k-3..k+3; forward stresses after the step
    // txz/tzz rows k-6..k, txx rows k-3..k; L1(v) rows k-3..k+3
    float vx[7] = {}, vz[7] = {}, xz[7] = {}, zz[7] = {}, xx[4] = {}, ux[7] = {}, uz[7] = {};
    float e1[4] = {}, e3[4] = {}, e2[7] = {}, e4[7] = {};
    float p1[4] = {}, q3[4] = {}, p4[7] = {}, q2[7] = {};
    float p3[ANISO ? 4 : 1] = {}, q1[ANISO ? 4 : 1] = {}, p2[ANISO ? 7 : 1] = {}, q4[ANISO ? 7 : 1] = {};
    float ltzz1 = 0.f, ltxz0 = 0.f, ltxz1 = 0.f;     // stress adjoints of rows NN-2, NN-1
    float wz0 = 0.f, wz1 = 0.f, wx1 = 0.f;           // damped velocity adjoints of rows NN-2, NN-1

    struct VRow {
      float fv[2], lv[2], b[2], dm;
    };
    auto vload = [&](int rr, VRow &R) {
      const unsigned o = rb(rr);
      R.fv[0] = bld(rFv, vc, F(VX) + o); R.fv[1] = bld(rFv, vc, F(VZ) + o);
      R.lv[0] = bld(rLi, vc, F(VX) + o); R.lv[1] = bld(rLi, vc, F(VZ) + o);
      R.b[0] = bld(rC, vc, BX * planeb + o); R.b[1] = bld(rC, vc, BZ * planeb + o);
      R.dm = bld(rD, vc, o);
    };
    auto vpush = [&](int rr, const VRow &R) {
      const bool in_grid = rr >= 0 && rr <= zmax;
      float lx = in_grid ? R.lv[0] * R.dm : 0.f, lz = in_grid ? R.lv[1] * R.dm : 0.f;   // 5'
      if (d.fsurf) {                                                                     // 4'
        if (rr == nn - 2) { wz0 = lz; if (cwin) lz = 0.f; }
        if (rr == nn - 1) { wz1 = lz; wx1 = lx; if (cwin) { lz = 0.f; lx = 0.f; } }
        if (rr == nn) {
          const float wx1m = shfl(wx1, lane - 1);
          if (cwin) lz = lz + (((wz1 + wz0) - 2.0f * wx1) + (jm1 ? 2.0f * wx1m : 0.0f));
        }
        if (rr == nn + 1 && cwin) lx = lx + wx1;
      }
      float ea = 0.f, eb = 0.f, ec = 0.f, ed = 0.f;
      if (rr >= nn && rr <= zw1 && cwin) {
        ea = lx * ((dt * R.b[0]) / dx);
        eb = lx * ((dt * R.b[0]) / dz);
        ec = lz * ((dt * R.b[1]) / dx);
        ed = lz * ((dt * R.b[1]) / dz);
      }
      push(vx, in_grid ? R.fv[0] : 0.f);
      push(vz, in_grid ? R.fv[1] : 0.f);
      push(ux, lx);
      push(uz, lz);
      push(e1, ea);
      push(e2, eb);
      push(e3, ec);
      push(e4, ed);
    };
    for (int rr = z0 - 6; rr <= z0 - 1; ++rr) {
      VRow R;
      vload(rr, R);
      vpush(rr, R);
    }
    struct SRow {
      float ls[3], cm[ANISO ? 6 : 4], fs[3];
    };
    auto sload = [&](int k, SRow &R) {
      const unsigned o = rb(k);
      R.ls[0] = bld(rLi, vc, F(TXX) + o); R.ls[1] = bld(rLi, vc, F(TZZ) + o); R.ls[2] = bld(rLi, vc, F(TXZ) + o);
      R.cm[0] = bld(rC, vc, C11 * planeb + o); R.cm[1] = bld(rC, vc, C13 * planeb + o);
      R.cm[2] = bld(rC, vc, C33 * planeb + o); R.cm[3] = bld(rC, vc, C55 * planeb + o);
      if constexpr (ANISO) {
        R.cm[4] = bld(rC, vc, C15 * planeb + o); R.cm[5] = bld(rC, vc, C35 * planeb + o);
      }
      R.fs[0] = bld(rFs, vc, F(TXX) + o); R.fs[1] = bld(rFs, vc, F(TZZ) + o); R.fs[2] = bld(rFs, vc, F(TXZ) + o);
    };
    VRow vcur;
    SRow scur;
    vload(z0, vcur);
    sload(z0 - 3, scur);
#pragma unroll 1
    for (int kk = 0; kk < nk; ++kk) {
      const int k = z0 - 3 + kk;
      VRow vnxt;
      SRow snxt;
      vload(k + 4, vnxt);
      sload(k + 1, snxt);
      vpush(k + 3, vcur);
      // 3'. stress adjoints of row k from e (e1/e3 row k = entry 0; e2/e4 row k = entry 3)
      float lt_xx = D.dxfT(e1[0], lane);
      float lt_xz = D.dzbT(e2, 3) + D.dxbT(e3[0], lane);
      float lt_zz = D.dzfT(e4, 3);
      // 2'. stress free surface transposed
      if (d.fsurf) {
        if (k == nn - 2) ltxz0 = lt_xz;
        if (k == nn - 1) { ltzz1 = lt_zz; ltxz1 = lt_xz; }
        if (k == nn) lt_xz = lt_xz - ltxz1;
        if (k == nn + 1) { lt_zz = lt_zz - ltzz1; lt_xz = lt_xz - ltxz0; }
      }
      const float dxb_vx = D.dxb(vx[3], lane), dxf_vz = D.dxf(vz[3], lane);
      float dxb_vz = 0.f, dxf_vx = 0.f;
      if constexpr (ANISO) {
        dxb_vz = D.dxb(vz[3], lane);
        dxf_vx = D.dxf(vx[3], lane);
      }
      // 1'. stress row k
      float P1 = 0.f, Q2 = 0.f, Q3 = 0.f, P4 = 0.f, P2 = 0.f, P3 = 0.f, Q1 = 0.f, Q4 = 0.f;
      if (k >= nn && k <= zw1 && cwin) {
        const unsigned o = rb(k);
        const float lxx = scur.ls[0] + lt_xx;
        const float lzz = (d.fsurf && k == nn) ? 0.0f : scur.ls[1] + lt_zz;
        const float lxz = scur.ls[2] + lt_xz;
        const bool own_k = out_col && k >= z0 && k < end_row;
        if (own_k && k == zs && c == xs && gsrc) {
          const float gm = (mxx * lxx + mzz * lzz) + mxz * lxz;
          if (d.src_fz) gmt = gm;        // the force term of row zs follows three rows later
          else gsrc[(long long)s * d.nt + t] = gm;
        }
        if (own_k) {
          bst(lxx, rLo, vs, F(TXX) + o); bst(lzz, rLo, vs, F(TZZ) + o); bst(lxz, rLo, vs, F(TXZ) + o);
        }
        const float axx_x = dtdx * lxx, axx_z = dtdz * lxx;
        const float azz_x = dtdx * lzz, azz_z = dtdz * lzz;
        const float axz_x = dtdx * lxz, axz_z = dtdz * lxz;
        const float C11v = scur.cm[0], C13v = scur.cm[1], C33v = scur.cm[2], C55v = scur.cm[3];
        P1 = C11v * axx_x + C13v * azz_x;
        Q2 = C13v * axx_z + C33v * azz_z;
        Q3 = C55v * axz_x;
        P4 = C55v * axz_z;
        if constexpr (ANISO) {
          const float C15v = scur.cm[4], C35v = scur.cm[5];
          P2 = C15v * axx_z + C35v * azz_z;
          P3 = C15v * axz_x;
          Q1 = C15v * axx_x + C35v * azz_x;
          Q4 = C35v * axz_z;
        }
        if (own_k) {
          const float dzb_vz = D.dzb(vz, 3), dzf_vx = D.dzf(vx, 3);
          if constexpr (ANISO) {
            const float dzb_vx = D.dzb(vx, 3), dzf_vz = D.dzf(vz, 3);
            acc(6, k, axx_x * dxb_vz + axx_z * dzb_vx + axz_x * dxf_vx);
            acc(7, k, azz_x * dxb_vz + azz_z * dzb_vx + axz_z * dzf_vz);
          }
          acc(0, k, axx_x * dxb_vx);
          acc(1, k, axx_z * dzb_vz + azz_x * dxb_vx);
          acc(2, k, azz_z * dzb_vz);
          acc(3, k, axz_x * dxf_vz + axz_z * dzf_vx);
        }
      }
      push(p1, P1);
      push(q3, Q3);
      push(p4, P4);
      push(q2, Q2);
      if constexpr (ANISO) {
        push(p3, P3);
        push(q1, Q1);
        push(p2, P2);
        push(q4, Q4);
      }
      // forward stresses of row k after the step (stored with the free surface applied)
      {
        const bool in_grid = k >= 0 && k <= zmax;
        push(xx, in_grid ? scur.fs[0] : 0.f);
        push(zz, in_grid ? scur.fs[1] : 0.f);
        push(xz, in_grid ? scur.fs[2] : 0.f);
      }
      // velocity adjoints of row r = k-3: L1(v) + D^T(products)
      const int r = k - 3;
      float lvx = D.dxbT(p1[0], lane) + D.dzfT(p4, 3);
      float lvz = D.dzbT(q2, 3) + D.dxfT(q3[0], lane);
      if constexpr (ANISO) {
        lvx = lvx + (D.dzbT(p2, 3) + D.dxfT(p3[0], lane));
        lvz = lvz + (D.dxbT(q1[0], lane) + D.dzfT(q4, 3));
      }
      const float dxf_txx = D.dxf(xx[0], lane), dxb_txz = D.dxb(xz[3], lane);
      const bool own_r = out_col && ((r >= z0 && r < end_row) || (first && r >= nn - 2 && r < nn));
      if (own_r) {
        const unsigned o = rb(r);
        if (d.src_fz && gsrc && r == zs && c == xs)   // vertical force entered vz before the taper
          gsrc[(long long)s * d.nt + t] = gmt + d.src_fz[s] * uz[0];
        bst(ux[0] + lvx, rLo, vs, F(VX) + o);
        bst(uz[0] + lvz, rLo, vs, F(VZ) + o);
        if (r >= z0) {
          const float dzb_txz = D.dzb(xz, 3), dzf_tzz = D.dzf(zz, 3);
          acc(4, r, ux[0] * (dt * (dxf_txx / dx + dzb_txz / dz)));
          acc(5, r, uz[0] * (dt * (dxb_txz / dx + dzf_tzz / dz)));
        }
      }
      vcur = vnxt;
      scur = snxt;
    }
  }
  __syncthreads();
  if (!out_col) return;
  const long long planef = (long long)d.Z * d.pitch;
  float *g = gacc + (long long)blockIdx.y * NCOEF * planef + c;
  const float *sb = slab + lane;
  for (int f = wv; f < NGF; f += EWPB)
    for (int row = z0; row < end_row; ++row) {
      const int o = (f * ETZA + (row - z0)) * WAVE;
      const float v = (sb[o] + sb[NGF * ETZA * WAVE + o]) + (sb[2 * NGF * ETZA * WAVE + o] + sb[3 * NGF * ETZA * WAVE + o]);
      float *p = g + (long long)f * planef + (long long)row * d.pitch;
      *p = *p + v;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
int el_layout(const adfwi_elastic_desc *d, adfwi_layout *L) {
  if (!d) { adfwi::set_error("null descriptor"); return ADFWI_E_ARG; }
  if (d->abi_version != ADFWI_ABI_VERSION || d->struct_size != (int32_t)sizeof(adfwi_elastic_desc)) {
    adfwi::set_error("elastic descriptor version %d size %d, library expects %d / %d", d->abi_version,
                     d->struct_size, ADFWI_ABI_VERSION, (int)sizeof(adfwi_elastic_desc));
    return ADFWI_E_VERSION;
  }
  ADFWI_CHECK_ARG(d->nshots >= 1, "nshots must be >= 1 (got %d)", d->nshots);
  const int nn = d->nn == 0 ? 2 : d->nn;
  ADFWI_CHECK_ARG(nn >= 2 && nn <= 5, "stencil half-order nn %d outside 2..5 (fd_order 4..10)", d->nn);
  ADFWI_CHECK_ARG(nn == 2 || d->tile_z == 0 || d->tile_z >= 8,
                  "tile_z %d: the three-term scheme needs strips of at least 8 rows", d->tile_z);
  ADFWI_CHECK_ARG(d->nz >= 2 * nn + 5 && d->nx >= 2 * nn + 5, "padded grid %d x %d too small (need >= %d x %d)",
                  d->nz, d->nx, 2 * nn + 5, 2 * nn + 5);
  ADFWI_CHECK_ARG(d->pitch >= d->nx, "pitch %d < nx %d", d->pitch, d->nx);
  ADFWI_CHECK_ARG(d->nt >= 1 && d->nrec >= 0, "nt %d / nrec %d invalid", d->nt, d->nrec);
  ADFWI_CHECK_ARG(d->shots_per_wave >= 0 && (d->tile_z == 0 || (d->tile_z >= 4 && d->tile_z <= ETZ_MAX)),
                  "shots_per_wave %d / tile_z %d invalid (tile_z: 0 or 4..%d)", d->shots_per_wave, d->tile_z,
                  ETZ_MAX);
  const int rows = d->nz - 2 * nn;   // window rows [nn, Z-nn-1]
  L->zbase = nn;
  L->tile_x = nn == 2 ? ESX : ESX3;
  L->ntiles_x = (d->nx - 2 * nn + L->tile_x - 1) / L->tile_x;
  // strips of 16 rows (shorter for small grids); shots per wavefront so that there are
  // about 4096 wavefronts, as few shot groups as that allows (fewer gradient planes)
  int tz = d->tile_z;
  if (tz == 0) {
    const int tz_min = nn == 2 ? 4 : 8;
    tz = tz_min;
    for (int cand = 16; cand >= tz_min; cand >>= 1)
      if ((long long)L->ntiles_x * ((rows + cand - 1) / cand) * d->nshots >= 4096) { tz = cand; break; }
    // the 4th-order PML forward re-reads fewer halo rows in 32-row strips, which pays (measured
    // 5-8 % per step on 242 x 680 / 32 shots and 362 x 1040 / 48 shots) while 2048 strip-shots
    // still fill the chip; taller strips or the other schemes measured slower
    if (nn == 2 && !d->damping && (long long)L->ntiles_x * ((rows + 31) / 32) * d->nshots >= 2048) tz = 32;
  }
  L->tile_z = tz;
  L->ntiles_z = (rows + tz - 1) / tz;
  L->ntiles = L->ntiles_z * L->ntiles_x;
  int spw = d->shots_per_wave;
  if (spw == 0) {
    const long long strips = L->ntiles;
    spw = 1;
    while (spw < d->nshots && strips * EWPB * ((d->nshots + EWPB * spw * 2 - 1) / (EWPB * spw * 2)) >= 4096)
      spw *= 2;
  }
  L->shots_per_block = spw * EWPB;
  L->ngroups = (d->nshots + L->shots_per_block - 1) / L->shots_per_block;
  ADFWI_CHECK_ARG(d->damping == 0 || d->damping == 1, "damping must be 0 (PML) or 1 (got %d)", d->damping);
  const int nf = d->damping ? NFIELD5 : NFIELD;
  L->nfields = nf;
  L->reserved = spw;
  L->field_floats = (int64_t)d->nshots * d->nz * d->pitch;
  ADFWI_CHECK_ARG(L->field_floats * 4 * nf < ((int64_t)1 << 32) - 1,
                  "state of %d shots x %d x %d exceeds the 4 GiB of one buffer descriptor; "
                  "advance the shots in smaller launches", d->nshots, d->nz, d->pitch);
  // the adjoint's accesses to a workgroup's gradient planes point the lanes that own no point
  // past the planes (dropped buffer accesses), which needs the planes under 2 GiB
  ADFWI_CHECK_ARG((int64_t)NCOEF * d->nz * d->pitch * 4 < ((int64_t)1 << 31),
                  "grid of %d x %d (pitch %d) exceeds the 2 GiB of one set of gradient planes",
                  d->nz, d->nx, d->pitch);
  L->state_floats = nf * L->field_floats;
  L->accum_floats = (int64_t)L->ngroups * NCOEF * d->nz * d->pitch;
  return ADFWI_OK;
}

ElDev el_dev(const adfwi_elastic_desc *d, const adfwi_layout &L) {
  ElDev v;
  v.S = d->nshots; v.Z = d->nz; v.X = d->nx; v.pitch = d->pitch; v.fsurf = d->free_surface ? 1 : 0;
  v.nt = d->nt; v.nrec = d->nrec; v.spw = L.reserved; v.tz = L.tile_z; v.nsx = L.ntiles_x;
  v.nstrips = L.ntiles;
  v.dt = d->dt; v.dt_dx = d->dt_dx; v.dt_dz = d->dt_dz; v.dx = d->dx; v.dz = d->dz; v.c0 = d->c0; v.c1 = d->c1;
  v.nn = d->nn == 0 ? 2 : d->nn;
  v.c2 = v.nn > 2 ? d->c2 : 0.0f;
  v.plane = (long long)d->nz * d->pitch;
  v.field = (long long)d->nshots * v.plane;
  v.coef = d->coef; v.pml_x = d->pml_x; v.pml_z = d->pml_z; v.src = d->src; v.src_m = d->src_m;
  v.src_fz = d->src_fz;
  v.damp = d->damp;
  v.src_z = d->src_z; v.src_x = d->src_x; v.rcv_z = d->rcv_z; v.rcv_x = d->rcv_x;
  return v;
}

}  // namespace

extern "C" {

int adfwi_elastic_get_layout(const adfwi_elastic_desc *d, adfwi_layout *out) {
  ADFWI_CHECK_ARG(out != nullptr, "null layout pointer");
  return el_layout(d, out);
}

int adfwi_elastic_forward(const adfwi_elastic_desc *d, adfwi_elastic_fwd_args *a, adfwi_stream_t stream) {
  if (int rc = adfwi::fault_injected("adfwi_elastic_forward")) return rc;
  adfwi_layout L;
  int rc = el_layout(d, &L);
  if (rc) return rc;
  ADFWI_CHECK_ARG(a != nullptr, "null forward args");
  ADFWI_CHECK_ARG(d->coef && d->src && d->src_m && d->src_z && d->src_x, "null coefficient/source pointer");
  ADFWI_CHECK_ARG(d->damping ? d->damp != nullptr : (d->pml_x && d->pml_z),
                  d->damping ? "null damping plane" : "null PML factor table");
  const bool recs = a->rec[0] || a->rec[1] || a->rec[2] || a->rec[3] || a->rec[4];
  ADFWI_CHECK_ARG(!recs || (a->rec[0] && a->rec[1] && a->rec[2] && a->rec[3] && a->rec[4]),
                  "records must be all set or all NULL");
  if (recs && d->nrec > 0) ADFWI_CHECK_ARG(d->rcv_z && d->rcv_x, "null receiver table pointer");
  ADFWI_CHECK_ARG(a->state_in && (a->trace || (a->work[0] && a->work[1] && a->work[0] != a->work[1])),
                  "state_in and two distinct work buffers (or a trace buffer) are required");
  ADFWI_CHECK_ARG(!(a->trace && a->ckpt), "trace and ckpt are exclusive");
  ADFWI_CHECK_ARG(a->t0 >= 0 && a->nsteps >= 0 && a->t0 + a->nsteps <= d->nt,
                  "steps [%d, %d) outside [0, %d)", a->t0, a->t0 + a->nsteps, d->nt);
  ADFWI_CHECK_ARG(!a->ckpt || a->ckpt_every >= 1, "ckpt_every must be >= 1 with a checkpoint buffer");
  ElDev v = el_dev(d, L);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const dim3 grid(L.ntiles, L.ngroups), block(ENT);
  const long long ng = (long long)d->nshots * d->nrec;
  const dim3 ggrid((unsigned)((ng + 255) / 256)), gblock(256);
  // the 4th-order PML step takes its records itself when the receiver map is given
  const bool fused = recs && ng > 0 && d->rcv_map && !d->damping && v.nn == 2;
  ElRec rk{};
  if (fused) {
    rk.map = d->rcv_map;
    for (int q = 0; q < 5; ++q) rk.r[q] = a->rec[q];
  }
  const float *cur = a->state_in;
  for (int k = 0; k < a->nsteps; ++k) {
    const int t = a->t0 + k;
    float *nxt;
    if (a->trace) {
      nxt = a->trace + (long long)k * L.state_floats;
    } else if (a->ckpt && (t + 1) % a->ckpt_every == 0 && t + 1 < d->nt) {
      nxt = a->ckpt + (long long)((t + 1) / a->ckpt_every - 1) * L.state_floats;
    } else {
      nxt = (cur == a->work[0]) ? a->work[1] : a->work[0];
    }
    if (d->damping && v.nn > 2) {
      if (d->aniso)
        hipLaunchKernelGGL(el3_abl_fwd<true>, grid, block, 0, st, v, cur, nxt, t);
      else
        hipLaunchKernelGGL(el3_abl_fwd<false>, grid, block, 0, st, v, cur, nxt, t);
      if ((rc = adfwi::check_launch("el3_abl_fwd"))) return rc;
    } else if (d->damping) {
      if (d->aniso)
        hipLaunchKernelGGL(el_abl_fwd<true>, grid, block, 0, st, v, cur, nxt, t);
      else
        hipLaunchKernelGGL(el_abl_fwd<false>, grid, block, 0, st, v, cur, nxt, t);
      if ((rc = adfwi::check_launch("el_abl_fwd"))) return rc;
    } else if (v.nn > 2) {
      if (d->aniso)
        hipLaunchKernelGGL(el3_fwd_step<true>, grid, block, 0, st, v, cur, nxt, t);
      else
        hipLaunchKernelGGL(el3_fwd_step<false>, grid, block, 0, st, v, cur, nxt, t);
      if ((rc = adfwi::check_launch("el3_fwd_step"))) return rc;
    } else {
      if (d->aniso && fused)
        hipLaunchKernelGGL((el_fwd_step<true, true>), grid, block, 0, st, v, cur, nxt, t, rk);
      else if (d->aniso)
        hipLaunchKernelGGL((el_fwd_step<true, false>), grid, block, 0, st, v, cur, nxt, t, rk);
      else if (fused)
        hipLaunchKernelGGL((el_fwd_step<false, true>), grid, block, 0, st, v, cur, nxt, t, rk);
      else
        hipLaunchKernelGGL((el_fwd_step<false, false>), grid, block, 0, st, v, cur, nxt, t, rk);
      if ((rc = adfwi::check_launch("el_fwd_step"))) return rc;
    }
    if (recs && ng > 0 && !fused) {
      if (d->damping)
        hipLaunchKernelGGL(el_gather5, ggrid, gblock, 0, st, v, nxt, a->rec[0], a->rec[1], a->rec[2], a->rec[3],
                           a->rec[4], t);
      else
        hipLaunchKernelGGL(el_gather, ggrid, gblock, 0, st, v, nxt, a->rec[0], a->rec[1], a->rec[2], a->rec[3],
                           a->rec[4], t);
      if ((rc = adfwi::check_launch("el_gather"))) return rc;
    }
    cur = nxt;
  }
  a->state_out = cur;
  return ADFWI_OK;
}

int adfwi_elastic_adjoint(const adfwi_elastic_desc *d, adfwi_elastic_adj_args *a, adfwi_stream_t stream) {
  if (int rc = adfwi::fault_injected("adfwi_elastic_adjoint")) return rc;
  adfwi_layout L;
  int rc = el_layout(d, &L);
  if (rc) return rc;
  ADFWI_CHECK_ARG(a != nullptr, "null adjoint args");
  ADFWI_CHECK_ARG(d->coef && d->src_m && d->src_z && d->src_x, "null coefficient/source pointer");
  ADFWI_CHECK_ARG(d->damping ? d->damp != nullptr : (d->pml_x && d->pml_z),
                  d->damping ? "null damping plane" : "null PML factor table");
  ADFWI_CHECK_ARG(a->work[0] && a->work[1] && a->work[0] != a->work[1], "two distinct work buffers are required");
  ADFWI_CHECK_ARG(a->lam_in == a->work[0] || a->lam_in == a->work[1], "lam_in must be work[0] or work[1]");
  ADFWI_CHECK_ARG(a->fwd0 && a->trace && a->grad_coef, "null forward state, trace or gradient buffer");
  ADFWI_CHECK_ARG(a->nsteps >= 0 && a->t_hi - a->nsteps >= 0 && a->t_hi <= d->nt,
                  "steps [%d, %d) outside [0, %d)", a->t_hi - a->nsteps, a->t_hi, d->nt);
  const bool recs = a->grad_rec[0] || a->grad_rec[1] || a->grad_rec[2] || a->grad_rec[3] || a->grad_rec[4];
  if (recs && d->nrec > 0) ADFWI_CHECK_ARG(d->rcv_z && d->rcv_x, "null receiver table pointer");
  ElDev v = el_dev(d, L);
  // the free-surface transposes need rows 0..3 inside the first strip (every strip height here
  // is at least 8)
  // (damping scheme: strips of ETZA rows, so that no strip but the first reaches the
  // free-surface rows, whose velocity adjoints are state there)
  // (three-term scheme: strips of ETZA rows, so that only the first strip reaches the
  // free-surface rows NN-2 .. NN+1)
  // The 4th-order PML adjoint keeps no per-strip storage: the fewest strips per column (equal
  // heights, multiples of 4 rows, at most 64) that still give 2048 wavefronts; taller strips
  // re-read fewer halo rows.  Measured on 242 x 680 / 32 shots and 362 x 1040 / 48 shots, this
  // beats both 8-row strips and one round of 2048 wavefronts of the tallest strips.
  // ADFWI_ADJ_TILE_Z sets the height.
  const int rows = d->nz - 2 * v.nn;
  if (d->damping || v.nn > 2) {
    v.tz = ETZA;
  } else {
    const char *env = std::getenv("ADFWI_ADJ_TILE_Z");
    int tz = env && *env ? std::atoi(env) : 0;
    if (tz < 8 || tz > ETZ_MAX) {
      tz = 8;
      for (int n = 1; (rows + n - 1) / n >= 8; ++n) {
        const int cand = ((rows + n - 1) / n + 3) / 4 * 4;
        if (cand <= ETZ_MAX && (long long)n * v.nsx * L.ngroups * EWPB >= 2048) { tz = cand; break; }
      }
    }
    v.tz = tz;
  }
  v.nstrips = ((rows + v.tz - 1) / v.tz) * v.nsx;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const dim3 grid(v.nstrips, L.ngroups), block(ENT);
  const long long ng = (long long)d->nshots * d->nrec;
  const dim3 ggrid((unsigned)((ng + 255) / 256)), gblock(256);
  const int t_lo = a->t_hi - a->nsteps;
  // the 4th-order PML adjoint step adds the record adjoints itself when the receiver map is given
  const bool fused = recs && ng > 0 && d->rcv_map && !d->damping && v.nn == 2;
  ElInj ij{};
  if (fused) {
    ij.map = d->rcv_map;
    for (int q = 0; q < 5; ++q) ij.g[q] = a->grad_rec[q];
  }
  float *cur = a->lam_in;
  for (int i = a->nsteps - 1; i >= 0; --i) {
    const int t = t_lo + i;
    if (recs && ng > 0 && !fused) {
      if (d->damping)
        hipLaunchKernelGGL(el_inject5, ggrid, gblock, 0, st, v, cur, a->grad_rec[0], a->grad_rec[1],
                           a->grad_rec[2], a->grad_rec[3], a->grad_rec[4], t);
      else
        hipLaunchKernelGGL(el_inject, ggrid, gblock, 0, st, v, cur, a->grad_rec[0], a->grad_rec[1],
                           a->grad_rec[2], a->grad_rec[3], a->grad_rec[4], t);
      if ((rc = adfwi::check_launch("el_inject"))) return rc;
    }
    const float *fv = i == 0 ? a->fwd0 : a->trace + (long long)(i - 1) * L.state_floats;
    const float *fsn = a->trace + (long long)i * L.state_floats;
    float *nxt = (cur == a->work[0]) ? a->work[1] : a->work[0];
    if (d->damping && v.nn > 2) {
      if (d->aniso)
        hipLaunchKernelGGL(el3_abl_adj<true>, grid, block, 0, st, v, cur, nxt, fv, fsn, t, a->grad_coef, a->grad_src);
      else
        hipLaunchKernelGGL(el3_abl_adj<false>, grid, block, 0, st, v, cur, nxt, fv, fsn, t, a->grad_coef, a->grad_src);
      if ((rc = adfwi::check_launch("el3_abl_adj"))) return rc;
    } else if (d->damping) {
      if (d->aniso)
        hipLaunchKernelGGL(el_abl_adj<true>, grid, block, 0, st, v, cur, nxt, fv, fsn, t, a->grad_coef, a->grad_src);
      else
        hipLaunchKernelGGL(el_abl_adj<false>, grid, block, 0, st, v, cur, nxt, fv, fsn, t, a->grad_coef, a->grad_src);
      if ((rc = adfwi::check_launch("el_abl_adj"))) return rc;
    } else if (v.nn > 2) {
      if (d->aniso)
        hipLaunchKernelGGL(el3_adj_step<true>, grid, block, 0, st, v, cur, nxt, fv, fsn, t, a->grad_coef, a->grad_src);
      else
        hipLaunchKernelGGL(el3_adj_step<false>, grid, block, 0, st, v, cur, nxt, fv, fsn, t, a->grad_coef, a->grad_src);
      if ((rc = adfwi::check_launch("el3_adj_step"))) return rc;
    } else {
      float *gc = a->grad_coef, *gs = a->grad_src;
      if (d->aniso && fused)
        hipLaunchKernelGGL((el_adj_step<true, true>), grid, block, 0, st, v, cur, nxt, fv, fsn, t, gc, gs, ij);
      else if (d->aniso)
        hipLaunchKernelGGL((el_adj_step<true, false>), grid, block, 0, st, v, cur, nxt, fv, fsn, t, gc, gs, ij);
      else if (fused)
        hipLaunchKernelGGL((el_adj_step<false, true>), grid, block, 0, st, v, cur, nxt, fv, fsn, t, gc, gs, ij);
      else
        hipLaunchKernelGGL((el_adj_step<false, false>), grid, block, 0, st, v, cur, nxt, fv, fsn, t, gc, gs, ij);
      if ((rc = adfwi::check_launch("el_adj_step"))) return rc;
    }
    cur = nxt;
  }
  a->lam_out = cur;
  return ADFWI_OK;
}

}  // extern "C"
